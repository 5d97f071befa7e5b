"""The Fortran side of the boundary on the device.

* mpm3d-f90_amd/fortran/mpm3d_host: a Fortran host calling the C-ABI through
  ISO_C_BINDING in the reference driver's call order;
* oracle/_ref/mpm3d_dropin: the reference's own main program and input phase
  (built from the reference sources in this container) linked with
  update_step_hip.f90 in place of update_step.o -- the drop-in case.
Both are compared with the CPU oracle after the same number of steps.
"""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import PGROUPS, ROOT, deck_path, group_errors, load_package

pytestmark = pytest.mark.gpu

HOST = os.path.join(ROOT, "mpm3d-f90_amd", "fortran", "mpm3d_host")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "mpm3d_dropin")


def read_dump(path, n_expected):
    raw = open(path, "rb").read()
    n, istep = np.frombuffer(raw, dtype="<i4", count=2)
    sc = np.frombuffer(raw, dtype="<f8", count=8, offset=8)
    ps = np.frombuffer(raw, dtype="<f8", offset=72).reshape(-1, int(n))
    assert n == n_expected
    return int(istep), sc, ps


@pytest.mark.parametrize("name", ["bar_j2_gimp_musl", "slope_dp_bspline_usf", "plate_neoh_musl_qulo",
                                  "plate_neoh_sgmp_musl", "impact_jcf_eos2_gimp_musl",
                                  "blocks_tl_contact_bspline_musl"])
def test_fortran_host_matches_oracle(oracle_mod, name):
    if not os.path.exists(HOST):
        pytest.skip("Fortran host not built")
    mpm = load_package()
    tmp = tempfile.mkdtemp(prefix="fhost_")
    try:
        out = os.path.join(tmp, "dump.bin")
        r = subprocess.run([HOST, deck_path(name), "10", "fp64", out], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "Time per particle cycle" in r.stdout
        prob = mpm.read_deck(deck_path(name))
        istep, sc, ps = read_dump(out, prob.n)
        assert istep == 10
        o = oracle_mod.Oracle(prob)
        o.step(10)
        errs = group_errors(ps, o.particles(), mpm.PF_NAMES, PGROUPS)
        assert max(errs.values()) <= 1e-12, errs
        assert abs(sc[2] / o.scalars()[2] - 1) < 1e-13
    finally:
        shutil.rmtree(tmp)


@pytest.mark.parametrize("name", ["bar_j2_gimp_musl", "slope_dp_bspline_usf", "dam_null_sgmp_musl",
                                  "block_j2_sgmp_usf", "slope_smoo_bspline_musl", "block_smoo_sgmp_usf",
                                  "blocks_contact_bspline_musl", "beam_tl_linear_musl", "block_tl_sgmp_musl",
                                  "impact_jcf_eos2_gimp_musl", "block_john_sgmp_usf",
                                  "blocks_tl_contact_bspline_musl"])
def test_reference_driver_runs_the_device_step(oracle_mod, name):
    """The reference's MPM3D.f90 loop, unchanged, over update_step_hip."""
    if not os.path.exists(DROPIN):
        pytest.skip("drop-in executable not built (needs the reference sources at build time)")
    mpm = load_package()
    prob = mpm.read_deck(deck_path(name))
    tmp = tempfile.mkdtemp(prefix="dropin_")
    try:
        text = open(deck_path(name)).read()
        text = re.sub(r"(?mi)^endtime .*$", "endtime %.9e" % (6.5 * prob.dt), text)
        with open(os.path.join(tmp, name + ".mpm"), "w") as f:
            f.write(text)
        env = dict(os.environ, MPM3D_DUMP=os.path.join(tmp, "dump.bin"))
        r = subprocess.run([DROPIN, name], cwd=tmp, env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        assert "Time per particle cycle" in r.stdout
        istep, sc, ps = read_dump(os.path.join(tmp, "dump.bin"), prob.n)
        assert 5 <= istep <= 8
        o = oracle_mod.Oracle(prob)
        o.step(istep)
        groups = dict(PGROUPS)
        if prob.mats[0].type not in (2, 3):
            # Initial (DataIn.f90:1241) copies Yield0 into sig_y for every
            # material, but only pla1/pla2 set Yield0: elsewhere the reference
            # stores whatever the allocation held (the field is unused there)
            groups.pop("sig_y")
        errs = group_errors(ps, o.particles(), mpm.PF_NAMES, groups)
        assert max(errs.values()) <= 1e-12, errs
    finally:
        shutil.rmtree(tmp)


REFPROG = os.path.join(ROOT, "oracle", "_ref", "mpm3d_ref")


@pytest.mark.parametrize("name", ["bar_j2_gimp_musl", "block_dp_bspline_pull_musl", "dam_null_sgmp_musl",
                                  "cube_smoo_linear_usl", "blocks_contact_gimp_usl", "impact_jcf_eos2_gimp_musl"])
def test_dropin_writes_the_reference_output_files(name):
    """The reference program itself (oracle/_ref/mpm3d_ref: MPM3D.f90 with its
    own update_step.f90, on the CPU) and the drop-in (the same main program
    over the device step in the reference-order mode) on one deck with curve
    and TecPlot output every few steps.  The drop-in downloads only the curve
    particles per step and everything at the output times; its TecPlot file
    is the reference's, character for character, and so is its curve file
    but for the energy columns (the device's block sums against the
    reference's serial sums), which agree to the printed precision.  On the
    smoothing deck the ParaView frames (with the smoothed pressure the driver's
    SmoothStressbyGrid call leaves) are the reference's too, and on the
    contact deck the contact force file."""
    if not (os.path.exists(DROPIN) and os.path.exists(REFPROG)):
        pytest.skip("reference program / drop-in not built (needs the reference sources at build time)")
    mpm = load_package()
    prob = mpm.read_deck(deck_path(name))
    n, dt = prob.n, prob.dt
    ids = [5, n // 3, n - 1, 7]
    extra = ("outt %.9e\ntecp\noutr seqv\noutr velz\noutr epef\ncurv seqv %d\ncurv velz %d\ncurv epef %d\n"
             "curv engk\ncurv volu %d\n" % ((10.5 * dt,) + tuple(ids)))
    if prob.smooth_stress:
        extra += "para\n"
    text = open(deck_path(name)).read()
    text = re.sub(r"(?mi)^endtime .*$", "endtime %.9e" % (40.5 * dt), text)
    text = re.sub(r"(?mi)^endi\s*$", extra + "endi", text)
    out = {}
    tmp = tempfile.mkdtemp(prefix="outputs_")
    try:
        for prog, env in ((REFPROG, {}), (DROPIN, {"MPM3D_PRECISION": "fp64exact"})):
            d = os.path.join(tmp, os.path.basename(prog))
            os.makedirs(d)
            with open(os.path.join(d, name + ".mpm"), "w") as f:
                f.write(text)
            r = subprocess.run([prog, name], cwd=d, env=dict(os.environ, **env), capture_output=True, text=True,
                               timeout=300)
            assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
            out[prog] = {k: open(os.path.join(d, name + k)).read().splitlines() for k in ("_anim.dat", "_curv.dat")}
            out[prog].update({f: open(os.path.join(d, f)).read() for f in os.listdir(d) if f.endswith(".vtu")})
            out[prog]["ContforcPlot.dat"] = open(os.path.join(d, "ContforcPlot.dat")).read()
        ref, dev = out[REFPROG], out[DROPIN]
        assert len(ref["_anim.dat"]) > 4 * n and ref["_anim.dat"] == dev["_anim.dat"]
        vtu = sorted(f for f in ref if f.endswith(".vtu"))
        assert vtu == sorted(f for f in dev if f.endswith(".vtu"))
        assert len(vtu) >= (4 if prob.smooth_stress else 0)
        for f in vtu:
            assert ref[f] == dev[f], f
        assert ref["ContforcPlot.dat"] == dev["ContforcPlot.dat"]
        if prob.contact:
            assert ref["ContforcPlot.dat"].count("\n") >= 40
        rc, dc = ref["_curv.dat"], dev["_curv.dat"]
        assert len(rc) >= 40 and len(rc) == len(dc) and rc[0] == dc[0]
        # energy columns (the reference's two default curves and ours), by header name
        energy = [k for k, h in enumerate(rc[0].split()) if h.startswith(("engk", "engi"))]
        assert len(energy) == 3
        for a, b in zip(rc[1:], dc[1:]):
            ta, tb = a.split(), b.split()
            assert len(ta) == len(tb) == 8
            assert [t for k, t in enumerate(ta) if k not in energy] == [t for k, t in enumerate(tb) if k not in energy]
            for k in energy:
                assert abs(float(ta[k]) - float(tb[k])) <= 1e-4 * abs(float(ta[k])) + 1e-300
    finally:
        shutil.rmtree(tmp)
