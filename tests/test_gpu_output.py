"""The result writers (mpm3d-f90_amd/output.py) on a device context in the
reference-order FP64 mode, against the reference program's own files.

Particle state is the reference's bit for bit, so the TecPlot and ParaView
frames are equal character for character, and so is every curve column but
the energies.  Energies and momentum are the device's deterministic block
sums (the reference sums serially): they agree to the printed 4 digits up to
one unit in the last place where a value sits on a rounding boundary, and
momentum components that cancel to round-off (1e-13 of the momentum's
largest component, where the reference and the device differ by 1e-16) are
compared against the file's largest value, or against the bodies' summed
momentum magnitude where they cancel each other.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import load_package
from test_output import OPTION_SETS, REFPROG, ref_files, with_outputs, write_deck

pytestmark = pytest.mark.gpu


def numeric_rows(text, skip_header):
    rows = text.splitlines()[1 if skip_header else 0:]
    return np.array([[float(t) for t in r.split()] for r in rows])


@pytest.mark.parametrize("name", sorted(OPTION_SETS))
def test_device_writers_match_the_reference_program(name):
    if not os.path.exists(REFPROG):
        pytest.skip("reference program not built (oracle/build_ref.sh needs the reference sources)")
    mpm = load_package()
    import importlib
    out = importlib.import_module("mpm3d_f90_amd.output")
    text = with_outputs(name, OPTION_SETS[name])
    tmp = tempfile.mkdtemp(prefix="outputs_")
    try:
        dref, dours = os.path.join(tmp, "ref"), os.path.join(tmp, "ours")
        write_deck(dref, name, text)
        p = write_deck(dours, name, text)
        r = subprocess.run([REFPROG, name], cwd=dref, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        prob = mpm.read_deck(p)
        with mpm.Context(prob, "fp64exact") as ctx:
            out.run(ctx, out.read_outputs(p), name, dours)
        os.remove(p)
        ref, ours = ref_files(dref), ref_files(dours)
        assert sorted(ref) == sorted(ours)
        numeric = {name + "_curv.dat", "EnergyPlot.dat", "MomentumPlot.dat"}
        for f in ref:
            if f not in numeric:
                assert ref[f] == ours[f], f
        rc, oc = ref[name + "_curv.dat"].splitlines(), ours[name + "_curv.dat"].splitlines()
        assert rc[0] == oc[0] and len(rc) == len(oc)
        energy = [k for k, h in enumerate(rc[0].split()) if h.startswith(("engk", "engi"))]
        for a, b in zip(rc[1:], oc[1:]):
            ta, tb = a.split(), b.split()
            assert [t for k, t in enumerate(ta) if k not in energy] == [t for k, t in enumerate(tb) if k not in energy]
        # bodies moving against each other cancel their momenta: the file's
        # momentum is round-off of sum m |v| (the contact decks)
        pmag = float(np.sum(prob.field("mass") * np.abs(prob.pstate[[mpm.PF["vx"], mpm.PF["vy"], mpm.PF["vz"]]])))
        for f, header in ((name + "_curv.dat", True), ("EnergyPlot.dat", False), ("MomentumPlot.dat", False)):
            a, b = numeric_rows(ref[f], header), numeric_rows(ours[f], header)
            assert a.shape == b.shape
            scale = np.abs(a[:, 1:]).max()      # the file's largest energy or momentum component
            if f == "MomentumPlot.dat":
                scale = max(scale, pmag)
            assert np.all(np.abs(a - b) <= 1e-4 * np.abs(a) + 1e-12 * scale), f
    finally:
        shutil.rmtree(tmp)
