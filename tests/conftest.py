"""Shared test setup.

* registers the ``gpu`` marker (tests that need an MI355X);
* loads the package directory ``mpm3d-f90_amd/`` as ``mpm3d_f90_amd``;
* puts ``oracle/`` on sys.path for the CPU oracle (the checker).
"""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "mpm3d-f90_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
DECK_DIR = os.path.join(ROOT, "tests", "golden", "decks")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def load_package():
    if "mpm3d_f90_amd" in sys.modules:
        return sys.modules["mpm3d_f90_amd"]
    spec = importlib.util.spec_from_file_location(
        "mpm3d_f90_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mpm3d_f90_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)

DECKS = ["bar_j2_gimp_musl", "slope_dp_bspline_usf", "cube_elastic_linear_usl",
         "plate_neoh_musl_qulo", "dam_null_eos4_musl",
         # the remaining shape x scheme pairs of the fused kernels
         "cube_elastic_linear_usf", "bar_j2_gimp_usf", "slope_dp_bspline_musl", "slope_dp_bspline_usl",
         # Drucker-Prager blocks that yield on every branch of the return map
         # within 10 steps (sheared and squeezed: USF; pulled apart: MUSL)
         "block_dp_bspline_shear_usf", "block_dp_bspline_pull_musl",
         # the EOS models: simplified Johnson-Cook with the Mie-Gruneisen EOS
         # (Taylor impact) and the linear polynomial EOS (GIMP), with failure
         # (B-spline impact), HE burn with the JWL EOS and a detonation point,
         # and dynamic relaxation damping
         "taylor_sjc_eos2_musl", "cube_sjc_eos1_gimp_usl", "impact_sjcf_eos2_bspline_usf", "tnt_hiex_jwl_usl",
         "cube_drda_linear_usf",
         # SmoothStressbyGrid (smoo) on every shape family; the cube also
         # carries two grid components without contact
         "cube_smoo_linear_usl", "bar_smoo_gimp_usf", "slope_smoo_bspline_musl",
         # particles outside the grid: generated beyond the span, and leaving it
         "blocks_outside_linear_usl"]
# contact (cont lagr): two blocks pressed together on grids of their own
# (linear, USF, frictionless, averaged normal), sliding with friction
# (B-spline, MUSL, normal of body 1, a J2 block), and a third body on a third
# component (GIMP, USL, normal of body 2, gravity); reference-order FP64 only
CONTACT_DECKS = ["blocks_contact_linear_usf", "blocks_contact_bspline_musl", "blocks_contact_gimp_usl",
                 # with stress smoothing, which sums every body on component 1's grid (GIMP, MUSL, friction)
                 "blocks_contact_smoo_gimp_musl"]
# total-Lagrangian MPM (TLmp, neo-Hookean): a beam on a clamped base (linear,
# MUSL, smoothed pressure), two sheared bodies (B-spline USF; GIMP USL with
# gravity and smoothing; SGMP MUSL on a clamped floor); reference-order FP64 only
TL_DECKS = ["beam_tl_linear_musl", "block_tl_bspline_usf", "block_tl_gimp_usl", "block_tl_sgmp_musl"]
# total-Lagrangian MPM with contact: two blocks on grids of their own that
# share a node layer in the initial configuration (B-spline, MUSL, friction)
TL_CONTACT_DECKS = ["blocks_tl_contact_bspline_musl"]
# Johnson-Cook with heating (john: Taylor impact, linear, MUSL; under a block
# dropped on a pla1 pad, SGMP, USF) and with damage and failure (jcf with
# the Mie-Gruneisen EOS, hit by a pla2 cube: GIMP, MUSL); the particles that
# do not yield in a step heat (and damage) with the plastic strain increment
# the last yielding particle before them left; reference-order FP64 only
JC_DECKS = ["taylor_john_linear_musl", "impact_jcf_eos2_gimp_musl", "block_john_sgmp_usf"]
# auxiliary-grid (SGMP) decks: clamped plate (MUSL), water column with
# symmetric walls (MUSL), J2 block on a clamped floor (USF)
SGMP_DECKS = ["plate_neoh_sgmp_musl", "dam_null_sgmp_musl", "block_j2_sgmp_usf", "block_smoo_sgmp_usf"]
ALL_DECKS = DECKS + SGMP_DECKS


def deck_path(name):
    return os.path.join(DECK_DIR, name + ".mpm")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def mpm():
    return load_package()


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.build()
    return pyoracle


def normwise_rel(a, b):
    """max |a - b| / max |b| per field (rows); 0 where both are all zero."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    out = []
    for x, y in zip(a, b):
        ok = np.isfinite(y)
        den = np.max(np.abs(y[ok])) if ok.any() else 0.0
        num = np.max(np.abs(x[ok] - y[ok])) if ok.any() else 0.0
        out.append(0.0 if num == 0 else num / (den if den > 0 else 1e-300))
    return np.array(out)


# Physical field groups for the parity metric: the error of every field is
# measured against the largest magnitude of its group (a stress component that
# is pure round-off in a symmetric problem is judged on the stress scale).
PGROUPS = {
    "x": ["x", "y", "z"],
    "v": ["vx", "vy", "vz"],
    "stress": ["sm", "seqv", "sdxx", "sdyy", "sdzz", "sdyz", "sdxz", "sdxy", "q"],
    "vol": ["vol"],
    "sig_y": ["sig_y"],
    "epeff": ["epeff"],
    "ie": ["ie"],
    "fp": ["fp%d%d" % (i, j) for i in (1, 2, 3) for j in (1, 2, 3)],
    "spre": ["spre"],
    "temp": ["temp"],
}
GGROUPS = {"mg": ["mg"], "p": ["px", "py", "pz"], "f": ["fx", "fy", "fz"]}


def group_errors(a, b, names, groups, unit_scale=("fp",)):
    """dict group -> max |a-b| / max(|b|) over the group's fields (F - I is
    measured against 1, i.e. against F)."""
    idx = {n: i for i, n in enumerate(names)}
    out = {}
    for g, fields in groups.items():
        rows = [idx[f] for f in fields]
        A = a[rows]
        B = b[rows]
        ok = np.isfinite(B) & np.isfinite(A)
        bad_nan = int(np.sum(np.isfinite(B) != np.isfinite(A)))
        den = np.max(np.abs(B[ok])) if ok.any() else 0.0
        if g in unit_scale:
            den = max(den, 1.0)
        num = np.max(np.abs(A[ok] - B[ok])) if ok.any() else 0.0
        e = 0.0 if num == 0 else num / (den if den > 0 else 1e-300)
        out[g] = float("inf") if bad_nan else e
    return out
