"""Every input deck the reference ships (Data/*.mpm and the
VolumetricLockingEx decks: Taylor impact, slope failure, detonations, lead
hypervelocity impact, Cook membranes, dam breaks, free and clamped soft
beams -- every shape family, scheme, material model, total-Lagrangian and
smoothing decks among them) is read by the deck reader and stepped by the
oracle bit-exactly with the reference build (oracle/_ref/ref_driver): every
particle field and the grid after 1 and 3 steps.

Runs only where the reference and its build are present (this container);
the decks are read in place and not copied into the repository.
"""
import glob
import os
import shutil
import tempfile

import numpy as np
import pytest

from conftest import load_package

REF = "/root/reference"
SHIPPED = sorted(glob.glob(os.path.join(REF, "Data", "*.mpm")) +
                 glob.glob(os.path.join(REF, "VolumetricLockingEx", "**", "*.mpm"), recursive=True))


@pytest.mark.skipif(not SHIPPED, reason="the reference's decks are not on this machine")
@pytest.mark.parametrize("path", SHIPPED, ids=[os.path.relpath(p, REF) for p in SHIPPED])
def test_shipped_deck_oracle_matches_reference(oracle_mod, path):
    if not os.path.exists(oracle_mod.REF_DRIVER):
        pytest.skip("reference build oracle/_ref not present on this machine")
    mpm = load_package()
    name = os.path.splitext(os.path.basename(path))[0]
    tmp = tempfile.mkdtemp(prefix="shipped_")
    try:
        shutil.copy(path, tmp)
        oracle_mod.run_ref(os.path.join(tmp, name + ".mpm"), 3, os.path.join(tmp, "d"), [1, 3], cwd=tmp)
        prob = mpm.read_deck(path)
        o = oracle_mod.Oracle(prob)
        done = 0
        for s in (1, 3):
            o.step(s - done)
            done = s
            d = oracle_mod.read_ref_dump(os.path.join(tmp, "d_s%05d.bin" % s))
            ps = o.particles()
            for f, v in d["particles"].items():
                assert np.array_equal(ps[mpm.PF[f]], v, equal_nan=True), (s, f)
            assert np.array_equal(o.grid(), d["grid"], equal_nan=True), s
            assert np.array_equal(o.icell(), d["icell"]), s
    finally:
        shutil.rmtree(tmp)
