"""mpm_cell_order: InWhichCell of caller-owned device particle arrays and
the stable permutation by cell, through the C-ABI.

Expected values: icell from the oracle's ParticleInitial (the reference's
InWhichCell, bit-exact to it) on the same positions, including points
outside the grid and points on its upper faces (where InWhichCell wraps the
linear index); the order from numpy's stable argsort of the cell key
(outside the grid last)."""
import ctypes as C

import numpy as np
import pytest

from conftest import ALL_DECKS, deck_path, load_package

pytestmark = pytest.mark.gpu


def positions(prob, n, seed):
    """n points: most inside the grid, some outside, some on the upper faces"""
    rng = np.random.default_rng(seed)
    sp = np.array(prob.struct().span[:6]).reshape(3, 2)
    lo, hi = sp[:, 0], sp[:, 1]
    w = hi - lo
    X = lo + rng.uniform(-0.1, 1.1, (n, 3)) * w
    k = rng.choice(n, n // 20, replace=False)
    X[k, rng.integers(0, 3, len(k))] = hi[rng.integers(0, 3, len(k))]   # some exactly on an upper face
    X[k[: len(k) // 2]] = np.minimum(X[k[: len(k) // 2]], hi)
    return np.ascontiguousarray(X.T)


def oracle_icell(oracle_mod, mpm, prob, X):
    """icell of the reference's ParticleInitial for the points X (3, n), n <= prob.n"""
    o = oracle_mod.Oracle(prob)
    ps = o.particles()
    n = X.shape[1]
    ps[mpm.PF["x"], :n], ps[mpm.PF["y"], :n], ps[mpm.PF["z"], :n] = X
    o.set_particles(ps)
    o.phase("begin_step")
    o.phase("grid_momentum_initial")
    return o.icell()[:n]


def cell_order(mpm, prob, X, fn="mpm_cell_order"):
    from hipmem import DeviceArray
    n = X.shape[1]
    x, y, z = (DeviceArray(X[d]) for d in range(3))
    icell = DeviceArray(np.full(n, -7, np.int32))
    order = DeviceArray(np.full(n, -7, np.int32))
    s = prob.struct()
    rc = getattr(mpm._lib.lib(), fn)(C.byref(s), x.ptr, y.ptr, z.ptr, n, icell.ptr, order.ptr, None)
    mpm.check(rc, fn)
    return icell.get(), order.get()


def expected_order(prob, icell):
    s = prob.struct()
    key = np.where(icell < 0, s.ncx * s.ncy * s.ncz + 1, icell).astype(np.int64)
    return np.argsort(key, kind="stable").astype(np.int32)


@pytest.mark.parametrize("name", ALL_DECKS)
def test_cell_order_matches_particle_initial(name, oracle_mod):
    mpm = load_package()
    prob = mpm.read_deck(deck_path(name))
    X = positions(prob, prob.n, seed=len(name))
    ref = oracle_icell(oracle_mod, mpm, prob, X)
    icell, order = cell_order(mpm, prob, X)
    assert np.array_equal(icell, ref)
    assert np.array_equal(order, expected_order(prob, ref))


def test_cell_order_at_large_sizes():
    """Several sizes across the sort's tile and pass boundaries, against a
    restatement of InWhichCell in numpy (pinned by the test above)."""
    mpm = load_package()
    prob = mpm.read_deck(deck_path("slope_dp_bspline_usf"))
    s = prob.struct()
    sp = np.array(s.span[:6]).reshape(3, 2)
    for n in (1, 4095, 4096, 4097, 300001, 2000003):
        X = positions(prob, n, seed=n)
        inside = np.all((X >= sp[:, :1]) & (X <= sp[:, 1:]), axis=0)
        ijk = ((X - sp[:, :1]) / s.dcell).astype(np.int64) + 1
        c = (ijk[2] - 1) * s.ncx * s.ncy + (ijk[1] - 1) * s.ncx + ijk[0]
        ref = np.where(inside & (c >= 1) & (c <= s.ncx * s.ncy * s.ncz), c, -1).astype(np.int32)
        icell, order = cell_order(mpm, prob, X)
        assert np.array_equal(icell, ref), n
        assert np.array_equal(order, expected_order(prob, ref)), n
