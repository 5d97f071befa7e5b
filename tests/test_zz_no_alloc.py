"""mpm_cell_order_noalloc (no device allocation at call time; scratch in a
fixed device array) gives the same cell indices and order as
mpm_cell_order, bit for bit, on two calls in a row, the second with fewer
particles (the shared scratch keeps the first call's contents), and rejects
sizes beyond its maximum."""
import ctypes as C

import numpy as np
import pytest

from conftest import deck_path, load_package

pytestmark = pytest.mark.gpu


def _run(mpm, fn, prob, X):
    from hipmem import DeviceArray
    n = X.shape[1]
    x, y, z = (DeviceArray(X[d]) for d in range(3))
    icell = DeviceArray(np.full(n, -7, np.int32))
    order = DeviceArray(np.full(n, -7, np.int32))
    s = prob.struct()
    rc = getattr(mpm._lib.lib(), fn)(C.byref(s), x.ptr, y.ptr, z.ptr, n, icell.ptr, order.ptr, None)
    mpm.check(rc, fn)
    return icell.get(), order.get()


def _points(prob, n, seed):
    rng = np.random.default_rng(seed)
    sp = np.array(prob.struct().span[:6]).reshape(3, 2)
    return np.ascontiguousarray((sp[:, 0] + rng.uniform(-0.05, 1.05, (n, 3)) * (sp[:, 1] - sp[:, 0])).T)


def test_noalloc_cell_order_equals_the_allocating_one():
    mpm = load_package()
    prob = mpm.read_deck(deck_path("slope_dp_bspline_usf"))
    for n, seed in ((1_500_007, 1), (65_537, 2)):   # second call smaller
        X = _points(prob, n, seed)
        a = _run(mpm, "mpm_cell_order", prob, X)
        b = _run(mpm, "mpm_cell_order_noalloc", prob, X)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), n
        assert np.array_equal(np.sort(b[1]), np.arange(n))   # a permutation


def test_noalloc_rejects_sizes_beyond_its_maximum():
    mpm = load_package()
    prob = mpm.read_deck(deck_path("slope_dp_bspline_usf"))
    s = prob.struct()
    # rejected before any device access (null arrays are not touched)
    rc = mpm._lib.lib().mpm_cell_order_noalloc(C.byref(s), 8, 8, 8, (1 << 21) + 1, 8, 8, None)
    assert rc == -1   # MPM_ERR_ARG
