"""The node-box cache rules (mpm3d-f90_amd/csrc/boxcheck.h) on the CPU.

The particle kernels take a block's cached node box in place of reducing
its own when the rules accept it.  An entry may come from another block size
(the scatter launches and the other particle kernels differ), a smaller tile
capacity, an earlier particle order, an earlier occupant of the slots, or be
garbage; these tests compile the rules for the host (tests/box_cache_shim.cpp)
and check:

* safety: whenever the rules accept an entry for a block, every stencil node
  of every active particle lies in the instance's grid and maps to a slot of
  its side's box inside the tile, and every node the tile loads and flushes
  is a grid node -- for entries built by other block sizes, tile caps and
  steps, for entries from another grid, and for random garbage;
* the case a work-in-progress version of the cache faulted on (a single-box
  entry written by a 128-particle block read by a 256-particle block, after a
  re-sort): the upper particles are on side 0 and the entry is taken only if
  it holds them;
* an entry whose box reaches outside the grid is refused even where it holds
  every stencil (the grid rule);
* liveness: a block accepts the entry built from its own particles.
Entries are built like the kernels' box reduction (block_box_n and
block_box2 in kernels.hip: tight box of the stencils, clamped to the grid,
split at the largest jump when one box exceeds the tile).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
INT_MAX = 0x7fffffff
REACH = {0: (0, 1), 1: (-1, 2), 2: (-1, 1)}      # linear, GIMP, B-spline: cell + lo .. cell + hi


@pytest.fixture(scope="module")
def rules(tmp_path_factory):
    d = tmp_path_factory.mktemp("boxcheck")
    so = str(d / "box_cache_shim.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "box_cache_shim.cpp")],
                   check=True)
    L = C.CDLL(so)
    ip = C.POINTER(C.c_int)
    L.entry_consistent.argtypes = [ip, C.c_int, C.c_int, C.c_int, C.c_int]
    L.entry_holds.argtypes = [ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    return L


def box(cells, act, lo, hi, m, cap):
    """block_box_n: [x0,y0,z0,bx,by,bz,n,ok,inner]"""
    if not act.any():
        return [0, 0, 0, 0, 0, 0, 0, 1, 0]
    mn, mxv = cells[act].min(0), cells[act].max(0)
    inner = int(all(mn + lo >= 0) and all(mxv + hi <= m))
    x0 = np.maximum(mn + lo, 0)
    b = np.minimum(mxv + hi, m) - x0 + 1
    n = int(np.prod(b))
    ok = int(n <= cap)
    return [int(x0[0]), int(x0[1]), int(x0[2]), int(b[0]), int(b[1]), int(b[2]), n if ok else 0, ok, inner]


def entry(cells, act, shape, m, cap):
    """block_box2: the 22-int record"""
    lo, hi = REACH[shape]
    b0 = box(cells, act, lo, hi, m, cap)
    single = b0 + b0[:6] + [0] + b0[7:9] + [INT_MAX, b0[6], b0[7], b0[8]]
    if b0[7]:
        return single
    best, j = -1, -1
    for t in range(1, len(cells)):
        if act[t] and act[t - 1]:
            d = int(np.abs(cells[t] - cells[t - 1]).max())
            if d > best:
                best, j = d, t
    if j < 0:
        return single
    s1 = np.arange(len(cells)) >= j
    a, b = box(cells, act & ~s1, lo, hi, m, cap), box(cells, act & s1, lo, hi, m, cap)
    if a[7] and b[7] and a[8] and b[8] and a[6] + b[6] <= cap:
        return a + b + [j, a[6] + b[6], 1, 1]
    return single


def accepts(L, e, cells, act, shape, m, cap):
    arr = (C.c_int * 22)(*e)
    if not L.entry_consistent(arr, cap, int(m[0]), int(m[1]), int(m[2])):
        return False
    return all(L.entry_holds(arr, shape, j, int(c[0]), int(c[1]), int(c[2]))
               for j, (c, a) in enumerate(zip(cells, act)) if a)


def check_safe(e, cells, act, shape, m, cap):
    """what the kernels do with an accepted entry stays in the tile and grid"""
    boxes = np.array([e[0:9], e[9:18]])
    split, n = e[18], e[19]
    assert 0 < n <= cap
    lo, hi = REACH[shape]
    r = np.arange(lo, hi + 1)
    offs = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    j = np.nonzero(act)[0]
    k = (j >= split).astype(int)
    nodes = cells[j][:, None, :] + offs[None]                       # (particles, stencil, 3)
    assert (nodes >= 0).all() and (nodes <= m).all()
    b = boxes[k]                                                    # the particle's side box
    rel = nodes - b[:, None, 0:3]
    slot = rel[..., 0] + b[:, None, 3] * (rel[..., 1] + b[:, None, 4] * rel[..., 2])
    assert (rel >= 0).all() and (rel < b[:, None, 3:6]).all()
    assert (slot >= 0).all() and (slot < b[:, None, 6]).all()
    assert (slot + np.where(k, boxes[0, 6], 0)[:, None] < cap).all()
    for x0, y0, z0, bx, by, bz, nk, _, _ in boxes:     # tile loads and flushes, e < n of each side
        if nk:
            assert bx * by * bz == nk
            assert x0 >= 0 and y0 >= 0 and z0 >= 0
            assert x0 + bx - 1 <= m[0] and y0 + by - 1 <= m[1] and z0 + bz - 1 <= m[2]


def sorted_cells(rng, n, m, jumps=4):
    """cells of particles in a cell order: runs along x-rows with a few jumps
    (where a space-filling order leaves the occupied cells)"""
    out, c = [], rng.integers(0, m - 1)
    while len(out) < n:
        for _ in range(int(rng.integers(1, 8))):
            out.append(c.copy())
        c = c + np.array([1, 0, 0])
        if c[0] >= m[0] or rng.random() < 0.08:
            c = np.array([rng.integers(0, m[0]), (c[1] + 1) % m[1], c[2]])
        if rng.random() < jumps / n:
            c = rng.integers(0, m - 1)
    return np.array(out[:n])


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_accepted_entries_stay_in_tile_and_grid(rules, shape):
    rng = np.random.default_rng(1234 + shape)
    taken = 0
    for trial in range(60):
        m = rng.integers(8, 24, size=3)            # node index range [0, m] per axis
        n_slots = 512
        # mostly interior cells (the rules only take inner boxes), some at the faces
        cells = sorted_cells(rng, n_slots, m - 4) + 2
        if trial % 4 == 0:
            cells = sorted_cells(rng, n_slots, m)
        act = rng.random(n_slots) > 0.05
        caps = (512, 384, 256)
        # entries written by other block sizes, tile caps and steps (cells moved
        # by one), keyed by the first 64-slot group of a 256-particle block
        moved = np.clip(cells + rng.integers(-1, 2, size=cells.shape), 0, m - 1)
        for bs_w in (64, 128, 256):
            for cap_w in caps:
                for src in (cells, moved):
                    s0 = 64 * int(rng.integers(0, 4))
                    w = entry(src[s0:s0 + bs_w], act[s0:s0 + bs_w], shape, m, cap_w)
                    for bs_r in (64, 128, 256):
                        cap_r = caps[int(rng.integers(0, 3))]
                        blk = slice(s0, s0 + bs_r)
                        if accepts(rules, w, cells[blk], act[blk], shape, m, cap_r):
                            taken += 1
                            check_safe(w, cells[blk], act[blk], shape, m, cap_r)
        # an entry of another instance's grid, and garbage
        m2 = rng.integers(4, 40, size=3)
        other = entry(sorted_cells(rng, 256, m2), np.ones(256, bool), shape, m2, 512)
        junk = [list(rng.integers(-50, 600, size=22)) for _ in range(20)]
        for e in [other] + junk:
            e = list(e)
            if rng.random() < 0.5:
                e[7] = e[8] = e[16] = e[17] = e[20] = e[21] = 1
            blk = slice(0, 256)
            if accepts(rules, e, cells[blk], act[blk], shape, m, 512):
                taken += 1
                check_safe(e, cells[blk], act[blk], shape, m, 512)
    assert taken > 100, taken      # the rules do take entries (the check above is not vacuous)


def test_entry_of_a_smaller_block_read_by_a_larger_one(rules):
    """a single-box entry from a 128-particle block has split = INT_MAX, so a
    256-particle block puts particles 128-255 on side 0: taken only when that
    box holds their stencils (after a re-sort it does not)"""
    m = np.array([40, 12, 12])
    lo_cells = np.array([[2 + (i // 8) % 10, 3 + i // 80, 4] for i in range(128)])
    hi_cells = lo_cells + np.array([20, 0, 2])            # elsewhere after a re-sort
    act = np.ones(256, bool)
    e = entry(lo_cells, act[:128], 2, m, 512)
    assert e[18] == INT_MAX and e[9 + 6] == 0
    assert not accepts(rules, e, np.vstack([lo_cells, hi_cells]), act, 2, m, 512)
    assert accepts(rules, e, np.vstack([lo_cells, lo_cells[::-1]]), act, 2, m, 512)


def test_box_reaching_outside_the_grid_is_refused(rules):
    m = np.array([10, 10, 10])
    cells = np.array([[1, 1, 1]] * 64)
    act = np.ones(64, bool)
    e = entry(cells, act, 2, m, 512)
    assert accepts(rules, e, cells, act, 2, m, 512)
    bad = list(e)
    bad[0] -= 2                 # x0 = -2: the stencil is still inside the box
    bad[3] += 2
    bad[6] = bad[3] * bad[4] * bad[5]
    bad[19] = bad[6]
    assert not accepts(rules, bad, cells, act, 2, m, 512)
    zero = [0] * 22
    assert not accepts(rules, zero, cells, act, 2, m, 512)


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_own_entry_is_taken(rules, shape):
    rng = np.random.default_rng(99 + shape)
    for _ in range(50):
        m = rng.integers(8, 30, size=3)
        # interior cells, so the box is inner
        cells = np.clip(sorted_cells(rng, 256, m - 4), 0, None) + 2
        act = np.ones(256, bool)
        e = entry(cells, act, shape, m, 512)
        if e[20]:
            assert accepts(rules, e, cells, act, shape, m, 512)
