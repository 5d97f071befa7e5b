"""Parity of the MI355X path against the CPU oracle and the reference fixtures.

Tolerances (normwise per physical group, see conftest.group_errors):
* FP64 plain path: 1e-12 after 2 and 10 steps, 1e-10 after 50 steps (the
  round-off difference grows with the step count).  The scatter sums (P2G)
  are accumulated in a different order than the reference's ascending
  particle loop, so results agree to round-off rather than bit for bit.
  Exception: the weakly compressible water column (null material + EOS 4)
  whose pressure c^2 (rho - rho0) amplifies a 1e-16 volume change by
  rho/(rho - rho0) ~ 1e4: 1e-9 for its stress and energy.
* cell indices (icell) and the sort permutation: bit-exact.
* FP32 bulk path (FP64 positions, volume, energy, F, grid mass): 1e-5
  after 10 steps.
* Decks on which the reference's own algorithm turns round-off into O(1)
  changes are compared only bit for bit, in the reference-order mode
  (tests/test_gpu_exact.py), not here (GATED below; tests/test_oracle.py
  shows the oracle itself moving by more than 1e-5 under a one-ulp change
  of the initial velocities).  In the Taylor impact with the Mie-Gruneisen
  EOS the artificial bulk viscosity applies only where the pressure before
  the step is compressive (eos2, Constitution.f90:1428-1430); particles
  ahead of the compression wave carry a pressure of round-off size, so its
  sign, not its value, decides an O(1) viscous term.  The Johnson-Cook
  block deck is compared to step 10 (SHORT below).
"""
import os

import numpy as np
import pytest

from conftest import ALL_DECKS as DECKS, GOLDEN_DIR, PGROUPS, GGROUPS, deck_path, group_errors, load_package

pytestmark = pytest.mark.gpu

TOL64 = {name: 1e-12 for name in DECKS}
TOL64_LOOSE = {"dam_null_eos4_musl": {"stress": 1e-9, "ie": 1e-9},
               "dam_null_sgmp_musl": {"stress": 1e-9, "ie": 1e-9}}
WATER = ("dam_null_eos4_musl", "dam_null_sgmp_musl")
GATED = {"taylor_sjc_eos2_musl": "bulk viscosity gated on the sign of a round-off-sized pressure: compared bit "
                                 "for bit in the reference-order mode (test_gpu_exact.py)"}
TOL_DECKS = [n for n in DECKS if n not in GATED]
# compared to 10 steps only: by step 50 a particle yields by a round-off
# amount, and the Johnson-Cook rate factor 1 + C ln(max(rate, 0.001))
# (Constitution.f90:1077-1081) then cuts its yield stress by C ln 1000 (17%);
# the oracle jumps by the same 2e-8 under one-ulp velocity changes
# (tests/test_oracle.py)
SHORT = {"cube_sjc_eos1_gimp_usl": 10}


def perr(mpm, a, b):
    return group_errors(a, b, mpm.PF_NAMES, PGROUPS)


def gerr(mpm, a, b):
    return group_errors(a, b, mpm.GF_NAMES, GGROUPS)


def assert_close(errs, tol, loose=None, what=""):
    for g, e in errs.items():
        t = (loose or {}).get(g, tol)
        assert e <= t, "%s %s: %.3e > %.1e (%s)" % (what, g, e, t, errs)


@pytest.fixture(scope="module")
def mpm():
    return load_package()


@pytest.mark.parametrize("name", DECKS)
def test_icell_bitexact(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    ctx = mpm.Context(prob, "fp64")
    o = oracle_mod.Oracle(prob)
    for _ in range(3):
        ctx.step(1)
        o.step(1)
        assert np.array_equal(ctx.icell(), o.icell())
    ctx.close()


@pytest.mark.parametrize("name", TOL_DECKS)
def test_fp64_fused_steps_match_oracle(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    ctx = mpm.Context(prob, "fp64")
    o = oracle_mod.Oracle(prob)
    done = 0
    for s in (2, 10, 50):
        if s > SHORT.get(name, 50):
            break
        ctx.step(s - done)
        o.step(s - done)
        done = s
        tol = TOL64[name] if s <= 10 else 1e-10
        assert_close(perr(mpm, ctx.particles(), o.particles()), tol, TOL64_LOOSE.get(name), "%s s%d" % (name, s))
        assert_close(gerr(mpm, ctx.grid(), o.grid()), tol, TOL64_LOOSE.get(name), "%s grid s%d" % (name, s))
        sc, so = ctx.scalars(), o.scalars()
        assert sc[0] == so[0]
        assert abs(sc[1] / so[1] - 1) < 1e-13 and abs(sc[2] / so[2] - 1) < 1e-13
    ctx.close()


@pytest.mark.parametrize("name", TOL_DECKS)
def test_fp64_matches_reference_fixture(mpm, name):
    """Directly against the Fortran reference's own state after 10 steps."""
    g = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    prob = mpm.read_deck(deck_path(name))
    ctx = mpm.Context(prob, "fp64")
    ctx.step(10)
    ref = np.zeros((mpm.PF_COUNT, prob.n))
    for key in g.files:
        if key.startswith("s10_"):
            ref[mpm.PF[key[4:]]] = g[key]
    assert_close(perr(mpm, ctx.particles(), ref), TOL64[name], TOL64_LOOSE.get(name), name)
    assert_close(gerr(mpm, ctx.grid(), g["g10"]), TOL64[name], TOL64_LOOSE.get(name), name + " grid")
    assert np.array_equal(ctx.icell(), g["icell10"])
    sc = ctx.scalars()
    assert abs(sc[2] / g["scal10"][0] - 1) < 1e-13
    ctx.close()


PHASES = ["grid_momentum_initial", "apply_bc", "particle_stress_update", "grid_momentum_update",
          "integrate_momentum", "particle_position_update", "grid_momentum_musl"]
CTX_NAME = {"grid_momentum_initial": "GridMomentumInitial", "apply_bc": "ApplyBoundaryConditions",
            "particle_stress_update": "ParticleStressUpdate", "grid_momentum_update": "GridMomentumUpdate",
            "integrate_momentum": "IntegrateMomentum", "particle_position_update": "ParticlePositionUpdate",
            "grid_momentum_musl": "GridMomentumMUSL"}


@pytest.mark.parametrize("name", TOL_DECKS)
def test_each_phase_entry_point_matches_oracle(mpm, oracle_mod, name):
    """Drive both sides phase by phase in the reference's call order and
    compare particles and grid after every phase of steps 1 and 2."""
    prob = mpm.read_deck(deck_path(name))
    ctx = mpm.Context(prob, "fp64")
    o = oracle_mod.Oracle(prob)
    usf, musl = prob.scheme == 0, prob.scheme == 2
    seq = ["grid_momentum_initial"]
    if usf:
        seq += ["apply_bc", "particle_stress_update"]
    seq += ["grid_momentum_update", "integrate_momentum", "particle_position_update"]
    if musl:
        seq += ["grid_momentum_musl", "apply_bc"]
    if not usf:
        seq += ["particle_stress_update"]
    tol_loose = {"stress": 1e-6, "ie": 1e-6} if name in WATER else None
    for step in (1, 2):
        ctx.begin_step()
        o.phase("begin_step")
        for ph in seq:
            getattr(ctx, CTX_NAME[ph])()
            o.phase(ph)
            where = "%s step %d after %s" % (name, step, ph)
            assert_close(gerr(mpm, ctx.grid(), o.grid()), 1e-12, None, where)
            pe = perr(mpm, ctx.particles(), o.particles())
            if step == 1 or name in WATER:
                # stresses of the first steps are round-off of a zero state (the
                # water column's pressure builds up over several steps); the
                # fused-step test compares them from step 2 on
                pe.pop("stress", None)
                pe.pop("ie", None)
            assert_close(pe, 1e-12, tol_loose, where)
        ek, ei, mom = ctx.calcEnergy()
        o.phase("calc_energy")
        so = o.scalars()
        assert abs(ek - so[3]) <= 1e-13 * abs(so[3]) + 1e-300
        dt = ctx.UpdateDT()
        o.phase("update_dt")
        assert abs(dt / o.scalars()[2] - 1) < 1e-13
    ctx.close()


@pytest.mark.parametrize("name", TOL_DECKS)
def test_fp32_bulk_within_1e5_of_fp64_oracle(mpm, oracle_mod, name):
    """FP32 bulk after 10 steps.  The water columns are compared later: their
    pressure c^2 (m/V - rho0) starts from zero and during the first steps is
    the small difference of FP32 strain increments of an almost rigidly falling
    column, so its relative error is large while the pressure is small (GIMP
    column, stress: 7.9e-5 at step 10, 5.3e-6 at step 50, 8.4e-7 at step 200;
    SGMP column, internal energy: 3.3e-6, 4.2e-6, 6.8e-6 at steps 10, 50, 200,
    on MI355X; the particle mass of the density is kept in FP64)."""
    prob = mpm.read_deck(deck_path(name))
    ctx = mpm.Context(prob, "fp32")
    o = oracle_mod.Oracle(prob)
    nst = {"dam_null_eos4_musl": 200, "dam_null_sgmp_musl": 50}.get(name, 10)
    ctx.step(nst)
    o.step(nst)
    pe = perr(mpm, ctx.particles(), o.particles())
    assert_close(pe, 1e-5, None, name + " fp32")
    assert_close(gerr(mpm, ctx.grid(), o.grid()), 1e-5, None, name + " fp32 grid")
    ctx.close()


def test_batched_engine_matches_solo_runs(mpm):
    """All decks (four shape families, every scheme, SGMP included, every
    material model) in one engine: each instance equals its solo context run (to FP64
    round-off)."""
    probs = {n: mpm.read_deck(deck_path(n)) for n in TOL_DECKS}
    eng = mpm.Engine(max_particles=sum(p.n for p in probs.values()) + len(probs) * 256,
                     max_nodes=sum(p.nb_node for p in probs.values()) + len(probs) * 256,
                     max_instances=len(probs), max_materials=sum(len(p.mats) for p in probs.values()),
                     max_bodies=sum(len(p.bodies) for p in probs.values()))
    ids = {n: eng.add(p) for n, p in probs.items()}
    eng.step(10)
    for n, p in probs.items():
        ctx = mpm.Context(p, "fp64")
        ctx.step(10)
        assert_close(perr(mpm, eng.particles(ids[n]), ctx.particles()), 1e-12,
                     TOL64_LOOSE.get(n), n + " batched")
        assert np.array_equal(eng.icell(ids[n]), ctx.icell())
        ctx.close()
    eng.close()


def test_graph_replay_matches_plain_steps(mpm):
    p = mpm.read_deck(deck_path("bar_j2_gimp_musl"))
    a = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=2)
    b = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=2)
    ia, ib = a.add(p), b.add(p)
    a.step(12)
    for _ in range(3):            # replays with host work in between
        b.step(4, graph=True, graph_steps=2)
        b.sync()
        _ = b.scalars(ib)
    assert_close(perr(mpm, b.particles(ib), a.particles(ia)), 1e-12, None, "graph")
    assert b.scalars(ib)[0] == 12
    a.close()
    b.close()


@pytest.mark.parametrize("name,brick", [("bar_j2_gimp_musl", 0), ("slope_dp_bspline_usf", 0),
                                        ("dam_null_eos4_musl", 0), ("slope_dp_bspline_usf", 4),
                                        ("slope_dp_bspline_usf", 1), ("bar_j2_gimp_musl", 2),
                                        ("plate_neoh_musl_qulo", 8)])
def test_sort_is_the_stable_cell_order(mpm, oracle_mod, name, brick):
    p = mpm.read_deck(deck_path(name))
    eng = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=2)
    i = eng.add(p)
    eng.apply_tuning(sort_brick=brick)
    ref = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=2)
    j = ref.add(p)
    eng.step(3)
    ref.step(3)
    order0 = eng.order(i)
    ps = eng.particles(i)
    eng.sort()
    # brick-major cell key (bricks of B^3 cells), stable in the current slot order
    key = brick_key(p, oracle_cells(p, ps), brick)
    expect = order0[np.argsort(key[order0], kind="stable")]
    assert np.array_equal(eng.order(i), expect)
    eng.step(5)
    ref.step(5)
    assert_close(perr(mpm, eng.particles(i), ref.particles(j)), 1e-12, None, "sorted")
    eng.close()
    ref.close()


@pytest.mark.parametrize("name", ["bar_j2_gimp_musl", "slope_dp_bspline_usf", "cube_elastic_linear_usl",
                                  "dam_null_eos4_musl", "plate_neoh_musl_qulo"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_small_tiles_take_the_split_and_direct_paths(mpm, name, prec, tol):
    """A block whose node box does not fit its LDS tile is split in two boxes
    at the largest jump of its particles' cells, or scatters and gathers
    directly in global memory.  Small tile capacities send most blocks down
    these paths (64 nodes: single, split and direct blocks mixed; 24 nodes:
    every B-spline and GIMP block direct); the results equal the full-tile
    run up to the summation order of the scatters."""
    if prec == "fp32" and name == "dam_null_eos4_musl":
        pytest.skip("FP32 water column: its early near-zero pressure is not reproducible to 1e-5 across "
                    "summation orders (as in the FP32 parity test)")
    p = mpm.read_deck(deck_path(name))
    steps = 6 if name == "dam_null_eos4_musl" else 10
    ref = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=1, precision=prec)
    j = ref.add(p)
    ref.step(steps)
    want = ref.particles(j)
    ref.close()
    loose = {"stress": 1e-9, "ie": 1e-9} if (prec == "fp64" and name == "dam_null_eos4_musl") else {}
    for cap in (64, 24):
        eng = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=1, precision=prec)
        i = eng.add(p)
        eng.apply_tuning(tile_nodes=cap)
        assert eng.tuning()["tile_nodes"] == cap
        eng.step(steps)
        errs = group_errors(eng.particles(i), want, mpm.PF_NAMES, PGROUPS)
        for g, e in errs.items():
            assert e <= loose.get(g, tol), (name, prec, cap, g, e, errs)
        eng.close()


@pytest.mark.parametrize("name", ["dam_null_eos4_musl", "slope_dp_bspline_usf", "cube_elastic_linear_usl"])
@pytest.mark.parametrize("bp,bsc,tile", [(64, 256, 512), (256, 64, 512), (128, 256, 64), (256, 128, 512)])
def test_block_sizes_share_the_node_box_cache(mpm, oracle_mod, name, bp, bsc, tile):
    """The particle kernels keep each block's node box in a cache keyed by the
    block's first particle slot; the scatter launch and the other particle
    kernels may run with different block sizes (and so read entries written
    for other particle ranges), and a re-sort or a smaller tile capacity
    leaves stale entries.  Whatever the mix, every step equals the oracle to
    FP64 round-off."""
    p = mpm.read_deck(deck_path(name))
    eng = mpm.Engine(max_particles=p.n + 256, max_nodes=p.nb_node + 256, max_instances=1, precision="fp64")
    i = eng.add(p)
    eng.apply_tuning(block_particles=bp, block_scatter=bsc, sort_every=3)
    o = oracle_mod.Oracle(p)
    loose = TOL64_LOOSE.get(name)
    for k, n in enumerate((2, 3, 4)):
        if k == 1:
            eng.apply_tuning(tile_nodes=tile)
        eng.step(n)
        o.step(n)
        assert_close(perr(mpm, eng.particles(i), o.particles()), 1e-12, loose, "%s %d/%d/%d" % (name, bp, bsc, tile))
        assert np.array_equal(eng.icell(i), o.icell())
    eng.close()


def oracle_cells(p, ps):
    """InWhichCell (Grid.f90:123-155) of every particle, 1-based, -1 outside."""
    x, y, z = ps[0], ps[1], ps[2]
    s = p.span
    ix = ((x - s[0]) / p.dcell).astype(np.int64) + 1
    iy = ((y - s[2]) / p.dcell).astype(np.int64) + 1
    iz = ((z - s[4]) / p.dcell).astype(np.int64) + 1
    c = (iz - 1) * p.ncx * p.ncy + (iy - 1) * p.ncx + ix
    out = (x < s[0]) | (x > s[1]) | (y < s[2]) | (y > s[3]) | (z < s[4]) | (z > s[5])
    out |= (c > p.ncx * p.ncy * p.ncz) | (c <= 0)
    return np.where(out, -1, c)


def hilbert_key(x, y, z, bits):
    """Skilling's axes-to-transpose Hilbert transform (AIP Conf. Proc. 707,
    2004), vectorised, bits interleaved x-first."""
    X = [x.astype(np.int64), y.astype(np.int64), z.astype(np.int64)]
    Q = 1 << (bits - 1)
    while Q > 1:
        P = Q - 1
        for i in range(3):
            m = (X[i] & Q) != 0
            X[0] = np.where(m, X[0] ^ P, X[0])
            t = np.where(m, 0, (X[0] ^ X[i]) & P)
            X[0] = X[0] ^ t
            X[i] = X[i] ^ t
        Q >>= 1
    X[1] ^= X[0]
    X[2] ^= X[1]
    t = np.zeros_like(X[0])
    Q = 1 << (bits - 1)
    while Q > 1:
        t = np.where((X[2] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    h = np.zeros_like(X[0])
    for b in range(bits - 1, -1, -1):
        for i in range(3):
            h = (h << 1) | (((X[i] ^ t) >> b) & 1)
    return h


def test_hilbert_key_walks_face_neighbours():
    """The 3D Hilbert order visits every cell of a 2^b cube once, each step
    to a face neighbour."""
    b = 3
    g = np.arange(1 << b)
    x, y, z = [a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")]
    h = hilbert_key(x, y, z, b)
    assert np.array_equal(np.sort(h), np.arange(1 << (3 * b)))
    o = np.argsort(h)
    step = np.abs(np.diff(x[o])) + np.abs(np.diff(y[o])) + np.abs(np.diff(z[o]))
    assert np.all(step == 1)


@pytest.mark.parametrize("gap", [False, True])
def test_pool_sort_of_several_instances(mpm, gap):
    """The pool re-sort orders every instance's slots by its cell key, stably:
    one device-wide sort of (segment rank, key) when the instances tile the
    pool, the segmented sort when a removed instance leaves a gap."""
    names = ("slope_dp_bspline_usf", "bar_j2_gimp_musl", "dam_null_eos4_musl", "cube_elastic_linear_usl")
    probs = [mpm.read_deck(deck_path(n)) for n in names]
    eng = mpm.Engine(max_particles=sum(p.n for p in probs) + 16 * 256,
                     max_nodes=sum(p.nb_node for p in probs) + 16 * 256, max_instances=8)
    ids = [eng.add(p) for p in probs]
    if gap:
        eng.remove(ids[1])
        probs, ids = [probs[0]] + probs[2:], [ids[0]] + ids[2:]
    eng.apply_tuning(sort_brick=2)
    eng.step(3)
    before = [(eng.order(i), eng.particles(i)) for i in ids]
    eng.sort()
    for p, i, (order0, ps) in zip(probs, ids, before):
        key = brick_key(p, oracle_cells(p, ps), 2)
        assert np.array_equal(eng.order(i), order0[np.argsort(key[order0], kind="stable")])
        assert np.array_equal(eng.particles(i), ps)
    eng.close()


def brick_key(p, c, B=4):
    """Sort key of the cell order: B = 0 Hilbert (falls back to B = 4 above
    1024 cells per axis), B > 0 brick-major with B^3-cell bricks."""
    if B == 0:
        bits = max(1, int(np.ceil(np.log2(max(p.ncx, p.ncy, p.ncz)))))
        if bits <= 10:
            c0 = np.maximum(c, 1) - 1
            cz, r = np.divmod(c0, p.ncx * p.ncy)
            cy, cx = np.divmod(r, p.ncx)
            return np.where(c > 0, hilbert_key(cx, cy, cz, bits), 1 << (3 * bits))
        B = 4
    c0 = np.maximum(c, 1) - 1
    cz, r = np.divmod(c0, p.ncx * p.ncy)
    cy, cx = np.divmod(r, p.ncx)
    nbx, nby, nbz = -(-p.ncx // B), -(-p.ncy // B), -(-p.ncz // B)
    key = (((cz // B) * nby + cy // B) * nbx + cx // B) * B ** 3 + ((cz % B) * B + cy % B) * B + cx % B
    return np.where(c > 0, key, nbx * nby * nbz * B ** 3)


def test_autotune_leaves_the_state_and_picks_valid_shapes(mpm):
    """The autotuner times trial steps from a snapshot and restores it: the
    run continues exactly as an untuned engine's (to FP64 round-off, since
    the chosen launch shape changes the scatter order)."""
    names = ["slope_dp_bspline_usf", "bar_j2_gimp_musl"]
    probs = [mpm.read_deck(deck_path(n)) for n in names]
    np_ = sum(p.n for p in probs) + 8 * 256
    nn_ = sum(p.nb_node for p in probs) + 8 * 256
    a = mpm.Engine(max_particles=np_, max_nodes=nn_, max_instances=4)
    b = mpm.Engine(max_particles=np_, max_nodes=nn_, max_instances=4)
    ia = [a.add(p) for p in probs]
    ib = [b.add(p) for p in probs]
    a.step(3)
    b.step(3)
    before = [a.particles(x) for x in ia]
    t = a.autotune(trial_steps=2)
    assert t["block_particles"] in (64, 128, 256) and t["block_nodes"] in (64, 128, 256)
    assert t["sort_brick"] in (0, 2, 4) and t["interleave"] in (0, 1)
    assert t["block_scatter"] in (64, 128, 256)
    assert 1 <= t["sort_every"] <= 64 and t["trials"] == 18 + 2 + 2 + 1
    assert t["step_ms"] > 0 and t["sort_ms"] > 0
    assert a.tuning()["sort_brick"] == t["sort_brick"]
    for x, p0, p in zip(ia, before, probs):
        assert a.scalars(x)[0] == 3
        assert np.array_equal(a.particles(x), p0)
        # the pool is re-sorted in the chosen cell order when tuning ends
        key = brick_key(p, oracle_cells(p, p0), t["sort_brick"])
        o = a.order(x)
        assert np.all(np.diff(key[o]) >= 0)
    a.step(6, graph=True, graph_steps=2)
    b.step(6)
    for x, y, n in zip(ia, ib, names):
        assert a.scalars(x)[0] == 9
        assert_close(perr(mpm, a.particles(x), b.particles(y)), 1e-12, None, n + " tuned")
    a.close()
    b.close()


def test_instances_stop_at_their_own_endtime(mpm):
    p = mpm.read_deck(deck_path("cube_elastic_linear_usl"))
    q = p.copy()
    p.endtime = 5.5 * p.dt
    q.endtime = 2.5 * q.dt
    eng = mpm.Engine(max_particles=2 * p.n + 512, max_nodes=2 * p.nb_node + 512, max_instances=4)
    a, b = eng.add(p, stop_at_endtime=True), eng.add(q, stop_at_endtime=True)
    eng.step(20, graph=True, graph_steps=4)
    st = eng.status()
    assert st[a][0] == 1 and st[b][0] == 1
    assert st[b][1] < st[a][1] <= 8
    eng.close()


def test_cell_indices_across_two_to_the_22(mpm, oracle_mod, tmp_path):
    """Cell coordinates are split off the cell index with an FP32-reciprocal
    quotient below 2^22 and with integer division above.  The small B-spline
    slope stretched along y to 40,001 cells (800,020 cells per z layer)
    puts its base layer below 2^22 and its terraces above: both paths, FP64
    and FP32, against the oracle."""
    text = open(deck_path("slope_dp_bspline_usf")).read()
    assert text.count("spy -0.2 0.6") == 1
    path = tmp_path / "slope_wide.mpm"
    path.write_text(text.replace("spy -0.2 0.6", "spy -0.2 8000.0"))
    prob = mpm.read_deck(str(path))
    o = oracle_mod.Oracle(prob)
    a = mpm.Context(prob, "fp64")
    b = mpm.Context(prob, "fp32")
    o.step(1)
    a.step(1)
    b.step(1)
    ic = o.icell() - 1
    assert (ic >= 1 << 22).sum() > 50 and ((ic >= 0) & (ic < 1 << 22)).sum() > 50
    assert np.array_equal(a.icell(), o.icell()) and np.array_equal(b.icell(), o.icell())
    o.step(9)
    a.step(9)
    b.step(9)
    assert np.array_equal(a.icell(), o.icell())
    assert_close(perr(mpm, a.particles(), o.particles()), 1e-12, None, "wide fp64")
    assert_close(gerr(mpm, a.grid(), o.grid()), 1e-12, None, "wide fp64 grid")
    assert_close(perr(mpm, b.particles(), o.particles()), 1e-5, None, "wide fp32")
    a.close()
    b.close()
