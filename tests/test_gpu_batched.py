"""The batched and served forms against the CPU FP64 oracle, every instance.

* C3 form: a seeded ensemble of slope instances (sizes 2e4-3e4 particles,
  random angle, friction and cohesion; B-spline, Drucker-Prager, USF) in one
  engine, launch shape picked by the autotuner, steps replayed from captured
  graphs with re-sorts in between: every instance after 24 steps against its
  oracle run (FP64 1e-12, FP32 bulk 1e-5, reference-order FP64 bit for bit).
* C5 form: a mix of small slopes and Taylor bars served under Poisson
  arrivals with a device budget that holds only a few at a time (runs queue,
  join and leave while others are on the GPU); every completed run against
  its oracle run of the same step count.
* C1 form: a reduced Taylor bar (GIMP, J2, MUSL) after 50 steps.
Tolerances are normwise per physical group (conftest.group_errors)."""
import numpy as np
import pytest

from conftest import GGROUPS, PGROUPS, group_errors, load_package

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}


def errs(mpm, a, b):
    return group_errors(a, b, mpm.PF_NAMES, PGROUPS)


def check(mpm, got, ref, prec, what):
    if prec == "fp64exact":
        d = [n for k, n in enumerate(mpm.PF_NAMES) if not np.array_equal(got[k], ref[k], equal_nan=True)]
        assert not d, "%s: fields differ bitwise: %s" % (what, d)
        return
    e = errs(mpm, got, ref)
    assert max(e.values()) <= TOL[prec], (what, prec, e)


@pytest.fixture(scope="module")
def mods():
    mpm = load_package()
    from mpm3d_f90_amd import decks
    return mpm, decks


def test_slope_ensemble_autotuned_graphs_every_instance(mods, oracle_mod):
    mpm, decks = mods
    ens = decks.slope_ensemble(12, seed=3, n_min=2e4, n_max=3e4)
    probs = [decks.load(t, mpm, "ens%d" % k) for k, (t, n, m) in enumerate(ens)]
    refs = []
    for p in probs:
        o = oracle_mod.Oracle(p)
        o.step(24)
        refs.append(o.particles())
    tot_p = sum(p.n for p in probs)
    tot_n = sum(m["nodes"] for _, _, m in ens)
    for prec in ("fp64", "fp32", "fp64exact"):
        eng = mpm.Engine(max_particles=tot_p + 20 * 256, max_nodes=tot_n + 20 * 256, max_instances=12,
                         max_materials=12, max_bodies=sum(m["bodies"] for _, _, m in ens), precision=prec)
        ids = [eng.add(p) for p in probs]
        if prec != "fp64exact":
            eng.autotune(trial_steps=3)   # state restored afterwards
        eng.apply_tuning(sort_every=8)
        for _ in range(3):
            eng.step(8, graph=True, graph_steps=4)
            eng.sync()
        for k, iid in enumerate(ids):
            assert int(eng.scalars(iid)[0]) == 24
            check(mpm, eng.particles(iid), refs[k], prec, "instance %d (%d particles)" % (k, probs[k].n))
        eng.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_served_poisson_mix_every_run(mods, oracle_mod, prec):
    mpm, decks = mods
    from mpm3d_f90_amd import serve
    pool = serve.request_pool(mpm, decks, count=8, seed=11, steps=(16, 40), slope_n=(2e4, 3e4),
                              bar_radius=(2.0, 2.6), bar_length=8.0)
    eng = mpm.Engine(hbm_bytes=60e6 if prec == "fp32" else 90e6, max_instances=8, precision=prec)
    srv = serve.Server(eng, chunk=8, graph="lazy", keep_state=True)
    st = serve.run_poisson(srv, pool, rate=1000.0, duration=0.03, seed=5, max_wall=120)
    assert st["completed"] == st["arrived"] >= 8, st
    assert st["max_concurrent"] >= 2, st                 # runs overlapped on the device
    assert max(r.admitted - r.arrival for r in srv.done) > 1e-3   # and some waited for room
    cache = {}
    for r in srv.done:
        key = (id(r.problem), r.steps)
        if key not in cache:
            o = oracle_mod.Oracle(r.problem)
            o.step(r.steps)
            cache[key] = o.particles()
        assert r.final_step == r.steps
        check(mpm, r.result["particles"], cache[key], prec, "request %d (%s, %d steps)" % (r.rid, r.kind, r.steps))
    eng.close()


def test_reduced_taylor_bar(mods, oracle_mod):
    mpm, decks = mods
    text, n = decks.taylor_bar(radius=1.6, length=8.0)
    prob = decks.load(text, mpm, "bar")
    o = oracle_mod.Oracle(prob)
    o.step(50)
    ref, gref = o.particles(), o.grid()
    for prec in ("fp64", "fp32", "fp64exact"):
        ctx = mpm.Context(prob, prec)
        ctx.step(50)
        check(mpm, ctx.particles(), ref, prec, "Taylor bar, %d particles" % prob.n)
        if prec == "fp64":
            ge = group_errors(ctx.grid(), gref, mpm.GF_NAMES, GGROUPS)
            assert max(ge.values()) <= 1e-12, ge
        ctx.close()
