"""Device numeric flags and instance isolation.

A particle given a NaN velocity makes its instance's stress update
non-finite: the count is read back (mpm_numeric_flags,
mpm_engine_numeric_flags), mpm_update_dt returns MPM_ERR_NUMERIC, and the
serving loop evicts and reports that run.  The NaN stays in its instance's
own particle and node ranges: the instances batched with it are bit for bit
their solo runs (reference-order FP64 mode, which is reproducible bitwise)."""
import numpy as np
import pytest

from conftest import deck_path, load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mpm():
    return load_package()


def poisoned(mpm, prob, k=7):
    p = prob.copy()
    p.pstate[mpm.PF["vx"], k] = np.nan
    return p


def test_context_counts_non_finite_state(mpm):
    prob = mpm.read_deck(deck_path("cube_elastic_linear_usf"))
    clean = mpm.Context(prob, "fp64")
    clean.step(3)
    assert clean.numeric_flags() == (0, 0)
    assert clean.UpdateDT() > 0
    clean.close()
    ctx = mpm.Context(poisoned(mpm, prob), "fp64")
    ctx.step(3)
    negvol, nonfinite = ctx.numeric_flags()
    assert nonfinite > 0
    with pytest.raises(mpm.MpmError) as ei:
        ctx.UpdateDT()
    assert ei.value.code == -8   # MPM_ERR_NUMERIC
    ctx.close()


def test_nan_instance_leaves_its_batch_mates_unchanged(mpm):
    names = ["bar_j2_gimp_musl", "cube_elastic_linear_usf", "slope_dp_bspline_usf"]
    probs = [mpm.read_deck(deck_path(n)) for n in names]
    bad = poisoned(mpm, probs[1])
    batch = probs[:1] + [bad] + probs[2:]
    eng = mpm.Engine(max_particles=sum(p.n for p in batch) + 8 * 256,
                     max_nodes=sum(p.nb_node for p in batch) + 8 * 256, max_instances=4, precision="fp64exact")
    ids = [eng.add(p) for p in batch]
    eng.step(12, graph=True, graph_steps=4)
    flags = eng.numeric_flags(ids)
    assert flags[ids[1]][1] > 0
    for k in (0, 2):
        assert flags[ids[k]] == (0, 0)
        solo = mpm.Context(probs[k], "fp64exact")
        solo.step(12)
        assert np.array_equal(eng.particles(ids[k]), solo.particles(), equal_nan=True), names[k]
        assert np.array_equal(eng.grid(ids[k]), solo.grid(), equal_nan=True), names[k]
        solo.close()
    eng.close()


def test_server_evicts_and_reports_a_non_finite_run(mpm):
    from mpm3d_f90_amd import serve
    good = [mpm.read_deck(deck_path(n)) for n in ("bar_j2_gimp_usf", "slope_dp_bspline_usl")]
    bad = poisoned(mpm, mpm.read_deck(deck_path("cube_elastic_linear_usl")))
    eng = mpm.Engine(max_particles=20000, max_nodes=20000, max_instances=4, precision="fp64exact")
    srv = serve.Server(eng, chunk=4, graph="lazy", keep_state=True)
    for rid, p in enumerate([good[0], bad, good[1]]):
        srv.submit(serve.Request(rid, p, 16, arrival=0.0))
    srv.admit(0.0)
    assert len(srv.running) == 3
    while not srv.idle():
        srv.advance()
        srv.retire(lambda: 0.0)
    assert [r.rid for r in srv.failed] == [1] and srv.failed[0].result["nonfinite"] > 0
    assert sorted(r.rid for r in srv.done) == [0, 2]
    for r in srv.done:
        solo = mpm.Context(r.problem, "fp64exact")
        solo.step(16)
        assert np.array_equal(r.result["particles"], solo.particles(), equal_nan=True)
        solo.close()
    eng.close()
