"""Johnson-Cook with heating (john) and with damage and failure (jcf) on the
device, in the reference-order FP64 mode (Constitution.f90:224-237,
278-319).

The stress pass records, for every particle, the plastic strain increment
its update assigned; a scan over each instance's records in original
particle order gives the particles that assigned none the increment the last
one before them left (the reference's module variable depeff, carried
across steps per instance), and a tail pass heats (john, jcf) and damages or
fails (jcf) with it.  The device is bit-identical to:
* the reference's own state after 10 steps (fixtures);
* the oracle (with the correctly rounded log and pow it shares) after 50 and
  200 steps, and after every phase of the first two steps;
* each instance's own oracle run in an engine that batches two copies of the
  jcf deck with the john decks and decks of other plastic models (whose
  increments must not reach the Johnson-Cook instances), stepped through
  captured graphs with re-sorts;
and the other precisions refuse such problems.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR, JC_DECKS, deck_path, load_package
from test_gpu_exact import ORACLE, PHASES, assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mpm():
    return load_package()


@pytest.mark.parametrize("name", JC_DECKS)
def test_jc_matches_reference_fixture_bitwise(mpm, name):
    g = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    prob = mpm.read_deck(deck_path(name))
    with mpm.Context(prob, "fp64exact") as ctx:
        ctx.step(10)
        ps = ctx.particles()
        ref = ps.copy()
        for key in g.files:
            if key.startswith("s10_"):
                ref[mpm.PF[key[4:]]] = g[key]
        assert_bitwise(mpm, ps, ref, ctx.grid(), g["g10"], name)
        assert np.array_equal(ctx.icell(), g["icell10"])
        assert np.any(ps[mpm.PF["temp"]] != 293.0)


@pytest.mark.parametrize("name", JC_DECKS)
def test_jc_matches_oracle_bitwise_to_200_steps(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob, log_cr=True)
    with mpm.Context(prob, "fp64exact") as ctx:
        done = 0
        for s in (50, 200):
            ctx.step(s - done)
            o.step(s - done)
            done = s
            assert_bitwise(mpm, ctx.particles(), o.particles(), ctx.grid(), o.grid(), "%s step %d" % (name, s))
            assert np.array_equal(ctx.icell(), o.icell())
            sc, so = ctx.scalars(), o.scalars()
            assert sc[1] == so[1] and sc[2] == so[2]


@pytest.mark.parametrize("name", JC_DECKS)
def test_jc_each_phase_entry_point_bitwise(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob, log_cr=True)
    seq = PHASES[{0: "usf", 1: "usl", 2: "musl"}[prob.scheme]]
    with mpm.Context(prob, "fp64exact") as ctx:
        for step in (1, 2, 3):
            ctx.begin_step()
            o.phase("begin_step")
            moved = False
            for ph in seq:
                getattr(ctx, ph)()
                o.phase(ORACLE[ph])
                moved = moved or ph == "ParticlePositionUpdate"
                assert_bitwise(mpm, ctx.particles(), o.particles(), ctx.grid(), o.grid(),
                               "%s step %d after %s" % (name, step, ph), skip=() if moved else ("xx", "xy", "xz"))
            ctx.calcEnergy()
            o.phase("calc_energy")
            assert ctx.UpdateDT() == (o.phase("update_dt") or o.scalars()[2])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_jc_is_refused_outside_the_exact_mode(mpm, prec):
    prob = mpm.read_deck(deck_path(JC_DECKS[0]))
    with pytest.raises(mpm.MpmError, match="reference-order"):
        mpm.Context(prob, prec)


def test_jc_in_a_batched_engine_with_graphs_and_resorts(mpm, oracle_mod):
    names = JC_DECKS + ["impact_jcf_eos2_gimp_musl", "bar_j2_gimp_musl", "slope_dp_bspline_usf",
                        "impact_sjcf_eos2_bspline_usf"]
    probs = [(n, mpm.read_deck(deck_path(n))) for n in names]
    eng = mpm.Engine(max_particles=sum(p.n for _, p in probs) + len(probs) * 256,
                     max_nodes=sum(p.nb_node for _, p in probs) + len(probs) * 256,
                     max_instances=len(probs), max_materials=sum(len(p.mats) for _, p in probs),
                     max_bodies=sum(len(p.bodies) for _, p in probs), precision="fp64exact")
    eng.apply_tuning(sort_every=8)
    ids = [eng.add(p) for _, p in probs]
    for _ in range(3):
        eng.step(8, graph=True, graph_steps=4)
        eng.sync()
    for (n, p), i in zip(probs, ids):
        o = oracle_mod.Oracle(p, log_cr=True)
        o.step(24)
        assert_bitwise(mpm, eng.particles(i), o.particles(), eng.grid(i), o.grid(), n + " batched")
        assert np.array_equal(eng.icell(i), o.icell())
    eng.close()
