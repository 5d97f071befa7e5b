"""Contact between grid components (cont lagr, Lagr_NodContact,
update_step.f90:1241-1376) on the device, in the reference-order FP64 mode.

Every body maps to the grid of its component; after IntegrateMomentum the
contact pass moves force and momentum between components 1 and 2.  The
device is bit-identical to:
* the reference's own state after 10 steps (fixtures): particles, the grids
  of every component and the contact force tot_cont_for;
* the oracle after 50 and 200 steps, and after every phase (the contact pass
  included) of the first two steps through the per-phase entry points;
and the other precisions refuse contact problems.  With stress smoothing the
smoothed pressure of every body comes from component 1's grid, as the
reference's SmoothStressbyGrid computes it.
"""
import os

import numpy as np
import pytest

from conftest import CONTACT_DECKS, GOLDEN_DIR, deck_path, load_package
from test_gpu_exact import ORACLE, PHASES, assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mpm():
    return load_package()


def _grids_equal(ctx, ref_grids, what):
    for c, g in enumerate(ref_grids):
        assert np.array_equal(ctx.grid_component(c), g, equal_nan=True), "%s component %d" % (what, c + 1)


@pytest.mark.parametrize("name", CONTACT_DECKS)
def test_contact_matches_reference_fixture_bitwise(mpm, name):
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
        _grids_equal(ctx, [g["g10"]] + list(g["gc10"]), name)
        assert np.array_equal(ctx.contact_force(), g["cont10"]), (ctx.contact_force(), g["cont10"])
        assert np.array_equal(ctx.icell(), g["icell10"])


@pytest.mark.parametrize("name", CONTACT_DECKS)
def test_contact_matches_oracle_bitwise_to_200_steps(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob)
    with mpm.Context(prob, "fp64exact") as ctx:
        done = 0
        for s in (50, 200):
            ctx.step(s - done)
            o.step(s - done)
            done = s
            assert_bitwise(mpm, ctx.particles(), o.particles(), what="%s step %d" % (name, s))
            _grids_equal(ctx, [o.grid_component(c) for c in range(o.nb_component())], "%s step %d" % (name, s))
            assert np.array_equal(ctx.contact_force(), o.contact_force())
            sc, so = ctx.scalars(), o.scalars()
            assert sc[1] == so[1] and sc[2] == so[2]


@pytest.mark.parametrize("name", CONTACT_DECKS)
def test_contact_each_phase_entry_point_bitwise(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob)
    seq = list(PHASES[{0: "usf", 1: "usl", 2: "musl"}[prob.scheme]])
    seq.insert(seq.index("IntegrateMomentum") + 1, "LagrNodContact")
    with mpm.Context(prob, "fp64exact") as ctx:
        for step in (1, 2):
            ctx.begin_step()
            o.phase("begin_step")
            moved = False
            for ph in seq:
                getattr(ctx, ph)()
                o.phase("contact" if ph == "LagrNodContact" else ORACLE[ph])
                moved = moved or ph == "ParticlePositionUpdate"
                what = "%s step %d after %s" % (name, step, ph)
                assert_bitwise(mpm, ctx.particles(), o.particles(), what=what,
                               skip=() if moved else ("xx", "xy", "xz"))
                _grids_equal(ctx, [o.grid_component(c) for c in range(o.nb_component())], what)
            assert np.array_equal(ctx.contact_force(), o.contact_force())
            ctx.calcEnergy()
            o.phase("calc_energy")
            assert ctx.UpdateDT() == (o.phase("update_dt") or o.scalars()[2])
            if prob.smooth_stress:   # every body's pressure smoothed on component 1's grid
                ctx.SmoothStressbyGrid()
                o.phase("smooth_stress")
                assert_bitwise(mpm, ctx.particles(), o.particles(), what="%s step %d smoothed" % (name, step))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_contact_is_refused_outside_the_exact_mode(mpm, prec):
    prob = mpm.read_deck(deck_path(CONTACT_DECKS[0]))
    with pytest.raises(mpm.MpmError, match="reference-order"):
        mpm.Context(prob, prec)
