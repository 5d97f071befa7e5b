"""Total-Lagrangian MPM (TLmp, MPM3D.f90:162-250) on the device, in the
reference-order FP64 mode.

The cells (and SGMP's auxiliary coordinates) are those of the initial
positions, the nodal mass is the first step's and kept, the stress update
is TLConstitution's neo-Hookean first Piola-Kirchhoff model, and the force
is PK over the initial volume.  The device is bit-identical to the oracle
(with the correctly rounded ln J it shares) after 50 and 200 steps and after
every phase of the first two steps, within 1e-13 of the reference's own
10-step fixtures (its glibc logarithm, as for the updated-Lagrangian
neo-Hookean decks), and the other precisions refuse such problems.  With
contact (two components that meet at a node layer of the initial
configuration) every component's grid and the contact force are compared
too.
"""
import os

import numpy as np
import pytest

from conftest import GGROUPS, GOLDEN_DIR, PGROUPS, TL_CONTACT_DECKS, TL_DECKS, deck_path, group_errors, load_package
from test_gpu_exact import assert_bitwise

pytestmark = pytest.mark.gpu

TL_GROUPS = dict(PGROUPS, pk=["pk%d" % (i + 1) for i in range(9)], vol0=["vol0"], xo=["xo", "yo", "zo"])
DECKS = TL_DECKS + TL_CONTACT_DECKS


def _components_equal(ctx, o, what):
    for c in range(o.nb_component()):
        assert np.array_equal(ctx.grid_component(c), o.grid_component(c), equal_nan=True), "%s component %d" % (
            what, c + 1)
    assert np.array_equal(ctx.contact_force(), o.contact_force()), what


@pytest.fixture(scope="module")
def mpm():
    return load_package()


@pytest.mark.parametrize("name", DECKS)
def test_tl_within_1e13_of_reference_fixture(mpm, name):
    g = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    prob = mpm.read_deck(deck_path(name))
    with mpm.Context(prob, "fp64exact") as ctx:
        ctx.step(10)
        ps = ctx.particles()
        ref = ps.copy()
        for key in g.files:
            if key.startswith("s10_"):
                ref[mpm.PF[key[4:]]] = g[key]
        pe = group_errors(ps, ref, mpm.PF_NAMES, TL_GROUPS)
        ge = group_errors(ctx.grid(), g["g10"], mpm.GF_NAMES, GGROUPS)
        assert max(pe.values()) <= 1e-13 and max(ge.values()) <= 1e-13, (pe, ge)
        assert np.abs(ps[mpm.PF["pk1"]]).max() > 0
        assert np.array_equal(ctx.icell(), g["icell10"])
        if "gc10" in g.files:
            for c in range(1, prob.nb_component):
                ge = group_errors(ctx.grid_component(c), g["gc10"][c - 1], mpm.GF_NAMES, GGROUPS)
                assert max(ge.values()) <= 1e-13, (c, ge)
            f = ctx.contact_force()
            assert np.abs(f - g["cont10"]).max() <= 1e-13 * np.abs(g["cont10"]).max()


@pytest.mark.parametrize("name", DECKS)
def test_tl_matches_oracle_bitwise_to_200_steps(mpm, oracle_mod, name):
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
            _components_equal(ctx, o, "%s step %d" % (name, s))
            sc, so = ctx.scalars(), o.scalars()
            assert sc[1] == so[1] and sc[2] == so[2]


TL_PHASES = {"usf": [("GridMomentumInitial", "tl_grid_momentum_initial"), ("ApplyBoundaryConditions", "apply_bc"),
                     ("ParticleStressUpdate", "tl_particle_stress_update"),
                     ("GridMomentumUpdate", "tl_grid_momentum_update"), ("IntegrateMomentum", "integrate_momentum"),
                     ("ParticlePositionUpdate", "particle_position_update")],
             "usl": [("GridMomentumInitial", "tl_grid_momentum_initial"),
                     ("GridMomentumUpdate", "tl_grid_momentum_update"), ("IntegrateMomentum", "integrate_momentum"),
                     ("ParticlePositionUpdate", "particle_position_update"),
                     ("ParticleStressUpdate", "tl_particle_stress_update")],
             "musl": [("GridMomentumInitial", "tl_grid_momentum_initial"),
                      ("GridMomentumUpdate", "tl_grid_momentum_update"), ("IntegrateMomentum", "integrate_momentum"),
                      ("ParticlePositionUpdate", "particle_position_update"),
                      ("GridMomentumMUSL", "grid_momentum_musl"), ("ApplyBoundaryConditions", "apply_bc"),
                      ("ParticleStressUpdate", "tl_particle_stress_update")]}


@pytest.mark.parametrize("name", DECKS)
def test_tl_each_phase_entry_point_bitwise(mpm, oracle_mod, name):
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob, log_cr=True)
    o.phase("tl_grid_mass_initial")   # GridMassInitial, before the loop (MPM3D.f90:164)
    seq = list(TL_PHASES[{0: "usf", 1: "usl", 2: "musl"}[prob.scheme]])
    if prob.contact:   # Lagr_NodContact after IntegrateMomentum (MPM3D.f90:186-190)
        seq.insert(seq.index(("IntegrateMomentum", "integrate_momentum")) + 1, ("LagrNodContact", "contact"))
    with mpm.Context(prob, "fp64exact") as ctx:
        for step in (1, 2):
            ctx.begin_step()
            o.phase("begin_step")
            moved = False
            for dev, orc in seq:
                getattr(ctx, dev)()
                o.phase(orc)
                moved = moved or dev == "ParticlePositionUpdate"
                assert_bitwise(mpm, ctx.particles(), o.particles(), ctx.grid(), o.grid(),
                               "%s step %d after %s" % (name, step, dev), skip=() if moved else ("xx", "xy", "xz"))
                if prob.contact:
                    _components_equal(ctx, o, "%s step %d after %s" % (name, step, dev))
            ctx.calcEnergy()
            o.phase("calc_energy")
            assert ctx.UpdateDT() == (o.phase("update_dt") or o.scalars()[2])
            if prob.smooth_stress:
                ctx.SmoothStressbyGrid()
                o.phase("smooth_stress")
                assert_bitwise(mpm, ctx.particles(), o.particles(), what="%s step %d smoothed" % (name, step))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tl_is_refused_outside_the_exact_mode(mpm, prec):
    prob = mpm.read_deck(deck_path(TL_DECKS[0]))
    with pytest.raises(mpm.MpmError, match="reference-order"):
        mpm.Context(prob, prec)


def test_tl_and_contact_in_a_batched_engine_with_graphs_and_resorts(mpm, oracle_mod):
    """Total-Lagrangian, contact and updated-Lagrangian decks in one engine,
    stepped through captured graphs with a re-sort every 8 steps (the PK
    block and the component grids travel with their particles and
    instances): each instance is bitwise its own oracle run."""
    from conftest import CONTACT_DECKS
    names = TL_DECKS + TL_CONTACT_DECKS + CONTACT_DECKS + ["bar_j2_gimp_musl", "block_smoo_sgmp_usf"]
    probs = {n: mpm.read_deck(deck_path(n)) for n in names}
    comps = {n: (p.nb_component if p.contact else 1) for n, p in probs.items()}
    eng = mpm.Engine(max_particles=sum(p.n for p in probs.values()) + len(probs) * 256,
                     max_nodes=sum(p.nb_node * comps[n] for n, p in probs.items()) + len(probs) * 256,
                     max_instances=len(probs), max_materials=sum(len(p.mats) for p in probs.values()),
                     max_bodies=sum(len(p.bodies) for p in probs.values()), precision="fp64exact")
    eng.apply_tuning(sort_every=8)
    ids = {n: eng.add(p) for n, p in probs.items()}
    for _ in range(3):
        eng.step(8, graph=True, graph_steps=4)
        eng.sync()
    for n, p in probs.items():
        o = oracle_mod.Oracle(p, log_cr=True)
        o.step(24)
        assert_bitwise(mpm, eng.particles(ids[n]), o.particles(), eng.grid(ids[n]), o.grid(), n + " batched")
        assert np.array_equal(eng.icell(ids[n]), o.icell())
    eng.close()
