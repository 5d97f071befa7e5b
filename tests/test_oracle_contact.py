"""Contact (cont lagr: Lagr_NodContact, update_step.f90:1241-1376) in the oracle.

The oracle is bit-exact with the reference build on the contact decks
(tests/test_oracle.py: every particle field, the grids of all components and
the contact force after 1, 10, 50 and 200 steps).  Independently of the
reference:
* each deck takes the branches it was written for within 10 steps
  (frictionless contact; friction with sticking and with slipping nodes);
* the contact pass only moves momentum between the two components: their
  sum at every node is unchanged (to round-off);
* without friction the exchange is along the contact normal and leaves no
  relative normal velocity: (p1 m2 - p2 m1) . n = 0 afterwards, n the
  direction of the exchange.
"""
import numpy as np
import pytest

from conftest import CONTACT_DECKS, deck_path, load_package

BRANCHES = {"blocks_contact_linear_usf": (True, False, False),
            "blocks_contact_bspline_musl": (False, True, True),
            "blocks_contact_gimp_usl": (False, True, True),
            "blocks_contact_smoo_gimp_musl": (False, True, True)}


@pytest.mark.parametrize("name", CONTACT_DECKS)
def test_contact_decks_take_their_branches(name, oracle_mod):
    mpm = load_package()
    o = oracle_mod.Oracle(mpm.read_deck(deck_path(name)))
    o.step(10)
    got = tuple(bool(c > 0) for c in o.contact_counts())
    assert got == BRANCHES[name], o.contact_counts()
    assert np.abs(o.contact_force()).max() > 0


def _until_contact(o, scheme):
    """one step up to IntegrateMomentum (the contact pass comes next)"""
    o.phase("begin_step")
    o.phase("grid_momentum_initial")
    if scheme == 0:
        o.phase("apply_bc")
        o.phase("particle_stress_update")
    o.phase("grid_momentum_update")
    o.phase("integrate_momentum")


def _finish(o, scheme):
    o.phase("particle_position_update")
    if scheme == 2:
        o.phase("grid_momentum_musl")
        o.phase("apply_bc")
    if scheme != 0:
        o.phase("particle_stress_update")
    o.phase("calc_energy")
    o.phase("update_dt")


@pytest.mark.parametrize("name", CONTACT_DECKS)
def test_contact_pass_conserves_momentum(name, oracle_mod):
    mpm = load_package()
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob)
    free = np.all(o.fix() == 0, axis=0)
    checked = 0
    for _ in range(6):
        _until_contact(o, prob.scheme)
        a = [o.grid_component(c) for c in (0, 1)]
        o.phase("contact")
        b = [o.grid_component(c) for c in (0, 1)]
        scale = np.abs(a[0][1:4]).max() + np.abs(a[1][1:4]).max()
        for f in (1, 2, 3):    # PXg
            before, after = a[0][f] + a[1][f], b[0][f] + b[1][f]
            assert np.all(np.abs(after - before)[free] <= 1e-13 * scale)
        moved = free & np.any(b[0][1:4] != a[0][1:4], axis=0)
        checked += moved.sum()
        if not prob.fricfa:
            dp = b[0][1:4] - a[0][1:4]
            n = dp[:, moved] / np.linalg.norm(dp[:, moved], axis=0)
            m1, m2 = b[0][0][moved], b[1][0][moved]
            rel = b[0][1:4][:, moved] * m2 - b[1][1:4][:, moved] * m1
            approach = a[0][1:4][:, moved] * m2 - a[1][1:4][:, moved] * m1
            assert np.all(np.abs((rel * n).sum(0)) <= 1e-12 * np.abs((approach * n).sum(0)))
        _finish(o, prob.scheme)
    assert checked > 0
