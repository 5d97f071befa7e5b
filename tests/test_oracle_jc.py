"""Johnson-Cook with heating (john, M3DM4) and with damage and failure (jcf,
M3DM8 + EOS) in the oracle (Constitution.f90:224-237, 278-319, 969-1040,
1144-1212).

Both models read MaterialModel's module variable depeff after their own
update: a particle that yields assigns it, one that does not reads what the
last yielding particle before it (in the loop over bodies and particles,
across steps) left there.  The oracle keeps that variable per run and is
bit-exact with the reference build on the three decks (tests/test_oracle.py:
every field after 1, 10, 50 and 200 steps).  Independently of the reference:
* each deck takes the branches it was written for within 10 steps: updates
  that read another particle's increment, and jcf failures by melting, by
  the failure strain and by damage;
* the reference's fixtures tell the module variable apart from a fresh
  per-particle increment (the oracle with depeff reset to 0 at every update
  misses the fixtures' temperatures);
* only Johnson-Cook particles heat, never cool, and a yielding particle's
  temperature rise is seqv depeff / (rho c) with its own depeff (the plastic
  strain it added), computed from the particle's state;
* a failed jcf particle carries no deviatoric stress.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR, JC_DECKS, deck_path, load_package

# running counts after 10 steps must be positive: [stale reads, melting,
# failure strain, damage]
BRANCHES = {"taylor_john_linear_musl": (True, False, False, False),
            "impact_jcf_eos2_gimp_musl": (True, True, True, True),
            "block_john_sgmp_usf": (True, False, False, False)}


@pytest.mark.parametrize("name", JC_DECKS)
def test_jc_decks_take_their_branches(name, oracle_mod):
    mpm = load_package()
    o = oracle_mod.Oracle(mpm.read_deck(deck_path(name)))
    o.step(10)
    got = tuple(bool(c > 0) for c in o.jc_counts())
    assert got == BRANCHES[name], o.jc_counts()


@pytest.mark.parametrize("name", JC_DECKS)
def test_fixtures_pin_the_module_variable(name, oracle_mod):
    mpm = load_package()
    g = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    prob = mpm.read_deck(deck_path(name))
    a, b = oracle_mod.Oracle(prob), oracle_mod.Oracle(prob)
    b.set_fresh_depeff(True)
    a.step(10)
    b.step(10)
    k = mpm.PF["temp"]
    assert np.array_equal(a.particles()[k], g["s10_temp"])
    assert not np.array_equal(b.particles()[k], g["s10_temp"])


@pytest.mark.parametrize("name", JC_DECKS)
def test_heating_of_yielding_particles(name, oracle_mod):
    mpm = load_package()
    prob = mpm.read_deck(deck_path(name))
    o = oracle_mod.Oracle(prob)
    mtype = np.array([prob.mats[prob.bodies[b].mat - 1].type for b in range(len(prob.bodies))])
    bodies = prob.pbody
    jc = np.isin(mtype[bodies], (4, 7))   # john, jcf
    pf = mpm.PF
    checked = 0
    prev = o.particles()
    for _ in range(10):
        o.step(1)
        ps = o.particles()
        dT = ps[pf["temp"]] - prev[pf["temp"]]
        assert np.all(dT[~jc] == 0) and np.all(dT >= 0)
        dep = ps[pf["epeff"]] - prev[pf["epeff"]]
        alive = ps[pf["fail"]] == 0
        # yielded this step: epeff grew by more than the round-off of the
        # hardening modulus' epeff + 1e-4 - 1e-4 (and its increment is known
        # to 1e-10 from the difference of two epeff)
        own = jc & alive & (dep > 1e-6 * ps[pf["epeff"]])
        rho = ps[pf["mass"]] / ps[pf["vol"]]
        heat = np.array([prob.mats[prob.bodies[b].mat - 1].spec_heat for b in bodies])
        want = ps[pf["seqv"]][own] * dep[own] / rho[own] / heat[own]
        # dT is the difference of two temperatures near 300: a few ulp of 300
        assert np.allclose(dT[own], want, rtol=1e-9, atol=4 * np.spacing(512.0))
        checked += own.sum()
        failed = ps[pf["fail"]] != 0
        for f in ("sdxx", "sdyy", "sdzz", "sdyz", "sdxz", "sdxy"):
            assert np.all(ps[pf[f]][failed & jc & (prev[pf["fail"]] == 0)] == 0)
        prev = ps
    assert checked > 0
