"""The Drucker-Prager return map (M3DM9, Constitution.f90:1214-1330) in the
CPU oracle, pinned on every branch.

* the two block decks take every branch within the 10 steps their reference
  fixtures cover (the branch counts are oracle diagnostics), and their state
  after those steps is the reference's own bit for bit (test_oracle.py);
* after a shear return the shear stress is tau = |k_phi - q_phi sm| (on the
  cone F = tau + q_phi sm - k_phi = 0 while sm is below the apex k_phi/q_phi;
  beyond the apex the reference has no apex return and scales the deviator by
  the negative ratio (k_phi - q_phi sm) / tau_trial, Constitution.f90:1296-
  1307), after the tension cap sm = tenf, in the decks' particle states;
* on single stress points the returned state follows from the plastic flow
  rule alone: the deviatoric direction is kept, the cone is reached, and the
  mean and shear stress move in the ratio the flow rule fixes,
  d(sm) / d(tau) = K q_psi / G (the plastic multiplier cancels), so a wrong
  sign or a dropped term in either update fails.
"""
import math

import numpy as np
import pytest

from conftest import deck_path, load_package

DP_DECKS = {"block_dp_bspline_shear_usf": (1, 2), "block_dp_bspline_pull_musl": (2, 3)}
F32 = lambda v: float(np.float32(v))   # noqa: E731


def _tau(sd):
    sd = np.asarray(sd)
    return np.sqrt(0.5 * (sd[0] ** 2 + sd[1] ** 2 + sd[2] ** 2) + sd[3] ** 2 + sd[4] ** 2 + sd[5] ** 2)


def _dp_material(mpm, prob):
    m = prob.struct().mats[0]
    assert m.type == 10   # MPM_MAT_DP (dpm)
    return m


@pytest.mark.parametrize("name", sorted(DP_DECKS))
def test_dp_decks_take_their_branches_and_land_on_the_surface(name, oracle_mod):
    mpm = load_package()
    prob = mpm.read_deck(deck_path(name))
    mat = _dp_material(mpm, prob)
    o = oracle_mod.Oracle(prob)
    o.step(10)
    last, counts = o.dp_branches()
    for br in DP_DECKS[name]:
        assert counts[br] > 0, (name, counts)
    ps = o.particles()
    sd = ps[[mpm.PF[f] for f in ("sdxx", "sdyy", "sdzz", "sdyz", "sdxz", "sdxy")]]
    sm = ps[mpm.PF["sm"]]
    tau = _tau(sd)
    tenf = min(mat.ten_f, mat.k_fai / mat.q_fai)
    shear = (last == 1) | (last == 2)
    F = tau + mat.q_fai * sm - mat.k_fai
    scale = mat.k_fai + np.abs(mat.q_fai * sm) + tau
    if shear.any():
        cone = np.abs(mat.k_fai - mat.q_fai * sm)
        assert np.max(np.abs(tau - cone)[shear] / scale[shear]) < 1e-13
    tens = last == 3
    if tens.any():
        assert np.all(sm[tens] == tenf)
    # elastic particles lie inside the cone (or below the cap)
    el = last == 0
    assert np.all(F[el] <= 1e-13 * scale[el])
    # plastic strain only grows
    assert np.all(ps[mpm.PF["epeff"]] >= 0)


def _points(mat, G, K):
    """Trial states (sd, sm) aimed at each branch: elastic, shear with
    sm < tenf, shear with sm >= tenf, tension cap."""
    tenf = min(mat.ten_f, mat.k_fai / mat.q_fai)
    d = np.array([0.7, -0.2, -0.5, 0.3, -0.4, 0.25])
    d = d / _tau(d)
    return tenf, [
        (0, d * 0.3 * mat.k_fai, -0.5 * mat.k_fai),
        (1, d * 2.5 * mat.k_fai, -0.5 * mat.k_fai),
        (2, d * 1.5 * mat.k_fai, tenf + 0.05 * mat.k_fai),
        (3, d * 0.05 * mat.k_fai, tenf + 0.5 * mat.k_fai),
        # sm >= tenf, returned beyond the apex k_phi / q_phi (ratio < 0)
        (2, d * 4.0 * mat.k_fai, 2.5 * mat.k_fai / mat.q_fai),
    ]


@pytest.mark.parametrize("qpsi", [0.0, 0.2])
def test_dp_return_map_follows_the_flow_rule(qpsi, oracle_mod):
    mpm = load_package()
    prob = mpm.read_deck(deck_path("block_dp_bspline_pull_musl"))
    mat = _dp_material(mpm, prob)
    mat = type(mat).from_buffer_copy(mat)
    mat.q_psi = qpsi
    E, nu = mat.young, mat.poisson
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    tenf, pts = _points(mat, G, K)
    for want, sd, sm in pts:
        br, sd1, sm1, seqv, dep = oracle_mod.dp_point(mat, E, nu, sd, sm)
        assert br == want, (want, br)
        tau0, tau1 = _tau(sd), _tau(sd1)
        if br == 0:
            assert np.array_equal(sd1, sd) and sm1 == sm and dep == 0.0
            assert seqv == tau0 * F32(math.sqrt(3.0))
            continue
        if br in (1, 2):
            # shear stress of the cone at the new mean stress; the deviatoric
            # direction is kept (flipped beyond the apex)
            signed = mat.k_fai - mat.q_fai * sm1
            assert abs(tau1 - abs(signed)) < 1e-12 * (mat.k_fai + abs(mat.q_fai * sm1))
            np.testing.assert_allclose(sd1 / tau1, np.sign(signed) * sd / tau0, rtol=0, atol=1e-14)
            # flow rule: the signed shear stress drops by G dlam, sm by K q_psi dlam
            dlam = (tau0 - signed) / G
            assert dlam > 0
            assert abs((sm - sm1) - K * qpsi * dlam) <= 1e-9 * abs(sm) + 1e-12 * mat.k_fai
            # effective plastic strain of the plastic potential's gradient,
            # with the reference's single-precision 1/3 and 2/9
            want_dep = dlam * math.sqrt(F32(1.0 / 3.0) + F32(2.0 / 9.0) * qpsi ** 2)
            assert abs(dep - want_dep) <= 1e-9 * want_dep
            assert abs(seqv - signed * F32(math.sqrt(3.0))) <= 1e-12 * abs(seqv)
        else:
            # tension cap: mean stress on the cap, deviator unchanged
            assert sm1 == tenf
            assert np.array_equal(sd1, sd)
            want_dep = (sm - tenf) / K * F32(1.0 / 3.0) * F32(math.sqrt(2.0))
            assert abs(dep - want_dep) <= 1e-12 * want_dep
