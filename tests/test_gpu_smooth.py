"""SmoothStressbyGrid (smoo) on the device, against a closed form.

The oracle, the reference's fixtures and the reference build pin the
smoothed pressure on the smoothing decks (tests/test_oracle.py; the device is
compared with them in every precision by the parametrized parity tests).
Here the kernels are checked without the oracle: a body at rest under a
uniform mean stress -c keeps its stress through the first USF step (the
stress update runs before any grid velocity exists; the particles then move
under the unbalanced surface forces, which the GIMP and B-spline weights
follow), every node it maps gets Gpre = sum N c V / sum N V = c, and the
gather returns c again wherever
a particle's weights sum to one over mapped nodes (every particle for the
linear, GIMP and B-spline shapes; for SGMP the particles whose auxiliary
cells have all their nodes inside the body, where every cell centre is the
mean of 8 nodes with Mg > CutOff).
"""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

PRESSURE = 2.5e5


def _deck(tmp_path, shape_kw, scheme_kw):
    text = """mpm3d *** uniform pressure
nbco 1
nbbo 1
nbmp 1728
nmat 1
spx 0 2.4
spy 0 2.4
spz 0 2.4
dcell 0.2
{shape}
{scheme}
smoo
material
 1 elas 1000 1e6 0.25
particle bloc 1 1 1000 0.1 0.6 0.6 0.6 12 12 12
endi
""".format(shape=shape_kw, scheme=scheme_kw)
    p = tmp_path / "uniform.mpm"
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp64exact"])
@pytest.mark.parametrize("shape,scheme", [("", "usf on"), ("gimp", "usf on"), ("bimp", "usf on"),
                                          ("sgmp", "usf on")])
def test_uniform_pressure_is_smoothed_to_itself(tmp_path, prec, shape, scheme):
    mpm = load_package()
    prob = mpm.read_deck(_deck(tmp_path, shape, scheme))
    assert prob.smooth_stress == 1
    prob.pstate[mpm.PF["sm"]] = -PRESSURE
    with mpm.Context(prob, prec) as ctx:
        ctx.step(1)
        ps = ctx.particles()
    assert np.all(ps[mpm.PF["sm"]] == -PRESSURE)          # the stress update saw no grid velocity
    assert np.any(ps[mpm.PF["x"]] != prob.pstate[mpm.PF["x"]])   # and the particles moved
    sp = ps[mpm.PF["spre"]]
    tol = 1e-6 if prec == "fp32" else 1e-13
    if shape == "sgmp":
        x = ps[[mpm.PF["x"], mpm.PF["y"], mpm.PF["z"]]]
        inner = np.all((x > 0.6 + 0.4) & (x < 1.8 - 0.4), axis=0)
        assert inner.sum() >= 64
        sp = sp[inner]
    assert np.max(np.abs(sp - PRESSURE)) <= tol * PRESSURE, np.max(np.abs(sp - PRESSURE))


def test_smoothing_is_off_without_smoo(tmp_path):
    """a deck without 'smoo' leaves Spre as uploaded"""
    mpm = load_package()
    p = _deck(tmp_path, "bimp", "usf on")
    text = open(p).read().replace("smoo\n", "")
    open(p, "w").write(text)
    prob = mpm.read_deck(p)
    assert prob.smooth_stress == 0
    prob.pstate[mpm.PF["spre"]] = 7.0
    with mpm.Context(prob, "fp64") as ctx:
        ctx.step(2)
        ctx.SmoothStressbyGrid()   # a no-op for a problem without smoothing
        assert np.all(ctx.particles()[mpm.PF["spre"]] == 7.0)
