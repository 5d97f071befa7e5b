"""The .mpm reader reproduces the reference's input phase exactly
(FFI.f90 tokenizer, DataIn.f90 semantics, Grid.f90 SetGridData, SetDT).

The golden fixtures hold the reference's own state right after InputPara +
calcEnergy (written by tests/golden/make_golden.py from the Fortran build);
the reader must match it bit for bit."""
import os

import numpy as np
import pytest

from conftest import ALL_DECKS, CONTACT_DECKS, GOLDEN_DIR, TL_DECKS, deck_path, load_package

DECKS = ALL_DECKS + CONTACT_DECKS + TL_DECKS

S0_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "vol", "sig_y", "ie", "cp", "temp", "lt"]


@pytest.mark.parametrize("name", DECKS)
def test_reader_matches_reference_input_phase(name):
    mpm = load_package()
    p = mpm.read_deck(deck_path(name))
    g = np.load(os.path.join(GOLDEN_DIR, name + ".npz"))
    for f in S0_FIELDS + ["xo", "yo", "zo", "vol0"]:
        if f in S0_FIELDS or "s0_" + f in g.files:
            assert np.array_equal(p.field(f), g["s0_" + f]), f
    sc = g["scal0"]
    assert p.dt == sc[0]
    assert p.cutoff == sc[7]
    assert p.dcell == sc[8]
    assert np.array_equal(p.span, sc[9:15])
    assert p.dtscale == sc[15]
    assert (p.ncx, p.ncy, p.ncz) == tuple(g["ncell"])


def write(tmp_path, text, name="t.mpm"):
    path = tmp_path / name
    path.write_text(text)
    return str(path)


BASE = """mpm3d *** t
nbco 1
nbbo 1
nbmp 8
nmat 1
spx 0 1.0
spy 0 1.0
spz 0 1.0
dcell 0.25
material
 1 elas 1000 1e6 0.25
particle bloc 1 1 1000 0.25 0.25 0.25 0.25 2 2 2
endi
"""


def test_every_real_is_rounded_to_fp32(tmp_path):
    mpm = load_package()
    text = BASE.replace("dcell 0.25", "dcell 0.1").replace("spx 0 1.0", "spx 0 1.1")
    text = text.replace("endi", "dtscale 0.3\nendtime 1.7d-3\nendi")
    p = mpm.read_deck(write(tmp_path, text))
    assert p.dcell == float(np.float32(0.1))
    assert p.dtscale == float(np.float32(0.3))
    assert p.endtime == float(np.float32(1.7e-3))
    assert p.mats[0].young == float(np.float32(1e6))


def test_tokenizer_comments_tabs_commas_case(tmp_path):
    mpm = load_package()
    text = BASE.replace("spx 0 1.0", "SPX\t0 ,  1.0   ! comment, ignored 9 9 9")
    text = text.replace(" 1 elas 1000 1e6 0.25", " 1 ElAs 1000,1E6, 0.25")
    text = text.replace("dcell 0.25", "! full comment line\n\ndcel\t0.25")
    p = mpm.read_deck(write(tmp_path, text))
    assert p.span[1] == 1.0 and p.mats[0].young == 1e6 and p.mats[0].poisson == float(np.float32(0.25))


def test_values_may_continue_on_following_lines(tmp_path):
    mpm = load_package()
    text = BASE.replace("spx 0 1.0", "spx\n0\n 1.0")
    p = mpm.read_deck(write(tmp_path, text))
    assert p.span[0] == 0.0 and p.span[1] == 1.0


def test_default_scheme_is_musl_and_flags(tmp_path):
    mpm = load_package()
    p = mpm.read_deck(write(tmp_path, BASE))
    assert p.scheme == 2 and p.shape == 0
    p = mpm.read_deck(write(tmp_path, BASE.replace("endi", "usf on\nbimp\nendi")))
    assert p.scheme == 0 and p.shape == 2
    p = mpm.read_deck(write(tmp_path, BASE.replace("endi", "usl on\ngimp\nQuLo 0.5\nDamp 0.25\nendi")))
    assert p.scheme == 1 and p.shape == 1 and p.quasi_load and p.quasi_damp
    assert p.quasi_load_time == 0.5 and p.damp_factor == 0.25


@pytest.mark.parametrize("bad,msg", [
    ("cont lagr 0.1 1", "contact"),
    ("TLmp", "total-Lagrangian"),
    ("frob 1 2", "invalid keyword"),
    ("spx a 1", "non-numeric"),
])
def test_unsupported_or_bad_input_is_rejected(tmp_path, bad, msg):
    mpm = load_package()
    with pytest.raises(mpm.MpmError) as e:
        mpm.read_deck(write(tmp_path, BASE.replace("endi", bad + "\nendi")))
    assert e.value.code == -3
    assert msg in str(e.value)


@pytest.mark.parametrize("mat,extra,msg", [
    # models that call seleos need an EOS; JWL only with HE burn
    (" 1 jcf 8.93e-3 117e3 0.35 90 292 0.3 0.025 1.0 20 1800 385 1e-3 2.0", "", "equation of state"),
    (" 1 sjc 8.93e-3 117e3 0.35 157 425 1.0 0.0 1e-3", "", "equation of state"),
    (" 1 hiex 1.63e-3 6930", "", "equation of state"),
    (" 1 sjc 8.93e-3 117e3 0.35 157 425 1.0 0.0 1e-3", "seos 1 3 3.7e5 3.2e3 4.15 0.95 0.3 6993", "JWL"),
    (" 1 null 1000 2.2e9 1e-3", "seos 1 5 1", "EOS type"),
])
def test_material_and_eos_combinations_rejected(tmp_path, mat, extra, msg):
    mpm = load_package()
    text = BASE.replace(" 1 elas 1000 1e6 0.25", mat).replace("endi", extra + "\nendi")
    with pytest.raises(mpm.MpmError) as e:
        mpm.read_deck(write(tmp_path, text))
    assert msg in str(e.value)


def test_johnson_cook_parameters(tmp_path):
    """john and jcf (DataIn.f90:427-478): density, E, nu, A, B, n, C, m,
    room and melting temperature, specific heat, reference strain rate and
    (jcf) the failure strain, in that order, each read as a default REAL"""
    mpm = load_package()
    vals = [8.93e-3, 117e3, 0.35, 90.0, 292.0, 0.3, 0.025, 1.0, 20.0, 1800.0, 385.0, 1e-3]
    text = BASE.replace(" 1 elas 1000 1e6 0.25", " 1 john " + " ".join(repr(v) for v in vals))
    m = mpm.read_deck(write(tmp_path, text)).mats[0]
    got = [m.density, m.young, m.poisson, m.yield0, m.b_jc, m.n_jc, m.c_jc, m.m_jc, m.roomt, m.melt, m.spec_heat,
           m.epso]
    assert m.type == 4 and got == [float(np.float32(v)) for v in vals]
    text = BASE.replace(" 1 elas 1000 1e6 0.25", " 1 jcf " + " ".join(repr(v) for v in vals) + " 0.25")
    m = mpm.read_deck(write(tmp_path, text.replace("endi", "seos 1 2 3940 1.49 1.96 0\nendi"))).mats[0]
    assert m.type == 7 and m.epf == 0.25 and m.melt == 1800.0 and m.eos_type == 2 and m.epso == float(np.float32(1e-3))


def test_eos_constants_and_lighting_time(tmp_path):
    """SetEos's Gruneisen constants (DataIn.f90:1123-1138) from FP32-read
    c0, lambda, gamma0, and Initial's lighting time through a default REAL
    (DataIn.f90:1252-1260); particles start at 293 degrees"""
    mpm = load_package()
    text = BASE.replace(" 1 elas 1000 1e6 0.25", " 1 sjc 8.93e-3 117e3 0.35 157 425 1.0 0.0 1e-3")
    p = mpm.read_deck(write(tmp_path, text.replace("endi", "seos 1 2 3940 1.49 1.96 7\nendi")))
    m = p.mats[0]
    f = lambda v: float(np.float32(v))
    c1 = f(8.93e-3) * (f(3940) * f(3940))
    assert (m.eos_type, m.ceos[0], m.ceos[1]) == (2, c1, c1 * (2 * f(1.49) - 1))
    assert m.ceos[2] == c1 * (f(1.49) - 1) * (3 * f(1.49) - 1)
    assert (m.ceos[3], m.ceos[4], m.ceos[9]) == (f(8.93e-3) * f(1.96), f(1.96), 7.0)
    assert np.all(p.field("temp") == 293.0) and np.all(p.field("lt") == 1e7)
    text = BASE.replace(" 1 elas 1000 1e6 0.25", " 1 hiex 1.63e-3 6930")
    p = mpm.read_deck(write(tmp_path, text.replace("endi", "seos 1 3 3.7e5 3.2e3 4.15 0.95 0.3 6993\n"
                                                           "deto 0 0 0\ndeto 1 1 1\nendi")))
    x, y, z = p.field("x"), p.field("y"), p.field("z")
    d = np.minimum(np.sqrt(x * x + y * y + z * z), np.sqrt((x - 1) ** 2 + (y - 1) ** 2 + (z - 1) ** 2))
    lt = (d / f(6930)).astype(np.float32).astype(np.float64)
    assert np.array_equal(p.field("lt"), lt) and p.mats[0].det_speed == 6930.0


def test_too_many_particles_rejected(tmp_path):
    mpm = load_package()
    with pytest.raises(mpm.MpmError):
        mpm.read_deck(write(tmp_path, BASE.replace("nbmp 8", "nbmp 7")))


def test_generators_closed_form(tmp_path):
    """sphe / blod / cicl / poin particle sets against the generator rules
    of DataIn.f90:563-846, evaluated independently here."""
    mpm = load_package()
    f32 = lambda v: float(np.float32(v))
    dp, r = f32(0.1), 3
    text = """mpm3d *** gen
nbco 1
nbbo 4
nbmp {n}
nmat 1
spx -1 1
spy -1 1
spz -1 1
dcell 0.2
material
 1 elas 1000 1e6 0.25
particle sphe 1 1 1000 0.1 0.0 0.0 0.0 3
particle cicl 1 1 1000 0.1 0.0 0.0 0.5 3
particle blod 1 1 1000 0.1 -0.5 -0.5 -0.9 10 10 2 0.0 0.0 0.0 3
particle poin 1 2
 1000 1 0.5 0.1 0.2 0.3
 1001 1 0.25 -0.1 -0.2 -0.3
endi
"""
    k = (np.arange(1 - r, r + 1) - 1) * dp + 0.5 * dp
    X, Y, Z = np.meshgrid(k, k, k, indexing="ij")
    sphere = (X ** 2 + Y ** 2 + Z ** 2) < (r * dp) ** 2
    XX, YY = np.meshgrid(k, k, indexing="ij")
    disc = (XX ** 2 + YY ** 2) < (r * dp) ** 2
    bx = (np.arange(1, 11) - 1) * dp + (f32(-0.5) + dp * 0.5)
    BZ, BY, BX = np.meshgrid((np.arange(1, 3) - 1) * dp + (f32(-0.9) + dp * 0.5), bx, bx, indexing="ij")
    hole = (BX ** 2 + BY ** 2) >= (r * dp) ** 2
    n = int(sphere.sum() + disc.sum() + hole.sum() + 2)
    p = mpm.read_deck(write(tmp_path, text.format(n=n)))
    ns, nd = int(sphere.sum()), int(disc.sum())
    assert np.array_equal(p.field("x")[:ns], X[sphere])
    assert np.array_equal(p.field("z")[ns:ns + nd], np.full(nd, 0.5 + dp * 0.5))
    assert np.array_equal(p.field("x")[ns + nd:n - 2], BX[hole])
    assert p.field("mass")[-2] == 0.5 and p.field("z")[-1] == float(np.float32(-0.3))
    assert np.allclose(p.field("mass")[:ns], 1000 * dp ** 3, rtol=1e-6)
    assert list(p.pbody[[0, ns, ns + nd, n - 1]]) == [0, 1, 2, 3]
