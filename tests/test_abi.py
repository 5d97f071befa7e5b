"""The C-ABI library loads on a machine without a GPU and exports every
entry point the public headers declare (no compute calls here)."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT, load_package


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(mpm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_phases():
    names = declared_functions()
    for n in ["mpm_grid_momentum_initial", "mpm_apply_boundary_conditions", "mpm_particle_stress_update",
              "mpm_grid_momentum_update", "mpm_integrate_momentum", "mpm_particle_position_update",
              "mpm_grid_momentum_musl", "mpm_calc_energy", "mpm_update_dt", "mpm_deck_read"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    mpm = load_package()
    L = ctypes.CDLL(mpm._lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    mpm = load_package()
    L = mpm._lib.lib()
    for n in declared_functions():
        f = getattr(L, n)
        assert f.argtypes is not None, "no ctypes signature for %s" % n


def test_version_string():
    mpm = load_package()
    assert b"gfx950" in mpm._lib.lib().mpm_version()


def test_missing_deck_is_an_io_error(tmp_path):
    mpm = load_package()
    with pytest.raises(mpm.MpmError) as e:
        mpm.read_deck(str(tmp_path / "nope.mpm"))
    assert e.value.code == -2
