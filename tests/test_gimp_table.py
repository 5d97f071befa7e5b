"""The GIMP node lists of the device's reference-order path
(mpm3d-f90_amd/csrc/gimp_tab.h, a compile-time table) equal the oracle's
FindInflNode lists (Grid.f90:621-847) for all 27 axis cases, entry by entry.
The table is printed by a host program compiled from the header here."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HDR_DIR = os.path.join(ROOT, "mpm3d-f90_amd", "csrc")
PROG = r'''
#include <cstdio>
#include "gimp_tab.h"
int main() {
  constexpr GimpTab T = make_gimp_tab();
  const int cs[3] = {-1, 0, 2};
  for (int zi = 0; zi < 3; zi++)
    for (int yi = 0; yi < 3; yi++)
      for (int xi = 0; xi < 3; xi++) {
        const int k = gimp_case(cs[zi], cs[yi], cs[xi]);
        std::printf("%d %d %d %d", cs[zi], cs[yi], cs[xi], T.n[k]);
        for (int j = 0; j < T.n[k]; j++) std::printf(" %d %d %d", T.off[k][j][0], T.off[k][j][1], T.off[k][j][2]);
        std::printf("\n");
      }
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_gimp_node_lists_equal_the_oracle(tmp_path, oracle_mod):
    src = tmp_path / "tab.cpp"
    src.write_text(PROG)
    exe = tmp_path / "tab"
    subprocess.run(["g++", "-std=c++17", "-I" + HDR_DIR, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    L = oracle_mod.lib()
    L.oracle_gimp_offsets.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.oracle_gimp_offsets.restype = C.c_int
    seen = 0
    sizes = set()
    for line in out:
        if not line.strip():
            continue
        v = [int(t) for t in line.split()]
        zc, yc, xc, n = v[:4]
        dev = np.array(v[4:]).reshape(n, 3)
        buf = (C.c_int * 81)()
        m = L.oracle_gimp_offsets(zc, yc, xc, buf)
        ref = np.array(buf[:3 * m]).reshape(m, 3)
        assert n == m and np.array_equal(dev, ref), (zc, yc, xc)
        # every list holds distinct nodes of the 4x4x4 block around the cell
        assert len({tuple(r) for r in dev}) == n and dev.min() >= -1 and dev.max() <= 2
        sizes.add(n)
        seen += 1
    assert seen == 27
    assert sizes == {8, 12, 18, 27}
