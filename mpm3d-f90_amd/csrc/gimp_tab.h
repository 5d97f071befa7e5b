// gimp_tab.h -- the node lists of the GIMP influence search, as a
// compile-time table (host and device).
#pragma once

// FindInflNode's node lists (Grid.f90:621-847): for the case of each axis
// (particle within a quarter cell of the low node: -1, of the high node: 2,
// else 0) the offsets of the nodes from the cell's first node, in order.
struct GimpTab {
  signed char off[27][27][3];
  int n[27];
};

constexpr int gimp_case(int zc, int yc, int xc) {
  return ((zc == -1 ? 0 : zc == 0 ? 1 : 2) * 3 + (yc == -1 ? 0 : yc == 0 ? 1 : 2)) * 3 + (xc == -1 ? 0 : xc == 0 ? 1 : 2);
}

constexpr GimpTab make_gimp_tab() {
  GimpTab t{};
  const int cs[3] = {-1, 0, 2};
  for (int zi = 0; zi < 3; zi++)
    for (int yi = 0; yi < 3; yi++)
      for (int xi = 0; xi < 3; xi++) {
        const int zc = cs[zi], yc = cs[yi], xc = cs[xi];
        const int k = gimp_case(zc, yc, xc);
        int n = 0;
        auto add = [&](int a, int b, int c) {
          t.off[k][n][0] = (signed char)a;
          t.off[k][n][1] = (signed char)b;
          t.off[k][n][2] = (signed char)c;
          n++;
        };
        // the cell's own 8 nodes, CellsNode order
        add(0, 0, 0); add(1, 0, 0); add(1, 1, 0); add(0, 1, 0);
        add(0, 0, 1); add(1, 0, 1); add(1, 1, 1); add(0, 1, 1);
        const int e = yc == -1 ? -1 : 2;   // the y layer added (when yc /= 0)
        const int fx = xc == -1 ? -1 : 2;  // the x layer added (when xc /= 0)
        if (zc != 0) {
          add(0, 0, zc); add(1, 0, zc); add(1, 1, zc); add(0, 1, zc);
          if (yc == -1) {
            add(0, -1, zc); add(1, -1, zc); add(0, -1, 0); add(1, -1, 0); add(0, -1, 1); add(1, -1, 1);
          } else if (yc == 2) {
            add(1, 2, zc); add(0, 2, zc); add(1, 2, 0); add(0, 2, 0); add(1, 2, 1); add(0, 2, 1);
          }
          if (xc != 0) {
            if (yc != 0) {
              add(fx, 0, 0); add(fx, 1, 0); add(fx, 0, 1); add(fx, 1, 1); add(fx, 0, zc); add(fx, 1, zc);
              add(fx, e, zc); add(fx, e, 0); add(fx, e, 1);
            } else {
              add(fx, 0, zc); add(fx, 1, zc); add(fx, 0, 0); add(fx, 1, 0); add(fx, 0, 1); add(fx, 1, 1);
            }
          }
        } else {
          if (yc == -1) {
            add(0, -1, 0); add(1, -1, 0); add(0, -1, 1); add(1, -1, 1);
          } else if (yc == 2) {
            add(1, 2, 0); add(0, 2, 0); add(1, 2, 1); add(0, 2, 1);
          }
          if (xc != 0) {
            add(fx, 0, 0); add(fx, 1, 0); add(fx, 0, 1); add(fx, 1, 1);
            if (yc != 0) {
              add(fx, e, 0); add(fx, e, 1);
            }
          }
        }
        t.n[k] = n;
      }
  return t;
}

