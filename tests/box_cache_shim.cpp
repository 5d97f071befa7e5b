// Host build of the node-box cache rules (mpm3d-f90_amd/csrc/boxcheck.h) for
// tests/test_box_cache.py: an entry is the device's 22-int record.
#include <cstring>

#include "../mpm3d-f90_amd/csrc/boxcheck.h"

struct HBox {
  int x0, y0, z0, bx, by, bz, n, ok, inner;
};
struct HBox2 {
  HBox b[2];
  int split, n, ok, inner;
};
static_assert(sizeof(HBox2) == 22 * sizeof(int), "entry record");

static HBox2 load(const int* e) {
  HBox2 c;
  std::memcpy(&c, e, sizeof c);
  return c;
}

extern "C" int entry_consistent(const int* e, int cap, int mx, int my, int mz) {
  return box_entry_consistent(load(e), cap, mx, my, mz) ? 1 : 0;
}

// shape: 0 linear, 1 GIMP, 2 B-spline (and SGMP's node stencil)
extern "C" int entry_holds(const int* e, int shape, int jl, int cx, int cy, int cz) {
  const HBox2 c = load(e);
  if (shape == 0) return box_entry_holds<0, 1>(c, jl, cx, cy, cz) ? 1 : 0;
  if (shape == 1) return box_entry_holds<-1, 2>(c, jl, cx, cy, cz) ? 1 : 0;
  return box_entry_holds<-1, 1>(c, jl, cx, cy, cz) ? 1 : 0;
}
