/* boxcheck.h -- when a cached node box may stand in for a block's own.
 *
 * The particle kernels keep each block's node box (one or two boxes of grid
 * nodes that hold its particles' stencils, laid out in the block's LDS tile)
 * in a per-block cache entry, keyed by the block's first particle slot, and
 * take the entry in a later kernel or step instead of reducing the box
 * again.  The entry may have been written for another block size, another
 * tile capacity or an earlier particle order, so it is only taken when these
 * rules hold; then every stencil node of every active particle maps to a
 * tile slot of its side's box, inside the tile, and every node the tile
 * loads and flushes is a grid node:
 *
 *   entry: consistent (ok, inner, n = n0 + n1 <= tile capacity), each
 *          non-empty side box whole (n = bx by bz) and inside the node range
 *          [0, mx] x [0, my] x [0, mz] of the instance's grid;
 *   particle (active): its side (slot index jl >= split) has a non-empty box
 *          that holds its cell's stencil, cell + lo .. cell + hi per axis.
 *
 * Written for host and device: the kernels (kernels.hip) and a CPU check of
 * the rules (tests/test_box_cache.py) include the same code.  B2 is any type
 * with the device entry's fields: b[2] (x0, y0, z0, bx, by, bz, n, ok, inner),
 * split, n, ok, inner.
 */
#ifndef MPM_BOXCHECK_H
#define MPM_BOXCHECK_H

#ifndef MPM_HD
#if defined(__HIPCC__)
#define MPM_HD __host__ __device__
#else
#define MPM_HD
#endif
#endif

template <class B>
MPM_HD inline bool box_whole_in_grid(const B& b, int mx, int my, int mz) {
  if (b.bx <= 0 || b.by <= 0 || b.bz <= 0 || b.x0 < 0 || b.y0 < 0 || b.z0 < 0) return false;
  if ((long long)b.x0 + b.bx - 1 > mx || (long long)b.y0 + b.by - 1 > my || (long long)b.z0 + b.bz - 1 > mz) return false;
  return (long long)b.n == (long long)b.bx * b.by * b.bz;
}

// block-uniform part of the rule
template <class B2>
MPM_HD inline bool box_entry_consistent(const B2& C, int cap, int mx, int my, int mz) {
  if (!(C.ok && C.inner) || C.n <= 0 || C.n > cap || C.b[0].n <= 0 || C.b[1].n < 0) return false;
  if ((long long)C.b[0].n + C.b[1].n != C.n) return false;
  if (!box_whole_in_grid(C.b[0], mx, my, mz)) return false;
  return C.b[1].n == 0 || box_whole_in_grid(C.b[1], mx, my, mz);
}

// per-particle part: the particle's side box holds the stencil of cell
// (cx, cy, cz); LO, HI: the shape's reach (linear 0..1, B-spline -1..1,
// GIMP -1..2)
template <int LO, int HI, class B2>
MPM_HD inline bool box_entry_holds(const B2& C, int jl, int cx, int cy, int cz) {
  const bool s1 = jl >= C.split;   // selects without indexed access
  const int n = s1 ? C.b[1].n : C.b[0].n;
  const int x0 = s1 ? C.b[1].x0 : C.b[0].x0, y0 = s1 ? C.b[1].y0 : C.b[0].y0, z0 = s1 ? C.b[1].z0 : C.b[0].z0;
  const int bx = s1 ? C.b[1].bx : C.b[0].bx, by = s1 ? C.b[1].by : C.b[0].by, bz = s1 ? C.b[1].bz : C.b[0].bz;
  return n > 0 && cx + LO >= x0 && cy + LO >= y0 && cz + LO >= z0 && (long long)cx + HI < (long long)x0 + bx &&
         (long long)cy + HI < (long long)y0 + by && (long long)cz + HI < (long long)z0 + bz;
}

#endif
