// kernels.hip -- gfx950 kernels of the explicit MPM step.
//
// Every particle kernel runs one particle per lane over the granules of the
// active instances; every node kernel one node per lane.  Shape-function
// weights are built per axis in registers and multiplied out per node in
// the reference's association order, so FP64 results match the reference
// up to the summation order of the scatter (P2G) sums.
//
// Reference correspondences are cited per routine.  T is the bulk type
// (double: plain FP64 path; float: FP32 bulk with FP64 positions, FP64
// grid mass and FP64 momentum/mass division).
#include <hip/hip_runtime.h>

#include "boxcheck.h"
#include "crlog.h"
#include "engine.h"
#include "kernels.h"

// This file is compiled twice: MPMK_PART 1 instantiates the FP64 path (and
// the precision-independent kernels) without FMA contraction, so its
// per-particle arithmetic rounds like the reference's; MPMK_PART 2 the
// FP32-bulk path with contraction allowed.
#ifndef MPMK_PART
#define MPMK_PART 0
#endif

namespace mpmk {

// ------------------------------------------------------------ utilities

// Particle slot of this thread.  With ilv the block's slots are dealt out
// round-robin over the 32-lane halves of the waves (slot = (t % 32) * bs/32
// + t / 32): neighbouring slots, which share a cell after the sort, land in
// different lanes of different instructions, so scatter atomics into the
// LDS tile rarely hit one address from many lanes at once.
__device__ __forceinline__ long part_index(int ilv, const int* glist, int bpg, int& gran) {
  const int b = blockIdx.x;
  gran = glist[b / bpg];
  const int t = threadIdx.x;
  const int j = ilv ? (t & 31) * (int)(blockDim.x >> 5) + (t >> 5) : t;
  return (long)gran * PGRAN + (long)(b % bpg) * blockDim.x + j;
}

__device__ __forceinline__ long node_index(const int* glist, int bpg, int& gran) {
  const int b = blockIdx.x;
  gran = glist[b / bpg];
  return (long)gran * NGRAN + (long)(b % bpg) * blockDim.x + threadIdx.x;
}

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

// momentum / mass at a node: in the FP32 path the division is done in FP64
template <typename T>
__device__ __forceinline__ T div_mass(T p, double mg) {
  return (T)((double)p / mg);
}
template <>
__device__ __forceinline__ double div_mass<double>(double p, double mg) {
  return p / mg;
}

// ------------------------------------------------- cell location (a2)

// InWhichCell (Grid.f90:123-155).  Returns the 1-based cell or -1; the
// cell's coordinates are decomposed from the linear index so that the
// reference's wrap at x == SpanX(2) is reproduced.
// A non-finite coordinate is outside as well (the comparisons are written so
// that NaN fails them), which keeps the integer conversions below defined.
__device__ __forceinline__ int in_which_cell(const InstC& I, double x, double y, double z) {
  if (!(x >= I.span0[0] && x <= I.span1[0] && y >= I.span0[1] && y <= I.span1[1] && z >= I.span0[2] &&
        z <= I.span1[2]))
    return -1;
  const int ix = (int)((x - I.span0[0]) / I.dcell) + 1;
  const int iy = (int)((y - I.span0[1]) / I.dcell) + 1;
  const int iz = (int)((z - I.span0[2]) / I.dcell) + 1;
  int c = (iz - 1) * I.ncxy + (iy - 1) * I.ncx + ix;
  if (c > I.ncell || c <= 0) c = -1;
  return c;
}

// InWhichCell with the cell's coordinates: the axis indices of the division,
// carried the way the reference's linear index wraps at the upper faces
// (x == SpanX(2) gives the first cell of the next row), equal to
// cell_coords of the returned index without its two integer divisions.
__device__ __forceinline__ int in_which_cell3(const InstC& I, double x, double y, double z, int& cx, int& cy,
                                              int& cz) {
  if (!(x >= I.span0[0] && x <= I.span1[0] && y >= I.span0[1] && y <= I.span1[1] && z >= I.span0[2] &&
        z <= I.span1[2]))
    return -1;
  int ix = (int)((x - I.span0[0]) / I.dcell);
  int iy = (int)((y - I.span0[1]) / I.dcell);
  int iz = (int)((z - I.span0[2]) / I.dcell);
  const int c = iz * I.ncxy + iy * I.ncx + ix + 1;
  if (c > I.ncell || c <= 0) return -1;
  if (ix >= I.ncx) {
    ix -= I.ncx;
    iy++;
  }
  if (iy >= I.ncy) {
    iy -= I.ncy;
    iz++;
  }
  cx = ix;
  cy = iy;
  cz = iz;
  return c;
}

// Quotient and remainder of 0 <= a < 2^22 by b >= 1 from the FP32
// reciprocal, exact: a and b (b < 2^24; a larger b gives a quotient of 0
// either way) convert exactly, v_rcp_f32 is within 1 ulp, so a * rcp(b) is
// within 1.5 * 2^-23 * a / b < 0.4 of a / b for b >= 2 (exact for b = 1),
// and one correction step lands on the integer quotient.  About a third of
// the instructions of the 32-bit integer division sequence.
__device__ __forceinline__ int fdivmod(int a, int b, int& r) {
  int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
  r = a - q * b;
  if (r < 0) {
    q -= 1;
    r += b;
  } else if (r >= b) {
    q += 1;
    r -= b;
  }
  return q;
}

__device__ __forceinline__ void cell_coords(const InstC& I, int c, int& cx, int& cy, int& cz) {
  const int c0 = c - 1;
  if ((unsigned)c0 < (1u << 22)) {
    int r;
    cz = fdivmod(c0, I.ncxy, r);
    cy = fdivmod(r, I.ncx, cx);
    return;
  }
  cz = c0 / I.ncxy;
  const int r = c0 - cz * I.ncxy;
  cy = r / I.ncx;
  cx = r - cy * I.ncx;
}

// ------------------------------------------------ shape functions (a3-a5)

template <typename T>
struct Stencil {
  T w[3][3];   // [axis][slot] weights
  T d[3][3];   // [axis][slot] derivative factors
  T r[3];      // B-spline: shifted local coordinate per axis (w and d follow from it)
  int base[3]; // node coordinate of slot 0 per axis
  int act[3];  // active slots per axis (GIMP: 2 or 3)
  int node0;   // local linear index of the cell's first node
};

// quadratic B-spline weights and derivative factors of one axis from the
// shifted local coordinate r = (Xp - Xg)/DCell*0.2 + 0.4 (NShape_Bspline,
// Grid.f90:927-962, FP32 literals 0.2 and 0.4 as there)
template <typename T>
__device__ __forceinline__ void bspline_axis(T r, T idc, T* w, T* d) {
  const T f02 = (T)(double)0.2f;
  const T a = (T)3 - (T)5 * r, b = (T)5 * r - (T)2;
  w[0] = (T)0.5 * (a * a);
  w[1] = -((T)25 * (r * r)) + (T)25 * r - (T)5.5;
  w[2] = (T)0.5 * (b * b);
  d[0] = ((T)5 * r - (T)3) * (T)5 * f02 * idc;
  d[1] = (-((T)50 * r) + (T)25) * f02 * idc;
  d[2] = ((T)5 * r - (T)2) * (T)5 * f02 * idc;
}

// per-axis weights.  c: 0-based cell coordinate; s0: span start.
template <int SHAPE, typename T>
__device__ __forceinline__ void axis_weights(const InstC& I, int ax, double x, int c, Stencil<T>& S) {
  const double xg1 = (double)c * I.dcell + I.span0[ax];   // Grid.f90:300-302
  if (SHAPE == MPM_SHAPE_LINEAR) {
    // NShape (Grid.f90:461-501): XN = (Xp - Cxg)*iJacobi, Cxg = Xg1 + DCell/2
    const double cxg = xg1 + I.dcell / 2.0;
    const double xn = (x - cxg) * I.ijac;
    S.w[ax][0] = (T)(-xn + 1.0);
    S.w[ax][1] = (T)(xn + 1.0);
    S.d[ax][0] = (T)-1;
    S.d[ax][1] = (T)1;
    S.base[ax] = c;
    S.act[ax] = 2;
  } else if (SHAPE == MPM_SHAPE_GIMP) {
    // FindInflNode case choice + NShape_GIMP (Grid.f90:546-599, 621-847)
    const double r1 = (x - xg1) * I.idcell;
    int base = c, act = 2;
    if (r1 < 0.25) { base = c - 1; act = 3; }
    else if (r1 > 0.75) { act = 3; }
    S.base[ax] = base;
    S.act[ax] = act;
    const int nmax = (ax == 0 ? I.ncx : (ax == 1 ? I.ncy : I.ncz));
#pragma unroll
    for (int s = 0; s < 3; s++) {
      T sh = (T)0, dn = (T)0;
      const int n = base + s;
      if (s < act && n >= 0 && n <= nmax) {
        const double xg = (double)n * I.dcell + I.span0[ax];
        const T rpg = (T)((x - xg) * I.idcell);
        const T r = fabs(rpg);
        if (r < (T)0.25) {
          sh = ((T)7 - (T)16 * r * r) * (T)0.125;
          dn = (T)-4 * rpg;
        } else if (r < (T)0.75) {
          sh = (T)1 - r;
          dn = -copysign((T)1, rpg);
        } else if (r < (T)1.25) {
          const T t = ((T)5 - (T)4 * r) * (T)0.25;
          sh = t * t;
          dn = (T)2 * rpg - copysign((T)1, rpg) * (T)2.5;
        }
      }
      S.w[ax][s] = sh;
      S.d[ax][s] = dn;
    }
  } else {  // B-spline: BFindInflNode + NShape_Bspline (Grid.f90:852-1049)
    const double c02 = (double)0.2f, c04 = (double)0.4f;   // FP32 literals
    const T r = (T)((x - xg1) * I.idcell * c02 + c04);
    bspline_axis<T>(r, (T)I.idcell, S.w[ax], S.d[ax]);
    S.r[ax] = r;
    S.base[ax] = c - 1;
    S.act[ax] = 3;
  }
}

template <int SHAPE, typename T>
__device__ __forceinline__ void build_stencil(const InstC& I, double x, double y, double z, int c, Stencil<T>& S) {
  int cx, cy, cz;
  cell_coords(I, c, cx, cy, cz);
  axis_weights<SHAPE, T>(I, 0, x, cx, S);
  axis_weights<SHAPE, T>(I, 1, y, cy, S);
  axis_weights<SHAPE, T>(I, 2, z, cz, S);
  S.node0 = cz * I.ngxy + cy * I.ngx + cx;
}

// node order of NShape / CellsNode (Grid.f90:372-379)
__device__ constexpr int LIN_ORD[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                          {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// node order of BFindInflNode with per-axis slot (0: i-1, 1: i, 2: i+1)
__device__ constexpr int BSP_ORD[27][3] = {
    {1, 1, 1}, {2, 1, 1}, {2, 2, 1}, {1, 2, 1}, {1, 1, 2}, {2, 1, 2}, {2, 2, 2}, {1, 2, 2}, {1, 1, 0},
    {2, 1, 0}, {2, 2, 0}, {1, 2, 0}, {1, 0, 0}, {2, 0, 0}, {1, 0, 1}, {2, 0, 1}, {1, 0, 2}, {2, 0, 2},
    {0, 1, 1}, {0, 2, 1}, {0, 1, 2}, {0, 2, 2}, {0, 1, 0}, {0, 2, 0}, {0, 0, 0}, {0, 0, 1}, {0, 0, 2}};

// Visit the influence nodes of one particle:
//   f(local linear node, node x, node y, node z, N, dN/dx, dN/dy, dN/dz).
// The linear index follows the reference's index arithmetic (B-spline wraps
// at low faces); the 3D coordinates are the unwrapped ones.
// Linear and B-spline visit in the reference's node order; GIMP in z-y-x order.
template <int SHAPE, typename T, typename F>
__device__ __forceinline__ void for_nodes(const InstC& I, const Stencil<T>& S, F&& f) {
  if (SHAPE == MPM_SHAPE_LINEAR) {
    const T ij4 = (T)I.ijac4;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int a = LIN_ORD[k][0], b = LIN_ORD[k][1], c = LIN_ORD[k][2];
      const int node = S.node0 + a + b * I.ngx + c * I.ngxy;
      const T wxy = S.w[0][a] * S.w[1][b];
      const T N = wxy * S.w[2][c] * (T)0.125;
      const T dx = S.d[0][a] * S.w[1][b] * S.w[2][c] * ij4;
      const T dy = S.w[0][a] * S.d[1][b] * S.w[2][c] * ij4;
      const T dz = wxy * S.d[2][c] * ij4;
      f(node, S.base[0] + a, S.base[1] + b, S.base[2] + c, N, dx, dy, dz);
    }
  } else if (SHAPE == MPM_SHAPE_BSPLINE) {
#pragma unroll
    for (int k = 0; k < 27; k++) {
      const int a = BSP_ORD[k][0], b = BSP_ORD[k][1], c = BSP_ORD[k][2];
      // linear index arithmetic of BFindInflNode, including its wrap at low faces
      const int node = S.node0 + (a - 1) + (b - 1) * I.ngx + (c - 1) * I.ngxy;
      if (node < 0 || node >= I.nnode) continue;
      const T wxy = S.w[0][a] * S.w[1][b];
      const T N = wxy * S.w[2][c];
      const T dx = S.d[0][a] * S.w[1][b] * S.w[2][c];
      const T dy = S.w[0][a] * S.d[1][b] * S.w[2][c];
      const T dz = wxy * S.d[2][c];
      f(node, S.base[0] + a, S.base[1] + b, S.base[2] + c, N, dx, dy, dz);
    }
  } else {  // GIMP
    const T idc = (T)I.idcell;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (c >= S.act[2]) continue;
      const int nz = S.base[2] + c;
      if (nz < 0 || nz > I.ncz) continue;
#pragma unroll
      for (int b = 0; b < 3; b++) {
        if (b >= S.act[1]) continue;
        const int ny = S.base[1] + b;
        if (ny < 0 || ny > I.ncy) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          if (a >= S.act[0]) continue;
          const int nx = S.base[0] + a;
          if (nx < 0 || nx > I.ncx) continue;
          const int node = nz * I.ngxy + ny * I.ngx + nx;
          const T wxy = S.w[0][a] * S.w[1][b];
          const T N = wxy * S.w[2][c];
          const T dx = S.d[0][a] * S.w[1][b] * S.w[2][c] * idc;
          const T dy = S.w[0][a] * S.d[1][b] * S.w[2][c] * idc;
          const T dz = wxy * S.d[2][c] * idc;
          f(node, nx, ny, nz, N, dx, dy, dz);
        }
      }
    }
  }
}

// ------------------------------------------------ particle field access

template <typename T>
struct PV {
  const PoolPtrs<T>& P;
  long i;
  __device__ T& t(int f) const { return P.t[(long)f * P.np + i]; }
  __device__ double& d(int f) const { return P.d[(long)f * P.np + i]; }
  __device__ int& n(int f) const { return P.i[(long)f * P.np + i]; }
};

// external particle force (update_step.f90:888-903)
template <typename T>
__device__ __forceinline__ void external_force(const PoolPtrs<T>& P, const PV<T>& p, const InstC& I, double time,
                                               T fx[3]) {
  const int fl = I.flags;
  if (fl & IF_LOADS) {
    fx[0] = p.t(T_FX); fx[1] = p.t(T_FY); fx[2] = p.t(T_FZ);
    if (fl & IF_QUASI_LOAD) {
      if (!(time >= I.qlt)) {
        const T q = (T)I.qlt, tt = (T)time;
#pragma unroll
        for (int d = 0; d < 3; d++) fx[d] = fx[d] / q * tt;
      }
    }
  } else {
    fx[0] = fx[1] = fx[2] = (T)0;
  }
  if (fl & IF_QUASI_DAMP) {
    const T dm = (T)I.damp * p.t(T_MASS);
    fx[0] = fx[0] - dm * p.t(T_VX);
    fx[1] = fx[1] - dm * p.t(T_VY);
    fx[2] = fx[2] - dm * p.t(T_VZ);
  }
  if (fl & IF_GRAVITY) {
    const BodyC& b = P.bodies[p.n(I_BODY)];
    const T m = p.t(T_MASS);
    fx[0] = fx[0] + m * (T)b.grav[0];
    fx[1] = fx[1] + m * (T)b.grav[1];
    fx[2] = fx[2] + m * (T)b.grav[2];
  }
}

// ------------------------------------------------- energy block partials

// block-level deterministic reduction of (ke, ie, mx, my, mz, maxv)
// DPP lane move of a double within a row of 16 lanes (CTRL: quad_perm,
// row_half_mirror 0x141, row_mirror 0x140); every lane of the wave active
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum (or max) of v over the wave, in every lane: quads by DPP (xor 1, xor
// 2), then the mirrored quad and half-row (each lane already holds its
// group's value), then the other half-wave rows by lane permutes
template <bool MAX>
__device__ __forceinline__ double wave_reduce(double v) {
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
  v = op(v, dpp_f64<0xB1>(v));
  v = op(v, dpp_f64<0x4E>(v));
  v = op(v, dpp_f64<0x141>(v));
  v = op(v, dpp_f64<0x140>(v));
  v = op(v, __shfl_xor(v, 16));
  v = op(v, __shfl_xor(v, 32));
  return v;
}

// Energy partials of a block: every wave writes the totals of its 64
// particle slots to its own slot of the partials (granule * PGRAN/64 +
// first particle of the wave / 64), so no block-wide step is needed;
// k_finalize sums the PGRAN/64 slots of every granule.
__device__ __forceinline__ void block_energy(double* part, int gran, int sub, double v[6]) {
#pragma unroll
  for (int q = 0; q < 5; q++) v[q] = wave_reduce<false>(v[q]);
  v[5] = wave_reduce<true>(v[5]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < 6) {
    double mine = v[0];
#pragma unroll
    for (int q = 1; q < 6; q++) mine = lane == q ? v[q] : mine;
    part[((long)gran * (PGRAN / 64) + (long)sub * (blockDim.x >> 6) + wave) * 6 + lane] = mine;
  }
}

// calcEnergy terms of one particle (Particle.f90:528-535) and max |v_i| (update_step.f90:1852)
template <typename T>
__device__ __forceinline__ void particle_energy(const PV<T>& p, double v[6]) {
  const double m = (double)p.t(T_MASS);
  const double vx = (double)p.t(T_VX), vy = (double)p.t(T_VY), vz = (double)p.t(T_VZ);
  double dot = 0.0;
  dot = dot + vx * vx;
  dot = dot + vy * vy;
  dot = dot + vz * vz;
  v[0] = dot * m * 0.5;
  v[1] = p.d(D_IE);
  v[2] = m * vx;
  v[3] = m * vy;
  v[4] = m * vz;
  v[5] = fmax(fmax(fabs(vx), fabs(vy)), fabs(vz));
}

// ============================================================ LDS tiles
//
// A block's particles are spatially compact (generator order or cell sort),
// so the nodes they touch fit a small box.  The box is computed
// collectively; when it fits TILE_NODES nodes, gathers read the box from
// LDS (loaded once, coalesced) and scatters accumulate with LDS atomics and
// flush each node once with one global atomic per field.  Otherwise the
// block falls back to direct global access.  Nodes outside the box (the
// B-spline low-face wrap) always take the direct path.

// Nodes of a block's LDS tile.  A cell-sorted block of 256 particles
// touches 150-300 nodes (two bricks when it straddles a brick boundary);
// blocks whose box does not fit take the direct path.  Each kernel sizes its
// LDS array as TILE_NODES times its bytes per node.
#define TILE_NODES 512

struct Box {
  int x0, y0, z0, bx, by, bz, n, ok;
  int inner;   // box not clamped at a grid face: every stencil node is a box node
  __device__ __forceinline__ int idx(int nx, int ny, int nz) const {
    const int ex = nx - x0, ey = ny - y0, ez = nz - z0;
    if ((unsigned)ex >= (unsigned)bx || (unsigned)ey >= (unsigned)by || (unsigned)ez >= (unsigned)bz) return -1;
    return ex + bx * (ey + by * ez);
  }
  __device__ __forceinline__ int node(const InstC& I, int e) const {
    int ex, ey;   // e < box size <= tile capacity
    const int r = fdivmod(e, bx, ex), ez = fdivmod(r, by, ey);
    return (z0 + ez) * I.ngxy + (y0 + ey) * I.ngx + (x0 + ex);
  }
};

// Box of the points (cx..cx+hi, ...) from (cx+lo, ...) of the block's
// active threads, clamped to [0, mx] x [0, my] x [0, mz].
__device__ __forceinline__ Box block_box_n(bool act, int cx, int cy, int cz, int lo, int hi, int mx, int my, int mz,
                                           int cap) {
  __shared__ int sb[6][16];
  __shared__ Box res;
  int v[6];
  v[0] = act ? cx : 0x7fffffff; v[1] = act ? cy : 0x7fffffff; v[2] = act ? cz : 0x7fffffff;
  v[3] = act ? cx : -1; v[4] = act ? cy : -1; v[5] = act ? cz : -1;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 3; q++) v[q] = min(v[q], __shfl_xor(v[q], off));
#pragma unroll
    for (int q = 3; q < 6; q++) v[q] = max(v[q], __shfl_xor(v[q], off));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < 6; q++) sb[q][wave] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int w = 1; w < nw; w++) {
      for (int q = 0; q < 3; q++) sb[q][0] = min(sb[q][0], sb[q][w]);
      for (int q = 3; q < 6; q++) sb[q][0] = max(sb[q][0], sb[q][w]);
    }
    Box b;
    if (sb[3][0] < 0) {   // no particle in the grid
      b.x0 = b.y0 = b.z0 = 0;
      b.bx = b.by = b.bz = 0;
      b.n = 0;
      b.ok = 1;
      b.inner = 0;
    } else {
      b.inner = sb[0][0] + lo >= 0 && sb[1][0] + lo >= 0 && sb[2][0] + lo >= 0 && sb[3][0] + hi <= mx &&
                sb[4][0] + hi <= my && sb[5][0] + hi <= mz;
      b.x0 = max(sb[0][0] + lo, 0);
      b.y0 = max(sb[1][0] + lo, 0);
      b.z0 = max(sb[2][0] + lo, 0);
      b.bx = min(sb[3][0] + hi, mx) - b.x0 + 1;
      b.by = min(sb[4][0] + hi, my) - b.y0 + 1;
      b.bz = min(sb[5][0] + hi, mz) - b.z0 + 1;
      const long n = (long)b.bx * b.by * b.bz;
      b.ok = n <= cap;
      b.n = b.ok ? (int)n : 0;
    }
    res = b;
  }
  __syncthreads();
  return res;
}

// node box of a block (stencil reach per shape: linear 0..+1, B-spline
// -1..+1, GIMP -1..+2 cells around the particles' cells)
template <int SHAPE>
__device__ __forceinline__ Box block_box(const InstC& I, bool act, int cx, int cy, int cz, int cap) {
  const int lo = SHAPE == MPM_SHAPE_LINEAR ? 0 : -1;
  const int hi = SHAPE == MPM_SHAPE_GIMP ? 2 : 1;
  return block_box_n(act, cx, cy, cz, lo, hi, I.ncx, I.ncy, I.ncz, cap);
}

// Two node boxes of one block.  A block of consecutive particles in the
// cell order can straddle a jump of the space-filling curve (where the curve
// leaves the occupied cells and re-enters elsewhere); its single box then
// exceeds the tile and the block would fall back to global atomics.  Such a
// block is split at the largest jump between consecutive particles and each
// side gets a box of its own; the two share the tile (box 1 from slot
// b[0].n).  A single-box block has split = INT_MAX.
struct Box2 {
  Box b[2];
  int split, n, ok, inner;
  __device__ __forceinline__ int side(int j) const { return j >= split; }
  __device__ __forceinline__ Box get(int k) const { return k ? b[1] : b[0]; }   // no indexed access (registers)
  __device__ __forceinline__ int off(int k) const { return k ? b[0].n : 0; }
  __device__ __forceinline__ int node(const InstC& I, int e) const {
    return e < b[0].n ? b[0].node(I, e) : b[1].node(I, e - b[0].n);
  }
  // tile slot of a node for the per-node (non-inner) paths, which only
  // single-box blocks take
  __device__ __forceinline__ int idx(int nx, int ny, int nz) const { return b[0].idx(nx, ny, nz); }
};

// jl: the particle's index in the block (particle order); scratch: 4 ints
// per thread of LDS that is free until the function returns (the tile)
template <int SHAPE>
__device__ __forceinline__ Box2 block_box2(const InstC& I, bool act, int cx, int cy, int cz, int jl, int cap,
                                           int* scratch) {
  __shared__ int wbest[16];
  Box2 r;
  r.b[0] = block_box<SHAPE>(I, act, cx, cy, cz, cap);
  r.b[1] = r.b[0];
  r.b[1].n = 0;
  r.split = 0x7fffffff;   // every particle on side 0, whatever the block size of a later user
  r.n = r.b[0].n;
  r.ok = r.b[0].ok;
  r.inner = r.b[0].inner;
  if (r.ok) return r;   // block-uniform
  const int t = threadIdx.x;
  scratch[4 * jl] = cx;
  scratch[4 * jl + 1] = cy;
  scratch[4 * jl + 2] = cz;
  scratch[4 * jl + 3] = act ? 1 : 0;
  __syncthreads();
  int key = -1;   // (jump << 16) | (0xffff - j) at particle j = t: the largest jump, first on ties
  if (t >= 1 && scratch[4 * t + 3] && scratch[4 * t - 1]) {
    const int d = max(max(abs(scratch[4 * t] - scratch[4 * t - 4]), abs(scratch[4 * t + 1] - scratch[4 * t - 3])),
                      abs(scratch[4 * t + 2] - scratch[4 * t - 2]));
    key = (min(d, 32767) << 16) | (0xffff - t);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) key = max(key, __shfl_xor(key, o));
  if ((t & 63) == 0) wbest[t >> 6] = key;
  __syncthreads();
  key = wbest[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) key = max(key, wbest[w]);
  if (key < 0) return r;
  const int j = 0xffff - (key & 0xffff);
  const bool s1 = jl >= j;
  const Box b0 = block_box<SHAPE>(I, act && !s1, cx, cy, cz, cap);
  const Box b1 = block_box<SHAPE>(I, act && s1, cx, cy, cz, cap);
  if (b0.ok && b1.ok && b0.inner && b1.inner && b0.n + b1.n <= cap) {
    r.b[0] = b0;
    r.b[1] = b1;
    r.split = j;
    r.n = b0.n + b1.n;
    r.ok = 1;
    r.inner = 1;
  }
  return r;
}

// Box of a block through the per-block cache.  With BOX_LOAD the block's
// cached entry (the box its particles had in an earlier particle kernel of
// the step, or in an earlier step: cells move by at most one per step) is
// taken when it is an inner box that holds the stencil of every active
// particle (of its side) and fits the tile; any such box gives the same
// tile sums and gathers as the tight one, so a stale entry is either
// correct or rejected.  Otherwise the box is computed, and with BOX_STORE
// written back.  blk: the block's first particle slot / 64.
template <int SHAPE>
__device__ __forceinline__ Box2 block_box_cached(const InstC& I, bool act, int cx, int cy, int cz, int jl, int cap,
                                                 int* scratch, int* boxes, int mode, long blk) {
  static_assert(sizeof(Box2) <= BOX_INTS * sizeof(int), "box cache entry");
  static_assert(sizeof(Box2) == 22 * sizeof(int), "the 22-int record tests/box_cache_shim.cpp reads");
  Box2* entry = (Box2*)(boxes + blk * BOX_INTS);
  if (mode & BOX_LOAD) {
    __shared__ int wbad[16];
    const Box2 C = *entry;   // block-uniform
    // the rules of boxcheck.h: a consistent entry whose boxes lie in the grid,
    // and every active particle's stencil inside its side's box
    bool bad = !box_entry_consistent(C, cap, I.ncx, I.ncy, I.ncz);
    if (!bad && act) {
      constexpr int lo = SHAPE == MPM_SHAPE_LINEAR ? 0 : -1;
      constexpr int hi = SHAPE == MPM_SHAPE_GIMP ? 2 : 1;
      bad = !box_entry_holds<lo, hi>(C, jl, cx, cy, cz);
    }
    const unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63) == 0) wbad[threadIdx.x >> 6] = m != 0ull;
    __syncthreads();
    int any = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) any |= wbad[w];
    if (!any) return C;
  }
  const Box2 r = block_box2<SHAPE>(I, act, cx, cy, cz, jl, cap, scratch);
  if ((mode & BOX_STORE) && threadIdx.x == 0) *entry = r;
  return r;
}

// ===================================================== tile-frame slots
//
// In a block whose box is inner, every particle's influence nodes are the
// NS^3 slots e0 + a + b*bx + c*bx*by of the box (NS = 2 linear, 3 GIMP and
// B-spline; GIMP slots outside the particle's 2- or 3-node case carry zero
// weight, so they add exact zeros).  Weights factor per axis:
//   N = w[0][a] w[1][b] w[2][c],  dN/dx = d[0][a] w[1][b] w[2][c], ...
// so gathers are evaluated by sum factorisation (axis by axis) and the
// block-uniform strides keep the per-node index work to one add.  The sums
// associate differently from the reference's node-by-node loop (round-off
// level differences, like the scatter order).

template <int SHAPE>
struct NSlot {
  static constexpr int n = SHAPE == MPM_SHAPE_LINEAR ? 2 : 3;
};

template <typename T>
struct Slots {
  T w[3][3];
  T d[3][3];
  T r[3];      // B-spline only (see Stencil)
  int e0;
};

template <int SHAPE, typename T>
__device__ __forceinline__ void make_slots(const InstC& I, const Box& B, double x, double y, double z, int c,
                                           Slots<T>& S) {
  Stencil<T> st;
  build_stencil<SHAPE, T>(I, x, y, z, c, st);
  // linear: N = s_x s_y s_z / 8 with s = 1 -+ xi, dN/dx = +-s_y s_z ijac/8
  const T ws = SHAPE == MPM_SHAPE_LINEAR ? (T)0.5 : (T)1;
  const T ds = SHAPE == MPM_SHAPE_LINEAR ? (T)(4.0 * I.ijac4) : SHAPE == MPM_SHAPE_GIMP ? (T)I.idcell : (T)1;
#pragma unroll
  for (int ax = 0; ax < 3; ax++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      S.w[ax][k] = st.w[ax][k] * ws;
      S.d[ax][k] = st.d[ax][k] * ds;
    }
  S.e0 = (st.base[0] - B.x0) + B.bx * ((st.base[1] - B.y0) + B.by * (st.base[2] - B.z0));
  if (SHAPE == MPM_SHAPE_BSPLINE)
#pragma unroll
    for (int ax = 0; ax < 3; ax++) S.r[ax] = st.r[ax];
}

// gather tiles in the fast path: NF fields per node, node-major with a
// 4-element stride per group of up to 4 fields (one 16 B LDS read for FP32)
template <typename T, int NF>
struct AosTile {
  static constexpr int G = (NF + 3) / 4;   // 4-field groups per node
  static constexpr int S = 4 * G;          // elements per node
};

template <typename T, int NF, typename BX>
__device__ __forceinline__ void load_tile_aos(const InstC& I, const BX& B, T* tf, const T* const* src, long nn) {
  constexpr int S = AosTile<T, NF>::S;
  for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
    const int k = B.node(I, e);
#pragma unroll
    for (int f = 0; f < NF; f++) tf[e * S + f] = src[f / 3][(f % 3) * nn + k];
  }
}

// Nodal velocity (NF = 3) or velocity and acceleration (NF = 6) formed from
// the grid sums where a gather needs them, instead of by a node pass:
// PXg/Mg and FXg/Mg where Mg > CutOff, 0 elsewhere (the values
// IntegrateMomentum's pass would store, update_step.f90:1194-1239, 1506-1509);
// with FIX the momentum of fixed directions counts as 0 (ApplyBoundaryConditions,
// update_step.f90:1795-1830, ahead of the strain-rate gather).  k: pool node.
// The quotients are formed as products with one FP64 reciprocal of Mg per
// node (within an FP64 ulp of the quotient).
// With INT (NF = 6) the momentum is integrated first, with k_integrate's
// expressions: PXg + FXg DT (DT/2 in the first step), both 0 on fixed
// directions (IntegrateMomentum, update_step.f90:1194-1239).
template <typename T, int NF, bool FIX, bool INT = false>
__device__ __forceinline__ void node_fly(const PoolPtrs<T>& P, const InstC& I, long k, T* v, T dt = (T)0,
                                         bool half = false) {
  const long nn = P.nn;
  const double mg = P.mg[k];
  const bool heavy = mg > I.cutoff;
  const double rm = heavy ? 1.0 / mg : 0.0;
  const int fx = (FIX || INT) ? (int)P.fix[k] : 0;
  if (INT) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      T pg = P.pg[d * nn + k];
      T f = P.fg[d * nn + k];
      if (half) pg = pg + f * dt * (T)0.5;
      else pg = pg + f * dt;
      if (fx & (1 << d)) {
        pg = (T)0;
        f = (T)0;
      }
      v[d] = (T)((double)pg * rm);
      v[3 + d] = (T)((double)f * rm);
    }
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T pd = P.pg[d * nn + k];
    v[d] = (fx & (1 << d)) ? (T)0 : (T)((double)pd * rm);
  }
  if (NF == 6)
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T fd = P.fg[d * nn + k];
      v[3 + d] = (T)((double)fd * rm);
    }
}

// node_fly values of a block's node box into the tile, node-major (AOS, the
// fast-path layout) or field-major with stride cap
template <typename T, int NF, bool FIX, bool AOS, typename BX, bool INT = false>
__device__ __forceinline__ void load_tile_fly(const PoolPtrs<T>& P, const InstC& I, const BX& B, T* tf, int cap,
                                              T dt = (T)0, bool half = false) {
  constexpr int S = AosTile<T, NF>::S;
  for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
    T v[6];
    node_fly<T, NF, FIX, INT>(P, I, (long)I.node_off + B.node(I, e), v, dt, half);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (AOS) tf[e * S + f] = v[f];
      else tf[f * cap + e] = v[f];
    }
  }
}

template <typename T, int NF>
__device__ __forceinline__ void read_node(const T* tf, int e, T v[NF]) {
  constexpr int S = AosTile<T, NF>::S;
  const T* q = tf + e * S;
  if (sizeof(T) == 4) {
#pragma unroll
    for (int g = 0; g < (NF + 3) / 4; g++) {
      const float4 u = *(const float4*)(q + 4 * g);
      const float uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int f = 0; f < 4; f++)
        if (4 * g + f < NF) v[4 * g + f] = (T)uu[f];
    }
  } else {
#pragma unroll
    for (int g = 0; g < (NF + 1) / 2; g++) {
      const double2 u = *(const double2*)(q + 2 * g);
      if (2 * g < NF) v[2 * g] = (T)u.x;
      if (2 * g + 1 < NF) v[2 * g + 1] = (T)u.y;
    }
  }
}

// sum_n N_n v_n for NF fields
template <int NS, typename T, int NF>
__device__ __forceinline__ void gather_w(const Slots<T>& S, const T* tf, int sb, int sc, T out[NF]) {
#pragma unroll
  for (int f = 0; f < NF; f++) out[f] = (T)0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
    T t[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) t[f] = (T)0;
#pragma unroll
    for (int b = 0; b < NS; b++) {
      const int eb = S.e0 + b * sb + c * sc;
      T u[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) u[f] = (T)0;
#pragma unroll
      for (int a = 0; a < NS; a++) {
        T v[NF];
        read_node<T, NF>(tf, eb + a, v);
#pragma unroll
        for (int f = 0; f < NF; f++) u[f] += S.w[0][a] * v[f];
      }
#pragma unroll
      for (int f = 0; f < NF; f++) t[f] += S.w[1][b] * u[f];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) out[f] += S.w[2][c] * t[f];
  }
}

// velocity gradient L[i][j] = sum_n v_n,i dN_n/dx_j
template <int NS, typename T>
__device__ __forceinline__ void gather_grad(const Slots<T>& S, const T* tf, int sb, int sc, T L[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) L[i][j] = (T)0;
#pragma unroll
  for (int c = 0; c < NS; c++) {
    T tx[3] = {(T)0, (T)0, (T)0}, ty[3] = {(T)0, (T)0, (T)0}, tw[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int b = 0; b < NS; b++) {
      const int eb = S.e0 + b * sb + c * sc;
      T sx[3] = {(T)0, (T)0, (T)0}, sw[3] = {(T)0, (T)0, (T)0};
#pragma unroll
      for (int a = 0; a < NS; a++) {
        T v[3];
        read_node<T, 3>(tf, eb + a, v);
#pragma unroll
        for (int i = 0; i < 3; i++) {
          sx[i] += S.d[0][a] * v[i];
          sw[i] += S.w[0][a] * v[i];
        }
      }
#pragma unroll
      for (int i = 0; i < 3; i++) {
        tx[i] += S.w[1][b] * sx[i];
        ty[i] += S.d[1][b] * sw[i];
        tw[i] += S.w[1][b] * sw[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
      L[i][0] += S.w[2][c] * tx[i];
      L[i][1] += S.w[2][c] * ty[i];
      L[i][2] += S.d[2][c] * tw[i];
    }
  }
}

// scatter of mass (FP64) and NV vector fields q_f * N into FP64 tile
// accumulators (field-major, stride cap)
template <int NS, typename T, bool MASS, int NV, typename A>
__device__ __forceinline__ void scatter_w(const Slots<T>& S, int sb, int sc, T m, const T q[3], double* tm, A* tv,
                                          int cap) {
#pragma unroll
  for (int c = 0; c < NS; c++)
#pragma unroll
    for (int b = 0; b < NS; b++) {
      const T mw = m * (S.w[1][b] * S.w[2][c]);
      const int eb = S.e0 + b * sb + c * sc;
#pragma unroll
      for (int a = 0; a < NS; a++) {
        const T shm = S.w[0][a] * mw;
        const int e = eb + a;
        if (MASS) atomicAdd(&tm[e], (double)shm);
#pragma unroll
        for (int f = 0; f < NV; f++) atomicAdd(&tv[f * cap + e], (A)(q[f] * shm));
      }
    }
}

// nodal force f_n = -(sigma . grad N_n) vol + fext N_n into FP64 tile
// accumulators; sv[i][j] = -vol sigma_ij
template <int NS, typename T, typename A>
__device__ __forceinline__ void scatter_force(const Slots<T>& S, int sb, int sc, const T sv[3][3], const T fx[3],
                                              A* tv, int cap) {
#pragma unroll
  for (int c = 0; c < NS; c++)
#pragma unroll
    for (int b = 0; b < NS; b++) {
      const T wyz = S.w[1][b] * S.w[2][c], dyz = S.d[1][b] * S.w[2][c], wdz = S.w[1][b] * S.d[2][c];
      T Af[3], Q[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        Af[i] = sv[i][0] * wyz;
        Q[i] = sv[i][1] * dyz + sv[i][2] * wdz + fx[i] * wyz;
      }
      const int eb = S.e0 + b * sb + c * sc;
#pragma unroll
      for (int a = 0; a < NS; a++) {
        const int e = eb + a;
#pragma unroll
        for (int i = 0; i < 3; i++) atomicAdd(&tv[i * cap + e], (A)(S.d[0][a] * Af[i] + S.w[0][a] * Q[i]));
      }
    }
}

// ========================================================= run-summed scatters
//
// After the cell sort a block's particles come in runs that share a cell,
// so the same NS^3 tile slots.  Adding every particle's node values to the
// LDS tile one by one keeps the LDS busy (27 atomics per field and particle
// for B-spline, lanes of one instruction spread over the tile's banks).
// Instead the particles of an inner block stage their weights and payload
// in LDS, one thread per run sums the run's node values in registers, slab
// by slab, and only the run sums reach the tile.  The per-particle terms are
// the ones of scatter_w / scatter_force; only the summation order changes
// (as with any other order of the atomics).
//
// Stage layout (dynamic LDS): NG groups of 16 bytes per particle slot, the
// slot of block-local particle j at j + j/8 (run heads are about 8 apart
// after the sort; the skew puts their 16-byte reads in different banks),
// then the run key of every particle (its stencil's first tile slot, or -1)
// and the compacted list of run heads.

template <typename T>
struct V16;
template <>
struct V16<float> {
  using type = float4;
  static constexpr int n = 4;
  __device__ static type make(const float* v) { return make_float4(v[0], v[1], v[2], v[3]); }
  __device__ static void split(const type& u, float* v) { v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w; }
};
template <>
struct V16<double> {
  using type = double2;
  static constexpr int n = 2;
  __device__ static type make(const double* v) { return make_double2(v[0], v[1]); }
  __device__ static void split(const type& u, double* v) { v[0] = u.x; v[1] = u.y; }
};

// record of one staged particle: the per-axis weights w[3][NS] and their
// derivative factors d[3][NS] (force scatters) -- or, compact (B-spline),
// the three local coordinates they follow from --, mass and velocity (mass /
// momentum scatters), -vol*sigma (xx yy zz xy yz xz) and the external force
template <int NS, bool MQ, bool FRC, bool CR = false>
struct RunRec {
  static constexpr int W = 0;
  static constexpr int D = CR ? 0 : 3 * NS;                      // valid if FRC && !CR
  static constexpr int M = CR ? 3 : D + (FRC ? 3 * NS : 0);      // m, q0, q1, q2 if MQ
  static constexpr int S = M + (MQ ? 4 : 0);                     // sv[6], fx[3] if FRC
  static constexpr int NV = S + (FRC ? 9 : 0);
};

// weights (and derivative factors) of a staged record
template <int NS, typename T, bool CR, bool FRC, int OD>
__device__ __forceinline__ void run_rec_weights(const T* u, T idc, T (&w)[3][NS], T (&d)[3][NS]) {
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    if (CR) {
      T ww[3], dd[3];
      bspline_axis<T>(u[ax], idc, ww, dd);
#pragma unroll
      for (int a = 0; a < NS; a++) {
        w[ax][a] = ww[a];
        d[ax][a] = dd[a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < NS; a++) {
        w[ax][a] = u[ax * NS + a];
        d[ax][a] = FRC ? u[OD + ax * NS + a] : (T)0;
      }
    }
  }
}

__host__ __device__ constexpr int run_slots(int bs) { return bs + bs / 8; }

template <typename T, int NV>
__host__ __device__ constexpr int run_groups() { return (NV + 16 / (int)sizeof(T) - 1) / (16 / (int)sizeof(T)); }

template <typename T, int NV>
__host__ __device__ constexpr size_t run_stage_bytes(int bs) {
  return (size_t)run_groups<T, NV>() * run_slots(bs) * 16 + (size_t)(3 * bs + 4) * sizeof(int);
}

template <typename T, int NV>
struct RunStage {
  static constexpr int NG = run_groups<T, NV>();
  static constexpr int GV = V16<T>::n;
  typename V16<T>::type* g;
  int* key;
  int* head;   // first particle of every run (compacted)
  int* hend;   // one past its last particle
  int* nh;
  int slots, n;
  __device__ RunStage(char* mem, int bs) {
    slots = run_slots(bs);
    n = bs;
    g = (typename V16<T>::type*)mem;
    key = (int*)(mem + (size_t)NG * slots * 16);
    head = key + bs;
    hend = head + bs;
    nh = hend + bs;
  }
  __device__ void put(int j, const T (&v)[NG * GV]) const {
    const int s = j + (j >> 3);
#pragma unroll
    for (int q = 0; q < NG; q++) g[q * slots + s] = V16<T>::make(v + q * GV);
  }
  __device__ void get(int j, T (&v)[NG * GV]) const {
    const int s = j + (j >> 3);
#pragma unroll
    for (int q = 0; q < NG; q++) V16<T>::split(g[q * slots + s], v + q * GV);
  }
};

// fill a record's weight part from the particle's slots
template <int NS, typename T, int NVV, bool FRC, bool CR>
__device__ __forceinline__ void run_rec_slots(const Slots<T>& S, T (&v)[NVV]) {
  if (CR) {
#pragma unroll
    for (int ax = 0; ax < 3; ax++) v[ax] = S.r[ax];
    return;
  }
#pragma unroll
  for (int ax = 0; ax < 3; ax++)
#pragma unroll
    for (int a = 0; a < NS; a++) {
      v[ax * NS + a] = S.w[ax][a];
      if (FRC) v[3 * NS + ax * NS + a] = S.d[ax][a];
    }
}

// after the stage (keys and records) is written and the block synchronised:
// compact the run heads (first particle of every maximal run of equal keys)
// with their run ends.  A run's end is its first tail (last particle of a
// run) at or after the head in the wave's tail mask; only a run that reaches
// past the end of the wave scans the keys beyond it.
template <typename T, int NV>
__device__ __forceinline__ void run_heads(const RunStage<T, NV>& R) {
  const int j = threadIdx.x, n = R.n;
  const int kj = j < n ? R.key[j] : -1;
  const int kp = (j > 0 && j < n) ? R.key[j - 1] : -1;
  const int kn = j + 1 < n ? R.key[j + 1] : -1;
  const bool h = kj >= 0 && (j == 0 || kp != kj);
  const bool t = kj >= 0 && kn != kj;
  const unsigned long long m = __ballot(h), tm = __ballot(t);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && m) base = atomicAdd(R.nh, __popcll(m));
  base = __shfl(base, 0);
  if (h) {
    const unsigned long long rest = tm >> lane;
    int end;
    if (rest) {
      end = j + __ffsll((long long)rest);   // one past the tail
    } else {
      end = j - lane + 64;
      while (end < n && R.key[end] == kj) end++;
    }
    const int q = base + __popcll(m & ((1ull << lane) - 1));
    R.head[q] = j;
    R.hend[q] = end;
  }
}

// A work item is a (run, slab, part) triple: a thread sums one slab (the
// NS^2 nodes of one stencil layer c) over one part of the particles of one
// run, reading each particle's record once per slab, then adds the sums to
// the tile.  A block has few runs (one per occupied cell), so each run is
// cut into as many parts as keep the item count within the block: the work
// spreads over every wave of the block instead of the first one or two, and
// each thread's loop is that much shorter.
struct RunItem {
  int j0, j1, e0, c;   // particle range [j0, j1) of the part, run key, slab
  int run0;            // first particle of the whole run (its box side)
};

// number of parts per run for nh runs of NS slabs
template <int NS>
__device__ __forceinline__ int run_parts(int nh) {
  const int per = nh * NS;
  return per > 0 ? max(1, (int)blockDim.x / per) : 1;
}

template <int NS, typename T, int NV>
__device__ __forceinline__ RunItem run_item(const RunStage<T, NV>& R, int nh, int parts, int item) {
  RunItem it;
  int k, dump;   // item < nh * NS * parts <= 3 * blockDim.x; len * parts <= blockDim.x^2
  const int rest = fdivmod(item, nh, k);
  it.c = rest % NS;
  const int h = rest / NS;
  it.run0 = R.head[k];
  it.e0 = R.key[it.run0];
  const int len = R.hend[k] - it.run0;
  it.j0 = it.run0 + fdivmod(len * h, parts, dump);
  it.j1 = it.run0 + fdivmod(len * (h + 1), parts, dump);
  return it;
}

// weight (and derivative factor) of stencil slot c along one axis of a staged
// record, without branching on c: B-spline from the local coordinate
// (bspline_axis), otherwise a register select
template <int NS, typename T, bool CR>
__device__ __forceinline__ T slot_w(const T* u, int ax, int c) {
  if (CR) {
    const T r = u[ax];
    const T a = (T)3 - (T)5 * r, b = (T)5 * r - (T)2;
    const T wo = (T)0.5 * (c == 0 ? a * a : b * b);
    const T wm = -((T)25 * (r * r)) + (T)25 * r - (T)5.5;
    return c == 1 ? wm : wo;
  }
  const T* w = u + ax * NS;
  return NS == 2 ? (c ? w[1] : w[0]) : (c == 0 ? w[0] : (c == 1 ? w[1] : w[2]));
}

template <int NS, typename T, bool CR, int OD>
__device__ __forceinline__ T slot_d(const T* u, T idc, int ax, int c) {
  if (CR) {
    const T f02 = (T)(double)0.2f, r = u[ax];
    const T dm = (-((T)50 * r) + (T)25) * f02 * idc;
    const T t = c == 0 ? (T)5 * r - (T)3 : (T)5 * r - (T)2;
    const T dout = t * (T)5 * f02 * idc;
    return c == 1 ? dm : dout;
  }
  const T* d = u + OD + ax * NS;
  return NS == 2 ? (c ? d[1] : d[0]) : (c == 0 ? d[0] : (c == 1 ? d[1] : d[2]));
}

// run sums of the mass and NQ velocity-weighted mass components; a part's
// mass sum is kept in T (a few terms) and added to the FP64 tile
template <int NS, typename T, int NV, int OM, bool MASS, int NQ, bool CR, typename A>
__device__ __forceinline__ void run_scatter_w(const RunStage<T, NV>& R, const Box2& B, T idc, double* tm, A* tv,
                                              int cap) {
  constexpr int NVV = RunStage<T, NV>::NG * RunStage<T, NV>::GV;
  const int nh = *R.nh;
  const int parts = run_parts<NS>(nh);
  for (int item = threadIdx.x; item < nh * NS * parts; item += blockDim.x) {
    const RunItem it = run_item<NS>(R, nh, parts, item);
    if (it.j0 >= it.j1) continue;
    const Box Bk = B.get(B.side(it.run0));
    const int sb = Bk.bx, sc = Bk.bx * Bk.by;
    T am[NS][NS];
    T aq[NS][NS][NQ > 0 ? NQ : 1];
#pragma unroll
    for (int b = 0; b < NS; b++)
#pragma unroll
      for (int a = 0; a < NS; a++) {
        am[b][a] = (T)0;
#pragma unroll
        for (int f = 0; f < NQ; f++) aq[b][a][f] = (T)0;
      }
    for (int j = it.j0; j < it.j1; j++) {
      T u[NVV];
      R.get(j, u);
      T w[3][NS], d[3][NS];
      run_rec_weights<NS, T, CR, false, 0>(u, idc, w, d);
      const T m = u[OM], w2 = slot_w<NS, T, CR>(u, 2, it.c);
#pragma unroll
      for (int b = 0; b < NS; b++) {
        const T mw = m * (w[1][b] * w2);
#pragma unroll
        for (int a = 0; a < NS; a++) {
          const T shm = w[0][a] * mw;
          if (MASS) am[b][a] += shm;
#pragma unroll
          for (int f = 0; f < NQ; f++) aq[b][a][f] += u[OM + 1 + f] * shm;
        }
      }
    }
#pragma unroll
    for (int b = 0; b < NS; b++)
#pragma unroll
      for (int a = 0; a < NS; a++) {
        const int e = it.e0 + a + b * sb + it.c * sc;
        if (MASS) atomicAdd(&tm[e], (double)am[b][a]);
#pragma unroll
        for (int f = 0; f < NQ; f++) atomicAdd(&tv[f * cap + e], (A)aq[b][a][f]);
      }
  }
}

// run sums of the nodal force -(sigma . grad N) vol + fext N
template <int NS, typename T, int NV, int OD, int OS, bool CR, typename A>
__device__ __forceinline__ void run_scatter_force(const RunStage<T, NV>& R, const Box2& B, T idc, A* tv, int cap) {
  constexpr int NVV = RunStage<T, NV>::NG * RunStage<T, NV>::GV;
  const int nh = *R.nh;
  const int parts = run_parts<NS>(nh);
  for (int item = threadIdx.x; item < nh * NS * parts; item += blockDim.x) {
    const RunItem it = run_item<NS>(R, nh, parts, item);
    if (it.j0 >= it.j1) continue;
    const int c = it.c;
    const Box Bk = B.get(B.side(it.run0));
    const int sb = Bk.bx, sc = Bk.bx * Bk.by;
    T af[NS][NS][3];
#pragma unroll
    for (int b = 0; b < NS; b++)
#pragma unroll
      for (int a = 0; a < NS; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) af[b][a][i] = (T)0;
    for (int j = it.j0; j < it.j1; j++) {
      T u[NVV];
      R.get(j, u);
      const T sv[3][3] = {{u[OS + 0], u[OS + 3], u[OS + 5]},
                          {u[OS + 3], u[OS + 1], u[OS + 4]},
                          {u[OS + 5], u[OS + 4], u[OS + 2]}};
      T w[3][NS], d[3][NS];
      run_rec_weights<NS, T, CR, true, OD>(u, idc, w, d);
      const T w2 = slot_w<NS, T, CR>(u, 2, c), d2 = slot_d<NS, T, CR, OD>(u, idc, 2, c);
#pragma unroll
      for (int b = 0; b < NS; b++) {
        const T wyz = w[1][b] * w2, dyz = d[1][b] * w2, wdz = w[1][b] * d2;
        T Af[3], Q[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          Af[i] = sv[i][0] * wyz;
          Q[i] = sv[i][1] * dyz + sv[i][2] * wdz + u[OS + 6 + i] * wyz;
        }
#pragma unroll
        for (int a = 0; a < NS; a++)
#pragma unroll
          for (int i = 0; i < 3; i++) af[b][a][i] += d[0][a] * Af[i] + w[0][a] * Q[i];
      }
    }
#pragma unroll
    for (int b = 0; b < NS; b++)
#pragma unroll
      for (int a = 0; a < NS; a++) {
        const int e = it.e0 + a + b * sb + c * sc;
#pragma unroll
        for (int i = 0; i < 3; i++) atomicAdd(&tv[i * cap + e], (A)af[b][a][i]);
      }
  }
}

// ================================================================ kernels

// GridMomentumInitial / GridMomentumUpdate / GridMomentumMUSL scatters
// (update_step.f90:335-392, 881-988, 1705-1758) with optional fusion.
template <int SHAPE, typename T, int MODE>
__global__ void __launch_bounds__(256) k_p2g(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(P.ilv, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  constexpr bool MASS = (MODE & P2G_MASS) != 0, MOM = (MODE & P2G_MOM) != 0, FRC = (MODE & P2G_FORCE) != 0;
  constexpr int NT = (MOM ? 3 : 0) + (FRC ? 3 : 0);
  // tile accumulation in FP64 for both paths: ds_add_f64 measured several
  // times faster than ds_add_f32 for these scatter patterns on gfx950
  using A = double;
  constexpr int CAP = TILE_NODES;
  __shared__ __align__(16) char smem[CAP * ((MASS ? 8 : 0) + NT * (int)sizeof(A))];
  double* tm = (double*)smem;                       // CAP (mass)
  A* tf = (A*)(smem + (MASS ? 8 * CAP : 0));        // NT * CAP
  PV<T> p{P, i};
  const bool live = i < I.p_end;
  int c = -1, cx = 0, cy = 0, cz = 0;
  double x = 0, y = 0, z = 0;
  // every particle load is issued before the box reduction (two barriers),
  // so that their latencies overlap
  T m = (T)0, vx = (T)0, vy = (T)0, vz = (T)0;
  T fx[3] = {(T)0, (T)0, (T)0};
  T sxx = 0, syy = 0, szz = 0, sxy = 0, syz = 0, sxz = 0, vol = 0;
  if (live) {
    x = p.d(D_X); y = p.d(D_Y); z = p.d(D_Z);
    if (!(MODE & P2G_ICELL)) c = p.n(I_CELL);
    if (MASS || MOM) {
      m = p.t(T_MASS);
      vx = p.t(T_VX); vy = p.t(T_VY); vz = p.t(T_VZ);
    }
    if (FRC) {
      external_force(P, p, I, P.ts[ii].time, fx);
      const T smq = p.t(T_SM) - p.t(T_Q);
      sxx = smq + p.t(T_SD0); syy = smq + p.t(T_SD1); szz = smq + p.t(T_SD2);
      sxy = p.t(T_SD5); syz = p.t(T_SD3); sxz = p.t(T_SD4);
      vol = (T)p.d(D_VOL);
    }
    if (MODE & P2G_ICELL) {
      c = in_which_cell3(I, x, y, z, cx, cy, cz);
      p.n(I_CELL) = c;
    }
  } else if (MODE & P2G_ICELL) {
    p.n(I_CELL) = -1;
  }
  const bool act = c >= 0;
  if (act && !(MODE & P2G_ICELL)) cell_coords(I, c, cx, cy, cz);
  const int jl = (int)(i - ((long)gran * PGRAN + (long)(blockIdx.x % bpg) * blockDim.x));
  const Box2 B = block_box_cached<SHAPE>(I, act, cx, cy, cz, jl, min(P.tile_cap, CAP), (int*)smem, P.boxes,
                                         P.box_mode, (i - jl) / 64);
  const long nn = P.nn;
  if (B.ok) {
    for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
      if (MASS) tm[e] = 0.0;
#pragma unroll
      for (int f = 0; f < NT; f++) tf[f * CAP + e] = (A)0;
    }
    __syncthreads();
  }
  double* Mg = P.mg + I.node_off;
  T* Pg = ((MODE & P2G_TO_P2) ? P.p2 : P.pg) + I.node_off;
  T* Fg = P.fg + I.node_off;
  if (B.ok && B.inner) {   // block-uniform: staged run sums
    constexpr int NS = NSlot<SHAPE>::n;
    constexpr bool CR = SHAPE == MPM_SHAPE_BSPLINE;
    using RR = RunRec<NS, MASS || MOM, FRC, CR>;
    extern __shared__ __align__(16) char run_mem[];
    const RunStage<T, RR::NV> R(run_mem, blockDim.x);
    constexpr int NVV = RunStage<T, RR::NV>::NG * RunStage<T, RR::NV>::GV;
    int key = -1;
    if (act) {
      Slots<T> S;
      const int kb = B.side(jl);
      make_slots<SHAPE, T>(I, B.get(kb), x, y, z, c, S);
      S.e0 += B.off(kb);
      T v[NVV];
#pragma unroll
      for (int q = 0; q < NVV; q++) v[q] = (T)0;
      run_rec_slots<NS, T, NVV, FRC, CR>(S, v);
      if (MASS || MOM) {
        v[RR::M] = m; v[RR::M + 1] = vx; v[RR::M + 2] = vy; v[RR::M + 3] = vz;
      }
      if (FRC) {
        v[RR::S + 0] = -(sxx * vol); v[RR::S + 1] = -(syy * vol); v[RR::S + 2] = -(szz * vol);
        v[RR::S + 3] = -(sxy * vol); v[RR::S + 4] = -(syz * vol); v[RR::S + 5] = -(sxz * vol);
        v[RR::S + 6] = fx[0]; v[RR::S + 7] = fx[1]; v[RR::S + 8] = fx[2];
      }
      R.put(jl, v);
      key = S.e0;
    }
    R.key[jl] = key;
    if (threadIdx.x == 0) *R.nh = 0;
    __syncthreads();
    run_heads(R);
    __syncthreads();
    const T idc = (T)I.idcell;
    if (MASS || MOM) run_scatter_w<NS, T, RR::NV, RR::M, MASS, MOM ? 3 : 0, CR>(R, B, idc, tm, tf, CAP);
    if (FRC) run_scatter_force<NS, T, RR::NV, RR::D, RR::S, CR>(R, B, idc, tf + (MOM ? 3 : 0) * CAP, CAP);
  } else if (act) {
    Stencil<T> S;
    build_stencil<SHAPE, T>(I, x, y, z, c, S);
    for_nodes<SHAPE, T>(I, S, [&](int k, int nx, int ny, int nz, T N, T dX, T dY, T dZ) __attribute__((always_inline)) {
      const int e = B.ok ? B.idx(nx, ny, nz) : -1;
      T val[6];
      int q = 0;
      const T shm = N * m;
      if (MOM) { val[q++] = vx * shm; val[q++] = vy * shm; val[q++] = vz * shm; }
      if (FRC) {
        val[q++] = -(sxx * dX + sxy * dY + sxz * dZ) * vol + fx[0] * N;
        val[q++] = -(sxy * dX + syy * dY + syz * dZ) * vol + fx[1] * N;
        val[q++] = -(sxz * dX + syz * dY + szz * dZ) * vol + fx[2] * N;
      }
      if (e >= 0) {
        if (MASS) atomicAdd(&tm[e], (double)shm);
#pragma unroll
        for (int f = 0; f < NT; f++) atomicAdd(&tf[f * CAP + e], (A)val[f]);
      } else {
        if (MASS) atomic_add(&Mg[k], (double)shm);
        q = 0;
        if (MOM) { atomic_add(&Pg[k], val[0]); atomic_add(&Pg[nn + k], val[1]); atomic_add(&Pg[2 * nn + k], val[2]); q = 3; }
        if (FRC) { atomic_add(&Fg[k], val[q]); atomic_add(&Fg[nn + k], val[q + 1]); atomic_add(&Fg[2 * nn + k], val[q + 2]); }
      }
    });
  }
  if (B.ok && B.n) {
    __syncthreads();
    for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
      const int k = B.node(I, e);
      if (MASS) { const double v = tm[e]; if (v != 0.0) atomic_add(&Mg[k], v); }
      int f = 0;
      if (MOM)
        for (int d = 0; d < 3; d++, f++) { const A v = tf[f * CAP + e]; if (v != (A)0) atomic_add(&Pg[d * nn + k], (T)v); }
      if (FRC)
        for (int d = 0; d < 3; d++, f++) { const A v = tf[f * CAP + e]; if (v != (A)0) atomic_add(&Fg[d * nn + k], (T)v); }
    }
  }
}

// IntegrateMomentum (update_step.f90:1194-1239).  With VA the same pass
// also forms the nodal velocity PXg/Mg and acceleration FXg/Mg that
// ParticlePositionUpdate interpolates (update_step.f90:1506-1509), once per
// node instead of once per particle-node pair; both are 0 where
// Mg <= CutOff, which the particle sums then skip.
template <typename T>
__global__ void __launch_bounds__(256) k_integrate(PoolPtrs<T> P, const int* nglist, const int* nginst, int bpg,
                                                   int va) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  const int ii = nginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  const InstT ts = P.ts[ii];
  if (ts.done) return;
  if (j >= (long)I.node_off + I.nnode_all) return;
  const long nn = P.nn;
  const T dt = (T)ts.dt;
  const bool half = ts.istep == 1;
  const uint8_t fx = P.fix[j];
  const double mg = P.mg[j];
  const bool heavy = mg > I.cutoff;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T pg = P.pg[d * nn + j];
    T f = P.fg[d * nn + j];
    if (half) pg = pg + f * dt * (T)0.5;
    else pg = pg + f * dt;
    if (fx & (1 << d)) {
      pg = (T)0;
      f = (T)0;
      P.fg[d * nn + j] = (T)0;
    }
    P.pg[d * nn + j] = pg;
    if (va) {
      P.vg[d * nn + j] = heavy ? div_mass(pg, mg) : (T)0;
      P.ag[d * nn + j] = heavy ? div_mass(f, mg) : (T)0;
    }
  }
}

// IntegrateMomentum of one instance's nodes [node0, node0 + nnode), without
// the nodal V/A (the grid of a fused USF step whose g2p integrated on the fly)
template <typename T>
__global__ void __launch_bounds__(256) k_integrate_range(PoolPtrs<T> P, int ii, long node0, long nnode) {
  const long j = node0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= node0 + nnode) return;
  const long nn = P.nn;
  const InstT ts = P.ts[ii];
  if (!ts.integ_pending) return;
  const T dt = (T)ts.integ_dt;
  const bool half = ts.integ_half != 0;
  const uint8_t fx = P.fix[j];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T pg = P.pg[d * nn + j];
    T f = P.fg[d * nn + j];
    if (half) pg = pg + f * dt * (T)0.5;
    else pg = pg + f * dt;
    if (fx & (1 << d)) {
      pg = (T)0;
      P.fg[d * nn + j] = (T)0;
    }
    P.pg[d * nn + j] = pg;
  }
}

// ApplyBoundaryConditions (update_step.f90:1795-1830) on PXg (or the
// re-mapped momentum) and FXg; with VEL also the nodal velocity the
// strain-rate gather of ParticleStressUpdate reads (update_step.f90:586-587).
template <typename T>
__global__ void __launch_bounds__(256) k_bc(PoolPtrs<T> P, const int* nglist, const int* nginst, int bpg, int use_p2,
                                            int bc, int vel) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  const int ii = nginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  if (j >= (long)I.node_off + I.nnode_all) return;
  const uint8_t fx = bc ? P.fix[j] : 0;
  T* pg = use_p2 ? P.p2 : P.pg;
  const long nn = P.nn;
  const double mg = P.mg[j];
  const bool heavy = mg > I.cutoff;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T p = pg[d * nn + j];
    if (fx & (1 << d)) {
      p = (T)0;
      pg[d * nn + j] = (T)0;
      P.fg[d * nn + j] = (T)0;
    }
    if (vel) P.vg[d * nn + j] = heavy ? div_mass(p, mg) : (T)0;
  }
}

// nodal acceleration only (per-phase ParticlePositionUpdate after a host
// driven IntegrateMomentum)
template <typename T>
__global__ void __launch_bounds__(256) k_node_va(PoolPtrs<T> P, const int* nglist, const int* nginst, int bpg) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  const int ii = nginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  if (j >= (long)I.node_off + I.nnode_all) return;
  const long nn = P.nn;
  const double mg = P.mg[j];
  const bool heavy = mg > I.cutoff;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    P.vg[d * nn + j] = heavy ? div_mass(P.pg[d * nn + j], mg) : (T)0;
    P.ag[d * nn + j] = heavy ? div_mass(P.fg[d * nn + j], mg) : (T)0;
  }
}

// zero the grid sums of a step in one pass: Mg, PXg, FXg and, when given,
// the MUSL re-map buffer (a kernel, so that steps can be graph-captured)
template <typename T>
__global__ void __launch_bounds__(256) k_zero_grid(double* mg, T* pg, T* fg, T* p2, long nn, const int* nglist,
                                                   int bpg) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  mg[j] = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    pg[d * nn + j] = (T)0;
    fg[d * nn + j] = (T)0;
    if (p2) p2[d * nn + j] = (T)0;
  }
}

// fill kernel (used instead of hipMemsetAsync so steps can be graph-captured)
template <typename T>
__global__ void __launch_bounds__(256) k_zero_nodes(T* a, int ncomp, long stride, const int* nglist, int bpg) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  for (int d = 0; d < ncomp; d++) a[d * stride + j] = (T)0;
}

// Gather the block's node box (NT fields of T) into LDS.
template <typename T, int NT, typename BX>
__device__ __forceinline__ void load_tile(const InstC& I, const BX& B, T* tf, int cap, const T* const* src,
                                          long nn) {
  for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
    const int k = B.node(I, e);
#pragma unroll
    for (int f = 0; f < NT; f++) tf[f * cap + e] = src[f / 3][(f % 3) * nn + k];
  }
}

template <int NT, typename A, typename BX>
__device__ __forceinline__ void zero_tile(const BX& B, A* ta, int cap) {
  for (int e = threadIdx.x; e < B.n; e += blockDim.x)
#pragma unroll
    for (int f = 0; f < NT; f++) ta[f * cap + e] = (A)0;
}

// flush a 3-component accumulation tile into a grid field of type T
template <typename T, typename A, typename BX>
__device__ __forceinline__ void flush_tile3(const InstC& I, const BX& B, const A* ta, int cap, T* dst, long nn) {
  for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
    const int k = B.node(I, e);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const A v = ta[d * cap + e];
      if (v != (A)0) atomic_add(&dst[d * nn + k], (T)v);
    }
  }
}

// ParticlePositionUpdate (update_step.f90:1434-1522), optionally fused with
// the MUSL re-map of momentum (GridMomentumMUSL, :1705-1758) into p2 and
// with the energy partials (when it is the last particle pass).
// Waves per SIMD the particle-update kernel is compiled for: the FP32
// B-spline kernel fits eight waves (64 VGPRs) without spilling, the FP32
// GIMP kernel six (it spills at eight)
template <int SHAPE, typename T>
__host__ __device__ constexpr int g2p_waves() {
  return sizeof(T) != 4 || SHAPE == MPM_SHAPE_LINEAR ? 1 : (SHAPE == MPM_SHAPE_BSPLINE ? 8 : 6);
}

template <int SHAPE, typename T, int MODE>
__global__ void __launch_bounds__(256, (g2p_waves<SHAPE, T>())) k_g2p(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(P.ilv, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  const InstT ts = P.ts[ii];
  if (ts.done) return;
  constexpr int NS = NSlot<SHAPE>::n;
  using A = double;   // re-map accumulation type (see k_p2g)
  constexpr int NB = AosTile<T, 6>::S * (int)sizeof(T);   // V, A per node (fast-path layout)
  constexpr int CAP = TILE_NODES;
  __shared__ __align__(16) char smem[CAP * (NB > 3 * (int)sizeof(A) ? NB : 3 * (int)sizeof(A))];
  T* tf = (T*)smem;
  A* ta = (A*)smem;   // re-map accumulation, reuses the gather tile
  PV<T> p{P, i};
  const bool live = i < I.p_end;
  // the particle's loads are issued before the box reduction (two barriers)
  // and the tile load, so that their latencies overlap
  int c = -1;
  double x = 0.0, y = 0.0, z = 0.0;
  T v0[3] = {(T)0, (T)0, (T)0}, mp = (T)0;
  if (live) {
    c = p.n(I_CELL);
    x = p.d(D_X); y = p.d(D_Y); z = p.d(D_Z);
#pragma unroll
    for (int d = 0; d < 3; d++) v0[d] = p.t(T_VX + d);
    if (MODE & G2P_REMAP) mp = p.t(T_MASS);
  }
  const bool act = c >= 0;
  int cx = 0, cy = 0, cz = 0;
  if (act) cell_coords(I, c, cx, cy, cz);
  const int jl = (int)(i - ((long)gran * PGRAN + (long)(blockIdx.x % bpg) * blockDim.x));
  const Box2 B = block_box_cached<SHAPE>(I, act, cx, cy, cz, jl, min(P.tile_cap, CAP), (int*)smem, P.boxes,
                                         P.box_mode, (i - jl) / 64);
  const bool fast = B.ok && B.inner;
  const int kb = B.side(jl);
  const Box Bk = B.get(kb);
  const int sb = Bk.bx, sc = Bk.bx * Bk.by;
  const long nn = P.nn;
  constexpr bool FLY = (MODE & G2P_VAFLY) != 0;
  constexpr bool INT = (MODE & G2P_INTFLY) != 0;
  const T idt = (T)ts.dt;
  const bool ihalf = ts.istep == 1;
  const T* Vg = P.vg + I.node_off;
  const T* Ag = P.ag + I.node_off;
  if (B.ok && B.n) {
    const T* srcs[2] = {Vg, Ag};
    if (FLY) {
      if (fast) load_tile_fly<T, 6, false, true, Box2, INT>(P, I, B, tf, CAP, idt, ihalf);
      else load_tile_fly<T, 6, false, false, Box2, INT>(P, I, B, tf, CAP, idt, ihalf);
    } else {
      if (fast) load_tile_aos<T, 6>(I, B, tf, srcs, nn);
      else load_tile<T, 6>(I, B, tf, CAP, srcs, nn);
    }
    __syncthreads();
  }
  double ev[6] = {0, 0, 0, 0, 0, 0};
  Stencil<T> S;
  Slots<T> F;
  T vn[3] = {(T)0, (T)0, (T)0};
  if (act) {
    T vx[3] = {(T)0, (T)0, (T)0}, ax[3] = {(T)0, (T)0, (T)0};
    if (fast) {
      make_slots<SHAPE, T>(I, Bk, x, y, z, c, F);
      F.e0 += B.off(kb);
      T va[6];
      gather_w<NS, T, 6>(F, tf, sb, sc, va);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        vx[d] = va[d];
        ax[d] = va[3 + d];
      }
    } else {
      build_stencil<SHAPE, T>(I, x, y, z, c, S);
      for_nodes<SHAPE, T>(I, S, [&](int k, int nx, int ny, int nz, T N, T, T, T) __attribute__((always_inline)) {
        const int e = B.ok ? B.idx(nx, ny, nz) : -1;
        T g[6];
        if (FLY && e < 0) node_fly<T, 6, false, INT>(P, I, (long)I.node_off + k, g, idt, ihalf);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const T v = e >= 0 ? tf[d * CAP + e] : (FLY ? g[d] : Vg[d * nn + k]);
          const T a = e >= 0 ? tf[(3 + d) * CAP + e] : (FLY ? g[3 + d] : Ag[d * nn + k]);
          vx[d] = vx[d] + N * v;
          ax[d] = ax[d] + N * a;
        }
      });
    }
    const double dt = ts.dt;
    const T tdt = (T)dt;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T v = v0[d];
      vn[d] = (ts.istep == 1) ? v + ax[d] * tdt * (T)0.5 : v + ax[d] * tdt;
      p.t(T_VX + d) = vn[d];
      if (!(MODE & G2P_MOVE)) p.t(T_BX + d) = vx[d];   // XX = Xp + vbar DT, pending only without MOVE
    }
    if (MODE & G2P_MOVE) {   // USF: Xp = XX = Xp + vx*DT
      p.d(D_X) = x + (double)vx[0] * dt;
      p.d(D_Y) = y + (double)vx[1] * dt;
      p.d(D_Z) = z + (double)vx[2] * dt;
    }
  }
  if (MODE & G2P_REMAP) {
    T* P2 = P.p2 + I.node_off;
    if (B.ok && B.n) {
      __syncthreads();
      zero_tile<3>(B, ta, CAP);
      __syncthreads();
    }
    if (act) {
      const T m = mp;
      if (fast) {
        scatter_w<NS, T, false, 3>(F, sb, sc, m, vn, nullptr, ta, CAP);
      } else {
        for_nodes<SHAPE, T>(I, S, [&](int k, int nx, int ny, int nz, T N, T, T, T) __attribute__((always_inline)) {
          const T shm = N * m;
          const int e = B.ok ? B.idx(nx, ny, nz) : -1;
          if (e >= 0) {
#pragma unroll
            for (int d = 0; d < 3; d++) atomicAdd(&ta[d * CAP + e], (A)(vn[d] * shm));
          } else {
#pragma unroll
            for (int d = 0; d < 3; d++) atomic_add(&P2[d * nn + k], vn[d] * shm);
          }
        });
      }
    }
    if (B.ok && B.n) {
      __syncthreads();
      flush_tile3<T>(I, B, ta, CAP, P2, nn);
    }
  }
  if (MODE & G2P_ENERGY) {
    if (live) particle_energy(p, ev);
    block_energy(P.part, gran, blockIdx.x % bpg, ev);
  }
}

// ------------------------------------------------------- constitution

template <typename T>
struct Cst {
  T dinc[6];
  double vold, vol, dvol, iener;   // FP64 in both paths
  T sm, smold, sd[6], sig[6], sold[6], bqf;
  T seqv, epeff, sigy, ieinc, den0, den, vol0, cp, mp;
  double dep;   // MM_JC: the plastic strain increment the update assigned (jcf & JC_ASSIGNED)
  int jcf;      // MM_JC: JC_* record flags
};

// material models a kernel instance may meet (MM: bit 1 << model type)
__host__ __device__ constexpr bool mm_has(int MM, int type) { return ((MM >> type) & 1) != 0; }

// the particle state the constitutive update reads (loaded ahead of the
// strain-rate gather so that the loads overlap the block's other latencies).
// sig_y is read only by the hardening models and Seqv only kept by the
// null, HE burn and (failed) sjcf models (the others recompute it), so the
// loads depend on the models MM.
template <typename T, int MM>
__device__ __forceinline__ void cst_load(const PV<T>& p, Cst<T>& m) {
  m.vold = p.d(D_VOL);
  m.epeff = p.t(T_EPEFF);
  m.sigy = (mm_has(MM, MPM_MAT_PLA1) || mm_has(MM, MPM_MAT_PLA2) || mm_has(MM, MPM_MAT_SJC) ||
            mm_has(MM, MPM_MAT_SJCF) || (MM & MM_JC)) ? p.t(T_SIGY) : (T)0;
  m.seqv = (mm_has(MM, MPM_MAT_NULL) || mm_has(MM, MPM_MAT_HIEX) || mm_has(MM, MPM_MAT_SJCF) || (MM & MM_JC))
               ? p.t(T_SEQV) : (T)0;
  m.iener = p.d(D_IE);
  m.mp = p.t(T_MASS);
  m.cp = p.t(T_CP);
  m.sm = p.t(T_SM);
#pragma unroll
  for (int k = 0; k < 6; k++) m.sd[k] = p.t(T_SD0 + k);
}

template <typename T>
__device__ __forceinline__ T eq_stress(const Cst<T>& m) {   // EquivalentStress, FP32 result
  const T* sd = m.sd;
  T J2 = (T)0.5 * (sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]) + sd[3] * sd[3] + sd[4] * sd[4] + sd[5] * sd[5];
  J2 = J2 * (T)3.0;
  return (T)(float)sqrt(J2);
}

template <typename T>
__device__ __forceinline__ void sigrot(Cst<T>& m, const T vort[3]) {   // Constitution.f90:527-575
  T* sig = m.sig;
  const T q0 = (T)2 * sig[5] * vort[2];
  const T q1 = (T)2 * sig[4] * vort[1];
  const T q2 = (T)2 * sig[3] * vort[0];
  T rot[6];
  rot[0] = -q0 + q1;
  rot[1] = q0 - q2;
  rot[2] = -q1 + q2;
  rot[3] = vort[0] * (sig[1] - sig[2]) + vort[2] * sig[4] - vort[1] * sig[5];
  rot[4] = vort[1] * (sig[2] - sig[0]) + vort[0] * sig[5] - vort[2] * sig[3];
  rot[5] = vort[2] * (sig[0] - sig[1]) + vort[1] * sig[3] - vort[0] * sig[4];
#pragma unroll
  for (int k = 0; k < 6; k++) sig[k] = sig[k] + rot[k];
  m.sm = (sig[0] + sig[1] + sig[2]) / (T)3;
  m.sd[0] = sig[0] - m.sm;
  m.sd[1] = sig[1] - m.sm;
  m.sd[2] = sig[2] - m.sm;
  m.sd[3] = sig[3];
  m.sd[4] = sig[4];
  m.sd[5] = sig[5];
}

template <typename T>
__device__ __forceinline__ void elastic_devi(Cst<T>& m, T G2) {
  const T dem = (m.dinc[0] + m.dinc[1] + m.dinc[2]) / (T)3;
  const T G = (T)0.5 * G2;
  m.sd[0] = m.sd[0] + G2 * (m.dinc[0] - dem);
  m.sd[1] = m.sd[1] + G2 * (m.dinc[1] - dem);
  m.sd[2] = m.sd[2] + G2 * (m.dinc[2] - dem);
  m.sd[3] = m.sd[3] + G * m.dinc[3];
  m.sd[4] = m.sd[4] + G * m.dinc[4];
  m.sd[5] = m.sd[5] + G * m.dinc[5];
}

template <typename T>
__device__ __forceinline__ void elastic_p(Cst<T>& m, T K3) {
  const T dem = (m.dinc[0] + m.dinc[1] + m.dinc[2]) / (T)3;
  const T dsm = K3 * dem;
  m.sm = m.sm + dsm;
}

template <typename T>
__device__ __forceinline__ T sig_dot(const Cst<T>& m) {
  return m.sig[0] * m.dinc[0] + m.sig[1] * m.dinc[1] + m.sig[2] * m.dinc[2] + m.sig[3] * m.dinc[3] +
         m.sig[4] * m.dinc[4] + m.sig[5] * m.dinc[5];
}

template <typename T>
__device__ __forceinline__ T new_dot(const Cst<T>& m, T e) {
  return e + (m.sd[0] + m.sm) * m.dinc[0] + (m.sd[1] + m.sm) * m.dinc[1] + (m.sd[2] + m.sm) * m.dinc[2] +
         m.sd[3] * m.dinc[3] + m.sd[4] * m.dinc[4] + m.sd[5] * m.dinc[5];
}

template <typename T>
__device__ __forceinline__ void bulkq(Cst<T>& m, const InstC& I, T dt) {   // Constitution.f90:1543-1563
  T dd = (m.dinc[0] + m.dinc[1] + m.dinc[2]) / dt;
  dd = dd < (T)0 ? dd : (T)0;
  const T h = (T)I.dcell;
  m.bqf = m.den * h * h * (T)I.bq1 * dd * dd - (T)I.bq2 * m.den * h * m.cp * dd;
}

template <typename T>
__device__ __forceinline__ void lieupd(Cst<T>& m, const InstC& I, T dt) {   // :627-646
  if (m.sm < (T)0) bulkq(m, I, dt);
  m.ieinc = m.ieinc - (T)2 * m.bqf * (m.dinc[0] + m.dinc[1] + m.dinc[2]);
  const float vavg = (float)(m.vol + m.vold);
  m.iener = m.iener + (double)((T)0.25 * m.ieinc) * (double)vavg;
}

template <typename T>
__device__ __forceinline__ void hieupd(Cst<T>& m, bool failure) {   // :649-676
  if (failure) {
    m.ieinc = (T)m.dvol * (m.smold - (T)2 * m.bqf);
  } else {
    const float vavg = (float)(m.vol + m.vold);
    const T e = m.dinc[0] * (m.sold[0] + m.sd[0]) + m.dinc[1] * (m.sold[1] + m.sd[1]) +
                m.dinc[2] * (m.sold[2] + m.sd[2]) + m.dinc[3] * (m.sold[3] + m.sd[3]) +
                m.dinc[4] * (m.sold[4] + m.sd[4]) + m.dinc[5] * (m.sold[5] + m.sd[5]);
    m.ieinc = (T)0.25 * e * (T)vavg + (T)m.dvol * (m.smold - (T)2 * m.bqf);
  }
  m.iener = m.iener + m.ieinc;
}

// log, exp and real powers: in the FP64 paths the correctly rounded ones of
// crlog.h (reproducible on any libm, and what the reference's glibc returns
// in all but rare cases: a one-ulp difference in the Johnson-Cook yield
// stress can flip its return-map test), the device libm's in the FP32 path
template <typename T, int MM>
__device__ __forceinline__ bool m_cr() { return (MM & MM_CRLOG) || sizeof(T) == 8; }
template <typename T, int MM>
__device__ __forceinline__ double m_log(double x) { return m_cr<T, MM>() ? mpm_log_cr(x) : log(x); }
template <typename T, int MM>
__device__ __forceinline__ double m_exp(double x) { return m_cr<T, MM>() ? mpm_exp_cr(x) : exp(x); }
template <typename T, int MM>
__device__ __forceinline__ double m_pow(double x, double y) { return m_cr<T, MM>() ? mpm_pow_cr(x, y) : pow(x, y); }

// an update that assigns the reference's module variable depeff (every
// plastic model, Constitution.f90:902, 944, 1005, 1074, 1177, 1285-1322)
// records it for the Johnson-Cook particles after it
template <int MM, typename C>
__device__ __forceinline__ void jc_assign(C& m, double depeff) {
  if (MM & MM_JC) {
    m.jcf |= JC_ASSIGNED;
    m.dep = depeff;
  }
}

// M3DM5, simplified Johnson-Cook (Constitution.f90:1043-1100): the hardening
// modulus at epeff + 1e-4 and the rate-dependent yield stress in FP64 (the
// literals 0.0001 and 0.001 are default REALs), the radial return in T
template <typename T, int MM>
__device__ __forceinline__ void m3dm5(Cst<T>& m, const MatC& mc, T G2, double dtd) {
  const double e4 = (double)0.0001f, rmin = (double)0.001f;
  double ep = (double)m.epeff + e4;
  const T plamod = (T)(mc.bjc * mc.njc * m_pow<T, MM>(ep, mc.njc - 1));
  ep = ep - e4;
  m.epeff = (T)ep;
  elastic_devi(m, G2);
  m.seqv = eq_stress(m);
  if (m.seqv > m.sigy) {
    const T depeff = (m.seqv - m.sigy) / ((T)1.5 * G2 + plamod);
    m.epeff = m.epeff + depeff;
    double srate = (double)depeff / mc.epso / dtd;
    if (srate < rmin) srate = rmin;
    m.sigy = (T)((mc.yield0 + mc.bjc * m_pow<T, MM>((double)m.epeff, mc.njc)) * (1 + mc.cjc * m_log<T, MM>(srate)));
    const bool back = m.sigy > m.seqv;
    if (back) {
      m.epeff = m.epeff - depeff;
    } else {
      const T ratio = m.sigy / m.seqv;
#pragma unroll
      for (int k = 0; k < 6; k++) m.sd[k] = m.sd[k] * ratio;
      m.seqv = m.seqv * ratio;
    }
    jc_assign<MM>(m, back ? 0.0 : (double)depeff);
  }
}

// M3DM4 (john, Constitution.f90:969-1040) and M3DM8 (jcf, :1144-1212):
// M3DM5's return with the temperature factor 1 - tstar**mjc, tstar from the
// temperature the particle entered the step with (jcf: clamped to [0, 1],
// and a particle past its melting point fails); john also takes the energy
// terms and the elastic pressure, jcf gets its pressure from the EOS
template <typename T, int MM, bool JCF>
__device__ __forceinline__ void m3dm48(Cst<T>& m, const MatC& mc, T G2, T K3, double dtd, double tmprt,
                                       bool& failure) {
  const double e4 = (double)0.0001f, rmin = (double)0.001f;
  if (!JCF) m.ieinc = sig_dot(m);
  double ep = (double)m.epeff + e4;
  const T plamod = (T)(mc.bjc * mc.njc * m_pow<T, MM>(ep, mc.njc - 1));
  ep = ep - e4;
  m.epeff = (T)ep;
  elastic_devi(m, G2);
  m.seqv = eq_stress(m);
  if (m.seqv > m.sigy) {
    const T depeff = (m.seqv - m.sigy) / ((T)1.5 * G2 + plamod);
    m.epeff = m.epeff + depeff;
    double tstar = (tmprt - mc.roomt) / (mc.melt - mc.roomt);
    if (JCF) {
      if (tstar > 1.0) {
        tstar = 1.0;
        failure = true;
      } else if (tstar < 0.0) {
        tstar = 0.0;
      }
    }
    double srate = (double)depeff / mc.epso / dtd;
    if (srate < rmin) srate = rmin;
    m.sigy = (T)((mc.yield0 + mc.bjc * m_pow<T, MM>((double)m.epeff, mc.njc)) * (1 + mc.cjc * m_log<T, MM>(srate)) *
                 (1 - m_pow<T, MM>(tstar, mc.mjc)));
    const bool back = m.sigy > m.seqv;
    if (back) {
      m.epeff = m.epeff - depeff;
    } else {
      const T ratio = m.sigy / m.seqv;
#pragma unroll
      for (int k = 0; k < 6; k++) m.sd[k] = m.sd[k] * ratio;
      m.seqv = m.seqv * ratio;
    }
    jc_assign<MM>(m, back ? 0.0 : (double)depeff);
  }
  if (!JCF) {
    elastic_p(m, K3);
    m.ieinc = new_dot(m, m.ieinc);
  }
}

// seleos (Constitution.f90:1333-1541) in FP64 from the FP64 mass and volume
// (the reference's mu, relative volume and specific energy); EOS 1 and 2
// keep A, B and the pressure in default REALs, as the reference declares
// them; bfac: the HE burn fraction (EOS 3)
template <typename T, int MM>
__device__ __forceinline__ void seleos(Cst<T>& m, const PV<T>& p, const MatC& mc, const InstC& I, T dt,
                                       bool failure, double bfac) {
  const double mass = p.d(D_MASS);
  const double vol0 = mass / mc.density, den = mass / m.vol;
  const double mu = den / mc.density - 1, rv = m.vol / vol0;
  const double* c = mc.ce;
  switch (mc.eos) {
    case 1: {   // linear polynomial
      const float A = (float)(c[0] + mu * (c[1] + mu * (c[2] + mu * c[3])));
      const float B = (float)(c[4] + mu * (c[5] + mu * c[6]));
      hieupd(m, failure);
      const double specen = m.iener / vol0;
      float pnew = (float)((A + B * specen) / (1 + B * m.dvol / vol0));
      if (failure && pnew < 0.0f) pnew = 0.0f;
      m.ieinc = (T)((double)m.ieinc - m.dvol * pnew);
      m.iener = m.iener - m.dvol * pnew;
      m.sm = (T)(-(double)pnew);
      break;
    }
    case 2: {   // Mie-Gruneisen; mu**2 = mu*mu, mu**3 = (mu*mu)*mu
      float A, B;
      if (mu > 0) {
        A = (float)((c[0] * mu + c[1] * (mu * mu) + c[2] * ((mu * mu) * mu)) * (1 - c[3] * mu * 0.5 / den));
        B = (float)c[4];
      } else {
        A = (float)(c[0] * mu);
        B = 0.0f;
      }
      if (m.sm < (T)0) bulkq(m, I, dt);
      hieupd(m, failure);
      const double specen = m.iener / vol0;
      float pnew = (float)((A + B * specen) / (1 + B * m.dvol / vol0));
      if (failure && pnew < 0.0f) pnew = 0.0f;
      m.ieinc = (T)((double)m.ieinc - m.dvol * pnew);
      m.iener = m.iener - m.dvol * pnew;
      m.sm = (T)(-(double)pnew);
      break;
    }
    case 3: {   // JWL
      const double r1v = c[2] * rv, r2v = c[3] * rv;
      const double wr1v = c[0] * c[4] / r1v, wr2v = c[1] * c[4] / r2v;
      const double wdr1v = c[0] - wr1v, wdr2v = c[1] - wr2v;
      double A = wdr1v * m_exp<T, MM>(-r1v) + wdr2v * m_exp<T, MM>(-r2v) + (double)m.bqf;
      double B = c[4] / rv;
      hieupd(m, failure);
      const double specen = m.iener / vol0;
      A = A * bfac;
      B = B * bfac;
      double pnew = (A + B * specen) / (1 + B * m.dvol / vol0);
      if (failure && pnew < 0) pnew = 0.0;
      m.iener = m.iener - m.dvol * pnew;
      m.sm = (T)(-pnew);
      break;
    }
    default: {   // 4: weakly compressible
      if (m.sm < (T)0) bulkq(m, I, dt);
      hieupd(m, failure);
      double pnew = mc.c0 * mc.c0 * (den - mc.density);
      if (failure && pnew < 0) pnew = 0.0;
      m.iener = m.iener - m.dvol * pnew;
      m.sm = (T)(-pnew);
      break;
    }
  }
}

__device__ __forceinline__ double det3(const double F[3][3]) {
  return F[0][0] * F[1][1] * F[2][2] - F[0][0] * F[1][2] * F[2][1] - F[1][0] * F[0][1] * F[2][2] +
         F[1][0] * F[0][2] * F[2][1] + F[2][0] * F[0][1] * F[1][2] - F[2][0] * F[0][2] * F[1][1];
}

// M3DM7, the null material's viscous deviator (Constitution.f90:1124-1142)
template <typename T>
__device__ __forceinline__ void m3dm7(Cst<T>& m, const MatC& mc, T dt) {
  const T miu2 = (T)mc.miu2;
  const T tri = (T)0 * (m.dinc[0] / dt + m.dinc[1] / dt + m.dinc[2] / dt);   // 1/3 is integer 0
  m.sd[0] = miu2 * (m.dinc[0] / dt - tri);
  m.sd[1] = miu2 * (m.dinc[1] / dt - tri);
  m.sd[2] = miu2 * (m.dinc[2] / dt - tri);
  m.sd[3] = (T)0.5 * miu2 * m.dinc[3] / dt;
  m.sd[4] = (T)0.5 * miu2 * m.dinc[4] / dt;
  m.sd[5] = (T)0.5 * miu2 * m.dinc[5] / dt;
}

// Constitution (Constitution.f90:79-368) for the models on the path:
// elas (M3DM1), pla1 (M3DM2), pla2 (M3DM3), sjc and sjcf (M3DM5 + EOS),
// HE burn (M3DM6 + JWL), null (M3DM7 + EOS), Drucker-Prager (M3DM9),
// neo-Hookean (M3DM10); with MM_JC also john (M3DM4) and jcf (M3DM8 +
// EOS), whose temperature, damage and failure read the increment another
// particle may have assigned and follow in k_x_jc_tail.
// material models compiled into a stress kernel: bit (1 << type); a launch
// over a group whose instances use one model gets the kernel with only that
// model (fewer registers, higher occupancy), mixed groups the general one
template <typename T, int MM = MM_ALL>
__device__ __forceinline__ void constitution(const PV<T>& p, const InstC& I, const MatC& mc, T dt, double dtd,
                                             double time, const T de[6], const T vort[3], const double F[3][3],
                                             const double deF[3][3], Cst<T>& m) {
#pragma unroll
  for (int k = 0; k < 6; k++) m.dinc[k] = de[k];
  if (I.flags & IF_UVFP) {
    const double detF = det3(deF);
    m.vol = m.vold * detF;
  } else {
    m.vol = m.vold * (1.0 + (double)de[0] + (double)de[1] + (double)de[2]);
  }
  p.d(D_VOL) = m.vol;
  m.dvol = 0.5 * (m.vol - m.vold);
  const T mp = m.mp;
  m.den0 = (T)mc.density;
  m.vol0 = mp / m.den0;
  m.den = (T)((double)mp / m.vol);
  m.smold = m.sm;
#pragma unroll
  for (int k = 0; k < 6; k++) m.sold[k] = m.sd[k];
  m.sig[0] = m.sd[0] + m.sm;
  m.sig[1] = m.sd[1] + m.sm;
  m.sig[2] = m.sd[2] + m.sm;
  m.sig[3] = m.sd[3];
  m.sig[4] = m.sd[4];
  m.sig[5] = m.sd[5];
  m.bqf = (T)0;
  m.ieinc = (T)0;
  if (MM & MM_JC) m.jcf = 0;
  const T G2 = (T)mc.G2, K3 = (T)mc.K3;
  // a model outside MM cannot occur here (the engine picks MM from the group)
  const int mtype = MM == (1 << MPM_MAT_ELASTIC) ? MPM_MAT_ELASTIC : MM == (1 << MPM_MAT_PLA1) ? MPM_MAT_PLA1
                  : MM == (1 << MPM_MAT_PLA2) ? MPM_MAT_PLA2 : MM == (1 << MPM_MAT_NULL) ? MPM_MAT_NULL
                  : MM == (1 << MPM_MAT_DP) ? MPM_MAT_DP : MM == (1 << MPM_MAT_NEOH) ? MPM_MAT_NEOH : mc.type;
  switch (mtype) {
    case MPM_MAT_ELASTIC:
      sigrot(m, vort);
      m.ieinc = sig_dot(m);
      elastic_devi(m, G2);
      elastic_p(m, K3);
      m.seqv = eq_stress(m);
      m.ieinc = new_dot(m, m.ieinc);
      lieupd(m, I, dt);
      break;
    case MPM_MAT_PLA1: {
      sigrot(m, vort);
      m.ieinc = sig_dot(m);
      elastic_devi(m, G2);
      m.seqv = eq_stress(m);
      const T y0 = (T)mc.yield0;
      if (m.seqv > y0) {
        const T depeff = (m.seqv - m.sigy) / ((T)1.5 * G2);
        m.epeff = m.epeff + depeff;
        jc_assign<MM>(m, (double)depeff);
        const T ratio = y0 / m.seqv;
#pragma unroll
        for (int k = 0; k < 6; k++) m.sd[k] = m.sd[k] * ratio;
        m.seqv = m.seqv * ratio;
      }
      elastic_p(m, K3);
      m.ieinc = new_dot(m, m.ieinc);
      lieupd(m, I, dt);
      break;
    }
    case MPM_MAT_PLA2: {
      sigrot(m, vort);
      m.ieinc = sig_dot(m);
      elastic_devi(m, G2);
      m.seqv = eq_stress(m);
      if (m.seqv > m.sigy) {
        const T H = (T)mc.PlaMod;
        const T depeff = (m.seqv - m.sigy) / ((T)1.5 * G2 + H);
        m.epeff = m.epeff + depeff;
        jc_assign<MM>(m, (double)depeff);
        m.sigy = m.sigy + H * depeff;
        const T ratio = m.sigy / m.seqv;
#pragma unroll
        for (int k = 0; k < 6; k++) m.sd[k] = m.sd[k] * ratio;
        m.seqv = m.seqv * ratio;
      }
      elastic_p(m, K3);
      m.ieinc = new_dot(m, m.ieinc);
      lieupd(m, I, dt);
      break;
    }
    case MPM_MAT_NULL:
      m3dm7(m, mc, dt);
      seleos<T, MM>(m, p, mc, I, dt, false, 0.0);
      break;
    case MPM_MAT_SJC:
      sigrot(m, vort);
      m3dm5<T, MM>(m, mc, G2, dtd);
      seleos<T, MM>(m, p, mc, I, dt, false, 0.0);
      break;
    case MPM_MAT_JOHN:   // Constitution.f90:224-237: bulkq and its energy term twice (here and in lieupd)
      if (MM & MM_JC) {
        bool melted = false;
        sigrot(m, vort);
        m3dm48<T, MM, false>(m, mc, G2, K3, dtd, p.d(D_TEMP), melted);
        if (m.sm < (T)0) bulkq(m, I, dt);
        m.ieinc = m.ieinc - (T)2 * m.bqf * (m.dinc[0] + m.dinc[1] + m.dinc[2]);
        lieupd(m, I, dt);
        m.jcf |= JC_HEAT;   // the temperature rise needs the increment the particle reads
      }
      break;
    case MPM_MAT_JCF:    // Constitution.f90:278-319 up to the EOS; the rest reads depeff (k_x_jc_tail)
      if (MM & MM_JC) {
        sigrot(m, vort);
        bool failure = p.d(D_FAIL) != 0.0;
        if (!failure) {
          m3dm48<T, MM, true>(m, mc, G2, K3, dtd, p.d(D_TEMP), failure);
          seleos<T, MM>(m, p, mc, I, dt, failure, 0.0);
          m.jcf |= JC_TAIL | (failure ? JC_MELTED : 0);
        } else {   // a failed particle flows as the null material (M3DM7)
          m3dm7(m, mc, dt);
          seleos<T, MM>(m, p, mc, I, dt, true, 0.0);
        }
      }
      break;
    case MPM_MAT_SJCF: {   // Constitution.f90:245-276
      sigrot(m, vort);
      bool failure = p.d(D_FAIL) != 0.0;
      if (!failure) {
        m3dm5<T, MM>(m, mc, G2, dtd);
        seleos<T, MM>(m, p, mc, I, dt, false, 0.0);
        if ((double)m.epeff > mc.epf) {
          m.epeff = (T)(mc.epf + (double)0.0000001f);
          failure = true;
        }
        if (failure && m.sm < (T)0) {
          m.sm = (T)0;
          m.bqf = (T)0;
          p.d(D_VOL) = m.vold;
        }
        if (failure) {
#pragma unroll
          for (int k = 0; k < 6; k++) m.sd[k] = (T)0;
          m.iener = m.iener - (double)m.ieinc + m.dvol * (double)(m.smold + m.sm - (T)2 * m.bqf);
          p.d(D_FAIL) = 1.0;
        }
      } else {   // a failed particle flows as the null material (M3DM7)
        m3dm7(m, mc, dt);
        seleos<T, MM>(m, p, mc, I, dt, true, 0.0);
      }
      break;
    }
    case MPM_MAT_HIEX:   // M3DM6 (:1103-1121): burnt once the time passes the lighting time
      seleos<T, MM>(m, p, mc, I, dt, p.d(D_FAIL) != 0.0, time > p.d(D_LT) ? 1.0 : 0.0);
      break;
    case MPM_MAT_DP: {   // M3DM9 (:1214-1330)
      sigrot(m, vort);
      const T Gmod = (T)mc.Gmod, Kmod = (T)mc.Kmod;
      const T qfai = (T)mc.q_fai, kfai = (T)mc.k_fai, qpsi = (T)mc.q_psi, tenf = (T)mc.tenf;
      m.ieinc = sig_dot(m);
      elastic_devi(m, G2);
      elastic_p(m, K3);
      const T* sd = m.sd;
      const T J2 = (T)0.5 * (sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]) + sd[3] * sd[3] + sd[4] * sd[4] +
                   sd[5] * sd[5];
      const T Tau = sqrt(J2);
      m.seqv = Tau * (T)sqrtf(3.0f);
      const T dpFi = Tau + qfai * m.sm - kfai;
      const T dpsig = m.sm - tenf;
      const T third = (T)(1.0f / 3.0f), twoninth = (T)(2.0f / 9.0f);
      bool shear = false;
      if (dpsig < (T)0) {
        shear = dpFi > (T)0;
      } else {
        const T alphap = sqrt((T)1 + qfai * qfai) - qfai;
        const T Taup = kfai - qfai * tenf;
        const T dp_hfai = Tau - Taup - alphap * dpsig;
        if (dp_hfai > (T)0) {
          shear = true;
        } else {
          const T dlamd = (m.sm - tenf) / Kmod;
          m.sm = tenf;
          const T depeff = dlamd * third * (T)sqrtf(2.0f);
          m.epeff = m.epeff + depeff;
          jc_assign<MM>(m, (double)depeff);
        }
      }
      if (shear) {
        const T dlamd = dpFi / (Gmod + Kmod * qfai * qpsi);
        m.sm = m.sm - Kmod * qpsi * dlamd;
        const T newTau = kfai - qfai * m.sm;
        const T ratio = newTau / Tau;
#pragma unroll
        for (int k = 0; k < 6; k++) m.sd[k] = m.sd[k] * ratio;
        m.seqv = m.seqv * ratio;
        const T depeff = dlamd * sqrt(third + twoninth * (qpsi * qpsi));
        m.epeff = m.epeff + depeff;
        jc_assign<MM>(m, (double)depeff);
      }
      m.ieinc = new_dot(m, m.ieinc);
      lieupd(m, I, dt);
      break;
    }
    case MPM_MAT_NEOH: {   // M3DM10 (:725-780), evaluated in FP64
      const double G = 0.5 * mc.G2;
      const double J = det3(F);
      double B[6];
      B[0] = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[0][2] * F[0][2] - 1.0;
      B[1] = F[1][0] * F[1][0] + F[1][1] * F[1][1] + F[1][2] * F[1][2] - 1.0;
      B[2] = F[2][0] * F[2][0] + F[2][1] * F[2][1] + F[2][2] * F[2][2] - 1.0;
      B[3] = F[0][0] * F[1][0] + F[0][1] * F[1][1] + F[0][2] * F[1][2];
      B[4] = F[1][0] * F[2][0] + F[1][1] * F[2][1] + F[1][2] * F[2][2];
      B[5] = F[0][0] * F[2][0] + F[0][1] * F[2][1] + F[0][2] * F[2][2];
      m.ieinc = sig_dot(m);
      const double iJ = 1.0 / J;
      // ln J: the correctly rounded logarithm in the FP64 paths (see m_log)
      const double lnJ = m_log<T, MM>(J);
      const double lj = mc.lamda * lnJ * 1.0;
      double s[6];
      s[0] = iJ * (G * B[0] + lj);
      s[1] = iJ * (G * B[1] + lj);
      s[2] = iJ * (G * B[2] + lj);
      s[3] = iJ * G * B[4];
      s[4] = iJ * G * B[5];
      s[5] = iJ * G * B[3];
      const double smd = (s[0] + s[1] + s[2]) / 3.0;
      m.sm = (T)smd;
      m.sd[0] = (T)(s[0] - smd);
      m.sd[1] = (T)(s[1] - smd);
      m.sd[2] = (T)(s[2] - smd);
      m.sd[3] = (T)s[3];
      m.sd[4] = (T)s[4];
      m.sd[5] = (T)s[5];
      m.seqv = eq_stress(m);
      m.ieinc = new_dot(m, m.ieinc);
      lieupd(m, I, dt);
      break;
    }
    default:
      break;
  }
  p.t(T_SM) = m.sm;
  p.t(T_SEQV) = m.seqv;
#pragma unroll
  for (int k = 0; k < 6; k++) p.t(T_SD0 + k) = m.sd[k];
  // only pla2 and the Johnson-Cook models harden; cp never changes
  if (mm_has(MM, MPM_MAT_PLA2) || mm_has(MM, MPM_MAT_SJC) || mm_has(MM, MPM_MAT_SJCF) || (MM & MM_JC))
    p.t(T_SIGY) = m.sigy;
  p.t(T_EPEFF) = m.epeff;
  p.d(D_IE) = m.iener;
  p.t(T_Q) = m.bqf;
}

// ParticleStressUpdate after the gather (update_step.f90:566-642): strain
// and spin increments, F update (UVFp), Constitution, and Xp = XX for
// USL/MUSL.  L is the velocity gradient sum (deFp before the DT scaling).
// TLParticleStressUpdate's update of one particle (update_step.f90:785-826)
// with TLConstitution and M3DM11 (Constitution.f90:369-525, 782-882): F +=
// L DT (the gradient is material: no product), V = det(F) V0, the first
// Piola-Kirchhoff stress P = G (F - F^-T) + lambda ln J F^-T and its Cauchy
// stress P F^T / J, the internal energy from the mid-step P against the
// increment of F.  TLConstitution stores back sig_y and epeff from module
// variables no total-Lagrangian run sets (0), and q = 0.  ln J is the
// correctly rounded logarithm (the mode's).
template <typename T>
__device__ __noinline__ void tl_stress_update(const PoolPtrs<T>& P, const PV<T>& p, const InstC& I, double dtd,
                                              const T L[3][3], bool uvfp, const MatC& mc, InstT* tsp) {
  double dF[3][3], F[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      dF[a][b] = (double)L[a][b] * dtd;
      double f = p.d(D_F0 + 3 * a + b);
      if (uvfp) {
        f = f + dF[a][b];
        p.d(D_F0 + 3 * a + b) = f;
      }
      F[a][b] = f;
    }
#pragma unroll
  for (int a = 0; a < 3; a++) F[a][a] = F[a][a] + 1.0;
  const double vold = p.d(D_VOL);
  const double J = F[0][0] * F[1][1] * F[2][2] - F[0][0] * F[1][2] * F[2][1] - F[1][0] * F[0][1] * F[2][2] +
                   F[1][0] * F[0][2] * F[2][1] + F[2][0] * F[0][1] * F[1][2] - F[2][0] * F[0][2] * F[1][1];
  const double vol = J * p.d(D_VOL0);
  p.d(D_VOL) = vol;
  const double G = 0.5 * mc.G2, lamda = mc.lamda;
  // slots 1..9: 11 22 33 23 13 12 32 31 21
  const double f[9] = {F[0][0], F[1][1], F[2][2], F[1][2], F[0][2], F[0][1], F[2][1], F[2][0], F[1][0]};
  double fit[9];
  fit[0] = (f[1] * f[2] - f[3] * f[6]) / J;
  fit[1] = (f[0] * f[2] - f[4] * f[7]) / J;
  fit[2] = (f[0] * f[1] - f[5] * f[8]) / J;
  fit[3] = (f[5] * f[7] - f[0] * f[6]) / J;
  fit[4] = (f[8] * f[6] - f[1] * f[7]) / J;
  fit[5] = (f[3] * f[7] - f[2] * f[8]) / J;
  fit[6] = (f[4] * f[8] - f[3] * f[0]) / J;
  fit[7] = (f[3] * f[5] - f[4] * f[1]) / J;
  fit[8] = (f[4] * f[6] - f[5] * f[2]) / J;
  const double lnJ = mpm_log_cr(J);
  const long np = P.np, i = p.i;
  double pk[9], ieinc = 0.0;
  const double dfp[9] = {dF[0][0], dF[1][1], dF[2][2], dF[1][2], dF[0][2], dF[0][1], dF[2][1], dF[2][0], dF[1][0]};
#pragma unroll
  for (int q = 0; q < 9; q++) {
    pk[q] = G * (f[q] - fit[q]) + lamda * lnJ * fit[q];
    const double old = P.pk[q * np + i];
    ieinc = ieinc + 0.5 * (old + pk[q]) * dfp[q];
    P.pk[q * np + i] = pk[q];
  }
  const double ft[9] = {f[0], f[1], f[2], f[6], f[7], f[8], f[3], f[4], f[5]};
  double sig[6];
  sig[0] = (pk[0] * ft[0] + pk[5] * ft[8] + pk[4] * ft[7]) / J;
  sig[1] = (pk[8] * ft[5] + pk[1] * ft[1] + pk[3] * ft[6]) / J;
  sig[2] = (pk[7] * ft[4] + pk[6] * ft[3] + pk[2] * ft[2]) / J;
  sig[3] = (pk[8] * ft[4] + pk[1] * ft[3] + pk[3] * ft[2]) / J;
  sig[4] = (pk[0] * ft[4] + pk[5] * ft[3] + pk[4] * ft[2]) / J;
  sig[5] = (pk[0] * ft[5] + pk[5] * ft[1] + pk[4] * ft[6]) / J;
  Cst<T> m;
  m.sm = (T)((sig[0] + sig[1] + sig[2]) / 3.0);
  m.sd[0] = (T)(sig[0] - (double)m.sm);
  m.sd[1] = (T)(sig[1] - (double)m.sm);
  m.sd[2] = (T)(sig[2] - (double)m.sm);
  m.sd[3] = (T)sig[3];
  m.sd[4] = (T)sig[4];
  m.sd[5] = (T)sig[5];
  p.d(D_IE) = p.d(D_IE) + ieinc * 0.5 * (vol + vold);
  p.t(T_SM) = m.sm;
  p.t(T_SEQV) = eq_stress<T>(m);
#pragma unroll
  for (int k = 0; k < 6; k++) p.t(T_SD0 + k) = m.sd[k];
  p.t(T_SIGY) = (T)0;
  p.t(T_EPEFF) = (T)0;
  p.t(T_Q) = (T)0;
  if (vol < 0.0) atomicAdd(&tsp->negvol_count, 1);
  if (!isfinite(vol) || !isfinite((double)m.sm)) atomicAdd(&tsp->nan_count, 1);
}

template <typename T, int MM>
__device__ __forceinline__ void stress_tail(const PoolPtrs<T>& P, const PV<T>& p, const InstC& I, const InstT& ts,
                                            T de[6], T vort[3], const T L[3][3], bool uvfp, double x, double y,
                                            double z, bool move, const MatC& mc, Cst<T>& m, InstT* tsp) {
  if ((MM & MM_TL) && (I.flags & IF_TL)) {
    tl_stress_update<T>(P, p, I, ts.dt, L, uvfp, mc, tsp);
    if (move) {   // Xp = XX (update_step.f90:824)
      p.d(D_X) = x + (double)p.t(T_BX) * ts.dt;
      p.d(D_Y) = y + (double)p.t(T_BY) * ts.dt;
      p.d(D_Z) = z + (double)p.t(T_BZ) * ts.dt;
    }
    return;
  }
    const double dtd = ts.dt;
    const T dt = (T)dtd;
#pragma unroll
    for (int k = 0; k < 6; k++) de[k] = de[k] * dt;
#pragma unroll
    for (int k = 0; k < 3; k++) vort[k] = vort[k] * dt / (T)2;
    double F[3][3], deF[3][3];
    if (uvfp) {
      double F0[3][3], dF[3][3];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          dF[a][b] = (double)(L[a][b] * dt);
          F0[a][b] = p.d(D_F0 + 3 * a + b);
        }
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          const double f = F0[a][b] + dF[a][b] + dF[a][0] * F0[0][b] + dF[a][1] * F0[1][b] + dF[a][2] * F0[2][b];
          p.d(D_F0 + 3 * a + b) = f;
          F[a][b] = f;
          deF[a][b] = dF[a][b];
        }
#pragma unroll
      for (int a = 0; a < 3; a++) {
        deF[a][a] = deF[a][a] + 1.0;
        F[a][a] = F[a][a] + 1.0;
      }
    }
    constitution<T, MM>(p, I, mc, dt, dtd, ts.time, de, vort, F, deF, m);
    if ((MM & MM_JC) && (I.flags & IF_JC) && m.jcf) {   // the record the Johnson-Cook scan and tails read
      const long r = (long)I.p_begin + p.n(I_ORIG), np = P.np;
      P.jcf[r] = m.jcf;
      P.jc[r] = m.dep;
      P.jc[np + r] = m.vold;
      P.jc[2 * np + r] = (double)m.smold;
      P.jc[3 * np + r] = (double)m.ieinc;
    }
    // counted per instance: a negative volume (the reference writes a
    // warning and goes on, Constitution.f90:135-137) and a non-finite volume
    // or stress (NaN or inf, which the reference would carry on silently)
    if (m.vol < 0.0) atomicAdd(&tsp->negvol_count, 1);
    if (!isfinite(m.vol) || !isfinite((double)m.sm) || !isfinite((double)m.seqv)) atomicAdd(&tsp->nan_count, 1);
    if (move) {   // Xp = XX (update_step.f90:642)
      p.d(D_X) = x + (double)p.t(T_BX) * dtd;
      p.d(D_Y) = y + (double)p.t(T_BY) * dtd;
      p.d(D_Z) = z + (double)p.t(T_BZ) * dtd;
    }
}

// ParticleStressUpdate (update_step.f90:456-647): strain-rate gather, F
// update, constitution; optionally Xp = XX (USL/MUSL), the USF force
// scatter, and the energy partials.
// Waves per SIMD the stress kernel is compiled for: the fused gather,
// constitutive update and force scatter need many registers, and the
// B-spline and linear FP32 kernels (and the B-spline FP64 kernel, one wave
// fewer) fit the bound with little or no spilling, which lets another block
// stream its particles while one sums its force runs.  GIMP and linear FP64
// spill heavily under a bound and are left unbounded.
template <int SHAPE, typename T>
__host__ __device__ constexpr int stress_waves() {
  return SHAPE == MPM_SHAPE_GIMP ? 1 : (sizeof(T) == 4 ? 4 : (SHAPE == MPM_SHAPE_BSPLINE ? 3 : 1));
}

template <int SHAPE, typename T, int MODE, int MM>
__global__ void __launch_bounds__(256, (stress_waves<SHAPE, T>())) k_stress(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(P.ilv, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  const InstT ts = P.ts[ii];
  if (ts.done) return;
  constexpr int NS = NSlot<SHAPE>::n;
  using A = double;   // force accumulation type (see k_p2g)
  constexpr int NB = AosTile<T, 3>::S * (int)sizeof(T);   // nodal velocity per node (fast-path layout)
  constexpr int CAP = TILE_NODES;
  __shared__ __align__(16) char smem[CAP * (NB > 3 * (int)sizeof(A) ? NB : 3 * (int)sizeof(A))];
  T* tf = (T*)smem;
  A* ta = (A*)smem;   // force accumulation after the gather
  PV<T> p{P, i};
  const bool live = i < I.p_end;
  // the particle's loads are issued before the box reduction (two barriers)
  // and the tile load, so that their latencies overlap
  int c = -1, body = 0;
  double x = 0.0, y = 0.0, z = 0.0;
  Cst<T> cs;
  if (live) {
    c = p.n(I_CELL);
    x = p.d(D_X); y = p.d(D_Y); z = p.d(D_Z);
    body = p.n(I_BODY);
    cst_load<T, MM>(p, cs);
  }
  const bool act = c >= 0;
  int cx = 0, cy = 0, cz = 0;
  if (act) cell_coords(I, c, cx, cy, cz);
  const int jl = (int)(i - ((long)gran * PGRAN + (long)(blockIdx.x % bpg) * blockDim.x));
  const Box2 B = block_box_cached<SHAPE>(I, act, cx, cy, cz, jl, min(P.tile_cap, CAP), (int*)smem, P.boxes,
                                         P.box_mode, (i - jl) / 64);
  const bool fast = B.ok && B.inner;
  const int kb = B.side(jl);
  const Box Bk = B.get(kb);
  const int sb = Bk.bx, sc = Bk.bx * Bk.by;
  const long nn = P.nn;
  constexpr bool FLY = (MODE & ST_VFLY) != 0;
  const T* Vg = P.vg + I.node_off;
  if (B.ok && B.n) {
    const T* srcs[1] = {Vg};
    if (FLY) {
      if (fast) load_tile_fly<T, 3, true, true>(P, I, B, tf, CAP);
      else load_tile_fly<T, 3, true, false>(P, I, B, tf, CAP);
    } else {
      if (fast) load_tile_aos<T, 3>(I, B, tf, srcs, nn);
      else load_tile<T, 3>(I, B, tf, CAP, srcs, nn);
    }
    __syncthreads();
  }
  double ev[6] = {0, 0, 0, 0, 0, 0};
  Stencil<T> S;
  Slots<T> FS;
  if (act) {
    T de[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0}, vort[3] = {(T)0, (T)0, (T)0};
    T L[3][3] = {{(T)0, (T)0, (T)0}, {(T)0, (T)0, (T)0}, {(T)0, (T)0, (T)0}};
    const bool uvfp = (I.flags & IF_UVFP) != 0;
    if (fast) {
      make_slots<SHAPE, T>(I, Bk, x, y, z, c, FS);
      FS.e0 += B.off(kb);
      gather_grad<NS, T>(FS, tf, sb, sc, L);
      de[0] = L[0][0];
      de[1] = L[1][1];
      de[2] = L[2][2];
      de[3] = L[2][1] + L[1][2];
      de[4] = L[0][2] + L[2][0];
      de[5] = L[1][0] + L[0][1];
      vort[0] = L[2][1] - L[1][2];
      vort[1] = L[0][2] - L[2][0];
      vort[2] = L[1][0] - L[0][1];
    } else {
    build_stencil<SHAPE, T>(I, x, y, z, c, S);
    for_nodes<SHAPE, T>(I, S, [&](int k, int nx, int ny, int nz, T, T dX, T dY, T dZ) __attribute__((always_inline)) {
      const int e = B.ok ? B.idx(nx, ny, nz) : -1;
      T g[3];
      if (FLY && e < 0) node_fly<T, 3, true>(P, I, (long)I.node_off + k, g);
      const T v0 = e >= 0 ? tf[e] : (FLY ? g[0] : Vg[k]);
      const T v1 = e >= 0 ? tf[CAP + e] : (FLY ? g[1] : Vg[nn + k]);
      const T v2 = e >= 0 ? tf[2 * CAP + e] : (FLY ? g[2] : Vg[2 * nn + k]);
      de[0] = de[0] + dX * v0;
      de[1] = de[1] + dY * v1;
      de[2] = de[2] + dZ * v2;
      de[3] = de[3] + (dY * v2 + dZ * v1);
      de[4] = de[4] + (dZ * v0 + dX * v2);
      de[5] = de[5] + (dX * v1 + dY * v0);
      vort[0] = vort[0] + (dY * v2 - dZ * v1);
      vort[1] = vort[1] + (dZ * v0 - dX * v2);
      vort[2] = vort[2] + (dX * v1 - dY * v0);
      if (uvfp) {
        L[0][0] = L[0][0] + dX * v0; L[0][1] = L[0][1] + dY * v0; L[0][2] = L[0][2] + dZ * v0;
        L[1][0] = L[1][0] + dX * v1; L[1][1] = L[1][1] + dY * v1; L[1][2] = L[1][2] + dZ * v1;
        L[2][0] = L[2][0] + dX * v2; L[2][1] = L[2][1] + dY * v2; L[2][2] = L[2][2] + dZ * v2;
      }
    });
    }
    stress_tail<T, MM>(P, p, I, ts, de, vort, L, uvfp, x, y, z, (MODE & ST_MOVE) != 0, P.mats[P.bodies[body].mat],
                       cs, P.ts + ii);
  }
  if (MODE & ST_FORCE) {   // USF: GridMomentumUpdate scatter right after the update
    T* Fg = P.fg + I.node_off;
    constexpr bool CR = SHAPE == MPM_SHAPE_BSPLINE;
    using RR = RunRec<NS, false, true, CR>;
    extern __shared__ __align__(16) char run_mem[];
    const RunStage<T, RR::NV> R(run_mem, blockDim.x);
    constexpr int NVV = RunStage<T, RR::NV>::NG * RunStage<T, RR::NV>::GV;
    T fx[3] = {(T)0, (T)0, (T)0};
    T sxx = 0, syy = 0, szz = 0, sxy = 0, syz = 0, sxz = 0, vol = 0;
    if (act) {   // the updated stress, bulk viscosity and volume (as just stored)
      external_force(P, p, I, ts.time, fx);
      const T smq = cs.sm - cs.bqf;
      sxx = smq + cs.sd[0]; syy = smq + cs.sd[1]; szz = smq + cs.sd[2];
      sxy = cs.sd[5]; syz = cs.sd[3]; sxz = cs.sd[4];
      vol = (T)cs.vol;
    }
    if (fast) {   // stage the run records (the stage does not alias the tile)
      int key = -1;
      if (act) {
        T v[NVV];
#pragma unroll
        for (int q = 0; q < NVV; q++) v[q] = (T)0;
        run_rec_slots<NS, T, NVV, true, CR>(FS, v);
        v[RR::S + 0] = -(sxx * vol); v[RR::S + 1] = -(syy * vol); v[RR::S + 2] = -(szz * vol);
        v[RR::S + 3] = -(sxy * vol); v[RR::S + 4] = -(syz * vol); v[RR::S + 5] = -(sxz * vol);
        v[RR::S + 6] = fx[0]; v[RR::S + 7] = fx[1]; v[RR::S + 8] = fx[2];
        R.put(jl, v);
        key = FS.e0;
      }
      R.key[jl] = key;
      if (threadIdx.x == 0) *R.nh = 0;
    }
    if (B.ok && B.n) {
      __syncthreads();
      zero_tile<3>(B, ta, CAP);
      __syncthreads();
    }
    if (fast) {
      run_heads(R);
      __syncthreads();
      run_scatter_force<NS, T, RR::NV, RR::D, RR::S, CR>(R, B, (T)I.idcell, ta, CAP);
    } else if (act) {
      for_nodes<SHAPE, T>(I, S, [&](int k, int nx, int ny, int nz, T N, T dX, T dY, T dZ) __attribute__((always_inline)) {
        T f[3];
        f[0] = -(sxx * dX + sxy * dY + sxz * dZ) * vol + fx[0] * N;
        f[1] = -(sxy * dX + syy * dY + syz * dZ) * vol + fx[1] * N;
        f[2] = -(sxz * dX + syz * dY + szz * dZ) * vol + fx[2] * N;
        const int e = B.ok ? B.idx(nx, ny, nz) : -1;
        if (e >= 0) {
#pragma unroll
          for (int d = 0; d < 3; d++) atomicAdd(&ta[d * CAP + e], (A)f[d]);
        } else {
#pragma unroll
          for (int d = 0; d < 3; d++) atomic_add(&Fg[d * nn + k], f[d]);
        }
      });
    }
    if (B.ok && B.n) {
      __syncthreads();
      flush_tile3<T>(I, B, ta, CAP, Fg, nn);
    }
  }
  if (MODE & ST_ENERGY) {
    if (live) particle_energy(p, ev);
    block_energy(P.part, gran, blockIdx.x % bpg, ev);
  }
}

// ======================================================== SGMP (a6)
//
// The auxiliary grid has one node per cell centre (nb_centernode = NumCell,
// Grid.f90:264-268), stored per instance at the cell's index in the node
// range.  Particles map to the 8 centres around their nearest grid node with
// trilinear weights SGNShape(XXN) (Grid.f90:434-458).  Particle passes
// scatter to / gather from the centres; the spreads from centres to grid
// nodes (every node sums its adjacent cells in ascending cell order, which is
// the reference's summation order) and the node-to-centre averages are
// separate node- and cell-parallel kernels.

// auxiliary-node field slots (the three passes alias the same storage)
enum { AX_P = 0, AX_FE = 3, AX_FI = 6 };   // momentum, external and internal force
enum { AX_CO = 0, AX_CW = 6, AX_DF = 9 };  // strain rate, spin, velocity gradient
// The gradient and kinematics passes store their values centre-major (NV
// values per centre: AUX_COUNT for the 18 gradient values, 8 for velocity
// and acceleration, padded so that a centre is a few 16-byte words) so that
// the particles' 8-centre gathers are a few vector loads per centre instead
// of one load per field.
template <typename T, int NV>
__device__ __forceinline__ void centre_store(T* aux, long cell, const T (&v)[NV]) {
  using VT = typename V16<T>::type;
  VT* q = (VT*)(aux + cell * NV);
#pragma unroll
  for (int g = 0; g < NV / V16<T>::n; g++) q[g] = V16<T>::make(v + g * V16<T>::n);
}
template <typename T, int NV>
__device__ __forceinline__ void centre_load(const T* aux, long cell, T (&v)[NV]) {
  using VT = typename V16<T>::type;
  const VT* q = (const VT*)(aux + cell * NV);
#pragma unroll
  for (int g = 0; g < NV / V16<T>::n; g++) V16<T>::split(q[g], v + g * V16<T>::n);
}
constexpr int KIN_NV = 8;   // velocity (3), acceleration (3), padding
enum { AX_V = 0, AX_A = 3 };               // velocity, acceleration

__device__ constexpr int SGN[3][8] = {{-1, 1, 1, -1, -1, 1, 1, -1}, {-1, -1, 1, 1, -1, -1, 1, 1},
                                      {-1, -1, -1, -1, 1, 1, 1, 1}};

struct SgLoc {
  int cn1;        // first auxiliary node (1-based cell index) of the particle's auxiliary cell
  int bd;         // bits: nearest node is a boundary node (bd_type /= 0) along x, y, z
  double xxn[3];  // natural coordinates in the auxiliary cell (snapped on boundary axes)
};

// ParticleInitial, SGMP part (update_step.f90:432-449), with
// CenterInWhichCell (Grid.f90:157-190) and CenterCellNode (Grid.f90:386-403)
__device__ __forceinline__ bool sgmp_locate(const InstC& I, const uint8_t* fix, double x, double y, double z,
                                            SgLoc& L) {
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) return false;   // keeps the conversions below defined
  const double h = I.dcell;
  const double X[3] = {x, y, z};
  const int ix = (int)((x - I.span0[0] - h / 2.0) / h) + 1;
  const int iy = (int)((y - I.span0[1] - h / 2.0) / h) + 1;
  const int iz = (int)((z - I.span0[2] - h / 2.0) / h) + 1;
  const int ac = (iz - 1) * I.cncxy + (iy - 1) * I.cncx + ix;
  if (ac > I.cncell || ac <= 0) return false;
  const int a0 = ac - 1, az = a0 / I.cncxy, r = a0 - az * I.cncxy, ay = r / I.cncx, ax = r - ay * I.cncx + 1;
  L.cn1 = az * I.ncxy + ay * I.ncx + ax;
  const int jx = (int)((x - I.span0[0]) / h + 0.5) + 1;
  const int jy = (int)((y - I.span0[1]) / h + 0.5) + 1;
  const int jz = (int)((z - I.span0[2]) / h + 0.5) + 1;
  const int nd = (jz - 1) * I.ngxy + (jy - 1) * I.ngx + jx;
  L.bd = 0;
  if (nd < 1 || nd > I.nnode) {   // outside the grid (undefined in the reference)
    L.xxn[0] = L.xxn[1] = L.xxn[2] = 0.0;
    return true;
  }
  const int n0 = nd - 1, nz = n0 / I.ngxy, rr = n0 - nz * I.ngxy, ny = rr / I.ngx, nx = rr - ny * I.ngx;
  const double xg[3] = {(double)nx * h + I.span0[0], (double)ny * h + I.span0[1], (double)nz * h + I.span0[2]};
  const int fb = fix[I.node_off + n0];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double v = (X[d] - xg[d]) * I.ijac;
    if ((fb >> (3 + d)) & 1) {
      if (v > 0) v = 1;
      else if (v < 0) v = -1;
      L.bd |= 1 << d;
    }
    L.xxn[d] = v;
  }
  return true;
}

// 1-based cell index of auxiliary node n of the cell (0 when outside)
__device__ __forceinline__ int sgmp_node(const InstC& I, const SgLoc& L, int n) {
  const int off = (n & 3) == 1 ? 1 : (n & 3) == 2 ? 1 + I.ncx : (n & 3) == 3 ? I.ncx : 0;
  const int k = L.cn1 + off + (n >= 4 ? I.ncxy : 0);
  return (k <= I.ncell && k > 0) ? k : 0;
}

// SGNShape (Grid.f90:434-458)
template <typename T>
__device__ __forceinline__ void sgn_shape(const double xi[3], T shp[8]) {
  const T a = (T)xi[0], b = (T)xi[1], c = (T)xi[2];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const T sx = (T)SGN[0][n] * a + (T)1, sy = (T)SGN[1][n] * b + (T)1, sz = (T)SGN[2][n] * c + (T)1;
    shp[n] = sx * sy * sz * (T)0.125;
  }
}

// Particle scatters to the auxiliary nodes: mass and momentum
// (GridMomentumInitial, update_step.f90:318-345), forces (GridMomentumUpdate,
// :906-930), momentum re-map (GridMomentumMUSL, :1689-1707).
// LDS tile capacity (auxiliary nodes) of a scatter of nf FP64 fields (a
// 128-node tile sent most blocks of the Cook-plate batch to the per-particle
// global atomics and doubled the scatter time)
__host__ __device__ constexpr int sgp_cap(int nf) { return nf > 4 ? TILE_NODES / 2 : TILE_NODES; }
// staged record of a particle: natural coordinates (3), mass and velocity
// (4), external force (3) and stress times volume (6)
__host__ __device__ constexpr int sgp_nvr(int mode) {
  return 3 + ((mode & (P2G_MASS | P2G_MOM)) ? 4 : 0) + ((mode & P2G_FORCE) ? 9 : 0);
}
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_sgmp_p2g(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(P.ilv, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  constexpr bool MASS = (MODE & P2G_MASS) != 0, MOM = (MODE & P2G_MOM) != 0, FRC = (MODE & P2G_FORCE) != 0;
  // LDS accumulation tile over the block's box of auxiliary nodes (FP64):
  // mass, 3 momentum, 3 external and 6 internal force components
  constexpr int NF = (MASS ? 1 : 0) + (MOM ? 3 : 0) + (FRC ? 9 : 0);
  constexpr int CAP = sgp_cap(NF);
  __shared__ double tile[CAP * NF];
  PV<T> p{P, i};
  const bool live = i < I.p_end;
  if (MODE & P2G_ICELL) p.n(I_CELL) = live ? in_which_cell(I, p.d(D_X), p.d(D_Y), p.d(D_Z)) : -1;
  SgLoc L;
  const bool act = live && sgmp_locate(I, P.fix, p.d(D_X), p.d(D_Y), p.d(D_Z), L);
  // auxiliary coordinates of node 0 of the particle's auxiliary cell
  int ax = 0, ay = 0, az = 0;
  if (act) cell_coords(I, L.cn1, ax, ay, az);
  const Box B = block_box_n(act, ax, ay, az, 0, 1, I.ncx - 1, I.ncy - 1, I.ncz - 1, CAP);
  // with fewer than 2 cells along an axis the reference's linear index of
  // the +1 neighbour leaves the axis: direct path
  const bool fast = B.ok && B.inner && I.ncx > 1 && I.ncy > 1 && I.ncz > 1;
  if (fast) {
    for (int e = threadIdx.x; e < B.n; e += blockDim.x)
#pragma unroll
      for (int f = 0; f < NF; f++) tile[f * CAP + e] = 0.0;
    __syncthreads();
  }
  const long nn = P.nn;
  T* aux = P.aux + I.node_off;
  double* cmg = P.cmg + I.node_off;
  T m = (T)0, v[3] = {(T)0, (T)0, (T)0};
  // sg: sigma * vol, the factor of every node's Cfint term N sigma V
  T fx[3] = {(T)0, (T)0, (T)0}, sg[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
  if (act) {
    if (MASS || MOM) {
      m = p.t(T_MASS);
      v[0] = p.t(T_VX); v[1] = p.t(T_VY); v[2] = p.t(T_VZ);
    }
    if (FRC) {
      external_force(P, p, I, P.ts[ii].time, fx);
      const T smq = p.t(T_SM) - p.t(T_Q), vol = (T)p.d(D_VOL);
      sg[0] = (smq + p.t(T_SD0)) * vol; sg[1] = (smq + p.t(T_SD1)) * vol; sg[2] = (smq + p.t(T_SD2)) * vol;
      sg[3] = p.t(T_SD5) * vol; sg[4] = p.t(T_SD3) * vol; sg[5] = p.t(T_SD4) * vol;   // Cfint(4..6): sxy, syz, sxz
    }
  }
  if (fast) {
    // run sums: particles in the nearest-node order share their auxiliary
    // cell (and so its 8 nodes) in runs; a (run, node, part) item sums one
    // node's values over its part of the run in registers and adds the sums
    // to the tile, instead of one LDS atomic per particle, node and value
    constexpr int OM = 3, OF = OM + ((MASS || MOM) ? 4 : 0);
    constexpr int NVR = sgp_nvr(MODE);   // ..., fext (3), sigma vol (6)
    extern __shared__ __align__(16) char run_mem[];
    const RunStage<T, NVR> R(run_mem, blockDim.x);
    constexpr int NVV = RunStage<T, NVR>::NG * RunStage<T, NVR>::GV;
    const int jl = (int)(i - ((long)gran * PGRAN + (long)(blockIdx.x % bpg) * blockDim.x));
    int key = -1;
    if (act) {
      T u[NVV];
#pragma unroll
      for (int q = 0; q < NVV; q++) u[q] = (T)0;
#pragma unroll
      for (int d = 0; d < 3; d++) u[d] = (T)L.xxn[d];
      if (MASS || MOM) {
        u[OM] = m;
#pragma unroll
        for (int d = 0; d < 3; d++) u[OM + 1 + d] = v[d];
      }
      if (FRC) {
#pragma unroll
        for (int d = 0; d < 3; d++) u[OF + d] = fx[d];
#pragma unroll
        for (int k = 0; k < 6; k++) u[OF + 3 + k] = sg[k];
      }
      R.put(jl, u);
      key = (ax - B.x0) + B.bx * ((ay - B.y0) + B.by * (az - B.z0));
    }
    R.key[jl] = key;
    if (threadIdx.x == 0) *R.nh = 0;
    __syncthreads();
    run_heads(R);
    __syncthreads();
    const int nh = *R.nh;
    const int parts = run_parts<8>(nh);
    for (int item = threadIdx.x; item < nh * 8 * parts; item += blockDim.x) {
      const RunItem it = run_item<8>(R, nh, parts, item);
      if (it.j0 >= it.j1) continue;
      const int n = it.c;
      const T gx = (T)SGN[0][n], gy = (T)SGN[1][n], gz = (T)SGN[2][n];
      // each particle's term as in the per-particle path, summed in FP64
      double acc[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) acc[f] = 0.0;
      for (int j = it.j0; j < it.j1; j++) {
        T u[NVV];
        R.get(j, u);
        // SGNShape weight of node n (sgn_shape)
        const T N = (gx * u[0] + (T)1) * (gy * u[1] + (T)1) * (gz * u[2] + (T)1) * (T)0.125;
        int q = 0;
        if (MASS || MOM) {
          const T shm = N * u[OM];
          if (MASS) acc[q++] += (double)shm;
          if (MOM)
#pragma unroll
            for (int d = 0; d < 3; d++) acc[q++] += (double)(u[OM + 1 + d] * shm);
        }
        if (FRC) {
#pragma unroll
          for (int d = 0; d < 3; d++) acc[q++] += (double)(u[OF + d] * N);
#pragma unroll
          for (int k = 0; k < 6; k++) acc[q++] += (double)(N * u[OF + 3 + k]);
        }
      }
      // node n of the SNX/SNY/SNZ order: x + 1 for n & 3 in {1, 2}, y + 1 for {2, 3}, z + 1 for n >= 4
      const int e = it.e0 + ((n & 3) == 1 || (n & 3) == 2 ? 1 : 0) + ((n & 3) >= 2 ? B.bx : 0) +
                    (n >= 4 ? B.bx * B.by : 0);
#pragma unroll
      for (int f = 0; f < NF; f++) atomicAdd(&tile[f * CAP + e], acc[f]);
    }
  } else if (act) {
    T shp[8];
    sgn_shape<T>(L.xxn, shp);
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const T N = shp[n];
      T val[NF];
      int q = 0;
      if (MASS || MOM) {
        const T shm = N * m;
        if (MASS) val[q++] = shm;
        if (MOM)
#pragma unroll
          for (int d = 0; d < 3; d++) val[q++] = v[d] * shm;
      }
      if (FRC) {
#pragma unroll
        for (int d = 0; d < 3; d++) val[q++] = fx[d] * N;
#pragma unroll
        for (int k = 0; k < 6; k++) val[q++] = N * sg[k];
      }
      const int k = sgmp_node(I, L, n);
      if (!k) continue;
      const long c = k - 1;
      q = 0;
      if (MASS) atomic_add(&cmg[c], (double)val[q++]);
      if (MOM)
#pragma unroll
        for (int d = 0; d < 3; d++) atomic_add(&aux[(AX_P + d) * nn + c], val[q++]);
      if (FRC)
#pragma unroll
        for (int f = 0; f < 9; f++) atomic_add(&aux[(AX_FE + f) * nn + c], val[q++]);
    }
  }
  if (fast) {
    __syncthreads();
    for (int e = threadIdx.x; e < B.n; e += blockDim.x) {
      int ex, ey;
      const int r = fdivmod(e, B.bx, ex), ez = fdivmod(r, B.by, ey);
      const long c = (long)(B.z0 + ez) * I.ncxy + (long)(B.y0 + ey) * I.ncx + (B.x0 + ex);
      int q = 0;
      if (MASS) {
        const double v = tile[q++ * CAP + e];
        if (v != 0.0) atomic_add(&cmg[c], v);
      }
      if (MOM)
        for (int d = 0; d < 3; d++) {
          const double v = tile[q++ * CAP + e];
          if (v != 0.0) atomic_add(&aux[(AX_P + d) * nn + c], (T)v);
        }
      if (FRC)
        for (int f = 0; f < 9; f++) {
          const double v = tile[q++ * CAP + e];
          if (v != 0.0) atomic_add(&aux[(AX_FE + f) * nn + c], (T)v);
        }
    }
  }
}

// spreads from the auxiliary nodes to the grid nodes: SGN_MASSMOM
// (update_step.f90:394-405), SGN_FORCE (:990-1011), SGN_MOM (:1760-1768)

// index of node (a,b,c) of a cell in the CellsNode order (Grid.f90:372-379)
__device__ __forceinline__ int cell_local(int a, int b, int c) { return 4 * c + (b ? (a ? 2 : 3) : (a ? 1 : 0)); }

template <typename T, int WHAT>
__global__ void __launch_bounds__(256) k_sgmp_nodes(PoolPtrs<T> P, const int* nglist, const int* nginst, int bpg) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  const int ii = nginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  if (j >= (long)I.node_off + I.nnode) return;
  const int k = (int)(j - I.node_off);
  const int nz = k / I.ngxy, r = k - nz * I.ngxy, ny = r / I.ngx, nx = r - ny * I.ngx;
  const long nn = P.nn;
  const T* aux = P.aux + I.node_off;
  const double* cmg = P.cmg + I.node_off;
  const T ij4 = (T)I.ijac4;
  double mg = 0.0;
  T pv[3] = {(T)0, (T)0, (T)0}, fv[3] = {(T)0, (T)0, (T)0};
  for (int cz = nz - 1; cz <= nz; cz++) {
    if (cz < 0 || cz >= I.ncz) continue;
    for (int cy = ny - 1; cy <= ny; cy++) {
      if (cy < 0 || cy >= I.ncy) continue;
      for (int cx = nx - 1; cx <= nx; cx++) {
        if (cx < 0 || cx >= I.ncx) continue;
        const long c = (long)cz * I.ncxy + cy * I.ncx + cx;
        if (WHAT & SGN_MASSMOM) {
          if (cmg[c] > I.cutoff) {
            mg = mg + 0.125 * cmg[c];
#pragma unroll
            for (int d = 0; d < 3; d++) pv[d] = pv[d] + (T)0.125 * aux[(AX_P + d) * nn + c];
          }
        }
        if (WHAT & SGN_MOM) {
#pragma unroll
          for (int d = 0; d < 3; d++) pv[d] = pv[d] + (T)0.125 * aux[(AX_P + d) * nn + c];
        }
        if (WHAT & SGN_FORCE) {
          const int jl = cell_local(nx - cx, ny - cy, nz - cz);
          const T DX = (T)SGN[0][jl] * ij4, DY = (T)SGN[1][jl] * ij4, DZ = (T)SGN[2][jl] * ij4;
          T f0, f1, f2;
          if (I.flags & IF_TL) {   // CPKint (9, PK order xx yy zz yz xz xy zy zx yx), update_step.f90:1174-1178
            T q[9];
#pragma unroll
            for (int k = 0; k < 9; k++) q[k] = aux[(AX_FI + k) * nn + c];
            f0 = -(q[0] * DX + q[5] * DY + q[4] * DZ);
            f1 = -(q[8] * DX + q[1] * DY + q[3] * DZ);
            f2 = -(q[7] * DX + q[6] * DY + q[2] * DZ);
          } else {
            T fi[6];
#pragma unroll
            for (int q = 0; q < 6; q++) fi[q] = aux[(AX_FI + q) * nn + c];
            f0 = -(fi[0] * DX + fi[3] * DY + fi[5] * DZ);
            f1 = -(fi[3] * DX + fi[1] * DY + fi[4] * DZ);
            f2 = -(fi[5] * DX + fi[4] * DY + fi[2] * DZ);
          }
          fv[0] = fv[0] + (T)0.125 * aux[(AX_FE + 0) * nn + c] + f0;
          fv[1] = fv[1] + (T)0.125 * aux[(AX_FE + 1) * nn + c] + f1;
          fv[2] = fv[2] + (T)0.125 * aux[(AX_FE + 2) * nn + c] + f2;
        }
      }
    }
  }
  if (WHAT & SGN_MASSMOM) P.mg[j] = mg;
  if (WHAT & (SGN_MASSMOM | SGN_MOM)) {
    T* dst = (WHAT & SGN_TO_P2) ? P.p2 : P.pg;
#pragma unroll
    for (int d = 0; d < 3; d++) dst[d * nn + j] = pv[d];
  }
  if (WHAT & SGN_FORCE)
#pragma unroll
    for (int d = 0; d < 3; d++) P.fg[d * nn + j] = fv[d];
}

// cell-centre values from the grid nodes: SGC_GRAD (update_step.f90:497-525),
// SGC_KIN (:1405-1417 and BoundaryAuxiliaryNodeUpdate :1528-1666)

template <typename T, int WHAT>
__global__ void __launch_bounds__(256) k_sgmp_cells(PoolPtrs<T> P, const int* nglist, const int* nginst, int bpg) {
  int gran;
  const long j = node_index(nglist, bpg, gran);
  const int ii = nginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  const int c = (int)(j - I.node_off);
  if (c >= I.ncell) return;
  int cx, cy, cz;
  cell_coords(I, c + 1, cx, cy, cz);
  const long nn = P.nn;
  const long n1 = I.node_off + (long)cz * I.ngxy + (long)cy * I.ngx + cx;
  const long nodes[8] = {n1, n1 + 1, n1 + 1 + I.ngx, n1 + I.ngx,
                         n1 + I.ngxy, n1 + I.ngxy + 1, n1 + I.ngxy + 1 + I.ngx, n1 + I.ngxy + I.ngx};
  if (WHAT & SGC_GRAD) {
    const T ij4 = (T)I.ijac4;
    T co[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0}, cw[3] = {(T)0, (T)0, (T)0};
    T df[9] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
#pragma unroll
    for (int jl = 0; jl < 8; jl++) {
      const long k = nodes[jl];
      if (!(P.mg[k] > I.cutoff)) continue;
      const T v[3] = {P.vg[k], P.vg[nn + k], P.vg[2 * nn + k]};
      const T DX = (T)SGN[0][jl] * ij4, DY = (T)SGN[1][jl] * ij4, DZ = (T)SGN[2][jl] * ij4;
      co[0] = co[0] + DX * v[0];
      co[1] = co[1] + DY * v[1];
      co[2] = co[2] + DZ * v[2];
      co[3] = co[3] + DZ * v[1] + DY * v[2];
      co[4] = co[4] + DZ * v[0] + DX * v[2];
      co[5] = co[5] + DY * v[0] + DX * v[1];
      cw[0] = cw[0] + DY * v[2] - DZ * v[1];
      cw[1] = cw[1] + DZ * v[0] - DX * v[2];
      cw[2] = cw[2] + DX * v[1] - DY * v[0];
#pragma unroll
      for (int n = 0; n < 3; n++) {
        df[3 * n + 0] = df[3 * n + 0] + DX * v[n];
        df[3 * n + 1] = df[3 * n + 1] + DY * v[n];
        df[3 * n + 2] = df[3 * n + 2] + DZ * v[n];
      }
    }
    T g[AUX_COUNT];
#pragma unroll
    for (int q = 0; q < AUX_COUNT; q++) g[q] = (T)0;
#pragma unroll
    for (int q = 0; q < 6; q++) g[AX_CO + q] = co[q];
#pragma unroll
    for (int q = 0; q < 3; q++) g[AX_CW + q] = cw[q];
#pragma unroll
    for (int q = 0; q < 9; q++) g[AX_DF + q] = df[q];
    centre_store<T, AUX_COUNT>(P.aux, (long)I.node_off + c, g);
  }
  if (WHAT & SGC_KIN) {
    // a centre on a fixed or symmetric face takes the nodal values at its
    // face point; elsewhere the mean of its 8 nodes.  0.125 and the face
    // weights are powers of two (or 0), so (w PXg)/Mg = w (PXg/Mg) exactly.
    const int id[3] = {cx, cy, cz}, nc[3] = {I.ncx, I.ncy, I.ncz};
    bool face = false;
#pragma unroll
    for (int a = 0; a < 3; a++)
      if (nc[a] > 1 && ((id[a] == 0 && I.fixs[2 * a] != 0) || (id[a] == nc[a] - 1 && I.fixs[2 * a + 1] != 0)))
        face = true;
    T w[8];
    if (face) {
      double cxx[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        if (nc[a] <= 1) continue;
        if (id[a] == 0) cxx[a] = 1.0;
        if (id[a] == nc[a] - 1) cxx[a] = -1.0;
      }
      sgn_shape<T>(cxx, w);
    } else {
#pragma unroll
      for (int jl = 0; jl < 8; jl++) w[jl] = (T)0.125;
    }
    T cv[3] = {(T)0, (T)0, (T)0}, ca[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int jl = 0; jl < 8; jl++) {
      const long k = nodes[jl];
      if (!(P.mg[k] > I.cutoff)) continue;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        cv[d] = cv[d] + w[jl] * P.vg[d * nn + k];
        ca[d] = ca[d] + w[jl] * P.ag[d * nn + k];
      }
    }
    T g[KIN_NV];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      g[AX_V + d] = cv[d];
      g[AX_A + d] = ca[d];
    }
    g[6] = g[7] = (T)0;
    centre_store<T, KIN_NV>(P.aux, (long)I.node_off + c, g);
  }
}

// ParticleStressUpdate, SGMP: de, vort, deFp interpolated from the centres
// (update_step.f90:542-562), then the common update
template <typename T, int MODE, int MM>
__global__ void __launch_bounds__(256) k_sgmp_stress(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(0, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  const InstT ts = P.ts[ii];
  if (ts.done) return;
  PV<T> p{P, i};
  const bool live = i < I.p_end;
  SgLoc L;
  const double x = live ? p.d(D_X) : 0.0, y = live ? p.d(D_Y) : 0.0, z = live ? p.d(D_Z) : 0.0;
  // a total-Lagrangian instance keeps the cells of its initial positions (TLParticleInitial)
  const bool tl = (MM & MM_TL) && (I.flags & IF_TL);
  const bool act = live && (tl ? sgmp_locate(I, P.fix, p.d(D_XO), p.d(D_YO), p.d(D_ZO), L)
                               : sgmp_locate(I, P.fix, x, y, z, L));
  if (act) {
    T shp[8];
    sgn_shape<T>(L.xxn, shp);
    T de[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0}, vort[3] = {(T)0, (T)0, (T)0};
    T Lg[3][3] = {{(T)0, (T)0, (T)0}, {(T)0, (T)0, (T)0}, {(T)0, (T)0, (T)0}};
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const int k = sgmp_node(I, L, n);
      if (!k) continue;
      T g[AUX_COUNT];
      centre_load<T, AUX_COUNT>(P.aux, (long)I.node_off + k - 1, g);
#pragma unroll
      for (int q = 0; q < 6; q++) de[q] = de[q] + g[AX_CO + q] * shp[n];
#pragma unroll
      for (int q = 0; q < 3; q++) vort[q] = vort[q] + g[AX_CW + q] * shp[n];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Lg[a][b] = Lg[a][b] + shp[n] * g[AX_DF + 3 * a + b];
    }
    Cst<T> cs;
    cst_load<T, MM>(p, cs);
    const MatC& mc = P.mats[P.bodies[p.n(I_BODY)].mat];
    stress_tail<T, MM>(P, p, I, ts, de, vort, Lg, (I.flags & IF_UVFP) != 0, x, y, z, (MODE & ST_MOVE) != 0, mc, cs,
                       P.ts + ii);
  }
  if (MODE & ST_ENERGY) {
    double ev[6] = {0, 0, 0, 0, 0, 0};
    if (live) particle_energy(p, ev);
    block_energy(P.part, gran, blockIdx.x % bpg, ev);
  }
}

// ParticlePositionUpdate, SGMP (update_step.f90:1434-1525): particles whose
// nearest node is a boundary node interpolate at AuxiliaryGridSHP's
// coordinates (Grid.f90:193-220, its integer p_type read as /= 0)
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_sgmp_g2p(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(0, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  const InstT ts = P.ts[ii];
  if (ts.done) return;
  PV<T> p{P, i};
  const bool live = i < I.p_end;
  SgLoc L;
  const double x = live ? p.d(D_X) : 0.0, y = live ? p.d(D_Y) : 0.0, z = live ? p.d(D_Z) : 0.0;
  // a total-Lagrangian instance keeps the cells (and XN, XXN) of its initial positions
  const bool tl = (I.flags & IF_TL) != 0;
  const double lx = tl && live ? p.d(D_XO) : x, ly = tl && live ? p.d(D_YO) : y, lz = tl && live ? p.d(D_ZO) : z;
  const bool act = live && sgmp_locate(I, P.fix, lx, ly, lz, L);
  if (act) {
    double xi[3] = {L.xxn[0], L.xxn[1], L.xxn[2]};
    if (L.bd) {
      // XN = (Xp - Cxg(icell)) * iJacobi (update_step.f90:431); 0 outside the grid
      const int c = p.n(I_CELL);
      double xn[3] = {0.0, 0.0, 0.0};
      if (c > 0) {
        int cx, cy, cz;
        cell_coords(I, c, cx, cy, cz);
        const int cc[3] = {cx, cy, cz};
        const double X[3] = {lx, ly, lz};
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const double cxg = ((double)cc[d] * I.dcell + I.span0[d]) + I.dcell / 2.0;
          xn[d] = (X[d] - cxg) * I.ijac;
        }
      }
#pragma unroll
      for (int d = 0; d < 3; d++)
        if ((L.bd >> d) & 1) xi[d] = xn[d] < 0 ? 1 + 2 * xn[d] : 2 * xn[d] - 1;
    }
    T shp[8];
    sgn_shape<T>(xi, shp);
    T vx[3] = {(T)0, (T)0, (T)0}, ax[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const int k = sgmp_node(I, L, n);
      if (!k) continue;
      T g[KIN_NV];
      centre_load<T, KIN_NV>(P.aux, (long)I.node_off + k - 1, g);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        vx[d] = vx[d] + shp[n] * g[AX_V + d];
        ax[d] = ax[d] + shp[n] * g[AX_A + d];
      }
    }
    const double dt = ts.dt;
    const T tdt = (T)dt;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T v = p.t(T_VX + d);
      p.t(T_VX + d) = (ts.istep == 1) ? v + ax[d] * tdt * (T)0.5 : v + ax[d] * tdt;
      if (!(MODE & G2P_MOVE)) p.t(T_BX + d) = vx[d];   // XX = Xp + vbar DT, pending only without MOVE
    }
    if (MODE & G2P_MOVE) {
      p.d(D_X) = x + (double)vx[0] * dt;
      p.d(D_Y) = y + (double)vx[1] * dt;
      p.d(D_Z) = z + (double)vx[2] * dt;
    }
  }
  if (MODE & G2P_ENERGY) {
    double ev[6] = {0, 0, 0, 0, 0, 0};
    if (live) particle_energy(p, ev);
    block_energy(P.part, gran, blockIdx.x % bpg, ev);
  }
}

// calcEnergy partials alone (per-phase entry point)
template <typename T>
__global__ void __launch_bounds__(256) k_energy(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(P.ilv, glist, bpg, gran);
  const int ii = ginst[gran];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  if (P.ts[ii].done) return;
  double ev[6] = {0, 0, 0, 0, 0, 0};
  if (i < I.p_end) particle_energy(PV<T>{P, i}, ev);
  block_energy(P.part, gran, blockIdx.x % bpg, ev);
}

// the velocities of the particles of instances that k_finalize damped
// (dynamicRelaxationDamping, Particle.f90:190-197)
template <typename T>
__global__ void __launch_bounds__(256) k_damp(PoolPtrs<T> P, const int* glist, const int* ginst, int bpg) {
  int gran;
  const long i = part_index(P.ilv, glist, bpg, gran);
  const int ii = ginst[gran];
  if (!P.ts[ii].damped || i >= P.inst[ii].p_end) return;
  P.t[T_VX * P.np + i] = (T)0;
  P.t[T_VY * P.np + i] = (T)0;
  P.t[T_VZ * P.np + i] = (T)0;
}

// calcEnergy / UpdateDT finalisation: one block per active instance sums
// its block partials in a fixed order; DT from max |v| (update_step.f90:1845-1872);
// optional end-of-run test (MPM3D.f90:152-158).
template <typename T>
__global__ void __launch_bounds__(256) k_finalize(PoolPtrs<T> P, const int* ilist, int what) {
  constexpr int bpg = PGRAN / 64;   // energy partial slots per granule (one per 64 particles)
  const int ii = ilist[blockIdx.x];
  const InstC I = P.inst[ii];   // by value: read once, ahead of the kernel's stores
  InstT& ts = P.ts[ii];
  if (ts.done) return;
  const long g0 = I.p_begin / PGRAN;
  const long g1 = (I.p_end + PGRAN - 1) / PGRAN;
  const long nblk = (g1 - g0) * bpg;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (long b = threadIdx.x; b < nblk; b += blockDim.x) {
    const long g = g0 + b / bpg, sub = b % bpg;
    const double* o = P.part + (g * (PGRAN / 64) + sub) * 6;
    for (int q = 0; q < 5; q++) s[q] += o[q];
    s[5] = fmax(s[5], o[5]);
  }
  __shared__ double red[6][256];
  for (int q = 0; q < 6; q++) red[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
      red[5][threadIdx.x] = fmax(red[5][threadIdx.x], red[5][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (what & FIN_INTPEND) {   // ahead of the DT update: the step's own DT
    ts.integ_pending = 1;
    ts.integ_dt = ts.dt;
    ts.integ_half = ts.istep == 1;
  }
  if (what & FIN_ENERGY) {
    ts.ek_prev = ts.ek;
    ts.ek = red[0][0];
    ts.ei = red[1][0];
    ts.mom[0] = red[2][0];
    ts.mom[1] = red[3][0];
    ts.mom[2] = red[4][0];
  }
  ts.maxv = red[5][0];
  if (what & FIN_DT) {
    double dt = (double)1e6f;
    for (int m = 0; m < I.nmat; m++) {
      const double vv = P.mats[I.mat_off + m].csound + ts.maxv;
      dt = fmin(dt, I.dcell / vv);
    }
    ts.dt = dt * I.dtscale;
    if (I.flags & IF_DRDA) {   // dynamicRelaxationDamping (Particle.f90:172-202), after UpdateDT
      ts.damped = ts.ek - ts.ek_prev <= 0.0;
      if (ts.damped) ts.ek = 0.0;   // k_damp zeroes the velocities
    }
  }
  if ((what & FIN_DONE) && (I.flags & IF_ENDTIME)) {   // MPM3D.f90:152-158
    bool stop;
    if (I.flags & IF_QUASI_LOAD)
      stop = (ts.time > I.endtime && ts.ek / (ts.ek + ts.ei) < 0.005) && ts.time > 100.0;
    else
      stop = ts.time > I.endtime;
    if (stop) ts.done = 1;
  }
  if ((what & FIN_DONE) && I.max_steps > 0 && ts.istep >= I.max_steps) ts.done = 1;
}

#if MPMK_PART != 2   // defined once, in the FP64 object
// istep += 1; CurrentTime += DT (MPM3D.f90:72-74)
__global__ void k_begin_step(InstT* ts, const int* ilist, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  InstT& t = ts[ilist[k]];
  if (t.done) return;
  t.istep += 1;
  t.time = t.time + t.dt;
  t.ei = 0.0;
}
#endif

// gather one T field through a permutation (sort support)
template <typename S>
__global__ void __launch_bounds__(256) k_permute(S* dst, const S* src, const int* perm, const int* glist, int bpg,
                                                 long np, int nfield) {
  int gran;
  const long j = part_index(0, glist, bpg, gran);
  const long from = perm[j];
  for (int f = 0; f < nfield; f++) dst[(long)f * np + j] = src[(long)f * np + from];
}

#if MPMK_PART != 2
// Sort key of a particle: its cell in brick-major order (bricks of B^3
// cells, cells x-fastest inside a brick), so that a block of consecutive
// particles covers a compact box of nodes.  The cell is located from the
// current position (InWhichCell); particles outside the grid and padding
// slots get the largest key of their instance.  B = 1 is the plain cell order.
__device__ __forceinline__ unsigned brick_key_xyz(const InstC& I, int cx, int cy, int cz, int B) {
  const int nbx = (I.ncx + B - 1) / B, nby = (I.ncy + B - 1) / B;
  const unsigned brick = (unsigned)(((cz / B) * nby + cy / B) * nbx + cx / B);
  const unsigned local = (unsigned)(((cz % B) * B + cy % B) * B + cx % B);
  return brick * (unsigned)(B * B * B) + local;
}

__device__ __forceinline__ unsigned brick_key(const InstC& I, int c, int B) {
  int cx, cy, cz;
  cell_coords(I, c, cx, cy, cz);
  return brick_key_xyz(I, cx, cy, cz, B);
}

// Index of cell (x, y, z) along the 3D Hilbert curve of 2^bits cells per
// axis: J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707
// (2004) -- axes to transposed Gray code, then the bits interleaved (x most
// significant).  Consecutive cells along the curve are face neighbours, so a
// block of consecutive particles covers a compact box even at the ragged
// edges of a body, where brick orders jump between rows of bricks.
__device__ __forceinline__ unsigned hilbert_key(unsigned x, unsigned y, unsigned z, int bits) {
  unsigned X[3] = {x, y, z};
  const unsigned M = 1u << (bits - 1);
  for (unsigned Q = M; Q > 1; Q >>= 1) {
    const unsigned P = Q - 1;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const unsigned t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  unsigned t = 0;
  for (unsigned Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  unsigned h = 0;
  for (int b = bits - 1; b >= 0; b--)
#pragma unroll
    for (int i = 0; i < 3; i++) h = (h << 1) | (((X[i] ^ t) >> b) & 1u);
  return h;
}

// B = 0: Hilbert order; B > 0: brick-major order with B^3-cell bricks
__global__ void __launch_bounds__(256) k_sort_keys(const double* X, long np, const InstC* inst, const int* glist,
                                                   const int* ginst, int bpg, int B, unsigned* keys, int* vals) {
  int gran;
  const long j = part_index(0, glist, bpg, gran);
  const InstC& I = inst[ginst[gran]];
  // key coordinates: the cell, or for SGMP the auxiliary cell (the cell of
  // the nearest node in the centre grid, decomposed as sgmp_locate does), so
  // that particles sharing their 8 auxiliary nodes are consecutive
  bool in = false;
  int cx = 0, cy = 0, cz = 0;
  if (j < I.p_end) {
    if (I.shape == MPM_SHAPE_SGMP) {
      const double h = I.dcell;
      // a non-finite position sorts last (and keeps the conversions defined)
      const bool fin = isfinite(X[j]) && isfinite(X[np + j]) && isfinite(X[2 * np + j]);
      const int ix = (int)((X[j] - I.span0[0] - h / 2.0) / h);
      const int iy = (int)((X[np + j] - I.span0[1] - h / 2.0) / h);
      const int iz = (int)((X[2 * np + j] - I.span0[2] - h / 2.0) / h);
      const int ac = iz * I.cncxy + iy * I.cncx + ix + 1;
      if (fin && ac >= 1 && ac <= I.cncell) {
        const int a0 = ac - 1;
        cz = a0 / I.cncxy;
        const int r = a0 - cz * I.cncxy;
        cy = r / I.cncx;
        cx = r - cy * I.cncx;
        in = true;
      }
    } else {
      const int c = in_which_cell(I, X[j], X[np + j], X[2 * np + j]);
      if (c > 0) {
        cell_coords(I, c, cx, cy, cz);
        in = true;
      }
    }
  }
  const int hb = B == 0 ? hilbert_bits(I) : 0;
  if (hb > 0) {
    keys[j] = in ? hilbert_key((unsigned)cx, (unsigned)cy, (unsigned)cz, hb) : 1u << (3 * hb);
  } else {
    if (B == 0) B = 4;
    const int nbx = (I.ncx + B - 1) / B, nby = (I.ncy + B - 1) / B, nbz = (I.ncz + B - 1) / B;
    keys[j] = in ? brick_key_xyz(I, cx, cy, cz, B) : (unsigned)nbx * nby * nbz * (unsigned)(B * B * B);
  }
  vals[j] = (int)j;
}
#endif

// ---------------------------------------------------------- admission

// Scatter one admitted instance's particle state (field-major FP64 rows of
// mpm_pfield, n particles, uploaded raw) into its padded slot range [b0,
// b0 + pad) of the pool, converting the bulk fields to T; pad slots are zero.
template <typename T>
__global__ void __launch_bounds__(256) k_admit(PoolPtrs<T> P, const double* ps, const int* pbody, int n, long b0,
                                               long pad, int body0) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= pad) return;
  const bool in = j < n;
  const long s = b0 + j, np = P.np;
  auto F = [&](int f) { return in ? ps[(size_t)f * n + j] : 0.0; };
  T* t = P.t;
  double* d = P.d;
  t[T_VX * np + s] = (T)F(PF_VX);
  t[T_VY * np + s] = (T)F(PF_VY);
  t[T_VZ * np + s] = (T)F(PF_VZ);
  t[T_BX * np + s] = (T)0;
  t[T_BY * np + s] = (T)0;
  t[T_BZ * np + s] = (T)0;
  t[T_FX * np + s] = (T)F(PF_FX);
  t[T_FY * np + s] = (T)F(PF_FY);
  t[T_FZ * np + s] = (T)F(PF_FZ);
  t[T_MASS * np + s] = (T)F(PF_MASS);
  t[T_SM * np + s] = (T)F(PF_SM);
  t[T_SEQV * np + s] = (T)F(PF_SEQV);
  t[T_SD0 * np + s] = (T)F(PF_SDXX);
  t[T_SD1 * np + s] = (T)F(PF_SDYY);
  t[T_SD2 * np + s] = (T)F(PF_SDZZ);
  t[T_SD3 * np + s] = (T)F(PF_SDYZ);
  t[T_SD4 * np + s] = (T)F(PF_SDXZ);
  t[T_SD5 * np + s] = (T)F(PF_SDXY);
  t[T_SIGY * np + s] = (T)F(PF_SIGY);
  t[T_EPEFF * np + s] = (T)F(PF_EPEFF);
  t[T_Q * np + s] = (T)F(PF_Q);
  t[T_CP * np + s] = (T)F(PF_CP);
  d[D_X * np + s] = F(PF_X);
  d[D_Y * np + s] = F(PF_Y);
  d[D_Z * np + s] = F(PF_Z);
  d[D_VOL * np + s] = F(PF_VOL);
  d[D_IE * np + s] = F(PF_IE);
#pragma unroll
  for (int k = 0; k < 9; k++) d[(D_F0 + k) * np + s] = F(PF_FP + k);
  d[D_MASS * np + s] = F(PF_MASS);
  d[D_TEMP * np + s] = F(PF_TEMP);
  d[D_DMG * np + s] = F(PF_DMG);
  d[D_LT * np + s] = F(PF_LT);
  d[D_FAIL * np + s] = F(PF_FAIL);
  d[D_SPRE * np + s] = F(PF_SPRE);
  d[D_XO * np + s] = F(PF_XO);
  d[D_YO * np + s] = F(PF_YO);
  d[D_ZO * np + s] = F(PF_ZO);
  d[D_VOL0 * np + s] = F(PF_VOL0);
  if (P.pk)
    for (int k = 0; k < TL_PK_COUNT; k++) P.pk[k * np + s] = F(PF_PK + k);
  P.i[I_BODY * np + s] = in ? body0 + pbody[j] : body0;
  P.i[I_ORIG * np + s] = (int)j;
  P.i[I_CELL * np + s] = -1;
}

#if MPMK_PART != 2
// Fixity byte of every node of one instance (Grid.f90:291-359): bits 0-2
// fixed x/y/z, bits 3-5 SGMP boundary node (bd_type /= 0, Grid.f90:303-316).
// One outer layer of a face with FixS = 1 (fixed) or 2 (symmetry: normal
// component only); two layers for GIMP, B-spline and SGMP.
__global__ void __launch_bounds__(256) k_fixity(uint8_t* fix, const InstC* inst, int id, long nslots) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nslots) return;
  const InstC& c = inst[id];
  uint8_t v = 0;
  if (k < c.nnode_all) {   // every grid component has the fixities of the grid
    const long kb = k % c.nnode;
    const int NGx = c.ngx, NGy = c.ngy, NGz = c.ngz;
    const int ix = (int)(kb % NGx) + 1, iy = (int)((kb / NGx) % NGy) + 1, iz = (int)(kb / c.ngxy) + 1;
    const int* FixS = c.fixs;
    const bool wide = c.shape == MPM_SHAPE_GIMP || c.shape == MPM_SHAPE_BSPLINE || c.shape == MPM_SHAPE_SGMP;
    int bx = 0, by = 0, bz = 0;
    if (wide) {
      if (((ix == 1 || ix == 2) && FixS[0] == 1) || ((ix == NGx || ix == NGx - 1) && FixS[1] == 1) ||
          ((iy == 1 || iy == 2) && FixS[2] == 1) || ((iy == NGy || iy == NGy - 1) && FixS[3] == 1) ||
          ((iz == 1 || iz == 2) && FixS[4] == 1) || ((iz == NGz || iz == NGz - 1) && FixS[5] == 1))
        bx = by = bz = 1;
      if (((ix == 1 || ix == 2) && FixS[0] == 2) || ((ix == NGx || ix == NGx - 1) && FixS[1] == 2)) bx = 1;
      if (((iy == 1 || iy == 2) && FixS[2] == 2) || ((iy == NGy || iy == NGy - 1) && FixS[3] == 2)) by = 1;
      if (((iz == 1 || iz == 2) && FixS[4] == 2) || ((iz == NGz || iz == NGz - 1) && FixS[5] == 2)) bz = 1;
    }
    if ((ix == 1 && FixS[0] == 1) || (ix == NGx && FixS[1] == 1) || (iy == 1 && FixS[2] == 1) ||
        (iy == NGy && FixS[3] == 1) || (iz == 1 && FixS[4] == 1) || (iz == NGz && FixS[5] == 1))
      bx = by = bz = 1;
    if ((ix == 1 && FixS[0] == 2) || (ix == NGx && FixS[1] == 2)) bx = 1;
    if ((iy == 1 && FixS[2] == 2) || (iy == NGy && FixS[3] == 2)) by = 1;
    if ((iz == 1 && FixS[4] == 2) || (iz == NGz && FixS[5] == 2)) bz = 1;
    int qx = 0, qy = 0, qz = 0;
    if (c.shape == MPM_SHAPE_SGMP) {
      qx = ((ix == 1 || ix == 2) && FixS[0] != 0) || ((ix == NGx || ix == NGx - 1) && FixS[1] != 0);
      qy = ((iy == 1 || iy == 2) && FixS[2] != 0) || ((iy == NGy || iy == NGy - 1) && FixS[3] != 0);
      qz = ((iz == 1 || iz == 2) && FixS[4] != 0) || ((iz == NGz || iz == NGz - 1) && FixS[5] != 0);
    }
    v = (uint8_t)(bx | (by << 1) | (bz << 2) | (qx << 3) | (qy << 4) | (qz << 5));
  }
  fix[c.node_off + k] = v;
}
#endif

// ---------------------------------------------------------- dispatch

template <int SHAPE, typename T, int MODE>
constexpr size_t p2g_stage_bytes(int bs) {
  using RR = RunRec<NSlot<SHAPE>::n, (MODE & (P2G_MASS | P2G_MOM)) != 0, (MODE & P2G_FORCE) != 0,
                   SHAPE == MPM_SHAPE_BSPLINE>;
  return run_stage_bytes<T, RR::NV>(bs);
}

// LDS of one block (static tile + box reduction + dynamic run stage) is
// kept within 64 KiB by halving the block size of a scatter launch (the
// particle blocks of a granule are independent; no per-block partials)
#define RUN_LDS_MAX 65536
template <typename F>
static LaunchP fit_lds(const LaunchP& L0, size_t static_bytes, F&& stage_bytes) {
  LaunchP L = L0;
  while (L.bs > 64 && static_bytes + 1024 + stage_bytes(L.bs) > RUN_LDS_MAX) {
    const int ngran = L.nblk / L.bpg;
    L.bs /= 2;
    L.bpg = PGRAN / L.bs;
    L.nblk = ngran * L.bpg;
  }
  return L;
}

template <int SHAPE, typename T>
void launch_p2g_s(int mode, const PoolPtrs<T>& P, const LaunchP& L0, hipStream_t s) {
#define P2G_CASE(M)                                                                                    \
  case M: {                                                                                            \
    constexpr int NT = ((M) & P2G_MOM ? 3 : 0) + ((M) & P2G_FORCE ? 3 : 0);                            \
    const LaunchP L = fit_lds(L0, (size_t)TILE_NODES * (((M) & P2G_MASS ? 8 : 0) + 8 * NT),            \
                              [](int bs) { return p2g_stage_bytes<SHAPE, T, M>(bs); });                \
    const size_t sh = p2g_stage_bytes<SHAPE, T, M>(L.bs);                                              \
    hipLaunchKernelGGL((k_p2g<SHAPE, T, M>), dim3(L.nblk), dim3(L.bs), sh, s, P, L.glist, L.ginst, L.bpg); \
  } break;
  switch (mode) {
    P2G_CASE(P2G_ICELL | P2G_MASS | P2G_MOM)
    P2G_CASE(P2G_ICELL | P2G_MASS | P2G_MOM | P2G_FORCE)
    P2G_CASE(P2G_FORCE)
    P2G_CASE(P2G_MOM)
    default: break;
  }
#undef P2G_CASE
}

template <int SHAPE, typename T>
void launch_g2p_s(int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  dim3 g(L.nblk), b(L.bs);
#define G2P_CASE(M) \
  case M: hipLaunchKernelGGL((k_g2p<SHAPE, T, M>), g, b, 0, s, P, L.glist, L.ginst, L.bpg); break;
  switch (mode) {
    G2P_CASE(0)
    G2P_CASE(G2P_MOVE)
    G2P_CASE(G2P_MOVE | G2P_ENERGY | G2P_VAFLY)
    G2P_CASE(G2P_MOVE | G2P_ENERGY | G2P_VAFLY | G2P_INTFLY)
    G2P_CASE(G2P_REMAP)
    default: break;
  }
#undef G2P_CASE
}

// the fused-step modes get one kernel per single material model, the
// per-phase modes only the general one
template <int SHAPE, typename T, int MODE>
void launch_stress_m(int mm, const PoolPtrs<T>& P, const LaunchP& L0, hipStream_t s) {
  constexpr int NV = RunRec<NSlot<SHAPE>::n, false, true, SHAPE == MPM_SHAPE_BSPLINE>::NV;
  // only the force-scatter mode stages runs in dynamic LDS: it runs with
  // smaller blocks when the stage does not fit
  const LaunchP L = ((MODE & ~ST_VFLY) == ST_FORCE)
                        ? fit_lds(L0, (size_t)TILE_NODES * (4 * sizeof(T) > 24 ? 4 * sizeof(T) : 24),
                                  [](int bs) { return run_stage_bytes<T, NV>(bs); })
                        : L0;
  dim3 g(L.nblk), b(L.bs);
  const size_t sh = (MODE & ST_FORCE) ? run_stage_bytes<T, NV>(L.bs) : 0;
#define ST_M(M) \
  case M: hipLaunchKernelGGL((k_stress<SHAPE, T, MODE, M>), g, b, sh, s, P, L.glist, L.ginst, L.bpg); break;
  switch (mm) {
    ST_M(1 << MPM_MAT_ELASTIC)
    ST_M(1 << MPM_MAT_PLA1)
    ST_M(1 << MPM_MAT_PLA2)
    ST_M(1 << MPM_MAT_NULL)
    ST_M(1 << MPM_MAT_DP)
    ST_M(1 << MPM_MAT_NEOH)
    default: hipLaunchKernelGGL((k_stress<SHAPE, T, MODE, MM_ALL>), g, b, sh, s, P, L.glist, L.ginst, L.bpg); break;
  }
#undef ST_M
}

template <int SHAPE, typename T>
void launch_stress_s(int mode, int mm, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  dim3 g(L.nblk), b(L.bs);
  switch (mode) {
    case ST_FORCE | ST_VFLY: launch_stress_m<SHAPE, T, ST_FORCE | ST_VFLY>(mm, P, L, s); break;
    case ST_MOVE | ST_ENERGY: launch_stress_m<SHAPE, T, ST_MOVE | ST_ENERGY>(mm, P, L, s); break;
    case ST_MOVE: hipLaunchKernelGGL((k_stress<SHAPE, T, ST_MOVE, MM_ALL>), g, b, 0, s, P, L.glist, L.ginst, L.bpg); break;
    case 0: hipLaunchKernelGGL((k_stress<SHAPE, T, 0, MM_ALL>), g, b, 0, s, P, L.glist, L.ginst, L.bpg); break;
    default: break;
  }
}

template <typename T>
void launch_p2g(int shape, int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (!L.nblk) return;
  switch (shape) {
    case MPM_SHAPE_LINEAR: launch_p2g_s<MPM_SHAPE_LINEAR, T>(mode, P, L, s); break;
    case MPM_SHAPE_GIMP: launch_p2g_s<MPM_SHAPE_GIMP, T>(mode, P, L, s); break;
    default: launch_p2g_s<MPM_SHAPE_BSPLINE, T>(mode, P, L, s); break;
  }
}

template <typename T>
void launch_g2p(int shape, int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (!L.nblk) return;
  if (shape == MPM_SHAPE_SGMP) {
    dim3 g(L.nblk), b(L.bs);
    if (mode == G2P_MOVE) hipLaunchKernelGGL((k_sgmp_g2p<T, G2P_MOVE>), g, b, 0, s, P, L.glist, L.ginst, L.bpg);
    else if (mode == (G2P_MOVE | G2P_ENERGY))
      hipLaunchKernelGGL((k_sgmp_g2p<T, G2P_MOVE | G2P_ENERGY>), g, b, 0, s, P, L.glist, L.ginst, L.bpg);
    else hipLaunchKernelGGL((k_sgmp_g2p<T, 0>), g, b, 0, s, P, L.glist, L.ginst, L.bpg);
    return;
  }
  switch (shape) {
    case MPM_SHAPE_LINEAR: launch_g2p_s<MPM_SHAPE_LINEAR, T>(mode, P, L, s); break;
    case MPM_SHAPE_GIMP: launch_g2p_s<MPM_SHAPE_GIMP, T>(mode, P, L, s); break;
    default: launch_g2p_s<MPM_SHAPE_BSPLINE, T>(mode, P, L, s); break;
  }
}

// the fused SGMP modes get one kernel per single material model and for the
// neo-Hookean + null pair, the per-phase mode only the general one
template <typename T, int MODE>
void launch_sgmp_stress_m(int mm, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  dim3 g(L.nblk), b(L.bs);
#define SGS_M(M) \
  case M: hipLaunchKernelGGL((k_sgmp_stress<T, MODE, M>), g, b, 0, s, P, L.glist, L.ginst, L.bpg); break;
  switch (mm) {
    SGS_M(1 << MPM_MAT_ELASTIC)
    SGS_M(1 << MPM_MAT_PLA2)
    SGS_M(1 << MPM_MAT_NULL)
    SGS_M(1 << MPM_MAT_NEOH)
    SGS_M((1 << MPM_MAT_NEOH) | (1 << MPM_MAT_NULL))
    default: hipLaunchKernelGGL((k_sgmp_stress<T, MODE, MM_ALL>), g, b, 0, s, P, L.glist, L.ginst, L.bpg); break;
  }
#undef SGS_M
}

template <typename T>
void launch_stress(int shape, int mode, int mm, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (!L.nblk) return;
  if (shape == MPM_SHAPE_SGMP) {
    dim3 g(L.nblk), b(L.bs);
    if (mode == ST_MOVE) hipLaunchKernelGGL((k_sgmp_stress<T, ST_MOVE, MM_ALL>), g, b, 0, s, P, L.glist, L.ginst, L.bpg);
    else if (mode == (ST_MOVE | ST_ENERGY)) launch_sgmp_stress_m<T, ST_MOVE | ST_ENERGY>(mm, P, L, s);
    else launch_sgmp_stress_m<T, 0>(mm, P, L, s);
    return;
  }
  switch (shape) {
    case MPM_SHAPE_LINEAR: launch_stress_s<MPM_SHAPE_LINEAR, T>(mode, mm, P, L, s); break;
    case MPM_SHAPE_GIMP: launch_stress_s<MPM_SHAPE_GIMP, T>(mode, mm, P, L, s); break;
    default: launch_stress_s<MPM_SHAPE_BSPLINE, T>(mode, mm, P, L, s); break;
  }
}

template <typename T>
void launch_sgmp_p2g(int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (!L.nblk) return;
#define SG_CASE(M)                                                                                       \
  case M: {                                                                                              \
    constexpr int NF = ((M) & P2G_MASS ? 1 : 0) + ((M) & P2G_MOM ? 3 : 0) + ((M) & P2G_FORCE ? 9 : 0);   \
    constexpr int NVR = sgp_nvr(M);                                                                      \
    const LaunchP L2 = fit_lds(L, (size_t)sgp_cap(NF) * NF * 8,                                          \
                               [](int bs) { return run_stage_bytes<T, NVR>(bs); });                      \
    const size_t sh = run_stage_bytes<T, NVR>(L2.bs);                                                    \
    hipLaunchKernelGGL((k_sgmp_p2g<T, M>), dim3(L2.nblk), dim3(L2.bs), sh, s, P, L2.glist, L2.ginst, L2.bpg); \
  } break;
  switch (mode) {
    SG_CASE(P2G_ICELL | P2G_MASS | P2G_MOM)
    SG_CASE(P2G_ICELL | P2G_MASS | P2G_MOM | P2G_FORCE)
    SG_CASE(P2G_FORCE)
    SG_CASE(P2G_MOM)
    default: break;
  }
#undef SG_CASE
}

template <typename T>
void launch_sgmp_nodes(int what, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (!L.nblk) return;
  dim3 g(L.nblk), b(L.bs);
#define SN_CASE(M) \
  case M: hipLaunchKernelGGL((k_sgmp_nodes<T, M>), g, b, 0, s, P, L.glist, L.ginst, L.bpg); break;
  switch (what) {
    SN_CASE(SGN_MASSMOM)
    SN_CASE(SGN_MASSMOM | SGN_FORCE)
    SN_CASE(SGN_FORCE)
    SN_CASE(SGN_MOM)
    SN_CASE(SGN_MOM | SGN_TO_P2)
    default: break;
  }
#undef SN_CASE
}

template <typename T>
void launch_sgmp_cells(int what, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (!L.nblk) return;
  dim3 g(L.nblk), b(L.bs);
  if (what == SGC_GRAD) hipLaunchKernelGGL((k_sgmp_cells<T, SGC_GRAD>), g, b, 0, s, P, L.glist, L.ginst, L.bpg);
  else hipLaunchKernelGGL((k_sgmp_cells<T, SGC_KIN>), g, b, 0, s, P, L.glist, L.ginst, L.bpg);
}

template <typename T>
void launch_integrate(const PoolPtrs<T>& P, const LaunchP& L, int va, hipStream_t s) {
  if (L.nblk) hipLaunchKernelGGL(k_integrate<T>, dim3(L.nblk), dim3(L.bs), 0, s, P, L.glist, L.ginst, L.bpg, va);
}

template <typename T>
void launch_bc(const PoolPtrs<T>& P, const LaunchP& L, int use_p2, int bc, int vel, hipStream_t s) {
  if (L.nblk)
    hipLaunchKernelGGL(k_bc<T>, dim3(L.nblk), dim3(L.bs), 0, s, P, L.glist, L.ginst, L.bpg, use_p2, bc, vel);
}

template <typename T>
void launch_node_va(const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (L.nblk) hipLaunchKernelGGL(k_node_va<T>, dim3(L.nblk), dim3(L.bs), 0, s, P, L.glist, L.ginst, L.bpg);
}

template <typename T>
void launch_zero_nodes(T* a, int ncomp, long stride, const LaunchP& L, hipStream_t s) {
  if (L.nblk) hipLaunchKernelGGL(k_zero_nodes<T>, dim3(L.nblk), dim3(L.bs), 0, s, a, ncomp, stride, L.glist, L.bpg);
}

template <typename T>
void launch_zero_grid(const PoolPtrs<T>& P, bool p2, const LaunchP& L, hipStream_t s) {
  if (L.nblk)
    hipLaunchKernelGGL(k_zero_grid<T>, dim3(L.nblk), dim3(L.bs), 0, s, P.mg, P.pg, P.fg, p2 ? P.p2 : (T*)nullptr,
                       P.nn, L.glist, L.bpg);
}

template <typename T>
void launch_energy(const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (L.nblk) hipLaunchKernelGGL(k_energy<T>, dim3(L.nblk), dim3(L.bs), 0, s, P, L.glist, L.ginst, L.bpg);
}

template <typename T>
void launch_finalize(const PoolPtrs<T>& P, const int* ilist, int ninst, int what, hipStream_t s) {
  if (ninst) hipLaunchKernelGGL(k_finalize<T>, dim3(ninst), dim3(256), 0, s, P, ilist, what);
}

// (a template: each precision's object holds its own instance)
template <typename T>
__global__ void k_clear_integ(InstT* ts, int ii) { ts[ii].integ_pending = 0; }

template <typename T>
void launch_integrate_range(const PoolPtrs<T>& P, int inst, long node0, long nnode, hipStream_t s) {
  if (nnode > 0)
    hipLaunchKernelGGL(k_integrate_range<T>, dim3((unsigned)((nnode + 255) / 256)), dim3(256), 0, s, P, inst, node0,
                       nnode);
  hipLaunchKernelGGL(k_clear_integ<T>, dim3(1), dim3(1), 0, s, P.ts, inst);
}

template <typename T>
void launch_damp(const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s) {
  if (L.nblk) hipLaunchKernelGGL(k_damp<T>, dim3(L.nblk), dim3(L.bs), 0, s, P, L.glist, L.ginst, L.bpg);
}

template <typename S>
void launch_permute(S* dst, const S* src, const int* perm, const LaunchP& L, long np, int nfield, hipStream_t s) {
  if (L.nblk) hipLaunchKernelGGL(k_permute<S>, dim3(L.nblk), dim3(L.bs), 0, s, dst, src, perm, L.glist, L.bpg, np, nfield);
}

// the raw fields (bulk, FP64 and integer, as doubles) of the particles of
// instance `id` whose original index is in want[0..k) (ascending), one row
// of T_COUNT + D_COUNT + I_COUNT values per wanted particle
template <typename T>
__global__ void __launch_bounds__(256) k_pack_subset(PoolPtrs<T> P, int id, const int* want, int k, double* rows) {
  const InstC& I = P.inst[id];
  const long i = (long)I.p_begin + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I.p_end) return;
  const int o = P.i[(long)I_ORIG * P.np + i];
  int lo = 0, hi = k - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (want[mid] < o) lo = mid + 1;
    else hi = mid;
  }
  if (k == 0 || want[lo] != o) return;
  double* r = rows + (long)lo * (T_COUNT + D_COUNT + I_COUNT);
  for (int f = 0; f < T_COUNT; f++) r[f] = (double)P.t[(long)f * P.np + i];
  for (int f = 0; f < D_COUNT; f++) r[T_COUNT + f] = P.d[(long)f * P.np + i];
  for (int f = 0; f < I_COUNT; f++) r[T_COUNT + D_COUNT + f] = (double)P.i[(long)f * P.np + i];
}

template <typename T>
void launch_pack_subset(const PoolPtrs<T>& P, int id, long n, const int* want, int k, double* rows, hipStream_t s) {
  if (n > 0 && k > 0)
    hipLaunchKernelGGL(k_pack_subset<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, id, want, k, rows);
}

template <typename T>
void launch_admit(const PoolPtrs<T>& P, const double* ps, const int* pbody, int n, long b0, long pad, int body0,
                  hipStream_t s) {
  if (pad > 0)
    hipLaunchKernelGGL(k_admit<T>, dim3((unsigned)((pad + 255) / 256)), dim3(256), 0, s, P, ps, pbody, n, b0, pad,
                       body0);
}

#if MPMK_PART != 2
// zero n ints (box-cache invalidation; a kernel so that it may be captured)
__global__ void __launch_bounds__(256) k_clear_ints(int* a, long n) {
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) a[j] = 0;
}

// keys[j] |= (rank of the segment holding slot j) << shift, for j in
// [b0, b0 + n); begins: the segments' first slots, ascending
__global__ void __launch_bounds__(256) k_seg_prefix(unsigned* keys, const int* begins, int nseg, long b0, long n,
                                                    int shift) {
  const long j = b0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= b0 + n) return;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (begins[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  keys[j] |= (unsigned)lo << shift;
}

void launch_seg_prefix(unsigned* keys, const int* begins, int nseg, long b0, long n, int shift, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_seg_prefix, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keys, begins, nseg, b0, n,
                       shift);
}

// ParticleInitial's cell location (InWhichCell) of caller-owned particle
// arrays, the sort key of the stable cell order (particles outside the grid
// after every cell) and the identity as the sort values
__global__ void __launch_bounds__(256) k_cell_keys(InstC I, const double* x, const double* y, const double* z,
                                                   long n, int* icell, unsigned* key, int* val) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int c = in_which_cell(I, x[j], y[j], z[j]);
  icell[j] = c;
  key[j] = c < 0 ? (unsigned)I.ncell + 1u : (unsigned)c;
  val[j] = (int)j;
}

void launch_cell_keys(const InstC& I, const double* x, const double* y, const double* z, long n, int* icell,
                      unsigned* key, int* val, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_cell_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, I, x, y, z, n, icell, key,
                       val);
}

void launch_clear_ints(int* a, long n, hipStream_t s) {
  if (n > 0) {
    const long nb = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_clear_ints, dim3((unsigned)nb), dim3(256), 0, s, a, n);
  }
}

void launch_fixity(uint8_t* fix, const InstC* inst, int id, long nslots, hipStream_t s) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_fixity, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s, fix, inst, id, nslots);
}

void launch_begin_step(InstT* ts, const int* ilist, int n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_begin_step, dim3((n + 255) / 256), dim3(256), 0, s, ts, ilist, n);
}

void launch_sort_keys(const double* X, long np, const InstC* inst, const LaunchP& L, int brick, unsigned* keys,
                      int* vals, hipStream_t s) {
  if (L.nblk)
    hipLaunchKernelGGL(k_sort_keys, dim3(L.nblk), dim3(L.bs), 0, s, X, np, inst, L.glist, L.ginst, L.bpg, brick, keys,
                       vals);
}
#endif

#include "smooth.inc"

#if MPMK_PART != 2
#include "exact.inc"
#endif

#define INST(T)                                                                                     \
  template void launch_p2g<T>(int, int, const PoolPtrs<T>&, const LaunchP&, hipStream_t);          \
  template void launch_g2p<T>(int, int, const PoolPtrs<T>&, const LaunchP&, hipStream_t);          \
  template void launch_stress<T>(int, int, int, const PoolPtrs<T>&, const LaunchP&, hipStream_t);  \
  template void launch_integrate<T>(const PoolPtrs<T>&, const LaunchP&, int, hipStream_t);         \
  template void launch_bc<T>(const PoolPtrs<T>&, const LaunchP&, int, int, int, hipStream_t);      \
  template void launch_node_va<T>(const PoolPtrs<T>&, const LaunchP&, hipStream_t);                \
  template void launch_zero_nodes<T>(T*, int, long, const LaunchP&, hipStream_t);                  \
  template void launch_zero_grid<T>(const PoolPtrs<T>&, bool, const LaunchP&, hipStream_t);        \
  template void launch_energy<T>(const PoolPtrs<T>&, const LaunchP&, hipStream_t);                 \
  template void launch_finalize<T>(const PoolPtrs<T>&, const int*, int, int, hipStream_t);         \
  template void launch_damp<T>(const PoolPtrs<T>&, const LaunchP&, hipStream_t);                    \
  template void launch_integrate_range<T>(const PoolPtrs<T>&, int, long, long, hipStream_t);        \
  template void launch_permute<T>(T*, const T*, const int*, const LaunchP&, long, int, hipStream_t);  \
  template void launch_sgmp_p2g<T>(int, const PoolPtrs<T>&, const LaunchP&, hipStream_t);         \
  template void launch_sgmp_nodes<T>(int, const PoolPtrs<T>&, const LaunchP&, hipStream_t);       \
  template void launch_sgmp_cells<T>(int, const PoolPtrs<T>&, const LaunchP&, hipStream_t);                     \
  template void launch_admit<T>(const PoolPtrs<T>&, const double*, const int*, int, long, long, int, hipStream_t); \
  template void launch_pack_subset<T>(const PoolPtrs<T>&, int, long, const int*, int, double*, hipStream_t); \
  template void launch_smooth_save_x<T>(const PoolPtrs<T>&, const SmPtrs&, const LaunchP&, hipStream_t);    \
  template void launch_smooth_nodes<T>(int, int, const PoolPtrs<T>&, const SmPtrs&, const LaunchP&, hipStream_t); \
  template void launch_smooth<T>(int, const PoolPtrs<T>&, const SmPtrs&, const LaunchP&, const LaunchP&, hipStream_t);
#if MPMK_PART != 2
INST(double)
template void launch_permute<int>(int*, const int*, const int*, const LaunchP&, long, int, hipStream_t);
template void launch_zero_nodes<uint8_t>(uint8_t*, int, long, const LaunchP&, hipStream_t);
#endif
#if MPMK_PART != 1
INST(float)
#endif

}  // namespace mpmk
