// kernels.h -- launch interface of kernels.hip (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"
#include "mpm3d_problem.h"

// k_p2g modes
#define P2G_ICELL 1   // locate particles (ParticleInitial) and store icell
#define P2G_MASS 2    // Mg += N m
#define P2G_MOM 4     // PXg += v N m
#define P2G_FORCE 8   // FXg += f_int + f_ext
#define P2G_TO_P2 16  // momentum into the re-map buffer

// k_g2p modes
#define G2P_MOVE 1    // USF: Xp = XX
#define G2P_REMAP 2   // MUSL: re-map momentum into p2
#define G2P_ENERGY 4  // energy partials
#define G2P_VAFLY 8   // nodal velocity/acceleration formed from PXg, FXg, Mg in the tile load
#define G2P_INTFLY 16 // with VAFLY: IntegrateMomentum folded into the tile load (the grid stays un-integrated)

// k_stress modes
#define ST_MOVE 1     // USL/MUSL: Xp = XX
#define ST_FORCE 2    // USF: force scatter with the updated stress
#define ST_ENERGY 4   // energy partials
#define ST_VFLY 8     // nodal velocity formed from PXg, Mg and the fixities in the tile load

// SGMP: spreads from the auxiliary nodes to the grid nodes
#define SGN_MASSMOM 1   // Mg, PXg from Cmg, Cpxg (CutOff test)
#define SGN_FORCE 2     // FXg from Cfext, Cfint
#define SGN_MOM 4       // PXg from Cpxg (MUSL, no CutOff test)
#define SGN_TO_P2 8     // SGN_MOM into the re-map buffer
// SGMP: cell-centre values from the grid nodes
#define SGC_GRAD 1      // strain rate, spin, velocity gradient
#define SGC_KIN 2       // velocity, acceleration
// SGMP auxiliary-node fields per cell (the passes alias the storage); the
// force pass keeps Cfext, Cfint in slots 3..11, behind the momentum slots;
// the gradient pass stores its 18 values centre-major with a stride of 20
// (16-byte aligned in both precisions)
#define AUX_COUNT 20
#define AUX_FORCE_OFF 3

// reference-order scatter (exact FP64 mode): what k_x_scatter sums
#define XS_MASS 1     // Mg (SGMP: Cmg)
#define XS_MOM 2      // PXg (SGMP: Cpxg)
#define XS_FORCE 4    // FXg (SGMP: Cfext, Cfint)
#define XS_SGMP 8     // to the auxiliary nodes
#define XS_NDIR 16    // with XS_FORCE: the contact normals of contact instances (mass gradients)

// SmoothStressbyGrid node passes (launch_smooth_nodes)
#define SM_ZERO 1     // clear the sums and flags
#define SM_NODES 2    // nodal pressure (SGMP: through the cell centres, and back to them)

// all material models (general stress kernel)
#define MM_ALL 0xffff
// flag bit of the material mask: the neo-Hookean ln J is the correctly
// rounded logarithm (crlog.h) instead of the device libm's
#define MM_CRLOG (1 << 30)
// flag bit of the material mask: total-Lagrangian instances may be met
// (TLConstitution; reference-order mode only)
#define MM_TL (1 << 29)
// flag bit of the material mask: Johnson-Cook instances (john, jcf) may be
// met, and every model's update records the plastic strain increment it
// assigned (reference-order mode only)
#define MM_JC (1 << 28)

// Johnson-Cook record flags (PoolPtrs::jcf)
#define JC_ASSIGNED 1   // the update assigned a plastic strain increment (any plastic model)
#define JC_HEAT 2       // john: the temperature rise follows with the increment it reads
#define JC_TAIL 4       // jcf (not failed before): failure, damage and temperature follow
#define JC_MELTED 8     // jcf: failed by melting inside M3DM8

// k_finalize parts
#define FIN_ENERGY 1
#define FIN_DT 2
#define FIN_DONE 4
#define FIN_INTPEND 8   // the step's IntegrateMomentum was folded into g2p: mark the grid un-integrated

// a launch over a list of granules (particle or node)
struct LaunchP {
  const int* glist;   // device: granule ids
  const int* ginst;   // device: instance of every pool granule
  int nblk;           // blocks = granules * bpg
  int bpg;            // blocks per granule = GRAN / bs
  int bs;             // block size
};

namespace mpmk {
template <typename T> void launch_p2g(int shape, int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
template <typename T> void launch_g2p(int shape, int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
// mm: material models present in the launch's instances (bits 1 << type)
template <typename T>
void launch_stress(int shape, int mode, int mm, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
template <typename T> void launch_integrate(const PoolPtrs<T>& P, const LaunchP& L, int va, hipStream_t s);
template <typename T>
void launch_bc(const PoolPtrs<T>& P, const LaunchP& L, int use_p2, int bc, int vel, hipStream_t s);
template <typename T> void launch_node_va(const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
template <typename T> void launch_zero_nodes(T* a, int ncomp, long stride, const LaunchP& L, hipStream_t s);
template <typename T> void launch_zero_grid(const PoolPtrs<T>& P, bool p2, const LaunchP& L, hipStream_t s);
template <typename T> void launch_energy(const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
template <typename T>
void launch_integrate_range(const PoolPtrs<T>& P, int inst, long node0, long nnode, hipStream_t s);
template <typename T> void launch_damp(const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
template <typename T>
void launch_finalize(const PoolPtrs<T>& P, const int* ilist, int ninst, int what, hipStream_t s);
void launch_begin_step(InstT* ts, const int* ilist, int n, hipStream_t s);
template <typename S>
void launch_permute(S* dst, const S* src, const int* perm, const LaunchP& L, long np, int nfield, hipStream_t s);
// admission: raw FP64 particle rows -> pool slots [b0, b0+pad); node fixities of instance id
template <typename T>
void launch_admit(const PoolPtrs<T>& P, const double* ps, const int* pbody, int n, long b0, long pad, int body0,
                  hipStream_t s);
void launch_fixity(uint8_t* fix, const InstC* inst, int id, long nslots, hipStream_t s);
// raw fields of the particles of instance id with original index in want (ascending, k of them):
// T_COUNT + D_COUNT + I_COUNT doubles per particle; n: the instance's particle count
template <typename T>
void launch_pack_subset(const PoolPtrs<T>& P, int id, long n, const int* want, int k, double* rows, hipStream_t s);
void launch_clear_ints(int* a, long n, hipStream_t s);
// InWhichCell of n particles (x, y, z), their cell-order sort keys and identity values
void launch_cell_keys(const InstC& I, const double* x, const double* y, const double* z, long n, int* icell,
                      unsigned* key, int* val, hipStream_t s);
void launch_seg_prefix(unsigned* keys, const int* begins, int nseg, long b0, long n, int shift, hipStream_t s);
template <typename T> void launch_sgmp_p2g(int mode, const PoolPtrs<T>& P, const LaunchP& L, hipStream_t s);
template <typename T> void launch_sgmp_nodes(int what, const PoolPtrs<T>& P, const LaunchP& N, hipStream_t s);
template <typename T> void launch_sgmp_cells(int what, const PoolPtrs<T>& P, const LaunchP& N, hipStream_t s);
void launch_sort_keys(const double* X, long np, const InstC* inst, const LaunchP& L, int brick, unsigned* keys,
                      int* vals, hipStream_t s);

// stable LSD radix sort of (key, value) pairs over key bits [begin_bit,
// end_bit) (radix.hip); returns 0 when the result is in (k0, v0), 1 when in
// (k1, v1); scratch: radix_scratch_words(n) words
size_t radix_scratch_words(long n);
template <typename K>
int radix_sort_pairs(K* k0, K* k1, int* v0, int* v1, long n, int begin_bit, int end_bit, unsigned* scratch,
                     hipStream_t s);

// reference-order FP64 step (exact.inc)
void launch_x_fill_keys(unsigned* key, int* val, long m, hipStream_t s);
void launch_x_pairs(int shape, const PoolPtrs<double>& P, const XPtrs& X, const LaunchP& L, hipStream_t s);
void launch_x_segments(const unsigned* key, long m, int obits, int* seg0, int* seg1, hipStream_t s);
void launch_x_prec(int what, const PoolPtrs<double>& P, const XPtrs& X, const LaunchP& L, hipStream_t s);
void launch_x_scatter(int what, const PoolPtrs<double>& P, const XPtrs& X, const int* sval, const LaunchP& N,
                      hipStream_t s);
void launch_x_stress(int shape, int move, const PoolPtrs<double>& P, const XPtrs& X, const LaunchP& L,
                     hipStream_t s);
void launch_x_g2p(int shape, int move, const PoolPtrs<double>& P, const XPtrs& X, const LaunchP& L, hipStream_t s);
void launch_x_sgmp_stress(int move, const PoolPtrs<double>& P, const LaunchP& L, hipStream_t s);
// after a stress pass over Johnson-Cook instances: the increment every
// particle reads (last assigned before it, in original order) and the
// Johnson-Cook tails
void launch_x_jc(const PoolPtrs<double>& P, const LaunchP& L, const int* ilist, int ninst, hipStream_t s);
// Lagr_NodContact over the node list N, then each instance's tot_cont_for (ilist: the group's instances)
void launch_x_contact(const PoolPtrs<double>& P, const XPtrs& X, const LaunchP& N, const int* ilist, int ninst,
                      hipStream_t s);
// SmoothStressbyGrid (smooth.inc; exact mode: exact.inc)
template <typename T> void launch_smooth_save_x(const PoolPtrs<T>& P, const SmPtrs& S, const LaunchP& L, hipStream_t s);
template <typename T>
void launch_smooth_nodes(int shape, int what, const PoolPtrs<T>& P, const SmPtrs& S, const LaunchP& N, hipStream_t s);
template <typename T>
void launch_smooth(int shape, const PoolPtrs<T>& P, const SmPtrs& S, const LaunchP& L, const LaunchP& N,
                   hipStream_t s);
void launch_x_pairs_smooth(int shape, const PoolPtrs<double>& P, const XPtrs& X, const LaunchP& L, hipStream_t s);
void launch_x_smooth(int shape, const PoolPtrs<double>& P, const XPtrs& X, const SmPtrs& S, const int* sval,
                     const LaunchP& L, const LaunchP& N, hipStream_t s);
}  // namespace mpmk
