// engine.h -- device data layout of the MI355X engine (shared by the kernels
// and the host-side engine).
//
// Particles of all instances live in one structure-of-arrays pool.  Each
// instance owns a contiguous particle range padded to PGRAN slots and a
// contiguous node range padded to NGRAN nodes, so a thread block never
// straddles two instances: kernels read their instance from a per-block
// table with one scalar load.
#pragma once
#include <cstdint>

#define PGRAN 256
// node-box cache of the particle kernels: one entry per 64 particle slots
#define BOX_INTS 24
#define BOX_LOAD 1
#define BOX_STORE 2
#define NGRAN 256

// per-instance constants (device memory, read through the scalar cache)
struct InstC {
  int ncx, ncy, ncz, ncxy, ncell;
  int ngx, ngy, ngz, ngxy, nnode;
  int node_off;          // first node of the instance in the node pool
  int p_begin, p_end;    // particle range [p_begin, p_end) in the pool
  int shape, scheme;
  int flags;             // IF_* below
  int mat_off, nmat;     // material range in the material table
  int cell_off;          // sort-key offset of this instance's cells
  int max_steps;         // > 0: the instance stops itself after this many steps
  int cncx, cncxy, cncell;   // SGMP auxiliary cells: CenterNumCellx, ..xy, CenterNumCell
  int fixs[6];           // FixS of the deck (x0 xn y0 yn z0 zn)
  int ncomp;             // grid components (contact: nbco, else 1); component c's node k at c * nnode + k
  int nnode_all;         // ncomp * nnode: the instance's node range
  int contact;           // Contact_type: 1 Lagr_NodContact after IntegrateMomentum, 2 components only
  int normbody;          // contact normal: 0 both bodies, 1 component 1, 2 component 2
  double fricfa;         // contact friction coefficient
  double span0[3];       // SpanX(1), SpanY(1), SpanZ(1)
  double span1[3];       // adjusted upper ends
  double dcell, idcell, ijac, ijac4, cutoff;
  double qlt, damp, bq1, bq2, dtscale, endtime;
};

// bits per axis of the Hilbert sort order of an instance (cells per axis
// rounded up to a power of two), 0 when its grid is too large for a 30-bit
// key (then the brick order with 4^3-cell bricks is used)
__host__ __device__ inline int hilbert_bits(const InstC& I) {
  const int m = I.ncx > I.ncy ? (I.ncx > I.ncz ? I.ncx : I.ncz) : (I.ncy > I.ncz ? I.ncy : I.ncz);
  int b = 1;
  while ((1 << b) < m) b++;
  return b <= 10 ? b : 0;
}

enum InstFlags {
  IF_QUASI_LOAD = 1,
  IF_QUASI_DAMP = 2,
  IF_GRAVITY = 4,
  IF_UVFP = 8,
  IF_LOADS = 16,         // some particle carries a prescribed load FXp
  IF_ENDTIME = 32,       // instance stops itself when time > endtime
  IF_DRDA = 64,          // dynamic relaxation damping after UpdateDT (drda)
  IF_SMOOTH = 128,       // SmoothStressbyGrid at the end of every step (smoo)
  IF_TL = 256,           // total-Lagrangian step (TLmp): cells of the initial positions, mass of the first step
  IF_JC = 512            // Johnson-Cook models (john, jcf): the plastic strain increment carried across particles
};

// per-instance time state, advanced on the device (graph-capturable)
struct InstT {
  double time, dt;
  double ek, ei, mom[3], maxv;
  double ek_prev;
  int istep;
  int done;
  int nan_count;         // particles whose stress update produced NaN
  int negvol_count;      // particles with negative volume (warning in the reference)
  int damped;            // drda: this step's kinetic energy did not grow, velocities are zeroed
  int integ_pending;     // the last step's g2p integrated on the fly: PXg, FXg are not integrated yet
  int integ_half;        //   that step was the first (half step)
  double integ_dt;       //   and its DT
  double cont_force[3];  // tot_cont_for of the last contact pass
  double depeff;         // Johnson-Cook instances: the plastic strain increment the last yielding particle
                         // left (the reference's module variable), 0 before the first
};

// per-material constants, all derived exactly as the reference derives them
// per call (Constitution.f90:156-175) so the values are bit-identical
struct MatC {
  int type, eos;
  double density, young_eff, poisson, G2, K3, lamda, PlaMod, miu2, yield0;
  double Gmod, Kmod, q_fai, k_fai, q_psi, tenf;   // Drucker-Prager (tenf = min(ten_f, k/q))
  double c0;             // EOS 4 sound speed
  double csound;         // SetDT/UpdateDT sound speed max(c, D, Wavespd)
  double ce[7];          // EOS 1, 2: cEos(1..7) through a default REAL; EOS 3: A, B, R1, R2, omega
  double bjc, njc, cjc, epso, epf;   // (simplified) Johnson-Cook
  double mjc, roomt, melt, spec_heat;   // Johnson-Cook temperature exponent, room and melting temperature,
                                        // specific heat (john, jcf)
};

struct BodyC {
  int mat;               // global material index
  int com;               // grid component of the body, 0-based (0 without contact)
  double grav[3];
};

// particle fields kept in the bulk type T
enum TField {
  T_VX = 0, T_VY, T_VZ,
  T_BX, T_BY, T_BZ,      // G2P velocity (XX = Xp + b*DT applied by the stress pass)
  T_FX, T_FY, T_FZ,      // prescribed load FXp
  T_MASS, T_SM, T_SEQV,
  T_SD0, T_SD1, T_SD2, T_SD3, T_SD4, T_SD5,
  T_SIGY, T_EPEFF, T_Q, T_CP,
  T_COUNT
};

// particle fields always kept in FP64: position, volume and internal energy
// (accumulated every step), F (det F, ln J of the neo-Hookean model), the
// mass as read (the weakly compressible EOS takes the small difference
// m/V - rho0: an FP32 mass alone would offset it by 1e-5 of itself), and
// the state only some models change: temperature, damage, the HE lighting
// time and the failure flag (0/1)
enum DField {
  D_X = 0, D_Y, D_Z, D_VOL, D_IE, D_F0, D_F1, D_F2, D_F3, D_F4, D_F5, D_F6, D_F7, D_F8, D_MASS,
  D_TEMP, D_DMG, D_LT, D_FAIL,
  D_SPRE,                // smoothed pressure (SmoothStressbyGrid; output only)
  D_XO, D_YO, D_ZO,      // position at t = 0 (total-Lagrangian reference configuration)
  D_VOL0,                // initial volume
  D_COUNT
};

// first Piola-Kirchhoff stress of total-Lagrangian instances, PK(1..9) in
// the reference's order (xx yy zz yz xz xy zy zx yx): a field-major block of
// 9 doubles per particle slot, allocated with the first such instance
#define TL_PK_COUNT 9

// Johnson-Cook record fields (PoolPtrs::jc): the plastic strain increment,
// then the jcf tail's V and SM at the step's start and its energy increment
#define JC_REC_COUNT 4

// particle integer fields
enum IField { I_BODY = 0, I_ORIG, I_CELL, I_COUNT };

// Reference-order (exact FP64) scatter lists: the stencil entries of every
// particle slot (entry q = j * np + slot, j the position in the shape
// function's node list), their sort keys and, after the sort, the range of
// sorted entries of every pool node.
struct XPtrs {
  unsigned* key;             // m sort keys: the pool node, at (p_begin + original index) * 27 + j (unused: all ones)
  int* kval;                 // m sort values: the entry q of the key
  int* rnode;                // m: instance-local node (SGMP: auxiliary node) of entry q, -1 if none
  double* rw;                // 4 * m: N, dN/dx, dN/dy, dN/dz of entry q (entry-major, 32 B per entry)
  double* prec;              // 12 * np: per-slot particle records of the node scatters (k_x_prec)
  int* seg0;                 // nn: first sorted entry of every pool node
  int* seg1;                 // nn: one past its last
  double* ndir;              // 3 * nn: contact normals (mass gradients) of every pool node (contact only)
  double* cf;                // 3 * nn: contact force of every node of component 1 (contact only)
  long m;                    // entries: 27 * np
  int obits;                 // key bits below the node index
};

// SmoothStressbyGrid buffers, FP64 in every precision (allocated with the
// first instance that smooths): node sums of -SM V and V, the nodal pressure,
// the SGMP cell-centre sums (stored at the cell's index inside the node
// range), the mapped flags, and the step-start particle positions (the
// linear and SGMP shape functions of the general modes are those of the
// step's own ParticleInitial)
struct SmPtrs {
  double* gpre;   // nn: Gpre
  double* gv;     // nn: GV_all
  double* cgpre;  // nn: CGpre (SGMP)
  double* cgv;    // nn: CGV_all (SGMP)
  uint8_t* map;   // nn: mapped
  double* xs;     // 3 * np: positions at the start of the step (x, y, z field-major)
};

template <typename T>
struct PoolPtrs {
  T* t;          // T_COUNT * np
  double* d;     // D_COUNT * np
  int* i;        // I_COUNT * np
  long np;       // padded particle capacity (stride of every field)
  // grid (node pool, stride nn)
  double* mg;
  T* pg;         // 3 * nn  momentum PXg
  T* fg;         // 3 * nn  force FXg
  T* p2;         // 3 * nn  MUSL re-mapped momentum
  T* vg;         // 3 * nn  nodal velocity PXg/Mg (0 where Mg <= CutOff)
  T* ag;         // 3 * nn  nodal acceleration FXg/Mg (0 where Mg <= CutOff)
  uint8_t* fix;  // nn, bit0-2 fixity x/y/z; bit3-5 SGMP boundary node (bd_type /= 0) x/y/z
  T* aux;        // AUX_COUNT * nn  SGMP cell-centre (auxiliary node) fields, indexed by cell
  double* cmg;   // nn  SGMP auxiliary node mass
  long nn;
  const InstC* inst;
  InstT* ts;
  const MatC* mats;
  const BodyC* bodies;
  double* part;  // energy partials: 6 per 64 particle slots (one per wave of the particle kernels)
  int ilv;       // interleaved lane -> particle mapping in particle kernels
  int tile_cap;  // nodes a block's LDS tile may hold (<= TILE_NODES; smaller values exercise the fallbacks)
  int* boxes;    // node-box cache of the particle kernels: BOX_INTS ints per 64 particle slots
  double* pk;    // TL_PK_COUNT * np: PK of total-Lagrangian instances (null without any)
  // Johnson-Cook instances (null without any): every stress update's record,
  // at the pool index p_begin + original particle index: JC_* flags (jcf),
  // and jc = JC_REC_COUNT field-major doubles (the increment it assigned,
  // or after the scan the one it reads; V, SM and the energy increment of
  // the step's start for the jcf tail); flags are cleared by the tail
  int* jcf;
  double* jc;
  int box_mode;  // BOX_LOAD | BOX_STORE
};
