// engine.hip -- host side of the MI355X engine: device pool, instance
// tables, step sequences, graph capture, and the extern "C" entry points of
// include/mpm3d_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "mpm3d_hip.h"
#include "mpm3d_engine.h"

// ----------------------------------------------------------- error state

static thread_local std::string g_err;

struct MpmFail {
  int code;
};

[[noreturn]] static void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  throw MpmFail{code};
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) fail(MPM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));        \
  } while (0)

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return MPM_OK;
  } catch (const MpmFail& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host out of memory";
    return MPM_ERR_NOMEM;
  }
}

extern "C" const char* mpm_last_error(void) { return g_err.c_str(); }
extern "C" const char* mpm_version(void) { return "mpm3d-mi355x 0.1 (gfx950)"; }

// ------------------------------------------------------ range allocator

struct RangeAlloc {  // first-fit over [0, cap) in units
  long cap = 0;
  std::map<long, long> free_;  // start -> length
  void init(long c) {
    cap = c;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  long alloc(long n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        long s = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[s + n] = len - n;
        return s;
      }
    }
    return -1;
  }
  void release(long s, long n) {
    auto it = free_.emplace(s, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  long largest() const {
    long m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
};

template <typename T>
static T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) fail(MPM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return (T*)p;
}

static long round_up(long a, long b) { return (a + b - 1) / b * b; }

// ------------------------------------------------------------- engine

struct HostInst {
  bool live = false;
  InstC c{};
  long pg0 = 0, npg = 0;   // particle granules
  long ng0 = 0, nng = 0;   // node granules
  int mat0 = 0, nmat = 0, body0 = 0, nbody = 0;
  int n = 0;
  int group = -1;
  bool p_in_p2 = false;    // latest PXg lives in the re-map buffer
  int matmask = 0;         // material models used (bits 1 << type)
  bool xx_pending = false; // G2P done, Xp not yet moved (per-phase mode)
  bool integ_pending = false;   // fused USF step: g2p integrated on the fly, PXg/FXg not yet
};

struct Group {
  int shape = 0, scheme = 0;
  int matmask = 0;   // material models of the group's instances (bits 1 << type)
  bool drda = false; // some instance damps by dynamic relaxation
  bool smooth = false; // some instance smooths its stress by the grid (smoo)
  bool contact = false; // some instance has contact between grid components (Lagr_NodContact)
  bool jc = false;      // some instance has Johnson-Cook materials (john, jcf)
  std::vector<int> insts;
  std::vector<int> pgran, ngran;
  int* d_pgran = nullptr;
  int* d_ngran = nullptr;
  int* d_ilist = nullptr;
  size_t cap_p = 0, cap_n = 0, cap_i = 0;
};

struct mpm_engine {
  int prec = MPM_PREC_FP64;
  size_t tsz = 8;
  hipStream_t s = nullptr;
  int device = 0;
  long np = 0, nn = 0;            // capacities (slots)
  int max_inst = 0, max_mat = 0, max_body = 0;
  RangeAlloc pa, na, ma, ba;      // particle granules, node granules, materials, bodies
  std::vector<HostInst> inst;
  std::vector<int> free_ids;
  // device buffers
  void* t = nullptr;
  double* d = nullptr;
  int* ii = nullptr;
  double* mg = nullptr;
  void* pg = nullptr;
  void* fg = nullptr;
  void* p2 = nullptr;
  void* vg = nullptr;
  void* ag = nullptr;
  uint8_t* fix = nullptr;
  void* aux = nullptr;            // SGMP auxiliary-node fields (allocated with the first SGMP instance)
  SmPtrs sm{};                    // SmoothStressbyGrid buffers (allocated with the first instance that smooths)
  double* pk = nullptr;           // PK of total-Lagrangian instances (allocated with the first one)
  double* pk_alt = nullptr;       //   and its sort buffer
  int* jcf = nullptr;             // Johnson-Cook records (allocated with the first Johnson-Cook instance)
  double* jc = nullptr;
  double* cmg = nullptr;
  InstC* d_inst = nullptr;
  InstT* d_ts = nullptr;
  MatC* d_mats = nullptr;
  BodyC* d_bodies = nullptr;
  double* part = nullptr;
  int* d_ginst = nullptr;         // instance of every particle granule
  int* d_nginst = nullptr;        // instance of every node granule
  std::vector<int> h_ginst, h_nginst;
  int* d_seq = nullptr;           // 0, 1, 2, ...: granule range and instance lists without uploads
  // host-to-device staging: a ring of pinned buffers; every upload is a
  // copy ordered on the engine stream (so admissions and list changes never
  // wait for queued steps), and a buffer is reused only after its copies ran
  struct Pinned {
    char* h = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  std::vector<Pinned> ring;       // grows (up to 32) rather than wait for a busy buffer
  int ring_next = 0;
  char* d_up = nullptr;           // device staging of the raw particle rows of an admission
  size_t d_up_cap = 0;
  std::vector<std::array<int*, 3>> spare_lists;   // launch-list buffers of vanished groups (reused)
  InstT* h_snap = nullptr;        // pinned: time states as of the last status snapshot
  hipEvent_t snap_ev = nullptr;
  bool snap_pending = false;
  // launch lists
  std::vector<Group> groups;
  bool lists_dirty = true;
  int* d_ilist_all = nullptr;
  size_t cap_ilist = 0;
  int n_active = 0;
  // Launch shape.  The defaults are what the autotuner picks for the slope
  // decks (single slope and ensemble, FP32 and FP64) and the Taylor bar:
  // contexts and the Fortran drop-in, which never autotune, run with them.
  int bs_p = 256, bs_n = 256;     // block sizes (autotuned)
  int bs_s = 256;                 // block size of the particle-to-grid scatter launches
  int tile_cap = 512;             // nodes per LDS tile of the particle kernels (TILE_NODES)
  int* boxes = nullptr;           // node-box cache of the particle kernels (BOX_INTS per 64 slots)
  int box_mode = BOX_LOAD | BOX_STORE;
  int ilv = 0;                    // interleaved lane mapping in particle kernels
  int sort_every = 64;            // steps between re-sorts in mpm_step (0: never)
  int brick = 2;                  // sort order: 0 Hilbert, B > 0 bricks of B^3 cells
  long steps_done = 0;
  long sorts_done = 0;            // whole-pool re-sorts so far (admission sorts excluded)
  long since_sort = 1L << 40;     // steps since the last re-sort of the whole pool (none yet: sort first)
  // sort scratch (lazily allocated)
  unsigned* k_in = nullptr;
  unsigned* k_out = nullptr;
  int* v_in = nullptr;
  int* v_out = nullptr;
  unsigned* rs_scr = nullptr;     // radix-sort scratch for a sort over the whole pool
  void* t_alt = nullptr;
  double* d_alt = nullptr;
  int* i_alt = nullptr;
  int* d_seg = nullptr;           // segment begin/end slots, 2 x max_inst
  std::vector<int> h_seg;         // host copy of d_seg (kept alive for the async upload)
  // graph: a captured step sequence bakes in the particle buffers'
  // addresses, and the re-sort swaps the buffers with their sort twins, so
  // up to two graphs are kept, one per buffer set (gkey: the T buffer)
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  const void* gkey[2] = {nullptr, nullptr};
  int gnext = 0;                  // slot the next capture replaces when both hold a graph
  int graph_steps = 0;
  bool graph_dirty = true;
  bool graph_plain_once = false;  // lazy replay: the changed lists already ran once without a graph
  // reference-order FP64 mode (MPM_PREC_FP64_EXACT): stencil entries, their
  // sort double buffers and the radix scratch, allocated before the first
  // exact step
  bool exact = false;
  unsigned* x_key[2] = {nullptr, nullptr};
  int* x_val[2] = {nullptr, nullptr};
  int* x_rnode = nullptr;
  double* x_rw = nullptr;
  double* x_prec = nullptr;   // per-slot particle records of the node scatters
  int* x_seg0 = nullptr;
  int* x_seg1 = nullptr;
  double* x_ndir = nullptr;       // contact normals (allocated with the first contact instance)
  double* x_cf = nullptr;         // contact force per node
  unsigned* x_scr = nullptr;
  long x_m = 0;
  int x_obits = 0, x_nbits = 0;
  int x_sorted = 0;               // buffer holding the step's sorted entries
  // subset downloads (a few particles per step for the drop-in's curves)
  int* d_sub_ids = nullptr;
  double* d_sub_rows = nullptr;
  int sub_cap = 0;
  // profiling mode: HIP events around every launch of mpm_step / engine steps
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_kid;
  double kt_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long kt_n[8] = {0, 0, 0, 0, 0, 0, 0, 0};

  template <typename T>
  PoolPtrs<T> ptrs() const {
    PoolPtrs<T> P;
    P.t = (T*)t;
    P.d = d;
    P.i = ii;
    P.np = np;
    P.mg = mg;
    P.pg = (T*)pg;
    P.fg = (T*)fg;
    P.p2 = (T*)p2;
    P.vg = (T*)vg;
    P.ag = (T*)ag;
    P.fix = fix;
    P.aux = (T*)aux;
    P.cmg = cmg;
    P.nn = nn;
    P.inst = d_inst;
    P.ts = d_ts;
    P.mats = d_mats;
    P.bodies = d_bodies;
    P.part = part;
    P.ilv = ilv;
    P.tile_cap = tile_cap;
    P.boxes = boxes;
    P.pk = pk;
    P.jcf = jcf;
    P.jc = jc;
    P.box_mode = box_mode;
    return P;
  }
};

static void free_engine(mpm_engine* e) {
  if (!e) return;
  if (e->s) (void)hipStreamSynchronize(e->s);   // queued work may still use any buffer below
  for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
  for (int k = 0; k < 2; k++) {
    if (e->gexec[k]) (void)hipGraphExecDestroy(e->gexec[k]);
    if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
  }
  void* bufs[] = {e->t, e->d, e->ii, e->mg, e->pg, e->fg, e->p2, e->vg, e->ag, e->fix, e->d_inst, e->d_ts, e->d_mats,
                  e->d_bodies, e->part, e->d_ginst, e->d_nginst, e->d_ilist_all, e->k_in, e->k_out, e->v_in,
                  e->v_out, e->rs_scr, e->t_alt, e->d_alt, e->i_alt, e->d_seg, e->aux, e->cmg, e->boxes,
                  e->x_key[0], e->x_key[1], e->x_val[0], e->x_val[1], e->x_rnode, e->x_rw, e->x_prec, e->x_seg0, e->x_seg1,
                  e->x_scr, e->d_sub_ids, e->d_sub_rows, e->sm.gpre, e->sm.gv, e->sm.cgpre, e->sm.cgv, e->sm.map,
                  e->sm.xs, e->x_ndir, e->x_cf, e->pk, e->pk_alt, e->jcf, e->jc};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto& g : e->groups) {
    if (g.d_pgran) (void)hipFree(g.d_pgran);
    if (g.d_ngran) (void)hipFree(g.d_ngran);
    if (g.d_ilist) (void)hipFree(g.d_ilist);
  }
  for (auto& l : e->spare_lists)
    for (int* b : l)
      if (b) (void)hipFree(b);
  for (auto& r : e->ring) {
    if (r.h) (void)hipHostFree(r.h);
    if (r.ev) (void)hipEventDestroy(r.ev);
  }
  if (e->d_up) (void)hipFree(e->d_up);
  if (e->h_snap) (void)hipHostFree(e->h_snap);
  if (e->snap_ev) (void)hipEventDestroy(e->snap_ev);
  if (e->d_seq) (void)hipFree(e->d_seq);
  if (e->s) (void)hipStreamDestroy(e->s);
  delete e;
}

// bytes of device memory per particle slot / node for a precision
static size_t bytes_per_particle(int prec) {
  size_t tsz = prec == MPM_PREC_FP32 ? 4 : 8;
  return T_COUNT * tsz + D_COUNT * 8 + I_COUNT * 4;
}
static size_t bytes_per_node(int prec) {
  size_t tsz = prec == MPM_PREC_FP32 ? 4 : 8;
  return 8 + 15 * tsz + 1;
}

static mpm_engine* new_engine(long np_slots, long nn_slots, int max_inst, int max_mat, int max_body, int prec) {
  std::unique_ptr<mpm_engine, void (*)(mpm_engine*)> e(new mpm_engine, free_engine);
  e->exact = prec == MPM_PREC_FP64_EXACT;
  if (e->exact) prec = MPM_PREC_FP64;
  e->prec = prec;
  e->tsz = prec == MPM_PREC_FP32 ? 4 : 8;
  HIPCHK(hipGetDevice(&e->device));
  HIPCHK(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
  e->np = round_up(std::max(np_slots, (long)PGRAN), PGRAN);
  e->nn = round_up(std::max(nn_slots, (long)NGRAN), NGRAN);
  e->max_inst = max_inst;
  e->max_mat = max_mat;
  e->max_body = max_body;
  e->pa.init(e->np / PGRAN);
  e->na.init(e->nn / NGRAN);
  e->ma.init(max_mat);
  e->ba.init(max_body);
  e->t = dalloc<char>((size_t)T_COUNT * e->np * e->tsz);
  e->d = dalloc<double>((size_t)D_COUNT * e->np);
  e->ii = dalloc<int>((size_t)I_COUNT * e->np);
  e->mg = dalloc<double>(e->nn);
  e->pg = dalloc<char>(3 * e->nn * e->tsz);
  e->fg = dalloc<char>(3 * e->nn * e->tsz);
  e->p2 = dalloc<char>(3 * e->nn * e->tsz);
  e->vg = dalloc<char>(3 * e->nn * e->tsz);
  e->ag = dalloc<char>(3 * e->nn * e->tsz);
  e->fix = dalloc<uint8_t>(e->nn);
  e->d_inst = dalloc<InstC>(max_inst);
  e->d_ts = dalloc<InstT>(max_inst);
  e->d_mats = dalloc<MatC>(max_mat);
  e->d_bodies = dalloc<BodyC>(max_body);
  e->part = dalloc<double>((size_t)(e->np / PGRAN) * (PGRAN / 64) * 6);
  e->boxes = dalloc<int>((size_t)(e->np / 64) * BOX_INTS);
  e->d_ginst = dalloc<int>(e->np / PGRAN);
  e->d_nginst = dalloc<int>(e->nn / NGRAN);
  e->h_ginst.assign(e->np / PGRAN, 0);
  e->h_nginst.assign(e->nn / NGRAN, 0);
  {
    const long nseq = std::max({e->np / PGRAN, e->nn / NGRAN, (long)max_inst});
    std::vector<int> seq(nseq);
    for (long k = 0; k < nseq; k++) seq[k] = (int)k;
    e->d_seq = dalloc<int>(nseq);
    HIPCHK(hipMemcpy(e->d_seq, seq.data(), nseq * sizeof(int), hipMemcpyHostToDevice));
  }
  e->inst.resize(max_inst);
  for (int k = max_inst - 1; k >= 0; k--) e->free_ids.push_back(k);
  HIPCHK(hipMemsetAsync(e->t, 0, (size_t)T_COUNT * e->np * e->tsz, e->s));
  HIPCHK(hipMemsetAsync(e->d, 0, (size_t)D_COUNT * e->np * 8, e->s));
  HIPCHK(hipMemsetAsync(e->mg, 0, e->nn * 8, e->s));
  HIPCHK(hipMemsetAsync(e->pg, 0, 3 * e->nn * e->tsz, e->s));
  HIPCHK(hipMemsetAsync(e->fg, 0, 3 * e->nn * e->tsz, e->s));
  HIPCHK(hipMemsetAsync(e->p2, 0, 3 * e->nn * e->tsz, e->s));
  HIPCHK(hipMemsetAsync(e->vg, 0, 3 * e->nn * e->tsz, e->s));
  HIPCHK(hipMemsetAsync(e->ag, 0, 3 * e->nn * e->tsz, e->s));
  HIPCHK(hipMemsetAsync(e->boxes, 0, (size_t)(e->np / 64) * BOX_INTS * sizeof(int), e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  return e.release();
}

// material constants exactly as the reference derives them per call
static MatC make_mat(const mpm_material& m) {
  MatC c{};
  c.type = m.type;
  c.eos = m.eos_type;
  c.density = m.density;
  const double DMG = 0.0;
  double young = m.young * (1.0 - DMG + 0x1p-52);   // Constitution.f90:171
  double poisson = m.poisson;
  c.young_eff = young;
  c.poisson = poisson;
  c.G2 = young / (1 + poisson);
  c.K3 = young / (1 - 2 * poisson);
  c.lamda = (young * poisson) / ((1 + poisson) * (1 - 2 * poisson));
  c.PlaMod = young * m.tangmod / (young - m.tangmod);
  c.miu2 = 2 * m.miu;
  c.yield0 = m.yield0;
  c.Gmod = young / (2 * (1 + poisson));            // Constitution.f90:1232-1233
  c.Kmod = young / (3 * (1 - 2 * poisson));
  c.q_fai = m.q_fai;
  c.k_fai = m.k_fai;
  c.q_psi = m.q_psi;
  if (m.q_fai == 0.0) {
    c.tenf = 0.0;
  } else {
    double tmax = m.k_fai / m.q_fai;
    c.tenf = m.ten_f < tmax ? m.ten_f : tmax;
  }
  c.c0 = m.ceos[0];
  for (int k = 0; k < 7; k++)   // eos1/eos2 hold their constants in default REALs
    c.ce[k] = (m.eos_type == 1 || m.eos_type == 2) ? (double)(float)m.ceos[k] : m.ceos[k];
  c.bjc = m.b_jc;
  c.njc = m.n_jc;
  c.cjc = m.c_jc;
  c.epso = m.epso;
  c.epf = m.epf;
  c.mjc = m.m_jc;
  c.roomt = m.roomt;
  c.melt = m.melt;
  c.spec_heat = m.spec_heat;
  // SetDT / UpdateDT sound speed (DataIn.f90:1285-1290, update_step.f90:1861-1866)
  double E = m.young, nu = m.poisson, ro = m.density;
  double cp = std::sqrt(E * (1 - nu) / (1 + nu) / (1 - 2 * nu) / ro);
  cp = std::fmax(cp, m.det_speed);
  cp = std::fmax(cp, m.wavespd);
  c.csound = cp;
  return c;
}

static void check_problem(const mpm_problem* p) {
  if (!p || p->nb_particle <= 0 || p->nb_mat <= 0 || p->nb_body <= 0 || !p->pstate || !p->pbody || !p->mats ||
      !p->bodies)
    fail(MPM_ERR_ARG, "incomplete problem description");
  if (p->shape < 0 || p->shape > 3 || p->scheme < 0 || p->scheme > 2) fail(MPM_ERR_ARG, "bad shape/scheme");
  for (int m = 0; m < p->nb_mat; m++) {
    const int t = p->mats[m].type, et = p->mats[m].eos_type;
    if (t < MPM_MAT_ELASTIC || t > MPM_MAT_NEOH) fail(MPM_ERR_ARG, "material type %d", t);
    // models that call seleos need an EOS; JWL only with HE burn (its burn fraction)
    const bool needs_eos = t == MPM_MAT_SJC || t == MPM_MAT_SJCF || t == MPM_MAT_JCF || t == MPM_MAT_HIEX ||
                           t == MPM_MAT_NULL;
    if (needs_eos && (et < 1 || et > 4)) fail(MPM_ERR_ARG, "material %d needs an equation of state", m + 1);
    if (et == 3 && t != MPM_MAT_HIEX) fail(MPM_ERR_UNSUPPORTED, "material %d: JWL EOS without HE burn", m + 1);
  }
  for (int b = 0; b < p->nb_body; b++)
    if (p->bodies[b].mat < 1 || p->bodies[b].mat > p->nb_mat) fail(MPM_ERR_ARG, "body %d has no material", b);
  if (p->ncx <= 0 || p->ncy <= 0 || p->ncz <= 0 || !(p->dcell > 0)) fail(MPM_ERR_ARG, "bad grid");
  if (p->contact) {
    if (p->contact != 1 && p->contact != 2) fail(MPM_ERR_ARG, "contact type %d", p->contact);
    if (p->nb_component < 2) fail(MPM_ERR_ARG, "contact needs at least 2 grid components");
    if (p->normbody < 0 || p->normbody > 2) fail(MPM_ERR_ARG, "contact normal %d", p->normbody);
    for (int b = 0; b < p->nb_body; b++)
      if (p->bodies[b].com < 1 || p->bodies[b].com > p->nb_component)
        fail(MPM_ERR_ARG, "body %d: component %d outside 1..%d", b + 1, p->bodies[b].com, p->nb_component);
    if (p->shape == MPM_SHAPE_SGMP) fail(MPM_ERR_UNSUPPORTED, "contact with SGMP");
    if (p->smooth_stress && p->nb_component > 8)
      fail(MPM_ERR_UNSUPPORTED, "stress smoothing (smoo) with contact over more than 8 grid components");
  }
  if (p->tlmpm) {   // TLConstitution has the neo-Hookean model only (Constitution.f90:487-497)
    for (int m = 0; m < p->nb_mat; m++)
      if (p->mats[m].type != MPM_MAT_NEOH)
        fail(MPM_ERR_UNSUPPORTED, "total-Lagrangian MPM with material type %d (neo-Hookean only)", p->mats[m].type);
  }
  long ncell = (long)p->ncx * p->ncy * p->ncz;
  long nnode = (long)(p->ncx + 1) * (p->ncy + 1) * (p->ncz + 1);
  if (nnode > 0x7fffffffL || ncell > 0x7fffffffL) fail(MPM_ERR_ARG, "grid too large");
}

// A pinned staging buffer of at least `bytes` that no queued copy still
// reads (waits for the copies of the ring slot's previous use).
static char* stage_get(mpm_engine* e, size_t bytes, int& slot) {
  const int nr = (int)e->ring.size();
  slot = -1;
  for (int k = 0; k < nr && slot < 0; k++) {   // a buffer whose copies have run
    const int q = (e->ring_next + k) % nr;
    mpm_engine::Pinned& r = e->ring[q];
    if (!r.pending) {
      slot = q;
    } else {
      const hipError_t st = hipEventQuery(r.ev);
      if (st == hipSuccess) {
        r.pending = false;
        slot = q;
      } else if (st != hipErrorNotReady) {
        HIPCHK(st);
      }
    }
  }
  if (slot < 0 && nr < 32) {
    e->ring.emplace_back();
    slot = nr;
  }
  if (slot < 0) {   // all 32 busy: wait for the next in turn
    slot = e->ring_next % nr;
    HIPCHK(hipEventSynchronize(e->ring[slot].ev));
    e->ring[slot].pending = false;
  }
  e->ring_next = (slot + 1) % (int)e->ring.size();
  mpm_engine::Pinned& r = e->ring[slot];
  if (!r.ev) HIPCHK(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
  if (r.cap < bytes) {
    if (r.h) HIPCHK(hipHostFree(r.h));
    r.h = nullptr;
    r.cap = 0;
    const size_t cap = std::max<size_t>(bytes, (size_t)1 << 20);
    HIPCHK(hipHostMalloc((void**)&r.h, cap, hipHostMallocDefault));
    r.cap = cap;
  }
  return r.h;
}

// the copies out of a ring slot are enqueued: mark it busy until they ran
static void stage_done(mpm_engine* e, int slot) {
  HIPCHK(hipEventRecord(e->ring[slot].ev, e->s));
  e->ring[slot].pending = true;
}

static size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

template <typename T>
static void upload_particles_T(mpm_engine* e, const HostInst& h, const double* ps, int n, bool first) {
  HIPCHK(hipStreamSynchronize(e->s));   // the copies below are not ordered on the engine stream
  const long np = e->np, pad = h.npg * PGRAN, b0 = h.pg0 * PGRAN;
  std::vector<T> tb((size_t)T_COUNT * pad, (T)0);
  std::vector<double> db((size_t)D_COUNT * pad, 0.0);
  auto P = [&](int f, int i) { return ps[(size_t)f * n + i]; };
  for (int i = 0; i < n; i++) {
    tb[(size_t)T_VX * pad + i] = (T)P(PF_VX, i);
    tb[(size_t)T_VY * pad + i] = (T)P(PF_VY, i);
    tb[(size_t)T_VZ * pad + i] = (T)P(PF_VZ, i);
    tb[(size_t)T_FX * pad + i] = (T)P(PF_FX, i);
    tb[(size_t)T_FY * pad + i] = (T)P(PF_FY, i);
    tb[(size_t)T_FZ * pad + i] = (T)P(PF_FZ, i);
    tb[(size_t)T_MASS * pad + i] = (T)P(PF_MASS, i);
    db[(size_t)D_VOL * pad + i] = P(PF_VOL, i);
    tb[(size_t)T_SM * pad + i] = (T)P(PF_SM, i);
    tb[(size_t)T_SEQV * pad + i] = (T)P(PF_SEQV, i);
    tb[(size_t)T_SD0 * pad + i] = (T)P(PF_SDXX, i);
    tb[(size_t)T_SD1 * pad + i] = (T)P(PF_SDYY, i);
    tb[(size_t)T_SD2 * pad + i] = (T)P(PF_SDZZ, i);
    tb[(size_t)T_SD3 * pad + i] = (T)P(PF_SDYZ, i);
    tb[(size_t)T_SD4 * pad + i] = (T)P(PF_SDXZ, i);
    tb[(size_t)T_SD5 * pad + i] = (T)P(PF_SDXY, i);
    tb[(size_t)T_SIGY * pad + i] = (T)P(PF_SIGY, i);
    tb[(size_t)T_EPEFF * pad + i] = (T)P(PF_EPEFF, i);
    db[(size_t)D_IE * pad + i] = P(PF_IE, i);
    tb[(size_t)T_Q * pad + i] = (T)P(PF_Q, i);
    tb[(size_t)T_CP * pad + i] = (T)P(PF_CP, i);
    db[(size_t)D_X * pad + i] = P(PF_X, i);
    db[(size_t)D_Y * pad + i] = P(PF_Y, i);
    db[(size_t)D_Z * pad + i] = P(PF_Z, i);
    for (int k = 0; k < 9; k++) db[(size_t)(D_F0 + k) * pad + i] = P(PF_FP + k, i);
    db[(size_t)D_MASS * pad + i] = P(PF_MASS, i);
    db[(size_t)D_TEMP * pad + i] = P(PF_TEMP, i);
    db[(size_t)D_DMG * pad + i] = P(PF_DMG, i);
    db[(size_t)D_LT * pad + i] = P(PF_LT, i);
    db[(size_t)D_FAIL * pad + i] = P(PF_FAIL, i);
    db[(size_t)D_SPRE * pad + i] = P(PF_SPRE, i);
    for (int k = 0; k < 3; k++) db[(size_t)(D_XO + k) * pad + i] = P(PF_XO + k, i);
    db[(size_t)D_VOL0 * pad + i] = P(PF_VOL0, i);
  }
  if (!first) {
    // upload keeps the device permutation: scatter host rows to their slots
    std::vector<int> orig(pad);
    HIPCHK(hipMemcpy(orig.data(), e->ii + (size_t)I_ORIG * np + b0, pad * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<T> tb2(tb.size());
    std::vector<double> db2(db.size());
    for (long j = 0; j < pad; j++) {
      int o = (j < n) ? orig[j] : (int)j;
      for (int f = 0; f < T_COUNT; f++) tb2[(size_t)f * pad + j] = tb[(size_t)f * pad + o];
      for (int f = 0; f < D_COUNT; f++) db2[(size_t)f * pad + j] = db[(size_t)f * pad + o];
    }
    tb.swap(tb2);
    db.swap(db2);
  }
  for (int f = 0; f < T_COUNT; f++)
    HIPCHK(hipMemcpy((char*)e->t + ((size_t)f * np + b0) * sizeof(T), &tb[(size_t)f * pad], pad * sizeof(T),
                     hipMemcpyHostToDevice));
  for (int f = 0; f < D_COUNT; f++)
    HIPCHK(hipMemcpy(e->d + (size_t)f * np + b0, &db[(size_t)f * pad], pad * sizeof(double), hipMemcpyHostToDevice));
  if (e->pk && (h.c.flags & IF_TL)) {   // PK of a total-Lagrangian instance, in the slots' order
    std::vector<int> orig(pad);
    HIPCHK(hipMemcpy(orig.data(), e->ii + (size_t)I_ORIG * np + b0, pad * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double> kb((size_t)TL_PK_COUNT * pad, 0.0);
    for (long j = 0; j < pad; j++) {
      const long o = first ? j : (j < n ? orig[j] : -1);
      if (o < 0 || o >= n) continue;
      for (int k = 0; k < TL_PK_COUNT; k++) kb[(size_t)k * pad + j] = P(PF_PK + k, (int)o);
    }
    for (int k = 0; k < TL_PK_COUNT; k++)
      HIPCHK(hipMemcpy(e->pk + (size_t)k * np + b0, &kb[(size_t)k * pad], pad * sizeof(double),
                       hipMemcpyHostToDevice));
  }
}

static void upload_particles(mpm_engine* e, const HostInst& h, const double* ps, int n, bool first) {
  if (e->prec == MPM_PREC_FP32) upload_particles_T<float>(e, h, ps, n, first);
  else upload_particles_T<double>(e, h, ps, n, first);
}

// add one instance to the pool; returns instance id
// the launch lists or launch shape changed: the captured step graph is stale
static void mark_graph_dirty(mpm_engine* e) {
  e->graph_dirty = true;
  e->graph_plain_once = false;
}

static void ensure_aux(mpm_engine* e) {
  if (e->aux) return;
  e->aux = dalloc<char>((size_t)AUX_COUNT * e->nn * e->tsz);
  e->cmg = dalloc<double>(e->nn);
  mark_graph_dirty(e);
}

static void ensure_smooth(mpm_engine* e) {
  if (e->sm.gpre) return;
  e->sm.gpre = dalloc<double>(e->nn);
  e->sm.gv = dalloc<double>(e->nn);
  e->sm.cgpre = dalloc<double>(e->nn);
  e->sm.cgv = dalloc<double>(e->nn);
  e->sm.map = dalloc<uint8_t>(e->nn);
  e->sm.xs = dalloc<double>(3 * e->np);
  mark_graph_dirty(e);
}

static void ensure_tl(mpm_engine* e) {
  if (e->pk) return;
  e->pk = dalloc<double>((size_t)TL_PK_COUNT * e->np);
  if (e->k_in) e->pk_alt = dalloc<double>((size_t)TL_PK_COUNT * e->np);   // sort buffers exist already
  mark_graph_dirty(e);
}

// Johnson-Cook records: flags cleared once (the tails clear what the stress
// passes set), the doubles are written before they are read
static void ensure_jc(mpm_engine* e) {
  if (e->jcf) return;
  e->jcf = dalloc<int>(e->np);
  HIPCHK(hipMemsetAsync(e->jcf, 0, sizeof(int) * e->np, e->s));
  e->jc = dalloc<double>((size_t)JC_REC_COUNT * e->np);
  mark_graph_dirty(e);
}

static void ensure_contact(mpm_engine* e) {
  if (e->x_ndir) return;
  e->x_ndir = dalloc<double>(3 * e->nn);
  e->x_cf = dalloc<double>(3 * e->nn);
  mark_graph_dirty(e);
}

static int add_instance(mpm_engine* e, const mpm_problem* p, int flags_extra) {
  check_problem(p);
  if (p->contact && !e->exact)
    fail(MPM_ERR_UNSUPPORTED, "contact between grid components runs in the reference-order FP64 mode "
                              "(MPM_PREC_FP64_EXACT)");
  if (p->tlmpm && !e->exact)
    fail(MPM_ERR_UNSUPPORTED, "total-Lagrangian MPM runs in the reference-order FP64 mode (MPM_PREC_FP64_EXACT)");
  bool jc = false;
  for (int m = 0; m < p->nb_mat; m++) jc = jc || p->mats[m].type == MPM_MAT_JOHN || p->mats[m].type == MPM_MAT_JCF;
  if (jc && !e->exact)
    fail(MPM_ERR_UNSUPPORTED, "the Johnson-Cook models john and jcf run in the reference-order FP64 mode "
                              "(MPM_PREC_FP64_EXACT): each particle that does not yield reads the plastic strain "
                              "increment of the last particle before it that did");
  if (p->tlmpm) ensure_tl(e);   // before the admission kernel, which fills it
  if (jc) ensure_jc(e);
  if (e->free_ids.empty()) fail(MPM_ERR_NOMEM, "no free instance slot");
  const int n = p->nb_particle;
  const long npg = (n + PGRAN - 1) / PGRAN;
  const long nnode = (long)(p->ncx + 1) * (p->ncy + 1) * (p->ncz + 1);
  const int ncomp = p->contact ? p->nb_component : 1;
  if (nnode * ncomp > 0x7fffffffL) fail(MPM_ERR_ARG, "grid components too large");
  const long nng = (nnode * ncomp + NGRAN - 1) / NGRAN;
  long pg0 = e->pa.alloc(npg);
  if (pg0 < 0) fail(MPM_ERR_NOMEM, "particle pool full (%d particles requested)", n);
  long ng0 = e->na.alloc(nng);
  if (ng0 < 0) {
    e->pa.release(pg0, npg);
    fail(MPM_ERR_NOMEM, "node pool full (%ld nodes requested)", nnode);
  }
  long m0 = e->ma.alloc(p->nb_mat);
  long b0 = m0 >= 0 ? e->ba.alloc(p->nb_body) : -1;
  if (m0 < 0 || b0 < 0) {
    if (m0 >= 0) e->ma.release(m0, p->nb_mat);
    e->pa.release(pg0, npg);
    e->na.release(ng0, nng);
    fail(MPM_ERR_NOMEM, "material/body table full");
  }
  const int id = e->free_ids.back();
  e->free_ids.pop_back();
  HostInst& h = e->inst[id];
  h = HostInst{};
  h.live = true;
  h.pg0 = pg0;
  h.npg = npg;
  h.ng0 = ng0;
  h.nng = nng;
  h.mat0 = (int)m0;
  h.nmat = p->nb_mat;
  h.body0 = (int)b0;
  h.nbody = p->nb_body;
  h.n = n;
  for (int m = 0; m < p->nb_mat; m++) h.matmask |= 1 << p->mats[m].type;
  InstC& c = h.c;
  c.ncx = p->ncx; c.ncy = p->ncy; c.ncz = p->ncz;
  c.ncxy = c.ncx * c.ncy;
  c.ncell = c.ncxy * c.ncz;
  c.ngx = c.ncx + 1; c.ngy = c.ncy + 1; c.ngz = c.ncz + 1;
  c.ngxy = c.ngx * c.ngy;
  c.nnode = (int)nnode;
  c.ncomp = ncomp;
  c.nnode_all = (int)(nnode * ncomp);
  c.contact = p->contact;
  c.normbody = p->normbody;
  c.fricfa = p->fricfa;
  c.node_off = (int)(ng0 * NGRAN);
  c.p_begin = (int)(pg0 * PGRAN);
  c.p_end = c.p_begin + n;
  c.shape = p->shape;
  c.scheme = p->scheme;
  c.flags = (p->quasi_load ? IF_QUASI_LOAD : 0) | (p->quasi_damp ? IF_QUASI_DAMP : 0) |
            (p->gravity ? IF_GRAVITY : 0) | (p->update_v_by_fp ? IF_UVFP : 0) | (p->dr_damping ? IF_DRDA : 0) |
            (p->smooth_stress ? IF_SMOOTH : 0) | (p->tlmpm ? IF_TL : 0) | (jc ? IF_JC : 0) | flags_extra;
  bool loads = false;
  for (int i = 0; i < n && !loads; i++)
    for (int d = 0; d < 3; d++)
      if (p->pstate[(size_t)(PF_FX + d) * n + i] != 0.0) loads = true;
  if (loads) c.flags |= IF_LOADS;
  c.mat_off = (int)m0;
  c.nmat = p->nb_mat;
  c.cell_off = 0;
  c.max_steps = 0;
  // SGMP auxiliary cells (Grid.f90:254-268)
  c.cncx = std::max(c.ncx - 1, 1);
  c.cncxy = c.cncx * std::max(c.ncy - 1, 1);
  c.cncell = c.cncxy * std::max(c.ncz - 1, 1);
  for (int k = 0; k < 6; k++) c.fixs[k] = p->fix[k];
  if (p->shape == MPM_SHAPE_SGMP) ensure_aux(e);
  if (p->smooth_stress) ensure_smooth(e);
  if (p->contact == 1) ensure_contact(e);
  for (int k = 0; k < 3; k++) {
    c.span0[k] = p->span[2 * k];
    c.span1[k] = p->span[2 * k + 1];
  }
  c.dcell = p->dcell;
  c.idcell = 1.0 / p->dcell;      // Grid.f90:416
  c.ijac = 2.0 / p->dcell;        // :414
  c.ijac4 = c.ijac * 0.125;       // :415
  c.cutoff = p->cutoff;
  c.qlt = p->quasi_load_time;
  c.damp = p->damp_factor;
  c.bq1 = p->bq1;
  c.bq2 = p->bq2;
  c.dtscale = p->dtscale;
  c.endtime = p->endtime;
  // materials and bodies
  std::vector<MatC> mats(p->nb_mat);
  for (int m = 0; m < p->nb_mat; m++) mats[m] = make_mat(p->mats[m]);
  std::vector<BodyC> bodies(p->nb_body);
  for (int b = 0; b < p->nb_body; b++) {
    bodies[b].mat = (int)m0 + p->bodies[b].mat - 1;
    bodies[b].com = p->contact ? p->bodies[b].com - 1 : 0;
    for (int k = 0; k < 3; k++) bodies[b].grav[k] = p->bodies[b].grav[k];
  }
  // time state: calcEnergy of the initial state is done on the device later
  InstT ts{};
  ts.dt = p->dt;
  for (long g = 0; g < npg; g++) e->h_ginst[pg0 + g] = id;
  for (long g = 0; g < nng; g++) e->h_nginst[ng0 + g] = id;
  // Everything goes through one pinned staging buffer as copies ordered on
  // the engine stream, behind any queued steps (which never touch the new
  // ranges): constants, tables, granule maps, and the raw particle rows and
  // body ids, which a kernel converts into the pool (k_admit) while another
  // derives the node fixities (k_fixity).
  const size_t o_m = align16(sizeof(InstC));
  const size_t o_b = align16(o_m + sizeof(MatC) * p->nb_mat);
  const size_t o_t = align16(o_b + sizeof(BodyC) * p->nb_body);
  const size_t o_gi = align16(o_t + sizeof(InstT));
  const size_t o_ng = align16(o_gi + sizeof(int) * npg);
  const size_t o_pb = align16(o_ng + sizeof(int) * nng);
  const size_t rows_pb = align16(sizeof(int) * (size_t)n);
  const size_t rows = rows_pb + sizeof(double) * PF_COUNT * (size_t)n;
  int slot;
  char* st = stage_get(e, o_pb + rows, slot);
  memcpy(st, &c, sizeof(InstC));
  memcpy(st + o_m, mats.data(), sizeof(MatC) * p->nb_mat);
  memcpy(st + o_b, bodies.data(), sizeof(BodyC) * p->nb_body);
  memcpy(st + o_t, &ts, sizeof(InstT));
  memcpy(st + o_gi, &e->h_ginst[pg0], sizeof(int) * npg);
  memcpy(st + o_ng, &e->h_nginst[ng0], sizeof(int) * nng);
  if (n) {
    memcpy(st + o_pb, p->pbody, sizeof(int) * (size_t)n);
    memcpy(st + o_pb + rows_pb, p->pstate, sizeof(double) * PF_COUNT * (size_t)n);
  }
  if (rows > e->d_up_cap) {   // grows with the largest admission so far
    HIPCHK(hipStreamSynchronize(e->s));
    if (e->d_up) HIPCHK(hipFree(e->d_up));
    e->d_up = nullptr;
    e->d_up_cap = 0;
    e->d_up = dalloc<char>(rows);
    e->d_up_cap = rows;
  }
  const hipMemcpyKind H2D = hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(e->d_inst + id, st, sizeof(InstC), H2D, e->s));
  HIPCHK(hipMemcpyAsync(e->d_mats + m0, st + o_m, sizeof(MatC) * p->nb_mat, H2D, e->s));
  HIPCHK(hipMemcpyAsync(e->d_bodies + b0, st + o_b, sizeof(BodyC) * p->nb_body, H2D, e->s));
  HIPCHK(hipMemcpyAsync(e->d_ts + id, st + o_t, sizeof(InstT), H2D, e->s));
  HIPCHK(hipMemcpyAsync(e->d_ginst + pg0, st + o_gi, sizeof(int) * npg, H2D, e->s));
  HIPCHK(hipMemcpyAsync(e->d_nginst + ng0, st + o_ng, sizeof(int) * nng, H2D, e->s));
  if (n) HIPCHK(hipMemcpyAsync(e->d_up, st + o_pb, rows, H2D, e->s));
  stage_done(e, slot);
  const int* d_pbody = (const int*)e->d_up;
  const double* d_rows = (const double*)(e->d_up + rows_pb);
  if (e->prec == MPM_PREC_FP32)
    mpmk::launch_admit<float>(e->ptrs<float>(), d_rows, d_pbody, n, pg0 * PGRAN, npg * PGRAN, (int)b0, e->s);
  else
    mpmk::launch_admit<double>(e->ptrs<double>(), d_rows, d_pbody, n, pg0 * PGRAN, npg * PGRAN, (int)b0, e->s);
  mpmk::launch_fixity(e->fix, e->d_inst, id, nng * NGRAN, e->s);
  // the slots' cached node boxes belong to earlier occupants
  mpmk::launch_clear_ints(e->boxes + pg0 * (PGRAN / 64) * BOX_INTS, npg * (PGRAN / 64) * BOX_INTS, e->s);
  HIPCHK(hipGetLastError());
  e->lists_dirty = true;
  mark_graph_dirty(e);
  return id;
}

static void remove_instance(mpm_engine* e, int id) {
  if (id < 0 || id >= e->max_inst || !e->inst[id].live) fail(MPM_ERR_ARG, "no live instance %d", id);
  // no wait: queued steps may still run over the instance; whatever is
  // later written into its ranges is ordered behind them on the stream
  HostInst& h = e->inst[id];
  e->pa.release(h.pg0, h.npg);
  e->na.release(h.ng0, h.nng);
  e->ma.release(h.mat0, h.nmat);
  e->ba.release(h.body0, h.nbody);
  h.live = false;
  e->free_ids.push_back(id);
  e->lists_dirty = true;
  mark_graph_dirty(e);
}

// Launch lists live in device buffers of the largest size they can take
// (so they are never reallocated under queued launches) and are refilled by
// copies ordered on the engine stream: launches already queued read the old
// lists, later ones the new.  Buffers of groups that vanish are kept for
// reuse until the engine is destroyed.
static void rebuild_lists(mpm_engine* e) {
  if (!e->lists_dirty) return;
  std::vector<Group> old;
  old.swap(e->groups);
  for (auto& g : old)
    if (g.d_pgran) e->spare_lists.push_back({g.d_pgran, g.d_ngran, g.d_ilist});
  std::map<std::pair<int, int>, int> gid;
  std::vector<int> all;
  for (int id = 0; id < e->max_inst; id++) {
    const HostInst& h = e->inst[id];
    if (!h.live) continue;
    all.push_back(id);
    auto key = std::make_pair(h.c.shape, h.c.scheme);
    auto it = gid.find(key);
    int g;
    if (it == gid.end()) {
      g = (int)e->groups.size();
      gid[key] = g;
      e->groups.emplace_back();
      Group& G = e->groups[g];
      G.shape = h.c.shape;
      G.scheme = h.c.scheme;
      if (!e->spare_lists.empty()) {
        const auto l = e->spare_lists.back();
        e->spare_lists.pop_back();
        G.d_pgran = l[0];
        G.d_ngran = l[1];
        G.d_ilist = l[2];
      } else {
        G.d_pgran = dalloc<int>(e->np / PGRAN);
        G.d_ngran = dalloc<int>(e->nn / NGRAN);
        G.d_ilist = dalloc<int>(e->max_inst);
      }
      G.cap_p = e->np / PGRAN;
      G.cap_n = e->nn / NGRAN;
      G.cap_i = e->max_inst;
    } else {
      g = it->second;
    }
    Group& G = e->groups[g];
    G.insts.push_back(id);
    G.matmask |= h.matmask;
    G.drda = G.drda || (h.c.flags & IF_DRDA) != 0;
    G.smooth = G.smooth || (h.c.flags & IF_SMOOTH) != 0;
    G.contact = G.contact || h.c.contact == 1;
    G.jc = G.jc || (h.c.flags & IF_JC) != 0;
    for (long k = 0; k < h.npg; k++) G.pgran.push_back((int)(h.pg0 + k));
    for (long k = 0; k < h.nng; k++) G.ngran.push_back((int)(h.ng0 + k));
  }
  if (!e->d_ilist_all) {
    e->d_ilist_all = dalloc<int>(e->max_inst);
    e->cap_ilist = e->max_inst;
  }
  size_t bytes = align16(all.size() * sizeof(int));
  for (auto& G : e->groups)
    bytes += align16(G.pgran.size() * sizeof(int)) + align16(G.ngran.size() * sizeof(int)) +
             align16(G.insts.size() * sizeof(int));
  int slot;
  char* st = stage_get(e, bytes, slot);
  size_t off = 0;
  auto put = [&](int* dst, const std::vector<int>& v) {
    if (v.empty()) return;
    memcpy(st + off, v.data(), v.size() * sizeof(int));
    HIPCHK(hipMemcpyAsync(dst, st + off, v.size() * sizeof(int), hipMemcpyHostToDevice, e->s));
    off += align16(v.size() * sizeof(int));
  };
  for (auto& G : e->groups) {
    put(G.d_pgran, G.pgran);
    put(G.d_ngran, G.ngran);
    put(G.d_ilist, G.insts);
  }
  put(e->d_ilist_all, all);
  stage_done(e, slot);
  e->n_active = (int)all.size();
  e->lists_dirty = false;
}

static LaunchP plist(const mpm_engine* e, const Group& G, int bs = 0) {
  LaunchP L;
  L.glist = G.d_pgran;
  L.ginst = e->d_ginst;
  L.bs = bs > 0 ? bs : e->bs_p;
  L.bpg = PGRAN / L.bs;
  L.nblk = (int)G.pgran.size() * L.bpg;
  return L;
}

static LaunchP nlist(const mpm_engine* e, const Group& G) {
  LaunchP L;
  L.glist = G.d_ngran;
  L.ginst = e->d_nginst;
  L.bs = e->bs_n;
  L.bpg = NGRAN / e->bs_n;
  L.nblk = (int)G.ngran.size() * L.bpg;
  return L;
}

// --------------------------------------------------------- phases

enum Phase {
  PH_BEGIN, PH_GMI, PH_BC, PH_STRESS, PH_GMU, PH_INTEGRATE, PH_G2P, PH_MUSL, PH_ENERGY, PH_DT, PH_SMOOTH,
  PH_CONTACT
};

// the smoothing shape functions of the general modes need the step-start
// positions for the linear and SGMP shapes (their XN, XXN)
static bool smooth_saves_x(const Group& G) {
  return G.smooth && (G.shape == MPM_SHAPE_LINEAR || G.shape == MPM_SHAPE_SGMP);
}

// ---- reference-order FP64 mode (MPM_PREC_FP64_EXACT)

static int bits_for(long v) {
  int b = 0;
  while ((1L << b) < v) b++;
  return b;
}

// stencil-entry buffers of the whole pool: 27 entries per particle slot
static void ensure_exact(mpm_engine* e) {
  if (!e->exact || e->x_key[0]) return;
  const long m = 27 * e->np;
  if (m >= (1L << 31))
    fail(MPM_ERR_UNSUPPORTED, "reference-order mode: %ld particle slots exceed 2^31 stencil entries", e->np);
  // keys are pool nodes, below the all-ones key of unused places
  const int obits = 0, nbits = bits_for(e->nn + 1);
  e->x_m = m;
  e->x_obits = obits;
  e->x_nbits = nbits;
  for (int k = 0; k < 2; k++) {
    e->x_key[k] = dalloc<unsigned>(m);
    e->x_val[k] = dalloc<int>(m);
  }
  e->x_rnode = dalloc<int>(m);
  e->x_rw = dalloc<double>(4 * m);
  e->x_prec = dalloc<double>(12 * e->np);
  e->x_seg0 = dalloc<int>(e->nn);
  e->x_seg1 = dalloc<int>(e->nn);
  e->x_scr = dalloc<unsigned>(mpmk::radix_scratch_words(m));
}

static XPtrs xptrs(const mpm_engine* e) {
  XPtrs X;
  X.key = e->x_key[0];
  X.kval = e->x_val[0];
  X.rnode = e->x_rnode;
  X.rw = e->x_rw;
  X.prec = e->x_prec;
  X.seg0 = e->x_seg0;
  X.seg1 = e->x_seg1;
  X.ndir = e->x_ndir;
  X.cf = e->x_cf;
  X.m = e->x_m;
  X.obits = e->x_obits;
  return X;
}

// GridMomentumInitial of every group: ParticleInitial and the stencil
// entries placed in the original particle order, one stable sort of them by
// node (each node's entries then run in ascending original particle index,
// then list position), the node ranges, then every node's sums in that order
// (update_step.f90:288-406)
static void exact_gmi(mpm_engine* e) {
  ensure_exact(e);
  const PoolPtrs<double> P = e->ptrs<double>();
  const XPtrs X = xptrs(e);
  hipStream_t s = e->s;
  mpmk::launch_x_fill_keys(e->x_key[0], e->x_val[0], e->x_m, s);
  for (Group& G : e->groups) mpmk::launch_x_pairs(G.shape, P, X, plist(e, G), s);
  e->x_sorted = mpmk::radix_sort_pairs<unsigned>(e->x_key[0], e->x_key[1], e->x_val[0], e->x_val[1],
                                                           e->x_m, 0, e->x_nbits, e->x_scr, s);
  mpmk::launch_clear_ints(e->x_seg0, e->nn, s);
  mpmk::launch_clear_ints(e->x_seg1, e->nn, s);
  mpmk::launch_x_segments(e->x_key[e->x_sorted], e->x_m, e->x_obits, e->x_seg0, e->x_seg1, s);
  const int* sval = e->x_val[e->x_sorted];
  for (Group& G : e->groups) {
    const LaunchP N = nlist(e, G);
    mpmk::launch_x_prec(XS_MASS | XS_MOM, P, X, plist(e, G), s);
    if (G.shape == MPM_SHAPE_SGMP) {
      mpmk::launch_x_scatter(XS_SGMP | XS_MASS | XS_MOM, P, X, sval, N, s);
      mpmk::launch_sgmp_nodes<double>(SGN_MASSMOM, P, N, s);
    } else {
      mpmk::launch_x_scatter(XS_MASS | XS_MOM, P, X, sval, N, s);
    }
    for (int id : G.insts) e->inst[id].p_in_p2 = false;
  }
}

// one later phase of one group in the reference-order mode: the particle
// scatters reuse the step's sorted entries, the gathers walk each
// particle's entries in list order; node passes and the SGMP centre <->
// node passes are the general kernels
static void exact_group_phase(mpm_engine* e, Group& G, int ph) {
  const PoolPtrs<double> P = e->ptrs<double>();
  const XPtrs X = xptrs(e);
  hipStream_t s = e->s;
  const LaunchP L = plist(e, G), N = nlist(e, G);
  const bool sg = G.shape == MPM_SHAPE_SGMP, usf = G.scheme == MPM_USF;
  const int* sval = e->x_val[e->x_sorted];
  switch (ph) {
    case PH_BC:
      mpmk::launch_bc<double>(P, N, 0, 1, 0, s);
      break;
    case PH_STRESS:
      if (sg) {
        mpmk::launch_bc<double>(P, N, 0, 0, 1, s);
        mpmk::launch_sgmp_cells<double>(SGC_GRAD, P, N, s);
        mpmk::launch_x_sgmp_stress(usf ? 0 : 1, P, L, s);
      } else {
        mpmk::launch_x_stress(G.shape, usf ? 0 : 1, P, X, L, s);
      }
      if (G.jc) mpmk::launch_x_jc(P, L, G.d_ilist, (int)G.insts.size(), s);
      for (int id : G.insts) e->inst[id].xx_pending = false;
      break;
    case PH_GMU:
      mpmk::launch_x_prec(XS_FORCE, P, X, L, s);
      if (sg) {
        mpmk::launch_x_scatter(XS_SGMP | XS_FORCE, P, X, sval, N, s);
        mpmk::launch_sgmp_nodes<double>(SGN_FORCE, P, N, s);
      } else {
        mpmk::launch_x_scatter(G.contact ? XS_FORCE | XS_NDIR : XS_FORCE, P, X, sval, N, s);
      }
      break;
    case PH_CONTACT:
      if (G.contact) mpmk::launch_x_contact(P, X, N, G.d_ilist, (int)G.insts.size(), s);
      break;
    case PH_INTEGRATE:
      mpmk::launch_integrate<double>(P, N, 0, s);
      break;
    case PH_G2P:
      if (sg) {
        mpmk::launch_node_va<double>(P, N, s);
        mpmk::launch_sgmp_cells<double>(SGC_KIN, P, N, s);
        mpmk::launch_g2p<double>(MPM_SHAPE_SGMP, usf ? G2P_MOVE : 0, P, L, s);
      } else {
        mpmk::launch_x_g2p(G.shape, usf ? 1 : 0, P, X, L, s);
      }
      for (int id : G.insts) e->inst[id].xx_pending = !usf;
      break;
    case PH_MUSL:
      mpmk::launch_x_prec(XS_MOM, P, X, L, s);
      if (sg) {
        mpmk::launch_x_scatter(XS_SGMP | XS_MOM, P, X, sval, N, s);
        mpmk::launch_sgmp_nodes<double>(SGN_MOM, P, N, s);
      } else {
        mpmk::launch_x_scatter(XS_MOM, P, X, sval, N, s);
      }
      mpmk::launch_bc<double>(P, N, 0, 1, 0, s);
      break;
    default:
      break;
  }
}

// SmoothStressbyGrid of every smoothing group in the reference-order mode,
// after all of the step's other phases: linear and SGMP take the step's
// entries (GridMomentumInitial's XN, XXN), B-spline rewrites its entries'
// weights at the current position (same node list, so the sorted order
// stands), GIMP rebuilds its entries at the current position (its case may
// change) and sorts them again -- last, since that sort replaces every
// group's entries
static void exact_smooth(mpm_engine* e) {
  const PoolPtrs<double> P = e->ptrs<double>();
  const XPtrs X = xptrs(e);
  hipStream_t s = e->s;
  bool gimp = false;
  for (Group& G : e->groups) {
    if (!G.smooth) continue;
    if (G.shape == MPM_SHAPE_GIMP) {
      gimp = true;
      continue;
    }
    if (G.shape == MPM_SHAPE_BSPLINE) mpmk::launch_x_pairs_smooth(G.shape, P, X, plist(e, G), s);
    mpmk::launch_x_smooth(G.shape, P, X, e->sm, e->x_val[e->x_sorted], plist(e, G), nlist(e, G), s);
  }
  if (!gimp) return;
  mpmk::launch_x_fill_keys(e->x_key[0], e->x_val[0], e->x_m, s);
  for (Group& G : e->groups)
    if (G.smooth && G.shape == MPM_SHAPE_GIMP) mpmk::launch_x_pairs_smooth(G.shape, P, X, plist(e, G), s);
  e->x_sorted = mpmk::radix_sort_pairs<unsigned>(e->x_key[0], e->x_key[1], e->x_val[0], e->x_val[1],
                                                           e->x_m, 0, e->x_nbits, e->x_scr, s);
  mpmk::launch_clear_ints(e->x_seg0, e->nn, s);
  mpmk::launch_clear_ints(e->x_seg1, e->nn, s);
  mpmk::launch_x_segments(e->x_key[e->x_sorted], e->x_m, e->x_obits, e->x_seg0, e->x_seg1, s);
  for (Group& G : e->groups)
    if (G.smooth && G.shape == MPM_SHAPE_GIMP)
      mpmk::launch_x_smooth(G.shape, P, X, e->sm, e->x_val[e->x_sorted], plist(e, G), nlist(e, G), s);
}

// the grid state IntegrateMomentum leaves, for instances whose last fused
// USF step integrated on the fly inside g2p (before any read of the grid or
// any per-phase call): the host flag says "maybe", the instance's device
// flag (set by that step's finalize, so not for instances that had already
// stopped) decides
template <typename T>
static void settle_integrate(mpm_engine* e, int only = -1) {
  for (size_t id = 0; id < e->inst.size(); id++) {
    HostInst& h = e->inst[id];
    if (!h.live || !h.integ_pending || (only >= 0 && (int)id != only)) continue;
    mpmk::launch_integrate_range<T>(e->ptrs<T>(), (int)id, h.ng0 * NGRAN, h.c.nnode_all, e->s);
    h.integ_pending = false;
  }
}

template <typename T>
static void run_phase(mpm_engine* e, int ph) {
  settle_integrate<T>(e);
  rebuild_lists(e);
  const PoolPtrs<T> P = e->ptrs<T>();
  hipStream_t s = e->s;
  if (ph == PH_BEGIN) {
    mpmk::launch_begin_step(e->d_ts, e->d_ilist_all, e->n_active, s);
    return;
  }
  if (e->exact && ph != PH_ENERGY && ph != PH_DT) {
    if (ph == PH_GMI) exact_gmi(e);
    else if (ph == PH_SMOOTH) exact_smooth(e);
    else
      for (Group& G : e->groups) exact_group_phase(e, G, ph);
    HIPCHK(hipGetLastError());
    return;
  }
  for (Group& G : e->groups) {
    const LaunchP L = plist(e, G), N = nlist(e, G), LS = plist(e, G, e->bs_s);
    const bool sg = G.shape == MPM_SHAPE_SGMP;
    switch (ph) {
      case PH_GMI:
        if (sg) {   // particles -> auxiliary nodes -> grid nodes
          mpmk::launch_zero_nodes<double>(e->cmg, 1, e->nn, N, s);
          mpmk::launch_zero_nodes<T>((T*)e->aux, 3, e->nn, N, s);
          mpmk::launch_sgmp_p2g<T>(P2G_ICELL | P2G_MASS | P2G_MOM, P, LS, s);
          mpmk::launch_sgmp_nodes<T>(SGN_MASSMOM, P, N, s);
        } else {
          mpmk::launch_zero_nodes<double>(e->mg, 1, e->nn, N, s);
          mpmk::launch_zero_nodes<T>((T*)e->pg, 3, e->nn, N, s);
          mpmk::launch_p2g<T>(G.shape, P2G_ICELL | P2G_MASS | P2G_MOM, P, LS, s);
        }
        if (smooth_saves_x(G)) mpmk::launch_smooth_save_x<T>(P, e->sm, L, s);
        for (int id : G.insts) e->inst[id].p_in_p2 = false;
        break;
      case PH_BC:
        mpmk::launch_bc<T>(P, N, 0, 1, 0, s);
        break;
      case PH_STRESS:
        mpmk::launch_bc<T>(P, N, 0, 0, 1, s);   // nodal velocity of the current PXg
        if (sg) mpmk::launch_sgmp_cells<T>(SGC_GRAD, P, N, s);
        mpmk::launch_stress<T>(G.shape, G.scheme == MPM_USF ? 0 : ST_MOVE, MM_ALL, P, L, s);
        for (int id : G.insts) e->inst[id].xx_pending = false;
        break;
      case PH_GMU:
        if (sg) {
          mpmk::launch_zero_nodes<T>((T*)e->aux + AUX_FORCE_OFF * e->nn, 9, e->nn, N, s);
          mpmk::launch_sgmp_p2g<T>(P2G_FORCE, P, LS, s);
          mpmk::launch_sgmp_nodes<T>(SGN_FORCE, P, N, s);
        } else {
          mpmk::launch_zero_nodes<T>((T*)e->fg, 3, e->nn, N, s);
          mpmk::launch_p2g<T>(G.shape, P2G_FORCE, P, LS, s);
        }
        break;
      case PH_INTEGRATE:
        mpmk::launch_integrate<T>(P, N, 0, s);
        break;
      case PH_G2P:
        mpmk::launch_node_va<T>(P, N, s);
        if (sg) mpmk::launch_sgmp_cells<T>(SGC_KIN, P, N, s);
        mpmk::launch_g2p<T>(G.shape, G.scheme == MPM_USF ? G2P_MOVE : 0, P, L, s);
        for (int id : G.insts) e->inst[id].xx_pending = G.scheme != MPM_USF;
        break;
      case PH_MUSL:
        if (sg) {
          mpmk::launch_zero_nodes<T>((T*)e->aux, 3, e->nn, N, s);
          mpmk::launch_sgmp_p2g<T>(P2G_MOM, P, LS, s);
          mpmk::launch_sgmp_nodes<T>(SGN_MOM, P, N, s);
        } else {
          mpmk::launch_zero_nodes<T>((T*)e->pg, 3, e->nn, N, s);
          mpmk::launch_p2g<T>(G.shape, P2G_MOM, P, LS, s);
        }
        mpmk::launch_bc<T>(P, N, 0, 1, 0, s);
        break;
      case PH_ENERGY:
        mpmk::launch_energy<T>(P, L, s);
        mpmk::launch_finalize<T>(P, G.d_ilist, (int)G.insts.size(), FIN_ENERGY, s);
        break;
      case PH_DT:
        mpmk::launch_energy<T>(P, L, s);
        mpmk::launch_finalize<T>(P, G.d_ilist, (int)G.insts.size(), FIN_DT, s);
        if (G.drda) mpmk::launch_damp<T>(P, L, s);
        break;
      case PH_SMOOTH:
        if (G.smooth) mpmk::launch_smooth<T>(G.shape, P, e->sm, L, N, s);
        break;
    }
  }
  HIPCHK(hipGetLastError());
}

static void phase(mpm_engine* e, int ph) {
  if (e->prec == MPM_PREC_FP32) run_phase<float>(e, ph);
  else run_phase<double>(e, ph);
}

// kernel classes timed by the profiling mode
enum KernelId { KI_ZERO = 0, KI_P2G, KI_BC, KI_STRESS, KI_INTEGRATE, KI_G2P, KI_FINALIZE, KI_BEGIN, KI_COUNT };
static const char* KI_NAMES[KI_COUNT] = {"zero", "p2g", "bc", "stress", "integrate", "g2p", "finalize", "begin"};

// HIP-event bracket around one launch when the engine is profiling
template <typename F>
static void timed(mpm_engine* e, int kid, F&& launch) {
  if (!e->prof) {
    launch();
    return;
  }
  if (e->ev_used + 2 > e->ev_pool.size()) {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    e->ev_pool.push_back(a);
    e->ev_pool.push_back(b);
  }
  hipEvent_t a = e->ev_pool[e->ev_used], b = e->ev_pool[e->ev_used + 1];
  e->ev_used += 2;
  HIPCHK(hipEventRecord(a, e->s));
  launch();
  HIPCHK(hipEventRecord(b, e->s));
  e->ev_kid.push_back(kid);
}

// one step of an SGMP group: particle passes through the auxiliary nodes,
// with the centre <-> node passes between them (timed with their phase)
template <typename T>
static void enqueue_sgmp(mpm_engine* e, Group& G, const PoolPtrs<T>& P, const LaunchP& L, const LaunchP& LS,
                         const LaunchP& N) {
  hipStream_t s = e->s;
  const long nn = e->nn;
  T* aux = (T*)e->aux;
  auto p2g = [&](int mode) {
    timed(e, KI_P2G, [&] {
      mpmk::launch_zero_nodes<double>(e->cmg, 1, nn, N, s);
      mpmk::launch_zero_nodes<T>(aux, (mode & P2G_FORCE) ? 12 : 3, nn, N, s);
      mpmk::launch_sgmp_p2g<T>(mode, P, LS, s);
      mpmk::launch_sgmp_nodes<T>((mode & P2G_FORCE) ? SGN_MASSMOM | SGN_FORCE : SGN_MASSMOM, P, N, s);
    });
  };
  if (G.scheme == MPM_USF) {
    p2g(P2G_ICELL | P2G_MASS | P2G_MOM);
    timed(e, KI_BC, [&] { mpmk::launch_bc<T>(P, N, 0, 1, 1, s); });
    timed(e, KI_STRESS, [&] {
      mpmk::launch_sgmp_cells<T>(SGC_GRAD, P, N, s);
      mpmk::launch_stress<T>(MPM_SHAPE_SGMP, 0, G.matmask, P, L, s);
    });
    timed(e, KI_P2G, [&] {
      mpmk::launch_zero_nodes<T>(aux + AUX_FORCE_OFF * nn, 9, nn, N, s);
      mpmk::launch_sgmp_p2g<T>(P2G_FORCE, P, LS, s);
      mpmk::launch_sgmp_nodes<T>(SGN_FORCE, P, N, s);
    });
    timed(e, KI_INTEGRATE, [&] { mpmk::launch_integrate<T>(P, N, 1, s); });
    timed(e, KI_G2P, [&] {
      mpmk::launch_sgmp_cells<T>(SGC_KIN, P, N, s);
      mpmk::launch_g2p<T>(MPM_SHAPE_SGMP, G2P_MOVE | G2P_ENERGY, P, L, s);
    });
  } else {
    p2g(P2G_ICELL | P2G_MASS | P2G_MOM | P2G_FORCE);
    timed(e, KI_INTEGRATE, [&] { mpmk::launch_integrate<T>(P, N, 1, s); });
    timed(e, KI_G2P, [&] {
      mpmk::launch_sgmp_cells<T>(SGC_KIN, P, N, s);
      mpmk::launch_g2p<T>(MPM_SHAPE_SGMP, 0, P, L, s);
    });
    if (G.scheme == MPM_MUSL) {
      timed(e, KI_P2G, [&] {
        mpmk::launch_zero_nodes<T>(aux, 3, nn, N, s);
        mpmk::launch_sgmp_p2g<T>(P2G_MOM, P, LS, s);
        mpmk::launch_sgmp_nodes<T>(SGN_MOM | SGN_TO_P2, P, N, s);
      });
      timed(e, KI_BC, [&] { mpmk::launch_bc<T>(P, N, 1, 1, 1, s); });
    }
    timed(e, KI_STRESS, [&] {
      mpmk::launch_sgmp_cells<T>(SGC_GRAD, P, N, s);
      mpmk::launch_stress<T>(MPM_SHAPE_SGMP, ST_MOVE | ST_ENERGY, G.matmask, P, L, s);
    });
  }
}

// one step of every active instance in the reference-order mode: the
// phases of update_step.f90 one by one, in the call order of
// MPM3D.f90:69-113 (no fusion: each scatter is its own sorted-order pass)
static void enqueue_step_exact(mpm_engine* e, int fin_flags) {
  const PoolPtrs<double> P = e->ptrs<double>();
  hipStream_t s = e->s;
  timed(e, KI_BEGIN, [&] { mpmk::launch_begin_step(e->d_ts, e->d_ilist_all, e->n_active, s); });
  timed(e, KI_P2G, [&] { exact_gmi(e); });
  for (Group& G : e->groups) {
    const LaunchP L = plist(e, G);
    auto ph = [&](int kid, int p) { timed(e, kid, [&] { exact_group_phase(e, G, p); }); };
    if (G.scheme == MPM_USF) {
      ph(KI_BC, PH_BC);
      ph(KI_STRESS, PH_STRESS);
    }
    ph(KI_P2G, PH_GMU);
    ph(KI_INTEGRATE, PH_INTEGRATE);
    if (G.contact) ph(KI_INTEGRATE, PH_CONTACT);   // Lagr_NodContact (MPM3D.f90:94-96)
    ph(KI_G2P, PH_G2P);
    if (G.scheme == MPM_MUSL) ph(KI_P2G, PH_MUSL);
    if (G.scheme != MPM_USF) ph(KI_STRESS, PH_STRESS);
  }
  // SmoothStressbyGrid (timed with finalize) before finalize, which may
  // mark an instance done
  bool smooth = false;
  for (Group& G : e->groups) smooth = smooth || G.smooth;
  if (smooth) timed(e, KI_FINALIZE, [&] { exact_smooth(e); });
  for (Group& G : e->groups) {
    const LaunchP L = plist(e, G);
    timed(e, KI_FINALIZE, [&] {
      mpmk::launch_energy<double>(P, L, s);
      mpmk::launch_finalize<double>(P, G.d_ilist, (int)G.insts.size(), FIN_ENERGY | FIN_DT | fin_flags, s);
      if (G.drda) mpmk::launch_damp<double>(P, L, s);
    });
  }
}

// one fused step of every active instance (MPM3D.f90:69-113)
template <typename T>
static void enqueue_step(mpm_engine* e, int fin_flags) {
  if (e->exact) {
    enqueue_step_exact(e, fin_flags);
    return;
  }
  const PoolPtrs<T> P = e->ptrs<T>();
  hipStream_t s = e->s;
  timed(e, KI_BEGIN, [&] { mpmk::launch_begin_step(e->d_ts, e->d_ilist_all, e->n_active, s); });
  for (Group& G : e->groups) {
    const LaunchP L = plist(e, G), N = nlist(e, G), LS = plist(e, G, e->bs_s);
    timed(e, KI_ZERO, [&] { mpmk::launch_zero_grid<T>(P, G.scheme == MPM_MUSL, N, s); });
    if (smooth_saves_x(G)) timed(e, KI_FINALIZE, [&] { mpmk::launch_smooth_save_x<T>(P, e->sm, L, s); });
    if (G.shape == MPM_SHAPE_SGMP) {
      enqueue_sgmp<T>(e, G, P, L, LS, N);
    } else if (G.scheme == MPM_USF) {
      // ApplyBoundaryConditions ahead of the stress update and the nodal
      // velocity/acceleration of ParticlePositionUpdate are formed in the
      // tile loads of the two gathers (the bc pass and IntegrateMomentum's
      // V, A stores are not needed: integrate zeroes the fixed directions
      // of PXg and FXg itself, and V, A are not grid state)
      // IntegrateMomentum is folded into g2p's tile load as well (the node
      // pass that stored the integrated momentum is gone from the step; the
      // grid is integrated only when it is read, settle_integrate)
      timed(e, KI_P2G, [&] { mpmk::launch_p2g<T>(G.shape, P2G_ICELL | P2G_MASS | P2G_MOM, P, LS, s); });
      timed(e, KI_STRESS, [&] { mpmk::launch_stress<T>(G.shape, ST_FORCE | ST_VFLY, G.matmask, P, L, s); });
      timed(e, KI_G2P, [&] {
        mpmk::launch_g2p<T>(G.shape, G2P_MOVE | G2P_ENERGY | G2P_VAFLY | G2P_INTFLY, P, L, s);
      });
      for (int id : G.insts) e->inst[id].integ_pending = true;   // host: maybe (the device flag decides)
    } else if (G.scheme == MPM_USL) {
      timed(e, KI_P2G, [&] { mpmk::launch_p2g<T>(G.shape, P2G_ICELL | P2G_MASS | P2G_MOM | P2G_FORCE, P, LS, s); });
      timed(e, KI_INTEGRATE, [&] { mpmk::launch_integrate<T>(P, N, 1, s); });
      timed(e, KI_G2P, [&] { mpmk::launch_g2p<T>(G.shape, 0, P, L, s); });
      timed(e, KI_STRESS, [&] { mpmk::launch_stress<T>(G.shape, ST_MOVE | ST_ENERGY, G.matmask, P, L, s); });
    } else {
      timed(e, KI_P2G, [&] { mpmk::launch_p2g<T>(G.shape, P2G_ICELL | P2G_MASS | P2G_MOM | P2G_FORCE, P, LS, s); });
      timed(e, KI_INTEGRATE, [&] { mpmk::launch_integrate<T>(P, N, 1, s); });
      timed(e, KI_G2P, [&] { mpmk::launch_g2p<T>(G.shape, G2P_REMAP, P, L, s); });
      timed(e, KI_BC, [&] { mpmk::launch_bc<T>(P, N, 1, 1, 1, s); });
      timed(e, KI_STRESS, [&] { mpmk::launch_stress<T>(G.shape, ST_MOVE | ST_ENERGY, G.matmask, P, L, s); });
    }
    const int intpend = (G.shape != MPM_SHAPE_SGMP && G.scheme == MPM_USF) ? FIN_INTPEND : 0;
    // SmoothStressbyGrid (after UpdateDT in the reference; it reads none of
    // finalize's results, and finalize may mark an instance done)
    if (G.smooth) timed(e, KI_FINALIZE, [&] { mpmk::launch_smooth<T>(G.shape, P, e->sm, L, N, s); });
    timed(e, KI_FINALIZE, [&] {
      mpmk::launch_finalize<T>(P, G.d_ilist, (int)G.insts.size(), FIN_ENERGY | FIN_DT | fin_flags | intpend, s);
      if (G.drda) mpmk::launch_damp<T>(P, L, s);
    });
  }
}

// fold the recorded event pairs into per-kernel totals (synchronising)
static void collect_times(mpm_engine* e) {
  if (e->ev_kid.empty()) return;
  HIPCHK(hipStreamSynchronize(e->s));
  for (size_t k = 0; k < e->ev_kid.size(); k++) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e->ev_pool[2 * k], e->ev_pool[2 * k + 1]));
    e->kt_ms[e->ev_kid[k]] += ms;
    e->kt_n[e->ev_kid[k]] += 1;
  }
  e->ev_kid.clear();
  e->ev_used = 0;
}

static void mark_after_step(mpm_engine* e) {
  for (auto& G : e->groups)
    for (int id : G.insts) {
      e->inst[id].p_in_p2 = G.scheme == MPM_MUSL && !e->exact;   // the exact step re-maps into PXg
      e->inst[id].xx_pending = false;
    }
}

template <typename T>
static void permute_pool(mpm_engine* e);

static void sort_particles(mpm_engine* e);

static void step_plain(mpm_engine* e, int nsteps, int fin_flags) {
  rebuild_lists(e);
  for (int k = 0; k < nsteps; k++) {
    if (e->sort_every > 0 && e->since_sort >= e->sort_every) sort_particles(e);
    if (e->prec == MPM_PREC_FP32) enqueue_step<float>(e, fin_flags);
    else enqueue_step<double>(e, fin_flags);
    e->steps_done++;
    e->since_sort++;
  }
  HIPCHK(hipGetLastError());
  mark_after_step(e);
}

// graph of `per` fused steps, replayed; instances finish on their own.
// lazy: after the launch lists change, the first call launches plainly and
// only a second call with the same lists captures (continuous batching
// changes the instance set every few steps, and a capture costs several
// plain steps)
static void drop_graphs(mpm_engine* e) {
  for (int k = 0; k < 2; k++) {
    if (e->gexec[k]) HIPCHK(hipGraphExecDestroy(e->gexec[k]));
    if (e->graph[k]) HIPCHK(hipGraphDestroy(e->graph[k]));
    e->gexec[k] = nullptr;
    e->graph[k] = nullptr;
    e->gkey[k] = nullptr;
  }
  e->gnext = 0;
}

// the graph of `per` steps over the current particle buffers, captured on
// first use
static hipGraphExec_t graph_for_buffers(mpm_engine* e, int per, int fin_flags) {
  for (int k = 0; k < 2; k++)
    if (e->gexec[k] && e->gkey[k] == e->t) return e->gexec[k];
  const int k = !e->gexec[0] ? 0 : !e->gexec[1] ? 1 : e->gnext;
  e->gnext = 1 - k;
  if (e->gexec[k]) HIPCHK(hipGraphExecDestroy(e->gexec[k]));
  if (e->graph[k]) HIPCHK(hipGraphDestroy(e->graph[k]));
  e->gexec[k] = nullptr;
  e->graph[k] = nullptr;
  const bool was_prof = e->prof;
  e->prof = false;
  ensure_exact(e);   // no allocation inside the capture
  HIPCHK(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
  for (int q = 0; q < per; q++) {
    if (e->prec == MPM_PREC_FP32) enqueue_step<float>(e, fin_flags);
    else enqueue_step<double>(e, fin_flags);
  }
  HIPCHK(hipStreamEndCapture(e->s, &e->graph[k]));
  e->prof = was_prof;
  HIPCHK(hipGraphInstantiate(&e->gexec[k], e->graph[k], nullptr, nullptr, 0));
  e->gkey[k] = e->t;
  return e->gexec[k];
}

static void step_graph(mpm_engine* e, int nsteps, int per, int fin_flags, bool lazy) {
  rebuild_lists(e);
  if (per < 1) per = 1;
  if (e->graph_dirty || e->graph_steps != per) {
    if (lazy && !e->graph_plain_once) {
      step_plain(e, nsteps, fin_flags);
      e->graph_plain_once = true;
      return;
    }
    drop_graphs(e);
    e->graph_steps = per;
    e->graph_dirty = false;
    e->graph_plain_once = false;
  }
  int done = 0;
  while (done + per <= nsteps) {
    const long se = e->sort_every;
    if (se > 0 && e->since_sort + per > se) sort_particles(e);   // at most one period between re-sorts
    HIPCHK(hipGraphLaunch(graph_for_buffers(e, per, fin_flags), e->s));
    done += per;
    e->steps_done += per;
    e->since_sort += per;
  }
  if (done < nsteps) step_plain(e, nsteps - done, fin_flags);
  mark_after_step(e);
}

// ---------------------------------------------------------- sorting

// largest sort key of an instance (the key of particles outside the grid)
static unsigned max_sort_key(const InstC& c, int brick) {
  const int hb = brick == 0 ? hilbert_bits(c) : 0;
  if (hb > 0) return 1u << (3 * hb);
  const unsigned B = brick == 0 ? 4u : (unsigned)brick;
  return ((c.ncx + B - 1) / B) * ((c.ncy + B - 1) / B) * ((c.ncz + B - 1) / B) * B * B * B;
}

// Stable sort of every instance's particles by cell key (Hilbert or brick order)
// (segmented LSD radix sort: equal keys keep their order, so the
// permutation is the one a stable CPU sort of the same keys gives), then
// gather all fields.
static void ensure_sort_buffers(mpm_engine* e) {
  const long np = e->np;
  if (!e->k_in) {
    e->k_in = dalloc<unsigned>(np);
    e->k_out = dalloc<unsigned>(np);
    e->v_in = dalloc<int>(np);
    e->v_out = dalloc<int>(np);
    e->t_alt = dalloc<char>((size_t)T_COUNT * np * e->tsz);
    e->d_alt = dalloc<double>((size_t)D_COUNT * np);
    e->i_alt = dalloc<int>((size_t)I_COUNT * np);
    if (e->pk) e->pk_alt = dalloc<double>((size_t)TL_PK_COUNT * np);
    // scratch of a radix sort over the whole pool, which covers the
    // single-instance sorts at admission (no reallocation there)
    e->rs_scr = dalloc<unsigned>(mpmk::radix_scratch_words(np));
  }
}

// Sort one newly admitted instance (its slot range only), so that it joins
// the resident ones in cell order instead of in generator order.
static void sort_instance(mpm_engine* e, int id) {
  ensure_sort_buffers(e);
  const HostInst& h = e->inst[id];
  const long np = e->np, b0 = h.pg0 * PGRAN, pad = h.npg * PGRAN;
  LaunchP L;
  L.glist = e->d_seq + h.pg0;   // the instance's granules pg0, pg0 + 1, ...
  L.ginst = e->d_ginst;
  L.bs = 256;
  L.bpg = PGRAN / 256;
  L.nblk = (int)(h.npg * L.bpg);
  mpmk::launch_sort_keys(e->d, np, e->d_inst, L, e->brick, e->k_in, e->v_in, e->s);
  int end_bit = 1;
  const unsigned maxkey = max_sort_key(h.c, e->brick);
  while (end_bit < 32 && (1u << end_bit) <= maxkey) end_bit++;
  const int par = mpmk::radix_sort_pairs<unsigned>(e->k_in + b0, e->k_out + b0, e->v_in + b0, e->v_out + b0, pad, 0,
                                                   end_bit, e->rs_scr, e->s);
  const int* perm = par ? e->v_out : e->v_in;
  if (e->prec == MPM_PREC_FP32)
    mpmk::launch_permute<float>((float*)e->t_alt, (const float*)e->t, perm, L, np, T_COUNT, e->s);
  else
    mpmk::launch_permute<double>((double*)e->t_alt, (const double*)e->t, perm, L, np, T_COUNT, e->s);
  mpmk::launch_permute<double>(e->d_alt, e->d, perm, L, np, D_COUNT, e->s);
  mpmk::launch_permute<int>(e->i_alt, e->ii, perm, L, np, I_COUNT, e->s);
  if (e->pk) {
    mpmk::launch_permute<double>(e->pk_alt, e->pk, perm, L, np, TL_PK_COUNT, e->s);
    for (int f = 0; f < TL_PK_COUNT; f++)
      HIPCHK(hipMemcpyAsync(e->pk + (size_t)f * np + b0, e->pk_alt + (size_t)f * np + b0, pad * sizeof(double),
                            hipMemcpyDeviceToDevice, e->s));
  }
  for (int f = 0; f < T_COUNT; f++)
    HIPCHK(hipMemcpyAsync((char*)e->t + ((size_t)f * np + b0) * e->tsz, (char*)e->t_alt + ((size_t)f * np + b0) * e->tsz,
                          pad * e->tsz, hipMemcpyDeviceToDevice, e->s));
  for (int f = 0; f < D_COUNT; f++)
    HIPCHK(hipMemcpyAsync(e->d + (size_t)f * np + b0, e->d_alt + (size_t)f * np + b0, pad * sizeof(double),
                          hipMemcpyDeviceToDevice, e->s));
  for (int f = 0; f < I_COUNT; f++)
    HIPCHK(hipMemcpyAsync(e->ii + (size_t)f * np + b0, e->i_alt + (size_t)f * np + b0, pad * sizeof(int),
                          hipMemcpyDeviceToDevice, e->s));
}

static void sort_particles(mpm_engine* e) {
  rebuild_lists(e);
  e->since_sort = 0;
  e->sorts_done++;
  const long np = e->np;
  ensure_sort_buffers(e);
  if (!e->d_seg) e->d_seg = dalloc<int>(2 * (size_t)e->max_inst);
  std::vector<int> seg;
  unsigned maxkey = 1;
  for (auto& G : e->groups) {
    const LaunchP L = plist(e, G);
    mpmk::launch_sort_keys(e->d, np, e->d_inst, L, e->brick, e->k_in, e->v_in, e->s);
    for (int id : G.insts) {
      const HostInst& h = e->inst[id];
      seg.push_back((int)(h.pg0 * PGRAN));
      maxkey = std::max<unsigned>(maxkey, max_sort_key(h.c, e->brick));
    }
  }
  const int nseg = (int)seg.size();
  if (!nseg) return;
  for (auto& G : e->groups)
    for (int id : G.insts) {
      const HostInst& h = e->inst[id];
      seg.push_back((int)((h.pg0 + h.npg) * PGRAN));
    }
  // segments in pool order (begins, then ends)
  {
    std::vector<std::pair<int, int>> sr(nseg);
    for (int k = 0; k < nseg; k++) sr[k] = {seg[k], seg[nseg + k]};
    std::sort(sr.begin(), sr.end());
    for (int k = 0; k < nseg; k++) {
      seg[k] = sr[k].first;
      seg[nseg + k] = sr[k].second;
    }
  }
  int end_bit = 1;
  while (end_bit < 32 && (1u << end_bit) <= maxkey) end_bit++;
  // When the segments tile one slot range without gaps, one device-wide
  // radix sort of (segment rank, cell key) gives the permutation of the
  // segmented sort (both are stable), at a fraction of its time: the
  // segmented sort gives each segment (an instance) little parallelism.
  bool tiled = true;
  for (int k = 1; k < nseg; k++) tiled = tiled && seg[k] == seg[nseg + k - 1];
  int seg_bits = 0;
  while ((1 << seg_bits) < nseg) seg_bits++;
  const bool global = nseg > 1 && tiled && end_bit + seg_bits <= 32;
  if (seg != e->h_seg) {
    // the previous upload may still read h_seg: wait before replacing it
    HIPCHK(hipStreamSynchronize(e->s));
    e->h_seg = seg;
    HIPCHK(hipMemcpyAsync(e->d_seg, e->h_seg.data(), seg.size() * sizeof(int), hipMemcpyHostToDevice, e->s));
  }
  // one instance, or tiled segments: one radix sort of the slot range
  // (the segment rank above the cell key); otherwise one sort per segment
  // (every segment takes the same number of passes, so all results land in
  // the same buffer)
  int par = 0;
  if (nseg == 1 || global) {
    const int b0 = seg[0], n0 = seg[2 * nseg - 1] - seg[0];
    const int nbits = global ? end_bit + seg_bits : end_bit;
    if (global) mpmk::launch_seg_prefix(e->k_in, e->d_seg, nseg, b0, n0, end_bit, e->s);
    par = mpmk::radix_sort_pairs<unsigned>(e->k_in + b0, e->k_out + b0, e->v_in + b0, e->v_out + b0, n0, 0, nbits,
                                           e->rs_scr, e->s);
  } else {
    for (int k = 0; k < nseg; k++) {
      const int b0 = seg[k], n0 = seg[nseg + k] - seg[k];
      par = mpmk::radix_sort_pairs<unsigned>(e->k_in + b0, e->k_out + b0, e->v_in + b0, e->v_out + b0, n0, 0, end_bit,
                                             e->rs_scr, e->s);
    }
  }
  const int* perm = par ? e->v_out : e->v_in;
  for (auto& G : e->groups) {
    const LaunchP L = plist(e, G);
    if (e->prec == MPM_PREC_FP32)
      mpmk::launch_permute<float>((float*)e->t_alt, (const float*)e->t, perm, L, np, T_COUNT, e->s);
    else
      mpmk::launch_permute<double>((double*)e->t_alt, (const double*)e->t, perm, L, np, T_COUNT, e->s);
    mpmk::launch_permute<double>(e->d_alt, e->d, perm, L, np, D_COUNT, e->s);
    mpmk::launch_permute<int>(e->i_alt, e->ii, perm, L, np, I_COUNT, e->s);
    if (e->pk) mpmk::launch_permute<double>(e->pk_alt, e->pk, perm, L, np, TL_PK_COUNT, e->s);
  }
  // the cached node boxes describe the old particle order
  mpmk::launch_clear_ints(e->boxes, (e->np / 64) * BOX_INTS, e->s);
  // the sorted twins become the particle arrays (every live granule was
  // permuted; free granules are never read); captured step graphs are kept
  // per buffer set
  std::swap(e->t, e->t_alt);
  std::swap(e->d, e->d_alt);
  std::swap(e->ii, e->i_alt);
  if (e->pk) std::swap(e->pk, e->pk_alt);
}

// --------------------------------------------------------- autotuner

// Device copy of the resident particle state and per-instance step state,
// so that tuning trials leave the simulation where it was.
struct Snapshot {
  void* t = nullptr;
  double* d = nullptr;
  int* i = nullptr;
  InstT* ts = nullptr;
  double* pk = nullptr;
  long steps = 0, since_sort = 0;
  std::vector<std::pair<bool, bool>> flags;
};

static void snap_free(Snapshot& S) {
  void* b[] = {S.t, S.d, S.i, S.ts, S.pk};
  for (void* p : b)
    if (p) (void)hipFree(p);
  S = Snapshot();
}

static void snap_take(mpm_engine* e, Snapshot& S) {
  const long np = e->np;
  S.t = dalloc<char>((size_t)T_COUNT * np * e->tsz);
  S.d = dalloc<double>((size_t)D_COUNT * np);
  S.i = dalloc<int>((size_t)I_COUNT * np);
  S.ts = dalloc<InstT>(e->max_inst);
  HIPCHK(hipMemcpyAsync(S.t, e->t, (size_t)T_COUNT * np * e->tsz, hipMemcpyDeviceToDevice, e->s));
  HIPCHK(hipMemcpyAsync(S.d, e->d, (size_t)D_COUNT * np * sizeof(double), hipMemcpyDeviceToDevice, e->s));
  HIPCHK(hipMemcpyAsync(S.i, e->ii, (size_t)I_COUNT * np * sizeof(int), hipMemcpyDeviceToDevice, e->s));
  HIPCHK(hipMemcpyAsync(S.ts, e->d_ts, sizeof(InstT) * e->max_inst, hipMemcpyDeviceToDevice, e->s));
  if (e->pk) {
    S.pk = dalloc<double>((size_t)TL_PK_COUNT * np);
    HIPCHK(hipMemcpyAsync(S.pk, e->pk, (size_t)TL_PK_COUNT * np * sizeof(double), hipMemcpyDeviceToDevice, e->s));
  }
  S.steps = e->steps_done;
  S.since_sort = e->since_sort;
  S.flags.clear();
  for (const HostInst& h : e->inst) S.flags.emplace_back(h.p_in_p2, h.xx_pending);
}

static void snap_restore(mpm_engine* e, const Snapshot& S) {
  const long np = e->np;
  HIPCHK(hipMemcpyAsync(e->t, S.t, (size_t)T_COUNT * np * e->tsz, hipMemcpyDeviceToDevice, e->s));
  HIPCHK(hipMemcpyAsync(e->d, S.d, (size_t)D_COUNT * np * sizeof(double), hipMemcpyDeviceToDevice, e->s));
  HIPCHK(hipMemcpyAsync(e->ii, S.i, (size_t)I_COUNT * np * sizeof(int), hipMemcpyDeviceToDevice, e->s));
  HIPCHK(hipMemcpyAsync(e->d_ts, S.ts, sizeof(InstT) * e->max_inst, hipMemcpyDeviceToDevice, e->s));
  if (S.pk)
    HIPCHK(hipMemcpyAsync(e->pk, S.pk, (size_t)TL_PK_COUNT * np * sizeof(double), hipMemcpyDeviceToDevice, e->s));
  e->steps_done = S.steps;
  e->since_sort = S.since_sort;
  for (size_t k = 0; k < e->inst.size() && k < S.flags.size(); k++) {
    e->inst[k].p_in_p2 = S.flags[k].first;
    e->inst[k].xx_pending = S.flags[k].second;
  }
}

static void apply_tuning(mpm_engine* e, const mpm_tuning& t) {
  auto ok = [](int b) { return b == 64 || b == 128 || b == 256; };
  if (!ok(t.block_particles) || !ok(t.block_nodes) || t.sort_every < 0 ||
      !(t.sort_brick == 0 || t.sort_brick == 1 || t.sort_brick == 2 || t.sort_brick == 4 || t.sort_brick == 8))
    fail(MPM_ERR_ARG, "bad tuning values");
  e->bs_p = t.block_particles;
  e->bs_n = t.block_nodes;
  e->sort_every = t.sort_every;
  e->brick = t.sort_brick;
  e->ilv = t.interleave != 0;
  if (t.block_scatter) {
    if (!ok(t.block_scatter)) fail(MPM_ERR_ARG, "bad tuning values");
    e->bs_s = t.block_scatter;
  }
  if (t.tile_nodes) {
    if (t.tile_nodes < 8 || t.tile_nodes > 512) fail(MPM_ERR_ARG, "bad tuning values");
    e->tile_cap = t.tile_nodes;
  }
  mark_graph_dirty(e);
}

static mpm_tuning current_tuning(const mpm_engine* e) {
  mpm_tuning t{};
  t.block_particles = e->bs_p;
  t.block_nodes = e->bs_n;
  t.sort_every = e->sort_every;
  t.sort_brick = e->brick;
  t.interleave = e->ilv;
  t.block_scatter = e->bs_s;
  t.tile_nodes = e->tile_cap;
  return t;
}

// mean milliseconds of one step over n steps, no re-sort (plain launches:
// the trials would otherwise capture and instantiate a graph per shape, and
// the launch overhead is not what is being compared)
static double time_trial(mpm_engine* e, int n, std::vector<float>* each = nullptr) {
  const int se = e->sort_every;
  e->sort_every = 0;
  step_plain(e, 1, FIN_DONE);   // warm-up
  std::vector<hipEvent_t> ev(each ? n + 1 : 2);
  for (auto& x : ev) HIPCHK(hipEventCreate(&x));
  HIPCHK(hipEventRecord(ev[0], e->s));
  for (int k = 0; k < n; k++) {
    step_plain(e, 1, FIN_DONE);
    if (each) HIPCHK(hipEventRecord(ev[k + 1], e->s));
  }
  if (!each) HIPCHK(hipEventRecord(ev[1], e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev[0], ev.back()));
  if (each) {
    each->resize(n);
    for (int k = 0; k < n; k++) HIPCHK(hipEventElapsedTime(&(*each)[k], ev[k], ev[k + 1]));
  }
  for (auto& x : ev) HIPCHK(hipEventDestroy(x));
  e->sort_every = se;
  return ms / n;
}

static double time_sort(mpm_engine* e) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, e->s));
  sort_particles(e);
  HIPCHK(hipEventRecord(b, e->s));
  HIPCHK(hipStreamSynchronize(e->s));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  HIPCHK(hipEventDestroy(a));
  HIPCHK(hipEventDestroy(b));
  return ms;
}

// Coordinate search over the launch shape on the resident workload:
// (1) sort order (Hilbert, 2^3 or 4^3 bricks) x particle block size x lane
// mapping, timed right
// after a re-sort; (2) node block size; (3) block size of the
// particle-to-grid scatters; (4) re-sort period, from the
// measured cost of one sort and the measured growth of the step time as
// particles drift from their sorted order: the period P minimising
// (sort + sum of the first P step times) / P.  Every trial starts from a
// snapshot of the state, which is restored at the end.
static mpm_tuning autotune(mpm_engine* e, int trial_steps) {
  rebuild_lists(e);
  if (e->n_active == 0) fail(MPM_ERR_ARG, "autotune: no instances");
  if (trial_steps < 1) trial_steps = 1;
  Snapshot S;
  mpm_tuning best = current_tuning(e);
  const bool was_prof = e->prof;
  e->prof = false;
  int trials = 0;
  try {
    snap_take(e, S);
    double best_ms = 1e300;
    auto trial = [&](const mpm_tuning& t) {
      snap_restore(e, S);
      apply_tuning(e, t);
      sort_particles(e);
      const double ms = time_trial(e, trial_steps);
      trials++;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    };
    mpm_tuning t = best;
    t.block_nodes = 256;
    for (int B : {0, 2, 4})
      for (int bs : {64, 128, 256})
        for (int il : {0, 1}) {
          t.sort_brick = B;
          t.block_particles = bs;
          t.interleave = il;
          trial(t);
        }
    t = best;
    for (int bn : {64, 128}) {
      t.block_nodes = bn;
      trial(t);
    }
    t = best;
    const int bsc0 = best.block_scatter;   // already timed
    for (int bsc : {64, 128, 256}) {   // particle-to-grid scatters: bigger blocks flush fewer tile nodes
      if (bsc == bsc0) continue;
      t.block_scatter = bsc;
      trial(t);
    }
    // re-sort period
    snap_restore(e, S);
    apply_tuning(e, best);
    const double sort_ms = time_sort(e);
    const int K = 64;
    std::vector<float> each;
    time_trial(e, K, &each);
    double acc = sort_ms, best_cost = 1e300;
    int best_p = K;
    for (int P = 1; P <= K; P++) {
      acc += each[P - 1];
      if ((P & (P - 1)) == 0 || P == 10 || P == 20 || P == 40) {   // 1,2,4,8,10,16,20,32,40,64
        const double c = acc / P;
        if (c < best_cost) {
          best_cost = c;
          best_p = P;
        }
      }
    }
    best.sort_every = best_p;
    best.step_ms = best_ms;
    best.sort_ms = sort_ms;
    best.trials = trials + 1;
    snap_restore(e, S);
    apply_tuning(e, best);
    // the restored particles are in the order of the last sort before
    // tuning: re-sort in the chosen order now (this also starts a full
    // re-sort period and clears the node-box cache of the trials)
    sort_particles(e);
    HIPCHK(hipStreamSynchronize(e->s));
  } catch (...) {
    e->prof = was_prof;
    (void)hipStreamSynchronize(e->s);
    snap_free(S);
    throw;
  }
  e->prof = was_prof;
  snap_free(S);
  return best;
}

// --------------------------------------------------------- downloads

// XX is ParticlePositionUpdate's scratch (update_step.f90:1378-1526), which
// skips particles outside the grid (this step's cell index < 0): one that
// has been outside since the start never had it set (the reference reads 0
// there), one that left the grid keeps the position it moved to, its Xp.
// Particles outside the grid do not move, so "outside and still at Xo" is
// "never set".  x, xo: the particle's X and Xo fields, stride apart.
static bool xx_never_set(int icell, const double* x, const double* xo, long stride) {
  return icell < 0 && x[0] == xo[0] && x[stride] == xo[stride] && x[2 * stride] == xo[2 * stride];
}

template <typename T>
static void download_particles_T(mpm_engine* e, int id, double* out) {
  HostInst& h = e->inst[id];
  const long np = e->np, pad = h.npg * PGRAN, b0 = h.pg0 * PGRAN;
  const int n = h.n;
  std::vector<T> tb((size_t)T_COUNT * pad);
  std::vector<double> db((size_t)D_COUNT * pad);
  std::vector<int> ib((size_t)I_COUNT * pad);
  HIPCHK(hipStreamSynchronize(e->s));
  for (int f = 0; f < T_COUNT; f++)
    HIPCHK(hipMemcpy(&tb[(size_t)f * pad], (char*)e->t + ((size_t)f * np + b0) * sizeof(T), pad * sizeof(T),
                     hipMemcpyDeviceToHost));
  for (int f = 0; f < D_COUNT; f++)
    HIPCHK(hipMemcpy(&db[(size_t)f * pad], e->d + (size_t)f * np + b0, pad * sizeof(double), hipMemcpyDeviceToHost));
  for (int f = 0; f < I_COUNT; f++)
    HIPCHK(hipMemcpy(&ib[(size_t)f * pad], e->ii + (size_t)f * np + b0, pad * sizeof(int), hipMemcpyDeviceToHost));
  InstT ts;
  HIPCHK(hipMemcpy(&ts, e->d_ts + id, sizeof ts, hipMemcpyDeviceToHost));
  std::vector<double> pkb;   // PK: total-Lagrangian instances only (0 otherwise, as the reference keeps it)
  if (e->pk && (h.c.flags & IF_TL)) {
    pkb.resize((size_t)TL_PK_COUNT * pad);
    for (int k = 0; k < TL_PK_COUNT; k++)
      HIPCHK(hipMemcpy(&pkb[(size_t)k * pad], e->pk + (size_t)k * np + b0, pad * sizeof(double),
                       hipMemcpyDeviceToHost));
  }
  for (long j = 0; j < n; j++) {
    const int o = ib[(size_t)I_ORIG * pad + j];
    auto O = [&](int f) -> double& { return out[(size_t)f * n + o]; };
    auto Tv = [&](int f) { return (double)tb[(size_t)f * pad + j]; };
    O(PF_X) = db[(size_t)D_X * pad + j];
    O(PF_Y) = db[(size_t)D_Y * pad + j];
    O(PF_Z) = db[(size_t)D_Z * pad + j];
    O(PF_VX) = Tv(T_VX); O(PF_VY) = Tv(T_VY); O(PF_VZ) = Tv(T_VZ);
    O(PF_FX) = Tv(T_FX); O(PF_FY) = Tv(T_FY); O(PF_FZ) = Tv(T_FZ);
    O(PF_MASS) = db[(size_t)D_MASS * pad + j]; O(PF_SM) = Tv(T_SM); O(PF_SEQV) = Tv(T_SEQV);
    O(PF_SDXX) = Tv(T_SD0); O(PF_SDYY) = Tv(T_SD1); O(PF_SDZZ) = Tv(T_SD2);
    O(PF_SDYZ) = Tv(T_SD3); O(PF_SDXZ) = Tv(T_SD4); O(PF_SDXY) = Tv(T_SD5);
    O(PF_SIGY) = Tv(T_SIGY); O(PF_EPEFF) = Tv(T_EPEFF); O(PF_Q) = Tv(T_Q); O(PF_CP) = Tv(T_CP);
    for (int k = 0; k < 9; k++) O(PF_FP + k) = db[(size_t)(D_F0 + k) * pad + j];
    O(PF_VOL) = db[(size_t)D_VOL * pad + j];
    O(PF_IE) = db[(size_t)D_IE * pad + j];
    O(PF_TEMP) = db[(size_t)D_TEMP * pad + j];
    O(PF_DMG) = db[(size_t)D_DMG * pad + j];
    O(PF_LT) = db[(size_t)D_LT * pad + j];
    O(PF_FAIL) = db[(size_t)D_FAIL * pad + j];
    O(PF_SPRE) = db[(size_t)D_SPRE * pad + j];
    for (int k = 0; k < 3; k++) O(PF_XO + k) = db[(size_t)(D_XO + k) * pad + j];
    O(PF_VOL0) = db[(size_t)D_VOL0 * pad + j];
    for (int k = 0; k < TL_PK_COUNT; k++) O(PF_PK + k) = pkb.empty() ? 0.0 : pkb[(size_t)k * pad + j];
    const int icell = ib[(size_t)I_CELL * pad + j];
    const bool never = xx_never_set(icell, &db[(size_t)D_X * pad + j], &db[(size_t)D_XO * pad + j], pad);
    for (int k = 0; k < 3; k++) {
      const double x = db[(size_t)(D_X + k) * pad + j];
      O(PF_XX + k) = never ? 0.0 : (h.xx_pending && icell >= 0) ? x + Tv(T_BX + k) * ts.dt : x;
    }
  }
}

static void download_particles(mpm_engine* e, int id, double* out) {
  if (e->prec == MPM_PREC_FP32) download_particles_T<float>(e, id, out);
  else download_particles_T<double>(e, id, out);
}

// PF_COUNT x k doubles (field-major) of the particles with 1-based original
// indices ids[0..k): one gather kernel and one small copy, not the whole pool
template <typename T>
static void download_subset_T(mpm_engine* e, int id, int k, const int32_t* ids, double* out) {
  HostInst& h = e->inst[id];
  std::vector<int> u(k);
  for (int q = 0; q < k; q++) {
    if (ids[q] < 1 || ids[q] > h.n) fail(MPM_ERR_ARG, "particle %d outside 1..%d", ids[q], h.n);
    u[q] = ids[q] - 1;
  }
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  const int nu = (int)u.size();
  constexpr int NF = T_COUNT + D_COUNT + I_COUNT;
  HIPCHK(hipStreamSynchronize(e->s));
  if (nu > e->sub_cap) {
    if (e->d_sub_ids) HIPCHK(hipFree(e->d_sub_ids));
    if (e->d_sub_rows) HIPCHK(hipFree(e->d_sub_rows));
    e->d_sub_ids = nullptr;
    e->d_sub_rows = nullptr;
    e->sub_cap = 0;
    e->d_sub_ids = dalloc<int>(nu);
    e->d_sub_rows = dalloc<double>((size_t)nu * NF);
    e->sub_cap = nu;
  }
  HIPCHK(hipMemcpy(e->d_sub_ids, u.data(), sizeof(int) * nu, hipMemcpyHostToDevice));
  mpmk::launch_pack_subset<T>(e->ptrs<T>(), id, h.n, e->d_sub_ids, nu, e->d_sub_rows, e->s);
  HIPCHK(hipGetLastError());
  std::vector<double> rows((size_t)nu * NF);
  HIPCHK(hipStreamSynchronize(e->s));
  HIPCHK(hipMemcpy(rows.data(), e->d_sub_rows, sizeof(double) * rows.size(), hipMemcpyDeviceToHost));
  InstT ts;
  HIPCHK(hipMemcpy(&ts, e->d_ts + id, sizeof ts, hipMemcpyDeviceToHost));
  for (int q = 0; q < k; q++) {
    const double* r = &rows[(size_t)(std::lower_bound(u.begin(), u.end(), ids[q] - 1) - u.begin()) * NF];
    auto O = [&](int f) -> double& { return out[(size_t)f * k + q]; };
    auto Tv = [&](int f) { return r[f]; };
    auto Dv = [&](int f) { return r[T_COUNT + f]; };
    O(PF_X) = Dv(D_X); O(PF_Y) = Dv(D_Y); O(PF_Z) = Dv(D_Z);
    O(PF_VX) = Tv(T_VX); O(PF_VY) = Tv(T_VY); O(PF_VZ) = Tv(T_VZ);
    O(PF_FX) = Tv(T_FX); O(PF_FY) = Tv(T_FY); O(PF_FZ) = Tv(T_FZ);
    O(PF_MASS) = Dv(D_MASS); O(PF_SM) = Tv(T_SM); O(PF_SEQV) = Tv(T_SEQV);
    O(PF_SDXX) = Tv(T_SD0); O(PF_SDYY) = Tv(T_SD1); O(PF_SDZZ) = Tv(T_SD2);
    O(PF_SDYZ) = Tv(T_SD3); O(PF_SDXZ) = Tv(T_SD4); O(PF_SDXY) = Tv(T_SD5);
    O(PF_SIGY) = Tv(T_SIGY); O(PF_EPEFF) = Tv(T_EPEFF); O(PF_Q) = Tv(T_Q); O(PF_CP) = Tv(T_CP);
    for (int c = 0; c < 9; c++) O(PF_FP + c) = Dv(D_F0 + c);
    O(PF_VOL) = Dv(D_VOL);
    O(PF_IE) = Dv(D_IE);
    O(PF_TEMP) = Dv(D_TEMP);
    O(PF_DMG) = Dv(D_DMG);
    O(PF_LT) = Dv(D_LT);
    O(PF_FAIL) = Dv(D_FAIL);
    O(PF_SPRE) = Dv(D_SPRE);
    for (int c = 0; c < 3; c++) O(PF_XO + c) = Dv(D_XO + c);
    O(PF_VOL0) = Dv(D_VOL0);
    for (int c = 0; c < TL_PK_COUNT; c++) O(PF_PK + c) = 0.0;   // not gathered (mpm_download_particles has it)
    const int icell = (int)r[T_COUNT + D_COUNT + I_CELL];
    const bool never = xx_never_set(icell, &r[T_COUNT + D_X], &r[T_COUNT + D_XO], 1);
    for (int c = 0; c < 3; c++)
      O(PF_XX + c) = never ? 0.0 : (h.xx_pending && icell >= 0) ? Dv(D_X + c) + Tv(T_BX + c) * ts.dt : Dv(D_X + c);
  }
}

template <typename T>
static void download_grid_T(mpm_engine* e, int id, double* out, int comp = 0) {
  HostInst& h = e->inst[id];
  if (comp < 0 || comp >= h.c.ncomp) fail(MPM_ERR_ARG, "grid component %d outside 0..%d", comp, h.c.ncomp - 1);
  const long nn = h.c.nnode, n0 = h.ng0 * NGRAN + (long)comp * nn;
  settle_integrate<T>(e, id);
  HIPCHK(hipStreamSynchronize(e->s));
  std::vector<double> m(nn);
  std::vector<T> v(nn);
  HIPCHK(hipMemcpy(m.data(), e->mg + n0, nn * sizeof(double), hipMemcpyDeviceToHost));
  for (long k = 0; k < nn; k++) out[k] = m[k];
  const T* psrc = (const T*)(h.p_in_p2 ? e->p2 : e->pg);
  for (int c = 0; c < 3; c++) {
    HIPCHK(hipMemcpy(v.data(), psrc + (size_t)c * e->nn + n0, nn * sizeof(T), hipMemcpyDeviceToHost));
    for (long k = 0; k < nn; k++) out[(size_t)(GF_PX + c) * nn + k] = (double)v[k];
    HIPCHK(hipMemcpy(v.data(), (const T*)e->fg + (size_t)c * e->nn + n0, nn * sizeof(T), hipMemcpyDeviceToHost));
    for (long k = 0; k < nn; k++) out[(size_t)(GF_FX + c) * nn + k] = (double)v[k];
  }
}

static void download_grid(mpm_engine* e, int id, double* out, int comp = 0) {
  if (e->prec == MPM_PREC_FP32) download_grid_T<float>(e, id, out, comp);
  else download_grid_T<double>(e, id, out, comp);
}

static void download_icell(mpm_engine* e, int id, int32_t* out) {
  HostInst& h = e->inst[id];
  const long pad = h.npg * PGRAN, b0 = h.pg0 * PGRAN;
  std::vector<int> c(pad), o(pad);
  HIPCHK(hipStreamSynchronize(e->s));
  HIPCHK(hipMemcpy(c.data(), e->ii + (size_t)I_CELL * e->np + b0, pad * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(o.data(), e->ii + (size_t)I_ORIG * e->np + b0, pad * sizeof(int), hipMemcpyDeviceToHost));
  for (int j = 0; j < h.n; j++) out[o[j]] = c[j];
}

static void get_scalars(mpm_engine* e, int id, double* out) {
  InstT ts;
  HIPCHK(hipStreamSynchronize(e->s));
  HIPCHK(hipMemcpy(&ts, e->d_ts + id, sizeof ts, hipMemcpyDeviceToHost));
  out[0] = ts.istep;
  out[1] = ts.time;
  out[2] = ts.dt;
  out[3] = ts.ek;
  out[4] = ts.ei;
  out[5] = ts.mom[0];
  out[6] = ts.mom[1];
  out[7] = ts.mom[2];
}

static mpm_engine* check_ctx(mpm_engine* e) {
  if (!e) fail(MPM_ERR_ARG, "null context");
  return e;
}

static int live_id(mpm_engine* e, int id) {
  if (id < 0 || id >= e->max_inst || !e->inst[id].live) fail(MPM_ERR_ARG, "no live instance %d", id);
  return id;
}

// calcEnergy of one freshly added instance (MPM3D.f90:42)
static void initial_energy(mpm_engine* e, int id) {
  const HostInst& h = e->inst[id];
  const int* d_il = e->d_seq + id;   // the one-instance list {id}
  LaunchP L;
  L.glist = e->d_seq + h.pg0;
  L.ginst = e->d_ginst;
  L.bs = e->bs_p;
  L.bpg = PGRAN / e->bs_p;
  L.nblk = (int)h.npg * L.bpg;
  if (e->prec == MPM_PREC_FP32) {
    mpmk::launch_energy<float>(e->ptrs<float>(), L, e->s);
    mpmk::launch_finalize<float>(e->ptrs<float>(), d_il, 1, FIN_ENERGY, e->s);
  } else {
    mpmk::launch_energy<double>(e->ptrs<double>(), L, e->s);
    mpmk::launch_finalize<double>(e->ptrs<double>(), d_il, 1, FIN_ENERGY, e->s);
  }
  HIPCHK(hipGetLastError());
}

// ================================================================ C-ABI

static void numeric_flags(mpm_engine* e, int n, const int* ids, int* negvol, int* nonfinite);

struct mpm_ctx {
  mpm_engine* e;
};

extern "C" int mpm_ctx_create(const mpm_problem* prob, int precision, mpm_ctx** out) {
  return guarded([&] {
    if (!out) fail(MPM_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (precision != MPM_PREC_FP64 && precision != MPM_PREC_FP32 && precision != MPM_PREC_FP64_EXACT)
      fail(MPM_ERR_ARG, "bad precision");
    check_problem(prob);
    const long nnode = (long)(prob->ncx + 1) * (prob->ncy + 1) * (prob->ncz + 1) * (prob->contact ? prob->nb_component : 1);
    mpm_engine* e = new_engine(prob->nb_particle, nnode, 1, prob->nb_mat, prob->nb_body, precision);
    // a problem below ~1M particles fills the GPU with fewer 256-particle
    // blocks than it needs to hide latency: 128 (what the autotuner picks
    // for the 1.3e5-particle Taylor bar; 256 for the 1.3e6-particle slope)
    if (prob->nb_particle < (1 << 20)) e->bs_p = 128;
    // the auxiliary-grid scatter of SGMP runs 2.2x faster in Hilbert order
    // than in 2^3 bricks (measured on the Cook-plate batch, which the
    // autotuner also sorts in Hilbert order)
    if (prob->shape == MPM_SHAPE_SGMP) e->brick = 0;
    try {
      add_instance(e, prob, 0);
      rebuild_lists(e);
      initial_energy(e, 0);   // MPM3D.f90:42 calcEnergy after the input phase
    } catch (...) {
      free_engine(e);
      throw;
    }
    *out = new mpm_ctx{e};
  });
}

extern "C" int mpm_ctx_destroy(mpm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->e) (void)hipStreamSynchronize(ctx->e->s);
    free_engine(ctx->e);
    delete ctx;
  });
}

#define CTX_PHASE(fname, ph)                      \
  extern "C" int fname(mpm_ctx* ctx) {            \
    return guarded([&] {                          \
      if (!ctx) fail(MPM_ERR_ARG, "null context"); \
      phase(ctx->e, ph);                          \
    });                                           \
  }
CTX_PHASE(mpm_begin_step, PH_BEGIN)
CTX_PHASE(mpm_grid_momentum_initial, PH_GMI)
CTX_PHASE(mpm_apply_boundary_conditions, PH_BC)
CTX_PHASE(mpm_particle_stress_update, PH_STRESS)
CTX_PHASE(mpm_grid_momentum_update, PH_GMU)
CTX_PHASE(mpm_integrate_momentum, PH_INTEGRATE)
CTX_PHASE(mpm_particle_position_update, PH_G2P)
CTX_PHASE(mpm_grid_momentum_musl, PH_MUSL)
CTX_PHASE(mpm_smooth_stress_by_grid, PH_SMOOTH)
CTX_PHASE(mpm_lagr_nod_contact, PH_CONTACT)
#undef CTX_PHASE

extern "C" int mpm_calc_energy(mpm_ctx* ctx, double* ek, double* ei, double* mom3) {
  return guarded([&] {
    if (!ctx) fail(MPM_ERR_ARG, "null context");
    phase(ctx->e, PH_ENERGY);
    double s[8];
    get_scalars(ctx->e, 0, s);
    if (ek) *ek = s[3];
    if (ei) *ei = s[4];
    if (mom3) { mom3[0] = s[5]; mom3[1] = s[6]; mom3[2] = s[7]; }
  });
}

extern "C" int mpm_update_dt(mpm_ctx* ctx, double* dt) {
  return guarded([&] {
    if (!ctx) fail(MPM_ERR_ARG, "null context");
    phase(ctx->e, PH_DT);
    double s[8];
    get_scalars(ctx->e, 0, s);
    if (dt) *dt = s[2];
    const int id = 0;
    int nonfinite = 0;
    numeric_flags(ctx->e, 1, &id, nullptr, &nonfinite);
    if (nonfinite > 0)
      fail(MPM_ERR_NUMERIC, "%d particle stress updates gave a non-finite volume or stress (DT was still computed)",
           nonfinite);
  });
}

extern "C" int mpm_numeric_flags(mpm_ctx* ctx, int* negvol, int* nonfinite) {
  return guarded([&] {
    if (!ctx) fail(MPM_ERR_ARG, "null context");
    const int id = 0;
    numeric_flags(ctx->e, 1, &id, negvol, nonfinite);
  });
}

extern "C" int mpm_step(mpm_ctx* ctx, int nsteps) {
  return guarded([&] {
    if (!ctx || nsteps < 0) fail(MPM_ERR_ARG, "bad arguments");
    step_plain(ctx->e, nsteps, 0);
  });
}

extern "C" int mpm_download_particles(mpm_ctx* ctx, double* ps) {
  return guarded([&] {
    if (!ctx || !ps) fail(MPM_ERR_ARG, "bad arguments");
    download_particles(ctx->e, 0, ps);
  });
}

extern "C" int mpm_download_particle_subset(mpm_ctx* ctx, int k, const int32_t* ids, double* out) {
  return guarded([&] {
    if (!ctx || k < 0 || (k > 0 && (!ids || !out))) fail(MPM_ERR_ARG, "bad arguments");
    if (k == 0) return;
    if (ctx->e->prec == MPM_PREC_FP32) download_subset_T<float>(ctx->e, 0, k, ids, out);
    else download_subset_T<double>(ctx->e, 0, k, ids, out);
  });
}

extern "C" int mpm_upload_particles(mpm_ctx* ctx, const double* ps) {
  return guarded([&] {
    if (!ctx || !ps) fail(MPM_ERR_ARG, "bad arguments");
    HIPCHK(hipStreamSynchronize(ctx->e->s));
    upload_particles(ctx->e, ctx->e->inst[0], ps, ctx->e->inst[0].n, false);
  });
}

extern "C" int mpm_download_grid(mpm_ctx* ctx, double* g) {
  return guarded([&] {
    if (!ctx || !g) fail(MPM_ERR_ARG, "bad arguments");
    download_grid(ctx->e, 0, g);
  });
}

extern "C" int mpm_download_grid_component(mpm_ctx* ctx, int comp, double* g) {
  return guarded([&] {
    if (!ctx || !g) fail(MPM_ERR_ARG, "bad arguments");
    download_grid(ctx->e, 0, g, comp);
  });
}

extern "C" int mpm_get_contact_force(mpm_ctx* ctx, double* out3) {
  return guarded([&] {
    if (!ctx || !out3) fail(MPM_ERR_ARG, "bad arguments");
    mpm_engine* e = ctx->e;
    InstT ts;
    HIPCHK(hipStreamSynchronize(e->s));
    HIPCHK(hipMemcpy(&ts, e->d_ts, sizeof ts, hipMemcpyDeviceToHost));
    for (int d = 0; d < 3; d++) out3[d] = e->inst[0].c.contact == 1 ? ts.cont_force[d] : 0.0;
  });
}

extern "C" int mpm_download_icell(mpm_ctx* ctx, int32_t* c) {
  return guarded([&] {
    if (!ctx || !c) fail(MPM_ERR_ARG, "bad arguments");
    download_icell(ctx->e, 0, c);
  });
}

extern "C" int mpm_get_scalars(mpm_ctx* ctx, double* out8) {
  return guarded([&] {
    if (!ctx || !out8) fail(MPM_ERR_ARG, "bad arguments");
    get_scalars(ctx->e, 0, out8);
  });
}

extern "C" int mpm_nb_node(const mpm_ctx* ctx) { return ctx ? ctx->e->inst[0].c.nnode : -1; }

extern "C" int mpm_sync(mpm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) fail(MPM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(ctx->e->s));
  });
}

// ---------------------------------------------------- engine (batched) API

extern "C" int mpm_engine_create(long max_particles, long max_nodes, int max_instances, int max_materials,
                                 int max_bodies, int precision, mpm_engine** out) {
  return guarded([&] {
    if (!out || max_particles <= 0 || max_nodes <= 0 || max_instances <= 0 || max_materials <= 0 || max_bodies <= 0)
      fail(MPM_ERR_ARG, "bad arguments");
    if (precision != MPM_PREC_FP64 && precision != MPM_PREC_FP32 && precision != MPM_PREC_FP64_EXACT)
      fail(MPM_ERR_ARG, "bad precision");
    *out = new_engine(max_particles, max_nodes, max_instances, max_materials, max_bodies, precision);
  });
}

extern "C" int mpm_engine_create_budget(double hbm_bytes, int max_instances, int precision, mpm_engine** out) {
  return guarded([&] {
    if (!out || !(hbm_bytes > 0) || max_instances <= 0) fail(MPM_ERR_ARG, "bad arguments");
    // split the budget: particles with sort double-buffers (and 1.5 B of
    // node-box cache), nodes ~ particles/2
    const double bp = 2.0 * bytes_per_particle(precision) + 18, bn = bytes_per_node(precision);
    const long np = (long)(hbm_bytes / (bp + 0.5 * bn));
    const long nn = np / 2;
    // decks built from many particle generators have one body per generator
    // line (DataIn.f90:568-846): 32 B per body slot, so be generous
    *out = new_engine(np, nn, max_instances, 16 * max_instances, 2048 * max_instances, precision);
  });
}

extern "C" int mpm_engine_destroy(mpm_engine* e) {
  return guarded([&] {
    if (e) (void)hipStreamSynchronize(e->s);
    free_engine(e);
  });
}

extern "C" int mpm_engine_add(mpm_engine* e, const mpm_problem* p, int stop_at_endtime, int* id) {
  return guarded([&] {
    check_ctx(e);
    if (!id) fail(MPM_ERR_ARG, "id is NULL");
    *id = add_instance(e, p, stop_at_endtime ? IF_ENDTIME : 0);
    e->lists_dirty = true;    // rebuilt before the next step
    if (e->sort_every > 0) sort_instance(e, *id);
    initial_energy(e, *id);   // MPM3D.f90:42
  });
}

extern "C" int mpm_engine_set_step_limit(mpm_engine* e, int id, int max_steps) {
  return guarded([&] {
    check_ctx(e);
    HostInst& h = e->inst[live_id(e, id)];
    if (max_steps < 0) fail(MPM_ERR_ARG, "max_steps < 0");
    h.c.max_steps = max_steps;
    // ordered behind queued work (kernels read the constants at run time,
    // so a captured graph stays valid)
    int slot;
    char* st = stage_get(e, sizeof(InstC), slot);
    memcpy(st, &h.c, sizeof(InstC));
    HIPCHK(hipMemcpyAsync(e->d_inst + id, st, sizeof(InstC), hipMemcpyHostToDevice, e->s));
    stage_done(e, slot);
  });
}

extern "C" int mpm_engine_remove(mpm_engine* e, int id) {
  return guarded([&] {
    check_ctx(e);
    remove_instance(e, id);
  });
}

extern "C" int mpm_engine_step(mpm_engine* e, int nsteps, int use_graph, int graph_steps) {
  return guarded([&] {
    check_ctx(e);
    if (nsteps < 0) fail(MPM_ERR_ARG, "nsteps < 0");
    if (use_graph < 0 || use_graph > 2) fail(MPM_ERR_ARG, "use_graph must be 0, 1 or 2");
    if (use_graph) step_graph(e, nsteps, graph_steps, FIN_DONE, use_graph == 2);
    else step_plain(e, nsteps, FIN_DONE);
  });
}

extern "C" int mpm_engine_sync(mpm_engine* e) {
  return guarded([&] {
    check_ctx(e);
    HIPCHK(hipStreamSynchronize(e->s));
  });
}

extern "C" int mpm_engine_sort(mpm_engine* e) {
  return guarded([&] {
    check_ctx(e);
    sort_particles(e);
  });
}

extern "C" int mpm_engine_counters(mpm_engine* e, long* steps_done, long* sorts_done) {
  return guarded([&] {
    if (!e) fail(MPM_ERR_ARG, "null engine");
    if (steps_done) *steps_done = e->steps_done;
    if (sorts_done) *sorts_done = e->sorts_done;
  });
}

extern "C" int mpm_engine_set_interleave(mpm_engine* e, int enable) {
  return guarded([&] {
    check_ctx(e);
    e->ilv = enable != 0;
    mark_graph_dirty(e);
  });
}

extern "C" int mpm_engine_set_tuning(mpm_engine* e, int block_particles, int block_nodes, int sort_every) {
  return guarded([&] {
    check_ctx(e);
    auto ok = [](int b) { return b == 64 || b == 128 || b == 256; };
    if (!ok(block_particles) || !ok(block_nodes) || sort_every < 0) fail(MPM_ERR_ARG, "bad tuning values");
    e->bs_p = block_particles;
    e->bs_n = block_nodes;
    e->sort_every = sort_every;
    mark_graph_dirty(e);
  });
}

extern "C" int mpm_engine_get_tuning(mpm_engine* e, mpm_tuning* out) {
  return guarded([&] {
    check_ctx(e);
    if (!out) fail(MPM_ERR_ARG, "out is NULL");
    *out = current_tuning(e);
  });
}

extern "C" int mpm_engine_apply_tuning(mpm_engine* e, const mpm_tuning* t) {
  return guarded([&] {
    check_ctx(e);
    if (!t) fail(MPM_ERR_ARG, "tuning is NULL");
    apply_tuning(e, *t);
  });
}

extern "C" int mpm_engine_autotune(mpm_engine* e, int trial_steps, mpm_tuning* best) {
  return guarded([&] {
    check_ctx(e);
    const mpm_tuning t = autotune(e, trial_steps);
    if (best) *best = t;
  });
}

extern "C" int mpm_engine_download_particles(mpm_engine* e, int id, double* ps) {
  return guarded([&] {
    check_ctx(e);
    download_particles(e, live_id(e, id), ps);
  });
}

extern "C" int mpm_engine_download_grid(mpm_engine* e, int id, double* g) {
  return guarded([&] {
    check_ctx(e);
    download_grid(e, live_id(e, id), g);
  });
}

extern "C" int mpm_engine_download_icell(mpm_engine* e, int id, int32_t* c) {
  return guarded([&] {
    check_ctx(e);
    download_icell(e, live_id(e, id), c);
  });
}

extern "C" int mpm_engine_profile(mpm_engine* e, int enable) {
  return guarded([&] {
    check_ctx(e);
    collect_times(e);
    e->prof = enable != 0;
    for (int k = 0; k < KI_COUNT; k++) {
      e->kt_ms[k] = 0;
      e->kt_n[k] = 0;
    }
  });
}

extern "C" int mpm_engine_kernel_times(mpm_engine* e, double* total_ms8, long* launches8) {
  return guarded([&] {
    check_ctx(e);
    collect_times(e);
    for (int k = 0; k < KI_COUNT; k++) {
      if (total_ms8) total_ms8[k] = e->kt_ms[k];
      if (launches8) launches8[k] = e->kt_n[k];
    }
  });
}

extern "C" const char* mpm_engine_kernel_name(int kid) { return (kid >= 0 && kid < KI_COUNT) ? KI_NAMES[kid] : ""; }

extern "C" int mpm_engine_download_order(mpm_engine* e, int id, int32_t* orig) {
  return guarded([&] {
    check_ctx(e);
    const HostInst& h = e->inst[live_id(e, id)];
    HIPCHK(hipStreamSynchronize(e->s));
    HIPCHK(hipMemcpy(orig, e->ii + (size_t)I_ORIG * e->np + h.pg0 * PGRAN, h.n * sizeof(int), hipMemcpyDeviceToHost));
  });
}

extern "C" int mpm_engine_get_scalars(mpm_engine* e, int id, double* out8) {
  return guarded([&] {
    check_ctx(e);
    get_scalars(e, live_id(e, id), out8);
  });
}

extern "C" int mpm_engine_status(mpm_engine* e, int n, const int* ids, int* done, int* istep) {
  return guarded([&] {
    check_ctx(e);
    HIPCHK(hipStreamSynchronize(e->s));
    std::vector<InstT> ts(e->max_inst);
    HIPCHK(hipMemcpy(ts.data(), e->d_ts, sizeof(InstT) * e->max_inst, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++) {
      int id = ids[k];
      if (id < 0 || id >= e->max_inst || !e->inst[id].live) {
        done[k] = -1;
        if (istep) istep[k] = -1;
        continue;
      }
      done[k] = ts[id].done;
      if (istep) istep[k] = ts[id].istep;
    }
  });
}

// per-instance counts of particles with a negative or non-finite state
static void numeric_flags(mpm_engine* e, int n, const int* ids, int* negvol, int* nonfinite) {
  HIPCHK(hipStreamSynchronize(e->s));
  std::vector<InstT> ts(e->max_inst);
  HIPCHK(hipMemcpy(ts.data(), e->d_ts, sizeof(InstT) * e->max_inst, hipMemcpyDeviceToHost));
  for (int k = 0; k < n; k++) {
    const int id = ids[k];
    const bool ok = id >= 0 && id < e->max_inst && e->inst[id].live;
    if (negvol) negvol[k] = ok ? ts[id].negvol_count : -1;
    if (nonfinite) nonfinite[k] = ok ? ts[id].nan_count : -1;
  }
}

extern "C" int mpm_engine_numeric_flags(mpm_engine* e, int n, const int* ids, int* negvol, int* nonfinite) {
  return guarded([&] {
    check_ctx(e);
    if (n < 0 || (n > 0 && !ids)) fail(MPM_ERR_ARG, "bad arguments");
    numeric_flags(e, n, ids, negvol, nonfinite);
  });
}

extern "C" int mpm_engine_snapshot_status(mpm_engine* e) {
  return guarded([&] {
    check_ctx(e);
    if (!e->h_snap) {
      HIPCHK(hipHostMalloc((void**)&e->h_snap, sizeof(InstT) * e->max_inst, hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&e->snap_ev, hipEventDisableTiming));
    }
    if (e->snap_pending) HIPCHK(hipEventSynchronize(e->snap_ev));   // the buffer is being filled
    HIPCHK(hipMemcpyAsync(e->h_snap, e->d_ts, sizeof(InstT) * e->max_inst, hipMemcpyDeviceToHost, e->s));
    HIPCHK(hipEventRecord(e->snap_ev, e->s));
    e->snap_pending = true;
  });
}

extern "C" int mpm_engine_snapshot_read(mpm_engine* e, int n, const int* ids, int* done, double* out8) {
  return guarded([&] {
    check_ctx(e);
    if (!e->h_snap || !e->snap_pending) fail(MPM_ERR_STATE, "no status snapshot taken");
    HIPCHK(hipEventSynchronize(e->snap_ev));
    for (int k = 0; k < n; k++) {
      const int id = ids[k];
      const bool ok = id >= 0 && id < e->max_inst && e->inst[id].live;
      done[k] = ok ? e->h_snap[id].done : -1;
      if (out8) {
        double* o = out8 + 8 * k;
        const InstT& t = e->h_snap[ok ? id : 0];
        o[0] = ok ? t.istep : -1;
        o[1] = t.time;
        o[2] = t.dt;
        o[3] = t.ek;
        o[4] = t.ei;
        o[5] = t.mom[0];
        o[6] = t.mom[1];
        o[7] = t.mom[2];
      }
    }
  });
}

extern "C" int mpm_engine_capacity(mpm_engine* e, long* free_particles, long* free_nodes, int* free_instances) {
  return guarded([&] {
    check_ctx(e);
    if (free_particles) *free_particles = e->pa.largest() * PGRAN;
    if (free_nodes) *free_nodes = e->na.largest() * NGRAN;
    if (free_instances) *free_instances = (int)e->free_ids.size();
  });
}

extern "C" long mpm_engine_device_bytes(mpm_engine* e) {
  if (!e) return -1;
  return (long)(e->np * (2 * bytes_per_particle(e->prec) + 16) + e->np / 64 * BOX_INTS * 4 + e->nn * bytes_per_node(e->prec) +
                (e->aux ? e->nn * (AUX_COUNT * e->tsz + 8) : 0));
}

extern "C" int mpm_ctx_engine(mpm_ctx* ctx, mpm_engine** out) {
  return guarded([&] {
    if (!ctx || !out) fail(MPM_ERR_ARG, "bad arguments");
    *out = ctx->e;
  });
}

// ------------------------------------------------ particles by cell (C-ABI)

// the grid constants InWhichCell reads, from a problem's grid description
static InstC cell_grid(const mpm_problem* p) {
  if (!p || p->ncx <= 0 || p->ncy <= 0 || p->ncz <= 0 || !(p->dcell > 0))
    fail(MPM_ERR_ARG, "bad grid description");
  InstC c{};
  c.ncx = p->ncx;
  c.ncy = p->ncy;
  c.ncz = p->ncz;
  c.ncxy = c.ncx * c.ncy;
  c.ncell = c.ncxy * c.ncz;
  for (int k = 0; k < 3; k++) {
    c.span0[k] = p->span[2 * k];
    c.span1[k] = p->span[2 * k + 1];
  }
  c.dcell = p->dcell;
  return c;
}

// scratch words of one cell-order call over n particles: two key buffers,
// one value buffer (the other is the caller's order array) and the radix
// sort's counts
static size_t cell_order_words(long n) { return 3 * (size_t)n + mpmk::radix_scratch_words(n); }

static void check_cell_order_args(const mpm_problem* p, const double* x, const double* y, const double* z, long n,
                                  const int32_t* icell, const int32_t* order) {
  if (!p || n < 0) fail(MPM_ERR_ARG, "bad arguments");
  if (n > 0 && (!x || !y || !z || !icell || !order)) fail(MPM_ERR_ARG, "null device array");
  if (n > (1L << 30)) fail(MPM_ERR_ARG, "at most 2^30 particles per call");
}

// InWhichCell keys, then the stable radix sort of (key, index); the value
// buffers are arranged so that the last pass writes the caller's order array
static void cell_order_run(const InstC& g, const double* x, const double* y, const double* z, long n, int32_t* icell,
                           int32_t* order, unsigned* scr, hipStream_t s) {
  unsigned* k0 = scr;
  unsigned* k1 = scr + n;
  int* vt = (int*)(scr + 2 * n);
  unsigned* rs = scr + 3 * n;
  const int nbits = bits_for((long)g.ncell + 2);
  const int odd = ((nbits + 7) / 8) % 2;   // passes of 8 bits: the result lands in v1 when odd
  int* v0 = odd ? vt : order;
  int* v1 = odd ? order : vt;
  mpmk::launch_cell_keys(g, x, y, z, n, icell, k0, v0, s);
  const int par = mpmk::radix_sort_pairs<unsigned>(k0, k1, v0, v1, n, 0, nbits, rs, s);
  if (par != odd) fail(MPM_ERR_STATE, "cell order: unexpected sort parity");
  HIPCHK(hipGetLastError());
}

// per-device scratch of mpm_cell_order, grown on demand and kept
struct CellOrderScratch {
  unsigned* p = nullptr;
  size_t words = 0;
};
static std::mutex g_cell_order_mu;
static std::map<int, CellOrderScratch> g_cell_order_scr;

extern "C" int mpm_cell_order(const mpm_problem* prob, const double* x, const double* y, const double* z, long n,
                              int32_t* icell, int32_t* order, void* stream) {
  return guarded([&] {
    check_cell_order_args(prob, x, y, z, n, icell, order);
    const InstC g = cell_grid(prob);
    if (n == 0) return;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_cell_order_mu);
    CellOrderScratch& sc = g_cell_order_scr[dev];
    const size_t need = cell_order_words(n);
    if (sc.words < need) {
      if (sc.p) {
        HIPCHK(hipDeviceSynchronize());   // earlier calls may still use it
        HIPCHK(hipFree(sc.p));
        sc.p = nullptr;
        sc.words = 0;
      }
      sc.p = dalloc<unsigned>(need);
      sc.words = need;
    }
    cell_order_run(g, x, y, z, n, icell, order, sc.p, (hipStream_t)stream);
  });
}

// Scratch of mpm_cell_order_noalloc: one fixed device array, no allocation
// at call time.  Every GPU the library runs on holds its own copy (at its
// own address) until the process ends.  Sized for MPM_CELL_ORDER_MAX
// particles: two key buffers and one value buffer of that many entries and
// the radix sort's counts (256 per 4096-particle tile, plus 256 totals).
constexpr long CO_MAX = MPM_CELL_ORDER_MAX;
constexpr size_t CO_WORDS = 3 * (size_t)CO_MAX + 256 * (size_t)((CO_MAX + 4095) / 4096) + 256;
static __device__ __align__(256) unsigned scratch[CO_WORDS];

// the device address of `scratch` on the current device, looked up once
// per device (the host-side name is a stand-in, not a device address)
static std::mutex g_scratch_mu;
static std::map<int, unsigned*> g_scratch_ptr;

static unsigned* scratch_ptr() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || !p) fail(MPM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(rc));
  g_scratch_ptr[dev] = (unsigned*)p;
  return (unsigned*)p;
}

extern "C" int mpm_cell_order_noalloc(const mpm_problem* prob, const double* x, const double* y, const double* z,
                                      long n, int32_t* icell, int32_t* order, void* stream) {
  return guarded([&] {
    check_cell_order_args(prob, x, y, z, n, icell, order);
    if (n > CO_MAX) fail(MPM_ERR_ARG, "mpm_cell_order_noalloc: %ld particles exceed its maximum of %ld", n, CO_MAX);
    const InstC g = cell_grid(prob);
    if (n == 0) return;
    if (cell_order_words(n) > CO_WORDS) fail(MPM_ERR_STATE, "cell order scratch size");
    cell_order_run(g, x, y, z, n, icell, order, scratch_ptr(), (hipStream_t)stream);
  });
}
