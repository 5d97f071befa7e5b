/*
 * mpm3d_hip.h -- C-ABI of libmpm3d_hip.so, the MI355X (gfx950) engine for
 * the MPM3D explicit updated-Lagrangian time step.
 *
 * The reference calls its step as argument-less external subroutines over
 * module globals (update_step.f90, called from MPM3D.f90:69-113).  Each of
 * those routines has one entry point here taking an explicit context; the
 * context owns device-resident structure-of-arrays particle and grid state
 * for the whole run.  A Fortran host binds these through ISO_C_BINDING
 * (mpm3d-f90_amd/fortran/mpm_hip.f90).
 *
 * Conventions: every function returns 0 (MPM_OK) or a negative MPM_ERR_*
 * code; mpm_last_error() returns a message for the calling thread's last
 * failure.  No C++ exceptions cross this boundary.  All calls on one context
 * are ordered on that context's HIP stream; functions that return host data
 * synchronise that stream.  A context may be driven by one host thread at a
 * time; distinct contexts are independent.
 */
#ifndef MPM3D_HIP_H
#define MPM3D_HIP_H

#include <stdint.h>
#include "mpm3d_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_ERR_ARG (-1)          /* bad argument */
#define MPM_ERR_IO (-2)           /* file could not be opened */
#define MPM_ERR_DECK (-3)         /* deck syntax / semantics */
#define MPM_ERR_UNSUPPORTED (-4)  /* feature outside the explicit UL step */
#define MPM_ERR_HIP (-5)          /* HIP runtime error */
#define MPM_ERR_NOMEM (-6)        /* device memory / HBM budget exhausted */
#define MPM_ERR_STATE (-7)        /* call out of order / context unusable */
#define MPM_ERR_NUMERIC (-8)      /* device flagged NaN / negative volume */

/* arithmetic of the bulk particle/grid state */
#define MPM_PREC_FP64 0           /* plain FP64 everywhere */
#define MPM_PREC_FP32 1           /* FP32 bulk, FP64 positions and small-mass division */
/* plain FP64 with every grid sum taken in the reference's order (ascending
 * particle index per node, update_step.f90:335-392, 946-988, 1705-1758) and
 * every gather in its node-list order: bit-identical to the reference and
 * reproducible from run to run (no floating-point atomics) */
#define MPM_PREC_FP64_EXACT 2

typedef struct mpm_ctx mpm_ctx;

const char* mpm_last_error(void);
const char* mpm_version(void);

/* ---- input deck (replaces InputPara, DataIn.f90:34-103) ---------------- */
/* Reads a .mpm deck; *out is allocated by the library (mpm_problem_free). */
int mpm_deck_read(const char* path, mpm_problem** out, char* errbuf, int errlen);
void mpm_problem_free(mpm_problem* p);

/* Output options of a deck (keywords mpm3 (title), endt/outt, rptt, tecp,
 * para, curv, curx, outr; DataIn.f90:130-330, defaults of DataOut.f90:39-59),
 * for hosts that write the reference's result files themselves
 * (mpm3d-f90_amd/output.py); the drop-in runs the reference's own writers. */
#define MPM_MAX_CURVES 15
#define MPM_MAX_ANIM 10
typedef struct mpm_outputs {
  char title[257];          /* the deck's first line as FFI's ReadLine leaves it: 256 characters, blank-padded */
  double out_time;          /* OutTime: outt, else EndTime */
  double report_time;       /* ReportTime (rptt) */
  int32_t tecplot;          /* tecp */
  int32_t paraview;         /* para */
  int32_t n_curves;         /* 2 defaults (engk, engi of particle 1), then curv / curx lines */
  int32_t curve_option[MPM_MAX_CURVES];   /* result option, 1-based: seqv epef mat pres volu engk engi */
  int32_t curve_point[MPM_MAX_CURVES];    /*   velx vely velz cels fail sspd damg stsx disx disy disz; particle, 1-based */
  int32_t n_anim;           /* 1 default (mat), then outr lines */
  int32_t anim_option[MPM_MAX_ANIM];
  int32_t plot2d;           /* pt2d given: TecPlot frames leave out particles outside plot2d_box */
  int32_t skip_first;       /*   that were generated after it (0-based index of the first such particle) */
  double plot2d_box[6];     /*   x0 x1 y0 y1 z0 z1 (a default REAL array in the reference) */
} mpm_outputs;
int mpm_deck_read_outputs(const char* path, mpm_outputs* out, char* errbuf, int errlen);

/* ---- context -------------------------------------------------------------
 * Uploads the problem (particles as SoA, grid description, materials,
 * flags) to the current HIP device.  The device keeps DT, time and step
 * count; the host copy of the problem may be freed afterwards.  Contact
 * ('cont lagr'), total-Lagrangian ('TLmp') and Johnson-Cook ('john', 'jcf')
 * problems run in MPM_PREC_FP64_EXACT only; the other modes return
 * MPM_ERR_UNSUPPORTED for them. */
int mpm_ctx_create(const mpm_problem* prob, int precision, mpm_ctx** out);
int mpm_ctx_destroy(mpm_ctx* ctx);

/* ---- one entry point per reference phase ------------------------------- */
/* istep += 1; CurrentTime += DT (MPM3D.f90:72-74) */
int mpm_begin_step(mpm_ctx* ctx);
/* GridMomentumInitial + ParticleInitial (update_step.f90:288-453) */
int mpm_grid_momentum_initial(mpm_ctx* ctx);
/* ApplyBoundaryConditions (update_step.f90:1795-1830) */
int mpm_apply_boundary_conditions(mpm_ctx* ctx);
/* ParticleStressUpdate (update_step.f90:456-647, Constitution.f90:79-368);
 * with john/jcf materials it also resolves the plastic strain increment each
 * particle reads from the one before it (Constitution.f90:237, 290, 314) */
int mpm_particle_stress_update(mpm_ctx* ctx);
/* GridMomentumUpdate (update_step.f90:831-1013) */
int mpm_grid_momentum_update(mpm_ctx* ctx);
/* IntegrateMomentum (update_step.f90:1194-1239) */
int mpm_integrate_momentum(mpm_ctx* ctx);
/* Lagr_NodContact (update_step.f90:1241-1376, called after IntegrateMomentum
 * when the deck has 'cont lagr', MPM3D.f90:94-96): contact forces between
 * grid components 1 and 2; a no-op without contact.  Contact problems run in
 * MPM_PREC_FP64_EXACT (mpm_ctx_create refuses them in the other modes). */
int mpm_lagr_nod_contact(mpm_ctx* ctx);
/* ParticlePositionUpdate (update_step.f90:1378-1526) */
int mpm_particle_position_update(mpm_ctx* ctx);
/* GridMomentumMUSL (update_step.f90:1668-1793) */
int mpm_grid_momentum_musl(mpm_ctx* ctx);
/* SmoothStressbyGrid (update_step.f90:1876-2055, called after UpdateDT when
 * the deck has 'smoo', MPM3D.f90:114): the smoothed pressure Spre (PF_SPRE)
 * of every particle; a no-op for a problem without smooth_stress.  mpm_step
 * runs it inside each step of such a problem. */
int mpm_smooth_stress_by_grid(mpm_ctx* ctx);
/* calcEnergy (Particle.f90:204-230); any output pointer may be NULL */
int mpm_calc_energy(mpm_ctx* ctx, double* eng_kinetic, double* eng_internal, double* momentum3);
/* UpdateDT (update_step.f90:1833-1874); *dt may be NULL.  Returns
 * MPM_ERR_NUMERIC (DT still written) once a stress update has produced a
 * non-finite volume or stress. */
int mpm_update_dt(mpm_ctx* ctx, double* dt);
/* Particles whose stress update gave a negative volume (the reference warns
 * and goes on, Constitution.f90:135-137) and a non-finite volume or stress,
 * counted over all steps so far (synchronising; either pointer may be NULL).
 * mpm_step does not synchronise: its flags are read here or by the next
 * mpm_update_dt. */
int mpm_numeric_flags(mpm_ctx* ctx, int* negvol, int* nonfinite);

/* nsteps whole steps in the reference order (MPM3D.f90:69-113), fused
 * kernels, no host synchronisation inside */
int mpm_step(mpm_ctx* ctx, int nsteps);

/* ---- particles by cell -------------------------------------------------
 * The cell of every particle as ParticleInitial finds it (InWhichCell,
 * update_step.f90:408-431, Grid.f90:123-155: 1-based, -1 outside the grid)
 * and the stable order of the particles by cell (ascending cell, equal
 * cells in ascending index, particles outside the grid last) from the
 * engine's radix sort, which reorders particle arrays for locality where
 * the reference walks them in index order.  x, y, z (FP64), icell and order
 * are device arrays of n entries owned by the caller; the grid is prob's
 * (span, dcell, ncx, ncy, ncz; its particles are not read).  Enqueued on
 * stream (a hipStream_t; NULL: the null stream); returns when enqueued.
 * Scratch of 12 B per particle plus the sort's counts is held per device by
 * the library, grown on demand (hipMalloc) and kept; calls on one device
 * share it, so calls that overlap in time on different streams race. */
int mpm_cell_order(const mpm_problem* prob, const double* x, const double* y, const double* z, long n,
                   int32_t* icell, int32_t* order, void* stream);

/* The same computation without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is one fixed device array of the library (about 25 MiB, held by
 * every GPU the library runs on until the process ends), so n is at most
 * MPM_CELL_ORDER_MAX (larger n: MPM_ERR_ARG).  The caller passes inputs and
 * outputs as above.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU race (order them on one stream). */
#define MPM_CELL_ORDER_MAX (1L << 21)
int mpm_cell_order_noalloc(const mpm_problem* prob, const double* x, const double* y, const double* z, long n,
                           int32_t* icell, int32_t* order, void* stream);

/* ---- state exchange (synchronising) ---------------------------------- */
/* pstate: PF_COUNT * nb_particle doubles, original particle order */
int mpm_download_particles(mpm_ctx* ctx, double* pstate);
/* the state of k particles chosen by their 1-based original index (as the
 * reference's CurvePoint): out holds PF_COUNT * k doubles, field-major
 * (out[f * k + q] for ids[q]); a gather on the device and one small copy,
 * for per-step curve output without downloading every particle */
int mpm_download_particle_subset(mpm_ctx* ctx, int k, const int32_t* ids, double* out);
/* gstate: GF_COUNT * nb_node doubles, node order of the reference */
int mpm_download_grid(mpm_ctx* ctx, double* gstate);
/* the grid of component comp (0-based; 0 is what mpm_download_grid gives) of
 * a contact problem, same layout */
int mpm_download_grid_component(mpm_ctx* ctx, int comp, double* gstate);
/* tot_cont_for, the contact force of the last Lagr_NodContact summed over the
 * nodes (DataOut.f90:148-151 writes it every step); 0 without contact */
int mpm_get_contact_force(mpm_ctx* ctx, double* out3);
/* icell: nb_particle, 1-based cell index of the last GridMomentumInitial, -1 outside */
int mpm_download_icell(mpm_ctx* ctx, int32_t* icell);
/* overwrite particle state (e.g. a host-side restart); same layout as download */
int mpm_upload_particles(mpm_ctx* ctx, const double* pstate);
/* out[0..7] = istep, CurrentTime, DT, EngKinetic, EngInternal, Momentum(3) */
int mpm_get_scalars(mpm_ctx* ctx, double* out8);
int mpm_nb_node(const mpm_ctx* ctx);
int mpm_sync(mpm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* MPM3D_HIP_H */
