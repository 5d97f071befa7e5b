/*
 * mpm_oracle.h -- CPU FP64 restatement of the MPM3D explicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the device path is compared
 * against; it is never linked into the product library.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 *
 * It follows the reference statement by statement, in the reference's
 * summation and evaluation order (ascending particle index, reference node
 * order per shape function), so that it is bit-exact against the Fortran
 * build in oracle/_ref on the same deck.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>
#include "../include/mpm3d_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

oracle* oracle_create(const mpm_problem* p);
void oracle_destroy(oracle* o);
int oracle_nb_node(const oracle* o);
int oracle_nb_particle(const oracle* o);

/* MPM3D.f90:72-74 */
void oracle_begin_step(oracle* o);
/* update_step.f90 phases */
void oracle_grid_momentum_initial(oracle* o);     /* :288-406 (+ ParticleInitial :408-453) */
void oracle_apply_bc(oracle* o);                  /* :1795-1830 */
void oracle_particle_stress_update(oracle* o);    /* :456-647 */
void oracle_grid_momentum_update(oracle* o);      /* :831-1013 */
void oracle_integrate_momentum(oracle* o);        /* :1194-1239 */
void oracle_contact(oracle* o);                   /* :1241-1376 Lagr_NodContact */
void oracle_particle_position_update(oracle* o);  /* :1378-1526 */
void oracle_grid_momentum_musl(oracle* o);        /* :1668-1793 */
void oracle_calc_energy(oracle* o);               /* Particle.f90:204-230 */
void oracle_update_dt(oracle* o);                 /* update_step.f90:1833-1874 */
void oracle_smooth_stress(oracle* o);             /* :1876-2055 SmoothStressbyGrid (Spre); TL: :2057-2236 */
/* total-Lagrangian phases (TLmp) */
void oracle_tl_grid_mass_initial(oracle* o);      /* :22-172 GridMassInitial + TLParticleInitial, once */
void oracle_tl_grid_momentum_initial(oracle* o);  /* :174-286 */
void oracle_tl_particle_stress_update(oracle* o); /* :649-828, Constitution.f90:369-525, 782-882 */
void oracle_tl_grid_momentum_update(oracle* o);   /* :1015-1191 */
/* whole steps in the order of MPM3D.f90:69-113 */
void oracle_step(oracle* o, int nsteps);

void oracle_get_particles(const oracle* o, double* out);   /* PF_COUNT * nb_particle */
void oracle_get_icell(const oracle* o, int32_t* out);      /* nb_particle, 1-based, -1 outside */
void oracle_get_grid(const oracle* o, double* out);        /* GF_COUNT * nb_node, component 1 */
/* grid of component c (0-based) of a contact problem, GF_COUNT * nb_node */
void oracle_get_grid_component(const oracle* o, int c, double* out);
int oracle_nb_component(const oracle* o);
/* tot_cont_for of the last Lagr_NodContact (DataOut.f90:148-151) */
void oracle_get_contact_force(const oracle* o, double* out3);
/* diagnostics, not reference state: contact nodes over all steps so far,
 * frictionless / sticking (full tangential force) / slipping (capped) */
void oracle_get_contact_counts(const oracle* o, long* out3);
void oracle_get_fix(const oracle* o, int8_t* out);         /* 3 * nb_node */
/* out[0..7] = istep, CurrentTime, DT, EngKinetic, EngInternal, Momentum(3) */
void oracle_get_scalars(const oracle* o, double* out);
void oracle_set_particles(oracle* o, const double* in);

/* ln J of the neo-Hookean model (M3DM10): on = 0 the C library's log (the
 * default, as the reference's Fortran runtime), 1 the correctly rounded
 * logarithm that the device's reference-order mode uses; oracle_log_cr is
 * that logarithm (test entry) */
void oracle_set_log_cr(oracle* o, int on);
double oracle_log_cr(double x);
double oracle_exp_cr(double x);
double oracle_pow_cr(double x, double y);

/* Diagnostics of the Drucker-Prager return map (M3DM9, Constitution.f90:
 * 1263-1321), not reference state: the branch each particle took in its
 * last stress update, and running counts per branch over all updates. */
enum { DP_BR_ELASTIC = 0, DP_BR_SHEAR = 1, DP_BR_SHEAR_TENSILE = 2, DP_BR_TENSION = 3 };
void oracle_get_dp_branches(const oracle* o, int8_t* last, long* counts4);
/* Diagnostics of the Johnson-Cook models (Constitution.f90:224-237,
 * 278-319), not reference state: running counts of john/jcf updates that
 * read the depeff another particle assigned, and of jcf failures by melting,
 * by the failure strain and by damage. */
enum { JC_STALE = 0, JC_MELT = 1, JC_EPF = 2, JC_DMG = 3 };
void oracle_get_jc_counts(const oracle* o, long* counts4);
/* Test switch, not the reference: on = 1 starts every particle's update
 * from depeff = 0 instead of the last update's value. */
void oracle_set_fresh_depeff(oracle* o, int on);
/* M3DM9 on one stress point with a zero strain increment (trial = input):
 * returns the branch; sd[6], *sm updated, *seqv and *depeff written. */
/* FindInflNode's node list for the axis cases zc, yc, xc (-1, 0 or 2):
 * writes n (dx, dy, dz) offsets from the cell's first node, returns n. */
int oracle_gimp_offsets(int zc, int yc, int xc, int* off);
int oracle_dp_point(const mpm_material* mat, double young, double poisson, double sd[6], double* sm, double* seqv,
                    double* depeff);

#ifdef __cplusplus
}
#endif
#endif
