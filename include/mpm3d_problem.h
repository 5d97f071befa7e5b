/*
 * mpm3d_problem.h -- plain-C description of one MPM3D problem instance as it
 * stands after the reference's input phase (InputPara: GetData, Initial,
 * SetGridData, SetDT; DataIn.f90:34-103).
 *
 * The same struct is produced by the .mpm deck reader (mpm_deck_read), by the
 * synthetic deck builders, and consumed by the device context
 * (mpm_ctx_create) and by the CPU oracle.  Everything is host memory, FP64,
 * structure-of-arrays.
 */
#ifndef MPM3D_PROBLEM_H
#define MPM3D_PROBLEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material model ids, identical to the reference's MatType values
 * (DataIn.f90:406-507, Constitution.f90:204-347). */
enum mpm_mat_type {
  MPM_MAT_ELASTIC = 1,   /* elas  -> M3DM1 */
  MPM_MAT_PLA1 = 2,      /* pla1  -> M3DM2, elastic-perfectly plastic */
  MPM_MAT_PLA2 = 3,      /* pla2  -> M3DM3, J2 with linear hardening */
  MPM_MAT_JOHN = 4,      /* john  -> M3DM4, Johnson-Cook with heating (reference-order FP64 mode only) */
  MPM_MAT_SJC = 5,       /* sjc   -> M3DM5 + EOS, simplified Johnson-Cook */
  MPM_MAT_SJCF = 6,      /* sjcf  -> M3DM5 + EOS with failure strain */
  MPM_MAT_JCF = 7,       /* jcf   -> M3DM8 + EOS with damage (reference-order FP64 mode only) */
  MPM_MAT_HIEX = 8,      /* hiex  -> M3DM6 + EOS (JWL), high explosive burn */
  MPM_MAT_NULL = 9,      /* null  -> M3DM7 + EOS */
  MPM_MAT_DP = 10,       /* dpm   -> M3DM9, Drucker-Prager */
  MPM_MAT_NEOH = 11      /* neoh  -> M3DM10, neo-Hookean */
};

/* Shape function family (Particle.f90:91-93 flags GIMP / Bspline / SGMP). */
enum mpm_shape { MPM_SHAPE_LINEAR = 0, MPM_SHAPE_GIMP = 1, MPM_SHAPE_BSPLINE = 2, MPM_SHAPE_SGMP = 3 };

/* Update scheme (Particle.f90:88-90; MUSL is the default, DataIn.f90:364). */
enum mpm_scheme { MPM_USF = 0, MPM_USL = 1, MPM_MUSL = 2 };

typedef struct mpm_material {
  int32_t type;           /* enum mpm_mat_type */
  int32_t eos_type;       /* 0 none; seos: 1 linear polynomial, 2 Mie-Gruneisen, 3 JWL, 4 weakly compressible */
  double density;
  double young, poisson;
  double yield0, tangmod;
  double miu;             /* dynamic viscosity (null) */
  double bulk_modulus;    /* null */
  double wavespd;         /* null: sqrt(K/rho) */
  double q_fai, k_fai, q_psi, ten_f;   /* Drucker-Prager */
  double ceos[10];        /* cEos(1..10) as SetEos stores them (Gruneisen: C1..C5 derived; E0 in ceos[9]) */
  double det_speed;       /* D: detonation velocity (hiex; SetDT/UpdateDT take it as a sound speed) */
  double b_jc, n_jc, c_jc, m_jc;   /* Johnson-Cook B, n, C, m */
  double epso;            /* reference strain rate */
  double epf;             /* failure strain (sjcf, jcf) */
  double roomt, melt, spec_heat;   /* john, jcf: room and melting temperature, specific heat */
} mpm_material;

typedef struct mpm_body {
  int32_t mat;            /* 1-based material id */
  int32_t com;            /* component id */
  int32_t par_begin;      /* 0-based first particle */
  int32_t par_end;        /* 0-based one past last particle */
  double grav[3];
} mpm_body;

/* Particle state fields of mpm_problem.pstate, each an array of
 * nb_particle doubles.  Deviatoric stress uses the reference's Voigt order
 * (SDxx, SDyy, SDzz, SDyz, SDxz, SDxy) = sd(1..6), Constitution.f90:184-189.
 * Fp stores F - I row-major: PF_FP + 3*i + j is Fp(i+1, j+1). */
enum mpm_pfield {
  PF_X = 0, PF_Y, PF_Z,
  PF_VX, PF_VY, PF_VZ,
  PF_FX, PF_FY, PF_FZ,          /* prescribed particle load FXp */
  PF_MASS, PF_VOL,
  PF_SM, PF_SEQV,
  PF_SDXX, PF_SDYY, PF_SDZZ, PF_SDYZ, PF_SDXZ, PF_SDXY,
  PF_SIGY, PF_EPEFF, PF_IE, PF_Q, PF_CP,
  PF_FP,                        /* 9 entries */
  PF_XX = PF_FP + 9, PF_XY, PF_XZ,   /* XX: position at t+1 */
  PF_TEMP,                      /* celsius_t: temperature (293 at start, Particle.f90:131) */
  PF_DMG,                       /* damage */
  PF_LT,                        /* lighting time of high explosive (Initial, DataIn.f90:1252-1260) */
  PF_FAIL,                      /* failure flag, 0 or 1 */
  PF_SPRE,                      /* Spre: smoothed pressure of SmoothStressbyGrid (update_step.f90:1876-2055) */
  PF_XO, PF_YO, PF_ZO,          /* Xo: position at t = 0 (DataIn.f90:859-863), TLMPM's reference configuration */
  PF_VOL0,                      /* VOL0: initial volume (Initial, DataIn.f90:1240) */
  PF_PK,                        /* 9 entries: TLMPM's first Piola-Kirchhoff stress PK(1..9), the reference's */
                                /* order xx yy zz yz xz xy zy zx yx (Constitution.f90:806-815) */
  PF_COUNT = PF_PK + 9
};

/* Grid node fields of a grid state array (each nb_node doubles). */
enum mpm_gfield { GF_MG = 0, GF_PX, GF_PY, GF_PZ, GF_FX, GF_FY, GF_FZ, GF_COUNT };

typedef struct mpm_problem {
  /* sizes */
  int32_t nb_particle;
  int32_t nb_body;
  int32_t nb_mat;
  int32_t shape;          /* enum mpm_shape */
  int32_t scheme;         /* enum mpm_scheme */
  int32_t quasi_load;     /* QuLo */
  int32_t quasi_damp;     /* Damp */
  int32_t gravity;        /* any 'grav' load given */
  int32_t update_v_by_fp; /* UVFp, forced by neoh */
  int32_t fix[6];         /* FixS(1..6): x0 xn y0 yn z0 zn */
  /* grid, after SetGridData (Grid.f90:250-263) */
  int32_t ncx, ncy, ncz;  /* NumCellx/y/z */
  int32_t dr_damping;     /* drda: dynamic relaxation damping after UpdateDT (Particle.f90:172-202) */
  double span[6];         /* SpanX(1), SpanX(2), SpanY(1), ... (upper ends adjusted) */
  double dcell;
  double cutoff;          /* grid mass cut-off (DataIn.f90:517-524) */
  double dt;              /* initial DT (SetDT, DataIn.f90:1269-1297) */
  double dtscale;
  double endtime;
  double quasi_load_time;
  double damp_factor;
  double bq1, bq2;        /* artificial bulk viscosity */
  /* arrays (host memory) */
  mpm_material* mats;     /* nb_mat */
  mpm_body* bodies;       /* nb_body */
  double* pstate;         /* PF_COUNT * nb_particle, field-major */
  int32_t* pbody;         /* nb_particle, 0-based body id */
  /* appended after the fields above (whose layout stays as it was) */
  int32_t smooth_stress;  /* smoo: SmoothStressbyGrid after UpdateDT every step (MPM3D.f90:114) */
  int32_t tlmpm;          /* TLmp: total-Lagrangian step sequence (MPM3D.f90:166-250) */
  int32_t contact;        /* cont lagr: Contact_type 1, Lagr_NodContact after IntegrateMomentum */
  int32_t normbody;       /* contact normal: 0 from both bodies, 1 from component 1, 2 from component 2 */
  int32_t nb_component;   /* nbco: grid components (bodies carry their component in com) */
  int32_t reserved0;
  double fricfa;          /* contact friction coefficient */
} mpm_problem;

#ifdef __cplusplus
}
#endif
#endif /* MPM3D_PROBLEM_H */
