/*
 * sph_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the reference
 * SPH particle loops (Artemisaturn/GRAPHICS, Source/).  It is the checker for
 * the HIP path and, in its FP32 instantiation, the timed CPU baseline
 * ("port": the reference's own OpenMP code cannot be built, its external
 * headers are not vendored and several of its translation units are stale).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links it.
 */
#ifndef SPH_ORACLE_H_
#define SPH_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_WCSPH = 0, ORC_DFSPH = 1, ORC_IISPH = 2, ORC_PCISPH = 3, ORC_IISPH_DF = 4 };

/* Mirrors ThresholdsBase m_TH (SphBase.h:546-583) plus the solver settings
 * that the reference hard-codes (DFSPH.cpp:225-230, 762-766; IISph_M.cpp:285,
 * 341; WcSph.h:34). */
typedef struct orc_params {
    int   solver;              /* ORC_* */
    float spacing_r;           /* SphBase.h:548 */
    float smooth_radius_h;     /* SphBase.h:549, cell size and 2x kernel h */
    float space_min[3];        /* SphBase.h:559 */
    float space_max[3];        /* SphBase.h:560 */
    float dt;                  /* fixed dt, or dt_max when cfl > 0 */
    float gravity[3];
    float sound_speed;         /* m_TH.soundSpeed_cs (WCSPH, multiphase viscosity) */
    float gamma;               /* WcSph m_DensityPower_gamma (WcSph.h:34) */
    float nu;                  /* m_TH.nu: DFSPH/IISPH viscosity (DFSPH.cpp:1265) */
    float cfl;                 /* 0: fixed dt;  >0: dt = min(cfl*r/max|v|, dt) */
    float dt_min;
    float density_tol;         /* DFSPH: relative (DFSPH.cpp:225); IISPH: eta/rho0 (IISph_M.cpp:285) */
    float divergence_tol;      /* DFSPH: absolute kg m^-3 s^-1 (DFSPH.cpp:762) */
    int   min_iter_density;    /* DFSPH.cpp:229 (2); IISph_M.cpp:286 (2) */
    int   max_iter_density;    /* none in the reference (DFSPH.cpp:228 unused) */
    int   min_iter_div;        /* DFSPH.cpp:765 (1) */
    int   max_iter_div;        /* DFSPH.cpp:764 (30, loop breaks when it > 30) */
    float omega;               /* IISPH relaxed Jacobi (IISph_M.cpp:341) */
    int   warm_start;          /* 0 none; 1 divergence loop warm-started from the previous step */
    float warm_scale;          /* scale of the previous step's stiffness in the guess */
    float boundary_density;    /* m_TH.boundaryDensity (SphBase.h:576), IISph DF_* solvers */
    int   adjust_density;      /* m_TH.adjustDensity (SphBase.h:575), IISph DF_* solvers */
    int   multiphase_consistent; /* IISPH, several phases: 1 = the solve's d_ij, d_ji follow the
                                  number-density force between phases (m_j^2/m_i instead of m_i) */
    int   reserved[2];
} orc_params;

typedef struct orc_stats {
    int    iter_density;       /* DFSPH density / IISPH pressure iterations */
    int    iter_divergence;    /* DFSPH divergence iterations */
    double density_err;        /* last solver density error (relative) */
    double divergence_err;     /* last divergence error (absolute) */
    double rho_mean_err;       /* mean(rho/rho0 - 1) over fluid */
    double rho_max_err;        /* max(rho/rho0 - 1) */
    double kinetic_energy;     /* sum 1/2 m |v|^2 */
    float  dt;                 /* dt used by this step */
    int    zero_dis;           /* m_EC.zeroDis increments this step */
    int    too_many_neigb;     /* m_EC.tooManyNeigb increments this step */
    int    removed;            /* particles removed outside [spaceMin,spaceMax] */
    int    max_iter_hit;       /* a capped solver loop hit its cap */
    int    nan_count;          /* non-finite fluid positions/velocities */
} orc_stats;

/* fluid field selectors for orc_get_fluid */
enum {
    ORC_F_POS = 0,      /* 3 */
    ORC_F_VEL = 1,      /* 3 */
    ORC_F_RHO = 2,      /* 1  density of the last density pass */
    ORC_F_PRESSURE = 3, /* 1  WCSPH/IISPH pressure */
    ORC_F_ACC = 4,      /* 3  WCSPH acceleration, IISPH pressure acceleration */
    ORC_F_ALPHA = 5,    /* 1  DFSPH alpha factor */
    ORC_F_KAPPA = 6,    /* 1  DFSPH last kappa */
    ORC_F_ID = 7,       /* 1  original id (as double) */
    ORC_F_AADV = 8,     /* 3  non-pressure acceleration */
    ORC_F_KAPPA_SUM = 9,/* 1  accumulated kappa*dt^2 of the density solve */
    ORC_F_AII = 10,     /* 1  IISPH a_ii */
    ORC_F_RHO_ADV = 11, /* 1  IISPH rho_adv; DFSPH rho* of the last density iteration */
    ORC_F_KAPPAV = 12,  /* 1  DFSPH kappaV of the last divergence iteration */
    ORC_F_DIVDEV = 13,  /* 1  DFSPH clamped Drho/Dt of the last divergence iteration */
    ORC_F_VADV = 14     /* 3  DFSPH v* / IISPH v_adv */
};

typedef struct orc_sim orc_sim;

orc_sim* orc_create(int precision_bits, const orc_params* p);
void     orc_destroy(orc_sim* s);
int      orc_add_fluid(orc_sim* s, int n, const float* xyz, const float* vel,
                       const int* ids, float rest_density, float visc_alpha);
/* appends n particles (given ids) to fluid k between steps (an emitter in
 * stepEvent, SphWithIO::addFluidParts) */
int      orc_append_fluid(orc_sim* s, int k, int n, const float* xyz, const float* vel, const int* ids);
/* per-particle rest-density factor beta of fluid k (IISph DF_* solvers,
 * m = rho0 beta d^3, IISph.cpp:110-115); default 1 */
int      orc_set_fluid_beta(orc_sim* s, int k, int n, const float* beta);
int      orc_add_solid(orc_sim* s, int n, const float* xyz, int dynamic, float visc_alpha);
/* dynamic solid k: OpenGL column-major transform of its body-frame samples,
 * centre of mass, linear and angular velocity (SphSolid::updatePaticles) */
int      orc_set_solid_pose(orc_sim* s, int k, const float* T16, const float* com, const float* v,
                            const float* w);
/* SphSolid::rigidForceJoin of the last step's reaction forces */
int      orc_solid_force_torque(orc_sim* s, int k, double* f, double* t);
int      orc_init(orc_sim* s);                       /* prepareSolidPartWeight */
int      orc_step(orc_sim* s, orc_stats* st);        /* one sphStep + runOneStep bookkeeping */
/* the same with a host callback at the point of SphWithBullet::updateSolids
 * (WcSph.cpp:85, IISph_M.cpp:64, DFSPH.cpp:54); poses set with
 * orc_set_solid_pose take effect at that point (or at orc_init) */
int      orc_step_cb(orc_sim* s, orc_stats* st, int (*cb)(void*), void* user);
int      orc_step_dt(const orc_sim* s, double* dt);     /* dt of the current (or last) step */
int      orc_pci_delta(const orc_sim* s, double* delta); /* PCISPH delta (before / dt^2) */
int      orc_num_fluid(const orc_sim* s, int k);
/* Overwrites the positions (and, with vel != NULL, the velocities) of fluid
 * k's particles in their current order, at the oracle's precision (no FP32
 * rounding in the 64-bit oracle): the hook of the finite-difference pins. */
int      orc_set_fluid_state(orc_sim* s, int k, const double* xyz, const double* vel);
int      orc_get_fluid(const orc_sim* s, int k, int field, double* out);
int      orc_get_boundary_volume(const orc_sim* s, int k, double* out);
/* IISph DF_* solvers: boundary particle fields of solid k, field 0 density,
 * 1 alpha, 2 kappa, 3 predicted density of the last density iteration,
 * 4 kappaV; returns the count */
int      orc_get_boundary_field(const orc_sim* s, int k, int field, double* out);
/* Runs gridSetup + neighbourSearch on the current state and dumps the fluid
 * neighbour lists: per fluid particle (current order) up to `cap` entries of
 * (kind 0 fluid / 1 boundary, id, dis).  Returns total overflow count. */
int      orc_neighbours(orc_sim* s, int k, int cap, int* counts, int* kinds,
                        int* ids, float* dis);
/* Runs neighbourSearch + updateSolidPartWeight + DF_constantDensitySolverFirst
 * on the current state (density and alpha fields), without stepping. */
int      orc_compute_density(orc_sim* s);
/* Evaluates the cubic spline W(r) and |dW/dr| (SphBase.h:855-906) with kernel
 * h at n radii, in the given precision. */
void     orc_kernel_eval(int precision_bits, double h, int n, const double* r, double* W, double* dW);
void     orc_set_threads(int n);
int      orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
