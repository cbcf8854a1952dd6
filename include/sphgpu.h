/*
 * sphgpu.h -- C-ABI of the MI355X SPH particle loops (libsphgpu.so).
 *
 * The library replaces the hot protected methods of the reference's solver
 * classes (Artemisaturn/GRAPHICS, Source/):
 *   SphBase::neighbourSearch / gridSetup        SphBase.h:673, SphBase.cpp:849-1089
 *   SphBase::prepare/updateSolidPartWeight      SphBase.h:477-481, SphBase.cpp:766-841
 *   SphBase::updateFluids                       SphBase.h:414, SphBase.cpp:377-402
 *   WcSph::wc_computePressure / wc_computeForce WcSph.h:72, 91
 *   DFSPH::DF_constantDensitySolverFirst/Second, DF_divergenceFreeSolver,
 *         DF_forceExceptPressure                DFSPH.h:58-68
 *   IISph_M::ii_computePressure, ii_forceExceptPressure,
 *         ii_computeForcePressure               IISph_M.h:56-90
 * and the per-step driver SphBase::runOneStep -> sphStep (SphBase.cpp:348-354).
 * A host class (graphics_amd/host/SphGpu.h) keeps the reference class names
 * and calls these entry points from its sphStep() override.
 *
 * Conventions: plain C types only; vec3 arrays are AoS float[3*n] like vec_t
 * (SphBase.h:41); the caller owns host buffers (copied in/out); the library
 * owns device memory.  Every entry point returns 0 on success and a negative
 * SPHGPU_E* code on error, with a message in sphgpu_last_error().  A handle
 * is not thread-safe (like SphBase, which is not re-entrant); it owns one HIP
 * stream.  No entry point here falls back to a CPU implementation: without a
 * usable gfx950 device the compute entry points return SPHGPU_E_DEVICE.
 */
#ifndef SPHGPU_H_
#define SPHGPU_H_

#ifdef __cplusplus
extern "C" {
#endif

#define SPHGPU_ABI_VERSION 4
#define SPHGPU_MAX_ITER 10000   /* largest accepted solver-loop cap */

enum sphgpu_status {
    SPHGPU_OK = 0,
    SPHGPU_E_ARG = -1,      /* bad argument (null pointer, bad index, bad size) */
    SPHGPU_E_STATE = -2,    /* call not valid in the current state */
    SPHGPU_E_DEVICE = -3,   /* HIP runtime / device error, or no device */
    SPHGPU_E_CAPACITY = -4, /* a fixed capacity would be exceeded */
    SPHGPU_E_DIVERGED = -5, /* non-finite particle state detected */
    SPHGPU_E_CALLBACK = -6  /* the update-solids callback returned nonzero */
};

enum sphgpu_solver {
    SPHGPU_WCSPH = 0,       /* WcSph::sphStep            WcSph.cpp:23-96   */
    SPHGPU_DFSPH = 1,       /* DFSPH::sphStep            DFSPH.cpp:9-75    */
    SPHGPU_IISPH = 2,       /* IISph_M::sphStep (Ihmsen)  IISph_M.cpp:10-75 */
    SPHGPU_PCISPH = 3,      /* PciSph::sphStep (SP09)     PciSph.cpp:31-95; eta = density_tol,
                               min/max iterations = min/max_iter_density (reference 0.01, 3, 10) */
    SPHGPU_IISPH_DF = 4     /* IISph class, DF_* solvers  IISph.cpp:100-288, 816-1217: per-particle mass
                               rho0 beta d^3 (sphgpu_set_fluid_beta), adjust_density, boundary_density,
                               boundary particles' own kappa; divergence-free solve, then the constant-
                               density solve (reference: density_tol 1e-4 relative, 3..100 iterations;
                               divergence_tol 1e-3 on mean(Drho/Dt) dt, 1..30); loops break when the
                               iteration count exceeds max_iter_*.  No warm start, no graph capture. */
};

enum sphgpu_warm_start {
    SPHGPU_WARM_NONE = 0,     /* plain solver (the reference) */
    SPHGPU_WARM_PREVIOUS = 1, /* previous step's accumulated kappa (Bender & Koschier) */
    SPHGPU_WARM_LEARNED = 2   /* per-particle MLP prediction (sphgpu_mlp_*) */
};

/* Mirrors ThresholdsBase m_TH (SphBase.h:546-583) plus the solver settings the
 * reference hard-codes.  Field meanings and defaults: see sphgpu_default_params. */
typedef struct sphgpu_params {
    int   solver;              /* sphgpu_solver */
    float spacing_r;           /* m_TH.spacing_r: particle spacing and diameter d */
    float smooth_radius_h;     /* m_TH.smoothRadius_h: cell size; kernel h = smooth_radius_h/2 */
    float space_min[3];        /* m_TH.spaceMin: scene 0 domain (particles outside are removed) */
    float space_max[3];        /* m_TH.spaceMax */
    float dt;                  /* m_TH.dt: fixed step, or the upper bound when cfl > 0 */
    float gravity[3];          /* m_TH.gravity_g */
    float sound_speed;         /* m_TH.soundSpeed_cs (WCSPH EOS and artificial viscosity) */
    float gamma;               /* WcSph::m_DensityPower_gamma (WcSph.h:34) */
    float nu;                  /* m_TH.nu: DFSPH/IISPH viscosity (DFSPH.cpp:1265) */
    float cfl;                 /* 0: fixed dt; >0: dt = clamp(cfl*spacing/max|v|, dt_min, dt) */
    float dt_min;
    float density_tol;         /* DFSPH avg (rho*-rho0)/rho0 (DFSPH.cpp:225); IISPH eta/rho0 (IISph_M.cpp:285) */
    float divergence_tol;      /* DFSPH avg Drho/Dt in kg m^-3 s^-1 (DFSPH.cpp:762) */
    int   min_iter_density;    /* 2 (DFSPH.cpp:229, IISph_M.cpp:286) */
    int   max_iter_density;    /* cap, 1..SPHGPU_MAX_ITER (the reference DFSPH loop has none) */
    int   min_iter_div;        /* 1 (DFSPH.cpp:765) */
    int   max_iter_div;        /* 30: the loop breaks when it > max (DFSPH.cpp:772); 0..SPHGPU_MAX_ITER */
    float omega;               /* IISPH relaxation, 0.5 (IISph_M.cpp:341) */
    int   warm_start;          /* sphgpu_warm_start */
    int   num_scenes;          /* >= 1 independent scenes advanced together */
    int   learn_steps;         /* learned warm start: steps of online training before use */
    int   learn_iters;         /* Adam iterations per training step */
    int   learn_batch;         /* samples per Adam iteration */
    float learn_rate;          /* Adam learning rate */
    float learn_over_weight;   /* loss weight of over-predictions (>= 1 biases the guess low) */
    float warm_scale;          /* PREVIOUS (and LEARNED while it trains): the guess is warm_scale x the
                                  previous stiffness; default 0.7 (Bender & Koschier use 0.5) */
    int   use_graph;           /* 1: WCSPH/IISPH steps are captured once into HIP graphs and replayed */
    float boundary_density;    /* m_TH.boundaryDensity (SphBase.h:576, 1000): IISph DF_* boundary terms and
                                  the boundary particles' rest density */
    int   adjust_density;      /* m_TH.adjustDensity (SphBase.h:575, 0): IISph DF_* number density */
    int   fluid_capacity;      /* fluid particles the handle can hold, counting later appends (emitters);
                                  0 = the particles added before init */
    float learn_scale;         /* LEARNED once trained, DFSPH divergence solve: the guess is learn_scale x
                                  the bounded prediction (the IISPH pressure start takes it whole) */
    int   multiphase_consistent; /* IISPH with several phases.  0 (default): IISph_M's pressure solve, which
                                  counts every fluid neighbour with the particle's own phase mass
                                  (IISph_M.cpp:262, 302, 329) while its pressure force between phases takes
                                  the number-density form (IISph_M.cpp:526-546).  1: the solve's d_ij and
                                  d_ji follow that force (d_ij = dt^2 m_j^2/(m_i rho_j^2) gradW), so the
                                  solved pressure is the one the force applies; same phase: no change */
} sphgpu_params;

typedef struct sphgpu_stats {
    int    iter_density;       /* DFSPH density / IISPH pressure iterations (max over scenes) */
    int    iter_divergence;    /* DFSPH divergence iterations (max over scenes) */
    double density_err;        /* last solver density error, relative (max over scenes) */
    double divergence_err;     /* last divergence error, kg m^-3 s^-1 (max over scenes) */
    double rho_mean_err;       /* mean(rho/rho0 - 1) over fluid */
    double rho_max_err;        /* max(rho/rho0 - 1) */
    double kinetic_energy;     /* sum 1/2 m |v|^2 */
    float  dt;                 /* dt of this step (scene 0) */
    int    zero_dis;           /* m_EC.zeroDis increments (WcSph.cpp:160) */
    int    too_many_neigb;     /* m_EC.tooManyNeigb: neighbours beyond the 48 of the reference (SphBase.h:630) */
    int    removed;            /* fluid particles removed outside [spaceMin,spaceMax] (SphBase.cpp:397) */
    int    max_iter_hit;       /* a capped solver loop reached its cap */
    int    nan_count;          /* non-finite fluid positions/velocities */
    long long num_fluid;       /* live fluid particles after the step */
    int    list_overflow;      /* neighbours dropped because a list was full (GPU capacity) */
    int    reserved[5];
} sphgpu_stats;

/* fluid fields for sphgpu_get_fluid (values in ascending original-id order) */
enum sphgpu_field {
    SPHGPU_F_POS = 0,          /* 3 floats  position */
    SPHGPU_F_VEL = 1,          /* 3  velocity */
    SPHGPU_F_RHO = 2,          /* 1  density of the last density pass */
    SPHGPU_F_PRESSURE = 3,     /* 1  WCSPH / IISPH pressure */
    SPHGPU_F_ACC = 4,          /* 3  WCSPH acceleration / IISPH pressure acceleration */
    SPHGPU_F_ALPHA = 5,        /* 1  DFSPH alpha factor */
    SPHGPU_F_KAPPA_SUM = 6,    /* 1  accumulated kappa*dt^2 of the last density solve */
    SPHGPU_F_PHASE = 7,        /* 1  fluid phase index (as float) */
    SPHGPU_F_KAPPAV_SUM = 8    /* 1  accumulated kappaV*dt of the last divergence solve */
};

typedef struct sphgpu_handle sphgpu_handle;

/* ---- lifecycle ---------------------------------------------------------- */
void        sphgpu_default_params(sphgpu_params* p);
int         sphgpu_abi_version(void);
/* Replaces SphBase::SphBase (SphBase.cpp:210-248): parameter defaults + state. */
int         sphgpu_create(const sphgpu_params* p, sphgpu_handle** out);
void        sphgpu_destroy(sphgpu_handle* h);
const char* sphgpu_last_error(const sphgpu_handle* h);
/* 1 when a gfx9 HIP device is visible to this process, else 0 (never initialises a context twice). */
int         sphgpu_device_available(void);

/* ---- scene state (before the first step) ------------------------------- */
/* Domain of scene s (m_TH.spaceMin/spaceMax).  Scene 0 defaults to params. */
int sphgpu_set_scene_domain(sphgpu_handle* h, int scene, const float mn[3], const float mx[3]);
/* SphFluid(restDensity, viscosity, tension) (SphBase.h:284): a fluid phase.
 * Returns the phase index in *phase (at most 8 phases). */
int sphgpu_add_phase(sphgpu_handle* h, float rest_density, float visc_alpha, int* phase);
/* Appends n fluid particles of `phase` to scene `scene`
 * (SphWithIO::addFluidCuboid output, SphWithIO.cpp:149-194).  Original ids are
 * assigned consecutively in call order; *first_id receives the first. */
int sphgpu_add_fluid(sphgpu_handle* h, int scene, int phase, long long n,
                     const float* xyz, const float* vxyz, long long* first_id);
/* Appends a static solid of n boundary particles to scene `scene`
 * (SphWithBullet::addSolid -> SphSolid, SphBase.h:315-365, rigid type).
 * visc_alpha is SphSolid::viscosity_alpha. */
/* Appends fluid particles after sphgpu_init, between steps: the emitters a
 * stepEvent runs (SphWithIO::addFluidParts, bangzi.h:193-196 inside
 * generateWater1, 129-208).  Same arguments and ids as sphgpu_add_fluid
 * (which forwards here once the handle is initialized); the next step's
 * neighbour search takes the new particles in (a DFSPH step then searches
 * before its density solve, since that solve reads the previous search's
 * lists).  SPHGPU_E_CAPACITY beyond sphgpu_params.fluid_capacity. */
int sphgpu_append_fluid(sphgpu_handle* h, int scene, int phase, long long n, const float* xyz,
                        const float* vxyz, long long* first_id);
/* Per-particle rest-density factor beta (FluidParticle::beta, SphBase.h:98)
 * of the fluid particles [first, first + n) in add order, for the IISph DF_*
 * solvers (m = rho0 beta d^3, IISph.cpp:110-115); default 1.  Before
 * sphgpu_init. */
int sphgpu_set_fluid_beta(sphgpu_handle* h, long long first, long long n, const float* beta);
int sphgpu_add_solid(sphgpu_handle* h, int scene, long long n, const float* xyz,
                     float visc_alpha, int* solid);
/* Appends a dynamic rigid body of n boundary particles given in body
 * coordinates (SphWithBullet::addSolid with a rigid body, SphBase.h:315-365;
 * innerPositions).  Its world positions follow the pose set with
 * sphgpu_set_solid_pose (identity until then).  Before sphgpu_init only. */
int sphgpu_add_rigid(sphgpu_handle* h, int scene, long long n, const float* local_xyz,
                     float visc_alpha, int* solid);
/* Pose of a dynamic solid from the host's rigid-body step
 * (SphSolid::updatePaticles, SphBase.cpp:95-128): T is the OpenGL column-major
 * 4x4 transform of the body-frame samples, com the centre of mass, v and w the
 * linear and angular velocity (particle velocity v + w x (x - com)).  Takes
 * effect at the next solids update point of sphgpu_step (see
 * sphgpu_set_update_solids); set before sphgpu_init it places the body. */
int sphgpu_set_solid_pose(sphgpu_handle* h, int solid, const float T[16], const float com[3],
                          const float v[3], const float w[3]);
/* Total force and torque about the centre of mass that the fluid exerted on a
 * dynamic solid (SphSolid::rigidForceJoin, SphBase.cpp:131-143), summed in
 * double in a fixed order: WCSPH and IISPH the reaction of the last step
 * (WcSph.cpp:438-460, IISph_M.cpp:660-681); DFSPH the pressure impulse of the
 * last constant-density solve over dt (DFSPH.cpp:674-690, 723-737).  Zero for
 * static solids. */
int sphgpu_get_solid_force_torque(sphgpu_handle* h, int solid, double f[3], double t[3]);
/* Host rigid-body step inside sphgpu_step, at the point where the reference
 * calls SphWithBullet::updateSolids (SphWithBullet.cpp:180-224): after the
 * fluid update for WCSPH and IISPH (WcSph.cpp:85, IISph_M.cpp:64), between
 * the constant-density solve and the next neighbour search for DFSPH
 * (DFSPH.cpp:54).  The callback may call sphgpu_get_solid_force_torque and
 * sphgpu_set_solid_pose; poses it sets take effect before the step goes on.
 * A nonzero return makes sphgpu_step return SPHGPU_E_CALLBACK once the step
 * has finished (with the poses set so far).  fn = NULL removes
 * it; without a callback, poses set between steps take effect at the next
 * update point (for DFSPH: after the next constant-density solve). */
typedef int (*sphgpu_update_solids_fn)(sphgpu_handle* h, void* user);
int sphgpu_set_update_solids(sphgpu_handle* h, sphgpu_update_solids_fn fn, void* user);
/* Uploads, sorts the boundary, and computes the Akinci boundary volumes
 * (SphBase::prepareSolidPartWeight, SphBase.cpp:766-813).  Called by the first
 * step if not called before. */
int sphgpu_init(sphgpu_handle* h);

/* ---- stepping ---------------------------------------------------------- */
/* One SphBase::runOneStep (SphBase.cpp:348-354) of the configured solver.
 * With st == NULL nothing is read back at the end of the step (the host does
 * not wait for the device unless a solver loop must); the counters then
 * accumulate until the next step that passes a stats struct. */
int sphgpu_step(sphgpu_handle* h, sphgpu_stats* st);
/* Per-scene statistics of the last step (batched scenes). */
int sphgpu_scene_stats(const sphgpu_handle* h, int scene, sphgpu_stats* st);
/* Time step of scene `scene` in the current step (inside the update-solids
 * callback) or the last one: m_TH.dt after the CFL adaptation, which the
 * reference's updateSolids hands to Bullet (SphWithBullet.cpp:214). */
int sphgpu_get_step_dt(sphgpu_handle* h, int scene, float* dt);
/* Waits for all work of the handle's stream. */
int sphgpu_synchronize(sphgpu_handle* h);

/* ---- readback ---------------------------------------------------------- */
long long sphgpu_num_fluid(const sphgpu_handle* h);
/* Copies a field of all live fluid particles, ascending original id.
 * `out` holds num_fluid * components floats. */
int sphgpu_get_fluid(sphgpu_handle* h, int field, float* out);
/* Original ids of the live fluid particles, ascending. */
int sphgpu_get_fluid_ids(sphgpu_handle* h, long long* out);
/* Akinci boundary volume of every boundary particle, in add order. */
int sphgpu_get_boundary_volume(sphgpu_handle* h, float* out);
/* Test dump of the neighbour sets at the current positions (re-runs the grid
 * search).  For live fluid particle k (ascending id) writes counts[k] and up
 * to `cap` entries (kind 0 fluid / 1 boundary, original id, FP32 distance
 * computed as SphBase.cpp:885-886), sorted by (kind, id). */
int sphgpu_get_neighbours(sphgpu_handle* h, int cap, int* counts, int* kinds,
                          long long* ids, float* dis);

/* ---- frame output (fUtility.h wire formats) ---------------------------- */
/* Components of a .pos record, in the reference's order (fUtility.h:14-21;
 * color and dt are not kept on the device). */
enum {
    SPHGPU_POS_POSITION = 1,
    SPHGPU_POS_VELOCITY = 2,
    SPHGPU_POS_DENSITY = 8,
    SPHGPU_POS_PRESURE = 16,       /* reference spelling */
    SPHGPU_POS_ACCELERATION = 32
};
/* Starts an asynchronous capture of the fluid state (positions, velocities,
 * densities, pressures, accelerations, ids) at the current point of the
 * stream: a device-side copy on the handle's stream, then the transfer to
 * pinned host memory on a second stream, overlapping the steps issued after
 * it.  The frame number is recorded with it.  One capture at a time: a new
 * one first waits for the previous. */
int sphgpu_snapshot_begin(sphgpu_handle* h);
/* Waits for the capture started last. */
int sphgpu_snapshot_wait(sphgpu_handle* h);
/* write_fluid_particles (fUtility.h:254-303): four text lines (log-file
 * name, component mask, frame number and fluid index, count), then per
 * particle the selected components as raw float32 (vec_t = 3 floats), fluid
 * phase `phase` in ascending original id, from the last capture (one is
 * taken synchronously when none exists).  components: SPHGPU_POS_* bits. */
int sphgpu_write_pos(sphgpu_handle* h, const char* path, int phase, unsigned components,
                     const char* log_name);
/* write_solid_particles (fUtility.h:459-484): positions of solid `solid`'s
 * boundary particles in add order (dynamic solids at their current pose). */
int sphgpu_write_solid_pos(sphgpu_handle* h, const char* path, int solid, const char* log_name);
/* write_fluid_particles_json (fUtility.h:390-427): {"num", "position":
 * [{x, y, z}], "spacing_r", "temp_beta": [{value}], "velocity": [{x, y,
 * z}]} with y and z swapped as the reference writes them, members sorted as
 * jsoncpp's styled writer emits them; spacing_r is the particle's d
 * (smoothRadius_h / 2), temp_beta 0 (no heat transfer here). */
int sphgpu_write_json(sphgpu_handle* h, const char* path, int phase);
/* read_fluid_particles (fUtility.h:486-510), host only: reads a .pos file;
 * *n = particle count, data (optional, cap floats) the raw records, mask
 * (optional, mask_len bytes) the component line. */
int sphgpu_read_pos(const char* path, float* data, long long cap, long long* n, long long* frame,
                    char* mask, int mask_len);

/* ---- surface reconstruction (Source/re.c, surface_field, marching_cubes) */
/* The reference's offline surfacer: Yu & Turk anisotropic kernels
 * (surface_field::surface_aniso, surface_field.cpp:83-176) or the isotropic
 * field (surface_isotro, 45-75) on a node grid, then marching cubes
 * (marchingCubes, marching_cubes.cpp:45-120).  Independent of a simulation
 * handle (re.c reads .pos files). */
typedef struct sphgpu_surfacer sphgpu_surfacer;
/* surface_field::surface_setup (surface_field.cpp:8-43): field grid from
 * min - 0.75 r_h with int((max - min + 1.5 r_h) / cell) + 1 nodes per axis;
 * particle volume r_p^3; anisotropy radius 2 r_h; k_r 4, k_s 1.6 /
 * (0.16 (2 r_h)^2), k_n 0.95, N_e 25, lambda 1.  re.c uses r_h = 2 r_p and
 * cell = r_p / 4.  Without a device the parameters are still queryable and
 * the compute calls return SPHGPU_E_DEVICE. */
int sphgpu_surfacer_create(const float min[3], const float max[3], float radius_p, float radius_h,
                           float cell_size, sphgpu_surfacer** out);
void sphgpu_surfacer_destroy(sphgpu_surfacer* s);
const char* sphgpu_surfacer_last_error(const sphgpu_surfacer* s);
/* get_grid_min / get_grid_num / get_cell_size (surface_field.h:43-46) */
int sphgpu_surfacer_grid(const sphgpu_surfacer* s, double grid_min[3], int grid_num[3], double* cell_size);
/* Field from n particle positions (xyz, float32): surface_aniso / surface_isotro. */
int sphgpu_surfacer_aniso(sphgpu_surfacer* s, long long n, const float* xyz);
int sphgpu_surfacer_isotro(sphgpu_surfacer* s, long long n, const float* xyz);
/* After sphgpu_surfacer_aniso, per particle in input order (each optional):
 * ellipsoid centre x_bar (3), G row-major (9), det G, neighbours within 2 r_h. */
int sphgpu_surfacer_particles(sphgpu_surfacer* s, float* xbar, float* G, float* det, int* neighbours);
/* The node field (get_field), x fastest: grid_num[0] * grid_num[1] * grid_num[2]
 * floats; set_field replaces it (marching cubes on a caller's field). */
int sphgpu_surfacer_get_field(sphgpu_surfacer* s, float* out);
int sphgpu_surfacer_set_field(sphgpu_surfacer* s, const float* in);
/* marchingCubes: triangles of the iso-surface, 9 floats each (three
 * vertices), in cube order (x fastest, then y, then z); *ntris = the count
 * (tris NULL: count only; more than cap: SPHGPU_E_CAPACITY after writing
 * cap).  Corners below iso are outside; normals point towards them. */
int sphgpu_surfacer_marching_cubes(sphgpu_surfacer* s, float iso, float* tris, long long cap,
                                   long long* ntris);
/* The marching-cubes case table the surfacer uses (host only, no device):
 * per case (bit k = corner k below iso; corners and edges numbered as in
 * marching_cubes.cpp:1-30, 126-156) the crossed-edge mask (256), the
 * triangle count (256) and up to 5 edge triples, -1 terminated (256 x 16). */
int sphgpu_marching_cubes_table(unsigned short* edges, unsigned char* ntris, signed char* tris);

/* ---- measurement ------------------------------------------------------- */
/* Selects one kernel (by name, NULL to stop) whose launches are bracketed by
 * HIP events on the handle's stream; sphgpu_kernel_timing returns the launch
 * count, total milliseconds and the algorithmic bytes per launch summed over
 * launches. */
int sphgpu_profile_kernel(sphgpu_handle* h, const char* name);
/* Calibration of byte counters: copies `bytes` (rounded down to 16) between
 * two device buffers with a plain streaming copy kernel (k_copy_probe), twice;
 * *ms receives the second launch's duration.  No handle needed. */
int sphgpu_copy_probe(long long bytes, double* ms);
/* Choices of the autotuner, which runs before the first step on the scene
 * itself: choice[0] = tile mapping (1: contiguous tile runs per XCD, 0: round
 * robin), choice[1] = LDS staging depth of the list builder (6 or 8 groups);
 * ms (optional, 4 floats) = best times of the candidates (mapping 0, 1;
 * depth 6, 8).  SPHGPU_E_STATE before the first step. */
int sphgpu_get_tuning(const sphgpu_handle* h, int* choice, float* ms);
/* Fixes autotuner choices instead of timing them (no counterpart in the
 * reference; for reproducible set-ups and for tests): tile_map 0/1,
 * stage_groups 6/8, stream_lists 0/1, -1 leaves a choice to the timing.
 * Every choice gives bit-identical results.  Before the first step only
 * (SPHGPU_E_STATE after). */
int sphgpu_set_tuning(sphgpu_handle* h, int tile_map, int stage_groups, int stream_lists);
/* LDS windows of the solver-loop passes (no counterpart in the reference):
 * a pass with its window on stages each packed tile's fluid neighbour
 * records in LDS and reads neighbours from there; with it off every lane
 * gathers them from memory.  Both give bit-identical results.  By default
 * the autotuner times each DFSPH and IISPH iteration pass both ways before
 * the first step and keeps the faster; on = 1 / 0 instead turns every
 * pass's window on / off (a switch for measurements and tests). */
int sphgpu_set_lds_windows(sphgpu_handle* h, int on);
/* The passes whose window is on: *mask is a sum of SPHGPU_WIN_* bits;
 * ms (10 floats, may be NULL) receives per pass, in bit order, the timed
 * ms per launch with gathers (the best occupancy tried) and with the
 * window (0 where not timed); blocks_per_cu (5 ints, may be NULL) the
 * blocks per CU each gathering pass is capped at (0: not capped, or the
 * pass uses its window).  The gather build of a pass is timed at its
 * register-limited occupancy and capped at 6, 5, 4 and 3 blocks per CU
 * (by dynamic LDS it does not use); both builds give the same bits. */
#define SPHGPU_WIN_DENSITY 1     /* DFSPH density pass (DFSPH.cpp:572-643) */
#define SPHGPU_WIN_DIVERGENCE 2  /* DFSPH divergence pass (DFSPH.cpp:780-836) */
#define SPHGPU_WIN_VELOCITY 4    /* DFSPH velocity update of both loops (DFSPH.cpp:647-693, 840-883) */
#define SPHGPU_WIN_SUMDIJ 8      /* IISPH sum_j d_ij p_j (IISph_M.cpp:288-309) */
#define SPHGPU_WIN_PRESSURE 16   /* IISPH relaxed Jacobi update (IISph_M.cpp:312-351) */
int sphgpu_get_lds_windows(const sphgpu_handle* h, int* mask, float* ms, int* blocks_per_cu);
/* Form of the neighbour lists of the last step: *tiles = tiles holding
 * particles, *packed_tiles = those whose lists use 16-bit entries (the rest
 * use 32-bit entries; both hold the same neighbours in the same order);
 * *streamed (optional) = 1 if the passes read the lists with streaming
 * loads (the autotuner's choice: faster where lists and records exceed the
 * last-level cache).  Diagnostic; SPHGPU_E_STATE before lists were first
 * built. */
int sphgpu_get_list_form(sphgpu_handle* h, int* packed_tiles, int* tiles, int* streamed);
int sphgpu_kernel_timing(sphgpu_handle* h, long long* launches, double* total_ms,
                         double* total_alg_bytes);

/* ---- learned warm start ------------------------------------------------ */
/* Per-particle MLP 8 -> 64 -> 64 -> 64 -> 1 (tanh-form GELU), BF16 weights on
 * the matrix cores with FP32 accumulation.  In a DFSPH handle with
 * warm_start = SPHGPU_WARM_LEARNED it predicts the converged stiffness of the
 * divergence-free solve from 8 local features (density deviation, velocity
 * divergence, neighbour count, previous stiffness, alpha factor, |v|,
 * boundary distance, dt); it is trained online on the solver's own converged
 * values during the first learn_steps steps.  It is only the initial guess:
 * the solver still iterates to the same tolerance.  No counterpart in the
 * reference. */
#define SPHGPU_MLP_PARAMS 8961   /* W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[1][64] b4[1] */
typedef struct sphgpu_mlp sphgpu_mlp;
/* standalone network (tests, offline training); capacity = max samples per call */
int  sphgpu_mlp_create(int capacity, unsigned long long seed, sphgpu_mlp** out);
void sphgpu_mlp_destroy(sphgpu_mlp* m);
int  sphgpu_mlp_set_params(sphgpu_mlp* m, const float* params);
int  sphgpu_mlp_get_params(sphgpu_mlp* m, float* params);
/* out[i] = MLP(feat[8i .. 8i+7]) */
int  sphgpu_mlp_forward(sphgpu_mlp* m, long long n, const float* feat, float* out);
/* gradient of mean((MLP(feat) - target)^2) w.r.t. the parameters; *loss = that mean */
int  sphgpu_mlp_gradient(sphgpu_mlp* m, long long n, const float* feat, const float* target,
                         float* grad, double* loss);
/* loss of sphgpu_mlp_gradient / _train: mean(w e^2), e = out - target,
 * w = over_weight where e > 0 and 1 elsewhere (default 1: mean squared error) */
int  sphgpu_mlp_set_loss(sphgpu_mlp* m, float over_weight);
/* `iters` Adam steps on the full set (n <= capacity); *loss = final mean squared error */
int  sphgpu_mlp_train(sphgpu_mlp* m, long long n, const float* feat, const float* target,
                      int iters, float lr, double* loss);
/* Gradient of the loss of sphgpu_mlp_gradient on caller-provided DEVICE
 * buffers, allocating nothing: params (SPHGPU_MLP_PARAMS FP32 master
 * weights), feat (n x 8), target (n) in; grad (SPHGPU_MLP_PARAMS) and loss
 * (one double, the weighted mean squared error) out; over_weight as in
 * sphgpu_mlp_set_loss (1 = mean squared error).  Work is queued on `stream`
 * (a hipStream_t; NULL = the default stream) and the call returns without
 * waiting.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training
 * loops).  Its scratch (weight fragments, activations, partial sums; 8.2
 * MB) is one static array per GPU, held until the process ends, so
 * n <= SPHGPU_MLP_DEVICE_MAX (else SPHGPU_E_CAPACITY), and calls that
 * overlap on one GPU (on different streams, or from different threads)
 * race on it; calls on one stream run one after another.  Same kernels and
 * the same results as sphgpu_mlp_gradient. */
#define SPHGPU_MLP_DEVICE_MAX 4096
int  sphgpu_mlp_gradient_device(const float* params, long long n, const float* feat, const float* target,
                                float over_weight, float* grad, double* loss, void* stream);
/* network of a simulation handle (SPHGPU_MLP_PARAMS floats) */
int  sphgpu_get_mlp_params(sphgpu_handle* h, float* params);
int  sphgpu_set_mlp_params(sphgpu_handle* h, const float* params);

/* ---- scene generation (host only; the reference's samplers) ------------ */
/* Cubic lattice block of nx*ny*nz particles, x fastest (SphWithIO.cpp:164-192),
 * origin + i*spacing, plus uniform jitter in [-jitter, +jitter]*spacing per axis
 * drawn from PCG32(seed, stream) in lattice order.  Returns the count. */
long long sphgpu_scene_lattice(const int n[3], const float origin[3], float spacing,
                               float jitter, unsigned long long seed,
                               unsigned long long stream, float* xyz);
/* Surface samples of an axis-aligned box (SphWithIO::makeCubeSampleParts,
 * SphWithIO.cpp:489-522) centred at `center`; have_top = 0 leaves the +y face
 * open.  xyz may be NULL to count.  Returns the count. */
long long sphgpu_scene_box(const float center[3], const float half[3], float spacing,
                           int have_top, float* xyz);

#ifdef __cplusplus
}
#endif
#endif /* SPHGPU_H_ */
