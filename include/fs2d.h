/*
 * fs2d.h -- C ABI of the MI355X (gfx950) implementation of fluid-sim's 2D
 * Eulerian step.
 *
 * The reference exposes this path as a C++ value type, FluidSim2D
 * (include/FluidSim2D.h:65-176, src/FluidSim2D.cpp), with public fields and
 * stage methods; there is no plugin or FFI layer.  Each entry point below
 * replaces one member or method of that type; the comment above it cites it.
 * fluid-sim_amd/cpp/FluidSim2D.hpp rebuilds the reference's method names on
 * top of these calls.
 *
 * Conventions
 *   - Arrays are row-major, element (i, j) at [j * width + i], with the
 *     reference's shapes (Array2D.h:41-44, MACGrid2D.h:19-25):
 *       u, newu : (size_x + 1) x size_y        v, newv : size_x x (size_y + 1)
 *       p, phi, cell, density : size_x x size_y
 *     particle arrays are (x, y) pairs, particle order as created.
 *   - Every call returns an fs2d_status; fs2d_last_error() explains the last
 *     non-OK status of a handle.  No call exits the process (the reference
 *     calls exit() on allocation failure and on NaN particles).
 *   - One HIP stream per handle; calls on different handles may run on
 *     different threads.  A handle is not thread-safe.
 *   - Device state is FP32 (velocities, pressure, particles, density) except
 *     the level set, which is FP64 as in the reference; dot products and
 *     residual accumulations of the PCG are FP64.
 */
#ifndef FS2D_H
#define FS2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs2d_sim fs2d_sim;

typedef enum {
    FS2D_OK = 0,
    FS2D_WARN_MAXITER = 1,   /* PCG hit max_iters (reference: log_error, FluidSim2D.cpp:464-466) */
    FS2D_WARN_CFL = 2,       /* particle CFL > 5 (reference: log_warn, FluidSim2D.cpp:582-585) */
    FS2D_ERR_NAN = -1,       /* NaN particle position (reference: exit(1), FluidSim2D.cpp:598-601) */
    FS2D_ERR_OOM = -2,       /* allocation failed (reference: exit, Array2D.h:69-71) */
    FS2D_ERR_CONFIG = -3,    /* invalid configuration */
    FS2D_ERR_ARG = -4,       /* invalid argument (field id, size, stage) */
    FS2D_ERR_DEVICE = -5,    /* HIP runtime error */
    FS2D_ERR_HANG = -6       /* a wavefront kernel's bounded wait expired */
} fs2d_status;

/* FluidSimMode (include/FluidSim2D.h:40-42). */
typedef enum { FS2D_SEMILAGRANGIAN = 0, FS2D_PICFLIP = 1 } fs2d_mode;

/* FluidCellType (include/FluidSim2D.h:44-46). */
typedef enum { FS2D_EMPTY = 0, FS2D_FLUID = 1, FS2D_SOLID = 2 } fs2d_cell;

/* FluidSim2D::StageType (include/FluidSim2D.h:93-97). */
typedef enum {
    FS2D_STAGE_INIT = 0,
    FS2D_STAGE_CREATE_WATER_LEVEL_SET = 1,
    FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID = 2,
    FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION = 3,
    FS2D_STAGE_APPLY_GRAVITY = 4,
    FS2D_STAGE_CREATE_SOLID_LEVEL_SET = 5,
    FS2D_STAGE_APPLY_PROJECTION = 6,
    FS2D_STAGE_UPDATE_VELOCITY = 7,
    FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES = 8,
    FS2D_STAGE_APPLY_ADVECTION = 9
} fs2d_stage;

/* Field ids for fs2d_get_field / fs2d_set_field. */
typedef enum {
    FS2D_FIELD_U = 0,        /* mac.u    (FluidSim2D.h:71) */
    FS2D_FIELD_V = 1,        /* mac.v */
    FS2D_FIELD_NEWU = 2,     /* newMac.u (FluidSim2D.h:72) */
    FS2D_FIELD_NEWV = 3,     /* newMac.v */
    FS2D_FIELD_P = 4,        /* p        (FluidSim2D.h:73) */
    FS2D_FIELD_PHI = 5,      /* waterLevelSet.phi (FluidSim2D.h:78) */
    FS2D_FIELD_DENSITY = 6,  /* passive density (no reference counterpart) */
    FS2D_FIELD_CELL = 7,     /* cell     (FluidSim2D.h:74), uint8 only */
    FS2D_FIELD_PARTICLE_POS = 8, /* particles    (FluidSim2D.h:75), 2 per particle */
    FS2D_FIELD_PARTICLE_VEL = 9, /* particleVels (FluidSim2D.h:76), 2 per particle */
    /* Read-only, FS2D_I32, 1 per particle: each particle's cell key
     * y_cell * size_x + x_cell, with x_cell = (int)(x / dx) as the reference's
     * LevelSet::constructFromParticles computes it (FluidSim2D.cpp:766), and
     * the particle indices sorted by that key (ties by index), the order in
     * which P2G and the level set gather particles.  Reading either sorts
     * the current positions if they changed since the last sort. */
    FS2D_FIELD_PARTICLE_CELL = 10,
    FS2D_FIELD_PARTICLE_ORDER = 11
} fs2d_field;

typedef enum { FS2D_F32 = 0, FS2D_F64 = 1, FS2D_U8 = 2, FS2D_I32 = 3 } fs2d_dtype;

/* Switches.  STATIC_CELLS and DENSITY have no reference counterpart; the
 * other two select the reference's compile-time variants
 * (src/FluidSim2D.cpp:15-16). */
enum {
    FS2D_FLAG_STATIC_CELLS = 1,     /* keep the created cell flags; no level set ("all-fluid" mode) */
    FS2D_FLAG_DENSITY = 2,          /* advect a passive cell-centred density in the SL stage */
    FS2D_FLAG_PARTICLE_CELLS = 4,   /* USE_LEVEL_SET undefined: cells from particle positions and a
                                       bubble fill (FluidSim2D.cpp:665-706); needs NO_GHOST_PRESSURE */
    FS2D_FLAG_NO_GHOST_PRESSURE = 8, /* USE_GHOST_PRESSURE undefined (FluidSim2D.cpp:304-331, 486-530) */
    /* Options of the 512^2 PIC/FLIP dam-break workload, no reference counterpart:
     * seed each of the k^2 particles of a FLUID cell uniformly over the whole cell
     * (the reference jitters one per sub-cell, FluidSim2D.cpp:52-64), and advect
     * particles by the midpoint rule (RK2) instead of applyAdvection's RK3
     * (FluidSim2D.cpp:588-603). */
    FS2D_FLAG_UNIFORM_SEEDING = 16,
    FS2D_FLAG_RK2 = 32
};

/*
 * FluidSim2DConfig (include/FluidSim2D.h:48-63), plus the solver constants the
 * reference hard-codes (FluidSim2D.cpp:369-370, 429, 453) as parameters.
 * fs2d_config_init() fills the reference's values: tau 0.999, sigma 0.25,
 * pcg_tol 1e-12 (relative infinity norm), pcg_max_iters 200, seed 1 (glibc's
 * default rand() seed, which create() uses through randf, FluidSim2D.cpp:19-21).
 */
typedef struct {
    int size_x;
    int size_y;
    int particles_per_cell_sqrt;
    double dt;
    double dx;
    double rho;
    double gravity_x;
    double gravity_y;
    int mode;                       /* fs2d_mode */
    double pic_flip_alpha;          /* PIC weight: 1 = pure PIC (FluidSim2D.cpp:564) */
    const uint8_t* initial_values;  /* size_x*size_y flags, row-major; border must be SOLID */
    double tau;
    double sigma;
    double pcg_tol;
    int pcg_max_iters;
    unsigned seed;
    int flags;                      /* FS2D_FLAG_* */
} fs2d_config;

/* Per-step observables: FluidSim2D::waterVolume/totalEnergy/particleTotalEnergy
 * (FluidSim2D.h:106-114, computed at FluidSim2D.cpp:709-731), currentTime
 * (FluidSim2D.h:87) and the solver's state. */
typedef struct {
    double water_volume;
    double total_energy;
    double particle_total_energy;
    double orig_water_volume;
    double current_time;
    double pcg_residual;      /* |r|_inf / |rhs|_inf at exit of the last solve */
    double cfl_max;           /* max (vx+vy)*dt/dx over particles (FluidSim2D.cpp:572-581) */
    int pcg_iters;            /* iterations of the last solve */
    int pcg_hit_max;
    long long n_particles;
    long long steps;
} fs2d_stats;

/* Fill *cfg with the reference's defaults (FluidSim2D.h:80-91 and the
 * constants above); size, dt, dx, mode and flags still have to be set. */
fs2d_status fs2d_config_init(fs2d_config* cfg);

/* FluidSim2D::create (src/FluidSim2D.cpp:23-76): allocate device state, copy
 * the flags, seed k^2 jittered particles per FLUID cell. */
fs2d_status fs2d_create(const fs2d_config* cfg, fs2d_sim** out);

/* FluidSim2D::free (src/FluidSim2D.cpp:78-92). */
void fs2d_destroy(fs2d_sim* sim);

/* FluidSim2D::runFrame / update (src/FluidSim2D.cpp:94-142): one step of the
 * member dt; stage order by mode.  Returns the worst status of the stages. */
fs2d_status fs2d_step(fs2d_sim* sim);

/* One stage method (FluidSim2D.h:124-140), e.g. applyProjection(). */
fs2d_status fs2d_run_stage(fs2d_sim* sim, int stage);

/* Wait for all queued work of the handle. */
fs2d_status fs2d_synchronize(fs2d_sim* sim);

/* Public field access (FluidSim2D.h:71-78): copy between device state and a
 * host buffer of `count` elements of `dtype` (conversion on the host side). */
size_t fs2d_field_count(const fs2d_sim* sim, int field);
fs2d_status fs2d_get_field(fs2d_sim* sim, int field, int dtype, void* host, size_t count);
fs2d_status fs2d_set_field(fs2d_sim* sim, int field, int dtype, const void* host, size_t count);
/* Replacing the particle set changes its size; velocities are zeroed. */
fs2d_status fs2d_set_particles(fs2d_sim* sim, int dtype, const void* pos_xy, size_t n);

/* Device pointer of a field (FP32 arrays, uint8 cell flags; particle fields
 * are SoA: x then y, each n long).  For zero-copy interop on the handle's stream. */
fs2d_status fs2d_device_field(fs2d_sim* sim, int field, void** dptr);
void* fs2d_stream(fs2d_sim* sim);

/* Runtime-mutable fields: gravity, picFlipAlpha, dt (FluidSim2D.h:80, 84, 89),
 * written by demo/FluidRenderer2D.cpp:322, 382-388. */
fs2d_status fs2d_set_gravity(fs2d_sim* sim, double gx, double gy);
fs2d_status fs2d_set_pic_flip_alpha(fs2d_sim* sim, double alpha);
fs2d_status fs2d_set_dt(fs2d_sim* sim, double dt);
/* Solver constants of applyProjection (FluidSim2D.cpp:369-370, 429, 453). */
fs2d_status fs2d_set_solver(fs2d_sim* sim, double tau, double sigma, double tol, int max_iters);

/* waterVolume, totalEnergy, particleTotalEnergy, currentTime and solver stats. */
fs2d_status fs2d_get_stats(fs2d_sim* sim, fs2d_stats* out);

/* FluidSim2D::avgPressure / avgPressureInFluid / maxVelocity
 * (src/FluidSim2D.cpp:607-638). */
fs2d_status fs2d_avg_pressure(fs2d_sim* sim, int fluid_only, double* out);
fs2d_status fs2d_max_velocity(fs2d_sim* sim, double* out);

/* Step by replaying captured HIP graphs (no reference counterpart; runFrame,
 * src/FluidSim2D.cpp:94-138, is what is captured).  fs2d_step then replays
 * one graph for the stages up to a fixed number of PCG iterations (kernels
 * of iterations after convergence return at once), checks convergence once on
 * the host, continues the solve with queued chunks if the budget fell short,
 * and replays a second graph for the rest of the step.  The budget follows
 * the previous solve's iteration count; a step is captured again when a
 * parameter, the particle count, the budget or a buffer it uses changed (at
 * most four captures are kept).  Results are bit-identical to the uncaptured
 * step.  Timing (fs2d_set_timing) runs uncaptured steps. */
fs2d_status fs2d_set_graph(fs2d_sim* sim, int enable);

/* Per-kernel timing with HIP events on the handle's stream (no reference
 * counterpart; PerformanceCounter, src/PerformanceCounter.cpp:9-74, timed
 * stages with a host clock).  Kernel classes are listed by fs2d_kernel_name. */
fs2d_status fs2d_set_timing(fs2d_sim* sim, int enable);
/* Times only the kernel classes whose bit is set in class_mask (bit k: class k
 * of fs2d_kernel_name; 0 turns timing off).  Each timed launch records two
 * events on the stream, which costs the step time: timing one class measures
 * that class while leaving the others untouched.  No reference counterpart. */
fs2d_status fs2d_set_timing_classes(fs2d_sim* sim, unsigned long long class_mask);
int fs2d_kernel_count(void);
const char* fs2d_kernel_name(int kernel);
fs2d_status fs2d_get_kernel_time(fs2d_sim* sim, int kernel, double* total_ms, long long* launches);
fs2d_status fs2d_reset_kernel_times(fs2d_sim* sim);

const char* fs2d_last_error(const fs2d_sim* sim);
const char* fs2d_status_string(int status);

/*
 * Launch shapes that change no result, and an autotuner for them.  No
 * reference counterpart (its PCG loop, src/FluidSim2D.cpp:431-463, is what
 * they shape).
 *   pcg_grid: physical workgroups of the PCG SpMV and axpy launches, a
 *     multiple of 8; 0 (single handles only) launches one per partial sum.
 *     The loops run a fixed list of virtual workgroups whatever the launch
 *     size, so every value gives the same bits.  Unused by fused_pcg
 *     iterations, whose launches size themselves to the device.
 *   sweep_prefetch: (ensembles) the batched triangular solves' loaders issue
 *     their first loads before testing whether their member has converged.
 *   fused_pcg: (single handles) 1, the default for solver layouts of at least
 *     2^18 entries (ceil(size_y / 64) * 64 * ceil((size_x + 64) / 32) * 32;
 *     512 x 512 has 294912; below it the separate launches measured 3-8%
 *     faster): a PCG iteration is two launches instead of four: the p/r update with the forward solve, whose
 *     strips start as soon as the update has passed them, and the backward
 *     solve with the next iteration's SpMV, whose workgroups take each region
 *     as soon as the solve's strips around it have finished (the stopping
 *     test moves from the forward to the backward launch); 0: separate
 *     launches.  Ensembles (default 0): 1 runs the p/r update of every
 *     member beside the forward solves in one launch (the SpMV stays
 *     separate); it pays only when those solves leave CUs idle.
 * fs2d_autotune / fs2d_batch_autotune time the candidates on the handle's
 * current state (call them after a first step): `iters` PCG iterations of the
 * projection each (tolerance 0), best of three after a warm-up; the fastest is
 * set and appended to the text file `cache_path` (NULL: none) under a key of
 * the device and the problem, from which later calls with the same key take it
 * without timing (*from_cache = 1).  Tuning overwrites the pressure and the
 * solver's scratch, not the state a step starts from.
 */
typedef struct {
    int pcg_grid;
    int sweep_prefetch;
    int fused_pcg;
} fs2d_tuning;
fs2d_status fs2d_get_tuning(fs2d_sim* sim, fs2d_tuning* out);
fs2d_status fs2d_set_tuning(fs2d_sim* sim, const fs2d_tuning* t);
fs2d_status fs2d_autotune(fs2d_sim* sim, int iters, const char* cache_path, fs2d_tuning* out, int* from_cache);

/*
 * applyProjection without allocation.  No reference counterpart: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It runs the
 * ghost-fluid assembly, the MIC(0) factor and the PCG of
 * fs2d_run_stage(sim, FS2D_STAGE_APPLY_PROJECTION) (src/FluidSim2D.cpp:252-467),
 * with four launches per PCG iteration (fused_pcg 0, whatever the grid), on the
 * caller's device arrays and gives the same pressure bit for bit.  Every other array
 * lives in a static device buffer of about 28 MB per GPU, held from the
 * library's load until the process ends; grids whose solver layout has at most
 * 2^19 entries fit: ceil(size_y / 64) * 64 * ceil((size_x + 64) / 32) * 32
 * <= 524288 (512 x 512 does, 700 x 700 does not).  Larger grids are rejected
 * with FS2D_ERR_ARG before any device work.  Calls on one GPU share that
 * buffer: calls that overlap on one GPU (two threads, or two streams) race.
 * The call returns after the solve has finished on `stream`; the host reads
 * the convergence flag between chunks of iterations with synchronising
 * copies into its own stack.  Failures are explained by fs2d_last_error(NULL).
 */
typedef struct {
    int size_x;
    int size_y;
    double dt;
    double dx;
    double rho;
    double tau;
    double sigma;
    double pcg_tol;
    int pcg_max_iters;
    int flags;              /* 0 or FS2D_FLAG_NO_GHOST_PRESSURE */
    const uint8_t* cell;    /* device, size_x * size_y flags (SOLID border) */
    const double* phi;      /* device, size_x * size_y level set */
    const float* u;         /* device, (size_x + 1) * size_y */
    const float* v;         /* device, size_x * (size_y + 1) */
    float* p;               /* device, size_x * size_y: the pressure out */
    void* stream;           /* hipStream_t to run on; NULL: the default stream */
} fs2d_projection_args;
/* Iterations and |r|_inf / |rhs|_inf at exit go to *iters and *residual when
 * they are not NULL; FS2D_WARN_MAXITER when pcg_max_iters was reached. */
fs2d_status fs2d_project_noalloc(const fs2d_projection_args* args, int* iters, double* residual);

/* Per-frame render buffers of the demo renderer, built on the device from the
 * handle's state (FluidRenderer2D::updateBuffers, demo/FluidRenderer2D.cpp:
 * 435-486), for a renderer that maps its vertex buffers into device memory
 * (GL/Vulkan interop): the state never goes to the host.  Every pointer is a
 * device address and may be NULL to skip that buffer; all are FP32, 2 floats
 * per point, in the reference's push order (row-major cells, j outer).
 *   water_cells, solid_cells  (i dx, j dx) of each FLUID / SOLID cell
 *                             (capacity: size_x * size_y points each)
 *   cell_vel_lines            per cell (i, j) with i < size_x - 1, j < size_y - 1:
 *                             c = ((i + 1/2) dx, (j + 1/2) dx) and c + dt velInterp(c)
 *   pressure_cells, pressure_values  (i dx, j dx) and sigmoid(0.01 p) of each
 *                             cell with p != 0 (capacity: size_x * size_y)
 *   particle_lines            per particle: x and x + dt velInterp(x)
 *   phi_values                per cell: sigmoid(100 phi)
 * sigmoid(x) = 1 / (1 + e^-x).  The counts go to *out (synchronises the
 * handle's stream). */
typedef struct {
    float* water_cells;
    float* solid_cells;
    float* cell_vel_lines;
    float* pressure_cells;
    float* pressure_values;
    float* particle_lines;
    float* phi_values;
} fs2d_render_targets;
typedef struct {
    long water_cells, solid_cells, pressure_cells, particles, cell_vel_lines, phi_values;
} fs2d_render_counts;
fs2d_status fs2d_render_buffers(fs2d_sim* sim, const fs2d_render_targets* targets, fs2d_render_counts* out);

/*
 * Ensembles: many independent simulations of different sizes stepped
 * together.  No reference counterpart (the reference steps one FluidSim2D per
 * runFrame, src/FluidSim2D.cpp:94-138); each member steps exactly as
 * fs2d_step steps it alone (the same kernels, bit-identical results), but
 * every kernel of the step is launched once for all members and the PCG loop
 * masks members that have converged.  Members share mode, flags and the
 * solver parameters (tau, sigma, pcg_tol, pcg_max_iters); sizes, dt, dx,
 * gravity and pic_flip_alpha may differ.
 */
typedef struct fs2d_batch fs2d_batch;

/* Create n members from cfgs[0..n) (each as fs2d_create would). */
fs2d_status fs2d_batch_create(const fs2d_config* cfgs, int n, fs2d_batch** out);
/* Destroys the members too. */
void fs2d_batch_destroy(fs2d_batch* batch);
int fs2d_batch_size(const fs2d_batch* batch);
/* Member i's handle, for fields, statistics and set-up calls (owned by the
 * batch; do not destroy it).  Calls on a member's handle complete before they
 * return, so they can be interleaved with batch steps. */
fs2d_sim* fs2d_batch_sim(fs2d_batch* batch, int i);
/* One fs2d_step of every member.  Returns the worst status. */
fs2d_status fs2d_batch_step(fs2d_batch* batch);
/* One stage method (FluidSim2D.h:124-140) on every member. */
fs2d_status fs2d_batch_run_stage(fs2d_batch* batch, int stage);
/* Step by replaying captured HIP graphs, as fs2d_set_graph does for one
 * simulation: every member's stages and PCG iterations up to a fixed budget in
 * one graph (converged members' work returns at once on the device), one host
 * check of convergence, and the rest of the step in a second graph.  Results
 * are bit-identical to the uncaptured batch step. */
fs2d_status fs2d_batch_set_graph(fs2d_batch* batch, int enable);
/* Tuning of the batched launches (see fs2d_tuning). */
fs2d_status fs2d_batch_get_tuning(fs2d_batch* batch, fs2d_tuning* out);
fs2d_status fs2d_batch_set_tuning(fs2d_batch* batch, const fs2d_tuning* t);
fs2d_status fs2d_batch_autotune(fs2d_batch* batch, int iters, const char* cache_path, fs2d_tuning* out,
                                int* from_cache);
/* Per-kernel timing of the batched launches (kernel classes of fs2d_kernel_name). */
fs2d_status fs2d_batch_set_timing(fs2d_batch* batch, int enable);
fs2d_status fs2d_batch_get_kernel_time(fs2d_batch* batch, int kernel, double* total_ms, long long* launches);
fs2d_status fs2d_batch_reset_kernel_times(fs2d_batch* batch);
const char* fs2d_batch_last_error(const fs2d_batch* batch);

/*
 * Adjoint (reverse-mode gradient) of a rollout of the all-fluid
 * semi-Lagrangian step.  No reference counterpart: the reference has no
 * adjoint (its step is runFrame, src/FluidSim2D.cpp:94-115, whose stages
 * applySemiLagrangianAdvection :206-235, applyGravity :237-250,
 * applyProjection :252-467 and updateVelocity :469-550 are differentiated
 * here).  Scope: FS2D_SEMILAGRANGIAN, FS2D_FLAG_STATIC_CELLS, no EMPTY cell,
 * no particles (k = 0); FS2D_FLAG_DENSITY adds the passive density to the
 * differentiated state (u, v, density).  Errors are reported through
 * fs2d_last_error of the simulation handle.  The adjoint handle must be
 * destroyed before its simulation.
 */
typedef struct fs2d_adjoint fs2d_adjoint;

/* Allocate checkpoints for rollouts of up to max_steps steps, keeping the
 * state every checkpoint_every steps (1 keeps every state: no recomputation). */
fs2d_status fs2d_adjoint_create(fs2d_sim* sim, int checkpoint_every, int max_steps, fs2d_adjoint** out);
void fs2d_adjoint_destroy(fs2d_adjoint* adj);

/* Run nsteps steps (fs2d_step) from the simulation's state, recording
 * checkpoints.  The simulation ends at the rollout's final state. */
fs2d_status fs2d_adjoint_forward(fs2d_adjoint* adj, int nsteps);

/* Reverse pass over the last forward(): the adjoint fields hold the gradient
 * of a scalar loss with respect to the final state (set them with
 * fs2d_adjoint_set_field) and receive its gradient with respect to the initial
 * state.  Recomputes each segment from its checkpoint; the simulation is left
 * at the rollout's final state.  May be called again with new seeds. */
fs2d_status fs2d_adjoint_backward(fs2d_adjoint* adj);

/* One vector-Jacobian product at the simulation's current state: the adjoint
 * fields hold the adjoints of the step's outputs and receive those of its
 * inputs.  The state is not changed. */
fs2d_status fs2d_adjoint_step(fs2d_adjoint* adj);

/* Adjoint fields FS2D_FIELD_U, FS2D_FIELD_V, FS2D_FIELD_DENSITY (shapes and
 * layout of the state fields; FP32 on the device). */
fs2d_status fs2d_adjoint_get_field(fs2d_adjoint* adj, int field, int dtype, void* host, size_t count);
fs2d_status fs2d_adjoint_set_field(fs2d_adjoint* adj, int field, int dtype, const void* host, size_t count);
fs2d_status fs2d_adjoint_device_field(fs2d_adjoint* adj, int field, void** dptr);

#ifdef __cplusplus
}
#endif

#endif /* FS2D_H */
