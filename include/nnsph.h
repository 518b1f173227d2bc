/*
 * nnsph.h -- C-ABI of libnnsph.so, the MI355X-native SPH step of chunleili/nonNewton.
 *
 * The reference has no FFI: its boundary is the Python object surface of
 * SPH/main.py (README.md:25-33).  Each entry point below replaces one piece of
 * that surface; the citation names it.  The Python mirror nonnewton_amd/sph.py
 * binds these with ctypes and keeps the reference's names.
 *
 * Conventions
 *  - Every function returns an int32 status: 0 = OK, negative = error (NN_E_*),
 *    NN_W_NOCONV (positive) = a solver hit its iteration limit.
 *    nn_last_error() gives the text of the last error on the calling thread.
 *  - The library owns all device memory; it allocates only in nn_create,
 *    nn_upload, nn_pool_* admission/tick and nn_surface_create/run*.  Host
 *    pointers are borrowed for the duration of the call.
 *  - Internally particles are kept in cell order; every array crossing this
 *    boundary is in the caller's original particle order.
 *  - A context is not thread-safe; distinct contexts are independent.
 *  - Arithmetic is FP32 per particle; solver-termination reductions and CG
 *    scalars are FP64.
 */
#ifndef NNSPH_H
#define NNSPH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NNSPH_ABI_VERSION 1

/* status codes */
#define NN_OK 0
#define NN_E_ARG (-1)
#define NN_E_HIP (-2)
#define NN_E_OOM (-3)
#define NN_E_NEIGHBOR_OVERFLOW (-4) /* > 60 neighbors (max_num_neighbors, SPH/main.py:777) */
#define NN_E_CELL_OVERFLOW (-5)     /* > 50 particles in a cell (:778) */
#define NN_E_OUT_OF_GRID (-6)       /* particle outside the grid (pos_to_index, :805-807) */
#define NN_E_NAN (-7)               /* non-finite position or velocity after a step */
#define NN_E_UNSUPPORTED (-8)
#define NN_E_BUSY (-9)              /* pool admission: over budget */
#define NN_E_STATE (-10)            /* call out of order (e.g. step before init) */
#define NN_E_REST_OVERFLOW (-11)    /* > 50 rest neighbors of an elastic particle (:1192) */
#define NN_W_NOCONV 1               /* a solver stopped at its iteration limit */

/* particle kinds (material + solid type, SPH/main.py:27-40, :258-272) */
#define NN_KIND_FLUID 0
#define NN_KIND_STATIC 1
#define NN_KIND_ELASTIC 2
/* NN_KIND_RIGID + b: a particle of rigid body b (solidType RIGID, RigidBodySolver :1099-1158);
 * bodies are numbered 0 .. B-1 without gaps, B <= 4096 */
#define NN_KIND_RIGID 3

/* Scene parameters: Parameter (SPH/main.py:598-621), DFSPHSolver.__init__
 * (:1520-1544), NonNewton.initParameters (:358-371), SurfaceTension (:1382),
 * Elasticity.__init__ (:1188-1189), Viscosity_Weiler2018.__init__ (:1480-1481). */
typedef struct nn_params {
    double domain_start[3];
    double domain_end[3];
    double particle_radius;
    double density0;
    double gravity[3];
    double boundary_viscosity;
    double dt;
    double sticky_coeff;
    double collision_coeff;
    double viscosity;
    int32_t max_iterations_v;
    int32_t max_iterations;
    double max_error_v;
    double max_error;
    double eps;
    int32_t enable_divergence_solver;
    int32_t use_surface_tension;
    int32_t use_viscosity;
    int32_t use_non_newtonian;
    int32_t viscosity_method; /* 0 explicit (ViscosityModel), 1 Weiler 2018 implicit */
    int32_t non_newton_method; /* 0..6, NonNewton.NON_NEWTON_METHOD (:377-384) */
    double surface_tension;
    double power_index;
    double consistency_index;
    double viscosity0;
    double viscosity_inf;
    double critical_strain_rate;
    double mu_c;
    double yield_stress;
    double max_viscosity;
    int32_t weiler_max_iter;
    int32_t compensated; /* Kahan sums for density and boundary volume */
    double weiler_tol;
    double youngs_modulus;
    double poisson_ratio;
    double particle_mass; /* <= 0: volume0 * 1000 (:743) */
    int32_t warm_start;   /* bit 0: the Weiler CG starts from v + (x - v) of the previous solve
                           * (Viscosity_Weiler2018's m_vDiff, :1486-1487); other bits must be 0 */
    int32_t coupling_interval; /* rigid-body substeps per step (Parameter.coupling_interval, :616) */
} nn_params;

/* Per-step statistics (what the reference prints each step, :916,1548,1704,1785). */
typedef struct nn_step_stats {
    int32_t iterations_v;  /* divergence-solver iterations run */
    int32_t iterations_p;  /* pressure-solver iterations run */
    int32_t iterations_cg; /* Weiler CG iterations */
    int32_t status;        /* NN status bits of the step (1 nbr, 2 cell, 4 grid, 8 nan, 16 rest) */
    double avg_density_err_v;
    double avg_density_err_p;
    double cg_rel_residual;
    double max_density;
    int64_t num_pairs;       /* sum of neighbor counts of fluid, elastic and rigid particles (the
                              * lists a step reads; static boundary lists are built on demand
                              * by nn_neighbors and the NN_FIELD_NUM_NEIGHBORS download) */
    int64_t num_pairs_fluid; /* sum of neighbor counts of fluid particles */
} nn_step_stats;

/* field ids for nn_download / nn_upload_field (ParticleData, :701-723) */
#define NN_FIELD_X 0            /* float[n][3] */
#define NN_FIELD_V 1            /* float[n][3] */
#define NN_FIELD_ACCELERATION 2 /* float[n][3] */
#define NN_FIELD_VOLUME 3       /* float[n] */
#define NN_FIELD_DENSITY 4      /* float[n] */
#define NN_FIELD_DFSPH_FACTOR 5 /* float[n], as the reference holds it after a step */
#define NN_FIELD_DENSITY_ADV 6  /* float[n] */
#define NN_FIELD_STRAIN_NORM 7  /* float[n] (NonNewton.strainRateNorm, :340) */
#define NN_FIELD_VISCOSITY 8    /* float[n] (NonNewton.nonNewtonViscosity, :342) */
#define NN_FIELD_ROTATION 9     /* float[n][9] (Elasticity.m_rotations, :1199) */
#define NN_FIELD_STRESS 10      /* float[n][6] (Elasticity.m_stress, :1200) */
#define NN_FIELD_REST_VOLUME 12 /* float[n] (Elasticity.m_restVolumes, :1198) */
#define NN_FIELD_X0 13          /* float[n][3] (ParticleData.x_0) */
#define NN_FIELD_KIND 14        /* int32[n] */
#define NN_FIELD_NUM_NEIGHBORS 15 /* int32[n] (NeighborhoodSearchSparse.num_neighbors) */
#define NN_FIELD_VISC_DIFF 16   /* float[n][3]: x - v of the previous Weiler solve, the warm
                                 * start's direction (Viscosity_Weiler2018.m_vDiff, :1486-1487);
                                 * with x and v it is the state a run continues from */

typedef struct nn_ctx nn_ctx;
typedef struct nn_pool nn_pool;

/* Library identity: ABI version and the device it runs on. */
int32_t nn_version(void);
const char *nn_last_error(void);
int32_t nn_device_info(char *name, int32_t name_len, int64_t *hbm_bytes, int32_t *num_cus);

/* Context for one scene -- replaces initialize() + ParticleData(N) +
 * NeighborhoodSearchSparse(...) (SPH/main.py:1870-1880, :701-723, :766-791). */
int32_t nn_create(const nn_params *params, int64_t max_particles, nn_ctx **out);
int32_t nn_destroy(nn_ctx *ctx);

/* Particle upload -- replaces initialize_particles (SPH/main.py:735-758):
 * x, v float[n][3], kind int32[n] (NN_KIND_*); volume = volume0, m = mass. */
int32_t nn_upload(nn_ctx *ctx, const float *x, const float *v, const int32_t *kind, int64_t n);

/* Replaces Elasticity(N).initValues (:1216-1236) + SPHBase.initialize (:957-960). */
int32_t nn_init(nn_ctx *ctx);

/* Replaces SPHBase.step (:915-940), run nsteps times; stats of the last step. */
int32_t nn_step(nn_ctx *ctx, int32_t nsteps, nn_step_stats *out);

/* Replaces meta.ns.run_search() (:822-828) on the current positions. */
int32_t nn_search(nn_ctx *ctx);

/* Replaces meta.pd.<field>.to_numpy() / .from_numpy() (:740-741, :940, :1548).
 * bytes must equal n * width * 4.  nn_upload_field takes NN_FIELD_X, NN_FIELD_V
 * and NN_FIELD_VISC_DIFF. */
int32_t nn_download(nn_ctx *ctx, int32_t field, void *host, size_t bytes);
int32_t nn_upload_field(nn_ctx *ctx, int32_t field, const void *host, size_t bytes);

/* Replaces meta.ns.neighbors / num_neighbors (:780-781): nbr int32[n][cap],
 * -1 padded, original indices, in the reference's serial list order. */
int32_t nn_neighbors(nn_ctx *ctx, int32_t *nbr, int32_t *count, int32_t cap);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The lists of nn_neighbors -- meta.ns.run_search() +
 * meta.ns.neighbors (:815-847), same sets and serial order -- for n positions
 * already in device memory (x: n x 3 float, caller order), written to device
 * memory: nbr int32[n][cap] (-1 padded, original indices), count int32[n],
 * status int32[1] (bits of nn_step_stats.status: 1 neighbor overflow, 2 cell
 * overflow, 4 out of grid; an out-of-grid particle gets no list and is nobody's
 * neighbor).  Only params' domain and particle_radius are read.  Enqueued on
 * `stream` (a hipStream_t; NULL = the default stream) and asynchronous: the
 * outputs, status included, are valid once the stream reaches that point.
 * Allocates nothing: the search's scratch is one fixed __device__ array of
 * 42 MB per GPU, held from the first call on that GPU until the process ends;
 * n above 1048576 or a grid above 4194304 cells is NN_E_ARG.  Every call on a
 * GPU shares that GPU's scratch: calls that overlap on one GPU (on different
 * streams or threads) race. */
int32_t nn_neighbors_noalloc(const nn_params *params, const float *x, int64_t n, int32_t *nbr, int32_t *count,
                             int32_t cap, int32_t *status, void *stream);

/* Replaces the 0-d fields meta.parm.dt[None] / viscosity[None] (:611-612,
 * :620-621) that the reference mutates at run time (:1905).  key: "dt",
 * "viscosity", "boundary_viscosity", "max_error", "max_error_v". */
int32_t nn_set_param(nn_ctx *ctx, const char *key, double value);

/* Number of particles and of fluid particles (meta.particle_max_num,
 * meta.fluid_particle_num, :1873). */
int32_t nn_counts(nn_ctx *ctx, int64_t *n, int64_t *n_fluid);

/* ---------------------------------------------------------------------------
 * Batched / serving form (no counterpart in the reference, whose scenes are
 * single; this is the multi-scene driver over the same step).  A pool holds
 * many independent scenes resident in HBM in one concatenated, offset-indexed
 * particle array and advances all of them with one launch sequence per tick.
 * ------------------------------------------------------------------------- */
int32_t nn_pool_create(int64_t budget_bytes, nn_pool **out);
/* nn_pool_create with flags.  NN_POOL_OWN_QUEUE: the pool's stream gets a
 * hardware queue of its own (a stream whose CU mask covers every CU), so that
 * pools ticked concurrently from several threads (lanes) never share a queue;
 * with plain streams the device maps streams onto a few shared hardware
 * queues, and two pools on one queue serialize. */
#define NN_POOL_OWN_QUEUE 1
int32_t nn_pool_create_ex(int64_t budget_bytes, int32_t flags, nn_pool **out);
int32_t nn_pool_destroy(nn_pool *pool);
/* Bytes of device memory a scene of n particles (n_elastic elastic) occupies. */
int64_t nn_pool_footprint(int64_t n, int64_t n_elastic, const nn_params *params);
/* Admit a scene: uploaded and initialized now (on a side stream, overlapping a
 * running tick), joins the batch at the next tick.  NN_E_BUSY if over budget
 * (the bytes of scenes retired since the last tick count as free). */
int32_t nn_pool_admit(nn_pool *pool, const nn_params *params, const float *x, const float *v,
                      const int32_t *kind, int64_t n, int32_t *slot);
/* One step of every resident scene; per-slot stats into stats[slot] when given
 * (array of nn_pool_capacity() entries).  Returns the number of scenes stepped
 * through *stepped. */
int32_t nn_pool_tick(nn_pool *pool, nn_step_stats *stats, int32_t stats_len, int32_t *stepped);
/* The same tick started on a library thread: returns at once, so the caller
 * can admit, retire or prepare the next requests while the GPU steps the pool.
 * Admission, retirement, nn_pool_steps and nn_pool_info run during a tick (on
 * any thread, synchronous ticks included); nn_pool_read and the launch-shape
 * and profiling calls, which touch the batch, wait for it; nn_pool_tick
 * returns NN_E_BUSY.  nn_pool_sync waits for it and returns what nn_pool_tick
 * would have: its status, the stats of the slots it stepped, the count
 * stepped.  NN_E_BUSY if a tick is already in flight; nn_pool_sync without
 * one is NN_E_STATE.  (The reference steps one scene synchronously; no
 * counterpart.) */
int32_t nn_pool_tick_async(nn_pool *pool);
int32_t nn_pool_sync(nn_pool *pool, nn_step_stats *stats, int32_t stats_len, int32_t *stepped);
int32_t nn_pool_retire(nn_pool *pool, int32_t slot);
/* Number of ticks the scene in `slot` has been stepped (0 while it is still
 * being admitted), -1 for a free slot. */
int64_t nn_pool_steps(nn_pool *pool, int32_t slot);
int32_t nn_pool_read(nn_pool *pool, int32_t slot, int32_t field, void *host, size_t bytes);
int32_t nn_pool_info(nn_pool *pool, int32_t *resident, int32_t *pending, int64_t *bytes_in_use);
int32_t nn_pool_capacity(void);
/* Launch shape used by the sweep kernels (block size, power of two 64..1024);
 * the autotuner sets it.  0 = library default. */
int32_t nn_pool_set_block(nn_pool *pool, int32_t block);
int32_t nn_set_block(nn_ctx *ctx, int32_t block);
/* Workgroup size of one kernel (ids of nn_kernel_name; 0 = the default above).
 * The sweeps' results do not depend on it: per-particle arithmetic and the
 * grouping of the partial sums are fixed by the wavefront, not the workgroup. */
int32_t nn_pool_set_kernel_block(nn_pool *pool, int32_t kernel, int32_t block);
int32_t nn_set_kernel_block(nn_ctx *ctx, int32_t kernel, int32_t block);
/* Gather depth of one kernel: how many neighbor-list entries a lane loads
 * before using them (0 = the kernel's default; 1..16, a power of two).  The
 * neighbor sweeps take 4, 8 or 16 (default 8), cg_matvec 1, 2, 4 or 8 (default
 * 2); other values select the default.  Results do not depend on it. */
int32_t nn_pool_set_kernel_depth(nn_pool *pool, int32_t kernel, int32_t depth);
int32_t nn_set_kernel_depth(nn_ctx *ctx, int32_t kernel, int32_t depth);

/* Per-kernel timing with HIP events recorded around launches on the stream the
 * kernel runs on (no counterpart in the reference).  enable != 0 turns it on
 * for every launch of every kernel and clears the totals.  The _mask forms
 * time only the kernel ids whose bit is set (0 = off), and of those only every
 * stride-th launch, so that the other launches run without events between
 * them.  nn_*_kernel_times copies the accumulated milliseconds and the number
 * of timed launches of kernel ids 0..n-1 (names from nn_kernel_name,
 * NN_KERNEL_COUNT ids); nn_*_kernel_launches the number of launches of every
 * kernel made while timing was on, timed or not, masked or not. */
#define NN_KERNEL_COUNT 21
const char *nn_kernel_name(int32_t id);
int32_t nn_profile(nn_ctx *ctx, int32_t enable);
int32_t nn_profile_mask(nn_ctx *ctx, uint32_t mask, int32_t stride);
int32_t nn_kernel_times(nn_ctx *ctx, double *ms, int64_t *launches, int32_t n);
int32_t nn_kernel_launches(nn_ctx *ctx, int64_t *seen, int32_t n);
int32_t nn_pool_profile(nn_pool *pool, int32_t enable);
int32_t nn_pool_profile_mask(nn_pool *pool, uint32_t mask, int32_t stride);
int32_t nn_pool_kernel_times(nn_pool *pool, double *ms, int64_t *launches, int32_t n);
int32_t nn_pool_kernel_launches(nn_pool *pool, int64_t *seen, int32_t n);
/* The launch and event latency contained in a timed launch's interval (no
 * counterpart in the reference): the mean interval, in ms, between the start
 * and stop events of an empty kernel bracketed as the timer brackets a timed
 * launch, over `reps` launches on a stream of its own.  A kernel's duration is
 * its timed interval minus this. */
int32_t nn_event_overhead(int32_t reps, double *ms);

/* ------------------------------------------------------------------------
 * Surface reconstruction: the particle-to-mesh path of SRTool
 * (SPH/surf_recon/srtool.py), a dual-contouring mesh of the particles'
 * anisotropic smooth field.
 *
 * The field is accumulated in 2^-32 fixed point, so a run's results do not
 * depend on the order the GPU adds the particles' contributions; mesh vertices
 * are numbered in grid order of their cubes (x slowest, z fastest), triangles
 * in grid order of the cube that emits them.  The library allocates device
 * memory in nn_surface_run* when the grid grows.  An nn_surface is not
 * thread-safe; distinct ones are independent, each on its own stream. 
 * ------------------------------------------------------------------------ */
typedef struct nn_surface nn_surface;

typedef struct nn_surface_params {
  double origin[3];          /* bounding_box_extent[0] (SRTool.__init__, srtool.py:887-899) */
  double voxel_size;         /* voxel_size (:889) */
  double particle_radius;    /* particle_radius (:890): volume = 4/3 pi r^3 (:945) */
  double smooth_radius;      /* dual_contouring(smooth_radius) (:1264-1271) */
  double isolevel;           /* dual_contouring(isolevel) */
  int32_t anisotropic;       /* use_smooth_fluid_field (default 1): add the anisotropic field */
  int32_t smooth_iterations; /* smooth_sdf(iteration) (:1425), 1 in dual_contouring (:1289); 0..64 */
  int32_t normals;           /* record_normals (:891): per-vertex normals (process_normal, :1384-1410) */
  int32_t velocity;          /* record_velocity (:892): velocity channels (:1160-1190, :1352-1381) */
} nn_surface_params;

/* SRTool(...) (srtool.py:887-991).  NN_E_UNSUPPORTED when the anisotropic box
 * of 4 * smooth_radius / voxel_size + 2 vertices exceeds 32 per axis. */
int32_t nn_surface_create(const nn_surface_params *params, nn_surface **out);
int32_t nn_surface_destroy(nn_surface *s);
/* Cap on the device bytes a run's grids may take (0 = none): NN_E_OOM above it. */
int32_t nn_surface_set_budget(nn_surface *s, int64_t bytes);
/* SRTool.dual_contouring(isolevel, smooth_radius, num_particles, pos)
 * (srtool.py:1264-1304) on n host positions (x, y, z float triples).
 * NN_E_OUT_OF_GRID: a position lies below the box origin's voxel or is not
 * finite; NN_E_CELL_OVERFLOW: more than 64 particles in a voxel
 * (max_num_particle_per_voxel, :685); NN_E_NAN: a field contribution is not
 * finite or exceeds 2^20. */
int32_t nn_surface_run(nn_surface *s, const float *pos, int64_t n);
/* The same on positions already in device memory (n x 3 float, caller order). */
int32_t nn_surface_run_device(nn_surface *s, const float *device_pos, int64_t n);
/* The same on the fluid particles of a context, where its last step left
 * them in device memory: the mesh of the simulation's current state without a
 * round trip through the host (the reference passes meta.pd.x to SRTool).
 * Fluid particles must be numbered 0 .. n_fluid - 1, as the reference numbers
 * them (SPH/main.py:627); NN_E_UNSUPPORTED otherwise.  With params.velocity
 * the fluid's velocities feed the velocity channels. */
int32_t nn_surface_run_context(nn_surface *s, nn_ctx *ctx);
/* dual_contouring(..., velocity) with record_velocity (params.velocity set):
 * the particles' velocities (n x 3) are splatted with the poly6 weight times
 * voxel_size^3 (generate_smooth_fluid_field_with_velocity_impl, :1160-1190) and
 * interpolated at the mesh vertices (process_velocity, :1378-1381). */
int32_t nn_surface_run_velocity(nn_surface *s, const float *pos, const float *vel, int64_t n);
/* mesh_velocity of the last nn_surface_run_velocity (export_mesh's vx, vy, vz). */
int32_t nn_surface_mesh_velocity(nn_surface *s, float *out);
/* num_vertices / num_indices of the last run and its grid (srtool.py:917-918):
 * out[0..2] first grid vertex, out[3..5] grid dims (vertices), out[6] vertices,
 * out[7] triangles, out[8] written 32^3 blocks, out[9..11] first voxel of the
 * particle hash, out[12..14] its dims, out[15] the particles meshed; *sd = the
 * smoothing's sd (:1437-1456).  out holds 16 entries. */
int32_t nn_surface_info(nn_surface *s, int64_t *out, double *sd);
/* mesh_vertex, mesh_index as triangles, mesh_normal (export_mesh, :1802-1838);
 * normals may be NULL, and must be NULL unless params.normals was set. */
int32_t nn_surface_mesh(nn_surface *s, float *vertices, int32_t *faces, float *normals);
/* voxel_vertex_value over the grid (dims from nn_surface_info, z fastest):
 * which 0 = after the two field passes, 1 = after smoothing (what the
 * contouring reads); 0 outside written blocks. */
int32_t nn_surface_field(nn_surface *s, int32_t which, float *out, size_t count);
/* G (row-major 3x3) and x_bar per particle of the last anisotropic run
 * (pre_process_data, srtool.py:1007-1112). */
int32_t nn_surface_aniso(nn_surface *s, float *G, float *xbar, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* NNSPH_H */
