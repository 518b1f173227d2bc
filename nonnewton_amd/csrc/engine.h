// engine.h -- device data layout and kernel interfaces of libnnsph.
//
// Layout in HBM (one concatenated, offset-indexed array for all scenes of a batch):
//  * Particle slots [0, P): scene s owns [p_begin, p_begin + n), p_begin a multiple
//    of 64 so every wavefront belongs to exactly one scene; slots past n up to the
//    next multiple of 64 are padding.  Inside a scene particles are kept in cell
//    order (cell key, then original index), re-sorted at every neighbor search.
//  * pos = float4 (x, y, z, signed volume): +V for fluid, -V for solids, so one
//    16-byte gather per neighbor yields position, volume and material.
//  * vel = float4 (vx, vy, vz, 0); acc = float4.
//  * info = kind << 28 | original index (scene-local); -1 on padding slots.
//  * Neighbor lists in groups of four positions: position k < 60 of particle i
//    is word 4 ((k / 4) P + i) + k % 4 of nbr (nb_word), so a group is one
//    16-byte int4 and particles i, i+1 have adjacent groups: a lane reading
//    its list 4 positions per load, and the four lanes of a particle in the
//    four-lane sweeps (lane q takes positions q, q+4, ...), read contiguous
//    memory, and the list builder writes whole groups.  nnbr[i] holds the
//    count.
//  * Original-index ("id") space: inv[p_begin + id] = local slot; x0 (rest
//    position, w = rest volume) and the stress-set map live there.
//  * Cells: scene s owns global cell keys [cell_begin, cell_begin + gn0*gn1*gn2],
//    key = cell_begin + (cx * gn1 + cy) * gn2 + cz; the extra last key parks
//    particles that left the grid (reported as an error, never searched).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nn {

constexpr int MAX_NB = 60;    // NeighborhoodSearchSparse.max_num_neighbors, SPH/main.py:777
constexpr int MAX_PIC = 50;   // max_num_particles_in_grid, SPH/main.py:778
constexpr int MAX_REST = 50;  // Elasticity.max_num_neighbors, SPH/main.py:1192
constexpr int WAVE = 64;
// Workgroup sizes of the per-scene reductions.  Smaller workgroups find room
// on a CU sooner while other pools' sweeps occupy the device (lanes): with 8
// waves per CG-update workgroup and 256 threads per scene finish the mixed
// batch ticks in 3.11 ms in three lanes against 3.23 with 16 waves and 1,024
// threads (3.54 against 3.59 in one pool), with the same bits.
#ifndef NN_UPD_WAVES
#define NN_UPD_WAVES 8
#endif
#ifndef NN_FINISH_THREADS
#define NN_FINISH_THREADS 256
#endif
// waves of one CG-update workgroup (its cg_blocks entry covers that many
// waves of one scene)
constexpr int CG_UPD_WAVES = NN_UPD_WAVES;

enum { KIND_FLUID = 0, KIND_STATIC = 1, KIND_ELASTIC = 2, KIND_RIGID = 3 };
constexpr int MAX_BODIES = 4096;  // rigid bodies per scene
enum {
  ST_NEIGHBOR_OVERFLOW = 1,
  ST_CELL_OVERFLOW = 2,
  ST_OUT_OF_GRID = 4,
  ST_NAN = 8,
  ST_REST_OVERFLOW = 16,
  ST_NOCONV = 32
};
enum {
  SF_SURFACE_TENSION = 1,
  SF_VISCOSITY = 2,
  SF_NON_NEWTONIAN = 4,
  SF_WEILER = 8,
  SF_DIVERGENCE = 16,
  SF_COMPENSATED = 32,
  SF_HAS_FLUID = 64,
  SF_HAS_ELASTIC = 128,
  SF_WARM_CG = 256,  // Weiler CG starts from v + (x - v) of the previous solve
  SF_HAS_RIGID = 512
};

// Per-scene constants (device table).  Floats are the reference's FP32
// constants, derived in double on the host and rounded once.
struct SceneDev {
  int p_begin, n, n_fluid, n_elastic;
  int cell_begin, gn0, gn1, gn2;
  int shift;  // p_begin minus the unpadded particle prefix
  int wave_begin, wave_end;
  int s_begin, ns;  // stress-set range
  int flags, nn_method;
  int nbody, body_begin, coupling;  // rigid bodies: count, first entry of d.bodies, substeps per step
  int max_iter_v, max_iter_p, max_iter_cg;
  float h, thr2, kW, kG, W0;
  float dt, inv_dt, inv_dt2, rho0, mass;
  float nu, bnu, gx, gy, gz;
  float lo0, lo1, lo2, hi0, hi1, hi2;
  float diam2, Wd, visc_eps, sigma, opc, sticky, eps;
  float C_diag, C_off, C_shear;
  float nn_n, nn_K, nn_e0, nn_einf, nn_gc, nn_muc, nn_tauy, nn_max;
  float ten_nu, ten_bnu, sig_m;  // 10*nu, 10*bnu, (sigma/m)*m
  float vol0;  // volume0 = 0.8 (2r)^3 (:607)
  double eta_v, eta_p, cg_tol2, nf_d, rho0_d;
};

// A rigid body (RigidBodySolver, SPH/main.py:1099-1158).  Its particles are
// the rigid-set slots of waves [wave_begin, wave_end) (original ids in r_pid,
// -1 on padding); `rank` is its index among its scene's bodies.
struct BodyDev {
  int scene, rank, wave_begin, wave_end;
  float rest_cm[3];  // centre of mass at construction (:1105)
  float cm[3];       // of the current projection
  float R[9];        // rotation of the current projection, row-major
  float pad;
};

// Per-scene mutable solver state.
struct SceneState {
  // active_cg: the scene's CG runs in this solve (set by the start reduction);
  // cg_stop_at: the update k that stopped it (0 while it iterates), written by
  // the scene's first workgroup of that update
  int active_v, active_p, active_cg, cg_stop_at;
  int iter_v, iter_p, iter_cg, status;
  unsigned counter, max_rho_bits;
  unsigned long long pairs, pairs_fluid;
  double err_v, err_p;
  double cg_bb, cg_rr, cg_theta, pad_d;
  // CG scalars gamma = r.u and alpha, kept by iteration parity: the update of
  // iteration k reads slot (k - 1) & 1 while the scene's first workgroup
  // writes slot k & 1 in the same launch
  double cg_gamma2[2], cg_alpha2[2];
};

// Device pointers of a batch (passed by value to every kernel).
struct XU {
  float3 p;
  float3 u;
};

struct Dev {
  int P;        // particle slots allocated (multiple of 64; the stride of per-pair arrays)
  int W;        // waves allocated = P / 64
  int Pn;       // slots in use: sweeps are launched over [0, Pn)
  int C;        // total cells
  int NS;       // stress-set slots allocated (each scene's range padded to 64; the rest-list stride)
  int SW;       // stress-set waves allocated = NS / 64
  int NSn;      // stress-set slots in use
  int NR;       // rigid-set slots allocated (each body's range padded to 64)
  int NRn;      // rigid-set slots in use
  int nbodies;  // entries of bodies in use
  // A wave (or stress-set wave) that belongs to no scene -- the range of a
  // retired scene, or slack past the last one -- has scene index -1 and is
  // skipped by every sweep.
  int nscenes;
  const SceneDev* scenes;
  SceneState* state;
  const int* wave_scene;
  const int* s_wave_scene;  // stress-set wave -> scene
  float4* pos;  // current
  float4* pos_alt;
  float4* vel;  // current
  float4* vel_alt;
  int* info;
  int* info_alt;
  float4* acc;
  float* rho;
  // (src, alpha): density_adv of the reference (rho-dot in the divergence
  // solve) and the DFSPH factor, side by side: the solver updates gather both
  float2* sa;
  int* nbr;
  int* nnbr;
  int* key;
  int* rank;
  int* inv;
  int* bucket;
  int* cell_count;
  int* cell_start;
  int* scan_tmp;
  double* partials;  // 8 arrays of 4W: wave partials (per 64 particles, or per 16 for four-lane sweeps)
  const int2* cg_blocks;  // CG update workgroups: (scene, first wave), up to CG_UPD_WAVES waves of one scene each
  int* cg_list;           // 2 x W: waves of the scenes still iterating, by iteration parity
  int* cg_list_n;         // 2 counts
  int* n_active;     // [0] divergence, [1] pressure, [2] cg
  // Weiler CG: 3-vectors packed at 12 bytes (the sweeps stream them and the
  // matvec gathers u and the position of every neighbor)
  float3* wx;  // iterate x
  float3* wr;  // residual r
  float3* ww;  // w = A u
  float3* wp;  // direction p
  float3* ws;  // s = A p
  // what the matvec gathers of a neighbor, in one 24-byte record: its position
  // during the solve (written by cg_init for every particle of a Weiler scene)
  // and u = M r, the preconditioned residual (fluid particles)
  XU* xu;
  float4* wd0;  // inverse diagonal block: (d00, d01, d02, d11)
  float2* wd1;  //                          (d12, d22)
  float* spair; // per-pair scalar s_ij of the Weiler operator, at the list word of the pair (nb_word)
  float4* vdiff; // x - v of the previous Weiler solve (initial guess; moves with the particle)
  float4* vdiff_alt;  // its sort buffer
  float* mu;    // dynamic viscosity per particle
  float* gam;   // strain-rate norm
  float* eta;   // kinematic viscosity from the model
  // elasticity (id space and stress set)
  float4* x0;   // rest position, w = rest volume
  int* s_of;    // id space -> stress-set index (scene-local) or -1
  int* s_pid;   // stress-set -> local id (-1 on padding)
  int* rest;    // rest[k * NS + q]: local id of the k-th rest neighbor
  int* rest_cnt;
  float4* rs;   // 4 per member: rows of R with stress 0..2 in w, then (s3, s4, s5, 0)
  // rigid bodies (rigid set: each body's particles by original id, padded to 64)
  const int* r_pid;        // rigid-set slot -> local id (-1 on padding)
  const int* r_wave_body;  // rigid-set wave -> entry of bodies (-1: none)
  BodyDev* bodies;
  double* rpartials;       // 9 arrays of NR / 64: rigid-set wave partials
};

// ---------------------------------------------------------------------------
// host-side launchers (kernels_*.hip)
// ---------------------------------------------------------------------------
// What a launcher needs of a batch: its Dev, which the kernels take by value.
// (A pointer to a device-resident copy would shrink the kernel arguments from
// about 600 bytes to 8, but the compiler then cannot tell that the pointers
// read from it address global memory and emits flat loads; measured on this
// device, the sweeps ran about 10% slower that way.)
struct DevRef {
  const Dev& h;
};

// Start and stop events for the next kernel launch, set by the per-kernel
// timer around a timed launch.  The launch carries them (hipExtLaunchKernel
// stamps them with the dispatch's own start and end times), so timing puts no
// packets between kernels; the first launch after they are set takes them.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  bool launched = false;
};
LaunchEvents& launch_events();  // of the calling thread
// Flags of the timing events.  They only stamp times: without the
// system-scope fence an event record would otherwise add (an L2 write-back and
// invalidate on every XCD), a timed launch perturbs the launches around it less
// (measured: about 2 us less idle per timed launch, 1% of a pooled tick).
constexpr unsigned TIMING_EVENT_FLAGS = hipEventDisableSystemFence;

#define NN_LAUNCH(kern, grid, blk, shm, st, ...)                                                   \
  do {                                                                                          \
    LaunchEvents& _le = launch_events();                                                        \
    if (_le.start) {                                                                            \
      hipExtLaunchKernelGGL(kern, grid, blk, shm, st, _le.start, _le.stop, 0u, __VA_ARGS__);   \
      _le.start = _le.stop = nullptr;                                                           \
      _le.launched = true;                                                                      \
    } else {                                                                                    \
      hipLaunchKernelGGL(kern, grid, blk, shm, st, __VA_ARGS__);                                \
    }                                                                                           \
  } while (0)

void launch_fill_i32(int* p, long n, int v, hipStream_t st);
void launch_fill_f4(float4* p, long n, float4 v, hipStream_t st);
void launch_scan(const DevRef& d, int n, hipStream_t st);  // cell_start = exclusive scan(cell_count)
// exclusive scan of in[0..n) into out; tmp holds ceil(n / 2048) ints
void launch_scan_i32(const int* in, int n, int* tmp, int* out, hipStream_t st);
constexpr int SCAN_TILE_ITEMS = 2048;

void launch_cell_count(const DevRef& d, int block, hipStream_t st);
void launch_bucket(const DevRef& d, int block, hipStream_t st);
void launch_place(const DevRef& d, int block, hipStream_t st);
void launch_build_lists(const DevRef& d, int block, int all, hipStream_t st);  // all = 0: no static lists
void launch_bvol(const DevRef& d, int block, int dynamic, hipStream_t st);

void launch_density_alpha(const DevRef& d, int block, int depth, hipStream_t st);
void launch_div_update(const DevRef& d, int block, int depth, hipStream_t st);
void launch_div_source(const DevRef& d, int block, int depth, hipStream_t st);
void launch_solver_begin(const DevRef& d, int which, hipStream_t st);
void launch_strain(const DevRef& d, int block, int depth, hipStream_t st);
void launch_cg_init(const DevRef& d, int block, hipStream_t st);
void launch_cg_guess(const DevRef& d, int block, hipStream_t st);
void launch_cg_matvec(const DevRef& d, int block, int depth, int start, hipStream_t st);  // depth 1, 2, 4, 8
void launch_cg_matvec_list(const DevRef& d, int depth, int k, int grid, hipStream_t st);  // iteration k
void launch_cg_update(const DevRef& d, int nblocks, int k, hipStream_t st);  // iteration k >= 1
void launch_nonpressure(const DevRef& d, int block, int depth, hipStream_t st);
void launch_press_source(const DevRef& d, int block, int depth, int reduce, hipStream_t st);
void launch_press_update(const DevRef& d, int block, int depth, hipStream_t st);
void launch_clear_acc(const DevRef& d, int block, hipStream_t st);
void launch_elastic_init(const DevRef& d, int block, hipStream_t st);
void launch_elastic_rot_stress(const DevRef& d, int block, hipStream_t st);
void launch_elastic_force(const DevRef& d, int block, hipStream_t st);
// rigid body r of every scene, substep c (RigidBodySolver.substep, :1112-1115):
// ground collision, centre of mass, A and its rotation, projection
void launch_rigid_ground(const DevRef& d, int r, int c, hipStream_t st);
void launch_rigid_sum(const DevRef& d, int pass, int r, int c, hipStream_t st);
void launch_body_finish(const DevRef& d, int pass, int r, int c, hipStream_t st);
void launch_rigid_place(const DevRef& d, int r, int c, hipStream_t st);
void launch_advect(const DevRef& d, int block, hipStream_t st);
void launch_step_begin(const DevRef& d, hipStream_t st);
void launch_scene_finish(const DevRef& d, int which, hipStream_t st);  // 0 div, 1 press, 2 cg start, 3 cg iteration, 4 cg guess scale

}  // namespace nn
