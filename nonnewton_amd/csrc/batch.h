// batch.h -- host side of libnnsph: scenes, their device layout and the step
// driver.  A Batch holds one or many independent scenes in one concatenated
// particle array (see engine.h); the single-scene context of the C-ABI is a
// Batch of one, the pool of the serving form is a Batch whose membership
// changes between ticks.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../include/nnsph.h"
#include "engine.h"

namespace nn {

void set_error(const std::string& msg);
class Batch;
// surface reconstruction of scene k's fluid particles, as they lie in the
// batch's device arrays (surface.hip)
int surface_run_scene(nn_surface* s, Batch& b, int k);
int hip_fail(hipError_t e, const char* what);

struct SceneHost {
  nn_params params{};
  int n = 0, n_fluid = 0, n_elastic = 0, n_rigid = 0;
  // rigid bodies: original ids of each body in increasing order (body b's ids
  // are body_ids[body_off[b] .. body_off[b + 1])), rest centres of mass
  int nbody = 0;
  std::vector<int> body_ids, body_off;
  std::vector<float> rest_cm;  // 3 per body
  int r_begin = 0, nr = 0;     // rigid-set range in its batch (each body padded to 64)
  int body_begin = 0;          // first entry of the batch's body table
  int gn[3] = {0, 0, 0};
  long ncells = 0;  // grid cells (one more is reserved for out-of-grid particles)
  // layout inside its current batch
  int p_begin = 0, cell_begin = 0, s_begin = 0, ns = 0;
  SceneDev dev{};
  // stress set in original ids (built at init from the first search)
  std::vector<int> s_pid, rest, rest_cnt;  // rest: ns * MAX_REST ids, row-major per member
  std::vector<int> s_of_host;               // id -> stress-set index (scene-local) or -1
  bool initialized = false;
  bool stepped = false;
  int last_iter[3] = {1, 1, 1};
  nn_step_stats last{};
};

// derive the FP32 constants of a scene from its parameters (no layout fields)
int derive_scene(SceneHost& s, std::string& err);
// rest centres of mass of the rigid bodies from the initial positions x
// (after derive_scene and the kinds' body tables)
void rigid_rest_centres(SceneHost& s, const float* x);
// least float T with sqrt_rn(T) >= h
float neighbor_threshold(float h);
// device bytes of one scene of n particles, ns stress-set members, c cells
long long scene_footprint(long long n, long long ns, long long cells);

// Per-kernel durations from HIP events recorded around each launch on the
// batch stream (enabled on request; read after the step's final sync).
enum KernelId {
  K_CELL_COUNT, K_SCAN, K_BUCKET, K_PLACE, K_BUILD_LISTS, K_BVOL, K_CLEAR_ACC, K_DENSITY_ALPHA,
  K_DIV_UPDATE, K_DIV_SOURCE, K_STRAIN, K_CG_INIT, K_CG_MATVEC, K_CG_UPDATE, K_NONPRESSURE,
  K_PRESS_SOURCE, K_PRESS_UPDATE, K_ELASTIC_ROT_STRESS, K_ELASTIC_FORCE, K_ADVECT, K_CG_GUESS, K_COUNT
};
const char* kernel_name(int id);

struct KernelTimer {
  bool on = false;
  unsigned mask = ~0u;  // kernel ids timed while on
  int stride = 1;       // time every stride-th launch of each timed kernel
  bool open = false;    // a start event was recorded for the launch in flight
  long long seen[K_COUNT] = {};  // launches of each kernel while on (timed or not)
  double total_ms[K_COUNT] = {};
  long long count[K_COUNT] = {};
  std::vector<hipEvent_t> events;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> pending;  // kernel id, index of its start event
  ~KernelTimer();
  int begin(int id, hipStream_t st);
  int end(hipStream_t st);
  int collect();  // after a stream sync
  void reset();
};

class Batch {
 public:
  explicit Batch(hipStream_t stream);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // membership: scenes are owned by the caller (shared_ptr)
  std::vector<std::shared_ptr<SceneHost>> scenes;

  // Lay the current scenes out and allocate device memory.  When `from` is
  // given, every scene already present in `from` (same SceneHost object) has
  // its persistent state copied from there; other scenes must be uploaded.
  // The arrays get `slack` times the room the scenes need (and room for
  // `scene_capacity` scenes), so that append_in_place can add scenes later.
  int layout(const std::vector<const Batch*>& from, double slack = 1.0, int scene_capacity = 0);
  // Change membership without moving any scene: `members` lists the scenes
  // kept, in their present order, followed by new scenes, which are placed
  // after the last used slot, cell and stress slot and get their state from
  // `from`.  Retired scenes leave unused ranges whose waves map to no scene.
  // Returns LAYOUT_NEEDED (nothing changed) when the order or the room does
  // not allow it.
  static constexpr int LAYOUT_NEEDED = 2;
  int append_in_place(const std::vector<std::shared_ptr<SceneHost>>& members, const std::vector<const Batch*>& from);
  // When set, arrays replaced by layout() are not freed but handed out by
  // take_deferred(), to be freed off the thread that steps the batch.
  bool defer_free_ = false;
  void take_deferred(std::vector<void*>& out);
  int upload(int k, const float* x, const float* v, const int* kind);
  int init_scenes(const std::vector<int>& which);  // search, rest state, boundary volumes
  int search();
  int step(std::vector<nn_step_stats>* stats);
  int download(int k, int field, void* host, size_t bytes);
  int upload_field(int k, int field, const void* host, size_t bytes);
  int neighbors(int k, int* nbr, int* cnt, int cap);
  int refresh_scene_table();  // after a parameter change
  int synchronize();
  DevRef dref();  // the current Dev for a launch
  long long bytes() const { return bytes_; }
  int block = 256;             // default workgroup size of the sweeps
  int kblock[K_COUNT] = {};    // per-kernel workgroup size (0 = block), set by the autotuner
  int kdepth[K_COUNT] = {};    // per-kernel gather depth (0 = the kernel's default), set by the autotuner
  int blk(int id) const { return kblock[id] ? kblock[id] : block; }
  int dep(int id) const { return kdepth[id]; }
  hipStream_t stream;
  Dev d{};
  KernelTimer timer;

 private:
  struct Alloc {
    void* p;
    size_t bytes;
  };
  std::vector<Alloc> allocs_;
  long long bytes_ = 0;
  int C_total_ = 0;  // cells in use
  int C_cap_ = 0;    // cells allocated
  int scene_cap_ = 0;
  int n_cg_blocks_ = 0;  // entries of d.cg_blocks
  std::vector<void*> deferred_;
  int scan_blocks_ = 0;
  SceneState* h_state_ = nullptr;  // pinned, h_state_cap_ entries
  int h_state_cap_ = 0;
  int* h_active_ = nullptr;        // pinned
  bool any_fluid_ = false, any_elastic_ = false, any_weiler_ = false, any_strain_ = false, any_div_ = false;
  bool any_warm_cg_ = false, any_rigid_ = false;
  int max_body_ = 0, max_coupling_ = 0;  // over the scenes with rigid bodies
  int NB_cap_ = 0;                       // body-table entries allocated
  int rigid_step();
  int max_iter_[3] = {1, 1, 1};

  int alloc_arena(std::vector<std::pair<void**, size_t>>& req);
  void release();
  int search_impl(int all = 1);
  int complete_lists();
  // neighbor lists: 0 none valid (new layout), 1 built by a step's search
  // (static particles without lists; that search's positions in pos_alt),
  // 2 complete
  int lists_state_ = 0;
  int solver_loop(int which);
  int cg_loop();
  int build_stress_set(const std::vector<int>& which);
  int gather_state();
  void swap_sort_buffers();
  void derive_members();
  int upload_tables(const Dev& nd);
  int copy_scene_state(const Dev& nd, const SceneHost& s, const Batch* from, int ob, int osb, int ons);
};

}  // namespace nn
