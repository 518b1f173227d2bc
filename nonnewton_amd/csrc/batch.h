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
int hip_fail(hipError_t e, const char* what);

struct SceneHost {
  nn_params params{};
  int n = 0, n_fluid = 0, n_elastic = 0;
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
  bool open = false;    // a start event was recorded for the launch in flight
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
  int layout(const std::vector<const Batch*>& from);
  int upload(int k, const float* x, const float* v, const int* kind);
  int init_scenes(const std::vector<int>& which);  // search, rest state, boundary volumes
  int search();
  int step(std::vector<nn_step_stats>* stats);
  int download(int k, int field, void* host, size_t bytes);
  int upload_field(int k, int field, const void* host, size_t bytes);
  int neighbors(int k, int* nbr, int* cnt, int cap);
  int refresh_scene_table();  // after a parameter change
  int synchronize();
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
  int C_total_ = 0;
  int scan_blocks_ = 0;
  SceneState* h_state_ = nullptr;  // pinned
  int* h_active_ = nullptr;        // pinned
  bool any_fluid_ = false, any_elastic_ = false, any_weiler_ = false, any_strain_ = false, any_div_ = false;
  bool any_warm_cg_ = false;
  int max_iter_[3] = {1, 1, 1};

  template <class T>
  int alloc(T*& p, size_t count);
  void release();
  int search_impl();
  int solver_loop(int which);
  int build_stress_set(const std::vector<int>& which);
  int gather_state();
  void swap_sort_buffers();
};

}  // namespace nn
