// capi.hip -- extern "C" entry points of libnnsph.so (declared in include/nnsph.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include <hip/hip_ext.h>

#include "batch.h"

namespace nn {
const char* last_error_text();
}

using nn::Batch;
using nn::SceneHost;

struct nn_ctx {
  hipStream_t stream = nullptr;
  std::unique_ptr<Batch> batch;
  std::shared_ptr<SceneHost> scene;
  int64_t max_particles = 0;
  bool uploaded = false;
};

// Kinds and rigid bodies of a scene: kind NN_KIND_RIGID + b is a particle of
// body b; bodies are numbered 0 .. B-1 without gaps.
static int count_kinds(SceneHost& s, const int32_t* kind, int64_t n) {
  s.n = (int)n;
  s.n_fluid = s.n_elastic = s.n_rigid = 0;
  s.nbody = 0;
  for (int64_t i = 0; i < n; i++) {
    const int k = kind[i];
    if (k == NN_KIND_FLUID) s.n_fluid++;
    else if (k == NN_KIND_ELASTIC) s.n_elastic++;
    else if (k >= NN_KIND_RIGID && k < NN_KIND_RIGID + nn::MAX_BODIES) {
      s.n_rigid++;
      s.nbody = std::max(s.nbody, k - NN_KIND_RIGID + 1);
    } else if (k != NN_KIND_STATIC) {
      nn::set_error("kind must be NN_KIND_FLUID, NN_KIND_STATIC, NN_KIND_ELASTIC or NN_KIND_RIGID + body (body < 4096)");
      return NN_E_ARG;
    }
  }
  s.body_off.assign(s.nbody + 1, 0);
  for (int64_t i = 0; i < n; i++)
    if (kind[i] >= NN_KIND_RIGID) s.body_off[kind[i] - NN_KIND_RIGID + 1]++;
  s.nr = 0;
  for (int b = 0; b < s.nbody; b++) {
    if (s.body_off[b + 1] == 0) {
      nn::set_error("rigid bodies must be numbered 0 .. B-1 without gaps");
      return NN_E_ARG;
    }
    s.nr += (s.body_off[b + 1] + 63) / 64 * 64;
    s.body_off[b + 1] += s.body_off[b];
  }
  s.body_ids.assign(s.n_rigid, 0);
  std::vector<int> fill(s.body_off.begin(), s.body_off.end() - (s.nbody > 0 ? 1 : 0));
  for (int64_t i = 0; i < n; i++)
    if (kind[i] >= NN_KIND_RIGID) s.body_ids[fill[kind[i] - NN_KIND_RIGID]++] = (int)i;
  return NN_OK;
}


extern "C" {

int32_t nn_version(void) { return NNSPH_ABI_VERSION; }

const char* nn_last_error(void) { return nn::last_error_text(); }

int32_t nn_device_info(char* name, int32_t name_len, int64_t* hbm_bytes, int32_t* num_cus) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nn::hip_fail(e, "hipGetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return nn::hip_fail(e, "hipGetDeviceProperties");
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
  }
  if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
  if (num_cus) *num_cus = prop.multiProcessorCount;
  return NN_OK;
}

int32_t nn_create(const nn_params* params, int64_t max_particles, nn_ctx** out) {
  if (!params || !out || max_particles <= 0 || max_particles > (1LL << 28)) {
    nn::set_error("nn_create: bad arguments");
    return NN_E_ARG;
  }
  *out = nullptr;
  auto ctx = new (std::nothrow) nn_ctx();
  if (!ctx) return NN_E_OOM;
  ctx->scene = std::make_shared<SceneHost>();
  ctx->scene->params = *params;
  std::string err;
  int rc = nn::derive_scene(*ctx->scene, err);
  if (rc < 0) {
    nn::set_error(err);
    delete ctx;
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return nn::hip_fail(e, "hipStreamCreate");
  }
  ctx->batch.reset(new Batch(ctx->stream));
  ctx->max_particles = max_particles;
  *out = ctx;
  return NN_OK;
}

int32_t nn_destroy(nn_ctx* ctx) {
  if (!ctx) return NN_E_ARG;
  ctx->batch.reset();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return NN_OK;
}

int32_t nn_upload(nn_ctx* ctx, const float* x, const float* v, const int32_t* kind, int64_t n) {
  if (!ctx || !x || !kind || n <= 0) {
    nn::set_error("nn_upload: bad arguments");
    return NN_E_ARG;
  }
  if (n > ctx->max_particles) {
    nn::set_error("nn_upload: more particles than max_particles");
    return NN_E_ARG;
  }
  SceneHost& s = *ctx->scene;
  int rc = count_kinds(s, kind, n);
  if (rc < 0) return rc;
  std::string err;
  rc = nn::derive_scene(s, err);
  if (rc < 0) {
    nn::set_error(err);
    return rc;
  }
  nn::rigid_rest_centres(s, x);
  s.ns = 0;
  s.s_pid.clear();
  s.rest.clear();
  s.rest_cnt.clear();
  ctx->batch->scenes = {ctx->scene};
  rc = ctx->batch->layout({});
  if (rc < 0) return rc;
  rc = ctx->batch->upload(0, x, v, kind);
  if (rc < 0) return rc;
  ctx->uploaded = true;
  return NN_OK;
}

int32_t nn_init(nn_ctx* ctx) {
  if (!ctx || !ctx->uploaded) {
    nn::set_error("nn_init before nn_upload");
    return NN_E_STATE;
  }
  return ctx->batch->init_scenes({0});
}

int32_t nn_step(nn_ctx* ctx, int32_t nsteps, nn_step_stats* out) {
  if (!ctx || !ctx->uploaded) {
    nn::set_error("nn_step before nn_upload/nn_init");
    return NN_E_STATE;
  }
  int rc = NN_OK;
  std::vector<nn_step_stats> st;
  for (int k = 0; k < nsteps; k++) {
    rc = ctx->batch->step(&st);
    if (rc < 0) break;
  }
  if (out && !st.empty()) *out = st[0];
  return rc;
}

int32_t nn_search(nn_ctx* ctx) {
  if (!ctx || !ctx->uploaded) return NN_E_STATE;
  return ctx->batch->search();
}

int32_t nn_download(nn_ctx* ctx, int32_t field, void* host, size_t bytes) {
  if (!ctx || !ctx->uploaded || !host) return NN_E_STATE;
  return ctx->batch->download(0, field, host, bytes);
}

int32_t nn_upload_field(nn_ctx* ctx, int32_t field, const void* host, size_t bytes) {
  if (!ctx || !ctx->uploaded || !host) return NN_E_STATE;
  return ctx->batch->upload_field(0, field, host, bytes);
}

int32_t nn_neighbors(nn_ctx* ctx, int32_t* nbr, int32_t* count, int32_t cap) {
  if (!ctx || !ctx->uploaded || !nbr || !count) return NN_E_STATE;
  return ctx->batch->neighbors(0, nbr, count, cap);
}

static int apply_param(SceneHost& s, const char* key, double value) {
  nn_params& p = s.params;
  if (!key) return NN_E_ARG;
  const std::string k(key);
  if (k == "dt") p.dt = value;
  else if (k == "viscosity") p.viscosity = value;
  else if (k == "boundary_viscosity") p.boundary_viscosity = value;
  else if (k == "max_error") p.max_error = value;
  else if (k == "max_error_v") p.max_error_v = value;
  else if (k == "surface_tension") p.surface_tension = value;
  else if (k == "max_iterations") p.max_iterations = (int32_t)value;
  else if (k == "max_iterations_v") p.max_iterations_v = (int32_t)value;
  else if (k == "weiler_tol") p.weiler_tol = value;
  else if (k == "weiler_max_iter") p.weiler_max_iter = (int32_t)value;
  else {
    nn::set_error("nn_set_param: unknown or immutable key " + k);
    return NN_E_ARG;
  }
  std::string err;
  const nn::SceneDev keep = s.dev;
  int rc = nn::derive_scene(s, err);
  if (rc < 0) {
    nn::set_error(err);
    return rc;
  }
  // layout fields are not derived from the parameters
  s.dev.p_begin = keep.p_begin;
  s.dev.shift = keep.shift;
  s.dev.wave_begin = keep.wave_begin;
  s.dev.wave_end = keep.wave_end;
  s.dev.cell_begin = keep.cell_begin;
  s.dev.s_begin = keep.s_begin;
  s.dev.ns = keep.ns;
  s.dev.body_begin = keep.body_begin;
  return NN_OK;
}

int32_t nn_set_param(nn_ctx* ctx, const char* key, double value) {
  if (!ctx) return NN_E_ARG;
  int rc = apply_param(*ctx->scene, key, value);
  if (rc < 0) return rc;
  if (ctx->uploaded) return ctx->batch->refresh_scene_table();
  return NN_OK;
}

int32_t nn_counts(nn_ctx* ctx, int64_t* n, int64_t* n_fluid) {
  if (!ctx) return NN_E_ARG;
  if (n) *n = ctx->scene->n;
  if (n_fluid) *n_fluid = ctx->scene->n_fluid;
  return NN_OK;
}

static int copy_times(const nn::KernelTimer& t, double* ms, int64_t* launches, int32_t n) {
  for (int k = 0; k < n && k < nn::K_COUNT; k++) {
    if (ms) ms[k] = t.total_ms[k];
    if (launches) launches[k] = t.count[k];
  }
  return NN_OK;
}

static int copy_seen(const nn::KernelTimer& t, int64_t* seen, int32_t n) {
  if (!seen) return NN_E_ARG;
  for (int k = 0; k < n && k < nn::K_COUNT; k++) seen[k] = t.seen[k];
  return NN_OK;
}

const char* nn_kernel_name(int32_t id) { return nn::kernel_name(id); }

__global__ void k_empty(int) {}

// The launch and event latency inside a timed launch's interval: an empty
// kernel launched behind another one and bracketed by a start and a stop
// event exactly as the per-kernel timer brackets a timed launch, `reps`
// times on a stream of its own; the mean interval in ms.
int32_t nn_event_overhead(int32_t reps, double* ms) {
  if (reps <= 0 || reps > 100000 || !ms) return NN_E_ARG;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&e0, nn::TIMING_EVENT_FLAGS);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&e1, nn::TIMING_EVENT_FLAGS);
  double total = 0.0;
  for (int r = -8; r < reps && e == hipSuccess; r++) {  // 8 untimed rounds first
    hipLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, st, 0);
    hipExtLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, st, e0, e1, 0u, 1);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (r >= 0) total += t;
  }
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_event_overhead");
  *ms = total / reps;
  return NN_OK;
}

int32_t nn_profile(nn_ctx* ctx, int32_t enable) { return nn_profile_mask(ctx, enable ? 0xFFFFFFFFu : 0u, 1); }

int32_t nn_profile_mask(nn_ctx* ctx, uint32_t mask, int32_t stride) {
  if (!ctx || stride < 1) return NN_E_ARG;
  ctx->batch->timer.on = mask != 0;
  ctx->batch->timer.mask = mask;
  ctx->batch->timer.stride = stride;
  ctx->batch->timer.reset();
  return NN_OK;
}

int32_t nn_kernel_times(nn_ctx* ctx, double* ms, int64_t* launches, int32_t n) {
  if (!ctx) return NN_E_ARG;
  return copy_times(ctx->batch->timer, ms, launches, n);
}

int32_t nn_kernel_launches(nn_ctx* ctx, int64_t* seen, int32_t n) {
  if (!ctx) return NN_E_ARG;
  return copy_seen(ctx->batch->timer, seen, n);
}

int32_t nn_set_block(nn_ctx* ctx, int32_t block) {
  if (!ctx) return NN_E_ARG;
  if (block == 0) block = 256;
  if (block < 64 || block > 1024 || (block & (block - 1))) return NN_E_ARG;
  ctx->batch->block = block;
  return NN_OK;
}

static bool depth_ok(int32_t depth) { return depth == 0 || (depth >= 1 && depth <= 16 && !(depth & (depth - 1))); }

int32_t nn_set_kernel_depth(nn_ctx* ctx, int32_t kernel, int32_t depth) {
  if (!ctx || kernel < 0 || kernel >= nn::K_COUNT || !depth_ok(depth)) return NN_E_ARG;
  ctx->batch->kdepth[kernel] = depth;
  return NN_OK;
}

int32_t nn_set_kernel_block(nn_ctx* ctx, int32_t kernel, int32_t block) {
  if (!ctx || kernel < 0 || kernel >= nn::K_COUNT) return NN_E_ARG;
  if (block != 0 && (block < 64 || block > 1024 || (block & (block - 1)))) return NN_E_ARG;
  ctx->batch->kblock[kernel] = block;
  return NN_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// pool: many resident scenes stepped together
// ---------------------------------------------------------------------------
static constexpr int POOL_CAPACITY = 4096;

struct PoolSlot {
  std::shared_ptr<SceneHost> scene;
  std::unique_ptr<Batch> staging;  // set while the scene waits to join the batch
  long long bytes = 0;
  long long steps = 0;  // ticks this scene took part in
  bool used = false, retired = false;
};

// One staging stream per device for every pool's admissions (upload and
// initialization of a new scene), created at the first admission and kept for
// the life of the process.  The device maps streams onto a few hardware
// queues in creation order; pools create their tick streams first, so this
// one does not queue behind a pool's ticks, as a stream per admission slot
// would once there are more streams than hardware queues.
static int staging_stream(hipStream_t* out) {
  static std::mutex mu;
  static std::vector<std::pair<int, hipStream_t>> streams;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nn::hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(mu);
  for (auto& ds : streams)
    if (ds.first == dev) {
      *out = ds.second;
      return NN_OK;
    }
  hipStream_t st = nullptr;
  e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return nn::hip_fail(e, "hipStreamCreate");
  streams.push_back({dev, st});
  *out = st;
  return NN_OK;
}

struct nn_pool {
  hipStream_t stream = nullptr;
  int64_t budget = 0;
  long long in_use = 0;
  std::unique_ptr<Batch> main;
  std::vector<PoolSlot> slots;
  // tick_mu is held for a whole tick (membership change and step): taken by
  // what touches the batch (downloads, launch shapes, profiling).  mu guards
  // the slots, the budget and membership: admission and retirement take only
  // mu, so they run while a tick steps the batch.  Order: tick_mu, then mu.
  std::mutex tick_mu;
  std::mutex mu;
  int block = 256;
  int kblock[nn::K_COUNT] = {};
  int kdepth[nn::K_COUNT] = {};
  std::vector<int> member_slot;  // batch index -> slot
  // device memory no longer used by the batch, freed by the next admission
  // (hipFree waits for the device, so not on the thread that ticks)
  std::vector<std::unique_ptr<Batch>> graveyard;
  std::vector<void*> dead_arrays;
  // a tick in flight (nn_pool_tick_async): its thread, and what nn_pool_sync returns
  std::thread async;
  bool async_on = false;
  int async_rc = 0, async_stepped = 0;
  std::string async_err;
  std::vector<nn_step_stats> async_stats;
};

// frees what the ticks left behind; called outside the pool lock
static void pool_free_dead(nn_pool* p) {
  std::vector<std::unique_ptr<Batch>> dead;
  std::vector<void*> arrays;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    dead.swap(p->graveyard);
    arrays.swap(p->dead_arrays);
  }
  dead.clear();
  for (void* a : arrays) (void)hipFree(a);
}

extern "C" {

int32_t nn_surface_run_context(nn_surface* s, nn_ctx* ctx) {
  if (!s || !ctx || !ctx->batch) {
    nn::set_error("nn_surface_run_context: bad arguments");
    return NN_E_ARG;
  }
  if (!ctx->uploaded || ctx->batch->scenes.empty()) {
    nn::set_error("nn_surface_run_context: the context holds no particles (nn_upload first)");
    return NN_E_STATE;
  }
  return nn::surface_run_scene(s, *ctx->batch, 0);
}

int32_t nn_pool_capacity(void) { return POOL_CAPACITY; }

// A stream on a hardware queue of its own: a CU mask is a property of the
// queue, so a stream created with one never shares its queue.  The mask is the
// full one a plain stream reports (every CU of the device).
static hipError_t own_queue_stream(hipStream_t* out) {
  hipStream_t tmp = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&tmp, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  uint32_t mask[32] = {};
  e = hipExtStreamGetCUMask(tmp, 32, mask);
  (void)hipStreamDestroy(tmp);
  if (e != hipSuccess) return e;
  uint32_t words = 32;
  while (words > 1 && mask[words - 1] == 0) words--;
  return hipExtStreamCreateWithCUMask(out, words, mask);
}

int32_t nn_pool_create(int64_t budget_bytes, nn_pool** out) { return nn_pool_create_ex(budget_bytes, 0, out); }

int32_t nn_pool_create_ex(int64_t budget_bytes, int32_t flags, nn_pool** out) {
  if (!out || budget_bytes <= 0 || (flags & ~NN_POOL_OWN_QUEUE)) return NN_E_ARG;
  auto p = new (std::nothrow) nn_pool();
  if (!p) return NN_E_OOM;
  hipError_t e = (flags & NN_POOL_OWN_QUEUE) ? own_queue_stream(&p->stream)
                                             : hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete p;
    return nn::hip_fail(e, "hipStreamCreate");
  }
  p->budget = budget_bytes;
  p->main.reset(new Batch(p->stream));
  p->main->defer_free_ = true;
  p->slots.resize(POOL_CAPACITY);
  *out = p;
  return NN_OK;
}

int32_t nn_pool_destroy(nn_pool* p) {
  if (!p) return NN_E_ARG;
  if (p->async_on) {  // a tick in flight finishes first
    p->async.join();
    p->async_on = false;
  }
  if (p->main) p->main->take_deferred(p->dead_arrays);
  pool_free_dead(p);
  p->main.reset();
  for (auto& s : p->slots) {
    s.staging.reset();
  }
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return NN_OK;
}

int64_t nn_pool_footprint(int64_t n, int64_t n_elastic, const nn_params* params) {
  if (!params || n <= 0) return -1;
  SceneHost s;
  s.params = *params;
  s.n = (int)n;
  std::string err;
  if (nn::derive_scene(s, err) < 0) return -1;
  // the stress set is the elastic particles plus a one-particle-wide shell
  const long long ns = n_elastic > 0 ? n_elastic + n_elastic / 2 + 64 : 0;
  return nn::scene_footprint(n, ns, s.ncells);
}

int32_t nn_pool_admit(nn_pool* p, const nn_params* params, const float* x, const float* v, const int32_t* kind,
                      int64_t n, int32_t* slot) {
  if (!p || !params || !x || !kind || n <= 0 || !slot) return NN_E_ARG;
  pool_free_dead(p);
  auto sc = std::make_shared<SceneHost>();
  sc->params = *params;
  int rc = count_kinds(*sc, kind, n);
  if (rc < 0) return rc;
  std::string err;
  rc = nn::derive_scene(*sc, err);
  if (rc < 0) {
    nn::set_error(err);
    return rc;
  }
  nn::rigid_rest_centres(*sc, x);
  const long long ns_est = sc->n_elastic > 0 ? sc->n_elastic + sc->n_elastic / 2 + 64 : 0;
  const long long fp = nn::scene_footprint(n, ns_est, sc->ncells);
  int s_id = -1;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    // scenes retired since the last tick leave at the next membership change:
    // their bytes are already free for the budget
    long long retiring = 0;
    for (const PoolSlot& q : p->slots)
      if (q.used && q.retired) retiring += q.bytes;
    if (p->in_use - retiring + 2 * fp > p->budget) {  // the scene lives twice while it is merged
      nn::set_error("pool budget exceeded");
      return NN_E_BUSY;
    }
    for (int k = 0; k < POOL_CAPACITY; k++)
      if (!p->slots[k].used) {
        s_id = k;
        break;
      }
    if (s_id < 0) {
      nn::set_error("pool has no free slot");
      return NN_E_BUSY;
    }
    p->slots[s_id].used = true;
    p->slots[s_id].retired = false;
    p->slots[s_id].bytes = fp;
    p->slots[s_id].steps = 0;
    p->in_use += fp;
  }
  PoolSlot& ps = p->slots[s_id];
  auto fail = [&](int code) {
    std::lock_guard<std::mutex> lk(p->mu);
    ps.staging.reset();
    ps.scene.reset();
    ps.used = false;
    p->in_use -= fp;
    return code;
  };
  hipStream_t sst = nullptr;
  {
    const int rc2 = staging_stream(&sst);
    if (rc2 < 0) return fail(rc2);
  }
  std::unique_ptr<Batch> stg(new Batch(sst));
  stg->block = p->block;
  std::copy(p->kblock, p->kblock + nn::K_COUNT, stg->kblock);
  std::copy(p->kdepth, p->kdepth + nn::K_COUNT, stg->kdepth);
  stg->scenes = {sc};
  rc = stg->layout({});
  if (rc < 0) return fail(rc);
  rc = stg->upload(0, x, v, kind);
  if (rc < 0) return fail(rc);
  rc = stg->init_scenes({0});
  if (rc < 0) return fail(rc);
  {
    std::lock_guard<std::mutex> lk(p->mu);
    ps.scene = sc;
    ps.staging = std::move(stg);
  }
  *slot = s_id;
  return NN_OK;
}

int32_t nn_pool_retire(nn_pool* p, int32_t slot) {
  if (!p || slot < 0 || slot >= POOL_CAPACITY) return NN_E_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  if (!p->slots[slot].used) return NN_E_ARG;
  p->slots[slot].retired = true;
  return NN_OK;
}

// Membership changes are applied at tick boundaries: retired scenes leave,
// staged scenes join.  Normally nothing moves: the scenes kept stay where they
// are (a retired scene leaves an unused range) and new scenes are appended
// after the last used range, so a change costs the new scenes' state copies
// and the tables.  Only when the room runs out is the batch laid out afresh,
// packed, with slack for later appends.
static int pool_apply_membership(nn_pool* p) {
  std::vector<std::shared_ptr<SceneHost>> members;
  std::vector<int> mslot;
  std::vector<const Batch*> from{p->main.get()};
  bool changed = false;
  for (size_t b = 0; b < p->main->scenes.size(); b++) {  // present members, in batch order
    const int k = p->member_slot[b];
    if (p->slots[k].retired) {
      changed = true;
      continue;
    }
    members.push_back(p->slots[k].scene);
    mslot.push_back(k);
  }
  for (int k = 0; k < POOL_CAPACITY; k++) {  // staged scenes join at the end
    PoolSlot& s = p->slots[k];
    if (!s.used || s.retired || !s.scene || !s.staging) continue;
    changed = true;
    from.push_back(s.staging.get());
    members.push_back(s.scene);
    mslot.push_back(k);
  }
  for (int k = 0; k < POOL_CAPACITY; k++)  // retired before joining
    if (p->slots[k].used && p->slots[k].retired && p->slots[k].staging) changed = true;
  if (!changed) return NN_OK;
  p->main->block = p->block;
  std::copy(p->kblock, p->kblock + nn::K_COUNT, p->main->kblock);
  std::copy(p->kdepth, p->kdepth + nn::K_COUNT, p->main->kdepth);
  int rc = p->main->append_in_place(members, from);
  if (rc == Batch::LAYOUT_NEEDED) {
    long long live = 0;
    for (int k : mslot) live += p->slots[k].bytes;
    // room for appends, within the budget
    const double slack = std::min(1.5, std::max(1.0, 0.9 * (double)p->budget / (double)std::max(live, 1LL)));
    std::vector<std::shared_ptr<SceneHost>> before = p->main->scenes;
    p->main->scenes = members;
    rc = p->main->layout(from, slack, POOL_CAPACITY);
    if (rc < 0) p->main->scenes = before;  // layout() kept the old arrays
  }
  if (rc < 0) return rc;
  for (int k = 0; k < POOL_CAPACITY; k++) {
    PoolSlot& s = p->slots[k];
    if (!s.used) continue;
    if (s.retired) {
      if (s.staging) p->graveyard.push_back(std::move(s.staging));
      s.scene.reset();
      s.used = false;
      s.retired = false;
      p->in_use -= s.bytes;
      s.bytes = 0;
      continue;
    }
    if (s.staging && s.scene) p->graveyard.push_back(std::move(s.staging));
  }
  p->main->take_deferred(p->dead_arrays);
  p->member_slot = mslot;
  return NN_OK;
}

static int pool_tick_impl(nn_pool* p, nn_step_stats* stats, int32_t stats_len, int32_t* stepped) {
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::vector<int> members;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    int rc = pool_apply_membership(p);
    if (rc < 0) return rc;
    members = p->member_slot;
  }
  if (stepped) *stepped = (int32_t)members.size();
  if (members.empty()) return NN_OK;
  // the step runs without mu: admissions and retirements meanwhile touch only
  // slots outside the batch or flags read at the next membership change
  std::vector<nn_step_stats> st;
  const int rc = p->main->step(&st);
  std::lock_guard<std::mutex> lk(p->mu);
  for (size_t k = 0; k < st.size(); k++) {
    const int s = members[k];
    p->slots[s].steps++;
    if (stats && s < stats_len) stats[s] = st[k];
  }
  return rc;
}

int32_t nn_pool_tick(nn_pool* p, nn_step_stats* stats, int32_t stats_len, int32_t* stepped) {
  if (!p) return NN_E_ARG;
  if (p->async_on) {
    nn::set_error("nn_pool_tick: a tick is in flight (nn_pool_sync first)");
    return NN_E_BUSY;
  }
  return pool_tick_impl(p, stats, stats_len, stepped);
}

int32_t nn_pool_tick_async(nn_pool* p) {
  if (!p) return NN_E_ARG;
  if (p->async_on) {
    nn::set_error("nn_pool_tick_async: a tick is already in flight (nn_pool_sync first)");
    return NN_E_BUSY;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_pool_tick_async");
  nn_step_stats unset{};
  unset.iterations_v = -1;  // marks the slots the tick does not step
  p->async_stats.assign(POOL_CAPACITY, unset);
  p->async_err.clear();
  p->async_on = true;
  p->async = std::thread([p, dev] {
    int stepped = 0;
    hipError_t se = hipSetDevice(dev);
    const int rc = se != hipSuccess ? nn::hip_fail(se, "nn_pool_tick_async: device")
                                    : pool_tick_impl(p, p->async_stats.data(), POOL_CAPACITY, &stepped);
    if (rc != NN_OK) p->async_err = nn::last_error_text();
    p->async_stepped = stepped;
    p->async_rc = rc;
  });
  return NN_OK;
}

int32_t nn_pool_sync(nn_pool* p, nn_step_stats* stats, int32_t stats_len, int32_t* stepped) {
  if (!p) return NN_E_ARG;
  if (!p->async_on) {
    nn::set_error("nn_pool_sync: no tick in flight");
    return NN_E_STATE;
  }
  p->async.join();
  p->async_on = false;
  if (stats)
    for (int s = 0; s < std::min<int>(stats_len, POOL_CAPACITY); s++)
      if (p->async_stats[s].iterations_v >= 0) stats[s] = p->async_stats[s];
  if (stepped) *stepped = p->async_stepped;
  if (p->async_rc != NN_OK) nn::set_error(p->async_err);
  return p->async_rc;
}

int64_t nn_pool_steps(nn_pool* p, int32_t slot) {
  if (!p || slot < 0 || slot >= POOL_CAPACITY) return -1;
  std::lock_guard<std::mutex> lk(p->mu);
  return p->slots[slot].used ? p->slots[slot].steps : -1;
}

static Batch* pool_locate(nn_pool* p, int slot, int& index) {
  PoolSlot& s = p->slots[slot];
  if (!s.used || !s.scene) return nullptr;
  Batch* b = s.staging ? s.staging.get() : p->main.get();
  for (size_t k = 0; k < b->scenes.size(); k++)
    if (b->scenes[k].get() == s.scene.get()) {
      index = (int)k;
      return b;
    }
  return nullptr;
}

int32_t nn_pool_read(nn_pool* p, int32_t slot, int32_t field, void* host, size_t bytes) {
  if (!p || slot < 0 || slot >= POOL_CAPACITY || !host) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  int idx = 0;
  Batch* b = pool_locate(p, slot, idx);
  if (!b) {
    nn::set_error("slot not resident");
    return NN_E_ARG;
  }
  return b->download(idx, field, host, bytes);
}

int32_t nn_pool_info(nn_pool* p, int32_t* resident, int32_t* pending, int64_t* bytes_in_use) {
  if (!p) return NN_E_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  int r = 0, q = 0;
  for (auto& s : p->slots) {
    if (!s.used || !s.scene) continue;
    if (s.staging) q++;
    else r++;
  }
  if (resident) *resident = r;
  if (pending) *pending = q;
  if (bytes_in_use) *bytes_in_use = p->in_use;
  return NN_OK;
}

int32_t nn_pool_profile(nn_pool* p, int32_t enable) { return nn_pool_profile_mask(p, enable ? 0xFFFFFFFFu : 0u, 1); }

int32_t nn_pool_profile_mask(nn_pool* p, uint32_t mask, int32_t stride) {
  if (!p || stride < 1) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  p->main->timer.on = mask != 0;
  p->main->timer.mask = mask;
  p->main->timer.stride = stride;
  p->main->timer.reset();
  return NN_OK;
}

int32_t nn_pool_kernel_times(nn_pool* p, double* ms, int64_t* launches, int32_t n) {
  if (!p) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  return copy_times(p->main->timer, ms, launches, n);
}

int32_t nn_pool_kernel_launches(nn_pool* p, int64_t* seen, int32_t n) {
  if (!p) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  return copy_seen(p->main->timer, seen, n);
}

int32_t nn_pool_set_block(nn_pool* p, int32_t block) {
  if (!p) return NN_E_ARG;
  if (block == 0) block = 256;
  if (block < 64 || block > 1024 || (block & (block - 1))) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  p->block = block;
  p->main->block = block;
  return NN_OK;
}

int32_t nn_pool_set_kernel_depth(nn_pool* p, int32_t kernel, int32_t depth) {
  if (!p || kernel < 0 || kernel >= nn::K_COUNT || !depth_ok(depth)) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  p->kdepth[kernel] = depth;
  p->main->kdepth[kernel] = depth;
  return NN_OK;
}

int32_t nn_pool_set_kernel_block(nn_pool* p, int32_t kernel, int32_t block) {
  if (!p || kernel < 0 || kernel >= nn::K_COUNT) return NN_E_ARG;
  if (block != 0 && (block < 64 || block > 1024 || (block & (block - 1)))) return NN_E_ARG;
  std::lock_guard<std::mutex> tk(p->tick_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  p->kblock[kernel] = block;
  p->main->kblock[kernel] = block;
  return NN_OK;
}

}  // extern "C"
