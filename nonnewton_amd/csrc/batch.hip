// batch.hip -- host driver of the SPH step: derived constants, device layout of
// one or many scenes, upload, initialization, the per-step launch sequence with
// device-side solver termination, and downloads in original particle order.
#include "batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nn {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error_text() { return g_err.c_str(); }
int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return NN_E_HIP;
}

#define HIPCHK(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

#define CHK(expr)              \
  do {                         \
    int _rc = (expr);          \
    if (_rc < 0) return _rc;   \
  } while (0)

static inline int round_up64(int v) { return (v + 63) & ~63; }

// The neighbor lists of slots [pb, pb + n) into host memory (n * MAX_NB ints),
// in the device's grouped layout restricted to those slots: position m of
// slot pb + t is at host_nb(n, t, m).
static hipError_t copy_lists(int* host, const Dev& d, int pb, int n, hipStream_t st) {
  return hipMemcpy2DAsync(host, sizeof(int4) * n, d.nbr + 4L * pb, sizeof(int4) * d.P, sizeof(int4) * n, MAX_NB / 4,
                          hipMemcpyDeviceToHost, st);
}
static inline size_t host_nb(int n, int t, int m) { return 4 * ((size_t)(m >> 2) * n + t) + (m & 3); }
// After the copy has completed: list entries to scene-local slots, each
// checked to lie in [0, n).
static int lists_to_local(std::vector<int>& nb, const std::vector<int>& cnt, int pb, int n) {
  for (int t = 0; t < n; t++) {
    if (cnt[t] < 0 || cnt[t] > MAX_NB) {
      set_error("neighbor count out of range");
      return NN_E_STATE;
    }
    for (int m = 0; m < cnt[t]; m++) {
      int& j = nb[host_nb(n, t, m)];
      j -= pb;
      if (j < 0 || j >= n) {
        set_error("neighbor list entry outside its scene");
        return NN_E_STATE;
      }
    }
  }
  return NN_OK;
}

// cubic_kernel (SPH/main.py:859-880), host copy of the device function
static float host_kern_w(float h, float kW, float rn) {
  float res = 0.0f;
  float q = rn / h;
  if (q <= 1.0f) {
    if (q <= 0.5f) {
      float q2 = q * q;
      float q3 = q2 * q;
      res = kW * (6.0f * q3 - 6.0f * q2 + 1.0f);
    } else {
      float t = 1.0f - q;
      res = (kW * 2.0f) * (t * t * t);
    }
  }
  return res;
}

float neighbor_threshold(float h) {
  float t = h * h;
  while (t > 0.0f && sqrtf(t) >= h) t = nextafterf(t, 0.0f);
  while (sqrtf(t) < h) t = nextafterf(t, INFINITY);
  return t;
}

int derive_scene(SceneHost& s, std::string& err) {
  const nn_params& p = s.params;
  if (!(p.particle_radius > 0.0) || !(p.dt > 0.0) || !(p.density0 > 0.0)) {
    err = "particle_radius, dt and density0 must be positive";
    return NN_E_ARG;
  }
  if (p.warm_start & ~1) {
    err = "warm_start: only bit 0 (Weiler CG initial guess) is supported";
    return NN_E_UNSUPPORTED;
  }
  if (p.non_newton_method < 0 || p.non_newton_method > 6) {
    err = "non_newton_method must be 0..6";
    return NN_E_ARG;
  }
  const double r = p.particle_radius, dd = 2.0 * r, h = 4.0 * r;
  long cells = 1;
  for (int k = 0; k < 3; k++) {
    const double g = std::ceil((p.domain_end[k] - p.domain_start[k]) / h);  // :773-775
    if (!(g >= 1.0) || g > 1.0e6) {
      err = "domain must span at least one cell per axis";
      return NN_E_ARG;
    }
    s.gn[k] = (int)g;
    cells *= s.gn[k];
  }
  if (cells > (1L << 30)) {
    err = "grid too large";
    return NN_E_ARG;
  }
  s.ncells = cells;
  SceneDev& c = s.dev;
  c.gn0 = s.gn[0];
  c.gn1 = s.gn[1];
  c.gn2 = s.gn[2];
  c.h = (float)h;
  c.thr2 = neighbor_threshold(c.h);
  c.kW = (float)(8.0 / M_PI / (h * h * h));
  c.kG = (float)(6.0 * 8.0 / M_PI / (h * h * h));
  c.W0 = host_kern_w(c.h, c.kW, 0.0f);
  c.dt = (float)p.dt;
  c.inv_dt = (float)(1.0 / (double)c.dt);
  c.inv_dt2 = (float)(1.0 / ((double)c.dt * (double)c.dt));
  c.rho0 = (float)p.density0;
  const double vol0 = 0.8 * dd * dd * dd;  // :607
  c.mass = (float)(p.particle_mass > 0.0 ? p.particle_mass : vol0 * 1000.0);  // :743
  c.nu = (float)p.viscosity;
  c.bnu = (float)p.boundary_viscosity;
  c.gx = (float)p.gravity[0];
  c.gy = (float)p.gravity[1];
  c.gz = (float)p.gravity[2];
  c.lo0 = (float)(p.domain_start[0] + h);
  c.lo1 = (float)(p.domain_start[1] + h);
  c.lo2 = (float)(p.domain_start[2] + h);
  c.hi0 = (float)(p.domain_end[0] - h);
  c.hi1 = (float)(p.domain_end[1] - h);
  c.hi2 = (float)(p.domain_end[2] - h);
  c.diam2 = (float)(dd * dd);
  const float dr = (float)dd;
  c.Wd = host_kern_w(c.h, c.kW, sqrtf(dr * dr + 0.0f * 0.0f + 0.0f * 0.0f));  // :1414
  c.visc_eps = (float)(0.01 * h * h);
  c.sigma = (float)p.surface_tension;
  c.opc = (float)(1.0 + p.collision_coeff);
  c.sticky = (float)p.sticky_coeff;
  c.eps = (float)p.eps;
  const double E = p.youngs_modulus, nu = p.poisson_ratio;
  const double factor = E / ((1.0 + nu) * (1.0 - 2.0 * nu));  // :1277-1281
  c.C_diag = (float)(factor * (1.0 - nu));
  c.C_off = (float)(factor * nu);
  c.C_shear = (float)(factor * 0.5 * (1.0 - 2.0 * nu));
  c.nn_method = p.non_newton_method;
  c.nn_n = (float)p.power_index;
  c.nn_K = (float)p.consistency_index;
  c.nn_e0 = (float)p.viscosity0;
  c.nn_einf = (float)p.viscosity_inf;
  c.nn_gc = (float)p.critical_strain_rate;
  c.nn_muc = (float)p.mu_c;
  c.nn_tauy = (float)p.yield_stress;
  c.nn_max = (float)p.max_viscosity;
  c.ten_nu = 10.0f * c.nu;
  c.ten_bnu = 10.0f * c.bnu;
  c.sig_m = c.sigma / c.mass * c.mass;
  c.eta_v = 1.0 / (double)c.dt * p.max_error_v * 0.01 * p.density0;  // :1699
  c.eta_p = p.max_error * 0.01 * p.density0;                          // :1781
  c.cg_tol2 = p.weiler_tol * p.weiler_tol;
  c.rho0_d = p.density0;
  c.max_iter_v = p.max_iterations_v;
  c.max_iter_p = p.max_iterations;
  c.max_iter_cg = p.weiler_max_iter;
  c.vol0 = (float)vol0;
  c.nbody = s.nbody;
  c.coupling = std::max(0, p.coupling_interval);
  int f = 0;
  if (p.use_surface_tension) f |= SF_SURFACE_TENSION;
  if (p.use_viscosity) f |= SF_VISCOSITY;
  if (p.use_non_newtonian) f |= SF_NON_NEWTONIAN;
  if (p.viscosity_method == 1) f |= SF_WEILER;
  else if (p.viscosity_method != 0) {
    err = "viscosity_method must be 0 (explicit) or 1 (Weiler)";
    return NN_E_ARG;
  }
  if (p.enable_divergence_solver) f |= SF_DIVERGENCE;
  if (p.warm_start & 1) f |= SF_WARM_CG;
  if (p.compensated) f |= SF_COMPENSATED;
  if (s.n_fluid > 0) f |= SF_HAS_FLUID;
  if (s.n_elastic > 0) f |= SF_HAS_ELASTIC;
  if (s.nbody > 0) f |= SF_HAS_RIGID;
  c.flags = f;
  c.n = s.n;
  c.n_fluid = s.n_fluid;
  c.n_elastic = s.n_elastic;
  c.nf_d = (double)s.n_fluid;
  return NN_OK;
}

// Rest centres of mass of the rigid bodies (RigidBodySolver.__init__,
// :1100-1106) from the initial positions: FP64 sums of the FP32 terms m x in
// index order, rounded once.  Needs derive_scene (the mass) and the bodies of
// the kinds; runs before the scene is laid out.
void rigid_rest_centres(SceneHost& s, const float* x) {
  const float mass_r = s.dev.vol0 * s.dev.rho0;
  s.rest_cm.assign(3 * (size_t)s.nbody, 0.0f);
  for (int b = 0; b < s.nbody; b++) {
    double sm = 0.0, sx[3] = {0.0, 0.0, 0.0};
    for (int m = s.body_off[b]; m < s.body_off[b + 1]; m++) {
      const int id = s.body_ids[m];
      sm += (double)mass_r;
      for (int c = 0; c < 3; c++) sx[c] += (double)(mass_r * x[3 * id + c]);
    }
    for (int c = 0; c < 3; c++) s.rest_cm[3 * b + c] = (float)(sx[c] / sm);
  }
}

// bytes per particle slot: pos,pos_alt,vel,vel_alt,acc,wd0,x0,vdiff,vdiff_alt (9 float4),
// wx,wr,ww,wp,ws (5 float3), xu (2 float3), wd1 (float2), info,info_alt,key,rank,inv,bucket,nnbr,s_of (8 int),
// rho,mu,gam,eta (4 float), sa (float2), nbr and spair (60 int + 60 float), partials (16 doubles per 64
// slots), wave table
static constexpr long long BYTES_PER_SLOT = 9 * 16 + 7 * 12 + 8 + 8 * 4 + 6 * 4 + 2 * MAX_NB * 4 + 3;
static constexpr long long BYTES_PER_MEMBER = 4 + 4 + MAX_REST * 4 + 4 * 16 + 1;

long long scene_footprint(long long n, long long ns, long long cells) {
  const long long np = (n + 63) / 64 * 64, nsp = (ns + 63) / 64 * 64;
  return np * BYTES_PER_SLOT + nsp * BYTES_PER_MEMBER + (cells + 1) * 8 + 2048;
}

Batch::Batch(hipStream_t st) : stream(st) {}

Batch::~Batch() {
  release();
  if (h_state_) (void)hipHostFree(h_state_);
  if (h_active_) (void)hipHostFree(h_active_);
}

// The arrays of a layout are carved from one device allocation (256-byte
// aligned sub-ranges): one hipMalloc and, when the batch is released or laid
// out again, one hipFree, which synchronizes the device -- a pool admission
// creates and later frees a staging batch, and with one allocation per array
// it paid that synchronization some forty times.
int Batch::alloc_arena(std::vector<std::pair<void**, size_t>>& req) {
  size_t total = 0;
  for (auto& r : req) total += (r.second + 255) & ~(size_t)255;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, total);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return NN_E_OOM;
  }
  char* c = (char*)q;
  for (auto& r : req) {
    *r.first = c;
    c += (r.second + 255) & ~(size_t)255;
  }
  allocs_.push_back({q, total});
  bytes_ += (long long)total;
  return NN_OK;
}

void Batch::release() {
  for (auto& a : allocs_) (void)hipFree(a.p);
  allocs_.clear();
  bytes_ = 0;
}

DevRef Batch::dref() { return DevRef{d}; }

void Batch::swap_sort_buffers() {
  std::swap(d.pos, d.pos_alt);
  std::swap(d.vel, d.vel_alt);
  std::swap(d.info, d.info_alt);
  std::swap(d.vdiff, d.vdiff_alt);
}

// Per-scene fields that depend on the scenes' order and offsets (p_begin,
// cell_begin and s_begin must be set): the slot shift used by the sort, the
// wave range, and the batch-wide feature flags.
void Batch::derive_members() {
  long base = 0;
  any_fluid_ = any_elastic_ = any_weiler_ = any_strain_ = any_div_ = any_warm_cg_ = any_rigid_ = false;
  max_body_ = max_coupling_ = 0;
  max_iter_[0] = max_iter_[1] = max_iter_[2] = 1;
  for (auto& sp : scenes) {
    SceneHost& s = *sp;
    s.dev.p_begin = s.p_begin;
    s.dev.shift = (int)(s.p_begin - base);  // particles of the scenes before this one, in cell order
    s.dev.wave_begin = s.p_begin / 64;
    s.dev.wave_end = (s.p_begin + round_up64(s.n)) / 64;
    s.dev.cell_begin = s.cell_begin;
    s.dev.s_begin = s.s_begin;
    s.dev.ns = s.ns;
    s.dev.body_begin = s.body_begin;
    base += s.n;
    const int f = s.dev.flags;
    if (f & SF_HAS_FLUID) {
      any_fluid_ = true;
      if (f & SF_DIVERGENCE) any_div_ = true;
      if ((f & SF_VISCOSITY) && (f & SF_WEILER)) any_weiler_ = true;
      if ((f & SF_VISCOSITY) && (f & SF_WEILER) && (f & SF_WARM_CG)) any_warm_cg_ = true;
      if ((f & SF_NON_NEWTONIAN) || ((f & SF_VISCOSITY) && (f & SF_WEILER))) any_strain_ = true;
    }
    if (f & SF_HAS_ELASTIC) any_elastic_ = true;
    if (f & SF_HAS_RIGID) {
      any_rigid_ = true;
      max_body_ = std::max(max_body_, s.dev.nbody);
      max_coupling_ = std::max(max_coupling_, s.dev.coupling);
    }
    max_iter_[0] = std::max(max_iter_[0], s.dev.max_iter_v);
    max_iter_[1] = std::max(max_iter_[1], s.dev.max_iter_p);
    max_iter_[2] = std::max(max_iter_[2], s.dev.max_iter_cg);
  }
}

// Scene table and the wave -> scene maps of `nd` for the current scenes
// (waves of no scene map to -1).
int Batch::upload_tables(const Dev& nd) {
  std::vector<int> hw(std::max(nd.W, 1), -1), hsw(std::max(nd.SW, 1), -1);
  std::vector<SceneDev> hs(scenes.size());
  std::vector<int2> cgb;
  for (size_t k = 0; k < scenes.size(); k++) {
    SceneHost& s = *scenes[k];
    for (int w = s.dev.wave_begin; w < s.dev.wave_end; w++) hw[w] = (int)k;
    for (int w = s.s_begin / 64; w < (s.s_begin + round_up64(s.ns)) / 64; w++) hsw[w] = (int)k;
    hs[k] = s.dev;
    const int f = s.dev.flags;
    if ((f & SF_HAS_FLUID) && (f & SF_VISCOSITY) && (f & SF_WEILER))  // CG update workgroups
      for (int w = s.dev.wave_begin; w < s.dev.wave_end; w += CG_UPD_WAVES) cgb.push_back(make_int2((int)k, w));
  }
  n_cg_blocks_ = (int)cgb.size();
  // rigid set and body table, rebuilt from the scenes' host data (ranges of
  // retired scenes map to no body; their table entries get rank -1)
  if (nd.NRn > 0 || nd.nbodies > 0) {
    std::vector<int> rp(std::max(nd.NRn, 1), -1), rwb(std::max(nd.NRn / 64, 1), -1);
    std::vector<BodyDev> bt(std::max(nd.nbodies, 1));
    for (auto& b : bt) {
      b = BodyDev{};
      b.rank = -1;
    }
    for (size_t k = 0; k < scenes.size(); k++) {
      const SceneHost& s = *scenes[k];
      if (s.rest_cm.size() < 3 * (size_t)s.nbody || s.body_off.size() < (size_t)s.nbody + 1) {
        set_error("rigid bodies without rest centres");
        return NN_E_STATE;
      }
      int off = s.r_begin;
      for (int b = 0; b < s.nbody; b++) {
        const int lo = s.body_off[b], hi = s.body_off[b + 1];
        for (int m = lo; m < hi; m++) rp[off + m - lo] = s.body_ids[m];
        const int len = round_up64(hi - lo);
        for (int w = off / 64; w < (off + len) / 64; w++) rwb[w] = s.body_begin + b;
        BodyDev& e = bt[s.body_begin + b];
        e.scene = (int)k;
        e.rank = b;
        e.wave_begin = off / 64;
        e.wave_end = (off + len) / 64;
        for (int c = 0; c < 3; c++) e.rest_cm[c] = e.cm[c] = s.rest_cm[3 * b + c];
        for (int q = 0; q < 9; q++) e.R[q] = (q % 4) == 0 ? 1.0f : 0.0f;
        off += len;
      }
    }
    if (nd.NRn > 0) {
      HIPCHK(hipMemcpyAsync((void*)nd.r_pid, rp.data(), sizeof(int) * nd.NRn, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync((void*)nd.r_wave_body, rwb.data(), sizeof(int) * (nd.NRn / 64), hipMemcpyHostToDevice, stream));
    }
    if (nd.nbodies > 0)
      HIPCHK(hipMemcpyAsync((void*)nd.bodies, bt.data(), sizeof(BodyDev) * nd.nbodies, hipMemcpyHostToDevice, stream));
    // the host vectors must outlive the copies
    HIPCHK(hipStreamSynchronize(stream));
  }
  if (!cgb.empty())
    HIPCHK(hipMemcpyAsync((void*)nd.cg_blocks, cgb.data(), sizeof(int2) * cgb.size(), hipMemcpyHostToDevice, stream));
  if (!scenes.empty())
    HIPCHK(hipMemcpyAsync((void*)nd.scenes, hs.data(), sizeof(SceneDev) * hs.size(), hipMemcpyHostToDevice, stream));
  if (nd.W) HIPCHK(hipMemcpyAsync((void*)nd.wave_scene, hw.data(), sizeof(int) * nd.W, hipMemcpyHostToDevice, stream));
  if (nd.SW) HIPCHK(hipMemcpyAsync((void*)nd.s_wave_scene, hsw.data(), sizeof(int) * nd.SW, hipMemcpyHostToDevice, stream));
  // the host vectors must outlive the copies
  HIPCHK(hipStreamSynchronize(stream));
  return NN_OK;
}

// Persistent state of scene s (at its offsets in nd) from batch `from`, where
// it sits at slot ob and stress-set slot osb with a stress set of ons members.
int Batch::copy_scene_state(const Dev& nd, const SceneHost& s, const Batch* from, int ob, int osb, int ons) {
  const Dev& od = from->d;
  const int nb = s.p_begin, n = s.n;
  auto cp = [&](void* dst, const void* sp, size_t bytes) -> hipError_t {
    return bytes ? hipMemcpyAsync(dst, sp, bytes, hipMemcpyDeviceToDevice, stream) : hipSuccess;
  };
  HIPCHK(cp(nd.pos + nb, od.pos + ob, sizeof(float4) * n));
  HIPCHK(cp(nd.vel + nb, od.vel + ob, sizeof(float4) * n));
  HIPCHK(cp(nd.info + nb, od.info + ob, sizeof(int) * n));
  HIPCHK(cp(nd.inv + nb, od.inv + ob, sizeof(int) * n));
  HIPCHK(cp(nd.x0 + nb, od.x0 + ob, sizeof(float4) * n));
  HIPCHK(cp(nd.s_of + nb, od.s_of + ob, sizeof(int) * n));
  HIPCHK(cp(nd.rho + nb, od.rho + ob, sizeof(float) * n));
  HIPCHK(cp(nd.sa + nb, od.sa + ob, sizeof(float2) * n));
  HIPCHK(cp(nd.acc + nb, od.acc + ob, sizeof(float4) * n));
  HIPCHK(cp(nd.gam + nb, od.gam + ob, sizeof(float) * n));
  HIPCHK(cp(nd.eta + nb, od.eta + ob, sizeof(float) * n));
  HIPCHK(cp(nd.vdiff + nb, od.vdiff + ob, sizeof(float4) * n));
  if (s.ns > 0 && ons == s.ns) {  // a resized stress set is uploaded afresh
    const int ns_ = s.s_begin;
    HIPCHK(cp(nd.s_pid + ns_, od.s_pid + osb, sizeof(int) * s.ns));
    HIPCHK(cp(nd.rest_cnt + ns_, od.rest_cnt + osb, sizeof(int) * s.ns));
    HIPCHK(cp(nd.rs + 4L * ns_, od.rs + 4L * osb, sizeof(float4) * 4 * s.ns));
    HIPCHK(hipMemcpy2DAsync(nd.rest + ns_, sizeof(int) * nd.NS, od.rest + osb, sizeof(int) * od.NS,
                            sizeof(int) * s.ns, MAX_REST, hipMemcpyDeviceToDevice, stream));
  }
  return NN_OK;
}

int Batch::layout(const std::vector<const Batch*>& from, double slack, int scene_capacity) {
  struct Src {
    const Batch* b;
    int pb, sb, ns;  // offsets and stress-set size in the source batch
  };
  std::vector<Src> src(scenes.size(), Src{nullptr, 0, 0, 0});
  for (size_t k = 0; k < scenes.size(); k++) {
    for (const Batch* b : from) {
      if (!b) continue;
      for (auto& t : b->scenes)
        if (t.get() == scenes[k].get()) src[k] = Src{b, scenes[k]->p_begin, scenes[k]->s_begin, scenes[k]->dev.ns};
    }
  }
  // offsets, packed in member order (committed once the arrays exist)
  int P = 0, C = 0, NS = 0, NR = 0, NB = 0;
  std::vector<int> pb(scenes.size()), cb(scenes.size()), sb(scenes.size()), rb(scenes.size()), bb(scenes.size());
  for (size_t k = 0; k < scenes.size(); k++) {
    const SceneHost& s = *scenes[k];
    pb[k] = P;
    P += round_up64(s.n);
    cb[k] = C;
    C += (int)s.ncells + 1;
    sb[k] = NS;
    NS += round_up64(s.ns);
    rb[k] = NR;
    NR += s.nr;
    bb[k] = NB;
    NB += s.nbody;
  }
  // capacities: the extents, grown by `slack` so that scenes can be appended
  // later without moving the others
  slack = slack < 1.0 ? 1.0 : slack;
  const int Pc = round_up64((int)std::ceil(P * slack));
  const int Cc = (int)std::ceil(C * slack);
  const int NSc = round_up64((int)std::ceil(NS * slack));
  const int NRc = round_up64((int)std::ceil(NR * slack));
  const int NBc = (int)std::ceil(NB * slack);
  const int Sc = std::max<int>((int)scenes.size(), scene_capacity);
  // new arrays
  Dev nd{};
  std::vector<Alloc> old_allocs;
  old_allocs.swap(allocs_);
  long long old_bytes = bytes_;
  bytes_ = 0;
  nd.P = Pc;
  nd.W = Pc / 64;
  nd.Pn = P;
  nd.C = C;
  nd.NS = NSc;
  nd.SW = NSc / 64;
  nd.NSn = NS;
  nd.NR = NRc;
  nd.NRn = NR;
  nd.nbodies = NB;
  nd.nscenes = (int)scenes.size();
  const int ncell_entries = Cc + 1;
  const int nscan = (ncell_entries + 2047) / 2048;
  std::vector<std::pair<void**, size_t>> req;
  auto A = [&](auto*& p, size_t count) {
    using T = std::remove_reference_t<decltype(*p)>;
    req.push_back({(void**)&p, (count ? count : 1) * sizeof(T)});
  };
  SceneDev* scenes_dev = nullptr;
  A(scenes_dev, Sc);
  A(nd.state, Sc);
  int* ws = nullptr;
  A(ws, nd.W);
  int* sws = nullptr;
  A(sws, nd.SW);
  A(nd.pos, Pc);
  A(nd.pos_alt, Pc);
  A(nd.vel, Pc);
  A(nd.vel_alt, Pc);
  A(nd.info, Pc);
  A(nd.info_alt, Pc);
  A(nd.acc, Pc);
  A(nd.rho, Pc);
  A(nd.sa, Pc);
  A(nd.nbr, (size_t)Pc * MAX_NB);
  A(nd.nnbr, Pc);
  A(nd.key, Pc);
  A(nd.rank, Pc);
  A(nd.inv, Pc);
  A(nd.bucket, Pc);
  A(nd.cell_count, ncell_entries);
  A(nd.cell_start, ncell_entries);
  A(nd.scan_tmp, nscan);
  A(nd.partials, 32 * (size_t)nd.W);
  int2* cgb = nullptr;
  A(cgb, (size_t)nd.W / CG_UPD_WAVES + Sc + 1);
  A(nd.cg_list, 2 * (size_t)nd.W);
  A(nd.cg_list_n, 2);
  A(nd.n_active, 4);
  A(nd.wx, Pc);
  A(nd.wr, Pc);
  A(nd.ww, Pc);
  A(nd.wp, Pc);
  A(nd.ws, Pc);
  A(nd.xu, Pc);
  A(nd.vdiff_alt, Pc);
  A(nd.wd0, Pc);
  A(nd.wd1, Pc);
  A(nd.spair, (size_t)Pc * MAX_NB);
  A(nd.vdiff, Pc);
  A(nd.mu, Pc);
  A(nd.gam, Pc);
  A(nd.eta, Pc);
  A(nd.x0, Pc);
  A(nd.s_of, Pc);
  A(nd.s_pid, NSc);
  A(nd.rest, (size_t)NSc * MAX_REST);
  A(nd.rest_cnt, NSc);
  A(nd.rs, 4 * (size_t)NSc);
  int* rpid = nullptr;
  A(rpid, NRc);
  int* rwb = nullptr;
  A(rwb, NRc / 64);
  A(nd.rpartials, 9 * (size_t)(NRc / 64));
  A(nd.bodies, NBc);
  const int rc = alloc_arena(req);
  if (rc != NN_OK) {
    release();
    allocs_.swap(old_allocs);
    bytes_ = old_bytes;
    return rc;
  }
  nd.scenes = scenes_dev;
  nd.wave_scene = ws;
  nd.s_wave_scene = sws;
  nd.cg_blocks = cgb;
  nd.r_pid = rpid;
  nd.r_wave_body = rwb;
  for (size_t k = 0; k < scenes.size(); k++) {
    scenes[k]->p_begin = pb[k];
    scenes[k]->cell_begin = cb[k];
    scenes[k]->s_begin = sb[k];
    scenes[k]->r_begin = rb[k];
    scenes[k]->body_begin = bb[k];
  }
  derive_members();
  C_total_ = C;
  C_cap_ = Cc;
  NB_cap_ = NBc;
  scene_cap_ = Sc;
  scan_blocks_ = nscan;
  // pinned host copies of the scene states, for as many scenes as the layout
  // holds (a staging batch holds one; hipHostFree synchronizes the device)
  if (!h_state_ || h_state_cap_ < Sc) {
    if (h_state_) HIPCHK(hipHostFree(h_state_));
    h_state_ = nullptr;
    HIPCHK(hipHostMalloc((void**)&h_state_, sizeof(SceneState) * (size_t)std::max(Sc, 1), hipHostMallocDefault));
    h_state_cap_ = std::max(Sc, 1);
  }
  if (!h_active_) HIPCHK(hipHostMalloc((void**)&h_active_, sizeof(int) * 4, hipHostMallocDefault));
  CHK(upload_tables(nd));
  HIPCHK(hipMemsetAsync(nd.state, 0, sizeof(SceneState) * std::max<size_t>(Sc, 1), stream));
  launch_fill_i32(nd.info, Pc, -1, stream);
  launch_fill_i32(nd.info_alt, Pc, -1, stream);
  launch_fill_i32(nd.s_of, Pc, -1, stream);
  launch_fill_i32(nd.s_pid, NSc, -1, stream);
  launch_fill_i32(nd.rest_cnt, NSc, 0, stream);
  launch_fill_i32(nd.nnbr, Pc, 0, stream);
  launch_fill_f4(nd.pos, Pc, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  launch_fill_f4(nd.pos_alt, Pc, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  launch_fill_f4(nd.vel, Pc, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  launch_fill_f4(nd.vel_alt, Pc, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  launch_fill_f4(nd.x0, Pc, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  launch_fill_f4(nd.vdiff, Pc, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  HIPCHK(hipMemsetAsync(nd.rho, 0, sizeof(float) * std::max(Pc, 1), stream));
  HIPCHK(hipMemsetAsync(nd.sa, 0, sizeof(float2) * std::max(Pc, 1), stream));
  HIPCHK(hipMemsetAsync(nd.gam, 0, sizeof(float) * std::max(Pc, 1), stream));
  HIPCHK(hipMemsetAsync(nd.eta, 0, sizeof(float) * std::max(Pc, 1), stream));
  HIPCHK(hipMemsetAsync(nd.acc, 0, sizeof(float4) * std::max(Pc, 1), stream));
  // copy persistent state of relocated scenes
  for (size_t k = 0; k < scenes.size(); k++)
    if (src[k].b) CHK(copy_scene_state(nd, *scenes[k], src[k].b, src[k].pb, src[k].sb, src[k].ns));
  HIPCHK(hipStreamSynchronize(stream));
  // the old arrays are freed by the caller's choice: now, or later (off the
  // thread that steps the batch; hipFree waits for the device)
  if (defer_free_) {
    for (auto& al : old_allocs) deferred_.push_back(al.p);
  } else {
    for (auto& al : old_allocs) (void)hipFree(al.p);
  }
  d = nd;
  lists_state_ = 0;
  return NN_OK;
}

int Batch::append_in_place(const std::vector<std::shared_ptr<SceneHost>>& members, const std::vector<const Batch*>& from) {
  // the scenes kept must come first, in their present order
  std::vector<SceneHost*> cur;
  for (auto& sp : scenes) cur.push_back(sp.get());
  size_t nk = 0, pos = 0;
  for (; nk < members.size(); nk++) {
    auto it = std::find(cur.begin() + pos, cur.end(), members[nk].get());
    if (it == cur.end()) break;
    pos = (size_t)(it - cur.begin()) + 1;
  }
  for (size_t k = nk; k < members.size(); k++)
    if (std::find(cur.begin(), cur.end(), members[k].get()) != cur.end()) return LAYOUT_NEEDED;
  if ((int)members.size() > scene_cap_) return LAYOUT_NEEDED;
  // offsets of the new scenes, after the last slot, cell and stress slot used
  int P = d.Pn, C = C_total_, NS = d.NSn, NR = d.NRn, NB = d.nbodies;
  std::vector<std::pair<int, int>> place(members.size(), {0, 0});  // (p_begin, s_begin) of new scenes
  std::vector<int> cells(members.size(), 0), rb(members.size(), 0), bb(members.size(), 0);
  for (size_t k = nk; k < members.size(); k++) {
    const SceneHost& s = *members[k];
    place[k] = {P, NS};
    cells[k] = C;
    rb[k] = NR;
    bb[k] = NB;
    P += round_up64(s.n);
    C += (int)s.ncells + 1;
    NS += round_up64(s.ns);
    NR += s.nr;
    NB += s.nbody;
  }
  if (P > d.P || C > C_cap_ || NS > d.NS || NR > d.NR || NB > NB_cap_) return LAYOUT_NEEDED;
  // sources of the new scenes
  struct Src {
    const Batch* b;
    int pb, sb, ns;
  };
  std::vector<Src> src(members.size(), Src{nullptr, 0, 0, 0});
  for (size_t k = nk; k < members.size(); k++)
    for (const Batch* b : from) {
      if (!b || b == this) continue;
      for (auto& t : b->scenes)
        if (t.get() == members[k].get()) src[k] = Src{b, t->p_begin, t->s_begin, t->dev.ns};
    }
  for (size_t k = nk; k < members.size(); k++) {
    SceneHost& s = *members[k];
    s.p_begin = place[k].first;
    s.s_begin = place[k].second;
    s.cell_begin = cells[k];
    s.r_begin = rb[k];
    s.body_begin = bb[k];
  }
  scenes = members;
  derive_members();
  d.Pn = P;
  d.NSn = NS;
  d.NRn = NR;
  d.nbodies = NB;
  d.nscenes = (int)scenes.size();
  C_total_ = C;
  d.C = C;
  CHK(upload_tables(d));
  for (size_t k = nk; k < members.size(); k++)
    if (src[k].b) CHK(copy_scene_state(d, *members[k], src[k].b, src[k].pb, src[k].sb, src[k].ns));
  HIPCHK(hipStreamSynchronize(stream));
  if (nk < members.size()) lists_state_ = 0;  // the new scenes have no lists here yet
  return NN_OK;
}

void Batch::take_deferred(std::vector<void*>& out) {
  out.insert(out.end(), deferred_.begin(), deferred_.end());
  deferred_.clear();
}

int Batch::refresh_scene_table() {
  std::vector<SceneDev> hs(scenes.size());
  for (size_t k = 0; k < scenes.size(); k++) hs[k] = scenes[k]->dev;
  if (!hs.empty())
    HIPCHK(hipMemcpyAsync((void*)d.scenes, hs.data(), sizeof(SceneDev) * hs.size(), hipMemcpyHostToDevice, stream));
  max_iter_[0] = max_iter_[1] = max_iter_[2] = 1;
  for (auto& sp : scenes) {
    max_iter_[0] = std::max(max_iter_[0], sp->dev.max_iter_v);
    max_iter_[1] = std::max(max_iter_[1], sp->dev.max_iter_p);
    max_iter_[2] = std::max(max_iter_[2], sp->dev.max_iter_cg);
  }
  HIPCHK(hipStreamSynchronize(stream));
  return NN_OK;
}

int Batch::upload(int k, const float* x, const float* v, const int* kind) {
  SceneHost& s = *scenes[k];
  const int n = s.n, pb = s.p_begin;
  const float V0 = s.dev.vol0;  // :607, :741
  std::vector<float4> hp(n), hv(n);
  std::vector<int> hi(n), hinv(n);
  for (int i = 0; i < n; i++) {
    const int kd = kind[i] >= KIND_RIGID ? KIND_RIGID : kind[i];
    hp[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], kd == KIND_FLUID ? V0 : -V0);
    hv[i] = v ? make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
    hi[i] = (kd << 28) | i;
    hinv[i] = i;
  }
  HIPCHK(hipMemcpyAsync(d.pos + pb, hp.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(d.vel + pb, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(d.info + pb, hi.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(d.inv + pb, hinv.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
  launch_fill_f4(d.vdiff + pb, n, make_float4(0.f, 0.f, 0.f, 0.f), stream);
  HIPCHK(hipStreamSynchronize(stream));
  s.initialized = false;
  s.stepped = false;
  return NN_OK;
}

// ---------------------------------------------------------------------------
// kernel timer
// ---------------------------------------------------------------------------
static const char* const KERNEL_NAMES[K_COUNT] = {
    "cell_count", "scan", "bucket", "place", "build_lists", "bvol", "clear_acc", "density_alpha",
    "div_update", "div_source", "strain", "cg_init", "cg_matvec", "cg_update", "nonpressure",
    "press_source", "press_update", "elastic_rot_stress", "elastic_force", "advect", "cg_guess"};

const char* kernel_name(int id) { return (id >= 0 && id < K_COUNT) ? KERNEL_NAMES[id] : ""; }

KernelTimer::~KernelTimer() {
  for (auto e : events) (void)hipEventDestroy(e);
}

LaunchEvents& launch_events() {
  static thread_local LaunchEvents e;
  return e;
}

// A timed launch carries a start and a stop event (NN_LAUNCH): the dispatch
// stamps them itself, so the duration is the kernel's and the timing adds no
// packets to the stream.
int KernelTimer::begin(int id, hipStream_t st) {
  (void)st;
  open = false;
  if (!on) return NN_OK;
  const long long n = seen[id]++;  // every kernel's launches are counted while on
  if (!((mask >> id) & 1u) || n % stride != 0) return NN_OK;
  open = true;
  while (events.size() < used + 2) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, TIMING_EVENT_FLAGS));
    events.push_back(e);
  }
  pending.push_back({id, used});
  LaunchEvents& le = launch_events();
  le.start = events[used];
  le.stop = events[used + 1];
  le.launched = false;
  used += 2;
  return NN_OK;
}

int KernelTimer::end(hipStream_t st) {
  (void)st;
  if (!open) return NN_OK;
  open = false;
  LaunchEvents& le = launch_events();
  if (!le.launched) pending.pop_back();  // the launcher launched nothing (no work): no sample
  le = LaunchEvents{};
  return NN_OK;
}

int KernelTimer::collect() {
  for (auto& p : pending) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, events[p.second], events[p.second + 1]));
    total_ms[p.first] += ms;
    count[p.first]++;
  }
  pending.clear();
  used = 0;
  return NN_OK;
}

void KernelTimer::reset() {
  for (int k = 0; k < K_COUNT; k++) {
    total_ms[k] = 0.0;
    count[k] = 0;
    seen[k] = 0;
  }
}

#define TIMED(id, call)               \
  do {                                \
    CHK(timer.begin((id), stream));   \
    call;                             \
    CHK(timer.end(stream));           \
  } while (0)

// The counting sort and the lists.  all == 0 (inside a step) leaves static
// boundary particles without a list; lists_state_ records which kind of search
// ran last, so that list downloads can complete them first.
int Batch::search_impl(int all) {
  launch_fill_i32(d.cell_count, (long)C_total_ + 1, 0, stream);
  TIMED(K_CELL_COUNT, launch_cell_count(dref(), blk(K_CELL_COUNT), stream));
  TIMED(K_SCAN, launch_scan(dref(), C_total_ + 1, stream));
  TIMED(K_BUCKET, launch_bucket(dref(), blk(K_BUCKET), stream));
  TIMED(K_PLACE, launch_place(dref(), blk(K_PLACE), stream));
  swap_sort_buffers();
  TIMED(K_BUILD_LISTS, launch_build_lists(dref(), blk(K_BUILD_LISTS), all, stream));
  lists_state_ = all ? 2 : 1;
  return NN_OK;
}

int Batch::search() {
  if (scenes.empty()) return NN_OK;
  launch_step_begin(dref(), stream);
  CHK(search_impl());
  HIPCHK(hipGetLastError());
  CHK(gather_state());
  for (size_t k = 0; k < scenes.size(); k++) {
    const int stt = h_state_[k].status;
    if (stt & ST_OUT_OF_GRID) { set_error("particle outside the grid"); return NN_E_OUT_OF_GRID; }
    if (stt & ST_CELL_OVERFLOW) { set_error("more than 50 particles in a cell"); return NN_E_CELL_OVERFLOW; }
    if (stt & ST_NEIGHBOR_OVERFLOW) { set_error("more than 60 neighbors"); return NN_E_NEIGHBOR_OVERFLOW; }
  }
  return NN_OK;
}

int Batch::gather_state() {
  const size_t ns = scenes.size();
  if (ns > 4096) {
    set_error("too many scenes in one batch");
    return NN_E_ARG;
  }
  HIPCHK(hipMemcpyAsync(h_state_, d.state, sizeof(SceneState) * ns, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return NN_OK;
}

// stress set: elastic particles and their rest neighbors (first search), with
// rest lists capped at 50 (Elasticity.initValues_kernel, :1220-1236)
int Batch::build_stress_set(const std::vector<int>& which) {
  bool any = false;
  for (int k : which) {
    SceneHost& s = *scenes[k];
    s.s_pid.clear();
    s.rest.clear();
    s.rest_cnt.clear();
    s.ns = 0;
    if (s.n_elastic == 0) continue;
    any = true;
    const int n = s.n, pb = s.p_begin;
    std::vector<int> info(n), cnt(n), nb((size_t)n * MAX_NB);
    HIPCHK(hipMemcpyAsync(info.data(), d.info + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), d.nnbr + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    HIPCHK(copy_lists(nb.data(), d, pb, n, stream));
    HIPCHK(hipStreamSynchronize(stream));
    CHK(lists_to_local(nb, cnt, pb, n));
    std::vector<int> slot_of_id(n, -1);
    for (int t = 0; t < n; t++) slot_of_id[info[t] & 0x0FFFFFFF] = t;
    std::vector<char> mark(n, 0);
    for (int t = 0; t < n; t++) {
      if ((info[t] >> 28) != KIND_ELASTIC) continue;
      mark[info[t] & 0x0FFFFFFF] = 1;
      for (int m = 0; m < cnt[t]; m++) mark[info[nb[host_nb(n, t, m)]] & 0x0FFFFFFF] = 1;
    }
    std::vector<int> s_of(n, -1);
    for (int id = 0; id < n; id++)
      if (mark[id]) {
        s_of[id] = (int)s.s_pid.size();
        s.s_pid.push_back(id);
      }
    s.ns = (int)s.s_pid.size();
    s.rest.assign((size_t)s.ns * MAX_REST, -1);
    s.rest_cnt.assign(s.ns, 0);
    bool overflow = false;
    for (int q = 0; q < s.ns; q++) {
      const int t = slot_of_id[s.s_pid[q]];
      int c = cnt[t];
      if (c > MAX_REST) {
        overflow = true;
        c = MAX_REST;
      }
      s.rest_cnt[q] = c;
      for (int m = 0; m < c; m++) s.rest[(size_t)q * MAX_REST + m] = info[nb[host_nb(n, t, m)]] & 0x0FFFFFFF;
    }
    if (overflow) {
      set_error("an elastic particle has more than 50 rest neighbors");
      return NN_E_REST_OVERFLOW;
    }
    s_of.swap(s.s_of_host);  // uploaded after the relayout below
  }
  if (!any) return NN_OK;
  // re-layout with the new stress-set sizes, keeping all particle state
  CHK(layout({this}));
  for (int k : which) {
    SceneHost& s = *scenes[k];
    if (s.ns == 0) continue;
    const int sb = s.s_begin, pb = s.p_begin;
    std::vector<int> rest_t((size_t)s.ns * MAX_REST);
    for (int q = 0; q < s.ns; q++)
      for (int m = 0; m < MAX_REST; m++) rest_t[(size_t)m * s.ns + q] = s.rest[(size_t)q * MAX_REST + m];
    HIPCHK(hipMemcpyAsync(d.s_pid + sb, s.s_pid.data(), sizeof(int) * s.ns, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d.rest_cnt + sb, s.rest_cnt.data(), sizeof(int) * s.ns, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpy2DAsync(d.rest + sb, sizeof(int) * d.NS, rest_t.data(), sizeof(int) * s.ns, sizeof(int) * s.ns,
                            MAX_REST, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d.s_of + pb, s.s_of_host.data(), sizeof(int) * s.n, hipMemcpyHostToDevice, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  return NN_OK;
}

int Batch::init_scenes(const std::vector<int>& which) {
  if (scenes.empty()) return NN_OK;
  launch_step_begin(dref(), stream);
  CHK(search_impl());
  launch_elastic_init(dref(), block, stream);
  HIPCHK(hipGetLastError());
  CHK(gather_state());
  for (size_t k = 0; k < scenes.size(); k++) {
    const int stt = h_state_[k].status;
    if (stt & ST_OUT_OF_GRID) { set_error("particle outside the grid"); return NN_E_OUT_OF_GRID; }
    if (stt & ST_CELL_OVERFLOW) { set_error("more than 50 particles in a cell"); return NN_E_CELL_OVERFLOW; }
    if (stt & ST_NEIGHBOR_OVERFLOW) { set_error("more than 60 neighbors"); return NN_E_NEIGHBOR_OVERFLOW; }
  }
  CHK(build_stress_set(which));
  // the stress-set relayout does not carry neighbor lists (they hold absolute
  // slots); search again on the unchanged positions before the volume sweeps
  CHK(search_impl());
  launch_bvol(dref(), block, 0, stream);
  launch_bvol(dref(), block, 1, stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  for (int k : which) scenes[k]->initialized = true;
  return NN_OK;
}

// RigidBodySolver.substep for every scene's bodies (:929-931): body rank r
// of all scenes together, coupling_interval substeps each (a scene with fewer
// bodies or substeps sits the extra launches out).
int Batch::rigid_step() {
  for (int r = 0; r < max_body_; r++)
    for (int c = 0; c < max_coupling_; c++) {
      launch_rigid_ground(dref(), r, c, stream);
      launch_rigid_sum(dref(), 0, r, c, stream);
      launch_body_finish(dref(), 0, r, c, stream);
      launch_rigid_sum(dref(), 1, r, c, stream);
      launch_body_finish(dref(), 1, r, c, stream);
      launch_rigid_place(dref(), r, c, stream);
    }
  HIPCHK(hipGetLastError());
  return NN_OK;
}

// Runs solver iterations until no scene is active: the first chunk is the
// previous step's count, later chunks two at a time; each chunk ends with one
// 4-byte read of the active-scene counter.
int Batch::solver_loop(int which) {
  int pred = 1;
  for (auto& sp : scenes) pred = std::max(pred, sp->last_iter[which]);
  const int maxit = std::max(1, max_iter_[which]);
  int done = 0, chunk = std::min(pred, maxit);
  for (;;) {
    for (int c = 0; c < chunk; c++) {
      if (which == 0) {
        TIMED(K_DIV_UPDATE, launch_div_update(dref(), blk(K_DIV_UPDATE), dep(K_DIV_UPDATE), stream));
        TIMED(K_DIV_SOURCE, launch_div_source(dref(), blk(K_DIV_SOURCE), dep(K_DIV_SOURCE), stream));
        launch_scene_finish(dref(), 0, stream);
      } else if (which == 1) {
        TIMED(K_PRESS_UPDATE, launch_press_update(dref(), blk(K_PRESS_UPDATE), dep(K_PRESS_UPDATE), stream));
        TIMED(K_PRESS_SOURCE, launch_press_source(dref(), blk(K_PRESS_SOURCE), dep(K_PRESS_SOURCE), 1, stream));
        launch_scene_finish(dref(), 1, stream);
      }
    }
    done += chunk;
    if (done >= maxit) break;
    HIPCHK(hipMemcpyAsync(h_active_, d.n_active + which, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (*h_active_ <= 0) break;
    chunk = std::min(4, maxit - done);
  }
  return NN_OK;
}

// The Weiler CG iterations: update k (which first forms the scene sums of
// iteration k - 1 and stops converged scenes) and matvec k, for k = 1 ..
// limit, then update limit + 1, which only tests the last residual.  Enqueued
// in chunks -- first the previous step's iteration count plus two, then four
// -- with one 4-byte read of the active-scene counter per chunk; iterations
// past convergence are near-empty launches.
int Batch::cg_loop() {
  int pred = 1;
  long total_waves = 0;
  for (auto& sp : scenes) {
    pred = std::max(pred, sp->last_iter[2]);
    const int f = sp->dev.flags;
    if ((f & SF_HAS_FLUID) && (f & SF_VISCOSITY) && (f & SF_WEILER)) total_waves += sp->dev.wave_end - sp->dev.wave_begin;
  }
  // waves expected to still iterate at iteration k, from each scene's count
  // in the previous step (plus a margin): while they are many the matvec runs
  // on the full grid, once they are few on a grid sized to them (the list of
  // waves the update writes holds the exact set either way)
  auto expected_waves = [&](int k) {
    long w = 0;
    for (auto& sp : scenes) {
      const int f = sp->dev.flags;
      if ((f & SF_HAS_FLUID) && (f & SF_VISCOSITY) && (f & SF_WEILER) && sp->last_iter[2] + 2 >= k)
        w += sp->dev.wave_end - sp->dev.wave_begin;
    }
    return w;
  };
  const int maxit = std::max(1, max_iter_[2]);
  const int last = maxit + 1;
  int k = 1, chunk = std::min(pred + 3, last);
  for (;;) {
    for (int c = 0; c < chunk && k <= last; c++, k++) {
      TIMED(K_CG_UPDATE, launch_cg_update(dref(), n_cg_blocks_, k, stream));
      if (k > maxit) continue;
      const long ew = expected_waves(k);
      if (4 * ew >= total_waves) {
        TIMED(K_CG_MATVEC, launch_cg_matvec(dref(), blk(K_CG_MATVEC), dep(K_CG_MATVEC), 0, stream));
      } else {
        const int grid = (int)std::min<long>(std::max<long>(ew, 64), 2048);
        TIMED(K_CG_MATVEC, launch_cg_matvec_list(dref(), dep(K_CG_MATVEC), k, grid, stream));
      }
    }
    if (k > last) break;
    HIPCHK(hipMemcpyAsync(h_active_, d.n_active + 2, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (*h_active_ <= 0) break;
    chunk = 4;
  }
  return NN_OK;
}

int Batch::step(std::vector<nn_step_stats>* stats) {
  if (scenes.empty()) return NN_OK;
  for (auto& sp : scenes)
    if (!sp->initialized) {
      set_error("step before init");
      return NN_E_STATE;
    }
  launch_step_begin(dref(), stream);
  // static particles get no list inside a step, except in batches with rigid
  // bodies: their ground collision moves positions after the search, so the
  // search positions would not survive to build those lists on demand
  CHK(search_impl(any_rigid_ ? 1 : 0));                                // :917
  if (any_elastic_ || any_rigid_) TIMED(K_BVOL, launch_bvol(dref(), blk(K_BVOL), 1, stream));  // :918
  TIMED(K_CLEAR_ACC, launch_clear_acc(dref(), blk(K_CLEAR_ACC), stream));             // :920
  if (any_fluid_) {  // DFSPHSolver.substep, :1546-1555
    TIMED(K_DENSITY_ALPHA, launch_density_alpha(dref(), blk(K_DENSITY_ALPHA), dep(K_DENSITY_ALPHA), stream));
    if (any_div_) {
      launch_solver_begin(dref(), 0, stream);
      CHK(solver_loop(0));
    }
    if (any_strain_) TIMED(K_STRAIN, launch_strain(dref(), blk(K_STRAIN), dep(K_STRAIN), stream));
    if (any_weiler_) {
      launch_solver_begin(dref(), 2, stream);
      TIMED(K_CG_INIT, launch_cg_init(dref(), blk(K_CG_INIT), stream));
      if (any_warm_cg_) launch_scene_finish(dref(), 4, stream);
      TIMED(K_CG_GUESS, launch_cg_guess(dref(), blk(K_CG_GUESS), stream));
      TIMED(K_CG_MATVEC, launch_cg_matvec(dref(), blk(K_CG_MATVEC), dep(K_CG_MATVEC), 1, stream));
      launch_scene_finish(dref(), 2, stream);
      CHK(cg_loop());
    }
    TIMED(K_NONPRESSURE, launch_nonpressure(dref(), blk(K_NONPRESSURE), dep(K_NONPRESSURE), stream));
    std::swap(d.vel, d.vel_alt);
    launch_solver_begin(dref(), 1, stream);
    TIMED(K_PRESS_SOURCE, launch_press_source(dref(), blk(K_PRESS_SOURCE), dep(K_PRESS_SOURCE), 0, stream));
    CHK(solver_loop(1));
  }
  if (any_elastic_) {  // Elasticity.substep, :1241-1244
    TIMED(K_ELASTIC_ROT_STRESS, launch_elastic_rot_stress(dref(), blk(K_ELASTIC_ROT_STRESS), stream));
    TIMED(K_ELASTIC_FORCE, launch_elastic_force(dref(), blk(K_ELASTIC_FORCE), stream));
  }
  if (any_rigid_) CHK(rigid_step());  // :929-931
  TIMED(K_ADVECT, launch_advect(dref(), blk(K_ADVECT), stream));  // :933, :936-937
  std::swap(d.pos, d.pos_alt);  // pos_alt keeps the positions of this step's search
  HIPCHK(hipGetLastError());
  CHK(gather_state());
  CHK(timer.collect());
  int rc = NN_OK;
  if (stats) stats->resize(scenes.size());
  for (size_t k = 0; k < scenes.size(); k++) {
    SceneHost& s = *scenes[k];
    const SceneState& st = h_state_[k];
    nn_step_stats o{};
    o.iterations_v = st.iter_v;
    o.iterations_p = st.iter_p;
    o.iterations_cg = st.iter_cg;
    o.status = st.status;
    o.avg_density_err_v = st.err_v;
    o.avg_density_err_p = st.err_p;
    o.cg_rel_residual = st.cg_bb > 0.0 ? std::sqrt(st.cg_rr / st.cg_bb) : 0.0;
    float mx;
    unsigned bits = st.max_rho_bits;
    std::memcpy(&mx, &bits, 4);
    o.max_density = mx;
    o.num_pairs = (int64_t)st.pairs;
    o.num_pairs_fluid = (int64_t)st.pairs_fluid;
    s.last = o;
    s.stepped = true;
    s.last_iter[0] = std::max(1, st.iter_v);
    s.last_iter[1] = std::max(1, st.iter_p);
    s.last_iter[2] = std::max(1, st.iter_cg);
    if (stats) (*stats)[k] = o;
    int e = NN_OK;
    if (st.status & ST_NAN) e = NN_E_NAN;
    if (st.status & ST_NEIGHBOR_OVERFLOW) e = NN_E_NEIGHBOR_OVERFLOW;
    if (st.status & ST_CELL_OVERFLOW) e = NN_E_CELL_OVERFLOW;
    if (st.status & ST_OUT_OF_GRID) e = NN_E_OUT_OF_GRID;
    if (e < 0 && (rc >= 0 || e < rc)) rc = e;
    if (e == NN_OK && (st.status & ST_NOCONV) && rc == NN_OK) rc = NN_W_NOCONV;
  }
  if (rc < 0) set_error("step reported an error status (see nn_step_stats.status)");
  return rc;
}

int Batch::synchronize() {
  HIPCHK(hipStreamSynchronize(stream));
  return NN_OK;
}

// ---------------------------------------------------------------------------
// downloads in original order
// ---------------------------------------------------------------------------
static int field_width(int field) {
  switch (field) {
    case NN_FIELD_X:
    case NN_FIELD_V:
    case NN_FIELD_ACCELERATION:
    case NN_FIELD_X0:
    case NN_FIELD_VISC_DIFF:
      return 3;
    case NN_FIELD_ROTATION:
      return 9;
    case NN_FIELD_STRESS:
      return 6;
    case NN_FIELD_VOLUME:
    case NN_FIELD_DENSITY:
    case NN_FIELD_DFSPH_FACTOR:
    case NN_FIELD_DENSITY_ADV:
    case NN_FIELD_STRAIN_NORM:
    case NN_FIELD_VISCOSITY:
    case NN_FIELD_REST_VOLUME:
    case NN_FIELD_KIND:
    case NN_FIELD_NUM_NEIGHBORS:
      return 1;
    default:
      return 0;
  }
}

int Batch::download(int k, int field, void* host, size_t bytes) {
  SceneHost& s = *scenes[k];
  const int w = field_width(field);
  if (w == 0) {
    set_error("unknown field");
    return NN_E_ARG;
  }
  const int n = s.n, pb = s.p_begin;
  if (bytes != (size_t)n * w * 4) {
    set_error("buffer size does not match n * width * 4");
    return NN_E_ARG;
  }
  if (field == NN_FIELD_NUM_NEIGHBORS) CHK(complete_lists());
  std::vector<int> info(n);
  HIPCHK(hipMemcpyAsync(info.data(), d.info + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
  float* out = (float*)host;
  int* iout = (int*)host;
  auto slot_f4 = [&](const float4* p, std::vector<float4>& v) -> hipError_t {
    v.resize(n);
    return hipMemcpyAsync(v.data(), p + pb, sizeof(float4) * n, hipMemcpyDeviceToHost, stream);
  };
  auto slot_f = [&](const float* p, std::vector<float>& v) -> hipError_t {
    v.resize(n);
    return hipMemcpyAsync(v.data(), p + pb, sizeof(float) * n, hipMemcpyDeviceToHost, stream);
  };
  // one component of the (src, alpha) pairs
  auto slot_sa = [&](int comp, std::vector<float>& v) -> hipError_t {
    std::vector<float2> t(n);
    hipError_t e = hipMemcpyAsync(t.data(), d.sa + pb, sizeof(float2) * n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    v.resize(n);
    for (int k = 0; k < n; k++) v[k] = comp ? t[k].y : t[k].x;
    return e;
  };
  std::vector<float4> f4;
  std::vector<float> f1;
  std::vector<int> i1;
  switch (field) {
    case NN_FIELD_X:
    case NN_FIELD_V:
    case NN_FIELD_ACCELERATION:
    case NN_FIELD_VISC_DIFF:
    case NN_FIELD_VOLUME: {
      const float4* p = field == NN_FIELD_V ? d.vel : field == NN_FIELD_ACCELERATION ? d.acc
                        : field == NN_FIELD_VISC_DIFF ? d.vdiff : d.pos;
      HIPCHK(slot_f4(p, f4));
      HIPCHK(hipStreamSynchronize(stream));
      for (int t = 0; t < n; t++) {
        const int id = info[t] & 0x0FFFFFFF;
        if (field == NN_FIELD_VOLUME) out[id] = std::fabs(f4[t].w);
        else {
          out[3 * id] = f4[t].x;
          out[3 * id + 1] = f4[t].y;
          out[3 * id + 2] = f4[t].z;
        }
      }
      return NN_OK;
    }
    case NN_FIELD_X0:
    case NN_FIELD_REST_VOLUME: {
      f4.resize(n);
      HIPCHK(hipMemcpyAsync(f4.data(), d.x0 + pb, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      for (int id = 0; id < n; id++) {
        if (field == NN_FIELD_REST_VOLUME) out[id] = f4[id].w;
        else {
          out[3 * id] = f4[id].x;
          out[3 * id + 1] = f4[id].y;
          out[3 * id + 2] = f4[id].z;
        }
      }
      return NN_OK;
    }
    case NN_FIELD_DENSITY:
    case NN_FIELD_DFSPH_FACTOR:
    case NN_FIELD_DENSITY_ADV:
    case NN_FIELD_STRAIN_NORM:
    case NN_FIELD_VISCOSITY: {
      if (field == NN_FIELD_DFSPH_FACTOR || field == NN_FIELD_DENSITY_ADV) {
        HIPCHK(slot_sa(field == NN_FIELD_DFSPH_FACTOR ? 1 : 0, f1));
      } else {
        const float* p = field == NN_FIELD_DENSITY ? d.rho : field == NN_FIELD_STRAIN_NORM ? d.gam : d.eta;
        HIPCHK(slot_f(p, f1));
      }
      HIPCHK(hipStreamSynchronize(stream));
      const SceneDev& c = s.dev;
      for (int t = 0; t < n; t++) {
        const int id = info[t] & 0x0FFFFFFF;
        const bool fluid = (info[t] >> 28) == KIND_FLUID;
        float val = f1[t];
        if (field == NN_FIELD_DENSITY) val = (fluid && s.stepped) ? f1[t] : c.rho0;  // density.fill(density0), :744
        else if (!fluid || !s.stepped) val = 0.0f;
        else if (field == NN_FIELD_DFSPH_FACTOR) {
          // the value the reference holds after a step: alpha * 1/dt * dt * 1/dt^2 (:1688,1713,1771)
          val = (c.flags & SF_DIVERGENCE) ? ((f1[t] * c.inv_dt) * c.dt) * c.inv_dt2 : f1[t] * c.inv_dt2;
        }
        out[id] = val;
      }
      return NN_OK;
    }
    case NN_FIELD_KIND:
      HIPCHK(hipStreamSynchronize(stream));
      for (int t = 0; t < n; t++) iout[info[t] & 0x0FFFFFFF] = info[t] >> 28;
      for (int b = 0; b < s.nbody; b++)  // a rigid particle reads back as KIND_RIGID + its body
        for (int m = s.body_off[b]; m < s.body_off[b + 1]; m++) iout[s.body_ids[m]] = KIND_RIGID + b;
      return NN_OK;
    case NN_FIELD_NUM_NEIGHBORS: {
      i1.resize(n);
      HIPCHK(hipMemcpyAsync(i1.data(), d.nnbr + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      for (int t = 0; t < n; t++) iout[info[t] & 0x0FFFFFFF] = i1[t];
      return NN_OK;
    }
    case NN_FIELD_ROTATION:
    case NN_FIELD_STRESS: {
      HIPCHK(hipStreamSynchronize(stream));
      for (int id = 0; id < n; id++) {
        for (int a = 0; a < w; a++) out[(size_t)w * id + a] = (field == NN_FIELD_ROTATION && (a % 4) == 0) ? 1.0f : 0.0f;
      }
      if (s.ns > 0) {
        std::vector<float4> rs((size_t)4 * s.ns);
        HIPCHK(hipMemcpyAsync(rs.data(), d.rs + 4L * s.s_begin, sizeof(float4) * 4 * s.ns, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int q = 0; q < s.ns; q++) {
          const int id = s.s_pid[q];
          const float4* r = &rs[4 * (size_t)q];
          if (field == NN_FIELD_ROTATION) {
            const float R[9] = {r[0].x, r[0].y, r[0].z, r[1].x, r[1].y, r[1].z, r[2].x, r[2].y, r[2].z};
            if (s.stepped) std::memcpy(out + 9 * (size_t)id, R, sizeof(R));
          } else {
            const float S[6] = {r[0].w, r[1].w, r[2].w, r[3].x, r[3].y, r[3].z};
            if (s.stepped) std::memcpy(out + 6 * (size_t)id, S, sizeof(S));
          }
        }
      }
      return NN_OK;
    }
  }
  set_error("unknown field");
  return NN_E_ARG;
}

int Batch::upload_field(int k, int field, const void* host, size_t bytes) {
  SceneHost& s = *scenes[k];
  const int n = s.n, pb = s.p_begin;
  if ((field != NN_FIELD_X && field != NN_FIELD_V && field != NN_FIELD_VISC_DIFF) || bytes != (size_t)n * 12) {
    set_error("only x, v and the Weiler warm-start direction (float[n][3]) can be uploaded");
    return NN_E_ARG;
  }
  std::vector<int> info(n);
  std::vector<float4> f4(n);
  float4* p = field == NN_FIELD_X ? d.pos : field == NN_FIELD_V ? d.vel : d.vdiff;
  HIPCHK(hipMemcpyAsync(info.data(), d.info + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(f4.data(), p + pb, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  const float* in = (const float*)host;
  for (int t = 0; t < n; t++) {
    const int id = info[t] & 0x0FFFFFFF;
    f4[t].x = in[3 * id];
    f4[t].y = in[3 * id + 1];
    f4[t].z = in[3 * id + 2];
  }
  HIPCHK(hipMemcpyAsync(p + pb, f4.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return NN_OK;
}

// After a step, the static particles' lists are built on demand, from the
// positions and cell order of that step's search (kept in pos_alt): the lists
// the reference holds after a step.  (A step of a batch with rigid bodies
// builds them in its search.)
int Batch::complete_lists() {
  if (lists_state_ == 2) return NN_OK;
  if (lists_state_ == 0) return search_impl(1);  // laid out since: search the current positions
  Dev at_search = d;
  at_search.pos = d.pos_alt;
  launch_build_lists(DevRef{at_search}, blk(K_BUILD_LISTS), 1, stream);
  HIPCHK(hipGetLastError());
  lists_state_ = 2;
  return NN_OK;
}

int Batch::neighbors(int k, int* nbr, int* cnt, int cap) {
  CHK(complete_lists());
  SceneHost& s = *scenes[k];
  const int n = s.n, pb = s.p_begin;
  if (cap < MAX_NB) {
    set_error("cap must be at least 60");
    return NN_E_ARG;
  }
  std::vector<int> info(n), c(n), nb((size_t)n * MAX_NB);
  HIPCHK(hipMemcpyAsync(info.data(), d.info + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(c.data(), d.nnbr + pb, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(copy_lists(nb.data(), d, pb, n, stream));
  HIPCHK(hipStreamSynchronize(stream));
  CHK(lists_to_local(nb, c, pb, n));
  for (size_t t = 0; t < (size_t)n * cap; t++) nbr[t] = -1;
  for (int t = 0; t < n; t++) {
    const int id = info[t] & 0x0FFFFFFF;
    cnt[id] = c[t];
    for (int m = 0; m < c[t]; m++) nbr[(size_t)id * cap + m] = info[nb[host_nb(n, t, m)]] & 0x0FFFFFFF;
  }
  return NN_OK;
}

}  // namespace nn
