// kernels_dfsph.hip -- DFSPH sweeps over the cell-sorted particles
// (DFSPHSolver, SPH/main.py:1519-1848; SPHBase, :908-1093).
//
// One lane per particle; each lane walks its neighbor list in list order, so
// every per-particle sum is formed in the reference's serial order.  Solver
// iterations never return to the host: the source sweeps reduce the average
// density error per scene in FP64 (scene_reduce) and the last wavefront of a
// scene decides termination, clearing the scene's active flag; later launches
// of a finished scene return at once.
#include "device_common.h"

namespace nn {

// alpha as the pressure solve sees it: the reference scales the stored factor by
// 1/dt and dt around the divergence solve and then by 1/dt^2 (:1688,1713,1771).
__device__ __forceinline__ float alpha_div(const SceneDev& sc, float a) { return a * sc.inv_dt; }
__device__ __forceinline__ float alpha_press(const SceneDev& sc, float a) {
  return (sc.flags & SF_DIVERGENCE) ? ((a * sc.inv_dt) * sc.dt) * sc.inv_dt2 : a * sc.inv_dt2;
}

__global__ void k_step_begin(Dev d) {
  for (int s = threadIdx.x; s < d.nscenes; s += blockDim.x) {
    d.state[s].status = 0;
    d.state[s].max_rho_bits = 0u;
    d.state[s].counter = 0u;
    d.state[s].pairs = 0ull;
    d.state[s].pairs_fluid = 0ull;
  }
}

// clear_accelerations_kernel, :945-951
__global__ void k_clear_acc(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  const bool stat = kind_of(d.info[i]) == KIND_STATIC;
  d.acc[i] = stat ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(sc.gx, sc.gy, sc.gz, 0.f);
}

// compute_densities (:962-984), compute_DFSPH_factor (:1569-1608) and the first
// compute_density_change of the divergence solve (:1610-1645), fused: all three
// read the same neighbors and none reads what another writes.
template <int B>
__global__ void k_density_alpha(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (!(sc.flags & SF_HAS_FLUID)) return;
  float rho_out = 0.0f;
  if (i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID) {
    const float4 xi = d.pos[i];
    const float4 vi = d.vel[i];
    const int cnt = d.nnbr[i];
    const bool comp = (sc.flags & SF_COMPENSATED) != 0;
    KSum den{0.0f, 0.0f};
    float denp = 0.0f;
    float gix = 0.f, giy = 0.f, giz = 0.f, sk = 0.f, rd = 0.f;
    for_neighbors<B>(
        d, i, cnt, [&](int j) { return PosVel{d.pos[j], d.vel[j]}; },
        [&](int, const PosVel& nj) {
          const float4 xj = nj.x, vj = nj.v;
          const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
          const float rn = norm3(rx, ry, rz);
          const float Vj = fabsf(xj.w);
          const float t = Vj * kern_w(sc.h, sc.kW, rn);
          if (comp) den.add(t);
          else denp += t;
          const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, rn);
          const float gpx = -Vj * g.x, gpy = -Vj * g.y, gpz = -Vj * g.z;
          if (xj.w > 0.0f) sk += gpx * gpx + gpy * gpy + gpz * gpz;
          gix -= gpx;
          giy -= gpy;
          giz -= gpz;
          const float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
          rd += Vj * (dvx * g.x + dvy * g.y + dvz * g.z);
        });
    float rho = xi.w * sc.W0;
    rho += comp ? den.s : denp;
    rho *= sc.rho0;
    d.rho[i] = rho;
    rho_out = rho;
    sk += gix * gix + giy * giy + giz * giz;
    const float al = sk > 1e-6f ? -1.0f / sk : 0.0f;
    rd = rd > 0.0f ? rd : 0.0f;
    if (cnt < 20) rd = 0.0f;
    d.sa[i] = make_float2(rd, al);
  }
  const float m = wave_max(rho_out);
  if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(&d.state[s].max_rho_bits, __float_as_uint(m));
}

// which: 0 divergence, 1 pressure, 2 cg (cg_init activates scenes)
__global__ void k_solver_begin(Dev d, int which) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < d.nscenes; s += blockDim.x) {
    const SceneDev& sc = d.scenes[s];
    SceneState& st = d.state[s];
    int act = 0;
    if (which == 0) {
      act = (sc.flags & SF_HAS_FLUID) && (sc.flags & SF_DIVERGENCE) ? 1 : 0;
      st.active_v = act;
      st.iter_v = 0;
      st.err_v = 0.0;
    } else if (which == 1) {
      act = (sc.flags & SF_HAS_FLUID) ? 1 : 0;
      st.active_p = act;
      st.iter_p = 0;
      st.err_p = 0.0;
    } else {
      st.active_cg = 0;
      st.cg_stop_at = 0;
      st.iter_cg = 0;
      st.cg_rr = 0.0;
      st.cg_bb = 0.0;
      if (s == 0) d.cg_list_n[0] = d.cg_list_n[1] = 0;
    }
    st.counter = 0u;
    if (act) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) d.n_active[which] = cnt;
}

// divergence_solver_iteration_kernel, :1721-1760 (Jacobi: reads rho-dot and
// alpha, writes v).  Reactions on dynamic solids are not formed: the second
// clear_accelerations (:1552) erases them.
template <int B>
__global__ void k_div_update(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (!d.state[s].active_v) return;
  if (i >= sc.p_begin + sc.n || kind_of(d.info[i]) != KIND_FLUID) return;
  const float4 xi = d.pos[i];
  const float2 sai = d.sa[i];
  const float ki = sai.x * alpha_div(sc, sai.y);
  const int cnt = d.nnbr[i];
  float dvx = 0.f, dvy = 0.f, dvz = 0.f;
  for_neighbors<B>(
      d, i, cnt, [&](int j) { const float2 t = d.sa[j]; return PosSrc{d.pos[j], t.x, t.y}; },
      [&](int, const PosSrc& nj) {
        const float4 xj = nj.x;
        if (xj.w > 0.0f) {
          const float kj = nj.src * alpha_div(sc, nj.alpha);
          const float ksum = ki + kj;
          if (fabsf(ksum) > sc.eps) {
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
            const float f = sc.dt * ksum;
            const float Vj = xj.w;
            dvx -= f * (-Vj * g.x);
            dvy -= f * (-Vj * g.y);
            dvz -= f * (-Vj * g.z);
          }
        } else if (fabsf(ki) > sc.eps) {
          const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
          const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
          const float f = -sc.dt * ki;
          const float Vj = -xj.w;
          dvx += f * (-Vj * g.x);
          dvy += f * (-Vj * g.y);
          dvz += f * (-Vj * g.z);
        }
      });
  float4 v = d.vel[i];
  v.x += dvx;
  v.y += dvy;
  v.z += dvz;
  d.vel[i] = v;
}

// sum_j V_j (v_i - v_j) . gradW_ij over all neighbors, and the count
template <int B>
__device__ __forceinline__ float velocity_divergence(const Dev& d, const SceneDev& sc, int i, float4 xi, int& cnt) {
  const float4 vi = d.vel[i];
  cnt = d.nnbr[i];
  float acc = 0.f;
  for_neighbors<B>(
      d, i, cnt, [&](int j) { return PosVel{d.pos[j], d.vel[j]}; },
      [&](int, const PosVel& nj) {
        const float4 xj = nj.x, vj = nj.v;
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
        const float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
        acc += fabsf(xj.w) * (dvx * g.x + dvy * g.y + dvz * g.z);
      });
  return acc;
}

// compute_density_change (:1610-1645) + compute_density_error(0) (:1669-1675)
// and the loop test of divergence_solve (:1695-1703)
template <int B>
__global__ void k_div_source(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  SceneState* st = d.state + s;
  if (!st->active_v) return;
  double v[1] = {0.0};
  if (i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID) {
    int cnt;
    float rd = velocity_divergence<B>(d, sc, i, d.pos[i], cnt);
    rd = rd > 0.0f ? rd : 0.0f;
    if (cnt < 20) rd = 0.0f;
    d.sa[i].x = rd;
    v[0] = (double)sc.rho0 * (double)rd - 0.0;
  }
  wave_partials<1>(d, (int)(i >> 6), v);
}

// non-pressure forces (:1557-1567): surface tension (:1378-1414) and the
// explicit viscosity (:1417-1468) in one neighbor walk, the Weiler result when
// the implicit solver ran, then non_pressure_advect (:1843-1848).  The second
// clear_accelerations (:1552) is folded in.  v is double-buffered: the sweep
// reads vel and writes vel_alt.
template <int B>
__global__ void k_nonpressure(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  const int kind = kind_of(d.info[i]);
  const float4 vi = d.vel[i];
  if (!(sc.flags & SF_HAS_FLUID)) {
    d.vel_alt[i] = vi;
    return;
  }
  const int P = d.P;
  const bool vis_on = (sc.flags & SF_VISCOSITY) != 0;
  const bool weiler = vis_on && (sc.flags & SF_WEILER);
  const bool expl = vis_on && !weiler;
  const bool bnu_on = vis_on && sc.bnu != 0.0f;
  if (kind == KIND_FLUID) {
    const float4 xi = d.pos[i];
    const bool st_on = (sc.flags & SF_SURFACE_TENSION) != 0;
    const float rhoi = d.rho[i];
    float sx = 0.f, sy = 0.f, sz = 0.f, wvx = 0.f, wvy = 0.f, wvz = 0.f;
    if (st_on || expl || bnu_on) {
      const int cnt = d.nnbr[i];
      struct NJ {
        float4 x, v;
        float rho;
      };
      // surface tension reads positions only: the velocities and densities
      // are gathered only for a viscosity term formed here (two loops, so
      // that the position-only one holds no velocity registers)
      const bool need_v = expl || bnu_on;
      auto body = [&](int, const NJ& nj) {
            const float4 xj = nj.x, vj = nj.v;
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            if (xj.w > 0.0f) {
              if (st_on) {
                const float r2 = rx * rx + ry * ry + rz * rz;
                const float w = r2 > sc.diam2 ? kern_w(sc.h, sc.kW, norm3(rx, ry, rz)) : sc.Wd;
                sx -= sc.sig_m * rx * w;
                sy -= sc.sig_m * ry * w;
                sz -= sc.sig_m * rz * w;
              }
              if (expl) {
                const float vxy = (vi.x - vj.x) * rx + (vi.y - vj.y) * ry + (vi.z - vj.z) * rz;
                const float rn = norm3(rx, ry, rz);
                const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, rn);
                const float f = sc.ten_nu * (sc.mass / nj.rho) * vxy / (rn * rn + sc.visc_eps);
                wvx += f * g.x;
                wvy += f * g.y;
                wvz += f * g.z;
              }
            } else if (bnu_on) {
              const float vxy = (vi.x - vj.x) * rx + (vi.y - vj.y) * ry + (vi.z - vj.z) * rz;
              const float rn = norm3(rx, ry, rz);
              const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, rn);
              const float f = sc.ten_bnu * (sc.rho0 * (-xj.w) / rhoi) * vxy / (rn * rn + sc.visc_eps);
              wvx += f * g.x;
              wvy += f * g.y;
              wvz += f * g.z;
            }
          };
      if (need_v)
        for_neighbors<B>(d, i, cnt, [&](int j) { return NJ{d.pos[j], d.vel[j], d.rho[j]}; }, body);
      else
        for_neighbors<B>(d, i, cnt, [&](int j) { return NJ{d.pos[j], make_float4(0.f, 0.f, 0.f, 0.f), 0.0f}; }, body);
    }
    float ax = sc.gx, ay = sc.gy, az = sc.gz;
    if (st_on) {
      ax += sx;
      ay += sy;
      az += sz;
    }
    if (expl) {
      ax += wvx;
      ay += wvy;
      az += wvz;
    } else if (weiler) {
      const float3 x = d.wx[i];
      const float dx = x.x - vi.x, dy = x.y - vi.y, dz = x.z - vi.z;
      ax += dx / sc.dt;
      ay += dy / sc.dt;
      az += dz / sc.dt;
      if (sc.flags & SF_WARM_CG) d.vdiff[i] = make_float4(dx, dy, dz, 0.f);
      if (bnu_on) {
        ax += wvx;
        ay += wvy;
        az += wvz;
      }
    }
    d.acc[i] = make_float4(ax, ay, az, 0.f);
    d.vel_alt[i] = make_float4(vi.x + sc.dt * ax, vi.y + sc.dt * ay, vi.z + sc.dt * az, vi.w);
    return;
  }
  if (dyn_solid(kind)) {
    float ax = sc.gx, ay = sc.gy, az = sc.gz;
    if (bnu_on) {  // reaction of the boundary viscosity (:1458-1460), gathered
      const float4 xi = d.pos[i];
      const float Vi = -xi.w;
      const int cnt = d.nnbr[i];
      for (int k = 0; k < cnt; k++) {
        const int f = d.nbr[nb_word(P, i, k)];
        const float4 xf = d.pos[f];
        if (!(xf.w > 0.0f)) continue;
        const float4 vf = d.vel[f];
        const float rx = xf.x - xi.x, ry = xf.y - xi.y, rz = xf.z - xi.z;
        const float vxy = (vf.x - vi.x) * rx + (vf.y - vi.y) * ry + (vf.z - vi.z) * rz;
        const float rn = norm3(rx, ry, rz);
        const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, rn);
        const float ff = sc.ten_bnu * (sc.rho0 * Vi / d.rho[f]) * vxy / (rn * rn + sc.visc_eps);
        ax += -(ff * g.x);
        ay += -(ff * g.y);
        az += -(ff * g.z);
      }
    }
    d.acc[i] = make_float4(ax, ay, az, 0.f);
  } else {
    d.acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  d.vel_alt[i] = vi;
}

// compute_density_adv (:1647-1667) + compute_density_error(rho0) and the loop
// test of pressure_solve (:1778-1784)
template <int B>
__global__ void k_press_source(Dev d, int reduce) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  SceneState* st = d.state + s;
  if (!st->active_p) return;
  double v[1] = {0.0};
  if (i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID) {
    int cnt;
    const float delta = velocity_divergence<B>(d, sc, i, d.pos[i], cnt);
    float da = d.rho[i] / sc.rho0 + sc.dt * delta;
    da = da > 1.0f ? da : 1.0f;
    d.sa[i].x = da;
    v[0] = (double)sc.rho0 * (double)da - sc.rho0_d;
  }
  if (reduce) wave_partials<1>(d, (int)(i >> 6), v);
}

// Second stage of the per-scene reductions: one FINISH_THREADS-thread workgroup
// per scene adds the partials of its waves and applies the loop test of the
// solver.
//   which 0: divergence solve (:1695-1703), 1: pressure solve (:1778-1784),
//         2: CG start (after cg_guess and the first matvec),
//         4: scale of the initial guess (after cg_init, SF_WARM_CG scenes).
// (The CG's per-iteration reduction is done inside k_cg_update.)
// Thread t adds entries t, t + FINISH_THREADS, ... in order (two interleaved
// chains); the wavefront sums are combined by a fixed pairwise tree: a fixed
// order, the same bits every run.
constexpr int FINISH_THREADS = NN_FINISH_THREADS;
constexpr int FINISH_WAVES = FINISH_THREADS / 64;

__global__ __launch_bounds__(FINISH_THREADS) void k_scene_finish(Dev d, int which) {
  __shared__ double red[4][FINISH_WAVES];
  const int s = blockIdx.x;
  const int t = threadIdx.x;
  const SceneDev& sc = d.scenes[s];
  SceneState* st = d.state + s;
  const bool weiler = (sc.flags & SF_HAS_FLUID) && (sc.flags & SF_VISCOSITY) && (sc.flags & SF_WEILER);
  bool part;
  switch (which) {
    case 0: part = st->active_v != 0; break;
    case 1: part = st->active_p != 0; break;
    case 2: part = weiler; break;
    default: part = weiler && (sc.flags & SF_WARM_CG); break;  // 4
  }
  if (!part) return;  // uniform over the workgroup
  // partial arrays read and their granularity (4 = four-lane sweep, a
  // partial per 16 particles)
  const int nr = which == 2 ? 4 : (which == 4 ? 2 : 1);
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < nr; r++) {
    // the first matvec's partials (array 3) and cg_init's are per 16 particles
    const int g = which == 4 ? 4 : (which == 2 ? (r == 3 ? 4 : 1) : 1);
    const double* src = d.partials + (long)r * 4 * d.W;
    const int lo = g * sc.wave_begin, hi = g * sc.wave_end;
    double a0 = 0.0, a1 = 0.0;
    int k = lo + t;
    for (; k + FINISH_THREADS < hi; k += 2 * FINISH_THREADS) {
      a0 += src[k];
      a1 += src[k + FINISH_THREADS];
    }
    if (k < hi) a0 += src[k];
    const double w = wave_sum(a0 + a1);
    if ((t & 63) == 0) red[r][t >> 6] = w;
  }
  __syncthreads();
  if (t != 0) return;
  for (int r = 0; r < nr; r++) {  // fixed pairwise tree over the wave sums
    double w[FINISH_WAVES];
#pragma unroll
    for (int q = 0; q < FINISH_WAVES; q++) w[q] = red[r][q];
#pragma unroll
    for (int span = 1; span < FINISH_WAVES; span *= 2)
#pragma unroll
      for (int q = 0; q + span < FINISH_WAVES; q += 2 * span) w[q] += w[q + span];
    tot[r] = w[0];
  }
  if (which == 0) {
    const double err = tot[0] / sc.nf_d;
    const int ran = st->iter_v + 1;
    st->iter_v = ran;
    st->err_v = err;
    if (err <= sc.eta_v || ran >= sc.max_iter_v) {
      st->active_v = 0;
      if (!(err <= sc.eta_v)) st->status |= ST_NOCONV;
      atomicSub(&d.n_active[0], 1);
    }
  } else if (which == 1) {
    const double err = tot[0] / sc.nf_d;
    const int ran = st->iter_p + 1;
    st->iter_p = ran;
    st->err_p = err;
    if (err <= sc.eta_p || ran >= sc.max_iter_p) {
      st->active_p = 0;
      if (!(err <= sc.eta_p)) st->status |= ST_NOCONV;
      atomicSub(&d.n_active[1], 1);
    }
  } else if (which == 4) {
    // tot: e.r0, e.Ae
    st->cg_theta = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
  } else {
    // which 2, tot: b.b, r.r, r.u, u.w of the initial residual; the CG's
    // iteration 1 reads alpha and gamma from parity slot 1
    st->cg_bb = tot[0];
    st->cg_rr = tot[1];
    st->cg_gamma2[1] = tot[2];
    st->cg_alpha2[1] = tot[2] / tot[3];
    st->iter_cg = 0;
    const int act = (tot[0] > 0.0 && tot[1] > sc.cg_tol2 * tot[0]) ? 1 : 0;
    st->active_cg = act;
    st->cg_stop_at = 0;
    if (act) atomicAdd(&d.n_active[2], 1);
  }
}

// pressure_solve_iteration_kernel, :1799-1841.  Fluid lanes update v in place
// (nothing in the sweep reads v); dynamic-solid lanes gather the reaction
// -vel_change/dt of every fluid neighbor (:1835-1841) into their acceleration.
template <int B>
__global__ void k_press_update(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (!d.state[s].active_p) return;
  if (i >= sc.p_begin + sc.n) return;
  const int kind = kind_of(d.info[i]);
  const int P = d.P;
  const float4 xi = d.pos[i];
  const int cnt = d.nnbr[i];
  if (kind == KIND_FLUID) {
    const float2 sai = d.sa[i];
    const float ki = (sai.x - 1.0f) * alpha_press(sc, sai.y);
    float4 v = d.vel[i];
    for_neighbors<B>(
        d, i, cnt, [&](int j) { const float2 t = d.sa[j]; return PosSrc{d.pos[j], t.x, t.y}; },
        [&](int, const PosSrc& nj) {
          const float4 xj = nj.x;
          if (xj.w > 0.0f) {
            const float kj = (nj.src - 1.0f) * alpha_press(sc, nj.alpha);
            const float ksum = ki + kj;
            if (fabsf(ksum) > sc.eps) {
              const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
              const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
              const float f = sc.dt * ksum;
              const float Vj = xj.w;
              v.x -= f * (-Vj * g.x);
              v.y -= f * (-Vj * g.y);
              v.z -= f * (-Vj * g.z);
            }
          } else if (fabsf(ki) > sc.eps) {
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
            const float f = -sc.dt * ki;
            const float Vj = -xj.w;
            v.x += f * (-Vj * g.x);
            v.y += f * (-Vj * g.y);
            v.z += f * (-Vj * g.z);
          }
        });
    d.vel[i] = v;
  } else if (dyn_solid(kind)) {
    const float Vi = -xi.w;
    float4 a = d.acc[i];
    for (int k = 0; k < cnt; k++) {
      const int f = d.nbr[nb_word(P, i, k)];
      const float4 xf = d.pos[f];
      if (!(xf.w > 0.0f)) continue;
      const float2 saf = d.sa[f];
      const float kf = (saf.x - 1.0f) * alpha_press(sc, saf.y);
      if (fabsf(kf) > sc.eps) {
        const float rx = xf.x - xi.x, ry = xf.y - xi.y, rz = xf.z - xi.z;
        const float3 g = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
        const float fk = -sc.dt * kf;
        a.x += -(fk * (-Vi * g.x)) * 1.0f / sc.dt;
        a.y += -(fk * (-Vi * g.y)) * 1.0f / sc.dt;
        a.z += -(fk * (-Vi * g.z)) * 1.0f / sc.dt;
      }
    }
    d.acc[i] = a;
  }
}

// advect (:1086-1093) then enforce_boundary_3D for FLUID and SOLID (:1044-1080,
// :936-937), fused: each dynamic particle is touched once.
// New positions go to pos_alt (the host swaps the pair afterwards), so the
// positions the step's search used stay available until the next search.
__global__ void k_advect(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  const int kind = kind_of(d.info[i]);
  float4 x = d.pos[i];
  if (kind == KIND_STATIC) {
    d.pos_alt[i] = x;
    return;
  }
  float4 v = d.vel[i];
  if (dyn_solid(kind)) {
    const float4 a = d.acc[i];
    v.x += sc.dt * a.x;
    v.y += sc.dt * a.y;
    v.z += sc.dt * a.z;
  }
  x.x += sc.dt * v.x;
  x.y += sc.dt * v.y;
  x.z += sc.dt * v.z;
  const float px = x.x, py = x.y, pz = x.z;
  float nx = 0.f, ny = 0.f, nz = 0.f;
  if (px > sc.hi0) { nx += 1.0f; x.x = sc.hi0; }
  if (px <= sc.lo0) { nx += -1.0f; x.x = sc.lo0; }
  if (py > sc.hi1) { ny += 1.0f; x.y = sc.hi1; }
  if (py <= sc.lo1) { ny += -1.0f; x.y = sc.lo1; }
  if (pz > sc.hi2) { nz += 1.0f; x.z = sc.hi2; }
  if (pz <= sc.lo2) { nz += -1.0f; x.z = sc.lo2; }
  const float nl = norm3(nx, ny, nz);
  if (nl > 1e-6f) {
    const float hx = nx / nl, hy = ny / nl, hz = nz / nl;
    const float f = sc.opc * (v.x * hx + v.y * hy + v.z * hz);
    v.x -= f * hx;
    v.y -= f * hy;
    v.z -= f * hz;
    v.x *= sc.sticky;
    v.y *= sc.sticky;
    v.z *= sc.sticky;
  }
  if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) && isfinite(v.y) && isfinite(v.z)))
    atomicOr(&d.state[s].status, ST_NAN);
  d.pos_alt[i] = x;
  d.vel[i] = v;
}

static inline dim3 grid_of(const DevRef& d, int block) { return dim3((d.h.Pn + block - 1) / block); }

void launch_step_begin(const DevRef& d, hipStream_t st) {
  NN_LAUNCH(k_step_begin, dim3(1), dim3(256), 0, st, d.h);
}
void launch_clear_acc(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_clear_acc, grid_of(d, block), dim3(block), 0, st, d.h);
}
void launch_density_alpha(const DevRef& d, int block, int depth, hipStream_t st) {
  DEPTH_LAUNCH(k_density_alpha, depth, grid_of(d, block), dim3(block), st, d.h);
}
void launch_solver_begin(const DevRef& d, int which, hipStream_t st) {
  NN_LAUNCH(k_solver_begin, dim3(1), dim3(256), 0, st, d.h, which);
}
void launch_div_update(const DevRef& d, int block, int depth, hipStream_t st) {
  DEPTH_LAUNCH(k_div_update, depth, grid_of(d, block), dim3(block), st, d.h);
}
void launch_div_source(const DevRef& d, int block, int depth, hipStream_t st) {
  DEPTH_LAUNCH(k_div_source, depth, grid_of(d, block), dim3(block), st, d.h);
}
void launch_nonpressure(const DevRef& d, int block, int depth, hipStream_t st) {
  DEPTH_LAUNCH(k_nonpressure, depth, grid_of(d, block), dim3(block), st, d.h);
}
void launch_press_source(const DevRef& d, int block, int depth, int reduce, hipStream_t st) {
  DEPTH_LAUNCH(k_press_source, depth, grid_of(d, block), dim3(block), st, d.h, reduce);
}
void launch_press_update(const DevRef& d, int block, int depth, hipStream_t st) {
  DEPTH_LAUNCH(k_press_update, depth, grid_of(d, block), dim3(block), st, d.h);
}
void launch_advect(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_advect, grid_of(d, block), dim3(block), 0, st, d.h);
}
void launch_scene_finish(const DevRef& d, int which, hipStream_t st) {
  if (d.h.nscenes > 0) NN_LAUNCH(k_scene_finish, dim3(d.h.nscenes), dim3(FINISH_THREADS), 0, st, d.h, which);
}

}  // namespace nn
