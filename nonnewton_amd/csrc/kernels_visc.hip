// kernels_visc.hip -- non-Newtonian viscosity and the Weiler 2018 implicit
// viscosity solver.
//
// The reference computes a per-particle non-Newtonian viscosity that no force
// reads (NonNewton, SPH/main.py:332-492) and declares an implicit solver whose
// operator is empty (Viscosity_Weiler2018, :1471-1513).  Here the strain rate
// feeds the model, the model feeds the implicit solve:
//   D_i   = (0.5 / rho_i) sum_j m_j (v_ji (x) gradW_ij + gradW_ij (x) v_ji)
//   eta_i = model(||D_i||_F)                        (kinematic, :404-459)
//   mu_i  = rho0 * eta_i   (rho0 * viscosity when the model is off)
//   (A x)_i = x_i - (dt / rho_i) sum_{fluid j} (s_ij ((x_i - x_j) . r_ij)) r_ij
//   s_ij  = c_ij F(|r_ij|) / (|r_ij| h)   (gradW_ij = F r_ij / (|r_ij| h))
//   c_ij  = 10 mu_ij (m / rho_j) / (|r_ij|^2 + 0.01 h^2),  mu_ij = (mu_i + mu_j) / 2
// Weiler et al. 2018, "A Physically Consistent Implicit Viscosity Solver for
// SPH Fluids", solved by preconditioned conjugate gradients: block-Jacobi 3x3
// preconditioner M, FP64 dot products, stop when ||r|| <= tol ||b|| with b = v.
// A is symmetric (f_i s_ij = 10 dt mu_ij m F / (rho_i rho_j |r| h (|r|^2 +
// 0.01 h^2)) with mu_ij the mean), so CG applies; it runs in the
// Chronopoulos-Gear arrangement, which needs one reduction per iteration, and
// an iteration is two launches:
//   update sweep:  the scene's sums of the previous iteration -> test r.r,
//                  gamma = r.u, delta = u.w, beta = gamma / gamma_old,
//                  alpha = gamma / (delta - beta gamma / alpha_old); then
//                  p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s,
//                  u = M r; partials (r.r, r.u)
//   matvec sweep:  w = A u (gathers u of the neighbors); partial u.w
// The pair scalars s_ij are formed once per solve (cg_init) and stored next to
// the neighbor lists, so a CG matvec needs no sqrt or division.  cg_init and
// cg_matvec run four lanes per particle, each lane over every fourth list
// position; list sums are (S0 + S1) + (S2 + S3).
#include "device_common.h"

namespace nn {

__device__ __forceinline__ bool weiler_scene(const SceneDev& sc) {
  return (sc.flags & SF_HAS_FLUID) && (sc.flags & SF_VISCOSITY) && (sc.flags & SF_WEILER);
}

// NonNewton.calcStrainRate (:461-492, with the accumulation the reference
// omits) + the viscosity model; mu for the implicit solver.
template <int B>
__global__ void k_strain(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (!(sc.flags & SF_HAS_FLUID)) return;
  const bool nn_on = (sc.flags & SF_NON_NEWTONIAN) != 0;
  const bool weiler = weiler_scene(sc);
  if (!nn_on && !weiler) return;
  if (i >= sc.p_begin + sc.n || kind_of(d.info[i]) != KIND_FLUID) return;
  float eta = 0.0f;
  if (nn_on) {
    const float4 xi = d.pos[i];
    const float4 vi = d.vel[i];
    const int cnt = d.nnbr[i];
    float S[9];
#pragma unroll
    for (int q = 0; q < 9; q++) S[q] = 0.0f;
    for_neighbors<B>(
        d, i, cnt, [&](int j) { return PosVel{d.pos[j], d.vel[j]}; },
        [&](int, const PosVel& nj) {
          const float4 xj = nj.x, vj = nj.v;
          const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
          const float3 g3 = kern_grad(sc.h, sc.kG, rx, ry, rz, norm3(rx, ry, rz));
          const float g[3] = {g3.x, g3.y, g3.z};
          const float vji[3] = {vj.x - vi.x, vj.y - vi.y, vj.z - vi.z};
#pragma unroll
          for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) S[3 * a + b] += sc.mass * (vji[a] * g[b] + vji[b] * g[a]);
        });
    const float f = 0.5f / d.rho[i];
    float n2 = 0.0f;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const float t = f * S[q];
      n2 += t * t;
    }
    const float gam = sqrtf(n2);
    eta = nn_model(sc, gam);
    d.gam[i] = gam;
    d.eta[i] = eta;
  }
  if (weiler) d.mu[i] = (nn_on ? eta : sc.nu) * sc.rho0;
}

// Start of a solve, b = v.  Pair scalars s_ij (stored), the inverse diagonal
// block M, r0 = b - A b (into wr).  With SF_WARM_CG also A e for the previous
// solve's e = x - v (into ww) and partial sums of (e.r0, e.Ae), from which the
// scene finish forms theta = (e.r0) / (e.Ae); k_cg_guess then starts CG from
// x0 = b + theta e, r = r0 - theta A e.
__global__ void k_cg_init(Dev d) {
  int i, q, s, hw;
  if (!quad_slot_scene(d, i, q, s, hw)) return;
  const SceneDev& sc = d.scenes[s];
  if (!weiler_scene(sc)) return;
  double v[2] = {0.0, 0.0};
  const bool valid = i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID;
  const bool warm = (sc.flags & SF_WARM_CG) != 0;
  float ax = 0.f, ay = 0.f, az = 0.f;  // sum for A v
  float bx = 0.f, by = 0.f, bz = 0.f;  // sum for A e
  float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f, m5 = 0.f;
  float4 xi = make_float4(0.f, 0.f, 0.f, 0.f), vi = xi, ei = xi;
  if (valid) {
    xi = d.pos[i];
    vi = d.vel[i];
    if (warm) ei = d.vdiff[i];
    const float mui = d.mu[i];
    const int cnt = d.nnbr[i];
    const long P = d.P;
    // one neighbor's contribution, in list order
    auto pair_term = [&](int k, const float4& xj, const float4& vj, const float4& ej, float muj, float rhoj) {
      float sij = 0.0f;
      if (xj.w > 0.0f) {
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        sij = weiler_pair(sc, rx, ry, rz, mui, muj, rhoj);
        const float t = sij * ((vi.x - vj.x) * rx + (vi.y - vj.y) * ry + (vi.z - vj.z) * rz);
        ax += t * rx;
        ay += t * ry;
        az += t * rz;
        if (warm) {
          const float te = sij * ((ei.x - ej.x) * rx + (ei.y - ej.y) * ry + (ei.z - ej.z) * rz);
          bx += te * rx;
          by += te * ry;
          bz += te * rz;
        }
        m0 += sij * rx * rx;
        m1 += sij * rx * ry;
        m2 += sij * rx * rz;
        m3 += sij * ry * ry;
        m4 += sij * ry * rz;
        m5 += sij * rz * rz;
      }
      d.spair[nb_word(P, i, k)] = sij;
    };
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // two list positions of this lane per batch, their loads issued together
    for (int k = q; k < cnt; k += 8) {
      const bool two = k + 4 < cnt;
      const int j0 = d.nbr[nb_word(P, i, k)];
      const int j1 = two ? d.nbr[nb_word(P, i, k + 4)] : i;
      const float4 x0 = d.pos[j0], x1 = d.pos[j1];
      const float4 v0 = d.vel[j0], v1 = d.vel[j1];
      const float mu0 = d.mu[j0], mu1 = d.mu[j1];
      const float r0 = d.rho[j0], r1 = d.rho[j1];
      const float4 e0 = warm ? d.vdiff[j0] : zero4;
      const float4 e1 = warm ? d.vdiff[j1] : zero4;
      pair_term(k, x0, v0, e0, mu0, r0);
      if (two) pair_term(k + 4, x1, v1, e1, mu1, r1);
    }
  }
  ax = quad_sum(ax);
  ay = quad_sum(ay);
  az = quad_sum(az);
  if (warm) {
    bx = quad_sum(bx);
    by = quad_sum(by);
    bz = quad_sum(bz);
  }
  m0 = quad_sum(m0);
  m1 = quad_sum(m1);
  m2 = quad_sum(m2);
  m3 = quad_sum(m3);
  m4 = quad_sum(m4);
  m5 = quad_sum(m5);
  // the positions the CG gathers, for every particle of the scene
  if (q == 0 && i < sc.p_begin + sc.n) {
    const float4 p = valid ? xi : d.pos[i];
    d.xu[i].p = make_float3(p.x, p.y, p.z);
  }
  if (valid && q == 0) {
    const float f = sc.dt / d.rho[i];
    const float qx = vi.x - f * ax, qy = vi.y - f * ay, qz = vi.z - f * az;
    const float rrx = vi.x - qx, rry = vi.y - qy, rrz = vi.z - qz;
    const float d00 = 1.0f - f * m0, d01 = -f * m1, d02 = -f * m2;
    const float d11 = 1.0f - f * m3, d12 = -f * m4, d22 = 1.0f - f * m5;
    const float c00 = d11 * d22 - d12 * d12, c01 = d02 * d12 - d01 * d22, c02 = d01 * d12 - d02 * d11;
    const float c11 = d00 * d22 - d02 * d02, c12 = d01 * d02 - d00 * d12, c22 = d00 * d11 - d01 * d01;
    const float det = d00 * c00 + d01 * c01 + d02 * c02;
    float4 w0;
    float2 w1;
    if (det > 0.0f && det == det) {
      const float id = 1.0f / det;
      w0 = make_float4(c00 * id, c01 * id, c02 * id, c11 * id);
      w1 = make_float2(c12 * id, c22 * id);
    } else {
      w0 = make_float4(1.f, 0.f, 0.f, 1.f);
      w1 = make_float2(0.f, 1.f);
    }
    d.wd0[i] = w0;
    d.wd1[i] = w1;
    d.wr[i] = make_float3(rrx, rry, rrz);
    if (warm) {
      const float zx = ei.x - f * bx, zy = ei.y - f * by, zz = ei.z - f * bz;
      d.ww[i] = make_float3(zx, zy, zz);
      v[0] = (double)ei.x * (double)rrx + (double)ei.y * (double)rry + (double)ei.z * (double)rrz;
      v[1] = (double)ei.x * (double)zx + (double)ei.y * (double)zy + (double)ei.z * (double)zz;
    }
  }
  wave_partials<2>(d, hw, v);
}

// x0 = b + theta e, r = r0 - theta A e (x0 = b, r = r0 without SF_WARM_CG),
// u = M r; partials (b.b, r.r, r.u).
__global__ void k_cg_guess(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (!weiler_scene(sc)) return;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID) {
    const float4 b = d.vel[i];
    float3 r = d.wr[i];
    float3 x = make_float3(b.x, b.y, b.z);
    if (sc.flags & SF_WARM_CG) {
      const float th = (float)d.state[s].cg_theta;
      const float4 e = d.vdiff[i];
      const float3 z = d.ww[i];
      x = make_float3(b.x + th * e.x, b.y + th * e.y, b.z + th * e.z);
      r = make_float3(r.x - th * z.x, r.y - th * z.y, r.z - th * z.z);
    }
    const float4 w0 = d.wd0[i];
    const float2 w1 = d.wd1[i];
    const float ux = w0.x * r.x + w0.y * r.y + w0.z * r.z;
    const float uy = w0.y * r.x + w0.w * r.y + w1.x * r.z;
    const float uz = w0.z * r.x + w1.x * r.y + w1.y * r.z;
    d.wx[i] = x;
    d.wr[i] = r;
    d.xu[i].u = make_float3(ux, uy, uz);
    v[0] = (double)b.x * (double)b.x + (double)b.y * (double)b.y + (double)b.z * (double)b.z;
    v[1] = (double)r.x * (double)r.x + (double)r.y * (double)r.y + (double)r.z * (double)r.z;
    v[2] = (double)r.x * (double)ux + (double)r.y * (double)uy + (double)r.z * (double)uz;
  }
  wave_partials<3>(d, (int)(i >> 6), v);
}

// w = A u (u gathered from the neighbors); partial u.w into partial array 3
// (start: after cg_guess, every Weiler scene; read by the start reduction --
// otherwise: an iteration, active scenes; read by the next update).  D list
// positions of a lane per batch: their list words and pair scalars, then their
// neighbors' records (position and u, 24 bytes), are loaded before any of them is used (a
// shorter chain of dependent loads), and the next batch's list words and pair
// scalars are requested right after this batch's records, so that they arrive
// while this batch is summed; the terms are added in list order whatever D is.
// A pair with s_ij = 0 -- every neighbor that is not fluid --
// adds nothing and is skipped, so the neighbor's u is never read there.
template <int D>
__global__ void k_cg_matvec(Dev d, int start) {
  int i, q, s, hw;
  if (!quad_slot_scene(d, i, q, s, hw)) return;
  const SceneDev& sc = d.scenes[s];
  const SceneState* st = d.state + s;
  // the start product runs for every Weiler scene, an iteration's for the
  // scenes that no update has stopped
  if (!weiler_scene(sc) || !(start || (st->active_cg && st->cg_stop_at == 0))) return;
  double v[1] = {0.0};
  const bool valid = i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID;
  float ax = 0.f, ay = 0.f, az = 0.f;
  float3 ui = make_float3(0.f, 0.f, 0.f);
  if (valid) {
    const XU own = d.xu[i];
    const float3 xi = own.p;
    ui = own.u;
    const int cnt = d.nnbr[i];
    const long P = d.P;
    int jn[D];
    float sn[D];
#pragma unroll
    for (int e = 0; e < D; e++) {
      const bool in = q + 4 * e < cnt;
      jn[e] = in ? d.nbr[nb_word(P, i, q + 4 * e)] : i;
      sn[e] = in ? d.spair[nb_word(P, i, q + 4 * e)] : 0.0f;
    }
    for (int k = q; k < cnt; k += 4 * D) {
      int j[D];
      float sp[D];
      float3 xj[D], uj[D];
#pragma unroll
      for (int e = 0; e < D; e++) {
        j[e] = jn[e];
        sp[e] = sn[e];
      }
#pragma unroll
      for (int e = 0; e < D; e++) {
        const XU nb = d.xu[j[e]];
        xj[e] = nb.p;
        uj[e] = nb.u;
      }
      // the next batch's list words and pair scalars, in flight while this
      // batch's records arrive
#pragma unroll
      for (int e = 0; e < D; e++) {
        const int kk = k + 4 * D + 4 * e;
        const bool in = kk < cnt;
        jn[e] = in ? d.nbr[nb_word(P, i, kk)] : i;
        sn[e] = in ? d.spair[nb_word(P, i, kk)] : 0.0f;
      }
#pragma unroll
      for (int e = 0; e < D; e++) {
        if (sp[e] != 0.0f) {
          const float rx = xi.x - xj[e].x, ry = xi.y - xj[e].y, rz = xi.z - xj[e].z;
          const float t = sp[e] * ((ui.x - uj[e].x) * rx + (ui.y - uj[e].y) * ry + (ui.z - uj[e].z) * rz);
          ax += t * rx;
          ay += t * ry;
          az += t * rz;
        }
      }
    }
  }
  ax = quad_sum(ax);
  ay = quad_sum(ay);
  az = quad_sum(az);
  if (valid && q == 0) {
    const float f = sc.dt / d.rho[i];
    const float wx = ui.x - f * ax, wy = ui.y - f * ay, wz = ui.z - f * az;
    d.ww[i] = make_float3(wx, wy, wz);
    v[0] = (double)ui.x * (double)wx + (double)ui.y * (double)wy + (double)ui.z * (double)wz;
  }
  wave_partials<1>(d, hw, v, 3);
}

// Matvec k of the CG over the waves listed by update k (the scenes still
// iterating), for the late iterations when few scenes are left: a grid sized
// to the expected work, workgroup = 256 threads = the four lanes of one wave's
// 64 particles, looping over the list.  Same arithmetic and the same partials
// (one per 16 particles, array 3) as k_cg_matvec.
constexpr int MVL_THREADS = 256;

template <int D>
__global__ __launch_bounds__(MVL_THREADS) void k_cg_matvec_list(Dev d, int k) {
  const int count = d.cg_list_n[k & 1];
  const int* list = d.cg_list + (long)(k & 1) * d.W;
  const int q = threadIdx.x & 3;
  for (int u = blockIdx.x; u < count; u += gridDim.x) {
    const int w = list[u];
    const int s = uniform(d.wave_scene[w]);
    const SceneDev& sc = d.scenes[s];
    const int i = w * 64 + (threadIdx.x >> 2);
    const bool valid = i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID;
    float ax = 0.f, ay = 0.f, az = 0.f;
    float3 ui = make_float3(0.f, 0.f, 0.f);
    if (valid) {
      const XU own = d.xu[i];
      const float3 xi = own.p;
      ui = own.u;
      const int cnt = d.nnbr[i];
      const long P = d.P;
      int jn[D];
      float sn[D];
#pragma unroll
      for (int e = 0; e < D; e++) {
        const bool in = q + 4 * e < cnt;
        jn[e] = in ? d.nbr[nb_word(P, i, q + 4 * e)] : i;
        sn[e] = in ? d.spair[nb_word(P, i, q + 4 * e)] : 0.0f;
      }
      for (int kk = q; kk < cnt; kk += 4 * D) {
        int j[D];
        float sp[D];
        float3 xj[D], uj[D];
#pragma unroll
        for (int e = 0; e < D; e++) {
          j[e] = jn[e];
          sp[e] = sn[e];
        }
#pragma unroll
        for (int e = 0; e < D; e++) {
          const XU nb = d.xu[j[e]];
          xj[e] = nb.p;
          uj[e] = nb.u;
        }
#pragma unroll
        for (int e = 0; e < D; e++) {
          const int k2 = kk + 4 * D + 4 * e;
          const bool in = k2 < cnt;
          jn[e] = in ? d.nbr[nb_word(P, i, k2)] : i;
          sn[e] = in ? d.spair[nb_word(P, i, k2)] : 0.0f;
        }
#pragma unroll
        for (int e = 0; e < D; e++) {
          if (sp[e] != 0.0f) {
            const float rx = xi.x - xj[e].x, ry = xi.y - xj[e].y, rz = xi.z - xj[e].z;
            const float t = sp[e] * ((ui.x - uj[e].x) * rx + (ui.y - uj[e].y) * ry + (ui.z - uj[e].z) * rz);
            ax += t * rx;
            ay += t * ry;
            az += t * rz;
          }
        }
      }
    }
    ax = quad_sum(ax);
    ay = quad_sum(ay);
    az = quad_sum(az);
    double v[1] = {0.0};
    if (valid && q == 0) {
      const float f = sc.dt / d.rho[i];
      const float wx = ui.x - f * ax, wy = ui.y - f * ay, wz = ui.z - f * az;
      d.ww[i] = make_float3(wx, wy, wz);
      v[0] = (double)ui.x * (double)wx + (double)ui.y * (double)wy + (double)ui.z * (double)wz;
    }
    wave_partials<1>(d, 4 * w + (threadIdx.x >> 6), v, 3);  // the quarter-wave index k_cg_matvec uses
  }
}

// Iteration k of the CG with the scene's reduction of iteration k - 1 done in
// the same launch (no separate reduction kernel per iteration).  Workgroup =
// up to CG_UPD_WAVES waves of one scene (d.cg_blocks).  Every workgroup of a scene
// first adds the scene's partials -- r.r and r.u of update k - 1, u.w of
// matvec k - 1 -- in one fixed order, so all of them obtain the same bits and
// take the same decision:
//   converged (r.r <= tol^2 b.b) or k - 1 iterations reached the limit: the
//     scene leaves the solve (its first workgroup records it);
//   otherwise beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma /
//     alpha_old) (iteration 1 takes alpha from the start reduction, beta = 0)
// and then p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s,
// u = M r, with partials (r.r, r.u) into arrays 4 + 2 (k & 1), 5 + 2 (k & 1).
constexpr int UPD_THREADS = 64 * CG_UPD_WAVES;
constexpr int UPD_WAVES = UPD_THREADS / 64;

// Which scenes take part is read from state that no workgroup of this launch
// changes in a way that matters to the others: active_cg is set by the start
// reduction, and cg_stop_at by an earlier update -- or by this one's first
// workgroup of the scene, as k, which reads the same as 0 here.  (Had that
// workgroup cleared a flag that the others read, a workgroup whose waves read
// it on both sides of the store would split at the barrier below, some waves
// leaving and the rest reducing with wave sums never written.)  A workgroup of
// a scene stopping now reaches the same decision from the same sums and
// leaves without writing.
__global__ __launch_bounds__(UPD_THREADS) void k_cg_update(Dev d, int k) {
  __shared__ double red[3][UPD_WAVES];
  const int2 e = d.cg_blocks[blockIdx.x];
  const int s = e.x;
  const SceneDev& sc = d.scenes[s];
  SceneState* st = d.state + s;
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t == 0) d.cg_list_n[(k + 1) & 1] = 0;  // for update k + 1
  const int stop = st->cg_stop_at;
  if (!st->active_cg || (stop != 0 && stop < k)) return;  // uniform over the workgroup
  const bool lead = e.y == sc.wave_begin && t == 0;
  double al, be;
  if (k == 1) {
    al = st->cg_alpha2[1];
    be = 0.0;
  } else {
    const int q = (k - 1) & 1;
    const double* src[3] = {d.partials + (long)(4 + 2 * q) * 4 * d.W, d.partials + (long)(5 + 2 * q) * 4 * d.W,
                            d.partials + 3L * 4 * d.W};
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int g = r == 2 ? 4 : 1;  // the matvec's partials are per 16 particles
      const int lo = g * sc.wave_begin, hi = g * sc.wave_end;
      double a0 = 0.0, a1 = 0.0;
      int j = lo + t;
      for (; j + UPD_THREADS < hi; j += 2 * UPD_THREADS) {
        a0 += src[r][j];
        a1 += src[r][j + UPD_THREADS];
      }
      if (j < hi) a0 += src[r][j];
      const double w = wave_sum(a0 + a1);
      if ((t & 63) == 0) red[r][t >> 6] = w;
    }
    __syncthreads();
    double tot[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {  // fixed pairwise tree over the wave sums
      double w[UPD_WAVES];
#pragma unroll
      for (int v = 0; v < UPD_WAVES; v++) w[v] = red[r][v];
#pragma unroll
      for (int span = 1; span < UPD_WAVES; span *= 2)
#pragma unroll
        for (int v = 0; v + span < UPD_WAVES; v += 2 * span) w[v] += w[v + span];
      tot[r] = w[0];
    }
    const double rr = tot[0], ga = tot[1], de = tot[2];
    const int it = k - 1;
    const bool conv = rr <= sc.cg_tol2 * st->cg_bb;
    if (conv || it >= sc.max_iter_cg) {
      if (lead) {
        st->iter_cg = it;
        st->cg_rr = rr;
        st->cg_stop_at = k;
        if (!conv) st->status |= ST_NOCONV;
        atomicSub(&d.n_active[2], 1);
      }
      return;
    }
    be = ga / st->cg_gamma2[q];
    al = ga / (de - be * ga / st->cg_alpha2[q]);
    if (lead) {
      st->iter_cg = it;
      st->cg_rr = rr;
      st->cg_gamma2[k & 1] = ga;
      st->cg_alpha2[k & 1] = al;
    }
  }
  // the scene iterates on: its first workgroup lists its waves for matvec k
  if (e.y == sc.wave_begin) {
    __shared__ int base;
    const int nw = sc.wave_end - sc.wave_begin;
    if (t == 0) base = atomicAdd(&d.cg_list_n[k & 1], nw);
    __syncthreads();
    int* out = d.cg_list + (long)(k & 1) * d.W + base;
    for (int m = t; m < nw; m += UPD_THREADS) out[m] = sc.wave_begin + m;
  }
  const bool first = k == 1;
  const float alf = (float)al, bef = (float)be;
  const int w = e.y + (t >> 6);
  if (w >= sc.wave_end) return;  // whole waves past the scene
  const int i = w * 64 + (t & 63);
  double v[2] = {0.0, 0.0};
  if (i < sc.p_begin + sc.n && kind_of(d.info[i]) == KIND_FLUID) {
    const float3 u = d.xu[i].u;
    const float3 ww = d.ww[i];
    float3 p, sv;
    if (first) {
      p = u;
      sv = ww;
    } else {
      const float3 po = d.wp[i];
      const float3 so = d.ws[i];
      p = make_float3(u.x + bef * po.x, u.y + bef * po.y, u.z + bef * po.z);
      sv = make_float3(ww.x + bef * so.x, ww.y + bef * so.y, ww.z + bef * so.z);
    }
    float3 x = d.wx[i];
    float3 r = d.wr[i];
    x.x += alf * p.x;
    x.y += alf * p.y;
    x.z += alf * p.z;
    r.x -= alf * sv.x;
    r.y -= alf * sv.y;
    r.z -= alf * sv.z;
    const float4 w0 = d.wd0[i];
    const float2 w1 = d.wd1[i];
    const float zx = w0.x * r.x + w0.y * r.y + w0.z * r.z;
    const float zy = w0.y * r.x + w0.w * r.y + w1.x * r.z;
    const float zz = w0.z * r.x + w1.x * r.y + w1.y * r.z;
    d.wp[i] = p;
    d.ws[i] = sv;
    d.wx[i] = x;
    d.wr[i] = r;
    d.xu[i].u = make_float3(zx, zy, zz);
    v[0] = (double)r.x * (double)r.x + (double)r.y * (double)r.y + (double)r.z * (double)r.z;
    v[1] = (double)r.x * (double)zx + (double)r.y * (double)zy + (double)r.z * (double)zz;
  }
  wave_partials<2>(d, w, v, 4 + 2 * (k & 1));
}

static inline dim3 grid_of(const DevRef& d, int block) { return dim3((d.h.Pn + block - 1) / block); }
// four-lane sweeps: a workgroup covers the particles of a one-lane workgroup
// (fewer workgroups to dispatch when most scenes have left the solve)
static inline int quad_block(int block) { return block * 4 <= 1024 ? block * 4 : 1024; }
static inline dim3 grid4_of(const DevRef& d, int block) {
  const int qb = quad_block(block);
  return dim3((4 * d.h.Pn + qb - 1) / qb);
}

void launch_strain(const DevRef& d, int block, int depth, hipStream_t st) {
  DEPTH_LAUNCH(k_strain, depth, grid_of(d, block), dim3(block), st, d.h);
}
void launch_cg_init(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_cg_init, grid4_of(d, block), dim3(quad_block(block)), 0, st, d.h);
}
void launch_cg_guess(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_cg_guess, grid_of(d, block), dim3(block), 0, st, d.h);
}
void launch_cg_matvec(const DevRef& d, int block, int depth, int start, hipStream_t st) {
  const dim3 g = grid4_of(d, block), b = dim3(quad_block(block));
  switch (depth) {
    case 1: NN_LAUNCH(k_cg_matvec<1>, g, b, 0, st, d.h, start); break;
    case 4: NN_LAUNCH(k_cg_matvec<4>, g, b, 0, st, d.h, start); break;
    case 8: NN_LAUNCH(k_cg_matvec<8>, g, b, 0, st, d.h, start); break;
    default: NN_LAUNCH(k_cg_matvec<2>, g, b, 0, st, d.h, start); break;
  }
}
void launch_cg_matvec_list(const DevRef& d, int depth, int k, int grid, hipStream_t st) {
  const dim3 g = dim3(grid > 0 ? grid : 1), b = dim3(MVL_THREADS);
  switch (depth) {
    case 1: NN_LAUNCH(k_cg_matvec_list<1>, g, b, 0, st, d.h, k); break;
    case 4: NN_LAUNCH(k_cg_matvec_list<4>, g, b, 0, st, d.h, k); break;
    case 8: NN_LAUNCH(k_cg_matvec_list<8>, g, b, 0, st, d.h, k); break;
    default: NN_LAUNCH(k_cg_matvec_list<2>, g, b, 0, st, d.h, k); break;
  }
}
void launch_cg_update(const DevRef& d, int nblocks, int k, hipStream_t st) {
  if (nblocks > 0) NN_LAUNCH(k_cg_update, dim3(nblocks), dim3(UPD_THREADS), 0, st, d.h, k);
}

}  // namespace nn
