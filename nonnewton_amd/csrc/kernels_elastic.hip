// kernels_elastic.hip -- Becker et al. 2009 corotated SPH elasticity on frozen
// rest-state neighbor lists (Elasticity, SPH/main.py:1182-1375).
//
// The stress set S holds every elastic particle and every rest neighbor of one
// (the reference forms rotations and stresses for all particles, :1191, but the
// forces of elastic particles read only those of their rest neighbors).  S is
// padded per scene to 64 so a wavefront never spans two scenes.  Rest data live
// in original-index space; the current slot of particle `id` of scene s is
// p_begin + inv[p_begin + id].
#include "device_common.h"

namespace nn {

// Elasticity.initValues_kernel, :1220-1236: x_0 and the rest volume
// V0_i = m_i / (m_i W0 + sum_j m_j W_ij) on the lists of the first search.
__global__ void k_elastic_init(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  const float4 xi = d.pos[i];
  const int cnt = d.nnbr[i];
  const int P = d.P;
  float den = sc.mass * sc.W0;
  for (int k = 0; k < cnt; k++) {
    const int j = d.nbr[nb_word(P, i, k)];
    const float4 xj = d.pos[j];
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    den += sc.mass * kern_w(sc.h, sc.kW, norm3(rx, ry, rz));
  }
  d.x0[sc.p_begin + id_of(d.info[i])] = make_float4(xi.x, xi.y, xi.z, sc.mass / den);
}

__device__ __forceinline__ bool member(const Dev& d, int& q, int& s) {
  q = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = uniform(q >> 6);
  if (w >= d.SW) return false;
  s = uniform(d.s_wave_scene[w]);
  if (s < 0) return false;
  return d.s_pid[q] >= 0;
}

// computeRotations_kernel (:1249-1274) and computeStress_kernel (:1285-1323),
// fused: the stress of particle i uses only its own rotation.
__global__ void k_elastic_rot_stress(Dev d) {
  int q, s;
  if (!member(d, q, s)) return;
  const SceneDev& sc = d.scenes[s];
  const int pb = sc.p_begin;
  const int pid = d.s_pid[q];
  const int i = pb + d.inv[pb + pid];
  const float4 xi = d.pos[i];
  const float4 xi0 = d.x0[pb + pid];
  const int cnt = d.rest_cnt[q];
  const int NS = d.NS;
  float A[9];
#pragma unroll
  for (int k = 0; k < 9; k++) A[k] = 0.0f;
  for (int k = 0; k < cnt; k++) {
    const int jid = d.rest[(long)k * NS + q];
    const float4 xj = d.pos[pb + d.inv[pb + jid]];
    const float4 xj0 = d.x0[pb + jid];
    const float dd[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
    const float d0[3] = {xj0.x - xi0.x, xj0.y - xi0.y, xj0.z - xi0.z};
    const float w = sc.mass * kern_w(sc.h, sc.kW, norm3(d0[0], d0[1], d0[2]));
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) A[3 * a + b] += w * (dd[a] * d0[b]);
  }
  float R[9];
  polar3(A, R);
  float U[9];
#pragma unroll
  for (int k = 0; k < 9; k++) U[k] = 0.0f;
  for (int k = 0; k < cnt; k++) {
    const int jid = d.rest[(long)k * NS + q];
    const float4 xj = d.pos[pb + d.inv[pb + jid]];
    const float4 xj0 = d.x0[pb + jid];
    const float dd[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
    const float d0[3] = {xj0.x - xi0.x, xj0.y - xi0.y, xj0.z - xi0.z};
    float u[3];
#pragma unroll
    for (int a = 0; a < 3; a++) u[a] = R[a] * dd[0] + R[3 + a] * dd[1] + R[6 + a] * dd[2] - d0[a];
    const float3 g3 = kern_grad(sc.h, sc.kG, d0[0], d0[1], d0[2], norm3(d0[0], d0[1], d0[2]));
    const float g[3] = {g3.x, g3.y, g3.z};
    const float vj0 = xj0.w;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) U[3 * a + b] -= (vj0 * u[a]) * g[b];
  }
  const float e[6] = {U[0], U[4], U[8], 0.5f * (U[1] + U[3]), 0.5f * (U[2] + U[6]), 0.5f * (U[5] + U[7])};
  const float Cd = sc.C_diag, Co = sc.C_off, Cs = sc.C_shear;
  const float C[6][6] = {{Cd, Co, Co, 0.f, 0.f, 0.f}, {Co, Cd, Co, 0.f, 0.f, 0.f}, {Co, Co, Cd, 0.f, 0.f, 0.f},
                         {0.f, 0.f, 0.f, Cs, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f, Cs, 0.f}, {0.f, 0.f, 0.f, 0.f, 0.f, Cs}};
  float sg[6];
#pragma unroll
  for (int a = 0; a < 6; a++) {
    float t = 0.0f;
#pragma unroll
    for (int b = 0; b < 6; b++) t += C[a][b] * e[b];
    sg[a] = t;
  }
  float4* o = d.rs + 4L * q;
  o[0] = make_float4(R[0], R[1], R[2], sg[0]);
  o[1] = make_float4(R[3], R[4], R[5], sg[1]);
  o[2] = make_float4(R[6], R[7], R[8], sg[2]);
  o[3] = make_float4(sg[3], sg[4], sg[5], 0.f);
}

__device__ __forceinline__ void sym_mat_vec(const float* M, const float* v, float* out) {
  out[0] = M[0] * v[0] + M[3] * v[1] + M[4] * v[2];
  out[1] = M[3] * v[0] + M[1] * v[1] + M[5] * v[2];
  out[2] = M[4] * v[0] + M[5] * v[1] + M[2] * v[2];
}

// computeForces_kernel, :1334-1375, for elastic particles: a += f / m.
__global__ void k_elastic_force(Dev d) {
  int q, s;
  if (!member(d, q, s)) return;
  const SceneDev& sc = d.scenes[s];
  const int pb = sc.p_begin;
  const int pid = d.s_pid[q];
  const int i = pb + d.inv[pb + pid];
  if (kind_of(d.info[i]) != KIND_ELASTIC) return;
  const float4 xi0 = d.x0[pb + pid];
  const float vi0 = xi0.w;
  const float4* ri = d.rs + 4L * q;
  const float4 a0 = ri[0], a1 = ri[1], a2 = ri[2], a3 = ri[3];
  const float Ri[9] = {a0.x, a0.y, a0.z, a1.x, a1.y, a1.z, a2.x, a2.y, a2.z};
  const float si[6] = {a0.w, a1.w, a2.w, a3.x, a3.y, a3.z};
  const int cnt = d.rest_cnt[q];
  const int NS = d.NS;
  float fi[3] = {0.f, 0.f, 0.f};
  for (int k = 0; k < cnt; k++) {
    const int jid = d.rest[(long)k * NS + q];
    const int qj = sc.s_begin + d.s_of[pb + jid];
    const float4 xj0 = d.x0[pb + jid];
    const float4* rj = d.rs + 4L * qj;
    const float4 b0 = rj[0], b1 = rj[1], b2 = rj[2], b3 = rj[3];
    const float Rj[9] = {b0.x, b0.y, b0.z, b1.x, b1.y, b1.z, b2.x, b2.y, b2.z};
    const float sj[6] = {b0.w, b1.w, b2.w, b3.x, b3.y, b3.z};
    const float d0[3] = {xj0.x - xi0.x, xj0.y - xi0.y, xj0.z - xi0.z};
    const float3 g3 = kern_grad(sc.h, sc.kG, d0[0], d0[1], d0[2], norm3(d0[0], d0[1], d0[2]));
    const float g[3] = {g3.x, g3.y, g3.z};
    const float vj0 = xj0.w;
    float dji[3], dij[3], sdji[3], sdij[3], fij[3], fji[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      dji[c] = vi0 * g[c];
      dij[c] = -vj0 * g[c];
    }
    sym_mat_vec(sj, dji, sdji);
    sym_mat_vec(si, dij, sdij);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      fij[c] = -vj0 * sdji[c];
      fji[c] = -vi0 * sdij[c];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float t1 = Rj[3 * a] * fij[0] + Rj[3 * a + 1] * fij[1] + Rj[3 * a + 2] * fij[2];
      const float t2 = Ri[3 * a] * fji[0] + Ri[3 * a + 1] * fji[1] + Ri[3 * a + 2] * fji[2];
      fi[a] += t1 - t2;
    }
  }
  float4 acc = d.acc[i];
  fi[0] = 0.5f * fi[0];
  fi[1] = 0.5f * fi[1];
  fi[2] = 0.5f * fi[2];
  acc.x += fi[0] / sc.mass;
  acc.y += fi[1] / sc.mass;
  acc.z += fi[2] / sc.mass;
  d.acc[i] = acc;
}

void launch_elastic_init(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_elastic_init, dim3((d.h.Pn + block - 1) / block), dim3(block), 0, st, d.h);
}
void launch_elastic_rot_stress(const DevRef& d, int block, hipStream_t st) {
  if (d.h.NSn <= 0) return;
  NN_LAUNCH(k_elastic_rot_stress, dim3((d.h.NSn + block - 1) / block), dim3(block), 0, st, d.h);
}
void launch_elastic_force(const DevRef& d, int block, hipStream_t st) {
  if (d.h.NSn <= 0) return;
  NN_LAUNCH(k_elastic_force, dim3((d.h.NSn + block - 1) / block), dim3(block), 0, st, d.h);
}

}  // namespace nn
