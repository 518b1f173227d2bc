// device_common.h -- device helpers shared by the kernels of libnnsph.
//
// Every floating-point expression here is written in the operation order of the
// reference's Taichi code (cited per function) and compiled with
// -ffp-contract=off, so a result is rounded exactly as the reference's FP32
// expression would be.  Division and sqrt are IEEE correctly rounded (hipcc's
// default for single precision).
#pragma once
#include "engine.h"

namespace nn {

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Slot of this thread and the scene of its wavefront; returns false for waves
// past the batch (the whole wave leaves together).
__device__ __forceinline__ bool slot_scene(const Dev& d, int& i, int& s) {
  i = blockIdx.x * blockDim.x + threadIdx.x;
  int w = uniform(i >> 6);
  if (w >= d.W) return false;
  s = uniform(d.wave_scene[w]);
  return s >= 0;
}

__device__ __forceinline__ int kind_of(int info) { return info >> 28; }
__device__ __forceinline__ int id_of(int info) { return info & 0x0FFFFFFF; }
// dynamic solids (SOLID and is_dynamic, SPH/main.py:281-282): elastic and rigid particles
__device__ __forceinline__ bool dyn_solid(int kind) { return kind == KIND_ELASTIC || kind == KIND_RIGID; }

// ---------------------------------------------------------------------------
// SPH kernel, SPH/main.py:859-880 (cubic_kernel) and :883-905 (derivative)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float norm3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

__device__ __forceinline__ float kern_w(float h, float kW, float rn) {
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

// gradient of W at r with |r| = rn already known (rn = norm3(r) bit for bit)
__device__ __forceinline__ float3 kern_grad(float h, float kG, float rx, float ry, float rz, float rn) {
  float3 g = make_float3(0.0f, 0.0f, 0.0f);
  float q = rn / h;
  if (rn > 1e-5f && q <= 1.0f) {
    float den = rn * h;
    float f;
    if (q <= 0.5f) {
      f = (kG * q) * (3.0f * q - 2.0f);
    } else {
      float fa = 1.0f - q;
      f = kG * ((-fa) * fa);
    }
    g.x = f * (rx / den);
    g.y = f * (ry / den);
    g.z = f * (rz / den);
  }
  return g;
}

// Word of list position k of particle i in nbr (and of its pair scalar in
// spair): groups of four positions, one int4 per particle and group.
__device__ __forceinline__ long nb_word(long P, int i, int k) { return 4 * ((long)(k >> 2) * P + i) + (k & 3); }

// Walk the neighbor list of particle i in list order, B entries at a time
// (B a multiple of 4): the B list words are loaded together (B / 4 int4
// groups), then the B neighbors' data (load(j)), and only then body(j, data)
// runs for each in order.  The loads of a batch are independent, so a
// wavefront has B gathers in flight instead of one; the order of the
// per-neighbor terms (and so every sum) is unchanged.  Lanes past the end of
// their list load their own particle's data and skip the body.
template <int B, typename Load, typename Body>
__device__ __forceinline__ void for_neighbors(const Dev& d, int i, int cnt, Load load, Body body) {
  static_assert(B % 4 == 0, "whole groups of four list positions");
  using T = decltype(load(0));
  const long P = d.P;
  const int4* groups = reinterpret_cast<const int4*>(d.nbr);
  for (int k0 = 0; k0 < cnt; k0 += B) {
    int j[B];
    T v[B];
#pragma unroll
    for (int g = 0; g < B / 4; g++) {
      const int k = k0 + 4 * g;
      const int4 w = k < cnt ? groups[(long)(k >> 2) * P + i] : make_int4(i, i, i, i);
      j[4 * g + 0] = w.x;
      j[4 * g + 1] = k + 1 < cnt ? w.y : i;
      j[4 * g + 2] = k + 2 < cnt ? w.z : i;
      j[4 * g + 3] = k + 3 < cnt ? w.w : i;
    }
#pragma unroll
    for (int u = 0; u < B; u++) v[u] = load(j[u]);
#pragma unroll
    for (int u = 0; u < B; u++)
      if (k0 + u < cnt) body(j[u], v[u]);
  }
}

// Launch a sweep templated on its batch size B of for_neighbors (4, 8 or 16;
// anything else is the default 8).  B changes how many gathers are in flight,
// not the order of any sum.
#define DEPTH_LAUNCH(kern, depth, grid, blk, st, ...)                          \
  do {                                                                        \
    switch (depth) {                                                          \
      case 4: NN_LAUNCH(kern<4>, grid, blk, 0, st, __VA_ARGS__); break;        \
      case 16: NN_LAUNCH(kern<16>, grid, blk, 0, st, __VA_ARGS__); break;      \
      default: NN_LAUNCH(kern<8>, grid, blk, 0, st, __VA_ARGS__); break;       \
    }                                                                         \
  } while (0)

struct PosVel {
  float4 x, v;
};
struct PosSrc {
  float4 x;
  float src, alpha;
};

// Kahan (compensated) summation
struct KSum {
  float s, c;
  __device__ __forceinline__ void add(float y) {
    float t = y - c;
    float u = s + t;
    c = (u - s) - t;
    s = u;
  }
};

// ---------------------------------------------------------------------------
// wavefront reductions (fixed butterfly order: every lane ends with the same
// bits, and the result is the same on every run)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Per-scene FP64 reductions, first stage: every wavefront of a participating
// scene stores its NR partial sums into partial arrays base..base+NR-1 (plain
// stores, no fences).  The second stage is k_scene_finish, one workgroup per
// scene in the next launch, which adds the partials in a fixed order: the
// totals are the same on every run and do not depend on the order in which
// wavefronts ran.
template <int NR>
__device__ __forceinline__ void wave_partials(const Dev& d, int wave, double (&v)[NR], int base = 0) {
#pragma unroll
  for (int r = 0; r < NR; r++) v[r] = wave_sum(v[r]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int r = 0; r < NR; r++) d.partials[(long)(base + r) * 4 * d.W + wave] = v[r];
  }
}

// Four-lane sweeps: lanes 4i..4i+3 share particle slot i, lane q walks list
// positions k = q, q+4, ...; a wavefront covers 16 particles.  `hw` is the
// hardware wavefront index used for the partial sums.
__device__ __forceinline__ bool quad_slot_scene(const Dev& d, int& i, int& q, int& s, int& hw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  hw = uniform(t >> 6);
  i = t >> 2;
  q = t & 3;
  const int w = hw >> 2;
  if (w >= d.W) return false;
  s = uniform(d.wave_scene[w]);
  return s >= 0;
}

// (S0 + S1) + (S2 + S3) over the four lanes of a particle; all four lanes end
// with the same bits.
__device__ __forceinline__ float quad_sum(float v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  return v;
}

// Weiler operator pair scalar s_ij = c_ij F(|r|) / (|r| h), c_ij =
// 10 mu_ij (m / rho_j) / (|r|^2 + 0.01 h^2); 0 where the kernel gradient is 0.
__device__ __forceinline__ float weiler_pair(const SceneDev& sc, float rx, float ry, float rz, float mui, float muj,
                                             float rhoj) {
  const float rn = norm3(rx, ry, rz);
  const float q = rn / sc.h;
  if (!(rn > 1e-5f && q <= 1.0f)) return 0.0f;
  float F;
  if (q <= 0.5f) {
    F = (sc.kG * q) * (3.0f * q - 2.0f);
  } else {
    const float fa = 1.0f - q;
    F = sc.kG * ((-fa) * fa);
  }
  const float den = rn * sc.h;
  const float muij = 0.5f * (mui + muj);
  const float c = 10.0f * muij * (sc.mass / rhoj) / ((rx * rx + ry * ry + rz * rz) + sc.visc_eps);
  return c * F / den;
}

// ---------------------------------------------------------------------------
// Polar decomposition A = R S with det R = +1 (the factor ti.polar_decompose
// returns through its SVD, SPH/main.py:1273).  Cyclic Jacobi eigen-decomposition
// of A^T A (Golub & Van Loan, Matrix Computations, Alg. 8.5.3), columns sorted by
// decreasing eigenvalue with det V kept +1, then a Givens QR of A V (the
// structure of McAdams et al. 2011); R = U V^T.  A is row-major a[3*r + c].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void jacobi_rot(float* S, float* V, int p, int q) {
  float apq = S[3 * p + q];
  if (apq == 0.0f) return;
  float theta = (S[3 * q + q] - S[3 * p + p]) / (2.0f * apq);
  float tt = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
  float c = 1.0f / sqrtf(tt * tt + 1.0f);
  float sn = tt * c;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float skp = S[3 * k + p], skq = S[3 * k + q];
    S[3 * k + p] = c * skp - sn * skq;
    S[3 * k + q] = sn * skp + c * skq;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float spk = S[3 * p + k], sqk = S[3 * q + k];
    S[3 * p + k] = c * spk - sn * sqk;
    S[3 * q + k] = sn * spk + c * sqk;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float vkp = V[3 * k + p], vkq = V[3 * k + q];
    V[3 * k + p] = c * vkp - sn * vkq;
    V[3 * k + q] = sn * vkp + c * vkq;
  }
}

__device__ __forceinline__ void givens_rows(float* B, float* U, int p, int q, int col) {
  float x = B[3 * p + col], y = B[3 * q + col];
  float rn = sqrtf(x * x + y * y);
  float c = 1.0f, sn = 0.0f;
  if (rn > 0.0f) {
    c = x / rn;
    sn = y / rn;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float bp = B[3 * p + k], bq = B[3 * q + k];
    B[3 * p + k] = c * bp + sn * bq;
    B[3 * q + k] = -sn * bp + c * bq;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float up = U[3 * k + p], uq = U[3 * k + q];
    U[3 * k + p] = c * up + sn * uq;
    U[3 * k + q] = -sn * up + c * uq;
  }
}

__device__ __forceinline__ void polar3(const float* A, float* R) {
  float S[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) S[3 * a + b] = A[a] * A[b] + A[3 + a] * A[3 + b] + A[6 + a] * A[6 + b];
  float V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 6; sweep++) {
    jacobi_rot(S, V, 0, 1);
    jacobi_rot(S, V, 0, 2);
    jacobi_rot(S, V, 1, 2);
  }
  float ev[3] = {S[0], S[4], S[8]};
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2 - a; b++)
      if (ev[b] < ev[b + 1]) {
        float t = ev[b];
        ev[b] = ev[b + 1];
        ev[b + 1] = t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          float x = V[3 * k + b];
          V[3 * k + b] = V[3 * k + b + 1];
          V[3 * k + b + 1] = -x;
        }
      }
  float B[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) B[3 * a + b] = A[3 * a] * V[b] + A[3 * a + 1] * V[3 + b] + A[3 * a + 2] * V[6 + b];
  float U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  givens_rows(B, U, 0, 1, 0);
  givens_rows(B, U, 0, 2, 0);
  givens_rows(B, U, 1, 2, 1);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) R[3 * a + b] = U[3 * a] * V[3 * b] + U[3 * a + 1] * V[3 * b + 1] + U[3 * a + 2] * V[3 * b + 2];
}

// NonNewton viscosity models, SPH/main.py:404-459, clamped to [0, max].
__device__ __forceinline__ float nn_model(const SceneDev& c, float gam) {
  float eta;
  const float n = c.nn_n, K = c.nn_K, e0 = c.nn_e0, einf = c.nn_einf, gc = c.nn_gc;
  switch (c.nn_method) {
    case 1:
      eta = K * powf(gam, n - 1.0f);
      break;
    case 2:
      eta = einf + (e0 - einf) / (1.0f + powf(K * gam, n));
      break;
    case 3: {
      float t = sqrtf(c.nn_muc) + sqrtf(c.nn_tauy / gam);
      eta = t * t;
      break;
    }
    case 4:
      eta = einf + (e0 - einf) / (1.0f + powf(K * gam * gam, (1.0f - n) / 2.0f));
      break;
    case 5:
      if (gam < gc) eta = e0;
      else eta = gc * (e0 - einf) / gam + einf;
      break;
    case 6:
      if (gam < gc) eta = e0;
      else {
        float tau0 = e0 * gc - K * powf(gc, n);
        eta = tau0 / gam + K * powf(gam, n - 1.0f);
      }
      break;
    default:
      eta = e0;
      break;
  }
  const float mx = c.nn_max;
  if (!(eta >= 0.0f)) eta = (eta != eta) ? mx : 0.0f;
  if (eta > mx) eta = mx;
  return eta;
}

}  // namespace nn
