// kernels_rigid.hip -- rigid bodies by shape matching (RigidBodySolver,
// SPH/main.py:1099-1158), one substep of body rank r of every scene at a time:
//   ground      : every solid particle of the scene below y = padding is lifted
//                 to it (ground_collision, :1117-1122)
//   sum 0       : per rigid-set wave, FP64 partials of m x and m, m = volume0 *
//                 density0 (compute_com, :1120-1130; a solid keeps its initial
//                 density, :744)
//   finish 0    : per body, cm = sum(m x) / sum(m), rounded once
//   sum 1       : partials of A = sum m (x - cm)(x0 - cm0)^T (:1141-1146)
//   finish 1    : per body, R = polar(A), identity if every |R_ab| < 1e-6
//   place       : x += (cm + R (x0 - cm0) - x) * 1.0 (:1153-1156)
// The reference runs the bodies one after another, each coupling_interval
// times (:929-931); the host launches that sequence for rank r = 0, 1, ... with
// every scene's r-th body in the same launches.  Sums are FP64 partials per
// wavefront, added per body in a fixed order (the reference's FP32 atomics
// have no defined order).
#include "device_common.h"

namespace nn {

constexpr int RB_THREADS = 256;

// the body of rigid-set wave w when it is rank r and substep c runs for its scene
__device__ __forceinline__ int rigid_body_of(const Dev& d, int w, int r, int c) {
  const int b = uniform(d.r_wave_body[w]);
  if (b < 0) return -1;
  const BodyDev& B = d.bodies[b];
  if (B.rank != r) return -1;
  return c < d.scenes[B.scene].coupling ? b : -1;
}

__global__ void k_rigid_ground(Dev d, int r, int c) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (!(sc.flags & SF_HAS_RIGID) || r >= sc.nbody || c >= sc.coupling) return;
  if (i >= sc.p_begin + sc.n || kind_of(d.info[i]) == KIND_FLUID) return;
  float4 p = d.pos[i];
  if (p.y < sc.h) {  // padding = support radius (:615)
    p.y = sc.h;
    d.pos[i] = p;
  }
}

// pass 0: (m x, m y, m z, m); pass 1: the 9 entries of m (x - cm)(x0 - cm0)^T
__global__ void k_rigid_sum(Dev d, int pass, int r, int c) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = q >> 6;
  if (w >= d.NRn / 64) return;  // waves in use (NRn is a multiple of 64)
  const int b = rigid_body_of(d, w, r, c);
  if (b < 0) return;  // uniform over the wavefront
  const BodyDev& B = d.bodies[b];
  const SceneDev& sc = d.scenes[B.scene];
  const int id = d.r_pid[q];
  double v[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (id >= 0) {
    const float4 x = d.pos[sc.p_begin + d.inv[sc.p_begin + id]];
    const float m = sc.vol0 * sc.rho0;
    if (pass == 0) {
      v[0] = (double)(m * x.x);
      v[1] = (double)(m * x.y);
      v[2] = (double)(m * x.z);
      v[3] = (double)m;
    } else {
      const float4 x0 = d.x0[sc.p_begin + id];
      const float qv[3] = {x0.x - B.rest_cm[0], x0.y - B.rest_cm[1], x0.z - B.rest_cm[2]};
      const float pv[3] = {x.x - B.cm[0], x.y - B.cm[1], x.z - B.cm[2]};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int e = 0; e < 3; e++) v[3 * a + e] = (double)(m * (pv[a] * qv[e]));
    }
  }
  const int nv = pass == 0 ? 4 : 9;
  const long RW = d.NR / 64;
  for (int k = 0; k < nv; k++) {
    const double t = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) d.rpartials[k * RW + w] = t;
  }
}

// One workgroup per body: its waves' partials in a fixed order (lane t adds
// waves t, t + 256, ...; then the wave butterfly; then a pairwise tree over
// the 4 wave sums).
__global__ __launch_bounds__(RB_THREADS) void k_body_finish(Dev d, int pass, int r, int c) {
  __shared__ double red[9][RB_THREADS / 64];
  const int b = blockIdx.x;
  BodyDev& B = d.bodies[b];
  if (B.rank != r) return;  // uniform over the workgroup (entries of retired scenes have rank -1)
  if (c >= d.scenes[B.scene].coupling) return;
  const int t = threadIdx.x;
  const int nv = pass == 0 ? 4 : 9;
  const long RW = d.NR / 64;
  for (int k = 0; k < nv; k++) {
    double a = 0.0;
    for (int w = B.wave_begin + t; w < B.wave_end; w += RB_THREADS) a += d.rpartials[k * RW + w];
    const double s = wave_sum(a);
    if ((t & 63) == 0) red[k][t >> 6] = s;
  }
  __syncthreads();
  if (t != 0) return;
  double tot[9];
  for (int k = 0; k < nv; k++) tot[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
  if (pass == 0) {
#pragma unroll
    for (int a = 0; a < 3; a++) B.cm[a] = (float)(tot[a] / tot[3]);
  } else {
    float A[9], R[9];
#pragma unroll
    for (int k = 0; k < 9; k++) A[k] = (float)tot[k];
    polar3(A, R);
    bool tiny = true;
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (!(fabsf(R[k]) < 1e-6f)) tiny = false;
#pragma unroll
    for (int k = 0; k < 9; k++) B.R[k] = tiny ? ((k % 4) == 0 ? 1.0f : 0.0f) : R[k];
  }
}

__global__ void k_rigid_place(Dev d, int r, int c) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = q >> 6;
  if (w >= d.NRn / 64) return;  // waves in use (NRn is a multiple of 64)
  const int b = rigid_body_of(d, w, r, c);
  if (b < 0) return;
  const int id = d.r_pid[q];
  if (id < 0) return;
  const BodyDev& B = d.bodies[b];
  const SceneDev& sc = d.scenes[B.scene];
  const int i = sc.p_begin + d.inv[sc.p_begin + id];
  const float4 x0 = d.x0[sc.p_begin + id];
  const float qv[3] = {x0.x - B.rest_cm[0], x0.y - B.rest_cm[1], x0.z - B.rest_cm[2]};
  float4 x = d.pos[i];
  float xs[3] = {x.x, x.y, x.z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float goal = B.cm[a] + (B.R[3 * a] * qv[0] + B.R[3 * a + 1] * qv[1] + B.R[3 * a + 2] * qv[2]);
    const float corr = (goal - xs[a]) * 1.0f;
    xs[a] += corr;
  }
  x.x = xs[0];
  x.y = xs[1];
  x.z = xs[2];
  d.pos[i] = x;
}

static inline dim3 grid_of(const DevRef& d, int block) { return dim3((d.h.Pn + block - 1) / block); }
static inline dim3 rgrid(const DevRef& d) { return dim3((d.h.NRn + RB_THREADS - 1) / RB_THREADS); }

void launch_rigid_ground(const DevRef& d, int r, int c, hipStream_t st) {
  if (d.h.Pn > 0) hipLaunchKernelGGL(k_rigid_ground, grid_of(d, 256), dim3(256), 0, st, d.h, r, c);
}
void launch_rigid_sum(const DevRef& d, int pass, int r, int c, hipStream_t st) {
  if (d.h.NRn > 0) hipLaunchKernelGGL(k_rigid_sum, rgrid(d), dim3(RB_THREADS), 0, st, d.h, pass, r, c);
}
void launch_body_finish(const DevRef& d, int pass, int r, int c, hipStream_t st) {
  if (d.h.nbodies > 0) hipLaunchKernelGGL(k_body_finish, dim3(d.h.nbodies), dim3(RB_THREADS), 0, st, d.h, pass, r, c);
}
void launch_rigid_place(const DevRef& d, int r, int c, hipStream_t st) {
  if (d.h.NRn > 0) hipLaunchKernelGGL(k_rigid_place, rgrid(d), dim3(RB_THREADS), 0, st, d.h, r, c);
}

}  // namespace nn
