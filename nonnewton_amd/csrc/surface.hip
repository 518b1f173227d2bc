// surface.hip -- particle surface reconstruction on the GPU: the path
// SRTool.dual_contouring runs (SPH/surf_recon/srtool.py:1264-1304), behind the
// nn_surface_* entry points of include/nnsph.h.
//
// Stages (one or more launches each, all on the object's stream):
//   bounds      voxel range of the particles (integer min/max atomics)
//   hash        per-voxel particle lists, <= 64 each, sorted by particle index
//               (update_particles / add_particle, :1193-1196, :820-823)
//   aniso       x_bar, C_i, G_i per particle (pre_process_data, :1007-1112)
//   field       the blocks the particles' boxes write are marked; then one
//               workgroup per 8^3 vertex tile sums, per vertex, the poly6 values
//               (generate_fluid_field, :1239-1251) and the anisotropic values
//               (generate_smooth_fluid_field_impl, :1136-1157) of the particles
//               whose boxes reach it
//   stats       FP64 mean and sd of the negative values (density_field_stats, :1437-1456)
//   smooth      bilateral 27-point pass(es) (smooth_sdf_impl, :1459-1484)
//   contour     cubes with a crossing numbered by a scan, their dual-contouring
//               points (dual_contouring_impl, :1532-1596) and triangle counts,
//               a scan, the triangles (dual_contouring_polygen, :1708-1763)
//   normals     per mesh vertex (process_normal, :1384-1410)
//
// Layout in HBM: a dense vertex grid over the 32^3 blocks that can be written
// (vertex (x, y, z) at origin + (x, y, z) * voxel_size, z fastest), one byte per
// block for "written" (the reference's sparse leaf blocks: its loops visit
// exactly those, and an unwritten vertex reads 0).  The field is accumulated in
// 2^-32 fixed point (int64 atomics), so its value does not depend on the order
// in which the particles' contributions are added (a per-vertex gather here, a
// per-particle scatter in the reference), then rounded once to FP32.  The
// reference adds FP32 with atomics in no defined order.  Every other
// expression is FP32 in the reference's operation order (-ffp-contract=off).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "batch.h"

namespace nn {
namespace sr {

constexpr int BLOCK = 32;
constexpr int MAX_PER_VOXEL = 64;  // max_num_particle_per_voxel, srtool.py:685
constexpr double FIX_SCALE = 4294967296.0;
constexpr double FIX_INV = 1.0 / 4294967296.0;
constexpr float MAX_CONTRIB = 1048576.0f;  // |contribution| bound keeping the int64 sums exact
constexpr int SCAN_TILE = 1024;

enum { ERR_OUT_OF_GRID = 1, ERR_VOXEL_OVERFLOW = 2, ERR_RANGE = 4 };

struct K {
  float o[3], vs, sr, iso, vol, acoef, c48, c16, lam, lam1, kr, ks, kn, c2pi;
  int R, off, aniso, n;
  int vlo[3], vdim[3];
  int glo[3], gdim[3], bdim[3];
  long nvox, nvert, nblock;
  const float* pos;  // 3 per particle, caller order
  const float* vel;  // 3 per particle when the velocity channels are on, else null
  unsigned long long* vacc;  // 3 per vertex (velocity channels)
  float *vfield, *mvel;      // 3 per vertex, 3 per mesh vertex
  float cvs3;                // voxel_size^3 (contribution_coeff, :1165)
  int* pvox;         // 3 per particle: its voxel
  int *vcount, *vstart, *vslot, *vlist;
  float *G, *xbar, *detG;
  unsigned char* active;
  unsigned long long* acc;
  float *field0, *field, *tmp;
  int *flag, *cube_id, *tcount, *tstart;
  float *verts, *normals;
  int* faces;
  int* err;
  int* bounds;     // min xyz, max xyz
  double* stats;   // sum, count, mean (as float), sd (as float)
};

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ float norm3f(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
// a value every lane of the wavefront holds (lets the compiler use scalar loads)
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// exp for the bilateral weights (x <= 0 there).  Cody-Waite reduction
// x = k ln2 + r with ln2 split into a 16-bit high part (k ln2_hi exact) and a
// low part, the degree-7 Taylor polynomial of e^r in Horner form, exact
// scaling by 2^k; below 2^-125 the result is 0.
__device__ __forceinline__ float sr_expf(float x) {
  if (x != x) return x;
  if (x < -87.0f) return 0.0f;
  if (x > 88.0f) return INFINITY;
  const float k = rintf(x * 1.44269504f);
  const float r = (x - k * 0.693145751953125f) - k * 1.428606765330187e-06f;
  float p = 1.98412698e-4f;
  p = p * r + 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 1.66666667e-1f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexpf(p, (int)k);
}

// cyclic Jacobi eigen-decomposition of a symmetric 3x3 (Golub & Van Loan,
// Alg. 8.5.2), 6 sweeps; ev = the diagonal, V = the accumulated rotations
__device__ __forceinline__ void jrot(float* S, float* V, int p, int q) {
  const float apq = S[3 * p + q];
  if (apq == 0.0f) return;
  const float theta = (S[3 * q + q] - S[3 * p + p]) / (2.0f * apq);
  const float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
  const float c = 1.0f / sqrtf(t * t + 1.0f);
  const float s = t * c;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float a = S[3 * k + p], b = S[3 * k + q];
    S[3 * k + p] = c * a - s * b;
    S[3 * k + q] = s * a + c * b;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float a = S[3 * p + k], b = S[3 * q + k];
    S[3 * p + k] = c * a - s * b;
    S[3 * q + k] = s * a + c * b;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float a = V[3 * k + p], b = V[3 * k + q];
    V[3 * k + p] = c * a - s * b;
    V[3 * k + q] = s * a + c * b;
  }
}
__device__ __forceinline__ void eig3(const float* A, float* ev, float* V) {
  float S[9];
#pragma unroll
  for (int k = 0; k < 9; k++) S[k] = A[k];
#pragma unroll
  for (int k = 0; k < 9; k++) V[k] = (k % 4) == 0 ? 1.0f : 0.0f;
  for (int sweep = 0; sweep < 6; sweep++) {
    jrot(S, V, 0, 1);
    jrot(S, V, 0, 2);
    jrot(S, V, 1, 2);
  }
  ev[0] = S[0];
  ev[1] = S[4];
  ev[2] = S[8];
}

__device__ __forceinline__ long gidx(const K& k, int gx, int gy, int gz) {
  return ((long)gx * k.gdim[1] + gy) * k.gdim[2] + gz;
}
__device__ __forceinline__ long bidx(const K& k, int gx, int gy, int gz) {
  return ((long)(gx / BLOCK) * k.bdim[1] + gy / BLOCK) * k.bdim[2] + gz / BLOCK;
}
// value at vertex (x, y, z) (absolute vertex coordinates); 0 if unwritten
__device__ __forceinline__ float val_at(const K& k, const float* f, int x, int y, int z) {
  const int gx = x - k.glo[0], gy = y - k.glo[1], gz = z - k.glo[2];
  if (gx < 0 || gy < 0 || gz < 0 || gx >= k.gdim[0] || gy >= k.gdim[1] || gz >= k.gdim[2]) return 0.0f;
  if (!k.active[bidx(k, gx, gy, gz)]) return 0.0f;
  return f[gidx(k, gx, gy, gz)];
}
// get_mapped_position (:795-796)
__device__ __forceinline__ void vpos(const K& k, int x, int y, int z, float* p) {
  p[0] = (float)x * k.vs + k.o[0];
  p[1] = (float)y * k.vs + k.o[1];
  p[2] = (float)z * k.vs + k.o[2];
}
// get_voxel_index (:799-803): truncation toward zero
__device__ __forceinline__ void voxel_of(const K& k, const float* p, int* v) {
  v[0] = (int)((p[0] - k.o[0]) / k.vs);
  v[1] = (int)((p[1] - k.o[1]) / k.vs);
  v[2] = (int)((p[2] - k.o[2]) / k.vs);
}
// poly6_weight (:1213-1227)
__device__ __forceinline__ float poly6(const K& k, float rx, float ry, float rz, float h) {
  const float rn = norm3f(rx, ry, rz);
  float res = 0.0f;
  if (rn <= h) {
    float tmp = h * h;
    res = tmp - rn * rn;
    res = res * res * res;
    tmp = tmp * tmp * tmp * tmp * h;
    res = k.acoef * res / tmp;
  }
  return res;
}
// cubic_spline_kernel (:1115-1124), "+ 1" where the reference puts it
__device__ __forceinline__ float aniso_w(const K& k, const float* G, float det, float dx, float dy, float dz) {
  const float a = G[0] * dx + G[1] * dy + G[2] * dz;
  const float b = G[3] * dx + G[4] * dy + G[5] * dz;
  const float c = G[6] * dx + G[7] * dy + G[8] * dz;
  const float q = norm3f(a, b, c);
  float res = 0.0f;
  if (0.0f < q && q <= 0.5f)
    res = k.c48 * det * (q * q * q - q * q) + 1.0f;
  else if (0.5f < q && q <= 1.0f)
    res = k.c16 * det * (1.0f - q) * (1.0f - q) * (1.0f - q);
  return res;
}
__device__ __forceinline__ unsigned long long fix_of(float v) {
  return (unsigned long long)(long long)rint((double)v * FIX_SCALE);
}
__device__ __forceinline__ int mult_axis(int v, int c, int off, int R) {
  const int lo = c - off, hi = c - off + R - 1;
  return (v >= lo && v <= hi) + (v - 1 >= lo && v - 1 <= hi);
}

// ------------------------------------------------------------------ scan
// exclusive prefix sum of in[0, L) into out[0, L), total into out[L];
// tiles of 1024 scanned in LDS, tile sums scanned recursively, added back

__global__ __launch_bounds__(SCAN_TILE) void k_scan_tile(const int* in, int* out, int* tsum, long L) {
  __shared__ int s[SCAN_TILE];
  const long i = (long)blockIdx.x * SCAN_TILE + threadIdx.x;
  const int v = i < L ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int d = 1; d < SCAN_TILE; d <<= 1) {
    const int t = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < L) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == SCAN_TILE - 1) tsum[blockIdx.x] = s[SCAN_TILE - 1];
}
__global__ void k_scan_add(int* out, const int* toff, long L) {
  const long i = (long)blockIdx.x * SCAN_TILE + threadIdx.x;
  if (i < L) out[i] += toff[blockIdx.x];
}
__global__ void k_scan_total(const int* in, int* out, long L) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[L] = L > 0 ? out[L - 1] + in[L - 1] : 0;
}

// ------------------------------------------------------------------ stages

__global__ void k_fill_i32(int* p, int v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u8(unsigned char* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// voxel range of the particles: a workgroup min/max in LDS, then one atomic
// per bound per workgroup (integers: the order does not matter)
constexpr int BOUNDS_THREADS = 256;
__global__ __launch_bounds__(BOUNDS_THREADS) void k_bounds(K k) {
  __shared__ int red[6][BOUNDS_THREADS];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = threadIdx.x;
  int v[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, u[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  if (i < k.n) {
    const float* p = k.pos + 3 * (long)i;
    bool ok = true;
    for (int a = 0; a < 3; a++) {
      const float tt = (p[a] - k.o[a]) / k.vs;
      if (!(tt > -1.0f && tt < 1.0e8f)) ok = false;
    }
    if (ok) {
      int w[3];
      voxel_of(k, p, w);
      for (int a = 0; a < 3; a++)
        if (w[a] < 0) ok = false;
      if (ok)
        for (int a = 0; a < 3; a++) v[a] = u[a] = w[a];
    }
    if (!ok) atomicOr(k.err, ERR_OUT_OF_GRID);
  }
  for (int a = 0; a < 3; a++) {
    red[a][t] = v[a];
    red[3 + a][t] = u[a];
  }
  __syncthreads();
  for (int span = BOUNDS_THREADS / 2; span > 0; span >>= 1) {
    if (t < span)
      for (int a = 0; a < 3; a++) {
        red[a][t] = min(red[a][t], red[a][t + span]);
        red[3 + a][t] = max(red[3 + a][t], red[3 + a][t + span]);
      }
    __syncthreads();
  }
  if (t < 3 && red[t][0] != INT32_MAX) {
    atomicMin(k.bounds + t, red[t][0]);
    atomicMax(k.bounds + 3 + t, red[3 + t][0]);
  }
}

__device__ __forceinline__ long vkey(const K& k, const int* v) {
  return ((long)(v[0] - k.vlo[0]) * k.vdim[1] + (v[1] - k.vlo[1])) * k.vdim[2] + (v[2] - k.vlo[2]);
}

__global__ void k_vcount(K k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  int v[3];
  voxel_of(k, k.pos + 3 * (long)i, v);
  k.pvox[3 * (long)i] = v[0];
  k.pvox[3 * (long)i + 1] = v[1];
  k.pvox[3 * (long)i + 2] = v[2];
  const int slot = atomicAdd(k.vcount + vkey(k, v), 1);
  if (slot >= MAX_PER_VOXEL) atomicOr(k.err, ERR_VOXEL_OVERFLOW);
  k.vslot[i] = slot;
}

__global__ void k_vplace(K k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  int v[3];
  voxel_of(k, k.pos + 3 * (long)i, v);
  k.vlist[k.vstart[vkey(k, v)] + k.vslot[i]] = i;
}

// a voxel's list in increasing particle index (insertion sort, <= 64 entries)
__global__ void k_vsort(K k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvox) return;
  const int b = k.vstart[q], e = k.vstart[q + 1];
  for (int m = b + 1; m < e; m++) {
    const int x = k.vlist[m];
    int j = m - 1;
    while (j >= b && k.vlist[j] > x) {
      k.vlist[j + 1] = k.vlist[j];
      j--;
    }
    k.vlist[j + 1] = x;
  }
}

__device__ __forceinline__ void vrange(const K& k, int x, int y, int z, int& b, int& e) {
  const int a = x - k.vlo[0], c = y - k.vlo[1], d = z - k.vlo[2];
  if (a < 0 || c < 0 || d < 0 || a >= k.vdim[0] || c >= k.vdim[1] || d >= k.vdim[2]) {
    b = e = 0;
    return;
  }
  const long q = ((long)a * k.vdim[1] + c) * k.vdim[2] + d;
  b = k.vstart[q];
  e = k.vstart[q + 1];
}

// The particles a sweep of pre_process_data visits, in its order (voxel dx,
// dy, dz nested, z fastest, then the voxel's list), restricted to the voxels
// that can hold a particle closer than r2: a voxel's particles are contiguous
// in vlist and the voxels of one (x, y) column are consecutive, so each
// column is one range.  A skipped particle has w = 0 and adds +0 to every sum
// (rows and column ends are trimmed with a margin of 1e-3 voxel around each
// voxel and 1e-4 of r2), so the sums are those of the full loop bit for bit.
template <typename F>
__device__ __forceinline__ void for_reach(const K& k, const int* v, float x0, float x1, float x2, float r2, F f) {
  const int R = k.R, off = k.off;
  const float reach = r2 * 1.0001f;
  const float reach2 = reach * reach;
  const float m = 1e-3f * k.vs;
  for (int dx = 0; dx < R; dx++) {
    const int x = v[0] + dx - off;
    if (x < 0 || x < k.vlo[0] || x >= k.vlo[0] + k.vdim[0]) continue;
    const float lx = (float)x * k.vs + k.o[0] - m, hx = lx + k.vs + 2.0f * m;
    const float gx = fmaxf(0.0f, fmaxf(lx - x0, x0 - hx));
    for (int dy = 0; dy < R; dy++) {
      const int y = v[1] + dy - off;
      if (y < 0 || y < k.vlo[1] || y >= k.vlo[1] + k.vdim[1]) continue;
      const float ly = (float)y * k.vs + k.o[1] - m, hy = ly + k.vs + 2.0f * m;
      const float gy = fmaxf(0.0f, fmaxf(ly - x1, x1 - hy));
      const float rest = reach2 - (gx * gx + gy * gy);
      if (!(rest > 0.0f)) continue;
      const float zr = sqrtf(rest);
      int z0 = (int)floorf((x2 - zr - k.o[2]) / k.vs) - 1;
      int z1 = (int)floorf((x2 + zr - k.o[2]) / k.vs) + 1;
      z0 = max(max(z0, v[2] - off), max(0, k.vlo[2]));
      z1 = min(min(z1, v[2] - off + R - 1), k.vlo[2] + k.vdim[2] - 1);
      if (z0 > z1) continue;
      const long base = ((long)(x - k.vlo[0]) * k.vdim[1] + (y - k.vlo[1])) * k.vdim[2] - k.vlo[2];
      const int b = k.vstart[base + z0], e = k.vstart[base + z1 + 1];
      for (int mm = b; mm < e; mm++) f(k.vlist[mm]);
    }
  }
}

// pre_process_data (:1007-1112) for particle i >= 1 (particle 0 keeps G = 0,
// x_bar = 0: the reference's loop starts at 1 and reset() zeroes both)
__global__ void k_aniso(K k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  float* G = k.G + 9 * (long)i;
  float* xb = k.xbar + 3 * (long)i;
  if (i == 0) {
    for (int a = 0; a < 9; a++) G[a] = 0.0f;
    for (int a = 0; a < 3; a++) xb[a] = 0.0f;
    k.detG[0] = 0.0f;
    return;
  }
  const float* xi = k.pos + 3 * (long)i;
  const float x0 = xi[0], x1 = xi[1], x2 = xi[2];
  int v[3];
  voxel_of(k, xi, v);
  const float r2 = 2.0f * k.sr;
  float tw = 0.0f, w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
  for_reach(k, v, x0, x1, x2, r2, [&](int j) {
    const float* xj = k.pos + 3 * (long)j;
    const float y0 = xj[0], y1 = xj[1], y2 = xj[2];
    const float nr = norm3f(x0 - y0, x1 - y1, x2 - y2);
    float w = 0.0f;
    if (nr < r2) {
      const float t = nr / r2;
      w = 1.0f - t * t * t;
    }
    tw += w;
    w0 += w * y0;
    w1 += w * y1;
    w2 += w * y2;
  });
  if (tw != 0.0f) {
    w0 /= tw;
    w1 /= tw;
    w2 /= tw;
  }
  xb[0] = k.lam1 * x0 + k.lam * w0;
  xb[1] = k.lam1 * x1 + k.lam * w1;
  xb[2] = k.lam1 * x2 + k.lam * w2;
  int nn = 0;
  float C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for_reach(k, v, x0, x1, x2, r2, [&](int j) {
    const float* xj = k.pos + 3 * (long)j;
    const float y0 = xj[0], y1 = xj[1], y2 = xj[2];
    const float nr = norm3f(x0 - y0, x1 - y1, x2 - y2);
    float w = 0.0f;
    if (nr < r2) {
      const float t = nr / r2;
      w = 1.0f - t * t * t;
    }
    if (w > 0.0f) {
      nn++;
      const float dp[3] = {y0 - w0, y1 - w1, y2 - w2};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) C[3 * a + c] += w * (dp[a] * dp[c]);
    }
  });
  if (tw != 0.0f)
#pragma unroll
    for (int a = 0; a < 9; a++) C[a] /= tw;
  float ev[3], V[9];
  eig3(C, ev, V);
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2 - a; b++)
      if (ev[b] < ev[b + 1]) {
        const float t = ev[b];
        ev[b] = ev[b + 1];
        ev[b + 1] = t;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const float x = V[3 * c + b];
          V[3 * c + b] = V[3 * c + b + 1];
          V[3 * c + b + 1] = x;
        }
      }
  float g[9];
  if (tw == 0.0f) {
#pragma unroll
    for (int a = 0; a < 9; a++) g[a] = (a % 4) == 0 ? 1.0f : 0.0f;
  } else {
    float sg[3];
    if (nn > 25) {  // Nvep
      sg[0] = ev[0];
      sg[1] = fmaxf(ev[1], ev[0] / k.kr);
      sg[2] = fmaxf(ev[2], ev[0] / k.kr);
#pragma unroll
      for (int a = 0; a < 3; a++) sg[a] *= k.ks;
    } else {
      sg[0] = sg[1] = sg[2] = k.kn;
    }
    const float ih = 1.0f / k.sr;
    float P[9];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int c = 0; c < 3; c++) P[3 * a + c] = (ih * V[3 * a + c]) * (1.0f / sg[c]);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) g[3 * a + b] = P[3 * a] * V[3 * b] + P[3 * a + 1] * V[3 * b + 1] + P[3 * a + 2] * V[3 * b + 2];
  }
#pragma unroll
  for (int a = 0; a < 9; a++) G[a] = g[a];
  k.detG[i] = g[0] * (g[4] * g[8] - g[7] * g[5]) - g[3] * (g[1] * g[8] - g[7] * g[2]) + g[6] * (g[1] * g[5] - g[4] * g[2]);
}

// The blocks the two field passes write: every vertex >= 0 of the particle's
// 27-vertex poly6 box and of its anisotropic box [c - off, c - off + R]^3
// (the reference writes them all, zeros included).  Each box spans at most two
// blocks per axis (R + 1 <= 32 is checked on the host), so its 8 corners name
// every block it touches.
__global__ void k_mark(K k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  const int* v = k.pvox + 3 * (long)i;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = v[a] - 1;
    hi[a] = v[a] + 1;
    if (k.aniso) {
      lo[a] = min(lo[a], v[a] - k.off);
      hi[a] = max(hi[a], v[a] - k.off + k.R);
    }
    lo[a] = max(lo[a], 0);
  }
  for (int c = 0; c < 8; c++) {
    const int x = (c >> 2) ? hi[0] : lo[0], y = ((c >> 1) & 1) ? hi[1] : lo[1], z = (c & 1) ? hi[2] : lo[2];
    k.active[bidx(k, x - k.glo[0], y - k.glo[1], z - k.glo[2])] = 1;
  }
}

// The field at the vertices of one 8^3 tile, one thread per vertex, summed
// over every particle whose boxes can reach the tile: the particles of voxels
// c in [T - R + off, T + 7 + off] per axis (T = the tile's first vertex).  The
// workgroup lists the candidate voxel columns and their particle counts (a
// scan in LDS), then stages the candidates' data in LDS 512 at a time (each
// thread finds its candidate's column by bisection), and every thread adds
// what each staged particle gives its vertex: the poly6 value when
// |v - c_j| <= 1 per axis, the anisotropic value times the multiplicity of the
// reference's voxel-corner loop when v is in j's box.  The contributions are
// 2^-32 fixed-point integers, so the sum is the one any order of the
// reference's atomics would give exactly.  Tiles of unwritten blocks store 0.
constexpr int TILE = 8;
constexpr int GATHER_THREADS = TILE * TILE * TILE;
constexpr int MAX_COLS = 1536;  // (TILE + R)^2 columns, R + 1 <= 32
__global__ __launch_bounds__(GATHER_THREADS) void k_field(K k) {
  __shared__ int col_off[MAX_COLS + 1];
  __shared__ int col_begin[MAX_COLS];
  __shared__ float s_pos[3][GATHER_THREADS], s_xb[3][GATHER_THREADS], s_g[9][GATHER_THREADS], s_det[GATHER_THREADS];
  __shared__ int s_c[3][GATHER_THREADS];
  __shared__ float s_vel[3][GATHER_THREADS];
  const int tb = blockIdx.x;
  const int tdy = k.gdim[1] / TILE, tdz = k.gdim[2] / TILE;
  const int tx = tb / (tdy * tdz), ty = (tb / tdz) % tdy, tz = tb % tdz;
  const int t = threadIdx.x;
  const int gx = tx * TILE + (t >> 6), gy = ty * TILE + ((t >> 3) & 7), gz = tz * TILE + (t & 7);
  const long q = gidx(k, gx, gy, gz);
  if (!k.active[bidx(k, tx * TILE, ty * TILE, tz * TILE)]) {  // a tile lies in one block
    k.acc[q] = 0ull;
    if (k.vel)
      for (int c = 0; c < 3; c++) k.vacc[3 * q + c] = 0ull;
    return;
  }
  const int vx = gx + k.glo[0], vy = gy + k.glo[1], vz = gz + k.glo[2];
  float p[3];
  vpos(k, vx, vy, vz, p);
  const int R = k.R, off = k.off;
  const int T[3] = {tx * TILE + k.glo[0], ty * TILE + k.glo[1], tz * TILE + k.glo[2]};
  unsigned long long va[3] = {0ull, 0ull, 0ull};
  int clo[3], chi[3];
  for (int a = 0; a < 3; a++) {
    clo[a] = k.aniso ? T[a] - R + off : T[a] - 1;
    chi[a] = k.aniso ? T[a] + TILE - 1 + off : T[a] + TILE;
    clo[a] = max(clo[a], k.vlo[a]);
    chi[a] = min(chi[a], k.vlo[a] + k.vdim[a] - 1);
  }
  const int ncx = max(0, chi[0] - clo[0] + 1), ncy = max(0, chi[1] - clo[1] + 1);
  const int ncol = clo[2] <= chi[2] ? ncx * ncy : 0;  // <= MAX_COLS (checked on the host)
  // columns: particle counts, then an inclusive scan (Hillis-Steele over
  // GATHER_THREADS-wide rounds with a carry)
  int carry = 0;
  for (int c0 = 0; c0 < ncol; c0 += GATHER_THREADS) {
    const int c = c0 + t;
    int cnt = 0;
    if (c < ncol) {
      const int cx = clo[0] + c / ncy, cy = clo[1] + c % ncy;
      const long base = ((long)(cx - k.vlo[0]) * k.vdim[1] + (cy - k.vlo[1])) * k.vdim[2] - k.vlo[2];
      const int b = k.vstart[base + clo[2]], e = k.vstart[base + chi[2] + 1];
      col_begin[c] = b;
      cnt = e - b;
    }
    __shared__ int scan[GATHER_THREADS];
    scan[t] = cnt;
    __syncthreads();
    for (int d = 1; d < GATHER_THREADS; d <<= 1) {
      const int u = t >= d ? scan[t - d] : 0;
      __syncthreads();
      scan[t] += u;
      __syncthreads();
    }
    if (c < ncol) col_off[c + 1] = carry + scan[t];
    if (c0 == 0 && t == 0) col_off[0] = 0;
    carry += scan[GATHER_THREADS - 1];
    __syncthreads();
  }
  const int total = ncol > 0 ? carry : 0;
  unsigned long long acc = 0ull;
  bool bad = false;
  for (int s0 = 0; s0 < total; s0 += GATHER_THREADS) {
    const int idx = s0 + t;
    if (idx < total) {
      int lo = 0, hi = ncol - 1;  // the column c with col_off[c] <= idx < col_off[c + 1]
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (col_off[mid] <= idx) lo = mid;
        else hi = mid - 1;
      }
      const int j = k.vlist[col_begin[lo] + (idx - col_off[lo])];
      s_pos[0][t] = k.pos[3 * (long)j];
      s_pos[1][t] = k.pos[3 * (long)j + 1];
      s_pos[2][t] = k.pos[3 * (long)j + 2];
      s_c[0][t] = k.pvox[3 * (long)j];
      s_c[1][t] = k.pvox[3 * (long)j + 1];
      s_c[2][t] = k.pvox[3 * (long)j + 2];
      if (k.aniso) {
        s_xb[0][t] = k.xbar[3 * (long)j];
        s_xb[1][t] = k.xbar[3 * (long)j + 1];
        s_xb[2][t] = k.xbar[3 * (long)j + 2];
#pragma unroll
        for (int a = 0; a < 9; a++) s_g[a][t] = k.G[9 * (long)j + a];
        s_det[t] = k.detG[j];
        if (k.vel)
          for (int a = 0; a < 3; a++) s_vel[a][t] = k.vel[3 * (long)j + a];
      }
    }
    __syncthreads();
    const int nst = min(GATHER_THREADS, total - s0);
#pragma unroll 4
    for (int r = 0; r < nst; r++) {
      const int c0 = s_c[0][r], c1 = s_c[1][r], c2 = s_c[2][r];
      if (abs(vx - c0) <= 1 && abs(vy - c1) <= 1 && abs(vz - c2) <= 1) {
        const float w = k.vol * poly6(k, s_pos[0][r] - p[0], s_pos[1][r] - p[1], s_pos[2][r] - p[2], k.sr);
        if (!(fabsf(w) <= MAX_CONTRIB)) bad = true;
        else acc += fix_of(w);
      }
      if (!k.aniso) continue;
      const int mm = mult_axis(vx, c0, off, R) * mult_axis(vy, c1, off, R) * mult_axis(vz, c2, off, R);
      if (mm == 0) continue;
      float g[9];
#pragma unroll
      for (int a = 0; a < 9; a++) g[a] = s_g[a][r];
      const float w = k.vol * aniso_w(k, g, s_det[r], s_xb[0][r] - p[0], s_xb[1][r] - p[1], s_xb[2][r] - p[2]);
      if (!(fabsf(w) <= MAX_CONTRIB)) bad = true;
      else acc += (unsigned long long)mm * fix_of(w);
      if (k.vel) {  // generate_smooth_fluid_field_with_velocity_impl (:1160-1190)
        const float f = poly6(k, s_pos[0][r] - p[0], s_pos[1][r] - p[1], s_pos[2][r] - p[2], k.sr) * k.cvs3;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const float u = s_vel[c][r] * f;
          if (!(fabsf(u) <= MAX_CONTRIB)) bad = true;
          else va[c] += (unsigned long long)mm * fix_of(u);
        }
      }
    }
    __syncthreads();
  }
  k.acc[q] = acc;
  if (k.vel)
    for (int c = 0; c < 3; c++) k.vacc[3 * q + c] = va[c];
  if (bad) atomicOr(k.err, ERR_RANGE);
}

__global__ void k_convert(K k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvert) return;
  const int gz = (int)(q % k.gdim[2]), gy = (int)((q / k.gdim[2]) % k.gdim[1]), gx = (int)(q / ((long)k.gdim[1] * k.gdim[2]));
  float v = 0.0f;
  if (k.active[bidx(k, gx, gy, gz)]) v = (float)((double)(long long)k.acc[q] * FIX_INV);
  k.field0[q] = v;
  k.field[q] = v;
  if (k.vel) {
    const bool act = k.active[bidx(k, gx, gy, gz)];
    for (int c = 0; c < 3; c++) k.vfield[3 * q + c] = act ? (float)((double)(long long)k.vacc[3 * q + c] * FIX_INV) : 0.0f;
  }
}

// Blocked FP64 statistic of the negative values, one workgroup per block in
// block order (inactive blocks add 0): lane t adds vertices l = t + 256 j of
// the block (l = (lx * 32 + ly) * 32 + lz) in order, then a pairwise tree.
// which 0: sum of v, count; which 1: sum of (v - mean)^2 (v - mean in FP32).
constexpr int ST_THREADS = 256;
__global__ __launch_bounds__(ST_THREADS) void k_stats(K k, int which, double* part, long long* pcount) {
  __shared__ double lane[ST_THREADS];
  __shared__ int cnt;
  const long b = blockIdx.x;
  const int t = threadIdx.x;
  if (t == 0) cnt = 0;
  __syncthreads();
  if (!k.active[b]) {
    if (t == 0) {
      part[b] = 0.0;
      pcount[b] = 0;
    }
    return;
  }
  const int bz = (int)(b % k.bdim[2]), by = (int)((b / k.bdim[2]) % k.bdim[1]), bx = (int)(b / ((long)k.bdim[1] * k.bdim[2]));
  const float mean = (float)k.stats[2];
  double a = 0.0;
  int c = 0;
  for (int l = t; l < BLOCK * BLOCK * BLOCK; l += ST_THREADS) {
    const int gx = bx * BLOCK + (l >> 10), gy = by * BLOCK + ((l >> 5) & 31), gz = bz * BLOCK + (l & 31);
    const float v = k.field[gidx(k, gx, gy, gz)];
    if (v < 0.0f) {
      if (which == 0) {
        a += (double)v;
        c++;
      } else {
        const double d = (double)(v - mean);
        a += d * d;
      }
    }
  }
  lane[t] = a;
  if (c) atomicAdd(&cnt, c);
  __syncthreads();
  for (int span = ST_THREADS / 2; span > 0; span >>= 1) {
    if (t < span) lane[t] += lane[t + span];
    __syncthreads();
  }
  if (t == 0) {
    part[b] = lane[0];
    pcount[b] = cnt;
  }
}

// blocks added in block order; which 0 sets the mean, which 1 the sd
__global__ void k_stats_final(K k, int which, const double* part, const long long* pcount) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  long long c = 0;
  for (long b = 0; b < k.nblock; b++) {
    s += part[b];
    c += pcount[b];
  }
  if (which == 0) {
    k.stats[0] = s;
    k.stats[1] = (double)c;
    k.stats[2] = c > 0 ? (double)(float)(s / (double)c) : 0.0;
    if (c == 0) k.stats[3] = (double)NAN;
  } else if (k.stats[1] > 0.0) {
    k.stats[3] = (double)(float)sqrt(s / k.stats[1]);
  }
}

// smooth_sdf_impl (:1459-1484), radius 1
__global__ void k_smooth(K k, const float* in, float* out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvert) return;
  const int gz = (int)(q % k.gdim[2]), gy = (int)((q / k.gdim[2]) % k.gdim[1]), gx = (int)(q / ((long)k.gdim[1] * k.gdim[2]));
  if (!k.active[bidx(k, gx, gy, gz)]) {
    out[q] = 0.0f;
    return;
  }
  const float sd = (float)k.stats[3];
  const int i = gx + k.glo[0], j = gy + k.glo[1], l = gz + k.glo[2];
  const float vc = in[q];
  float avg = 0.0f, weight = 0.0f;
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        const int x = i + dx, y = j + dy, z = l + dz;
        if (x < 0 || y < 0 || z < 0) continue;
        const float vn = val_at(k, in, x, y, z);
        const float d = fabsf(vc - vn);
        const float tt = k.c2pi * sd * sd;
        const float gw = sr_expf(-d * d / (2.0f * sd * sd)) / (tt * sqrtf(tt));
        avg += gw * vn;
        weight += gw;
      }
  if (weight != 0.0f) avg /= weight;
  out[q] = avg;
}

// corner offsets (voxel_vertex_offset_map, :772-779) and edge corner pairs
// (interpolation_table, :925-936)
__constant__ int CORNER[8][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};
__constant__ int EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// cubes sharing edges 2, 3, 11 of a cube (get_edge_adjacent_cubes, :1613-1705)
__constant__ int ADJ_W[3] = {2, 3, 11};
__constant__ int ADJ[3][3][3] = {
    {{0, 0, -1}, {0, -1, -1}, {0, -1, 0}},
    {{0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}},
    {{0, -1, 0}, {-1, -1, 0}, {-1, 0, 0}},
};

// cube index bits (value < isolevel) << 12 | crossed-edge mask (the
// marching-cubes edge table, derived: an edge is crossed when its corners differ)
__device__ __forceinline__ int cube_setup(const K& k, int x, int y, int z, float* cv, float (*cp)[3]) {
  int ci = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    cv[j] = val_at(k, k.field, x + CORNER[j][0], y + CORNER[j][1], z + CORNER[j][2]);
    vpos(k, x + CORNER[j][0], y + CORNER[j][1], z + CORNER[j][2], cp[j]);
    if (cv[j] < k.iso) ci |= 1 << j;
  }
  int mask = 0;
#pragma unroll
  for (int w = 0; w < 12; w++)
    if (((ci >> EDGE[w][0]) & 1) != ((ci >> EDGE[w][1]) & 1)) mask |= 1 << w;
  return (ci << 12) | mask;
}

__global__ void k_cube_flag(K k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvert) return;
  const int gz = (int)(q % k.gdim[2]), gy = (int)((q / k.gdim[2]) % k.gdim[1]), gx = (int)(q / ((long)k.gdim[1] * k.gdim[2]));
  int f = 0;
  if (k.active[bidx(k, gx, gy, gz)]) {
    float cv[8], cp[8][3];
    f = (cube_setup(k, gx + k.glo[0], gy + k.glo[1], gz + k.glo[2], cv, cp) & 0xFFF) != 0;
  }
  k.flag[q] = f;
}

__global__ void k_cube_id(K k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvert) return;
  k.cube_id[q] = k.flag[q] ? k.cube_id[q] + 1 : 0;  // cube_id holds the exclusive scan of flag
}

// calc_vertex_normal (:1307-1321) with normalize (:1413-1417)
__device__ __forceinline__ void vertex_normal(const K& k, int x, int y, int z, float* n) {
  const float* f = k.field;
  n[0] = x == 0 ? -val_at(k, f, x + 1, y, z) : val_at(k, f, x - 1, y, z) - val_at(k, f, x + 1, y, z);
  n[1] = y == 0 ? -val_at(k, f, x, y + 1, z) : val_at(k, f, x, y - 1, z) - val_at(k, f, x, y + 1, z);
  n[2] = z == 0 ? -val_at(k, f, x, y, z + 1) : val_at(k, f, x, y, z - 1) - val_at(k, f, x, y, z + 1);
  const float nr = norm3f(n[0], n[1], n[2]);
  if (nr != 0.0f) {
    n[0] /= nr;
    n[1] /= nr;
    n[2] /= nr;
  } else {
    n[0] = n[1] = n[2] = 0.0f;
  }
}
// process_normal_at (:1336-1349)
__device__ __forceinline__ void normal_at(const K& k, const float* p, float* out) {
  int v[3];
  voxel_of(k, p, v);
  float base[3];
  vpos(k, v[0], v[1], v[2], base);
  const float xd = (p[0] - base[0]) / k.vs, yd = (p[1] - base[1]) / k.vs, zd = (p[2] - base[2]) / k.vs;
  float c[8][3];
#pragma unroll
  for (int j = 0; j < 8; j++) vertex_normal(k, v[0] + (j >> 2), v[1] + ((j >> 1) & 1), v[2] + (j & 1), c[j]);
  float r[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float c00 = c[0][a] * (1.0f - xd) + c[4][a] * xd;
    const float c01 = c[1][a] * (1.0f - xd) + c[5][a] * xd;
    const float c10 = c[2][a] * (1.0f - xd) + c[6][a] * xd;
    const float c11 = c[3][a] * (1.0f - xd) + c[7][a] * xd;
    const float c0 = c00 * (1.0f - yd) + c10 * yd;
    const float c1 = c01 * (1.0f - yd) + c11 * yd;
    r[a] = c0 * (1.0f - zd) + c1 * zd;
  }
  const float nr = norm3f(r[0], r[1], r[2]);
  if (nr != 0.0f) {
    out[0] = r[0] / nr;
    out[1] = r[1] / nr;
    out[2] = r[2] / nr;
  } else {
    out[0] = out[1] = out[2] = 0.0f;
  }
}
// vertex_interpolate (:1254-1262)
__device__ __forceinline__ void vinterp(float iso, const float* p1, const float* p2, float v1, float v2, float* out) {
  const float delta = v2 - v1;
  const float off = fabsf(delta) < 0.00001f ? 0.5f : (iso - v1) / delta;
#pragma unroll
  for (int a = 0; a < 3; a++) out[a] = p1[a] + off * (p2[a] - p1[a]);
}

// the QEF: rows folded into a 4x4 triangle by Givens rotations (incremental QR,
// Ju et al. 2002; the reference's Householder form holds 4 rows), the 3x3 block
// solved through its SVD with singular values clamped to sign(s) max(0.1, |s|)
// (solve_qef, :70-80), x = V diag(1 / s') U^T b; singular values up to 1e-6 of
// the largest count as 0 (the null space, where U is arbitrary, is dropped)
__device__ __forceinline__ void qef_add(float* R, float a0, float a1, float a2, float a3) {
  float a[4] = {a0, a1, a2, a3};
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const float xk = a[kk];
    if (xk == 0.0f) continue;
    const float rk = R[5 * kk];
    const float rho = sqrtf(rk * rk + xk * xk);
    const float c = rk / rho, sn = xk / rho;
#pragma unroll
    for (int j = kk; j < 4; j++) {
      const float u = R[4 * kk + j], w = a[j];
      R[4 * kk + j] = c * u + sn * w;
      a[j] = -sn * u + c * w;
    }
    a[kk] = 0.0f;
  }
}
__device__ __forceinline__ void givens_rows3(float* B, float* U, int p, int q, int col) {
  const float x = B[3 * p + col], y = B[3 * q + col];
  const float rn = sqrtf(x * x + y * y);
  float c = 1.0f, sn = 0.0f;
  if (rn > 0.0f) {
    c = x / rn;
    sn = y / rn;
  }
#pragma unroll
  for (int kk = 0; kk < 3; kk++) {
    const float bp = B[3 * p + kk], bq = B[3 * q + kk];
    B[3 * p + kk] = c * bp + sn * bq;
    B[3 * q + kk] = -sn * bp + c * bq;
  }
#pragma unroll
  for (int kk = 0; kk < 3; kk++) {
    const float up = U[3 * kk + p], uq = U[3 * kk + q];
    U[3 * kk + p] = c * up + sn * uq;
    U[3 * kk + q] = -sn * up + c * uq;
  }
}
// A = U diag(s) V^T: V = Jacobi eigenvectors of A^T A by decreasing
// eigenvalue, then a Givens QR of A V = U B, s = diag(B) (McAdams et al. 2011)
__device__ __forceinline__ void svd3(const float* A, float* U, float* sg, float* V) {
  float M[9], ev[3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) M[3 * a + b] = A[a] * A[b] + A[3 + a] * A[3 + b] + A[6 + a] * A[6 + b];
  eig3(M, ev, V);
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2 - a; b++)
      if (ev[b] < ev[b + 1]) {
        const float t = ev[b];
        ev[b] = ev[b + 1];
        ev[b + 1] = t;
#pragma unroll
        for (int kk = 0; kk < 3; kk++) {
          const float x = V[3 * kk + b];
          V[3 * kk + b] = V[3 * kk + b + 1];
          V[3 * kk + b + 1] = x;
        }
      }
  float B[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) B[3 * a + b] = A[3 * a] * V[b] + A[3 * a + 1] * V[3 + b] + A[3 * a + 2] * V[6 + b];
#pragma unroll
  for (int kk = 0; kk < 9; kk++) U[kk] = (kk % 4) == 0 ? 1.0f : 0.0f;
  givens_rows3(B, U, 0, 1, 0);
  givens_rows3(B, U, 0, 2, 0);
  givens_rows3(B, U, 1, 2, 1);
  sg[0] = B[0];
  sg[1] = B[4];
  sg[2] = B[8];
}
__device__ __forceinline__ void qef_solve(const float* R, float* x) {
  float A[9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) A[3 * a + b] = R[4 * a + b];
  float Us[9], sg[3], V[9];
  svd3(A, Us, sg, V);
  const float bb[3] = {R[3], R[7], R[11]};
  const float tiny = 1e-6f * fmaxf(fabsf(sg[0]), fmaxf(fabsf(sg[1]), fabsf(sg[2])));
  float z[3];
#pragma unroll
  for (int kk = 0; kk < 3; kk++) {
    const float bk = Us[kk] * bb[0] + Us[3 + kk] * bb[1] + Us[6 + kk] * bb[2];
    const float sc = (sg[kk] < 0.0f ? -1.0f : 1.0f) * fmaxf(0.1f, fabsf(sg[kk]));
    z[kk] = fabsf(sg[kk]) > tiny ? bk / sc : 0.0f;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) x[a] = V[3 * a] * z[0] + V[3 * a + 1] * z[1] + V[3 * a + 2] * z[2];
}

__device__ __forceinline__ int cube_vertex(const K& k, int x, int y, int z) {
  const int gx = x - k.glo[0], gy = y - k.glo[1], gz = z - k.glo[2];
  if (gx < 0 || gy < 0 || gz < 0 || gx >= k.gdim[0] || gy >= k.gdim[1] || gz >= k.gdim[2]) return 0;
  return k.cube_id[gidx(k, gx, gy, gz)];
}

// dual_contouring_impl (:1532-1596) for the cubes with a crossing, and the
// number of triangles each emits (dual_contouring_polygen, :1708-1763)
__global__ void k_dc_point(K k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvert) return;
  const int id = k.cube_id[q];
  if (!id) {
    k.tcount[q] = 0;
    return;
  }
  const int gz = (int)(q % k.gdim[2]), gy = (int)((q / k.gdim[2]) % k.gdim[1]), gx = (int)(q / ((long)k.gdim[1] * k.gdim[2]));
  const int x = gx + k.glo[0], y = gy + k.glo[1], z = gz + k.glo[2];
  float cv[8], cp[8][3];
  const int cm = cube_setup(k, x, y, z, cv, cp);
  const int mask = cm & 0xFFF;
  float mean[3] = {0.0f, 0.0f, 0.0f};
  int cnt = 0;
  for (int w = 0; w < 12; w++) {
    if (!(mask & (1 << w))) continue;
    float p[3];
    vinterp(k.iso, cp[EDGE[w][0]], cp[EDGE[w][1]], cv[EDGE[w][0]], cv[EDGE[w][1]], p);
    mean[0] += p[0];
    mean[1] += p[1];
    mean[2] += p[2];
    cnt++;
  }
  const float fc = (float)cnt;
  mean[0] /= fc;
  mean[1] /= fc;
  mean[2] /= fc;
  float R[16];
#pragma unroll
  for (int a = 0; a < 16; a++) R[a] = 0.0f;
  for (int w = 0; w < 12; w++) {
    if (!(mask & (1 << w))) continue;
    float p[3], n[3];
    vinterp(k.iso, cp[EDGE[w][0]], cp[EDGE[w][1]], cv[EDGE[w][0]], cv[EDGE[w][1]], p);
    normal_at(k, p, n);
    const float d0 = p[0] - mean[0], d1 = p[1] - mean[1], d2 = p[2] - mean[2];
    qef_add(R, n[0], n[1], n[2], n[0] * d0 + n[1] * d1 + n[2] * d2);
  }
  float r[3];
  qef_solve(R, r);
  const float t[3] = {r[0] + mean[0], r[1] + mean[1], r[2] + mean[2]};
  int v[3];
  voxel_of(k, t, v);
  float* out = k.verts + 3 * (long)(id - 1);
  if (v[0] != x || v[1] != y || v[2] != z) {
    out[0] = mean[0];
    out[1] = mean[1];
    out[2] = mean[2];
  } else {
    out[0] = r[0] + mean[0];
    out[1] = r[1] + mean[1];
    out[2] = r[2] + mean[2];
  }
  int nt = 0;
  for (int e = 0; e < 3; e++) {
    if (!(mask & (1 << ADJ_W[e]))) continue;
    if (cube_vertex(k, x + ADJ[e][0][0], y + ADJ[e][0][1], z + ADJ[e][0][2]) &&
        cube_vertex(k, x + ADJ[e][1][0], y + ADJ[e][1][1], z + ADJ[e][1][2]) &&
        cube_vertex(k, x + ADJ[e][2][0], y + ADJ[e][2][1], z + ADJ[e][2][2]))
      nt += 2;
  }
  k.tcount[q] = nt;
}

__global__ void k_faces(K k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= k.nvert) return;
  const int id = k.cube_id[q];
  if (!id) return;
  const int gz = (int)(q % k.gdim[2]), gy = (int)((q / k.gdim[2]) % k.gdim[1]), gx = (int)(q / ((long)k.gdim[1] * k.gdim[2]));
  const int x = gx + k.glo[0], y = gy + k.glo[1], z = gz + k.glo[2];
  float cv[8], cp[8][3];
  const int cm = cube_setup(k, x, y, z, cv, cp);
  const int mask = cm & 0xFFF, ci = cm >> 12;
  int* f = k.faces + 3 * (long)k.tstart[q];
  for (int e = 0; e < 3; e++) {
    const int w = ADJ_W[e];
    if (!(mask & (1 << w))) continue;
    const int i1 = id;
    const int i2 = cube_vertex(k, x + ADJ[e][0][0], y + ADJ[e][0][1], z + ADJ[e][0][2]);
    const int i3 = cube_vertex(k, x + ADJ[e][1][0], y + ADJ[e][1][1], z + ADJ[e][1][2]);
    const int i4 = cube_vertex(k, x + ADJ[e][2][0], y + ADJ[e][2][1], z + ADJ[e][2][2]);
    if (!(i2 && i3 && i4)) continue;
    const int flip = ((ci >> EDGE[w][0]) & 1) != (w == 11);
    if (!flip) {
      f[0] = i3 - 1; f[1] = i2 - 1; f[2] = i1 - 1;
      f[3] = i4 - 1; f[4] = i3 - 1; f[5] = i1 - 1;
    } else {
      f[0] = i1 - 1; f[1] = i2 - 1; f[2] = i3 - 1;
      f[3] = i1 - 1; f[4] = i3 - 1; f[5] = i4 - 1;
    }
    f += 6;
  }
}

// get_velocity_at (:1352-1359): 0 outside written blocks
__device__ __forceinline__ void vel_at(const K& k, int x, int y, int z, float* out) {
  const int gx = x - k.glo[0], gy = y - k.glo[1], gz = z - k.glo[2];
  if (gx < 0 || gy < 0 || gz < 0 || gx >= k.gdim[0] || gy >= k.gdim[1] || gz >= k.gdim[2] ||
      !k.active[bidx(k, gx, gy, gz)]) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const long q = gidx(k, gx, gy, gz);
  for (int c = 0; c < 3; c++) out[c] = k.vfield[3 * q + c];
}

// per mesh vertex: the normal (process_normal, :1384-1410) and the velocity
// (process_velocity_at with trilinear_interpolate, :1362-1375, :1324-1333)
__global__ void k_vertex_attrs(K k, long nv, int want_normals, int want_vel) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float p[3] = {k.verts[3 * v], k.verts[3 * v + 1], k.verts[3 * v + 2]};
  if (want_normals) normal_at(k, p, k.normals + 3 * v);
  if (!want_vel) return;
  float* out = k.mvel + 3 * v;
  if (!k.vel) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  int c3[3];
  voxel_of(k, p, c3);
  float base[3];
  vpos(k, c3[0], c3[1], c3[2], base);
  const float xd = (p[0] - base[0]) / k.vs, yd = (p[1] - base[1]) / k.vs, zd = (p[2] - base[2]) / k.vs;
  float c[8][3];
#pragma unroll
  for (int j = 0; j < 8; j++) vel_at(k, c3[0] + (j >> 2), c3[1] + ((j >> 1) & 1), c3[2] + (j & 1), c[j]);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float c00 = c[0][a] * (1.0f - xd) + c[4][a] * xd;
    const float c01 = c[1][a] * (1.0f - xd) + c[5][a] * xd;
    const float c10 = c[2][a] * (1.0f - xd) + c[6][a] * xd;
    const float c11 = c[3][a] * (1.0f - xd) + c[7][a] * xd;
    const float c0 = c00 * (1.0f - yd) + c10 * yd;
    const float c1 = c01 * (1.0f - yd) + c11 * yd;
    out[a] = c0 * (1.0f - zd) + c1 * zd;
  }
}

// ------------------------------------------------------------------ host

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n) {
    if (n <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t c = n > 0 ? n : 1;
    hipError_t e = hipMalloc(&p, c * sizeof(T));
    if (e == hipSuccess) cap = c;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Surface {
  nn_surface_params prm{};
  K k{};
  hipStream_t st = nullptr;
  DBuf<float> pos, vel, G, xbar, detG, field0, field, tmp, verts, normals, vfield, mvel;
  DBuf<int> vcount, vstart, vslot, vlist, pvox, flag, cube_id, tcount, tstart, faces, err, bounds, scratch;
  DBuf<unsigned char> active;
  DBuf<unsigned long long> acc, vacc;
  DBuf<double> stats, part;
  DBuf<long long> pcount;
  long nverts = 0, ntris = 0, nactive = 0;
  double sd = 0.0;
  bool have = false;
  int64_t budget = 0;  // device bytes a run may use (0 = no limit)

  ~Surface() {
    for (auto* b : {&pos, &vel, &G, &xbar, &detG, &field0, &field, &tmp, &verts, &normals, &vfield, &mvel}) b->release();
    for (auto* b : {&vcount, &vstart, &vslot, &vlist, &pvox, &flag, &cube_id, &tcount, &tstart, &faces, &err, &bounds, &scratch})
      b->release();
    active.release();
    acc.release();
    vacc.release();
    stats.release();
    part.release();
    pcount.release();
    if (st) (void)hipStreamDestroy(st);
  }
};

static inline unsigned grid1(long n, int b) { return (unsigned)((n + b - 1) / b); }

// exclusive scan of in[0, L) into out[0, L], total in out[L]
static hipError_t scan(Surface& S, const int* in, int* out, long L) {
  if (L <= 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S.st, out, 0, 1L);
    return hipGetLastError();
  }
  // levels: tile sums of each level kept in scratch
  std::vector<long> len;
  std::vector<long> offs;
  long l = L, total = 0;
  while (true) {
    const long tiles = (l + SCAN_TILE - 1) / SCAN_TILE;
    len.push_back(l);
    offs.push_back(total);
    total += tiles + 1;
    if (tiles == 1) break;
    l = tiles;
  }
  hipError_t e = S.scratch.ensure((size_t)total * 2);
  if (e != hipSuccess) return e;
  int* base = S.scratch.p;
  // level 0: in -> out; level j: tile sums of level j - 1 (at base + offs[j - 1]
  // in the first half) scanned into the second half
  const int* lin = in;
  int* lout = out;
  for (size_t j = 0; j < len.size(); j++) {
    const long lj = len[j];
    const long tiles = (lj + SCAN_TILE - 1) / SCAN_TILE;
    int* tsum = base + offs[j];
    hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)tiles), dim3(SCAN_TILE), 0, S.st, lin, lout, tsum, lj);
    lin = tsum;
    lout = base + total + offs[j];
  }
  // add back from the top: level j's output gets the scanned tile sums of level j
  for (size_t j = len.size() - 1; j-- > 0;) {
    const long lj = len[j];
    const long tiles = (lj + SCAN_TILE - 1) / SCAN_TILE;
    int* tgt = j == 0 ? out : base + total + offs[j - 1];
    const int* toff = base + total + offs[j];
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_TILE), 0, S.st, tgt, toff, lj);
  }
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, S.st, in, out, L);
  return hipGetLastError();
}

static int fail(hipError_t e, const char* what) { return hip_fail(e, what); }

#define SR_HIP(expr, what)                      \
  do {                                          \
    hipError_t e__ = (expr);                    \
    if (e__ != hipSuccess) return fail(e__, what); \
  } while (0)

static int device_error(Surface& S) {
  int err = 0;
  SR_HIP(hipMemcpyAsync(&err, S.err.p, sizeof(int), hipMemcpyDeviceToHost, S.st), "nn_surface_run: error flag");
  SR_HIP(hipStreamSynchronize(S.st), "nn_surface_run");
  if (err & ERR_OUT_OF_GRID) {
    set_error("nn_surface_run: a particle lies below the bounding box's first voxel, or its position is not finite");
    return NN_E_OUT_OF_GRID;
  }
  if (err & ERR_VOXEL_OVERFLOW) {
    set_error("nn_surface_run: more than 64 particles in one voxel (max_num_particle_per_voxel)");
    return NN_E_CELL_OVERFLOW;
  }
  if (err & ERR_RANGE) {
    set_error("nn_surface_run: a field contribution is not finite or exceeds 2^20");
    return NN_E_NAN;
  }
  return NN_OK;
}

static int run(Surface& S, const float* dpos, const float* dvel, long n) {
  K& k = S.k;
  S.have = false;
  k.n = (int)n;
  k.pos = dpos;
  k.vel = (S.prm.velocity && k.aniso) ? dvel : nullptr;  // the reference samples velocity in the smooth field only
  SR_HIP(S.err.ensure(1), "nn_surface_run: alloc");
  SR_HIP(S.bounds.ensure(6), "nn_surface_run: alloc");
  SR_HIP(S.stats.ensure(4), "nn_surface_run: alloc");
  k.err = S.err.p;
  k.bounds = S.bounds.p;
  k.stats = S.stats.p;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S.st, k.err, 0, 1L);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S.st, k.bounds, INT32_MAX, 3L);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S.st, k.bounds + 3, INT32_MIN, 3L);
  hipLaunchKernelGGL(k_bounds, dim3(grid1(n, BOUNDS_THREADS)), dim3(BOUNDS_THREADS), 0, S.st, k);
  int hb[6];
  SR_HIP(hipMemcpyAsync(hb, k.bounds, sizeof(hb), hipMemcpyDeviceToHost, S.st), "nn_surface_run: bounds");
  int rc = device_error(S);
  if (rc) return rc;
  // voxel range and the vertex grid (blocks that the field passes can write)
  long nvox = 1;
  for (int a = 0; a < 3; a++) {
    k.vlo[a] = hb[a];
    k.vdim[a] = hb[3 + a] - hb[a] + 1;
    nvox *= k.vdim[a];
  }
  long nvert = 1, nblock = 1;
  for (int a = 0; a < 3; a++) {
    const int vmin = k.vlo[a], vmax = k.vlo[a] + k.vdim[a] - 1;
    long lo = vmin - 1, hi = (long)vmax + 1;
    if (k.aniso) {
      lo = std::min<long>(lo, vmin - k.off);
      hi = std::max<long>(hi, (long)vmax - k.off + k.R);
    }
    if (lo < 0) lo = 0;
    const long glo = (lo / BLOCK) * BLOCK, ghi = (hi / BLOCK) * BLOCK + BLOCK - 1;
    k.glo[a] = (int)glo;
    k.gdim[a] = (int)(ghi - glo + 1);
    k.bdim[a] = k.gdim[a] / BLOCK;
    nvert *= k.gdim[a];
    nblock *= k.bdim[a];
  }
  k.nvox = nvox;
  k.nvert = nvert;
  k.nblock = nblock;
  const double need = (double)nvox * 8 + (double)nvert * (k.vel ? 72 : 36) + (double)n * 64 + (double)nblock * 20;
  // int32 scans: up to 6 triangles per cube must stay below 2^31
  if (nvert > (1L << 28) || nvox > (1L << 30) || (S.budget > 0 && need > (double)S.budget)) {
    set_error("nn_surface_run: the vertex grid over the particles' extent exceeds the device budget (" +
              std::to_string(nvert) + " vertices)");
    return NN_E_OOM;
  }
  SR_HIP(S.vcount.ensure(nvox), "nn_surface_run: alloc");
  SR_HIP(S.vstart.ensure(nvox + 1), "nn_surface_run: alloc");
  SR_HIP(S.vslot.ensure(n), "nn_surface_run: alloc");
  SR_HIP(S.vlist.ensure(n), "nn_surface_run: alloc");
  SR_HIP(S.pvox.ensure(3 * n), "nn_surface_run: alloc");
  SR_HIP(S.G.ensure(9 * n), "nn_surface_run: alloc");
  SR_HIP(S.xbar.ensure(3 * n), "nn_surface_run: alloc");
  SR_HIP(S.detG.ensure(n), "nn_surface_run: alloc");
  SR_HIP(S.active.ensure(nblock), "nn_surface_run: alloc");
  SR_HIP(S.acc.ensure(nvert), "nn_surface_run: alloc");
  SR_HIP(S.field0.ensure(nvert), "nn_surface_run: alloc");
  SR_HIP(S.field.ensure(nvert), "nn_surface_run: alloc");
  SR_HIP(S.tmp.ensure(nvert), "nn_surface_run: alloc");
  SR_HIP(S.flag.ensure(nvert), "nn_surface_run: alloc");
  SR_HIP(S.cube_id.ensure(nvert + 1), "nn_surface_run: alloc");
  SR_HIP(S.tcount.ensure(nvert), "nn_surface_run: alloc");
  SR_HIP(S.tstart.ensure(nvert + 1), "nn_surface_run: alloc");
  SR_HIP(S.part.ensure(nblock), "nn_surface_run: alloc");
  SR_HIP(S.pcount.ensure(nblock), "nn_surface_run: alloc");
  if (k.vel) {
    SR_HIP(S.vacc.ensure(3 * nvert), "nn_surface_run: alloc");
    SR_HIP(S.vfield.ensure(3 * nvert), "nn_surface_run: alloc");
    k.vacc = S.vacc.p;
    k.vfield = S.vfield.p;
  }
  k.vcount = S.vcount.p;
  k.vstart = S.vstart.p;
  k.vslot = S.vslot.p;
  k.vlist = S.vlist.p;
  k.pvox = S.pvox.p;
  k.G = S.G.p;
  k.xbar = S.xbar.p;
  k.detG = S.detG.p;
  k.active = S.active.p;
  k.acc = S.acc.p;
  k.field0 = S.field0.p;
  k.field = S.field.p;
  k.tmp = S.tmp.p;
  k.flag = S.flag.p;
  k.cube_id = S.cube_id.p;
  k.tcount = S.tcount.p;
  k.tstart = S.tstart.p;
  // hash
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1(nvox, 256)), dim3(256), 0, S.st, k.vcount, 0, nvox);
  hipLaunchKernelGGL(k_vcount, dim3(grid1(n, 256)), dim3(256), 0, S.st, k);
  SR_HIP(scan(S, k.vcount, k.vstart, nvox), "nn_surface_run: scan");
  hipLaunchKernelGGL(k_vplace, dim3(grid1(n, 256)), dim3(256), 0, S.st, k);
  hipLaunchKernelGGL(k_vsort, dim3(grid1(nvox, 256)), dim3(256), 0, S.st, k);
  if ((rc = device_error(S))) return rc;
  // anisotropic kernel and field
  if (k.aniso) hipLaunchKernelGGL(k_aniso, dim3(grid1(n, 64)), dim3(64), 0, S.st, k);
  hipLaunchKernelGGL(k_fill_u8, dim3(grid1(nblock, 256)), dim3(256), 0, S.st, k.active, nblock);
  hipLaunchKernelGGL(k_mark, dim3(grid1(n, 256)), dim3(256), 0, S.st, k);
  hipLaunchKernelGGL(k_field, dim3((unsigned)(nvert / GATHER_THREADS)), dim3(GATHER_THREADS), 0, S.st, k);
  hipLaunchKernelGGL(k_convert, dim3(grid1(nvert, 256)), dim3(256), 0, S.st, k);
  if ((rc = device_error(S))) return rc;
  // smoothing
  hipLaunchKernelGGL(k_stats, dim3((unsigned)nblock), dim3(ST_THREADS), 0, S.st, k, 0, S.part.p, S.pcount.p);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, S.st, k, 0, S.part.p, S.pcount.p);
  hipLaunchKernelGGL(k_stats, dim3((unsigned)nblock), dim3(ST_THREADS), 0, S.st, k, 1, S.part.p, S.pcount.p);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, S.st, k, 1, S.part.p, S.pcount.p);
  for (int it = 0; it < S.prm.smooth_iterations; it++) {
    hipLaunchKernelGGL(k_smooth, dim3(grid1(nvert, 256)), dim3(256), 0, S.st, k, k.field, k.tmp);
    std::swap(k.field, k.tmp);
  }
  // contour
  hipLaunchKernelGGL(k_cube_flag, dim3(grid1(nvert, 256)), dim3(256), 0, S.st, k);
  SR_HIP(scan(S, k.flag, k.cube_id, nvert), "nn_surface_run: scan");
  int nv = 0;
  SR_HIP(hipMemcpyAsync(&nv, k.cube_id + nvert, sizeof(int), hipMemcpyDeviceToHost, S.st), "nn_surface_run: count");
  hipLaunchKernelGGL(k_cube_id, dim3(grid1(nvert, 256)), dim3(256), 0, S.st, k);
  SR_HIP(hipStreamSynchronize(S.st), "nn_surface_run");
  SR_HIP(S.verts.ensure(3 * (size_t)nv), "nn_surface_run: alloc");
  k.verts = S.verts.p;
  hipLaunchKernelGGL(k_dc_point, dim3(grid1(nvert, 128)), dim3(128), 0, S.st, k);
  SR_HIP(scan(S, k.tcount, k.tstart, nvert), "nn_surface_run: scan");
  int nt = 0;
  SR_HIP(hipMemcpyAsync(&nt, k.tstart + nvert, sizeof(int), hipMemcpyDeviceToHost, S.st), "nn_surface_run: count");
  SR_HIP(hipStreamSynchronize(S.st), "nn_surface_run");
  SR_HIP(S.faces.ensure(3 * (size_t)nt), "nn_surface_run: alloc");
  k.faces = S.faces.p;
  hipLaunchKernelGGL(k_faces, dim3(grid1(nvert, 128)), dim3(128), 0, S.st, k);
  const int want_n = S.prm.normals ? 1 : 0, want_v = S.prm.velocity ? 1 : 0;
  if ((want_n || want_v) && nv > 0) {
    if (want_n) {
      SR_HIP(S.normals.ensure(3 * (size_t)nv), "nn_surface_run: alloc");
      k.normals = S.normals.p;
    }
    if (want_v) {
      SR_HIP(S.mvel.ensure(3 * (size_t)nv), "nn_surface_run: alloc");
      k.mvel = S.mvel.p;
    }
    hipLaunchKernelGGL(k_vertex_attrs, dim3(grid1(nv, 128)), dim3(128), 0, S.st, k, (long)nv, want_n, want_v);
  }
  double st4[4];
  SR_HIP(hipMemcpyAsync(st4, k.stats, sizeof(st4), hipMemcpyDeviceToHost, S.st), "nn_surface_run: stats");
  SR_HIP(hipStreamSynchronize(S.st), "nn_surface_run");
  SR_HIP(hipGetLastError(), "nn_surface_run: launch");
  // active-block count for the caller's statistics
  std::vector<unsigned char> act(nblock);
  SR_HIP(hipMemcpy(act.data(), k.active, nblock, hipMemcpyDeviceToHost), "nn_surface_run: active");
  long na = 0;
  for (unsigned char c : act) na += c;
  S.nactive = na;
  S.nverts = nv;
  S.ntris = nt;
  S.sd = st4[3];
  S.have = true;
  return NN_OK;
}

// a scene's fluid positions, from the batch's cell order into original-index
// order (the reference numbers the fluid first, so ids are 0 .. n_fluid - 1)
__global__ void k_fluid_positions(const float4* pos, const float4* vel, const int* info, int pb, int n, int nf, float* out,
                                  float* vout, int* err) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int inf = info[pb + t];
  if ((inf >> 28) != KIND_FLUID) return;
  const int id = inf & 0x0FFFFFFF;
  if (id >= nf) {
    atomicOr(err, 1);
    return;
  }
  const float4 p = pos[pb + t];
  out[3 * (long)id] = p.x;
  out[3 * (long)id + 1] = p.y;
  out[3 * (long)id + 2] = p.z;
  if (vout) {
    const float4 v = vel[pb + t];
    vout[3 * (long)id] = v.x;
    vout[3 * (long)id + 1] = v.y;
    vout[3 * (long)id + 2] = v.z;
  }
}

}  // namespace sr
}  // namespace nn

using nn::sr::Surface;

struct nn_surface {
  Surface impl;
};

int nn::surface_run_scene(nn_surface* h, Batch& b, int k) {
  Surface& S = h->impl;
  const SceneHost& sc = *b.scenes[k];
  const int nf = sc.n_fluid;
  if (nf <= 0) {
    set_error("nn_surface_run_context: the scene has no fluid particles");
    return NN_E_ARG;
  }
  hipError_t e = hipStreamSynchronize(b.stream);  // the step's positions are complete
  if (e == hipSuccess) e = S.pos.ensure(3 * (size_t)nf);
  if (e == hipSuccess) e = S.err.ensure(1);
  const bool with_vel = S.prm.velocity != 0;
  if (e == hipSuccess && with_vel) e = S.vel.ensure(3 * (size_t)nf);
  if (e != hipSuccess) return hip_fail(e, "nn_surface_run_context");
  hipLaunchKernelGGL(sr::k_fill_i32, dim3(1), dim3(64), 0, S.st, S.err.p, 0, 1L);
  hipLaunchKernelGGL(sr::k_fluid_positions, dim3((sc.n + 255) / 256), dim3(256), 0, S.st, b.d.pos, b.d.vel, b.d.info,
                     sc.p_begin, sc.n, nf, S.pos.p, with_vel ? S.vel.p : nullptr, S.err.p);
  int err = 0;
  e = hipMemcpyAsync(&err, S.err.p, sizeof(int), hipMemcpyDeviceToHost, S.st);
  if (e == hipSuccess) e = hipStreamSynchronize(S.st);
  if (e != hipSuccess) return hip_fail(e, "nn_surface_run_context");
  if (err) {
    set_error("nn_surface_run_context: fluid particles must be numbered 0 .. n_fluid - 1 (the reference numbers them first)");
    return NN_E_UNSUPPORTED;
  }
  return sr::run(S, S.pos.p, with_vel ? S.vel.p : nullptr, nf);
}

extern "C" {

int32_t nn_surface_create(const nn_surface_params* p, nn_surface** out) {
  if (!p || !out) {
    nn::set_error("nn_surface_create: bad arguments");
    return NN_E_ARG;
  }
  *out = nullptr;
  if (!(p->voxel_size > 0.0) || !(p->particle_radius > 0.0) || !(p->smooth_radius > 0.0) ||
      p->smooth_iterations < 0 || p->smooth_iterations > 64) {
    nn::set_error("nn_surface_create: voxel_size, particle_radius and smooth_radius must be positive, "
                  "smooth_iterations in [0, 64]");
    return NN_E_ARG;
  }
  nn_surface* H = new nn_surface();
  Surface* S = &H->impl;
  S->prm = *p;
  nn::sr::K& k = S->k;
  for (int a = 0; a < 3; a++) k.o[a] = (float)p->origin[a];
  k.vs = (float)p->voxel_size;
  const double r = p->particle_radius;
  k.sr = (float)p->smooth_radius;
  k.iso = (float)p->isolevel;
  k.vol = (float)(r * r * r * M_PI * 4 / 3);  // srtool.py:945
  k.acoef = (float)(315 / (64 * M_PI));        // :946
  k.c48 = (float)(48 / M_PI);
  k.c16 = (float)(16 / M_PI);
  k.lam = (float)0.9;  // akLambda, :929
  k.lam1 = (float)(1 - 0.9);
  k.kr = 4.0f;  // kr, ks, kn, :925-928
  k.ks = 1400.0f;
  k.kn = 0.5f;
  k.c2pi = (float)(2 * M_PI);
  k.cvs3 = (float)(p->voxel_size * p->voxel_size * p->voxel_size);
  k.aniso = p->anisotropic ? 1 : 0;
  k.R = (int)ceilf(4.0f * k.sr / k.vs) + 1;  // smooth_range, :1137
  k.off = k.R / 2;
  if (k.aniso && (k.R < 1 || k.R + 1 > nn::sr::BLOCK)) {
    nn::set_error("nn_surface_create: 4 * smooth_radius / voxel_size must be below 30 (the anisotropic box "
                  "spans at most two 32-vertex blocks per axis)");
    delete H;
    return NN_E_UNSUPPORTED;
  }
  hipError_t e = hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete H;
    return nn::hip_fail(e, "nn_surface_create: stream");
  }
  *out = H;
  return NN_OK;
}

int32_t nn_surface_destroy(nn_surface* s) {
  delete s;
  return NN_OK;
}

int32_t nn_surface_set_budget(nn_surface* s, int64_t bytes) {
  if (!s || bytes < 0) {
    nn::set_error("nn_surface_set_budget: bad arguments");
    return NN_E_ARG;
  }
  s->impl.budget = bytes;
  return NN_OK;
}

int32_t nn_surface_run_device(nn_surface* s, const float* dpos, int64_t n) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || !dpos || n <= 0 || n > (1L << 30)) {
    nn::set_error("nn_surface_run_device: bad arguments (n must be in [1, 2^30])");
    return NN_E_ARG;
  }
  return nn::sr::run(*S, dpos, nullptr, (long)n);
}

int32_t nn_surface_run(nn_surface* s, const float* pos, int64_t n) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || !pos || n <= 0 || n > (1L << 30)) {
    nn::set_error("nn_surface_run: bad arguments (n must be in [1, 2^30])");
    return NN_E_ARG;
  }
  hipError_t e = S->pos.ensure(3 * (size_t)n);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_run: alloc");
  e = hipMemcpyAsync(S->pos.p, pos, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, S->st);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_run: upload");
  return nn::sr::run(*S, S->pos.p, nullptr, (long)n);
}

int32_t nn_surface_run_velocity(nn_surface* s, const float* pos, const float* vel, int64_t n) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || !pos || !vel || n <= 0 || n > (1L << 30)) {
    nn::set_error("nn_surface_run_velocity: bad arguments (n must be in [1, 2^30])");
    return NN_E_ARG;
  }
  if (!S->prm.velocity) {
    nn::set_error("nn_surface_run_velocity: the velocity channels were not requested (nn_surface_params.velocity)");
    return NN_E_STATE;
  }
  hipError_t e = S->pos.ensure(3 * (size_t)n);
  if (e == hipSuccess) e = S->vel.ensure(3 * (size_t)n);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_run_velocity: alloc");
  e = hipMemcpyAsync(S->pos.p, pos, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, S->st);
  if (e == hipSuccess) e = hipMemcpyAsync(S->vel.p, vel, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, S->st);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_run_velocity: upload");
  return nn::sr::run(*S, S->pos.p, S->vel.p, (long)n);
}

int32_t nn_surface_mesh_velocity(nn_surface* s, float* out) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || (!out && S->nverts)) {
    nn::set_error("nn_surface_mesh_velocity: bad arguments");
    return NN_E_ARG;
  }
  if (!S->have || !S->prm.velocity) {
    nn::set_error("nn_surface_mesh_velocity: no completed run with the velocity channels");
    return NN_E_STATE;
  }
  if (S->nverts) {
    hipError_t e = hipMemcpy(out, S->k.mvel, sizeof(float) * 3 * S->nverts, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_mesh_velocity");
  }
  return NN_OK;
}

int32_t nn_surface_info(nn_surface* s, int64_t* out, double* sd) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || !out) {
    nn::set_error("nn_surface_info: bad arguments");
    return NN_E_ARG;
  }
  if (!S->have) {
    nn::set_error("nn_surface_info: no completed run");
    return NN_E_STATE;
  }
  const nn::sr::K& k = S->k;
  for (int a = 0; a < 3; a++) {
    out[a] = k.glo[a];
    out[3 + a] = k.gdim[a];
    out[9 + a] = k.vlo[a];
    out[12 + a] = k.vdim[a];
  }
  out[6] = S->nverts;
  out[7] = S->ntris;
  out[8] = S->nactive;
  out[15] = k.n;
  if (sd) *sd = S->sd;
  return NN_OK;
}

int32_t nn_surface_mesh(nn_surface* s, float* vertices, int32_t* faces, float* normals) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || (!vertices && S->nverts) || (!faces && S->ntris)) {
    nn::set_error("nn_surface_mesh: bad arguments");
    return NN_E_ARG;
  }
  if (!S->have) {
    nn::set_error("nn_surface_mesh: no completed run");
    return NN_E_STATE;
  }
  if (normals && !S->prm.normals) {
    nn::set_error("nn_surface_mesh: normals were not requested (nn_surface_params.normals)");
    return NN_E_STATE;
  }
  hipError_t e = hipSuccess;
  if (S->nverts) e = hipMemcpy(vertices, S->k.verts, sizeof(float) * 3 * S->nverts, hipMemcpyDeviceToHost);
  if (e == hipSuccess && S->ntris) e = hipMemcpy(faces, S->k.faces, sizeof(int) * 3 * S->ntris, hipMemcpyDeviceToHost);
  if (e == hipSuccess && normals && S->nverts)
    e = hipMemcpy(normals, S->k.normals, sizeof(float) * 3 * S->nverts, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_mesh");
  return NN_OK;
}

int32_t nn_surface_field(nn_surface* s, int32_t which, float* out, size_t count) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || !out || (which != 0 && which != 1)) {
    nn::set_error("nn_surface_field: bad arguments");
    return NN_E_ARG;
  }
  if (!S->have) {
    nn::set_error("nn_surface_field: no completed run");
    return NN_E_STATE;
  }
  if (count < (size_t)S->k.nvert) {
    nn::set_error("nn_surface_field: count below the grid's vertex count");
    return NN_E_ARG;
  }
  const float* src = which == 0 ? S->k.field0 : S->k.field;
  hipError_t e = hipMemcpy(out, src, sizeof(float) * S->k.nvert, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_field");
  return NN_OK;
}

int32_t nn_surface_aniso(nn_surface* s, float* G, float* xbar, int64_t n) {
  Surface* S = s ? &s->impl : nullptr;
  if (!S || !G || !xbar) {
    nn::set_error("nn_surface_aniso: bad arguments");
    return NN_E_ARG;
  }
  if (!S->have || !S->k.aniso) {
    nn::set_error("nn_surface_aniso: no completed anisotropic run");
    return NN_E_STATE;
  }
  if (n != S->k.n) {
    nn::set_error("nn_surface_aniso: n differs from the last run's particle count");
    return NN_E_ARG;
  }
  hipError_t e = hipMemcpy(G, S->k.G, sizeof(float) * 9 * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(xbar, S->k.xbar, sizeof(float) * 3 * n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return nn::hip_fail(e, "nn_surface_aniso");
  return NN_OK;
}

}  // extern "C"
