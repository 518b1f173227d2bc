// noalloc.hip -- nn_neighbors_noalloc: the neighbor lists of nn_neighbors
// (meta.ns.run_search() + meta.ns.neighbors, SPH/main.py:815-847) for
// positions already in device memory, written to device memory, with no
// device allocation: the search's scratch (cell counts, cell starts, scan
// partials, cell keys, the particles in cell order) lives in one file-scope
// __device__ array per device, at a fixed maximum size.
//
// Order and test are those of the context's search: cell = truncation of x / h
// (pos_to_index, :805-807), the 27 cells visited dx slowest and dz fastest,
// inside a cell increasing particle index (the serial fill order of
// update_grid, :815-820), and |x_i - x_j| < h evaluated as the FP32
// sqrt_rn(s) < h (s < the least float whose rounded square root is >= h).
#include <cmath>
#include <mutex>

#include "batch.h"

namespace nn {

constexpr long NA_MAX_PARTICLES = 1L << 20;
constexpr long NA_MAX_CELLS = 1L << 22;
constexpr long na_round(long w) { return (w + 63) / 64 * 64; }  // 256-byte sections
constexpr long NA_COUNTS = 0;                                                  // NA_MAX_CELLS + 1 ints
constexpr long NA_STARTS = NA_COUNTS + na_round(NA_MAX_CELLS + 1);             // NA_MAX_CELLS + 1 ints
constexpr long NA_SUMS = NA_STARTS + na_round(NA_MAX_CELLS + 1);               // scan partials
constexpr long NA_KEY = NA_SUMS + na_round((NA_MAX_CELLS + 1 + SCAN_TILE_ITEMS - 1) / SCAN_TILE_ITEMS);
constexpr long NA_SORTED = NA_KEY + na_round(NA_MAX_PARTICLES);                // particles in cell order
constexpr long NA_WORDS = NA_SORTED + na_round(NA_MAX_PARTICLES);              // 10.5M ints = 42 MB

// The scratch of nn_neighbors_noalloc.  Host code never names it directly:
// scratch_base() returns the device's address from hipGetSymbolAddress.
static __device__ __align__(256) int scratch[NA_WORDS];

__global__ void k_na_zero(int* p, long n, int* status) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
  if (status && i == 0) *status = 0;
}

// cell_of of the context's search on caller-layout positions
__device__ __forceinline__ int na_key(const float* x, int i, float h, int g0, int g1, int g2) {
  const float fx = x[3 * (long)i] / h, fy = x[3 * (long)i + 1] / h, fz = x[3 * (long)i + 2] / h;
  if (!(fabsf(fx) < 1.0e9f && fabsf(fy) < 1.0e9f && fabsf(fz) < 1.0e9f)) return -1;
  const int cx = (int)fx, cy = (int)fy, cz = (int)fz;
  if (cx < 0 || cx >= g0 || cy < 0 || cy >= g1 || cz < 0 || cz >= g2) return -1;
  return (cx * g1 + cy) * g2 + cz;
}

__global__ void k_na_count(const float* x, int n, float h, int g0, int g1, int g2, int* key, int* counts,
                           int* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = na_key(x, i, h, g0, g1, g2);
  key[i] = k;
  if (k < 0) atomicOr(status, ST_OUT_OF_GRID);
  else atomicAdd(&counts[k], 1);
}

// every particle into its cell's range (in arrival order; k_na_order sorts)
__global__ void k_na_bucket(const int* key, int n, const int* starts, int* cursor, int* sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = key[i];
  if (k >= 0) sorted[starts[k] + atomicAdd(&cursor[k], 1)] = i;
}

// one thread per cell: its particles in increasing index (insertion sort of
// at most MAX_PIC entries; a fuller cell is reported and left unsorted)
__global__ void k_na_order(const int* starts, int ncells, int* sorted, int* status) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int b = starts[c], e = starts[c + 1];
  if (e - b > MAX_PIC) {
    atomicOr(status, ST_CELL_OVERFLOW);
    return;
  }
  for (int a = b + 1; a < e; a++) {
    const int v = sorted[a];
    int q = a - 1;
    while (q >= b && sorted[q] > v) {
      sorted[q + 1] = sorted[q];
      q--;
    }
    sorted[q + 1] = v;
  }
}

// store_neighbors (:836-847) for particle i in caller order: the 9 runs of
// (cx+dx, cy+dy, cz-1..cz+1) cells in cell order, -1 padding up to cap
__global__ void k_na_lists(const float* x, int n, float h, float thr2, int g0, int g1, int g2, const int* key,
                           const int* starts, const int* sorted, int* nbr, int* count, int cap, int* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int* row = nbr + (long)i * cap;
  int cnt = 0;
  const int k = key[i];
  if (k >= 0) {
    const int cz = k % g2, cy = (k / g2) % g1, cx = k / (g1 * g2);
    const float xi = x[3 * (long)i], yi = x[3 * (long)i + 1], zi = x[3 * (long)i + 2];
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < g2 ? cz + 1 : g2 - 1;
    for (int dx = -1; dx <= 1; dx++) {
      const int a = cx + dx;
      if (a < 0 || a >= g0) continue;
      for (int dy = -1; dy <= 1; dy++) {
        const int bb = cy + dy;
        if (bb < 0 || bb >= g1) continue;
        const int base = (a * g1 + bb) * g2;
        const int j0 = starts[base + z0], j1 = starts[base + z1 + 1];
        for (int t = j0; t < j1; t++) {
          const int j = sorted[t];
          if (j == i) continue;
          const float ex = xi - x[3 * (long)j], ey = yi - x[3 * (long)j + 1], ez = zi - x[3 * (long)j + 2];
          const float ss = ex * ex + ey * ey + ez * ez;
          if (ss < thr2) {
            if (cnt < MAX_NB) row[cnt] = j;
            cnt++;
          }
        }
      }
    }
  }
  if (cnt > MAX_NB) {
    atomicOr(status, ST_NEIGHBOR_OVERFLOW);
    cnt = MAX_NB;
  }
  for (int m = cnt; m < cap; m++) row[m] = -1;
  count[i] = cnt;
}

// Address of this device's scratch, looked up once per device with the device
// current (each GPU holds its own copy of the array).
constexpr int NA_MAX_DEVICES = 64;
static std::mutex na_mu;
static int* na_base[NA_MAX_DEVICES] = {};

static int scratch_base(int** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= NA_MAX_DEVICES) {
    set_error("nn_neighbors_noalloc: device index beyond 63");
    return NN_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(na_mu);
  if (!na_base[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
    na_base[dev] = static_cast<int*>(p);
  }
  *out = na_base[dev];
  return NN_OK;
}

}  // namespace nn

extern "C" int32_t nn_neighbors_noalloc(const nn_params* params, const float* x, int64_t n, int32_t* nbr,
                                        int32_t* count, int32_t cap, int32_t* status, void* stream) {
  using namespace nn;
  if (!params || !status || n < 0 || (n > 0 && (!x || !nbr || !count))) {
    set_error("nn_neighbors_noalloc: bad arguments");
    return NN_E_ARG;
  }
  if (cap < MAX_NB) {
    set_error("nn_neighbors_noalloc: cap must be at least 60");
    return NN_E_ARG;
  }
  if (n > NA_MAX_PARTICLES) {
    set_error("nn_neighbors_noalloc: more than 1048576 particles (the fixed scratch's maximum)");
    return NN_E_ARG;
  }
  if (!(params->particle_radius > 0.0)) {
    set_error("nn_neighbors_noalloc: particle_radius must be positive");
    return NN_E_ARG;
  }
  const double h = 4.0 * params->particle_radius;  // Parameter.support_radius, :606
  long g[3];
  long cells = 1;
  for (int k = 0; k < 3; k++) {
    const double gk = std::ceil((params->domain_end[k] - params->domain_start[k]) / h);  // :773-775
    if (!(gk >= 1.0) || gk > (double)NA_MAX_CELLS) {
      set_error("nn_neighbors_noalloc: the domain must span 1 .. 4194304 cells per axis");
      return NN_E_ARG;
    }
    g[k] = (long)gk;
    cells *= g[k];
    if (cells > NA_MAX_CELLS) {
      set_error("nn_neighbors_noalloc: more than 4194304 grid cells (the fixed scratch's maximum)");
      return NN_E_ARG;
    }
  }
  int* base = nullptr;
  int rc = scratch_base(&base);
  if (rc < 0) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float hf = (float)h;
  const float thr2 = neighbor_threshold(hf);
  int* counts = base + NA_COUNTS;
  int* starts = base + NA_STARTS;
  int* sums = base + NA_SUMS;
  int* key = base + NA_KEY;
  int* sorted = base + NA_SORTED;
  const int nc = (int)cells;
  const int ni = (int)n;
  const int T = 256;
  // the scratch keeps the last call's contents: counts (and the status) start at zero
  hipLaunchKernelGGL(k_na_zero, dim3((unsigned)((nc + 1 + T - 1) / T)), dim3(T), 0, st, counts, (long)nc + 1, status);
  if (ni > 0) {
    hipLaunchKernelGGL(k_na_count, dim3((ni + T - 1) / T), dim3(T), 0, st, x, ni, hf, (int)g[0], (int)g[1], (int)g[2],
                       key, counts, status);
  }
  launch_scan_i32(counts, nc + 1, sums, starts, st);
  if (ni > 0) {
    hipLaunchKernelGGL(k_na_zero, dim3((unsigned)((nc + T - 1) / T)), dim3(T), 0, st, counts, (long)nc, nullptr);
    hipLaunchKernelGGL(k_na_bucket, dim3((ni + T - 1) / T), dim3(T), 0, st, key, ni, starts, counts, sorted);
    hipLaunchKernelGGL(k_na_order, dim3((nc + T - 1) / T), dim3(T), 0, st, starts, nc, sorted, status);
    hipLaunchKernelGGL(k_na_lists, dim3((ni + T - 1) / T), dim3(T), 0, st, x, ni, hf, thr2, (int)g[0], (int)g[1],
                       (int)g[2], key, starts, sorted, nbr, count, cap, status);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "nn_neighbors_noalloc launch");
  return NN_OK;
}
