// kernels_search.hip -- neighbor search as a counting sort into cell order plus
// compact neighbor lists (replaces NeighborhoodSearchSparse.run_search,
// SPH/main.py:815-847).
//
// Pipeline per step:
//   cell_count  : cell key of every particle, per-cell arrival counter (atomic)
//   scan        : exclusive prefix sum of the cell counts (block scans over LDS
//                 with wavefront shuffles, three passes)
//   bucket      : every particle into its cell's range of a bucket array
//   place       : final slot = cell start + number of cell-mates with a smaller
//                 original index; state is moved into the new order
//   build_lists : for every particle the 9 contiguous runs of (cx+dx, cy+dy,
//                 cz-1..cz+1) cells are scanned with the reference's FP32
//                 distance test, which yields the reference's serial list order
//                 (SPH/main.py:840-847: dx slowest, dz fastest, then slot order =
//                 original index order, :815-820)
#include "device_common.h"

namespace nn {

__global__ void k_fill_i32(int* p, long n, int v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_f4(float4* p, long n, float4 v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_i32(int* p, long n, int v, hipStream_t st) {
  if (n <= 0) return;
  NN_LAUNCH(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

void launch_fill_f4(float4* p, long n, float4 v, hipStream_t st) {
  if (n <= 0) return;
  NN_LAUNCH(k_fill_f4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

// ---------------------------------------------------------------------------
// exclusive scan over n ints: 256 threads x 8 items per block
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256;
constexpr int SCAN_I = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_I;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ int block_excl_scan(int v, int* lds, int& total) {
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int incl = wave_incl_scan(v);
  if (lane == 63) lds[wid] = incl;
  __syncthreads();
  int off = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < SCAN_T / 64; w++) {
    int t = lds[w];
    if (w < wid) off += t;
    total += t;
  }
  __syncthreads();
  return off + incl - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const int* in, int n, int* sums) {
  __shared__ int lds[SCAN_T / 64];
  long base = (long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_I; k++) s += (base + k < n) ? in[base + k] : 0;
  int total;
  block_excl_scan(s, lds, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_top(int* sums, int nb) {
  __shared__ int lds[SCAN_T / 64];
  int carry = 0;
  for (int base = 0; base < nb; base += SCAN_T) {
    int i = base + threadIdx.x;
    int v = i < nb ? sums[i] : 0;
    int total;
    int ex = block_excl_scan(v, lds, total);
    if (i < nb) sums[i] = ex + carry;
    carry += total;
  }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_down(const int* in, int n, const int* sums, int* out) {
  __shared__ int lds[SCAN_T / 64];
  long base = (long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
  int v[SCAN_I];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_I; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_excl_scan(s, lds, total) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_I; k++) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

// exclusive scan of in[0..n) into out, with ceil(n / SCAN_TILE) ints of tmp
void launch_scan_i32(const int* in, int n, int* tmp, int* out, hipStream_t st) {
  int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  NN_LAUNCH(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, st, in, n, tmp);
  NN_LAUNCH(k_scan_top, dim3(1), dim3(SCAN_T), 0, st, tmp, nb);
  NN_LAUNCH(k_scan_down, dim3(nb), dim3(SCAN_T), 0, st, in, n, tmp, out);
}

void launch_scan(const DevRef& d, int n, hipStream_t st) {
  launch_scan_i32(d.h.cell_count, n, d.h.scan_tmp, d.h.cell_start, st);
}

// ---------------------------------------------------------------------------
// counting sort into cell order
// ---------------------------------------------------------------------------
// pos_to_index, SPH/main.py:805-807: truncation of x / h, no domain offset.
__device__ __forceinline__ bool cell_of(const SceneDev& sc, float4 p, int& cx, int& cy, int& cz) {
  float fx = p.x / sc.h, fy = p.y / sc.h, fz = p.z / sc.h;
  if (!(fabsf(fx) < 1.0e9f && fabsf(fy) < 1.0e9f && fabsf(fz) < 1.0e9f)) return false;
  cx = (int)fx;
  cy = (int)fy;
  cz = (int)fz;
  return cx >= 0 && cx < sc.gn0 && cy >= 0 && cy < sc.gn1 && cz >= 0 && cz < sc.gn2;
}

// Cell key and a unique rank among the cell's particles.  Particles arrive in
// the previous step's cell order, so equal keys sit in runs of adjacent lanes:
// the first lane of each run adds the run length with one atomic and the run's
// lanes take consecutive ranks from it.  (Ranks only need to be unique; the
// final order inside a cell is fixed by k_place.)
__global__ void k_cell_count(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;  // uniform over the wavefront
  const SceneDev& sc = d.scenes[s];
  const bool valid = i < sc.p_begin + sc.n;
  int key = -1;
  if (valid) {
    int cx, cy, cz;
    if (cell_of(sc, d.pos[i], cx, cy, cz)) {
      key = sc.cell_begin + (cx * sc.gn1 + cy) * sc.gn2 + cz;
    } else {
      // out of grid: parked in the scene's extra cell, reported, never searched
      key = sc.cell_begin + sc.gn0 * sc.gn1 * sc.gn2;
      atomicOr(&d.state[s].status, ST_OUT_OF_GRID);
    }
  }
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool start = lane == 0 || prev != key;
  const unsigned long long starts = __ballot(start);
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all 64 for lane 63)
  const int leader = 63 - __clzll(starts & upto);
  const unsigned long long after = starts & ~upto;
  const int next = after ? __ffsll(after) - 1 : 64;
  int base = 0;
  if (start && key >= 0) base = atomicAdd(&d.cell_count[key], next - lane);
  base = __shfl(base, leader, 64);
  if (!valid) return;
  const int r = base + (lane - leader);
  if (r >= MAX_PIC) atomicOr(&d.state[s].status, ST_CELL_OVERFLOW);
  d.key[i] = key;
  d.rank[i] = r;
}

__global__ void k_bucket(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  d.bucket[d.cell_start[d.key[i]] + d.rank[i]] = i;
}

__global__ void k_place(Dev d) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  const int key = d.key[i];
  const int cs = d.cell_start[key], ce = d.cell_start[key + 1];
  const int inf = d.info[i];
  const int my = id_of(inf);
  int r = 0;
  for (int k = cs; k < ce; k++) r += id_of(d.info[d.bucket[k]]) < my ? 1 : 0;
  const int dest = sc.shift + cs + r;
  d.pos_alt[dest] = d.pos[i];
  d.vel_alt[dest] = d.vel[i];
  d.info_alt[dest] = inf;
  if (sc.flags & SF_WARM_CG) d.vdiff_alt[dest] = d.vdiff[i];
  d.inv[sc.p_begin + my] = dest - sc.p_begin;
}

// store_neighbors, SPH/main.py:836-847, over the sorted particles.  With
// all == 0 (the search inside a step) static boundary particles get no list:
// after the boundary volumes of the first search no sweep reads one.
__global__ void k_build_lists(Dev d, int all) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  int cnt = 0;
  int cx, cy, cz;
  const float4 xi = d.pos[i];
  const bool want = i < sc.p_begin + sc.n && (all || kind_of(d.info[i]) != KIND_STATIC);
  // list positions are collected in groups of four and each group is written
  // as one int4 when complete (the last one possibly partial, its unused words
  // never read)
  int4 grp = make_int4(0, 0, 0, 0);
  int4* groups = reinterpret_cast<int4*>(d.nbr);
  if (want && cell_of(sc, xi, cx, cy, cz)) {
    const float thr2 = sc.thr2;
    const long P = d.P;
    const int z0 = cz > 0 ? cz - 1 : 0;
    const int z1 = cz + 1 < sc.gn2 ? cz + 1 : sc.gn2 - 1;
    for (int dx = -1; dx <= 1; dx++) {
      const int a = cx + dx;
      if (a < 0 || a >= sc.gn0) continue;
      for (int dy = -1; dy <= 1; dy++) {
        const int b = cy + dy;
        if (b < 0 || b >= sc.gn1) continue;
        const int base = sc.cell_begin + (a * sc.gn1 + b) * sc.gn2;
        const int j0 = d.cell_start[base + z0] + sc.shift;
        const int j1 = d.cell_start[base + z1 + 1] + sc.shift;
        // candidates in order, 8 positions loaded together
        for (int jb = j0; jb < j1; jb += 8) {
          float4 xj[8];
#pragma unroll
          for (int u = 0; u < 8; u++) xj[u] = d.pos[jb + u < j1 ? jb + u : i];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int j = jb + u;
            if (j >= j1 || j == i) continue;
            const float ex = xi.x - xj[u].x, ey = xi.y - xj[u].y, ez = xi.z - xj[u].z;
            const float ss = ex * ex + ey * ey + ez * ez;
            // sqrt_rn(ss) < h  <=>  ss < thr2 (thr2: least float whose rounded sqrt >= h)
            if (ss < thr2) {
              if (cnt < MAX_NB) {
                const int w = cnt & 3;
                grp.x = w == 0 ? j : grp.x;
                grp.y = w == 1 ? j : grp.y;
                grp.z = w == 2 ? j : grp.z;
                grp.w = w == 3 ? j : grp.w;
                if (w == 3) groups[(long)(cnt >> 2) * P + i] = grp;
              }
              cnt++;
            }
          }
        }
      }
    }
    if (cnt > MAX_NB) {
      atomicOr(&d.state[s].status, ST_NEIGHBOR_OVERFLOW);
      cnt = MAX_NB;
    }
    if (cnt & 3) groups[(long)(cnt >> 2) * P + i] = grp;
  }
  // the step's pair counts cover the lists a step uses: fluid and dynamic
  // solid particles (static boundary lists serve only the boundary volumes)
  int cf = 0, tot = 0;
  if (i < sc.p_begin + sc.n) {
    d.nnbr[i] = cnt;
    const int kind = kind_of(d.info[i]);
    if (kind == KIND_FLUID) cf = cnt;
    if (kind != KIND_STATIC) tot = cnt;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    tot += __shfl_xor(tot, off, 64);
    cf += __shfl_xor(cf, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && tot > 0) atomicAdd(&d.state[s].pairs, (unsigned long long)tot);
  if ((threadIdx.x & 63) == 0 && cf > 0) atomicAdd(&d.state[s].pairs_fluid, (unsigned long long)cf);
}

// compute_static/moving_boundary_volume, SPH/main.py:1010-1035:
// V_b = 3 / (W0 + sum over solid neighbors of W); stored as -V (solid sign).
__global__ void k_bvol(Dev d, int dynamic) {
  int i, s;
  if (!slot_scene(d, i, s)) return;
  const SceneDev& sc = d.scenes[s];
  if (i >= sc.p_begin + sc.n) return;
  const int kind = kind_of(d.info[i]);
  if (dynamic ? !dyn_solid(kind) : kind != KIND_STATIC) return;
  const float4 xi = d.pos[i];
  const int cnt = d.nnbr[i];
  const int P = d.P;
  KSum ks{sc.W0, 0.0f};
  float plain = sc.W0;
  const bool comp = (sc.flags & SF_COMPENSATED) != 0;
  for (int k = 0; k < cnt; k++) {
    const int j = d.nbr[nb_word(P, i, k)];
    const float4 xj = d.pos[j];
    if (xj.w < 0.0f) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float w = kern_w(sc.h, sc.kW, norm3(rx, ry, rz));
      if (comp) ks.add(w);
      else plain += w;
    }
  }
  const float delta = comp ? ks.s : plain;
  d.pos[i].w = -(1.0f / delta * 3.0f);
}

void launch_cell_count(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_cell_count, dim3((d.h.Pn + block - 1) / block), dim3(block), 0, st, d.h);
}
void launch_bucket(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_bucket, dim3((d.h.Pn + block - 1) / block), dim3(block), 0, st, d.h);
}
void launch_place(const DevRef& d, int block, hipStream_t st) {
  NN_LAUNCH(k_place, dim3((d.h.Pn + block - 1) / block), dim3(block), 0, st, d.h);
}
void launch_build_lists(const DevRef& d, int block, int all, hipStream_t st) {
  NN_LAUNCH(k_build_lists, dim3((d.h.Pn + block - 1) / block), dim3(block), 0, st, d.h, all);
}
void launch_bvol(const DevRef& d, int block, int dynamic, hipStream_t st) {
  NN_LAUNCH(k_bvol, dim3((d.h.Pn + block - 1) / block), dim3(block), 0, st, d.h, dynamic);
}

}  // namespace nn
