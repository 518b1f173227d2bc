/*
 * surface_oracle.c -- CPU restatement of the particle surface reconstruction of
 * SPH/surf_recon/srtool.py, the path SRTool.dual_contouring runs
 * (srtool.py:1264-1304):
 *
 *   particle hash       update_particles / _SRVoxel.add_particle (:1193-1196, :820-823)
 *   isotropic field     generate_fluid_field, poly6 (:1239-1251, :1213-1227)
 *   anisotropic kernel  pre_process_data (:1007-1112): x_bar, C_i, G_i
 *   smooth field        generate_smooth_fluid_field_impl, cubic_spline_kernel (:1136-1157, :1115-1124)
 *   smoothing           smooth_sdf, density_field_stats, smooth_sdf_impl (:1425-1484)
 *   dual contouring     dual_contouring_impl, process_normal_at, solve_qef (:1532-1596, :1336-1349, :70-80)
 *   polygons            dual_contouring_polygen, get_edge_adjacent_cubes (:1708-1763, :1613-1705)
 *   normals             process_normal (:1384-1410)
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (nonnewton_amd/csrc/
 * surface.hip).  Only tests/ load it.
 *
 * The reference runs these as Taichi kernels with atomics on sparse fields; the
 * restatement fixes every order the reference leaves to its scheduler:
 *  - the vertex field is accumulated in 2^-32 fixed point (int64), so it is
 *    independent of the order of the contributions, then rounded once to FP32;
 *  - a voxel's particles are listed by increasing particle index;
 *  - the statistics of the smoothing are FP64 sums in a fixed blocked order;
 *  - mesh vertices are numbered in grid order of their cubes, triangles in
 *    grid order of the cube that emits them.
 * Every other expression is FP32 in the reference's operation order (compile
 * with -ffp-contract=off).  Deviations are named where they occur.
 *
 * Grid: vertex (x, y, z) >= 0 sits at origin + (x, y, z) * voxel_size.  Values are
 * stored for every vertex of the 32^3 vertex blocks that cover the particles'
 * footprints (the reference's leaf blocks, srtool.py:747-750); "active" blocks
 * are the blocks the two field passes write, exactly the blocks the reference's
 * sparse loops visit.  Reads outside active blocks return 0 (an inactive
 * sparse cell reads 0).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define SRO_BLOCK 32
#define SRO_MAX_PER_VOXEL 64 /* max_num_particle_per_voxel, srtool.py:685 */
#define SRO_FIX_SCALE 4294967296.0
#define SRO_FIX_INV (1.0 / 4294967296.0)
#define SRO_MAX_CONTRIB 1048576.0f /* |contribution| bound that keeps the int64 sums exact */

enum { SRO_OK = 0, SRO_E_ARG = -1, SRO_E_OOM = -3, SRO_E_VOXEL_OVERFLOW = -5, SRO_E_OUT_OF_GRID = -6, SRO_E_RANGE = -7 };

typedef struct sro {
  /* parameters, rounded to FP32 as the Taichi kernels see them */
  float o[3], vs, sr, iso, vol, acoef, c48, c16, lam, lam1, kr, ks, kn;
  int aniso, iters, want_normals, want_vel;
  float cvs3; /* voxel_size^3, contribution_coeff (:1165) */
  int n;
  float *pos, *vel;
  /* particle hash over the voxel range of the particles */
  int vlo[3], vdim[3];
  int *vstart, *vcount, *vlist;
  /* anisotropic kernel per particle */
  float *G, *xbar, *detG;
  /* vertex grid */
  int glo[3], gdim[3], bdim[3];
  long nvert, nblock;
  unsigned char *active;
  int64_t *acc, *vacc; /* vacc: 3 per vertex */
  float *field0, *field, *tmp, *vfield;
  double sd;
  int *cube_id;
  float *verts, *normals, *mvel;
  int *faces;
  long nverts, ntris;
} sro;

/* ---------------------------------------------------------------- helpers */

static inline float norm3f(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

/* exp for x <= 0 (the bilateral weights of smooth_sdf_impl, srtool.py:1433-1434).
 * Own method: Cody-Waite reduction x = k ln2 + r, |r| <= ln2 / 2, with ln2 split
 * in a 16-bit high part (k * ln2_hi exact) and a low part, then the degree-7
 * Taylor polynomial of e^r in Horner form (truncation error < 3e-9 relative)
 * and an exact scaling by 2^k.  Results below 2^-125 are returned as 0. */
float sro_expf(float x) {
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

/* one cyclic Jacobi rotation of the symmetric S (row-major) zeroing S[p][q];
 * V accumulates the rotations (Golub & Van Loan, Alg. 8.5.2) */
static void jacobi_rot(float *S, float *V, int p, int q) {
  const float apq = S[3 * p + q];
  if (apq == 0.0f) return;
  const float theta = (S[3 * q + q] - S[3 * p + p]) / (2.0f * apq);
  const float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
  const float c = 1.0f / sqrtf(t * t + 1.0f);
  const float s = t * c;
  for (int k = 0; k < 3; k++) {
    const float a = S[3 * k + p], b = S[3 * k + q];
    S[3 * k + p] = c * a - s * b;
    S[3 * k + q] = s * a + c * b;
  }
  for (int k = 0; k < 3; k++) {
    const float a = S[3 * p + k], b = S[3 * q + k];
    S[3 * p + k] = c * a - s * b;
    S[3 * q + k] = s * a + c * b;
  }
  for (int k = 0; k < 3; k++) {
    const float a = V[3 * k + p], b = V[3 * k + q];
    V[3 * k + p] = c * a - s * b;
    V[3 * k + q] = s * a + c * b;
  }
}

/* eigen-decomposition of a symmetric 3x3: 6 cyclic sweeps; ev = diagonal */
static void sym_eig3(const float *A, float *ev, float *V) {
  float S[9];
  for (int k = 0; k < 9; k++) S[k] = A[k];
  for (int k = 0; k < 9; k++) V[k] = (k % 4) == 0 ? 1.0f : 0.0f;
  for (int sweep = 0; sweep < 6; sweep++) {
    jacobi_rot(S, V, 0, 1);
    jacobi_rot(S, V, 0, 2);
    jacobi_rot(S, V, 1, 2);
  }
  ev[0] = S[0];
  ev[1] = S[4];
  ev[2] = S[8];
}

/* vertex grid */
static inline long gidx(const sro *s, int gx, int gy, int gz) { return ((long)gx * s->gdim[1] + gy) * s->gdim[2] + gz; }
static inline long bidx(const sro *s, int gx, int gy, int gz) {
  return ((long)(gx / SRO_BLOCK) * s->bdim[1] + gy / SRO_BLOCK) * s->bdim[2] + gz / SRO_BLOCK;
}
/* value of vertex (x, y, z) in absolute vertex coordinates; 0 outside active blocks */
static inline float val_at(const sro *s, const float *f, int x, int y, int z) {
  const int gx = x - s->glo[0], gy = y - s->glo[1], gz = z - s->glo[2];
  if (gx < 0 || gy < 0 || gz < 0 || gx >= s->gdim[0] || gy >= s->gdim[1] || gz >= s->gdim[2]) return 0.0f;
  if (!s->active[bidx(s, gx, gy, gz)]) return 0.0f;
  return f[gidx(s, gx, gy, gz)];
}
/* get_mapped_position (:795-796) */
static inline void vpos(const sro *s, int x, int y, int z, float *p) {
  p[0] = (float)x * s->vs + s->o[0];
  p[1] = (float)y * s->vs + s->o[1];
  p[2] = (float)z * s->vs + s->o[2];
}
/* get_voxel_index (:799-803): truncation toward zero */
static inline void voxel_of(const sro *s, const float *p, int *v) {
  v[0] = (int)((p[0] - s->o[0]) / s->vs);
  v[1] = (int)((p[1] - s->o[1]) / s->vs);
  v[2] = (int)((p[2] - s->o[2]) / s->vs);
}

/* poly6_weight (:1213-1227) */
static inline float poly6(const sro *s, float rx, float ry, float rz, float h) {
  const float rn = norm3f(rx, ry, rz);
  float res = 0.0f;
  if (rn <= h) {
    float tmp = h * h;
    res = tmp - rn * rn;
    res = res * res * res;
    tmp = tmp * tmp * tmp * tmp * h;
    res = s->acoef * res / tmp;
  }
  return res;
}

/* cubic_spline_kernel (:1115-1124) with the reference's placement of "+ 1" */
static inline float aniso_w(const sro *s, const float *G, float det, float dx, float dy, float dz) {
  const float a = G[0] * dx + G[1] * dy + G[2] * dz;
  const float b = G[3] * dx + G[4] * dy + G[5] * dz;
  const float c = G[6] * dx + G[7] * dy + G[8] * dz;
  const float q = norm3f(a, b, c);
  float res = 0.0f;
  if (0.0f < q && q <= 0.5f)
    res = s->c48 * det * (q * q * q - q * q) + 1.0f;
  else if (0.5f < q && q <= 1.0f)
    res = s->c16 * det * (1.0f - q) * (1.0f - q) * (1.0f - q);
  return res;
}

static inline int64_t fix_of(float v) { return (int64_t)llrint((double)v * SRO_FIX_SCALE); }

/* ---------------------------------------------------------------- lifecycle */

static void sro_free_run(sro *s) {
  free(s->pos); free(s->vel); free(s->vacc); free(s->vfield); free(s->mvel);
  s->vel = s->vfield = s->mvel = NULL;
  s->vacc = NULL;
  free(s->vstart); free(s->vcount); free(s->vlist);
  free(s->G); free(s->xbar); free(s->detG);
  free(s->active); free(s->acc); free(s->field0); free(s->field); free(s->tmp);
  free(s->cube_id); free(s->verts); free(s->normals); free(s->faces);
  s->pos = NULL; s->vstart = s->vcount = s->vlist = NULL;
  s->G = s->xbar = s->detG = NULL;
  s->active = NULL; s->acc = NULL; s->field0 = s->field = s->tmp = NULL;
  s->cube_id = NULL; s->verts = s->normals = NULL; s->faces = NULL;
  s->nverts = s->ntris = 0;
}

void sro_destroy(sro *s) {
  if (!s) return;
  sro_free_run(s);
  free(s);
}

/* dp: origin[3], voxel_size, particle_radius, smooth_radius, isolevel (Python
 * floats in the reference); ip: anisotropic, smooth iterations, normals */
sro *sro_create(const double *dp, const int *ip) {
  if (!(dp[3] > 0.0) || !(dp[4] > 0.0) || !(dp[5] > 0.0) || ip[1] < 0) return NULL;
  sro *s = (sro *)calloc(1, sizeof(sro));
  for (int a = 0; a < 3; a++) s->o[a] = (float)dp[a];
  s->vs = (float)dp[3];
  const double r = dp[4];
  s->sr = (float)dp[5];
  s->iso = (float)dp[6];
  s->vol = (float)(r * r * r * M_PI * 4 / 3);       /* :945 */
  s->acoef = (float)(315 / (64 * M_PI));            /* :946 */
  s->c48 = (float)(48 / M_PI);
  s->c16 = (float)(16 / M_PI);
  s->lam = (float)0.9;                              /* akLambda, :929 */
  s->lam1 = (float)(1 - 0.9);
  s->kr = 4.0f;                                     /* :925-928 */
  s->ks = 1400.0f;
  s->kn = 0.5f;
  s->aniso = ip[0];
  s->iters = ip[1];
  s->want_normals = ip[2];
  s->want_vel = ip[3];
  s->cvs3 = (float)(dp[3] * dp[3] * dp[3]);
  return s;
}

/* ---------------------------------------------------------------- stages */

/* update_particles: each particle's voxel, the voxel range, the lists in
 * increasing particle order */
static int build_hash(sro *s) {
  int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (int i = 0; i < s->n; i++) {
    for (int a = 0; a < 3; a++) {
      const float t = (s->pos[3 * i + a] - s->o[a]) / s->vs;
      if (!(t > -1.0f && t < 1.0e8f)) return SRO_E_OUT_OF_GRID; /* below the origin's voxel, huge or NaN */
    }
    int v[3];
    voxel_of(s, s->pos + 3 * i, v);
    for (int a = 0; a < 3; a++) {
      if (v[a] < lo[a]) lo[a] = v[a];
      if (v[a] > hi[a]) hi[a] = v[a];
    }
  }
  long nv = 1;
  for (int a = 0; a < 3; a++) {
    s->vlo[a] = lo[a];
    s->vdim[a] = hi[a] - lo[a] + 1;
    nv *= s->vdim[a];
  }
  s->vcount = (int *)calloc(nv, sizeof(int));
  s->vstart = (int *)calloc(nv + 1, sizeof(int));
  s->vlist = (int *)malloc(sizeof(int) * (size_t)(s->n > 0 ? s->n : 1));
  if (!s->vcount || !s->vstart || !s->vlist) return SRO_E_OOM;
  for (int i = 0; i < s->n; i++) {
    int v[3];
    voxel_of(s, s->pos + 3 * i, v);
    const long k = ((long)(v[0] - lo[0]) * s->vdim[1] + (v[1] - lo[1])) * s->vdim[2] + (v[2] - lo[2]);
    if (++s->vcount[k] > SRO_MAX_PER_VOXEL) return SRO_E_VOXEL_OVERFLOW;
  }
  for (long k = 0; k < nv; k++) s->vstart[k + 1] = s->vstart[k] + s->vcount[k];
  int *cur = (int *)calloc(nv, sizeof(int));
  for (int i = 0; i < s->n; i++) {
    int v[3];
    voxel_of(s, s->pos + 3 * i, v);
    const long k = ((long)(v[0] - lo[0]) * s->vdim[1] + (v[1] - lo[1])) * s->vdim[2] + (v[2] - lo[2]);
    s->vlist[s->vstart[k] + cur[k]++] = i;
  }
  free(cur);
  return SRO_OK;
}

/* particles of voxel (x, y, z): [*b, *e) in vlist */
static inline void voxel_range(const sro *s, int x, int y, int z, int *b, int *e) {
  const int a = x - s->vlo[0], c = y - s->vlo[1], d = z - s->vlo[2];
  if (a < 0 || c < 0 || d < 0 || a >= s->vdim[0] || c >= s->vdim[1] || d >= s->vdim[2]) {
    *b = *e = 0;
    return;
  }
  const long k = ((long)a * s->vdim[1] + c) * s->vdim[2] + d;
  *b = s->vstart[k];
  *e = s->vstart[k + 1];
}

static inline int smooth_range(const sro *s) { return (int)ceilf(4.0f * s->sr / s->vs) + 1; }

/* pre_process_data (:1007-1112) for particle i >= 1; particle 0 keeps G = 0,
 * x_bar = 0 (the loop starts at 1, :1013; reset() zeroes both, :1229-1236) */
static void aniso_particle(sro *s, int i) {
  const int R = smooth_range(s), off = R / 2;
  const float *xi = s->pos + 3 * i;
  int v[3];
  voxel_of(s, xi, v);
  const float r2 = 2.0f * s->sr;
  float tw = 0.0f, xw[3] = {0.0f, 0.0f, 0.0f};
  for (int dx = 0; dx < R; dx++)
    for (int dy = 0; dy < R; dy++)
      for (int dz = 0; dz < R; dz++) {
        const int x = v[0] + dx - off, y = v[1] + dy - off, z = v[2] + dz - off;
        if (x < 0 || y < 0 || z < 0) continue;
        int b, e;
        voxel_range(s, x, y, z, &b, &e);
        for (int m = b; m < e; m++) {
          const float *xj = s->pos + 3 * s->vlist[m];
          const float nr = norm3f(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]);
          float w = 0.0f;
          if (nr < r2) {
            const float t = nr / r2;
            w = 1.0f - t * t * t; /* isoweight (:994-999); pow(t, 3.0) as t * t * t */
          }
          tw += w;
          xw[0] += w * xj[0];
          xw[1] += w * xj[1];
          xw[2] += w * xj[2];
        }
      }
  if (tw != 0.0f)
    for (int a = 0; a < 3; a++) xw[a] /= tw;
  float *xb = s->xbar + 3 * i;
  for (int a = 0; a < 3; a++) xb[a] = s->lam1 * xi[a] + s->lam * xw[a];
  int nn = 0;
  float C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int dx = 0; dx < R; dx++)
    for (int dy = 0; dy < R; dy++)
      for (int dz = 0; dz < R; dz++) {
        const int x = v[0] + dx - off, y = v[1] + dy - off, z = v[2] + dz - off;
        if (x < 0 || y < 0 || z < 0) continue;
        int b, e;
        voxel_range(s, x, y, z, &b, &e);
        for (int m = b; m < e; m++) {
          const float *xj = s->pos + 3 * s->vlist[m];
          const float nr = norm3f(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]);
          float w = 0.0f;
          if (nr < r2) {
            const float t = nr / r2;
            w = 1.0f - t * t * t;
          }
          if (w > 0.0f) {
            nn++;
            const float dp[3] = {xj[0] - xw[0], xj[1] - xw[1], xj[2] - xw[2]};
            for (int a = 0; a < 3; a++)
              for (int c = 0; c < 3; c++) C[3 * a + c] += w * (dp[a] * dp[c]);
          }
        }
      }
  if (tw != 0.0f)
    for (int k = 0; k < 9; k++) C[k] /= tw;
  /* C = R Sigma R^T (ti.svd of the symmetric C, singular values descending);
   * here the symmetric eigen-decomposition, sorted */
  float ev[3], V[9];
  sym_eig3(C, ev, V);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2 - a; b++)
      if (ev[b] < ev[b + 1]) {
        const float t = ev[b];
        ev[b] = ev[b + 1];
        ev[b + 1] = t;
        for (int k = 0; k < 3; k++) {
          const float x = V[3 * k + b];
          V[3 * k + b] = V[3 * k + b + 1];
          V[3 * k + b + 1] = x;
        }
      }
  float *G = s->G + 9 * i;
  if (tw == 0.0f) {
    for (int k = 0; k < 9; k++) G[k] = (k % 4) == 0 ? 1.0f : 0.0f;
  } else {
    float sg[3];
    if (nn > 25) { /* Nvep */
      sg[0] = ev[0];
      sg[1] = fmaxf(ev[1], ev[0] / s->kr);
      sg[2] = fmaxf(ev[2], ev[0] / s->kr);
      for (int a = 0; a < 3; a++) sg[a] *= s->ks;
    } else {
      sg[0] = sg[1] = sg[2] = s->kn;
    }
    const float ih = 1.0f / s->sr;
    float P[9];
    for (int a = 0; a < 3; a++)
      for (int k = 0; k < 3; k++) P[3 * a + k] = (ih * V[3 * a + k]) * (1.0f / sg[k]);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) G[3 * a + b] = P[3 * a] * V[3 * b] + P[3 * a + 1] * V[3 * b + 1] + P[3 * a + 2] * V[3 * b + 2];
  }
  s->detG[i] = G[0] * (G[4] * G[8] - G[7] * G[5]) - G[3] * (G[1] * G[8] - G[7] * G[2]) + G[6] * (G[1] * G[5] - G[4] * G[2]);
}

/* the vertex-box multiplicity along one axis: how many voxels n in
 * [c - off, c - off + R - 1] have n == v or n + 1 == v (each voxel of the
 * smooth range visits its 8 corners, :1142-1155) */
static inline int mult_axis(int v, int c, int off, int R) {
  const int lo = c - off, hi = c - off + R - 1;
  return (v >= lo && v <= hi) + (v - 1 >= lo && v - 1 <= hi);
}

static int alloc_grid(sro *s) {
  const int R = smooth_range(s), off = R / 2;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    const int vmin = s->vlo[a], vmax = s->vlo[a] + s->vdim[a] - 1;
    int l = vmin - 1, h = vmax + 1; /* generate_fluid_field's 3^3 box */
    if (s->aniso) {
      if (vmin - off < l) l = vmin - off;
      if (vmax - off + R > h) h = vmax - off + R; /* corners up to n + 1 */
    }
    if (l < 0) l = 0;
    lo[a] = (l / SRO_BLOCK) * SRO_BLOCK;
    hi[a] = (h / SRO_BLOCK) * SRO_BLOCK + SRO_BLOCK - 1;
    s->glo[a] = lo[a];
    s->gdim[a] = hi[a] - lo[a] + 1;
    s->bdim[a] = s->gdim[a] / SRO_BLOCK;
  }
  s->nvert = (long)s->gdim[0] * s->gdim[1] * s->gdim[2];
  s->nblock = (long)s->bdim[0] * s->bdim[1] * s->bdim[2];
  s->active = (unsigned char *)calloc(s->nblock, 1);
  s->acc = (int64_t *)calloc(s->nvert, sizeof(int64_t));
  s->field0 = (float *)calloc(s->nvert, sizeof(float));
  s->field = (float *)calloc(s->nvert, sizeof(float));
  s->tmp = (float *)calloc(s->nvert, sizeof(float));
  s->cube_id = (int *)calloc(s->nvert, sizeof(int));
  if (!s->active || !s->acc || !s->field0 || !s->field || !s->tmp || !s->cube_id) return SRO_E_OOM;
  if (s->vel) {
    s->vacc = (int64_t *)calloc(3 * (size_t)s->nvert, sizeof(int64_t));
    s->vfield = (float *)calloc(3 * (size_t)s->nvert, sizeof(float));
    if (!s->vacc || !s->vfield) return SRO_E_OOM;
  }
  return SRO_OK;
}

static inline void mark(sro *s, int x, int y, int z) {
  s->active[bidx(s, x - s->glo[0], y - s->glo[1], z - s->glo[2])] = 1;
}

/* generate_fluid_field (:1239-1251) and generate_smooth_fluid_field_impl
 * (:1136-1157), both added into voxel_vertex_value */
static int splat(sro *s) {
  const int R = smooth_range(s), off = R / 2;
  const float h = s->sr;
  for (int i = 0; i < s->n; i++) {
    const float *xi = s->pos + 3 * i;
    int v[3];
    voxel_of(s, xi, v);
    for (int dx = 0; dx < 3; dx++)
      for (int dy = 0; dy < 3; dy++)
        for (int dz = 0; dz < 3; dz++) {
          const int x = v[0] + dx - 1, y = v[1] + dy - 1, z = v[2] + dz - 1;
          if (x < 0 || y < 0 || z < 0) continue;
          float p[3];
          vpos(s, x, y, z, p);
          const float w = s->vol * poly6(s, xi[0] - p[0], xi[1] - p[1], xi[2] - p[2], h);
          if (!(fabsf(w) <= SRO_MAX_CONTRIB)) return SRO_E_RANGE;
          mark(s, x, y, z);
          s->acc[gidx(s, x - s->glo[0], y - s->glo[1], z - s->glo[2])] += fix_of(w);
        }
    if (!s->aniso) continue;
    const float *G = s->G + 9 * i, *xb = s->xbar + 3 * i;
    const float det = s->detG[i];
    for (int x = v[0] - off; x <= v[0] - off + R; x++) {
      if (x < 0) continue;
      const int mx = mult_axis(x, v[0], off, R);
      for (int y = v[1] - off; y <= v[1] - off + R; y++) {
        if (y < 0) continue;
        const int my = mult_axis(y, v[1], off, R);
        for (int z = v[2] - off; z <= v[2] - off + R; z++) {
          if (z < 0) continue;
          const int m = mx * my * mult_axis(z, v[2], off, R);
          float p[3];
          vpos(s, x, y, z, p);
          const float w = s->vol * aniso_w(s, G, det, xb[0] - p[0], xb[1] - p[1], xb[2] - p[2]);
          if (!(fabsf(w) <= SRO_MAX_CONTRIB)) return SRO_E_RANGE;
          mark(s, x, y, z);
          const long q = gidx(s, x - s->glo[0], y - s->glo[1], z - s->glo[2]);
          s->acc[q] += (int64_t)m * fix_of(w);
          if (s->vel) { /* generate_smooth_fluid_field_with_velocity_impl (:1160-1190) */
            const float f = poly6(s, xi[0] - p[0], xi[1] - p[1], xi[2] - p[2], h) * s->cvs3;
            for (int c = 0; c < 3; c++) {
              const float u = s->vel[3 * i + c] * f;
              if (!(fabsf(u) <= SRO_MAX_CONTRIB)) return SRO_E_RANGE;
              s->vacc[3 * q + c] += (int64_t)m * fix_of(u);
            }
          }
        }
      }
    }
  }
  for (long k = 0; k < s->nvert; k++) s->field0[k] = (float)((double)s->acc[k] * SRO_FIX_INV);
  if (s->vel)
    for (long k = 0; k < 3 * s->nvert; k++) s->vfield[k] = (float)((double)s->vacc[k] * SRO_FIX_INV);
  /* inactive blocks read 0 whatever the accumulator holds (it holds 0) */
  return SRO_OK;
}

/* The blocked FP64 sum the GPU forms: per active block (in block order), 256
 * lanes each add the block's vertices l = lane + 256 k (local index
 * l = (lx * 32 + ly) * 32 + lz) in order, then a pairwise tree over the lanes;
 * blocks added in block order.  which = 0: sum of the values < 0 and their
 * count; which = 1: sum of (v - mean)^2 over them (v - mean in FP32). */
static void blocked_stats(const sro *s, const float *f, int which, float mean, double *sum, long *cnt) {
  double total = 0.0;
  long count = 0;
  double lane[256];
  for (int bx = 0; bx < s->bdim[0]; bx++)
    for (int by = 0; by < s->bdim[1]; by++)
      for (int bz = 0; bz < s->bdim[2]; bz++) {
        if (!s->active[((long)bx * s->bdim[1] + by) * s->bdim[2] + bz]) continue;
        for (int t = 0; t < 256; t++) {
          double a = 0.0;
          for (int l = t; l < SRO_BLOCK * SRO_BLOCK * SRO_BLOCK; l += 256) {
            const int gx = bx * SRO_BLOCK + (l >> 10), gy = by * SRO_BLOCK + ((l >> 5) & 31), gz = bz * SRO_BLOCK + (l & 31);
            const float v = f[gidx(s, gx, gy, gz)];
            if (v < 0.0f) {
              if (which == 0) {
                a += (double)v;
                count++;
              } else {
                const double d = (double)(v - mean);
                a += d * d;
              }
            }
          }
          lane[t] = a;
        }
        for (int span = 128; span > 0; span >>= 1)
          for (int t = 0; t < span; t++) lane[t] += lane[t + span];
        total += lane[0];
      }
  *sum = total;
  *cnt = count;
}

/* smooth_sdf (:1425-1430): sd from density_field_stats (:1437-1456) once, then
 * `iters` bilateral passes smooth_sdf_impl (:1459-1484) with radius 1 */
static void smooth(sro *s) {
  double sum;
  long cnt;
  blocked_stats(s, s->field, 0, 0.0f, &sum, &cnt);
  float sd;
  if (cnt == 0) {
    sd = NAN; /* mean /= 0 in the reference: every weight below is NaN */
  } else {
    const float mean = (float)(sum / (double)cnt);
    double var;
    long c2;
    blocked_stats(s, s->field, 1, mean, &var, &c2);
    sd = (float)sqrt(var / (double)cnt);
  }
  s->sd = sd;
  const float c2pi = (float)(2 * M_PI);
  for (int it = 0; it < s->iters; it++) {
    for (int gx = 0; gx < s->gdim[0]; gx++)
      for (int gy = 0; gy < s->gdim[1]; gy++)
        for (int gz = 0; gz < s->gdim[2]; gz++) {
          const long k = gidx(s, gx, gy, gz);
          if (!s->active[bidx(s, gx, gy, gz)]) {
            s->tmp[k] = 0.0f;
            continue;
          }
          const int i = gx + s->glo[0], j = gy + s->glo[1], l = gz + s->glo[2];
          const float vc = s->field[k];
          float avg = 0.0f, weight = 0.0f;
          for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dz = -1; dz <= 1; dz++) {
                const int x = i + dx, y = j + dy, z = l + dz;
                if (x < 0 || y < 0 || z < 0) continue;
                const float vn = val_at(s, s->field, x, y, z);
                const float d = fabsf(vc - vn);
                const float t = c2pi * sd * sd;
                const float gw = sro_expf(-d * d / (2.0f * sd * sd)) / (t * sqrtf(t));
                avg += gw * vn;
                weight += gw;
              }
          if (weight != 0.0f) avg /= weight;
          s->tmp[k] = avg;
        }
    float *t = s->field;
    s->field = s->tmp;
    s->tmp = t;
  }
}

/* calc_vertex_normal (:1307-1321) and its normalize (:1413-1417) */
static void vertex_normal(const sro *s, int x, int y, int z, float *n) {
  const float *f = s->field;
  n[0] = x == 0 ? -val_at(s, f, x + 1, y, z) : val_at(s, f, x - 1, y, z) - val_at(s, f, x + 1, y, z);
  n[1] = y == 0 ? -val_at(s, f, x, y + 1, z) : val_at(s, f, x, y - 1, z) - val_at(s, f, x, y + 1, z);
  n[2] = z == 0 ? -val_at(s, f, x, y, z + 1) : val_at(s, f, x, y, z - 1) - val_at(s, f, x, y, z + 1);
  const float nr = norm3f(n[0], n[1], n[2]);
  if (nr != 0.0f) {
    n[0] /= nr;
    n[1] /= nr;
    n[2] /= nr;
  } else {
    n[0] = n[1] = n[2] = 0.0f;
  }
}

/* process_normal_at (:1336-1349): trilinear blend of the 8 vertex normals */
static void normal_at(const sro *s, const float *p, float *out) {
  int v[3];
  voxel_of(s, p, v);
  float base[3];
  vpos(s, v[0], v[1], v[2], base);
  const float xd = (p[0] - base[0]) / s->vs, yd = (p[1] - base[1]) / s->vs, zd = (p[2] - base[2]) / s->vs;
  float c[8][3]; /* index (a << 2) | (b << 1) | c for offsets (a, b, c) */
  for (int k = 0; k < 8; k++) vertex_normal(s, v[0] + (k >> 2), v[1] + ((k >> 1) & 1), v[2] + (k & 1), c[k]);
  float r[3];
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

/* Givens rotation of rows p, q of B zeroing B[q][col]; U accumulates it in
 * columns p, q so that U B stays constant */
static void givens_rows(float *B, float *U, int p, int q, int col) {
  const float x = B[3 * p + col], y = B[3 * q + col];
  const float rn = sqrtf(x * x + y * y);
  float c = 1.0f, sn = 0.0f;
  if (rn > 0.0f) {
    c = x / rn;
    sn = y / rn;
  }
  for (int k = 0; k < 3; k++) {
    const float bp = B[3 * p + k], bq = B[3 * q + k];
    B[3 * p + k] = c * bp + sn * bq;
    B[3 * q + k] = -sn * bp + c * bq;
  }
  for (int k = 0; k < 3; k++) {
    const float up = U[3 * k + p], uq = U[3 * k + q];
    U[3 * k + p] = c * up + sn * uq;
    U[3 * k + q] = -sn * up + c * uq;
  }
}

/* A = U diag(s) V^T (3x3, row-major) by the structure of McAdams et al. 2011:
 * V = Jacobi eigenvectors of A^T A sorted by decreasing eigenvalue, then a
 * Givens QR of A V = U B; s = diag(B) (signed).  A small singular value comes
 * out at the rounding level of |A|, not of |A|^2. */
static void svd3(const float *A, float *U, float *sg, float *V) {
  float M[9], ev[3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) M[3 * a + b] = A[a] * A[b] + A[3 + a] * A[3 + b] + A[6 + a] * A[6 + b];
  sym_eig3(M, ev, V);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2 - a; b++)
      if (ev[b] < ev[b + 1]) {
        const float t = ev[b];
        ev[b] = ev[b + 1];
        ev[b + 1] = t;
        for (int k = 0; k < 3; k++) {
          const float x = V[3 * k + b];
          V[3 * k + b] = V[3 * k + b + 1];
          V[3 * k + b + 1] = x;
        }
      }
  float B[9];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) B[3 * a + b] = A[3 * a] * V[b] + A[3 * a + 1] * V[3 + b] + A[3 * a + 2] * V[6 + b];
  for (int k = 0; k < 9; k++) U[k] = (k % 4) == 0 ? 1.0f : 0.0f;
  givens_rows(B, U, 0, 1, 0);
  givens_rows(B, U, 0, 2, 0);
  givens_rows(B, U, 1, 2, 1);
  sg[0] = B[0];
  sg[1] = B[4];
  sg[2] = B[8];
}

/* The QEF of the cube's edge crossings, rows (n, n . (p - mean)).  The
 * reference stores them in a 4x4 matrix and triangularises it by Householder
 * (householder_decomp, :36-53), which holds only 4 rows; here every row is
 * folded into the 4x4 triangle by Givens rotations (the incremental QR of Ju et
 * al. 2002), identical up to row signs for 4 rows or fewer.  solve_qef (:70-80)
 * inverts the 3x3 block through its SVD with singular values clamped to
 * sign(s) max(0.1, |s|): x = V diag(1 / s') U^T b.  Singular values up to 1e-6
 * of the largest (the rounding level of a rank-deficient block) count as 0 and
 * their terms are dropped: there U is arbitrary, in the reference's SVD too. */
void sro_qef(const float *rows, int cnt, float *x) {
  float R[16] = {0};
  for (int r = 0; r < cnt; r++) {
    float a[4] = {rows[4 * r], rows[4 * r + 1], rows[4 * r + 2], rows[4 * r + 3]};
    for (int k = 0; k < 4; k++) {
      const float xk = a[k];
      if (xk == 0.0f) continue;
      const float rk = R[5 * k];
      const float rho = sqrtf(rk * rk + xk * xk);
      const float c = rk / rho, sn = xk / rho;
      for (int j = k; j < 4; j++) {
        const float u = R[4 * k + j], w = a[j];
        R[4 * k + j] = c * u + sn * w;
        a[j] = -sn * u + c * w;
      }
      a[k] = 0.0f;
    }
  }
  float A[9];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) A[3 * a + b] = R[4 * a + b];
  float Us[9], sg[3], V[9];
  svd3(A, Us, sg, V);
  const float bb[3] = {R[3], R[7], R[11]};
  const float tiny = 1e-6f * fmaxf(fabsf(sg[0]), fmaxf(fabsf(sg[1]), fabsf(sg[2])));
  float z[3];
  for (int k = 0; k < 3; k++) {
    const float bk = Us[k] * bb[0] + Us[3 + k] * bb[1] + Us[6 + k] * bb[2];
    const float sc = (sg[k] < 0.0f ? -1.0f : 1.0f) * fmaxf(0.1f, fabsf(sg[k]));
    z[k] = fabsf(sg[k]) > tiny ? bk / sc : 0.0f;
  }
  for (int a = 0; a < 3; a++) x[a] = V[3 * a] * z[0] + V[3 * a + 1] * z[1] + V[3 * a + 2] * z[2];
}

/* corner offsets of voxel_vertex_offset_map (:772-779) and the edge corner
 * pairs of interpolation_table (:925-936) */
static const int CORNER[8][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};
static const int EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

/* get_velocity_at (:1352-1359): 0 outside written blocks */
static void vel_at(const sro *s, int x, int y, int z, float *out) {
  const int gx = x - s->glo[0], gy = y - s->glo[1], gz = z - s->glo[2];
  if (gx < 0 || gy < 0 || gz < 0 || gx >= s->gdim[0] || gy >= s->gdim[1] || gz >= s->gdim[2] ||
      !s->active[bidx(s, gx, gy, gz)]) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const long q = gidx(s, gx, gy, gz);
  for (int c = 0; c < 3; c++) out[c] = s->vfield[3 * q + c];
}

/* process_velocity_at (:1362-1375) with trilinear_interpolate (:1324-1333) */
static void velocity_at(const sro *s, const float *p, float *out) {
  int v[3];
  voxel_of(s, p, v);
  float base[3];
  vpos(s, v[0], v[1], v[2], base);
  const float xd = (p[0] - base[0]) / s->vs, yd = (p[1] - base[1]) / s->vs, zd = (p[2] - base[2]) / s->vs;
  float c[8][3];
  for (int k = 0; k < 8; k++) vel_at(s, v[0] + (k >> 2), v[1] + ((k >> 1) & 1), v[2] + (k & 1), c[k]);
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

/* vertex_interpolate (:1254-1262) */
static inline void vinterp(float iso, const float *p1, const float *p2, float v1, float v2, float *out) {
  const float delta = v2 - v1;
  const float off = fabsf(delta) < 0.00001f ? 0.5f : (iso - v1) / delta;
  for (int a = 0; a < 3; a++) out[a] = p1[a] + off * (p2[a] - p1[a]);
}

/* cube (x, y, z) = the cube whose corner 3 is vertex (x, y, z): corner values,
 * positions, the cube index (bit j set: value < isolevel, :1535-1538) and the
 * crossed-edge mask (the marching-cubes edge table, derived: edge w is crossed
 * when its two corners differ) */
static int cube_setup(const sro *s, int x, int y, int z, float cv[8], float cp[8][3]) {
  int ci = 0;
  for (int j = 0; j < 8; j++) {
    cv[j] = val_at(s, s->field, x + CORNER[j][0], y + CORNER[j][1], z + CORNER[j][2]);
    vpos(s, x + CORNER[j][0], y + CORNER[j][1], z + CORNER[j][2], cp[j]);
    if (cv[j] < s->iso) ci |= 1 << j;
  }
  int mask = 0;
  for (int w = 0; w < 12; w++)
    if (((ci >> EDGE[w][0]) & 1) != ((ci >> EDGE[w][1]) & 1)) mask |= 1 << w;
  return (ci << 12) | mask;
}

/* dual_contouring_impl (:1532-1596) for one cube */
static void dc_point(const sro *s, int x, int y, int z, int mask, const float cv[8], const float cp[8][3], float *res) {
  float mean[3] = {0.0f, 0.0f, 0.0f};
  int cnt = 0;
  for (int w = 0; w < 12; w++) {
    if (!(mask & (1 << w))) continue;
    float q[3];
    vinterp(s->iso, cp[EDGE[w][0]], cp[EDGE[w][1]], cv[EDGE[w][0]], cv[EDGE[w][1]], q);
    for (int a = 0; a < 3; a++) mean[a] += q[a];
    cnt++;
  }
  for (int a = 0; a < 3; a++) mean[a] /= (float)cnt;
  float rows[48];
  int nr = 0;
  for (int w = 0; w < 12; w++) {
    if (!(mask & (1 << w))) continue;
    float q[3], n[3];
    vinterp(s->iso, cp[EDGE[w][0]], cp[EDGE[w][1]], cv[EDGE[w][0]], cv[EDGE[w][1]], q);
    normal_at(s, q, n);
    const float d[3] = {q[0] - mean[0], q[1] - mean[1], q[2] - mean[2]};
    rows[4 * nr] = n[0];
    rows[4 * nr + 1] = n[1];
    rows[4 * nr + 2] = n[2];
    rows[4 * nr + 3] = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
    nr++;
  }
  float r[3];
  sro_qef(rows, nr, r);
  const float t[3] = {r[0] + mean[0], r[1] + mean[1], r[2] + mean[2]};
  int v[3];
  voxel_of(s, t, v);
  if (v[0] != x || v[1] != y || v[2] != z) {
    for (int a = 0; a < 3; a++) res[a] = mean[a];
  } else {
    for (int a = 0; a < 3; a++) res[a] = r[a] + mean[a];
  }
}

/* the cubes sharing edge w of cube c0 (get_edge_adjacent_cubes, :1613-1705),
 * for the three edges at corner 3 that dual_contouring_polygen walks */
static const int ADJ_W[3] = {2, 3, 11};
static const int ADJ[3][3][3] = {
    {{0, 0, -1}, {0, -1, -1}, {0, -1, 0}},  /* edge 2 (corners 2-3, along x) */
    {{0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}},  /* edge 3 (corners 3-0, along y) */
    {{0, -1, 0}, {-1, -1, 0}, {-1, 0, 0}},  /* edge 11 (corners 3-7, along z) */
};

/* the DC vertex id + 1 of cube (x, y, z), 0 if it has none */
static inline int cube_vertex(const sro *s, int x, int y, int z) {
  const int gx = x - s->glo[0], gy = y - s->glo[1], gz = z - s->glo[2];
  if (gx < 0 || gy < 0 || gz < 0 || gx >= s->gdim[0] || gy >= s->gdim[1] || gz >= s->gdim[2]) return 0;
  return s->cube_id[gidx(s, gx, gy, gz)];
}

static int contour(sro *s) {
  /* pass 1: cubes with a crossing, numbered in grid order */
  long nv = 0;
  for (int gx = 0; gx < s->gdim[0]; gx++)
    for (int gy = 0; gy < s->gdim[1]; gy++)
      for (int gz = 0; gz < s->gdim[2]; gz++) {
        const long k = gidx(s, gx, gy, gz);
        s->cube_id[k] = 0;
        if (!s->active[bidx(s, gx, gy, gz)]) continue;
        float cv[8], cp[8][3];
        const int m = cube_setup(s, gx + s->glo[0], gy + s->glo[1], gz + s->glo[2], cv, cp) & 0xFFF;
        if (m) s->cube_id[k] = (int)(++nv);
      }
  s->nverts = nv;
  s->verts = (float *)malloc(sizeof(float) * 3 * (size_t)(nv > 0 ? nv : 1));
  if (!s->verts) return SRO_E_OOM;
  /* pass 2: the points and the triangles */
  long cap = 64, nt = 0;
  s->faces = (int *)malloc(sizeof(int) * 3 * cap);
  for (int gx = 0; gx < s->gdim[0]; gx++)
    for (int gy = 0; gy < s->gdim[1]; gy++)
      for (int gz = 0; gz < s->gdim[2]; gz++) {
        const long k = gidx(s, gx, gy, gz);
        const int id = s->cube_id[k];
        if (!id) continue;
        const int x = gx + s->glo[0], y = gy + s->glo[1], z = gz + s->glo[2];
        float cv[8], cp[8][3];
        const int cm = cube_setup(s, x, y, z, cv, cp);
        const int mask = cm & 0xFFF, ci = cm >> 12;
        dc_point(s, x, y, z, mask, cv, cp, s->verts + 3 * (long)(id - 1));
        for (int e = 0; e < 3; e++) {
          const int w = ADJ_W[e];
          if (!(mask & (1 << w))) continue;
          const int i1 = id;
          const int i2 = cube_vertex(s, x + ADJ[e][0][0], y + ADJ[e][0][1], z + ADJ[e][0][2]);
          const int i3 = cube_vertex(s, x + ADJ[e][1][0], y + ADJ[e][1][1], z + ADJ[e][1][2]);
          const int i4 = cube_vertex(s, x + ADJ[e][2][0], y + ADJ[e][2][1], z + ADJ[e][2][2]);
          if (!(i2 && i3 && i4)) continue;
          const int in0 = (ci >> EDGE[w][0]) & 1;
          const int flip = in0 != (w == 11);
          if (nt + 2 > cap) {
            cap *= 2;
            s->faces = (int *)realloc(s->faces, sizeof(int) * 3 * cap);
          }
          int *f = s->faces + 3 * nt;
          if (!flip) {
            f[0] = i3 - 1; f[1] = i2 - 1; f[2] = i1 - 1;
            f[3] = i4 - 1; f[4] = i3 - 1; f[5] = i1 - 1;
          } else {
            f[0] = i1 - 1; f[1] = i2 - 1; f[2] = i3 - 1;
            f[3] = i1 - 1; f[4] = i3 - 1; f[5] = i4 - 1;
          }
          nt += 2;
        }
      }
  s->ntris = nt;
  if (s->want_normals) {
    s->normals = (float *)malloc(sizeof(float) * 3 * (size_t)(nv > 0 ? nv : 1));
    for (long v = 0; v < nv; v++) normal_at(s, s->verts + 3 * v, s->normals + 3 * v);
  }
  if (s->want_vel) { /* process_velocity (:1378-1381) */
    s->mvel = (float *)calloc(3 * (size_t)(nv > 0 ? nv : 1), sizeof(float));
    if (s->vfield)
      for (long v = 0; v < nv; v++) velocity_at(s, s->verts + 3 * v, s->mvel + 3 * v);
  }
  return SRO_OK;
}

/* SRTool.dual_contouring (:1264-1304) on positions pos[3 n] and, when the
 * velocity channels are on and vel is given, velocities vel[3 n] */
int sro_run(sro *s, const float *pos, const float *vel, int n) {
  if (!s || n < 0 || (n > 0 && !pos)) return SRO_E_ARG;
  sro_free_run(s);
  s->n = n;
  s->pos = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
  if (n > 0) memcpy(s->pos, pos, sizeof(float) * 3 * (size_t)n);
  if (s->want_vel && vel && s->aniso && n > 0) {
    s->vel = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    memcpy(s->vel, vel, sizeof(float) * 3 * (size_t)n);
  }
  if (n == 0) return SRO_E_ARG;
  int rc = build_hash(s);
  if (rc) return rc;
  s->G = (float *)calloc(9 * (size_t)n, sizeof(float));
  s->xbar = (float *)calloc(3 * (size_t)n, sizeof(float));
  s->detG = (float *)calloc((size_t)n, sizeof(float));
  if (!s->G || !s->xbar || !s->detG) return SRO_E_OOM;
  if (s->aniso)
    for (int i = 1; i < n; i++) aniso_particle(s, i);
  if ((rc = alloc_grid(s))) return rc;
  if ((rc = splat(s))) return rc;
  memcpy(s->field, s->field0, sizeof(float) * (size_t)s->nvert);
  smooth(s);
  return contour(s);
}

/* ---------------------------------------------------------------- accessors */

void sro_info(const sro *s, int64_t *out) {
  /* glo[3], gdim[3], nverts, ntris, active blocks, vlo[3], vdim[3] */
  for (int a = 0; a < 3; a++) {
    out[a] = s->glo[a];
    out[3 + a] = s->gdim[a];
    out[9 + a] = s->vlo[a];
    out[12 + a] = s->vdim[a];
  }
  out[6] = s->nverts;
  out[7] = s->ntris;
  long na = 0;
  for (long b = 0; b < s->nblock; b++) na += s->active[b];
  out[8] = na;
}
double sro_sd(const sro *s) { return (double)s->sd; }
void sro_get_field(const sro *s, int which, float *out) {
  memcpy(out, which == 0 ? s->field0 : s->field, sizeof(float) * (size_t)s->nvert);
}
void sro_get_active(const sro *s, unsigned char *out) { memcpy(out, s->active, (size_t)s->nblock); }
void sro_get_aniso(const sro *s, float *G, float *xbar) {
  memcpy(G, s->G, sizeof(float) * 9 * (size_t)s->n);
  memcpy(xbar, s->xbar, sizeof(float) * 3 * (size_t)s->n);
}
void sro_get_velocity(const sro *s, float *vfield, float *mesh_vel) {
  if (vfield) {
    if (s->vfield) memcpy(vfield, s->vfield, sizeof(float) * 3 * (size_t)s->nvert);
    else memset(vfield, 0, sizeof(float) * 3 * (size_t)s->nvert);
  }
  if (mesh_vel && s->mvel && s->nverts) memcpy(mesh_vel, s->mvel, sizeof(float) * 3 * (size_t)s->nverts);
}

void sro_get_mesh(const sro *s, float *verts, int *faces, float *normals) {
  if (s->nverts) memcpy(verts, s->verts, sizeof(float) * 3 * (size_t)s->nverts);
  if (s->ntris) memcpy(faces, s->faces, sizeof(int) * 3 * (size_t)s->ntris);
  if (normals && s->normals && s->nverts) memcpy(normals, s->normals, sizeof(float) * 3 * (size_t)s->nverts);
}
