/*
 * sph_oracle.c -- CPU restatement of the SPH step pipeline of chunleili/nonNewton
 * (SPH/main.py, SPHBase.step :915-940 and everything it calls).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
 * (nonnewton_amd/) never links, calls or falls back to it.
 *
 * Build: oracle/build.sh compiles this file twice, `real` = float
 * (liboracle_f32.so) and `real` = double (-DORACLE_F64, liboracle_f64.so),
 * with -ffp-contract=off so every expression rounds as written.
 *
 * Conventions of this restatement (each one is also stated where it is used):
 *  - Particle order is the caller's (original) order, as in the reference
 *    (particle_id[i] = i, :729-732).
 *  - Neighbor lists are built in the reference's serial order: cells are filled
 *    in increasing particle index (:815-820 with one thread), the 27 cells are
 *    visited in ti.ndrange order, first offset slowest (:840), and every list is
 *    capped at 60 and every cell at 50 (:777-778); exceeding a cap is reported
 *    as a status bit instead of writing out of bounds.
 *  - The distance test is the reference's FP32 `(x_i - x_j).norm() < h` (:845)
 *    in both builds, so the FP64 build sees the same neighbor sets.
 *  - Constants that the reference derives in Python double (h, kernel factors,
 *    clamp planes, 1/dt, ...) are derived here in double and rounded once.
 *  - (1-q)^3 in the cubic kernel is evaluated as t*t*t (the reference calls
 *    ti.pow(1 - q, 3.0), :879).
 *  - The solver-termination reductions (compute_density_error, :1669-1675) are
 *    accumulated in FP64 in index order; the reference reduces in FP32 in an
 *    unspecified parallel order.
 *  - Reactions on dynamic solids (:1755-1760, :1835-1841, :1458-1460) are added
 *    per solid particle over its own neighbor list, in list order.  The reference
 *    scatters them with atomics from the fluid side, whose order is unspecified;
 *    each added term is identical.
 *  - Parts that the reference leaves dead or stubbed (non-Newtonian viscosity
 *    feeding a force, Weiler 2018 implicit viscosity, :332-492, :1471-1513) are
 *    restated from the published methods as documented at each function.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_F64
typedef double real;
#define RSQRT(x) sqrt(x)
#define RFABS(x) fabs(x)
#define RPOW(a, b) pow(a, b)
#else
typedef float real;
#define RSQRT(x) sqrtf(x)
#define RFABS(x) fabsf(x)
#define RPOW(a, b) powf(a, b)
#endif

#define MAX_NB 60   /* NeighborhoodSearchSparse.max_num_neighbors, :777 */
#define MAX_PIC 50  /* max_num_particles_in_grid, :778 */
#define MAX_REST 50 /* Elasticity.max_num_neighbors, :1192 */

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct or_sim {
    or_params p;
    int n, nf, ne;
    /* derived constants, rounded once from double (Parameter :598-621) */
    real h, kW, kG, W0, dt, inv_dt, inv_dt2, rho0, mass, vol0, nu, bnu;
    real g[3], lo[3], hi[3];
    real diam2, Wd, visc_eps, sigma, opc, sticky, eps;
    real C_diag, C_off, C_shear;
    float hf;
    int gn[3];
    long ncells;
    /* particle state, original order (ParticleData :701-723) */
    real *x, *v, *a, *vol, *rho, *alpha, *dadv, *x0, *rest_vol;
    int32_t *kind;
    /* grid and lists (NeighborhoodSearchSparse :766-791) */
    int32_t *gcount, *gslot, *nbr, *nnbr;
    int32_t status;
    /* elastic stress set S = elastic particles and their rest neighbors */
    int ns;
    int32_t *s_pid, *s_of, *rest, *rest_cnt;
    real *R, *stress, *Fm;
    /* non-Newtonian */
    real *gam, *eta;
    /* Weiler CG work vectors (fluid particles, 3 per particle) */
    real *wx, *wr, *wz, *wp, *wq, *ws, *wdinv, *wmu, *wpair;
    real *wvd; /* x - v of the previous solve (initial guess with warm_start bit 0) */
    /* rigid bodies: body index of each particle (-1 if none); per body the
     * rest centre of mass and the last projection's centre and rotation */
    int nbodies;
    int32_t *body;
    real *rb_rest_cm, *rb_cm, *rb_R;
    int initialized;
};

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */
/* ------------------------------------------------------------------------ */
static inline int is_fluid(const or_sim *s, int i) { return s->kind[i] == OR_KIND_FLUID; }
static inline int is_static_solid(const or_sim *s, int i) { return s->kind[i] == OR_KIND_STATIC; }
/* dynamic solids (SOLID and is_dynamic, :281-282): elastic and rigid particles */
static inline int is_dyn_solid(const or_sim *s, int i) { return s->kind[i] == OR_KIND_ELASTIC || s->kind[i] == OR_KIND_RIGID; }
static inline int is_elastic(const or_sim *s, int i) { return s->kind[i] == OR_KIND_ELASTIC; }

typedef struct { real s, c; } ksum;
/* Kahan (compensated) summation step. */
static inline void kadd(ksum *k, real y)
{
    real t = y - k->c;
    real u = k->s + t;
    k->c = (u - k->s) - t;
    k->s = u;
}

static inline real norm3(const real r[3]) { return RSQRT(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); }
static inline real dot3(const real a[3], const real b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* cubic_kernel, SPH/main.py:859-880: W(q) = k(6q^3 - 6q^2 + 1) for q <= 1/2,
 * 2k(1-q)^3 for 1/2 < q <= 1, k = 8/(pi h^3). */
static inline real kern_w(const or_sim *s, real rn)
{
    real res = 0;
    real q = rn / s->h;
    if (q <= (real)1.0) {
        if (q <= (real)0.5) {
            real q2 = q * q;
            real q3 = q2 * q;
            res = s->kW * ((real)6.0 * q3 - (real)6.0 * q2 + (real)1.0);
        } else {
            real t = (real)1.0 - q;
            res = (s->kW * (real)2.0) * (t * t * t);
        }
    }
    return res;
}

/* cubic_kernel_derivative, SPH/main.py:883-905 (k here is 6 * 8/(pi h^3)). */
static inline void kern_grad(const or_sim *s, const real r[3], real out[3])
{
    real rn = norm3(r);
    real q = rn / s->h;
    out[0] = out[1] = out[2] = 0;
    if (rn > (real)1e-5 && q <= (real)1.0) {
        real den = rn * s->h;
        real f;
        if (q <= (real)0.5) {
            f = (s->kG * q) * ((real)3.0 * q - (real)2.0);
        } else {
            real fa = (real)1.0 - q;
            f = s->kG * ((-fa) * fa);
        }
        for (int c = 0; c < 3; c++) out[c] = f * (r[c] / den);
    }
}

static inline void diff3(const real *a, const real *b, real out[3])
{
    out[0] = a[0] - b[0];
    out[1] = a[1] - b[1];
    out[2] = a[2] - b[2];
}

static void derive(or_sim *s)
{
    const or_params *p = &s->p;
    double r = p->particle_radius, d = 2.0 * r, h = 4.0 * r;
    s->h = (real)h;
    s->hf = (float)h;
    s->kW = (real)(8.0 / M_PI / (h * h * h));
    s->kG = (real)(6.0 * 8.0 / M_PI / (h * h * h));
    s->W0 = kern_w(s, (real)0.0);
    s->dt = (real)p->dt;
    s->inv_dt = (real)(1.0 / (double)s->dt);
    s->inv_dt2 = (real)(1.0 / ((double)s->dt * (double)s->dt));
    s->rho0 = (real)p->density0;
    double vol0 = 0.8 * d * d * d; /* :607 */
    s->vol0 = (real)vol0;
    s->mass = (real)(p->particle_mass > 0 ? p->particle_mass : vol0 * 1000.0); /* :743 */
    s->nu = (real)p->viscosity;
    s->bnu = (real)p->boundary_viscosity;
    for (int k = 0; k < 3; k++) {
        s->g[k] = (real)p->gravity[k];
        s->lo[k] = (real)(p->domain_start[k] + h); /* padding = support radius, :615 */
        s->hi[k] = (real)(p->domain_end[k] - h);
        s->gn[k] = (int)ceil((p->domain_end[k] - p->domain_start[k]) / h); /* :773-775 */
    }
    s->ncells = (long)s->gn[0] * s->gn[1] * s->gn[2];
    s->diam2 = (real)(d * d);
    real dr = (real)d;
    s->Wd = kern_w(s, RSQRT(dr * dr + (real)0.0 * (real)0.0 + (real)0.0 * (real)0.0)); /* :1414 */
    s->visc_eps = (real)(0.01 * h * h);
    s->sigma = (real)p->surface_tension;
    s->opc = (real)(1.0 + p->collision_coeff);
    s->sticky = (real)p->sticky_coeff;
    s->eps = (real)p->eps;
    double E = p->youngs_modulus, nu = p->poisson_ratio;
    double factor = E / ((1.0 + nu) * (1.0 - 2.0 * nu)); /* :1277-1281 */
    s->C_diag = (real)(factor * (1.0 - nu));
    s->C_off = (real)(factor * nu);
    s->C_shear = (real)(factor * 0.5 * (1.0 - 2.0 * nu));
}

/* ------------------------------------------------------------------------ */
/* neighbor search: NeighborhoodSearchSparse.run_search, :822-847            */
/* ------------------------------------------------------------------------ */
static inline int cell_of(const or_sim *s, const real *x, int c3[3])
{
    int ok = 1;
    for (int k = 0; k < 3; k++) {
        c3[k] = (int)((float)x[k] / s->hf); /* pos_to_index, :805-807 (truncation) */
        if (c3[k] < 0 || c3[k] >= s->gn[k]) ok = 0;
    }
    return ok;
}

static inline int within_h(const or_sim *s, const real *xi, const real *xj)
{
    /* (x_i - x_j).norm() < support_radius, FP32 (:845) */
    float dx = (float)xi[0] - (float)xj[0];
    float dy = (float)xi[1] - (float)xj[1];
    float dz = (float)xi[2] - (float)xj[2];
    float ss = dx * dx + dy * dy + dz * dz;
    return sqrtf(ss) < s->hf;
}

int32_t or_search(or_sim *s)
{
    int32_t st = 0;
    memset(s->gcount, 0, sizeof(int32_t) * s->ncells);
    /* update_grid (:815-820) in serial order */
    for (int i = 0; i < s->n; i++) {
        int c3[3];
        if (!cell_of(s, s->x + 3 * i, c3)) {
            st |= OR_ST_OUT_OF_GRID;
            continue;
        }
        long c = ((long)c3[0] * s->gn[1] + c3[1]) * s->gn[2] + c3[2];
        int k = s->gcount[c]++;
        if (k >= MAX_PIC) {
            st |= OR_ST_CELL_OVERFLOW;
            continue;
        }
        s->gslot[c * MAX_PIC + k] = i;
    }
    /* store_neighbors (:836-847) */
#pragma omp parallel for schedule(static) reduction(| : st)
    for (int i = 0; i < s->n; i++) {
        int c3[3];
        int cnt = 0;
        int32_t *out = s->nbr + (long)i * MAX_NB;
        for (int k = 0; k < MAX_NB; k++) out[k] = -1;
        if (!cell_of(s, s->x + 3 * i, c3)) {
            s->nnbr[i] = 0;
            continue;
        }
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    int a = c3[0] + dx, b = c3[1] + dy, e = c3[2] + dz;
                    if (a < 0 || a >= s->gn[0] || b < 0 || b >= s->gn[1] || e < 0 || e >= s->gn[2]) continue;
                    long c = ((long)a * s->gn[1] + b) * s->gn[2] + e;
                    int m = s->gcount[c] < MAX_PIC ? s->gcount[c] : MAX_PIC;
                    for (int k = 0; k < m; k++) {
                        int j = s->gslot[c * MAX_PIC + k];
                        if (j != i && within_h(s, s->x + 3 * i, s->x + 3 * j)) {
                            if (cnt < MAX_NB) out[cnt] = j;
                            else st |= OR_ST_NEIGHBOR_OVERFLOW;
                            cnt++;
                        }
                    }
                }
        s->nnbr[i] = cnt < MAX_NB ? cnt : MAX_NB;
    }
    s->status |= st;
    return st;
}

#define FOR_NB(s, i, k, j)                                                 \
    for (int k = 0, j = 0; k < (s)->nnbr[i] && ((j = (s)->nbr[(long)(i) * MAX_NB + k]), 1); k++)

/* ------------------------------------------------------------------------ */
/* boundary volume, :1010-1035 (V_b = 3 / (W0 + sum_solid W))                */
/* ------------------------------------------------------------------------ */
static void boundary_volume(or_sim *s, int dynamic)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (dynamic ? !is_dyn_solid(s, i) : !is_static_solid(s, i)) continue;
        real delta;
        ksum ks = {s->W0, 0};
        real plain = s->W0;
        FOR_NB(s, i, k, j)
        {
            if (!is_fluid(s, j)) {
                real r[3];
                diff3(s->x + 3 * i, s->x + 3 * j, r);
                real w = kern_w(s, norm3(r));
                if (s->p.compensated) kadd(&ks, w);
                else plain += w;
            }
        }
        delta = s->p.compensated ? ks.s : plain;
        s->vol[i] = (real)1.0 / delta * (real)3.0;
    }
}

/* clear_accelerations_kernel, :945-951 */
static void clear_acc(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++)
        for (int c = 0; c < 3; c++) s->a[3 * i + c] = is_static_solid(s, i) ? (real)0 : s->g[c];
}

/* compute_densities, :962-984 */
static void densities(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real d = s->vol[i] * s->W0;
        ksum ks = {0, 0};
        real den = 0;
        FOR_NB(s, i, k, j)
        {
            real r[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            real t = s->vol[j] * kern_w(s, norm3(r));
            if (s->p.compensated) kadd(&ks, t);
            else den += t;
        }
        d += s->p.compensated ? ks.s : den;
        d *= s->rho0;
        s->rho[i] = d;
    }
}

/* compute_DFSPH_factor, :1569-1608 */
static void dfsph_factor(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real gi[3] = {0, 0, 0}, sk = 0;
        FOR_NB(s, i, k, j)
        {
            real r[3], g[3], gp[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            kern_grad(s, r, g);
            for (int c = 0; c < 3; c++) gp[c] = -s->vol[j] * g[c];
            if (is_fluid(s, j)) sk += dot3(gp, gp);
            for (int c = 0; c < 3; c++) gi[c] -= gp[c];
        }
        sk += dot3(gi, gi);
        s->alpha[i] = sk > (real)1e-6 ? (real)-1.0 / sk : (real)0.0;
    }
}

/* compute_density_change, :1610-1645 */
static void density_change(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real acc = 0;
        int cnt = 0;
        FOR_NB(s, i, k, j)
        {
            real r[3], g[3], dv[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            kern_grad(s, r, g);
            diff3(s->v + 3 * i, s->v + 3 * j, dv);
            acc += s->vol[j] * dot3(dv, g);
            cnt++;
        }
        acc = acc > (real)0.0 ? acc : (real)0.0;
        if (cnt < 20) acc = 0; /* particle deficiency, :1623-1628 */
        s->dadv[i] = acc;
    }
}

/* compute_density_adv, :1647-1667 */
static void density_adv(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real delta = 0;
        FOR_NB(s, i, k, j)
        {
            real r[3], g[3], dv[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            kern_grad(s, r, g);
            diff3(s->v + 3 * i, s->v + 3 * j, dv);
            delta += s->vol[j] * dot3(dv, g);
        }
        real da = s->rho[i] / s->rho0 + s->dt * delta;
        s->dadv[i] = da > (real)1.0 ? da : (real)1.0;
    }
}

/* compute_density_error, :1669-1675, accumulated in FP64 in index order. */
static double density_error(const or_sim *s, double offset)
{
    double e = 0;
    for (int i = 0; i < s->n; i++)
        if (is_fluid(s, i)) e += (double)s->rho0 * (double)s->dadv[i] - offset;
    return e;
}

static void scale_alpha(or_sim *s, real f)
{
    for (int i = 0; i < s->n; i++)
        if (is_fluid(s, i)) s->alpha[i] *= f; /* multiply_time_step, :1677-1681 */
}

/* divergence_solver_iteration_kernel, :1721-1760.  Reactions on dynamic solids
 * are not formed: the second clear_accelerations (:1552) erases them. */
static void div_update(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real ki = s->dadv[i] * s->alpha[i];
        real dv[3] = {0, 0, 0};
        FOR_NB(s, i, k, j)
        {
            real r[3], g[3];
            if (is_fluid(s, j)) {
                real kj = s->dadv[j] * s->alpha[j];
                real ksum = ki + kj;
                if (RFABS(ksum) > s->eps) {
                    diff3(s->x + 3 * i, s->x + 3 * j, r);
                    kern_grad(s, r, g);
                    real f = s->dt * ksum;
                    for (int c = 0; c < 3; c++) dv[c] -= f * (-s->vol[j] * g[c]);
                }
            } else if (RFABS(ki) > s->eps) {
                diff3(s->x + 3 * i, s->x + 3 * j, r);
                kern_grad(s, r, g);
                real f = -s->dt * ki;
                for (int c = 0; c < 3; c++) dv[c] += f * (-s->vol[j] * g[c]);
            }
        }
        for (int c = 0; c < 3; c++) s->v[3 * i + c] += dv[c];
    }
}

/* pressure_solve_iteration_kernel, :1799-1841, with the reaction on dynamic
 * solids (:1835-1841) gathered per solid particle in list order. */
static void press_update(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (is_fluid(s, i)) {
            real ki = (s->dadv[i] - (real)1.0) * s->alpha[i];
            real vi[3] = {s->v[3 * i], s->v[3 * i + 1], s->v[3 * i + 2]};
            FOR_NB(s, i, k, j)
            {
                real r[3], g[3];
                if (is_fluid(s, j)) {
                    real kj = (s->dadv[j] - (real)1.0) * s->alpha[j];
                    real ksum = ki + kj;
                    if (RFABS(ksum) > s->eps) {
                        diff3(s->x + 3 * i, s->x + 3 * j, r);
                        kern_grad(s, r, g);
                        real f = s->dt * ksum;
                        for (int c = 0; c < 3; c++) vi[c] -= f * (-s->vol[j] * g[c]);
                    }
                } else if (RFABS(ki) > s->eps) {
                    diff3(s->x + 3 * i, s->x + 3 * j, r);
                    kern_grad(s, r, g);
                    real f = -s->dt * ki;
                    for (int c = 0; c < 3; c++) vi[c] += f * (-s->vol[j] * g[c]);
                }
            }
            for (int c = 0; c < 3; c++) s->v[3 * i + c] = vi[c];
        } else if (is_dyn_solid(s, i)) {
            real *aj = s->a + 3 * i;
            FOR_NB(s, i, k, f)
            {
                if (!is_fluid(s, f)) continue;
                real kf = (s->dadv[f] - (real)1.0) * s->alpha[f];
                if (RFABS(kf) > s->eps) {
                    real r[3], g[3];
                    diff3(s->x + 3 * f, s->x + 3 * i, r);
                    kern_grad(s, r, g);
                    real fk = -s->dt * kf;
                    for (int c = 0; c < 3; c++) {
                        real vc = fk * (-s->vol[i] * g[c]);
                        aj[c] += -vc * (real)1.0 / s->dt;
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* non-pressure forces, :1557-1567                                           */
/* ------------------------------------------------------------------------ */
/* SurfaceTension, :1378-1414 */
static void surface_tension(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real dv[3] = {0, 0, 0};
        real sm = s->sigma / s->mass * s->mass;
        FOR_NB(s, i, k, j)
        {
            if (!is_fluid(s, j)) continue;
            real r[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            real r2 = dot3(r, r);
            real w = r2 > s->diam2 ? kern_w(s, norm3(r)) : s->Wd;
            for (int c = 0; c < 3; c++) dv[c] -= sm * r[c] * w;
        }
        for (int c = 0; c < 3; c++) s->a[3 * i + c] += dv[c];
    }
}

/* ViscosityModel, :1417-1468 (explicit).  The boundary term and its reaction on
 * dynamic solids are active only when boundaryViscosity != 0 (:1446-1460). */
static void viscosity_explicit(or_sim *s)
{
    const real d10 = (real)10.0 * s->nu;
    const real b10 = (real)10.0 * s->bnu;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (is_fluid(s, i)) {
            real dv[3] = {0, 0, 0};
            FOR_NB(s, i, k, j)
            {
                real r[3], g[3], vij[3];
                diff3(s->x + 3 * i, s->x + 3 * j, r);
                diff3(s->v + 3 * i, s->v + 3 * j, vij);
                real vxy = dot3(vij, r);
                if (is_fluid(s, j)) {
                    real rn = norm3(r);
                    kern_grad(s, r, g);
                    real f = d10 * (s->mass / s->rho[j]) * vxy / (rn * rn + s->visc_eps);
                    for (int c = 0; c < 3; c++) dv[c] += f * g[c];
                } else if (s->bnu != (real)0.0) {
                    real rn = norm3(r);
                    kern_grad(s, r, g);
                    real f = b10 * (s->rho0 * s->vol[j] / s->rho[i]) * vxy / (rn * rn + s->visc_eps);
                    for (int c = 0; c < 3; c++) dv[c] += f * g[c];
                }
            }
            for (int c = 0; c < 3; c++) s->a[3 * i + c] += dv[c];
        } else if (is_dyn_solid(s, i) && s->bnu != (real)0.0) {
            FOR_NB(s, i, k, f)
            {
                if (!is_fluid(s, f)) continue;
                real r[3], g[3], vij[3];
                diff3(s->x + 3 * f, s->x + 3 * i, r);
                diff3(s->v + 3 * f, s->v + 3 * i, vij);
                real vxy = dot3(vij, r);
                real rn = norm3(r);
                kern_grad(s, r, g);
                real ff = b10 * (s->rho0 * s->vol[i] / s->rho[f]) * vxy / (rn * rn + s->visc_eps);
                for (int c = 0; c < 3; c++) s->a[3 * i + c] += -(ff * g[c]);
            }
        }
    }
}

/* NonNewton viscosity models, :404-459, given the strain-rate norm.  The result
 * is a kinematic viscosity (the unit of the `viscosity` key, :621); it is
 * clamped to [0, max_viscosity] because several models diverge at zero strain
 * rate (power law with n < 1, Casson). */
static real nn_model(const or_params *p, real gam)
{
    real eta;
    real n = (real)p->power_index, K = (real)p->consistency_index;
    real e0 = (real)p->viscosity0, einf = (real)p->viscosity_inf;
    real gc = (real)p->critical_strain_rate;
    switch (p->non_newton_method) {
    case 1: /* power law */
        eta = K * RPOW(gam, n - (real)1.0);
        break;
    case 2: /* Cross */
        eta = einf + (e0 - einf) / ((real)1.0 + RPOW(K * gam, n));
        break;
    case 3: { /* Casson */
        real t = RSQRT((real)p->mu_c) + RSQRT((real)p->yield_stress / gam);
        eta = t * t;
        break;
    }
    case 4: /* Carreau */
        eta = einf + (e0 - einf) / ((real)1.0 + RPOW(K * gam * gam, ((real)1.0 - n) / (real)2.0));
        break;
    case 5: /* Bingham */
        if (gam < gc) eta = e0;
        else eta = gc * (e0 - einf) / gam + einf;
        break;
    case 6: /* Herschel-Bulkley */
        if (gam < gc) eta = e0;
        else {
            real tau0 = e0 * gc - K * RPOW(gc, n);
            eta = tau0 / gam + K * RPOW(gam, n - (real)1.0);
        }
        break;
    default: /* Newtonian */
        eta = e0;
        break;
    }
    real mx = (real)p->max_viscosity;
    if (!(eta >= (real)0.0)) eta = eta != eta ? mx : (real)0.0; /* NaN -> max, negative -> 0 */
    if (eta > mx) eta = mx;
    return eta;
}

/* NonNewton.calcStrainRate, :461-492, with the accumulation the reference omits
 * (it overwrites the tensor with the last neighbor's term):
 *   D_i = (0.5 / rho_i) * sum_j m_j (v_ji (x) gradW_ij + gradW_ij (x) v_ji),
 * strain-rate norm = ||D_i||_F. */
static void strain_rate(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        FOR_NB(s, i, k, j)
        {
            real r[3], g[3], vji[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            kern_grad(s, r, g);
            diff3(s->v + 3 * j, s->v + 3 * i, vji);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) S[3 * a + b] += s->mass * (vji[a] * g[b] + vji[b] * g[a]);
        }
        real f = (real)0.5 / s->rho[i];
        real n2 = 0;
        for (int q = 0; q < 9; q++) {
            real d = f * S[q];
            n2 += d * d;
        }
        s->gam[i] = RSQRT(n2);
        s->eta[i] = nn_model(&s->p, s->gam[i]);
    }
}

/* Weiler et al. 2018, "A Physically Consistent Implicit Viscosity Solver for
 * SPH Fluids" (the reference's Viscosity_Weiler2018 is a stub, :1471-1513).
 * Unknown: the fluid velocity after viscosity, x.  Operator:
 *   (A x)_i = x_i - (dt/rho_i) sum_{fluid j} c_ij ((x_i - x_j) . r_ij) gradW_ij,
 *   c_ij = 10 mu_ij (m_j/rho_j) / (|r_ij|^2 + 0.01 h^2),  mu_ij = (mu_i + mu_j)/2,
 * mu_i = rho0 * (eta_i if the non-Newtonian model is on, else viscosity).
 * gradW_ij = F(|r|) r_ij / (|r| h) is parallel to r_ij, so each pair enters
 * through one scalar s_ij = c_ij F / (|r| h), formed once per solve:
 *   (A x)_i = x_i - (dt/rho_i) sum_j (s_ij ((x_i - x_j) . r_ij)) r_ij,
 * and the 3x3 diagonal block is I - (dt/rho_i) sum_j s_ij r_ij (x) r_ij.
 * Sums over a neighbor list are formed as four partial sums over the list
 * positions k = 0, 1, 2, 3 mod 4, combined as (S0 + S1) + (S2 + S3).
 * A is symmetric positive definite for a uniform particle mass.  RHS b = v.
 * Solved by conjugate gradients with the block-Jacobi preconditioner; dot
 * products and norms in FP64; stop when ||r|| <= tol ||b||. */
static inline real mu_of(const or_sim *s, int i)
{
    return s->p.use_non_newtonian ? s->eta[i] * s->rho0 : s->nu * s->rho0;
}

static inline real weiler_pair(const or_sim *s, const real r[3], real mui, real muj, real rhoj)
{
    real rn = norm3(r);
    real q = rn / s->h;
    if (!(rn > (real)1e-5 && q <= (real)1.0)) return (real)0.0;
    real F;
    if (q <= (real)0.5) {
        F = (s->kG * q) * ((real)3.0 * q - (real)2.0);
    } else {
        real fa = (real)1.0 - q;
        F = s->kG * ((-fa) * fa);
    }
    real den = rn * s->h;
    real muij = (real)0.5 * (mui + muj);
    real c = (real)10.0 * muij * (s->mass / rhoj) / (dot3(r, r) + s->visc_eps);
    return c * F / den;
}

static void weiler_apply(or_sim *s, const real *in, real *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real acc[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        FOR_NB(s, i, k, j)
        {
            if (!is_fluid(s, j)) continue;
            real r[3], pd[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            diff3(in + 3 * i, in + 3 * j, pd);
            real t = s->wpair[(long)i * MAX_NB + k] * dot3(pd, r);
            for (int q = 0; q < 3; q++) acc[k & 3][q] += t * r[q];
        }
        real f = s->dt / s->rho[i];
        for (int q = 0; q < 3; q++) {
            real a = (acc[0][q] + acc[1][q]) + (acc[2][q] + acc[3][q]);
            out[3 * i + q] = in[3 * i + q] - f * a;
        }
    }
}

/* per-pair scalars s_ij and the inverse diagonal blocks */
static void weiler_precond(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        real m4[4][6];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 6; b++) m4[a][b] = 0; /* 00 01 02 11 12 22 */
        real mui = s->wmu[i];
        FOR_NB(s, i, k, j)
        {
            real *sp = s->wpair + (long)i * MAX_NB + k;
            *sp = 0;
            if (!is_fluid(s, j)) continue;
            real r[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            real sij = weiler_pair(s, r, mui, s->wmu[j], s->rho[j]);
            *sp = sij;
            real *m = m4[k & 3];
            m[0] += sij * r[0] * r[0];
            m[1] += sij * r[0] * r[1];
            m[2] += sij * r[0] * r[2];
            m[3] += sij * r[1] * r[1];
            m[4] += sij * r[1] * r[2];
            m[5] += sij * r[2] * r[2];
        }
        real m[6];
        for (int b = 0; b < 6; b++) m[b] = (m4[0][b] + m4[1][b]) + (m4[2][b] + m4[3][b]);
        real f = s->dt / s->rho[i];
        real d00 = (real)1.0 - f * m[0], d01 = -f * m[1], d02 = -f * m[2];
        real d11 = (real)1.0 - f * m[3], d12 = -f * m[4], d22 = (real)1.0 - f * m[5];
        real c00 = d11 * d22 - d12 * d12, c01 = d02 * d12 - d01 * d22, c02 = d01 * d12 - d02 * d11;
        real c11 = d00 * d22 - d02 * d02, c12 = d01 * d02 - d00 * d12, c22 = d00 * d11 - d01 * d01;
        real det = d00 * c00 + d01 * c01 + d02 * c02;
        real *o = s->wdinv + 6 * i;
        if (det > (real)0.0 && det == det) {
            real id = (real)1.0 / det;
            o[0] = c00 * id; o[1] = c01 * id; o[2] = c02 * id;
            o[3] = c11 * id; o[4] = c12 * id; o[5] = c22 * id;
        } else {
            o[0] = 1; o[1] = 0; o[2] = 0; o[3] = 1; o[4] = 0; o[5] = 1;
        }
    }
}

static void weiler_precond_apply(or_sim *s, const real *in, real *out)
{
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        const real *m = s->wdinv + 6 * i;
        const real *a = in + 3 * i;
        out[3 * i + 0] = m[0] * a[0] + m[1] * a[1] + m[2] * a[2];
        out[3 * i + 1] = m[1] * a[0] + m[3] * a[1] + m[4] * a[2];
        out[3 * i + 2] = m[2] * a[0] + m[4] * a[1] + m[5] * a[2];
    }
}

static double fdot(const or_sim *s, const real *a, const real *b)
{
    double acc = 0;
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        for (int q = 0; q < 3; q++) acc += (double)a[3 * i + q] * (double)b[3 * i + q];
    }
    return acc;
}

static void weiler_solve(or_sim *s, or_stats *st)
{
    int n3 = 3 * s->n;
    for (int i = 0; i < s->n; i++) s->wmu[i] = is_fluid(s, i) ? mu_of(s, i) : (real)0.0;
    /* b = v; x0 = b, or with warm_start bit 0 x0 = b + theta e, e = (x - v) of
     * the previous solve (the guess of SPlisHSPlasH's Viscosity_Weiler2018,
     * whose m_vDiff fields the reference declares, SPH/main.py:1486-1487),
     * scaled by theta = (e, r0) / (e, A e), r0 = b - A b: the theta that
     * minimizes the A-norm error along e, so the start is never worse than
     * x0 = b in that norm.  r = r0 - theta A e. */
    memcpy(s->wx, s->v, sizeof(real) * n3);
    const int warm = s->p.warm_start & 1;
    weiler_precond(s);
    weiler_apply(s, s->v, s->wq);
    for (int i = 0; i < n3; i++) s->wr[i] = s->v[i] - s->wq[i];
    if (warm) {
        weiler_apply(s, s->wvd, s->ws);
        double num = fdot(s, s->wvd, s->wr), den = fdot(s, s->wvd, s->ws);
        real th = (real)(den > 0.0 ? num / den : 0.0);
        for (int i = 0; i < s->n; i++) {
            if (!is_fluid(s, i)) continue;
            for (int q = 0; q < 3; q++) {
                int k = 3 * i + q;
                s->wx[k] = s->v[k] + th * s->wvd[k];
                s->wr[k] = s->wr[k] - th * s->ws[k];
            }
        }
    }
    double bb = fdot(s, s->v, s->v);
    double rr = fdot(s, s->wr, s->wr);
    double tol2 = s->p.weiler_tol * s->p.weiler_tol * bb;
    int it = 0;
    /* Preconditioned CG in the Chronopoulos-Gear arrangement (one reduction of
     * three dot products per iteration); equal to textbook PCG in exact
     * arithmetic for the symmetric operator:
     *   u = M r, w = A u, gamma = (r, u), delta = (u, w)
     *   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
     *   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s
     * wz holds u, wq holds w, ws holds s = A p. */
    if (bb > 0.0 && rr > tol2) {
        weiler_precond_apply(s, s->wr, s->wz);
        double ga = fdot(s, s->wr, s->wz);
        weiler_apply(s, s->wz, s->wq);
        double de = fdot(s, s->wz, s->wq);
        double al = ga / de, be = 0.0;
        for (it = 1; it <= s->p.weiler_max_iter; it++) {
            real alf = (real)al, bef = (real)be;
            for (int i = 0; i < s->n; i++) {
                if (!is_fluid(s, i)) continue;
                for (int q = 0; q < 3; q++) {
                    int k = 3 * i + q;
                    if (it == 1) {
                        s->wp[k] = s->wz[k];
                        s->ws[k] = s->wq[k];
                    } else {
                        s->wp[k] = s->wz[k] + bef * s->wp[k];
                        s->ws[k] = s->wq[k] + bef * s->ws[k];
                    }
                    s->wx[k] += alf * s->wp[k];
                    s->wr[k] -= alf * s->ws[k];
                }
            }
            weiler_precond_apply(s, s->wr, s->wz);
            rr = fdot(s, s->wr, s->wr);
            if (rr <= tol2) break;
            double gn = fdot(s, s->wr, s->wz);
            weiler_apply(s, s->wz, s->wq);
            de = fdot(s, s->wz, s->wq);
            be = gn / ga;
            al = gn / (de - be * gn / al);
            ga = gn;
        }
        if (it > s->p.weiler_max_iter) it = s->p.weiler_max_iter;
    }
    st->iterations_cg = it;
    st->cg_rel_residual = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    for (int i = 0; i < s->n; i++) {
        if (!is_fluid(s, i)) continue;
        for (int q = 0; q < 3; q++) {
            real dv = s->wx[3 * i + q] - s->v[3 * i + q];
            s->a[3 * i + q] += dv / s->dt;
            if (warm) s->wvd[3 * i + q] = dv;
        }
    }
}

/* non_pressure_advect, :1843-1848 */
static void np_advect(or_sim *s)
{
    for (int i = 0; i < s->n; i++)
        if (is_fluid(s, i))
            for (int c = 0; c < 3; c++) s->v[3 * i + c] += s->dt * s->a[3 * i + c];
}

/* ------------------------------------------------------------------------ */
/* Becker 2009 corotated elasticity, :1182-1375                               */
/* ------------------------------------------------------------------------ */
/* Polar decomposition A = R S with R a proper rotation (det +1), the factor
 * that an SVD A = U diag(s) V^T with det U = det V = +1 yields as R = U V^T
 * (what ti.polar_decompose returns, :1273).  Method: cyclic Jacobi
 * eigen-decomposition of A^T A (Golub & Van Loan, Matrix Computations, 4th ed.,
 * Alg. 8.5.3), columns sorted by decreasing eigenvalue, then a Givens QR of A V
 * (the structure of McAdams et al. 2011, "Computing the SVD of 3x3 matrices
 * with minimal branching and elementary floating point operations"). */
static void polar3(const real A[9], real R[9])
{
    real S[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S[3 * a + b] = A[a] * A[b] + A[3 + a] * A[3 + b] + A[6 + a] * A[6 + b];
    real V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 6; sweep++) {
        for (int t = 0; t < 3; t++) {
            int p = PQ[t][0], q = PQ[t][1];
            real apq = S[3 * p + q];
            if (apq == (real)0.0) continue;
            real theta = (S[3 * q + q] - S[3 * p + p]) / ((real)2.0 * apq);
            real tt = (theta >= (real)0.0 ? (real)1.0 : (real)-1.0) / (RFABS(theta) + RSQRT(theta * theta + (real)1.0));
            real c = (real)1.0 / RSQRT(tt * tt + (real)1.0);
            real sn = tt * c;
            /* S <- J^T S J, J = identity with J_pp = J_qq = c, J_pq = s, J_qp = -s */
            for (int k = 0; k < 3; k++) { /* columns */
                real skp = S[3 * k + p], skq = S[3 * k + q];
                S[3 * k + p] = c * skp - sn * skq;
                S[3 * k + q] = sn * skp + c * skq;
            }
            for (int k = 0; k < 3; k++) { /* rows */
                real spk = S[3 * p + k], sqk = S[3 * q + k];
                S[3 * p + k] = c * spk - sn * sqk;
                S[3 * q + k] = sn * spk + c * sqk;
            }
            for (int k = 0; k < 3; k++) {
                real vkp = V[3 * k + p], vkq = V[3 * k + q];
                V[3 * k + p] = c * vkp - sn * vkq;
                V[3 * k + q] = sn * vkp + c * vkq;
            }
        }
    }
    /* sort eigenvalues descending; each swap negates one column so det V stays +1 */
    real ev[3] = {S[0], S[4], S[8]};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2 - a; b++)
            if (ev[b] < ev[b + 1]) {
                real t = ev[b];
                ev[b] = ev[b + 1];
                ev[b + 1] = t;
                for (int k = 0; k < 3; k++) {
                    real x = V[3 * k + b];
                    V[3 * k + b] = V[3 * k + b + 1];
                    V[3 * k + b + 1] = -x;
                }
            }
    /* B = A V */
    real B[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) B[3 * a + b] = A[3 * a] * V[b] + A[3 * a + 1] * V[3 + b] + A[3 * a + 2] * V[6 + b];
    /* Givens QR: zero B10, B20, B21; U accumulates the transposed rotations */
    real U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    static const int GR[3][3] = {{0, 1, 0}, {0, 2, 0}, {1, 2, 1}}; /* rows p,q, column */
    for (int t = 0; t < 3; t++) {
        int p = GR[t][0], q = GR[t][1], col = GR[t][2];
        real x = B[3 * p + col], y = B[3 * q + col];
        real rn = RSQRT(x * x + y * y);
        real c = (real)1.0, sn = (real)0.0;
        if (rn > (real)0.0) {
            c = x / rn;
            sn = y / rn;
        }
        for (int k = 0; k < 3; k++) {
            real bp = B[3 * p + k], bq = B[3 * q + k];
            B[3 * p + k] = c * bp + sn * bq;
            B[3 * q + k] = -sn * bp + c * bq;
        }
        for (int k = 0; k < 3; k++) {
            real up = U[3 * k + p], uq = U[3 * k + q];
            U[3 * k + p] = c * up + sn * uq;
            U[3 * k + q] = -sn * up + c * uq;
        }
    }
    /* R = U V^T */
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) R[3 * a + b] = U[3 * a] * V[3 * b] + U[3 * a + 1] * V[3 * b + 1] + U[3 * a + 2] * V[3 * b + 2];
}

/* Elasticity.initValues_kernel, :1220-1236: rest lists (capped at 50) for the
 * stress set, rest volumes V0_i = m_i / (m_i W0 + sum_j m_j W_ij) for all. */
static void elastic_init(or_sim *s)
{
    memcpy(s->x0, s->x, sizeof(real) * 3 * s->n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        real den = s->mass * s->W0;
        FOR_NB(s, i, k, j)
        {
            real r[3];
            diff3(s->x + 3 * i, s->x + 3 * j, r);
            den += s->mass * kern_w(s, norm3(r));
        }
        s->rest_vol[i] = s->mass / den;
    }
    /* stress set S: elastic particles and every rest neighbor of one */
    char *mark = (char *)calloc(s->n, 1);
    for (int i = 0; i < s->n; i++) {
        if (!is_elastic(s, i)) continue;
        mark[i] = 1;
        FOR_NB(s, i, k, j) mark[j] = 1;
    }
    s->ns = 0;
    for (int i = 0; i < s->n; i++) s->s_of[i] = mark[i] ? s->ns++ : -1;
    free(mark);
    free(s->s_pid);
    free(s->rest);
    free(s->rest_cnt);
    free(s->R);
    free(s->stress);
    free(s->Fm);
    int ns = s->ns > 0 ? s->ns : 1;
    s->s_pid = (int32_t *)malloc(sizeof(int32_t) * ns);
    s->rest = (int32_t *)malloc(sizeof(int32_t) * ns * MAX_REST);
    s->rest_cnt = (int32_t *)malloc(sizeof(int32_t) * ns);
    s->R = (real *)calloc((size_t)ns * 9, sizeof(real));
    s->stress = (real *)calloc((size_t)ns * 6, sizeof(real));
    s->Fm = (real *)calloc((size_t)ns * 9, sizeof(real));
    for (int i = 0; i < s->n; i++) {
        int q = s->s_of[i];
        if (q < 0) continue;
        s->s_pid[q] = i;
        int cnt = s->nnbr[i];
        if (cnt > MAX_REST) {
            s->status |= OR_ST_REST_OVERFLOW;
            cnt = MAX_REST;
        }
        s->rest_cnt[q] = cnt;
        for (int k = 0; k < cnt; k++) s->rest[(long)q * MAX_REST + k] = s->nbr[(long)i * MAX_NB + k];
        s->R[9 * q] = s->R[9 * q + 4] = s->R[9 * q + 8] = 1; /* fill_rotations, :1211-1214 */
    }
}

/* computeRotations_kernel (:1249-1274) and computeStress_kernel (:1285-1323) */
static void elastic_rot_stress(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int q = 0; q < s->ns; q++) {
        int i = s->s_pid[q];
        const real *xi = s->x + 3 * i, *xi0 = s->x0 + 3 * i;
        real A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int32_t *rl = s->rest + (long)q * MAX_REST;
        for (int k = 0; k < s->rest_cnt[q]; k++) {
            int j = rl[k];
            real d[3], d0[3];
            diff3(s->x + 3 * j, xi, d);
            diff3(s->x0 + 3 * j, xi0, d0);
            real w = s->mass * kern_w(s, norm3(d0));
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) A[3 * a + b] += w * (d[a] * d0[b]);
        }
        real *R = s->R + 9 * q;
        polar3(A, R);
        real U[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < s->rest_cnt[q]; k++) {
            int j = rl[k];
            real d[3], d0[3], g[3], u[3];
            diff3(s->x + 3 * j, xi, d);
            diff3(s->x0 + 3 * j, xi0, d0);
            for (int a = 0; a < 3; a++) u[a] = R[a] * d[0] + R[3 + a] * d[1] + R[6 + a] * d[2] - d0[a];
            kern_grad(s, d0, g);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) U[3 * a + b] -= (s->rest_vol[j] * u[a]) * g[b];
        }
        real *F = s->Fm + 9 * q;
        for (int a = 0; a < 9; a++) F[a] = U[a] + ((a % 4) == 0 ? (real)1.0 : (real)0.0);
        real e[6] = {U[0], U[4], U[8], (real)0.5 * (U[1] + U[3]), (real)0.5 * (U[2] + U[6]), (real)0.5 * (U[5] + U[7])};
        real C[6][6] = {{s->C_diag, s->C_off, s->C_off, 0, 0, 0},
                        {s->C_off, s->C_diag, s->C_off, 0, 0, 0},
                        {s->C_off, s->C_off, s->C_diag, 0, 0, 0},
                        {0, 0, 0, s->C_shear, 0, 0},
                        {0, 0, 0, 0, s->C_shear, 0},
                        {0, 0, 0, 0, 0, s->C_shear}};
        real *sg = s->stress + 6 * q;
        for (int a = 0; a < 6; a++) {
            real t = 0;
            for (int b = 0; b < 6; b++) t += C[a][b] * e[b];
            sg[a] = t;
        }
    }
}

static inline void sym_mat_vec(const real *M, const real v[3], real out[3])
{
    out[0] = M[0] * v[0] + M[3] * v[1] + M[4] * v[2];
    out[1] = M[3] * v[0] + M[1] * v[1] + M[5] * v[2];
    out[2] = M[4] * v[0] + M[5] * v[1] + M[2] * v[2];
}

/* computeForces_kernel, :1334-1375, for elastic particles (the forces the
 * reference also forms for other particles are never read). */
static void elastic_forces(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int q = 0; q < s->ns; q++) {
        int i = s->s_pid[q];
        if (!is_elastic(s, i)) continue;
        const real *xi0 = s->x0 + 3 * i;
        const int32_t *rl = s->rest + (long)q * MAX_REST;
        const real *Ri = s->R + 9 * q, *si = s->stress + 6 * q;
        real vi0 = s->rest_vol[i];
        real fi[3] = {0, 0, 0};
        for (int k = 0; k < s->rest_cnt[q]; k++) {
            int j = rl[k];
            int qj = s->s_of[j];
            const real *Rj = s->R + 9 * qj, *sj = s->stress + 6 * qj;
            real d0[3], g[3], dji[3], dij[3], sdji[3], sdij[3], fij[3], fji[3];
            diff3(s->x0 + 3 * j, xi0, d0);
            kern_grad(s, d0, g);
            real vj0 = s->rest_vol[j];
            for (int c = 0; c < 3; c++) {
                dji[c] = vi0 * g[c];
                dij[c] = -vj0 * g[c];
            }
            sym_mat_vec(sj, dji, sdji);
            sym_mat_vec(si, dij, sdij);
            for (int c = 0; c < 3; c++) {
                fij[c] = -vj0 * sdji[c];
                fji[c] = -vi0 * sdij[c];
            }
            for (int a = 0; a < 3; a++) {
                real t1 = Rj[3 * a] * fij[0] + Rj[3 * a + 1] * fij[1] + Rj[3 * a + 2] * fij[2];
                real t2 = Ri[3 * a] * fji[0] + Ri[3 * a + 1] * fji[1] + Ri[3 * a + 2] * fji[2];
                fi[a] += t1 - t2;
            }
        }
        for (int c = 0; c < 3; c++) {
            fi[c] = (real)0.5 * fi[c];
            s->a[3 * i + c] += fi[c] / s->mass;
        }
    }
}

/* advect, :1086-1093 */
static void advect(or_sim *s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; i++) {
        if (is_static_solid(s, i)) continue;
        if (is_dyn_solid(s, i))
            for (int c = 0; c < 3; c++) s->v[3 * i + c] += s->dt * s->a[3 * i + c];
        for (int c = 0; c < 3; c++) s->x[3 * i + c] += s->dt * s->v[3 * i + c];
    }
}

/* enforce_boundary_3D + simulate_collisions, :1037-1080, for FLUID then SOLID
 * dynamic particles (each particle is touched once, so one pass). */
static int32_t clamp_boundary(or_sim *s)
{
    int32_t st = 0;
#pragma omp parallel for schedule(static) reduction(| : st)
    for (int i = 0; i < s->n; i++) {
        if (is_static_solid(s, i)) continue;
        real *x = s->x + 3 * i, *v = s->v + 3 * i;
        real pos[3] = {x[0], x[1], x[2]};
        real nrm[3] = {0, 0, 0};
        for (int c = 0; c < 3; c++) {
            if (pos[c] > s->hi[c]) {
                nrm[c] += (real)1.0;
                x[c] = s->hi[c];
            }
            if (pos[c] <= s->lo[c]) {
                nrm[c] += (real)-1.0;
                x[c] = s->lo[c];
            }
        }
        real nl = norm3(nrm);
        if (nl > (real)1e-6) {
            real nh[3] = {nrm[0] / nl, nrm[1] / nl, nrm[2] / nl};
            real f = s->opc * dot3(v, nh);
            for (int c = 0; c < 3; c++) v[c] -= f * nh[c];
            for (int c = 0; c < 3; c++) v[c] *= s->sticky;
        }
        for (int c = 0; c < 3; c++)
            if (!isfinite((double)x[c]) || !isfinite((double)v[c])) st |= OR_ST_NAN;
    }
    s->status |= st;
    return st;
}

/* ------------------------------------------------------------------------ */
/* Rigid bodies: RigidBodySolver, :1099-1158                                  */
/* ------------------------------------------------------------------------ */
/* compute_com, :1120-1130: mass-weighted centre of body b, mass = volume0 *
 * density (a solid's density keeps its initial density0, :744; only fluid
 * densities are recomputed).  The reference adds the FP32 terms m x and m with
 * atomics in an unspecified order; here they are added in FP64 in index order
 * and the quotient is rounded once. */
static void rigid_com(const or_sim *s, int b, real cm[3])
{
    double sm = 0.0, sx[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < s->n; i++) {
        if (s->body[i] != b) continue;
        real m = s->vol0 * s->rho[i];
        sm += (double)m;
        for (int c = 0; c < 3; c++) sx[c] += (double)(m * s->x[3 * i + c]);
    }
    for (int c = 0; c < 3; c++) cm[c] = (real)(sx[c] / sm);
}

/* ground_collision, :1117-1122: every solid particle (static, elastic or
 * rigid, of any body) below y = padding is lifted to y = padding. */
static void rigid_ground(or_sim *s)
{
    for (int i = 0; i < s->n; i++)
        if (!is_fluid(s, i) && s->x[3 * i + 1] < s->h) s->x[3 * i + 1] = s->h;
}

/* solve_constraints, :1138-1158: A = sum m (x - cm)(x0 - cm0)^T (FP64 sums in
 * index order, rounded once), R = polar(A) (identity if every entry is below
 * 1e-6), then x += (cm + R (x0 - cm0) - x) * 1.0. */
static void rigid_project(or_sim *s, int b)
{
    real cm[3];
    rigid_com(s, b, cm);
    const real *c0 = s->rb_rest_cm + 3 * b;
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < s->n; i++) {
        if (s->body[i] != b) continue;
        real m = s->vol0 * s->rho[i];
        real q[3], p[3];
        for (int c = 0; c < 3; c++) {
            q[c] = s->x0[3 * i + c] - c0[c];
            p[c] = s->x[3 * i + c] - cm[c];
        }
        for (int a = 0; a < 3; a++)
            for (int e = 0; e < 3; e++) A[3 * a + e] += (double)(m * (p[a] * q[e]));
    }
    real Af[9], R[9];
    for (int k = 0; k < 9; k++) Af[k] = (real)A[k];
    polar3(Af, R);
    int tiny = 1;
    for (int k = 0; k < 9; k++)
        if (!(RFABS(R[k]) < (real)1e-6)) tiny = 0;
    if (tiny)
        for (int k = 0; k < 9; k++) R[k] = (k % 4) == 0 ? (real)1.0 : (real)0.0;
    for (int i = 0; i < s->n; i++) {
        if (s->body[i] != b) continue;
        real q[3];
        for (int c = 0; c < 3; c++) q[c] = s->x0[3 * i + c] - c0[c];
        for (int a = 0; a < 3; a++) {
            real goal = cm[a] + (R[3 * a] * q[0] + R[3 * a + 1] * q[1] + R[3 * a + 2] * q[2]);
            real corr = (goal - s->x[3 * i + a]) * (real)1.0;
            s->x[3 * i + a] += corr;
        }
    }
    for (int c = 0; c < 3; c++) s->rb_cm[3 * b + c] = cm[c];
    for (int k = 0; k < 9; k++) s->rb_R[9 * b + k] = R[k];
}

/* RigidBodySolver.__init__, :1100-1106: rest centre of mass from the initial
 * positions (x_0 is the copy of x taken at construction, :1106). */
static void rigid_init(or_sim *s)
{
    for (int b = 0; b < s->nbodies; b++) {
        rigid_com(s, b, s->rb_rest_cm + 3 * b);
        for (int c = 0; c < 3; c++) s->rb_cm[3 * b + c] = s->rb_rest_cm[3 * b + c];
        for (int k = 0; k < 9; k++) s->rb_R[9 * b + k] = (k % 4) == 0 ? (real)1.0 : (real)0.0;
    }
}

/* SPHBase.step, :929-931: each body in turn, coupling_interval substeps of
 * ground collision and shape matching. */
static void rigid_step(or_sim *s)
{
    for (int b = 0; b < s->nbodies; b++)
        for (int c = 0; c < s->p.coupling_interval; c++) {
            rigid_ground(s);
            rigid_project(s, b);
        }
}

/* ------------------------------------------------------------------------ */
/* public API                                                               */
/* ------------------------------------------------------------------------ */
or_sim *or_create(const or_params *p, int32_t n, const float *x, const float *v, const int32_t *kind)
{
    if (n <= 0 || !p || !x || !kind) return NULL;
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    s->p = *p;
    s->n = n;
    derive(s);
    if (s->gn[0] <= 0 || s->gn[1] <= 0 || s->gn[2] <= 0) {
        free(s);
        return NULL;
    }
    size_t n3 = (size_t)3 * n;
    s->x = (real *)calloc(n3, sizeof(real));
    s->v = (real *)calloc(n3, sizeof(real));
    s->a = (real *)calloc(n3, sizeof(real));
    s->x0 = (real *)calloc(n3, sizeof(real));
    s->vol = (real *)calloc(n, sizeof(real));
    s->rho = (real *)calloc(n, sizeof(real));
    s->alpha = (real *)calloc(n, sizeof(real));
    s->dadv = (real *)calloc(n, sizeof(real));
    s->rest_vol = (real *)calloc(n, sizeof(real));
    s->gam = (real *)calloc(n, sizeof(real));
    s->eta = (real *)calloc(n, sizeof(real));
    s->wx = (real *)calloc(n3, sizeof(real));
    s->wr = (real *)calloc(n3, sizeof(real));
    s->wz = (real *)calloc(n3, sizeof(real));
    s->wp = (real *)calloc(n3, sizeof(real));
    s->wq = (real *)calloc(n3, sizeof(real));
    s->ws = (real *)calloc(n3, sizeof(real));
    s->wvd = (real *)calloc(n3, sizeof(real));
    s->wdinv = (real *)calloc((size_t)6 * n, sizeof(real));
    s->wmu = (real *)calloc(n, sizeof(real));
    s->wpair = (real *)calloc((size_t)n * MAX_NB, sizeof(real));
    s->kind = (int32_t *)calloc(n, sizeof(int32_t));
    s->s_of = (int32_t *)calloc(n, sizeof(int32_t));
    s->gcount = (int32_t *)calloc(s->ncells, sizeof(int32_t));
    s->gslot = (int32_t *)calloc((size_t)s->ncells * MAX_PIC, sizeof(int32_t));
    s->nbr = (int32_t *)calloc((size_t)n * MAX_NB, sizeof(int32_t));
    s->nnbr = (int32_t *)calloc(n, sizeof(int32_t));
    s->body = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int i = 0; i < n; i++) {
        if (kind[i] < 0 || kind[i] >= OR_KIND_RIGID + OR_MAX_BODIES) {
            or_destroy(s);
            return NULL;
        }
        s->kind[i] = kind[i] >= OR_KIND_RIGID ? OR_KIND_RIGID : kind[i];
        s->body[i] = kind[i] >= OR_KIND_RIGID ? kind[i] - OR_KIND_RIGID : -1;
        if (s->body[i] >= s->nbodies) s->nbodies = s->body[i] + 1;
        if (kind[i] == OR_KIND_FLUID) s->nf++;
        if (kind[i] == OR_KIND_ELASTIC) s->ne++;
        for (int c = 0; c < 3; c++) {
            s->x[3 * i + c] = (real)x[3 * i + c];
            s->v[3 * i + c] = v ? (real)v[3 * i + c] : (real)0.0;
        }
        s->vol[i] = s->vol0;      /* :741 */
        s->rho[i] = s->rho0;      /* :744 */
        s->eta[i] = (real)0.0;
        s->s_of[i] = -1;
    }
    /* every body index below the largest must have particles */
    int nb = s->nbodies > 0 ? s->nbodies : 1;
    int32_t *bc = (int32_t *)calloc(nb, sizeof(int32_t));
    for (int i = 0; i < n; i++)
        if (s->body[i] >= 0) bc[s->body[i]]++;
    for (int b = 0; b < s->nbodies; b++)
        if (bc[b] == 0) {
            free(bc);
            or_destroy(s);
            return NULL;
        }
    free(bc);
    s->rb_rest_cm = (real *)calloc((size_t)3 * nb, sizeof(real));
    s->rb_cm = (real *)calloc((size_t)3 * nb, sizeof(real));
    s->rb_R = (real *)calloc((size_t)9 * nb, sizeof(real));
    return s;
}

void or_destroy(or_sim *s)
{
    if (!s) return;
    void *ptrs[] = {s->x, s->v, s->a, s->x0, s->vol, s->rho, s->alpha, s->dadv, s->rest_vol, s->gam, s->eta,
                    s->wx, s->wr, s->wz, s->wp, s->wq, s->ws, s->wvd, s->wdinv, s->wmu, s->wpair, s->kind, s->s_of, s->gcount, s->gslot,
                    s->nbr, s->nnbr, s->s_pid, s->rest, s->rest_cnt, s->R, s->stress, s->Fm,
                    s->body, s->rb_rest_cm, s->rb_cm, s->rb_R};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    free(s);
}

/* RigidBodySolver.__init__ (:1100-1106) and Elasticity.__init__ -> initValues
 * (:1216-1236), then SPHBase.initialize (:957-960): search, static boundary
 * volume, moving boundary volume. */
int32_t or_init(or_sim *s)
{
    s->status = 0;
    or_search(s);
    elastic_init(s); /* x0 = x for every particle */
    rigid_init(s);
    boundary_volume(s, 0);
    boundary_volume(s, 1);
    s->initialized = 1;
    return s->status;
}

/* SPHBase.step, :915-940, with DFSPHSolver.substep :1546-1555. */
int32_t or_step(or_sim *s, or_stats *st)
{
    or_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    s->status = 0;
    or_search(s);
    boundary_volume(s, 1);
    clear_acc(s);
    if (s->nf > 0) {
        densities(s);
        double mx = 0;
        for (int i = 0; i < s->n; i++)
            if (is_fluid(s, i) && (double)s->rho[i] > mx) mx = (double)s->rho[i];
        st->max_density = mx;
        dfsph_factor(s);
        if (s->p.enable_divergence_solver) {
            density_change(s);
            scale_alpha(s, s->inv_dt);
            int it = 0, ran = 0;
            double err = 0;
            double eta = 1.0 / (double)s->dt * s->p.max_error_v * 0.01 * s->p.density0;
            while (it < 1 || it < s->p.max_iterations_v) {
                div_update(s);
                density_change(s);
                err = density_error(s, 0.0) / s->nf;
                ran++;
                if (err <= eta) break;
                it++;
            }
            scale_alpha(s, s->dt);
            st->iterations_v = ran;
            st->avg_density_err_v = err;
        }
        clear_acc(s);
        if (s->p.use_surface_tension) surface_tension(s);
        if (s->p.use_non_newtonian) strain_rate(s);
        if (s->p.use_viscosity) {
            if (s->p.viscosity_method == 1) {
                weiler_solve(s, st);
                if (s->bnu != (real)0.0) {
                    /* explicit boundary term of the reference model, fluid-solid pairs only */
                    real save = s->nu;
                    s->nu = (real)0.0;
                    viscosity_explicit(s);
                    s->nu = save;
                }
            } else {
                viscosity_explicit(s);
            }
        }
        np_advect(s);
        density_adv(s);
        scale_alpha(s, s->inv_dt2);
        int it = 0, ran = 0;
        double err = 0;
        double eta = s->p.max_error * 0.01 * s->p.density0;
        while (it < 1 || it < s->p.max_iterations) {
            press_update(s);
            density_adv(s);
            err = density_error(s, s->p.density0) / s->nf;
            ran++;
            if (err <= eta) break;
            it++;
        }
        st->iterations_p = ran;
        st->avg_density_err_p = err;
    }
    if (s->ne > 0) {
        elastic_rot_stress(s);
        elastic_forces(s);
    }
    if (s->nbodies > 0) rigid_step(s);
    advect(s);
    clamp_boundary(s);
    /* pairs of the lists a step uses: fluid and elastic particles (the static
     * boundary's lists serve only the boundary volumes of the first search) */
    long long np = 0;
    for (int i = 0; i < s->n; i++)
        if (s->kind[i] != OR_KIND_STATIC) np += s->nnbr[i];
    st->num_pairs = np;
    st->status = s->status;
    return s->status;
}

int32_t or_get(const or_sim *s, int32_t field, double *out)
{
    int n = s->n;
    switch (field) {
    case OR_F_X: for (int i = 0; i < 3 * n; i++) out[i] = s->x[i]; break;
    case OR_F_V: for (int i = 0; i < 3 * n; i++) out[i] = s->v[i]; break;
    case OR_F_ACC: for (int i = 0; i < 3 * n; i++) out[i] = s->a[i]; break;
    case OR_F_X0: for (int i = 0; i < 3 * n; i++) out[i] = s->x0[i]; break;
    case OR_F_VISC_DIFF: for (int i = 0; i < 3 * n; i++) out[i] = s->wvd[i]; break;
    case OR_F_VOLUME: for (int i = 0; i < n; i++) out[i] = s->vol[i]; break;
    case OR_F_DENSITY: for (int i = 0; i < n; i++) out[i] = s->rho[i]; break;
    case OR_F_DFSPH_FACTOR: for (int i = 0; i < n; i++) out[i] = s->alpha[i]; break;
    case OR_F_DENSITY_ADV: for (int i = 0; i < n; i++) out[i] = s->dadv[i]; break;
    case OR_F_STRAIN_NORM: for (int i = 0; i < n; i++) out[i] = s->gam[i]; break;
    case OR_F_VISCOSITY: for (int i = 0; i < n; i++) out[i] = s->eta[i]; break;
    case OR_F_REST_VOLUME: for (int i = 0; i < n; i++) out[i] = s->rest_vol[i]; break;
    case OR_F_ROTATION:
    case OR_F_F:
        for (int i = 0; i < n; i++) {
            int q = s->s_of[i];
            for (int a = 0; a < 9; a++) {
                double idv = (a % 4) == 0 ? 1.0 : 0.0;
                out[9 * i + a] = q < 0 ? idv : (double)(field == OR_F_ROTATION ? s->R : s->Fm)[9 * q + a];
            }
        }
        break;
    case OR_F_STRESS:
        for (int i = 0; i < n; i++) {
            int q = s->s_of[i];
            for (int a = 0; a < 6; a++) out[6 * i + a] = q < 0 ? 0.0 : (double)s->stress[6 * q + a];
        }
        break;
    default: return -1;
    }
    return 0;
}

int32_t or_rigid_frame(const or_sim *s, int32_t b, double rest_cm[3], double cm[3], double R[9])
{
    if (b < 0 || b >= s->nbodies) return -1;
    for (int c = 0; c < 3; c++) {
        if (rest_cm) rest_cm[c] = (double)s->rb_rest_cm[3 * b + c];
        if (cm) cm[c] = (double)s->rb_cm[3 * b + c];
    }
    if (R)
        for (int k = 0; k < 9; k++) R[k] = (double)s->rb_R[9 * b + k];
    return 0;
}

/* overwrite positions, velocities or the warm start's direction (tests pose
 * deformed states after init, or continue a run from another's state) */
int32_t or_set(or_sim *s, int32_t field, const double *in)
{
    real *dst = field == OR_F_X ? s->x : field == OR_F_V ? s->v : field == OR_F_VISC_DIFF ? s->wvd : NULL;
    if (!dst) return -1;
    for (int i = 0; i < 3 * s->n; i++) dst[i] = (real)in[i];
    return 0;
}

int32_t or_neighbors(const or_sim *s, int32_t *nbr, int32_t *cnt)
{
    memcpy(nbr, s->nbr, sizeof(int32_t) * (size_t)s->n * MAX_NB);
    memcpy(cnt, s->nnbr, sizeof(int32_t) * (size_t)s->n);
    return 0;
}

int32_t or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int32_t n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* unit-level entry points                                                  */
/* ------------------------------------------------------------------------ */
double or_kernel_w(const or_params *p, double r)
{
    or_sim s;
    memset(&s, 0, sizeof(s));
    s.p = *p;
    derive(&s);
    return (double)kern_w(&s, (real)r);
}

void or_kernel_grad(const or_params *p, const double r[3], double out[3])
{
    or_sim s;
    memset(&s, 0, sizeof(s));
    s.p = *p;
    derive(&s);
    real rr[3] = {(real)r[0], (real)r[1], (real)r[2]}, g[3];
    kern_grad(&s, rr, g);
    for (int c = 0; c < 3; c++) out[c] = (double)g[c];
}

void or_polar(const double a[9], double r[9])
{
    real A[9], R[9];
    for (int k = 0; k < 9; k++) A[k] = (real)a[k];
    polar3(A, R);
    for (int k = 0; k < 9; k++) r[k] = (double)R[k];
}

double or_viscosity_model(const or_params *p, double g) { return (double)nn_model(p, (real)g); }

/* Weiler probe for tests on tiny scenes: runs the step's stages up to the
 * implicit viscosity solve (search, dynamic boundary volumes, densities, strain
 * rate) on the current state, then returns the dense 3nf x 3nf operator, the
 * right-hand side b = v and the CG solution, fluid particles in index order. */
int32_t or_weiler_probe(or_sim *s, double *mat, double *b, double *x, int32_t nmax, or_stats *st)
{
    int nf = s->nf;
    if (nf > nmax || nf == 0) return -1;
    or_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    or_search(s);
    boundary_volume(s, 1);
    densities(s);
    if (s->p.use_non_newtonian) strain_rate(s);
    int *fidx = (int *)malloc(sizeof(int) * s->n);
    int c = 0;
    for (int i = 0; i < s->n; i++) fidx[i] = is_fluid(s, i) ? c++ : -1;
    for (int i = 0; i < s->n; i++) s->wmu[i] = is_fluid(s, i) ? mu_of(s, i) : (real)0.0;
    weiler_precond(s);
    int m = 3 * nf;
    real *e = (real *)calloc((size_t)3 * s->n, sizeof(real));
    real *o = (real *)calloc((size_t)3 * s->n, sizeof(real));
    for (int i = 0; i < s->n; i++) {
        if (fidx[i] < 0) continue;
        for (int q = 0; q < 3; q++) {
            memset(e, 0, sizeof(real) * 3 * s->n);
            e[3 * i + q] = 1;
            weiler_apply(s, e, o);
            int col = 3 * fidx[i] + q;
            for (int k = 0; k < s->n; k++) {
                if (fidx[k] < 0) continue;
                for (int t = 0; t < 3; t++) mat[(long)(3 * fidx[k] + t) * m + col] = (double)o[3 * k + t];
            }
        }
    }
    for (int i = 0; i < 3 * s->n; i++) s->a[i] = 0;
    weiler_solve(s, st);
    for (int i = 0; i < s->n; i++) {
        if (fidx[i] < 0) continue;
        for (int t = 0; t < 3; t++) {
            b[3 * fidx[i] + t] = (double)s->v[3 * i + t];
            x[3 * fidx[i] + t] = (double)s->wx[3 * i + t];
        }
    }
    free(e);
    free(o);
    free(fidx);
    return 0;
}

/* Force probe for tests (see sph_oracle.h): one term on the current state. */
int32_t or_force_probe(or_sim *s, int32_t which, double *out)
{
    if (which < 0 || which > 2) return -1;
    if (which == 2 && s->ne > 0 && s->ns == 0) return -1;
    or_search(s);
    boundary_volume(s, 1);
    if (s->nf > 0) densities(s);
    for (int i = 0; i < 3 * s->n; i++) s->a[i] = 0;
    if (which == 0) viscosity_explicit(s);
    else if (which == 1) surface_tension(s);
    else if (s->ne > 0) {
        elastic_rot_stress(s);
        elastic_forces(s);
    }
    for (int i = 0; i < 3 * s->n; i++) out[i] = (double)s->a[i];
    return 0;
}
