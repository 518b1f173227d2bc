/*
 * sph_oracle.h -- CPU restatement of the SPH step of chunleili/nonNewton SPH/main.py.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * nonnewton_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links, calls or falls back to it.
 *
 * The same source is compiled twice: with `real` = float (liboracle_f32.so, the
 * reference's own FP32 arithmetic, which is also the CPU baseline) and with
 * `real` = double (liboracle_f64.so, the FP64 yardstick that errors are stated
 * against).  Neighbor search always uses the reference's FP32 distance test.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene parameters.  Field meaning and defaults follow SPH/main.py
 * Parameter (:594-621), DFSPHSolver.__init__ (:1520-1544), NonNewton.initParameters
 * (:358-371), Elasticity.__init__ (:1186-1189) and SurfaceTension (:1380-1382). */
typedef struct or_params {
    double domain_start[3];
    double domain_end[3];
    double particle_radius;
    double density0;
    double gravity[3];
    double boundary_viscosity;
    double dt;
    double sticky_coeff;
    double collision_coeff;
    double viscosity;             /* kinematic, used by the explicit model */
    int32_t max_iterations_v;
    int32_t max_iterations;
    double max_error_v;           /* percent */
    double max_error;             /* percent */
    double eps;                   /* m_eps = 1e-5 */
    int32_t enable_divergence_solver;
    int32_t use_surface_tension;
    int32_t use_viscosity;
    int32_t use_non_newtonian;
    int32_t viscosity_method;     /* 0 explicit (ViscosityModel), 1 Weiler 2018 implicit */
    int32_t non_newton_method;    /* 0..6, NonNewton.NON_NEWTON_METHOD */
    double surface_tension;
    double power_index;
    double consistency_index;
    double viscosity0;
    double viscosity_inf;
    double critical_strain_rate;
    double mu_c;
    double yield_stress;
    double max_viscosity;
    int32_t weiler_max_iter;
    int32_t compensated;          /* Kahan sums for density and boundary volume */
    double weiler_tol;
    double youngs_modulus;
    double poisson_ratio;
    double particle_mass;         /* <= 0: volume0 * 1000 (:743) */
    int32_t warm_start;           /* bit 0: Weiler CG starts from v + (x - v) of the previous solve */
    int32_t coupling_interval;    /* rigid-body substeps per step, Parameter.coupling_interval (:616) */
} or_params;

typedef struct or_stats {
    int32_t iterations_v;
    int32_t iterations_p;
    int32_t iterations_cg;
    int32_t status;
    double avg_density_err_v;
    double avg_density_err_p;
    double cg_rel_residual;
    double max_density;
    int64_t num_pairs;
} or_stats;

/* particle kinds; kind OR_KIND_RIGID + b is a particle of rigid body b */
#define OR_KIND_FLUID 0
#define OR_KIND_STATIC 1
#define OR_KIND_ELASTIC 2
#define OR_KIND_RIGID 3
#define OR_MAX_BODIES 4096

/* status bits */
#define OR_ST_NEIGHBOR_OVERFLOW 1
#define OR_ST_CELL_OVERFLOW 2
#define OR_ST_OUT_OF_GRID 4
#define OR_ST_NAN 8
#define OR_ST_REST_OVERFLOW 16

/* field ids for or_get */
enum {
    OR_F_X = 0, OR_F_V = 1, OR_F_ACC = 2, OR_F_VOLUME = 3, OR_F_DENSITY = 4,
    OR_F_DFSPH_FACTOR = 5, OR_F_DENSITY_ADV = 6, OR_F_STRAIN_NORM = 7,
    OR_F_VISCOSITY = 8, OR_F_ROTATION = 9, OR_F_STRESS = 10, OR_F_F = 11,
    OR_F_REST_VOLUME = 12, OR_F_X0 = 13,
    OR_F_VISC_DIFF = 16 /* x - v of the previous Weiler solve (the warm start's direction) */
};

typedef struct or_sim or_sim;

or_sim *or_create(const or_params *p, int32_t n, const float *x, const float *v, const int32_t *kind);
void or_destroy(or_sim *s);
int32_t or_init(or_sim *s);
int32_t or_search(or_sim *s);
int32_t or_step(or_sim *s, or_stats *st);
int32_t or_get(const or_sim *s, int32_t field, double *out);
/* Rigid body b: rest centre of mass, and the centre of mass and rotation of
 * its last shape-matching projection (identity before the first step).
 * Returns -1 if b is not a body of the scene. */
int32_t or_rigid_frame(const or_sim *s, int32_t b, double rest_cm[3], double cm[3], double R[9]);
int32_t or_neighbors(const or_sim *s, int32_t *nbr, int32_t *cnt);
int32_t or_set(or_sim *s, int32_t field, const double *in);
int32_t or_num_threads(void);
void or_set_threads(int32_t n);

/* unit-level entry points for tests */
double or_kernel_w(const or_params *p, double r);
void or_kernel_grad(const or_params *p, const double r[3], double out[3]);
void or_polar(const double a[9], double r[9]);
double or_viscosity_model(const or_params *p, double strain_rate_norm);
int32_t or_weiler_probe(or_sim *s, double *mat, double *b, double *x, int32_t nmax, or_stats *st);
/* One force term on the current state, for tests: search, dynamic boundary
 * volumes and densities as a step forms them, then the accelerations of that
 * term alone (out: n x 3, all other terms and gravity left out).
 * which: 0 ViscosityModel (fluid term, and the boundary term with its
 * reaction on dynamic solids when boundary_viscosity != 0), 1 SurfaceTension,
 * 2 Elasticity (rotations, stresses, forces; needs or_init). */
int32_t or_force_probe(or_sim *s, int32_t which, double *out);

#ifdef __cplusplus
}
#endif
#endif
