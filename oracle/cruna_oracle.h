/*
 * CPU oracle for the CRUNA time-stepping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran algorithm (cruna-toolkit/public, src/...) used as the checker for
 * the HIP library in public_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never calls it.
 *
 * Layout: every field is Fortran column-major q(n1,n2,n3,nv), i fastest,
 * variable index last (src/data/data_rhs.f90:16-41); state variables are
 * (rho, rho*u, rho*v, rho*w, p) (src/equations/ns_rho_rhou_p.f90:95-99).
 */
#ifndef CRUNA_ORACLE_H
#define CRUNA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PERIODIC = 0, ORC_OPEN = 1 };
enum { ORC_VISC_SUTHERLAND = 0, ORC_VISC_POWER = 1 };
#define ORC_MAX_BC 32

typedef struct orc_cfg {
  int n[3];               /* grid points per direction */
  double dx[3];           /* params%geom%dx1..3 */
  int scheme[3];          /* ORC_PERIODIC: o6_i_c_p, ORC_OPEN: o6_i_c_o (derivative and filter) */
  double filter_alpha;    /* filter.alpha */
  double gamma;           /* material.gamma */
  double R;               /* material.R */
  double W;               /* material.W */
  double Pr;              /* material.Pr */
  double mu0;             /* material.mu0 */
  double T0;              /* material.T0 */
  int visc_law;           /* ORC_VISC_SUTHERLAND (reference trafos) or ORC_VISC_POWER (case override) */
  double visc_exp;        /* power-law exponent */
  double dt;              /* time.dt */
  int filter_freq[3];     /* filter.diNfreq (0 = never) */
  int afilter_freq[3];    /* filter.aiNfreq (0 = never) */
  /* boundary_conditions.dat rows: type; i1s; i1e; i2s; i2e; i3s; i3e; k1; k2; k3 (1-based, inclusive) */
  int nbc;
  double bc[ORC_MAX_BC][10];
  /* Deformed grid (data_geom X_metric(n1,n2,n3,3,3), X_jacobian(n1,n2,n3)); when
   * X_metric is set the direct RHS is ns_rho_rhou_p_grid_transform's, as a
   * case selects it through governing_equations.  NULL: Cartesian. */
  const double *X_metric;
  const double *X_jacobian;
  /* derivative core per direction (the case's discretisation_x_core choice):
   * ORC_DIFF_O6_COMPACT: diff_pen_core_o6_i_c_{p,o}; ORC_DIFF_O4_EXPLICIT:
   * diff_pen_core_o4_e_c_{p,o}; ORC_DIFF_O4_PADE: diff_pen_core_o4_i_c_{p,o}_pade.
   * The filter is the o6 one either way. */
  int diff_kind[3];
  /* Volume penalisation (volume_penalization's penalization(:,:,:,1) = phi
   * and (:,:,:,2) = 1/phi as the case stores them); when set (and X_metric
   * is not) the direct RHS is ns_rho_rhou_p_ev_phi's.  NULL: plain NS. */
  const double *penal_phi;
  const double *penal_inv_phi;
  /* filter.dwi1freq..dwi3freq (0 = never): the cruna_hpc loop's weighted
   * filters with the o2_e_c_p core, weight penal_phi (or 1), strength 1 */
  int wfilter_freq[3];
  /* governing_equations: ORC_EQ_NS (ns_rho_rhou_p*) or ORC_EQ_EULER
   * (euler_rho_rhou_p, Cartesian) */
  int equations;
  /* kvector(n1,n2,n3,3): boundary normals per point (rows 202, 205) */
  const double *kvector;
} orc_cfg;

#define ORC_EQ_NS 0
#define ORC_EQ_EULER 1

#define ORC_DIFF_O6_COMPACT 0
#define ORC_DIFF_O4_EXPLICIT 1
#define ORC_DIFF_O4_PADE 2
#define ORC_DIFF_O2_EXPLICIT 3 /* diff_pen_core_o2_e_c_o (open lines only) */

/* LAPACK dgttrf/dgttrs restatement (no transpose), 1 RHS at a time. */
int orc_dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv);
void orc_dgttrs(int n, int nrhs, const double *dl, const double *d, const double *du,
                const double *du2, const int *ipiv, double *b, int ldb);

/* Pencil cores: q(n, nlines) in place, as diff_iN / filter_iN cores. */
void orc_diff_core(int scheme, int n, int nlines, double *q, double dx);
void orc_diff_core_o4(int scheme, int n, int nlines, double *q, double dx);
void orc_diff_core_o2(int n, int nlines, double *q, double dx);
void orc_diff_core_pade(int scheme, int n, int nlines, double *q, double dx);
void orc_filter_core(int scheme, int n, int nlines, double *q, double alpha);

/* 3-D wrappers (diff_3d / filter_4d): dir = 0,1,2. */
void orc_diff3d(const orc_cfg *c, int dir, double *out, const double *in, double h);
void orc_filter4d(const orc_cfg *c, int dir, double *q, int nv);

/* Thermodynamic closures on a single point. */
double orc_T(const orc_cfg *c, double rho, double p);
double orc_mu(const orc_cfg *c, double T);
double orc_lambda(const orc_cfg *c, double mu);

/* Metric of a deformed grid, src/geometry/grid_metric.f90:16-188: X is the
 * physical grid, X_calc the calculation space, both (n1,n2,n3,3); outputs
 * X_metric(n1,n2,n3,3,3) (metric without the Jacobian) and X_jacobian. */
void orc_calc_metric(const orc_cfg *c, const double *X, const double *X_calc, double *X_metric,
                     double *X_jacobian);
/* Euler RHS, src/equations/euler_rho_rhou_p.f90:30-122. */
void orc_rhs_direct_euler(const orc_cfg *c, double *rhs, const double *q_in, const double *qref,
                          const double *shape);
/* euler_rho_rhou_p_grid_transform.f90:33-140 on the deformed grid (Xm, Xj). */
void orc_rhs_direct_euler_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                             const double *Xj, const double *qref, const double *shape);
/* euler_rho_rhou_p_ev_phi_grid_transform.f90:33-160 (penalised, phi and 1/phi). */
void orc_rhs_direct_euler_ev_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                                const double *Xj, const double *phi, const double *inv_phi,
                                const double *qref, const double *shape);
/* Volume-penalised deformed-grid RHS, src/equations/ns_rho_rhou_p_ev_phi_grid_transform.f90. */
void orc_rhs_direct_ev_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                          const double *Xj, const double *phi, const double *inv_phi,
                          const double *qref, const double *shape);
/* filter_weighted_iN_mpi_cartesian::filter_4d with the o2_e_c_p core on
 * q(n1,n2,n3,nv) along dir; weight / strength null mean 1. */
void orc_filter_weighted4d(const orc_cfg *c, int dir, double *q, int nv, const double *weight,
                           const double *strength);
/* Direct RHS with volume penalisation, src/equations/ns_rho_rhou_p_ev_phi.f90:33-291. */
void orc_rhs_direct_ev(const orc_cfg *c, double *rhs, const double *q_in, const double *phi,
                       const double *inv_phi, const double *qref, const double *shape);

/* Direct RHS on a deformed grid, ns_rho_rhou_p_grid_transform.f90:31-330. */
void orc_rhs_direct_gt(const orc_cfg *c, double *rhs, const double *q_in,
                       const double *X_metric, const double *X_jacobian, const double *qref,
                       const double *shape);

/* Right-hand sides.  qref/shape may be NULL when no BC rows use them. */
void orc_rhs_direct(const orc_cfg *c, double *rhs, const double *q_in,
                    const double *qref, const double *shape, int heat_unit_spacing);
void orc_rhs_adjoint(const orc_cfg *c, double *rhs, const double *qs_in, const double *q,
                     const double *g, const double *shape);

/* Boundary conditions (boundary_conditions.f90). */
void orc_bc_q(const orc_cfg *c, double *q, const double *qref, const double *shape);
void orc_bc_qs(const orc_cfg *c, double *qs, const double *q, const double *shape);
void orc_bc_rhs(const orc_cfg *c, double *rhs, const double *q, const double *qref,
                const double *shape);
void orc_characteristic_rhs_k1d(int np, int sv, double *rhs, const double *q, const double *aim,
                                const double k[3], double gamma, double dt);
void orc_characteristic_q_k1d(int npts, int stride_v, double *q, const double *aim,
                              const double k[3], double gamma);
void orc_characteristic_qs_k1d(int npts, int stride_v, double *qs, const double *q,
                               const double k[3], double gamma);

/* Sponge shape (shape_functions.f90 init_general_shape_fct), quadratic or cos. */
void orc_sponge_shape(const orc_cfg *c, const double x0[3], const double x1[3],
                      const double len[3], double amplitude, int cos_shape,
                      const double *X1, const double *X2, const double *X3, double *shape);

/* RK4 (rk_o4_e.f90) and the step-loop bodies (computation_direct.f90 / computation_adjoint.f90). */
void orc_time_stepper_direct(const orc_cfg *c, double *q1, const double *q0,
                             const double *qref, const double *shape, int heat_unit_spacing);
void orc_time_stepper_adjoint(const orc_cfg *c, double *qs1, const double *qs0,
                              const double *q0, const double *g, const double *shape);
void orc_run_direct(const orc_cfg *c, double *q, int nt0, int nsteps, const double *qref,
                    const double *shape, int heat_unit_spacing);
/* Q holds nsteps direct states (state after step nt at Q + (nt-1)*N*5). */
/* run the reference's compiled o6 derivative / filter cores in place of the
 * restated ones (function pointers with the signatures of ref_diff and
 * ref_filter; NULL restores the restatement), and the oracle's thread count */
void orc_set_core_hooks(void *diff, void *filter);
void orc_set_threads(int n);
/* cruna_hpc's loop with the residual check of run.cfreq / run.ctol */
void orc_run_direct_conv(const orc_cfg *c, double *q, int nt0, int nsteps, int cfreq, double ctol,
                         const double *qref, const double *shape, int heat_unit_spacing,
                         double *res, int *nt_conv);
void orc_run_adjoint(const orc_cfg *c, double *qs, const double *Q, int nsteps,
                     const double *shape);

#ifdef __cplusplus
}
#endif
#endif
