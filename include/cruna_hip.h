/*
 * cruna_hip -- MI355X (gfx950) implementation of the CRUNA time-stepping hot
 * path, exported as a plain C ABI for the reference's Fortran host
 * (iso_c_binding) and for any other FFI (ctypes, ...).
 *
 * Data layout is the reference's: every 3-D field is Fortran column-major
 * f(n1,n2,n3) with i fastest, and a state is q(n1,n2,n3,5) holding
 * (rho, rho*u, rho*v, rho*w, p) with the variable index last
 * (src/data/data_rhs.f90:16-41, src/equations/ns_rho_rhou_p.f90:95-99).
 * Host pointers are plain double*; every call returns 0 on success or a
 * negative CRUNA_HIP_E* code, with a message in cruna_hip_last_error().
 *
 * A context is bound to one HIP device and one stream and is not
 * thread-safe, like the reference's module state (save variables).
 */
#ifndef CRUNA_HIP_H
#define CRUNA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRUNA_HIP_OK 0
#define CRUNA_HIP_EINVAL (-1)   /* bad argument / unsupported configuration */
#define CRUNA_HIP_EDEVICE (-2)  /* HIP runtime error */
#define CRUNA_HIP_ENOMEM (-3)   /* device allocation failed */

#define CRUNA_HIP_PERIODIC 0 /* diff/filter_pen_core_o6_i_c_p */
#define CRUNA_HIP_OPEN 1     /* diff/filter_pen_core_o6_i_c_o */

#define CRUNA_HIP_VISC_SUTHERLAND 0 /* ns_rho_rhou_p_trafos.f90:119-142 */
#define CRUNA_HIP_VISC_POWER 1      /* case-level power-law override */

#define CRUNA_HIP_FP64 0       /* all arithmetic in FP64 */
#define CRUNA_HIP_FP32_MIXED 1 /* stencil/filter/RHS in FP32, RK4 state in FP64 */

#define CRUNA_HIP_MAX_BC 32

#define CRUNA_HIP_DIFF_O6_COMPACT 0  /* src/diff/diff_pen_core_o6_i_c_{p,o}.f90 */
#define CRUNA_HIP_DIFF_O4_EXPLICIT 1 /* src/diff/diff_pen_core_o4_e_c_{p,o}.f90 */
#define CRUNA_HIP_DIFF_O4_PADE 2     /* src/diff/diff_pen_core_o4_i_c_{p,o}_pade.f90 */
#define CRUNA_HIP_DIFF_O2_EXPLICIT 3 /* src/diff/diff_pen_core_o2_e_c_o.f90 (open lines only) */

#define CRUNA_HIP_EQ_NS 0    /* src/equations/ns_rho_rhou_p*.f90 (default) */
#define CRUNA_HIP_EQ_EULER 1 /* src/equations/euler_rho_rhou_p.f90, _grid_transform.f90 on a metric */

typedef struct cruna_hip_cfg {
  int n[3];              /* params%geom%n1..n3 (block = whole grid on one GPU) */
  double dx[3];          /* params%geom%dx1..dx3 */
  int scheme[3];         /* CRUNA_HIP_PERIODIC / CRUNA_HIP_OPEN per direction */
  double filter_alpha;   /* filter.alpha (default 0.35) */
  double gamma;          /* material.gamma */
  double R;              /* material.R */
  double W;              /* material.W */
  double Pr;             /* material.Pr */
  double mu0;            /* material.mu0 */
  double T0;             /* material.T0 */
  int visc_law;          /* CRUNA_HIP_VISC_* */
  double visc_exp;       /* exponent of the power law */
  double dt;             /* time.dt */
  int filter_freq[3];    /* filter.di1freq..di3freq (0 = never) */
  int afilter_freq[3];   /* filter.ai1freq..ai3freq of the adjoint loop (0 = never) */
  int heat_unit_spacing; /* 1: heat-flux Dx without h, as ns_rho_rhou_p.f90:189-197 */
  int precision;         /* CRUNA_HIP_FP64 / CRUNA_HIP_FP32_MIXED */
  int tile_lines[3];     /* lines per work-group of the sweeps along each direction:
                            0 = library choice, else a power of two dividing the
                            line count with (n/8) * tile_lines <= 512 */
  /* boundary_conditions.dat rows: type; i1s; i1e; i2s; i2e; i3s; i3e; k1; k2; k3.
     Types 0, 100 (slip wall: column 8 names the variable whose RHS is zero),
     200 (characteristic toward the inner neighbour; box one point thick along
     the normal), 201, 204 (characteristic on q / on the RHS), 202, 205 (the
     same with the normals per point, cruna_hip_set_kvector), 203, 206 (zonal:
     the same over the whole domain, weighted by the sponge shape; box columns
     unused; a 203 row's column 8 is -1 / 1 / 0 for acting only after the RK
     step / only inside the RHS / in both, the last 203 row deciding), 400,
     401 (sponge on the RHS / on q).  The adjoint refuses rows 204-206 as the
     reference's set_boundary_condition_rhss stops on them */
  int nbc;
  double bc[CRUNA_HIP_MAX_BC][10];
  int diff_kind[3];      /* derivative core per direction (discretisation_x_core):
                            CRUNA_HIP_DIFF_O6_COMPACT (diff_pen_core_o6_i_c_*, default),
                            CRUNA_HIP_DIFF_O4_EXPLICIT (diff_pen_core_o4_e_c_*) or
                            CRUNA_HIP_DIFF_O4_PADE (diff_pen_core_o4_i_c_*_pade) or
                            CRUNA_HIP_DIFF_O2_EXPLICIT (diff_pen_core_o2_e_c_o, open only); the
                            filter is the o6 compact one in every case */
  int wfilter_freq[3];   /* filter.dwi1freq..dwi3freq of the cruna_hpc loop (0 = never):
                            the weighted filter (filter_weighted_iN with the
                            o2_e_c_p core, weight phi of the penalisation or 1,
                            strength 1) after the compact filters of a step */
  int equations;         /* governing_equations: CRUNA_HIP_EQ_NS or CRUNA_HIP_EQ_EULER
                            (euler_rho_rhou_p, or euler_rho_rhou_p_grid_transform
                            after cruna_hip_set_metric: the NS sweeps with
                            mu = lambda = 0; penalised only with a metric,
                            euler_rho_rhou_p_ev_phi_grid_transform, as the
                            reference has no Cartesian penalised Euler module;
                            its adjoint on a deformed grid returns EINVAL) */
} cruna_hip_cfg;

typedef struct cruna_hip_ctx cruna_hip_ctx;

/* Fill cfg with the reference defaults (filter.alpha 0.35, Pr 0.71, ...). */
void cruna_hip_default_cfg(cruna_hip_cfg *cfg);

/* Create / destroy a context: builds the line-operator tables
 * (init_left_implicit_diff_part / init_right_explicit_diff_part of
 * src/diff/diff_pen_core_o6_i_c_{p,o}.f90 and the filter equivalents) and
 * allocates the device-resident state. */
int cruna_hip_create(const cruna_hip_cfg *cfg, cruna_hip_ctx **out);
int cruna_hip_destroy(cruna_hip_ctx *ctx);
const char *cruna_hip_last_error(void);

/* Device-resident state q^n (host <-> device copies, layout q(n1,n2,n3,5)). */
int cruna_hip_set_state(cruna_hip_ctx *ctx, const double *q);
int cruna_hip_get_state(cruna_hip_ctx *ctx, double *q);
/* Reference state qref (src/data/data_reference.f90, qref = q0 at init) and
 * the sponge shape sigma(n1,n2,n3) (shape_functions.f90:69-150). */
int cruna_hip_set_qref(cruna_hip_ctx *ctx, const double *qref);
int cruna_hip_set_sponge_shape(cruna_hip_ctx *ctx, const double *shape);

/* kvector(n1,n2,n3,3): the outward boundary normal per point that rows of
 * types 202, 203, 205 and 206 use (characteristic_q_k4d / characteristic_rhs_k4d /
 * characteristic_qs_k4d, src/boundaries/nonreflecting.f90; the reference
 * loads it as the "kvector" field, io_helper.f90 get_fname "kvector"). */
int cruna_hip_set_kvector(cruna_hip_ctx *ctx, const double *kvector);

/* Deformed grid: X_metric(n1,n2,n3,3,3) and X_jacobian(n1,n2,n3) of data_geom
 * as grid_metric::calc_metric leaves them (src/geometry/grid_metric.f90:16-188);
 * from then on the direct RHS is that of ns_rho_rhou_p_grid_transform.f90:31-330,
 * as a case selects it through governing_equations.  Any metric: a diagonal
 * one whose entry (d,d) is constant along direction d (a tensor-product
 * stretch) runs the per-line sweeps, every other (general curvilinear grids)
 * the sweeps that read the full metric per point.  Non-finite entries return
 * EINVAL.  The adjoint RHS of ns_rho_rhou_p_grid_transform.f90:334-613 is
 * the Cartesian one (it uses neither X_metric nor X_jacobian), so the
 * adjoint entry points run unchanged on a deformed grid. */
int cruna_hip_set_metric(cruna_hip_ctx *ctx, const double *X_metric, const double *X_jacobian);

/* Volume penalisation: phi(n1,n2,n3) and inv_phi(n1,n2,n3), the case's
 * volume_penalization penalization(:,:,:,1) and (:,:,:,2) (porosity and the
 * stored 1/phi, src/immersed/volume_penalization.f90 and the case's
 * init_volume_penalization); from then on the direct RHS is that of
 * ns_rho_rhou_p_ev_phi.f90:33-291 (Cartesian) or of
 * ns_rho_rhou_p_ev_phi_grid_transform.f90 (with a metric), as a case selects
 * it through governing_equations, with the reference's no-op
 * apply_volume_penalization.  Both null: back to the unpenalised RHS.
 * EINVAL: non-finite values; a tensor-product metric set before (set the
 * penalisation first: the penalised deformed grid keeps the full metric per
 * point); and afterwards every adjoint entry point (the reference's ev_phi
 * adjoint stops, ns_rho_rhou_p_ev_phi.f90:295-309). */
int cruna_hip_set_penalization(cruna_hip_ctx *ctx, const double *phi, const double *inv_phi);

/* filter_weighted_iN_mpi_cartesian::filter_4d on nv fields q(n1,n2,n3,nv) in
 * place (src/filter/filter_weighted_i1_mpi_cartesian.f90:15-54) with the
 * filter_weighted_pen_core_o2_e_c_p core (periodic along dir):
 * q <- q - F(strength weight F(q)) / weight.  weight, strength: (n1,n2,n3)
 * fields, or null for 1. */
int cruna_hip_filter_weighted(cruna_hip_ctx *ctx, int dir, double *q, int nv, const double *weight,
                              const double *strength);

/* Dx1/Dx2/Dx3 on a 3-D field: diff_iN_mpi_cartesian::diff_3d
 * (src/diff/diff_i1_mpi_cartesian.f90:83-103).  dir = 0,1,2; h = spacing
 * (pass 1.0 for the optional-h default). */
int cruna_hip_diff(cruna_hip_ctx *ctx, int dir, double *q_x, const double *q, double h);

/* filter_i1/i2/i3 on a 4-D field in place: filter_iN_mpi_cartesian::filter_4d
 * (src/filter/filter_i1_mpi_cartesian.f90:83-97); nv variables. */
int cruna_hip_filter(cruna_hip_ctx *ctx, int dir, double *q, int nv);

/* right_hand_side_direct_3d(rhs,q_in,t) (src/equations/ns_rho_rhou_p.f90:31-253). */
int cruna_hip_rhs_direct(cruna_hip_ctx *ctx, double *rhs, const double *q, double t);

/* cruna_hip_rhs_direct without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, code between graph captures, harnesses that forbid
 * hipMalloc).  Same arguments and results as cruna_hip_rhs_direct, which
 * allocates its staging buffer on first use; this variant stages through a
 * static device array of 2^24 doubles (128 MiB per GPU, held until the
 * process ends) and returns CRUNA_HIP_EINVAL when 5 (N_device + N) words do
 * not fit it (N = n1 n2 n3, N_device the padded point count).  Calls that
 * overlap on one GPU share that array and race. */
int cruna_hip_rhs_direct_noalloc(cruna_hip_ctx *ctx, double *rhs, const double *q, double t);

/* Box averages over f x f x f blocks of the device state q^n and (rhs_box
 * not NULL) of its direct RHS, into host arrays (n1/f, n2/f, n3/f, 5): the
 * filtered fine fields from which a closure's training data are made.  No
 * counterpart in the reference (its solution output is whole fields). */
int cruna_hip_box_filter_state(cruna_hip_ctx *ctx, int f, double *q_box, double *rhs_box);

/* right_hand_side_adjoint_3d(rhs,qs_in,q,g,t) (ns_rho_rhou_p.f90:257-536); g may be NULL. */
int cruna_hip_rhs_adjoint(cruna_hip_ctx *ctx, double *rhs, const double *qs, const double *q,
                          const double *g, double t);

/* time_stepper_direct(q1,q0,t) (src/time/rk_o4_e.f90:17-58), host in/out. */
int cruna_hip_time_stepper_direct(cruna_hip_ctx *ctx, double *q1, const double *q0, double t);

/* time_stepper_adjoint(qs1,qs0,q0,g,t) (src/time/rk_o4_e.f90:62-105), host in/out;
 * g may be NULL.  Leaves qs1 also in the device adjoint state. */
int cruna_hip_time_stepper_adjoint(cruna_hip_ctx *ctx, double *qs1, const double *qs0,
                                   const double *q0, const double *g, double t);

/* Device-resident step loop body of cruna_hpc / calc_direct
 * (src/main/cruna_hpc.f90:171-229): for nt = nt0+1..nt0+nsteps:
 * RK4 -> BC on q (outside rhs) -> filter i1,i2,i3 where mod(nt,freq) == 0.
 * Asynchronous on the context stream; cruna_hip_synchronize waits. */
int cruna_hip_run_direct(cruna_hip_ctx *ctx, int nt0, int nsteps);
int cruna_hip_synchronize(cruna_hip_ctx *ctx);

/* Convergence check of the cruna_hpc loop (src/main/cruna_hpc.f90:185-208,
 * run.cfreq and run.ctol read at :81-82): at steps with mod(nt, cfreq) == 0,
 * residual = maxval(abs((qn - q1)/dt)) over every point and variable right
 * after the RK4 step; when residual <= ctol, cruna_hip_run_direct ends after
 * that step with the unfiltered q1 as the device state (the state the
 * reference stores as its snapshot before stop_cruna).  cfreq 0 (the
 * default) disables the check.  Each check waits for the step to finish (the
 * reference's allreduce).  get_residual: the last residual computed (-1 if
 * none), the step it was computed at, and the step at which the last
 * cruna_hip_run_direct reached the tolerance (0 if it did not). */
int cruna_hip_set_convergence(cruna_hip_ctx *ctx, int cfreq, double ctol);
int cruna_hip_get_residual(cruna_hip_ctx *ctx, double *residual, int *nt_checked, int *nt_converged);

/* Device-resident adjoint state qs^n (layout as q). */
int cruna_hip_set_adjoint_state(cruna_hip_ctx *ctx, const double *qs);
int cruna_hip_get_adjoint_state(cruna_hip_ctx *ctx, double *qs);

/* Forward trajectory Q(:,:,:,:,nt) of calc_direct / calc_adjoint
 * (src/main/computation_direct.f90:100-103, computation_adjoint.f90:69-73),
 * kept in device memory: slot nt (1-based) holds the state after forward
 * step nt.  alloc reserves nslots states (0 frees them); set/get copy one
 * slot from/to the host (the reference's load of data_direct); with
 * record on, cruna_hip_run_direct stores each step's state in its slot. */
int cruna_hip_trajectory_alloc(cruna_hip_ctx *ctx, int nslots);
int cruna_hip_trajectory_set(cruna_hip_ctx *ctx, int nt, const double *q);
int cruna_hip_trajectory_get(cruna_hip_ctx *ctx, int nt, double *q);
int cruna_hip_trajectory_record(cruna_hip_ctx *ctx, int on);

/* Subsets with checkpoints (time.subsets of calc_direct / calc_adjoint,
 * src/main/computation_direct.f90:55-133, computation_adjoint.f90:45-135):
 * the reference keeps each subset's forward trajectory Q(:,:,:,:,1..steps)
 * in a cache ('data_direct_subsets_cache') and loads it subset by subset,
 * last first, for the adjoint.  Here only the state at the start of every
 * subset is kept on the device (a checkpoint), and the adjoint recomputes
 * a subset's trajectory from its checkpoint into `steps` trajectory slots
 * right before running backward over it: subsets + 1 + steps states in HBM
 * instead of subsets * steps.
 * alloc: checkpoints for `subsets` and a trajectory of `steps` slots.
 * run_direct_subsets: for every subset, checkpoint q^n, then the step-loop
 * body for nt = 1..steps (filters by mod(nt, freq) within the subset, as
 * the reference counts nt); the final state is kept too.
 * run_adjoint_subsets: for s = subsets..1: recompute subset s, then the
 * adjoint loop nt = steps..1 over it; q^n is the final forward state again
 * afterwards. */
int cruna_hip_subsets_alloc(cruna_hip_ctx *ctx, int subsets, int steps);
int cruna_hip_run_direct_subsets(cruna_hip_ctx *ctx, int subsets, int steps);
int cruna_hip_run_adjoint_subsets(cruna_hip_ctx *ctx, int subsets, int steps);

/* Adjoint step loop of calc_adjoint (src/main/computation_adjoint.f90:62-99)
 * with g = 0: for nt = nt_end down to nt_end-nsteps+1, with q0 = Q(nt)
 * (or the only slot when one is allocated: the linear adjoint):
 * time_stepper_adjoint -> BC on qs (outside rhs) -> adjoint filters i1,i2,i3
 * where mod(nt, afilter_freq) == 0.  Asynchronous on the context stream. */
int cruna_hip_run_adjoint(cruna_hip_ctx *ctx, int nt_end, int nsteps);

/* Tile-shape autotuner: times `steps` RK4 steps (after one warm-up step)
 * on a saved copy of the state for each valid number of lines per
 * work-group, one direction at a time, keeps the fastest in
 * cfg.tile_lines and restores the state.  tile_lines (may be NULL) receives
 * the choice.  Results do not depend on the tile shape. */
int cruna_hip_autotune(cruna_hip_ctx *ctx, int steps, int tile_lines[3]);

/* Ensembles: independent members on one device -- unevenly sized boxes,
 * forward or adjoint, e.g. the seeds of a study or the direct and adjoint
 * sweeps of an optimisation.  Each member is a context with its own stream,
 * reachable through cruna_hip_ensemble_member for set_state, trajectories,
 * etc.  cruna_hip_ensemble_run advances every member by nsteps: the
 * launches of one period of steps of each member (the period after which
 * the filter frequencies repeat) are recorded once into a HIP graph per
 * member, and the members' graphs are replayed on their own streams so that
 * members run concurrently; steps that do not fill a period, and the first
 * period of a member, run eagerly.
 * Forward members run calc_direct's step body from their step counter;
 * adjoint members run calc_adjoint's backward from step nt over their
 * trajectory.  Members must not be used on their own while a run is in
 * progress, and cannot carry a learned closure (cruna_hip_set_closure). */
#define CRUNA_HIP_FORWARD 0
#define CRUNA_HIP_ADJOINT 1
typedef struct cruna_hip_ensemble cruna_hip_ensemble;
int cruna_hip_ensemble_create(const cruna_hip_cfg *cfgs, int nmembers, cruna_hip_ensemble **out);
int cruna_hip_ensemble_destroy(cruna_hip_ensemble *ens);
int cruna_hip_ensemble_member(cruna_hip_ensemble *ens, int i, cruna_hip_ctx **ctx);
/* role CRUNA_HIP_FORWARD: nt = steps already done; CRUNA_HIP_ADJOINT: nt =
 * the first (latest) step to run backward (needs the member's adjoint state
 * and trajectory set). */
int cruna_hip_ensemble_set_role(cruna_hip_ensemble *ens, int i, int role, int nt);
int cruna_hip_ensemble_run(cruna_hip_ensemble *ens, int nsteps);
/* Batched runs (on by default): when every member is FP64 on a Cartesian grid
 * without boundary rows, closure or weighted filter, the members share their
 * schemes and filter switches (frequency 0 or 1), and the adjoint members run
 * either all over whole trajectories or all over subset checkpoints of one
 * length from a subset boundary, cruna_hip_ensemble_run advances whole
 * periods (a step, or a subset) in batched launches: the members split into
 * `bins` of about equal points, and each bin runs one launch per sweep kind
 * over its members (a work list of member tiles), its period recorded once
 * into a HIP graph replayed on the bin's own stream; the bins run side by
 * side.  The subset members' recomputed trajectories step with the forward
 * members.  Results are those of the members' own runs, bit for bit.
 * No counterpart in the reference (one box per process,
 * src/diff/diff_pen_core_o6_i_c_p.f90:27-32).  set_batching(ens, bins):
 * -1 automatic (the default: 4 bins, at most one per member), 1 one launch
 * per sweep kind over all members, 0 the per-member graphs; batched_steps
 * reports the steps of the last run that went through batched launches. */
int cruna_hip_ensemble_set_batching(cruna_hip_ensemble *ens, int bins);
int cruna_hip_ensemble_batched_steps(const cruna_hip_ensemble *ens, int *steps);
/* cruna_hip_autotune on every member; tile_lines (may be NULL) receives 3
 * values per member. */
int cruna_hip_ensemble_autotune(cruna_hip_ensemble *ens, int steps, int *tile_lines);

/* Learned subfilter closure: a 3-D CNN (5 -> 32 -> 32 -> 32 -> 5 channels,
 * 3x3x3 periodic kernels, GELU after the hidden layers) on a periodic S^3
 * grid, trained and evaluated with BF16 matrix cores and FP32 accumulation.
 * It is the additive force term of the direct RHS (apply_force,
 * src/equations/ns_rho_rhou_p.f90:212 and src/opt/force.f90:28): with a
 * closure set, every direct RHS evaluation of the context adds
 * force = cnn(q) with the normalisation of cruna_hip_cnn_set_norm, the terms
 * of rho, rho u, rho v, rho w then shifted to zero domain mean (subfilter
 * terms of conservation laws add nothing to the domain totals).
 * Parameters (FP32) are laid out per layer as W[27][co][ci] then b[co] with
 * the channel counts padded to (16,32), (32,32), (32,32), (32,16); taps
 * t = 9 (dz+1) + 3 (dy+1) + (dx+1).  The closure's buffers are shared by
 * every context it is set on: use those contexts from one host thread. */
typedef struct cruna_hip_cnn cruna_hip_cnn;
int cruna_hip_cnn_create(int S, cruna_hip_cnn **out);
/* EINVAL while the closure is attached to a context (detach it, or destroy
 * the context, first) */
int cruna_hip_cnn_destroy(cruna_hip_cnn *cnn);
long cruna_hip_cnn_num_params(const cruna_hip_cnn *cnn);
int cruna_hip_cnn_init_params(cruna_hip_cnn *cnn, unsigned seed);
int cruna_hip_cnn_set_params(cruna_hip_cnn *cnn, const float *params);
int cruna_hip_cnn_get_params(cruna_hip_cnn *cnn, float *params);
/* inputs: x_c = (q_c - mean of q_c over the grid) / in_std[c]; outputs:
 * force_v = out_v out_std[v], the pressure equation's also times
 * mean(p) / p_ref when p_ref > 0 (its residual scales with the mean
 * pressure, i.e. 1 / Mach^2); defaults 1, 1 and 0 */
int cruna_hip_cnn_set_norm(cruna_hip_cnn *cnn, const double in_std[5], const double out_std[5], double p_ref);
/* scale of the closure term of each variable at inference (default 1;
 * training is unaffected) */
int cruna_hip_cnn_set_gain(cruna_hip_cnn *cnn, const double gain[5]);
/* closure (5 fields, FP64, field-major like q) of the state q */
int cruna_hip_cnn_forward(cruna_hip_cnn *cnn, const double *q, double *closure);
/* mean-square loss of the normalised output against target / out_std and
 * its gradient with respect to every parameter (grads may be NULL) */
int cruna_hip_cnn_gradients(cruna_hip_cnn *cnn, const double *q, const double *target, float *grads,
                            float *loss);
/* keep a (state, target) pair on the device for cruna_hip_cnn_train */
int cruna_hip_cnn_add_sample(cruna_hip_cnn *cnn, const double *q, const double *target);
int cruna_hip_cnn_num_samples(const cruna_hip_cnn *cnn);
/* iters Adam steps (beta 0.9, 0.999, eps 1e-8) on batches drawn with the
 * seed: `batch` windows of crop^3 points at random samples and offsets (the
 * network runs on each window with a 4-point halo, so a window's values are
 * those of the whole periodic sample), or, with crop 0, `batch` whole
 * samples; losses (may be NULL) receives the loss of every step */
int cruna_hip_cnn_train(cruna_hip_cnn *cnn, int iters, float lr, int batch, int crop, unsigned seed,
                        float *losses);
/* add the closure to the context's direct RHS (NULL removes it); the
 * context's grid must be the periodic S^3 grid of the closure */
int cruna_hip_set_closure(cruna_hip_ctx *ctx, cruna_hip_cnn *cnn);

/* Per-kernel timing with HIP events on the context stream: enable, run,
 * then read the summed milliseconds and launch counts per kernel id. */
int cruna_hip_profile_enable(cruna_hip_ctx *ctx, int on);
int cruna_hip_profile_read(cruna_hip_ctx *ctx, int kernel_id, double *ms, long *launches,
                           char *name, size_t name_len);
int cruna_hip_profile_count(void);

#ifdef __cplusplus
}
#endif
#endif
