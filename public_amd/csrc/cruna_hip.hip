// C-ABI implementation of include/cruna_hip.h for MI355X (gfx950).
#include "../../include/cruna_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "line_ops.hpp"
#include "boundary.hpp"
#include "sweeps.hpp"
#include "curvilinear.hpp"
#include "weighted_filter.hpp"
#include "adjoint.hpp"
#include "closure.hpp"
#include "residual.hpp"
#include "batch.hpp"

using namespace cruna;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(CRUNA_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

enum KernelId {
  KID_S1X = 0, KID_S2Y, KID_S3Z, KID_S4Z, KID_S5Y, KID_S6X, KID_FX, KID_FY, KID_FZ, KID_DIFF,
  KID_AX, KID_AY, KID_AZ, KID_CLOSURE, KID_S12Z, KID_COUNT
};
// KID_CLOSURE times one whole closure evaluation (field means, features,
// the four convolutions, the force and its mean projection)
const char *kKernelNames[KID_COUNT] = {"sweep1_x", "sweep1_y",    "sweep1_z", "sweep2_z",
                                       "sweep2_y", "sweep2_x_rk", "filter_x", "filter_y",
                                       "filter_z", "diff",        "asweep_x_rk", "asweep_y",
                                       "asweep_z", "closure",     "sweep12_z"};

template <typename T>
struct DevOp {
  T *buf = nullptr;
  LineOp<T> op{};
  int upload(const HostLineOp &h) {
    const size_t nI = h.I.size(), nM = h.M.size(), nL = h.L.size(), nR = h.R.size(),
                 nW = h.W.size();
    std::vector<T> all(nI + nM + nL + nR + nW);
    size_t o = 0;
    for (double v : h.I) all[o++] = T(v);
    for (double v : h.M) all[o++] = T(v);
    for (double v : h.L) all[o++] = T(v);
    for (double v : h.R) all[o++] = T(v);
    for (double v : h.W) all[o++] = T(v);
    HIP_TRY(hipMalloc(&buf, all.size() * sizeof(T)));
    HIP_TRY(hipMemcpy(buf, all.data(), all.size() * sizeof(T), hipMemcpyHostToDevice));
    op.I = buf;
    op.M = buf + nI;
    op.L = buf + nI + nM;
    op.R = buf + nI + nM + nL;
    op.W = buf + nI + nM + nL + nR;
    op.n = h.n;
    op.NB = h.NB;
    op.K = h.K;
    op.per = h.per;
    op.spec2 = h.spec2;
    return CRUNA_HIP_OK;
  }
  void release() {
    if (buf) (void)hipFree(buf);
    buf = nullptr;
  }
};

}  // namespace

struct cruna_hip_cnn;
// count of the contexts a closure is attached to (cruna_hip_set_closure)
static void cnn_user(cruna_hip_cnn *n, int delta);
struct cruna_hip_ctx {
  cruna_hip_cfg cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  // device layout: an open direction whose length is not a multiple of 8 is
  // stored padded to np[d] points with pf[d] of them in front of the real
  // ones (decoupled rows of the line operators); N counts the padded points
  int np[3] = {0, 0, 0}, pf[3] = {0, 0, 0};
  bool padded = false;
  long N = 0;
  SweepGeom geom[3]{};
  int wg[3] = {0, 0, 0};
  SweepGeom fgeom[3]{};  // filter passes (FP32: up to 32 y/z lines per work-group)
  int fwg[3] = {0, 0, 0};
  SweepGeom sgeom[3]{};  // Cartesian viscous sweeps (FP32: up to 32 y/z lines per work-group)
  int swg[3] = {0, 0, 0};
  DevOp<double> dop64[3], fop64[3];
  DevOp<float> dop32[3], fop32[3];
  Phys phys{};
  // device state
  double *qn = nullptr, *acc = nullptr, *qref = nullptr, *shape = nullptr, *tmp = nullptr;
  // boundary_conditions.dat rows in file order (types 201, 400, 401)
  std::vector<std::pair<int, BoxRow>> bc;
  bool has_bcq = false, has_bcrhs = false;
  // rows that act on the complete RHS (100 wall, 204 characteristic): the
  // last sweep then writes the RHS, the rows follow, then the RK4 update
  bool has_bclate = false, has_wall = false;
  // boundary normals per point, kvector(n1,n2,n3,3) (types 202, 203, 205, 206)
  bool needs_kvec = false;
  // zonal rows (203, 206) weight by the sponge shape over the whole domain
  bool needs_shape = false;
  // column 8 of the last 203 row (init_boundary_conditions,
  // boundary_conditions.f90:434-438): -1 acts only in the call after the RK
  // step, 1 only in the RHS's own call, 0 in both
  int zone_flag = 0;
  // rows 204-206 stop the reference's adjoint (set_boundary_condition_rhss)
  bool has_char_rhs = false;
  double *kvec = nullptr;
  // deformed grid (cruna_hip_set_metric): per-line metric entries and Jacobian
  void *met = nullptr;
  Metric<double> mt64{};
  Metric<float> mt32{};
  bool gt = false;
  // general curvilinear grid: the full metric and Jacobian per point in `met`
  // (gt is set as well), and the physical temperature gradient (3 fields)
  bool curv = false;
  CurvMetric<double> cm64{};
  CurvMetric<float> cm32{};
  void *tg = nullptr;
  // volume penalisation (cruna_hip_set_penalization): phi and 1/phi per
  // point in the compute precision
  bool ev = false;
  void *pen = nullptr;
  // bumped by every setter that changes device buffers, kernel choices or
  // launch shapes, so that an ensemble re-records graphs made before
  unsigned gen = 0;
  // FP64 phi, the weight of the weighted filter when a penalisation is set
  // (filter_weighted_pen_core_o2_e_c_p init_filter), and its output buffer
  double *wphi = nullptr, *wtmp = nullptr, *wbuf = nullptr;
  Penal<double> pe64{};
  Penal<float> pe32{};
  // adjoint state qs^n and the forward trajectory Q(:,:,:,:,nt) (slot nt-1)
  double *qa = nullptr, *traj = nullptr, *gobj = nullptr;
  // device cells: the forward state of the current adjoint step and the
  // backward step counter (k_adjoint_slot)
  const double **q0p = nullptr;
  int *ntp = nullptr;
  // subset counter of an ensemble member running the subset adjoint loop
  // inside recorded graphs: the checkpoint to restore next
  int *sckp = nullptr;
  int traj_slots = 0;
  bool record = false;
  // subset checkpoints: the state at the start of every subset and the final one
  double *ckpt = nullptr;
  int ckpt_n = 0;
  int sub_steps = 0;  // steps per subset (cruna_hip_subsets_alloc)
  void *qs = nullptr, *rhs = nullptr, *G = nullptr;  // compute precision
  // learned closure added to the direct RHS (cruna_hip_set_closure)
  cruna_hip_cnn *cnn = nullptr;
  void *force = nullptr;  // its term, 5 fields in the compute precision
  // qs holds qn in the compute precision (FP32-mixed: written by the step's
  // last filter pass, read by the next step's first stage instead of qn)
  bool qs_fresh = false;
  // convergence check of the cruna_hpc loop (run.cfreq, run.ctol): the state
  // before the step (q0 copy, 5 fields), partial maxima and the residual
  int cfreq = 0;
  double ctol = 0.0;
  double *q0copy = nullptr, *res_buf = nullptr;
  double last_res = -1.0;
  int res_nt = 0, conv_nt = 0;
  // profiling
  bool prof = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> pool;
  double prof_ms[KID_COUNT] = {};
  long prof_n[KID_COUNT] = {};
};

namespace {

// Lines per work-group: TL lines of NB threads each (NB * TL <= wg_max); y/z
// tiles are TL neighbouring x-positions, so their rows are TL elements long.
SweepGeom make_geom(const int *np, int d, int TL) {
  SweepGeom g{};
  const int n1 = np[0], n2 = np[1], n3 = np[2];
  g.dir = d;
  g.n = np[d];
  g.NB = np[d] / PPT;
  g.n1 = n1;
  g.n2 = n2;
  g.n3 = n3;
  g.TL = TL;
  if (d == 0) {
    g.PS = 1;
    g.ntiles = (int)((long)n2 * n3 / TL);
    g.ntiles_in = 1;
  } else {
    g.PS = d == 1 ? n1 : (long)n1 * n2;
    g.ntiles_in = n1 / TL;
    g.ntiles = g.ntiles_in * (d == 1 ? n3 : n2);
  }
  return g;
}

int auto_tl(const int *np, int d, int wg_max, int tl_max) {
  const long lines_contig = d == 0 ? (long)np[1] * np[2] : np[0];
  const int NB = np[d] / PPT;
  int TL = 1;
  while (NB * TL * 2 <= wg_max && lines_contig % (TL * 2) == 0 && TL * 2 <= tl_max) TL *= 2;
  return TL;
}

int build_geom(cruna_hip_ctx *c, int d) {
  const cruna_hip_cfg &cfg = c->cfg;
  const int *np = c->np;
  const long lines_contig = d == 0 ? (long)np[1] * np[2] : np[0];
  const int NB = np[d] / PPT;
  int TL = cfg.tile_lines[d];
  if (TL == 0) {
    // x-lines are contiguous: up to 256 threads; y/z tiles take up to 16
    // neighbouring x-lines (128-byte rows in FP64) in up to 512 threads
    TL = d == 0 ? auto_tl(np, d, 256, 1 << 20) : auto_tl(np, d, NT, 16);
  } else if (TL < 0 || (TL & (TL - 1)) != 0 || lines_contig % TL != 0 || NB * TL > NT) {
    return fail(CRUNA_HIP_EINVAL, "tile_lines[" + std::to_string(d) + "] = " + std::to_string(TL) +
                                      ": needs a power of two dividing " +
                                      std::to_string(lines_contig) + " with (n/8)*tile_lines <= " +
                                      std::to_string(NT));
  }
  c->geom[d] = make_geom(np, d, TL);
  c->wg[d] = NB * TL;
  // filter passes: FP32 y/z tiles up to 32 lines (128-byte rows of floats)
  int FTL = TL;
  if (d != 0 && cfg.precision == CRUNA_HIP_FP32_MIXED) FTL = auto_tl(np, d, NT_FILTER32, 32);
  c->fgeom[d] = make_geom(np, d, FTL);
  c->fwg[d] = NB * FTL;
  // Cartesian viscous sweeps: in FP32, y/z tiles of up to 32 lines as well,
  // unless the tile size is given (the autotuner's choice holds for all);
  // measured on the 256^3 step, the first-derivative y sweep is faster with
  // the 16-line tiles and the viscous y and z sweeps with 32
  int STL = TL;
  if (d != 0 && cfg.precision == CRUNA_HIP_FP32_MIXED && cfg.tile_lines[d] == 0)
    STL = auto_tl(np, d, NT_SWEEP32, 32);
  c->sgeom[d] = make_geom(np, d, STL);
  c->swg[d] = NB * STL;
  return CRUNA_HIP_OK;
}

hipEvent_t get_event(cruna_hip_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

struct ProfScope {
  cruna_hip_ctx *c;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(cruna_hip_ctx *c_, int k) : c(c_), kid(k) {
    if (c->prof) {
      a = get_event(c);
      b = get_event(c);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~ProfScope() {
    if (c->prof) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({kid, {a, b}});
    }
  }
};

int collect_profile(cruna_hip_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    c->prof_ms[p.first] += ms;
    c->prof_n[p.first] += 1;
    c->pool.push_back(p.second.first);
    c->pool.push_back(p.second.second);
  }
  c->pending.clear();
  return CRUNA_HIP_OK;
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CRUNA_HIP_EDEVICE, std::string("launch: ") + hipGetErrorString(e));
  return CRUNA_HIP_OK;
}

// Launch KERNEL<..., DIR, PER> for a runtime direction d with the periodic/open
// specialisation of that direction.
#define DISPATCH_DIR(c, d, KER, TARGS, ...)                                                 \
  do {                                                                                      \
    const dim3 grid_((c)->geom[d].ntiles), block_((c)->wg[d]);                              \
    const bool per_ = (c)->cfg.scheme[d] == CRUNA_HIP_PERIODIC;                            \
    if (d == 0) {                                                                           \
      if (per_) hipLaunchKernelGGL((KER<TARGS, 0, true>), grid_, block_, 0, (c)->stream, __VA_ARGS__); \
      else hipLaunchKernelGGL((KER<TARGS, 0, false>), grid_, block_, 0, (c)->stream, __VA_ARGS__);     \
    } else if (d == 1) {                                                                    \
      if (per_) hipLaunchKernelGGL((KER<TARGS, 1, true>), grid_, block_, 0, (c)->stream, __VA_ARGS__); \
      else hipLaunchKernelGGL((KER<TARGS, 1, false>), grid_, block_, 0, (c)->stream, __VA_ARGS__);     \
    } else {                                                                                \
      if (per_) hipLaunchKernelGGL((KER<TARGS, 2, true>), grid_, block_, 0, (c)->stream, __VA_ARGS__); \
      else hipLaunchKernelGGL((KER<TARGS, 2, false>), grid_, block_, 0, (c)->stream, __VA_ARGS__);     \
    }                                                                                       \
  } while (0)

template <typename T>
DevOp<T> *dops(cruna_hip_ctx *c);
template <>
DevOp<double> *dops<double>(cruna_hip_ctx *c) { return c->dop64; }
template <>
DevOp<float> *dops<float>(cruna_hip_ctx *c) { return c->dop32; }
template <typename T>
DevOp<T> *fops(cruna_hip_ctx *c);
template <>
DevOp<double> *fops<double>(cruna_hip_ctx *c) { return c->fop64; }
template <>
DevOp<float> *fops<float>(cruna_hip_ctx *c) { return c->fop32; }

template <typename T>
Metric<T> metric(cruna_hip_ctx *c);
template <>
Metric<double> metric<double>(cruna_hip_ctx *c) { return c->mt64; }
template <>
Metric<float> metric<float>(cruna_hip_ctx *c) { return c->mt32; }

template <typename T>
CurvMetric<T> curv_metric(cruna_hip_ctx *c);
template <>
CurvMetric<double> curv_metric<double>(cruna_hip_ctx *c) { return c->cm64; }
template <>
CurvMetric<float> curv_metric<float>(cruna_hip_ctx *c) { return c->cm32; }

template <typename T>
Penal<T> penal(cruna_hip_ctx *c);
template <>
Penal<double> penal<double>(cruna_hip_ctx *c) { return c->pe64; }
template <>
Penal<float> penal<float>(cruna_hip_ctx *c) { return c->pe32; }

// ns_rho_rhou_p_ev_phi.f90:295-309: the penalised module's adjoint RHS stops
// with "not yet implemented"
// the reference's penalised Euler module exists for deformed grids only
// (euler_rho_rhou_p_ev_phi_grid_transform; there is no euler_rho_rhou_p_ev_phi)
const char *kEulerCartesian =
    "the penalised Euler equations need a metric (the reference has no Cartesian "
    "euler_rho_rhou_p_ev_phi module)";
// euler_rho_rhou_p_grid_transform.f90:158-509 has its own metric adjoint,
// which the device does not build (the direct RHS is the NS one with mu = 0)
const char *kNoEulerGtAdjoint = "the Euler adjoint on a deformed grid is not built on the device";
// set_boundary_condition_rhss (boundary_conditions.f90:339-353) has its
// cases 204-206 commented out, so the adjoint reaches case default and stops
const char *kNoCharRhsAdjoint =
    "rows of types 204-206 stop the reference's adjoint (set_boundary_condition_rhss has no case for them)";
const char *kNoEvAdjoint =
    "the adjoint of ns_rho_rhou_p_ev_phi is not implemented (the reference stops there too)";

// Configurations the adjoint entry points refuse, checked before anything is
// launched: the penalised adjoint, rows 204-206, the Euler adjoint on a
// deformed grid, and inputs that rows of the adjoint need (the sponge shape
// of 400/401 rows, whose Euler-adjoint RHS term and adjoint-state relaxation
// read it; the normals per point of rows 202/203).
int adjoint_checks(cruna_hip_ctx *c) {
  if (c->ev) return fail(CRUNA_HIP_EINVAL, kNoEvAdjoint);
  if (c->has_char_rhs) return fail(CRUNA_HIP_EINVAL, kNoCharRhsAdjoint);
  if (c->cfg.equations == CRUNA_HIP_EQ_EULER && c->gt) return fail(CRUNA_HIP_EINVAL, kNoEulerGtAdjoint);
  for (auto &r : c->bc) {
    const bool euler_sponge = r.first == 400 && c->cfg.equations == CRUNA_HIP_EQ_EULER;
    if ((euler_sponge || r.first == 401 || r.first == 203) && !c->shape)
      return fail(CRUNA_HIP_EINVAL, "sponge rows need a shape (cruna_hip_set_sponge_shape)");
    if ((r.first == 202 || r.first == 203) && !c->kvec)
      return fail(CRUNA_HIP_EINVAL, "rows of types 202/203/205/206 need the normals per point (cruna_hip_set_kvector)");
  }
  return CRUNA_HIP_OK;
}

// X_jacobian in the compute precision on a deformed grid, else null
template <typename T>
const T *jacobian(cruna_hip_ctx *c) {
  if (!c->gt) return nullptr;
  return c->curv ? curv_metric<T>(c).J : metric<T>(c).J;
}

int need_bc_inputs(cruna_hip_ctx *c) {
  if (c->needs_kvec && !c->kvec)
    return fail(CRUNA_HIP_EINVAL, "rows of types 202/203/205/206 need the normals per point (cruna_hip_set_kvector)");
  if (c->needs_shape && !c->shape)
    return fail(CRUNA_HIP_EINVAL, "zonal rows (203/206) need a shape (cruna_hip_set_sponge_shape)");
  for (auto &r : c->bc) {
    if (!c->qref) return fail(CRUNA_HIP_EINVAL, "boundary rows need qref (cruna_hip_set_qref)");
    if ((r.first == 400 || r.first == 401) && !c->shape)
      return fail(CRUNA_HIP_EINVAL, "sponge rows need a shape (cruna_hip_set_sponge_shape)");
  }
  return CRUNA_HIP_OK;
}

// does a 203 row act in this call (outside: the one after the RK step)?
bool zone_applies(const cruna_hip_ctx *c, bool outside) {
  return outside ? c->zone_flag != 1 : c->zone_flag != -1;
}

// set_boundary_condition_q (boundary_conditions.f90:52-136) on q, rows in
// order; outside: the call with outside_rhs=.true. after the RK step.
template <typename Q>
int bc_q(cruna_hip_ctx *c, Q *q, bool outside = false) {
  if (int rc = need_bc_inputs(c)) return rc;
  for (auto &r : c->bc) {
    if (r.first == 201) {
      const BoxRow &b = r.second;
      const long np = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
      hipLaunchKernelGGL((k_bc_characteristic_q<Q, double>), dim3((np + 255) / 256), dim3(256), 0,
                         c->stream, q, (const double *)c->qref, b, c->np[0], c->np[1], c->N,
                         c->cfg.gamma, (const double *)nullptr, (const double *)nullptr);
    } else if (r.first == 202 || r.first == 203) {
      if (r.first == 203 && !zone_applies(c, outside)) continue;
      const BoxRow &b = r.second;
      const long np = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
      hipLaunchKernelGGL((k_bc_characteristic_q<Q, double>), dim3((np + 255) / 256), dim3(256), 0,
                         c->stream, q, (const double *)c->qref, b, c->np[0], c->np[1], c->N,
                         c->cfg.gamma, (const double *)c->kvec,
                         r.first == 203 ? (const double *)c->shape : (const double *)nullptr);
    } else if (r.first == 200) {
      const BoxRow &b = r.second;
      const long np = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
      hipLaunchKernelGGL((k_bc_characteristic_q<Q, Q>), dim3((np + 255) / 256), dim3(256), 0,
                         c->stream, q, (const Q *)q, b, c->np[0], c->np[1], c->N, c->cfg.gamma,
                         (const double *)nullptr, (const double *)nullptr);
    } else if (r.first == 401) {
      hipLaunchKernelGGL((k_bc_sponge_q<Q>), dim3((c->N + 255) / 256), dim3(256), 0, c->stream, q,
                         (const double *)c->qref, (const double *)c->shape, c->N);
    }
    if (int rc = check_launch()) return rc;
  }
  return CRUNA_HIP_OK;
}

// set_boundary_condition_rhs (boundary_conditions.f90:228-298), rows in order.
// Without 100/204 rows only the sponge (400) is there; it is linear and runs
// before the last sweep, its share pre-divided by what that sweep multiplies
// in.  With them (late), every row acts on the complete RHS.
template <typename T, typename Q>
int bc_rhs(cruna_hip_ctx *c, T *rhs, const Q *q, bool late = false) {
  if (int rc = need_bc_inputs(c)) return rc;
  const int n1 = c->np[0], n2 = c->np[1];
  for (auto &r : c->bc) {
    const BoxRow &b = r.second;
    const long np = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
    if (r.first == 400) {
      hipLaunchKernelGGL((k_sponge_rhs<T, Q>), dim3((c->N + 255) / 256), dim3(256), 0, c->stream,
                         rhs, q, (const double *)c->qref, (const double *)c->shape, c->N,
                         late ? 1.0 : c->cfg.gamma - 1.0, late ? nullptr : jacobian<T>(c));
    } else if (late && (r.first == 204 || r.first == 205 || r.first == 206)) {
      hipLaunchKernelGGL((k_bc_characteristic_rhs<T, Q>), dim3((np + 255) / 256), dim3(256), 0,
                         c->stream, rhs, q, (const double *)c->qref, b, n1, n2, c->N, c->cfg.gamma,
                         c->cfg.dt, r.first == 204 ? (const double *)nullptr : (const double *)c->kvec,
                         r.first == 206 ? (const double *)c->shape : (const double *)nullptr);
    } else if (late && r.first == 100) {
      hipLaunchKernelGGL((k_bc_wall_rhs<T>), dim3((np + 255) / 256), dim3(256), 0, c->stream, rhs, b,
                         (int)b.k[0], n1, n2, c->N);
    } else {
      continue;
    }
    if (int rc = check_launch()) return rc;
  }
  return CRUNA_HIP_OK;
}

// RK4 stage update of all five variables from the complete RHS in rhs, point
// by point: rk_update's arithmetic (the accumulator of type T: FP64 the
// partial new state, FP32 the increment alone), for stages whose RHS rows
// need the complete RHS first.
template <typename T>
__global__ __launch_bounds__(256) void k_rk_stage(int mode, const T *__restrict__ rhs,
                                                   double *__restrict__ qn, T *__restrict__ acc,
                                                   T *__restrict__ qs_out, double dt, long N) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= N) return;
  const double base = sizeof(T) == 8 ? 1.0 : 0.0;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const long p = v * N + o;
    const double k = double(rhs[p]);
    if (mode == 3) {
      qn[p] = sizeof(T) == 8 ? double(acc[p]) + dt / 6.0 * k : qn[p] + (double(acc[p]) + dt / 6.0 * k);
      continue;
    }
    const double q0 = qn[p];
    const double a = mode == 0 ? base * q0 + dt / 6.0 * k : double(acc[p]) + dt / 3.0 * k;
    acc[p] = T(a);
    const double c_stage = mode == 2 ? dt : dt * 0.5;
    qs_out[p] = T(q0 + c_stage * k);
  }
}

template <typename T, typename Q>
int closure_term(cruna_hip_ctx *c, const Q *in);

template <typename T, typename Q, int DIR, int FORM>
void launch_sweep1_form(cruna_hip_ctx *c, const Q *in) {
  const dim3 grid(c->geom[DIR].ntiles), block(c->wg[DIR]);
  T *rhs = (T *)c->rhs, *G = (T *)c->G;
  const LineOp<T> &op = dops<T>(c)[DIR].op;
  if (c->cfg.scheme[DIR] == CRUNA_HIP_PERIODIC)
    hipLaunchKernelGGL((k_sweep1<T, Q, DIR, true, FORM>), grid, block, 0, c->stream, in, rhs, G, op,
                       c->geom[DIR], c->phys, c->N, metric<T>(c), penal<T>(c));
  else
    hipLaunchKernelGGL((k_sweep1<T, Q, DIR, false, FORM>), grid, block, 0, c->stream, in, rhs, G, op,
                       c->geom[DIR], c->phys, c->N, metric<T>(c), penal<T>(c));
}

template <typename T, typename Q, int DIR>
void launch_curv_sweep1(cruna_hip_ctx *c, const Q *in) {
  const dim3 grid(c->geom[DIR].ntiles), block(c->wg[DIR]);
  T *rhs = (T *)c->rhs, *G = (T *)c->G, *TG = (T *)c->tg;
  const LineOp<T> &op = dops<T>(c)[DIR].op;
  const bool per = c->cfg.scheme[DIR] == CRUNA_HIP_PERIODIC;
  if (per && c->ev)
    hipLaunchKernelGGL((k_curv_sweep1<T, Q, DIR, true, true>), grid, block, 0, c->stream, in, rhs, G, TG, op,
                       c->geom[DIR], c->phys, c->N, curv_metric<T>(c), penal<T>(c));
  else if (per)
    hipLaunchKernelGGL((k_curv_sweep1<T, Q, DIR, true, false>), grid, block, 0, c->stream, in, rhs, G, TG, op,
                       c->geom[DIR], c->phys, c->N, curv_metric<T>(c), penal<T>(c));
  else if (c->ev)
    hipLaunchKernelGGL((k_curv_sweep1<T, Q, DIR, false, true>), grid, block, 0, c->stream, in, rhs, G, TG, op,
                       c->geom[DIR], c->phys, c->N, curv_metric<T>(c), penal<T>(c));
  else
    hipLaunchKernelGGL((k_curv_sweep1<T, Q, DIR, false, false>), grid, block, 0, c->stream, in, rhs, G, TG, op,
                       c->geom[DIR], c->phys, c->N, curv_metric<T>(c), penal<T>(c));
}

// the last first-derivative sweep and the first viscous sweep (both along z)
// as one kernel over the same z-line tiles (k_sweep12_z)
template <typename T, typename Q, int FORM>
void launch_sweep12z_form(cruna_hip_ctx *c, const Q *in) {
  const dim3 grid(c->geom[2].ntiles), block(c->wg[2]);
  T *rhs = (T *)c->rhs, *G = (T *)c->G;
  const LineOp<T> &op = dops<T>(c)[2].op;
  if (c->cfg.scheme[2] == CRUNA_HIP_PERIODIC)
    hipLaunchKernelGGL((k_sweep12_z<T, Q, true, FORM>), grid, block, 0, c->stream, in, rhs, G, op, c->geom[2],
                       c->phys, c->N, metric<T>(c), penal<T>(c));
  else
    hipLaunchKernelGGL((k_sweep12_z<T, Q, false, FORM>), grid, block, 0, c->stream, in, rhs, G, op, c->geom[2],
                       c->phys, c->N, metric<T>(c), penal<T>(c));
}

template <typename T, typename Q>
int launch_sweep12z(cruna_hip_ctx *c, const Q *in) {
  ProfScope p(c, KID_S12Z);
  if (c->gt)
    launch_sweep12z_form<T, Q, RHS_GT>(c, in);
  else if (c->ev)
    launch_sweep12z_form<T, Q, RHS_EV>(c, in);
  else
    launch_sweep12z_form<T, Q, RHS_CART>(c, in);
  return check_launch();
}

template <typename T, typename Q, int DIR>
int launch_sweep1(cruna_hip_ctx *c, const Q *in) {
  ProfScope p(c, DIR == 0 ? KID_S1X : (DIR == 1 ? KID_S2Y : KID_S3Z));
  if (c->curv)
    launch_curv_sweep1<T, Q, DIR>(c, in);
  else if (c->gt)
    launch_sweep1_form<T, Q, DIR, RHS_GT>(c, in);
  else if (c->ev)
    launch_sweep1_form<T, Q, DIR, RHS_EV>(c, in);
  else
    launch_sweep1_form<T, Q, DIR, RHS_CART>(c, in);
  return check_launch();
}

template <typename T, int DIR, int FORM>
void launch_sweep2_form(cruna_hip_ctx *c, int mode, bool fuse_filter) {
  const dim3 grid(c->sgeom[DIR].ntiles), block(c->swg[DIR]);
  T *rhs = (T *)c->rhs, *qs = (T *)c->qs;
  const T *G = (const T *)c->G;
  const LineOp<T> &op = dops<T>(c)[DIR].op, &fop = fops<T>(c)[DIR].op;
  const T *force = (const T *)(DIR == 0 && c->cnn ? c->force : nullptr);
  const bool per = c->cfg.scheme[DIR] == CRUNA_HIP_PERIODIC;
  // the fused x filter exists on the last stage of the x sweep only
  if constexpr (DIR == 0) {
    if (fuse_filter && mode == 3) {
      if (per)
        hipLaunchKernelGGL((k_sweep2<T, DIR, true, FORM, true>), grid, block, 0, c->stream, rhs, G, op, fop,
                           c->sgeom[DIR], c->phys, c->N, mode, c->qn, (T *)c->acc, qs, c->cfg.dt, metric<T>(c), force,
                           penal<T>(c));
      else
        hipLaunchKernelGGL((k_sweep2<T, DIR, false, FORM, true>), grid, block, 0, c->stream, rhs, G, op, fop,
                           c->sgeom[DIR], c->phys, c->N, mode, c->qn, (T *)c->acc, qs, c->cfg.dt, metric<T>(c), force,
                           penal<T>(c));
      return;
    }
  }
  if (per) {
    hipLaunchKernelGGL((k_sweep2<T, DIR, true, FORM, false>), grid, block, 0, c->stream, rhs, G, op, fop,
                       c->sgeom[DIR], c->phys, c->N, mode, c->qn, (T *)c->acc, qs, c->cfg.dt, metric<T>(c), force,
                           penal<T>(c));
  } else {
    hipLaunchKernelGGL((k_sweep2<T, DIR, false, FORM, false>), grid, block, 0, c->stream, rhs, G, op, fop,
                       c->sgeom[DIR], c->phys, c->N, mode, c->qn, (T *)c->acc, qs, c->cfg.dt, metric<T>(c), force,
                           penal<T>(c));
  }
}

template <typename T, int DIR, bool PER, bool FF>
void launch_curv_sweep2_k(cruna_hip_ctx *c, int mode) {
  const dim3 grid(c->geom[DIR].ntiles), block(c->wg[DIR]);
  const T *force = (const T *)(DIR == 0 && c->cnn ? c->force : nullptr);
  if (c->ev)
    hipLaunchKernelGGL((k_curv_sweep2<T, DIR, PER, FF, true>), grid, block, 0, c->stream, (T *)c->rhs,
                       (const T *)c->G, (const T *)c->tg, dops<T>(c)[DIR].op, fops<T>(c)[DIR].op,
                       c->geom[DIR], c->phys, c->N, mode, c->qn, (T *)c->acc, (T *)c->qs, c->cfg.dt,
                       curv_metric<T>(c), force, penal<T>(c));
  else
    hipLaunchKernelGGL((k_curv_sweep2<T, DIR, PER, FF, false>), grid, block, 0, c->stream, (T *)c->rhs,
                       (const T *)c->G, (const T *)c->tg, dops<T>(c)[DIR].op, fops<T>(c)[DIR].op,
                       c->geom[DIR], c->phys, c->N, mode, c->qn, (T *)c->acc, (T *)c->qs, c->cfg.dt,
                       curv_metric<T>(c), force, penal<T>(c));
}

template <typename T, int DIR>
void launch_curv_sweep2(cruna_hip_ctx *c, int mode, bool fuse_filter) {
  const bool per = c->cfg.scheme[DIR] == CRUNA_HIP_PERIODIC;
  // the fused x filter exists on the last stage of the x sweep only
  if constexpr (DIR == 0) {
    if (fuse_filter && mode == 3) {
      if (per)
        launch_curv_sweep2_k<T, DIR, true, true>(c, mode);
      else
        launch_curv_sweep2_k<T, DIR, false, true>(c, mode);
      return;
    }
  }
  if (per)
    launch_curv_sweep2_k<T, DIR, true, false>(c, mode);
  else
    launch_curv_sweep2_k<T, DIR, false, false>(c, mode);
}

template <typename T, int DIR>
int launch_sweep2(cruna_hip_ctx *c, int mode, bool fuse_filter) {
  ProfScope p(c, DIR == 2 ? KID_S4Z : (DIR == 1 ? KID_S5Y : KID_S6X));
  if (c->curv)
    launch_curv_sweep2<T, DIR>(c, mode, fuse_filter);
  else if (c->gt)
    launch_sweep2_form<T, DIR, RHS_GT>(c, mode, fuse_filter);
  else if (c->ev)
    launch_sweep2_form<T, DIR, RHS_EV>(c, mode, fuse_filter);
  else
    launch_sweep2_form<T, DIR, RHS_CART>(c, mode, fuse_filter);
  return check_launch();
}

// One RHS evaluation (mode 4) or RK stage (mode 0..3) from stage input `in`:
// first-derivative sweeps x, y, z, viscous sweeps z, y, x (the last one
// finishes the stage).  fuse_filter: apply filter_i1 inside the final stage.
template <typename T, typename Q>
int rhs_stage(cruna_hip_ctx *c, const Q *in, int mode, bool fuse_filter = false) {
  if (c->ev && !c->gt && c->cfg.equations == CRUNA_HIP_EQ_EULER)
    return fail(CRUNA_HIP_EINVAL, kEulerCartesian);
  if (int rc = launch_sweep1<T, Q, 0>(c, in)) return rc;
  if (int rc = launch_sweep1<T, Q, 1>(c, in)) return rc;
  if (c->curv) {
    if (int rc = launch_sweep1<T, Q, 2>(c, in)) return rc;
    if (int rc = launch_sweep2<T, 2>(c, mode, false)) return rc;
  } else {
    if (int rc = launch_sweep12z<T, Q>(c, in)) return rc;
  }
  if (int rc = launch_sweep2<T, 1>(c, mode, false)) return rc;
  if (c->has_bcrhs && !c->has_bclate)
    if (int rc = bc_rhs<T, Q>(c, (T *)c->rhs, in)) return rc;
  if (c->cnn)
    if (int rc = closure_term<T, Q>(c, in)) return rc;
  if (!c->has_bclate) return launch_sweep2<T, 0>(c, mode, fuse_filter);
  // complete RHS, the rows on it, then the stage update
  if (int rc = launch_sweep2<T, 0>(c, 4, false)) return rc;
  if (int rc = bc_rhs<T, Q>(c, (T *)c->rhs, in, true)) return rc;
  if (mode == 4) return CRUNA_HIP_OK;
  hipLaunchKernelGGL((k_rk_stage<T>), dim3((c->N + 255) / 256), dim3(256), 0, c->stream, mode,
                     (const T *)c->rhs, c->qn, (T *)c->acc, (T *)c->qs, c->cfg.dt, c->N);
  return check_launch();
}

// A stage from the FP64 state `from` (stage 0, or a single RHS evaluation) or,
// with from == nullptr, from the stage input already in qs.  With q-type
// boundary rows the RHS works on a private copy (ns_rho_rhou_p.f90:89-92).
template <typename T>
int run_stage(cruna_hip_ctx *c, const double *from, int mode, bool fuse_filter = false) {
  T *qs = (T *)c->qs;
  const bool fresh = from == c->qn && c->qs_fresh;
  c->qs_fresh = false;  // the stage overwrites qs with the next stage input
  if (fresh) from = nullptr;
  if (from && !c->has_bcq) return rhs_stage<T, double>(c, from, mode, fuse_filter);
  if (from) {
    const long n = 5 * c->N;
    hipLaunchKernelGGL((k_copy_state<T>), dim3((n + 255) / 256), dim3(256), 0, c->stream, qs, from, n);
    if (int rc = check_launch()) return rc;
  }
  if (c->has_bcq)
    if (int rc = bc_q<T>(c, qs)) return rc;
  return rhs_stage<T, T>(c, (const T *)qs, mode, fuse_filter);
}

// RK4 step on qn; with fuse_filter the last stage also applies filter_i1.
template <typename T>
int rk4_step(cruna_hip_ctx *c, bool fuse_filter = false) {
  if (int rc = run_stage<T>(c, c->qn, 0)) return rc;
  for (int st = 1; st < 4; ++st)
    if (int rc = run_stage<T>(c, nullptr, st, st == 3 && fuse_filter)) return rc;
  return CRUNA_HIP_OK;
}

// copy: also store the filtered state in the compute precision there (or null)
template <typename T, int DIR, typename S>
void launch_filter(cruna_hip_ctx *c, const S *src, double *q64, int nv, T *copy) {
  const dim3 grid(c->fgeom[DIR].ntiles), block(c->fwg[DIR]);
  const LineOp<T> &op = fops<T>(c)[DIR].op;
  if (c->cfg.scheme[DIR] == CRUNA_HIP_PERIODIC)
    hipLaunchKernelGGL((k_filter<T, DIR, true, S>), grid, block, 0, c->stream, src, q64, op, c->fgeom[DIR],
                       c->N, nv, copy);
  else
    hipLaunchKernelGGL((k_filter<T, DIR, false, S>), grid, block, 0, c->stream, src, q64, op, c->fgeom[DIR],
                       c->N, nv, copy);
}

// one filter pass along d: from src (FP64 or the compute precision) into the
// FP64 fields q64 and/or the compute-precision fields copy
template <typename T, typename S>
int filter_pass(cruna_hip_ctx *c, const S *src, double *q64, int d, int nv, T *copy) {
  ProfScope p(c, KID_FX + d);
  if (d == 0) launch_filter<T, 0, S>(c, src, q64, nv, copy);
  if (d == 1) launch_filter<T, 1, S>(c, src, q64, nv, copy);
  if (d == 2) launch_filter<T, 2, S>(c, src, q64, nv, copy);
  return check_launch();
}

// the FP64 fields q filtered in place (and copied to `copy`)
template <typename T>
int filter_dir(cruna_hip_ctx *c, double *q, int d, int nv, T *copy = nullptr) {
  return filter_pass<T, double>(c, q, q, d, nv, copy);
}

// filter_weighted_iN on nv fields of the FP64 state q in place (through out)
int filter_weighted_dir(cruna_hip_ctx *c, double *q, double *out, int d, int nv, const double *w,
                        const double *s) {
  const dim3 grid((c->N + 255) / 256), block(256);
  const int n1 = c->np[0], n2 = c->np[1], n3 = c->np[2];
  if (d == 0)
    hipLaunchKernelGGL(k_filter_weighted<0>, grid, block, 0, c->stream, q, out, nv, w, s, n1, n2, n3, c->N);
  else if (d == 1)
    hipLaunchKernelGGL(k_filter_weighted<1>, grid, block, 0, c->stream, q, out, nv, w, s, n1, n2, n3, c->N);
  else
    hipLaunchKernelGGL(k_filter_weighted<2>, grid, block, 0, c->stream, q, out, nv, w, s, n1, n2, n3, c->N);
  if (int rc = check_launch()) return rc;
  HIP_TRY(hipMemcpyAsync(q, out, (size_t)nv * c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return CRUNA_HIP_OK;
}

}  // namespace
// the direct step loop of cruna_hpc (weighted filters) or computation_direct (none)
int run_direct_loop(cruna_hip_ctx *c, int nt0, int nsteps, bool weighted);
namespace {

// weighted: the cruna_hpc loop's filter_weighted_i1..i3 after the filters
// (cruna_hpc.f90:231-242); the subset loop (computation_direct.f90) has none.
// converged (may be null): the convergence check of the cruna_hpc loop
// (cruna_hpc.f90:185-208) at steps with mod(nt, cfreq) == 0; when the
// residual max|qn - q1|/dt is <= ctol the step ends with the unfiltered q1 in
// qn (the state the reference stores before it stops) and *converged is set.
int step_once(cruna_hip_ctx *c, int nt, bool weighted = true, bool *converged = nullptr) {
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  const bool check = converged && c->cfreq > 0 && nt % c->cfreq == 0;
  const size_t sbytes = 5 * (size_t)c->N * sizeof(double);
  if (check) HIP_TRY(hipMemcpyAsync(c->q0copy, c->qn, sbytes, hipMemcpyDeviceToDevice, c->stream));
  // filter_i1 directly follows the RK4 update unless q-type boundary rows
  // come between them (or the residual needs the unfiltered new state);
  // then it runs inside the update's x sweep
  const bool fuse_fx = !check && !c->has_bcq && !c->has_bclate && c->cfg.filter_freq[0] > 0 &&
                       nt % c->cfg.filter_freq[0] == 0;
  int rc = mixed ? rk4_step<float>(c, fuse_fx) : rk4_step<double>(c, fuse_fx);
  if (rc) return rc;
  if (check) {
    const long M = 5 * c->N;
    const int parts = (int)std::min<long>(RES_MAX_PARTS, (M / 2 + RES_THREADS - 1) / RES_THREADS);
    hipLaunchKernelGGL(k_absdiff_max, dim3(std::max(parts, 1)), dim3(RES_THREADS), 0, c->stream,
                       (const double *)c->q0copy, (const double *)c->qn, M, c->res_buf + 1);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(RES_THREADS), 0, c->stream,
                       (const double *)(c->res_buf + 1), std::max(parts, 1), c->cfg.dt, c->res_buf);
    if ((rc = check_launch())) return rc;
    double res = 0.0;
    HIP_TRY(hipMemcpyAsync(&res, c->res_buf, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->last_res = res;
    c->res_nt = nt;
    if (res <= c->ctol) {
      c->conv_nt = nt;
      c->qs_fresh = false;
      *converged = true;
      return CRUNA_HIP_OK;
    }
  }
  // set_boundary_condition_q(qn, outside_rhs=.true.) (cruna_hpc.f90:216)
  if (c->has_bcq)
    if ((rc = bc_q<double>(c, c->qn, true))) return rc;
  int last = -1;
  for (int d = fuse_fx ? 1 : 0; d < 3; ++d) {
    const int fq = c->cfg.filter_freq[d];
    if (fq > 0 && nt % fq == 0) last = d;
  }
  // FP32-mixed: a pass's output is FP32 by construction (FP32 arithmetic), so
  // a pass that another pass of the step follows leaves it in FP32 only (in
  // qs) and the next pass reads it from there; the step's last pass writes
  // the FP64 state and, for the next step's first stage, its FP32 copy
  bool in32 = false;
  for (int d = fuse_fx ? 1 : 0; d < 3; ++d) {
    const int fq = c->cfg.filter_freq[d];
    if (fq > 0 && nt % fq == 0) {
      if (mixed) {
        float *q32 = (float *)c->qs;
        double *q64 = d == last ? c->qn : nullptr;
        rc = in32 ? filter_pass<float, float>(c, q32, q64, d, 5, q32)
                  : filter_pass<float, double>(c, c->qn, q64, d, 5, q32);
        in32 = true;
      } else {
        rc = filter_dir<double>(c, c->qn, d, 5);
      }
      if (rc) return rc;
    }
  }
  c->qs_fresh = mixed && last >= 0;
  // weight phi of the penalisation (or 1), strength 1: init_filter of
  // filter_weighted_pen_core_o2_e_c_p.f90:18-37
  for (int d = 0; d < 3 && weighted; ++d) {
    const int fq = c->cfg.wfilter_freq[d];
    if (fq > 0 && nt % fq == 0) {
      if (!c->wtmp) HIP_TRY(hipMalloc(&c->wtmp, 5 * c->N * sizeof(double)));
      if ((rc = filter_weighted_dir(c, c->qn, c->wtmp, d, 5, c->ev ? c->wphi : nullptr, nullptr)))
        return rc;
      c->qs_fresh = false;  // the FP32 copy of the compact filter pass is stale now
    }
  }
  return CRUNA_HIP_OK;
}

// Host arrays hold the reference's layout f(n1,n2,n3[,nv]); device fields the
// padded one, real point (i,j,k) at (i+pf1, j+pf2, k+pf3) of (np1,np2,np3).
// The padding points hold a uniform state, which the line operators keep
// away from the real points and which stays constant (a zero RHS).
const double kRestState[5] = {1.0, 0.0, 0.0, 0.0, 1.0};

// nv host fields in the padded layout: the real points copied, the padding
// filled with fill[v] (fill null: 0) or, with replicate, with the value of
// the nearest real point (keeps a field constant along lines constant).
std::vector<double> pad_host(const cruna_hip_ctx *c, const double *h, int nv, const double *fill,
                             bool replicate = false) {
  const long n1 = c->cfg.n[0], n2 = c->cfg.n[1], n3 = c->cfg.n[2];
  const long p1 = c->np[0], p2 = c->np[1], p3 = c->np[2], N = c->N, Nr = n1 * n2 * n3;
  std::vector<double> out((size_t)nv * N);
  for (int v = 0; v < nv; ++v)
    for (long k = 0; k < p3; ++k)
      for (long j = 0; j < p2; ++j)
        for (long i = 0; i < p1; ++i) {
          long ri = i - c->pf[0], rj = j - c->pf[1], rk = k - c->pf[2];
          const bool real = ri >= 0 && ri < n1 && rj >= 0 && rj < n2 && rk >= 0 && rk < n3;
          double val;
          if (real || replicate) {
            ri = std::min(std::max(ri, 0L), n1 - 1);
            rj = std::min(std::max(rj, 0L), n2 - 1);
            rk = std::min(std::max(rk, 0L), n3 - 1);
            val = h[v * Nr + ri + n1 * (rj + n2 * rk)];
          } else {
            val = fill ? fill[v] : 0.0;
          }
          out[(size_t)v * N + i + p1 * (j + p2 * k)] = val;
        }
  return out;
}

// the real points of nv padded device-layout fields into a host array
void unpad_host(const cruna_hip_ctx *c, const std::vector<double> &d, double *h, int nv) {
  const long n1 = c->cfg.n[0], n2 = c->cfg.n[1], n3 = c->cfg.n[2];
  const long p1 = c->np[0], p2 = c->np[1], N = c->N, Nr = n1 * n2 * n3;
  for (int v = 0; v < nv; ++v)
    for (long k = 0; k < n3; ++k)
      for (long j = 0; j < n2; ++j)
        for (long i = 0; i < n1; ++i)
          h[v * Nr + i + n1 * (j + n2 * k)] =
              d[(size_t)v * N + (i + c->pf[0]) + p1 * ((j + c->pf[1]) + p2 * (k + c->pf[2]))];
}

// host -> device fields (asynchronous on the context stream when unpadded,
// as the plain copies were; waits for the staging copy when padded)
int put_fields(cruna_hip_ctx *c, double *dev, const double *host, int nv, const double *fill) {
  const size_t bytes = (size_t)nv * c->N * sizeof(double);
  if (!c->padded) {
    HIP_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
    return CRUNA_HIP_OK;
  }
  std::vector<double> v = pad_host(c, host, nv, fill);
  HIP_TRY(hipMemcpyAsync(dev, v.data(), bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

// device -> host fields (the real points)
int get_fields(cruna_hip_ctx *c, double *host, const double *dev, int nv) {
  const size_t bytes = (size_t)nv * c->N * sizeof(double);
  if (!c->padded) {
    HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    return CRUNA_HIP_OK;
  }
  std::vector<double> v((size_t)nv * c->N);
  HIP_TRY(hipMemcpyAsync(v.data(), dev, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  unpad_host(c, v, host, nv);
  return CRUNA_HIP_OK;
}

// Staging area of the host-pointer entry points: two states.
int ensure_tmp(cruna_hip_ctx *c) {
  if (!c->tmp) HIP_TRY(hipMalloc(&c->tmp, 10 * c->N * sizeof(double)));
  return CRUNA_HIP_OK;
}

// The RHS (compute precision) back to a host FP64 array.
int copy_rhs_out(cruna_hip_ctx *c, double *rhs) {
  if (c->cfg.precision == CRUNA_HIP_FP32_MIXED) {
    std::vector<float> h(5 * c->N);
    HIP_TRY(hipMemcpyAsync(h.data(), c->rhs, 5 * c->N * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->padded) {
      for (long i = 0; i < 5 * c->N; ++i) rhs[i] = h[i];
    } else {
      std::vector<double> d(h.begin(), h.end());
      unpad_host(c, d, rhs, 5);
    }
  } else {
    if (int rc = get_fields(c, rhs, (const double *)c->rhs, 5)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return CRUNA_HIP_OK;
}

// ---------------------------------------------------------------------------
// Adjoint

// set_boundary_condition_qs (boundary_conditions.f90:140-224) on qs, rows in
// order: 201 characteristic_qs toward zero with the forward state q, 401
// sponge_qs; 400 has no adjoint-state part.
template <typename Q>
int bc_qs(cruna_hip_ctx *c, Q *qs, bool outside = false) {
  for (auto &r : c->bc) {
    if (r.first == 201 || r.first == 200 || r.first == 202 || r.first == 203) {
      if (r.first == 203 && !zone_applies(c, outside)) continue;
      const BoxRow &b = r.second;
      const long np = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
      if (r.first >= 202 && !c->kvec)
        return fail(CRUNA_HIP_EINVAL, "rows of types 202/203/205/206 need the normals per point (cruna_hip_set_kvector)");
      if (r.first == 203 && !c->shape)
        return fail(CRUNA_HIP_EINVAL, "zonal rows (203/206) need a shape (cruna_hip_set_sponge_shape)");
      hipLaunchKernelGGL((k_bc_characteristic_qs<Q>), dim3((np + 255) / 256), dim3(256), 0,
                         c->stream, qs, (const double *const *)c->q0p, b, c->np[0], c->np[1],
                         c->N, c->cfg.gamma,
                         r.first >= 202 ? (const double *)c->kvec : (const double *)nullptr,
                         r.first == 203 ? (const double *)c->shape : (const double *)nullptr);
    } else if (r.first == 401) {
      if (!c->shape)
        return fail(CRUNA_HIP_EINVAL, "sponge rows need a shape (cruna_hip_set_sponge_shape)");
      hipLaunchKernelGGL((k_bc_sponge_qs<Q>), dim3((c->N + 255) / 256), dim3(256), 0, c->stream,
                         qs, (const double *)c->shape, c->N);
    } else {
      continue;
    }
    if (int rc = check_launch()) return rc;
  }
  return CRUNA_HIP_OK;
}

template <typename T, typename Q, int DIR>
int launch_asweep(cruna_hip_ctx *c, const Q *in, const double *g, int mode) {
  ProfScope p(c, DIR == 0 ? KID_AX : (DIR == 1 ? KID_AY : KID_AZ));
  const dim3 grid(c->geom[DIR].ntiles), block(c->wg[DIR]);
  T *rhs = (T *)c->rhs, *qs = (T *)c->qs;
  const double dt = -c->cfg.dt;  // backward in time (rk_o4_e.f90:62-105)
  if (c->cfg.scheme[DIR] == CRUNA_HIP_PERIODIC)
    hipLaunchKernelGGL((k_asweep<T, Q, DIR, true>), grid, block, 0, c->stream, in,
                       (const double *const *)c->q0p, rhs, g,
                       dops<T>(c)[DIR].op, c->geom[DIR], c->cfg.gamma, c->phys.inv_h[DIR], c->N,
                       mode, c->qa, (T *)c->acc, qs, dt);
  else
    hipLaunchKernelGGL((k_asweep<T, Q, DIR, false>), grid, block, 0, c->stream, in,
                       (const double *const *)c->q0p, rhs, g,
                       dops<T>(c)[DIR].op, c->geom[DIR], c->cfg.gamma, c->phys.inv_h[DIR], c->N,
                       mode, c->qa, (T *)c->acc, qs, dt);
  return check_launch();
}

// right_hand_side_adjoint_3d from stage input `in` with the forward state in
// the device cell q0p: sweeps z, y, x; the x sweep adds g and finishes the
// stage (mode as rhs_stage).
// With wall rows (100), or sponge rows (400) in the Euler adjoint, the x
// sweep writes the complete RHS (g added) and set_boundary_condition_rhss
// (boundary_conditions.f90:302-377) applies the rows to it in order: 100
// zeroes its variable, 400 adds sponge_rhss (the Euler adjoint passes the
// forward state, euler_rho_rhou_p.f90:384; the NS adjoint does not, and
// there 400 is the empty sponge_rhss); then the stage update runs point-wise.
template <typename T, typename Q>
int arhs_stage(cruna_hip_ctx *c, const Q *in, const double *g, int mode) {
  if (int rc = launch_asweep<T, Q, 2>(c, in, g, mode)) return rc;
  if (int rc = launch_asweep<T, Q, 1>(c, in, g, mode)) return rc;
  const bool euler_sponge = c->cfg.equations == CRUNA_HIP_EQ_EULER && c->has_bcrhs;
  if (!c->has_wall && !euler_sponge) return launch_asweep<T, Q, 0>(c, in, g, mode);
  if (int rc = launch_asweep<T, Q, 0>(c, in, g, 4)) return rc;
  const int n1 = c->np[0], n2 = c->np[1];
  for (auto &r : c->bc) {
    if (r.first == 400 && euler_sponge) {
      if (!c->shape) return fail(CRUNA_HIP_EINVAL, "sponge rows need a shape (cruna_hip_set_sponge_shape)");
      hipLaunchKernelGGL((k_sponge_rhss<T, Q>), dim3((c->N + 255) / 256), dim3(256), 0, c->stream,
                         (T *)c->rhs, in, (const double *const *)c->q0p, (const double *)c->shape,
                         c->cfg.gamma, c->N);
    } else if (r.first == 100) {
      const BoxRow &b = r.second;
      const long np = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
      hipLaunchKernelGGL((k_bc_wall_rhs<T>), dim3((np + 255) / 256), dim3(256), 0, c->stream,
                         (T *)c->rhs, b, (int)b.k[0], n1, n2, c->N);
    } else {
      continue;
    }
    if (int rc = check_launch()) return rc;
  }
  if (mode == 4) return CRUNA_HIP_OK;
  // backward in time, as the sweeps' own update (rk_o4_e.f90:62-105)
  hipLaunchKernelGGL((k_rk_stage<T>), dim3((c->N + 255) / 256), dim3(256), 0, c->stream, mode,
                     (const T *)c->rhs, c->qa, (T *)c->acc, (T *)c->qs, -c->cfg.dt, c->N);
  return check_launch();
}

// An adjoint stage from the FP64 state `from` or, with from == nullptr, from
// the stage input in qs.  With 201/401 rows the RHS works on a private copy
// to which set_boundary_condition_qs is applied (ns_rho_rhou_p.f90:288-290).
template <typename T>
int arun_stage(cruna_hip_ctx *c, const double *from, const double *g, int mode) {
  T *qs = (T *)c->qs;
  c->qs_fresh = false;
  if (from && !c->has_bcq) return arhs_stage<T, double>(c, from, g, mode);
  if (from) {
    const long n = 5 * c->N;
    hipLaunchKernelGGL((k_copy_state<T>), dim3((n + 255) / 256), dim3(256), 0, c->stream, qs, from, n);
    if (int rc = check_launch()) return rc;
  }
  if (c->has_bcq)
    if (int rc = bc_qs<T>(c, qs)) return rc;
  return arhs_stage<T, T>(c, (const T *)qs, g, mode);
}

// time_stepper_adjoint (rk_o4_e.f90:62-105) on the device adjoint state qa.
template <typename T>
int ark4_step(cruna_hip_ctx *c, const double *g) {
  if (int rc = arun_stage<T>(c, c->qa, g, 0)) return rc;
  for (int st = 1; st < 4; ++st)
    if (int rc = arun_stage<T>(c, nullptr, g, st)) return rc;
  return CRUNA_HIP_OK;
}

// One step of the calc_adjoint loop (computation_adjoint.f90:72-99) with g = 0:
// stepper -> BC on qs (outside rhs) -> adjoint filters where mod(nt, freq) == 0.
int astep_once(cruna_hip_ctx *c, int nt) {
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  // q0 = Q(:,:,:,:,nt) into the device cell; the counter then holds nt - 1
  hipLaunchKernelGGL(k_adjoint_slot, dim3(1), dim3(64), 0, c->stream, c->q0p, c->ntp,
                     (const double *)c->traj, 5 * c->N, c->traj_slots);
  if (int rc = check_launch()) return rc;
  int rc = mixed ? ark4_step<float>(c, nullptr) : ark4_step<double>(c, nullptr);
  if (rc) return rc;
  if (c->has_bcq)
    if ((rc = bc_qs<double>(c, c->qa, true))) return rc;
  for (int d = 0; d < 3; ++d) {
    const int fq = c->cfg.afilter_freq[d];
    if (fq > 0 && nt % fq == 0) {
      rc = mixed ? filter_dir<float>(c, c->qa, d, 5) : filter_dir<double>(c, c->qa, d, 5);
      if (rc) return rc;
    }
  }
  return CRUNA_HIP_OK;
}

int ensure_adjoint(cruna_hip_ctx *c) {
  if (!c->qa) HIP_TRY(hipMalloc(&c->qa, 5 * c->N * sizeof(double)));
  if (!c->q0p) {
    void *cells = nullptr;
    HIP_TRY(hipMalloc(&cells, 256));  // one pointer, two counters
    c->q0p = (const double **)cells;
    c->ntp = (int *)((char *)cells + 128);
    c->sckp = (int *)((char *)cells + 160);
  }
  return CRUNA_HIP_OK;
}

// The forward state of the next adjoint stages is the host-supplied state q.
int point_q0(cruna_hip_ctx *c, const double *q) {
  hipLaunchKernelGGL(k_set_state_ptr, dim3(1), dim3(64), 0, c->stream, c->q0p, q);
  return check_launch();
}

double *traj_slot(cruna_hip_ctx *c, int nt) { return c->traj + (size_t)(nt - 1) * 5 * c->N; }



template <typename T>
int upload_metric(cruna_hip_ctx *c, const std::vector<double> &m1, const std::vector<double> &m2,
                  const std::vector<double> &m3, const double *Xj, Metric<T> &out) {
  const size_t n1 = m1.size(), n2 = m2.size(), n3 = m3.size(), N = c->N;
  std::vector<T> h(n1 + n2 + n3 + N);
  size_t o = 0;
  for (double v : m1) h[o++] = T(v);
  for (double v : m2) h[o++] = T(v);
  for (double v : m3) h[o++] = T(v);
  for (size_t i = 0; i < N; ++i) h[o++] = T(Xj[i]);
  if (c->met) HIP_TRY(hipFree(c->met));
  c->met = nullptr;
  HIP_TRY(hipMalloc(&c->met, h.size() * sizeof(T)));
  HIP_TRY(hipMemcpy(c->met, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  T *b = (T *)c->met;
  out.m[0] = b;
  out.m[1] = b + n1;
  out.m[2] = b + n1 + n2;
  out.J = b + n1 + n2 + n3;
  return CRUNA_HIP_OK;
}

// The full metric (9 fields, X_metric's order) and the Jacobian per point in
// the compute precision, and the temperature-gradient fields of the
// curvilinear sweeps.
template <typename T>
int upload_curv_metric(cruna_hip_ctx *c, const double *Xm, const double *Xj, CurvMetric<T> &out) {
  const size_t N = c->N;
  std::vector<T> h(10 * N);
  for (size_t i = 0; i < 9 * N; ++i) h[i] = T(Xm[i]);
  for (size_t i = 0; i < N; ++i) h[9 * N + i] = T(Xj[i]);
  if (c->met) HIP_TRY(hipFree(c->met));
  c->met = nullptr;
  HIP_TRY(hipMalloc(&c->met, h.size() * sizeof(T)));
  HIP_TRY(hipMemcpy(c->met, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  if (!c->tg) HIP_TRY(hipMalloc(&c->tg, 3 * N * sizeof(T)));
  out.M = (const T *)c->met;
  out.J = (const T *)c->met + 9 * N;
  return CRUNA_HIP_OK;
}

}  // namespace

extern "C" {

void cruna_hip_default_cfg(cruna_hip_cfg *cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  for (int d = 0; d < 3; ++d) {
    cfg->n[d] = 64;
    cfg->dx[d] = 1.0;
    cfg->scheme[d] = CRUNA_HIP_PERIODIC;
    cfg->filter_freq[d] = 1;
    cfg->afilter_freq[d] = 0;
  }
  cfg->filter_alpha = 0.35;
  cfg->gamma = 1.4;
  cfg->R = 8.3144598;
  cfg->W = 0.0289647;
  cfg->Pr = 0.71;
  cfg->mu0 = 0.0;
  cfg->T0 = 273.0;
  cfg->visc_law = CRUNA_HIP_VISC_SUTHERLAND;
  cfg->visc_exp = 0.76;
  cfg->dt = 1e-3;
  cfg->heat_unit_spacing = 1;
  cfg->precision = CRUNA_HIP_FP64;
}

const char *cruna_hip_last_error(void) { return g_last_error.c_str(); }

int cruna_hip_create(const cruna_hip_cfg *cfg, cruna_hip_ctx **out) {
  if (!cfg || !out) return fail(CRUNA_HIP_EINVAL, "null argument");
  *out = nullptr;
  for (int d = 0; d < 3; ++d) {
    if (cfg->scheme[d] != CRUNA_HIP_PERIODIC && cfg->scheme[d] != CRUNA_HIP_OPEN)
      return fail(CRUNA_HIP_EINVAL, "unknown scheme");
    const bool per = cfg->scheme[d] == CRUNA_HIP_PERIODIC;
    if (per && (cfg->n[d] < 16 || cfg->n[d] % PPT != 0))
      return fail(CRUNA_HIP_EINVAL, "n" + std::to_string(d + 1) + " = " + std::to_string(cfg->n[d]) +
                                        ": a periodic direction needs a multiple of 8, >= 16");
    if (!per && cfg->n[d] < PPT + 1)
      return fail(CRUNA_HIP_EINVAL, "n" + std::to_string(d + 1) + " = " + std::to_string(cfg->n[d]) +
                                        ": an open direction needs at least 9 points");
    if (padded_length(per, cfg->n[d]) / PPT > 256)
      return fail(CRUNA_HIP_EINVAL, "grid size above 2048 points per line is not supported");
    if (cfg->diff_kind[d] != CRUNA_HIP_DIFF_O6_COMPACT && cfg->diff_kind[d] != CRUNA_HIP_DIFF_O4_EXPLICIT &&
        cfg->diff_kind[d] != CRUNA_HIP_DIFF_O4_PADE && cfg->diff_kind[d] != CRUNA_HIP_DIFF_O2_EXPLICIT)
      return fail(CRUNA_HIP_EINVAL, "unknown derivative core");
    if (cfg->diff_kind[d] == CRUNA_HIP_DIFF_O2_EXPLICIT && cfg->scheme[d] != CRUNA_HIP_OPEN)
      return fail(CRUNA_HIP_EINVAL, "the 2nd-order explicit core (diff_pen_core_o2_e_c_o) is for open lines only");
    if (!(cfg->dx[d] > 0)) return fail(CRUNA_HIP_EINVAL, "dx must be positive");
    if (cfg->wfilter_freq[d] < 0) return fail(CRUNA_HIP_EINVAL, "wfilter_freq must be >= 0");
    if (cfg->wfilter_freq[d] > 0 && padded_length(per, cfg->n[d]) != cfg->n[d])
      return fail(CRUNA_HIP_EINVAL, "the weighted filter's core is periodic: its direction needs a multiple of 8 points");
  }
  if (cfg->equations != CRUNA_HIP_EQ_NS && cfg->equations != CRUNA_HIP_EQ_EULER)
    return fail(CRUNA_HIP_EINVAL, "unknown equations");
  {
    double pts = 1.0;
    for (int d = 0; d < 3; ++d) pts *= padded_length(cfg->scheme[d] == CRUNA_HIP_PERIODIC, cfg->n[d]);
    if (pts * sizeof(double) >= 2147483648.0)
      return fail(CRUNA_HIP_EINVAL, "one FP64 field must fit in 2 GiB (the sweeps' buffer extent)");
  }
  if (cfg->precision != CRUNA_HIP_FP64 && cfg->precision != CRUNA_HIP_FP32_MIXED)
    return fail(CRUNA_HIP_EINVAL, "unknown precision");
  if (cfg->nbc < 0 || cfg->nbc > CRUNA_HIP_MAX_BC) return fail(CRUNA_HIP_EINVAL, "bad nbc");
  for (int r = 0; r < cfg->nbc; ++r) {
    const int t = (int)lround(cfg->bc[r][0]);
    if (t != 0 && t != 100 && (t < 200 || t > 206) && t != 400 && t != 401)
      return fail(CRUNA_HIP_EINVAL,
                  "boundary type " + std::to_string(t) + " not implemented on the device");
  }
  // partial contexts are released through cruna_hip_destroy on every error return
  std::unique_ptr<cruna_hip_ctx, int (*)(cruna_hip_ctx *)> c(new cruna_hip_ctx(), cruna_hip_destroy);
  c->cfg = *cfg;
  HIP_TRY(hipGetDevice(&c->device));
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int d = 0; d < 3; ++d) {
    const bool per = cfg->scheme[d] == CRUNA_HIP_PERIODIC;
    c->np[d] = padded_length(per, cfg->n[d]);
    c->pf[d] = front_padding(per, cfg->n[d]);
    c->padded = c->padded || c->np[d] != cfg->n[d];
  }
  c->N = (long)c->np[0] * c->np[1] * c->np[2];
  for (int r = 0; r < cfg->nbc; ++r) {
    const int t = (int)lround(cfg->bc[r][0]);
    if (t == 0) continue;
    BoxRow b;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      // 1-based inclusive indices, clamped to the grid (modify_boundary_conditions_array)
      lo[d] = std::max(1, (int)lround(cfg->bc[r][1 + 2 * d])) - 1;
      hi[d] = std::min(cfg->n[d], (int)lround(cfg->bc[r][2 + 2 * d])) - 1;
      b.k[d] = cfg->bc[r][7 + d];
    }
    if (t == 203 || t == 206) {
      // zonal rows act on the whole domain; their box columns are unused
      // (boundary_conditions.f90:93-108, 276-279)
      for (int d = 0; d < 3; ++d) {
        lo[d] = 0;
        hi[d] = cfg->n[d] - 1;
        b.k[d] = 0.0;
      }
      if (t == 203) {
        const int f = (int)lround(cfg->bc[r][7]);
        if (f < -1 || f > 1) return fail(CRUNA_HIP_EINVAL, "type 203 row: column 8 must be -1, 0 or 1");
        c->zone_flag = f;
      }
      c->needs_shape = true;
    }
    if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) continue;
    // device indices of the padded layout
    for (int d = 0; d < 3; ++d) {
      lo[d] += c->pf[d];
      hi[d] += c->pf[d];
    }
    b.i0 = lo[0]; b.i1 = hi[0]; b.j0 = lo[1]; b.j1 = hi[1]; b.k0 = lo[2]; b.k1 = hi[2];
    b.aim_shift = 0;
    if (t == 200) {
      // characteristic_zero_grad_init (nonreflecting.f90:665-728): the aim is
      // the inner neighbour, one point against each nonzero normal component
      // (its sign decides); the box must be one point thick along them so
      // that no aim point is itself updated
      const double kn = std::sqrt(b.k[0] * b.k[0] + b.k[1] * b.k[1] + b.k[2] * b.k[2]);
      if (!(kn > 0)) return fail(CRUNA_HIP_EINVAL, "type 200 row needs a nonzero normal");
      long shift[3];
      for (int d = 0; d < 3; ++d) {
        const double kd = b.k[d] / kn;
        shift[d] = kd > 0 ? -1 : (kd < 0 ? 1 : 0);
        if (shift[d] != 0 && lo[d] != hi[d])
          return fail(CRUNA_HIP_EINVAL, "type 200 row: the box must be one point thick along its normal");
        if (lo[d] - c->pf[d] + shift[d] < 0 || hi[d] - c->pf[d] + shift[d] > cfg->n[d] - 1)
          return fail(CRUNA_HIP_EINVAL, "type 200 row: inner neighbour outside the grid");
      }
      b.aim_shift = shift[0] + (long)c->np[0] * (shift[1] + (long)c->np[1] * shift[2]);
    }
    if (t == 100) {  // the variable whose RHS is zero (1-based, in column 8)
      const int v = (int)lround(cfg->bc[r][7]);
      if (v < 1 || v > 5) return fail(CRUNA_HIP_EINVAL, "wall row (100): column 8 must name a variable 1..5");
      b.k[0] = v - 1;
    }
    c->bc.push_back({t, b});
    if (t == 200 || t == 201 || t == 202 || t == 203 || t == 401) c->has_bcq = true;
    if (t == 202 || t == 203 || t == 205 || t == 206) c->needs_kvec = true;
    if (t == 400) c->has_bcrhs = true;
    if (t == 100 || t == 204 || t == 205 || t == 206) c->has_bclate = true;
    if (t == 204 || t == 205 || t == 206) c->has_char_rhs = true;
    if (t == 100) c->has_wall = true;
  }
  for (int d = 0; d < 3; ++d) {
    if (int rc = build_geom(c.get(), d)) return rc;
    try {
      const OpKind dk = cfg->diff_kind[d] == CRUNA_HIP_DIFF_O4_EXPLICIT ? OP_DIFF_O4E
                        : cfg->diff_kind[d] == CRUNA_HIP_DIFF_O4_PADE   ? OP_DIFF_O4P
                        : cfg->diff_kind[d] == CRUNA_HIP_DIFF_O2_EXPLICIT ? OP_DIFF_O2E
                                                                        : OP_DIFF;
      HostLineOp dd = build_line_op(dk, cfg->scheme[d] == CRUNA_HIP_PERIODIC, cfg->n[d], 0.0);
      HostLineOp ff =
          build_line_op(OP_FILTER, cfg->scheme[d] == CRUNA_HIP_PERIODIC, cfg->n[d], cfg->filter_alpha);
      if (int rc = c->dop64[d].upload(dd)) return rc;
      if (int rc = c->fop64[d].upload(ff)) return rc;
      if (int rc = c->dop32[d].upload(dd)) return rc;
      if (int rc = c->fop32[d].upload(ff)) return rc;
    } catch (const std::exception &e) {
      return fail(CRUNA_HIP_EINVAL, e.what());
    }
  }
  Phys &ph = c->phys;
  ph.gamma = cfg->gamma;
  ph.gm1 = cfg->gamma - 1.0;
  ph.gt = cfg->gamma / (cfg->gamma - 1.0);
  ph.w_inv_r = cfg->W / cfg->R;
  // euler_rho_rhou_p (governing_equations): the NS sweeps without viscosity
  // and conduction (mu = lambda = 0 make every viscous and heat term zero)
  ph.mu0 = cfg->equations == CRUNA_HIP_EQ_EULER ? 0.0 : cfg->mu0;
  ph.T0 = cfg->T0;
  ph.visc_exp = cfg->visc_exp;
  ph.visc_law = cfg->visc_law;
  ph.cp = cfg->gamma * (cfg->R / cfg->W) / (cfg->gamma - 1.0);
  ph.Pr = cfg->Pr;
  ph.lam_coef = ph.cp / cfg->Pr;
  ph.inv_T0 = 1.0 / cfg->T0;
  for (int d = 0; d < 3; ++d) {
    ph.inv_h[d] = 1.0 / cfg->dx[d];
    ph.heat_inv_h[d] = cfg->heat_unit_spacing ? 1.0 : 1.0 / cfg->dx[d];
  }
  const size_t N = c->N;
  const size_t ts = cfg->precision == CRUNA_HIP_FP32_MIXED ? sizeof(float) : sizeof(double);
  if (hipMalloc(&c->qn, 5 * N * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->acc, 5 * N * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->qs, 5 * N * ts) != hipSuccess || hipMalloc(&c->rhs, 5 * N * ts) != hipSuccess ||
      hipMalloc(&c->G, 7 * N * ts) != hipSuccess ||
      // the weighted filter's output buffer, allocated here so that a step
      // loop under graph capture allocates nothing
      ((cfg->wfilter_freq[0] > 0 || cfg->wfilter_freq[1] > 0 || cfg->wfilter_freq[2] > 0) &&
       hipMalloc(&c->wtmp, 5 * N * sizeof(double)) != hipSuccess)) {
    (void)hipGetLastError();
    cruna_hip_destroy(c.release());
    return fail(CRUNA_HIP_ENOMEM, "device allocation failed");
  }
  *out = c.release();
  return CRUNA_HIP_OK;
}

int cruna_hip_destroy(cruna_hip_ctx *c) {
  if (!c) return CRUNA_HIP_OK;
  (void)hipStreamSynchronize(c->stream);
  if (c->cnn) cnn_user(c->cnn, -1);
  for (int d = 0; d < 3; ++d) {
    c->dop64[d].release();
    c->fop64[d].release();
    c->dop32[d].release();
    c->fop32[d].release();
  }
  for (void *p : {(void *)c->qn, (void *)c->acc, (void *)c->qref, (void *)c->shape, (void *)c->tmp,
                  (void *)c->qa, (void *)c->traj, (void *)c->gobj, c->met, (void *)c->q0p, c->force,
                  (void *)c->ckpt, c->tg, c->pen, (void *)c->wphi, (void *)c->wtmp, (void *)c->wbuf,
                  (void *)c->kvec, (void *)c->q0copy, (void *)c->res_buf,
                  c->qs, c->rhs, c->G})
    if (p) (void)hipFree(p);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CRUNA_HIP_OK;
}

int cruna_hip_set_state(cruna_hip_ctx *c, const double *q) {
  if (!c || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  c->qs_fresh = false;
  if (int rc = put_fields(c, c->qn, q, 5, kRestState)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_get_state(cruna_hip_ctx *c, double *q) {
  if (!c || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = get_fields(c, q, c->qn, 5)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_set_qref(cruna_hip_ctx *c, const double *qref) {
  if (!c || !qref) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (!c->qref) HIP_TRY(hipMalloc(&c->qref, 5 * c->N * sizeof(double)));
  if (int rc = put_fields(c, c->qref, qref, 5, kRestState)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_set_kvector(cruna_hip_ctx *c, const double *kvector) {
  if (!c || !kvector) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  for (long i = 0; i < 3 * c->N; ++i)
    if (!std::isfinite(kvector[i])) return fail(CRUNA_HIP_EINVAL, "kvector is not finite");
  if (!c->kvec) HIP_TRY(hipMalloc(&c->kvec, 3 * c->N * sizeof(double)));
  const double kfill[3] = {1.0, 0.0, 0.0};
  if (int rc = put_fields(c, c->kvec, kvector, 3, kfill)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_set_sponge_shape(cruna_hip_ctx *c, const double *shape) {
  if (!c || !shape) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (!c->shape) HIP_TRY(hipMalloc(&c->shape, c->N * sizeof(double)));
  if (int rc = put_fields(c, c->shape, shape, 1, nullptr)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_set_metric(cruna_hip_ctx *c, const double *X_metric, const double *X_jacobian) {
  if (!c || !X_metric || !X_jacobian) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;

  // the padded layout replicates the nearest real point's metric, which
  // keeps a metric that is constant along lines constant along them
  std::vector<double> xm_pad, xj_pad;
  if (c->padded) {
    xm_pad = pad_host(c, X_metric, 9, nullptr, true);
    xj_pad = pad_host(c, X_jacobian, 1, nullptr, true);
    X_metric = xm_pad.data();
    X_jacobian = xj_pad.data();
  }
  const long n1 = c->np[0], n2 = c->np[1], n3 = c->np[2], N = c->N;
  auto M = [&](int k, int j, long i, long jj, long kk) {
    return X_metric[(size_t)N * (k + 3 * j) + i + n1 * (jj + n2 * kk)];
  };
  double scale = 0.0;
  for (size_t p = 0; p < 9 * (size_t)N; ++p) {
    if (!std::isfinite(X_metric[p])) return fail(CRUNA_HIP_EINVAL, "X_metric is not finite");
    scale = std::max(scale, std::fabs(X_metric[p]));
  }
  for (long p = 0; p < N; ++p)
    if (!std::isfinite(X_jacobian[p])) return fail(CRUNA_HIP_EINVAL, "X_jacobian is not finite");
  if (!(scale > 0.0)) return fail(CRUNA_HIP_EINVAL, "zero metric");
  // A diagonal metric whose entries M(d,d) are constant along the lines of
  // direction d (every tensor-product stretch) takes the per-line sweeps;
  // any other takes the general curvilinear sweeps with the metric per point.
  bool tensor = true;
  for (int k = 0; k < 3 && tensor; ++k)
    for (int j = 0; j < 3 && tensor; ++j) {
      if (k == j) continue;
      for (long p = 0; p < N; ++p)
        if (std::fabs(X_metric[(size_t)N * (k + 3 * j) + p]) > 1e-12 * scale) {
          tensor = false;
          break;
        }
    }
  std::vector<double> m1((size_t)n2 * n3), m2((size_t)n1 * n3), m3((size_t)n1 * n2);
  for (long kk = 0; kk < n3 && tensor; ++kk)
    for (long jj = 0; jj < n2 && tensor; ++jj)
      for (long i = 0; i < n1; ++i) {
        const double a = M(0, 0, i, jj, kk), b = M(1, 1, i, jj, kk), e = M(2, 2, i, jj, kk);
        const double a0 = M(0, 0, 0, jj, kk), b0 = M(1, 1, i, 0, kk), e0 = M(2, 2, i, jj, 0);
        if (std::fabs(a - a0) > 1e-13 * scale || std::fabs(b - b0) > 1e-13 * scale ||
            std::fabs(e - e0) > 1e-13 * scale) {
          tensor = false;
          break;
        }
        m1[jj + n2 * kk] = a0;
        m2[i + n1 * kk] = b0;
        m3[i + n1 * jj] = e0;
      }
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  c->gt = false;
  c->curv = false;
  if (c->ev) tensor = false;  // the penalised form runs on the per-point sweeps
  if (tensor) {
    const int rc = mixed ? upload_metric<float>(c, m1, m2, m3, X_jacobian, c->mt32)
                         : upload_metric<double>(c, m1, m2, m3, X_jacobian, c->mt64);
    if (rc) return rc;
  } else {
    const int rc = mixed ? upload_curv_metric<float>(c, X_metric, X_jacobian, c->cm32)
                         : upload_curv_metric<double>(c, X_metric, X_jacobian, c->cm64);
    if (rc) return rc;
    c->curv = true;
  }
  c->gt = true;
  return CRUNA_HIP_OK;
}

int cruna_hip_set_penalization(cruna_hip_ctx *c, const double *phi, const double *inv_phi) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (!phi && !inv_phi) {  // back to ns_rho_rhou_p
    (void)hipStreamSynchronize(c->stream);
    if (c->pen) (void)hipFree(c->pen);
    if (c->wphi) (void)hipFree(c->wphi);
    c->pen = nullptr;
    c->wphi = nullptr;
    c->ev = false;
    c->pe64 = Penal<double>{};
    c->pe32 = Penal<float>{};
    return CRUNA_HIP_OK;
  }
  if (!phi || !inv_phi) return fail(CRUNA_HIP_EINVAL, "phi and inv_phi go together");
  // on a deformed grid the penalised RHS (ns_rho_rhou_p_ev_phi_grid_transform)
  // runs on the sweeps with the full metric per point
  if (c->gt && !c->curv)
    return fail(CRUNA_HIP_EINVAL,
                "set the penalisation before the metric: a penalised deformed grid keeps the full "
                "metric per point, which a tensor-product metric set before it did not");
  std::vector<double> ph_pad, ip_pad;
  if (c->padded) {
    const double one = 1.0;
    ph_pad = pad_host(c, phi, 1, &one);
    ip_pad = pad_host(c, inv_phi, 1, &one);
    phi = ph_pad.data();
    inv_phi = ip_pad.data();
  }
  const size_t N = c->N;
  for (size_t i = 0; i < N; ++i)
    if (!std::isfinite(phi[i]) || !std::isfinite(inv_phi[i]))
      return fail(CRUNA_HIP_EINVAL, "phi or inv_phi is not finite");
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  const size_t ts = mixed ? sizeof(float) : sizeof(double);
  std::vector<char> h(2 * N * ts);
  for (size_t i = 0; i < N; ++i) {
    if (mixed) {
      ((float *)h.data())[i] = float(phi[i]);
      ((float *)h.data())[N + i] = float(inv_phi[i]);
    } else {
      ((double *)h.data())[i] = phi[i];
      ((double *)h.data())[N + i] = inv_phi[i];
    }
  }
  if (!c->pen) HIP_TRY(hipMalloc(&c->pen, h.size()));
  HIP_TRY(hipMemcpy(c->pen, h.data(), h.size(), hipMemcpyHostToDevice));
  if (!c->wphi) HIP_TRY(hipMalloc(&c->wphi, N * sizeof(double)));
  HIP_TRY(hipMemcpy(c->wphi, phi, N * sizeof(double), hipMemcpyHostToDevice));
  if (mixed)
    c->pe32 = Penal<float>{(const float *)c->pen, (const float *)c->pen + N};
  else
    c->pe64 = Penal<double>{(const double *)c->pen, (const double *)c->pen + N};
  c->ev = true;
  return CRUNA_HIP_OK;
}

int cruna_hip_filter_weighted(cruna_hip_ctx *c, int dir, double *q, int nv, const double *weight,
                              const double *strength) {
  if (!c || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (dir < 0 || dir > 2) return fail(CRUNA_HIP_EINVAL, "dir must be 0, 1 or 2");
  if (nv < 1 || nv > 5) return fail(CRUNA_HIP_EINVAL, "nv must be 1..5");
  if (c->np[dir] != c->cfg.n[dir])
    return fail(CRUNA_HIP_EINVAL, "the weighted filter's core is periodic: its direction needs a multiple of 8 points");
  if (int rc = ensure_tmp(c)) return rc;
  const size_t N = c->N;
  if (!c->wbuf) HIP_TRY(hipMalloc(&c->wbuf, 2 * N * sizeof(double)));
  double *in = c->tmp, *out = c->tmp + 5 * N;
  const double one = 1.0;
  if (int rc = put_fields(c, in, q, nv, nullptr)) return rc;
  const double *w = nullptr, *s = nullptr;
  if (weight) {
    if (int rc = put_fields(c, c->wbuf, weight, 1, &one)) return rc;
    w = c->wbuf;
  }
  if (strength) {
    if (int rc = put_fields(c, c->wbuf + N, strength, 1, &one)) return rc;
    s = c->wbuf + N;
  }
  if (int rc = filter_weighted_dir(c, in, out, dir, nv, w, s)) return rc;
  if (int rc = get_fields(c, q, in, nv)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_diff(cruna_hip_ctx *c, int dir, double *q_x, const double *q, double h) {
  if (!c || !q_x || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (dir < 0 || dir > 2) return fail(CRUNA_HIP_EINVAL, "dir must be 0, 1 or 2");
  if (!(h != 0.0)) return fail(CRUNA_HIP_EINVAL, "h must be nonzero");
  if (int rc = ensure_tmp(c)) return rc;
  double *in = c->tmp, *out = c->tmp + c->N;
  if (int rc = put_fields(c, in, q, 1, nullptr)) return rc;
  {
    ProfScope p(c, KID_DIFF);
    DISPATCH_DIR(c, dir, k_diff, double, (const double *)in, out, c->dop64[dir].op, c->geom[dir],
                 1.0 / h);
  }
  if (int rc = check_launch()) return rc;
  if (int rc = get_fields(c, q_x, out, 1)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_filter(cruna_hip_ctx *c, int dir, double *q, int nv) {
  if (!c || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (dir < 0 || dir > 2) return fail(CRUNA_HIP_EINVAL, "dir must be 0, 1 or 2");
  if (nv < 1 || nv > 5) return fail(CRUNA_HIP_EINVAL, "nv must be 1..5");
  if (int rc = ensure_tmp(c)) return rc;
  if (int rc = put_fields(c, c->tmp, q, nv, nullptr)) return rc;
  int rc = c->cfg.precision == CRUNA_HIP_FP32_MIXED ? filter_dir<float>(c, c->tmp, dir, nv)
                                                    : filter_dir<double>(c, c->tmp, dir, nv);
  if (rc) return rc;
  if ((rc = get_fields(c, q, c->tmp, nv))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_rhs_direct(cruna_hip_ctx *c, double *rhs, const double *q, double t) {
  (void)t;  // the direct RHS of the reference does not depend on t
  if (!c || !rhs || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  c->qs_fresh = false;
  if (int rc = ensure_tmp(c)) return rc;
  if (int rc = put_fields(c, c->tmp, q, 5, kRestState)) return rc;
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  int rc = mixed ? run_stage<float>(c, c->tmp, 4) : run_stage<double>(c, c->tmp, 4);
  if (rc) return rc;
  return copy_rhs_out(c, rhs);
}

int cruna_hip_box_filter_state(cruna_hip_ctx *c, int f, double *q_box, double *rhs_box) {
  if (!c || !q_box) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (f < 1) return fail(CRUNA_HIP_EINVAL, "box factor must be >= 1");
  if (c->padded) return fail(CRUNA_HIP_EINVAL, "box filtering needs every grid size a multiple of 8");
  for (int d = 0; d < 3; ++d)
    if (c->cfg.n[d] % f) return fail(CRUNA_HIP_EINVAL, "box factor must divide every grid size");
  if (int rc = ensure_tmp(c)) return rc;
  const long M = c->N / ((long)f * f * f);
  const unsigned nb = (unsigned)((5 * M + 255) / 256);
  const int *n = c->cfg.n;
  hipLaunchKernelGGL((k_box_average<double>), dim3(nb), dim3(256), 0, c->stream, (const double *)c->qn, n[0],
                     n[1], n[2], f, c->tmp);
  if (rhs_box) {
    const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
    if (int rc = mixed ? run_stage<float>(c, c->qn, 4) : run_stage<double>(c, c->qn, 4)) return rc;
    if (mixed)
      hipLaunchKernelGGL((k_box_average<float>), dim3(nb), dim3(256), 0, c->stream, (const float *)c->rhs, n[0],
                         n[1], n[2], f, c->tmp + 5 * M);
    else
      hipLaunchKernelGGL((k_box_average<double>), dim3(nb), dim3(256), 0, c->stream, (const double *)c->rhs,
                         n[0], n[1], n[2], f, c->tmp + 5 * M);
  }
  if (int rc = check_launch()) return rc;
  HIP_TRY(hipMemcpyAsync(q_box, c->tmp, 5 * M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (rhs_box)
    HIP_TRY(hipMemcpyAsync(rhs_box, c->tmp + 5 * M, 5 * M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

// Host inputs of the adjoint entry points: qs -> qa, q -> second half of
// tmp, g -> gobj.
int stage_adjoint_inputs(cruna_hip_ctx *c, const double *qs, const double *q, const double *g) {
  if (int rc = ensure_tmp(c)) return rc;
  if (int rc = ensure_adjoint(c)) return rc;
  if (int rc = put_fields(c, c->qa, qs, 5, nullptr)) return rc;
  if (int rc = put_fields(c, c->tmp + 5 * c->N, q, 5, kRestState)) return rc;
  if (g) {
    if (!c->gobj) HIP_TRY(hipMalloc(&c->gobj, 5 * c->N * sizeof(double)));
    if (int rc = put_fields(c, c->gobj, g, 5, nullptr)) return rc;
  }
  return CRUNA_HIP_OK;
}

int cruna_hip_rhs_adjoint(cruna_hip_ctx *c, double *rhs, const double *qs, const double *q,
                          const double *g, double t) {
  (void)t;  // the adjoint RHS of the reference does not depend on t
  if (!c || !rhs || !qs || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = adjoint_checks(c)) return rc;
  if (int rc = stage_adjoint_inputs(c, qs, q, g)) return rc;
  const double *gd = g ? c->gobj : nullptr;
  if (int rc = point_q0(c, c->tmp + 5 * c->N)) return rc;
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  int rc = mixed ? arun_stage<float>(c, c->qa, gd, 4) : arun_stage<double>(c, c->qa, gd, 4);
  if (rc) return rc;
  return copy_rhs_out(c, rhs);
}

int cruna_hip_time_stepper_direct(cruna_hip_ctx *c, double *q1, const double *q0, double t) {
  (void)t;
  if (!c || !q1 || !q0) return fail(CRUNA_HIP_EINVAL, "null argument");
  c->qs_fresh = false;
  if (int rc = put_fields(c, c->qn, q0, 5, kRestState)) return rc;
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  int rc = mixed ? rk4_step<float>(c) : rk4_step<double>(c);
  if (rc) return rc;
  if ((rc = get_fields(c, q1, c->qn, 5))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_time_stepper_adjoint(cruna_hip_ctx *c, double *qs1, const double *qs0,
                                   const double *q0, const double *g, double t) {
  (void)t;
  if (!c || !qs1 || !qs0 || !q0) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = adjoint_checks(c)) return rc;
  if (int rc = stage_adjoint_inputs(c, qs0, q0, g)) return rc;
  const double *gd = g ? c->gobj : nullptr;
  if (int rc = point_q0(c, c->tmp + 5 * c->N)) return rc;
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  int rc = mixed ? ark4_step<float>(c, gd) : ark4_step<double>(c, gd);
  if (rc) return rc;
  if (int rc = get_fields(c, qs1, c->qa, 5)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_set_adjoint_state(cruna_hip_ctx *c, const double *qs) {
  if (!c || !qs) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = ensure_adjoint(c)) return rc;
  if (int rc = put_fields(c, c->qa, qs, 5, nullptr)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_get_adjoint_state(cruna_hip_ctx *c, double *qs) {
  if (!c || !qs) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (!c->qa) return fail(CRUNA_HIP_EINVAL, "no adjoint state (cruna_hip_set_adjoint_state)");
  if (int rc = get_fields(c, qs, c->qa, 5)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_trajectory_alloc(cruna_hip_ctx *c, int nslots) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (nslots < 0) return fail(CRUNA_HIP_EINVAL, "nslots must be >= 0");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->traj) HIP_TRY(hipFree(c->traj));
  c->traj = nullptr;
  c->traj_slots = 0;
  if (nslots == 0) return CRUNA_HIP_OK;
  if (hipMalloc(&c->traj, (size_t)nslots * 5 * c->N * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    c->traj = nullptr;
    return fail(CRUNA_HIP_ENOMEM, "trajectory of " + std::to_string(nslots) +
                                      " states does not fit in device memory");
  }
  c->traj_slots = nslots;
  return CRUNA_HIP_OK;
}

int cruna_hip_trajectory_set(cruna_hip_ctx *c, int nt, const double *q) {
  if (!c || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (nt < 1 || nt > c->traj_slots) return fail(CRUNA_HIP_EINVAL, "nt outside the trajectory");
  if (int rc = put_fields(c, traj_slot(c, nt), q, 5, kRestState)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_trajectory_get(cruna_hip_ctx *c, int nt, double *q) {
  if (!c || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (nt < 1 || nt > c->traj_slots) return fail(CRUNA_HIP_EINVAL, "nt outside the trajectory");
  if (int rc = get_fields(c, q, traj_slot(c, nt), 5)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_trajectory_record(cruna_hip_ctx *c, int on) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (on && c->traj_slots == 0) return fail(CRUNA_HIP_EINVAL, "no trajectory (cruna_hip_trajectory_alloc)");
  c->record = on != 0;
  return CRUNA_HIP_OK;
}

int cruna_hip_run_adjoint(cruna_hip_ctx *c, int nt_end, int nsteps) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = adjoint_checks(c)) return rc;
  if (nsteps < 0) return fail(CRUNA_HIP_EINVAL, "nsteps must be >= 0");
  if (!c->qa) return fail(CRUNA_HIP_EINVAL, "no adjoint state (cruna_hip_set_adjoint_state)");
  if (c->traj_slots == 0) return fail(CRUNA_HIP_EINVAL, "no trajectory (cruna_hip_trajectory_alloc)");
  // one slot: the linear adjoint about a fixed state (computation_adjoint.f90:69-73)
  if (c->traj_slots > 1 && nsteps > 0 && (nt_end > c->traj_slots || nt_end - nsteps + 1 < 1))
    return fail(CRUNA_HIP_EINVAL, "steps outside the trajectory");
  HIP_TRY(hipMemcpyAsync(c->ntp, &nt_end, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));  // nt_end lives on this stack frame
  for (int nt = nt_end; nt > nt_end - nsteps; --nt)
    if (int rc = astep_once(c, nt)) return rc;
  return CRUNA_HIP_OK;
}

int cruna_hip_run_direct(cruna_hip_ctx *c, int nt0, int nsteps) {
  return run_direct_loop(c, nt0, nsteps, true);
}

int cruna_hip_set_convergence(cruna_hip_ctx *c, int cfreq, double ctol) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (cfreq < 0) return fail(CRUNA_HIP_EINVAL, "cfreq must be >= 0");
  ++c->gen;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (cfreq > 0 && !c->q0copy) {
    if (hipMalloc(&c->q0copy, 5 * (size_t)c->N * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->res_buf, (1 + RES_MAX_PARTS) * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(CRUNA_HIP_ENOMEM, "convergence buffers do not fit in device memory");
    }
  }
  c->cfreq = cfreq;
  c->ctol = ctol;
  c->last_res = -1.0;
  c->res_nt = 0;
  c->conv_nt = 0;
  return CRUNA_HIP_OK;
}

int cruna_hip_get_residual(cruna_hip_ctx *c, double *residual, int *nt_checked, int *nt_converged) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (residual) *residual = c->last_res;
  if (nt_checked) *nt_checked = c->res_nt;
  if (nt_converged) *nt_converged = c->conv_nt;
  return CRUNA_HIP_OK;
}

extern "C++" int run_direct_loop(cruna_hip_ctx *c, int nt0, int nsteps, bool weighted) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (nsteps < 0) return fail(CRUNA_HIP_EINVAL, "nsteps must be >= 0");
  // the convergence check belongs to the cruna_hpc loop (weighted filters);
  // calc_direct's subset loop has none
  bool conv = false;
  if (weighted) c->conv_nt = 0;
  for (int nt = nt0 + 1; nt <= nt0 + nsteps; ++nt) {
    if (int rc = step_once(c, nt, weighted, weighted ? &conv : nullptr)) return rc;
    if (conv) break;  // stop_cruna after the snapshot of q1 (cruna_hpc.f90:190-206)
    // Q(:,:,:,:,nt) = qn (computation_direct.f90:100-103)
    if (c->record && nt >= 1 && nt <= c->traj_slots)
      HIP_TRY(hipMemcpyAsync(traj_slot(c, nt), c->qn, 5 * c->N * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  }
  return CRUNA_HIP_OK;
}

int cruna_hip_subsets_alloc(cruna_hip_ctx *c, int subsets, int steps) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (subsets < 0 || steps < 0) return fail(CRUNA_HIP_EINVAL, "subsets and steps must be >= 0");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->ckpt) HIP_TRY(hipFree(c->ckpt));
  c->ckpt = nullptr;
  c->ckpt_n = 0;
  c->sub_steps = 0;
  if (subsets == 0) return cruna_hip_trajectory_alloc(c, 0);
  if (hipMalloc(&c->ckpt, (size_t)(subsets + 1) * 5 * c->N * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    c->ckpt = nullptr;
    return fail(CRUNA_HIP_ENOMEM, "subset checkpoints do not fit in device memory");
  }
  c->ckpt_n = subsets + 1;
  c->sub_steps = steps;
  return cruna_hip_trajectory_alloc(c, steps);
}

int cruna_hip_run_direct_subsets(cruna_hip_ctx *c, int subsets, int steps) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (subsets < 1 || steps < 0) return fail(CRUNA_HIP_EINVAL, "subsets >= 1 and steps >= 0");
  if (c->ckpt_n < subsets + 1) return fail(CRUNA_HIP_EINVAL, "not enough subset checkpoints (cruna_hip_subsets_alloc)");
  const size_t bytes = 5 * (size_t)c->N * sizeof(double);
  const bool record = c->record;
  c->record = false;
  int rc = CRUNA_HIP_OK;
  for (int s = 0; s < subsets && !rc; ++s) {
    if (hipMemcpyAsync(c->ckpt + (size_t)s * 5 * c->N, c->qn, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      rc = fail(CRUNA_HIP_EDEVICE, "checkpoint copy failed");
    if (!rc) rc = run_direct_loop(c, 0, steps, false);  // nt = 1..steps within the subset
  }
  if (!rc && hipMemcpyAsync(c->ckpt + (size_t)subsets * 5 * c->N, c->qn, bytes, hipMemcpyDeviceToDevice,
                            c->stream) != hipSuccess)
    rc = fail(CRUNA_HIP_EDEVICE, "checkpoint copy failed");
  c->record = record;
  return rc;
}

int cruna_hip_run_adjoint_subsets(cruna_hip_ctx *c, int subsets, int steps) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = adjoint_checks(c)) return rc;
  if (subsets < 1 || steps < 1) return fail(CRUNA_HIP_EINVAL, "subsets and steps must be >= 1");
  if (c->ckpt_n < subsets + 1) return fail(CRUNA_HIP_EINVAL, "not enough subset checkpoints (cruna_hip_subsets_alloc)");
  if (c->traj_slots < steps) return fail(CRUNA_HIP_EINVAL, "the trajectory holds fewer slots than steps");
  if (!c->qa) return fail(CRUNA_HIP_EINVAL, "no adjoint state (cruna_hip_set_adjoint_state)");
  const size_t bytes = 5 * (size_t)c->N * sizeof(double);
  const bool record = c->record;
  int rc = CRUNA_HIP_OK;
  for (int s = subsets - 1; s >= 0 && !rc; --s) {
    // recompute the subset's forward trajectory from its checkpoint
    if (hipMemcpyAsync(c->qn, c->ckpt + (size_t)s * 5 * c->N, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      rc = fail(CRUNA_HIP_EDEVICE, "checkpoint copy failed");
    c->qs_fresh = false;
    c->record = true;
    if (!rc) rc = run_direct_loop(c, 0, steps, false);
    c->record = false;
    if (!rc) rc = cruna_hip_run_adjoint(c, steps, steps);
  }
  // the forward state is the final one again
  if (!rc && hipMemcpyAsync(c->qn, c->ckpt + (size_t)subsets * 5 * c->N, bytes, hipMemcpyDeviceToDevice,
                            c->stream) != hipSuccess)
    rc = fail(CRUNA_HIP_EDEVICE, "checkpoint copy failed");
  c->qs_fresh = false;
  c->record = record;
  return rc;
}

int cruna_hip_autotune(cruna_hip_ctx *c, int steps, int tile_lines[3]) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (steps < 1) return fail(CRUNA_HIP_EINVAL, "steps must be >= 1");
  if (int rc = ensure_tmp(c)) return rc;
  // the trial steps run on a saved copy of the state, restored afterwards
  const size_t bytes = 5 * (size_t)c->N * sizeof(double);
  HIP_TRY(hipMemcpyAsync(c->tmp, c->qn, bytes, hipMemcpyDeviceToDevice, c->stream));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const bool prof = c->prof;
  c->prof = false;
  int best[3];
  for (int d = 0; d < 3; ++d) best[d] = c->geom[d].TL;
  int rc = CRUNA_HIP_OK;
  auto trial = [&](float *ms) -> int {
    for (int d = 0; d < 3; ++d) c->cfg.tile_lines[d] = best[d];
    for (int d = 0; d < 3; ++d)
      if (int r = build_geom(c, d)) return r;
    HIP_TRY(hipMemcpyAsync(c->qn, c->tmp, bytes, hipMemcpyDeviceToDevice, c->stream));
    c->qs_fresh = false;
    if (int r = step_once(c, 1)) return r;  // warm-up
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int s = 0; s < steps; ++s)
      if (int r = step_once(c, 2 + s)) return r;
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    return CRUNA_HIP_OK;
  };
  // coordinate search: one direction at a time over the valid line counts
  for (int d = 0; d < 3 && !rc; ++d) {
    const long lines = d == 0 ? (long)c->np[1] * c->np[2] : c->np[0];
    const int NB = c->np[d] / PPT;
    int arg = best[d];
    float tbest = 0.f;
    bool first = true;
    for (int TL = 1; NB * TL <= NT && TL <= 64; TL *= 2) {
      if (lines % TL != 0 || NB * TL < 64) continue;  // at least one full wavefront
      best[d] = TL;
      float ms = 0.f;
      if ((rc = trial(&ms))) break;
      if (first || ms < tbest) {
        tbest = ms;
        arg = TL;
        first = false;
      }
    }
    best[d] = arg;
  }
  c->prof = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int d = 0; d < 3; ++d) c->cfg.tile_lines[d] = best[d];
  for (int d = 0; d < 3 && !rc; ++d) rc = build_geom(c, d);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(c->qn, c->tmp, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->qs_fresh = false;
  if (tile_lines)
    for (int d = 0; d < 3; ++d) tile_lines[d] = best[d];
  return CRUNA_HIP_OK;
}

int cruna_hip_synchronize(cruna_hip_ctx *c) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->prof || !c->pending.empty()) return collect_profile(c);
  return CRUNA_HIP_OK;
}

int cruna_hip_profile_enable(cruna_hip_ctx *c, int on) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (int rc = collect_profile(c)) return rc;
  c->prof = on != 0;
  for (int k = 0; k < KID_COUNT; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return CRUNA_HIP_OK;
}

int cruna_hip_profile_read(cruna_hip_ctx *c, int kid, double *ms, long *launches, char *name,
                           size_t name_len) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (kid < 0 || kid >= KID_COUNT) return fail(CRUNA_HIP_EINVAL, "bad kernel id");
  if (ms) *ms = c->prof_ms[kid];
  if (launches) *launches = c->prof_n[kid];
  if (name && name_len) {
    std::strncpy(name, kKernelNames[kid], name_len - 1);
    name[name_len - 1] = 0;
  }
  return CRUNA_HIP_OK;
}

int cruna_hip_profile_count(void) { return KID_COUNT; }

// ---------------------------------------------------------------------------
// Ensembles: independent members (forward or adjoint runs of any size) on one
// device.  Each member keeps its own stream; a run records one period of
// steps of each member into a HIP graph and replays the members' graphs on
// their own streams, so that members run concurrently on the device's
// hardware queues (a single graph with one branch per member replays its
// branches one after the other here, and small members alone leave most of
// the device idle: 48^3 runs at ~290 Mpts/s, 128^3 at ~1350).

}  // extern "C"

// One bin of a batched ensemble (host side): work-groups and work-group size
// of each launch, its adjoint members and their filter switches, its graph.
struct BatchBin {
  int blocks[2][2][3] = {}, wg[2][2][3] = {};  // [group][kind][dir]
  int nadj = 0;
  int afreq[3] = {0, 0, 0};
  hipGraphExec_t exec = nullptr;
};

struct cruna_hip_ensemble {
  std::vector<cruna_hip_ctx *> m;
  std::vector<int> role;  // CRUNA_HIP_FORWARD / CRUNA_HIP_ADJOINT
  std::vector<int> nt;    // forward: steps done; adjoint: next step to run backward
  // adjoint members with subset checkpoints (time.subsets): the subset loop
  // of calc_adjoint, each subset's trajectory recomputed from its checkpoint;
  // pos = backward steps left in the recomputed subset (0: recompute next)
  std::vector<bool> subm;
  std::vector<int> pos;
  std::vector<bool> warm;
  std::vector<hipGraphExec_t> exec;  // one period of steps of each member
  std::vector<unsigned> gen;         // each member's setter generation when recorded
  int period = 0;
  bool dirty = true;  // the recorded graph no longer matches the members
  // batched launches (batch.hpp): the members split into bins; a bin runs one
  // launch per sweep kind over its members, recorded into a graph that is
  // replayed on the bin's stream, and the bins run side by side
  int batching = -1;            // bins; 0: per-member graphs; -1: automatic
  int batched_steps = 0;        // steps of the last run done by batched launches
  BatchMember *bmem = nullptr;  // device: one descriptor per member
  BatchList *blist = nullptr;   // device: per bin, lists [group: stepping forward, adjoint][kind: sweep, filter][dir]
  int blist_cap = 0;            // lists allocated
  std::vector<BatchBin> bins;
  std::vector<hipStream_t> bstreams;  // one per bin (kept while the ensemble lives)
  bool bper[3] = {true, true, true};
  int bperiod = 0;              // steps of one recorded batched period
  std::vector<unsigned> bgen;   // members' setter generations when recorded
  std::vector<int> brole;       // members' roles (2: adjoint over subsets) when recorded
};

namespace {

int gcd_i(int a, int b) { return b ? gcd_i(b, a % b) : a; }
int lcm_i(int a, int b) { return a / gcd_i(a, b) * b; }

// Steps after which a member's launch sequence repeats: its filter
// frequencies (direct: the compact and the weighted filters; adjoint: the
// adjoint filters) decide which passes run at a step; a subset adjoint
// member repeats with every subset (step numbers restart at 1 in each).
int member_period(const cruna_hip_ctx *c, int role, bool subset) {
  if (role == CRUNA_HIP_ADJOINT && subset) return c->sub_steps;
  int p = 1;
  for (int d = 0; d < 3; ++d) {
    const int f = role == CRUNA_HIP_ADJOINT ? c->cfg.afilter_freq[d] : c->cfg.filter_freq[d];
    if (f > 0) p = lcm_i(p, f);
    if (role == CRUNA_HIP_FORWARD && c->cfg.wfilter_freq[d] > 0) p = lcm_i(p, c->cfg.wfilter_freq[d]);
  }
  return p;
}

// One backward step of a subset adjoint member: at the start of a subset
// restore its checkpoint (named by the device counter, which then counts
// down) and recompute the subset's trajectory (calc_direct's subset loop,
// nt = 1..steps, no weighted filters), then the adjoint step over it.
int subset_astep(cruna_hip_ensemble *e, size_t i) {
  cruna_hip_ctx *c = e->m[i];
  const int n = c->sub_steps;
  if (e->pos[i] == 0) {
    const long M = 5 * c->N;
    hipLaunchKernelGGL(k_restore_ckpt, dim3((unsigned)std::min<long>((M + 255) / 256, 4096)), dim3(256), 0,
                       c->stream, c->qn, (const double *)c->ckpt, (const int *)c->sckp, M);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->sckp, INT_MIN, -1);
    if (int rc = check_launch()) return rc;
    c->qs_fresh = false;
    const bool rec = c->record;
    c->record = true;
    int rc = run_direct_loop(c, 0, n, false);
    c->record = rec;
    if (rc) return rc;
    // backward from the subset's last step, or from within it right after
    // set_role named such a step
    const int p = (e->nt[i] - 1) % n + 1;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->ntp, p, 0);
    if ((rc = check_launch())) return rc;
    e->pos[i] = p;
  }
  if (int rc = astep_once(c, e->pos[i])) return rc;
  --e->pos[i];
  --e->nt[i];
  return CRUNA_HIP_OK;
}

int member_steps(cruna_hip_ensemble *e, size_t i, int k) {
  cruna_hip_ctx *c = e->m[i];
  for (int s = 0; s < k; ++s) {
    if (e->role[i] == CRUNA_HIP_FORWARD) {
      if (int rc = step_once(c, ++e->nt[i])) return rc;
    } else if (e->subm[i]) {
      if (int rc = subset_astep(e, i)) return rc;
    } else {
      if (int rc = astep_once(c, e->nt[i]--)) return rc;
    }
  }
  return CRUNA_HIP_OK;
}

void drop_graphs(cruna_hip_ensemble *e) {
  for (auto &x : e->exec)
    if (x) (void)hipGraphExecDestroy(x);
  e->exec.clear();
}

// ---------------------------------------------------------------------------
// Batched ensemble steps (batch.hpp)

void drop_batch(cruna_hip_ensemble *e) {
  for (auto &b : e->bins)
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
  e->bins.clear();
  e->bperiod = 0;
}

// Whether the members can run as one batch, and the period of the batched
// launch sequence: every member FP64 on a Cartesian grid without boundary
// rows, closure or weighted filter, all with the same schemes per direction
// and the same filter switches (frequency 0 or 1, so that every step runs
// the same passes); the adjoint members all over subset checkpoints of the
// same length, each at a subset boundary (period: one subset), or all over
// whole trajectories (period: one step).
int batch_period(const cruna_hip_ensemble *e, int nsteps) {
  const size_t M = e->m.size();
  if (M > (size_t)BATCH_MAX) return 0;
  const cruna_hip_ctx *c0 = e->m[0];
  int sub = -1;  // -1: no adjoint member yet; 0: whole trajectories; S: subsets of S steps
  for (size_t i = 0; i < M; ++i) {
    const cruna_hip_ctx *c = e->m[i];
    if (c->cfg.precision != CRUNA_HIP_FP64 || c->gt || c->curv || c->ev || c->padded || c->cnn ||
        c->has_bcq || c->has_bcrhs || c->has_bclate || !c->bc.empty())
      return 0;
    for (int d = 0; d < 3; ++d) {
      if (c->cfg.scheme[d] != c0->cfg.scheme[d] || c->cfg.wfilter_freq[d] > 0) return 0;
      const int f = c->cfg.filter_freq[d], f0 = c0->cfg.filter_freq[d];
      if ((f > 1) || (f0 > 1) || ((f > 0) != (f0 > 0))) return 0;
    }
    if (e->role[i] == CRUNA_HIP_ADJOINT) {
      const int si = e->subm[i] ? c->sub_steps : 0;
      if (sub >= 0 && si != sub) return 0;
      sub = si;
      if (e->subm[i] && (e->pos[i] != 0 || e->nt[i] % c->sub_steps != 0)) return 0;
      if (!e->subm[i] && c->traj_slots > 1 && e->nt[i] - nsteps < 0) return 0;
      for (int d = 0; d < 3; ++d) {
        const int f = c->cfg.afilter_freq[d];
        if (f > 1) return 0;
        for (size_t j = 0; j < i; ++j)
          if (e->role[j] == CRUNA_HIP_ADJOINT && (e->m[j]->cfg.afilter_freq[d] > 0) != (f > 0)) return 0;
      }
    }
  }
  return sub > 0 ? sub : 1;
}

// Descriptors and work lists of the batch on the device: the members split
// into nb bins of about equal points (largest member first, each to the bin
// with the fewest points so far).  Group 0 of a bin steps forward (forward
// members and, with subsets, the adjoint members' recomputed trajectories),
// group 1 runs the bin's adjoint members backward.
int batch_upload(cruna_hip_ensemble *e, bool subsets, int nb) {
  const size_t M = e->m.size();
  std::vector<BatchMember> hm(M);
  for (size_t i = 0; i < M; ++i) {
    cruna_hip_ctx *c = e->m[i];
    BatchMember &b = hm[i];
    b.qn = c->qn;
    b.acc = c->acc;
    b.qs = (double *)c->qs;
    b.rhs = (double *)c->rhs;
    b.G = (double *)c->G;
    b.qa = c->qa;
    b.q0p = c->q0p;
    b.ntp = c->ntp;
    b.sckp = c->sckp;
    b.traj = c->traj;
    b.ckpt = c->ckpt;
    b.traj_slots = c->traj_slots;
    for (int d = 0; d < 3; ++d) {
      b.op[d] = c->dop64[d].op;
      b.fop[d] = c->fop64[d].op;
    }
    b.ph = c->phys;
    b.N = c->N;
    b.dt = c->cfg.dt;
    b.gamma = c->cfg.gamma;
  }
  // tiles of as many lines as fill a work-group of 256 threads (measured on
  // the 32-member workload: 5 % faster than 512 along y and z), where a y/z
  // tile's lines are whole 64-byte segments of x-rows (a multiple of 8
  // lines) when that leaves it at least 8 lines
  for (size_t i = 0; i < M; ++i)
    for (int d = 0; d < 3; ++d) {
      const int *np = e->m[i]->np;
      const int NB = np[d] / PPT;
      int TL = std::max(1, 256 / NB);
      if (d != 0 && TL >= 8) TL -= TL % 8;
      SweepGeom g = make_geom(np, d, 1);
      g.TL = TL;
      g.lines = d == 0 ? (long)np[1] * np[2] : (long)np[0] * np[3 - d];
      g.ntiles = (int)((g.lines + TL - 1) / TL);
      g.ntiles_in = 0;
      hm[i].geom[d] = g;
      hm[i].fgeom[d] = g;
    }
  std::vector<int> order(M);
  for (size_t i = 0; i < M; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hm[a].N > hm[b].N; });
  std::vector<std::vector<int>> bm(nb);
  std::vector<long> load(nb, 0);
  for (int i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    bm[k].push_back(i);
    load[k] += hm[i].N;
  }
  e->bins.assign(nb, BatchBin{});
  std::vector<BatchList> hl((size_t)nb * 12);
  for (int k = 0; k < nb; ++k) {
    std::sort(bm[k].begin(), bm[k].end());
    BatchBin &B = e->bins[k];
    for (int i : bm[k])
      if (e->role[i] == CRUNA_HIP_ADJOINT) {
        ++B.nadj;
        for (int d = 0; d < 3; ++d) B.afreq[d] = e->m[i]->cfg.afilter_freq[d];
      }
    for (int grp = 0; grp < 2; ++grp) {
      std::vector<int> idx;
      for (int i : bm[k])
        if (grp == 0 ? (e->role[i] == CRUNA_HIP_FORWARD || subsets) : e->role[i] == CRUNA_HIP_ADJOINT)
          idx.push_back(i);
      for (int kind = 0; kind < 2; ++kind)
        for (int d = 0; d < 3; ++d) {
          BatchList &L = hl[(((size_t)k * 2 + grp) * 2 + kind) * 3 + d];
          L.nm = (int)idx.size();
          int blocks = 0, wg = 0;
          for (size_t j = 0; j < idx.size(); ++j) {
            const SweepGeom &g = kind == 0 ? hm[idx[j]].geom[d] : hm[idx[j]].fgeom[d];
            L.mem[j] = idx[j];
            L.start[j] = blocks;
            blocks += g.ntiles;
            wg = std::max(wg, g.TL * g.NB);
          }
          L.start[idx.size()] = blocks;
          B.blocks[grp][kind][d] = blocks;
          B.wg[grp][kind][d] = wg;
        }
    }
  }
  if (!e->bmem) HIP_TRY(hipMalloc(&e->bmem, BATCH_MAX * sizeof(BatchMember)));
  if (e->blist_cap < nb * 12) {
    if (e->blist) HIP_TRY(hipFree(e->blist));
    e->blist = nullptr;
    e->blist_cap = 0;
    HIP_TRY(hipMalloc(&e->blist, hl.size() * sizeof(BatchList)));
    e->blist_cap = nb * 12;
  }
  HIP_TRY(hipMemcpy(e->bmem, hm.data(), M * sizeof(BatchMember), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->blist, hl.data(), hl.size() * sizeof(BatchList), hipMemcpyHostToDevice));
  for (int d = 0; d < 3; ++d) e->bper[d] = e->m[0]->cfg.scheme[d] == CRUNA_HIP_PERIODIC;
  return CRUNA_HIP_OK;
}

// The launches of one bin go to stream s.
struct BatchSlot {
  const cruna_hip_ensemble *e;
  int bin;
  hipStream_t s;
  const BatchBin &bb() const { return e->bins[bin]; }
  const BatchList *list(int grp, int kind, int d) const {
    return e->blist + (((size_t)bin * 2 + grp) * 2 + kind) * 3 + d;
  }
  const BatchMember *mem() const { return e->bmem; }
};

// launch of a batched kernel over list (grp, kind, d) of the bin; nothing for
// an empty list
#define BATCH_LAUNCH(S, grp, kind, d, KER, ...)                                                          \
  do {                                                                                                   \
    if ((S).bb().blocks[grp][kind][d] > 0)                                                               \
      hipLaunchKernelGGL(KER, dim3((S).bb().blocks[grp][kind][d]), dim3((S).bb().wg[grp][kind][d]), 0,   \
                         (S).s, (S).mem(), (S).list(grp, kind, d), __VA_ARGS__);                         \
  } while (0)

template <int DIR>
void batch_sweep1(const BatchSlot &S, int from_qs) {
  if (S.e->bper[DIR]) BATCH_LAUNCH(S, 0, 0, DIR, (kb_sweep1<DIR, true>), from_qs);
  else BATCH_LAUNCH(S, 0, 0, DIR, (kb_sweep1<DIR, false>), from_qs);
}

void batch_sweep12z(const BatchSlot &S, int from_qs) {
  if (S.e->bper[2]) BATCH_LAUNCH(S, 0, 0, 2, (kb_sweep12_z<true>), from_qs);
  else BATCH_LAUNCH(S, 0, 0, 2, (kb_sweep12_z<false>), from_qs);
}

template <int DIR, bool FF>
void batch_sweep2(const BatchSlot &S, int mode) {
  if (S.e->bper[DIR]) BATCH_LAUNCH(S, 0, 0, DIR, (kb_sweep2<DIR, true, FF>), mode);
  else BATCH_LAUNCH(S, 0, 0, DIR, (kb_sweep2<DIR, false, FF>), mode);
}

template <int DIR>
void batch_filter(const BatchSlot &S, int grp) {
  if (S.e->bper[DIR]) BATCH_LAUNCH(S, grp, 1, DIR, (kb_filter<DIR, true>), grp);
  else BATCH_LAUNCH(S, grp, 1, DIR, (kb_filter<DIR, false>), grp);
}

template <int DIR>
void batch_asweep(const BatchSlot &S, int mode, int from_qs) {
  if (S.e->bper[DIR]) BATCH_LAUNCH(S, 1, 0, DIR, (kb_asweep<DIR, true>), mode, from_qs);
  else BATCH_LAUNCH(S, 1, 0, DIR, (kb_asweep<DIR, false>), mode, from_qs);
}

// One forward step of the bin's group 0 (step_once of each member: RK4 with
// the x filter fused into the last stage, then the y and z filters); with
// `record` the bin's subset members store Q(:,:,:,:,slot).
int batch_forward_step(const BatchSlot &S, int record, int slot) {
  const cruna_hip_cfg &cf = S.e->m[0]->cfg;
  const bool fuse = cf.filter_freq[0] > 0;
  for (int st = 0; st < 4; ++st) {
    batch_sweep1<0>(S, st > 0);
    batch_sweep1<1>(S, st > 0);
    batch_sweep12z(S, st > 0);
    batch_sweep2<1, false>(S, st);
    if (st == 3 && fuse) batch_sweep2<0, true>(S, st);
    else batch_sweep2<0, false>(S, st);
  }
  if (cf.filter_freq[1] > 0) batch_filter<1>(S, 0);
  if (cf.filter_freq[2] > 0) batch_filter<2>(S, 0);
  if (record && S.bb().nadj > 0)
    hipLaunchKernelGGL(kb_record, dim3(256, S.bb().nadj), dim3(256), 0, S.s, S.mem(), S.list(1, 0, 0), slot);
  return check_launch();
}

// One backward step of the bin's adjoint members (astep_once of each).
int batch_adjoint_step(const BatchSlot &S) {
  const BatchBin &B = S.bb();
  hipLaunchKernelGGL(kb_adjoint_slot, dim3(1, B.nadj), dim3(64), 0, S.s, S.mem(), S.list(1, 0, 0));
  for (int st = 0; st < 4; ++st) {
    batch_asweep<2>(S, st, st > 0);
    batch_asweep<1>(S, st, st > 0);
    batch_asweep<0>(S, st, st > 0);
  }
  if (B.afreq[0] > 0) batch_filter<0>(S, 1);
  if (B.afreq[1] > 0) batch_filter<1>(S, 1);
  if (B.afreq[2] > 0) batch_filter<2>(S, 1);
  return check_launch();
}

// The launches of one batched period of a bin: with subsets, the adjoint
// members restore their subset's checkpoint, step it forward with the forward
// members recording its trajectory, then run backward over it; without, one
// forward step of the forward members and one backward step of the adjoint
// members.
int batch_record_period(const BatchSlot &S, int P, bool subsets) {
  const int nadj = S.bb().nadj;
  if (subsets) {
    if (nadj > 0) {
      hipLaunchKernelGGL(kb_restore, dim3(256, nadj), dim3(256), 0, S.s, S.mem(), S.list(1, 0, 0));
      hipLaunchKernelGGL(kb_count_down, dim3(1, nadj), dim3(64), 0, S.s, S.mem(), S.list(1, 0, 0));
      if (int rc = check_launch()) return rc;
    }
    for (int k = 1; k <= P; ++k)
      if (int rc = batch_forward_step(S, 1, k)) return rc;
    if (nadj > 0) {
      hipLaunchKernelGGL(kb_set_counter, dim3(1, nadj), dim3(64), 0, S.s, S.mem(), S.list(1, 0, 0), P);
      if (int rc = check_launch()) return rc;
      for (int k = 0; k < P; ++k)
        if (int rc = batch_adjoint_step(S)) return rc;
    }
    return CRUNA_HIP_OK;
  }
  if (int rc = batch_forward_step(S, 0, 0)) return rc;
  if (nadj > 0)
    if (int rc = batch_adjoint_step(S)) return rc;
  return CRUNA_HIP_OK;
}

// Bins of a batched run: automatic (nb < 0) takes one bin per hardware queue
// of the process (4), at most one per member.
int batch_bins(const cruna_hip_ensemble *e) {
  const int M = (int)e->m.size();
  const int nb = e->batching < 0 ? 4 : e->batching;
  return std::max(1, std::min(nb, M));
}

// Run `periods` batched periods of P steps (each bin's period recorded into
// one graph once).
int batch_run(cruna_hip_ensemble *e, int P, int periods) {
  const size_t M = e->m.size();
  const int nb = batch_bins(e);
  bool subsets = false;
  for (size_t i = 0; i < M; ++i) subsets = subsets || e->subm[i];
  std::vector<int> roles(M);
  for (size_t i = 0; i < M; ++i) roles[i] = e->subm[i] ? 2 : e->role[i];
  bool stale = e->bins.size() != (size_t)nb || e->bperiod != P || e->bgen.size() != M || roles != e->brole;
  for (size_t i = 0; i < M && !stale; ++i) stale = e->m[i]->gen != e->bgen[i];
  while (e->bstreams.size() < (size_t)nb) {
    hipStream_t st;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    e->bstreams.push_back(st);
  }
  for (auto *c : e->m) HIP_TRY(hipStreamSynchronize(c->stream));
  if (stale) {
    drop_batch(e);
    if (int rc = batch_upload(e, subsets, nb)) return rc;
    for (int k = 0; k < nb; ++k) {
      const BatchSlot S{e, k, e->bstreams[k]};
      HIP_TRY(hipStreamBeginCapture(S.s, hipStreamCaptureModeThreadLocal));
      const int rc = batch_record_period(S, P, subsets);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(S.s, &graph);
      if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        drop_batch(e);
        return rc ? rc : fail(CRUNA_HIP_EDEVICE, std::string("capture: ") + hipGetErrorString(ec));
      }
      const hipError_t ei = hipGraphInstantiate(&e->bins[k].exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        e->bins[k].exec = nullptr;
        drop_batch(e);
        return fail(CRUNA_HIP_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
      }
    }
    e->bperiod = P;
    e->bgen.resize(M);
    for (size_t i = 0; i < M; ++i) e->bgen[i] = e->m[i]->gen;
    e->brole = roles;
  }
  for (int r = 0; r < periods; ++r)
    for (int k = 0; k < nb; ++k) HIP_TRY(hipGraphLaunch(e->bins[k].exec, e->bstreams[k]));
  for (int k = 0; k < nb; ++k) HIP_TRY(hipStreamSynchronize(e->bstreams[k]));
  for (size_t i = 0; i < M; ++i) {
    cruna_hip_ctx *c = e->m[i];
    c->qs_fresh = false;
    e->nt[i] += e->role[i] == CRUNA_HIP_FORWARD ? P * periods : -P * periods;
  }
  return CRUNA_HIP_OK;
}

}  // namespace

extern "C" {

int cruna_hip_ensemble_create(const cruna_hip_cfg *cfgs, int nmembers, cruna_hip_ensemble **out) {
  if (!cfgs || !out || nmembers < 1) return fail(CRUNA_HIP_EINVAL, "bad ensemble arguments");
  *out = nullptr;
  std::unique_ptr<cruna_hip_ensemble> e(new cruna_hip_ensemble());
  for (int i = 0; i < nmembers; ++i) {
    cruna_hip_ctx *c = nullptr;
    if (int rc = cruna_hip_create(&cfgs[i], &c)) {
      for (auto *x : e->m) cruna_hip_destroy(x);
      return rc;
    }
    e->m.push_back(c);
    e->role.push_back(CRUNA_HIP_FORWARD);
    e->nt.push_back(0);
    e->subm.push_back(false);
    e->pos.push_back(0);
    e->warm.push_back(false);
  }
  *out = e.release();
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_destroy(cruna_hip_ensemble *e) {
  if (!e) return CRUNA_HIP_OK;
  for (auto *c : e->m) (void)hipStreamSynchronize(c->stream);
  for (auto st : e->bstreams) (void)hipStreamSynchronize(st);
  drop_graphs(e);
  drop_batch(e);
  if (e->bmem) (void)hipFree(e->bmem);
  if (e->blist) (void)hipFree(e->blist);
  for (auto st : e->bstreams) (void)hipStreamDestroy(st);
  for (auto *c : e->m) cruna_hip_destroy(c);
  delete e;
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_member(cruna_hip_ensemble *e, int i, cruna_hip_ctx **ctx) {
  if (!e || !ctx || i < 0 || i >= (int)e->m.size()) return fail(CRUNA_HIP_EINVAL, "bad member");
  *ctx = e->m[i];
  e->dirty = true;  // the caller may change the member (state, trajectory, role)
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_set_role(cruna_hip_ensemble *e, int i, int role, int nt) {
  if (!e || i < 0 || i >= (int)e->m.size()) return fail(CRUNA_HIP_EINVAL, "bad member");
  if (role != CRUNA_HIP_FORWARD && role != CRUNA_HIP_ADJOINT) return fail(CRUNA_HIP_EINVAL, "bad role");
  cruna_hip_ctx *c = e->m[i];
  if (role == CRUNA_HIP_ADJOINT) {
    if (int rc = adjoint_checks(c)) return rc;
    if (!c->qa || c->traj_slots == 0)
      return fail(CRUNA_HIP_EINVAL, "an adjoint member needs its adjoint state and trajectory first");
    if (c->ckpt_n >= 2) {
      // subset checkpoints: nt counts the backward steps over all subsets
      const int total = (c->ckpt_n - 1) * c->sub_steps;
      if (nt < 1 || nt > total) return fail(CRUNA_HIP_EINVAL, "nt outside the subsets' steps");
      const int sidx = (nt - 1) / c->sub_steps;
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipMemcpy(c->sckp, &sidx, sizeof(int), hipMemcpyHostToDevice));
    } else if (c->traj_slots > 1 && (nt < 1 || nt > c->traj_slots)) {
      return fail(CRUNA_HIP_EINVAL, "nt outside the trajectory");
    }
  }
  if (role == CRUNA_HIP_FORWARD && c->record)
    return fail(CRUNA_HIP_EINVAL, "forward members do not record trajectories");
  e->role[i] = role;
  e->nt[i] = nt;
  e->subm[i] = role == CRUNA_HIP_ADJOINT && c->ckpt_n >= 2;
  e->pos[i] = 0;
  e->warm[i] = false;
  e->dirty = true;
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_run(cruna_hip_ensemble *e, int nsteps) {
  if (!e) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (nsteps < 0) return fail(CRUNA_HIP_EINVAL, "nsteps must be >= 0");
  const size_t M = e->m.size();
  int P = 1;
  for (size_t i = 0; i < M; ++i) {
    P = lcm_i(P, member_period(e->m[i], e->role[i], e->subm[i]));
    if (e->role[i] == CRUNA_HIP_ADJOINT && (e->m[i]->traj_slots > 1 || e->subm[i]) && e->nt[i] - nsteps < 0)
      return fail(CRUNA_HIP_EINVAL, "an adjoint member would run past its first step");
    // a closure's work buffers can move when it trains, under a recorded graph
    if (e->m[i]->cnn) return fail(CRUNA_HIP_EINVAL, "ensemble members cannot carry a learned closure");
  }
  if (P > 64) return fail(CRUNA_HIP_EINVAL, "filter frequencies repeat only after more than 64 steps");
  // adjoint step counters live on the device; set them for this run (the
  // counters of subset members are kept by their own kernels)
  for (size_t i = 0; i < M; ++i) {
    cruna_hip_ctx *c = e->m[i];
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (e->role[i] == CRUNA_HIP_ADJOINT && !e->subm[i])
      HIP_TRY(hipMemcpy(c->ntp, &e->nt[i], sizeof(int), hipMemcpyHostToDevice));
    c->prof = false;
  }
  int left = nsteps;
  // batched launches over all members for whole periods, when the members
  // allow them (the rest of the steps as below)
  e->batched_steps = 0;
  if (e->batching != 0) {
    const int BP = batch_period(e, nsteps);
    if (BP > 0 && left >= BP) {
      const int periods = left / BP;
      if (int rc = batch_run(e, BP, periods)) return rc;
      e->batched_steps = BP * periods;
      left -= BP * periods;
      if (left == 0) return CRUNA_HIP_OK;
      for (size_t i = 0; i < M; ++i) {
        cruna_hip_ctx *c = e->m[i];
        if (e->role[i] == CRUNA_HIP_ADJOINT && !e->subm[i])
          HIP_TRY(hipMemcpy(c->ntp, &e->nt[i], sizeof(int), hipMemcpyHostToDevice));
      }
    }
  }
  // eager steps: members not yet in their repeating state, and the remainder
  // that does not fill a period
  bool all_warm = true;
  for (size_t i = 0; i < M; ++i) all_warm = all_warm && e->warm[i];
  int eager = all_warm ? left % P : std::min(left, P + left % P);
  for (size_t i = 0; i < M; ++i) {
    if (int rc = member_steps(e, i, eager)) return rc;
    if (eager >= P) e->warm[i] = true;
  }
  left -= eager;
  if (left == 0) {
    for (auto *c : e->m) HIP_TRY(hipStreamSynchronize(c->stream));
    return CRUNA_HIP_OK;
  }
  // a member changed through a setter since the recording (its buffers,
  // kernels or launch shapes may differ): record again
  if (e->gen.size() != M) e->dirty = true;
  for (size_t i = 0; i < M && !e->dirty; ++i) e->dirty = e->m[i]->gen != e->gen[i];
  // record one period of steps of each member (the recording runs nothing)
  if (e->dirty || e->exec.size() != M || e->period != P) {
    drop_graphs(e);
    e->exec.assign(M, nullptr);
    std::vector<int> nt_save = e->nt, pos_save = e->pos;
    for (size_t i = 0; i < M; ++i) {
      cruna_hip_ctx *c = e->m[i];
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = member_steps(e, i, P);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        e->nt = nt_save;
        e->pos = pos_save;
        drop_graphs(e);
        return rc ? rc : fail(CRUNA_HIP_EDEVICE, std::string("capture: ") + hipGetErrorString(ec));
      }
      const hipError_t ei = hipGraphInstantiate(&e->exec[i], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        e->exec[i] = nullptr;
        e->nt = nt_save;
        e->pos = pos_save;
        drop_graphs(e);
        return fail(CRUNA_HIP_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
      }
    }
    e->nt = nt_save;
    e->pos = pos_save;
    e->period = P;
    e->gen.resize(M);
    for (size_t i = 0; i < M; ++i) e->gen[i] = e->m[i]->gen;
    e->dirty = false;
  }
  // replay: members interleaved so that their streams overlap
  for (int r = 0; r < left / P; ++r)
    for (size_t i = 0; i < M; ++i) HIP_TRY(hipGraphLaunch(e->exec[i], e->m[i]->stream));
  for (auto *c : e->m) HIP_TRY(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < M; ++i) e->nt[i] += e->role[i] == CRUNA_HIP_FORWARD ? left : -left;
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_set_batching(cruna_hip_ensemble *e, int bins) {
  if (!e) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (bins < -1) return fail(CRUNA_HIP_EINVAL, "bins: -1 (automatic), 0 (per-member graphs) or a count");
  e->batching = bins;
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_batched_steps(const cruna_hip_ensemble *e, int *steps) {
  if (!e || !steps) return fail(CRUNA_HIP_EINVAL, "null argument");
  *steps = e->batched_steps;
  return CRUNA_HIP_OK;
}

int cruna_hip_ensemble_autotune(cruna_hip_ensemble *e, int steps, int *tile_lines) {
  if (!e) return fail(CRUNA_HIP_EINVAL, "null argument");
  for (size_t i = 0; i < e->m.size(); ++i)
    if (int rc = cruna_hip_autotune(e->m[i], steps, tile_lines ? tile_lines + 3 * i : nullptr))
      return rc;
  e->dirty = true;
  return CRUNA_HIP_OK;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// Learned subfilter closure (closure.hpp): a 3-D CNN on a periodic S^3 grid.

struct cruna_hip_cnn {
  int S = 0;
  long N = 0;  // S^3
  hipStream_t stream = nullptr;
  static constexpr int L = 4;
  static constexpr int HALO = 4;  // reach of the network: one point per layer
  const int ci[L] = {16, 32, 32, 32}, co[L] = {32, 32, 32, 16};
  long woff[L] = {}, boff[L] = {}, wpk[L] = {};
  long nparams = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  bf16 *wf = nullptr, *wt = nullptr;  // packed BF16 operands of all layers
  CnnNorm hn = {};              // normalisation (host copy)
  CnnNorm *dn = nullptr;        // its device copy
  double *fmean = nullptr;      // field means of the sample in act[0]
  double *stage = nullptr;      // host inputs of the staging calls (10 S^3)
  long adam_t = 0;
  // per-point work buffers, sized for `cap` points
  long cap = 0;
  bf16 *act[L] = {}, *z[L - 1] = {}, *dz_a = nullptr, *dz_b = nullptr;
  float *out32 = nullptr, *tgt = nullptr, *lpart = nullptr, *part = nullptr, *bpart = nullptr;
  long part_cap = 0;
  CropSpec *spec = nullptr;
  long spec_cap = 0;
  std::vector<bf16 *> ds_feat;
  std::vector<float *> ds_tgt;
  int users = 0;              // contexts the closure is attached to
  std::vector<void *> fixed;  // parameter buffers
  std::vector<void *> work;   // per-point buffers
};

static void cnn_user(cruna_hip_cnn *n, int delta) { n->users += delta; }

namespace {

constexpr int WGRAD_CHUNK = 8192;

// launch of a convolution (k_conv_tile): persistent work-groups, as many as
// the LDS lets reside on the MI355X's 256 CUs (160 KB each), a multiple of
// the XCD count
constexpr long CONV_CUS = 256, CONV_LDS_PER_CU = 163840;
template <int CK, int CN, int MODE>
void launch_conv(hipStream_t s, const ConvGeom &geo, long npts, const bf16 *in, const bf16 *W, const float *bias,
                 float *out32, bf16 *out16, bf16 *zsave, const bf16 *zin) {
  constexpr long lds = 27L * CN * CK * 2 + (long)CT_HALO * CK * 2;
  const long box = (long)geo.X * geo.Y * geo.Z;
  const long ntiles = npts / box * ((geo.X + CT_X - 1) / CT_X) * ((geo.Y + CT_Y - 1) / CT_Y) *
                      ((geo.Z + CT_Z - 1) / CT_Z);
  long G = std::min(ntiles, CONV_CUS * std::max(1L, CONV_LDS_PER_CU / lds));
  G = (G + CONV_XCDS - 1) / CONV_XCDS * CONV_XCDS;
  hipLaunchKernelGGL((k_conv_tile<CK, CN, MODE>), dim3((unsigned)G), dim3(CT_THREADS), 0, s, in, W, bias, geo, npts, out32,
                     out16, zsave, zin);
}

ConvGeom conv_geom(int X, int Y, int Z, bool periodic) {
  ConvGeom g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  g.periodic = periodic ? 1 : 0;
  g.inv_x = 1.0 / X;
  g.inv_xy = 1.0 / ((double)X * Y);
  g.inv_xyz = 1.0 / ((double)X * Y * Z);
  return g;
}

template <typename X>
int cnn_alloc(std::vector<void *> &list, X **p, size_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, count * sizeof(X)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(CRUNA_HIP_ENOMEM, "device allocation failed");
  }
  list.push_back(q);
  *p = (X *)q;
  return CRUNA_HIP_OK;
}

// work buffers for npts points (kept; grown when a larger batch comes)
int cnn_reserve(cruna_hip_cnn *n, long npts) {
  if (npts <= n->cap) return CRUNA_HIP_OK;
  if (npts * 32 >= (1L << 31)) return fail(CRUNA_HIP_EINVAL, "closure batch too large");
  HIP_TRY(hipStreamSynchronize(n->stream));
  for (void *q : n->work) (void)hipFree(q);
  n->work.clear();
  n->cap = 0;
  n->part_cap = 0;
  int rc = CRUNA_HIP_OK;
  for (int l = 0; l < n->L && !rc; ++l) rc = cnn_alloc(n->work, &n->act[l], npts * n->ci[l]);
  for (int l = 0; l < n->L - 1 && !rc; ++l) rc = cnn_alloc(n->work, &n->z[l], npts * n->co[l]);
  if (!rc) rc = cnn_alloc(n->work, &n->dz_a, npts * 32);
  if (!rc) rc = cnn_alloc(n->work, &n->dz_b, npts * 32);
  if (!rc) rc = cnn_alloc(n->work, &n->out32, npts * 16);
  if (!rc) rc = cnn_alloc(n->work, &n->tgt, npts * 5);
  if (!rc) rc = cnn_alloc(n->work, &n->lpart, (npts + 255) / 256);
  const long nch = (npts + WGRAD_CHUNK - 1) / WGRAD_CHUNK;
  if (!rc) rc = cnn_alloc(n->work, &n->part, (size_t)nch * 27 * 32 * 32);
  if (!rc) rc = cnn_alloc(n->work, &n->bpart, (size_t)nch * 32);
  if (rc) {
    for (void *q : n->work) (void)hipFree(q);
    n->work.clear();
    return rc;
  }
  n->cap = npts;
  n->part_cap = nch;
  return CRUNA_HIP_OK;
}

int cnn_pack(cruna_hip_cnn *n, hipStream_t s) {
  for (int l = 0; l < n->L; ++l) {
    const long e = 27L * n->co[l] * n->ci[l];
    hipLaunchKernelGGL(k_pack_weights, dim3((e + 255) / 256), dim3(256), 0, s, n->params + n->woff[l],
                       n->co[l], n->ci[l], n->wf + n->wpk[l], n->wt + n->wpk[l]);
  }
  return check_launch();
}

// forward pass from act[0] to out32 over npts points of geometry geo; with
// `train` the pre-activations are kept for the backward pass
int cnn_forward(cruna_hip_cnn *n, hipStream_t s, const ConvGeom &geo, long npts, bool train) {
  if (npts % 64 != 0 || npts > n->cap) return fail(CRUNA_HIP_EINVAL, "closure batch size");
  const float *P = n->params;
  launch_conv<16, 32, CONV_GELU>(s, geo, npts, n->act[0], n->wf + n->wpk[0], P + n->boff[0], nullptr, n->act[1],
                                 train ? n->z[0] : nullptr, nullptr);
  launch_conv<32, 32, CONV_GELU>(s, geo, npts, n->act[1], n->wf + n->wpk[1], P + n->boff[1], nullptr, n->act[2],
                                 train ? n->z[1] : nullptr, nullptr);
  launch_conv<32, 32, CONV_GELU>(s, geo, npts, n->act[2], n->wf + n->wpk[2], P + n->boff[2], nullptr, n->act[3],
                                 train ? n->z[2] : nullptr, nullptr);
  launch_conv<32, 16, CONV_LINEAR>(s, geo, npts, n->act[3], n->wf + n->wpk[3], P + n->boff[3], n->out32, nullptr,
                                   nullptr, nullptr);
  return check_launch();
}

template <int CI, int CO>
void cnn_wgrad(cruna_hip_cnn *n, hipStream_t s, int l, const ConvGeom &geo, long npts, const bf16 *in,
               const bf16 *dz) {
  const int nch = (int)((npts + WGRAD_CHUNK - 1) / WGRAD_CHUNK);
  hipLaunchKernelGGL((k_conv_wgrad<CI, CO>), dim3(nch, 27), dim3(256), 0, s, in, dz, geo, npts, WGRAD_CHUNK,
                     n->part, n->bpart);
  const long e = 27L * CO * CI;
  hipLaunchKernelGGL(k_sum_chunks, dim3((e + 255) / 256), dim3(256), 0, s, (const float *)n->part, nch, e,
                     n->grads + n->woff[l]);
  hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(256), 0, s, (const float *)n->bpart, nch, (long)CO,
                     n->grads + n->boff[l]);
}

// loss over the points at least `halo` from their box's faces (count of
// them: counted) and the gradients of all parameters, after cnn_forward
// with `train`; *loss (if set) receives the mean-square loss
int cnn_backward(cruna_hip_cnn *n, hipStream_t s, const ConvGeom &geo, long npts, int halo, long counted,
                 float *loss) {
  const int nb = (int)((npts + 255) / 256);
  const float scale = (float)(2.0 / ((double)counted * 5));
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, (const float *)n->out32, (const float *)n->tgt, npts,
                     5, geo, halo, scale, n->dz_a, n->lpart);
  // layer 3 (32 -> 16, linear)
  cnn_wgrad<32, 16>(n, s, 3, geo, npts, n->act[3], n->dz_a);
  launch_conv<16, 32, CONV_BACKWARD>(s, geo, npts, n->dz_a, n->wt + n->wpk[3], nullptr, nullptr, n->dz_b, nullptr,
                                     n->z[2]);
  // layer 2
  cnn_wgrad<32, 32>(n, s, 2, geo, npts, n->act[2], n->dz_b);
  launch_conv<32, 32, CONV_BACKWARD>(s, geo, npts, n->dz_b, n->wt + n->wpk[2], nullptr, nullptr, n->dz_a, nullptr,
                                     n->z[1]);
  // layer 1
  cnn_wgrad<32, 32>(n, s, 1, geo, npts, n->act[1], n->dz_a);
  launch_conv<32, 32, CONV_BACKWARD>(s, geo, npts, n->dz_a, n->wt + n->wpk[1], nullptr, nullptr, n->dz_b, nullptr,
                                     n->z[0]);
  // layer 0 (16 -> 32)
  cnn_wgrad<16, 32>(n, s, 0, geo, npts, n->act[0], n->dz_b);
  if (int rc = check_launch()) return rc;
  if (loss) {
    std::vector<float> h(nb);
    HIP_TRY(hipMemcpyAsync(h.data(), n->lpart, nb * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double sum = 0.0;
    for (float x : h) sum += x;
    *loss = (float)(sum / ((double)counted * 5));
  }
  return CRUNA_HIP_OK;
}

// closure evaluation on stream s: features of the state q (device, 5
// fields of S^3), forward, force (5 fields of T) when `force` is set
template <typename T, typename Q>
int closure_eval(cruna_hip_cnn *n, hipStream_t s, const Q *q, T *force) {
  const unsigned nb = (unsigned)((n->N + 255) / 256);
  hipLaunchKernelGGL((k_field_partials<Q>), dim3(MEAN_BLOCKS), dim3(256), 0, s, q, n->N, n->fmean + 8);
  hipLaunchKernelGGL(k_field_means, dim3(1), dim3(256), 0, s, (const double *)(n->fmean + 8), MEAN_BLOCKS, n->N,
                     n->fmean);
  hipLaunchKernelGGL((k_pack_features<Q>), dim3(nb), dim3(256), 0, s, q, n->N, (const double *)n->fmean,
                     (const CnnNorm *)n->dn, n->act[0]);
  if (int rc = check_launch()) return rc;
  if (int rc = cnn_forward(n, s, conv_geom(n->S, n->S, n->S, true), n->N, false)) return rc;
  if (force) {
    hipLaunchKernelGGL((k_unpack_closure<T>), dim3(nb), dim3(256), 0, s, (const float *)n->out32, n->N,
                       (const CnnNorm *)n->dn, (const double *)n->fmean, force);
    // domain means of the force (after the state's means are consumed)
    hipLaunchKernelGGL((k_field_partials<T>), dim3(MEAN_BLOCKS), dim3(256), 0, s, (const T *)force, n->N,
                       n->fmean + 8);
    hipLaunchKernelGGL(k_field_means, dim3(1), dim3(256), 0, s, (const double *)(n->fmean + 8), MEAN_BLOCKS, n->N,
                       n->fmean);
    hipLaunchKernelGGL((k_remove_mean<T>), dim3(nb), dim3(256), 0, s, force, n->N, (const double *)n->fmean);
  }
  return check_launch();
}

template <typename T, typename Q>
int closure_term(cruna_hip_ctx *c, const Q *in) {
  ProfScope p(c, KID_CLOSURE);
  return closure_eval<T, Q>(c->cnn, c->stream, in, (T *)c->force);
}

// what cnn_stage does with a state from the host
enum StageWhat { STAGE_FEATURES, STAGE_TRAIN, STAGE_FORCE };

// stage a state (5 fields, FP64, host) as features in act[0] and the
// normalised target, if given, in tgt; STAGE_TRAIN then runs the forward
// pass keeping what the backward pass needs, STAGE_FORCE the closure term
// (into the second half of `stage`)
int cnn_stage(cruna_hip_cnn *n, const double *q, const double *target, StageWhat what) {
  if (int rc = cnn_reserve(n, n->N)) return rc;
  if (!n->stage) HIP_TRY(hipMalloc(&n->stage, 10 * n->N * sizeof(double)));
  const unsigned nb = (unsigned)((n->N + 255) / 256);
  double *qd = n->stage, *td = n->stage + 5 * n->N;
  HIP_TRY(hipMemcpyAsync(qd, q, 5 * n->N * sizeof(double), hipMemcpyHostToDevice, n->stream));
  if (target) HIP_TRY(hipMemcpyAsync(td, target, 5 * n->N * sizeof(double), hipMemcpyHostToDevice, n->stream));
  if (what == STAGE_FORCE) {
    if (int rc = closure_eval<double, double>(n, n->stream, qd, td)) return rc;
  } else {
    hipLaunchKernelGGL((k_field_partials<double>), dim3(MEAN_BLOCKS), dim3(256), 0, n->stream, (const double *)qd,
                       n->N, n->fmean + 8);
    hipLaunchKernelGGL(k_field_means, dim3(1), dim3(256), 0, n->stream, (const double *)(n->fmean + 8), MEAN_BLOCKS,
                       n->N, n->fmean);
    hipLaunchKernelGGL((k_pack_features<double>), dim3(nb), dim3(256), 0, n->stream, (const double *)qd, n->N,
                       (const double *)n->fmean, (const CnnNorm *)n->dn, n->act[0]);
    if (int rc = check_launch()) return rc;
    if (what == STAGE_TRAIN)
      if (int rc = cnn_forward(n, n->stream, conv_geom(n->S, n->S, n->S, true), n->N, true)) return rc;
  }
  if (target)
    hipLaunchKernelGGL(k_scale_targets, dim3(nb), dim3(256), 0, n->stream, (const double *)td, n->N,
                       (const CnnNorm *)n->dn, (const double *)n->fmean, n->tgt);
  return check_launch();
}

}  // namespace

extern "C" {

int cruna_hip_cnn_create(int S, cruna_hip_cnn **out) {
  if (!out) return fail(CRUNA_HIP_EINVAL, "null argument");
  *out = nullptr;
  if (S < 16 || S > 512 || (S & (S - 1)) != 0) return fail(CRUNA_HIP_EINVAL, "S must be a power of two in [16, 512]");
  std::unique_ptr<cruna_hip_cnn> n(new cruna_hip_cnn());
  n->S = S;
  n->N = (long)S * S * S;
  long o = 0, pk = 0;
  for (int l = 0; l < n->L; ++l) {
    n->woff[l] = o;
    o += 27L * n->co[l] * n->ci[l];
    n->boff[l] = o;
    o += n->co[l];
    n->wpk[l] = pk;
    pk += 27L * n->co[l] * n->ci[l];
  }
  n->nparams = o;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(CRUNA_HIP_EDEVICE, "stream creation failed");
  int rc = cnn_alloc(n->fixed, &n->params, o);
  if (!rc) rc = cnn_alloc(n->fixed, &n->grads, o);
  if (!rc) rc = cnn_alloc(n->fixed, &n->m, o);
  if (!rc) rc = cnn_alloc(n->fixed, &n->v, o);
  if (!rc) rc = cnn_alloc(n->fixed, &n->wf, pk);
  if (!rc) rc = cnn_alloc(n->fixed, &n->wt, pk);
  if (!rc) rc = cnn_alloc(n->fixed, &n->dn, 1);
  if (!rc) rc = cnn_alloc(n->fixed, &n->fmean, 8 + 5 * MEAN_BLOCKS);  // means, then partial sums
  if (!rc) rc = cnn_reserve(n.get(), n->N);
  for (int k = 0; k < 5; ++k) n->hn.inv_std[k] = n->hn.out_std[k] = n->hn.out_inv_std[k] = n->hn.gain[k] = 1.f;
  n->hn.p_ref = 0.0;
  if (!rc && (hipMemcpy(n->dn, &n->hn, sizeof n->hn, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemset(n->params, 0, o * sizeof(float)) != hipSuccess ||
              hipMemset(n->m, 0, o * sizeof(float)) != hipSuccess ||
              hipMemset(n->v, 0, o * sizeof(float)) != hipSuccess))
    rc = fail(CRUNA_HIP_EDEVICE, "initialisation failed");
  if (!rc) rc = cnn_pack(n.get(), n->stream);
  if (rc) {
    cruna_hip_cnn *raw = n.release();
    for (void *q : raw->fixed) (void)hipFree(q);
    for (void *q : raw->work) (void)hipFree(q);
    (void)hipStreamDestroy(raw->stream);
    delete raw;
    return rc;
  }
  *out = n.release();
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_destroy(cruna_hip_cnn *n) {
  if (!n) return CRUNA_HIP_OK;
  if (n->users > 0)
    return fail(CRUNA_HIP_EINVAL, "the closure is still attached to a context (cruna_hip_set_closure(ctx, NULL))");
  (void)hipStreamSynchronize(n->stream);
  for (void *q : n->fixed) (void)hipFree(q);
  for (void *q : n->work) (void)hipFree(q);
  if (n->spec) (void)hipFree(n->spec);
  if (n->stage) (void)hipFree(n->stage);
  for (auto *q : n->ds_feat) (void)hipFree(q);
  for (auto *q : n->ds_tgt) (void)hipFree(q);
  (void)hipStreamDestroy(n->stream);
  delete n;
  return CRUNA_HIP_OK;
}

long cruna_hip_cnn_num_params(const cruna_hip_cnn *n) { return n ? n->nparams : 0; }

int cruna_hip_cnn_set_params(cruna_hip_cnn *n, const float *params) {
  if (!n || !params) return fail(CRUNA_HIP_EINVAL, "null argument");
  HIP_TRY(hipMemcpyAsync(n->params, params, n->nparams * sizeof(float), hipMemcpyHostToDevice, n->stream));
  HIP_TRY(hipMemsetAsync(n->m, 0, n->nparams * sizeof(float), n->stream));
  HIP_TRY(hipMemsetAsync(n->v, 0, n->nparams * sizeof(float), n->stream));
  n->adam_t = 0;
  if (int rc = cnn_pack(n, n->stream)) return rc;
  HIP_TRY(hipStreamSynchronize(n->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_get_params(cruna_hip_cnn *n, float *params) {
  if (!n || !params) return fail(CRUNA_HIP_EINVAL, "null argument");
  HIP_TRY(hipMemcpyAsync(params, n->params, n->nparams * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_init_params(cruna_hip_cnn *n, unsigned seed) {
  if (!n) return fail(CRUNA_HIP_EINVAL, "null argument");
  const int real_in[4] = {5, 32, 32, 32}, real_out[4] = {32, 32, 32, 5};
  std::vector<float> h(n->nparams, 0.f);
  std::mt19937 gen(seed);
  for (int l = 0; l < n->L; ++l) {
    // He initialisation; the last layer starts small so that an untrained
    // closure is a small perturbation of the RHS
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * real_in[l])) * (l == 3 ? 0.1f : 1.f));
    for (int t = 0; t < 27; ++t)
      for (int o = 0; o < real_out[l]; ++o)
        for (int i = 0; i < real_in[l]; ++i) h[n->woff[l] + ((long)t * n->co[l] + o) * n->ci[l] + i] = nd(gen);
  }
  return cruna_hip_cnn_set_params(n, h.data());
}

int cruna_hip_cnn_set_norm(cruna_hip_cnn *n, const double in_std[5], const double out_std[5], double p_ref) {
  if (!n || !in_std || !out_std) return fail(CRUNA_HIP_EINVAL, "null argument");
  for (int k = 0; k < 5; ++k)
    if (!(in_std[k] > 0) || !(out_std[k] > 0)) return fail(CRUNA_HIP_EINVAL, "scales must be positive");
  if (!(p_ref >= 0)) return fail(CRUNA_HIP_EINVAL, "p_ref must be >= 0");
  for (int k = 0; k < 5; ++k) {
    n->hn.inv_std[k] = (float)(1.0 / in_std[k]);
    n->hn.out_std[k] = (float)out_std[k];
    n->hn.out_inv_std[k] = (float)(1.0 / out_std[k]);
  }
  n->hn.p_ref = p_ref;
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(n->dn, &n->hn, sizeof n->hn, hipMemcpyHostToDevice));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_set_gain(cruna_hip_cnn *n, const double gain[5]) {
  if (!n || !gain) return fail(CRUNA_HIP_EINVAL, "null argument");
  for (int k = 0; k < 5; ++k)
    if (!(gain[k] >= 0.0)) return fail(CRUNA_HIP_EINVAL, "gains must be >= 0");
  for (int k = 0; k < 5; ++k) n->hn.gain[k] = (float)gain[k];
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(n->dn, &n->hn, sizeof n->hn, hipMemcpyHostToDevice));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_forward(cruna_hip_cnn *n, const double *q, double *closure) {
  if (!n || !q || !closure) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = cnn_stage(n, q, nullptr, STAGE_FORCE)) return rc;
  HIP_TRY(hipMemcpyAsync(closure, n->stage + 5 * n->N, 5 * n->N * sizeof(double), hipMemcpyDeviceToHost,
                         n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_gradients(cruna_hip_cnn *n, const double *q, const double *target, float *grads,
                            float *loss) {
  if (!n || !q || !target) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = cnn_stage(n, q, target, STAGE_TRAIN)) return rc;
  const ConvGeom geo = conv_geom(n->S, n->S, n->S, true);
  if (int rc = cnn_backward(n, n->stream, geo, n->N, 0, n->N, loss)) return rc;
  if (grads)
    HIP_TRY(hipMemcpyAsync(grads, n->grads, n->nparams * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_add_sample(cruna_hip_cnn *n, const double *q, const double *target) {
  if (!n || !q || !target) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (int rc = cnn_stage(n, q, target, STAGE_FEATURES)) return rc;
  bf16 *f = nullptr;
  float *t = nullptr;
  HIP_TRY(hipMalloc(&f, n->N * 16 * sizeof(bf16)));
  if (hipMalloc(&t, n->N * 5 * sizeof(float)) != hipSuccess) {
    (void)hipFree(f);
    return fail(CRUNA_HIP_ENOMEM, "device allocation failed");
  }
  n->ds_feat.push_back(f);
  n->ds_tgt.push_back(t);
  HIP_TRY(hipMemcpyAsync(f, n->act[0], n->N * 16 * sizeof(bf16), hipMemcpyDeviceToDevice, n->stream));
  HIP_TRY(hipMemcpyAsync(t, n->tgt, n->N * 5 * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_cnn_num_samples(const cruna_hip_cnn *n) { return n ? (int)n->ds_feat.size() : 0; }

int cruna_hip_cnn_train(cruna_hip_cnn *n, int iters, float lr, int batch, int crop, unsigned seed,
                        float *losses) {
  if (!n) return fail(CRUNA_HIP_EINVAL, "null argument");
  if (n->ds_feat.empty()) return fail(CRUNA_HIP_EINVAL, "no training samples (cruna_hip_cnn_add_sample)");
  if (iters < 0 || batch < 1 || batch > 64) return fail(CRUNA_HIP_EINVAL, "iters >= 0 and batch in [1, 64]");
  if (crop < 0 || crop > n->S || (crop && crop % 4)) return fail(CRUNA_HIP_EINVAL, "crop: 0 or a multiple of 4 up to S");
  // crop 0: whole periodic samples; else windows crop^3 with a halo of the
  // network's reach, the loss taken on the crop^3 centre
  const int H = crop ? n->HALO : 0;
  const int X = crop ? crop + 2 * H : n->S;
  const long per = (long)X * X * X;
  const long npts = per * batch;
  if (npts % 64 != 0) return fail(CRUNA_HIP_EINVAL, "batch points must be a multiple of 64");
  const long counted = crop ? (long)crop * crop * crop * batch : npts;
  const ConvGeom geo = conv_geom(X, X, X, crop == 0);
  if (int rc = cnn_reserve(n, npts)) return rc;
  // every step's batch drawn up front: sample and window offset per entry
  std::mt19937 gen(seed);
  std::uniform_int_distribution<int> pick(0, (int)n->ds_feat.size() - 1), off(0, n->S - 1);
  std::vector<CropSpec> h((size_t)iters * batch);
  for (auto &sp : h) {
    const int s = pick(gen);
    sp.feat = n->ds_feat[s];
    sp.tgt = n->ds_tgt[s];
    sp.ox = crop ? (off(gen) - H) & (n->S - 1) : 0;
    sp.oy = crop ? (off(gen) - H) & (n->S - 1) : 0;
    sp.oz = crop ? (off(gen) - H) & (n->S - 1) : 0;
    sp.pad = 0;
  }
  if ((long)h.size() > n->spec_cap) {
    if (n->spec) (void)hipFree(n->spec);
    n->spec = nullptr;
    n->spec_cap = 0;
    HIP_TRY(hipMalloc(&n->spec, std::max<size_t>(h.size(), 1) * sizeof(CropSpec)));
    n->spec_cap = (long)h.size();
  }
  if (!h.empty())
    HIP_TRY(hipMemcpyAsync(n->spec, h.data(), h.size() * sizeof(CropSpec), hipMemcpyHostToDevice, n->stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const unsigned nbp = (unsigned)((n->nparams + 255) / 256);
  const unsigned nbg = (unsigned)((npts + 255) / 256);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_gather_crops, dim3(nbg), dim3(256), 0, n->stream,
                       (const CropSpec *)(n->spec + (long)it * batch), n->S, geo, npts, n->act[0], n->tgt);
    if (int rc = cnn_forward(n, n->stream, geo, npts, true)) return rc;
    float loss = 0.f;
    if (int rc = cnn_backward(n, n->stream, geo, npts, H, counted, losses ? &loss : nullptr)) return rc;
    if (losses) losses[it] = loss;
    ++n->adam_t;
    const double bc1 = 1.0 - std::pow((double)b1, (double)n->adam_t);
    const double bc2 = 1.0 - std::pow((double)b2, (double)n->adam_t);
    const float lr_t = (float)(lr * std::sqrt(bc2) / bc1);
    const float eps_t = (float)(eps * std::sqrt(bc2));
    hipLaunchKernelGGL(k_adam, dim3(nbp), dim3(256), 0, n->stream, n->params, (const float *)n->grads, n->m, n->v,
                       n->nparams, lr_t, b1, b2, eps_t);
    if (int rc = cnn_pack(n, n->stream)) return rc;
  }
  HIP_TRY(hipStreamSynchronize(n->stream));
  return CRUNA_HIP_OK;
}

int cruna_hip_set_closure(cruna_hip_ctx *c, cruna_hip_cnn *n) {
  if (!c) return fail(CRUNA_HIP_EINVAL, "null argument");
  ++c->gen;
  if (n) {
    for (int d = 0; d < 3; ++d)
      if (c->cfg.n[d] != n->S || c->cfg.scheme[d] != CRUNA_HIP_PERIODIC)
        return fail(CRUNA_HIP_EINVAL, "the closure needs the periodic S^3 grid it was built for");
    if (int rc = cnn_reserve(n, n->N)) return rc;
    if (!c->force) {
      const size_t ts = c->cfg.precision == CRUNA_HIP_FP32_MIXED ? sizeof(float) : sizeof(double);
      HIP_TRY(hipMalloc(&c->force, 5 * c->N * ts));
    }
    HIP_TRY(hipStreamSynchronize(n->stream));
  }
  if (c->cnn != n) {
    if (c->cnn) cnn_user(c->cnn, -1);
    if (n) cnn_user(n, +1);
  }
  c->cnn = n;
  return CRUNA_HIP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free direct RHS (cruna_hip_rhs_direct_noalloc).
//
// Its staging memory is one file-scope device array, `scratch`, of
// NOALLOC_WORDS doubles (128 MiB per GPU, held until the process ends): the
// input state in the context's device layout (5 N words) and a contiguous
// copy of the real points for the host transfers (5 N_real words).  Host code
// never uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, asked once per device with that device current, and
// is kept per device.  Every call on a GPU shares that GPU's array, so calls
// that overlap on one GPU race; each call writes every word it reads.
constexpr long NOALLOC_WORDS = 1L << 24;
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on device dev (looked up once, with dev current)
int scratch_on(int dev, double **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(CRUNA_HIP_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess)
      return fail(CRUNA_HIP_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return CRUNA_HIP_OK;
}

// Real points (n1, n2, n3; contiguous, Fortran order) of nv fields into the
// padded device layout (np1, np2, np3, front padding pf), padding points set
// to fill[v].
__global__ void k_scatter_real(const double *__restrict__ src, double *__restrict__ dst, int nv, int n1, int n2,
                               int n3, int p1, int p2, int p3, int f1, int f2, int f3, double fill0,
                               double fill4) {
  const long N = (long)p1 * p2 * p3, Nr = (long)n1 * n2 * n3;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nv * N; t += (long)gridDim.x * blockDim.x) {
    const int v = (int)(t / N);
    const long p = t - v * N;
    const int i = (int)(p % p1) - f1, j = (int)((p / p1) % p2) - f2, k = (int)(p / ((long)p1 * p2)) - f3;
    const bool real = i >= 0 && i < n1 && j >= 0 && j < n2 && k >= 0 && k < n3;
    dst[t] = real ? src[v * Nr + i + (long)n1 * (j + (long)n2 * k)] : (v == 0 ? fill0 : (v == 4 ? fill4 : 0.0));
  }
}

// The real points of nv padded device-layout fields (type T) into a
// contiguous FP64 array.
template <typename T>
__global__ void k_gather_real(const T *__restrict__ src, double *__restrict__ dst, int nv, int n1, int n2, int n3,
                              int p1, int p2, int p3, int f1, int f2, int f3) {
  const long N = (long)p1 * p2 * p3, Nr = (long)n1 * n2 * n3;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nv * Nr; t += (long)gridDim.x * blockDim.x) {
    const int v = (int)(t / Nr);
    const long r = t - v * Nr;
    const int i = (int)(r % n1), j = (int)((r / n1) % n2), k = (int)(r / ((long)n1 * n2));
    dst[t] = double(src[v * N + (i + f1) + (long)p1 * ((j + f2) + (long)p2 * (k + f3))]);
  }
}

}  // namespace

extern "C" {

int cruna_hip_rhs_direct_noalloc(cruna_hip_ctx *c, double *rhs, const double *q, double t) {
  (void)t;  // the direct RHS of the reference does not depend on t
  if (!c || !rhs || !q) return fail(CRUNA_HIP_EINVAL, "null argument");
  const int *n = c->cfg.n;
  const long Nr = (long)n[0] * n[1] * n[2];
  if (5 * (c->N + Nr) > NOALLOC_WORDS)
    return fail(CRUNA_HIP_EINVAL, "cruna_hip_rhs_direct_noalloc: 5 (N_device + N) = " +
                                      std::to_string(5 * (c->N + Nr)) + " words exceed its static scratch of " +
                                      std::to_string(NOALLOC_WORDS) + " doubles");
  double *s = nullptr;
  if (int rc = scratch_on(c->device, &s)) return rc;
  double *stage = s, *real = s + 5 * c->N;
  c->qs_fresh = false;
  const int *np = c->np, *pf = c->pf;
  const unsigned blocks = (unsigned)std::min<long>(8192, (5 * c->N + 255) / 256);
  // the input state: contiguous host copy, then the device layout
  HIP_TRY(hipMemcpyAsync(real, q, 5 * Nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_scatter_real, dim3(blocks), dim3(256), 0, c->stream, (const double *)real, stage, 5, n[0],
                     n[1], n[2], np[0], np[1], np[2], pf[0], pf[1], pf[2], kRestState[0], kRestState[4]);
  if (int rc = check_launch()) return rc;
  const bool mixed = c->cfg.precision == CRUNA_HIP_FP32_MIXED;
  int rc = mixed ? run_stage<float>(c, stage, 4) : run_stage<double>(c, stage, 4);
  if (rc) return rc;
  // the RHS: real points in FP64, then to the host
  if (mixed)
    hipLaunchKernelGGL((k_gather_real<float>), dim3(blocks), dim3(256), 0, c->stream, (const float *)c->rhs, real,
                       5, n[0], n[1], n[2], np[0], np[1], np[2], pf[0], pf[1], pf[2]);
  else
    hipLaunchKernelGGL((k_gather_real<double>), dim3(blocks), dim3(256), 0, c->stream, (const double *)c->rhs,
                       real, 5, n[0], n[1], n[2], np[0], np[1], np[2], pf[0], pf[1], pf[2]);
  if (int rc2 = check_launch()) return rc2;
  HIP_TRY(hipMemcpyAsync(rhs, real, 5 * Nr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CRUNA_HIP_OK;
}

}  // extern "C"
