// Fused direct-RHS sweeps, filter sweep and single-derivative kernel.
//
// One RK stage of right_hand_side_direct_3d (src/equations/ns_rho_rhou_p.f90:31-253)
// is evaluated in six line sweeps; each sweep applies every compact
// derivative that runs along its direction, so the 42 Dx calls of the
// reference become 42 line solves on register-resident blocks:
//   S1 x : u_x v_x w_x, d/dx of (rho u), (rho u u_i), p, (p u), T and lambda T_x
//   S2 y : the same along y                                  (accumulates rhs)
//   S3 z : the same along z
//   S4 z : d/dz tau_{iz}
//   S5 y : d/dy tau_{iy}
//   S6 x : d/dx tau_{ix} and tau:grad(u), then the RK4 stage update
//          (src/time/rk_o4_e.f90:45-50)
// The pressure equation is accumulated as the bracket of ns_rho_rhou_p.f90:200-206
// and multiplied by (gamma - 1) once, as the reference does.
//
//
// Field layout in HBM: every field is f(n1,n2,n3) Fortran order; a state holds
// 5 fields at a stride of N = n1*n2*n3 elements.  The viscous stress and the
// dissipation use the velocity gradient only through its diagonal and the
// symmetric sums, so G holds 6 fields, and a seventh with the viscosity:
//   G0 = u_x, G1 = v_y, G2 = w_z, G3 = u_y + v_x, G4 = u_z + w_x, G5 = v_z + w_y,
//   G6 = mu(T) of the stage input
// (each sum is completed by the second sweep that contributes to it; the
// result equals the reference's u_y + v_x etc. bit for bit).  S1 x evaluates
// the viscosity law once per point and stage; S2, S3 and S4-S6 read it.
#pragma once

#include "line_solve.hpp"

namespace cruna {

struct Phys {
  double gamma;
  double gm1;        // gamma - 1
  double gt;         // gamma / (gamma - 1)
  double w_inv_r;    // W / R     (give_T)
  double mu0, T0;
  double visc_exp;
  int visc_law;      // 0 Sutherland, 1 power law
  double cp;         // gamma (R/W) / (gamma - 1)
  double Pr;
  double lam_coef;   // cp / Pr  (lambda = mu cp / Pr)
  double inv_T0;     // 1 / T0
  double inv_h[3];       // 1/dx_d
  double heat_inv_h[3];  // 1/dx_d, or 1 for the reference's unit-spacing heat flux
};

// Deformed grid with a diagonal metric whose entries M(d,d) are constant along
// the lines of direction d (every tensor-product stretch of a Cartesian grid:
// ns_rho_rhou_p_grid_transform.f90 with X_metric from grid_metric.f90).
// Stored per line of each direction and the Jacobian per point, in the
// compute precision.
template <typename T>
struct Metric {
  const T *m[3];  // M(d,d) of the line through (j,k) / (i,k) / (i,j): index ThreadPos::lidx
  const T *J;     // X_jacobian(n1,n2,n3)
};

// Volume penalisation (ns_rho_rhou_p_ev_phi.f90): the porosity phi and the
// stored 1/phi (volume_penalization's penalization(:,:,:,1:2)) per point.
template <typename T>
struct Penal {
  const T *phi;
  const T *inv_phi;
};

// Right-hand-side form of the sweeps: Cartesian (ns_rho_rhou_p), deformed
// grid with a tensor-product metric (ns_rho_rhou_p_grid_transform), or
// volume-penalised (ns_rho_rhou_p_ev_phi).
enum RhsForm { RHS_CART = 0, RHS_GT = 1, RHS_EV = 2 };

// give_mu, ns_rho_rhou_p_trafos.f90:119-142: mu0*(T/T0)**1.5*((T0+S1)/(T+S1)),
// or the power law mu0*(T/T0)**exp of the case-level override.
template <typename T>
__device__ __forceinline__ T viscosity(const Phys &ph, T Tk) {
  const T r = Tk * T(ph.inv_T0);
  // r > 0, so r**e = 2**(e log2 r) without pow's special-case branches
  if (ph.visc_law == 1) return T(ph.mu0) * exp2(T(ph.visc_exp) * log2(r));
  const T S1 = T(110.4);
  return T(ph.mu0) * (r * sqrt(r)) * ((T(ph.T0) + S1) / (Tk + S1));
}

// Issue point of a group of global loads: the compiler may not sink them below
// this line (it otherwise moves each point's loads down to its first use and
// serialises their latencies).
__device__ __forceinline__ void issue_loads() { asm volatile("" ::: "memory"); }

// Materialise a value here: the compiler otherwise sinks a computation to its
// (late) use and keeps all of its operands live across the line solves between.
template <typename T>
__device__ __forceinline__ void materialize(T (&a)[PPT]) {
#pragma unroll
  for (int m = 0; m < PPT; ++m) asm volatile("" : "+v"(a[m]));
}

// give_T (:57) and give_lambda (:169): T = (W/R) p / rho, lambda = mu cp / Pr,
// with 1/rho formed once per point by the caller
template <typename T>
__device__ __forceinline__ T temperature(const Phys &ph, T irho, T p) {
  return T(ph.w_inv_r) * p * irho;
}
template <typename T>
__device__ __forceinline__ T lambda_of(const Phys &ph, T mu) {
  return mu * T(ph.lam_coef);
}

// A thread's block is PPT consecutive points of its line.  Along x those are
// contiguous and 64-byte aligned (FP64), so they move as 16-byte vectors;
// along y/z they are PS elements apart.  Field accesses are raw buffer loads
// and stores: the field base in a buffer resource (SGPRs) and 32-bit byte
// offsets (the host checks that a field fits in 2 GiB).  Along x and y a
// block costs one offset VGPR: the distance m*PS between its points goes in
// the scalar offset.  Along z that distance is a multiple of the plane size
// (512 KiB at 256^2 FP64), and measured on the 256^3 step the z sweeps run
// 5-40 % faster with the point offsets in VGPRs and a zero scalar offset.
// Buffer accesses count against vmcnt only, so the line engine's LDS waits
// never wait for them.
typedef unsigned int bu2 __attribute__((ext_vector_type(2)));
typedef unsigned int bu4 __attribute__((ext_vector_type(4)));

// Buffer extent: 2 GiB from the field base (the host keeps one field below
// that).  An idle thread of a batched launch (IDLE_BASE) has the extent
// itself as its FP64 offset, so that its loads return zero and its stores are
// dropped by the buffer range check (which tests the per-thread offset); along
// z its point offsets stay below 4 GiB, as a field's plane distance times 7
// stays below 2 GiB.
constexpr unsigned FIELD_EXTENT = 0x80000000u;
static_assert(IDLE_BASE * 8 == FIELD_EXTENT, "idle FP64 offset is the buffer extent");

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p) {
  // raw buffer (stride 0); word 3 is the gfx9 format word of a plain dword
  // buffer
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, int(FIELD_EXTENT), 0x00020000);
}

template <int DIR, typename T, typename Q>
__device__ __forceinline__ void ld(const Q *__restrict__ p, long base, long PS, T (&v)[PPT]) {
  const __amdgpu_buffer_rsrc_t r = rsrc(p);
  const unsigned vo = unsigned(base) * unsigned(sizeof(Q));
  if (DIR == 0) {
#pragma unroll
    for (int k = 0; k < PPT * int(sizeof(Q)) / 16; ++k) {
      const bu4 t = __builtin_amdgcn_raw_buffer_load_b128(r, vo, k * 16, 0);
      if constexpr (sizeof(Q) == 8) {
        v[2 * k] = T(__builtin_bit_cast(double, bu2{t[0], t[1]}));
        v[2 * k + 1] = T(__builtin_bit_cast(double, bu2{t[2], t[3]}));
      } else {
        // each element through a scalar: a bit cast of a vector element
        // lvalue reads the vector's first element
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned e = t[j];
          v[4 * k + j] = T(__builtin_bit_cast(float, e));
        }
      }
    }
  } else {
    const unsigned ps = unsigned(PS) * unsigned(sizeof(Q));
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const unsigned vm = DIR == 2 ? vo + m * ps : vo, sm = DIR == 2 ? 0u : m * ps;
      if constexpr (sizeof(Q) == 8)
        v[m] = T(__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vm, sm, 0)));
      else
        v[m] = T(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vm, sm, 0)));
    }
  }
}

template <int DIR, typename T, typename Q>
__device__ __forceinline__ void st(Q *__restrict__ p, long base, long PS, const T (&v)[PPT]) {
  const __amdgpu_buffer_rsrc_t r = rsrc(p);
  const unsigned vo = unsigned(base) * unsigned(sizeof(Q));
  if (DIR == 0) {
#pragma unroll
    for (int k = 0; k < PPT * int(sizeof(Q)) / 16; ++k) {
      bu4 t;
      if constexpr (sizeof(Q) == 8) {
        const bu2 a = __builtin_bit_cast(bu2, double(v[2 * k])), b = __builtin_bit_cast(bu2, double(v[2 * k + 1]));
        t = bu4{a[0], a[1], b[0], b[1]};
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = __builtin_bit_cast(unsigned, float(v[4 * k + j]));
      }
      __builtin_amdgcn_raw_buffer_store_b128(t, r, vo, k * 16, 0);
    }
  } else {
    const unsigned ps = unsigned(PS) * unsigned(sizeof(Q));
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const unsigned vm = DIR == 2 ? vo + m * ps : vo, sm = DIR == 2 ? 0u : m * ps;
      if constexpr (sizeof(Q) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bu2, double(v[m])), r, vm, sm, 0);
      else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, float(v[m])), r, vm, sm, 0);
    }
  }
}

// Work-group size limit of the Cartesian viscous sweeps: FP32 tiles take up
// to 32 y/z lines so that rows of floats are 128 bytes, as FP64 rows of 16.
template <typename T>
constexpr int sweep_nt() { return sizeof(T) == 4 ? NT_SWEEP32 : NT; }

// ----------------------------------------------------------------------------
// S1 / S2 / S3: first-derivative sweeps along DIR.
//   qs  : stage input state (5 fields, type Q)
//   rhs : 5 accumulator fields (S1 writes, S2/S3 read-modify-write)
//   G   : velocity-gradient fields (layout above)
// The solves are ordered so that each state field is live only while needed:
// the two transverse momentum components first (velocity gradient and
// momentum flux of each), then the DIR component, mass, pressure gradient,
// energy and heat flux with rho, rho u_DIR and p only.  Read-modify-write
// operands are fetched just before the line solve whose result they meet, so
// their latency hides behind the solve (the engine's barriers do not wait on
// global loads).
// GT: deformed grid (ns_rho_rhou_p_grid_transform.f90:31-330).  Every
// derivative along DIR is scaled by the line's metric entry M(DIR,DIR)
// (physical gradients, contravariant fluxes and the projected stresses all
// reduce to that for a diagonal metric); the heat flux takes the temperature
// gradient with h and lambda J inside, the outer derivative without h.
// EV: volume penalisation (ns_rho_rhou_p_ev_phi.f90:33-291): every flux
// carries phi inside its derivative and its divergence is multiplied by
// the stored 1/phi, distributed over the directions (the pressure
// gradient and u.grad(p) are not); both heat-flux derivatives take no h.
// What the fused z kernel's viscous part takes from its first-derivative
// part (k_sweep12_z, Cartesian form): the block's velocity gradients that the
// z sweep completes (s13 = u_z + w_x, s23 = v_z + w_y, w_z) and the
// viscosity in registers; the momentum RHS of the block in LDS (the thread's
// own slots) instead of a store and a reload through HBM.
template <typename T>
struct ZCarry {
  T s13[PPT], s23[PPT], wz[PPT], mu[PPT];
  T (*mom)[PPT][NT];  // LDS: [3][PPT][NT] momentum RHS (x, y, z)
};

template <typename T, typename Q, int DIR, bool PER, int FORM, bool FUSE = false>
__device__ __forceinline__ void sweep1_body(const ThreadPos &tp, const Q *__restrict__ qs, T *__restrict__ rhs,
                                            T *__restrict__ G, const LineOp<T> &op, const SweepGeom &g,
                                            const Phys &ph, long N, const Metric<T> &mt, const Penal<T> &pe,
                                            LineLds<T> &lds, ZCarry<T> *zc = nullptr) {
  constexpr bool GT = FORM == RHS_GT, EV = FORM == RHS_EV;
  static_assert(!FUSE || (DIR == 2 && FORM == RHS_CART), "the fused z sweep is Cartesian");
  const long B = tp.base, PS = g.PS;
  const T mline = GT ? mt.m[DIR][tp.lidx] : T(1);
  const T ih = GT ? T(ph.inv_h[DIR]) * mline : T(ph.inv_h[DIR]);
  // heat flux: inner (temperature) and outer derivative scales
  const T hin = GT ? ih : (EV ? T(1) : T(ph.heat_inv_h[DIR]));
  const T hout = GT ? mline : (EV ? T(1) : T(ph.heat_inv_h[DIR]));
  // penalisation: phi inside the fluxes, the stored 1/phi on their divergence
  T PH[PPT], IP[PPT];
  if (EV) {
    ld<DIR>(pe.phi, B, PS, PH);
    ld<DIR>(pe.inv_phi, B, PS, IP);
  }
  auto div = [&](int m, T v) -> T { return EV ? IP[m] * v : v; };
  // G slot of d(u_i)/dx_DIR and whether it completes a symmetric sum begun by
  // an earlier sweep: x writes u_x, v_x, w_x; y adds u_y to v_x and writes
  // v_y, w_y; z adds u_z to w_x, v_z to w_y and writes w_z
  auto gslot = [](int i) -> int {
    return DIR == 0 ? (i == 0 ? 0 : (i == 1 ? 3 : 4))
                    : (DIR == 1 ? (i == 0 ? 3 : (i == 1 ? 1 : 5)) : (i == 0 ? 4 : (i == 1 ? 5 : 2)));
  };
  auto gsum = [](int i) -> bool { return (DIR == 1 && i == 0) || (DIR == 2 && i < 2); };
  T irho[PPT], mom[PPT], p[PPT], f[PPT], x[PPT], r[PPT];
  ld<DIR>(qs, B, PS, f);
  ld<DIR>(qs + (1 + DIR) * N, B, PS, mom);
#pragma unroll
  for (int m = 0; m < PPT; ++m) irho[m] = T(1) / f[m];

  // transverse components i != DIR: d(u_i)/dx_DIR (ns_rho_rhou_p.f90:102-112)
  // and the momentum flux -d/dx_DIR (rho u_DIR u_i) (:138-172)
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int i = k == 0 ? (DIR == 0 ? 1 : 0) : (DIR == 2 ? 1 : 2);
    T ui[PPT], gs[PPT];
    asm volatile("" ::: "memory");
    ld<DIR>(qs + (1 + i) * N, B, PS, ui);
    if (gsum(i)) ld<DIR>(G + gslot(i) * N, B, PS, gs);
    issue_loads();
#pragma unroll
    for (int m = 0; m < PPT; ++m) ui[m] = ui[m] * irho[m];
    line_solve<T, PER>(op, lds, tp, ui, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      x[m] = x[m] * ih;
      if (gsum(i)) x[m] = x[m] + gs[m];
    }
    st<DIR>(G + gslot(i) * N, B, PS, x);
    if constexpr (FUSE) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) (k == 0 ? zc->s13 : zc->s23)[m] = x[m];
    }
    if (DIR != 0) ld<DIR>(rhs + (1 + i) * N, B, PS, r);
    if (k == 1) ld<DIR>(qs + 4 * N, B, PS, p);  // in flight for the pressure terms
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = EV ? (PH[m] * mom[m]) * ui[m] : mom[m] * ui[m];
    line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = DIR == 0 ? div(m, -(x[m] * ih)) : r[m] + div(m, -(x[m] * ih));
    if constexpr (FUSE) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) zc->mom[i][m][threadIdx.x] = r[m];
    } else {
      st<DIR>(rhs + (1 + i) * N, B, PS, r);
    }
  }

  // d(u_DIR)/dx_DIR
  T ud[PPT];
#pragma unroll
  for (int m = 0; m < PPT; ++m) ud[m] = mom[m] * irho[m];
  line_solve<T, PER>(op, lds, tp, ud, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) x[m] = x[m] * ih;
  st<DIR>(G + DIR * N, B, PS, x);
  if constexpr (FUSE) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) zc->wz[m] = x[m];
  }

  // mass: -d/dx_DIR (rho u_DIR) (:131-135)
  if (DIR != 0) ld<DIR>(rhs, B, PS, r);
  if (EV) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = PH[m] * mom[m];
    line_solve<T, PER>(op, lds, tp, f, x);
  } else {
    line_solve<T, PER>(op, lds, tp, mom, x);
  }
#pragma unroll
  for (int m = 0; m < PPT; ++m) r[m] = DIR == 0 ? div(m, -(x[m] * ih)) : r[m] + div(m, -(x[m] * ih));
  st<DIR>(rhs, B, PS, r);

  // pressure gradient p_DIR
  T pd[PPT];
  line_solve<T, PER>(op, lds, tp, p, pd);
#pragma unroll
  for (int m = 0; m < PPT; ++m) pd[m] = pd[m] * ih;

  // momentum DIR: -d/dx_DIR (rho u_DIR u_DIR) - p_DIR
  if (DIR != 0) ld<DIR>(rhs + (1 + DIR) * N, B, PS, r);
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = EV ? (PH[m] * mom[m]) * ud[m] : mom[m] * ud[m];
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) {
    r[m] = DIR == 0 ? div(m, -(x[m] * ih)) : r[m] + div(m, -(x[m] * ih));
    r[m] = r[m] - pd[m];
  }
  if constexpr (FUSE) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) zc->mom[DIR][m][threadIdx.x] = r[m];
  } else {
    st<DIR>(rhs + (1 + DIR) * N, B, PS, r);
  }

  // pressure-equation bracket: gamma/(gamma-1) (-d(p u_DIR)) + u_DIR p_DIR (:180-187)
  T e[PPT];
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = EV ? (PH[m] * p[m]) * ud[m] : p[m] * ud[m];
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) e[m] = div(m, T(ph.gt) * (-(x[m] * ih))) + ud[m] * pd[m];

  // heat flux d/dx_DIR (lambda d/dx_DIR T) (:189-199); the x sweep evaluates
  // the viscosity law and leaves mu in G6 for the later sweeps of the stage
  T mu[PPT], Jv[PPT];
  if (DIR != 0) {
    ld<DIR>(rhs + 4 * N, B, PS, r);
    ld<DIR>(G + 6 * N, B, PS, mu);
  }
  if constexpr (FUSE) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) zc->mu[m] = mu[m];
  }
  if (GT) ld<DIR>(mt.J, B, PS, Jv);
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = temperature(ph, irho[m], p[m]);
  if (DIR == 0) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) mu[m] = viscosity(ph, f[m]);
    st<DIR>(G + 6 * N, B, PS, mu);
  }
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) {
    const T lam = GT ? lambda_of(ph, mu[m]) * Jv[m]
                     : (EV ? PH[m] * lambda_of(ph, mu[m]) : lambda_of(ph, mu[m]));
    f[m] = lam * (x[m] * hin);
  }
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) {
    e[m] = e[m] + div(m, x[m] * hout);
    if (DIR != 0) e[m] = r[m] + e[m];
  }
  st<DIR>(rhs + 4 * N, B, PS, e);
}

template <typename T, typename Q, int DIR, bool PER, int FORM>
__global__ __launch_bounds__(NT) void k_sweep1(const Q *__restrict__ qs, T *__restrict__ rhs,
                                                          T *__restrict__ G, LineOp<T> op, SweepGeom g,
                                                          Phys ph, long N, Metric<T> mt, Penal<T> pe) {
  __shared__ LineLds<T> lds;
  sweep1_body<T, Q, DIR, PER, FORM>(thread_pos<DIR>(g), qs, rhs, G, op, g, ph, N, mt, pe, lds);
}

// RK4 stage update of variable v from its complete RHS k (rk_o4_e.f90:45-50,
// and :90-95 for the adjoint with dt -> -dt): stage inputs q0 + dt/2 k1,
// q0 + dt/2 k2, q0 + dt k3, and q1 = q0 + dt/6 (k1 + 2 k2 + 2 k3 + k4).
// The accumulator has the arithmetic type T.  FP64: it carries the partial
// new state q0 + dt/6 k1 + dt/3 k2 + dt/3 k3, so the last stage reads it
// alone (not q0 as well).  FP32-mixed: it carries the increment
// dt/6 k1 + dt/3 k2 + dt/3 k3 alone, in FP32 (half the bytes of an FP64
// accumulator in three stages out of four), and the last stage adds q0 in
// FP64; the increment is a step's change of the state, so its FP32 rounding
// is far below that of the FP32 stage inputs.  Either sum is associated
// differently from the reference's, a rounding-level change.
// mode 4 writes k into rhs instead (a single RHS evaluation).
template <int DIR, typename T>
__device__ __forceinline__ void rk_new_state(int v, const T (&k)[PPT], long B, long PS, long N,
                                             const double *__restrict__ qn, const T *__restrict__ acc,
                                             double dt, double (&q1)[PPT]) {
  if constexpr (sizeof(T) == 8) {
    ld<DIR>(acc + v * N, B, PS, q1);
#pragma unroll
    for (int m = 0; m < PPT; ++m) q1[m] = q1[m] + dt / 6.0 * double(k[m]);
  } else {
    T a[PPT];
    ld<DIR>(qn + v * N, B, PS, q1);
    ld<DIR>(acc + v * N, B, PS, a);
#pragma unroll
    for (int m = 0; m < PPT; ++m) q1[m] = q1[m] + (double(a[m]) + dt / 6.0 * double(k[m]));
  }
}

template <int DIR, typename T>
__device__ __forceinline__ void rk_update(int mode, int v, const T (&k)[PPT], long B, long PS,
                                          long N, T *rhs, double *__restrict__ qn,
                                          T *__restrict__ acc, T *qs_out, double dt) {
  if (mode == 4) {
    st<DIR>(rhs + v * N, B, PS, k);
    return;
  }
  double a[PPT];
  if (mode == 3) {
    rk_new_state<DIR, T>(v, k, B, PS, N, qn, acc, dt, a);
    st<DIR>(qn + v * N, B, PS, a);
    return;
  }
  double q0[PPT];
  ld<DIR>(qn + v * N, B, PS, q0);
  // the accumulator's value after this stage
  const double base = sizeof(T) == 8 ? 1.0 : 0.0;  // FP32: the increment alone
  if (mode == 0) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) a[m] = base * q0[m] + dt / 6.0 * double(k[m]);
  } else {
    ld<DIR>(acc + v * N, B, PS, a);
#pragma unroll
    for (int m = 0; m < PPT; ++m) a[m] = a[m] + dt / 3.0 * double(k[m]);
  }
  st<DIR>(acc + v * N, B, PS, a);
  const double c_stage = mode == 2 ? dt : dt * 0.5;
#pragma unroll
  for (int m = 0; m < PPT; ++m) a[m] = q0[m] + c_stage * double(k[m]);
  st<DIR>(qs_out + v * N, B, PS, a);
}

// ----------------------------------------------------------------------------
// S4 (z) / S5 (y) / S6 (x): viscous fluxes d/dx_d tau_{i d}; the x sweep also
// adds tau:grad(u) (ns_rho_rhou_p.f90:202-205) and finishes the stage.
// mode 0..3: RK4 stage update; mode 4: write the complete RHS into rhs.
// FF (x sweep of the last stage, mode 3, only): also apply filter_i1 to
// the new state before storing it (the x filter of the step loop, fused
// because the x sweep holds complete x-lines); fop is that filter.  A
// separate instance, so that the other stages do not carry the filter
// solve's registers.
// The viscous sweeps read only G and rhs, so the next stage input (qs_out)
// may overwrite the current one in place.
// GT: deformed grid: d/dxi_d (M(d,d) tau_id) = M(d,d) d/dxi_d tau_id, and the
// last sweep multiplies every equation by X_jacobian and the dissipation once
// more (ns_rho_rhou_p_grid_transform.f90:190-276).
// EV: d/dx_d (phi tau_id), times the stored 1/phi (ns_rho_rhou_p_ev_phi.f90:161-201).
template <typename T, int DIR, bool PER, int FORM, bool FF>
__device__ __forceinline__ void sweep2_body(const ThreadPos &tp, T *__restrict__ rhs, const T *__restrict__ G,
                                            const LineOp<T> &op, const LineOp<T> &fop, const SweepGeom &g,
                                            const Phys &ph, long N, int mode, double *__restrict__ qn,
                                            T *__restrict__ acc, T *__restrict__ qs_out, double dt,
                                            const Metric<T> &mt, const T *__restrict__ force,
                                            const Penal<T> &pe) {
  constexpr bool GT = FORM == RHS_GT, EV = FORM == RHS_EV;
  constexpr bool LAST = DIR == 0;
  __shared__ LineLds<T, sweep_nt<T>()> lds;
  const long B = tp.base, PS = g.PS;
  const T ih = GT ? T(ph.inv_h[DIR]) * mt.m[DIR][tp.lidx] : T(ph.inv_h[DIR]);
  T t1[PPT], t2[PPT], t3[PPT], diss[PPT];
  {
    T mu[PPT], ux[PPT], vy[PPT], wz[PPT], sa[PPT], sb[PPT];
    ld<DIR>(G + 6 * N, B, PS, mu);
    ld<DIR>(G + 0 * N, B, PS, ux);
    ld<DIR>(G + 1 * N, B, PS, vy);
    ld<DIR>(G + 2 * N, B, PS, wz);
    // x: s12 = u_y + v_x, s13 = u_z + w_x;  y: s12, s23 = v_z + w_y;  z: s13, s23
    ld<DIR>(G + (DIR == 2 ? 4 : 3) * N, B, PS, sa);
    ld<DIR>(G + (DIR == 0 ? 4 : 5) * N, B, PS, sb);
    issue_loads();
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T md = -2.0 / 3.0 * mu[m] * (ux[m] + vy[m] + wz[m]);
      const T t11 = md + 2 * mu[m] * ux[m], t22 = md + 2 * mu[m] * vy[m], t33 = md + 2 * mu[m] * wz[m];
      if (DIR == 0) {
        t1[m] = t11;            // tau_11
        t2[m] = mu[m] * sa[m];  // tau_21
        t3[m] = mu[m] * sb[m];  // tau_31
        // tau_ij du_i/dx_j = tau_11 u_x + tau_22 v_y + tau_33 w_z + mu (s12^2 + s13^2 + s23^2)
        diss[m] = t11 * ux[m] + t22 * vy[m] + t33 * wz[m] + t2[m] * sa[m] + t3[m] * sb[m];
      } else if (DIR == 1) {
        t1[m] = mu[m] * sa[m];  // tau_12
        t2[m] = t22;            // tau_22
        t3[m] = mu[m] * sb[m];  // tau_32
      } else {
        t1[m] = mu[m] * sa[m];  // tau_13
        t2[m] = mu[m] * sb[m];  // tau_23
        t3[m] = t33;            // tau_33
      }
    }
    // the stresses now, so that mu and the gradients are not kept across
    // the solves below
    materialize(t1);
    materialize(t2);
    materialize(t3);
    if (LAST) {
      ld<DIR>(G + 5 * N, B, PS, sa);
#pragma unroll
      for (int m = 0; m < PPT; ++m) diss[m] = diss[m] + (mu[m] * sa[m]) * sa[m];
      materialize(diss);
    }
  }
  T Jv[PPT];
  if (LAST && GT) ld<DIR>(mt.J, B, PS, Jv);
  // the last sweep's complete RHS of variable v: RK4 update, and on the
  // final stage optionally the x filter of the new state
  auto finish = [&](int v, T(&k)[PPT]) {
    if constexpr (FF) {
      double a[PPT];
      rk_new_state<DIR, T>(v, k, B, PS, N, qn, acc, dt, a);
      T f[PPT], x[PPT];
#pragma unroll
      for (int m = 0; m < PPT; ++m) f[m] = T(a[m]);
      line_solve<T, PER>(fop, lds, tp, f, x);
      st<DIR>(qn + v * N, B, PS, x);
    } else {
      rk_update<DIR, T>(mode, v, k, B, PS, N, rhs, qn, acc, qs_out, dt);
    }
  };
  // each solve's result is consumed at once (z/y: read-modify-write of the
  // momentum accumulators; x: the complete RHS of that variable and its RK4
  // update), so only the remaining stress fields stay live across a solve;
  // the fences keep that traffic below the solve it depends on
  // force (apply_force, ns_rho_rhou_p.f90:212): an additive term of the
  // complete RHS, here the learned closure; null when there is none
  auto add_force = [&](int v, T(&k)[PPT]) {
    if (LAST && force) {
      T fv[PPT];
      ld<DIR>(force + v * N, B, PS, fv);
#pragma unroll
      for (int m = 0; m < PPT; ++m) k[m] = k[m] + fv[m];
    }
  };
  if (LAST && (mode != 4 || GT || force)) {
    // k_1 = rhs_1 (times J) is complete already: its stage update goes first
    T k[PPT];
    ld<DIR>(rhs, B, PS, k);
    if (GT) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) k[m] = Jv[m] * k[m];
    }
    add_force(0, k);
    finish(0, k);
  }
  if (LAST) {
    // pressure: (gamma - 1) (bracket + tau:grad(u)) (ns_rho_rhou_p.f90:200-209),
    // complete before the momentum solves: finishing it here frees the
    // dissipation's registers for them
    asm volatile("" ::: "memory");
    T k[PPT];
    ld<DIR>(rhs + 4 * N, B, PS, k);
#pragma unroll
    for (int m = 0; m < PPT; ++m)
      k[m] = GT ? (Jv[m] * T(ph.gm1)) * (k[m] + Jv[m] * diss[m]) : T(ph.gm1) * (k[m] + diss[m]);
    add_force(4, k);
    finish(4, k);
    asm volatile("" ::: "memory");
  }
  T IP[PPT];
  if (EV) {  // phi tau_id inside the derivative, 1/phi on it
    T PH[PPT];
    ld<DIR>(pe.phi, B, PS, PH);
    ld<DIR>(pe.inv_phi, B, PS, IP);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      t1[m] = PH[m] * t1[m];
      t2[m] = PH[m] * t2[m];
      t3[m] = PH[m] * t3[m];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T x[PPT];
    line_solve<T, PER>(op, lds, tp, i == 0 ? t1 : (i == 1 ? t2 : t3), x);
    asm volatile("" ::: "memory");
    T k[PPT];
    ld<DIR>(rhs + (1 + i) * N, B, PS, k);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      k[m] = k[m] + (EV ? IP[m] * (x[m] * ih) : x[m] * ih);
      if (LAST && GT) k[m] = Jv[m] * k[m];
    }
    if (!LAST) {
      st<DIR>(rhs + (1 + i) * N, B, PS, k);
    } else {
      add_force(1 + i, k);
      finish(1 + i, k);
    }
  }
}

template <typename T, int DIR, bool PER, int FORM, bool FF>
__global__ __launch_bounds__(sweep_nt<T>()) void k_sweep2(T *__restrict__ rhs, const T *__restrict__ G,
                                                          LineOp<T> op, LineOp<T> fop, SweepGeom g, Phys ph,
                                                          long N, int mode, double *__restrict__ qn,
                                                          T *__restrict__ acc, T *__restrict__ qs_out,
                                                          double dt, Metric<T> mt, const T *__restrict__ force,
                                                          Penal<T> pe) {
  sweep2_body<T, DIR, PER, FORM, FF>(thread_pos<DIR>(g), rhs, G, op, fop, g, ph, N, mode, qn, acc, qs_out, dt, mt,
                                     force, pe);
}

// S3 + S4: the z first-derivative sweep and the z viscous sweep as one
// kernel over the same z-line tiles.  The z sweep is the last
// first-derivative sweep of a stage, so once a thread has completed the
// velocity gradients of its block (w_z, u_z + w_x, v_z + w_y) the stresses
// of the block are known and the viscous z derivatives can follow at once.
// Cartesian form: the gradients and the viscosity stay in registers and the
// momentum RHS in LDS between the two parts (ZCarry), so the viscous part
// reads only u_x and v_y and writes the momentum RHS once; the deformed-grid
// and penalised forms run the two sweep bodies one after the other.
template <typename T, bool PER>
__device__ __forceinline__ void viscous_z_carry(const ThreadPos &tp, T *__restrict__ rhs, const T *__restrict__ G,
                                                const LineOp<T> &op, const SweepGeom &g, const Phys &ph, long N,
                                                LineLds<T> &lds, const ZCarry<T> &zc) {
  const long B = tp.base, PS = g.PS;
  const T ih = T(ph.inv_h[2]);
  T t1[PPT], t2[PPT], t3[PPT];
  {
    T ux[PPT], vy[PPT];
    ld<2>(G + 0 * N, B, PS, ux);
    ld<2>(G + 1 * N, B, PS, vy);
    issue_loads();
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T mu = zc.mu[m], wz = zc.wz[m];
      const T md = -2.0 / 3.0 * mu * (ux[m] + vy[m] + wz);
      t1[m] = mu * zc.s13[m];  // tau_13
      t2[m] = mu * zc.s23[m];  // tau_23
      t3[m] = md + 2 * mu * wz;  // tau_33
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T x[PPT], k[PPT];
    line_solve<T, PER>(op, lds, tp, i == 0 ? t1 : (i == 1 ? t2 : t3), x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) k[m] = zc.mom[i][m][threadIdx.x] + x[m] * ih;
    st<2>(rhs + (1 + i) * N, B, PS, k);
  }
}

template <typename T, typename Q, bool PER, int FORM>
__global__ __launch_bounds__(NT) void k_sweep12_z(const Q *__restrict__ qs, T *__restrict__ rhs,
                                                   T *__restrict__ G, LineOp<T> op, SweepGeom g, Phys ph,
                                                   long N, Metric<T> mt, Penal<T> pe) {
  const ThreadPos tp = thread_pos<2>(g);
  __shared__ LineLds<T> lds;
  if constexpr (FORM == RHS_CART) {
    __shared__ T mom[3][PPT][NT];
    ZCarry<T> zc;
    zc.mom = mom;
    sweep1_body<T, Q, 2, PER, FORM, true>(tp, qs, rhs, G, op, g, ph, N, mt, pe, lds, &zc);
    viscous_z_carry<T, PER>(tp, rhs, G, op, g, ph, N, lds, zc);
  } else {
    sweep1_body<T, Q, 2, PER, FORM>(tp, qs, rhs, G, op, g, ph, N, mt, pe, lds);
    asm volatile("" ::: "memory");
    sweep2_body<T, 2, PER, FORM, false>(tp, rhs, G, op, op, g, ph, N, 0, nullptr, nullptr, nullptr, 0.0, mt,
                                        nullptr, pe);
  }
}


// ----------------------------------------------------------------------------
// Single derivative (Dx1/2/3) of one field: out = (B^-1 A in) / h.
template <typename T, int DIR, bool PER>
__global__ __launch_bounds__(NT) void k_diff(const T *__restrict__ in, T *__restrict__ out,
                                             LineOp<T> op, SweepGeom g, double inv_h) {
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos<DIR>(g);
  T f[PPT], x[PPT];
  ld<DIR>(in, tp.base, g.PS, f);
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) x[m] = x[m] * T(inv_h);
  st<DIR>(out, tp.base, g.PS, x);
}

// Filter of nv fields in place (filter_4d): q = B_f^-1 A_f q per field.
template <typename T>
constexpr int filter_nt() { return sizeof(T) == 4 ? NT_FILTER32 : NT; }

// src: the fields to filter (FP64, or the compute precision); q64: where
// the filtered FP64 fields go (null: nowhere); copy: the filtered fields in
// the compute precision (null: nowhere).  src may be q64 or copy (in place:
// every thread loads its block before the line solve that the stores follow).
template <typename T, int DIR, bool PER, typename S>
__device__ __forceinline__ void filter_body(const ThreadPos &tp, const S *src, double *q64, const LineOp<T> &op,
                                            const SweepGeom &g, long N, int nv, T *copy) {
  __shared__ LineLds<T, filter_nt<T>()> lds;
  double a[PPT];
  ld<DIR>(src, tp.base, g.PS, a);
  for (int v = 0; v < nv; ++v) {
    T f[PPT], x[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = T(a[m]);
    // the next field's loads fly while this one is filtered
    if (v + 1 < nv) ld<DIR>(src + (v + 1) * N, tp.base, g.PS, a);
    line_solve<T, PER>(op, lds, tp, f, x);
    if (q64) st<DIR>(q64 + v * N, tp.base, g.PS, x);
    // the filtered state in the compute precision, when the next RK stage
    // or filter pass reads it (it equals the FP64 value converted back)
    if (copy) st<DIR>(copy + v * N, tp.base, g.PS, x);
  }
}

template <typename T, int DIR, bool PER, typename S>
__global__ __launch_bounds__(filter_nt<T>()) void k_filter(const S *src, double *q64, LineOp<T> op, SweepGeom g,
                                                           long N, int nv, T *copy) {
  filter_body<T, DIR, PER, S>(thread_pos<DIR>(g), src, q64, op, g, N, nv, copy);
}

}  // namespace cruna
