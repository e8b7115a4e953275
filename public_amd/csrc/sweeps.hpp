// Fused direct-RHS sweeps, filter sweep and single-derivative kernel.
//
// One RK stage of right_hand_side_direct_3d (src/equations/ns_rho_rhou_p.f90:31-253)
// is evaluated in five line sweeps; each sweep applies every compact
// derivative that runs along its direction, so the 42 Dx calls of the
// reference become 42 line solves on register-resident blocks:
//   S1 x : u_x v_x w_x, d/dx of (rho u), (rho u u_i), p, (p u), T and lambda T_x
//   S2 y : the same along y                                  (accumulates rhs)
//   S3 z : the same along z, plus d/dz tau_{iz} and tau:grad(u)
//   S4 x : d/dx tau_{ix}
//   S5 y : d/dy tau_{iy}, then the RK4 stage update (src/time/rk_o4_e.f90:45-50)
// The pressure equation is accumulated as the bracket of ns_rho_rhou_p.f90:200-206
// and multiplied by (gamma - 1) once, as the reference does.
//
// Each sweep is written as a sequence of phases that keep only the operands of
// the next line solve live; inputs used by several phases are re-read (the
// tile was just fetched, so those reads are served by L2).
//
// Field layout in HBM: every field is f(n1,n2,n3) Fortran order; a state holds
// 5 fields at a stride of N = n1*n2*n3 elements; G holds the 9 velocity
// gradients, G[3*i + d] = d u_i / d x_d.
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
  double inv_h[3];       // 1/dx_d
  double heat_inv_h[3];  // 1/dx_d, or 1 for the reference's unit-spacing heat flux
};

// give_mu, ns_rho_rhou_p_trafos.f90:119-142: mu0*(T/T0)**1.5*((T0+S1)/(T+S1)),
// or the power law mu0*(T/T0)**exp of the case-level override.
template <typename T>
__device__ __forceinline__ T viscosity(const Phys &ph, T Tk) {
  const T r = Tk / T(ph.T0);
  if (ph.visc_law == 1) return T(ph.mu0) * pow(r, T(ph.visc_exp));
  const T S1 = T(110.4);
  return T(ph.mu0) * (r * sqrt(r)) * ((T(ph.T0) + S1) / (Tk + S1));
}

// give_T (:57) and give_lambda (:169): T = (W/R) p / rho, lambda = mu cp / Pr
template <typename T>
__device__ __forceinline__ T temperature(const Phys &ph, T rho, T p) {
  return T(ph.w_inv_r) * p / rho;
}
template <typename T>
__device__ __forceinline__ T lambda_of(const Phys &ph, T mu) {
  return mu * T(ph.cp) / T(ph.Pr);
}

template <typename T, typename Q>
__device__ __forceinline__ void ld(const Q *__restrict__ p, long base, long PS, T (&v)[PPT]) {
#pragma unroll
  for (int m = 0; m < PPT; ++m) v[m] = T(p[base + m * PS]);
}

template <typename T, typename Q>
__device__ __forceinline__ void st(Q *__restrict__ p, long base, long PS, const T (&v)[PPT]) {
#pragma unroll
  for (int m = 0; m < PPT; ++m) p[base + m * PS] = Q(v[m]);
}

// ----------------------------------------------------------------------------
// S1 / S2 / S3: first-derivative sweeps along DIR.
//   qs  : stage input state (5 fields, type Q)
//   rhs : 5 accumulator fields (S1 writes, S2/S3 read-modify-write)
//   G   : velocity gradients; S3 additionally reads the x/y components
template <typename T, typename Q, int DIR, bool PER>
__global__ __launch_bounds__(NT) void k_sweep1(const Q *__restrict__ qs, T *__restrict__ rhs,
                                               T *__restrict__ G, LineOp<T> op, SweepGeom g,
                                               Phys ph, long N) {
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos(g);
  const long B = tp.base, PS = g.PS;
  const T ih = T(ph.inv_h[DIR]), hih = T(ph.heat_inv_h[DIR]);
  const Q *rho_ = qs, *p_ = qs + 4 * N, *md_ = qs + (1 + DIR) * N;
  T a[PPT], f[PPT], x[PPT], r[PPT];

  // velocity gradients u_d, v_d, w_d (ns_rho_rhou_p.f90:102-112)
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ld(rho_, B, PS, a);
    ld(qs + (1 + c) * N, B, PS, f);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = f[m] / a[m];
    line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) x[m] = x[m] * ih;
    st(G + (3 * c + DIR) * N, B, PS, x);
  }

  // mass: -d/dx_d (rho u_d) (:131-135)
  ld(md_, B, PS, f);
  line_solve<T, PER>(op, lds, tp, f, x);
  if (DIR == 0) {
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = -(x[m] * ih);
  } else {
    ld(rhs, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = r[m] - x[m] * ih;
  }
  st(rhs, B, PS, r);

  // pressure gradient p_d
  T pd[PPT];
  ld(p_, B, PS, f);
  line_solve<T, PER>(op, lds, tp, f, pd);
#pragma unroll
  for (int m = 0; m < PPT; ++m) pd[m] = pd[m] * ih;

  // momentum: -d/dx_d (rho u_d u_i) - p_i (:138-172)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ld(rho_, B, PS, a);
    ld(qs + (1 + i) * N, B, PS, f);
    ld(md_, B, PS, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = x[m] * (f[m] / a[m]);
    line_solve<T, PER>(op, lds, tp, f, x);
    if (DIR == 0) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) r[m] = -(x[m] * ih);
    } else {
      ld(rhs + (1 + i) * N, B, PS, r);
#pragma unroll
      for (int m = 0; m < PPT; ++m) r[m] = r[m] - x[m] * ih;
    }
    if (i == DIR) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) r[m] = r[m] - pd[m];
    }
    st(rhs + (1 + i) * N, B, PS, r);
  }

  // pressure-equation bracket: gamma/(gamma-1) (-d(p u_d)) + u_d p_d (:180-187)
  T e[PPT];
  ld(rho_, B, PS, a);
  ld(md_, B, PS, x);
  ld(p_, B, PS, f);
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = f[m] * (x[m] / a[m]);
  line_solve<T, PER>(op, lds, tp, f, x);
  ld(rho_, B, PS, a);
  ld(md_, B, PS, f);
#pragma unroll
  for (int m = 0; m < PPT; ++m) e[m] = T(ph.gt) * (-(x[m] * ih)) + (f[m] / a[m]) * pd[m];

  // heat flux d/dx_d (lambda d/dx_d T) (:189-199)
  ld(rho_, B, PS, a);
  ld(p_, B, PS, f);
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = temperature(ph, a[m], f[m]);
  line_solve<T, PER>(op, lds, tp, f, x);
  ld(rho_, B, PS, a);
  ld(p_, B, PS, f);
#pragma unroll
  for (int m = 0; m < PPT; ++m)
    f[m] = lambda_of(ph, viscosity(ph, temperature(ph, a[m], f[m]))) * (x[m] * hih);
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) e[m] = e[m] + x[m] * hih;

  if (DIR == 2) {
    // viscous fluxes d/dz tau_{iz} (:158-177) and tau:grad(u) (:202-205)
    T mu[PPT];
    ld(rho_, B, PS, a);
    ld(p_, B, PS, f);
#pragma unroll
    for (int m = 0; m < PPT; ++m) mu[m] = viscosity(ph, temperature(ph, a[m], f[m]));
    // tau_13 = mu (u_z + w_x)
    ld(G + 2 * N, B, PS, a);
    ld(G + 6 * N, B, PS, f);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = mu[m] * (a[m] + f[m]);
    line_solve<T, PER>(op, lds, tp, f, x);
    ld(rhs + 1 * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = r[m] + x[m] * ih;
    st(rhs + 1 * N, B, PS, r);
    // tau_23 = mu (v_z + w_y)
    ld(G + 5 * N, B, PS, a);
    ld(G + 7 * N, B, PS, f);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = mu[m] * (a[m] + f[m]);
    line_solve<T, PER>(op, lds, tp, f, x);
    ld(rhs + 2 * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = r[m] + x[m] * ih;
    st(rhs + 2 * N, B, PS, r);
    // tau_33 = -2/3 mu div(u) + 2 mu w_z
    ld(G + 0 * N, B, PS, a);
    ld(G + 4 * N, B, PS, f);
    ld(G + 8 * N, B, PS, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T md = -2.0 / 3.0 * mu[m] * (a[m] + f[m] + x[m]);
      f[m] = md + 2 * mu[m] * x[m];
    }
    line_solve<T, PER>(op, lds, tp, f, x);
    ld(rhs + 3 * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = r[m] + x[m] * ih;
    st(rhs + 3 * N, B, PS, r);
    // dissipation tau_ij du_i/dx_j, summed in the reference's order
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const long o = B + m * PS;
      const T ux = G[o], uy = G[N + o], uz = G[2 * N + o];
      const T vx = G[3 * N + o], vy = G[4 * N + o], vz = G[5 * N + o];
      const T wx = G[6 * N + o], wy = G[7 * N + o], wz = G[8 * N + o];
      const T md = -2.0 / 3.0 * mu[m] * (ux + vy + wz);
      const T t11 = md + 2 * mu[m] * ux, t22 = md + 2 * mu[m] * vy, t33 = md + 2 * mu[m] * wz;
      const T t12 = mu[m] * (uy + vx), t13 = mu[m] * (uz + wx), t23 = mu[m] * (vz + wy);
      e[m] = e[m] + t11 * ux + t12 * uy + t13 * uz + t12 * vx + t22 * vy + t23 * vz + t13 * wx +
             t23 * wy + t33 * wz;
    }
  }

  if (DIR != 0) {
    ld(rhs + 4 * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) e[m] = r[m] + e[m];
  }
  st(rhs + 4 * N, B, PS, e);
}

// ----------------------------------------------------------------------------
// S4 (x) / S5 (y): viscous fluxes d/dx_d tau_{i d}; S5 then finishes the stage.
// mode 0..3: RK4 stage update; mode 4: write the complete RHS into rhs.
template <typename T, typename Q, int DIR, bool PER>
__global__ __launch_bounds__(NT) void k_sweep2(const Q *__restrict__ qs, T *__restrict__ rhs,
                                               const T *__restrict__ G, LineOp<T> op, SweepGeom g,
                                               Phys ph, long N, int mode, double *__restrict__ qn,
                                               double *__restrict__ acc, T *__restrict__ qs_out,
                                               double dt) {
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos(g);
  const long B = tp.base, PS = g.PS;
  const T ih = T(ph.inv_h[DIR]);
  T a[PPT], b[PPT], c[PPT], f[PPT];

  T mu[PPT];
  ld(qs, B, PS, a);
  ld(qs + 4 * N, B, PS, b);
#pragma unroll
  for (int m = 0; m < PPT; ++m) mu[m] = viscosity(ph, temperature(ph, a[m], b[m]));

  T x1[PPT], x2[PPT], x3[PPT];
  // first component: tau_11 (x) or tau_12 (y)
  if (DIR == 0) {
    ld(G + 0 * N, B, PS, a);
    ld(G + 4 * N, B, PS, b);
    ld(G + 8 * N, B, PS, c);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T md = -2.0 / 3.0 * mu[m] * (a[m] + b[m] + c[m]);
      f[m] = md + 2 * mu[m] * a[m];
    }
  } else {
    ld(G + 1 * N, B, PS, a);
    ld(G + 3 * N, B, PS, b);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = mu[m] * (a[m] + b[m]);
  }
  line_solve<T, PER>(op, lds, tp, f, x1);
  // second component: tau_21 (x) or tau_22 (y)
  if (DIR == 0) {
    ld(G + 1 * N, B, PS, a);
    ld(G + 3 * N, B, PS, b);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = mu[m] * (a[m] + b[m]);
  } else {
    ld(G + 0 * N, B, PS, a);
    ld(G + 4 * N, B, PS, b);
    ld(G + 8 * N, B, PS, c);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T md = -2.0 / 3.0 * mu[m] * (a[m] + b[m] + c[m]);
      f[m] = md + 2 * mu[m] * b[m];
    }
  }
  line_solve<T, PER>(op, lds, tp, f, x2);
  // third component: tau_31 = mu (u_z + w_x) (x) or tau_32 = mu (v_z + w_y) (y)
  ld(G + (DIR == 0 ? 2 : 5) * N, B, PS, a);
  ld(G + (DIR == 0 ? 6 : 7) * N, B, PS, b);
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = mu[m] * (a[m] + b[m]);
  line_solve<T, PER>(op, lds, tp, f, x3);

  // momentum components with the viscous flux added
#pragma unroll
  for (int m = 0; m < PPT; ++m) {
    x1[m] = x1[m] * ih;
    x2[m] = x2[m] * ih;
    x3[m] = x3[m] * ih;
  }
  if (DIR == 0) {
    ld(rhs + 1 * N, B, PS, a);
#pragma unroll
    for (int m = 0; m < PPT; ++m) a[m] = a[m] + x1[m];
    st(rhs + 1 * N, B, PS, a);
    ld(rhs + 2 * N, B, PS, a);
#pragma unroll
    for (int m = 0; m < PPT; ++m) a[m] = a[m] + x2[m];
    st(rhs + 2 * N, B, PS, a);
    ld(rhs + 3 * N, B, PS, a);
#pragma unroll
    for (int m = 0; m < PPT; ++m) a[m] = a[m] + x3[m];
    st(rhs + 3 * N, B, PS, a);
    return;
  }
  // complete RHS k_v and the RK4 update (rk_o4_e.f90:45-50): stage inputs
  // q0 + dth k1, q0 + dth k2, q0 + dt k3; q1 = q0 + dt/6 (k1 + 2 k2 + 2 k3 + k4)
  const double c_stage = mode == 2 ? dt : dt * 0.5;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    T k[PPT];
    ld(rhs + v * N, B, PS, k);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      if (v == 1) k[m] = k[m] + x1[m];
      if (v == 2) k[m] = k[m] + x2[m];
      if (v == 3) k[m] = k[m] + x3[m];
      if (v == 4) k[m] = T(ph.gm1) * k[m];
    }
    if (mode == 4) {
      if (v > 0) st(rhs + v * N, B, PS, k);
      continue;
    }
    double q0[PPT], s[PPT];
    ld(qn + v * N, B, PS, q0);
    if (mode == 3) {
      ld(acc + v * N, B, PS, s);
#pragma unroll
      for (int m = 0; m < PPT; ++m) q0[m] = q0[m] + dt / 6.0 * (s[m] + double(k[m]));
      st(qn + v * N, B, PS, q0);
    } else {
      if (mode == 0) {
#pragma unroll
        for (int m = 0; m < PPT; ++m) s[m] = double(k[m]);
      } else {
        ld(acc + v * N, B, PS, s);
#pragma unroll
        for (int m = 0; m < PPT; ++m) s[m] = s[m] + 2 * double(k[m]);
      }
      st(acc + v * N, B, PS, s);
#pragma unroll
      for (int m = 0; m < PPT; ++m) s[m] = q0[m] + c_stage * double(k[m]);
      st(qs_out + v * N, B, PS, s);
    }
  }
}

// ----------------------------------------------------------------------------
// Single derivative (Dx1/2/3) of one field: out = (B^-1 A in) / h.
template <typename T, bool PER>
__global__ __launch_bounds__(NT) void k_diff(const T *__restrict__ in, T *__restrict__ out,
                                             LineOp<T> op, SweepGeom g, double inv_h) {
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos(g);
  T f[PPT], x[PPT];
  ld(in, tp.base, g.PS, f);
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) x[m] = x[m] * T(inv_h);
  st(out, tp.base, g.PS, x);
}

// Filter of nv fields in place (filter_4d): q = B_f^-1 A_f q per field.
template <typename T, bool PER>
__global__ __launch_bounds__(NT) void k_filter(double *__restrict__ q, LineOp<T> op, SweepGeom g,
                                               long N, int nv) {
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos(g);
  for (int v = 0; v < nv; ++v) {
    T f[PPT], x[PPT];
    ld(q + v * N, tp.base, g.PS, f);
    line_solve<T, PER>(op, lds, tp, f, x);
    st(q + v * N, tp.base, g.PS, x);
  }
}

}  // namespace cruna
