// Direct RHS on a general curvilinear grid: ns_rho_rhou_p_grid_transform.f90:31-330
// with a full metric X_metric(n1,n2,n3,3,3) and Jacobian X_jacobian(n1,n2,n3)
// (grid_metric.f90:16-188).  Every derivative along xi_d contributes to all
// three physical components through the metric row M(d,1..3), so a sweep
// along d reads that row (3 fields) and accumulates:
//   G1 x/y/z : a_d(u), a_d(v), a_d(w) into the six physical velocity-gradient
//              slots, a_d(T) into the three temperature-gradient fields,
//              -a_d(rho u_hat_d), -a_d(rho u_i u_hat_d) - M(d,i) a_d(p) into
//              the mass and momentum accumulators and
//              gamma/(gamma-1) (-a_d(p u_hat_d)) + u_hat_d a_d(p) into the
//              pressure accumulator (u . grad p = sum_d u_hat_d a_d(p), with
//              u_hat_d = sum_j M(d,j) u_j)
//   G2 z/y/x : a_d(sum_j M(d,j) tau_ij) into momentum i, and the heat flux
//              sum_j M(d,j) D_d(lambda J T_j) (outer derivative without h)
//              into the pressure accumulator; the x sweep adds
//              J tau:grad(u), multiplies every equation by X_jacobian (the
//              pressure also by gamma - 1) and finishes the RK4 stage.
// Fields (compute precision T): G0 = u_x, G1 = v_y, G2 = w_z, G3 = u_y + v_x,
// G4 = u_z + w_x, G5 = v_z + w_y, G6 = mu(T) as in the Cartesian sweeps (the
// gradients here are the reference's "without Jacobian" ones), TG0..2 =
// T_x, T_y, T_z; metric M(k,j) at M + N (k + 3 j) and J, both per point.
#pragma once

#include "sweeps.hpp"

namespace cruna {

template <typename T>
struct CurvMetric {
  const T *M;  // X_metric(:,:,:,k,j) at M + N (k + 3 j), k, j = 0..2
  const T *J;  // X_jacobian
};

// EV: volume penalisation (ns_rho_rhou_p_ev_phi_grid_transform.f90): phi
// inside every flux derivative, the stored 1/phi on each direction's share of
// the divergences (not on the pressure gradient, u.grad(p) or tau:grad(u)).
template <typename T, typename Q, int DIR, bool PER, bool EV>
__global__ __launch_bounds__(NT) void k_curv_sweep1(const Q *__restrict__ qs, T *__restrict__ rhs,
                                                    T *__restrict__ G, T *__restrict__ TG,
                                                    LineOp<T> op, SweepGeom g, Phys ph, long N,
                                                    CurvMetric<T> mt, Penal<T> pe) {
  constexpr bool FIRST = DIR == 0;
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos<DIR>(g);
  const long B = tp.base, PS = g.PS;
  const T ih = T(ph.inv_h[DIR]);
  // In FP64, rho, and in the periodic x sweep the momenta, stay in
  // registers from their first read to the convective fluxes: those sweeps
  // have the registers for them at two waves per SIMD (the open x sweep
  // would spill, FP32 sweeps would lose waves); re-reading them there
  // missed the caches
  constexpr bool KEEPR = sizeof(T) == 8, KEEPM = KEEPR && DIR == 0 && PER;
  T M0[PPT], M1[PPT], M2[PPT], rho[PPT], irho[PPT], uh[PPT], x[PPT], f[PPT];
  T mk[KEEPM ? 3 : 1][PPT];
  ld<DIR>(mt.M + (DIR + 0) * N, B, PS, M0);
  ld<DIR>(mt.M + (DIR + 3) * N, B, PS, M1);
  ld<DIR>(mt.M + (DIR + 6) * N, B, PS, M2);
  ld<DIR>(qs, B, PS, rho);
#pragma unroll
  for (int m = 0; m < PPT; ++m) irho[m] = T(1) / rho[m];
  // accumulate a value into field slot `fld` (the first sweep writes it)
  auto acc = [&](T *fld, T(&v)[PPT]) {
    if (!FIRST) {
      T o[PPT];
      ld<DIR>(fld, B, PS, o);
#pragma unroll
      for (int m = 0; m < PPT; ++m) v[m] = o[m] + v[m];
    }
    st<DIR>(fld, B, PS, v);
  };

  // physical velocity gradients (:114-135) and the contravariant velocity
  // u_hat_DIR (:139-141); the symmetric sums collect two solves each
  T s3[PPT], s4[PPT], s5[PPT];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T ui[PPT];
    ld<DIR>(qs + (1 + i) * N, B, PS, ui);
    if constexpr (KEEPM) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) mk[i][m] = ui[m];
    }
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      ui[m] = ui[m] * irho[m];
      const T c = i == 0 ? M0[m] : (i == 1 ? M1[m] : M2[m]);
      uh[m] = i == 0 ? c * ui[m] : uh[m] + c * ui[m];
    }
    line_solve<T, PER>(op, lds, tp, ui, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) x[m] = x[m] * ih;
    T d[PPT];
    if (i == 0) {  // a(u): u_x, u_y (+v_x), u_z (+w_x)
#pragma unroll
      for (int m = 0; m < PPT; ++m) {
        d[m] = M0[m] * x[m];
        s3[m] = M1[m] * x[m];
        s4[m] = M2[m] * x[m];
      }
      acc(G + 0 * N, d);
    } else if (i == 1) {  // a(v): v_x (+u_y), v_y, v_z (+w_y)
#pragma unroll
      for (int m = 0; m < PPT; ++m) {
        d[m] = M1[m] * x[m];
        s3[m] = s3[m] + M0[m] * x[m];
        s5[m] = M2[m] * x[m];
      }
      acc(G + 1 * N, d);
      acc(G + 3 * N, s3);
    } else {  // a(w): w_x (+u_z), w_y (+v_z), w_z
#pragma unroll
      for (int m = 0; m < PPT; ++m) {
        d[m] = M2[m] * x[m];
        s4[m] = s4[m] + M0[m] * x[m];
        s5[m] = s5[m] + M1[m] * x[m];
      }
      acc(G + 2 * N, d);
      acc(G + 4 * N, s4);
      acc(G + 5 * N, s5);
    }
  }

  // temperature, its physical gradient (:168-175) and, in the first sweep,
  // the viscosity of the stage input
  T p[PPT];
  ld<DIR>(qs + 4 * N, B, PS, p);
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = temperature(ph, irho[m], p[m]);
  if (FIRST) {
    T mu[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) mu[m] = viscosity(ph, f[m]);
    st<DIR>(G + 6 * N, B, PS, mu);
  }
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    T d[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) d[m] = (j == 0 ? M0[m] : (j == 1 ? M1[m] : M2[m])) * (x[m] * ih);
    acc(TG + j * N, d);
  }

  // pressure gradient along DIR: -M(DIR,i) a(p) in momentum i (:194 etc.),
  // u_hat a(p) in the pressure bracket
  T ap[PPT];
  line_solve<T, PER>(op, lds, tp, p, ap);
#pragma unroll
  for (int m = 0; m < PPT; ++m) ap[m] = ap[m] * ih;

  // convective fluxes with the contravariant velocity (:184-247)
  T PH[PPT], IP[PPT];
  if (EV) {
    ld<DIR>(pe.phi, B, PS, PH);
    ld<DIR>(pe.inv_phi, B, PS, IP);
  }
  auto div = [&](int m, T v) -> T { return EV ? IP[m] * v : v; };
  {  // mass
    if constexpr (!KEEPR) ld<DIR>(qs, B, PS, rho);
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = EV ? (PH[m] * rho[m]) * uh[m] : rho[m] * uh[m];
    line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) x[m] = div(m, -(x[m] * ih));
    acc(rhs, x);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T mom[PPT];
    if constexpr (KEEPM) {
#pragma unroll
      for (int m = 0; m < PPT; ++m) mom[m] = mk[i][m];
    } else {
      ld<DIR>(qs + (1 + i) * N, B, PS, mom);
    }
#pragma unroll
    for (int m = 0; m < PPT; ++m) f[m] = EV ? (PH[m] * mom[m]) * uh[m] : mom[m] * uh[m];
    line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T c = i == 0 ? M0[m] : (i == 1 ? M1[m] : M2[m]);
      x[m] = div(m, -(x[m] * ih)) - c * ap[m];
    }
    acc(rhs + (1 + i) * N, x);
  }
  // pressure bracket: gamma/(gamma-1) (-a(p u_hat)) + u_hat a(p) (:245-250)
#pragma unroll
  for (int m = 0; m < PPT; ++m) f[m] = EV ? (PH[m] * p[m]) * uh[m] : p[m] * uh[m];
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) x[m] = div(m, T(ph.gt) * (-(x[m] * ih))) + uh[m] * ap[m];
  acc(rhs + 4 * N, x);
}

template <typename T, int DIR, bool PER, bool FF, bool EV>
__global__ __launch_bounds__(NT) void k_curv_sweep2(T *__restrict__ rhs, const T *__restrict__ G,
                                                    const T *__restrict__ TG, LineOp<T> op,
                                                    LineOp<T> fop, SweepGeom g, Phys ph, long N,
                                                    int mode, double *__restrict__ qn,
                                                    T *__restrict__ acc,
                                                    T *__restrict__ qs_out, double dt,
                                                    CurvMetric<T> mt, const T *__restrict__ force,
                                                    Penal<T> pe) {
  constexpr bool LAST = DIR == 0;
  __shared__ LineLds<T> lds;
  const ThreadPos tp = thread_pos<DIR>(g);
  const long B = tp.base, PS = g.PS;
  const T ih = T(ph.inv_h[DIR]);
  T M0[PPT], M1[PPT], M2[PPT], mu[PPT];
  ld<DIR>(mt.M + (DIR + 0) * N, B, PS, M0);
  ld<DIR>(mt.M + (DIR + 3) * N, B, PS, M1);
  ld<DIR>(mt.M + (DIR + 6) * N, B, PS, M2);
  ld<DIR>(G + 6 * N, B, PS, mu);
  // heat flux along DIR first (the metric row and mu are live for it and the
  // stresses only): sum_j M(DIR,j) D_DIR(lambda J T_j), no h (:252-268).
  // J, phi and 1/phi are read where they are used (again later in the
  // sweep, from cache) rather than held across it: holding them spilled the
  // open and penalised forms.
  T hs[PPT];
  {
    T lj[PPT];  // (phi) lambda J
    ld<DIR>(mt.J, B, PS, lj);
    if (EV) {
      T PH[PPT];
      ld<DIR>(pe.phi, B, PS, PH);
#pragma unroll
      for (int m = 0; m < PPT; ++m) lj[m] = (PH[m] * lambda_of(ph, mu[m])) * lj[m];
    } else {
#pragma unroll
      for (int m = 0; m < PPT; ++m) lj[m] = lambda_of(ph, mu[m]) * lj[m];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T f[PPT], x[PPT];
      ld<DIR>(TG + j * N, B, PS, f);
#pragma unroll
      for (int m = 0; m < PPT; ++m) f[m] = lj[m] * f[m];
      line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
      for (int m = 0; m < PPT; ++m) {
        const T c = j == 0 ? M0[m] : (j == 1 ? M1[m] : M2[m]);
        hs[m] = j == 0 ? c * x[m] : hs[m] + c * x[m];
      }
    }
  }
  if (EV) {
    T IP[PPT];
    ld<DIR>(pe.inv_phi, B, PS, IP);
#pragma unroll
    for (int m = 0; m < PPT; ++m) hs[m] = IP[m] * hs[m];
  }
  if (!LAST) {
    T k[PPT];
    ld<DIR>(rhs + 4 * N, B, PS, k);
#pragma unroll
    for (int m = 0; m < PPT; ++m) k[m] = k[m] + hs[m];
    st<DIR>(rhs + 4 * N, B, PS, k);
  }
  // projected stresses tau_proj_{i,DIR} = sum_j M(DIR,j) tau_ij (:195-238)
  // (the diagonal first, then one symmetric pair at a time, so that only
  // one gradient field is live besides the partial sums)
  T t1[PPT], t2[PPT], t3[PPT], diss[PPT];
  {
    T ux[PPT], vy[PPT], wz[PPT];
    ld<DIR>(G + 0 * N, B, PS, ux);
    ld<DIR>(G + 1 * N, B, PS, vy);
    ld<DIR>(G + 2 * N, B, PS, wz);
    issue_loads();
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T md = -2.0 / 3.0 * mu[m] * (ux[m] + vy[m] + wz[m]);
      const T t11 = md + 2 * mu[m] * ux[m], t22 = md + 2 * mu[m] * vy[m], t33 = md + 2 * mu[m] * wz[m];
      t1[m] = M0[m] * t11;
      t2[m] = M1[m] * t22;
      t3[m] = M2[m] * t33;
      // tau_ij du_i/dx_j (:270-275): the symmetric sums carry the off-diagonal pairs
      if (LAST) diss[m] = t11 * ux[m] + t22 * vy[m] + t33 * wz[m];
    }
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {  // s12 = u_y + v_x, s13 = u_z + w_x, s23 = v_z + w_y
    T sv[PPT];
    ld<DIR>(G + (3 + s) * N, B, PS, sv);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T t = mu[m] * sv[m];
      if (s == 0) {
        t1[m] = t1[m] + M1[m] * t;
        t2[m] = t2[m] + M0[m] * t;
      } else if (s == 1) {
        t1[m] = t1[m] + M2[m] * t;
        t3[m] = t3[m] + M0[m] * t;
      } else {
        t2[m] = t2[m] + M2[m] * t;
        t3[m] = t3[m] + M1[m] * t;
      }
      if (LAST) diss[m] = diss[m] + t * sv[m];
    }
  }
  auto finish = [&](int v, T(&k)[PPT]) {
    if (LAST && force) {
      T fv[PPT];
      ld<DIR>(force + v * N, B, PS, fv);
#pragma unroll
      for (int m = 0; m < PPT; ++m) k[m] = k[m] + fv[m];
    }
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
  if (EV) {  // phi tau_proj inside the derivative
    T PH[PPT];
    ld<DIR>(pe.phi, B, PS, PH);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      t1[m] = PH[m] * t1[m];
      t2[m] = PH[m] * t2[m];
      t3[m] = PH[m] * t3[m];
    }
  }
  T Jv[PPT];
  if (LAST) ld<DIR>(mt.J, B, PS, Jv);
  if (LAST) {  // pressure: J (gamma - 1) (bracket + heat + J tau:grad(u)) (:270-276)
    T k[PPT];
    ld<DIR>(rhs + 4 * N, B, PS, k);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      k[m] = k[m] + hs[m];
      k[m] = (Jv[m] * T(ph.gm1)) * (k[m] + Jv[m] * diss[m]);
    }
    finish(4, k);
  }
  if (LAST) {  // mass (:189-190)
    T k[PPT];
    ld<DIR>(rhs, B, PS, k);
#pragma unroll
    for (int m = 0; m < PPT; ++m) k[m] = Jv[m] * k[m];
    finish(0, k);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T x[PPT];
    line_solve<T, PER>(op, lds, tp, i == 0 ? t1 : (i == 1 ? t2 : t3), x);
    asm volatile("" ::: "memory");
    T k[PPT];
    ld<DIR>(rhs + (1 + i) * N, B, PS, k);
    if (EV) {
      T IP[PPT];
      ld<DIR>(pe.inv_phi, B, PS, IP);
#pragma unroll
      for (int m = 0; m < PPT; ++m) x[m] = IP[m] * x[m];
    }
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      k[m] = k[m] + x[m] * ih;
      if (LAST) k[m] = Jv[m] * k[m];
    }
    if (LAST)
      finish(1 + i, k);
    else
      st<DIR>(rhs + (1 + i) * N, B, PS, k);
  }
}

}  // namespace cruna
