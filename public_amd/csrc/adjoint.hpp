// Adjoint right-hand side and adjoint boundary conditions on the device.
//
// right_hand_side_adjoint_3d (src/equations/ns_rho_rhou_p.f90:257-536) is the
// continuous adjoint of the Euler part in the variables (rho, rho u, rho v,
// rho w, p / (gamma - 1)).  Along direction d it takes seven compact
// derivatives -- the five adjoint variables qs1..qs5, -u_d qs5 and the
// forward pressure p -- and combines them point-wise with the forward state
// q.  One sweep per direction applies all seven on register-resident blocks.
//
// With A = D qs1, B_m = D qs_{m+2} (m = 0,1,2), E = D qs5, F = D(-u_d qs5),
// P = D p, gdp = gamma p / (rho (gamma - 1)) and
//   S = sum_m u_m B_m + gdp E + qs5 P / rho
// the contributions of direction d (ns_rho_rhou_p.f90:317-372, 407-445,
// 475-514) are
//   r1        += u_d S
//   r_{d+2}   -= A + u_d B_d + S
//   r_{m+2}   -= u_d B_m            (m != d)
//   r5        -= (gamma - 1) (B_d + F) + gamma u_d E
// (the reference spells S out term by term; the sums agree to rounding).
// The sweeps run z, y, x; the x sweep adds the objective gradient g and
// finishes the RK4 stage of time_stepper_adjoint (rk_o4_e.f90:62-105,
// dt -> -dt).
#pragma once

#include <climits>

#include "boundary.hpp"
#include "sweeps.hpp"

namespace cruna {

// Hand the accumulated variable v on: intermediate sweeps store it; the last
// sweep adds g and applies the RK4 stage update.
template <int DIR, typename T, bool LAST>
__device__ __forceinline__ void finish(int v, T (&r)[PPT], long B, long PS, long N, T *rhs,
                                       const double *__restrict__ gobj, int mode,
                                       double *__restrict__ qn, T *__restrict__ acc,
                                       T *qs_out, double dt) {
  if (!LAST) {
    st<DIR>(rhs + v * N, B, PS, r);
    return;
  }
  if (gobj) {
    double gg[PPT];
    ld<DIR>(gobj + v * N, B, PS, gg);
#pragma unroll
    for (int m = 0; m < PPT; ++m) r[m] = r[m] + T(gg[m]);
  }
  rk_update<DIR, T>(mode, v, r, B, PS, N, rhs, qn, acc, qs_out, dt);
}

// One adjoint sweep along DIR.
//   qs   : adjoint stage input (5 fields, type Q)
//   q0p  : device cell holding the forward state of the step (5 fields, FP64),
//          frozen over the stages; the cell lets one recorded launch sequence
//          serve every step of a backward loop (the slot is set on the device)
//   rhs  : 5 accumulator fields (written by the first sweep, z)
//   gobj : objective gradient g (5 fields) or null; added by the x sweep
// The x sweep (LAST) completes the stage: mode 0..3 RK4 update with step dt
// (negative for the backward integration), mode 4 writes the RHS.
template <typename T, typename Q, int DIR, bool PER>
__device__ __forceinline__ void asweep_body(const ThreadPos &tp, const Q *qs,
                                            const double *const *__restrict__ q0p, T *__restrict__ rhs,
                                            const double *__restrict__ gobj, const LineOp<T> &op,
                                            const SweepGeom &g, double gamma, double inv_h, long N, int mode,
                                            double *__restrict__ qn, T *__restrict__ acc, T *qs_out,
                                            double dt) {
  constexpr bool FIRST = DIR == 2, LAST = DIR == 0;
  const double *__restrict__ q = *q0p;
  __shared__ LineLds<T> lds;
  const long B = tp.base, PS = g.PS;
  const T ih = T(inv_h), gam = T(gamma), gm1 = T(gamma - 1.0), ggm = T(gamma / (gamma - 1.0));
  T ud[PPT], irho[PPT], S[PPT], R5[PPT], f[PPT], x[PPT], r[PPT];
  {
    T gdp[PPT], qs5[PPT];
    {
      T rho[PPT], md[PPT];
      ld<DIR>(q, B, PS, rho);
      ld<DIR>(q + (1 + DIR) * N, B, PS, md);
      ld<DIR>(q + 4 * N, B, PS, f);  // p
      ld<DIR>(qs + 4 * N, B, PS, qs5);
      issue_loads();
#pragma unroll
      for (int m = 0; m < PPT; ++m) {
        irho[m] = T(1) / rho[m];
        ud[m] = md[m] * irho[m];
        gdp[m] = ggm * f[m] * irho[m];
      }
    }
    // P = D p
    line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) S[m] = qs5[m] * irho[m] * (x[m] * ih);
    // E = D qs5
    line_solve<T, PER>(op, lds, tp, qs5, x);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T E = x[m] * ih;
      S[m] = S[m] + gdp[m] * E;
      R5[m] = gam * ud[m] * E;
      f[m] = -ud[m] * qs5[m];
    }
  }
  // F = D(-u_d qs5)
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) R5[m] = R5[m] + gm1 * (x[m] * ih);
  // B_d = D qs_{d+2} completes r5 (the fences keep each group of loads next
  // to the solve it feeds, and the accumulators and forward momenta that a
  // solve's result meets are loaded after it, so that they take no registers
  // across it)
  T Ta[PPT];
  {
    asm volatile("" ::: "memory");
    ld<DIR>(qs + (1 + DIR) * N, B, PS, f);
    line_solve<T, PER>(op, lds, tp, f, x);
    asm volatile("" ::: "memory");
    if (!FIRST) ld<DIR>(rhs + 4 * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T Bd = x[m] * ih;
      S[m] = S[m] + ud[m] * Bd;
      Ta[m] = ud[m] * Bd;
      R5[m] = R5[m] + gm1 * Bd;
      r[m] = FIRST ? -R5[m] : r[m] - R5[m];
    }
    finish<DIR, T, LAST>(4, r, B, PS, N, rhs, gobj, mode, qn, acc, qs_out, dt);
  }
  // A = D qs1; A + u_d B_d waits in LDS (the thread's own slots) until the
  // end, which keeps the x sweep within 256 registers without spilling
  __shared__ T stash[PPT][NT];
  asm volatile("" ::: "memory");
  ld<DIR>(qs, B, PS, f);
  line_solve<T, PER>(op, lds, tp, f, x);
#pragma unroll
  for (int m = 0; m < PPT; ++m) stash[m][threadIdx.x] = Ta[m] + x[m] * ih;
  // B_m = D qs_{m+2}, m != d
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i == DIR) continue;
    T um[PPT];
    asm volatile("" ::: "memory");
    ld<DIR>(qs + (1 + i) * N, B, PS, f);
    line_solve<T, PER>(op, lds, tp, f, x);
    asm volatile("" ::: "memory");
    ld<DIR>(q + (1 + i) * N, B, PS, um);
    if (!FIRST) ld<DIR>(rhs + (1 + i) * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T Bm = x[m] * ih;
      S[m] = S[m] + (um[m] * irho[m]) * Bm;
      r[m] = FIRST ? -(ud[m] * Bm) : r[m] - ud[m] * Bm;
    }
    finish<DIR, T, LAST>(1 + i, r, B, PS, N, rhs, gobj, mode, qn, acc, qs_out, dt);
  }
  // r1 += u_d S, r_{d+2} -= A + u_d B_d + S
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const int var = v == 0 ? 0 : 1 + DIR;
    asm volatile("" ::: "memory");
    if (!FIRST) ld<DIR>(rhs + var * N, B, PS, r);
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
      const T c = v == 0 ? ud[m] * S[m] : -(stash[m][threadIdx.x] + S[m]);
      r[m] = FIRST ? c : r[m] + c;
    }
    finish<DIR, T, LAST>(var, r, B, PS, N, rhs, gobj, mode, qn, acc, qs_out, dt);
  }
}

template <typename T, typename Q, int DIR, bool PER>
__global__ __launch_bounds__(NT) void k_asweep(const Q *qs, const double *const *__restrict__ q0p,
                                               T *__restrict__ rhs, const double *__restrict__ gobj,
                                               LineOp<T> op, SweepGeom g, double gamma, double inv_h, long N,
                                               int mode, double *__restrict__ qn, T *__restrict__ acc,
                                               T *qs_out, double dt) {
  asweep_body<T, Q, DIR, PER>(thread_pos<DIR>(g), qs, q0p, rhs, gobj, op, g, gamma, inv_h, N, mode, qn, acc, qs_out,
                              dt);
}

// characteristic_qs_k1d (src/boundaries/nonreflecting.f90:453-554) on the
// points of one box: removes the incoming adjoint characteristics of qs.
// shape (type 203, characteristic_qs_k4d, :596-601): the change toward zero
// is scaled by shape at the point.
template <typename Q>
__global__ __launch_bounds__(256) void k_bc_characteristic_qs(Q *__restrict__ qs,
                                                               const double *const *__restrict__ q0p,
                                                               BoxRow row, int n1, int n2, long N,
                                                               double gamma,
                                                               const double *__restrict__ kfield,
                                                               const double *__restrict__ shape) {
  // forward state at the point (201) or at its inner neighbour (200)
  const double *__restrict__ q = *q0p + row.aim_shift;
  const long nx = row.i1 - row.i0 + 1, ny = row.j1 - row.j0 + 1, nz = row.k1 - row.k0 + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny * nz) return;
  const long i = row.i0 + t % nx, j = row.j0 + (t / nx) % ny, k = row.k0 + t / (nx * ny);
  const long o = i + (long)n1 * (j + (long)n2 * k);
  const double s1 = double(qs[o]), s2 = double(qs[N + o]), s3 = double(qs[2 * N + o]),
               s4 = double(qs[3 * N + o]), s5 = double(qs[4 * N + o]);
  // the row's normal, or the point's (202: characteristic_qs_k4d)
  const double kx = -(kfield ? kfield[o] : row.k[0]), ky = -(kfield ? kfield[N + o] : row.k[1]),
               kz = -(kfield ? kfield[2 * N + o] : row.k[2]);
  const double gm1 = gamma - 1.0;
  const double c = sqrt(gamma * (q[4 * N + o] / q[o]));
  const double g_c = gm1 / c;
  const double u = q[N + o] / q[o], v = q[2 * N + o] / q[o], w = q[3 * N + o] / q[o];
  // increments toward zero adjoint in the incoming waves
  const double sh = shape ? shape[o] : 1.0;
  const double d1 = -s1 * sh, d2 = -s2 * sh, d3 = -s3 * sh, d4 = -s4 * sh, d5 = -s5 * sh;
  const double ku = kx * u + ky * v + kz * w;
  const double kus = kx * d2 + ky * d3 + kz * d4;
  const double ruus = (d1 + u * d2 + v * d3 + w * d4) / c;
  const double rups = ruus + d5 / g_c;
  double drs = 0, dus = 0, dvs = 0, dws = 0, dps = 0;
  if (ku < 0.0) {  // entropy and vorticity waves enter
    drs = d1 + ku * kus;
    dus = d2 - kx * kus;
    dvs = d3 - ky * kus;
    dws = d4 - kz * kus;
    dps = -g_c * ruus;
  }
  if (ku + c < 0.0) {  // acoustic wave ku + c enters
    const double a = 0.5 * (kus + rups);
    drs = drs - ku * a;
    dus = dus + kx * a;
    dvs = dvs + ky * a;
    dws = dws + kz * a;
    dps = dps + g_c * a;
  }
  if (ku - c < 0.0) {  // acoustic wave ku - c enters
    const double a = 0.5 * (kus - rups);
    drs = drs - ku * a;
    dus = dus + kx * a;
    dvs = dvs + ky * a;
    dws = dws + kz * a;
    dps = dps - g_c * a;
  }
  qs[o] = Q(s1 + drs);
  qs[N + o] = Q(s2 + dus);
  qs[2 * N + o] = Q(s3 + dvs);
  qs[3 * N + o] = Q(s4 + dws);
  qs[4 * N + o] = Q(s5 + dps);
}

// sponge_rhss_correct (src/boundaries/sponge_layer_core.f90:107-151), the
// adjoint sponge of rows 400 when set_boundary_condition_rhss gets the
// forward state (the Euler adjoint, euler_rho_rhou_p.f90:384):
// rhss += sigma (qs1 - u.qs(2:4)/rho, qs(2:4)/rho, (gamma - 1) qs5), with qs
// the adjoint state the RHS was computed from and q the forward state.
template <typename T, typename Q>
__global__ __launch_bounds__(256) void k_sponge_rhss(T *__restrict__ rhs, const Q *__restrict__ qs,
                                                      const double *const *__restrict__ q0p,
                                                      const double *__restrict__ sigma, double gamma,
                                                      long N) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= N) return;
  const double *__restrict__ q = *q0p;
  const double rho = q[o];
  const double u1 = q[N + o] / rho, u2 = q[2 * N + o] / rho, u3 = q[3 * N + o] / rho;
  const double s1 = double(qs[o]), s2 = double(qs[N + o]), s3 = double(qs[2 * N + o]),
               s4 = double(qs[3 * N + o]), s5 = double(qs[4 * N + o]);
  const double sg = sigma[o];
  const double r1 = 1.0 * s1 - u1 / rho * s2 - u2 / rho * s3 - u3 / rho * s4;
  rhs[o] = T(double(rhs[o]) + sg * r1);
  rhs[N + o] = T(double(rhs[N + o]) + sg * (1.0 / rho * s2));
  rhs[2 * N + o] = T(double(rhs[2 * N + o]) + sg * (1.0 / rho * s3));
  rhs[3 * N + o] = T(double(rhs[3 * N + o]) + sg * (1.0 / rho * s4));
  rhs[4 * N + o] = T(double(rhs[4 * N + o]) + sg * ((gamma - 1.0) * s5));
}

// sponge_qs (src/boundaries/sponge_layer_core.f90:155-167): qs <- qs - sigma qs.
template <typename Q>
__global__ __launch_bounds__(256) void k_bc_sponge_qs(Q *__restrict__ qs,
                                                       const double *__restrict__ sigma, long N) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= N) return;
  const double s = sigma[o];
  if (s == 0.0) return;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double x = double(qs[v * N + o]);
    qs[v * N + o] = Q(x - s * x);
  }
}

// The forward state of the next backward step: the step counter *ntp (set
// by the host before the loop) selects trajectory slot nt (or the only slot
// of a one-slot trajectory), then counts down.  One thread writes.
__global__ void k_adjoint_slot(const double **q0p, int *ntp, const double *traj, long stride,
                               int nslots) {
  if (threadIdx.x != 0) return;
  const int nt = *ntp;
  q0p[0] = traj + (long)(nslots > 1 ? nt - 1 : 0) * stride;
  ntp[0] = nt - 1;
}

// Subset checkpoints inside a recorded period: the subset counter *sp
// (device cell) names the checkpoint to restore; every thread copies its
// share of that state into q (grid-stride, 16-byte pairs; the stride of a
// state, 5 N, is even or handled by the tail thread).
__global__ __launch_bounds__(256) void k_restore_ckpt(double *__restrict__ q, const double *__restrict__ ckpt,
                                                      const int *__restrict__ sp, long M) {
  const int s = *sp;
  const double *src = ckpt + (long)s * M;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < M; i += (long)gridDim.x * 256) q[i] = src[i];
}

// Set an int cell (counter) from the stream, so that a recorded graph can
// reset it without a memset node.
__global__ void k_set_int(int *p, int v, int delta) {
  if (threadIdx.x == 0) p[0] = v == INT_MIN ? p[0] + delta : v;
}

// Point the forward-state cell at a given state.  One thread writes.
__global__ void k_set_state_ptr(const double **q0p, const double *q) {
  if (threadIdx.x == 0) q0p[0] = q;
}

}  // namespace cruna
