// Boundary conditions of src/boundaries/boundary_conditions.f90 used by the
// hot path: types 200-203 (characteristic, nonreflecting toward qref or the
// inner neighbour), 204-206 (the same on the RHS), 100 (slip wall on the
// RHS), 400 (sponge on the RHS) and 401 (sponge on q).  Point-wise kernels, one thread per
// point; the characteristic decomposition always runs in FP64, also in the
// FP32-mixed mode.
#pragma once

#include <hip/hip_runtime.h>

namespace cruna {

struct BoxRow {
  int i0, i1, j0, j1, k0, k1;  // 0-based inclusive
  double k[3];                 // outward normal (kvector), as given in the row
  long aim_shift;              // type 200: offset of the inner neighbour (zero-gradient aim)
};

// characteristic_q_k1d (src/boundaries/nonreflecting.f90:46-137) on the
// points of one box: q <- q + (incoming part of prim(aim) - prim(q)).
// Type 201: aim = qref at the same point (row.aim_shift 0); type 200: aim =
// q at the inner neighbour (characteristic_zero_grad_init, :665-728), a point
// outside the one-point-thick box, so the in-place update reads no value it
// writes.
// kfield (types 202/205, characteristic_*_k4d): the normal per point,
// kvector(n1,n2,n3,3), in place of the row's; null: the row's (201/204).
// shape (type 203, characteristic_q_k4d with shape, :180-185): the change
// toward aim is scaled by shape at the point before the decomposition.
template <typename Q, typename A>
__global__ __launch_bounds__(256) void k_bc_characteristic_q(Q *q, const A *aim_base, BoxRow row,
                                                              int n1, int n2, long N, double gamma,
                                                              const double *__restrict__ kfield,
                                                              const double *__restrict__ shape) {
  const long nx = row.i1 - row.i0 + 1, ny = row.j1 - row.j0 + 1, nz = row.k1 - row.k0 + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny * nz) return;
  const long i = row.i0 + t % nx, j = row.j0 + (t / nx) % ny, k = row.k0 + t / (nx * ny);
  const long o = i + (long)n1 * (j + (long)n2 * k);
  const A *aim = aim_base + row.aim_shift;
  double r = double(q[o]);
  double u1 = double(q[N + o]) / double(q[o]);
  double u2 = double(q[2 * N + o]) / double(q[o]);
  double u3 = double(q[3 * N + o]) / double(q[o]);
  double p = double(q[4 * N + o]);
  const double c = sqrt(gamma * double(q[4 * N + o]) / double(q[o]));
  double dq1 = double(aim[o]) - r;
  double dq2 = double(aim[N + o]) / double(aim[o]) - u1;
  double dq3 = double(aim[2 * N + o]) / double(aim[o]) - u2;
  double dq4 = double(aim[3 * N + o]) / double(aim[o]) - u3;
  double dq5 = double(aim[4 * N + o]) - p;
  if (shape) {
    const double sh = shape[o];
    dq1 *= sh;
    dq2 *= sh;
    dq3 *= sh;
    dq4 *= sh;
    dq5 *= sh;
  }
  const double kx = -(kfield ? kfield[o] : row.k[0]), ky = -(kfield ? kfield[N + o] : row.k[1]),
               kz = -(kfield ? kfield[2 * N + o] : row.k[2]);
  // primitive -> characteristic waves (:97-101)
  double R1 = kx * dq1 + kz * dq3 - ky * dq4 - kx * dq5 / (c * c);
  double R2 = ky * dq1 - kz * dq2 + kx * dq4 - ky * dq5 / (c * c);
  double R3 = kz * dq1 + ky * dq2 - kx * dq3 - kz * dq5 / (c * c);
  double R4 = kx * dq2 + ky * dq3 + kz * dq4 + dq5 / (r * c);
  double R5 = -kx * dq2 - ky * dq3 - kz * dq4 + dq5 / (r * c);
  // drop outgoing waves (:104-114)
  const double lamS = kx * u1 + ky * u2 + kz * u3;
  if (lamS < 0.0) R1 = R2 = R3 = 0.0;
  if (lamS + c < 0.0) R4 = 0.0;
  if (lamS - c < 0.0) R5 = 0.0;
  // back to primitive increments (:118-122) and conservative state (:125-135)
  dq1 = kx * R1 + ky * R2 + kz * R3 + r / (2.0 * c) * R4 + r / (2.0 * c) * R5;
  dq2 = -kz * R2 + ky * R3 + 0.5 * kx * R4 - 0.5 * kx * R5;
  dq3 = kz * R1 - kx * R3 + 0.5 * ky * R4 - 0.5 * ky * R5;
  dq4 = -ky * R1 + kx * R2 + 0.5 * kz * R4 - 0.5 * kz * R5;
  dq5 = +0.5 * r * c * R4 + 0.5 * r * c * R5;
  r = r + dq1;
  u1 = u1 + dq2;
  u2 = u2 + dq3;
  u3 = u3 + dq4;
  p = p + dq5;
  q[o] = Q(r);
  q[N + o] = Q(r * u1);
  q[2 * N + o] = Q(r * u2);
  q[3 * N + o] = Q(r * u3);
  q[4 * N + o] = Q(p);
}

// characteristic_rhs_k1d (src/boundaries/nonreflecting.f90:237-341, type 204)
// on the complete RHS of the points of one box: the change that the RHS
// makes over dt plus the change toward aim, in primitive variables, keeps
// only its incoming waves and becomes the RHS again.  With shape (type 206,
// characteristic_rhs_k4d, :431-439) the result is blended with the RHS as
// (1 - shape) rhs + shape (new rhs).
template <typename T, typename Q>
__global__ __launch_bounds__(256) void k_bc_characteristic_rhs(T *__restrict__ rhs,
                                                                const Q *__restrict__ q,
                                                                const double *__restrict__ aim,
                                                                BoxRow row, int n1, int n2, long N,
                                                                double gamma, double dt,
                                                                const double *__restrict__ kfield,
                                                                const double *__restrict__ shape) {
  const long nx = row.i1 - row.i0 + 1, ny = row.j1 - row.j0 + 1, nz = row.k1 - row.k0 + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny * nz) return;
  const long i = row.i0 + t % nx, j = row.j0 + (t / nx) % ny, k = row.k0 + t / (nx * ny);
  const long o = i + (long)n1 * (j + (long)n2 * k);
  const double inv_dt = 1.0 / dt;
  double r = double(q[o]);
  const double u1 = double(q[N + o]) / double(q[o]);
  const double u2 = double(q[2 * N + o]) / double(q[o]);
  const double u3 = double(q[3 * N + o]) / double(q[o]);
  const double c = sqrt(gamma * double(q[4 * N + o]) / double(q[o]));
  // change required by aim (:274-278), plus the RHS over dt (:280-285)
  double dq1 = aim[o] - r;
  double dq2 = aim[N + o] / aim[o] - u1;
  double dq3 = aim[2 * N + o] / aim[o] - u2;
  double dq4 = aim[3 * N + o] / aim[o] - u3;
  double dq5 = aim[4 * N + o] - double(q[4 * N + o]);
  double tmp = dt / r;
  const double f1 = double(rhs[o]), f2 = double(rhs[N + o]), f3 = double(rhs[2 * N + o]),
               f4 = double(rhs[3 * N + o]), f5 = double(rhs[4 * N + o]);
  dq1 = dq1 + f1 * dt;
  dq2 = dq2 + f2 * tmp;
  dq3 = dq3 + f3 * tmp;
  dq4 = dq4 + f4 * tmp;
  dq5 = dq5 + f5 * dt;
  const double kx = -(kfield ? kfield[o] : row.k[0]), ky = -(kfield ? kfield[N + o] : row.k[1]),
               kz = -(kfield ? kfield[2 * N + o] : row.k[2]);
  double R1 = kx * dq1 + kz * dq3 - ky * dq4 - kx * dq5 / (c * c);
  double R2 = ky * dq1 - kz * dq2 + kx * dq4 - ky * dq5 / (c * c);
  double R3 = kz * dq1 + ky * dq2 - kx * dq3 - kz * dq5 / (c * c);
  double R4 = kx * dq2 + ky * dq3 + kz * dq4 + dq5 / (r * c);
  double R5 = -kx * dq2 - ky * dq3 - kz * dq4 + dq5 / (r * c);
  const double lamS = kx * u1 + ky * u2 + kz * u3;
  if (lamS < 0.0) R1 = R2 = R3 = 0.0;
  if (lamS + c < 0.0) R4 = 0.0;
  if (lamS - c < 0.0) R5 = 0.0;
  dq1 = kx * R1 + ky * R2 + kz * R3 + r / (2.0 * c) * R4 + r / (2.0 * c) * R5;
  dq2 = -kz * R2 + ky * R3 + 0.5 * kx * R4 - 0.5 * kx * R5;
  dq3 = kz * R1 - kx * R3 + 0.5 * ky * R4 - 0.5 * ky * R5;
  dq4 = -ky * R1 + kx * R2 + 0.5 * kz * R4 - 0.5 * kz * R5;
  dq5 = +0.5 * r * c * R4 + 0.5 * r * c * R5;
  r = r + dq1;
  if (shape) {  // :431-439
    const double sh = shape[o], keep = 1.0 - sh;
    tmp = sh * inv_dt;
    rhs[o] = T(keep * f1 + dq1 * tmp);
    rhs[4 * N + o] = T(keep * f5 + dq5 * tmp);
    tmp = tmp * r;
    rhs[N + o] = T(keep * f2 + dq2 * tmp);
    rhs[2 * N + o] = T(keep * f3 + dq3 * tmp);
    rhs[3 * N + o] = T(keep * f4 + dq4 * tmp);
    return;
  }
  // back to the RHS (:321-339, no shape)
  rhs[o] = T(dq1 * inv_dt);
  tmp = inv_dt * r;
  rhs[N + o] = T(dq2 * tmp);
  rhs[2 * N + o] = T(dq3 * tmp);
  rhs[3 * N + o] = T(dq4 * tmp);
  rhs[4 * N + o] = T(dq5 * inv_dt);
}

// wall_slip_no_thermodynamics (src/boundaries/wall_boundary_conditions.f90:13-23,
// type 100): the RHS of one variable (v, 0-based) is zero on the box.
template <typename T>
__global__ __launch_bounds__(256) void k_bc_wall_rhs(T *__restrict__ rhs, BoxRow row, int v, int n1,
                                                      int n2, long N) {
  const long nx = row.i1 - row.i0 + 1, ny = row.j1 - row.j0 + 1, nz = row.k1 - row.k0 + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny * nz) return;
  const long i = row.i0 + t % nx, j = row.j0 + (t / nx) % ny, k = row.k0 + t / (nx * ny);
  rhs[(long)v * N + i + (long)n1 * (j + (long)n2 * k)] = T(0);
}

// sponge_q with qref (src/boundaries/sponge_layer_core.f90:70-103):
// q <- q - sigma (q - qref) for all five variables where sigma != 0.
template <typename Q>
__global__ __launch_bounds__(256) void k_bc_sponge_q(Q *__restrict__ q,
                                                      const double *__restrict__ qref,
                                                      const double *__restrict__ sigma, long N) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= N) return;
  const double s = sigma[o];
  if (s == 0.0) return;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double x = double(q[v * N + o]);
    q[v * N + o] = Q(x - s * (x - qref[v * N + o]));
  }
}

// sponge_rhs with qref (sponge_layer_core.f90:34-66): rhs -= sigma (q - qref).
// The accumulator is finished by the last sweep, which multiplies the
// pressure slot by (gamma - 1) and, on a deformed grid, every slot by the
// Jacobian J; the sponge acts on the finished RHS, so its share enters
// divided by those factors (J null: Cartesian grid).
template <typename T, typename Q>
__global__ __launch_bounds__(256) void k_sponge_rhs(T *__restrict__ rhs, const Q *__restrict__ q,
                                                     const double *__restrict__ qref,
                                                     const double *__restrict__ sigma, long N,
                                                     double gm1, const T *__restrict__ J) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= N) return;
  const double s = sigma[o];
  if (s == 0.0) return;
  const double jac = J ? double(J[o]) : 1.0;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double d = s * (double(q[v * N + o]) - qref[v * N + o]);
    if (v == 4) d = d / gm1;
    if (J) d = d / jac;
    rhs[v * N + o] = T(double(rhs[v * N + o]) - d);
  }
}

// Stage input with boundary conditions: a private copy of the state
// (q = q_in, ns_rho_rhou_p.f90:89) in the compute precision.
template <typename T>
__global__ __launch_bounds__(256) void k_copy_state(T *__restrict__ dst, const double *__restrict__ src,
                                                     long n) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n) dst[o] = T(src[o]);
}

}  // namespace cruna
