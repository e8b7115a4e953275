// Weighted filter: filter_weighted_iN_mpi_cartesian::filter_4d
// (src/filter/filter_weighted_i1_mpi_cartesian.f90:15-54) with the
// filter_weighted_pen_core_o2_e_c_p core, the reference's only core with an
// effect (src/filter/filter_weighted_pen_core_o2_e_c_p.f90):
//   F(g)_i = a0 g_i + a1 (g_{i+1} + g_{i-1}),  a0 = -2/4, a1 = 1/4, periodic
//   q <- q - F(strength weight F(q)) / weight
// along one direction.  F(q) is formed at the point and its two neighbours
// (five reads of q along the line, cache-resident), so one pass per
// direction reads q once from HBM and writes the result out of place.
#pragma once

#include <hip/hip_runtime.h>

namespace cruna {

template <int DIR>
__global__ __launch_bounds__(256) void k_filter_weighted(const double *__restrict__ q,
                                                         double *__restrict__ out, int nv,
                                                         const double *__restrict__ weight,
                                                         const double *__restrict__ strength,
                                                         int n1, int n2, int n3, long N) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int n = DIR == 0 ? n1 : (DIR == 1 ? n2 : n3);
  const long st = DIR == 0 ? 1 : (DIR == 1 ? (long)n1 : (long)n1 * n2);
  const int i = DIR == 0 ? int(p % n1) : (DIR == 1 ? int((p / n1) % n2) : int(p / ((long)n1 * n2)));
  // periodic neighbours along DIR (filter_main's wrap of rows 1 and n)
  auto nb = [&](int k) -> long {
    int j = i + k;
    j = j < 0 ? j + n : (j >= n ? j - n : j);
    return p + (long)(j - i) * st;
  };
  const long pm2 = nb(-2), pm1 = nb(-1), pp1 = nb(1), pp2 = nb(2);
  const double a0 = -2.0 * 0.25, a1 = 1.0 * 0.25;
  auto sw = [&](long o) -> double {
    const double s = strength ? strength[o] : 1.0, w = weight ? weight[o] : 1.0;
    return s * w;
  };
  const double swm = sw(pm1), sw0 = sw(p), swp = sw(pp1);
  const double w0 = weight ? weight[p] : 1.0;
  for (int v = 0; v < nv; ++v) {
    const double *x = q + (long)v * N;
    const double xm2 = x[pm2], xm1 = x[pm1], x0 = x[p], xp1 = x[pp1], xp2 = x[pp2];
    const double fm = swm * (a0 * xm1 + a1 * (x0 + xm2));
    const double f0 = sw0 * (a0 * x0 + a1 * (xp1 + xm1));
    const double fp = swp * (a0 * xp1 + a1 * (xp2 + x0));
    const double g = a0 * f0 + a1 * (fp + fm);
    out[(long)v * N + p] = x0 - g / w0;
  }
}

}  // namespace cruna
