// Host-side construction of the line operators used by the device sweeps.
//
// Every compact derivative or filter of the reference is  x = B^-1 (A f)
// along one grid line:
//   derivative (Lele 1992, 6th order):
//     periodic  src/diff/diff_pen_core_o6_i_c_p.f90:87-228
//     open      src/diff/diff_pen_core_o6_i_c_o.f90:85-199
//   derivative, 4th-order explicit (B = I):
//     periodic  src/diff/diff_pen_core_o4_e_c_p.f90:53-82
//     open      src/diff/diff_pen_core_o4_e_c_o.f90:53-82
//   filter (Gaitonde-Visbal 2000, 6th order):
//     periodic  src/filter/filter_pen_core_o6_i_c_p.f90:84-249
//     open      src/filter/filter_pen_core_o6_i_c_o.f90:79-240
// The reference factors B with LAPACK dgttrf (plus a Sherman-Morrison
// correction for the cyclic case) and solves line by line with dgttrs.  On
// the GPU a line of n points is split into n/8 blocks of 8 consecutive points,
// one block per thread:
//   1. y_b = M_b f_b   with M_b = T_b^-1 A_b, T_b the 8x8 diagonal block of B
//      and f_b the block's 8 values plus 3 halo values on each side;
//   2. the block-boundary unknowns z_b = (x_first, x_last) satisfy the 2x2
//      block-tridiagonal system  z_b = Y_b + A_b z_{b-1} + C_b z_{b+1},
//      solved by parallel cyclic reduction over blocks with precomputed
//      2x2 weights (cyclic wrap for periodic lines);
//   3. x = y_b + L_b x_last(b-1) + R_b x_first(b+1)  (the "spikes" L, R).
// In exact arithmetic this is the same solution as the LU solve; the reduced
// system is truncated once its couplings fall below 1e-18 (periodic) or
// solved exactly (open).  The tables are built here in FP64.
// Open lines of any length n >= 9 are carried as lines of n_pad (a multiple
// of 8) points: pf padding points, the n real points, then padding to the
// end.  The padding rows are decoupled: B is the identity there and A is
// zero (derivative) or the identity (filter).  The open closures never reach
// past the line's ends and B's first and last real rows have no coupling
// beyond them, so the real values of the solution are exactly those of the
// n-point line; the padding points keep a derivative of 0 and an unfiltered
// value.  pf (0..4) puts the last real point at position 3..7 of its block,
// so that the one-sided closure rows (which reach up to 6 points inward) find
// their points within one block and its 3-point halo.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace cruna {

constexpr int PPT = 8;           // points per thread (block length)
constexpr int HALO = 3;          // explicit-stencil half width carried in M
constexpr int FE = PPT + 2 * HALO;  // 14 = extended block
constexpr int MAXK = 12;         // max reduced-system stages

enum OpKind { OP_DIFF = 0, OP_FILTER = 1, OP_DIFF_O4E = 2, OP_DIFF_O4P = 3, OP_DIFF_O2E = 4 };

// M, L, R are stored for four block types: 0 = first block, 1 = interior
// blocks, 2 = last block, 3 = the block before the last when rows of the
// end closure reach into it (a padded open line); all equal on a periodic
// line.  W is stored per
// block for open lines ([K][NB][12]) and once per stage for periodic lines
// ([K][1][12]): W0, Wm, Wp as row-major 2x2 matrices.
// I holds the interior-block form of M (an 8x14 product) as the explicit
// stencil plus a local Thomas solve, which needs 21 numbers instead of 112:
//   I[0..3] = c0..c3, I[4] = s:  r_i = c0 f_i + sum_k c_k (f_{i+k} + s f_{i-k})
//   I[5..11]  = l_1..l_7   forward elimination  y_i = r_i - l_i y_{i-1}
//   I[12..19] = 1/d_0..1/d_7, I[20] = u:  x_i = (y_i - u x_{i+1}) / d_i
constexpr int NI = 21;

constexpr int NTYPES = 4;

struct HostLineOp {
  int n = 0, NB = 0, K = 0, per = 1, nw = 1;
  int n_real = 0, pf = 0;  // real points and the padding in front of them
  int spec2 = 0;          // block NB-2 has its own dense operator (type 3)
  std::vector<double> I;  // [NI] interior blocks
  std::vector<double> M;  // [NTYPES][PPT][FE]
  std::vector<double> L;  // [NTYPES][PPT]
  std::vector<double> R;  // [NTYPES][PPT]
  std::vector<double> W;  // [K][nw][12]
};

namespace detail {

inline void inverse(std::vector<double> &a, int n) {
  // Gauss-Jordan with partial pivoting, in place (n x n row-major)
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
    if (a[piv * n + c] == 0.0) throw std::runtime_error("singular local block");
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[c * n + k], a[piv * n + k]);
        std::swap(inv[c * n + k], inv[piv * n + k]);
      }
    const double d = a[c * n + c];
    for (int k = 0; k < n; ++k) {
      a[c * n + k] /= d;
      inv[c * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r * n + k] -= f * a[c * n + k];
        inv[r * n + k] -= f * inv[c * n + k];
      }
    }
  }
  a = inv;
}

struct M2 {
  double a[4] = {0, 0, 0, 0};
};
inline M2 mul(const M2 &x, const M2 &y) {
  M2 r;
  r.a[0] = x.a[0] * y.a[0] + x.a[1] * y.a[2];
  r.a[1] = x.a[0] * y.a[1] + x.a[1] * y.a[3];
  r.a[2] = x.a[2] * y.a[0] + x.a[3] * y.a[2];
  r.a[3] = x.a[2] * y.a[1] + x.a[3] * y.a[3];
  return r;
}
inline M2 neg(const M2 &x) {
  M2 r;
  for (int i = 0; i < 4; ++i) r.a[i] = -x.a[i];
  return r;
}
inline M2 inv2(const M2 &x) {
  const double det = x.a[0] * x.a[3] - x.a[1] * x.a[2];
  if (det == 0.0) throw std::runtime_error("singular reduced block");
  M2 r;
  r.a[0] = x.a[3] / det;
  r.a[1] = -x.a[1] / det;
  r.a[2] = -x.a[2] / det;
  r.a[3] = x.a[0] / det;
  return r;
}
inline double maxabs(const M2 &x) {
  double m = 0;
  for (int i = 0; i < 4; ++i) m = std::fmax(m, std::fabs(x.a[i]));
  return m;
}

// Explicit part A (unit spacing) as sparse rows: for row i, (column, coefficient).
inline void explicit_rows(OpKind kind, bool per, int n, double alpha,
                          std::vector<std::vector<std::pair<int, double>>> &A) {
  A.assign(n, {});
  auto add = [&](int i, int j, double c) {
    if (per) j = ((j % n) + n) % n;
    if (j < 0 || j >= n) return;
    A[i].push_back({j, c});
  };
  if (kind == OP_DIFF_O4E) {
    // diff_pen_core_o4_e_c_{p,o}.f90 diff_main: (8 (f+1 - f-1) - (f+2 - f-2)) / 12,
    // one-sided 5-point rows 1, 2, n-1, n on open lines
    const double s = 1.0 / 12.0;
    const int i0 = per ? 0 : 2, i1 = per ? n - 1 : n - 3;
    for (int i = i0; i <= i1; ++i) {
      add(i, i + 1, 8.0 * s);
      add(i, i - 1, -8.0 * s);
      add(i, i + 2, -1.0 * s);
      add(i, i - 2, 1.0 * s);
    }
    if (!per) {
      const double r0[5] = {-25.0, 48.0, -36.0, 16.0, -3.0}, r1[5] = {-3.0, -10.0, 18.0, -6.0, 1.0};
      for (int k = 0; k < 5; ++k) {
        add(0, k, r0[k] * s);
        add(1, k, r1[k] * s);
        add(n - 2, n - 1 - k, -r1[k] * s);
        add(n - 1, n - 1 - k, -r0[k] * s);
      }
    }
  } else if (kind == OP_DIFF_O2E) {
    // diff_pen_core_o2_e_c_o.f90 diff_main (open lines only): (f+1 - f-1) / 2,
    // one-sided rows 1, 2, n-1, n
    for (int i = 2; i <= n - 3; ++i) {
      add(i, i + 1, 0.5);
      add(i, i - 1, -0.5);
    }
    add(0, 0, -1.5); add(0, 1, 2.0); add(0, 2, -0.5);
    add(1, 0, -0.5); add(1, 2, 0.5);
    add(n - 2, n - 3, -0.5); add(n - 2, n - 1, 0.5);
    add(n - 1, n - 3, 0.5); add(n - 1, n - 2, -2.0); add(n - 1, n - 1, 1.5);
  } else if (kind == OP_DIFF_O4P) {
    // diff_pen_core_o4_i_c_{p,o}_pade.f90 diff_main: (f+1 - f-1) / 2, with the
    // one-sided rows (f2 - f1) / 2 and (fn - f(n-1)) / 2 on open lines
    const double a = 0.5;
    const int i0 = per ? 0 : 1, i1 = per ? n - 1 : n - 2;
    for (int i = i0; i <= i1; ++i) {
      add(i, i + 1, a);
      add(i, i - 1, -a);
    }
    if (!per) {
      add(0, 0, -a); add(0, 1, a);
      add(n - 1, n - 1, a); add(n - 1, n - 2, -a);
    }
  } else if (kind == OP_DIFF) {
    // diff_pen_core_o6_i_c_{p,o}.f90 init_right_explicit_diff_part
    const double a = 14.00 / 9.0 / 2.0, b = 1.00 / 9.0 / 4.0;
    const int i0 = per ? 0 : 2, i1 = per ? n - 1 : n - 3;
    for (int i = i0; i <= i1; ++i) {
      add(i, i + 1, a);
      add(i, i - 1, -a);
      add(i, i + 2, b);
      add(i, i - 2, -b);
    }
    if (!per) {
      const double c = 17.00 / 6.0, d = 1.50, e = 1.00 / 6.0, f = 0.75;
      add(0, 0, -c); add(0, 1, d); add(0, 2, d); add(0, 3, -e);
      add(1, 0, -f); add(1, 2, f);
      add(n - 2, n - 1, f); add(n - 2, n - 3, -f);
      add(n - 1, n - 1, c); add(n - 1, n - 2, -d); add(n - 1, n - 3, -d); add(n - 1, n - 4, e);
    }
  } else {
    // filter_pen_core_o6_i_c_{p,o}.f90 init_right_explicit_filter_part
    double a0 = 11.0 / 16.0 + 5.0 * alpha / 8.0;
    double a1 = (15.0 / 32.0 + 17.0 * alpha / 16.0) / 2.0;
    double a2 = (-3.0 / 16.0 + 3.0 * alpha / 8.0) / 2.0;
    double a3 = (1.0 / 32.0 - 1.0 * alpha / 16.0) / 2.0;
    const int i0 = per ? 0 : 3, i1 = per ? n - 1 : n - 4;
    for (int i = i0; i <= i1; ++i) {
      add(i, i, a0);
      add(i, i + 1, a1); add(i, i - 1, a1);
      add(i, i + 2, a2); add(i, i - 2, a2);
      add(i, i + 3, a3); add(i, i - 3, a3);
    }
    if (!per) {
      const double cb[2][7] = {
          {1.0 / 64.0 + 31.0 * alpha / 32.0, 29.0 / 32.0 + 3.0 * alpha / 16.0,
           15.0 / 64.0 + 17.0 * alpha / 32.0, -5.0 / 16.0 + 5.0 * alpha / 8.0,
           15.0 / 64.0 - 15.0 * alpha / 32.0, -3.0 / 32.0 + 3.0 * alpha / 16.0,
           1.0 / 64.0 - 1.0 * alpha / 32.0},
          {-1.0 / 64.0 + 1.0 * alpha / 32.0, 3.0 / 32.0 + 13.0 * alpha / 16.0,
           49.0 / 64.0 + 15.0 * alpha / 32.0, 5.0 / 16.0 + 3.0 * alpha / 8.0,
           -15.0 / 64.0 + 15.0 * alpha / 32.0, 3.0 / 32.0 - 3.0 * alpha / 16.0,
           -1.0 / 64.0 + 1.0 * alpha / 32.0}};
      add(0, 0, 1.0);
      add(n - 1, n - 1, 1.0);
      for (int k = 0; k < 7; ++k) {
        add(1, k, cb[0][k]);
        add(n - 2, n - 1 - k, cb[0][k]);
        add(2, k, cb[1][k]);
        add(n - 3, n - 1 - k, cb[1][k]);
      }
    }
  }
}

// Left matrix B as tridiagonal rows (sub, diag, super); cyclic corners for periodic.
inline void implicit_rows(OpKind kind, bool per, int n, double alpha, std::vector<double> &lo,
                          std::vector<double> &di, std::vector<double> &up) {
  if (kind == OP_DIFF_O4P) {
    // diff_pen_core_o4_i_c_p_pade.f90 (tridiag(1/6, 2/3, 1/6), cyclic) and
    // diff_pen_core_o4_i_c_o_pade.f90 (the same with dc(1) = dc(n) = 1/3)
    lo.assign(n, 1.0 / 6.0);
    di.assign(n, 2.0 / 3.0);
    up.assign(n, 1.0 / 6.0);
    if (!per) {
      lo[0] = 0.0;
      up[n - 1] = 0.0;
      di[0] = 1.0 / 3.0;
      di[n - 1] = 1.0 / 3.0;
    }
    return;
  }
  const bool expl = kind == OP_DIFF_O4E || kind == OP_DIFF_O2E;
  const double al = kind == OP_DIFF ? 1.0 / 3.0 : (expl ? 0.0 : alpha);
  lo.assign(n, al);
  di.assign(n, 1.0);
  up.assign(n, al);
  if (expl) return;  // explicit: B = I
  if (!per) {
    lo[0] = 0.0;
    up[n - 1] = 0.0;
    if (kind == OP_DIFF) {
      // diff_pen_core_o6_i_c_o.f90:187-192 (du(1)=3, du(2)=1/4, du(n-1)=1/4,
      // dl(1)=1/4, dl(n-1)=3, dl(n-2)=1/4 in LAPACK band storage)
      up[0] = 3.00;
      up[1] = 0.25;
      up[n - 2] = 0.25;
      lo[1] = 0.25;
      lo[n - 1] = 3.00;
      lo[n - 2] = 0.25;
    } else {
      // filter_pen_core_o6_i_c_o.f90:232-233 (du(1)=0, dl(n-1)=0)
      up[0] = 0.0;
      lo[n - 1] = 0.0;
    }
  }
}

}  // namespace detail

// Padding in front of an open line of n points (periodic lines: none), and
// the padded length.
inline int front_padding(bool per, int n) {
  if (per || n % PPT == 0) return 0;
  for (int pf = 0; pf <= 4; ++pf)
    if ((pf + n - 1) % PPT >= 3) return pf;
  return 0;  // unreachable: 5 consecutive residues always meet {3..7}
}
inline int padded_length(bool per, int n) {
  return per ? n : (front_padding(per, n) + n + PPT - 1) / PPT * PPT;
}

// Build the block-partitioned operator for a line of n_real points: periodic
// lines need n_real % 8 == 0 and n_real >= 16, open lines n_real >= 9 (padded
// to padded_length(false, n_real) points as described above).
inline HostLineOp build_line_op(OpKind kind, bool per, int n_real, double alpha, double tol = 1e-18) {
  using namespace detail;
  if (per && (n_real % PPT != 0 || n_real < 2 * PPT))
    throw std::runtime_error("periodic line length " + std::to_string(n_real) +
                             " must be a multiple of 8 and at least 16");
  if (!per && n_real < PPT + 1)
    throw std::runtime_error("open line length " + std::to_string(n_real) + " must be at least 9");
  const int n = padded_length(per, n_real);
  const int pf = front_padding(per, n_real);
  HostLineOp op;
  op.n = n;
  op.n_real = n_real;
  op.pf = pf;
  op.NB = n / PPT;
  op.per = per ? 1 : 0;
  const int NB = op.NB;
  op.nw = per ? 1 : NB;
  std::vector<std::vector<std::pair<int, double>>> Ar, A(n);
  explicit_rows(kind, per, n_real, alpha, Ar);
  std::vector<double> lr, dr, ur, lo(n, 0.0), di(n, 1.0), up(n, 0.0);
  implicit_rows(kind, per, n_real, alpha, lr, dr, ur);
  // real rows shifted by pf; decoupled padding rows: B = I, A = 0 (derivative)
  // or I (filter)
  for (int i = 0; i < n_real; ++i) {
    for (auto &jc : Ar[i]) A[i + pf].push_back({jc.first + pf, jc.second});
    lo[i + pf] = lr[i];
    di[i + pf] = dr[i];
    up[i + pf] = ur[i];
  }
  for (int i = 0; i < n; ++i)
    if ((i < pf || i >= pf + n_real) && kind == OP_FILTER) A[i].push_back({i, 1.0});
  // rows that differ from the interior form (the open closures, the padding)
  std::vector<bool> special(n, false);
  if (!per) {
    const int mid = pf + n_real / 2;
    for (int i = 0; i < n; ++i) {
      bool sp = i < pf || i >= pf + n_real || lo[i] != lo[mid] || di[i] != di[mid] || up[i] != up[mid] ||
                A[i].size() != A[mid].size();
      for (size_t k = 0; !sp && k < A[i].size(); ++k)
        sp = A[i][k].first - i != A[mid][k].first - mid || A[i][k].second != A[mid][k].second;
      special[i] = sp;
    }
  }
  for (int b = 1; b + 2 < NB; ++b)
    for (int r = 0; r < PPT; ++r)
      if (special[b * PPT + r]) throw std::runtime_error("closure rows beyond the two end blocks");
  op.spec2 = 0;
  if (!per && NB >= 3)
    for (int r = 0; r < PPT; ++r) op.spec2 |= special[(NB - 2) * PPT + r] ? 1 : 0;

  op.I.assign(NI, 0.0);
  op.M.assign(NTYPES * PPT * FE, 0.0);
  op.L.assign(NTYPES * PPT, 0.0);
  op.R.assign(NTYPES * PPT, 0.0);
  std::vector<double> Mint(PPT * FE), Lint(PPT), Rint(PPT);
  std::vector<M2> Pm(NB), Qm(NB);
  for (int b = 0; b < NB; ++b) {
    const int p0 = b * PPT;
    // local block T_b and its inverse
    std::vector<double> T(PPT * PPT, 0.0);
    for (int r = 0; r < PPT; ++r) {
      T[r * PPT + r] = di[p0 + r];
      if (r > 0) T[r * PPT + r - 1] = lo[p0 + r];
      if (r < PPT - 1) T[r * PPT + r + 1] = up[p0 + r];
    }
    inverse(T, PPT);
    // couplings to the neighbouring blocks
    const double aL = (per || b > 0) ? lo[p0] : 0.0;
    const double cR = (per || b < NB - 1) ? up[p0 + PPT - 1] : 0.0;
    double Lb[PPT], Rb[PPT];
    for (int r = 0; r < PPT; ++r) {
      Lb[r] = -aL * T[r * PPT + 0];
      Rb[r] = -cR * T[r * PPT + PPT - 1];
    }
    // local explicit operator on the extended block (columns p0-3 .. p0+10)
    double Aloc[PPT][FE] = {};
    for (int r = 0; r < PPT; ++r)
      for (auto &jc : A[p0 + r]) {
        int off = jc.first - (p0 + r);
        if (per) off = ((off + n / 2) % n + n) % n - n / 2;
        const int jj = r + off + HALO;
        if (jj < 0 || jj >= FE) throw std::runtime_error("explicit stencil wider than the halo");
        Aloc[r][jj] += jc.second;
      }
    // block type: first, interior, last or the special block before the
    // last (a periodic line has one type)
    const int ty = per ? -1 : (b == 0 ? 0 : (b == NB - 1 ? 2 : (b == NB - 2 && op.spec2 ? 3 : 1)));
    double Mb[PPT * FE];
    for (int r = 0; r < PPT; ++r)
      for (int j = 0; j < FE; ++j) {
        double s = 0.0;
        for (int k = 0; k < PPT; ++k) s += T[r * PPT + k] * Aloc[k][j];
        Mb[r * FE + j] = s;
      }
    for (int t = 0; t < NTYPES; ++t) {
      const bool take = per ? (b == 0) : (t == ty && (t != 1 || b == 1));
      if (!take) continue;
      for (int r = 0; r < PPT; ++r) {
        for (int j = 0; j < FE; ++j) op.M[(t * PPT + r) * FE + j] = Mb[r * FE + j];
        op.L[t * PPT + r] = Lb[r];
        op.R[t * PPT + r] = Rb[r];
      }
    }
    if (!per && ty == 1) {
      // every interior block must share the type-1 operator
      if (b == 1) {
        for (int k = 0; k < PPT * FE; ++k) Mint[k] = Mb[k];
        for (int k = 0; k < PPT; ++k) { Lint[k] = Lb[k]; Rint[k] = Rb[k]; }
      } else {
        for (int k = 0; k < PPT * FE; ++k)
          if (std::fabs(Mint[k] - Mb[k]) > 1e-14 * (1.0 + std::fabs(Mb[k])))
            throw std::runtime_error("interior blocks differ");
      }
    }
    // reduced system  z_b + P_b z_{b-1} + Q_b z_{b+1} = Y_b,  z = (x_first, x_last)
    Pm[b].a[0] = 0.0; Pm[b].a[1] = -Lb[0]; Pm[b].a[2] = 0.0; Pm[b].a[3] = -Lb[PPT - 1];
    Qm[b].a[0] = -Rb[0]; Qm[b].a[1] = 0.0; Qm[b].a[2] = -Rb[PPT - 1]; Qm[b].a[3] = 0.0;
  }
  // interior-block stencil and Thomas factors (rows of a block away from the
  // ends); an open line of two blocks has no interior block
  if (per || NB > 2) {
    const int ri = per ? 0 : PPT + HALO;  // a row of an interior block
    double c[7] = {0, 0, 0, 0, 0, 0, 0};
    for (auto &jc : A[ri]) {
      int off = jc.first - ri;
      if (per) off = ((off + n / 2) % n + n) % n - n / 2;
      if (off < -3 || off > 3) throw std::runtime_error("interior stencil wider than 3");
      c[off + 3] += jc.second;
    }
    const double sgn = (c[4] == 0.0 && c[2] == 0.0) ? 1.0 : (c[2] == -c[4] ? -1.0 : 1.0);
    for (int k = 1; k <= 3; ++k)
      if (std::fabs(c[3 + k] - (sgn > 0 ? c[3 - k] : -c[3 - k])) > 1e-15 * (1 + std::fabs(c[3 + k])))
        throw std::runtime_error("interior stencil is neither symmetric nor antisymmetric");
    if (sgn < 0 && c[3] != 0.0) throw std::runtime_error("antisymmetric stencil with a centre term");
    op.I[0] = c[3];
    op.I[1] = c[4];
    op.I[2] = c[5];
    op.I[3] = c[6];
    op.I[4] = sgn;
    const double a = lo[ri], u = up[ri], dg = di[ri];
    double d = dg;
    op.I[12] = 1.0 / d;
    for (int i = 1; i < PPT; ++i) {
      const double l = a / d;
      d = dg - l * u;
      op.I[4 + i] = l;
      op.I[12 + i] = 1.0 / d;
    }
    op.I[20] = u;
  }
  // parallel cyclic reduction over blocks
  int s = 1;
  for (;;) {
    if (op.K >= MAXK) throw std::runtime_error("reduced system did not converge");
    std::vector<M2> nP(NB), nQ(NB);
    std::vector<double> Wk(op.nw * 12, 0.0);
    for (int b = 0; b < NB; ++b) {
      const int bm = ((b - s) % NB + NB) % NB, bp = (b + s) % NB;
      const bool hasm = per || b - s >= 0, hasp = per || b + s < NB;
      M2 Pb = hasm ? Pm[b] : M2(), Qb = hasp ? Qm[b] : M2();
      M2 S;
      S.a[0] = 1.0; S.a[3] = 1.0;
      M2 t1 = mul(Pb, Qm[bm]), t2 = mul(Qb, Pm[bp]);
      for (int i = 0; i < 4; ++i) S.a[i] -= t1.a[i] + t2.a[i];
      M2 Si = inv2(S);
      M2 W0 = Si, Wm = neg(mul(Si, Pb)), Wp = neg(mul(Si, Qb));
      if (b < op.nw)
        for (int i = 0; i < 4; ++i) {
          Wk[b * 12 + i] = W0.a[i];
          Wk[b * 12 + 4 + i] = Wm.a[i];
          Wk[b * 12 + 8 + i] = Wp.a[i];
        }
      nP[b] = neg(mul(Si, mul(Pb, Pm[bm])));
      nQ[b] = neg(mul(Si, mul(Qb, Qm[bp])));
      if (!per && b - 2 * s < 0) nP[b] = M2();
      if (!per && b + 2 * s >= NB) nQ[b] = M2();
    }
    op.W.insert(op.W.end(), Wk.begin(), Wk.end());
    op.K++;
    Pm = nP;
    Qm = nQ;
    s *= 2;
    double cpl = 0.0;
    for (int b = 0; b < NB; ++b) cpl = std::fmax(cpl, std::fmax(maxabs(Pm[b]), maxabs(Qm[b])));
    if (cpl < tol) break;
    if (!per && s >= NB) break;
  }
  return op;
}

}  // namespace cruna
