// CPU emulation of the device line engine (public_amd/csrc/line_solve.hpp)
// driven by the host tables of public_amd/csrc/line_ops.hpp, compared with the
// oracle's LU-based cores (oracle/cruna_oracle.c) on random lines.
// Open lines of any length run as padded lines (decoupled rows beyond n).
// Prints one line per case: kind scheme n K max_rel_err.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../oracle/cruna_oracle.h"
#include "../../public_amd/csrc/line_ops.hpp"

using namespace cruna;

static int block_type(const HostLineOp &op, int b) {
  if (op.per) return 1;
  const int NB = op.NB;
  return b == 0 ? 0 : (b == NB - 1 ? 2 : (b == NB - 2 && op.spec2 ? 3 : 1));
}

// f holds the real points; the padding points of an open line are zero
// (op.pf of them in front); the result is in the padded layout
static std::vector<double> emulate(const HostLineOp &op, const std::vector<double> &f_real) {
  const int n = op.n, NB = op.NB;
  std::vector<double> f(n, 0.0);
  for (size_t i = 0; i < f_real.size(); ++i) f[i + op.pf] = f_real[i];
  std::vector<double> y(n), x(n);
  for (int b = 0; b < NB; ++b) {
    double fe[FE];
    for (int j = 0; j < FE; ++j) {
      int p = b * PPT - HALO + j;
      if (op.per) p = ((p % n) + n) % n;
      fe[j] = (p >= 0 && p < n) ? f[p] : 0.0;
    }
    const bool bnd = block_type(op, b) != 1;
    if (!bnd) {
      const double *I = op.I.data();
      double r[PPT];
      for (int i = 0; i < PPT; ++i) {
        const double d1 = fe[HALO + i + 1] + I[4] * fe[HALO + i - 1];
        const double d2 = fe[HALO + i + 2] + I[4] * fe[HALO + i - 2];
        const double d3 = fe[HALO + i + 3] + I[4] * fe[HALO + i - 3];
        r[i] = I[0] * fe[HALO + i] + I[1] * d1 + I[2] * d2 + I[3] * d3;
      }
      for (int i = 1; i < PPT; ++i) r[i] = r[i] - I[4 + i] * r[i - 1];
      r[PPT - 1] = r[PPT - 1] * I[12 + PPT - 1];
      for (int i = PPT - 2; i >= 0; --i) r[i] = (r[i] - I[20] * r[i + 1]) * I[12 + i];
      for (int i = 0; i < PPT; ++i) y[b * PPT + i] = r[i];
    } else {
      const int ty = block_type(op, b);
      for (int r = 0; r < PPT; ++r) {
        double acc = 0.0;
        for (int j = 0; j < FE; ++j) acc += op.M[(ty * PPT + r) * FE + j] * fe[j];
        y[b * PPT + r] = acc;
      }
    }
  }
  std::vector<double> z0(NB), z1(NB);
  for (int b = 0; b < NB; ++b) {
    z0[b] = y[b * PPT];
    z1[b] = y[b * PPT + PPT - 1];
  }
  for (int k = 0; k < op.K; ++k) {
    const int st = 1 << k;
    std::vector<double> n0(NB), n1(NB);
    for (int b = 0; b < NB; ++b) {
      const bool hm = op.per || b - st >= 0, hp = op.per || b + st < NB;
      const int bm = ((b - st) % NB + NB) % NB, bp = (b + st) % NB;
      const double m0 = hm ? z0[bm] : 0, m1 = hm ? z1[bm] : 0, p0 = hp ? z0[bp] : 0, p1 = hp ? z1[bp] : 0;
      const double *w = &op.W[(k * op.nw + (op.per ? 0 : b)) * 12];
      n0[b] = w[0] * z0[b] + w[1] * z1[b] + w[4] * m0 + w[5] * m1 + w[8] * p0 + w[9] * p1;
      n1[b] = w[2] * z0[b] + w[3] * z1[b] + w[6] * m0 + w[7] * m1 + w[10] * p0 + w[11] * p1;
    }
    z0 = n0;
    z1 = n1;
  }
  for (int b = 0; b < NB; ++b) {
    const bool hl = op.per || b > 0, hr = op.per || b < NB - 1;
    const double xl = hl ? z1[(b - 1 + NB) % NB] : 0, xf = hr ? z0[(b + 1) % NB] : 0;
    const int ty = block_type(op, b);
    for (int r = 0; r < PPT; ++r)
      x[b * PPT + r] = y[b * PPT + r] + op.L[ty * PPT + r] * xl + op.R[ty * PPT + r] * xf;
  }
  return x;
}

int main() {
  std::mt19937_64 rng(7);
  std::normal_distribution<double> nd;
  const int ns[] = {16, 24, 48, 64, 96, 128, 256, 512};
  // open lines of other lengths (padded): 2^k + 1 and the reference's
  // regression geometry 60 x 65 x 70
  const int ns_open[] = {9, 13, 17, 20, 33, 60, 65, 70, 129, 257};
  const OpKind kinds[] = {OP_DIFF, OP_FILTER, OP_DIFF_O4E, OP_DIFF_O4P, OP_DIFF_O2E};
  const char *names[] = {"diff", "filter", "diffo4", "diffpade", "diffo2"};
  std::vector<std::pair<int, int>> cases;  // (periodic, n)
  for (int per = 1; per >= 0; --per)
    for (int n : ns) cases.push_back({per, n});
  for (int n : ns_open) cases.push_back({0, n});
  for (int kind = 0; kind < 5; ++kind)
    for (auto pc : cases) {
        const int per = pc.first, n = pc.second;
        if (kinds[kind] == OP_DIFF_O2E && per) continue;  // open lines only
        const double alpha = 0.35;
        HostLineOp op = build_line_op(kinds[kind], per, n, alpha);
        double worst = 0.0;
        for (int t = 0; t < 4; ++t) {
          std::vector<double> f(n);
          for (auto &v : f) v = nd(rng);
          std::vector<double> ref = f;
          const int sch = per ? ORC_PERIODIC : ORC_OPEN;
          if (kind == 0)
            orc_diff_core(sch, n, 1, ref.data(), 1.0);
          else if (kind == 1)
            orc_filter_core(sch, n, 1, ref.data(), alpha);
          else if (kind == 2)
            orc_diff_core_o4(sch, n, 1, ref.data(), 1.0);
          else if (kind == 3)
            orc_diff_core_pade(sch, n, 1, ref.data(), 1.0);
          else
            orc_diff_core_o2(n, 1, ref.data(), 1.0);
          std::vector<double> x = emulate(op, f);
          double num = 0, den = 0;
          for (int i = 0; i < n; ++i) {
            num = std::fmax(num, std::fabs(x[i + op.pf] - ref[i]));
            den = std::fmax(den, std::fabs(ref[i]));
          }
          // the padding rows stay decoupled: zero derivative, unfiltered zero
          for (int i = 0; i < op.n; ++i)
            if (i < op.pf || i >= op.pf + n) num = std::fmax(num, std::fabs(x[i]));
          worst = std::fmax(worst, num / den);
        }
        std::printf("%s %s %d %d %.3e\n", names[kind], per ? "periodic" : "open", n,
                    op.K, worst);
      }
  return 0;
}
