// Device side of the block-partitioned line solve (tables: line_ops.hpp).
//
// A work-group owns TL complete grid lines along one direction; each thread
// owns one block of PPT = 8 consecutive points of one line (register
// resident).  Cross-thread traffic goes through LDS only for the explicit
// stencil halo (3 values per side) and the 2-unknown-per-block reduced
// system, so a derivative costs ~4 LDS accesses per point instead of the
// ~30 of a point-wise parallel cyclic reduction.  On periodic lines every
// table entry is the same for all threads and is read with scalar loads.
#pragma once

#include <hip/hip_runtime.h>

#include "line_ops.hpp"

namespace cruna {

constexpr int NT = 512;  // max threads per work-group of the RHS sweeps
// the FP32 filter's and Cartesian sweeps' work-groups take up to 32 y/z lines
// (128-byte rows of floats)
constexpr int NT_FILTER32 = 1024;
constexpr int NT_SWEEP32 = 1024;

template <typename T>
struct LineOp {
  const T *I;  // [NI] interior blocks: stencil + local Thomas factors
  const T *M;  // [NTYPES][PPT][FE]  dense local operators: first, interior, last, before last
  const T *L;  // [NTYPES][PPT]
  const T *R;  // [NTYPES][PPT]
  const T *W;  // [K][nw][12]
  int n, NB, K, per;
  int spec2;   // open line: block NB-2 uses its own dense operator (type 3)
};

// Sweep geometry: which lines a work-group owns and where they live in HBM.
struct SweepGeom {
  int dir;        // 0: x (contiguous lines), 1: y, 2: z
  int n;          // line length
  int NB;         // blocks per line (n / PPT)
  int TL;         // lines per work-group
  int n1, n2, n3;
  long PS;        // distance between consecutive line points in memory
  int ntiles;     // work-groups in the grid
  int ntiles_in;  // (y/z) tiles along i
  long lines;     // batched launches: lines of the direction, taken TL at a time
};

// Thread placement inside a tile.  The LDS slot of a (line, block) pair is the
// thread index; BS / LS are the slot strides between blocks / lines.
struct ThreadPos {
  int l, b;        // line in tile, block in line
  int BS, LS;
  long base;       // element offset of this thread's first point
  long lidx;       // index of the line in the plane across it: j + n2 k (x), i + n1 k (y), i + n1 j (z)
  bool wave_line;  // every line of the tile lies inside one wavefront
};

// Work-groups are dealt round-robin to the 8 XCDs; map consecutive tiles to the
// same XCD so that tiles sharing cache lines meet in the same L2.
__device__ inline int xcd_tile(int bid, int ntiles) {
  if (ntiles % 8 != 0) return bid;
  return (bid % 8) * (ntiles / 8) + bid / 8;
}

// Element offset of an idle thread's block: 2^28 FP64 elements is the 2 GiB
// buffer extent of the sweeps' field accesses (the host keeps every field
// below it), so an idle thread's loads return zero and its stores are dropped
// by the buffer range check.  FP64 fields only (the batched launches).
constexpr long IDLE_BASE = 1L << 28;

// Placement in tile `tile` of a sweep of geometry g (the launch has exactly
// TL x NB threads per work-group).
// BATCH (a batched launch over members of several sizes): tile t holds lines
// t*TL .. t*TL+TL-1 of the direction in line order (x: j + n2 k; y: i + n1 k;
// z: i + n1 j), so that a tile of any TL fills its work-group; the work-group
// may be larger than the tile, and the threads beyond it and the lines past
// the member's last line are idle: their block lies at IDLE_BASE and their
// exchange slots are those of block 0 of line 0, which they only read.
template <int DIR, bool BATCH = false>
__device__ inline ThreadPos thread_pos_tile(const SweepGeom &g, int tile) {
  ThreadPos t;
  const int tid = threadIdx.x;
  if (DIR == 0) {
    t.b = tid % g.NB;
    t.l = tid / g.NB;
    t.BS = 1;
    t.LS = g.NB;
    // x-lines take NB consecutive threads: with NB dividing the wavefront
    // size no line straddles two wavefronts
    t.wave_line = (64 % g.NB) == 0;
  } else {
    t.l = tid % g.TL;
    t.b = tid / g.TL;
    t.BS = g.TL;
    t.LS = 1;
    t.wave_line = false;
  }
  const long line = (long)tile * g.TL + t.l;
  if (BATCH && (tid >= g.TL * g.NB || line >= g.lines)) {
    t.l = 0;
    t.b = 0;
    t.base = IDLE_BASE;
    t.lidx = 0;
    return t;
  }
  if (DIR == 0) {
    t.base = line * g.n1 + (long)t.b * PPT;
    t.lidx = line;
  } else {
    long i, o;  // o: k for y-lines, j for z-lines
    if (BATCH) {
      // a member's lines are fewer than 2^31
      i = int(line) % g.n1;
      o = int(line) / g.n1;
    } else {
      i = (long)(tile % g.ntiles_in) * g.TL + t.l;
      o = tile / g.ntiles_in;
    }
    const long off = (DIR == 1) ? i + (long)g.n1 * g.n2 * o : i + (long)g.n1 * o;
    t.base = off + (long)t.b * PPT * g.PS;
    t.lidx = i + (long)g.n1 * o;
  }
  return t;
}

template <int DIR>
__device__ inline ThreadPos thread_pos(const SweepGeom &g) {
  return thread_pos_tile<DIR>(g, xcd_tile(blockIdx.x, g.ntiles));
}

// Work-group barrier that orders LDS traffic only.  __syncthreads() also waits
// for every outstanding global load, which would drain the prefetches the
// sweeps issue ahead of a line solve; this barrier leaves them in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Synchronisation point of the line engine: when each line lives inside one
// wavefront, the LDS exchange only has to be ordered within the wavefront
// (its LDS operations complete in issue order), so no work-group barrier is
// needed.
__device__ __forceinline__ void line_sync(bool wave_line) {
  if (wave_line) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  } else {
    lds_barrier();
  }
}

// LDS scratch of one work-group for the line engine, for up to NTM threads,
// indexed [value][thread slot]: consecutive lanes touch consecutive words
// (a thread-major [slot][value] layout put lanes 6 or 2 values apart, which
// the bank-conflict counters showed stalling every exchange).
template <typename T, int NTM = NT>
struct LineLds {
  T halo[2 * HALO][NTM];
  T z[2][2][NTM];
};

// x = B^-1 A f on the thread's block (unit spacing; the caller scales by 1/h).
// All threads of the work-group must call it together (barriers inside).
// Opaque copy of a uniform pointer: keeps the compiler from hoisting the table
// reads of one solve out of the surrounding sweep (which would pin ~130 table
// values in registers for the whole kernel); each solve re-reads them through
// the scalar cache instead.
// Tables are never written while a kernel runs, so uniform table reads may go
// through the scalar data cache: a pointer into the constant address space
// tells the compiler so.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> launder_uniform(const T *p) {
  asm volatile("" : "+s"(p));
  return (cptr<T>)(p);
}

template <typename T, bool PER, int NTM>
__device__ __forceinline__ void line_solve(const LineOp<T> &op, LineLds<T, NTM> &s, const ThreadPos &tp,
                                           const T (&f)[PPT], T (&x)[PPT]) {
  const int tid = threadIdx.x;
  const int NB = op.NB;
  const int b = tp.b;
  // 1. halo exchange
#pragma unroll
  for (int m = 0; m < HALO; ++m) {
    s.halo[m][tid] = f[m];
    s.halo[HALO + m][tid] = f[PPT - HALO + m];
  }
  line_sync(tp.wave_line);
  const bool hasL = PER || b > 0;
  const bool hasR = PER || b < NB - 1;
  const int bl = b == 0 ? NB - 1 : b - 1;
  const int br = b == NB - 1 ? 0 : b + 1;
  const int sl = tp.l * tp.LS + bl * tp.BS;
  const int sr = tp.l * tp.LS + br * tp.BS;
  T fe[FE];
#pragma unroll
  for (int m = 0; m < HALO; ++m) {
    fe[m] = hasL ? s.halo[HALO + m][sl] : T(0);
    fe[HALO + PPT + m] = hasR ? s.halo[m][sr] : T(0);
  }
#pragma unroll
  for (int m = 0; m < PPT; ++m) fe[HALO + m] = f[m];
  // 2. local operator y = T_b^-1 A_b f: explicit stencil and an 8-point
  //    Thomas solve for interior blocks (every block of a periodic line); the
  //    two end blocks of an open line use their dense operator instead.
  T y[PPT];
  int ty = 1;
  {
    const cptr<T> I = launder_uniform(op.I);
    const T c0 = I[0], c1 = I[1], c2 = I[2], c3 = I[3], sg = I[4];
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const T d1 = fe[HALO + i + 1] + sg * fe[HALO + i - 1];
      const T d2 = fe[HALO + i + 2] + sg * fe[HALO + i - 2];
      const T d3 = fe[HALO + i + 3] + sg * fe[HALO + i - 3];
      y[i] = c0 * fe[HALO + i] + c1 * d1 + c2 * d2 + c3 * d3;
    }
#pragma unroll
    for (int i = 1; i < PPT; ++i) y[i] = y[i] - I[4 + i] * y[i - 1];
    const T up = I[20];
    y[PPT - 1] = y[PPT - 1] * I[12 + PPT - 1];
#pragma unroll
    for (int i = PPT - 2; i >= 0; --i) y[i] = (y[i] - up * y[i + 1]) * I[12 + i];
  }
  if (!PER) {
    // the end blocks of an open line: one pass per block type present in the
    // wavefront, with that type's operator read through the scalar cache
    // (uniform), so that its 14 coefficients per row take no vector registers
    ty = b == 0 ? 0 : (b == NB - 1 ? 2 : ((b == NB - 2 && op.spec2) ? 3 : 1));
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t == 1) continue;
      if (__builtin_amdgcn_ballot_w64(ty == t) == 0) continue;
      const cptr<T> Mb = launder_uniform(op.M + t * PPT * FE);
#pragma unroll 1
      for (int r = 0; r < PPT; ++r) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < FE; ++j) acc += Mb[r * FE + j] * fe[j];
#pragma unroll
        for (int q = 0; q < PPT; ++q)
          if (q == r && ty == t) y[q] = acc;
      }
    }
  }
  // 3. reduced 2x2 block system by cyclic reduction over blocks
  T z0 = y[0], z1 = y[PPT - 1];
  int cur = 0;
  s.z[0][0][tid] = z0;
  s.z[0][1][tid] = z1;
  for (int k = 0; k < op.K; ++k) {
    line_sync(tp.wave_line);
    const int st = 1 << k;
    const bool hm = PER || b - st >= 0;
    const bool hp = PER || b + st < NB;
    int bm = b - st, bp = b + st;
    if (PER) {
      const int sr = st % NB;  // wave-uniform
      bm = b - sr;
      bm = bm < 0 ? bm + NB : bm;
      bp = b + sr;
      bp = bp >= NB ? bp - NB : bp;
    } else {
      bm = hm ? bm : b;
      bp = hp ? bp : b;
    }
    const int sm = tp.l * tp.LS + bm * tp.BS;
    const int sp = tp.l * tp.LS + bp * tp.BS;
    const T m0 = hm ? s.z[cur][0][sm] : T(0), m1 = hm ? s.z[cur][1][sm] : T(0);
    const T p0 = hp ? s.z[cur][0][sp] : T(0), p1 = hp ? s.z[cur][1][sp] : T(0);
    T w[12];
    if (PER) {
      const cptr<T> opW = launder_uniform(op.W);
#pragma unroll
      for (int i = 0; i < 12; ++i) w[i] = opW[k * 12 + i];
    } else {
#pragma unroll
      for (int i = 0; i < 12; ++i) w[i] = op.W[(k * NB + b) * 12 + i];
    }
    const T n0 = w[0] * z0 + w[1] * z1 + w[4] * m0 + w[5] * m1 + w[8] * p0 + w[9] * p1;
    const T n1 = w[2] * z0 + w[3] * z1 + w[6] * m0 + w[7] * m1 + w[10] * p0 + w[11] * p1;
    z0 = n0;
    z1 = n1;
    cur ^= 1;
    s.z[cur][0][tid] = z0;
    s.z[cur][1][tid] = z1;
  }
  line_sync(tp.wave_line);
  const T xl = hasL ? s.z[cur][1][sl] : T(0);
  const T xf = hasR ? s.z[cur][0][sr] : T(0);
  if (PER || ty == 1) {
    const cptr<T> opL = launder_uniform(op.L);
    const cptr<T> opR = launder_uniform(op.R);
#pragma unroll
    for (int r = 0; r < PPT; ++r) x[r] = y[r] + opL[PPT + r] * xl + opR[PPT + r] * xf;
  } else {
#pragma unroll
    for (int r = 0; r < PPT; ++r) x[r] = y[r];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (__builtin_amdgcn_ballot_w64(ty == t) == 0) continue;
      const cptr<T> L = launder_uniform(op.L + t * PPT);
      const cptr<T> R = launder_uniform(op.R + t * PPT);
#pragma unroll
      for (int r = 0; r < PPT; ++r) {
        const T v = y[r] + L[r] * xl + R[r] * xf;
        if (ty == t) x[r] = v;
      }
    }
  }
}

}  // namespace cruna
