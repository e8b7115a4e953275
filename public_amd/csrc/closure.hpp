// Learned subfilter closure: a small 3-D CNN on a periodic S^3 grid, trained
// and evaluated with BF16 MFMA (FP32 accumulation).
//
// Network: 5 -> W -> W -> W -> 5 channels, 3x3x3 kernels with periodic
// padding, GELU (tanh form) after the three hidden layers.  Activations are
// stored channels-last, act[p][C] with p = i + S (j + S k) (+ S^3 b), and the
// channel counts are padded to multiples of 16 (zero weights and zero inputs
// in the padding), so every convolution is an implicit GEMM of 16x16 tiles:
//   forward       out[p][co] = b[co] + sum_t sum_ci in[p + o_t][ci] W[t][co][ci]
//   input grad    din[p][ci] = gelu'(z[p][ci]) sum_t sum_co dz[p - o_t][co] W[t][co][ci]
//   weight grad   dW[t][co][ci] = sum_p dz[p][co] in[p + o_t][ci]
// MFMA operand layouts (lane l of 64; both checked on the device):
// v_mfma_f32_16x16x16_bf16: A row l%16 holds K = 4(l/16)..+3, B column l%16
// holds K = 4(l/16)..+3; v_mfma_f32_16x16x32_bf16: the same with
// K = 8(l/16)..+7; both: D row 4(l/16)+r, column l%16, r = 0..3.
#pragma once

#include <hip/hip_runtime.h>

namespace cruna {

typedef unsigned short bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ bf16 to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((unsigned)h << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = c (z + a z^3), written with the
// logistic function s = 1 / (1 + exp(-2u)) (1 + tanh(u) = 2 s,
// 1 - tanh(u)^2 = 4 s (1 - s)) on the hardware exp2 and reciprocal:
//   gelu(z) = z s,   gelu'(z) = s + 2 z s (1 - s) c (1 + 3 a z^2)
__device__ __forceinline__ float gelu_sigma(float z) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float u = c * (z + a * z * z * z);
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-2.8853900817779268f * u));  // 2 log2(e)
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float s = gelu_sigma(z);
  return s + 2.f * z * s * (1.f - s) * c * (1.f + 3.f * a * z * z);
}

// Geometry of a batch of boxes X x Y x Z, stored one after the other with
// p = i + X (j + Y k) + XYZ b: periodic boxes (a full field, wrapped at its
// faces) or crops whose outside reads as zero (a crop with a halo as wide as
// the network's reach gives exactly the periodic field's values at its
// centre).  Coordinates are found once per point with exact FP64 reciprocal
// division ((p + 1/2) / C is never within 1/(2C) of an integer).
struct ConvGeom {
  int X, Y, Z, periodic;
  double inv_x, inv_xy, inv_xyz;
};

struct Coord {
  int i, j, k;
  long base;  // index of the box's first point
};

__device__ __forceinline__ int exact_div(long p, double inv) { return (int)(((double)p + 0.5) * inv); }

__device__ __forceinline__ Coord coord_of(long p, const ConvGeom &g) {
  const long XY = (long)g.X * g.Y, XYZ = XY * g.Z;
  const int b = exact_div(p, g.inv_xyz);
  const long r = p - (long)b * XYZ;
  const int k = exact_div(r, g.inv_xy);
  const long r2 = r - (long)k * XY;
  const int j = exact_div(r2, g.inv_x);
  return Coord{(int)(r2 - (long)j * g.X), j, k, (long)b * XYZ};
}

// neighbour of a point at tap t (t = 9 (dz+1) + 3 (dy+1) + (dx+1)), sign -1
// for the transposed operator; -1 when it falls outside a crop
__device__ __forceinline__ long tap_neighbour(const Coord &c, int t, int sign, const ConvGeom &g) {
  int i = c.i + sign * (t % 3 - 1), j = c.j + sign * ((t / 3) % 3 - 1), k = c.k + sign * (t / 9 - 1);
  if (g.periodic) {
    i = i < 0 ? i + g.X : (i >= g.X ? i - g.X : i);
    j = j < 0 ? j + g.Y : (j >= g.Y ? j - g.Y : j);
    k = k < 0 ? k + g.Z : (k >= g.Z ? k - g.Z : k);
  } else if (i < 0 || i >= g.X || j < 0 || j >= g.Y || k < 0 || k >= g.Z) {
    return -1;
  }
  return c.base + i + (long)g.X * (j + (long)g.Y * k);
}

enum ConvMode { CONV_GELU = 0, CONV_LINEAR = 1, CONV_BACKWARD = 2 };

// Per-point neighbour offsets along one direction: off[0..2] for d = -1, 0,
// +1 (wrapped on periodic boxes) and a validity bit per d (crops).
struct AxisOffsets {
  int off[3];
  unsigned valid;  // bit d
};
__device__ __forceinline__ AxisOffsets axis_offsets(int i, int X, int stride, bool periodic, int sign) {
  AxisOffsets a;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int ii = i + sign * (d - 1);
    bool ok = true;
    if (ii < 0) {
      ok = periodic;
      ii += X;
    } else if (ii >= X) {
      ok = periodic;
      ii -= X;
    }
    a.off[d] = (ii - i) * stride;
  }
  a.valid = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int ii = i + sign * (d - 1);
    if (periodic || (ii >= 0 && ii < X)) a.valid |= 1u << d;
  }
  return a;
}

// a[d] for a runtime d in 0..2 without indexing a register array
__device__ __forceinline__ int pick3(const int (&a)[3], int d) { return d == 0 ? a[0] : (d == 1 ? a[1] : a[2]); }

// One convolution as an implicit GEMM on BF16 MFMA: out[p][co] over the 27
// taps and CK input channels of each point's neighbours.
// With 32 input channels a lane loads 16 bytes (channels 8(l/16)..+7) of a
// neighbour row per tap, the K = 32 operand of one v_mfma_f32_16x16x32_bf16;
// with 16 it loads 8 bytes for v_mfma_f32_16x16x16_bf16.
// CK: input (reduction) channels, CN: output channels.
//   CONV_GELU     : z = conv + bias; zsave <- bf16(z) (if set); out16 <- bf16(gelu(z))
//   CONV_LINEAR   : out32 <- conv + bias
//   CONV_BACKWARD : taps mirrored, W given as [t][CN][CK] of the transposed
//                   operator; out16 <- bf16(conv * gelu'(zin)) (zin: the
//                   pre-activations of this layer's input, CN channels)
// npts is a multiple of 64 and npts * 32 < 2^31 (checked by the host).
constexpr int CONV_TILES = 4;      // 16-point MFMA tiles per wavefront
constexpr int CONV_XCDS = 8;       // XCDs of the MI355X, each with its own L2
// The convolution with weights and input served from LDS: a work-group
// stages the layer's weights in LDS once and then walks tiles of 16 x 4 x 8 output points of one
// box (x, y, z; partial tiles at the box's far faces), loads each tile's
// input with a one-point halo, 18 x 6 x 10 points of CK channels, into LDS
// once (wrapped on periodic boxes, zero outside a crop), and takes all 27
// taps from there: each input row is read from L2 once per tile instead of
// once per tap; the next tile's rows are in flight (in registers) while a
// tile computes.  Wavefront w owns z-plane w of the tile and its 4 MFMA
// tiles are the plane's 4 x-rows of 16 points; it reuses every weight
// fragment for the 4 tiles, and each accumulator takes the taps in the
// order t = 9 dz + 3 dy + dx.  The work-groups are persistent: work-group g
// takes a contiguous range of tiles, the ranges of the work-groups on one
// XCD (g % 8) adjacent, so that neighbouring tiles share their halo rows in
// that XCD's L2.
constexpr int CT_X = 16, CT_Y = 4, CT_Z = 8;
constexpr int CT_THREADS = 64 * CT_Z;  // one wavefront per z-plane of the tile
constexpr int CH_X = CT_X + 2, CH_Y = CT_Y + 2, CH_Z = CT_Z + 2;
constexpr int CT_HALO = CH_X * CH_Y * CH_Z;  // input points per tile

template <int CK, int CN, int MODE>
__global__ __launch_bounds__(CT_THREADS) void k_conv_tile(const bf16 *__restrict__ in, const bf16 *__restrict__ W,
                                                   const float *__restrict__ bias, ConvGeom geo, long npts,
                                                   float *__restrict__ out32, bf16 *__restrict__ out16,
                                                   bf16 *__restrict__ zsave, const bf16 *__restrict__ zin) {
  // weight rows of CK channels: a fragment read (16 rows, 4 lane groups)
  // covers one contiguous run of LDS, free of bank conflicts
  __shared__ __align__(16) bf16 wl[27 * CN * CK];
  __shared__ __align__(16) bf16 xt[CT_HALO * CK];
  {
    constexpr int PIECES = 27 * CN * CK / 8, BATCH = 4;
    for (int base = 0; base < PIECES; base += CT_THREADS * BATCH) {
      uint4 v[BATCH];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        const int e = base + threadIdx.x + CT_THREADS * u;
        v[u] = e < PIECES ? *reinterpret_cast<const uint4 *>(W + (long)e * 8) : uint4{0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        const int e = base + threadIdx.x + CT_THREADS * u;
        if (e < PIECES) *reinterpret_cast<uint4 *>(wl + (long)e * 8) = v[u];
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kg = lane >> 4;
  const bool per = geo.periodic != 0;
  const int X = geo.X, Y = geo.Y, Z = geo.Z;
  const long XY = (long)X * Y, XYZ = XY * Z;
  const int nbx = (X + CT_X - 1) / CT_X, nby = (Y + CT_Y - 1) / CT_Y, nbz = (Z + CT_Z - 1) / CT_Z;
  const int per_box = nbx * nby * nbz;
  const int ntiles = (int)(npts / XYZ) * per_box;
  // contiguous range of tiles of this work-group (see above)
  const int G = gridDim.x;
  const int slot = (blockIdx.x % CONV_XCDS) * ((G + CONV_XCDS - 1) / CONV_XCDS) + blockIdx.x / CONV_XCDS;
  const int chunk = (ntiles + G - 1) / G;
  const int t0 = slot * chunk, t1 = t0 + chunk < ntiles ? t0 + chunk : ntiles;
  constexpr int C8 = CK / 8;  // 16-byte pieces per input row
  constexpr int PIECES = CT_HALO * C8, PER_THREAD = (PIECES + CT_THREADS - 1) / CT_THREADS;
  // a tile's input rows: every load of the tile issued into registers
  // (the next tile's while the current one computes), then stored to LDS
  uint4 pre[PER_THREAD];
  auto fetch = [&](int tile) {
    const int b = tile / per_box, rem = tile % per_box;
    const int i0 = (rem % nbx) * CT_X, j0 = ((rem / nbx) % nby) * CT_Y, k0 = (rem / (nbx * nby)) * CT_Z;
    const long base = (long)b * XYZ;
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int e = threadIdx.x + CT_THREADS * u;
      const int hp = e / C8, c8 = e % C8;
      int gi = i0 + hp % CH_X - 1, gj = j0 + (hp / CH_X) % CH_Y - 1, gk = k0 + hp / (CH_X * CH_Y) - 1;
      bool ok = e < PIECES;
      if (per) {
        gi = gi < 0 ? gi + X : (gi >= X ? gi - X : gi);
        gj = gj < 0 ? gj + Y : (gj >= Y ? gj - Y : gj);
        gk = gk < 0 ? gk + Z : (gk >= Z ? gk - Z : gk);
      } else {
        ok = ok && gi >= 0 && gi < X && gj >= 0 && gj < Y && gk >= 0 && gk < Z;
      }
      // points of a partial tile beyond the far face read a wrapped or zero
      // row: their outputs are not stored
      gi = gi < X ? gi : X - 1;
      gj = gj < Y ? gj : Y - 1;
      gk = gk < Z ? gk : Z - 1;
      const long gp = base + gi + X * ((long)gj + (long)Y * gk);
      pre[u] = ok ? *reinterpret_cast<const uint4 *>(in + gp * CK + c8 * 8) : uint4{0u, 0u, 0u, 0u};
    }
  };
  auto put = [&](bf16 *buf) {
#pragma unroll
    for (int u = 0; u < PER_THREAD; ++u) {
      const int e = threadIdx.x + CT_THREADS * u;
      if (e < PIECES) *reinterpret_cast<uint4 *>(buf + (long)e * 8) = pre[u];
    }
  };
  if (t0 < t1) fetch(t0);
  for (int tile = t0; tile < t1; ++tile) {
    __syncthreads();  // the previous tile's reads of xt are done
    put(xt);
    __syncthreads();
    if (tile + 1 < t1) fetch(tile + 1);
    const bf16 *x = xt;
    f32x4 acc[CONV_TILES][CN / 16];
#pragma unroll
    for (int j = 0; j < CONV_TILES; ++j)
#pragma unroll
      for (int n = 0; n < CN / 16; ++n) acc[j][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    // the 9 taps of each dz unrolled: the LDS reads of later taps issue
    // while the matrix cores work on earlier ones
#pragma unroll 1
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const int t = 9 * dz + 3 * dy + dx;
          // neighbour (dx - 1, dy - 1, dz - 1), mirrored for the transposed operator
          const int ex = MODE == CONV_BACKWARD ? 2 - dx : dx, ey = MODE == CONV_BACKWARD ? 2 - dy : dy,
                    ez = MODE == CONV_BACKWARD ? 2 - dz : dz;
          const int row0 = ((wave + ez) * CH_Y + ey) * CH_X + r + ex;  // x-row j adds j * CH_X
          if constexpr (CK == 32) {
            uint4 a[CONV_TILES];
#pragma unroll
            for (int j = 0; j < CONV_TILES; ++j)
              a[j] = *reinterpret_cast<const uint4 *>(x + (row0 + j * CH_X) * CK + kg * 8);
#pragma unroll
            for (int n = 0; n < CN / 16; ++n) {
              const bf16x8 w = __builtin_bit_cast(
                  bf16x8, *reinterpret_cast<const uint4 *>(wl + (t * CN + n * 16 + r) * CK + kg * 8));
#pragma unroll
              for (int j = 0; j < CONV_TILES; ++j)
                acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[j]), w, acc[j][n],
                                                                    0, 0, 0);
            }
          } else {
            static_assert(CK == 16, "16 or 32 input channels");
            bf16x4 a[CONV_TILES];
#pragma unroll
            for (int j = 0; j < CONV_TILES; ++j)
              a[j] = *reinterpret_cast<const bf16x4 *>(x + (row0 + j * CH_X) * CK + kg * 4);
#pragma unroll
            for (int n = 0; n < CN / 16; ++n) {
              const bf16x4 w = *reinterpret_cast<const bf16x4 *>(wl + (t * CN + n * 16 + r) * CK + kg * 4);
#pragma unroll
              for (int j = 0; j < CONV_TILES; ++j)
                acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[j], w, acc[j][n], 0, 0, 0);
            }
          }
        }
    const int b = tile / per_box, rem = tile % per_box;
    const int i0 = (rem % nbx) * CT_X, j0 = ((rem / nbx) % nby) * CT_Y, k0 = (rem / (nbx * nby)) * CT_Z;
    const long base = (long)b * XYZ;
    const int kk = k0 + wave;
#pragma unroll
    for (int j = 0; j < CONV_TILES; ++j) {
      const int jj = j0 + j;
      if (jj >= Y || kk >= Z) continue;
#pragma unroll
      for (int n = 0; n < CN / 16; ++n) {
        const int ch = n * 16 + r;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int ii = i0 + 4 * kg + rr;
          if (ii >= X) continue;
          const long o = (base + ii + X * ((long)jj + (long)Y * kk)) * CN + ch;
          if (MODE == CONV_GELU) {
            const float z = acc[j][n][rr] + bias[ch];
            if (zsave) zsave[o] = to_bf16(z);
            out16[o] = to_bf16(gelu(z));
          } else if (MODE == CONV_LINEAR) {
            out32[o] = acc[j][n][rr] + bias[ch];
          } else {
            out16[o] = to_bf16(acc[j][n][rr] * gelu_grad(from_bf16(zin[o])));
          }
        }
      }
    }
  }
}

// Weight gradient partials of one tap: block (chunk, t) sums
// dz[p][co] in[p + o_t][ci] over the chunk's points into
// part[chunk][t][co][ci], as an MFMA GEMM with the points as reduction
// index: sub-chunks of WG_SUB points of dz and of the tap-shifted input are
// staged transposed in LDS ([channel][point]), so that every lane reads
// its 4 consecutive points of one channel as one 8-byte word.  Blocks of
// tap 0 also sum dz over their points into bpart[chunk][co] (bias
// gradient partials).  Wavefront w takes the (co, ci) tile w (at most 4).
constexpr int WG_SUB = 256;
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16 *__restrict__ in, const bf16 *__restrict__ dz,
                                                    ConvGeom geo, long npts, int chunk,
                                                    float *__restrict__ part, float *__restrict__ bpart) {
  constexpr int LD = WG_SUB + 8;
  __shared__ __align__(16) bf16 dzT[CO * LD];
  __shared__ __align__(16) bf16 inT[CI * LD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kg = lane >> 4;
  const int t = blockIdx.y;
  const int tx = t % 3, ty = (t / 3) % 3, tz = t / 9;
  const int c0 = blockIdx.x * chunk;
  const int c1 = c0 + chunk < npts ? c0 + chunk : (int)npts;
  constexpr int NT_CI = CI / 16, NTILE = (CO / 16) * NT_CI;
  static_assert(NTILE <= 4, "one tile per wavefront");
  const bool has_tile = wave < NTILE;
  const int mb = wave / NT_CI, nbk = wave % NT_CI;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const bool per = geo.periodic != 0;
  for (int sp = c0; sp < c1; sp += WG_SUB) {
    const int np = c1 - sp < WG_SUB ? c1 - sp : WG_SUB;  // a multiple of 16
    __syncthreads();
    for (int e = threadIdx.x; e < np * (CO / 8); e += 256) {
      const int q = e / (CO / 8), piece = e % (CO / 8);
      const uint4 v = *reinterpret_cast<const uint4 *>(dz + (long)(sp + q) * CO + piece * 8);
      const unsigned w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        dzT[(piece * 8 + 2 * u) * LD + q] = (bf16)(w4[u] & 0xffffu);
        dzT[(piece * 8 + 2 * u + 1) * LD + q] = (bf16)(w4[u] >> 16);
      }
    }
    for (int e = threadIdx.x; e < np * (CI / 8); e += 256) {
      const int q = e / (CI / 8), piece = e % (CI / 8);
      const Coord cc = coord_of(sp + q, geo);
      const AxisOffsets ax = axis_offsets(cc.i, geo.X, 1, per, 1);
      const AxisOffsets ay = axis_offsets(cc.j, geo.Y, geo.X, per, 1);
      const AxisOffsets az = axis_offsets(cc.k, geo.Z, geo.X * geo.Y, per, 1);
      const bool ok = per || ((ax.valid >> tx) & (ay.valid >> ty) & (az.valid >> tz) & 1u);
      const int nb = sp + q + pick3(ax.off, tx) + pick3(ay.off, ty) + pick3(az.off, tz);
      uint4 v = {0u, 0u, 0u, 0u};
      if (ok) v = *reinterpret_cast<const uint4 *>(in + (long)nb * CI + piece * 8);
      const unsigned w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        inT[(piece * 8 + 2 * u) * LD + q] = (bf16)(w4[u] & 0xffffu);
        inT[(piece * 8 + 2 * u + 1) * LD + q] = (bf16)(w4[u] >> 16);
      }
    }
    __syncthreads();
    if (has_tile)
      for (int k = 0; k < np; k += 16) {
        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(dzT + (mb * 16 + r) * LD + k + 4 * kg);
        const bf16x4 b = *reinterpret_cast<const bf16x4 *>(inT + (nbk * 16 + r) * LD + k + 4 * kg);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
      }
    if (t == 0 && (int)threadIdx.x < CO)
      for (int k = 0; k < np; ++k) bsum += from_bf16(dzT[threadIdx.x * LD + k]);
  }
  if (has_tile) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int co = mb * 16 + 4 * kg + rr, ci = nbk * 16 + r;
      part[(((long)blockIdx.x * 27 + t) * CO + co) * CI + ci] = acc[rr];
    }
  }
  if (t == 0 && (int)threadIdx.x < CO) bpart[(long)blockIdx.x * CO + threadIdx.x] = bsum;
}

// dW[e] = sum over chunks of part[chunk][e] (e over nelem entries)
__global__ void k_sum_chunks(const float *__restrict__ part, int nchunks, long nelem,
                             float *__restrict__ dW) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; ++c) s += part[c * nelem + e];
  dW[e] = s;
}

// Mean-square loss over the nout real channels of out[p][16] against
// target[c][p], counting only the points at least `halo` away from their
// box's faces (all points for halo 0); dout[p][16] <- bf16(2 (out - target)
// / count) (zero elsewhere), partial loss sums per block into lpart.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ out, const float *__restrict__ target,
                                             long npts, int nout, ConvGeom geo, int halo, float scale,
                                             bf16 *__restrict__ dout, float *__restrict__ lpart) {
  __shared__ float red[256];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  float s = 0.f;
  if (p < npts) {
    bool in = true;
    if (halo > 0) {
      const Coord cc = coord_of(p, geo);
      in = cc.i >= halo && cc.i < geo.X - halo && cc.j >= halo && cc.j < geo.Y - halo && cc.k >= halo &&
           cc.k < geo.Z - halo;
    }
    for (int c = 0; c < 16; ++c) {
      float g = 0.f;
      if (c < nout && in) {
        const float d = out[p * 16 + c] - target[c * npts + p];
        s += d * d;
        g = scale * d;
      }
      dout[p * 16 + c] = to_bf16(g);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// One batch entry: a box of the training set and the offset of the window
// taken from it (the window's first point is the sample point (ox, oy, oz)).
struct CropSpec {
  const bf16 *feat;   // sample features [S^3][16]
  const float *tgt;   // sample targets [5][S^3]
  int ox, oy, oz, pad;
};

// Batch of windows X^3 (X = crop + 2 halo) cut from periodic S^3 samples:
// features of every window point, targets of every point (the loss masks
// the halo).  One thread per window point.
__global__ __launch_bounds__(256) void k_gather_crops(const CropSpec *__restrict__ spec, int S, ConvGeom geo,
                                                      long npts, bf16 *__restrict__ feat,
                                                      float *__restrict__ tgt) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npts) return;
  const Coord cc = coord_of(p, geo);
  const CropSpec sp = spec[cc.base / ((long)geo.X * geo.Y * geo.Z)];
  const int msk = S - 1;
  const int i = (sp.ox + cc.i) & msk, j = (sp.oy + cc.j) & msk, k = (sp.oz + cc.k) & msk;
  const long src = i + (long)S * (j + (long)S * k);
  const long S3 = (long)S * S * S;
  const uint4 *f = reinterpret_cast<const uint4 *>(sp.feat + src * 16);
  uint4 *d = reinterpret_cast<uint4 *>(feat + p * 16);
  d[0] = f[0];
  d[1] = f[1];
#pragma unroll
  for (int c = 0; c < 5; ++c) tgt[c * npts + p] = sp.tgt[c * S3 + src];
}

// Adam on the FP32 master parameters (bias-corrected step sizes from the host)
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long n, float lr_t, float b1, float b2, float eps_t) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  w[e] = w[e] - lr_t * mm / (sqrtf(vv) + eps_t);
}

// BF16 copies of one layer's weights W[t][co][ci] (FP32, padded): the forward
// operand as is, and the transposed operand WT[t][ci][co] for the input grad.
__global__ void k_pack_weights(const float *__restrict__ w, int CO, int CI, bf16 *__restrict__ wf,
                               bf16 *__restrict__ wt) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 27L * CO * CI) return;
  const int t = (int)(e / ((long)CO * CI));
  const int co = (int)((e / CI) % CO), ci = (int)(e % CI);
  const bf16 h = to_bf16(w[e]);
  wf[e] = h;
  wt[((long)t * CI + ci) * CO + co] = h;
}

// Means of the 5 fields of a state, in a fixed summation order: partial
// sums of MEAN_BLOCKS grid-strided work-groups, then one work-group adds
// them.  The network sees every sample relative to its own means.
constexpr int MEAN_BLOCKS = 256;
template <typename Q>
__global__ __launch_bounds__(256) void k_field_partials(const Q *__restrict__ q, long N,
                                                        double *__restrict__ part) {
  __shared__ double red[256];
  for (int c = 0; c < 5; ++c) {
    double s = 0.0;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < N; p += (long)gridDim.x * 256) s += (double)q[c * N + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[c * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_field_means(const double *__restrict__ part, int nparts, long N,
                                                     double *__restrict__ mean) {
  __shared__ double red[256];
  for (int c = 0; c < 5; ++c) {
    red[threadIdx.x] = (int)threadIdx.x < nparts ? part[c * nparts + threadIdx.x] : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) mean[c] = red[0] / (double)N;
    __syncthreads();
  }
}

// Normalisation of a closure network (device copy, set by the host).
// Features: x_c = (q_c - mean_c) inv_std_c with the sample's field means;
// outputs: force_v = out_v out_std_v, times mean_p / p_ref for the pressure
// equation when p_ref > 0 (its residual scales with the mean pressure).
struct CnnNorm {
  float inv_std[5], out_std[5], out_inv_std[5];
  float gain[5];  // inference only: force_v = gain[v] x the network's term
  double p_ref;
};

__device__ __forceinline__ float pressure_scale(const CnnNorm &nm, const double *mean) {
  return nm.p_ref > 0.0 ? (float)(mean[4] / nm.p_ref) : 1.f;
}

// Network input from a state q (5 fields of N points, field-major): channel c
// of point p = bf16(f32(q[c][p] - mean[c]) inv_std[c]), channels 5..15 zero.
template <typename Q>
__global__ void k_pack_features(const Q *__restrict__ q, long N, const double *__restrict__ mean,
                                const CnnNorm *__restrict__ nm, bf16 *__restrict__ feat) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  unsigned h[3];
#pragma unroll
  for (int c = 0; c < 6; c += 2) {
    const unsigned lo = to_bf16((float)((double)q[c * N + p] - mean[c]) * nm->inv_std[c]);
    const unsigned hi = c + 1 < 5 ? to_bf16((float)((double)q[(c + 1) * N + p] - mean[c + 1]) * nm->inv_std[c + 1]) : 0u;
    h[c / 2] = lo | (hi << 16);
  }
  uint4 *d = reinterpret_cast<uint4 *>(feat + p * 16);
  d[0] = uint4{h[0], h[1], h[2], 0u};
  d[1] = uint4{0u, 0u, 0u, 0u};
}

// Closure term in the solver's layout: force[v][p] = out[p][v] out_std[v] (x pressure scale) x gain[v]
template <typename T>
__global__ void k_unpack_closure(const float *__restrict__ out, long N, const CnnNorm *__restrict__ nm,
                                 const double *__restrict__ mean, T *__restrict__ force) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const float ps = pressure_scale(*nm, mean);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const float s = v == 4 ? nm->out_std[4] * ps : nm->out_std[v];
    force[v * N + p] = T(out[p * 16 + v] * s * nm->gain[v]);
  }
}

// Zero domain mean for the closure terms of the conserved variables (rho,
// rho u, rho v, rho w): a subfilter term of a conservation law is the
// divergence of a subfilter flux and adds nothing to the domain total.
template <typename T>
__global__ void k_remove_mean(T *__restrict__ force, long N, const double *__restrict__ mean) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) force[v * N + p] = force[v * N + p] - T(mean[v]);
}

// Normalised targets: tn[c][p] = target[c][p] out_inv_std[c] (/ pressure scale)
__global__ void k_scale_targets(const double *__restrict__ target, long N, const CnnNorm *__restrict__ nm,
                                const double *__restrict__ mean, float *__restrict__ tn) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const float ps = pressure_scale(*nm, mean);
  for (int c = 0; c < 5; ++c) {
    const float s = c == 4 ? nm->out_inv_std[4] / ps : nm->out_inv_std[c];
    tn[c * N + p] = (float)target[c * N + p] * s;
  }
}

// Box average of a state over f^3 blocks: out (m1, m2, m3, 5) with m = n / f.
template <typename T>
__global__ void k_box_average(const T *__restrict__ q, int n1, int n2, int n3, int f,
                              double *__restrict__ out) {
  const int m1 = n1 / f, m2 = n2 / f, m3 = n3 / f;
  const long M = (long)m1 * m2 * m3;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 5 * M) return;
  const int v = (int)(e / M);
  const long r = e - (long)v * M;
  const int i = (int)(r % m1), j = (int)((r / m1) % m2), k = (int)(r / ((long)m1 * m2));
  const long N = (long)n1 * n2 * n3;
  double s = 0.0;
  for (int c = 0; c < f; ++c)
    for (int b = 0; b < f; ++b)
      for (int a = 0; a < f; ++a)
        s += (double)q[v * N + (i * f + a) + (long)n1 * ((j * f + b) + (long)n2 * (k * f + c))];
  out[e] = s / ((double)f * f * f);
}

}  // namespace cruna
