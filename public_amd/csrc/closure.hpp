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
// v_mfma_f32_16x16x16_bf16 operand layout (lane l of 64): A row l%16 holds
// K = 4(l/16)..+3; B column l%16 holds K = 4(l/16)..+3; D row 4(l/16)+r,
// column l%16, r = 0..3.
#pragma once

#include <hip/hip_runtime.h>

namespace cruna {

typedef unsigned short bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ bf16 to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((unsigned)h << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(c (z + a z^3))), and its derivative
__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  return 0.5f * z * (1.f + tanhf(c * (z + a * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float t = tanhf(c * (z + a * z * z * z));
  return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * a * z * z);
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

// One convolution as an implicit GEMM; each wavefront computes 16 points x CN
// output channels.  CK: input (reduction) channels, CN: output channels.
//   CONV_GELU     : z = conv + bias; zsave <- bf16(z) (if set); out16 <- bf16(gelu(z))
//   CONV_LINEAR   : out32 <- conv + bias
//   CONV_BACKWARD : taps mirrored, W given as [t][CN][CK] of the transposed
//                   operator; out16 <- bf16(conv * gelu'(zin)) (zin: the
//                   pre-activations of this layer's input, CN channels)
// npts is a multiple of 64 (checked by the host).
template <int CK, int CN, int MODE>
__global__ __launch_bounds__(256) void k_conv(const bf16 *__restrict__ in, const bf16 *__restrict__ W,
                                              const float *__restrict__ bias, ConvGeom geo, long npts,
                                              float *__restrict__ out32, bf16 *__restrict__ out16,
                                              bf16 *__restrict__ zsave, const bf16 *__restrict__ zin) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long p0 = ((long)blockIdx.x * 4 + wave) * 16;
  if (p0 >= npts) return;
  const int r = lane & 15, kg = lane >> 4;
  const Coord pc = coord_of(p0 + r, geo);
  f32x4 acc[CN / 16];
#pragma unroll
  for (int n = 0; n < CN / 16; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int sign = MODE == CONV_BACKWARD ? -1 : 1;
  for (int t = 0; t < 27; ++t) {
    const long nb = tap_neighbour(pc, t, sign, geo);
#pragma unroll
    for (int kc = 0; kc < CK / 16; ++kc) {
      const bf16x4 a = nb >= 0 ? *reinterpret_cast<const bf16x4 *>(in + nb * CK + kc * 16 + kg * 4)
                               : bf16x4{0, 0, 0, 0};
#pragma unroll
      for (int n = 0; n < CN / 16; ++n) {
        const bf16x4 w =
            *reinterpret_cast<const bf16x4 *>(W + ((long)(t * CN + n * 16 + r) * CK + kc * 16 + kg * 4));
        acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc[n], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < CN / 16; ++n) {
    const int ch = n * 16 + r;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const long q = p0 + 4 * kg + rr;
      const long o = q * CN + ch;
      if (MODE == CONV_GELU) {
        const float z = acc[n][rr] + bias[ch];
        if (zsave) zsave[o] = to_bf16(z);
        out16[o] = to_bf16(gelu(z));
      } else if (MODE == CONV_LINEAR) {
        out32[o] = acc[n][rr] + bias[ch];
      } else {
        out16[o] = to_bf16(acc[n][rr] * gelu_grad(from_bf16(zin[o])));
      }
    }
  }
}

// Weight gradient partials: block (chunk, tap) sums dz[p][co] in[p + o_t][ci]
// over the chunk's points into part[chunk][t][co][ci].  Each wavefront takes
// 16x16 (co, ci) tiles in turn.
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16 *__restrict__ in, const bf16 *__restrict__ dz,
                                                    ConvGeom geo, long npts, int chunk,
                                                    float *__restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kg = lane >> 4;
  const int t = blockIdx.y;
  const long c0 = (long)blockIdx.x * chunk;
  const long c1 = c0 + chunk < npts ? c0 + chunk : npts;
  constexpr int NT_CO = CO / 16, NT_CI = CI / 16;
  for (int tile = wave; tile < NT_CO * NT_CI; tile += 4) {
    const int mb = tile / NT_CI, nbk = tile % NT_CI;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long pc = c0; pc < c1; pc += 16) {
      bf16x4 a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long q = pc + 4 * kg + j;
        const long nb = tap_neighbour(coord_of(q, geo), t, 1, geo);
        a[j] = (short)dz[q * CO + mb * 16 + r];
        b[j] = nb >= 0 ? (short)in[nb * CI + nbk * 16 + r] : (short)0;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int co = mb * 16 + 4 * kg + rr, ci = nbk * 16 + r;
      part[(((long)blockIdx.x * 27 + t) * CO + co) * CI + ci] = acc[rr];
    }
  }
}

// dW[e] = sum over chunks of part[chunk][e] (e over 27 CO CI entries)
__global__ void k_sum_chunks(const float *__restrict__ part, int nchunks, long nelem,
                             float *__restrict__ dW) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; ++c) s += part[c * nelem + e];
  dW[e] = s;
}

// db[co] = sum_p dz[p][co]: one block per channel, fixed reduction order
__global__ __launch_bounds__(256) void k_bias_grad(const bf16 *__restrict__ dz, long npts, int CO,
                                                   float *__restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (long p = threadIdx.x; p < npts; p += 256) s += from_bf16(dz[p * CO + co]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
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

// Means of the 5 fields of a state (one work-group, fixed summation order):
// the network sees every sample relative to its own means.
template <typename Q>
__global__ __launch_bounds__(1024) void k_field_means(const Q *__restrict__ q, long N, double *__restrict__ mean) {
  __shared__ double red[1024];
  for (int c = 0; c < 5; ++c) {
    double s = 0.0;
    for (long p = threadIdx.x; p < N; p += 1024) s += (double)q[c * N + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
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

// Closure term in the solver's layout: force[v][p] = out[p][v] out_std[v] (x pressure scale)
template <typename T>
__global__ void k_unpack_closure(const float *__restrict__ out, long N, const CnnNorm *__restrict__ nm,
                                 const double *__restrict__ mean, T *__restrict__ force) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const float ps = pressure_scale(*nm, mean);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const float s = v == 4 ? nm->out_std[4] * ps : nm->out_std[v];
    force[v * N + p] = T(out[p * 16 + v] * s);
  }
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
