// Streaming bandwidth of one MI355X: the practical HBM ceiling that the
// sweeps are measured against.  Four kernels on FP64 arrays far larger than
// the caches (default 2^27 doubles = 1 GiB each):
//   copy   b = a            (1 read, 1 write)
//   scale  b = s a          (1 read, 1 write)
//   triad  c = a + s b      (2 reads, 1 write)
//   read   sum of a         (1 read; one partial per work-group)
// plus copy and triad with four nontemporal 16-byte accesses per thread and
// iteration, the copy with only two work-groups per CU, and the sweeps'
// mixes of fields read and written per point (fields_rR_wW).
// 16-byte accesses, grid-stride over enough work-groups to fill every CU,
// timed with HIP events over `reps` launches after one warm-up.  Prints one
// JSON line: bytes moved per launch / average time for each kernel.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

constexpr int TPB = 256;

__global__ __launch_bounds__(TPB) void k_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long n2) {
  for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < n2; i += (long)gridDim.x * TPB) b[i] = a[i];
}

__global__ __launch_bounds__(TPB) void k_scale(const double2 *__restrict__ a, double2 *__restrict__ b, double s,
                                               long n2) {
  for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < n2; i += (long)gridDim.x * TPB) {
    const double2 x = a[i];
    b[i] = make_double2(s * x.x, s * x.y);
  }
}

__global__ __launch_bounds__(TPB) void k_triad(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                               double2 *__restrict__ c, double s, long n2) {
  for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < n2; i += (long)gridDim.x * TPB) {
    const double2 x = a[i], y = b[i];
    c[i] = make_double2(x.x + s * y.x, x.y + s * y.y);
  }
}

__global__ __launch_bounds__(TPB) void k_read(const double2 *__restrict__ a, double *__restrict__ part, long n2) {
  double s = 0.0;
  for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < n2; i += (long)gridDim.x * TPB) {
    const double2 x = a[i];
    s += x.x + x.y;
  }
  __shared__ double w[TPB];
  w[threadIdx.x] = s;
  __syncthreads();
  for (int o = TPB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) w[threadIdx.x] += w[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = w[0];
}

// four independent 16-byte accesses per thread and iteration, nontemporal
// (streaming) loads and stores: more bytes in flight per wave
typedef double dv2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(TPB) void k_copy4(const dv2 *__restrict__ a, dv2 *__restrict__ b, long n2) {
  const long stride = (long)gridDim.x * TPB;
  long i = (long)blockIdx.x * TPB + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    dv2 x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(x[u], b + i + u * stride);
  }
  for (; i < n2; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(TPB) void k_triad4(const dv2 *__restrict__ a, const dv2 *__restrict__ b,
                                                dv2 *__restrict__ c, double s, long n2) {
  const long stride = (long)gridDim.x * TPB;
  long i = (long)blockIdx.x * TPB + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    dv2 x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[u] = __builtin_nontemporal_load(a + i + u * stride);
      y[u] = __builtin_nontemporal_load(b + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      __builtin_nontemporal_store(x[u] + s * y[u], c + i + u * stride);
  }
  for (; i < n2; i += stride) c[i] = a[i] + s * b[i];
}

// The sweeps' access mix: R fields read and W fields written per point
// (sweep1_x reads 5 and writes 9 FP64 fields, sweep1_y 12 and 8,
// sweep2_x_rk 17 and 11 in its middle stages), fields of F doubles each, the
// point's 16-byte pair of every field per thread and iteration
constexpr int MAXF = 28;
struct Fields {
  const double2 *in[MAXF];
  double2 *out[MAXF];
};

template <int R, int W>
__global__ __launch_bounds__(TPB) void k_fields(Fields f, long n2) {
  for (long i = (long)blockIdx.x * TPB + threadIdx.x; i < n2; i += (long)gridDim.x * TPB) {
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 x = f.in[r][i];
      s.x += x.x;
      s.y += x.y;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) f.out[w][i] = make_double2(s.x * (w + 1), s.y * (w + 1));
  }
}

__global__ void k_init(double *a, long n, double v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    a[i] = v + 1e-9 * double(i & 1023);
}

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

}  // namespace

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : (1L << 27);
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  if (n < 2 || n % 2 != 0 || reps < 1) {
    std::fprintf(stderr, "usage: stream_bench [doubles per array, even] [reps]\n");
    return 2;
  }
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long n2 = n / 2;
  // 16 work-groups of 256 threads per CU, capped by the array
  const int blocks = (int)std::min<long>(16L * cus, (n2 + TPB - 1) / TPB);
  double *a, *b, *c, *part;
  CHECK(hipMalloc(&a, n * sizeof(double)));
  CHECK(hipMalloc(&b, n * sizeof(double)));
  CHECK(hipMalloc(&c, n * sizeof(double)));
  CHECK(hipMalloc(&part, blocks * sizeof(double)));
  hipLaunchKernelGGL(k_init, dim3(4096), dim3(256), 0, 0, a, n, 1.0);
  hipLaunchKernelGGL(k_init, dim3(4096), dim3(256), 0, 0, b, n, 2.0);
  hipLaunchKernelGGL(k_init, dim3(4096), dim3(256), 0, 0, c, n, 0.0);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const double bytes = double(n) * sizeof(double);
  struct Res { const char *name; double moved; float ms; };
  std::vector<Res> res;
  auto time_it = [&](const char *name, double moved, auto launch) {
    launch();
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) launch();
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    res.push_back({name, moved, ms / reps});
  };
  const double2 *A = (const double2 *)a, *B = (const double2 *)b;
  time_it("copy", 2 * bytes, [&] { hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(TPB), 0, 0, A, (double2 *)c, n2); });
  time_it("scale", 2 * bytes,
          [&] { hipLaunchKernelGGL(k_scale, dim3(blocks), dim3(TPB), 0, 0, A, (double2 *)c, 0.5, n2); });
  time_it("triad", 3 * bytes,
          [&] { hipLaunchKernelGGL(k_triad, dim3(blocks), dim3(TPB), 0, 0, A, B, (double2 *)c, 0.5, n2); });
  time_it("read", bytes, [&] { hipLaunchKernelGGL(k_read, dim3(blocks), dim3(TPB), 0, 0, A, part, n2); });
  time_it("copy4_nt", 2 * bytes,
          [&] { hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(TPB), 0, 0, (const dv2 *)a, (dv2 *)c, n2); });
  time_it("triad4_nt", 3 * bytes,
          [&] { hipLaunchKernelGGL(k_triad4, dim3(blocks), dim3(TPB), 0, 0, (const dv2 *)a, (const dv2 *)b, (dv2 *)c, 0.5,
                                   n2); });
  // the copy with fewer work-groups (2 per CU): the sweeps' occupancy
  const int few = 2 * cus;
  time_it("copy_2wg_per_cu", 2 * bytes,
          [&] { hipLaunchKernelGGL(k_copy, dim3(few), dim3(TPB), 0, 0, A, (double2 *)c, n2); });
  // the sweeps' field mixes on fields of 2^24 doubles (a 256^3 field)
  {
    const long nf = 1L << 24, nf2 = nf / 2;
    std::vector<double *> fl(MAXF);
    for (auto &p : fl) {
      CHECK(hipMalloc(&p, nf * sizeof(double)));
      hipLaunchKernelGGL(k_init, dim3(4096), dim3(256), 0, 0, p, nf, 1.0);
    }
    CHECK(hipDeviceSynchronize());
    Fields f{};
    const int fb = (int)std::min<long>(16L * cus, (nf2 + TPB - 1) / TPB);
    auto mix = [&](const char *name, int R, int W, auto kern) {
      for (int r = 0; r < R; ++r) f.in[r] = (const double2 *)fl[r];
      for (int w = 0; w < W; ++w) f.out[w] = (double2 *)fl[R + w];
      time_it(name, double(R + W) * nf * sizeof(double),
              [&] { hipLaunchKernelGGL(kern, dim3(fb), dim3(TPB), 0, 0, f, nf2); });
    };
    mix("fields_r5_w9", 5, 9, k_fields<5, 9>);
    mix("fields_r12_w8", 12, 8, k_fields<12, 8>);
    mix("fields_r13_w8", 13, 8, k_fields<13, 8>);
    mix("fields_r17_w11", 17, 11, k_fields<17, 11>);
    mix("fields_r5_w5", 5, 5, k_fields<5, 5>);
    for (auto p : fl) CHECK(hipFree(p));
  }
  std::string out = "{\"array_bytes\": " + std::to_string((long)bytes) + ", \"reps\": " + std::to_string(reps) +
                    ", \"blocks\": " + std::to_string(blocks) + ", \"cus\": " + std::to_string(cus);
  for (auto &r : res) {
    char buf[160];
    std::snprintf(buf, sizeof buf, ", \"%s\": {\"ms\": %.4f, \"GBps\": %.1f}", r.name, r.ms, r.moved / (r.ms * 1e6));
    out += buf;
  }
  out += "}";
  std::printf("%s\n", out.c_str());
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  CHECK(hipFree(c));
  CHECK(hipFree(part));
  return 0;
}
