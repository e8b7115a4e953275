// Batched sweeps: one launch per sweep kind over every member of an
// ensemble.
//
// The reference runs one box per process (its cores keep one line length per
// direction in module state, src/diff/diff_pen_core_o6_i_c_p.f90:27-32); an
// ensemble of unevenly sized boxes is here one launch per sweep kind that
// walks a work list of (member, tile): member k of a list owns work-groups
// start[k] .. start[k+1]-1 and each work-group runs the member's own kernel
// body on one of its tiles, with the member's tables, geometry and fields
// read from a device array of descriptors.  Each member's tile takes as many
// of its lines as fill the list's work-group, whatever the line length
// (thread_pos_tile with BATCH), so that members of 48 or 96 points per line
// keep the work-group as busy as 64 or 128 do.  FP64 members on Cartesian
// grids without boundary rows.
#pragma once

#include "adjoint.hpp"

namespace cruna {

constexpr int BATCH_MAX = 64;

// The arguments the per-member kernels take, for one member.
struct BatchMember {
  double *qn, *acc, *qs, *rhs, *G;  // direct state, RK accumulator, stage input, RHS, gradients
  double *qa;                       // adjoint state
  const double **q0p;               // adjoint: cell with the forward state of the step
  int *ntp;                         // adjoint: backward step counter
  int *sckp;                        // subset adjoint: checkpoint counter
  const double *traj;               // forward trajectory slots
  const double *ckpt;               // subset checkpoints
  int traj_slots;
  LineOp<double> op[3], fop[3];
  SweepGeom geom[3], fgeom[3];
  Phys ph;
  long N;
  double dt, gamma;
};

// Work list of a batched launch.
struct BatchList {
  int nm;
  int mem[BATCH_MAX];        // descriptor of list entry k
  int start[BATCH_MAX + 1];  // first work-group of entry k; start[nm] = all
};

// The list entry whose work-groups include this one (uniform).
__device__ __forceinline__ int batch_entry(const BatchList *__restrict__ L) {
  const int b = blockIdx.x;
  int lo = 0, hi = L->nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L->start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// first-derivative sweep; stage input q^n (from_qs = 0) or qs
template <int DIR, bool PER>
__global__ __launch_bounds__(NT) void kb_sweep1(const BatchMember *__restrict__ M, const BatchList *__restrict__ L,
                                                int from_qs) {
  const int e = batch_entry(L);
  const BatchMember &b = M[L->mem[e]];
  const ThreadPos tp = thread_pos_tile<DIR, true>(b.geom[DIR], blockIdx.x - L->start[e]);
  __shared__ LineLds<double> lds;
  sweep1_body<double, double, DIR, PER, RHS_CART>(tp, from_qs ? b.qs : b.qn, b.rhs, b.G, b.op[DIR], b.geom[DIR],
                                                  b.ph, b.N, Metric<double>{}, Penal<double>{}, lds);
}

// viscous sweep; the x sweep finishes RK4 stage `mode` (FF: the last stage
// with the fused x filter)
template <int DIR, bool PER, bool FF>
__global__ __launch_bounds__(NT) void kb_sweep2(const BatchMember *__restrict__ M, const BatchList *__restrict__ L,
                                                int mode) {
  const int e = batch_entry(L);
  const BatchMember &b = M[L->mem[e]];
  const ThreadPos tp = thread_pos_tile<DIR, true>(b.geom[DIR], blockIdx.x - L->start[e]);
  sweep2_body<double, DIR, PER, RHS_CART, FF>(tp, b.rhs, b.G, b.op[DIR], b.fop[DIR], b.geom[DIR], b.ph, b.N, mode,
                                               b.qn, b.acc, b.qs, b.dt, Metric<double>{}, nullptr,
                                               Penal<double>{});
}

// the z first-derivative and viscous sweeps as one kernel (k_sweep12_z)
template <bool PER>
__global__ __launch_bounds__(NT) void kb_sweep12_z(const BatchMember *__restrict__ M,
                                                   const BatchList *__restrict__ L, int from_qs) {
  const int e = batch_entry(L);
  const BatchMember &b = M[L->mem[e]];
  const ThreadPos tp = thread_pos_tile<2, true>(b.geom[2], blockIdx.x - L->start[e]);
  __shared__ LineLds<double> lds;
  __shared__ double mom[3][PPT][NT];
  ZCarry<double> zc;
  zc.mom = mom;
  sweep1_body<double, double, 2, PER, RHS_CART, true>(tp, from_qs ? b.qs : b.qn, b.rhs, b.G, b.op[2], b.geom[2],
                                                      b.ph, b.N, Metric<double>{}, Penal<double>{}, lds, &zc);
  viscous_z_carry<double, PER>(tp, b.rhs, b.G, b.op[2], b.geom[2], b.ph, b.N, lds, zc);
}

// filter of the direct (adj = 0) or adjoint state in place
template <int DIR, bool PER>
__global__ __launch_bounds__(NT) void kb_filter(const BatchMember *__restrict__ M, const BatchList *__restrict__ L,
                                                int adj) {
  const int e = batch_entry(L);
  const BatchMember &b = M[L->mem[e]];
  const ThreadPos tp = thread_pos_tile<DIR, true>(b.fgeom[DIR], blockIdx.x - L->start[e]);
  double *q = adj ? b.qa : b.qn;
  filter_body<double, DIR, PER, double>(tp, q, q, b.fop[DIR], b.fgeom[DIR], b.N, 5, (double *)nullptr);
}

// adjoint sweep (backward in time: dt -> -dt); stage input qs^n (from_qs = 0) or qs
template <int DIR, bool PER>
__global__ __launch_bounds__(NT) void kb_asweep(const BatchMember *__restrict__ M, const BatchList *__restrict__ L,
                                                int mode, int from_qs) {
  const int e = batch_entry(L);
  const BatchMember &b = M[L->mem[e]];
  const ThreadPos tp = thread_pos_tile<DIR, true>(b.geom[DIR], blockIdx.x - L->start[e]);
  asweep_body<double, double, DIR, PER>(tp, from_qs ? b.qs : b.qa, (const double *const *)b.q0p, b.rhs, nullptr,
                                        b.op[DIR], b.geom[DIR], b.gamma, b.ph.inv_h[DIR], b.N, mode, b.qa, b.acc,
                                        b.qs, -b.dt);
}

// Per-member bookkeeping of a batched step, one work-group row per member
// (blockIdx.y = list entry):
//   the adjoint step's forward state, Q(:,:,:,:,nt) of the counter, which
//   then counts down (k_adjoint_slot for every member of the list)
__global__ void kb_adjoint_slot(const BatchMember *__restrict__ M, const BatchList *__restrict__ L) {
  if (threadIdx.x != 0) return;
  const BatchMember &b = M[L->mem[blockIdx.y]];
  const int nt = *b.ntp;
  *b.q0p = b.traj + (long)(b.traj_slots > 1 ? nt - 1 : 0) * 5 * b.N;
  *b.ntp = nt - 1;
}

//   the backward step counter of every member set to v
__global__ void kb_set_counter(const BatchMember *__restrict__ M, const BatchList *__restrict__ L, int v) {
  if (threadIdx.x != 0) return;
  *M[L->mem[blockIdx.y]].ntp = v;
}

//   Q(:,:,:,:,slot) = q^n (the recomputed subset trajectory)
__global__ __launch_bounds__(256) void kb_record(const BatchMember *__restrict__ M, const BatchList *__restrict__ L,
                                                 int slot) {
  const BatchMember &b = M[L->mem[blockIdx.y]];
  const long n = 5 * b.N;
  double *dst = const_cast<double *>(b.traj) + (long)(slot - 1) * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = b.qn[i];
}

//   q^n = the checkpoint the member's subset counter names; the counter
//   then counts down (every work-group reads it before the last one of
//   the row could change it: the count-down is a separate launch)
__global__ __launch_bounds__(256) void kb_restore(const BatchMember *__restrict__ M, const BatchList *__restrict__ L) {
  const BatchMember &b = M[L->mem[blockIdx.y]];
  const long n = 5 * b.N;
  const double *src = b.ckpt + (long)(*b.sckp) * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b.qn[i] = src[i];
}

__global__ void kb_count_down(const BatchMember *__restrict__ M, const BatchList *__restrict__ L) {
  if (threadIdx.x != 0) return;
  *M[L->mem[blockIdx.y]].sckp -= 1;
}

}  // namespace cruna
