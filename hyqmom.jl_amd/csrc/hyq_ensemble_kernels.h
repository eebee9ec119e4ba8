// Kernels of the ensemble step (hyq_ensemble.hip), instantiated per precision mode in
// hyq_ens_fp64.hip, hyq_ens_mixed.hip and hyq_ens_newton.hip so that the three compile in
// parallel.  The layout and the block table are described in hyq_ensemble.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "hyq_pipeline.h"
#include "hyq_update.h"

namespace hyq_ens {
using namespace hyq;

struct RunDev {
  long long off, N;
  int Nx, Ny, Nz, flag2D, positivity, pad;
  double Ma, Kn, CFL, dtmax, dx, dy, dz;
  double t, dt, tmax, active;
  unsigned long long vmax_bits;
  long long nn, nnmax;
  unsigned long long err;  // bit 0: a non-finite z Jacobian (where the reference throws)
};
struct Blk {
  int run, pad;
  long long first;
};

__device__ __forceinline__ unsigned long long ewave_max(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

// at least two waves per SIMD: as for the single-run kernels, latency hiding pays for the
// spills this forces
template <int B, typename FT, int EM>
__global__ __launch_bounds__(B, 2) void ke_pass1(const double* __restrict__ M, double* __restrict__ Mr,
                                              FT* __restrict__ F, FT* __restrict__ vp,
                                              uint32_t* __restrict__ flags, RunDev* R,
                                              const Blk* __restrict__ T, long long P) {
  const Blk b = T[blockIdx.x];
  RunDev* q = R + b.run;
  if (q->active == 0.0) return;
  const long long c = b.first + threadIdx.x;
  unsigned long long vb = 0;
  if (c < q->off + q->N) {
    double m[35], mr[35], v[6];
#pragma unroll
    for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * P + c];
    uint32_t fl = 0;
    pass1_cell<FT, EM>(m, q->flag2D, q->Ma, q->positivity != 0, mr, F + c, F + (size_t)35 * P + c,
                       F + (size_t)70 * P + c, (size_t)P, v, fl);
#pragma unroll
    for (int n = 0; n < 35; ++n) Mr[(size_t)n * P + c] = mr[n];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      vp[(size_t)d * P + c] = (FT)v[d];
      const unsigned long long bits = __double_as_longlong(fabs(v[d]));
      vb = bits > vb ? bits : vb;
    }
    flags[c] = fl;
    if (fl & F_NONFIN) atomicOr(&q->err, 1ull);
  }
  vb = ewave_max(vb);
  if ((threadIdx.x & 63) == 0) atomicMax(&q->vmax_bits, vb);
}

template <int B, typename FT>
__global__ __launch_bounds__(B) void ke_update(const double* __restrict__ Mr, const FT* __restrict__ F,
                                               const FT* __restrict__ vp, double* __restrict__ Mout,
                                               const RunDev* R, const Blk* __restrict__ T, long long P) {
  const Blk b = T[blockIdx.x];
  const RunDev* q = R + b.run;
  if (q->active == 0.0) return;
  const long long c = b.first + threadIdx.x;
  if (c >= q->off + q->N) return;
  const long long l = c - q->off;
  const int Nx = q->Nx, Ny = q->Ny;
  const int i = (int)(l % Nx), j = (int)((l / Nx) % Ny), k = (int)(l / ((long long)Nx * Ny));
  const double dt = q->dt;
  update_cell<FT>(Mr, F, vp, Mout, c, i, j, k, Nx, Ny, q->Nz, P, dt / q->dx, dt / q->dy, dt / q->dz);
}

template <int B, bool MIXED, int EM>
__global__ __launch_bounds__(B, 2) void ke_pass2(double* __restrict__ M, const RunDev* R,
                                              const Blk* __restrict__ T, long long P) {
  const Blk b = T[blockIdx.x];
  const RunDev* q = R + b.run;
  if (q->active == 0.0) return;
  const long long c = b.first + threadIdx.x;
  if (c >= q->off + q->N) return;
  double m[35], o[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * P + c];
  uint32_t fl = 0;
  pass2_cell<MIXED, EM>(m, q->flag2D, q->Ma, q->positivity != 0, q->dt, q->Kn, o, fl);
#pragma unroll
  for (int n = 0; n < 35; ++n) M[(size_t)n * P + c] = o[n];
}

// Device pointers of the ensemble's planes.
struct EnsPtrs {
  double *M, *Mr;
  void *F, *vp;
  uint32_t* flags;
  RunDev* R;
  long long P;
};

// Launch one of the three cell kernels (0 pass 1, 1 update, 2 pass 2) with block size b.
template <typename FT, int EM, bool MIXED>
void launch_cells(int which, int b, unsigned nblk, const Blk* T, const EnsPtrs& p, hipStream_t s) {
  FT* F = (FT*)p.F;
  FT* vp = (FT*)p.vp;
#define HYQ_EL(B)                                                                                    \
  do {                                                                                               \
    if (which == 0)                                                                                  \
      hipLaunchKernelGGL((ke_pass1<B, FT, EM>), dim3(nblk), dim3(B), 0, s, p.M, p.Mr, F, vp, p.flags, \
                         p.R, T, p.P);                                                               \
    else if (which == 1)                                                                             \
      hipLaunchKernelGGL((ke_update<B, FT>), dim3(nblk), dim3(B), 0, s, p.Mr, F, vp, p.M, p.R, T,     \
                         p.P);                                                                       \
    else                                                                                             \
      hipLaunchKernelGGL((ke_pass2<B, MIXED, EM>), dim3(nblk), dim3(B), 0, s, p.M, p.R, T, p.P);      \
  } while (0)
  if (b == 64) HYQ_EL(64);
  else if (b == 128) HYQ_EL(128);
  else HYQ_EL(256);
#undef HYQ_EL
}

// per-precision entry points (one translation unit each)
void launch_fp64(int which, int b, unsigned nblk, const Blk* T, const EnsPtrs& p, hipStream_t s);
void launch_mixed(int which, int b, unsigned nblk, const Blk* T, const EnsPtrs& p, hipStream_t s);
void launch_newton(int which, int b, unsigned nblk, const Blk* T, const EnsPtrs& p, hipStream_t s);

}  // namespace hyq_ens
