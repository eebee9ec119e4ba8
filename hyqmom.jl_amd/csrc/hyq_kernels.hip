// HIP kernels and the C ABI of libhyqmom_hip.so (declared in include/hyqmom.h).
//
// Device layout: the state is 35 structure-of-arrays moment planes of N = Nx*Ny*Nz
// doubles, cell index i + Nx*(j + Ny*k) -- the reference's own (Nx,Ny,Nz,35) Julia layout,
// so host upload/download is a straight copy.  One step is four launches on the context
// stream, none of which waits on the host:
//   k_step_begin  : active = (t < tmax); nn += active                      (1 thread)
//   k_pass1       : per cell: closure -> x/y/z wave speeds -> closure + fluxes -> 1D
//                   speeds; writes Mr, Fx/Fy/Fz, vp, flags; wave max |vp| -> atomic max
//   k_dt          : dt = min(CFL*min(dx,dy,dz)/vmax, min(dtmax,Kn), tmax-t); t += dt
//   k_update      : per cell: x/y/z HLL updates with the reference's boundary rules,
//                   combined as Mx+My+Mz-2M
//   k_pass2       : per cell: closure -> x/y wave-speed corrections -> closure -> BGK
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hyqmom.h"
#include "hyq_pipeline.h"
#include "hyq_update.h"

using namespace hyq;

extern "C" int hyq_mlp_infer_async(hyq_mlp* m, const float* X, int64_t n, float* P,
                                   hipStream_t stream);
extern "C" int hyq_mlp_infer_masked_async(hyq_mlp* m, const float* X, int64_t n, float* P,
                                          hipStream_t stream);
extern "C" int hyq_mlp_infer_bf16_masked_async(hyq_mlp* m, const short* Xb, int64_t n, float* P,
                                               hipStream_t stream);
extern "C" int hyq_mlp_infer_bf16_segments_async(hyq_mlp* m, const short* Xb, const uint32_t* idx,
                                                 const uint32_t* seg, int64_t nseg, int32_t seg_rows,
                                                 float* P, hipStream_t stream);
extern "C" const float* hyq_mlp_input_norm(hyq_mlp* m);

namespace {

thread_local std::string g_err;

int fail(int code, const char* what, hipError_t e = hipSuccess) {
  g_err = what;
  if (e != hipSuccess) {
    g_err += ": ";
    g_err += hipGetErrorString(e);
  }
  return code;
}
#define HIPCHK(x, msg)                                   \
  do {                                                   \
    hipError_t _e = (x);                                 \
    if (_e != hipSuccess) return fail(HYQ_EHIP, msg, _e); \
  } while (0)

constexpr int kBlock = 128;
// Minimum waves per SIMD the register allocator must allow for the pass kernels: more waves
// hide more latency at the price of spills to scratch.  Measured on the 120x120x60 workload,
// 2 beats 1 for every pass kernel, and the fast kernels are slower at 3 or 4.
constexpr int kPassWaves = 2, kFastWaves = 2;
// The fast modes' projection kernels (k_pass1a, k_pass2_relax) need about 310 registers:
// one wave per SIMD holds them without scratch, where two waves spill 180-190 B per lane;
// measured 5-15 us faster each at 120x120x60 (k_pass2_check, 196 B of spill at two waves,
// was 35 us slower at one).
constexpr int kProjWaves = 1;

constexpr int kNev = 5;  // timing events per step
constexpr const char* kNonfiniteMsg =
    "non-finite Jacobian in the z wave-speed solve (eigenvalues6z_hyperbolic_3D.jl:108,121 throws)";

struct DevScalars {  // device-resident step scalars
  double t, dt, tmax, active;
  unsigned long long vmax_bits;
  long long nn, nnmax;
  unsigned long long err;  // bit 0: a non-finite z Jacobian (where the reference throws)
  // fast modes, this step: cells on the full pass 1 / pass 2 (nhard), cells with pending
  // planes (nslot), pending planes (nitem), and cells a pending plane sent to the full path
  // (ncplx: a complex pair, the eigenvalue correction)
  unsigned int nhard1, nhard2, nslot1, nslot2, nitem1, nitem2, ncplx1, ncplx2;
};

// ---- pass 1 -------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

template <typename FT, int EM>
__global__ __launch_bounds__(kBlock, EM == 0 ? kPassWaves : kFastWaves) void k_pass1(const double* __restrict__ M,
                                                  double* __restrict__ Mr, FT* __restrict__ F,
                                                  FT* __restrict__ vp, uint32_t* __restrict__ flags,
                                                  DevScalars* sc, long long N, int flag2D,
                                                  double Ma, int positivity) {
  if (sc != nullptr && sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long vb = 0;
  if (c < N) {
    double m[35], mr[35], v[6];
#pragma unroll
    for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * N + c];
    uint32_t fl = 0;
    pass1_cell<FT, EM>(m, flag2D, Ma, positivity != 0, mr, F + c, F + (size_t)35 * N + c,
                       F + (size_t)70 * N + c, (size_t)N, v, fl);
#pragma unroll
    for (int n = 0; n < 35; ++n) Mr[(size_t)n * N + c] = mr[n];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      vp[(size_t)d * N + c] = (FT)v[d];
      const unsigned long long b = __double_as_longlong(fabs(v[d]));
      vb = b > vb ? b : vb;  // |vp| >= 0: bit order = value order; NaN sorts above +Inf
    }
    if (flags) flags[c] = fl;
    if (sc != nullptr && (fl & F_NONFIN)) atomicOr(&sc->err, 1ull);
  }
  if (sc != nullptr) {
    vb = wave_max_u64(vb);
    if ((threadIdx.x & 63) == 0) atomicMax(&sc->vmax_bits, vb);
  }
}

// ---- dt (simulation_runner.jl:402-433, :496) -------------------------------------------
__global__ void k_step_begin(DevScalars* sc) {
  const bool act = (sc->t < sc->tmax) && (sc->nn < sc->nnmax);
  sc->active = act ? 1.0 : 0.0;
  if (act) sc->nn += 1;
  sc->vmax_bits = 0ull;
  sc->nhard1 = 0u;
  sc->nhard2 = 0u;
  sc->nslot1 = sc->nslot2 = sc->nitem1 = sc->nitem2 = sc->ncplx1 = sc->ncplx2 = 0u;
}
__global__ void k_dt(DevScalars* sc, double CFL, double hmin, double dtmax, double Kn) {
  if (sc->active == 0.0) return;
  const double vmax = __longlong_as_double((long long)sc->vmax_bits);
  double dt = jmin(CFL * hmin / vmax, jmin(dtmax, Kn));
  dt = jmin(dt, sc->tmax - sc->t);
  sc->dt = dt;
  sc->t += dt;
}

// ---- HLL update (the line logic is in hyq_update.h) ----------------------------------
// FT = double: FP64 fluxes and arithmetic; FT = float: the mixed-precision difference form.
constexpr int kUpdBlock = 256;  // 128: 2-3% slower; 64: 5-9%; 512: 3% (120x120x60)
HD inline unsigned update_grid(long long n) { return (unsigned)((n + kUpdBlock - 1) / kUpdBlock); }
template <typename FT>
__global__ __launch_bounds__(kUpdBlock) void k_update(const double* __restrict__ Mr,
                                                   const FT* __restrict__ F,
                                                   const FT* __restrict__ vp,
                                                   double* __restrict__ Mout, const DevScalars* sc,
                                                   int Nx, int Ny, int Nz, double dx, double dy,
                                                   double dz) {
  if (sc->active == 0.0) return;
  const long long N = (long long)Nx * Ny * Nz;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int i = (int)(c % Nx);
  const int j = (int)((c / Nx) % Ny);
  const int k = (int)(c / ((long long)Nx * Ny));
  const double dt = sc->dt;
  update_cell<FT>(Mr, F, vp, Mout, c, i, j, k, Nx, Ny, Nz, N, dt / dx, dt / dy, dt / dz);
}

// ---- pass 2 + BGK ----------------------------------------------------------------------
template <bool MIXED, int EM>
__global__ __launch_bounds__(kBlock, EM == 0 ? kPassWaves : kFastWaves) void k_pass2(double* __restrict__ M, const DevScalars* sc,
                                                  long long N, int flag2D, double Ma,
                                                  int positivity, double Kn) {
  if (sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double m[35], o[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * N + c];
  uint32_t fl = 0;
  pass2_cell<MIXED, EM>(m, flag2D, Ma, positivity != 0, sc->dt, Kn, o, fl);
#pragma unroll
  for (int n = 0; n < 35; ++n) M[(size_t)n * N + c] = o[n];
}

// ---- batch kernels ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_closure_batch(const double* __restrict__ M, long long N,
                                                          double Ma, int flag2D, double* Mr,
                                                          double* Fx, double* Fy, double* Fz,
                                                          uint32_t* flags) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double m[35], mr[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * N + c];
  uint32_t fl = 0;
  if (Fx != nullptr && Fy != nullptr && Fz != nullptr) {
    flux_closure35<true>(m, flag2D, Ma, mr, Fx + c, Fy + c, Fz + c, (size_t)N, fl);
  } else {
    flux_closure35<false, double>(m, flag2D, Ma, mr, nullptr, nullptr, nullptr, 0, fl);
  }
  if (Mr)
#pragma unroll
    for (int n = 0; n < 35; ++n) Mr[(size_t)n * N + c] = mr[n];
  if (flags) flags[c] = fl;
}

// Moments5_3D of n independent cells: M [35][n] in, M5, C5, S5 [21][n] out.
__global__ __launch_bounds__(kBlock) void k_moments5_batch(const double* __restrict__ M,
                                                           long long N, double* __restrict__ M5,
                                                           double* __restrict__ C5,
                                                           double* __restrict__ S5) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double m[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * N + c];
  moments5_cell(m, M5 + c, C5 + c, S5 + c, (size_t)N);
}

__global__ __launch_bounds__(kBlock) void k_collision_batch(const double* __restrict__ M,
                                                            long long N, double dt, double Kn,
                                                            int positivity, double* Mo) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double m[35], o[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * N + c];
  collision_cell(m, dt, Kn, positivity != 0, o);
#pragma unroll
  for (int n = 0; n < 35; ++n) Mo[(size_t)n * N + c] = o[n];
}

// closure_and_eigenvalues speeds of n 5-moment sets ([5][n]).
__global__ __launch_bounds__(kBlock) void k_closure1d_batch(const double* __restrict__ m5,
                                                            long long N, double* vmin,
                                                            double* vmax) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double a, b;
  closure1d_speeds(m5[c], m5[N + c], m5[2 * N + c], m5[3 * N + c], m5[4 * N + c], a, b);
  vmin[c] = a;
  vmax[c] = b;
}

// pas_HLL on L independent lines of np cells (cell s of line l at l*np + s), FP64: the
// x-direction part of the step's update (hyq_update.h) on its own.
__global__ __launch_bounds__(kBlock) void k_hll_lines(const double* __restrict__ M,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ vmin,
                                                      const double* __restrict__ vmax,
                                                      long long N, int np, double dtdx,
                                                      double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int s = (int)(c % np);
  const UpdDir<double> d = upd_dir(vmin + c, vmax + c, s, np, 1);
  const long long l0 = c - s;  // the line's first cell: offsets within the line
  const UpdOff o = upd_off<double, double>(d, (unsigned)s);
  for (int n = 0; n < 35; ++n)
    out[(size_t)n * N + c] =
        upd_value(d, upd_load(o, M + (size_t)n * N + l0, F + (size_t)n * N + l0), dtdx);
}

// eigvals of jacobian6 for a 15-moment slice (jacobian6.jl argument order), sorted by real
// part; computed on the standardized plane and mapped back.
__global__ __launch_bounds__(kBlock) void k_eig6_batch(const double* __restrict__ m15, long long N,
                                                       double* wr_out, double* wi_out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double r[15];
#pragma unroll
  for (int t = 0; t < 15; ++t) r[t] = m15[(size_t)t * N + c];
  // raw plane moments R[i][j] = m_ij / m00 (i: flux direction)
  double R[5][5];
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int b = 0; b < 5; ++b) R[a][b] = 0.0;
  const double i0 = 1.0 / r[0];
  R[0][0] = 1.0; R[0][1] = r[1] * i0; R[0][2] = r[2] * i0; R[0][3] = r[3] * i0; R[0][4] = r[4] * i0;
  R[1][0] = r[5] * i0; R[1][1] = r[6] * i0; R[1][2] = r[7] * i0; R[1][3] = r[8] * i0;
  R[2][0] = r[9] * i0; R[2][1] = r[10] * i0; R[2][2] = r[11] * i0; R[3][0] = r[12] * i0;
  R[3][1] = r[13] * i0; R[4][0] = r[14] * i0;
  // embed as a 3D packed array with the third axis empty and shift to central moments
  double cc[35];
#pragma unroll
  for (int t = 0; t < 35; ++t) cc[t] = 0.0;
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int b = 0; a + b < 5; ++b) cc[tpos(a, b, 0)] = R[a][b];
  const double a = R[1][0], b = R[0][1];
  taylor_shift<0, 4>(cc, -a);
  taylor_shift<1, 4>(cc, -b);
  Plane p = plane<0, 1>(cc, a, b, 0.0);
  double H[36];
  jacobian_std(p, H);
  double wr[6], wi[6];
  bool fin = p.ok;
#pragma unroll
  for (int t = 0; t < 36; ++t) fin = fin && (fabs(H[t]) < 1e300);
  if (fin) {
    balance_n<6>(H);
    hessenberg_n<6>(H);
    fin = hqr_n<6>(H, wr, wi);
  }
  for (int t = 0; t < 6; ++t) {
    if (!fin) { wr[t] = NAN; wi[t] = NAN; }
    else { wr[t] = p.a + p.sig * wr[t]; wi[t] = p.sig * wi[t]; }
  }
  // insertion sort by real part
  for (int x = 1; x < 6; ++x)
    for (int y = x; y > 0 && wr[y] < wr[y - 1]; --y) {
      double tt = wr[y]; wr[y] = wr[y - 1]; wr[y - 1] = tt;
      tt = wi[y]; wi[y] = wi[y - 1]; wi[y - 1] = tt;
    }
  for (int t = 0; t < 6; ++t) {
    wr_out[(size_t)t * N + c] = wr[t];
    wi_out[(size_t)t * N + c] = wi[t];
  }
}

// ---- learned warm start: plane features and the split pass 1 -------------------------
// The five distinct plane Jacobians of a cell, in warm-start slot order: UV, UW, VU, VW, WU.
// Feature row of a plane: its 15 standardized moments in jacobian6.jl argument order
// (S00 S01 S02 S03 S04 S10 S11 S12 S13 S20 S21 S22 S30 S31 S40), padded to 16.
__device__ void plane_feature(const Plane& p, float* x) {
  x[0] = 1.f; x[1] = 0.f; x[2] = 1.f; x[3] = (float)p.S03; x[4] = (float)p.S04;
  x[5] = 0.f; x[6] = (float)p.S11; x[7] = (float)p.S12; x[8] = (float)p.S13;
  x[9] = 1.f; x[10] = (float)p.S21; x[11] = (float)p.S22; x[12] = (float)p.S30;
  x[13] = (float)p.S31; x[14] = (float)p.S40;
  // column 15 is not a network input (normalization zeroes it): it marks the planes the
  // step's masked inference must evaluate (the others take the Gaussian seed)
  x[15] = plane_near_gaussian(p) ? 0.f : 1.f;
}
__device__ void cell_planes(const double* M, Plane pl[5]) {
  double c[35], M000, um, vm, wm;
  raw_to_central4(M, c, M000, um, vm, wm);
  pl[0] = plane<0, 1>(c, um, vm, wm);
  pl[1] = plane<0, 2>(c, um, vm, wm);
  pl[2] = plane<1, 0>(c, um, vm, wm);
  pl[3] = plane<1, 2>(c, um, vm, wm);
  pl[4] = plane<2, 0>(c, um, vm, wm);
}
// Features (and, with L != null, labels from the full FP64 solve: standardized lmin, lmax and
// a validity mask -- finite and real) of the five planes of n cells; sample = plane*n + cell.
__global__ __launch_bounds__(kBlock) void k_plane_dataset(const double* __restrict__ M, long long n,
                                                          float* __restrict__ X, float* L,
                                                          uint8_t* valid) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double m[35];
#pragma unroll
  for (int t = 0; t < 35; ++t) m[t] = M[(size_t)t * n + c];
  Plane pl[5];
  cell_planes(m, pl);
  for (int q = 0; q < 5; ++q) {
    float x[16];
    plane_feature(pl[q], x);
    x[15] = 0.f;
    float4* dst = reinterpret_cast<float4*>(X + ((size_t)q * n + c) * 16);
#pragma unroll
    for (int t = 0; t < 4; ++t) dst[t] = make_float4(x[4 * t], x[4 * t + 1], x[4 * t + 2], x[4 * t + 3]);
    if (L) {
      uint32_t fl = 0;
      Plane ps = pl[q];
      ps.a = 0.0;
      ps.sig = 1.0;  // standardized eigenvalues
      const EigRes r = plane_eigs(ps, fl);
      L[((size_t)q * n + c) * 2] = (float)r.vmin;
      L[((size_t)q * n + c) * 2 + 1] = (float)r.vmax;
      valid[(size_t)q * n + c] = (r.ok && !r.cplx) ? 1 : 0;
    }
  }
}
// Pass 1, first half: projection (state written to A) and, with FEAT, the network's work.
// FEAT: a plane whose standardized moments are within kNetTol of a Gaussian's takes the
// Gaussian's extreme eigenvalues, -sqrt(3), +sqrt(3), as its seeds (written to P directly;
// NaN for a plane that is not finite); the others are handed to the network as compacted
// rows, normalized with `norm` (16 means, 16 standard deviations) and rounded to bf16 as the
// network's first layer takes them (the same FP32 operations and rounding as its own
// normalization; column 15 is 1).  Each wave owns a segment of kSegRows rows of Xb, one row
// per lane and plane: its rows, plane by plane in lane order, go to the front of its
// segment with their plane indices q N + c in idx, and the wave writes its row count to
// seg[wave].  No atomics: the network's kernel reads each segment's count.
constexpr unsigned kSegRows = 320;  // 64 lanes x 5 planes
__device__ __forceinline__ short to_bf16(float f) { return __builtin_bit_cast(short, (__bf16)f); }
template <bool FEAT>
__global__ __launch_bounds__(kBlock, kProjWaves) void k_pass1a(const double* __restrict__ M, double* __restrict__ A,
                                                   short* __restrict__ Xb, unsigned* __restrict__ idx,
                                                   unsigned* __restrict__ seg, float* __restrict__ P,
                                                   uint32_t* __restrict__ flags, DevScalars* sc,
                                                   long long N, int flag2D, double Ma,
                                                   const float* __restrict__ norm) {
  if (sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;  // whole waves past N never run; the last wave's lane 0 is inside
  double m[35], a[35];
#pragma unroll
  for (int t = 0; t < 35; ++t) m[t] = HYQ_LOAD_STREAM(&M[(size_t)t * N + c]);
  uint32_t fl = 0;
  pass1a_cell(m, flag2D, Ma, a, fl);
#pragma unroll
  for (int t = 0; t < 35; ++t) HYQ_STORE_STREAM(&A[(size_t)t * N + c], a[t]);
  flags[c] = fl;
  if (!FEAT) return;
  Plane pl[5];
  cell_planes(a, pl);
  const unsigned lane = threadIdx.x & 63;
  const long long wave = c / 64;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned used = 0u;  // rows of the earlier planes (the same on every lane)
  for (int q = 0; q < 5; ++q) {
    const bool need = pl[q].ok && !(plane_gauss_dev(pl[q]) < kNetTol);
    const unsigned long long want = __ballot(need);  // the wave's active lanes
    if (need) {
      const size_t row = (size_t)wave * kSegRows + used + (unsigned)__popcll(want & below);
      float x[16];
      plane_feature(pl[q], x);
      typedef short s8v __attribute__((ext_vector_type(8)));
      s8v h[2];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const float v = t < 15 ? (x[t] - norm[t]) / norm[16 + t] : 1.0f;
        h[t / 8][t % 8] = to_bf16(v);
      }
      s8v* dst = reinterpret_cast<s8v*>(Xb + row * 16);
      dst[0] = h[0];
      dst[1] = h[1];
      idx[row] = (unsigned)((long long)q * N + c);
    } else {
      const float g = pl[q].ok ? 1.7320508075688772f : NAN;
      P[((size_t)q * N + c) * 2] = -g;
      P[((size_t)q * N + c) * 2 + 1] = g;
    }
    used += (unsigned)__popcll(want);
  }
  if (lane == 0) seg[wave] = used;
}

// Spread of a deferred-cell list of n entries over every wave of the grid: each wave takes
// `per` consecutive entries (per = ceil(n / waves), at most 64) on its first `per` lanes, and
// the grid strides by waves * per.  Few cells per wave keep the union of their branch paths,
// which a wave executes, short; the deferred cells are the rare, branch-heavy ones.
// LANES lanes per entry (1, or 2: a lane pair that shares one cell's work, see PairSolve).
template <unsigned LANES = 1>
struct ListSpread {
  unsigned n, per, wave, lane, stride;
  __device__ explicit ListSpread(unsigned n_) : n(n_) {
    const unsigned waves = gridDim.x * (blockDim.x / 64);
    per = (n + waves - 1) / waves;
    per = per < 1 ? 1 : (per > 64 / LANES ? 64 / LANES : per);
    wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    lane = (threadIdx.x & 63) / LANES;
    stride = waves * per;
  }
  __device__ bool idle() const { return wave * per >= n; }  // the same for every lane of a wave
  __device__ unsigned first() const { return wave * per + lane; }
  __device__ bool mine(unsigned i) const { return lane < per && i < n; }  // same on a pair
};
// Waves per SIMD of the list kernels (the full path's stages, the slot solver): their lists
// are short, so each kernel's time is one wave's latency through its stage, which is
// shortest with the whole register file (1 wave per SIMD: 256 VGPRs and 256 AGPRs, almost
// no scratch), not with more waves.  Grid: one wave per SIMD (256 CUs x 4 SIMDs), so that
// every wave with work runs at once, in 128-thread blocks.
constexpr int kFullWaves = 1, kSolveWaves = 1;
constexpr unsigned kListBlocks = 256 * 4 * kFullWaves / 2;
// Pass 2's slot solver (which fits two waves per SIMD: it needs no extremes) and fix-up
// kernel, which sorts out the slots (up to a quarter of the cells in the middle of a run),
// run on a grid of two waves per SIMD, with grid-stride loops over the slots.  (Pass 1's
// fix-up, which only combines speeds, was slower on this grid: more waves, each with its
// atomic wave maximum.)
constexpr unsigned kFixupBlocks = 256 * 4 * kFastWaves / 2;

// ---- the fast step (mixed_newton / mixed_learned modes) ------------------------------------
// A cell whose planes the fast solve proves hyperbolic finishes in the fast kernel.  A plane
// the small fast kernel cannot decide (a nearly double eigenvalue, whose pair the 2x2 block
// test decides, or an inconclusive Sturm chain, which the QR decides) leaves its cell
// pending: the fast kernel still computes the cell's closure, fluxes and 1D speeds, and
// stores the projected state and the proven planes' extremes in a slot.  A slot kernel then
// decides the pending planes, and a fix-up kernel combines the cell's speeds, unless a
// plane has a complex pair: the eigenvalue correction changes the state, and the cell takes
// the full path, as do cells whose z path clamps the density or re-projects and cells with
// a non-finite plane.  Every cell gives the values the full path would: the same functions
// on the same inputs.  The heavy deciding code stays out of the kernels that every cell
// runs, which keeps their registers (and spills) down.
struct FastLists {
  int* cells;          // [N] cells for the full path
  int* slot_cell;      // [N] cell of each slot
  unsigned* slot_mask; // [N] pending planes (bits 0..4)
  unsigned* slot_bad;  // [N] kSlotBad: a pending plane has a complex pair or is not finite
  uint32_t* slot_fl;   // [N] the cell's flags
  double* slot_v;      // [16][N] pass 1: plane extremes (10), 1D speeds (6)
  double* slot_A;      // [35][N] the projected state the planes come from
  double* full_S;      // [35][N] the state of each full-path cell between its stages
  double* full_v;      // [6][N] its x/y/z speeds
  uint32_t* full_fl;   // [N] its flags
  unsigned* full_known;  // [N] pass 2: the decisions of its x/y planes on full_S (known_of)
};
// A slot whose pending planes include a complex pair or a non-finite plane: slot_bad holds
// this bit; its low bits hold the decisions the plane solves reached (known_of).
constexpr unsigned kSlotBad = 0x80000000u;
// Appends cell c with its projected state A to the full path's list; returns its index.
__device__ __forceinline__ unsigned new_hard(const FastLists& L, unsigned* count, long long c,
                                             const double* A, uint32_t fl, long long N) {
  const unsigned i = atomicAdd(count, 1u);
  L.cells[i] = (int)c;
#pragma unroll
  for (int t = 0; t < 35; ++t) L.full_S[(size_t)t * N + i] = A[t];
  L.full_fl[i] = fl;
  return i;
}
// Moves slot s (cell c) to the full path's list with the projected state it holds.
__device__ __forceinline__ unsigned slot_to_hard(const FastLists& L, unsigned* count, unsigned s,
                                                 long long c, uint32_t fl, long long N) {
  const unsigned i = atomicAdd(count, 1u);
  L.cells[i] = (int)c;
#pragma unroll
  for (int t = 0; t < 35; ++t) L.full_S[(size_t)t * N + i] = L.slot_A[(size_t)t * N + s];
  L.full_fl[i] = fl;
  return i;
}
__device__ __forceinline__ unsigned new_slot(const FastLists& L, unsigned* count, long long c,
                                             unsigned pend, const double* A, long long N) {
  const unsigned slot = atomicAdd(count, 1u);
  L.slot_cell[slot] = (int)c;
  L.slot_mask[slot] = pend;
  L.slot_bad[slot] = 0u;
#pragma unroll
  for (int t = 0; t < 35; ++t) L.slot_A[(size_t)t * N + slot] = A[t];
  return slot;
}

// Pass 1 from the projected state that k_pass1a left in AMr (and, in the learned mode, the
// network's seeds in P).
// The common cells' wave speeds: the five fast plane solves of the projected state A; a
// proven cell's six plane extremes (min and max over its two planes per direction) go to
// vp, for k_pass1_close to combine with the 1D speeds; a cell with pending planes gets a
// slot, one for the full path a list entry.  The closure and the fluxes of every cell are
// k_pass1_close's, after the deferred cells: this kernel holds the plane solves' registers
// only (96 B of scratch per lane at two waves per SIMD, against 568 B with the closure).
template <typename FT>
__global__ __launch_bounds__(kBlock, kFastWaves) void k_pass1_fast(
    const double* __restrict__ A, FT* __restrict__ vp, uint32_t* __restrict__ flags,
    const float* __restrict__ P, FastLists L, DevScalars* sc, long long N, int positivity) {
  if (sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double a[35];
#pragma unroll
  for (int t = 0; t < 35; ++t) a[t] = A[(size_t)t * N + c];
  const uint32_t fl0 = flags[c];
  uint32_t fl = fl0;
  Warm w = no_warm();
  if (P != nullptr) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      w.w[2 * q] = P[((size_t)q * N + c) * 2];
      w.w[2 * q + 1] = P[((size_t)q * N + c) * 2 + 1];
    }
  }
  double v[10];
  unsigned pend;
  const int st = fast_cell_planes<false, true>(a, positivity != 0, w, v, pend, fl);
  if (st == kPlaneFull) {
    new_hard(L, &sc->nhard1, c, a, fl0, N);  // the full path starts from the projection
    return;
  }
  if (st == kPlaneSolve) {
    const unsigned slot = new_slot(L, &sc->nslot1, c, pend, a, N);
    L.slot_fl[slot] = fl;
#pragma unroll
    for (int t = 0; t < 10; ++t) L.slot_v[(size_t)t * N + slot] = v[t];
    return;
  }
  double v6[6];
  combine_planes(v, v6);
#pragma unroll
  for (int d = 0; d < 6; ++d) vp[(size_t)d * N + c] = (FT)v6[d];
  flags[c] = fl;
}

// Pass 1's closure and fluxes of every cell, after the deferred cells: A (the projected
// state, or the full path's final state) to its closure Mr in place, the fluxes, the 1D
// speeds of Mr combined with the plane extremes the speed kernels left in vp (min and max
// commute with rounding to FT, so combining the rounded extremes gives the same FT speeds),
// the flags, the wave maximum |vp|.  One wave per SIMD: the closure with fluxes needs about
// 400 registers.
template <typename FT>
__global__ __launch_bounds__(kBlock, kProjWaves) void k_pass1_close(
    double* __restrict__ AMr, FT* __restrict__ F, FT* __restrict__ vp, uint32_t* __restrict__ flags,
    DevScalars* sc, long long N, int flag2D, double Ma) {
  if (sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long vb = 0;
  if (c < N) {
    double a[35], mr[35], s1[6], v6[6], vd[6];
#pragma unroll
    for (int t = 0; t < 35; ++t) a[t] = HYQ_LOAD_STREAM(&AMr[(size_t)t * N + c]);
    uint32_t fl = flags[c];
    flux_closure35<true, FT>(a, flag2D, Ma, mr, F + c, F + (size_t)35 * N + c,
                             F + (size_t)70 * N + c, (size_t)N, fl);
    closure1d_all(mr, s1);
#pragma unroll
    for (int d = 0; d < 6; ++d) v6[d] = (double)vp[(size_t)d * N + c];
    combine_vp(s1, v6, vd);
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      vp[(size_t)d * N + c] = (FT)vd[d];
      const unsigned long long bb = __double_as_longlong(fabs(vd[d]));
      vb = bb > vb ? bb : vb;
    }
#pragma unroll
    for (int t = 0; t < 35; ++t) HYQ_STORE_STREAM(&AMr[(size_t)t * N + c], mr[t]);
    flags[c] = fl;
    if (fl & F_NONFIN) atomicOr(&sc->err, 1ull);
  }
  vb = wave_max_u64(vb);
  if ((threadIdx.x & 63) == 0) atomicMax(&sc->vmax_bits, vb);
}

// The pending planes of every slot: the fast solve with the 2x2 block test of a nearly double
// pair, the QR where that does not decide.  One item per (slot, plane), items spread over
// every wave of the grid (ListSpread), one plane per lane: the kernel's time is one plane
// solve's latency, not that of a cell's pending planes in turn.  A plane's results go to
// the slot: its extremes to slot_v (pass 1), its flags and iterations to slot_fl by atomic
// OR and add (the iteration field, bits 16-27, holds at most 5 x 2 x 40 per pass and the
// fast kernel's as many, far below its 4095 cap), its decision (known_of) to slot_bad, whose
// kSlotBad bit marks a complex pair or a non-finite plane.  VALS: pass 1 (extremes; P: the
// network's seeds or null); pass 2 only needs the decisions.
template <bool VALS>
__global__ __launch_bounds__(kBlock, kSolveWaves) void k_slot_solve(FastLists L, const unsigned* count,
                                                                    unsigned* nplanes,
                                                                    const float* __restrict__ P,
                                                                    DevScalars* sc, long long N) {
  if (sc->active == 0.0) return;
  constexpr unsigned NP = VALS ? 5u : 4u;  // planes per slot: pass 2 checks x and y only
  const ListSpread<> ls(*count * NP);
  if (ls.idle()) return;
  for (unsigned i = ls.first(); i < ls.n; i += ls.stride) {
    if (!ls.mine(i)) continue;
    const unsigned s = i / NP, q = i % NP;
    const unsigned pend = L.slot_mask[s];
    if (q == 0) atomicAdd(nplanes, (unsigned)__popc(pend));
    if (!(pend & (1u << q))) continue;
    double A[35];
#pragma unroll
    for (int t = 0; t < 35; ++t) A[t] = L.slot_A[(size_t)t * N + s];
    float wmin = NAN, wmax = NAN;
    if (VALS && P != nullptr) {
      const long long c = L.slot_cell[s];
      wmin = P[((size_t)q * N + c) * 2];
      wmax = P[((size_t)q * N + c) * 2 + 1];
    }
    uint32_t fl = 0;
    double v0 = 0.0, v1 = 0.0;
    const unsigned d = solve_plane(A, (int)q, VALS, wmin, wmax, v0, v1, fl);
    if (d != kKnownReal) {
      atomicOr(&L.slot_bad[s], kSlotBad | (d << (2 * q)));
      continue;
    }
    if (NP == 4) atomicOr(&L.slot_bad[s], d << (2 * q));  // pass 2: the decision for the full path
    if (VALS) {
      L.slot_v[(size_t)(2 * q) * N + s] = v0;
      L.slot_v[(size_t)(2 * q + 1) * N + s] = v1;
      atomicOr(&L.slot_fl[s], fl & 0xFFFFu);
      atomicAdd(&L.slot_fl[s], fl & 0xFFFF0000u);
    }
  }
}

// Pass 1 of the cells with plane solves: their speeds, or the full path after a complex pair.
template <typename FT>
__global__ __launch_bounds__(kBlock) void k_pass1_fixup(FastLists L, FT* __restrict__ vp,
                                                        uint32_t* __restrict__ flags, DevScalars* sc,
                                                        long long N) {
  if (sc->active == 0.0) return;
  const unsigned n = sc->nslot1;
  for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const long long c = L.slot_cell[s];
    if (L.slot_bad[s] & kSlotBad) {
      atomicAdd(&sc->ncplx1, 1u);
      slot_to_hard(L, &sc->nhard1, s, c, flags[c], N);  // flags[c]: those of the projection
      continue;
    }
    double v[10], v6[6];
#pragma unroll
    for (int t = 0; t < 10; ++t) v[t] = L.slot_v[(size_t)t * N + s];
    combine_planes(v, v6);
#pragma unroll
    for (int d = 0; d < 6; ++d) vp[(size_t)d * N + c] = (FT)v6[d];
    flags[c] = L.slot_fl[s];
  }
}

// The full pass 1 of the listed cells, in stages: x, y and z wave speeds (with their
// eigenvalue corrections, the z density clamp and re-projection) of the projected state; the
// closure with the fluxes is k_pass1_close's, with every other cell.  Each stage is one
// kernel over the list, the cell's state kept in full_S between stages: the same functions,
// on the same values, as pass1_cell, so the results are the same bits, but each kernel
// holds one stage's registers instead of the whole path's, which keeps its latency (the
// kernel's time on a short list) down.  P: the network's seeds (learned mode) or null.
__device__ __forceinline__ Warm list_warm(const float* P, long long c, long long N) {
  Warm w = no_warm();
  if (P != nullptr) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      w.w[2 * q] = P[((size_t)q * N + c) * 2];
      w.w[2 * q + 1] = P[((size_t)q * N + c) * 2 + 1];
    }
  }
  return w;
}
// ---- lane pairs: the two plane solves of an axis side by side -------------------------------
// A full-path cell's stage is a chain of latencies (one wave per SIMD, a few cells per
// wave), and each axis solves its two planes one after the other.  In pass 1's stage
// kernels a lane pair shares a cell: both lanes hold the same state and run the same code, except
// that lane `half` = 0 solves the axis's first plane and lane 1 its second, and the two
// exchange their results (and the flags their solves set) by a lane swap.  The values, the
// decisions and the flags are those of eig6_axis / eig6_z: the same plane_solve on the same
// planes, the flags OR-ed, the root-polishing iterations added (the field cannot reach its
// 4095 cap: at most 5 planes x 2 solves x 80 per pass, and as many in the fast kernel).
__device__ __forceinline__ EigRes swap_res(const EigRes& r) {
  EigRes o;
  o.vmin = __shfl_xor(r.vmin, 1);
  o.vmax = __shfl_xor(r.vmax, 1);
  o.cplx = __shfl_xor((int)r.cplx, 1) != 0;
  o.ok = __shfl_xor((int)r.ok, 1) != 0;
  return o;
}
HD uint32_t merge_flags(uint32_t fl, uint32_t fa, uint32_t fb) {
  const uint32_t it = (fl >> 16) + (fa >> 16) + (fb >> 16);
  return ((fl | fa | fb) & 0xFFFFu) | ((it < 0xFFFu ? it : 0xFFFu) << 16);
}
// Planes a and b solved on the two lanes of a pair; A, B: both results on both lanes.
template <bool CORRECTS>
__device__ __forceinline__ void pair_solve(const Plane& pa, const Plane& pb, float wa0, float wa1,
                                           float wb0, float wb1, bool need_vals, uint32_t& fl,
                                           int half, EigRes& A, EigRes& B) {
  uint32_t f = 0;
  const EigRes r = plane_solve<1, CORRECTS>(half ? pb : pa, half ? wb0 : wa0, half ? wb1 : wa1,
                                            need_vals, f);
  const EigRes o = swap_res(r);
  const uint32_t fo = (uint32_t)__shfl_xor((int)f, 1);
  A = half ? o : r;
  B = half ? r : o;
  fl = merge_flags(fl, half ? fo : f, half ? f : fo);
}
// eig6_axis<AXIS, 1> on a lane pair.
template <int AXIS>
__device__ bool eig6_axis_pair(const double* M, int flag2D, double Ma, double& vmin, double& vmax,
                               double* Mr, uint32_t& fl, const Warm& wm_, bool need_vals, int half) {
  double c[35];
  double M000, um, vm, wm;
  raw_to_central4(M, c, M000, um, vm, wm);
  const int ia = AXIS == 0 ? 0 : 2, ib = AXIS == 0 ? 1 : 3;  // Warm slots: UV UW VU VW WU
  EigRes A, B;
  pair_solve<true>(AXIS == 0 ? plane<0, 1>(c, um, vm, wm) : plane<1, 0>(c, um, vm, wm),
                   AXIS == 0 ? plane<0, 2>(c, um, vm, wm) : plane<1, 2>(c, um, vm, wm),
                   wm_.w[2 * ia], wm_.w[2 * ia + 1], wm_.w[2 * ib], wm_.w[2 * ib + 1], need_vals, fl,
                   half, A, B);
  combine_guarded(A, B, vmin, vmax);
  if (A.cplx || B.cplx) {
    axis_correct<AXIS>(c, M000, um, vm, wm, A.cplx, B.cplx, flag2D, Ma, Mr, fl);
    if (!need_vals) return true;
    raw_to_central4(Mr, c, M000, um, vm, wm);
    pair_solve<false>(AXIS == 0 ? plane<0, 1>(c, um, vm, wm) : plane<1, 0>(c, um, vm, wm),
                      AXIS == 0 ? plane<0, 2>(c, um, vm, wm) : plane<1, 2>(c, um, vm, wm),
                      wm_.w[2 * ia], wm_.w[2 * ia + 1], wm_.w[2 * ib], wm_.w[2 * ib + 1], need_vals,
                      fl, half, A, B);
    combine_guarded(A, B, vmin, vmax);
    return true;
  }
#pragma unroll
  for (int n = 0; n < 35; ++n) Mr[n] = M[n];
  return false;
}
// eig6_z<1> on a lane pair: the density clamp and the re-projection on both lanes, the WU
// and VW solves side by side.
__device__ void eig6_z_pair(const double* M, int flag2D, double Ma, bool positivity, double& vmin,
                            double& vmax, double* Mr, uint32_t& fl, const Warm& wm_, int half) {
#pragma unroll
  for (int n = 0; n < 35; ++n) Mr[n] = M[n];
  if (positivity && M[0] <= 0.0) {
    Mr[0] = jmax(M[0], 1e-14);
    fl |= F_RHO;
  }
  double W[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) W[n] = M[n];
  W[0] = Mr[0];
  if (positivity && W[0] > 1e-15) {
    const double m0 = W[0];
    const double vx = M[2] / m0 - (M[1] / m0) * (M[1] / m0);
    const double vy = M[9] / m0 - (M[5] / m0) * (M[5] / m0);
    const double vz = M[19] / m0 - (M[15] / m0) * (M[15] / m0);
    if (vx < 0.0 || vy < 0.0 || vz < 0.0) {
      fl |= F_ZNEG;
      flux_closure35<false, double>(M, flag2D, Ma, Mr, nullptr, nullptr, nullptr, 0, fl);
#pragma unroll
      for (int n = 0; n < 35; ++n) W[n] = Mr[n];
    }
  }
  double c[35];
  double M000, um, vm, wm;
  raw_to_central4(W, c, M000, um, vm, wm);
  EigRes A, B;
  pair_solve<true>(plane<2, 0>(c, um, vm, wm), plane<1, 2>(c, um, vm, wm), wm_.w[8], wm_.w[9],
                   wm_.w[6], wm_.w[7], true, fl, half, A, B);
  if (!A.ok || !B.ok) fl |= F_NONFIN;  // the reference's eigvals throws here
  vmin = jmin(A.vmin, B.vmin);
  vmax = jmax(A.vmax, B.vmax);
  if (A.cplx || B.cplx) {
    fl |= F_ECZ;
    raw_to_central4(M, c, M000, um, vm, wm);  // corrections start from the input M
    double cn[35];
#pragma unroll
    for (int t = 0; t < 35; ++t) cn[t] = c[t];
    if (A.cplx) correct_plane<0, 2>(c, cn);
    if (B.cplx) correct_plane<1, 2>(c, cn);
    central_to_raw<4>(cn, M000, um, vm, wm);
    double Mh[35];
    pack35(cn, Mh);
    flux_closure35<false, double>(Mh, flag2D, Ma, Mr, nullptr, nullptr, nullptr, 0, fl);
    raw_to_central4(Mr, c, M000, um, vm, wm);
    pair_solve<false>(plane<2, 0>(c, um, vm, wm), plane<1, 2>(c, um, vm, wm), wm_.w[8], wm_.w[9],
                      wm_.w[6], wm_.w[7], true, fl, half, A, B);
    vmin = jmin(A.vmin, B.vmin);
    vmax = jmax(A.vmax, B.vmax);
  }
}
// The list holds each cell's projected state (pass1a_cell, computed once by k_pass1a) and
// the flags of that projection.  STAGE 1: x; 2: y; 3: z.
template <int STAGE, typename FT>
__global__ __launch_bounds__(kBlock, kFullWaves) void k_full1_stage(
    const float* __restrict__ P, FastLists L, DevScalars* sc, long long N, int flag2D, double Ma,
    int positivity, double* __restrict__ AMr, FT* __restrict__ vp, uint32_t* __restrict__ flags) {
  if (sc->active == 0.0) return;
  const ListSpread<2> ls(sc->nhard1);  // a lane pair per cell
  if (ls.idle()) return;  // wave-uniform
  const int half = threadIdx.x & 1;
  for (unsigned i = ls.first(); i < ls.n; i += ls.stride) {
    if (!ls.mine(i)) continue;
    const long long c = L.cells[i];
    double a[35], b[35];
#pragma unroll
    for (int t = 0; t < 35; ++t) a[t] = L.full_S[(size_t)t * N + i];
    uint32_t fl = L.full_fl[i];
    const Warm w = list_warm(P, c, N);
    double mn, mx;
    if (STAGE == 1) eig6_axis_pair<0>(a, flag2D, Ma, mn, mx, b, fl, w, true, half);
    if (STAGE == 2) eig6_axis_pair<1>(a, flag2D, Ma, mn, mx, b, fl, w, true, half);
    if (STAGE == 3) eig6_z_pair(a, flag2D, Ma, positivity != 0, mn, mx, b, fl, w, half);
    if (half) continue;  // both lanes hold the same results
    if (STAGE < 3) {
      L.full_v[(size_t)(2 * (STAGE - 1)) * N + i] = mn;
      L.full_v[(size_t)(2 * (STAGE - 1) + 1) * N + i] = mx;
#pragma unroll
      for (int t = 0; t < 35; ++t) L.full_S[(size_t)t * N + i] = b[t];
      L.full_fl[i] = fl;
      continue;
    }
    // the cell's final state, speeds and flags, where k_pass1_close takes every cell's
#pragma unroll
    for (int d = 0; d < 4; ++d) vp[(size_t)d * N + c] = (FT)L.full_v[(size_t)d * N + i];
    vp[(size_t)4 * N + c] = (FT)mn;
    vp[(size_t)5 * N + c] = (FT)mx;
#pragma unroll
    for (int t = 0; t < 35; ++t) AMr[(size_t)t * N + c] = b[t];
    flags[c] = fl;
  }
}
// The full pass 1's stages on `stream`, then the closure and fluxes of every cell.
template <typename FT>
static void enqueue_full1(hipStream_t stream, double* AMr, FT* F, FT* vp, uint32_t* flags,
                          const float* P, const FastLists& L, DevScalars* sc, long long N,
                          int flag2D, double Ma, int positivity) {
  hipLaunchKernelGGL((k_full1_stage<1, FT>), dim3(kListBlocks), dim3(kBlock), 0, stream, P, L, sc, N,
                     flag2D, Ma, positivity, AMr, vp, flags);
  hipLaunchKernelGGL((k_full1_stage<2, FT>), dim3(kListBlocks), dim3(kBlock), 0, stream, P, L, sc, N,
                     flag2D, Ma, positivity, AMr, vp, flags);
  hipLaunchKernelGGL((k_full1_stage<3, FT>), dim3(kListBlocks), dim3(kBlock), 0, stream, P, L, sc, N,
                     flag2D, Ma, positivity, AMr, vp, flags);
  hipLaunchKernelGGL(k_pass1_close<FT>, dim3((unsigned)((N + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     stream, AMr, F, vp, flags, sc, N, flag2D, Ma);
}

// The fast modes' pass 2: projection and the x/y hyperbolicity checks (projected state to
// A; A[0] = NaN marks a cell that k_pass2_relax leaves alone), the plane solves, the fix-up
// of the cells with plane solves (those proven hyperbolic back into A, the others to the
// full path), BGK of every proven cell, and the full pass 2 of the rest.
__global__ __launch_bounds__(kBlock, kFastWaves) void k_pass2_check(const double* __restrict__ M,
                                                                   double* __restrict__ A,
                                                                   FastLists L, DevScalars* sc,
                                                                   long long N, int flag2D,
                                                                   double Ma) {
  if (sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double m[35], a[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = HYQ_LOAD_STREAM(&M[(size_t)n * N + c]);
  uint32_t fl = 0;
  unsigned pend, known;
  const int st = pass2_check_planes<false>(m, flag2D, Ma, a, pend, fl, &known);
  if (st == kPlaneProven) {
#pragma unroll
    for (int n = 0; n < 35; ++n) HYQ_STORE_STREAM(&A[(size_t)n * N + c], a[n]);
    return;
  }
  A[c] = NAN;
  if (st == kPlaneFull)
    L.full_known[new_hard(L, &sc->nhard2, c, a, 0u, N)] = known;
  else
    new_slot(L, &sc->nslot2, c, pend, a, N);
}

__global__ __launch_bounds__(kBlock, kProjWaves) void k_pass2_relax(double* __restrict__ M,
                                                                   const double* __restrict__ A,
                                                                   DevScalars* sc, long long N,
                                                                   int flag2D, double Ma,
                                                                   int positivity, double Kn) {
  if (sc->active == 0.0) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double a0 = A[c];
  if (a0 != a0) return;
  double a[35], o[35];
  a[0] = a0;
#pragma unroll
  for (int n = 1; n < 35; ++n) a[n] = A[(size_t)n * N + c];
  uint32_t fl = 0;
  pass2_relax<true>(a, flag2D, Ma, positivity != 0, sc->dt, Kn, o, fl);
#pragma unroll
  for (int n = 0; n < 35; ++n) M[(size_t)n * N + c] = o[n];
}

// Pass 2 of the cells with plane solves: back into A for k_pass2_relax, or the full path
// after a complex pair.
__global__ __launch_bounds__(kBlock, kFastWaves) void k_pass2_fixup(double* __restrict__ A,
                                                                   FastLists L, DevScalars* sc,
                                                                   long long N) {
  if (sc->active == 0.0) return;
  const unsigned n = sc->nslot2;
  for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const long long c = L.slot_cell[s];
    const unsigned bad = L.slot_bad[s];
    if (bad & kSlotBad) {
      atomicAdd(&sc->ncplx2, 1u);
      const unsigned i = slot_to_hard(L, &sc->nhard2, s, c, 0u, N);
      // the planes that were not pending were proven hyperbolic by the check
      unsigned known = bad & 0xFFu;
      const unsigned pend = L.slot_mask[s];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (!(pend & (1u << q))) known |= kKnownReal << (2 * q);
      L.full_known[i] = known;
      continue;
    }
    // proven hyperbolic: back into the relaxation kernel's input A, which then relaxes it
    // with the cells the check proved
#pragma unroll
    for (int t = 0; t < 35; ++t) A[(size_t)t * N + c] = L.slot_A[(size_t)t * N + s];
  }
}

// The full pass 2 of the listed cells, in stages like the full pass 1: the x and y
// corrections of the projected state the check left in the list (with the plane decisions the
// check and the plane solves reached on it).  Stage 2 hands the corrected state to
// k_pass2_relax through A, which relaxes it with every other cell (the same BGK on the same
// state); a state whose density is not a number, which would read as A's "not proven"
// mark, is relaxed here instead.  STAGE 1: x; 2: y.
template <int STAGE>
__global__ __launch_bounds__(kBlock, kFullWaves) void k_full2_stage(FastLists L, DevScalars* sc,
                                                                   long long N, int flag2D,
                                                                   double Ma, double* __restrict__ A,
                                                                   double* __restrict__ M,
                                                                   int positivity, double Kn) {
  if (sc->active == 0.0) return;
  const ListSpread<> ls(sc->nhard2);
  if (ls.idle()) return;
  for (unsigned i = ls.first(); i < ls.n; i += ls.stride) {
    if (!ls.mine(i)) continue;
    double a[35], b[35];
    uint32_t fl = 0;
#pragma unroll
    for (int t = 0; t < 35; ++t) a[t] = L.full_S[(size_t)t * N + i];
    const unsigned known = L.full_known[i];
    // the wave speeds of pass 2 are not used, only its corrections; after an x correction
    // the y decisions no longer hold
    const bool corrected = STAGE == 1 ? eig6_axis_decided<0>(a, flag2D, Ma, b, fl, known)
                                      : eig6_axis_decided<1>(a, flag2D, Ma, b, fl, known);
    if (STAGE == 1) {
      if (corrected) L.full_known[i] = known & 0x0Fu;
#pragma unroll
      for (int t = 0; t < 35; ++t) L.full_S[(size_t)t * N + i] = b[t];
      continue;
    }
    const long long c = L.cells[i];
    if (b[0] == b[0]) {
#pragma unroll
      for (int t = 0; t < 35; ++t) A[(size_t)t * N + c] = b[t];
    } else {
      double o[35];
      pass2_relax<true>(b, flag2D, Ma, positivity != 0, sc->dt, Kn, o, fl);
#pragma unroll
      for (int t = 0; t < 35; ++t) M[(size_t)t * N + c] = o[t];
    }
  }
}

// ---- snapshot fields ----------------------------------------------------------------
// compute_central_field / compute_standardized_field (src/utils/compute_standardized_field.jl:
// 16-62): M2CS4_35 (src/moments/M2CS4_35.jl) per cell.  STD = 0: the central moments C4
// (c000 = 1, first order 0); STD = 1: the standardized S4, C4[ijk] / (s_x^i s_y^j s_z^k)
// with s = sqrt(max(C2, eps)) per axis, 1 for the order-0 and pure second-order entries, 0
// for the first order.  Out: [35][N] planes.
template <int STD>
__global__ __launch_bounds__(kBlock) void k_snapshot_field(const double* __restrict__ M,
                                                           long long N, double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double m[35], cm[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) m[n] = M[(size_t)n * N + c];
  double M000, u, v, w;
  raw_to_central4(m, cm, M000, u, v, w);
  const double sx = sqrt(jmax(cm[tpos(2, 0, 0)], EPS));
  const double sy = sqrt(jmax(cm[tpos(0, 2, 0)], EPS));
  const double sz = sqrt(jmax(cm[tpos(0, 0, 2)], EPS));
#pragma unroll
  for (int n = 0; n < 35; ++n) {
    const int i = e35().e[n][0], j = e35().e[n][1], k = e35().e[n][2];
    const double cv = cm[T35(n)];
    double r = cv;
    if (STD) {
      const int o = i + j + k;
      if (o == 0) r = 1.0;
      else if (o == 1) r = 0.0;
      else if (o == 2 && (i == 2 || j == 2 || k == 2)) r = 1.0;
      else {
        double d = 1.0;
        for (int t = 0; t < i; ++t) d *= sx;
        for (int t = 0; t < j; ++t) d *= sy;
        for (int t = 0; t < k; ++t) d *= sz;
        r = cv / d;
      }
    }
    out[(size_t)n * N + c] = r;
  }
}

// ---- runner diagnostics -------------------------------------------------------------
// Symmetry test of simulation_runner.jl:565-628 on the k = 1 slice: the diagonal cells
// (d, d), d < Nd = min(Nx, Ny), against their mirrors (Nd-1-d, Nd-1-d) for the first five
// moments, differences for moments 1, 3, 5 and sums for 2, 4 (as the reference pairs them);
// per moment max(Diff) / (|Diff|_2 + 1), and the largest magnitude of the five.  One block.
__global__ __launch_bounds__(256) void k_symmetry(const double* __restrict__ M, long long N,
                                                  int Nx, int Ny, double* __restrict__ out) {
  __shared__ double smax[5][256], ssq[5][256];
  const int tid = threadIdx.x;
  const int nd = Nx < Ny ? Nx : Ny;
  double mx[5], sq[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    mx[n] = -INFINITY;
    sq[n] = 0.0;
  }
  for (int d = tid; d < nd; d += 256) {
    const long long a = d + (long long)Nx * d, b = (nd - 1 - d) + (long long)Nx * (nd - 1 - d);
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      const double va = M[(size_t)n * N + a], vb = M[(size_t)n * N + b];
      const double df = (n & 1) ? va + vb : va - vb;
      mx[n] = df > mx[n] ? df : mx[n];
      sq[n] += df * df;
    }
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    smax[n][tid] = mx[n];
    ssq[n][tid] = sq[n];
  }
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        smax[n][tid] = smax[n][tid + h] > smax[n][tid] ? smax[n][tid + h] : smax[n][tid];
        ssq[n][tid] += ssq[n][tid + h];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    double r = 0.0;
    for (int n = 0; n < 5; ++n) {
      const double v = smax[n][0] / (sqrt(ssq[n][0]) + 1.0);
      out[n] = v;
      r = fabs(v) > r ? fabs(v) : r;
    }
    out[5] = r;
  }
}

// First-step diagnostics of simulation_runner.jl:404-494 on the last step's wave speeds
// vp [6][N] (x min, x max, y min, y max, z min, z max): bits[0..2] the largest |speed| per
// direction (as ordered bits of non-negative doubles; a NaN speed ranks above every number,
// as the reference's maximum propagates it), bits[4] the largest |speed| over the cells
// without a NaN speed, then the cell that reaches it first in the reference's search order
// (k, then i, then j; its `>` test never takes a NaN cell): bits[3] its key
// (k Nx + i) Ny + j, minimum over the cells that reach bits[4].
template <typename FT>
__global__ __launch_bounds__(kBlock) void k_vmax_dirs(const FT* __restrict__ vp, long long N,
                                                      unsigned long long* __restrict__ bits) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long b[3] = {0ull, 0ull, 0ull}, fin = 0ull;
  if (c < N) {
    bool nan = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double lo = fabs((double)vp[(size_t)(2 * d) * N + c]);
      const double hi = fabs((double)vp[(size_t)(2 * d + 1) * N + c]);
      const bool dn = lo != lo || hi != hi;
      nan = nan || dn;
      b[d] = dn ? 0x7FF8000000000000ull : __double_as_longlong(hi > lo ? hi : lo);
    }
    if (!nan) {
      fin = b[0] > b[1] ? b[0] : b[1];
      fin = b[2] > fin ? b[2] : fin;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const unsigned long long w = wave_max_u64(b[d]);
    if ((threadIdx.x & 63) == 0) atomicMax(&bits[d], w);
  }
  const unsigned long long wf = wave_max_u64(fin);
  if ((threadIdx.x & 63) == 0) atomicMax(&bits[4], wf);
}
template <typename FT>
__global__ __launch_bounds__(kBlock) void k_vmax_worst(const FT* __restrict__ vp, long long N,
                                                       int Nx, int Ny,
                                                       unsigned long long* __restrict__ bits) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const unsigned long long g = bits[4];
  unsigned long long m = 0ull;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const double v = fabs((double)vp[(size_t)t * N + c]);
    if (v != v) return;
    const unsigned long long vb = __double_as_longlong(v);
    m = vb > m ? vb : m;
  }
  if (m != g) return;
  const long long i = c % Nx, j = (c / Nx) % Ny, k = c / ((long long)Nx * Ny);
  atomicMin(&bits[3], (unsigned long long)((k * Nx + i) * Ny + j));
}

unsigned grid_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------------------
// Context
// Steps that hyq_run queues before it reads the device scalars back: a run stops within one
// chunk of reaching tmax, however large nnmax is.
constexpr int kChunk = 64;

struct hyq_ctx {
  hyq_params p;
  long long N;
  double dx, dy, dz;
  hipStream_t stream;
  double* M;       // [35][N]
  double* Mr;      // [35][N]
  void* F;         // [105][N] double (FP64) or float (mixed)
  void* vp;        // [6][N]
  uint32_t* flags; // [N]
  DevScalars* sc;
  hipEvent_t ev[kNev * kChunk];  // per step: pass 1 | dt | update | pass 2 (a ring of kChunk steps)
  double ms_pass1, ms_update, ms_pass2;
  long long timed_steps;
  hyq_mlp* mlp;  // learned mode: warm-start network (not owned)
  float* X;        // learned mode: the network's input rows, a segment of kSegRows rows of
                   // 16 bf16 per wave of 64 cells
  unsigned* Xidx;  // [segments x kSegRows] plane index q N + c of each row
  unsigned* Xseg;  // [ceil(N / 64)] rows in each wave's segment
  float* P;        // [5N][2] seeds of each plane: the network's predictions or the Gaussian's
  FastLists L;             // fast modes: full-path list, slots and plane items
  double* diag;            // [8] results of the diagnostics
};

extern "C" {

const char* hyq_last_error(void) { return g_err.c_str(); }
const char* hyq_version(void) { return "hyqmom-mi355x 0.2 (gfx950)"; }

int hyq_create(const hyq_params* params, hyq_ctx** out) {
  if (!params || !out) return fail(HYQ_EINVAL, "null argument");
  const hyq_params& p = *params;
  if (p.Nmom != 35) return fail(HYQ_EINVAL, "Nmom must be 35");
  if (p.Nx < 1 || p.Ny < 1 || p.Nz < 1) return fail(HYQ_EINVAL, "grid sizes must be >= 1");
  if ((long long)p.Nx * p.Ny * p.Nz >= kMaxUpdateCells)
    return fail(HYQ_EINVAL, "more than 2^29 - 1 cells");
  if (p.precision < HYQ_PREC_FP64 || p.precision > HYQ_PREC_MIXED_NEWTON)
    return fail(HYQ_EINVAL, "unknown precision mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(HYQ_ENODEV, "no HIP device");
  hyq_ctx* c = new hyq_ctx();
  std::memset(c, 0, sizeof(*c));
  c->p = p;
  c->N = (long long)p.Nx * p.Ny * p.Nz;
  c->dx = (p.xmax - p.xmin) / p.Nx;  // simulation_runner.jl:133-139
  c->dy = (p.ymax - p.ymin) / p.Ny;
  c->dz = (p.zmax - p.zmin) / p.Nz;
  const size_t fsz = p.precision == HYQ_PREC_FP64 ? sizeof(double) : sizeof(float);
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    delete c;
    return fail(HYQ_EHIP, "hipStreamCreate", e);
  }
  bool ok = hipMalloc(&c->M, sizeof(double) * 35 * c->N) == hipSuccess &&
            hipMalloc(&c->Mr, sizeof(double) * 35 * c->N) == hipSuccess &&
            hipMalloc(&c->F, fsz * 105 * c->N) == hipSuccess &&
            hipMalloc(&c->vp, fsz * 6 * c->N) == hipSuccess &&
            hipMalloc(&c->flags, sizeof(uint32_t) * c->N) == hipSuccess &&
            hipMalloc(&c->sc, sizeof(DevScalars)) == hipSuccess &&
            hipMalloc(&c->diag, sizeof(double) * 8) == hipSuccess;
  if (ok && p.precision == HYQ_PREC_MIXED_LEARNED)
    ok = hipMalloc(&c->X, sizeof(short) * 16 * kSegRows * ((c->N + 63) / 64)) == hipSuccess &&
         hipMalloc(&c->Xidx, sizeof(unsigned) * kSegRows * ((c->N + 63) / 64)) == hipSuccess &&
         hipMalloc(&c->Xseg, sizeof(unsigned) * ((c->N + 63) / 64)) == hipSuccess &&
         hipMalloc(&c->P, sizeof(float) * 2 * 5 * c->N) == hipSuccess &&
         // no rows until a step's projection has listed some: the network's kernel then
         // reads counts of 0 (it runs, as every kernel of the step, also in a step that has
         // nothing left to do)
         hipMemset(c->Xseg, 0, sizeof(unsigned) * ((c->N + 63) / 64)) == hipSuccess;
  if (ok && (params->precision == HYQ_PREC_MIXED_NEWTON || params->precision == HYQ_PREC_MIXED_LEARNED)) {
    FastLists& L = c->L;
    ok = hipMalloc(&L.cells, sizeof(int) * c->N) == hipSuccess &&
         hipMalloc(&L.slot_cell, sizeof(int) * c->N) == hipSuccess &&
         hipMalloc(&L.slot_mask, sizeof(unsigned) * c->N) == hipSuccess &&
         hipMalloc(&L.slot_bad, sizeof(unsigned) * c->N) == hipSuccess &&
         hipMalloc(&L.slot_fl, sizeof(uint32_t) * c->N) == hipSuccess &&
         hipMalloc(&L.slot_v, sizeof(double) * 16 * c->N) == hipSuccess &&
         hipMalloc(&L.slot_A, sizeof(double) * 35 * c->N) == hipSuccess &&
         hipMalloc(&L.full_S, sizeof(double) * 35 * c->N) == hipSuccess &&
         hipMalloc(&L.full_v, sizeof(double) * 6 * c->N) == hipSuccess &&
         hipMalloc(&L.full_fl, sizeof(uint32_t) * c->N) == hipSuccess &&
         hipMalloc(&L.full_known, sizeof(unsigned) * c->N) == hipSuccess;
  }
  if (!ok) {
    hyq_destroy(c);
    return fail(HYQ_ENOMEM, "device allocation failed");
  }
  for (int i = 0; i < kNev * kChunk; ++i) {
    if ((e = hipEventCreate(&c->ev[i])) != hipSuccess) {
      hyq_destroy(c);
      return fail(HYQ_EHIP, "hipEventCreate", e);
    }
  }
  *out = c;
  return HYQ_OK;
}

int hyq_destroy(hyq_ctx* c) {
  if (!c) return HYQ_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  hipFree(c->M);
  hipFree(c->Mr);
  hipFree(c->F);
  hipFree(c->vp);
  hipFree(c->flags);
  hipFree(c->sc);
  hipFree(c->diag);
  hipFree(c->X);
  hipFree(c->Xidx);
  hipFree(c->Xseg);
  hipFree(c->P);
  hipFree(c->L.cells);
  hipFree(c->L.slot_cell);
  hipFree(c->L.slot_bad);
  hipFree(c->L.slot_fl);
  hipFree(c->L.slot_v);
  hipFree(c->L.slot_mask);
  hipFree(c->L.slot_A);
  hipFree(c->L.full_S);
  hipFree(c->L.full_v);
  hipFree(c->L.full_fl);
  hipFree(c->L.full_known);
  for (int i = 0; i < kNev * kChunk; ++i)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return HYQ_OK;
}

int hyq_set_state(hyq_ctx* c, const double* M) {
  if (!c || !M) return fail(HYQ_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(c->M, M, sizeof(double) * 35 * c->N, hipMemcpyHostToDevice, c->stream),
         "set_state copy");
  HIPCHK(hipStreamSynchronize(c->stream), "set_state sync");
  return HYQ_OK;
}

int hyq_get_state(hyq_ctx* c, double* M) {
  if (!c || !M) return fail(HYQ_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(M, c->M, sizeof(double) * 35 * c->N, hipMemcpyDeviceToHost, c->stream),
         "get_state copy");
  HIPCHK(hipStreamSynchronize(c->stream), "get_state sync");
  return HYQ_OK;
}

// One step on the context stream; ev (or null): kNev events bracketing its parts.
static int enqueue_step(hyq_ctx* c, hipEvent_t* ev) {
  const hyq_params& p = c->p;
  const unsigned g = grid_for(c->N);
  const double hmin = jmin(jmin(c->dx, c->dy), c->dz);
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, c->stream, c->sc);
  if (ev) hipEventRecord(ev[0], c->stream);
  const bool fast = p.precision == HYQ_PREC_MIXED_NEWTON || p.precision == HYQ_PREC_MIXED_LEARNED;
  if (p.precision == HYQ_PREC_FP64) {
    hipLaunchKernelGGL((k_pass1<double, 0>), dim3(g), dim3(kBlock), 0, c->stream, c->M, c->Mr,
                       (double*)c->F, (double*)c->vp, c->flags, c->sc, c->N, p.flag2D, p.Ma,
                       p.positivity);
  } else if (p.precision == HYQ_PREC_MIXED) {
    hipLaunchKernelGGL((k_pass1<float, 0>), dim3(g), dim3(kBlock), 0, c->stream, c->M, c->Mr,
                       (float*)c->F, (float*)c->vp, c->flags, c->sc, c->N, p.flag2D, p.Ma,
                       p.positivity);
  } else {
    // fast modes: projection (+ the network's input rows and warm starts in the learned
    // mode), the common cells, then the deferred ones
    const bool learned = p.precision == HYQ_PREC_MIXED_LEARNED && c->mlp != nullptr;
    if (learned) {
      hipLaunchKernelGGL(k_pass1a<true>, dim3(g), dim3(kBlock), 0, c->stream, c->M, c->Mr,
                         reinterpret_cast<short*>(c->X), c->Xidx, c->Xseg, c->P, c->flags, c->sc,
                         c->N, p.flag2D, p.Ma, hyq_mlp_input_norm(c->mlp));
      const int rc = hyq_mlp_infer_bf16_segments_async(
          c->mlp, reinterpret_cast<const short*>(c->X), c->Xidx, c->Xseg, (c->N + 63) / 64,
          (int32_t)kSegRows, c->P, c->stream);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL(k_pass1a<false>, dim3(g), dim3(kBlock), 0, c->stream, c->M, c->Mr,
                         (short*)nullptr, (unsigned*)nullptr, (unsigned*)nullptr, (float*)nullptr,
                         c->flags, c->sc, c->N, p.flag2D, p.Ma, (const float*)nullptr);
    }
    const float* P = learned ? (const float*)c->P : (const float*)nullptr;
    hipLaunchKernelGGL(k_pass1_fast<float>, dim3(g), dim3(kBlock), 0, c->stream, (const double*)c->Mr,
                       (float*)c->vp, c->flags, P, c->L, c->sc, c->N, p.positivity);
    hipLaunchKernelGGL(k_slot_solve<true>, dim3(kListBlocks), dim3(kBlock), 0, c->stream, c->L,
                       (const unsigned*)&c->sc->nslot1, &c->sc->nitem1, P, c->sc, c->N);
    hipLaunchKernelGGL(k_pass1_fixup<float>, dim3(kListBlocks), dim3(kBlock), 0, c->stream, c->L,
                       (float*)c->vp, c->flags, c->sc, c->N);
    enqueue_full1<float>(c->stream, c->Mr, (float*)c->F, (float*)c->vp, c->flags, P, c->L,
                         c->sc, c->N, p.flag2D, p.Ma, p.positivity);
  }
  if (ev) hipEventRecord(ev[1], c->stream);
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, c->stream, c->sc, p.CFL, hmin, p.dtmax, p.Kn);
  if (ev) hipEventRecord(ev[2], c->stream);
  if (p.precision == HYQ_PREC_FP64) {
    hipLaunchKernelGGL(k_update<double>, dim3(update_grid(c->N)), dim3(kUpdBlock), 0, c->stream, c->Mr,
                       (const double*)c->F, (const double*)c->vp, c->M, c->sc, p.Nx, p.Ny, p.Nz,
                       c->dx, c->dy, c->dz);
  } else {
    hipLaunchKernelGGL(k_update<float>, dim3(update_grid(c->N)), dim3(kUpdBlock), 0, c->stream, c->Mr,
                       (const float*)c->F, (const float*)c->vp, c->M, c->sc, p.Nx, p.Ny, p.Nz,
                       c->dx, c->dy, c->dz);
  }
  if (ev) hipEventRecord(ev[3], c->stream);
  if (p.precision == HYQ_PREC_FP64) {
    hipLaunchKernelGGL((k_pass2<false, 0>), dim3(g), dim3(kBlock), 0, c->stream, c->M, c->sc,
                       c->N, p.flag2D, p.Ma, p.positivity, p.Kn);
  } else if (!fast) {
    hipLaunchKernelGGL((k_pass2<true, 0>), dim3(g), dim3(kBlock), 0, c->stream, c->M, c->sc,
                       c->N, p.flag2D, p.Ma, p.positivity, p.Kn);
  } else {
    hipLaunchKernelGGL(k_pass2_check, dim3(g), dim3(kBlock), 0, c->stream, (const double*)c->M,
                       c->Mr, c->L, c->sc, c->N, p.flag2D, p.Ma);
    hipLaunchKernelGGL(k_slot_solve<false>, dim3(kFixupBlocks), dim3(kBlock), 0, c->stream, c->L,
                       (const unsigned*)&c->sc->nslot2, &c->sc->nitem2, (const float*)nullptr, c->sc,
                       c->N);
    hipLaunchKernelGGL(k_pass2_fixup, dim3(kFixupBlocks), dim3(kBlock), 0, c->stream, c->Mr, c->L,
                       c->sc, c->N);
    hipLaunchKernelGGL(k_full2_stage<1>, dim3(kListBlocks), dim3(kBlock), 0, c->stream, c->L,
                       c->sc, c->N, p.flag2D, p.Ma, c->Mr, c->M, p.positivity, p.Kn);
    hipLaunchKernelGGL(k_full2_stage<2>, dim3(kListBlocks), dim3(kBlock), 0, c->stream, c->L,
                       c->sc, c->N, p.flag2D, p.Ma, c->Mr, c->M, p.positivity, p.Kn);
    hipLaunchKernelGGL(k_pass2_relax, dim3(g), dim3(kBlock), 0, c->stream, c->M,
                       (const double*)c->Mr, c->sc, c->N, p.flag2D, p.Ma, p.positivity, p.Kn);
  }
  if (ev) hipEventRecord(ev[4], c->stream);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(HYQ_EHIP, "kernel launch", e);
  return HYQ_OK;
}

static int load_scalars(hyq_ctx* c, double t, long long nn, double tmax, long long nnmax) {
  DevScalars s;
  std::memset(&s, 0, sizeof(s));
  s.t = t;
  s.tmax = tmax;
  s.nn = nn;
  s.nnmax = nnmax;
  HIPCHK(hipMemcpyAsync(c->sc, &s, sizeof(s), hipMemcpyHostToDevice, c->stream), "scalars");
  return HYQ_OK;
}

static int accumulate_times(hyq_ctx* c, hipEvent_t* ev) {
  float a = 0, b = 0, d = 0;
  HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]), "event time");
  HIPCHK(hipEventElapsedTime(&b, ev[2], ev[3]), "event time");
  HIPCHK(hipEventElapsedTime(&d, ev[3], ev[4]), "event time");
  c->ms_pass1 += a;
  c->ms_update += b;
  c->ms_pass2 += d;
  c->timed_steps += 1;
  return HYQ_OK;
}

int hyq_step(hyq_ctx* c, double* t_inout, double* dt_out, int64_t* nn_inout) {
  if (!c || !t_inout || !nn_inout) return fail(HYQ_EINVAL, "null argument");
  // one pass of the loop body; a no-op if t >= tmax (the caller owns the loop condition)
  int rc = load_scalars(c, *t_inout, *nn_inout, c->p.tmax, (long long)*nn_inout + 1);
  if (rc) return rc;
  rc = enqueue_step(c, c->ev);
  if (rc) {
    hipStreamSynchronize(c->stream);
    return rc;
  }
  DevScalars s;
  HIPCHK(hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, c->stream), "scalars back");
  HIPCHK(hipStreamSynchronize(c->stream), "step sync");
  if (s.active != 0.0) {
    rc = accumulate_times(c, c->ev);
    if (rc) return rc;
  }
  *t_inout = s.t;
  if (dt_out) *dt_out = s.active != 0.0 ? s.dt : 0.0;
  *nn_inout = s.nn;
  if (s.err) return fail(HYQ_ENONFINITE, kNonfiniteMsg);
  return HYQ_OK;
}

int hyq_run(hyq_ctx* c, double tmax, int64_t nnmax, double* t_inout, int64_t* nn_inout) {
  if (!c || !t_inout || !nn_inout) return fail(HYQ_EINVAL, "null argument");
  int rc = load_scalars(c, *t_inout, *nn_inout, tmax, nnmax);
  if (rc) return rc;
  // Steps are queued in chunks of kChunk; a step after t >= tmax (or nn >= nnmax) is a no-op
  // on the device.  After each chunk the scalars come back, and the run ends once a chunk
  // left it inactive, or on an error.  HIP events from a ring of kChunk steps bracket pass 1,
  // the update and pass 2 of every step.
  long long nn0 = *nn_inout;
  DevScalars s;
  std::memset(&s, 0, sizeof(s));
  s.t = *t_inout;
  s.nn = nn0;
  bool more = s.t < tmax && s.nn < (long long)nnmax;
  while (more) {
    const long long left = (long long)nnmax - s.nn;
    const int todo = (int)(left < kChunk ? left : kChunk);
    for (int q = 0; q < todo && rc == HYQ_OK; ++q) rc = enqueue_step(c, &c->ev[kNev * q]);
    if (rc) {
      hipStreamSynchronize(c->stream);
      return rc;
    }
    HIPCHK(hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, c->stream), "scalars back");
    HIPCHK(hipStreamSynchronize(c->stream), "run sync");
    const long long ran = s.nn - nn0;  // active steps of this chunk (they come first)
    for (long long q = 0; q < ran && rc == HYQ_OK; ++q) rc = accumulate_times(c, &c->ev[kNev * q]);
    if (rc) return rc;
    nn0 = s.nn;
    if (s.err) break;
    more = s.active != 0.0 && s.t < tmax && s.nn < (long long)nnmax;
  }
  *t_inout = s.t;
  *nn_inout = s.nn;
  if (s.err) return fail(HYQ_ENONFINITE, kNonfiniteMsg);
  return HYQ_OK;
}

int hyq_get_wave_speeds(hyq_ctx* c, double* vp) {
  if (!c || !vp) return fail(HYQ_EINVAL, "null argument");
  if (c->p.precision == HYQ_PREC_FP64) {
    HIPCHK(hipMemcpy(vp, c->vp, sizeof(double) * 6 * c->N, hipMemcpyDeviceToHost), "vp copy");
  } else {
    float* tmp = new float[6 * c->N];
    const hipError_t e = hipMemcpy(tmp, c->vp, sizeof(float) * 6 * c->N, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      for (long long i = 0; i < 6 * c->N; ++i) vp[i] = tmp[i];
    delete[] tmp;
    HIPCHK(e, "vp copy");
  }
  return HYQ_OK;
}

int hyq_snapshot_fields(hyq_ctx* c, double* C, double* S) {
  if (!c) return fail(HYQ_EINVAL, "null argument");
  // Mr is free between steps (each step's pass 1 rewrites it before reading it)
  const unsigned g = grid_for(c->N);
  for (int which = 0; which < 2; ++which) {
    double* dst = which == 0 ? C : S;
    if (!dst) continue;
    if (which == 0)
      hipLaunchKernelGGL(k_snapshot_field<0>, dim3(g), dim3(kBlock), 0, c->stream,
                         (const double*)c->M, c->N, c->Mr);
    else
      hipLaunchKernelGGL(k_snapshot_field<1>, dim3(g), dim3(kBlock), 0, c->stream,
                         (const double*)c->M, c->N, c->Mr);
    HIPCHK(hipGetLastError(), "snapshot field launch");
    HIPCHK(hipMemcpyAsync(dst, c->Mr, sizeof(double) * 35 * c->N, hipMemcpyDeviceToHost, c->stream),
           "snapshot field copy");
    HIPCHK(hipStreamSynchronize(c->stream), "snapshot field sync");
  }
  return HYQ_OK;
}

int hyq_symmetry_maxdiff(hyq_ctx* c, double* maxdiff, double* per_moment) {
  if (!c || !maxdiff) return fail(HYQ_EINVAL, "null argument");
  double h[6];
  hipLaunchKernelGGL(k_symmetry, dim3(1), dim3(256), 0, c->stream, (const double*)c->M, c->N,
                     c->p.Nx, c->p.Ny, c->diag);
  HIPCHK(hipGetLastError(), "symmetry launch");
  HIPCHK(hipMemcpyAsync(h, c->diag, sizeof(h), hipMemcpyDeviceToHost, c->stream), "symmetry copy");
  HIPCHK(hipStreamSynchronize(c->stream), "symmetry sync");
  *maxdiff = h[5];
  if (per_moment)
    for (int n = 0; n < 5; ++n) per_moment[n] = h[n];
  return HYQ_OK;
}

int hyq_vmax_breakdown(hyq_ctx* c, double* vdir, int64_t* worst) {
  if (!c || !vdir) return fail(HYQ_EINVAL, "null argument");
  unsigned long long h[5] = {0ull, 0ull, 0ull, ~0ull, 0ull};
  unsigned long long* d = reinterpret_cast<unsigned long long*>(c->diag);
  HIPCHK(hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, c->stream), "vmax init");
  const unsigned g = grid_for(c->N);
  if (c->p.precision == HYQ_PREC_FP64) {
    hipLaunchKernelGGL(k_vmax_dirs<double>, dim3(g), dim3(kBlock), 0, c->stream,
                       (const double*)c->vp, c->N, d);
    hipLaunchKernelGGL(k_vmax_worst<double>, dim3(g), dim3(kBlock), 0, c->stream,
                       (const double*)c->vp, c->N, c->p.Nx, c->p.Ny, d);
  } else {
    hipLaunchKernelGGL(k_vmax_dirs<float>, dim3(g), dim3(kBlock), 0, c->stream,
                       (const float*)c->vp, c->N, d);
    hipLaunchKernelGGL(k_vmax_worst<float>, dim3(g), dim3(kBlock), 0, c->stream,
                       (const float*)c->vp, c->N, c->p.Nx, c->p.Ny, d);
  }
  HIPCHK(hipGetLastError(), "vmax launch");
  HIPCHK(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream), "vmax copy");
  HIPCHK(hipStreamSynchronize(c->stream), "vmax sync");
  for (int k = 0; k < 3; ++k) {
    double v;
    std::memcpy(&v, &h[k], sizeof(v));
    vdir[k] = v;
  }
  if (worst && h[3] == ~0ull) {  // every cell has a NaN speed
    worst[0] = worst[1] = worst[2] = -1;
  } else if (worst) {
    const long long key = (long long)h[3], Nx = c->p.Nx, Ny = c->p.Ny;
    worst[0] = (key / Ny) % Nx;      // i
    worst[1] = key % Ny;             // j
    worst[2] = key / (Ny * Nx);      // k
  }
  return HYQ_OK;
}

int hyq_get_flags(hyq_ctx* c, uint32_t* flags) {
  if (!c || !flags) return fail(HYQ_EINVAL, "null argument");
  HIPCHK(hipMemcpy(flags, c->flags, sizeof(uint32_t) * c->N, hipMemcpyDeviceToHost), "flags copy");
  return HYQ_OK;
}

int hyq_kernel_times(hyq_ctx* c, double* ms_pass1, double* ms_update, double* ms_pass2,
                     int64_t* nsteps) {
  if (!c) return fail(HYQ_EINVAL, "null argument");
  const double n = c->timed_steps > 0 ? (double)c->timed_steps : 1.0;
  if (ms_pass1) *ms_pass1 = c->ms_pass1 / n;
  if (ms_update) *ms_update = c->ms_update / n;
  if (ms_pass2) *ms_pass2 = c->ms_pass2 / n;
  if (nsteps) *nsteps = c->timed_steps;
  return HYQ_OK;
}

int hyq_deferred_cells(hyq_ctx* c, int64_t* pass1, int64_t* pass2) {
  if (!c) return fail(HYQ_EINVAL, "null argument");
  DevScalars s;
  HIPCHK(hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, c->stream), "scalars back");
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  if (pass1) *pass1 = (int64_t)s.nhard1 + (int64_t)s.nslot1 - (int64_t)s.ncplx1;
  if (pass2) *pass2 = (int64_t)s.nhard2 + (int64_t)s.nslot2 - (int64_t)s.ncplx2;
  return HYQ_OK;
}

int hyq_fast_path_counts(hyq_ctx* c, int64_t* counts) {
  if (!c || !counts) return fail(HYQ_EINVAL, "null argument");
  DevScalars s;
  HIPCHK(hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, c->stream), "scalars back");
  HIPCHK(hipStreamSynchronize(c->stream), "sync");
  const unsigned v[8] = {s.nslot1, s.nitem1, s.ncplx1, s.nhard1 - s.ncplx1,
                         s.nslot2, s.nitem2, s.ncplx2, s.nhard2 - s.ncplx2};
  for (int k = 0; k < 8; ++k) counts[k] = (int64_t)v[k];
  return HYQ_OK;
}

int hyq_reset_kernel_times(hyq_ctx* c) {
  if (!c) return fail(HYQ_EINVAL, "null argument");
  c->ms_pass1 = c->ms_update = c->ms_pass2 = 0.0;
  c->timed_steps = 0;
  return HYQ_OK;
}

int hyq_closure_batch(const double* M, int64_t n, double Ma, int32_t flag2D, double* Mr,
                      double* Fx, double* Fy, double* Fz, uint32_t* flags) {
  if (!M || n < 0) return fail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_closure_batch, dim3(grid_for(n)), dim3(kBlock), 0, 0, M, (long long)n, Ma,
                     flag2D, Mr, Fx, Fy, Fz, flags);
  HIPCHK(hipGetLastError(), "closure launch");
  HIPCHK(hipDeviceSynchronize(), "closure sync");
  return HYQ_OK;
}

int hyq_moments5_batch(const double* M, int64_t n, double* M5, double* C5, double* S5) {
  if (!M || !M5 || !C5 || !S5 || n < 0) return fail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_moments5_batch, dim3(grid_for(n)), dim3(kBlock), 0, 0, M, (long long)n, M5,
                     C5, S5);
  HIPCHK(hipGetLastError(), "moments5 launch");
  HIPCHK(hipDeviceSynchronize(), "moments5 sync");
  return HYQ_OK;
}

static int batch_args(const double* M, int64_t n, int32_t precision, const double* Mr,
                      const double* Fx, const double* Fy, const double* Fz, const double* vp,
                      const uint32_t* flags) {
  if (!M || !Mr || !Fx || !Fy || !Fz || !vp || n < 0) return fail(HYQ_EINVAL, "bad argument");
  if (precision != HYQ_PREC_FP64 && precision != HYQ_PREC_MIXED_NEWTON)
    return fail(HYQ_EINVAL, "batch wave speeds: precision must be FP64 or MIXED_NEWTON");
  if (precision == HYQ_PREC_MIXED_NEWTON && !flags)
    return fail(HYQ_EINVAL, "the fast batch needs the flags output");
  if (n > 0x7FFFFFFFLL / 5) return fail(HYQ_EINVAL, "batch too large");
  // the three flux outputs must be one [105][n] block for the pass-1 kernels
  if (n > 0 && (Fy != Fx + (size_t)35 * n || Fz != Fx + (size_t)70 * n))
    return fail(HYQ_EINVAL, "Fx, Fy, Fz must be consecutive [35][n] blocks");
  return HYQ_OK;
}

// The fast batch's scratch for n cells, in one block: the step scalars, then the full-path
// list, the slots and the plane items of FastLists.
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
static size_t fast_batch_bytes(long long n) {
  const size_t un = (size_t)n;
  return align256(sizeof(DevScalars)) + 7 * align256(4 * un) + align256(16 * 8 * un) +
         2 * align256(35 * 8 * un) + align256(6 * 8 * un);
}
static void fast_batch_carve(char* base, long long n, DevScalars** sc, FastLists* L) {
  const size_t un = (size_t)n;
  char* p = base;
  *sc = reinterpret_cast<DevScalars*>(p); p += align256(sizeof(DevScalars));
  L->cells = reinterpret_cast<int*>(p); p += align256(4 * un);
  L->slot_cell = reinterpret_cast<int*>(p); p += align256(4 * un);
  L->slot_mask = reinterpret_cast<unsigned*>(p); p += align256(4 * un);
  L->slot_bad = reinterpret_cast<unsigned*>(p); p += align256(4 * un);
  L->slot_fl = reinterpret_cast<uint32_t*>(p); p += align256(4 * un);
  L->slot_v = reinterpret_cast<double*>(p); p += align256(16 * 8 * un);
  L->slot_A = reinterpret_cast<double*>(p); p += align256(35 * 8 * un);
  L->full_S = reinterpret_cast<double*>(p); p += align256(35 * 8 * un);
  L->full_v = reinterpret_cast<double*>(p); p += align256(6 * 8 * un);
  L->full_fl = reinterpret_cast<uint32_t*>(p); p += align256(4 * un);
  L->full_known = reinterpret_cast<unsigned*>(p);
}
// Scalars of a batch: active, everything else zero.  A kernel rather than a memset, so that
// the launch sequence can be captured in a graph.
__global__ void k_batch_begin(DevScalars* sc) {
  sc->t = sc->dt = sc->tmax = 0.0;
  sc->active = 1.0;
  sc->vmax_bits = 0ull;
  sc->nn = sc->nnmax = 0;
  sc->err = 0ull;
  sc->nhard1 = sc->nhard2 = sc->nslot1 = sc->nslot2 = 0u;
  sc->nitem1 = sc->nitem2 = sc->ncplx1 = sc->ncplx2 = 0u;
}
// Pass 1 of the fast step on a batch (FP64 outputs) with its scratch at `scr`.
static void enqueue_fast_batch(const double* M, long long n, double Ma, int flag2D, int positivity,
                               double* Mr, double* F, double* vp, uint32_t* flags, char* scr) {
  DevScalars* sc;
  FastLists L;
  fast_batch_carve(scr, n, &sc, &L);
  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(1), 0, 0, sc);
  hipLaunchKernelGGL(k_pass1a<false>, dim3(g), dim3(kBlock), 0, 0, M, Mr, (short*)nullptr,
                     (unsigned*)nullptr, (unsigned*)nullptr, (float*)nullptr, flags, sc, n, flag2D,
                     Ma, (const float*)nullptr);
  hipLaunchKernelGGL(k_pass1_fast<double>, dim3(g), dim3(kBlock), 0, 0, (const double*)Mr, vp, flags,
                     (const float*)nullptr, L, sc, n, positivity);
  hipLaunchKernelGGL(k_slot_solve<true>, dim3(kListBlocks), dim3(kBlock), 0, 0, L,
                     (const unsigned*)&sc->nslot1, &sc->nitem1, (const float*)nullptr, sc, n);
  hipLaunchKernelGGL(k_pass1_fixup<double>, dim3(kListBlocks), dim3(kBlock), 0, 0, L, vp, flags, sc, n);
  enqueue_full1<double>(0, Mr, F, vp, flags, (const float*)nullptr, L, sc, n, flag2D, Ma,
                        positivity);
}

int hyq_wavespeed_batch(const double* M, int64_t n, double Ma, int32_t flag2D,
                        int32_t positivity, int32_t precision, double* Mr, double* Fx,
                        double* Fy, double* Fz, double* vp, uint32_t* flags) {
  const int rc = batch_args(M, n, precision, Mr, Fx, Fy, Fz, vp, flags);
  if (rc) return rc;
  if (n == 0) return HYQ_OK;
  if (precision == HYQ_PREC_FP64) {
    hipLaunchKernelGGL((k_pass1<double, 0>), dim3(grid_for(n)), dim3(kBlock), 0, 0, M, Mr, Fx, vp,
                       flags, (DevScalars*)nullptr, (long long)n, flag2D, Ma, positivity);
    HIPCHK(hipGetLastError(), "pass1 launch");
    HIPCHK(hipDeviceSynchronize(), "pass1 sync");
    return HYQ_OK;
  }
  void* scr = nullptr;
  HIPCHK(hipMalloc(&scr, fast_batch_bytes(n)), "fast batch scratch");
  enqueue_fast_batch(M, n, Ma, flag2D, positivity, Mr, Fx, vp, flags, static_cast<char*>(scr));
  hipError_t e = hipGetLastError();
  const hipError_t e2 = hipDeviceSynchronize();
  if (e == hipSuccess) e = e2;
  hipFree(scr);
  if (e != hipSuccess) return fail(HYQ_EHIP, "fast batch", e);
  return HYQ_OK;
}

// ---- the same without allocation ---------------------------------------------------------
// A fixed device array holds the fast batch's scratch, for callers that cannot allocate
// (graph capture, serving loops).  Each device holds its own copy until the process ends;
// every call on a device shares it, so calls that overlap on one device race.
constexpr long long kNoAllocMaxCells = 1LL << 16;
static __device__ __align__(256) char scratch[(7 * 4 + 16 * 8 + 2 * 35 * 8 + 6 * 8) * kNoAllocMaxCells + 13 * 256];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[256];  // per device: the array's device address, once looked up

static int scratch_address(char** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 256) return fail(HYQ_ENODEV, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<char*>(g_scratch_ptr[dev]);
  return HYQ_OK;
}

int hyq_wavespeed_batch_noalloc(const double* M, int64_t n, double Ma, int32_t flag2D,
                                int32_t positivity, int32_t precision, double* Mr, double* Fx,
                                double* Fy, double* Fz, double* vp, uint32_t* flags) {
  int rc = batch_args(M, n, precision, Mr, Fx, Fy, Fz, vp, flags);
  if (rc) return rc;
  if (n > kNoAllocMaxCells)
    return fail(HYQ_EINVAL, "hyq_wavespeed_batch_noalloc: at most 65536 cells per call");
  if (n == 0) return HYQ_OK;
  if (precision == HYQ_PREC_FP64) {  // no scratch on this path
    hipLaunchKernelGGL((k_pass1<double, 0>), dim3(grid_for(n)), dim3(kBlock), 0, 0, M, Mr, Fx, vp,
                       flags, (DevScalars*)nullptr, (long long)n, flag2D, Ma, positivity);
  } else {
    if (fast_batch_bytes(n) > sizeof(scratch)) return fail(HYQ_EINVAL, "scratch too small");
    char* base = nullptr;
    rc = scratch_address(&base);
    if (rc) return rc;
    enqueue_fast_batch(M, n, Ma, flag2D, positivity, Mr, Fx, vp, flags, base);
  }
  HIPCHK(hipGetLastError(), "batch launch");
  HIPCHK(hipDeviceSynchronize(), "batch sync");
  return HYQ_OK;
}

int hyq_eig6_batch(const double* m15, int64_t n, double* wr, double* wi) {
  if (!m15 || !wr || !wi || n < 0) return fail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_eig6_batch, dim3(grid_for(n)), dim3(kBlock), 0, 0, m15, (long long)n, wr, wi);
  HIPCHK(hipGetLastError(), "eig6 launch");
  HIPCHK(hipDeviceSynchronize(), "eig6 sync");
  return HYQ_OK;
}

int hyq_set_mlp(hyq_ctx* c, hyq_mlp* mlp) {
  if (!c) return fail(HYQ_EINVAL, "null argument");
  if (mlp && c->p.precision != HYQ_PREC_MIXED_LEARNED)
    return fail(HYQ_EINVAL, "a warm-start network needs precision HYQ_PREC_MIXED_LEARNED");
  c->mlp = mlp;
  return HYQ_OK;
}

int hyq_plane_dataset(const double* M, int64_t n, float* X, float* L, uint8_t* valid) {
  if (!M || !X || n < 0 || ((L == nullptr) != (valid == nullptr)))
    return fail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_plane_dataset, dim3(grid_for(n)), dim3(kBlock), 0, 0, M, (long long)n, X, L,
                     valid);
  HIPCHK(hipGetLastError(), "dataset launch");
  HIPCHK(hipDeviceSynchronize(), "dataset sync");
  return HYQ_OK;
}

int hyq_closure1d_batch(const double* m5, int64_t n, double* vmin, double* vmax) {
  if (!m5 || !vmin || !vmax || n < 0) return fail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_closure1d_batch, dim3(grid_for(n)), dim3(kBlock), 0, 0, m5, (long long)n,
                     vmin, vmax);
  HIPCHK(hipGetLastError(), "closure1d launch");
  HIPCHK(hipDeviceSynchronize(), "closure1d sync");
  return HYQ_OK;
}

int hyq_hll_lines(const double* M, const double* F, const double* vmin, const double* vmax,
                  int64_t nlines, int32_t np, double dt, double dx, double* Mout) {
  if (!M || !F || !vmin || !vmax || !Mout || nlines < 0 || np < 1 || !(dx > 0))
    return fail(HYQ_EINVAL, "bad argument");
  if (np >= kMaxUpdateCells) return fail(HYQ_EINVAL, "line longer than 2^29 - 1 cells");
  const long long n = (long long)nlines * np;
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_hll_lines, dim3(grid_for(n)), dim3(kBlock), 0, 0, M, F, vmin, vmax, n,
                     (int)np, dt / dx, Mout);
  HIPCHK(hipGetLastError(), "hll launch");
  HIPCHK(hipDeviceSynchronize(), "hll sync");
  return HYQ_OK;
}

int hyq_collision_batch(const double* M, int64_t n, double dt, double Kn, int32_t positivity,
                        double* Mout) {
  if (!M || !Mout || n < 0) return fail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  hipLaunchKernelGGL(k_collision_batch, dim3(grid_for(n)), dim3(kBlock), 0, 0, M, (long long)n, dt,
                     Kn, positivity, Mout);
  HIPCHK(hipGetLastError(), "collision launch");
  HIPCHK(hipDeviceSynchronize(), "collision sync");
  return HYQ_OK;
}

}  // extern "C"
