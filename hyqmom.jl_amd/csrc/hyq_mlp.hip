// Learned warm start: a 15 -> H -> H -> H -> 2 MLP (tanh-form GELU; H = HYQ_MLP_HID, 32 by
// default) that predicts the extreme eigenvalues of the standardized flux Jacobian from the 15
// standardized moments of a plane.
// Matrix products run on MFMA (v_mfma_f32_16x16x32_bf16, the double-rate bf16 instruction of
// gfx950: bf16 operands, FP32 accumulate);
// weights are kept as FP32 master copies and rounded to bf16 as they are loaded.  Training
// (forward, backward, Adam) and inference use the kernels of this file only.
//
// Parameter block (FP32, this order): W1 [16][H] (input padded to 16), b1 [H], W2 [H][H],
// b2 [H], W3 [H][H], b3 [H], W4 [H][16] (outputs padded to 16, columns >= 2 held at zero),
// b4 [16], then input mean/std [16]+[16] and output mean/std [2]+[2] used to normalize
// features and labels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hyqmom.h"

namespace {

thread_local std::string g_mlp_err;
int mfail(int code, const char* what, hipError_t e = hipSuccess) {
  g_mlp_err = what;
  if (e != hipSuccess) {
    g_mlp_err += ": ";
    g_mlp_err += hipGetErrorString(e);
  }
  return code;
}
#define MCHK(x, msg)                                       \
  do {                                                     \
    hipError_t _e = (x);                                   \
    if (_e != hipSuccess) return mfail(HYQ_EHIP, msg, _e); \
  } while (0)

// The hidden width is fixed per compilation: this file is built once per supported width,
// each build with its own HYQ_MLP_HID and namespace HYQ_MLP_NS for its entry points, and
// hyq_mlp_api.hip dispatches the C ABI to the width a network was created with.
#ifndef HYQ_MLP_HID
#define HYQ_MLP_HID 32
#endif
#ifndef HYQ_MLP_NS
#define HYQ_MLP_NS mlp_h32
#endif
constexpr int IN = 16, HID = HYQ_MLP_HID, OUT = 16;
static_assert(HID % 32 == 0 && HID >= 32 && HID <= 128, "hidden width: 32, 64, 96 or 128");
constexpr int NTH = HID / 16;  // 16-column tiles of a hidden layer
constexpr int KSH = HID / 32;  // 32-deep k steps over a hidden layer
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * HID, OFF_W2 = OFF_B1 + HID,
                 OFF_B2 = OFF_W2 + HID * HID, OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * HID,
                 OFF_W4 = OFF_B3 + HID, OFF_B4 = OFF_W4 + HID * OUT, NPARAM = OFF_B4 + OUT;
constexpr size_t OFF_XMU = NPARAM, OFF_XSD = OFF_XMU + 16, OFF_YMU = OFF_XSD + 16,
                 OFF_YSD = OFF_YMU + 2, NBLOB = OFF_YSD + 2;

typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// 16x16x32 bf16 MFMA: lane l gives A row l%16 and B column l%16 with eight k values each
// (this file uses k = 8(l/16)..8(l/16)+7 of the 32-wide k step), and receives D rows
// 4(l/16)..4(l/16)+3 of column l%16.
__device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a),
                                                 __builtin_bit_cast(bf8v, b), c, 0, 0, 0);
}

__device__ __forceinline__ short f2bf(float f) {  // round to nearest even (v_cvt_pk_bf16_f32)
  return __builtin_bit_cast(short, (__bf16)f);
}

// Activation: the tanh form of GELU, z * sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3)), with the
// hardware exp and reciprocal (its inputs and outputs are rounded to bf16 around the MFMAs,
// so the few-ulp error of the fast exp and reciprocal is invisible); used identically by
// training, the layer kernels and the fused inference.
__device__ __forceinline__ float act_u(float z) { return 1.5957691216057308f * (z + 0.044715f * z * z * z); }
__device__ __forceinline__ float act_sig(float u) { return __builtin_amdgcn_rcpf(1.0f + __expf(-u)); }
__device__ __forceinline__ float gelu(float z) { return z * act_sig(act_u(z)); }
__device__ __forceinline__ float dgelu(float z) {
  const float s = act_sig(act_u(z));
  return s + z * s * (1.0f - s) * 1.5957691216057308f * (1.0f + 0.134145f * z * z);
}

// C[M x N] = A[M x K] . B[K x N] (+ bias), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].
// One wave computes rows [16 r, 16 r + 16) x columns [16 NT c, 16 NT (c+1)); lane l feeds
// A rows l%16 and B columns l%16 with the eight k values 8(l/16)..+7 of each 32-wide k step,
// and receives D rows 4(l/16)..+3 of column l%16.
// Epilogue: act 0 = none, 1 = GELU; Z gets the pre-activation, H the activation (either may
// be null); stores only rows < M and columns < Nst.
template <int NT>
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, long sam, long sak,
                                              const float* __restrict__ B, long sbk, long sbn,
                                              int M, int N, int K, const float* __restrict__ bias,
                                              float* Z, float* H, long ldc, int Nst, int act) {
  const int lane = threadIdx.x & 63;
  const long rt = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int ct = blockIdx.y;
  if (rt * 16 >= M) return;
  const int r = lane & 15, q = lane >> 4;
  const long row = rt * 16 + r;
  f4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    s8 a;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + 8 * q + e;
      const float v = (row < M && k < K) ? A[row * sam + (long)k * sak] : 0.f;
      a[e] = f2bf(v);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const long col = (long)(ct * NT + t) * 16 + r;
      s8 b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + 8 * q + e;
        const float v = (col < N && k < K) ? B[(long)k * sbk + col * sbn] : 0.f;
        b[e] = f2bf(v);
      }
      acc[t] = mfma32(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const long col = (long)(ct * NT + t) * 16 + r;
    if (col >= Nst) continue;
    const float bv = bias ? bias[col] : 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const long orow = rt * 16 + 4 * q + e;
      if (orow >= M) continue;
      const float z = acc[t][e] + bv;
      if (Z) Z[orow * ldc + col] = z;
      if (H) H[orow * ldc + col] = act == 1 ? gelu(z) : z;
    }
  }
}

// dZ = dH * GELU'(Z)
__global__ void k_gelu_bwd(const float* __restrict__ Z, float* __restrict__ dH, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dH[i] *= dgelu(Z[i]);
}
// normalized features: Xn[b][j] = (X[b][j] - mu[j]) / sd[j]  (j < 15), 0 for j = 15
__global__ void k_normalize(const float* __restrict__ X, const float* __restrict__ stat,
                            float* __restrict__ Xn, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * 16) return;
  const int j = (int)(i & 15);
  Xn[i] = j < 15 ? (X[i] - stat[j]) / stat[16 + j] : 0.0f;
}
// loss gradient of L = 1/(2B) sum (Y - Tn)^2 on the two outputs; Y rows of 16.
__global__ void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ T,
                           const float* __restrict__ ystat, float* __restrict__ dY, long n,
                           float* loss) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < n) {
    for (int o = 0; o < 16; ++o) dY[b * 16 + o] = 0.f;
    for (int o = 0; o < 2; ++o) {
      const float tn = (T[b * 2 + o] - ystat[o]) / ystat[2 + o];
      const float d = Y[b * 16 + o] - tn;
      dY[b * 16 + o] = d / (float)n;
      l += 0.5f * d * d / (float)n;
    }
  }
  // wave reduction then one atomic per wave
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}
__global__ void k_colsum(const float* __restrict__ D, long n, int ncol, float* __restrict__ out) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (long b = threadIdx.x; b < n; b += blockDim.x) s += D[b * ncol + c];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ float red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[c] = t;
  }
}
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// outputs columns >= 2 of layer 4 are held at zero
__global__ void k_mask_w4(float* __restrict__ p) {
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i >= HID * OUT + OUT) return;
  if (i < HID * OUT) {
    if ((i % OUT) >= 2) p[OFF_W4 + i] = 0.f;
  } else if (i - HID * OUT >= 2) {
    p[OFF_B4 + (i - HID * OUT)] = 0.f;
  }
}
// de-normalized predictions (lmin, lmax) from the padded output rows
__global__ void k_denorm(const float* __restrict__ Y, const float* __restrict__ ystat,
                         float* __restrict__ P, long n) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  P[2 * b] = Y[b * 16] * ystat[2] + ystat[0];
  P[2 * b + 1] = Y[b * 16 + 1] * ystat[3] + ystat[1];
}

// ---- fused inference ------------------------------------------------------------------
// All four layers in one kernel: the bf16 weights, pre-arranged in MFMA B-operand order, sit
// in LDS for the whole block; each wave takes 16-row tiles, keeps a tile's activations in a
// per-wave LDS buffer between layers (bias, GELU and bf16 rounding applied on the way), and
// writes only the two de-normalized outputs.  HBM traffic per sample: 64 B in, 8 B out.
//
// Packed weight layout: for k-step ks (32 rows of W) and n-tile t (16 columns), chunk
// (ks * NT + t) holds, for lane l, the 8 bf16 W[32 ks + 8 (l / 16) + e][16 t + l % 16],
// e = 0..7 (zero for rows past the layer's K) -- one 16-byte LDS read per lane per MFMA.
constexpr int PK_W1 = 0;                        // 1 k-step (16 rows, padded) x NTH tiles
constexpr int PK_W2 = PK_W1 + 1 * NTH * 512;    // KSH x NTH chunks of 512 bf16
constexpr int PK_W3 = PK_W2 + KSH * NTH * 512;
constexpr int PK_W4 = PK_W3 + KSH * NTH * 512;  // KSH k-steps x 1 tile
constexpr int PK_N = PK_W4 + KSH * 1 * 512;     // bf16 elements
constexpr int ACT_LD = HID + 8;               // bf16 row pitch of the activation tile (16 B aligned)

__global__ void k_pack(const float* __restrict__ p, short* __restrict__ pk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= PK_N) return;
  int base, K, NT;
  const float* W;
  if (i < PK_W2) { base = PK_W1; K = IN; NT = NTH; W = p + OFF_W1; }
  else if (i < PK_W3) { base = PK_W2; K = HID; NT = NTH; W = p + OFF_W2; }
  else if (i < PK_W4) { base = PK_W3; K = HID; NT = NTH; W = p + OFF_W3; }
  else { base = PK_W4; K = HID; NT = 1; W = p + OFF_W4; }
  const int o = i - base, chunk = o / 512, r = o % 512, lane = r / 8, e = r % 8;
  const int ks = chunk / NT, t = chunk % NT;
  const int row = 32 * ks + 8 * (lane / 16) + e, col = 16 * t + lane % 16;
  pk[i] = row < K ? f2bf(W[row * (NT * 16) + col]) : (short)0;
}


// WAVES waves per block share the weights; each wave takes RT 16-row tiles at a time, so
// every B operand read from LDS feeds RT MFMAs.  A wave's activations live in its own LDS
// tile, so after the weights are loaded the waves run independently (no block barrier in the
// loop; a wavefront fence orders a wave's activation stores before its loads).  MASK: rows
// whose column 15 is 0 need no prediction; a wave whose rows all are such writes NaN
// predictions and skips the work.  BIN: the rows are already normalized and rounded to bf16
// (16 bf16 per row, column 15 the need flag), as the step's projection kernel writes them;
// the kernel then feeds them to the first MFMA as loaded.  Both forms give the same bits.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// Waves that share one segment of the step's rows (a divisor of every launch shape's waves).
constexpr int kSegWaves = 4;

template <int WAVES, int RT, bool MASK, bool BIN, bool SEG = false>
__global__ __launch_bounds__(64 * WAVES) void k_mlp_fused(const void* __restrict__ Xv, long n,
                                                          const float* __restrict__ p,
                                                          const short* __restrict__ pk,
                                                          float* __restrict__ P,
                                                          const unsigned* __restrict__ seg,
                                                          const unsigned* __restrict__ idx,
                                                          int seg_rows) {
  // SEG: n rows in segments of seg_rows rows (a multiple of 16 RT), of which the first
  // seg[s] of segment s are rows to predict; row i's prediction to P[idx[i]].  Otherwise
  // row i's prediction to P[i].  cnt: the row's segment's count (SEG).
  auto valid = [&](long row, long cnt) { return SEG ? (row % seg_rows) < cnt : row < n; };
  const float* X = reinterpret_cast<const float*>(Xv);
  const s8* Xb = reinterpret_cast<const s8*>(Xv);  // BIN: row b is Xb[2 b], Xb[2 b + 1]
  __shared__ s8 wts[PK_N / 8];
  __shared__ float bias[3 * HID + 2];
  __shared__ __align__(16) short act[WAVES][RT * 16 * ACT_LD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const s8* pk8 = reinterpret_cast<const s8*>(pk);
  for (int i = tid; i < PK_N / 8; i += 64 * WAVES) wts[i] = pk8[i];
  for (int i = tid; i < HID; i += 64 * WAVES) {
    bias[i] = p[OFF_B1 + i];
    bias[HID + i] = p[OFF_B2 + i];
    bias[2 * HID + i] = p[OFF_B3 + i];
  }
  if (tid < 2) bias[3 * HID + tid] = p[OFF_B4 + tid];
  __syncthreads();
  const int r = lane & 15, q = lane >> 4;
  // this lane's bias values: column 16 t + r of each hidden layer
  float bl[3][NTH];
#pragma unroll
  for (int L = 0; L < 3; ++L)
#pragma unroll
    for (int t = 0; t < NTH; ++t) bl[L][t] = bias[HID * L + 16 * t + r];
  const long ngroups = (n + 16 * RT - 1) / (16 * RT);  // RT-tile groups
  const long gstride = (long)gridDim.x * WAVES;
  // normalization of this lane's input columns k = 8q + e (lanes q < 2; column 15 is the
  // need flag and enters as 0)
  float xmu[8], xsd[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * q + e;
    xmu[e] = k < 15 ? p[OFF_XMU + k] : 0.f;
    xsd[e] = k < 15 ? p[OFF_XSD + k] : 1.f;
  }
  // raw feature rows of the next group, loaded while the current one is computed: lane
  // (r, q < 2) holds columns 8q..8q+7 of row 16u + r of the group, column 15 included;
  // SEG: with the destinations of the rows 16u + 4q + e whose outputs the lane writes
  float xn[RT][8];
  s8 bn[RT];
  unsigned dn[RT][4];
  auto load_group = [&](long g, long cnt) {
#pragma unroll
    for (int u = 0; u < RT; ++u) {
      const long row = g * 16 * RT + 16 * u + r;
      if (BIN) {
        bn[u] = (q < 2 && row < n && valid(row, cnt)) ? Xb[row * 2 + q] : s8{0, 0, 0, 0, 0, 0, 0, 0};
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) xn[u][e] = (q < 2 && row < n) ? X[row * 16 + 8 * q + e] : 0.f;
      }
      if (SEG) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const long ro = g * 16 * RT + 16 * u + 4 * q + e;
          dn[u][e] = (r < 2 && ro < n && valid(ro, cnt)) ? idx[ro] : 0u;
        }
      }
    }
  };
  short* A = act[w];
  // Group order.  Plain: group W, W + gstride, ... for wave W.  SEG: kSegWaves waves share
  // a segment: wave W takes groups j, j + kSegWaves, ... (j = W mod kSegWaves) of the groups
  // that hold rows of segments W / kSegWaves, + sstride, ... (a segment's seg[s] rows fill
  // its first ceil(seg[s] / (16 RT)) groups), so that a wave never visits a group past a
  // count, and a full segment's groups are computed by several waves side by side.
  constexpr long J = SEG ? kSegWaves : 1;
  const long wid = (long)blockIdx.x * WAVES + w;
  const long sub = wid % J, sstride = gstride / J;
  const long gps = SEG ? seg_rows / (16 * RT) : 1;  // groups per segment
  const long nseg = SEG ? n / seg_rows : 0;
  // first group of this wave's in the first of segments t, t + sstride, ... that has one;
  // gc, rc: that segment's groups and rows
  auto first_in = [&](long t, long& gc, long& rc) {
    gc = 0;
    rc = 0;
    while (t < nseg) {
      rc = (long)seg[t];
      gc = (rc + 16 * RT - 1) / (16 * RT);
      if (sub < gc) break;
      t += sstride;
    }
    return t < nseg ? t * gps + sub : ngroups;
  };
  long gc = 0, rc = 0;  // SEG: groups and rows of the current group's segment
  long grp = SEG ? first_in(wid / J, gc, rc) : wid;
  long nxt_g = ngroups, nxt_c = 0, nxt_r = 0;
  if (grp < ngroups) load_group(grp, rc);
  for (; grp < ngroups; grp = nxt_g, gc = nxt_c, rc = nxt_r) {
    const long row0 = grp * 16 * RT;
    float xv[RT][8];
    s8 bv8[RT];
    unsigned dv[RT][4];
#pragma unroll
    for (int u = 0; u < RT; ++u) {
      if (BIN) {
        bv8[u] = bn[u];
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[u][e] = xn[u][e];
      }
      if (SEG) {
#pragma unroll
        for (int e = 0; e < 4; ++e) dv[u][e] = dn[u][e];
      }
    }
    if (SEG) {
      nxt_c = gc;
      nxt_r = rc;
      nxt_g = (grp % gps) + J < gc ? grp + J : first_in(grp / gps + sstride, nxt_c, nxt_r);
    } else {
      nxt_g = grp + gstride;
    }
    if (nxt_g < ngroups) load_group(nxt_g, nxt_r);
    if (MASK) {
      // column 15 of row 16u + r sits in lane (r, q = 1), element 7
      bool need = false;
#pragma unroll
      for (int u = 0; u < RT; ++u) {
        const bool flag = BIN ? bv8[u][7] != 0 : xv[u][7] != 0.f;
        need = need || (q == 1 && row0 + 16 * u + r < n && flag);
      }
      if (__ballot(need) == 0ull) {
#pragma unroll
        for (int u = 0; u < RT; ++u) {
          const long row = row0 + 16 * u + r;
          if (q == 0 && row < n) {
            P[row * 2] = NAN;
            P[row * 2 + 1] = NAN;
          }
        }
        continue;
      }
    }
    f4 acc[RT][NTH];
    // layer 1: normalized input (k = 8q..8q+7 of the 32-wide step; features 0..14) times W1
#pragma unroll
    for (int u = 0; u < RT; ++u) {
      const long row = row0 + 16 * u + r;
      s8 a;
      if (BIN) {
        a = bv8[u];
        if (q == 1) a[7] = 0;  // the need flag does not enter the network
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = 8 * q + e;
          float v = 0.f;
          if (row < n && k < 15) v = (xv[u][e] - xmu[e]) / xsd[e];
          a[e] = f2bf(v);
        }
      }
#pragma unroll
      for (int t = 0; t < NTH; ++t)
        acc[u][t] = mfma32(a, wts[(PK_W1 / 8) + t * 64 + lane], f4{0.f, 0.f, 0.f, 0.f});
    }
#pragma unroll
    for (int u = 0; u < RT; ++u)
#pragma unroll
      for (int t = 0; t < NTH; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          A[(16 * u + 4 * q + e) * ACT_LD + 16 * t + r] = f2bf(gelu(acc[u][t][e] + bl[0][t]));
    wave_lds_fence();
    // layers 2 and 3
#pragma unroll
    for (int L = 0; L < 2; ++L) {
      const int wb = (L == 0 ? PK_W2 : PK_W3) / 8;
#pragma unroll
      for (int u = 0; u < RT; ++u)
#pragma unroll
        for (int t = 0; t < NTH; ++t) acc[u][t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KSH; ++ks) {
        s8 av[RT];
#pragma unroll
        for (int u = 0; u < RT; ++u)
          av[u] = *reinterpret_cast<const s8*>(&A[(16 * u + r) * ACT_LD + 32 * ks + 8 * q]);
#pragma unroll
        for (int t = 0; t < NTH; ++t) {
          const s8 bv = wts[wb + (ks * NTH + t) * 64 + lane];
#pragma unroll
          for (int u = 0; u < RT; ++u) acc[u][t] = mfma32(av[u], bv, acc[u][t]);
        }
      }
      wave_lds_fence();  // every lane has read the tile before it is overwritten
#pragma unroll
      for (int u = 0; u < RT; ++u)
#pragma unroll
        for (int t = 0; t < NTH; ++t)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            A[(16 * u + 4 * q + e) * ACT_LD + 16 * t + r] = f2bf(gelu(acc[u][t][e] + bl[L + 1][t]));
      wave_lds_fence();
    }
    // layer 4: one 16-column tile, outputs 0 and 1 kept
#pragma unroll
    for (int u = 0; u < RT; ++u) {
      f4 o = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KSH; ++ks) {
        const s8 av = *reinterpret_cast<const s8*>(&A[(16 * u + r) * ACT_LD + 32 * ks + 8 * q]);
        o = mfma32(av, wts[(PK_W4 / 8) + ks * 64 + lane], o);
      }
      if (r < 2) {
        const float mu = p[OFF_YMU + r], sd = p[OFF_YSD + r], b = bias[3 * HID + r];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const long row = row0 + 16 * u + 4 * q + e;
          if (row < n && valid(row, rc)) P[(SEG ? (long)dv[u][e] : row) * 2 + r] = (o[e] + b) * sd + mu;
        }
      }
    }
    wave_lds_fence();  // the next group overwrites A
  }
}

// Launch shapes of the fused inference (waves per block, row tiles per wave): 16 x 1, 8 x 2
// (each B operand feeding two MFMAs) and 8 x 1 (smaller blocks, so that more of them share a
// CU).  They compute the same values in the same order; the first large inference times them
// and keeps the fastest.
constexpr int kFusedShapes = 3;
template <bool MASK, bool BIN, bool SEG = false>
void launch_fused_t(int shape, const void* X, long n, const float* p, const short* pk, float* P,
                    hipStream_t s, const unsigned* seg = nullptr, const unsigned* idx = nullptr,
                    int seg_rows = 0) {
  const int waves = shape == 0 ? 16 : 8;
  const int rt = shape == 1 ? 2 : 1;
  const long groups = (n + 16 * rt - 1) / (16 * rt);
  long blocks = (groups + waves - 1) / waves;
  if (blocks > 2048) blocks = 2048;
  // segments: kSegWaves waves for each (the grid's waves are a multiple of kSegWaves)
  if (SEG) blocks = std::min<long>((n / seg_rows * kSegWaves + waves - 1) / waves, 1l << 16);
  const dim3 g((unsigned)blocks);
  if (shape == 0)
    hipLaunchKernelGGL((k_mlp_fused<16, 1, MASK, BIN, SEG>), g, dim3(64 * 16), 0, s, X, n, p, pk, P,
                       seg, idx, seg_rows);
  else if (shape == 1)
    hipLaunchKernelGGL((k_mlp_fused<8, 2, MASK, BIN, SEG>), g, dim3(64 * 8), 0, s, X, n, p, pk, P,
                       seg, idx, seg_rows);
  else
    hipLaunchKernelGGL((k_mlp_fused<8, 1, MASK, BIN, SEG>), g, dim3(64 * 8), 0, s, X, n, p, pk, P,
                       seg, idx, seg_rows);
}
// X: float rows (bin = false) or normalized bf16 rows (bin = true), 16 values each.
void launch_fused(int shape, bool mask, bool bin, const void* X, long n, const float* p,
                  const short* pk, float* P, hipStream_t s) {
  if (mask && bin) launch_fused_t<true, true>(shape, X, n, p, pk, P, s);
  else if (mask) launch_fused_t<true, false>(shape, X, n, p, pk, P, s);
  else if (bin) launch_fused_t<false, true>(shape, X, n, p, pk, P, s);
  else launch_fused_t<false, false>(shape, X, n, p, pk, P, s);
}

unsigned g1(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <int NT>
void gemm(hipStream_t s, const float* A, long sam, long sak, const float* B, long sbk, long sbn,
          long M, int N, long K, const float* bias, float* Z, float* H, long ldc, int Nst, int act) {
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 16 * NT - 1) / (16 * NT)));
  hipLaunchKernelGGL((k_gemm<NT>), grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, (int)M, N,
                     (int)K, bias, Z, H, ldc, Nst, act);
}

struct Net {
  float* blob;   // parameters + normalization (device)
  float* grad;   // NPARAM
  float* m1;     // Adam moments
  float* m2;
  long cap;      // training activation capacity (samples)
  float *Xn, *Z1, *H1, *Z2, *H2, *Z3, *H3, *Y, *dY, *dA, *dB;
  long icap;     // inference capacity: normalized input, two hidden ping-pong buffers, output
  float *iX, *iA, *iB, *iY;
  float* loss;
  long step;
  hipStream_t stream;
  short* packed;  // bf16 weights in MFMA operand order (fused inference)
  int shape;      // fused launch shape chosen by timing (-1: not yet)
  hipEvent_t ev[2];
  long version, packed_version;  // blob changes / version the packed copy was made from
};

static int ensure_icap(Net* m, long n) {
  if (n <= m->icap) return HYQ_OK;
  float** bufs[] = {&m->iX, &m->iA, &m->iB, &m->iY};
  const long width[] = {IN, HID, HID, OUT};
  for (int i = 0; i < 4; ++i) {
    (void)hipFree(*bufs[i]);
    *bufs[i] = nullptr;
    if (hipMalloc(bufs[i], sizeof(float) * width[i] * n) != hipSuccess)
      return mfail(HYQ_ENOMEM, "mlp inference allocation");
  }
  m->icap = n;
  return HYQ_OK;
}

static int ensure_cap(Net* m, long n) {
  if (n <= m->cap) return HYQ_OK;
  float** bufs[] = {&m->Xn, &m->Z1, &m->H1, &m->Z2, &m->H2, &m->Z3, &m->H3, &m->Y, &m->dY, &m->dA, &m->dB};
  const long width[] = {IN, HID, HID, HID, HID, HID, HID, OUT, OUT, HID, HID};
  for (int i = 0; i < 11; ++i) {
    (void)hipFree(*bufs[i]);
    *bufs[i] = nullptr;
    if (hipMalloc(bufs[i], sizeof(float) * width[i] * n) != hipSuccess)
      return mfail(HYQ_ENOMEM, "mlp activation allocation");
  }
  m->cap = n;
  return HYQ_OK;
}

// forward on normalized inputs m->Xn (n rows); fills Z*, H*, Y
static void forward(Net* m, long n) {
  const float* p = m->blob;
  gemm<NTH>(m->stream, m->Xn, IN, 1, p + OFF_W1, HID, 1, n, HID, IN, p + OFF_B1, m->Z1, m->H1, HID, HID, 1);
  gemm<NTH>(m->stream, m->H1, HID, 1, p + OFF_W2, HID, 1, n, HID, HID, p + OFF_B2, m->Z2, m->H2, HID, HID, 1);
  gemm<NTH>(m->stream, m->H2, HID, 1, p + OFF_W3, HID, 1, n, HID, HID, p + OFF_B3, m->Z3, m->H3, HID, HID, 1);
  gemm<1>(m->stream, m->H3, HID, 1, p + OFF_W4, OUT, 1, n, OUT, HID, p + OFF_B4, nullptr, m->Y, OUT, OUT, 0);
}

// gradients of the loss into m->grad (forward must have run)
static void backward(Net* m, long n) {
  const float* p = m->blob;
  float* g = m->grad;
  hipStream_t s = m->stream;
  // layer 4
  gemm<1>(s, m->H3, 1, HID, m->dY, OUT, 1, HID, OUT, n, nullptr, nullptr, g + OFF_W4, OUT, OUT, 0);
  hipLaunchKernelGGL(k_colsum, dim3(OUT), dim3(256), 0, s, m->dY, n, OUT, g + OFF_B4);
  gemm<NTH>(s, m->dY, OUT, 1, p + OFF_W4, 1, OUT, n, HID, OUT, nullptr, nullptr, m->dA, HID, HID, 0);
  hipLaunchKernelGGL(k_gelu_bwd, dim3(g1(n * HID)), dim3(256), 0, s, m->Z3, m->dA, n * HID);
  // layer 3
  gemm<NTH>(s, m->H2, 1, HID, m->dA, HID, 1, HID, HID, n, nullptr, nullptr, g + OFF_W3, HID, HID, 0);
  hipLaunchKernelGGL(k_colsum, dim3(HID), dim3(256), 0, s, m->dA, n, HID, g + OFF_B3);
  gemm<NTH>(s, m->dA, HID, 1, p + OFF_W3, 1, HID, n, HID, HID, nullptr, nullptr, m->dB, HID, HID, 0);
  hipLaunchKernelGGL(k_gelu_bwd, dim3(g1(n * HID)), dim3(256), 0, s, m->Z2, m->dB, n * HID);
  // layer 2
  gemm<NTH>(s, m->H1, 1, HID, m->dB, HID, 1, HID, HID, n, nullptr, nullptr, g + OFF_W2, HID, HID, 0);
  hipLaunchKernelGGL(k_colsum, dim3(HID), dim3(256), 0, s, m->dB, n, HID, g + OFF_B2);
  gemm<NTH>(s, m->dB, HID, 1, p + OFF_W2, 1, HID, n, HID, HID, nullptr, nullptr, m->dA, HID, HID, 0);
  hipLaunchKernelGGL(k_gelu_bwd, dim3(g1(n * HID)), dim3(256), 0, s, m->Z1, m->dA, n * HID);
  // layer 1
  gemm<NTH>(s, m->Xn, 1, IN, m->dA, HID, 1, IN, HID, n, nullptr, nullptr, g + OFF_W1, HID, HID, 0);
  hipLaunchKernelGGL(k_colsum, dim3(HID), dim3(256), 0, s, m->dA, n, HID, g + OFF_B1);
}

}  // namespace

// Entry points of this width's network (the C ABI in hyq_mlp_api.hip calls them); a network
// is passed as the void* that create returned.
namespace HYQ_MLP_NS {

const char* last_error() { return g_mlp_err.c_str(); }
int64_t blob_size() { return (int64_t)NBLOB; }
int64_t param_count() { return (int64_t)NPARAM; }

int create(const float* blob_host, void** out) {
  if (!out) return mfail(HYQ_EINVAL, "null argument");
  Net* m = new Net();
  std::memset(m, 0, sizeof(*m));
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return mfail(HYQ_EHIP, "mlp stream");
  }
  bool ok = hipMalloc(&m->blob, sizeof(float) * NBLOB) == hipSuccess &&
            hipMalloc(&m->grad, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc(&m->m1, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc(&m->m2, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc(&m->loss, sizeof(float)) == hipSuccess;
  if (!ok) {
    delete m;
    return mfail(HYQ_ENOMEM, "mlp allocation");
  }
  if (blob_host) {
    MCHK(hipMemcpy(m->blob, blob_host, sizeof(float) * NBLOB, hipMemcpyHostToDevice), "mlp blob");
  } else {
    MCHK(hipMemset(m->blob, 0, sizeof(float) * NBLOB), "mlp blob");
  }
  MCHK(hipMemset(m->m1, 0, sizeof(float) * NPARAM), "adam m");
  MCHK(hipMemset(m->m2, 0, sizeof(float) * NPARAM), "adam v");
  MCHK(hipMalloc(&m->packed, sizeof(short) * PK_N), "mlp packed weights");
  m->shape = -1;
  MCHK(hipEventCreate(&m->ev[0]), "mlp event");
  MCHK(hipEventCreate(&m->ev[1]), "mlp event");
  m->version = 1;
  m->packed_version = 0;
  *out = m;
  return HYQ_OK;
}

int destroy(void* h) {
  Net* m = static_cast<Net*>(h);
  if (!m) return HYQ_OK;
  (void)hipStreamSynchronize(m->stream);
  float* bufs[] = {m->blob, m->grad, m->m1, m->m2, m->loss, m->Xn, m->Z1, m->H1, m->Z2,
                   m->H2, m->Z3, m->H3, m->Y, m->dY, m->dA, m->dB, m->iX, m->iA, m->iB, m->iY};
  for (float* b : bufs) (void)hipFree(b);
  (void)hipFree(m->packed);
  for (hipEvent_t e : m->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(m->stream);
  delete m;
  return HYQ_OK;
}

int get_blob(void* h, float* blob_host) {
  Net* m = static_cast<Net*>(h);
  if (!m || !blob_host) return mfail(HYQ_EINVAL, "null argument");
  MCHK(hipStreamSynchronize(m->stream), "mlp sync");
  MCHK(hipMemcpy(blob_host, m->blob, sizeof(float) * NBLOB, hipMemcpyDeviceToHost), "mlp blob");
  return HYQ_OK;
}
int set_blob(void* h, const float* blob_host) {
  Net* m = static_cast<Net*>(h);
  if (!m || !blob_host) return mfail(HYQ_EINVAL, "null argument");
  MCHK(hipStreamSynchronize(m->stream), "mlp sync");
  MCHK(hipMemcpy(m->blob, blob_host, sizeof(float) * NBLOB, hipMemcpyHostToDevice), "mlp blob");
  m->step = 0;
  m->version += 1;
  MCHK(hipMemset(m->m1, 0, sizeof(float) * NPARAM), "adam m");
  MCHK(hipMemset(m->m2, 0, sizeof(float) * NPARAM), "adam v");
  return HYQ_OK;
}

int infer_fused(Net* m, const void* X, int64_t n, float* P, hipStream_t stream,
                       bool mask, bool bin = false) {
  if (!m || !X || !P || n < 0) return mfail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  if (m->packed_version != m->version) {
    // the parameters may have been written on the network's own stream
    MCHK(hipStreamSynchronize(m->stream), "mlp sync");
    hipLaunchKernelGGL(k_pack, dim3(g1(PK_N)), dim3(256), 0, stream, m->blob, m->packed);
    m->packed_version = m->version;
  }
  if (m->shape < 0 && n >= 65536) {
    // pick the fastest launch shape: every row computed (no mask), best of three runs each,
    // on this call's own rows (the real launch below overwrites the trial predictions)
    float best[kFusedShapes];
    for (int sh = 0; sh < kFusedShapes; ++sh) best[sh] = 1e30f;
    for (int rep = 0; rep < 3; ++rep)
      for (int sh = 0; sh < kFusedShapes; ++sh) {
        float ms = 0.f;
        MCHK(hipEventRecord(m->ev[0], stream), "mlp event");
        launch_fused(sh, false, bin, X, (long)n, m->blob, m->packed, P, stream);
        MCHK(hipEventRecord(m->ev[1], stream), "mlp event");
        MCHK(hipEventSynchronize(m->ev[1]), "mlp event");
        MCHK(hipEventElapsedTime(&ms, m->ev[0], m->ev[1]), "mlp event");
        best[sh] = ms < best[sh] ? ms : best[sh];
      }
    int pick = 0;
    for (int sh = 1; sh < kFusedShapes; ++sh)
      if (best[sh] < best[pick]) pick = sh;
    m->shape = pick;
  }
  launch_fused(m->shape < 0 ? 0 : m->shape, mask, bin, X, (long)n, m->blob, m->packed, P, stream);
  MCHK(hipGetLastError(), "mlp infer launch");
  return HYQ_OK;
}

// Predicted (lmin, lmax) for n feature rows X [n][16] (device) into P [n][2] (device),
// enqueued on `stream` without synchronizing: the fused kernel (the packed weights are
// refreshed first when the parameters changed).
int infer_async(void* h, const float* X, int64_t n, float* P, hipStream_t stream) {
  Net* m = static_cast<Net*>(h);
  return infer_fused(m, X, n, P, stream, false);
}
// The step's inference: rows whose column 15 is 0 (near-Gaussian planes) need no
// prediction; 16-row tiles made only of such rows are skipped and get NaN.
int infer_masked_async(void* h, const float* X, int64_t n, float* P, hipStream_t stream) {
  Net* m = static_cast<Net*>(h);
  return infer_fused(m, X, n, P, stream, true);
}

// The same from normalized bf16 rows Xb [n][16] (column 15 the need flag, 0 or 1), as the
// step's projection kernel writes them with the network's input normalization
// (hyq_mlp_input_norm); the predictions are the same bits as from the float rows.
int infer_bf16_masked_async(void* h, const short* Xb, int64_t n, float* P,
                            hipStream_t stream) {
  Net* m = static_cast<Net*>(h);
  return infer_fused(m, Xb, n, P, stream, true, true);
}
// The step's inference on segmented rows: nseg segments of seg_rows rows of Xb (normalized
// bf16 rows, as the step's projection kernel writes them), of which the first seg[s] of
// segment s (read on the device) are rows to predict; row i's prediction to P[idx[i]].
// seg_rows must be a multiple of 32 (whole row groups of every launch shape).  Every
// listed row is computed, with the launch shape the timing trial chose (8 waves x 1 tile
// until one was chosen): the same predictions, bit for bit, as the unsegmented forms give
// for the same rows.
int infer_bf16_segments_async(void* h, const short* Xb, const uint32_t* idx, const uint32_t* seg,
                              int64_t nseg, int32_t seg_rows, float* P, hipStream_t stream) {
  Net* m = static_cast<Net*>(h);
  if (!m || !Xb || !idx || !seg || !P || nseg < 0 || seg_rows <= 0 || seg_rows % 32 != 0)
    return mfail(HYQ_EINVAL, "bad argument");
  if (nseg == 0) return HYQ_OK;
  if (m->packed_version != m->version) {
    MCHK(hipStreamSynchronize(m->stream), "mlp sync");
    hipLaunchKernelGGL(k_pack, dim3(g1(PK_N)), dim3(256), 0, stream, m->blob, m->packed);
    m->packed_version = m->version;
  }
  launch_fused_t<false, true, true>(m->shape < 0 ? 2 : m->shape, Xb, (long)nseg * seg_rows, m->blob,
                                    m->packed, P, stream, seg, idx, seg_rows);
  MCHK(hipGetLastError(), "mlp infer launch");
  return HYQ_OK;
}
// Device pointer to the input normalization: 16 means, then 16 standard deviations.
const float* input_norm(void* h) {
  Net* m = static_cast<Net*>(h);
  return m ? m->blob + OFF_XMU : nullptr;
}

int set_shape(void* h, int32_t shape) {
  Net* m = static_cast<Net*>(h);
  if (!m || shape < -1 || shape >= kFusedShapes) return mfail(HYQ_EINVAL, "bad argument");
  m->shape = shape;
  return HYQ_OK;
}

// The same prediction through the layer-by-layer GEMM kernels of training (a cross-check
// of the fused kernel).
int infer_layers(void* h, const float* X, int64_t n, float* P) {
  Net* m = static_cast<Net*>(h);
  if (!m || !X || !P || n < 0) return mfail(HYQ_EINVAL, "bad argument");
  if (n == 0) return HYQ_OK;
  int rc = ensure_icap(m, n);
  if (rc) return rc;
  const float* p = m->blob;
  hipStream_t stream = m->stream;
  hipLaunchKernelGGL(k_normalize, dim3(g1(n * 16)), dim3(256), 0, stream, X, p + OFF_XMU, m->iX, (long)n);
  gemm<NTH>(stream, m->iX, IN, 1, p + OFF_W1, HID, 1, n, HID, IN, p + OFF_B1, nullptr, m->iA, HID, HID, 1);
  gemm<NTH>(stream, m->iA, HID, 1, p + OFF_W2, HID, 1, n, HID, HID, p + OFF_B2, nullptr, m->iB, HID, HID, 1);
  gemm<NTH>(stream, m->iB, HID, 1, p + OFF_W3, HID, 1, n, HID, HID, p + OFF_B3, nullptr, m->iA, HID, HID, 1);
  gemm<1>(stream, m->iA, HID, 1, p + OFF_W4, OUT, 1, n, OUT, HID, p + OFF_B4, nullptr, m->iY, OUT, OUT, 0);
  hipLaunchKernelGGL(k_denorm, dim3(g1(n)), dim3(256), 0, stream, m->iY, p + OFF_YMU, P, (long)n);
  MCHK(hipGetLastError(), "mlp infer launch");
  MCHK(hipStreamSynchronize(stream), "mlp infer sync");
  return HYQ_OK;
}
int infer(void* h, const float* X, int64_t n, float* P) {
  Net* m = static_cast<Net*>(h);
  if (!m) return mfail(HYQ_EINVAL, "null argument");
  int rc = infer_async(m, X, n, P, m->stream);
  if (rc) return rc;
  MCHK(hipStreamSynchronize(m->stream), "mlp infer sync");
  return HYQ_OK;
}

// Loss and parameter gradients on one batch (X [n][16], T [n][2], device); grad_host may
// be null.  No update.
int grad(void* h, const float* X, const float* T, int64_t n, float* loss_out,
         float* grad_host) {
  Net* m = static_cast<Net*>(h);
  if (!m || !X || !T || n <= 0) return mfail(HYQ_EINVAL, "bad argument");
  int rc = ensure_cap(m, n);
  if (rc) return rc;
  MCHK(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream), "loss reset");
  hipLaunchKernelGGL(k_normalize, dim3(g1(n * 16)), dim3(256), 0, m->stream, X, m->blob + OFF_XMU,
                     m->Xn, (long)n);
  forward(m, n);
  hipLaunchKernelGGL(k_mse_grad, dim3(g1(n)), dim3(256), 0, m->stream, m->Y, T, m->blob + OFF_YMU,
                     m->dY, (long)n, m->loss);
  backward(m, n);
  MCHK(hipGetLastError(), "mlp grad launch");
  MCHK(hipStreamSynchronize(m->stream), "mlp grad sync");
  if (loss_out) MCHK(hipMemcpy(loss_out, m->loss, sizeof(float), hipMemcpyDeviceToHost), "loss");
  if (grad_host)
    MCHK(hipMemcpy(grad_host, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost), "grad");
  return HYQ_OK;
}

// One Adam step on a batch.
int train_step(void* h, const float* X, const float* T, int64_t n, float lr,
               float* loss_out) {
  Net* m = static_cast<Net*>(h);
  int rc = grad(m, X, T, n, loss_out, nullptr);
  if (rc) return rc;
  m->step += 1;
  m->version += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
  hipLaunchKernelGGL(k_adam, dim3(g1(NPARAM)), dim3(256), 0, m->stream, m->blob, m->grad, m->m1,
                     m->m2, (long)NPARAM, lr, b1, b2, 1e-8f, c1, c2);
  hipLaunchKernelGGL(k_mask_w4, dim3(g1(HID * OUT + OUT)), dim3(256), 0, m->stream, m->blob);
  MCHK(hipGetLastError(), "adam launch");
  return HYQ_OK;
}

}  // namespace HYQ_MLP_NS
