// Device moment algebra for the HyQMOM 35-moment state (FP64, one cell per lane).
//
// Moments up to total order 5 are held in a packed "tetrahedral" array of 56 doubles:
// (i,j,k) with n = i+j+k lives at n(n+1)(n+2)/6 + r(r+1)/2 + k, r = j+k.  All loops run
// over compile-time bounds so that every index is a constant after unrolling and the
// arrays stay in VGPRs.
//
// Raw <-> central conversion is done as three 1D Taylor shifts (x, then y, then z) instead
// of the reference's expanded symbolic expressions (M4toC4_3D.jl, C4toM4_3D.jl,
// C5toM5_3D.jl): m_{i..} = sum_p C(i,p) u^(i-p) c_{p..} along one axis at a time.
#pragma once
#include <math.h>
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HD __host__ __device__ __forceinline__
#else  // host build of the cell numerics (CPU test harness only)
#define HD inline
#endif
// Keeps the instruction scheduler from moving instructions across this point (device code;
// nothing on the host): loads placed before it are all issued before any use after it.
#if defined(__HIP_DEVICE_COMPILE__)
#define HYQ_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#else
#define HYQ_SCHED_BARRIER() ((void)0)
#endif
// A store that the writing kernel never reads back, marked non-temporal on the device so
// that it does not displace the lines the kernel still reads from the caches (a plain
// store on the host).
#if defined(__HIP_DEVICE_COMPILE__)
#define HYQ_STORE_STREAM(ptr, v) __builtin_nontemporal_store((v), (ptr))
#else
#define HYQ_STORE_STREAM(ptr, v) (*(ptr) = (v))
#endif
// A load of a value that the reading kernel reads once and no other lane reads again,
// likewise marked non-temporal on the device.
#if defined(__HIP_DEVICE_COMPILE__)
#define HYQ_LOAD_STREAM(ptr) __builtin_nontemporal_load(ptr)
#else
#define HYQ_LOAD_STREAM(ptr) (*(ptr))
#endif

namespace hyq {

constexpr double EPS = 2.220446049250313e-16;

HD constexpr int tpos(int i, int j, int k) {
  return (i + j + k) * (i + j + k + 1) * (i + j + k + 2) / 6 + (j + k) * (j + k + 1) / 2 + k;
}
HD constexpr int binom(int n, int k) {
  int r = 1;
  for (int t = 1; t <= k; ++t) r = r * (n - k + t) / t;
  return r;
}

template <int N>
HD double pw(double x) {
  double r = 1.0;
#pragma unroll
  for (int t = 0; t < N; ++t) r *= x;
  return r;
}

// Julia NaN-propagating min/max (Base.min / Base.max).
HD double jmax(double a, double b) { return (a != a || b != b) ? (a + b) : (a > b ? a : b); }
HD double jmin(double a, double b) { return (a != a || b != b) ? (a + b) : (a < b ? a : b); }
HD double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

// Host/device constexpr copy for compile-time unrolled loops.
struct E35T {
  int e[35][3];
};
HD constexpr E35T e35() {
  return E35T{{{0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {3, 0, 0}, {4, 0, 0}, {0, 1, 0}, {1, 1, 0},
               {2, 1, 0}, {3, 1, 0}, {0, 2, 0}, {1, 2, 0}, {2, 2, 0}, {0, 3, 0}, {1, 3, 0},
               {0, 4, 0}, {0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}, {0, 0, 2}, {1, 0, 2},
               {2, 0, 2}, {0, 0, 3}, {1, 0, 3}, {0, 0, 4}, {0, 1, 1}, {1, 1, 1}, {2, 1, 1},
               {0, 2, 1}, {1, 2, 1}, {0, 3, 1}, {0, 1, 2}, {1, 1, 2}, {0, 1, 3}, {0, 2, 2}}};
}
#define T35(n) tpos(e35().e[n][0], e35().e[n][1], e35().e[n][2])

// In-place Taylor shift along axis AX of a packed moment array of order <= ORD:
// a_{..i..} <- sum_p C(i,p) s^(i-p) a_{..p..}.
template <int AX, int ORD>
HD void taylor_shift(double* a, double s) {
  double sp[ORD + 1];
  sp[0] = 1.0;
#pragma unroll
  for (int t = 1; t <= ORD; ++t) sp[t] = sp[t - 1] * s;
#pragma unroll
  for (int n = ORD; n >= 1; --n) {
#pragma unroll
    for (int j = 0; j <= ORD; ++j)
#pragma unroll
      for (int k = 0; k <= ORD; ++k) {
        // the other two exponents (o1, o2) fixed, shifted exponent from n down
        if (j + k + n > ORD) continue;
        double acc = 0.0;
#pragma unroll
        for (int p = 0; p <= n; ++p) {
          const int pi = AX == 0 ? tpos(p, j, k) : (AX == 1 ? tpos(j, p, k) : tpos(j, k, p));
          acc += binom(n, p) * sp[n - p] * a[pi];
        }
        const int ni = AX == 0 ? tpos(n, j, k) : (AX == 1 ? tpos(j, n, k) : tpos(j, k, n));
        a[ni] = acc;
      }
  }
}

// Central moments (order <= 4) of a 35-moment raw state.  c[tpos] is filled for order <= 4;
// c000 = 1 and the first-order central moments are exactly 0 (as M4toC4_3D sets them).
HD void raw_to_central4(const double* M, double* c, double& M000, double& u, double& v,
                        double& w) {
  M000 = M[0];
  const double r = 1.0 / M000;
  u = M[1] * r;
  v = M[5] * r;
  w = M[15] * r;
#pragma unroll
  for (int n = 0; n < 35; ++n) c[T35(n)] = M[n] * r;
  taylor_shift<0, 4>(c, -u);
  taylor_shift<1, 4>(c, -v);
  taylor_shift<2, 4>(c, -w);
  c[tpos(0, 0, 0)] = 1.0;
  c[tpos(1, 0, 0)] = 0.0;
  c[tpos(0, 1, 0)] = 0.0;
  c[tpos(0, 0, 1)] = 0.0;
}

// Raw moments (order <= ORD) from central moments: in place on the packed array.
template <int ORD>
HD void central_to_raw(double* c, double M000, double u, double v, double w) {
  c[tpos(0, 0, 0)] = 1.0;
  c[tpos(1, 0, 0)] = 0.0;
  c[tpos(0, 1, 0)] = 0.0;
  c[tpos(0, 0, 1)] = 0.0;
  taylor_shift<2, ORD>(c, w);
  taylor_shift<1, ORD>(c, v);
  taylor_shift<0, ORD>(c, u);
  constexpr int NN = (ORD + 1) * (ORD + 2) * (ORD + 3) / 6;
#pragma unroll
  for (int t = 0; t < NN; ++t) c[t] *= M000;
}

}  // namespace hyq
