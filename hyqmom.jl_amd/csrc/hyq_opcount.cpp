// Operation-counting host build of the device cell numerics (libhyqmom_opcount.so, built
// with the host compiler from the same headers as the kernels): every FP64 add, subtract,
// multiply, divide, sqrt, cbrt, exp and trigonometric call on the path a cell takes is
// counted once.  It measures the algorithmic FLOPs per cell of pass 1 and pass 2 on real
// cell states, the numerator of the kernels' FLOP-rate roofline figures.
#include <cmath>
#include <cstdint>
#include <cstring>

static thread_local unsigned long long g_flops = 0;

struct CD {
  double v;
  constexpr CD() : v(0.0) {}
  constexpr CD(double x) : v(x) {}
  explicit operator float() const { return (float)v; }
};
inline CD operator+(CD a, CD b) { ++g_flops; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++g_flops; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { ++g_flops; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++g_flops; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline CD& operator+=(CD& a, CD b) { a = a + b; return a; }
inline CD& operator-=(CD& a, CD b) { a = a - b; return a; }
inline CD& operator*=(CD& a, CD b) { a = a * b; return a; }
inline CD& operator/=(CD& a, CD b) { a = a / b; return a; }
inline bool operator<(CD a, CD b) { return a.v < b.v; }
inline bool operator>(CD a, CD b) { return a.v > b.v; }
inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
inline bool operator==(CD a, CD b) { return a.v == b.v; }
inline bool operator!=(CD a, CD b) { return a.v != b.v; }
inline CD sqrt(CD a) { ++g_flops; return CD(std::sqrt(a.v)); }
inline CD cbrt(CD a) { ++g_flops; return CD(std::cbrt(a.v)); }
inline CD exp(CD a) { ++g_flops; return CD(std::exp(a.v)); }
inline CD acos(CD a) { ++g_flops; return CD(std::acos(a.v)); }
inline CD cos(CD a) { ++g_flops; return CD(std::cos(a.v)); }
inline CD hypot(CD a, CD b) { g_flops += 3; return CD(std::hypot(a.v, b.v)); }
inline CD fabs(CD a) { return CD(std::fabs(a.v)); }
inline CD fmin(CD a, CD b) { return CD(std::fmin(a.v, b.v)); }
inline CD fmax(CD a, CD b) { return CD(std::fmax(a.v, b.v)); }

#define double CD
#include "hyq_pipeline.h"
#undef double

using namespace hyq;

template <int EM>
static void count_range(const double* M, long long n, int flag2D, double Ma, double dt, double Kn,
                        double* p1, double* p2) {
  double s1 = 0.0, s2 = 0.0;
  for (long long c = 0; c < n; ++c) {
    CD m[35], mr[35], fx[35], fy[35], fz[35], vp[6], mo[35];
    for (int t = 0; t < 35; ++t) m[t] = CD(M[(size_t)t * n + c]);
    uint32_t fl = 0;
    g_flops = 0;
    pass1_cell<CD, EM>(m, flag2D, Ma, true, mr, fx, fy, fz, 1, vp, fl);
    s1 += (double)g_flops;
    g_flops = 0;
    pass2_cell<false, EM>(mr, flag2D, Ma, true, dt, Kn, mo, fl);
    s2 += (double)g_flops;
  }
  *p1 = s1;
  *p2 = s2;
}

extern "C" {
// Sum over the n cells of M ([35][n], host) of the FLOPs of pass 1 (closure, wave speeds,
// fluxes) and of pass 2 (closure, eigen checks, BGK); em = 0 full eigen solve, 1 fast (Sturm + Laguerre)
// wave speeds from the analytic bounds.  Returns 0, or -1 on bad arguments.
int hyq_opcount(const double* M, long long n, int flag2D, double Ma, double dt, double Kn, int em,
                double* flops_pass1, double* flops_pass2) {
  if (!M || n < 0 || !flops_pass1 || !flops_pass2 || (em != 0 && em != 1)) return -1;
  if (em == 0)
    count_range<0>(M, n, flag2D, Ma, dt, Kn, flops_pass1, flops_pass2);
  else
    count_range<1>(M, n, flag2D, Ma, dt, Kn, flops_pass1, flops_pass2);
  return 0;
}
}
