// Operation-counting host build of the device cell numerics: every FP64 add, subtract,
// multiply, divide, sqrt, cbrt, exp and trigonometric call is counted.  Used to measure
// the algorithmic FLOPs per cell-step on real cell states (tests and the roofline figure).
#include <cmath>
#include <cstdint>
#include <cstring>

static thread_local unsigned long long g_flops = 0;

struct CD {
  double v;
  constexpr CD() : v(0.0) {}
  constexpr CD(double x) : v(x) {}
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
#include "../../hyqmom.jl_amd/csrc/hyq_pipeline.h"
#undef double

using namespace hyq;
extern "C" {
// FLOPs of pass 1 and of pass 2 (+BGK) for one cell, and of the HLL update (per cell,
// 35 moments x 3 directions x 2 faces).
void count_cell(const double* Min, int flag2D, double Ma, double dt, double Kn, double* out3) {
  CD M[35], Mr[35], Fx[35], Fy[35], Fz[35], vp[6], Mo[35];
  for (int n = 0; n < 35; ++n) M[n] = CD(Min[n]);
  uint32_t fl = 0;
  g_flops = 0;
  pass1_cell<CD>(M, flag2D, Ma, true, Mr, Fx, Fy, Fz, 1, vp, fl);
  out3[0] = (double)g_flops;
  g_flops = 0;
  pass2_cell(Mr, flag2D, Ma, true, dt, Kn, Mo, fl);
  out3[1] = (double)g_flops;
  out3[2] = (double)fl;
}
}
