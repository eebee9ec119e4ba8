// Fast wave speeds for the mixed_newton / mixed_learned modes: the characteristic polynomial of the
// standardized flux Jacobian, a Sturm-sequence hyperbolicity test, and Laguerre refinement
// of the extreme eigenvalues from outside the root interval (analytic bound or a verified
// warm start).  Cells where the test is not conclusive, or where the iteration does not
// converge in a bounded number of steps, fall back to the full FP64 eigen solve
// (hyq_eig.h).  Roots that coincide to rounding level (a Maxwellian plane has a double
// root at its mean) count as one multiple real root; a complex pair that close to the real
// axis is below what the full solve resolves either.
#pragma once
#include "hyq_eig.h"

namespace hyq {

// Characteristic polynomial det(l I - J) of the sparse Jacobian (rows 1, 2, 4 shift rows):
// monic degree 6, c[k] = coefficient of l^k.
HD void charpoly(const double* J, double c[7]) {
  const double a1 = J[12], a2 = J[13], a3 = J[14], a4 = J[15];
  const double b1 = J[24], b2 = J[25], b3 = J[26], b4 = J[27], b5 = J[28], b6 = J[29];
  const double c1 = J[30], c2 = J[31], c4 = J[33], c5 = J[34], c6 = J[35];
  c[6] = 1.0;
  c[5] = -a3 - b5 - c6;
  c[4] = -a2 + a3 * b5 + a3 * c6 - b4 + b5 * c6 - b6 * c5;
  c[3] = -a1 + a2 * b5 + a2 * c6 + a3 * b4 - a3 * b5 * c6 + a3 * b6 * c5 - a4 * b3 + b4 * c6 - b6 * c4;
  c[2] = a1 * b5 + a1 * c6 + a2 * b4 - a2 * b5 * c6 + a2 * b6 * c5 - a3 * b4 * c6 + a3 * b6 * c4 -
         a4 * b2 + a4 * b3 * c6;
  c[1] = a1 * b4 - a1 * b5 * c6 + a1 * b6 * c5 - a2 * b4 * c6 + a2 * b6 * c4 - a4 * b1 + a4 * b2 * c6 -
         a4 * b6 * c2;
  c[0] = -a1 * b4 * c6 + a1 * b6 * c4 + a4 * b1 * c6 - a4 * b6 * c1;
}

// Sturm sequence of the degree-6 characteristic polynomial, held in registers: member j has
// degree 6 - j (the generic case) and fixed-size storage, so every index is a compile-time
// constant.  Members are normalized to unit max-abs coefficient (signs kept).  A remainder at
// rounding level (max coefficient <= 1e-15 of its unit-scale dividend) ends the sequence at
// gcd(p, p'): numerically multiple roots.  A remainder that is small but resolved (<= 1e-9),
// or a degree drop, is ill-conditioned (ok = false) and left to the full solve.
struct Sturm {
  double p0[7], p1[6], p2[5], p3[4], p4[3], p5[2], p6[1];
  int n;    // members present, 2..7; the last one is gcd(p, p') when n < 7
  bool ok;
};

// -(a mod b) for deg a = DA, deg b = DA - 1 (two synthetic-division steps), normalized into
// out (degree DA - 2).  Returns 0: member made; 1: remainder at rounding level (gcd reached);
// 2: ill-conditioned.
template <int DA>
HD int sturm_rem(const double* a, const double* b, double* out) {
  double r[DA + 1];
#pragma unroll
  for (int t = 0; t <= DA; ++t) r[t] = a[t];
  const double lb = b[DA - 1];
  double q = r[DA] / lb;
#pragma unroll
  for (int t = 0; t < DA; ++t) r[t + 1] -= q * b[t];
  q = r[DA - 1] / lb;
#pragma unroll
  for (int t = 0; t < DA; ++t) r[t] -= q * b[t];
  double m = 0.0;
#pragma unroll
  for (int t = 0; t <= DA - 2; ++t) m = fmax(m, fabs(r[t]));
  if (!(m > 1e-15)) return 1;
  if (!(m > 1e-9) || !(fabs(r[DA - 2]) > 1e-12 * m)) return 2;
  const double inv = -1.0 / m;
#pragma unroll
  for (int t = 0; t <= DA - 2; ++t) out[t] = r[t] * inv;
  return 0;
}

template <int N>
HD void normalize_poly(double* c, bool& ok) {
  double m = 0.0;
#pragma unroll
  for (int t = 0; t < N; ++t) m = fmax(m, fabs(c[t]));
  ok = ok && (m > 0.0) && (fabs(c[N - 1]) > 0.0);
  const double inv = m > 0.0 ? 1.0 / m : 1.0;
#pragma unroll
  for (int t = 0; t < N; ++t) c[t] *= inv;
}

HD void sturm_build(const double c[7], Sturm& S) {
  S.ok = true;
#pragma unroll
  for (int t = 0; t < 7; ++t) S.p0[t] = c[t];
#pragma unroll
  for (int t = 0; t < 6; ++t) S.p1[t] = (t + 1) * c[t + 1];
  normalize_poly<7>(S.p0, S.ok);
  normalize_poly<6>(S.p1, S.ok);
  S.n = 2;
  if (!S.ok) return;
  int st = sturm_rem<6>(S.p0, S.p1, S.p2);
  if (st == 0) { S.n = 3; st = sturm_rem<5>(S.p1, S.p2, S.p3); }
  if (st == 0) { S.n = 4; st = sturm_rem<4>(S.p2, S.p3, S.p4); }
  if (st == 0) { S.n = 5; st = sturm_rem<3>(S.p3, S.p4, S.p5); }
  if (st == 0) { S.n = 6; st = sturm_rem<2>(S.p4, S.p5, S.p6); }
  if (st == 0) S.n = 7;
  S.ok = st != 2;
}

// degree of the last member (0 unless the sequence ended at a nontrivial gcd)
HD int sturm_gcd_degree(const Sturm& S) { return 7 - S.n; }

HD int sign_step(double v, double& prev) {  // one step of a sign-change count (zeros skipped)
  int ch = 0;
  if (v != 0.0) {
    if (prev != 0.0 && ((v > 0.0) != (prev > 0.0))) ch = 1;
    prev = v;
  }
  return ch;
}
template <int D>
HD double horner(const double* a, double x) {
  double acc = a[D];
#pragma unroll
  for (int t = D - 1; t >= 0; --t) acc = acc * x + a[t];
  return acc;
}
// Sign changes of the sequence at x (at_inf: +1 / -1 evaluates the leading terms at
// +inf / -inf, 0 evaluates at x).
HD int sturm_changes(const Sturm& S, double x, int at_inf) {
  double v[7];
  if (at_inf == 0) {
    v[0] = horner<6>(S.p0, x); v[1] = horner<5>(S.p1, x); v[2] = horner<4>(S.p2, x);
    v[3] = horner<3>(S.p3, x); v[4] = horner<2>(S.p4, x); v[5] = horner<1>(S.p5, x);
    v[6] = S.p6[0];
  } else {
    const double s = at_inf > 0 ? 1.0 : -1.0;  // sign of x^d at -inf alternates with d
    v[0] = S.p0[6]; v[1] = s * S.p1[5]; v[2] = S.p2[4]; v[3] = s * S.p3[3];
    v[4] = S.p4[2]; v[5] = s * S.p5[1]; v[6] = S.p6[0];
  }
  double prev = 0.0;
  int ch = 0;
#pragma unroll
  for (int k = 0; k < 7; ++k) ch += (k < S.n) ? sign_step(v[k], prev) : 0;
  return ch;
}
HD int sturm_real_roots(const Sturm& S) {  // distinct real roots on the whole line
  return sturm_changes(S, 0.0, -1) - sturm_changes(S, 0.0, 1);
}
HD int sturm_roots_above(const Sturm& S, double x) {  // distinct roots in (x, +inf)
  return sturm_changes(S, x, 0) - sturm_changes(S, 0.0, 1);
}

// Laguerre's iteration on the monic degree-6 p from x outside all roots: for a polynomial
// with only real roots it converges monotonically, and cubically for a simple root, to the
// nearest root.  It stops after a correction below 1e-14 (relative), or, when the roots are
// known to be simple, after applying one below 1e-6: the cubic rate leaves ~1e-18 then,
// which saves the confirming iteration.  Returns the iterations used, or -1 without
// convergence within kmax.
HD int laguerre_extreme(const double c[7], double& x, int kmax, bool simple) {
  const double tol = simple ? 1e-6 : 1e-14;
  const double n = 6.0;
  for (int it = 1; it <= kmax; ++it) {
    double p = 1.0, dp = 0.0, d2p = 0.0;
#pragma unroll
    for (int t = 5; t >= 0; --t) {
      d2p = d2p * x + 2.0 * dp;
      dp = dp * x + p;
      p = p * x + c[t];
    }
    if (p == 0.0) return it;
    const double G = dp / p;
    const double H = G * G - d2p / p;
    const double sq = sqrt(fmax((n - 1.0) * (n * H - G * G), 0.0));
    const double den = G >= 0.0 ? G + sq : G - sq;
    const double dx = n / den;
    x -= dx;
    if (!(fabs(dx) > tol * (1.0 + fabs(x)))) return it;
  }
  return -1;
}

// A warm start w is moved outward by this margin (relative, plus absolute) and used only if
// the Sturm count proves the moved point lies outside all roots.
constexpr double kWarmMargin = 0.005;

struct FastEig {
  double lmin, lmax;  // standardized extreme eigenvalues
  int iters;          // root-polishing iterations (both ends)
  bool real;          // proven: all six eigenvalues real
};
// Standardized fast solve.  wmin/wmax: warm starts (NaN = none).  need_vals=false only
// proves hyperbolicity (pass 2).
HD FastEig fast_eigs_std(const double* J, double wmin, double wmax, bool need_vals) {
  FastEig r;
  r.real = false;
  r.iters = 0;
  r.lmin = r.lmax = NAN;
  double c[7];
  charpoly(J, c);
  for (int t = 0; t < 7; ++t)
    if (!(fabs(c[t]) < 1e300)) return r;
  Sturm S;
  sturm_build(c, S);
  if (!S.ok) return r;
  // all six roots real <=> distinct real roots = 6 - deg gcd(p, p')
  const int nreal = sturm_real_roots(S);
  if (nreal != 6 - sturm_gcd_degree(S)) return r;
  r.real = true;
  if (!need_vals) return r;
  // Samuelson bounds: all roots in mu +- sqrt(5) s
  const double mu = -c[5] / 6.0;
  const double s2 = fmax((c[5] * c[5] - 2.0 * c[4]) / 6.0 - mu * mu, 0.0);
  const double half = sqrt(5.0 * s2) * 1.0001 + 1e-12 * (1.0 + fabs(mu));
  double x0 = mu - half, x1 = mu + half;
  if (wmin == wmin) {
    const double g = wmin - kWarmMargin * (1.0 + fabs(wmin));
    if (g > x0 && sturm_roots_above(S, g) == nreal) x0 = g;
  }
  if (wmax == wmax) {
    const double g = wmax + kWarmMargin * (1.0 + fabs(wmax));
    if (g < x1 && sturm_roots_above(S, g) == 0) x1 = g;
  }
  const bool simple = sturm_gcd_degree(S) == 0;  // no multiple root anywhere
  const int k0 = laguerre_extreme(c, x0, 40, simple), k1 = laguerre_extreme(c, x1, 40, simple);
  r.iters = (k0 < 0 ? 40 : k0) + (k1 < 0 ? 40 : k1);
  if (k0 < 0 || k1 < 0) {
    r.real = false;  // not converged: the caller falls back to the full solve
    return r;
  }
  r.lmin = x0;
  r.lmax = x1;
  return r;
}

}  // namespace hyq
