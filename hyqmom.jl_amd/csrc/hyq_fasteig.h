// Fast wave speeds for the learned / Newton modes: the characteristic polynomial of the
// standardized flux Jacobian, a Sturm-sequence hyperbolicity test, and Newton refinement of
// the extreme eigenvalues from a warm start.  Cells where the test is not conclusive, or
// where Newton does not converge in a bounded number of iterations, fall back to the full
// FP64 eigen solve (hyq_eig.h).  Roots that coincide to rounding level (a Maxwellian plane
// has a double root at its mean) count as one multiple real root; a complex pair that
// close to the real axis is below what the full solve resolves either.
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

// Sturm sequence of a degree-6 polynomial: s[k] holds polynomial k (coefficients of l^0..),
// deg[k] its degree.  Each member is normalized to unit max-abs coefficient (signs kept).
struct Sturm {
  double s[7][7];
  int deg[7];
  int n;      // number of members; deg[n-1] > 0 when the sequence ends at gcd(p, p')
  bool ok;    // false: ill-conditioned (a remainder small but not at rounding level)
};
HD void sturm_build(const double c[7], Sturm& S) {
  for (int k = 0; k < 7; ++k)
    for (int t = 0; t < 7; ++t) S.s[k][t] = 0.0;
  for (int t = 0; t < 7; ++t) S.s[0][t] = c[t];
  S.deg[0] = 6;
  for (int t = 0; t < 6; ++t) S.s[1][t] = (t + 1) * c[t + 1];
  S.deg[1] = 5;
  S.n = 2;
  S.ok = true;
  // normalize
  for (int k = 0; k < 2; ++k) {
    double m = 0;
    for (int t = 0; t <= S.deg[k]; ++t) m = fmax(m, fabs(S.s[k][t]));
    if (!(m > 0)) { S.ok = false; return; }
    for (int t = 0; t <= S.deg[k]; ++t) S.s[k][t] /= m;
  }
  while (S.deg[S.n - 1] > 0) {
    const int a = S.n - 2, b = S.n - 1;
    double r[7];
    for (int t = 0; t < 7; ++t) r[t] = S.s[a][t];
    int dr = S.deg[a];
    const int db = S.deg[b];
    const double lb = S.s[b][db];
    while (dr >= db) {
      const double q = r[dr] / lb;
      for (int t = 0; t <= db; ++t) r[dr - db + t] -= q * S.s[b][t];
      r[dr] = 0.0;
      --dr;
    }
    // trim and test for a vanishing remainder
    double m = 0;
    for (int t = 0; t <= dr; ++t) m = fmax(m, fabs(r[t]));
    // members have unit scale: a remainder at rounding level means (numerically) multiple
    // roots, and the last member is then the gcd of p and p'; a small but resolved one is
    // ill-conditioned and left to the full solve
    if (!(m > 1e-12)) break;
    if (!(m > 1e-9)) { S.ok = false; return; }
    while (dr > 0 && fabs(r[dr]) <= 1e-12 * m) --dr;
    const int k = S.n;
    for (int t = 0; t < 7; ++t) S.s[k][t] = 0.0;
    for (int t = 0; t <= dr; ++t) S.s[k][t] = -r[t] / m;
    S.deg[k] = dr;
    S.n = k + 1;
    if (S.n == 7) break;
  }
}
HD int sign_changes(const double* v, int n) {
  int ch = 0;
  double prev = 0.0;
  for (int k = 0; k < n; ++k) {
    if (v[k] == 0.0) continue;
    if (prev != 0.0 && ((v[k] > 0) != (prev > 0))) ++ch;
    prev = v[k];
  }
  return ch;
}
HD int sturm_real_roots(const Sturm& S) {  // distinct real roots on the whole line
  double vp[7], vm[7];
  for (int k = 0; k < S.n; ++k) {
    const double lc = S.s[k][S.deg[k]];
    vp[k] = lc;
    vm[k] = (S.deg[k] & 1) ? -lc : lc;
  }
  return sign_changes(vm, S.n) - sign_changes(vp, S.n);
}
HD int sturm_roots_above(const Sturm& S, double x) {  // roots in (x, +inf)
  double v[7], vp[7];
  for (int k = 0; k < S.n; ++k) {
    double acc = 0;
    for (int t = S.deg[k]; t >= 0; --t) acc = acc * x + S.s[k][t];
    v[k] = acc;
    vp[k] = S.s[k][S.deg[k]];
  }
  return sign_changes(v, S.n) - sign_changes(vp, S.n);
}

// Newton on p from x0 (outside all roots, on side dir = -1 left / +1 right).  Returns the
// number of iterations used, or -1 without convergence within kmax.
HD int newton_extreme(const double c[7], double& x, int kmax) {
  for (int it = 1; it <= kmax; ++it) {
    double p = 1.0, dp = 0.0;
#pragma unroll
    for (int t = 5; t >= 0; --t) {
      dp = dp * x + p;
      p = p * x + c[t];
    }
    const double dx = p / dp;
    x -= dx;
    if (!(fabs(dx) > 1e-14 * (1.0 + fabs(x)))) return it;
  }
  return -1;
}

struct FastEig {
  double lmin, lmax;  // standardized extreme eigenvalues
  int iters;          // Newton iterations (both ends)
  bool real;          // proven: six distinct real eigenvalues
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
  if (nreal != 6 - S.deg[S.n - 1]) return r;
  r.real = true;
  if (!need_vals) return r;
  // Samuelson bounds: all roots in mu +- sqrt(5) s
  const double mu = -c[5] / 6.0;
  const double s2 = fmax((c[5] * c[5] - 2.0 * c[4]) / 6.0 - mu * mu, 0.0);
  const double half = sqrt(5.0 * s2) * 1.0001 + 1e-12 * (1.0 + fabs(mu));
  double x0 = mu - half, x1 = mu + half;
  if (wmin == wmin) {
    const double g = wmin - (0.02 + 0.02 * fabs(wmin));
    if (g > x0 && sturm_roots_above(S, g) == nreal) x0 = g;
  }
  if (wmax == wmax) {
    const double g = wmax + (0.02 + 0.02 * fabs(wmax));
    if (g < x1 && sturm_roots_above(S, g) == 0) x1 = g;
  }
  const int k0 = newton_extreme(c, x0, 60), k1 = newton_extreme(c, x1, 60);
  r.iters = (k0 < 0 ? 60 : k0) + (k1 < 0 ? 60 : k1);
  if (k0 < 0 || k1 < 0) {
    r.real = false;  // not converged: the caller falls back to the full solve
    return r;
  }
  r.lmin = x0;
  r.lmax = x1;
  return r;
}

}  // namespace hyq
