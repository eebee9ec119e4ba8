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
  bool pair;  // not ok only because the last remainder (the constant p6) is ill-conditioned:
              // p0..p5 are resolved, and p has a nearly double root near the root of p5
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
  S.pair = false;
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
  S.pair = st == 2 && S.n == 6;
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
// The same count for a chain whose last member is the constant s6 (S.pair chains).
HD int sturm_real_roots_p6(const Sturm& S, double s6) {
  Sturm T = S;
  T.p6[0] = s6;
  T.n = 7;
  return sturm_real_roots(T);
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

// The nearly double eigenvalue of a plane Jacobian whose Sturm chain is resolved up to p5.
// A characteristic polynomial cannot separate two roots closer than about sqrt(eps) (its
// coefficients carry rounding of order eps), but the matrix can: inverse subspace
// iteration with J - zI, z the root of p5, converges to the invariant subspace of the pair
// (the other four eigenvalues are resolved by the chain and lie far from z), and the 2x2
// block B = Q^T J Q on that subspace has the pair as its eigenvalues, to the rounding of
// J's entries.  Returns +1: a real pair; 0: a complex pair with |imag| (raw scale, sig
// times the standardized value) clearly above the reference's 1000 eps test
// (eigenvalues6_hyperbolic_3D.jl:35); -1: not decided here (within the rounding bound of
// either case, or a singular shift); 2: a complex pair whose |imag| is clearly below the
// 1000 eps test (treated as real by the reference's decision).
HD int pair_block_decision(const double* J, double z, double sig) {
  double A[36];
#pragma unroll
  for (int t = 0; t < 36; ++t) A[t] = J[t];
#pragma unroll
  for (int i = 0; i < 6; ++i) A[i * 7] -= z;
  int piv[6];
  // LU with partial pivoting, in place
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int pk = k;
    double best = fabs(A[k * 6 + k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i)
      if (fabs(A[i * 6 + k]) > best) {
        best = fabs(A[i * 6 + k]);
        pk = i;
      }
    piv[k] = pk;
    if (!(best > 0.0)) return -1;
    // exchange rows k and pk by selects over the candidate rows: every index stays a
    // compile-time constant, so A stays in registers (an A[pk * 6 + j] with a run-time pk
    // would put the matrix in scratch memory)
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const bool sw = i == pk;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double a = A[k * 6 + j], b = A[i * 6 + j];
        A[k * 6 + j] = sw ? b : a;
        A[i * 6 + j] = sw ? a : b;
      }
    }
    const double inv = 1.0 / A[k * 6 + k];
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double l = A[i * 6 + k] * inv;
      A[i * 6 + k] = l;
#pragma unroll
      for (int j = k + 1; j < 6; ++j) A[i * 6 + j] -= l * A[k * 6 + j];
    }
  }
  // two generic start vectors, three inverse iterations, Gram-Schmidt after each
  double X[2][6] = {{1.0, 0.8, 0.6, 0.4, 0.2, 0.1}, {0.1, -0.3, 0.5, -0.7, 0.9, -1.1}};
  for (int it = 0; it < 3; ++it) {
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      double* x = X[v];
#pragma unroll
      for (int k = 0; k < 6; ++k) {  // apply the row exchanges, then L, then U
        const double t = x[k];
        double tk = x[k];
#pragma unroll
        for (int q = 0; q < 6; ++q)
          if (q == piv[k]) tk = x[q];
#pragma unroll
        for (int q = 0; q < 6; ++q)
          if (q == piv[k]) x[q] = t;
        x[k] = tk;
      }
#pragma unroll
      for (int i = 1; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) x[i] -= A[i * 6 + j] * x[j];
#pragma unroll
      for (int i = 5; i >= 0; --i) {
#pragma unroll
        for (int j = i + 1; j < 6; ++j) x[i] -= A[i * 6 + j] * x[j];
        x[i] /= A[i * 6 + i];
      }
    }
    double n0 = 0.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) n0 += X[0][t] * X[0][t];
    n0 = 1.0 / sqrt(n0);
    double d = 0.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      X[0][t] *= n0;
      d += X[0][t] * X[1][t];
    }
    double n1 = 0.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      X[1][t] -= d * X[0][t];
      n1 += X[1][t] * X[1][t];
    }
    if (!(n1 > 0.0) || !(n1 < 1e300)) return -1;
    n1 = 1.0 / sqrt(n1);
#pragma unroll
    for (int t = 0; t < 6; ++t) X[1][t] *= n1;
  }
  // B = Q^T J Q and its discriminant
  double JQ[2][6];
  double jn = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      a0 += J[i * 6 + j] * X[0][j];
      a1 += J[i * 6 + j] * X[1][j];
      jn += J[i * 6 + j] * J[i * 6 + j];
    }
    JQ[0][i] = a0;
    JQ[1][i] = a1;
  }
  double b[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) acc += X[r][i] * JQ[q][i];
      b[r][q] = acc;
    }
  // residual R = J Q - Q B: Q spans an invariant subspace of J - R Q^T, whose block is B,
  // so B's eigenvalues are those of J perturbed by |R|.  Inverse iteration converges slowly
  // on a pair that is nearly a Jordan block; such a subspace is not trusted.
  double rn = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const double t = JQ[q][i] - (X[0][i] * b[0][q] + X[1][i] * b[1][q]);
      rn += t * t;
    }
  rn = sqrt(rn);
  if (!(rn <= 1e-10 * sqrt(jn))) return -1;
  const double dd = b[0][0] - b[1][1];
  const double disc = dd * dd + 4.0 * b[0][1] * b[1][0];
  // bound on the error of B's entries: LAPACK's backward error (a generous multiple of eps
  // times the Frobenius norm of J) plus the residual, amplified for the subspace's coupling
  const double e = 256.0 * EPS * sqrt(jn) + 16.0 * rn;
  const double err = 4.0 * fabs(dd) * e + 8.0 * (fabs(b[0][1]) + fabs(b[1][0])) * e + 16.0 * e * e;
  if (!(disc == disc)) return -1;
  if (disc > err) return 1;
  // |imag| of the pair within sig/2 sqrt(max(0, -disc -+ err))
  if (sig * 0.5 * sqrt(fmax(err - disc, 0.0)) < 0.25 * 1000 * EPS) return 2;
  if (disc < -err && sig * 0.5 * sqrt(-disc - err) > 4.0 * 1000 * EPS) return 0;
  return -1;
}

// A warm start w is moved outward by this margin (relative, plus absolute) and used only if
// the Sturm count proves the moved point lies outside all roots.
constexpr double kWarmMargin = 0.005;

struct FastEig {
  double lmin, lmax;  // standardized extreme eigenvalues
  int iters;          // root-polishing iterations (both ends)
  bool real;          // proven: no eigenvalue with |imag| above the reference's test
  bool cplx;          // proven: a complex pair above the reference's test
};
// Standardized fast solve.  wmin/wmax: warm starts (NaN = none).  need_vals=false only
// proves hyperbolicity (pass 2).  sig: the plane's standard deviation (the raw scale of the
// imaginary parts).  PAIR: decide a nearly double pair with the 2x2 block test; without it
// such a plane is left undecided (kernels that must stay small hand it on).
template <bool PAIR = true>
HD FastEig fast_eigs_std(const double* J, double wmin, double wmax, bool need_vals,
                         double sig = 1.0) {
  FastEig r;
  r.real = false;
  r.cplx = false;
  r.iters = 0;
  r.lmin = r.lmax = NAN;
  double c[7];
  charpoly(J, c);
  for (int t = 0; t < 7; ++t)
    if (!(fabs(c[t]) < 1e300)) return r;
  Sturm S;
  sturm_build(c, S);
  int nreal;
  bool pair = false;
  if (S.ok) {
    // all six roots real <=> distinct real roots = 6 - deg gcd(p, p')
    nreal = sturm_real_roots(S);
    if (nreal != 6 - sturm_gcd_degree(S)) {
      // a complex pair in a resolved chain: every remainder is above 1e-9 of its dividend,
      // so no two roots are nearly double and the pair's imaginary part is of the order of
      // sqrt(1e-9) of the roots' spread, orders of magnitude above the reference's
      // 1000 eps test (eigenvalues6_hyperbolic_3D.jl:35): the eigenvalue correction
      r.cplx = true;
      return r;
    }
  } else {
    if (!PAIR || !S.pair) return r;
    // p0..p5 resolved: four simple real roots and a nearly double pair near the root of p5,
    // real when the unresolved p6 takes one sign and complex with the other
    const int np = sturm_real_roots_p6(S, 1.0), nm = sturm_real_roots_p6(S, -1.0);
    if (!((np == 6 && nm == 4) || (np == 4 && nm == 6))) return r;
    if (!(fabs(S.p5[1]) > 0.0)) return r;
    const int d = pair_block_decision(J, -S.p5[0] / S.p5[1], sig);
    if (d < 0) return r;
    if (d == 0) {
      r.cplx = true;
      return r;
    }
    pair = true;
    nreal = 6;
  }
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
  const bool simple = S.ok && sturm_gcd_degree(S) == 0;  // no multiple root anywhere
  const int k0 = laguerre_extreme(c, x0, 40, simple), k1 = laguerre_extreme(c, x1, 40, simple);
  r.iters = (k0 < 0 ? 40 : k0) + (k1 < 0 ? 40 : k1);
  if (k0 < 0 || k1 < 0) {
    r.real = false;  // not converged: the caller falls back to the full solve
    return r;
  }
  r.lmin = x0;
  r.lmax = x1;
  if (pair) {
    // the pair must be interior: an extreme root nearly double is left to the full solve
    const double z = -S.p5[0] / S.p5[1];
    if (!(z > x0 + 1e-3 * (1.0 + fabs(x0)) && z < x1 - 1e-3 * (1.0 + fabs(x1)))) r.real = false;
  }
  return r;
}

}  // namespace hyq
