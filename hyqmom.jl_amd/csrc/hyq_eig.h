// Wave speeds: eigenvalues of the 6x6 flux Jacobian (jacobian6.jl) and of the 1D HyQMOM
// Jacobi matrix (closure_and_eigenvalues.jl).  FP64, one cell per lane.
//
// The Jacobian of a plane (flux direction f, other direction g) is shift- and
// scale-covariant: with a = mean of f, sigma = sqrt(c20) and S_ij the standardized central
// moments of the plane,  eig(J(m)) = a + sigma * eig(J(S)).  J(S) is sparse (rows 1, 2, 4
// are shift rows; the other entries are short polynomials in S), so the device builds it
// from the standardized moments and runs the full FP64 eigen solve on it: balancing,
// Householder reduction to Hessenberg form, Francis double-shift QR with standardized 2x2
// blocks (the LAPACK dgeev family that the reference's eigvals calls).
#pragma once
#include "hyq_cell.h"

namespace hyq {

struct Plane {  // standardized central moments of one plane (c20 = c02 = 1)
  double S11, S30, S21, S12, S03, S40, S31, S22, S13, S04;
  double a, sig;  // mean and standard deviation along the flux direction
  bool ok;        // false if a variance is not positive / not finite
};

// Plane (f, g) from packed central moments c (order <= 4) and the three means.
template <int F, int G>
HD Plane plane(const double* c, double um, double vm, double wm) {
  auto C = [&](int i, int j) {
    int e[3] = {0, 0, 0};
    e[F] += i;
    e[G] += j;
    return c[tpos(e[0], e[1], e[2])];
  };
  Plane p;
  const double c20 = C(2, 0), c02 = C(0, 2);
  p.a = F == 0 ? um : (F == 1 ? vm : wm);
  p.ok = (c20 > 0) && (c02 > 0) && (c20 < 1e300) && (c02 < 1e300);
  const double sf = sqrt(c20), sg = sqrt(c02);
  p.sig = sf;
  const double f1 = 1.0 / sf, g1 = 1.0 / sg;
  const double f2 = f1 * f1, g2 = g1 * g1;
  p.S11 = C(1, 1) * f1 * g1;
  p.S30 = C(3, 0) * f2 * f1;
  p.S21 = C(2, 1) * f2 * g1;
  p.S12 = C(1, 2) * f1 * g2;
  p.S03 = C(0, 3) * g2 * g1;
  p.S40 = C(4, 0) * f2 * f2;
  p.S31 = C(3, 1) * f2 * f1 * g1;
  p.S22 = C(2, 2) * f2 * g2;
  p.S13 = C(1, 3) * f1 * g2 * g1;
  p.S04 = C(0, 4) * g2 * g2;
  return p;
}

// Standardized flux Jacobian (row-major 6x6).
HD void jacobian_std(const Plane& p, double* J) {
#pragma unroll
  for (int t = 0; t < 36; ++t) J[t] = 0.0;
  const double S30 = p.S30, S03 = p.S03, S11 = p.S11, S21 = p.S21, S12 = p.S12, S40 = p.S40,
               S04 = p.S04, S31 = p.S31, S13 = p.S13, S22 = p.S22;
  J[0 * 6 + 1] = 1.0;
  J[1 * 6 + 2] = 1.0;
  J[3 * 6 + 4] = 1.0;
  const double J34 = -(2 * S11 * S40 - 3 * S11 * S30 * S30 - 2 * S31 + 3 * S21 * S30) / 2;
  J[2 * 6 + 0] = -S30 / 2 - S03 * J34;
  J[2 * 6 + 1] = S40 - 1.5 * S30 * S30;
  J[2 * 6 + 2] = 1.5 * S30;
  J[2 * 6 + 3] = J34;
  J[4 * 6 + 0] = (3 * S03 * S12 * S30 - 3 * S03 * S22 + 2 * S04 * S11 * S30 - 2 * S04 * S21 -
                  6 * S03 * S03 * S11 * S30 + 6 * S03 * S03 * S21) / 2;
  J[4 * 6 + 1] = -1.5 * S30 * S03;
  J[4 * 6 + 2] = 1.5 * S03;
  J[4 * 6 + 3] = -(1 + 3 * S12 * S30 - 3 * S22 - 6 * S03 * S11 * S30 + 6 * S03 * S21) / 2;
  J[4 * 6 + 4] = S30;
  J[4 * 6 + 5] = S21 - S11 * S30;
  J[5 * 6 + 0] = (-6 * S12 - 8 * S03 * S13 - 10 * S04 * S12 + 30 * S03 * S03 * S12 +
                  16 * S03 * S04 * S11 - 27 * S03 * S03 * S03 * S11) / 4;
  J[5 * 6 + 1] = -(6 - 10 * S04 + 15 * S03 * S03) / 4;
  J[5 * 6 + 3] = -(-4 * S11 - 8 * S13 + 30 * S03 * S12 + 8 * S04 * S11 - 27 * S03 * S03 * S11) / 4;
  J[5 * 6 + 4] = 2 * S03;
  J[5 * 6 + 5] = (5 * S12 - 4 * S03 * S11) / 2;
}

// --- general real eigenvalues, n <= 6 (row-major A, destroyed) -----------------------
template <int N>
HD void balance_n(double* A) {
  bool done = false;
  for (int sweep = 0; sweep < 32 && !done; ++sweep) {
    done = true;
    for (int i = 0; i < N; ++i) {
      double r = 0, c = 0;
      for (int j = 0; j < N; ++j)
        if (j != i) { c += fabs(A[j * N + i]); r += fabs(A[i * N + j]); }
      if (c != 0.0 && r != 0.0) {
        double g = r / 2.0, f = 1.0;
        const double s = c + r;
        while (c < g) { f *= 2.0; c *= 4.0; }
        g = r * 2.0;
        while (c > g) { f /= 2.0; c /= 4.0; }
        if ((c + r) / f < 0.95 * s) {
          done = false;
          const double gi = 1.0 / f;
          for (int j = 0; j < N; ++j) A[i * N + j] *= gi;
          for (int j = 0; j < N; ++j) A[j * N + i] *= f;
        }
      }
    }
  }
}
template <int N>
HD void hessenberg_n(double* A) {
  for (int k = 0; k < N - 2; ++k) {
    double alpha = 0;
    for (int i = k + 1; i < N; ++i) alpha += A[i * N + k] * A[i * N + k];
    double v[N];
    for (int i = 0; i < N; ++i) v[i] = 0.0;
    for (int i = k + 1; i < N; ++i) v[i] = A[i * N + k];
    if (alpha < 1e-280) {
      // squares at the underflow threshold (subnormal entries): the reflector is that of
      // the column scaled to unit max-abs entry, as LAPACK's dlarfg rescales tiny columns
      double amax = 0.0;
      for (int i = k + 1; i < N; ++i) amax = fmax(amax, fabs(v[i]));
      if (amax == 0.0) continue;
      alpha = 0.0;
      for (int i = k + 1; i < N; ++i) {
        v[i] /= amax;
        alpha += v[i] * v[i];
      }
    }
    alpha = sqrt(alpha);
    if (alpha == 0.0) continue;
    if (A[(k + 1) * N + k] > 0) alpha = -alpha;
    v[k + 1] -= alpha;
    double vn = 0;
    for (int i = k + 1; i < N; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double tv = 2.0 / vn;
    for (int j = 0; j < N; ++j) {
      double d = 0;
      for (int i = k + 1; i < N; ++i) d += v[i] * A[i * N + j];
      d *= tv;
      for (int i = k + 1; i < N; ++i) A[i * N + j] -= d * v[i];
    }
    for (int i = 0; i < N; ++i) {
      double d = 0;
      for (int j = k + 1; j < N; ++j) d += A[i * N + j] * v[j];
      d *= tv;
      for (int j = k + 1; j < N; ++j) A[i * N + j] -= d * v[j];
    }
    for (int i = k + 2; i < N; ++i) A[i * N + k] = 0.0;
  }
}
HD void eig2x2(double a, double b, double c, double d, double& r1, double& i1, double& r2,
               double& i2) {
  const double p = 0.5 * (a - d), disc = p * p + b * c;
  if (disc >= 0) {
    const double sq = sqrt(disc), z = p + (p >= 0 ? sq : -sq);
    r1 = d + z;
    r2 = (z != 0.0) ? d - (b * c) / z : d;
    i1 = i2 = 0.0;
  } else {
    r1 = r2 = 0.5 * (a + d);
    i1 = sqrt(-disc);
    i2 = -i1;
  }
}
HD void house(double x, double y, double z, bool three, double& v1, double& v2, double& beta) {
  const double sig = y * y + (three ? z * z : 0.0);
  v1 = y;
  v2 = three ? z : 0.0;
  if (sig == 0.0) { beta = 0.0; return; }
  const double mu = sqrt(x * x + sig);
  const double v0 = (x <= 0) ? x - mu : -sig / (x + mu);
  beta = 2.0 * v0 * v0 / (sig + v0 * v0);
  v1 = y / v0;
  v2 = three ? z / v0 : 0.0;
}
// Francis double-shift QR on an upper Hessenberg matrix, on any 2D accessor H(i, j); each
// deflated eigenvalue (or conjugate pair) goes to sink(index, re, im).  Returns false if it
// stalled.
template <int N, typename Mat, typename Sink>
HD bool hqr_core(Mat& H, Sink& sink) {
  int hi = N - 1, its = 0;
  while (hi >= 0) {
    int lo = hi;
    while (lo > 0) {
      double s = fabs(H(lo - 1, lo - 1)) + fabs(H(lo, lo));
      if (s == 0.0) s = 1.0;
      if (fabs(H(lo, lo - 1)) <= EPS * s) { H(lo, lo - 1) = 0.0; break; }
      --lo;
    }
    if (lo == hi) { sink(hi, H(hi, hi), 0.0); --hi; its = 0; continue; }
    if (lo == hi - 1) {
      double r1, i1, r2, i2;
      eig2x2(H(hi - 1, hi - 1), H(hi - 1, hi), H(hi, hi - 1), H(hi, hi), r1, i1, r2, i2);
      sink(hi - 1, r1, i1);
      sink(hi, r2, i2);
      hi -= 2; its = 0; continue;
    }
    if (its >= 100) return false;
    ++its;
    double s, t;
    if (its % 10 == 0) {
      const double w = fabs(H(hi, hi - 1)) + fabs(H(hi - 1, hi - 2));
      s = 1.5 * w;
      t = w * w;
    } else {
      s = H(hi - 1, hi - 1) + H(hi, hi);
      t = H(hi - 1, hi - 1) * H(hi, hi) - H(hi - 1, hi) * H(hi, hi - 1);
    }
    double x = H(lo, lo) * H(lo, lo) + H(lo, lo + 1) * H(lo + 1, lo) - s * H(lo, lo) + t;
    double y = H(lo + 1, lo) * (H(lo, lo) + H(lo + 1, lo + 1) - s);
    double z = H(lo + 1, lo) * H(lo + 2, lo + 1);
    for (int k = lo - 1; k <= hi - 3; ++k) {
      double v1, v2, beta;
      house(x, y, z, true, v1, v2, beta);
      const int q = lo > k ? lo : k;
      for (int j = q; j <= hi; ++j) {
        const double d = beta * (H(k + 1, j) + v1 * H(k + 2, j) + v2 * H(k + 3, j));
        H(k + 1, j) -= d;
        H(k + 2, j) -= d * v1;
        H(k + 3, j) -= d * v2;
      }
      const int r = k + 4 < hi ? k + 4 : hi;
      for (int i = lo; i <= r; ++i) {
        const double d = beta * (H(i, k + 1) + v1 * H(i, k + 2) + v2 * H(i, k + 3));
        H(i, k + 1) -= d;
        H(i, k + 2) -= d * v1;
        H(i, k + 3) -= d * v2;
      }
      x = H(k + 2, k + 1);
      y = H(k + 3, k + 1);
      if (k < hi - 3) z = H(k + 4, k + 1);
    }
    double v1, v2, beta;
    house(x, y, 0.0, false, v1, v2, beta);
    for (int j = hi - 2; j <= hi; ++j) {
      const double d = beta * (H(hi - 1, j) + v1 * H(hi, j));
      H(hi - 1, j) -= d;
      H(hi, j) -= d * v1;
    }
    for (int i = lo; i <= hi; ++i) {
      const double d = beta * (H(i, hi - 1) + v1 * H(i, hi));
      H(i, hi - 1) -= d;
      H(i, hi) -= d * v1;
    }
  }
  return true;
}

// The same Francis iteration for N = 6 on a row-major array H[36] that stays in registers:
// every index is a compile-time constant.  One double-shift step on the active window
// [LO, HI] is instantiated per window (10 windows of size >= 3), the deflation search and
// the reads at the dynamic bottom index are unrolled over all positions with predicates.
// The arithmetic, its order and the sink calls are those of hqr_core<6>, so both give the
// same bits.
template <int LO, int HI>
HD void francis_step6(double* H, int its) {
  auto A = [H](int i, int j) -> double& { return H[i * 6 + j]; };
  double s, t;
  if (its % 10 == 0) {
    const double w = fabs(A(HI, HI - 1)) + fabs(A(HI - 1, HI - 2));
    s = 1.5 * w;
    t = w * w;
  } else {
    s = A(HI - 1, HI - 1) + A(HI, HI);
    t = A(HI - 1, HI - 1) * A(HI, HI) - A(HI - 1, HI) * A(HI, HI - 1);
  }
  double x = A(LO, LO) * A(LO, LO) + A(LO, LO + 1) * A(LO + 1, LO) - s * A(LO, LO) + t;
  double y = A(LO + 1, LO) * (A(LO, LO) + A(LO + 1, LO + 1) - s);
  double z = A(LO + 1, LO) * A(LO + 2, LO + 1);
#pragma unroll
  for (int k = LO - 1; k <= HI - 3; ++k) {
    double v1, v2, beta;
    house(x, y, z, true, v1, v2, beta);
    const int q = LO > k ? LO : k;
#pragma unroll
    for (int j = q; j <= HI; ++j) {
      const double d = beta * (A(k + 1, j) + v1 * A(k + 2, j) + v2 * A(k + 3, j));
      A(k + 1, j) -= d;
      A(k + 2, j) -= d * v1;
      A(k + 3, j) -= d * v2;
    }
    const int r = k + 4 < HI ? k + 4 : HI;
#pragma unroll
    for (int i = LO; i <= r; ++i) {
      const double d = beta * (A(i, k + 1) + v1 * A(i, k + 2) + v2 * A(i, k + 3));
      A(i, k + 1) -= d;
      A(i, k + 2) -= d * v1;
      A(i, k + 3) -= d * v2;
    }
    x = A(k + 2, k + 1);
    y = A(k + 3, k + 1);
    if (k < HI - 3) z = A(k + 4, k + 1);
  }
  double v1, v2, beta;
  house(x, y, 0.0, false, v1, v2, beta);
#pragma unroll
  for (int j = HI - 2; j <= HI; ++j) {
    const double d = beta * (A(HI - 1, j) + v1 * A(HI, j));
    A(HI - 1, j) -= d;
    A(HI, j) -= d * v1;
  }
#pragma unroll
  for (int i = LO; i <= HI; ++i) {
    const double d = beta * (A(i, HI - 1) + v1 * A(i, HI));
    A(i, HI - 1) -= d;
    A(i, HI) -= d * v1;
  }
}

template <typename Sink>
HD bool hqr6_reg(double* H, Sink& sink) {
  int hi = 5, its = 0;
  while (hi >= 0) {
    int lo = 0;
    bool found = false;
#pragma unroll
    for (int q = 5; q >= 1; --q) {
      if (!found && q <= hi) {
        double s = fabs(H[(q - 1) * 7]) + fabs(H[q * 7]);
        if (s == 0.0) s = 1.0;
        if (fabs(H[q * 7 - 1]) <= EPS * s) {
          H[q * 7 - 1] = 0.0;
          lo = q;
          found = true;
        }
      }
    }
    if (lo == hi) {
      double d = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (q == hi) d = H[q * 7];
      sink(hi, d, 0.0);
      --hi;
      its = 0;
      continue;
    }
    if (lo == hi - 1) {
      double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
#pragma unroll
      for (int q = 1; q < 6; ++q)
        if (q == hi) {
          a = H[(q - 1) * 7];
          b = H[(q - 1) * 7 + 1];
          c = H[q * 7 - 1];
          d = H[q * 7];
        }
      double r1, i1, r2, i2;
      eig2x2(a, b, c, d, r1, i1, r2, i2);
      sink(hi - 1, r1, i1);
      sink(hi, r2, i2);
      hi -= 2;
      its = 0;
      continue;
    }
    if (its >= 100) return false;
    ++its;
    switch (lo * 8 + hi) {
      case 0 * 8 + 2: francis_step6<0, 2>(H, its); break;
      case 0 * 8 + 3: francis_step6<0, 3>(H, its); break;
      case 1 * 8 + 3: francis_step6<1, 3>(H, its); break;
      case 0 * 8 + 4: francis_step6<0, 4>(H, its); break;
      case 1 * 8 + 4: francis_step6<1, 4>(H, its); break;
      case 2 * 8 + 4: francis_step6<2, 4>(H, its); break;
      case 0 * 8 + 5: francis_step6<0, 5>(H, its); break;
      case 1 * 8 + 5: francis_step6<1, 5>(H, its); break;
      case 2 * 8 + 5: francis_step6<2, 5>(H, its); break;
      default: francis_step6<3, 5>(H, its); break;
    }
  }
  return true;
}

// Row-major N x N array accessor.
template <int N>
struct ArrMat {
  double* a;
  HD double& operator()(int i, int j) { return a[i * N + j]; }
};
// Sinks: all eigenvalues into arrays, or only the extreme real parts and max |imag|.
struct ArraySink {
  double *wr, *wi;
  HD void operator()(int k, double re, double im) { wr[k] = re; wi[k] = im; }
};
struct ExtremeSink {
  double mn, mx, mi;
  bool first;
  HD void operator()(int, double re, double im) {
    mn = first ? re : fmin(mn, re);
    mx = first ? re : fmax(mx, re);
    mi = first ? fabs(im) : fmax(mi, fabs(im));
    first = false;
  }
};

template <int N>
HD bool hqr_n(double* H, double* wr, double* wi) {
  ArrMat<N> m{H};
  ArraySink s{wr, wi};
  return hqr_core<N>(m, s);
}

struct EigRes {
  double vmin, vmax;  // extreme real parts (raw scale), NaN if not finite
  bool cplx;          // some |imag| > 1000 eps (raw scale)
  bool ok;
};
// Eigenvalues of the plane Jacobian mapped back to the raw scale.
HD EigRes plane_eigs(const Plane& p, uint32_t& fl) {
  EigRes r;
  double H[36];
  jacobian_std(p, H);
  bool fin = p.ok;
#pragma unroll
  for (int t = 0; t < 36; ++t) fin = fin && (fabs(H[t]) < 1e300);
  if (!fin) {
    r.vmin = r.vmax = NAN;
    r.cplx = false;
    r.ok = false;
    return r;
  }
  balance_n<6>(H);
  hessenberg_n<6>(H);
  ExtremeSink ex{0.0, 0.0, 0.0, true};
  const bool conv = hqr6_reg(H, ex);
  if (!conv) {
    fl |= F_NOCONV;
    r.vmin = r.vmax = NAN;
    r.cplx = false;
    r.ok = false;
    return r;
  }
  r.vmin = p.a + p.sig * ex.mn;
  r.vmax = p.a + p.sig * ex.mx;
  r.cplx = p.sig * ex.mi > 1000 * EPS;
  r.ok = true;
  return r;
}

// closure_and_eigenvalues (src/numerics/closure_and_eigenvalues.jl:17-78) for N = 2: the
// Chebyshev recurrence gives a1, a2, b2, b3; the closure sets a3 = (a1+a2)/2 and scales b3
// by 5/2; the eigenvalues of the 3x3 Jacobi matrix are the roots of the real cubic
// (l-a1)(l-a2)(l-a3) - b2 (l-a3) - b3 (l-a1).  Returns min/max real parts.
HD void closure1d_speeds(double m0, double m1, double m2, double m3, double m4, double& vmin,
                         double& vmax) {
  const double a1 = m1 / m0;
  const double s33 = m2 - a1 * m1, s34 = m3 - a1 * m2, s35 = m4 - a1 * m3;
  const double a2 = s34 / s33 - m1 / m0;
  const double b2 = s33 / m0;
  const double s44 = s35 - a2 * s34 - b2 * m2;
  double b3 = s44 / s33;
  const double a3 = (a1 + a2) / 2;
  b3 = b3 * 5.0 / 2.0;
  if (!(fabs(a1) < 1e300 && fabs(a2) < 1e300 && fabs(b2) < 1e300 && fabs(b3) < 1e300)) {
    vmin = vmax = NAN;
    return;
  }
  // monic cubic l^3 + B l^2 + C l + D
  const double B = -(a1 + a2 + a3);
  const double C = a1 * a2 + a1 * a3 + a2 * a3 - b2 - b3;
  const double D = -a1 * a2 * a3 + b2 * a3 + b3 * a1;
  const double sh = -B / 3.0;
  const double p = C - B * B / 3.0;
  const double q = 2.0 * B * B * B / 27.0 - B * C / 3.0 + D;
  const double disc = (q / 2.0) * (q / 2.0) + (p / 3.0) * (p / 3.0) * (p / 3.0);
  if (disc <= 0.0 && p < 0.0) {  // three real roots (trigonometric form)
    const double r = 2.0 * sqrt(-p / 3.0);
    double arg = 3.0 * q / (p * r);
    arg = arg > 1.0 ? 1.0 : (arg < -1.0 ? -1.0 : arg);
    const double th = acos(arg) / 3.0;
    const double t0 = r * cos(th), t1 = r * cos(th - 2.0943951023931957),
                 t2 = r * cos(th - 4.1887902047863905);
    vmin = sh + fmin(t0, fmin(t1, t2));
    vmax = sh + fmax(t0, fmax(t1, t2));
  } else {
    const double sq = sqrt(fmax(disc, 0.0));
    const double u = cbrt(-q / 2.0 + sq), v = cbrt(-q / 2.0 - sq);
    const double t0 = u + v, tr = -(u + v) / 2.0;
    vmin = sh + fmin(t0, tr);
    vmax = sh + fmax(t0, tr);
  }
}

}  // namespace hyq
