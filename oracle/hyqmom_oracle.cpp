// CPU FP64 oracle for the HyQMOM.jl 3D moment-solver time step.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py may load this library, and only as the checker.  The product path
// (hyqmom.jl_amd/) never links or calls it.
//
// Every routine restates one function of the reference (comp-physics/HyQMOM.jl) and
// cites the file:line it follows.  The straight-line "autogen" functions of the
// reference are symbolic-toolbox outputs; here they are recomputed from the
// mathematical definitions they encode (verified numerically against the reference
// formulas, see tests/golden/autogen_vectors.json):
//   * raw <-> central moments: binomial expansion about the mean;
//   * delta2star3D / lower_bound_S220 / bound_minor1: the Schur complement of the
//     10x10 standardized moment (Gram) matrix over the basis
//     (1, x, y, z, x^2, xy, xz, y^2, yz, z^2) with respect to its leading 4x4 block;
//   * delta2starchol_L3: (1 - s11^2) * det of the 2D Schur complement;
//   * rootsR / rootsR_X_Y: Cardano roots (real sqrt, as under Julia @fastmath) of the
//     cubic determinant polynomial;
//   * jacobian6: the flux Jacobian expressed in central moments;
//   * LAPACK eigvals / det: balanced Hessenberg + Francis double-shift QR, and LU with
//     partial pivoting (the same algorithm families Julia's LinearAlgebra calls).
//
// Parity status: every function is pinned by at least one of the reference's MATLAB
// golden files, the autogen vectors, or a closed form (see tests/test_oracle_*.py).

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/hyqmom.h"

namespace orc {

constexpr double EPS = 2.220446049250313e-16;
static const double NaN = std::numeric_limits<double>::quiet_NaN();

// Julia semantics helpers ---------------------------------------------------------
static inline double jmax(double a, double b) {  // Base.max: NaN-propagating
  if (std::isnan(a) || std::isnan(b)) return NaN;
  return a > b ? a : b;
}
static inline double jmin(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NaN;
  return a < b ? a : b;
}
static inline double jsign(double x) {
  if (std::isnan(x)) return x;
  return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0);
}

// 35-moment ordering, Flux_closure35_and_realizable_3D.jl:18-20 --------------------
static const int E35[35][3] = {
    {0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {3, 0, 0}, {4, 0, 0}, {0, 1, 0}, {1, 1, 0},
    {2, 1, 0}, {3, 1, 0}, {0, 2, 0}, {1, 2, 0}, {2, 2, 0}, {0, 3, 0}, {1, 3, 0},
    {0, 4, 0}, {0, 0, 1}, {1, 0, 1}, {2, 0, 1}, {3, 0, 1}, {0, 0, 2}, {1, 0, 2},
    {2, 0, 2}, {0, 0, 3}, {1, 0, 3}, {0, 0, 4}, {0, 1, 1}, {1, 1, 1}, {2, 1, 1},
    {0, 2, 1}, {1, 2, 1}, {0, 3, 1}, {0, 1, 2}, {1, 1, 2}, {0, 1, 3}, {0, 2, 2}};

struct Cube {  // moments indexed by exponents, total order <= 5
  double v[6][6][6];
  double& operator()(int i, int j, int k) { return v[i][j][k]; }
  double operator()(int i, int j, int k) const { return v[i][j][k]; }
};

static double binom(int n, int k) {
  static const double B[7][7] = {{1}, {1, 1}, {1, 2, 1}, {1, 3, 3, 1}, {1, 4, 6, 4, 1},
                                 {1, 5, 10, 10, 5, 1}, {1, 6, 15, 20, 15, 6, 1}};
  return B[n][k];
}
static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}

// M4toC4_3D (src/autogen/M4toC4_3D.jl:14-27): central moments up to order 4.
static void raw_to_central4(const double* M, Cube& C) {
  std::memset(&C, 0, sizeof(C));
  Cube R;
  std::memset(&R, 0, sizeof(R));
  for (int n = 0; n < 35; ++n) R(E35[n][0], E35[n][1], E35[n][2]) = M[n] / M[0];
  const double u = M[1] / M[0], v = M[5] / M[0], w = M[15] / M[0];
  for (int i = 0; i <= 4; ++i)
    for (int j = 0; i + j <= 4; ++j)
      for (int k = 0; i + j + k <= 4; ++k) {
        double s = 0.0;
        for (int p = 0; p <= i; ++p)
          for (int q = 0; q <= j; ++q)
            for (int r = 0; r <= k; ++r)
              s += binom(i, p) * binom(j, q) * binom(k, r) * ipow(-u, i - p) * ipow(-v, j - q) *
                   ipow(-w, k - r) * R(p, q, r);
        C(i, j, k) = s;
      }
  C(0, 0, 0) = 1.0;
  C(1, 0, 0) = C(0, 1, 0) = C(0, 0, 1) = 0.0;
}

// C4toM4_3D / C5toM5_3D (src/autogen/C4toM4_3D.jl:9-63, C5toM5_3D.jl:16-28):
// raw moments up to `order` from central moments (C000=1, C100=C010=C001=0 implied).
static void central_to_raw(double M000, double u, double v, double w, const Cube& Cin, int order,
                           Cube& M) {
  Cube C = Cin;
  C(0, 0, 0) = 1.0;
  C(1, 0, 0) = C(0, 1, 0) = C(0, 0, 1) = 0.0;
  std::memset(&M, 0, sizeof(M));
  for (int i = 0; i <= order; ++i)
    for (int j = 0; i + j <= order; ++j)
      for (int k = 0; i + j + k <= order; ++k) {
        double s = 0.0;
        for (int p = 0; p <= i; ++p)
          for (int q = 0; q <= j; ++q)
            for (int r = 0; r <= k; ++r)
              s += binom(i, p) * binom(j, q) * binom(k, r) * ipow(u, i - p) * ipow(v, j - q) *
                   ipow(w, k - r) * C(p, q, r);
        M(i, j, k) = M000 * s;
      }
}

static void cube_to35(const Cube& C, double* out) {
  for (int n = 0; n < 35; ++n) out[n] = C(E35[n][0], E35[n][1], E35[n][2]);
}

// M2CS4_35 (src/moments/M2CS4_35.jl:15-98)
static void m2cs4(const double* M, double* C4, double* S4) {
  Cube C;
  raw_to_central4(M, C);
  const double s2 = std::sqrt(jmax(C(2, 0, 0), EPS));
  const double s0 = std::sqrt(jmax(C(0, 2, 0), EPS));
  const double sz = std::sqrt(jmax(C(0, 0, 2), EPS));
  cube_to35(C, C4);
  for (int n = 0; n < 35; ++n) {
    const int i = E35[n][0], j = E35[n][1], k = E35[n][2];
    const int o = i + j + k;
    if (o == 0) S4[n] = 1.0;
    else if (o == 1) S4[n] = 0.0;
    else if (o == 2 && (i == 2 || j == 2 || k == 2)) S4[n] = 1.0;
    else S4[n] = C(i, j, k) / (ipow(s2, i) * ipow(s0, j) * ipow(sz, k));
  }
}

// Small dense linear algebra (LAPACK getrf-style LU with partial pivoting) -----------
static double det_lu(const double* Ain, int n) {
  double A[36];
  std::memcpy(A, Ain, sizeof(double) * n * n);
  double d = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::fabs(A[c * n + c]);
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > best) { best = std::fabs(A[r * n + c]); p = r; }
    if (p != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[p * n + k]);
      d = -d;
    }
    const double piv = A[c * n + c];
    d *= piv;
    if (piv == 0.0) continue;
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / piv;
      for (int k = c + 1; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
    }
  }
  return d;
}
// det of the leading k x k block of a row-major 6x6 matrix (Julia det(E[1:k,1:k]))
static double det_lead(const double E[6][6], int k) {
  double A[36];
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) A[i * k + j] = E[i][j];
  return det_lu(A, k);
}
// Solve A X = B for n x n A and n x m B (row-major), partial pivoting.
static void solve_lu(const double* Ain, int n, double* B, int m) {
  double A[100];
  std::memcpy(A, Ain, sizeof(double) * n * n);
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::fabs(A[c * n + c]);
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > best) { best = std::fabs(A[r * n + c]); p = r; }
    if (p != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[p * n + k]);
      for (int k = 0; k < m; ++k) std::swap(B[c * m + k], B[p * m + k]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c + 1; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      for (int k = 0; k < m; ++k) B[r * m + k] -= f * B[c * m + k];
    }
  }
  for (int c = n - 1; c >= 0; --c)
    for (int k = 0; k < m; ++k) {
      double s = B[c * m + k];
      for (int j = c + 1; j < n; ++j) s -= A[c * n + j] * B[j * m + k];
      B[c * m + k] = s / A[c * n + c];
    }
}

// Standardized 28-moment set in the reference's positional order --------------------
struct S28 {
  double S300, S400, S110, S210, S310, S120, S220, S030, S130, S040, S101, S201, S301, S102,
      S202, S003, S103, S004, S011, S111, S211, S021, S121, S031, S012, S112, S013, S022;
};

// Standardized moments as a cube, including the implied S000=1, S1=0, S2 diagonal=1.
static void s28_cube(const S28& s, Cube& c) {
  std::memset(&c, 0, sizeof(c));
  c(0, 0, 0) = 1.0;
  c(2, 0, 0) = c(0, 2, 0) = c(0, 0, 2) = 1.0;
  c(3, 0, 0) = s.S300; c(4, 0, 0) = s.S400; c(1, 1, 0) = s.S110; c(2, 1, 0) = s.S210;
  c(3, 1, 0) = s.S310; c(1, 2, 0) = s.S120; c(2, 2, 0) = s.S220; c(0, 3, 0) = s.S030;
  c(1, 3, 0) = s.S130; c(0, 4, 0) = s.S040; c(1, 0, 1) = s.S101; c(2, 0, 1) = s.S201;
  c(3, 0, 1) = s.S301; c(1, 0, 2) = s.S102; c(2, 0, 2) = s.S202; c(0, 0, 3) = s.S003;
  c(1, 0, 3) = s.S103; c(0, 0, 4) = s.S004; c(0, 1, 1) = s.S011; c(1, 1, 1) = s.S111;
  c(2, 1, 1) = s.S211; c(0, 2, 1) = s.S021; c(1, 2, 1) = s.S121; c(0, 3, 1) = s.S031;
  c(0, 1, 2) = s.S012; c(1, 1, 2) = s.S112; c(0, 1, 3) = s.S013; c(0, 2, 2) = s.S022;
}

static const int BASIS10[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                   {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};

// Z = V^T A^{-1} V and Q of the Gram matrix; Delta2* = Q - Z.
static void gram_schur(const Cube& c, double Q[6][6], double Z[6][6]) {
  double A[16], V[24];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      A[a * 4 + b] = c(BASIS10[a][0] + BASIS10[b][0], BASIS10[a][1] + BASIS10[b][1],
                       BASIS10[a][2] + BASIS10[b][2]);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 6; ++b)
      V[a * 6 + b] = c(BASIS10[a][0] + BASIS10[4 + b][0], BASIS10[a][1] + BASIS10[4 + b][1],
                       BASIS10[a][2] + BASIS10[4 + b][2]);
  double X[24];
  std::memcpy(X, V, sizeof(X));
  solve_lu(A, 4, X, 6);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      Q[a][b] = c(BASIS10[4 + a][0] + BASIS10[4 + b][0], BASIS10[4 + a][1] + BASIS10[4 + b][1],
                  BASIS10[4 + a][2] + BASIS10[4 + b][2]);
      double z = 0.0;
      for (int k = 0; k < 4; ++k) z += V[k * 6 + a] * X[k * 6 + b];
      Z[a][b] = z;
    }
}

// delta2star3D (src/autogen/delta2star3D.jl:11-296): positional arguments as the reference.
static void delta2star3D(const S28& s, double E[6][6]) {
  Cube c;
  s28_cube(s, c);
  double Q[6][6], Z[6][6];
  gram_schur(c, Q, Z);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) E[a][b] = Q[a][b] - Z[a][b];
}
static S28 mk28(double a0, double a1, double a2, double a3, double a4, double a5, double a6,
                double a7, double a8, double a9, double a10, double a11, double a12, double a13,
                double a14, double a15, double a16, double a17, double a18, double a19, double a20,
                double a21, double a22, double a23, double a24, double a25, double a26,
                double a27) {
  return S28{a0,  a1,  a2,  a3,  a4,  a5,  a6,  a7,  a8,  a9,  a10, a11, a12, a13,
             a14, a15, a16, a17, a18, a19, a20, a21, a22, a23, a24, a25, a26, a27};
}
// The six axis permutations of delta2star3D_permutation.jl:26-61 (argument lists as there).
static S28 perm2(const S28& s) {  // ikj
  return mk28(s.S300, s.S400, s.S101, s.S201, s.S301, s.S102, s.S202, s.S003, s.S103, s.S004,
              s.S110, s.S210, s.S310, s.S120, s.S220, s.S030, s.S130, s.S040, s.S011, s.S111,
              s.S211, s.S012, s.S112, s.S013, s.S021, s.S121, s.S031, s.S022);
}
static S28 perm3(const S28& s) {  // jik
  return mk28(s.S030, s.S040, s.S110, s.S120, s.S130, s.S210, s.S220, s.S300, s.S310, s.S400,
              s.S011, s.S021, s.S031, s.S012, s.S022, s.S003, s.S013, s.S004, s.S101, s.S111,
              s.S121, s.S201, s.S211, s.S301, s.S102, s.S112, s.S103, s.S202);
}
static S28 perm4(const S28& s) {  // jki
  return mk28(s.S003, s.S004, s.S101, s.S102, s.S103, s.S201, s.S202, s.S300, s.S301, s.S400,
              s.S011, s.S012, s.S013, s.S021, s.S022, s.S030, s.S031, s.S040, s.S110, s.S111,
              s.S112, s.S210, s.S211, s.S310, s.S120, s.S121, s.S130, s.S220);
}
static S28 perm5(const S28& s) {  // kij
  return mk28(s.S030, s.S040, s.S011, s.S021, s.S031, s.S012, s.S022, s.S003, s.S013, s.S004,
              s.S110, s.S120, s.S130, s.S210, s.S220, s.S300, s.S310, s.S400, s.S101, s.S111,
              s.S121, s.S102, s.S112, s.S103, s.S201, s.S211, s.S301, s.S202);
}
static S28 perm6(const S28& s) {  // kji
  return mk28(s.S003, s.S004, s.S011, s.S012, s.S013, s.S021, s.S022, s.S030, s.S031, s.S040,
              s.S101, s.S102, s.S103, s.S201, s.S202, s.S300, s.S301, s.S400, s.S110, s.S111,
              s.S112, s.S120, s.S121, s.S130, s.S210, s.S211, s.S310, s.S220);
}

// lower_bound_S220 (src/autogen/lower_bound_S220.jl:14-37): values of S220, S202, S022
// that zero the xy, xz, yz diagonal entries of Delta2*.
static void lower_bound_S220(const S28& s, double out[3]) {
  Cube c;
  s28_cube(s, c);
  double Q[6][6], Z[6][6];
  gram_schur(c, Q, Z);
  out[0] = Z[1][1];
  out[1] = Z[2][2];
  out[2] = Z[4][4];
}
// bound_minor1 (src/autogen/bound_minor1.jl:14-60): values of the 4th-order moments that
// zero the off-diagonal Delta2* entries (basis x2,xy,xz,y2,yz,z2).
static void bound_minor1(const S28& s, double out[12]) {
  Cube c;
  s28_cube(s, c);
  double Q[6][6], Z[6][6];
  gram_schur(c, Q, Z);
  static const int P[12][2] = {{0, 1}, {0, 2}, {1, 3}, {2, 5}, {3, 4}, {4, 5},
                               {0, 3}, {0, 5}, {3, 5}, {0, 4}, {2, 3}, {1, 5}};
  for (int t = 0; t < 12; ++t) out[t] = Z[P[t][0]][P[t][1]];
}

// Polynomials (degree <= 4) with determinant by cofactor expansion ---------------------
struct Poly {
  double c[5] = {0, 0, 0, 0, 0};  // c[k] * x^k
};
static Poly pmul(const Poly& a, const Poly& b) {
  Poly r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; i + j < 5; ++j) r.c[i + j] += a.c[i] * b.c[j];
  return r;
}
static Poly padd(const Poly& a, const Poly& b, double sb) {
  Poly r;
  for (int i = 0; i < 5; ++i) r.c[i] = a.c[i] + sb * b.c[i];
  return r;
}
static Poly pdet(const Poly* A, int n) {  // row-major n x n, n <= 4
  if (n == 1) return A[0];
  Poly acc;
  for (int col = 0; col < n; ++col) {
    Poly sub[9];
    int t = 0;
    for (int r = 1; r < n; ++r)
      for (int cc = 0; cc < n; ++cc)
        if (cc != col) sub[t++] = A[r * n + cc];
    const Poly term = pmul(A[col], pdet(sub, n - 1));
    acc = padd(acc, term, (col % 2) ? -1.0 : 1.0);
  }
  return acc;
}

// Cardano roots of a x^3 + b x^2 + c x + d with a real square root of the discriminant:
// a negative discriminant (three real roots) yields NaN, as the @fastmath real sqrt does
// in rootsR.jl:275 / rootsR_X_Y.jl:60.
static void cardano(double a, double b, double c, double d, std::complex<double> R[3]) {
  const double p = (3.0 * a * c - b * b) / (3.0 * a * a);
  const double q = (2.0 * b * b * b - 9.0 * a * b * c + 27.0 * a * a * d) / (27.0 * a * a * a);
  const double disc = (q / 2.0) * (q / 2.0) + (p / 3.0) * (p / 3.0) * (p / 3.0);
  const double sq = disc < 0 ? NaN : std::sqrt(disc);
  const double u = std::cbrt(-q / 2.0 + sq);
  const double sh = -b / (3.0 * a);
  const double t = u - p / (3.0 * u);
  const double im = std::sqrt(3.0) / 2.0 * (u + p / (3.0 * u));
  R[0] = {sh + t, 0.0};
  R[1] = {sh - 0.5 * t, -im};
  R[2] = {sh - 0.5 * t, im};
}
// Julia: maximum(abs.(imag(R))) / maximum(abs.(R)) < 1000*eps()
static bool roots_real(const std::complex<double> R[3]) {
  double mi = 0, ma = 0;
  bool nan = false;
  for (int k = 0; k < 3; ++k) {
    const double ai = std::fabs(R[k].imag()), ar = std::abs(R[k]);
    if (std::isnan(ai) || std::isnan(ar)) nan = true;
    mi = std::max(mi, ai);
    ma = std::max(ma, ar);
  }
  if (nan) return false;
  return mi / ma < 1000 * EPS;
}
static void sort_real3(const std::complex<double> R[3], double Rr[3]) {
  for (int k = 0; k < 3; ++k) Rr[k] = R[k].real();
  // Julia sort: NaN last
  std::sort(Rr, Rr + 3, [](double x, double y) {
    if (std::isnan(x)) return false;
    if (std::isnan(y)) return true;
    return x < y;
  });
}

// rootsR_X_Y (src/autogen/rootsR_X_Y.jl:6-77): roots in X of
// det([[e11,e12,e13,X+ex],[e12,X,e23,e24],[e13,e23,Y,e34],[X+ex,e24,e34,e44]]).
static void rootsR_X_Y(double Y, double e11, double e12, double e13, double e23, double e24,
                       double e34, double e44, double ex, std::complex<double> R[3]) {
  Poly A[16];
  auto cst = [](double v) { Poly p; p.c[0] = v; return p; };
  Poly X;
  X.c[1] = 1.0;
  Poly Xe;
  Xe.c[0] = ex;
  Xe.c[1] = 1.0;
  A[0] = cst(e11); A[1] = cst(e12); A[2] = cst(e13); A[3] = Xe;
  A[4] = cst(e12); A[5] = X; A[6] = cst(e23); A[7] = cst(e24);
  A[8] = cst(e13); A[9] = cst(e23); A[10] = cst(Y); A[11] = cst(e34);
  A[12] = Xe; A[13] = cst(e24); A[14] = cst(e34); A[15] = cst(e44);
  const Poly d = pdet(A, 4);
  cardano(d.c[3], d.c[2], d.c[1], d.c[0], R);
}

// 2D standardized Gram Schur complement over (x2, xy, y2) given S40 and S04.
static void schur2d(double s03, double s04, double s11, double s12, double s13, double s21,
                    double s22, double s30, double s31, double s40, double E[3][3]) {
  double m[5][5] = {{0}};
  m[0][0] = 1; m[2][0] = 1; m[0][2] = 1; m[1][1] = s11; m[3][0] = s30; m[0][3] = s03;
  m[2][1] = s21; m[1][2] = s12; m[4][0] = s40; m[0][4] = s04; m[3][1] = s31; m[1][3] = s13;
  m[2][2] = s22;
  static const int B[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};
  double A[9], V[9], X[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      A[a * 3 + b] = m[B[a][0] + B[b][0]][B[a][1] + B[b][1]];
      V[a * 3 + b] = m[B[a][0] + B[3 + b][0]][B[a][1] + B[3 + b][1]];
    }
  std::memcpy(X, V, sizeof(X));
  solve_lu(A, 3, X, 3);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double z = 0;
      for (int k = 0; k < 3; ++k) z += V[k * 3 + a] * X[k * 3 + b];
      E[a][b] = m[B[3 + a][0] + B[3 + b][0]][B[3 + a][1] + B[3 + b][1]] - z;
    }
}
// delta2starchol_L3 (src/autogen/delta2starchol_L3.jl:6-37)
static double delta2starchol_L3(double s03, double s04, double s11, double s12, double s13,
                                double s21, double s22, double s30, double s31, double s40) {
  double E[3][3];
  schur2d(s03, s04, s11, s12, s13, s21, s22, s30, s31, s40, E);
  double A[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) A[a * 3 + b] = E[a][b];
  return (1.0 - s11 * s11) * det_lu(A, 3);
}
// rootsR (src/autogen/rootsR.jl:7-292): roots in S22 of the 2D Schur determinant.  The
// S22 dependence of the Schur complement is exactly E(S22) = E(0) + S22 * D with D the
// indicator of the (x2,y2), (y2,x2) and (xy,xy) entries.
static void rootsR(double Del1, double H02, double H20, double s03, double s11, double s12,
                   double s13, double s21, double s30, double s31, std::complex<double> R[3]) {
  (void)Del1;
  double E0[3][3];
  schur2d(s03, H02 + s03 * s03 + 1.0, s11, s12, s13, s21, 0.0, s30, s31, H20 + s30 * s30 + 1.0,
          E0);
  Poly A[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      A[a * 3 + b].c[0] = E0[a][b];
      const bool lin = (a == 0 && b == 2) || (a == 2 && b == 0) || (a == 1 && b == 1);
      A[a * 3 + b].c[1] = lin ? 1.0 : 0.0;
    }
  const Poly d = pdet(A, 3);
  cardano(d.c[3], d.c[2], d.c[1], d.c[0], R);
}

// realizability sub-projections ----------------------------------------------------------

// enforce_univariate (src/moments/enforce_univariate.jl:19-39)
static void enforce_univariate(double& S3, double& S4, double& H, double h2min, double s3max) {
  H = S4 - S3 * S3 - 1;
  if (H <= h2min) {
    H = h2min;
    S4 = H + S3 * S3 + 1;
  }
  if (S3 < -s3max) {
    S3 = -s3max;
    S4 = H + S3 * S3 + 1;
  } else if (S3 > s3max) {
    S3 = s3max;
    S4 = H + S3 * S3 + 1;
  }
}
// realizablity_S220 (src/realizability/realizablity_S220.jl:19-31)
static double realizablity_S220(double S110, double S220, double A220) {
  double r = S220;
  const double lo = jmax(S110 * S110, 1 - A220), hi = 1 + A220;
  if (S220 < lo) r = lo;
  else if (S220 > hi) r = hi;
  return r;
}
// realizability_S2 + find_zero_bisection (src/realizability/realizability_S2.jl:18-66)
static void realizability_S2(double& S110, double& S101, double& S011, double& S2r) {
  const double S2 = 1 + 2 * S110 * S101 * S011 - (S110 * S110 + S101 * S101 + S011 * S011);
  double xr = 1.0;
  if (S2 < 0) {
    auto Y = [&](double x) {
      return 1 + 2 * S110 * S101 * S011 * x * x * x -
             (S110 * S110 + S101 * S101 + S011 * S011) * x * x;
    };
    double a = 0.0, b = 1.0, fa = Y(a);
    const double tol = 1e-12;
    bool done = false;
    for (int it = 0; it < 100; ++it) {
      const double c = (a + b) / 2, fc = Y(c);
      if (std::fabs(fc) < tol || (b - a) / 2 < tol) { xr = c; done = true; break; }
      if (jsign(fc) == jsign(fa)) { a = c; fa = fc; }
      else { b = c; }
    }
    if (!done) xr = (a + b) / 2;
  }
  xr = 0.9999 * xr;
  S110 *= xr;
  S101 *= xr;
  S011 *= xr;
  S2r = 1 + 2 * S110 * S101 * S011 - (S110 * S110 + S101 * S101 + S011 * S011);
}
// 2x2 symmetric positive semi-definite square root via its eigen-decomposition.
static void sym2_sqrt(double a, double b, double d, double U[2][2]) {
  const double tr = a + d, df = a - d;
  const double r = std::sqrt(df * df / 4 + b * b);
  const double l1 = tr / 2 + r, l2 = tr / 2 - r;
  double c, s;
  if (b == 0.0) {
    if (a >= d) { c = 1; s = 0; }
    else { c = 0; s = 1; }
  } else {
    const double vx = l1 - d, vy = b;
    const double nn = std::hypot(vx, vy);
    c = vx / nn;
    s = vy / nn;
  }
  const double q1 = l1 < 0 ? NaN : std::sqrt(std::max(0.0, l1));
  const double q2 = l2 < -std::fabs(l1) * 2 * EPS ? NaN : std::sqrt(std::max(0.0, l2));
  U[0][0] = c * c * q1 + s * s * q2;
  U[0][1] = c * s * (q1 - q2);
  U[1][0] = U[0][1];
  U[1][1] = s * s * q1 + c * c * q2;
}
// realizability_S210 (src/realizability/realizability_S210.jl:20-44)
static void realizability_S210(double S110, double S101, double S011, double S300, double& S210,
                               double& S201, double H200, double beta) {
  double xr = 1.0;
  const double X0 = S210 - S110 * S300, X1 = S201 - S101 * S300;
  const double d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  double U[2][2];
  sym2_sqrt(d11, d12, d22, U);
  const double V0 = U[0][0] * X0 + U[0][1] * X1, V1 = U[1][0] * X0 + U[1][1] * X1;
  const double L = jmax(0.0, V0 * V0 + V1 * V1);
  const double D1[4] = {d11, d12, d12, d22};
  const double dD1 = jmax(0.0, det_lu(D1, 2));
  const double R = H200 * dD1;
  if (R <= 0 || X0 * X0 + X1 * X1 < 1000 * EPS) xr = 0.0;
  else if (L > R) xr = std::sqrt(R / L);
  double B[2] = {beta * xr * V0, beta * xr * V1};
  const double Um[4] = {U[0][0], U[0][1], U[1][0], U[1][1]};
  solve_lu(Um, 2, B, 1);
  S210 = B[0] + S110 * S300;
  S201 = B[1] + S101 * S300;
}
// realizability_S111 (src/realizability/realizability_S111.jl:18-36)
static double realizability_S111(double S110, double S101, double S011, double S210, double S201,
                                 double S120, double S021, double S102, double S012, double S111) {
  double r = S111;
  const double A110 = ((S101 - S011 * S110) * S210 + (S011 - S101 * S110) * S120) / (1 - S110 * S110);
  const double A101 = ((S110 - S011 * S101) * S201 + (S011 - S110 * S101) * S102) / (1 - S101 * S101);
  const double A011 = ((S110 - S101 * S011) * S021 + (S101 - S110 * S011) * S012) / (1 - S011 * S011);
  const double Rmin = jmin(jmin(A110, A101), A011), Rmax = jmax(jmax(A110, A101), A011);
  if (S111 > Rmax) r = Rmax;
  else if (S111 < Rmin) r = Rmin;
  return r;
}
static double quad_D2(double S110, double S101, double S011, double v0, double v1, double v2) {
  const double D[3][3] = {{1 - S011 * S011, S011 * S101 - S110, S011 * S110 - S101},
                          {S011 * S101 - S110, 1 - S101 * S101, S101 * S110 - S011},
                          {S011 * S110 - S101, S101 * S110 - S011, 1 - S110 * S110}};
  const double v[3] = {v0, v1, v2};
  double s = 0;
  for (int i = 0; i < 3; ++i) {
    double t = 0;
    for (int j = 0; j < 3; ++j) t += D[i][j] * v[j];
    s += v[i] * t;
  }
  return s;
}
// realizability_S310 (src/realizability/realizability_S310.jl:20-69); S220 is returned as is.
static double realizability_S310(double S110, double S101, double S011, double S300, double S210,
                                 double S201, double S120, double S111, double S310, double S220,
                                 double H200, double beta) {
  double r = S310;
  const double b310 =
      S111 * ((1 - S110 * S110) * S201 + (S101 * S110 - S011) * S210 + (S011 * S110 - S101) * S300) +
      S120 * ((1 - S101 * S101) * S210 + (S011 * S101 - S110) * S300 + (S101 * S110 - S011) * S201) +
      S210 * ((1 - S011 * S011) * S300 + (S011 * S101 - S110) * S210 + (S011 * S110 - S101) * S201);
  const double d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  const double D1[4] = {d11, d12, d12, d22};
  const double dD1 = det_lu(D1, 2);
  const double V0 = S210 - S110 * S300, V1 = S201 - S101 * S300;
  const double L1 = (V0 * (d11 * V0 + d12 * V1) + V1 * (d12 * V0 + d22 * V1)) / dD1;
  double G310b = H200 - L1;
  if (G310b < 0) G310b = 0.0;
  const double L2 = quad_D2(S110, S101, S011, S210, S120, S111) / dD1;
  double G310a = S220 - S110 * S110 - L2 + 1000 * EPS;
  if (G310a < 0 || L2 < 0) G310a = 0.0;
  const double sG = beta * std::sqrt(G310a * G310b);
  const double lo = S110 + b310 / dD1 - sG, hi = S110 + b310 / dD1 + sG;
  if (S310 <= lo) r = lo;
  else if (S310 >= hi) r = hi;
  return r;
}
// realizability_S310_220 (src/realizability/realizability_S310_220.jl:18-37)
static double realizability_S310_220(double S110, double S101, double S011, double S210,
                                     double S120, double S111, double S220) {
  double r = S220;
  const double d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  const double D1[4] = {d11, d12, d12, d22};
  const double dD1 = det_lu(D1, 2);
  const double L2 = quad_D2(S110, S101, S011, S210, S120, S111) / dD1;
  const double mn = S110 * S110 + L2 + 1000 * EPS;
  if (mn > r) r = mn;
  return r;
}
// realizability_S211 (src/realizability/realizability_S211.jl:19-34)
static double realizability_S211(double e11, double e22, double e33, double e12, double e13,
                                 double d23, double S211, double beta) {
  double r = S211;
  const double b = e12 * e13;
  const double G = jmax(0.0, (e11 * e22 - e13 * e13) * (e11 * e33 - e12 * e12));
  const double sG = beta * std::sqrt(G);
  const double lo = d23 + (b - sG) / e11, hi = d23 + (b + sG) / e11;
  if (S211 <= lo) r = lo;
  else if (S211 >= hi) r = hi;
  return r;
}

// realizable_2D (src/realizability/realizable_2D.jl:24-118)
static void realizable_2D(double S30, double S40, double S11, double& S21, double& S31,
                          double& S12, double& S22, double S03, double& S13, double S04,
                          uint32_t* flags) {
  const double Del1 = jmax(0.0, 1 - S11 * S11);
  const double H20 = jmax(EPS, S40 - S30 * S30 - 1);
  const double H02 = jmax(EPS, S04 - S03 * S03 - 1);
  double G1 = std::sqrt(Del1 * H02);
  double lo = S11 * S03 - G1, hi = S11 * S03 + G1;
  if (S12 <= lo) S12 = lo;
  else if (S12 >= hi) S12 = hi;
  G1 = std::sqrt(Del1 * H20);
  lo = S11 * S30 - G1;
  hi = S11 * S30 + G1;
  if (S21 <= lo) S21 = lo;
  else if (S21 >= hi) S21 = hi;
  const double G22 = std::sqrt((H20 + S30 * S30) * (H02 + S03 * S03));
  const double s22min = jmax(S11 * S11, 1 - G22), s22max = 1 + G22;
  if (S22 < s22min) S22 = s22min;
  else if (S22 > s22max) S22 = s22max;
  const double base = Del1 * S22 - Del1 * S11 * S11 - S12 * S12 + 2 * S11 * S12 * S21 - S21 * S21;
  double G31 = base * (Del1 * H20 - (S21 - S11 * S30) * (S21 - S11 * S30));
  double G13 = base * (Del1 * H02 - (S12 - S11 * S03) * (S12 - S11 * S03));
  if (G31 < 0 || G13 < 0) {
    G31 = 0.0;
    G13 = 0.0;
  } else {
    G13 = std::sqrt(G13);
    G31 = std::sqrt(G31);
  }
  const double n31 = S12 * S21 + S21 * S30 - S11 * S21 * S21 - S11 * S12 * S30;
  lo = S11 + (n31 - G31) / (Del1 + EPS);
  hi = S11 + (n31 + G31) / (Del1 + EPS);
  if (S31 <= lo) S31 = lo;
  else if (S31 >= hi) S31 = hi;
  const double n13 = S12 * S21 + S12 * S03 - S11 * S12 * S12 - S11 * S21 * S03;
  lo = S11 + (n13 - G13) / (Del1 + EPS);
  hi = S11 + (n13 + G13) / (Del1 + EPS);
  if (S13 <= lo) S13 = lo;
  else if (S13 >= hi) S13 = hi;
  const double L3 = delta2starchol_L3(S03, S04, S11, S12, S13, S21, S22, S30, S31, S40);
  if (L3 < 0) {
    std::complex<double> R[3];
    rootsR(Del1, H02, H20, S03, S11, S12, S13, S21, S30, S31, R);
    double Rr[3];
    sort_real3(R, Rr);
    double mx = jmin(s22max, Rr[2]);
    double mn = jmax(s22min, Rr[1]);
    if (mx < mn) mn = mx;
    double S22a = S22;
    if (S22 > mx) S22a = mx;
    else if (S22 < mn) S22a = mn;
    // complex roots (or NaN from the real sqrt): S22 unchanged
    double mi = 0, ma = 0;
    bool nan = false;
    for (int k = 0; k < 3; ++k) {
      const double ai = std::fabs(R[k].imag()), ar = std::abs(R[k]);
      if (std::isnan(ai) || std::isnan(ar)) nan = true;
      mi = std::max(mi, ai);
      ma = std::max(ma, ar);
    }
    const bool cplx = nan ? false : (mi / ma > 1000 * EPS);
    if (!cplx && !nan) S22 = S22a;
    // NaN ratio: "NaN > x" is false, so Julia takes the else-branch and assigns S22a,
    // which is itself NaN-free only if the bounds were; follow Julia exactly:
    if (nan) S22 = S22a;
    if (flags) *flags |= HYQ_FLAG_ROOTS2D;
  }
}

// check2D (src/utils/diagnostics.jl:17-42); returns the 10 values in reference order.
static void check2D(double S30, double S40, double S11, double S21, double S31, double S12,
                    double S22, double S03, double S13, double S04, double out[10],
                    uint32_t* flags) {
  const double h2min = 1000 * EPS;
  if (std::fabs(S11) >= 1) {
    S11 = jsign(S11);
    const double s3m = std::sqrt(std::fabs(S30 * S03));
    double s4m = std::sqrt(S40 * S04);
    const double H2m = jmax(h2min, s4m - s3m * s3m - 1);
    s4m = H2m + s3m * s3m + 1;
    S30 = jsign(S30) * s3m;
    S03 = S11 * S30;
    S40 = s4m;
    S04 = s4m;
    S12 = S11 * S03;
    S21 = S11 * S30;
    S13 = S11 * S04;
    S31 = S11 * S40;
    S22 = s4m;
  } else {
    realizable_2D(S30, S40, S11, S21, S31, S12, S22, S03, S13, S04, flags);
  }
  const double o[10] = {S30, S40, S11, S21, S31, S12, S22, S03, S13, S04};
  std::memcpy(out, o, sizeof(o));
}

// realizable_3D (src/realizability/realizable_3D.jl:39-624)
static int realizable_3D(S28& s, uint32_t* flags) {
  int flag220 = 0;
  const double diagmin = 1.0e-10, h2min = 1.0e-10, S2min = 1.0e-12;
  double &S300 = s.S300, &S400 = s.S400, &S110 = s.S110, &S210 = s.S210, &S310 = s.S310,
         &S120 = s.S120, &S220 = s.S220, &S030 = s.S030, &S130 = s.S130, &S040 = s.S040,
         &S101 = s.S101, &S201 = s.S201, &S301 = s.S301, &S102 = s.S102, &S202 = s.S202,
         &S003 = s.S003, &S103 = s.S103, &S004 = s.S004, &S011 = s.S011, &S111 = s.S111,
         &S211 = s.S211, &S021 = s.S021, &S121 = s.S121, &S031 = s.S031, &S012 = s.S012,
         &S112 = s.S112, &S013 = s.S013, &S022 = s.S022;
  const double H200 = jmax(EPS, S400 - S300 * S300 - 1);
  const double H020 = jmax(EPS, S040 - S030 * S030 - 1);
  const double H002 = jmax(EPS, S004 - S003 * S003 - 1);
  const double A220 = std::sqrt((H200 + S300 * S300) * (H020 + S030 * S030));
  const double S220max = realizablity_S220(S110, S220, A220);
  const double A202 = std::sqrt((H200 + S300 * S300) * (H002 + S003 * S003));
  const double S202max = realizablity_S220(S101, S202, A202);
  const double A022 = std::sqrt((H020 + S030 * S030) * (H002 + S003 * S003));
  const double S022max = realizablity_S220(S011, S022, A022);
  S220 = jmin(S220, S220max);
  S202 = jmin(S202, S202max);
  S022 = jmin(S022, S022max);
  double S2;
  realizability_S2(S110, S101, S011, S2);
  const double R110 = std::sqrt(jmax(0.0, 1 - S110 * S110));
  const double R101 = std::sqrt(jmax(0.0, 1 - S101 * S101));
  const double R011 = std::sqrt(jmax(0.0, 1 - S011 * S011));

  if (S2 <= S2min) {  // faces (:79-142)
    if (flags) *flags |= HYQ_FLAG_FACES;
    const double Rmax = jmax(jmax(R110, R101), R011);
    if (R110 == Rmax) {
      const double g1 = (S101 - S110 * S011) / (1 - S110 * S110);
      const double g2 = (S011 - S110 * S101) / (1 - S110 * S110);
      S021 = g2 * S030 + g1 * S120;
      S201 = g2 * S210 + g1 * S300;
      S031 = g2 * S040 + g1 * S130;
      S301 = g2 * S310 + g1 * S400;
      S111 = g2 * S120 + g1 * S210;
      S211 = g2 * S220 + g1 * S310;
      S121 = g2 * S130 + g1 * S220;
      S012 = g2 * g2 * S030 + 2 * g1 * g2 * S120 + g1 * g1 * S210;
      S102 = g2 * g2 * S120 + 2 * g1 * g2 * S210 + g1 * g1 * S300;
      S022 = g2 * g2 * S040 + 2 * g1 * g2 * S130 + g1 * g1 * S220;
      S202 = g2 * g2 * S220 + 2 * g1 * g2 * S310 + g1 * g1 * S400;
      S112 = g2 * g2 * S130 + 2 * g1 * g2 * S220 + g1 * g1 * S310;
      S003 = ipow(g2, 3) * S030 + 3 * g1 * g2 * g2 * S120 + 3 * g1 * g1 * g2 * S210 + ipow(g1, 3) * S300;
      S013 = ipow(g2, 3) * S040 + 3 * g1 * g2 * g2 * S130 + 3 * g1 * g1 * g2 * S220 + ipow(g1, 3) * S310;
      S103 = ipow(g2, 3) * S130 + 3 * g1 * g2 * g2 * S220 + 3 * g1 * g1 * g2 * S310 + ipow(g1, 3) * S400;
      S004 = ipow(g2, 4) * S040 + 4 * g1 * ipow(g2, 3) * S130 + 6 * g1 * g1 * g2 * g2 * S220 +
             4 * ipow(g1, 3) * g2 * S310 + ipow(g1, 4) * S400;
    } else if (R101 == Rmax) {
      const double g3 = (S110 - S101 * S011) / (1 - S101 * S101);
      const double g4 = (S011 - S110 * S101) / (1 - S101 * S101);
      S012 = g4 * S003 + g3 * S102;
      S210 = g4 * S201 + g3 * S300;
      S013 = g4 * S004 + g3 * S103;
      S310 = g4 * S301 + g3 * S400;
      S111 = g4 * S102 + g3 * S201;
      S211 = g4 * S202 + g3 * S301;
      S112 = g4 * S103 + g3 * S202;
      S021 = g4 * g4 * S003 + 2 * g3 * g4 * S102 + g3 * g3 * S201;
      S120 = g4 * g4 * S102 + 2 * g3 * g4 * S201 + g3 * g3 * S300;
      S022 = g4 * g4 * S004 + 2 * g3 * g4 * S103 + g3 * g3 * S202;
      S220 = g4 * g4 * S202 + 2 * g3 * g4 * S301 + g3 * g3 * S400;
      S121 = g4 * g4 * S103 + 2 * g3 * g4 * S202 + g3 * g3 * S301;
      S030 = ipow(g4, 3) * S003 + 3 * g3 * g4 * g4 * S102 + 3 * g3 * g3 * g4 * S201 + ipow(g3, 3) * S300;
      S031 = ipow(g4, 3) * S004 + 3 * g3 * g4 * g4 * S103 + 3 * g3 * g3 * g4 * S202 + ipow(g3, 3) * S301;
      S130 = ipow(g4, 3) * S103 + 3 * g3 * g4 * g4 * S202 + 3 * g3 * g3 * g4 * S301 + ipow(g3, 3) * S400;
      S040 = ipow(g4, 4) * S004 + 4 * g3 * ipow(g4, 3) * S103 + 6 * g3 * g3 * g4 * g4 * S202 +
             4 * ipow(g3, 3) * g4 * S301 + ipow(g3, 4) * S400;
    } else {
      const double g5 = (S101 - S110 * S011) / (1 - S011 * S011);
      const double g6 = (S110 - S101 * S011) / (1 - S011 * S011);
      S102 = g6 * S003 + g5 * S012;
      S120 = g6 * S021 + g5 * S030;
      S103 = g6 * S004 + g5 * S013;
      S130 = g6 * S031 + g5 * S040;
      S111 = g6 * S012 + g5 * S021;
      S121 = g6 * S022 + g5 * S031;
      S112 = g6 * S013 + g5 * S022;
      S201 = g6 * g6 * S003 + 2 * g5 * g6 * S012 + g5 * g5 * S021;
      S210 = g6 * g6 * S012 + 2 * g5 * g6 * S021 + g5 * g5 * S030;
      S202 = g6 * g6 * S004 + 2 * g5 * g6 * S013 + g5 * g5 * S022;
      S220 = g6 * g6 * S022 + 2 * g5 * g6 * S031 + g5 * g5 * S040;
      S211 = g6 * g6 * S013 + 2 * g5 * g6 * S022 + g5 * g5 * S031;
      S300 = ipow(g6, 3) * S003 + 3 * g5 * g6 * g6 * S012 + 3 * g5 * g5 * g6 * S021 + ipow(g5, 3) * S030;
      S301 = ipow(g6, 3) * S004 + 3 * g5 * g6 * g6 * S013 + 3 * g5 * g5 * g6 * S022 + ipow(g5, 3) * S031;
      S310 = ipow(g6, 3) * S013 + 3 * g5 * g6 * g6 * S022 + 3 * g5 * g5 * g6 * S031 + ipow(g5, 3) * S040;
      S400 = ipow(g6, 4) * S004 + 4 * g5 * ipow(g6, 3) * S013 + 6 * g5 * g5 * g6 * g6 * S022 +
             4 * ipow(g5, 3) * g6 * S031 + ipow(g5, 4) * S040;
    }
    return flag220;
  }

  // interior (:144-619)
  const double beta = 1.0;
  double E1[6][6];
  delta2star3D(s, E1);
  int fE11 = 0, fE22 = 0, fE33 = 0, fE44 = 0, fE55 = 0, fE66 = 0;
  if (E1[0][0] < 0) { realizability_S210(S110, S101, S011, S300, S210, S201, H200, beta); fE11 = 1; }
  if (E1[3][3] < 0) { realizability_S210(S110, S011, S101, S030, S120, S021, H020, beta); fE44 = 1; }
  if (E1[5][5] < 0) { realizability_S210(S011, S101, S110, S003, S012, S102, H002, beta); fE66 = 1; }
  {
    const double r = realizability_S111(S110, S101, S011, S210, S201, S120, S021, S102, S012, S111);
    if (r != S111) S111 = r;
  }
  const double S220_310 = realizability_S310_220(S110, S101, S011, S210, S120, S111, S220);
  const double S202_310 = realizability_S310_220(S101, S110, S011, S201, S102, S111, S202);
  const double S022_310 = realizability_S310_220(S011, S110, S101, S021, S012, S111, S022);
  double S220_diag = S220, S202_diag = S202, S022_diag = S022;
  double lb[3];
  lower_bound_S220(s, lb);
  if (S220 < lb[0]) { S220_diag = lb[0]; fE22 = 1; flag220 = 1; }
  if (S202 < lb[1]) { S202_diag = lb[1]; fE33 = 1; flag220 = 1; }
  if (S022 < lb[2]) { S022_diag = lb[2]; fE55 = 1; flag220 = 1; }
  delta2star3D(s, E1);
  if (E1[0][0] < diagmin) fE11 = 1;
  if (E1[1][1] < diagmin) fE22 = 1;
  if (E1[2][2] < diagmin) fE33 = 1;
  if (E1[3][3] < diagmin) fE44 = 1;
  if (E1[4][4] < diagmin) fE55 = 1;
  if (E1[5][5] < diagmin) fE66 = 1;
  if (fE11 || fE22 || fE33 || fE44 || fE55 || fE66) {
    double Mb[12];
    bound_minor1(s, Mb);
    const double S310b = Mb[0], S301b = Mb[1], S130b = Mb[2], S103b = Mb[3], S031b = Mb[4],
                 S013b = Mb[5], S211b = Mb[9], S121b = Mb[10], S112b = Mb[11];
    if (fE11) { S310 = S310b; S301 = S301b; S211 = S211b; }
    if (fE22) { S310 = S310b; S130 = S130b; S211 = S211b; S121 = S121b; S112 = S112b; }
    if (fE33) { S301 = S301b; S103 = S103b; S211 = S211b; S121 = S121b; S112 = S112b; }
    if (fE44) { S130 = S130b; S031 = S031b; S121 = S121b; }
    if (fE55) { S031 = S031b; S013 = S013b; S211 = S211b; S121 = S121b; S112 = S112b; }
    if (fE66) { S103 = S103b; S013 = S013b; S112 = S112b; }
  }
  if (fE11 && fE22 && fE33 && fE44 && fE55 && fE66) return flag220;

  // 2x2 minors (:284-378)
  double S220a = S220, S220b = S220, S202a = S202, S202b = S202, S022a = S022, S022b = S022;
  double P1[6][6], P2[6][6], P3[6][6], P4[6][6], P5[6][6], P6[6][6];
  delta2star3D(s, P1);
  delta2star3D(perm2(s), P2);
  delta2star3D(perm3(s), P3);
  delta2star3D(perm4(s), P4);
  delta2star3D(perm5(s), P5);
  delta2star3D(perm6(s), P6);
  int flag310 = 0;
  if (!fE11 && !fE44 && det_lead(P1, 2) < h2min) {
    S310 = realizability_S310(S110, S101, S011, S300, S210, S120, S201, S111, S310, S220, H200, beta);
    S220a = S220;
    flag310 = 1;
  }
  if (!fE22 && !fE44 && det_lead(P3, 2) < h2min) {
    S130 = realizability_S310(S110, S011, S101, S030, S120, S021, S210, S111, S130, S220, H020, beta);
    S220b = S220;
    flag310 = 1;
  }
  if (!fE11 && !fE33 && det_lead(P2, 2) < h2min) {
    S301 = realizability_S310(S101, S110, S011, S300, S201, S210, S102, S111, S301, S202, H200, beta);
    S202a = S202;
    flag310 = 1;
  }
  if (!fE33 && !fE66 && det_lead(P4, 2) < h2min) {
    S103 = realizability_S310(S101, S011, S110, S003, S102, S012, S201, S111, S103, S202, H002, beta);
    S202b = S202;
    flag310 = 1;
  }
  if (!fE44 && !fE55 && det_lead(P5, 2) < h2min) {
    S031 = realizability_S310(S011, S110, S101, S030, S021, S120, S012, S111, S031, S022, H020, beta);
    S022a = S022;
    flag310 = 1;
  }
  if (!fE55 && !fE66 && det_lead(P6, 2) < h2min) {
    S013 = realizability_S310(S011, S101, S110, S003, S012, S102, S021, S111, S013, S022, H002, beta);
    S022b = S022;
    flag310 = 1;
  }
  double S220r = jmax(S220a, S220b), S220_310r = S220_310;
  if (S220 < S220r) { S220_310r = S220r; flag220 = 1; }
  double S022r = jmax(S022a, S022b), S022_310r = S022_310;
  if (S022 < S022r) { S022_310r = S022r; flag220 = 1; }
  double S202r = jmax(S202a, S202b), S202_310r = S202_310;
  if (S202 < S202r) { S202_310r = S202r; flag220 = 1; }
  if (flag310) {
    delta2star3D(s, P1);
    delta2star3D(perm2(s), P2);
    delta2star3D(perm3(s), P3);
    delta2star3D(perm4(s), P4);
    delta2star3D(perm5(s), P5);
    delta2star3D(perm6(s), P6);
    double m = det_lead(P1, 2);
    m = jmin(m, det_lead(P2, 2));
    m = jmin(m, det_lead(P3, 2));
    m = jmin(m, det_lead(P4, 2));
    m = jmin(m, det_lead(P5, 2));
    m = jmin(m, det_lead(P6, 2));
    if (m < 0) flag220 = 1;
  }

  // 3x3 minors (:380-453)
  S220a = S220; S220b = S220; S202a = S202; double S202c = S202; S022b = S022; double S022c = S022;
  if (!fE11 && !fE22 && !fE33 && !fE55) {
    if (det_lead(P1, 3) <= h2min || det_lead(P2, 3) <= h2min) {
      const double e11 = P1[0][0], e12 = P1[0][1], e13 = P1[0][2], e22 = P1[1][1], e23 = P1[1][2],
                   e33 = P1[2][2];
      const double d23 = S211 - e23, d22 = S220 - e22, d33 = S202 - e33;
      if (e11 > diagmin) { S220a = d22 + e13 * e13 / e11; S202a = d33 + e12 * e12 / e11; }
      S211 = realizability_S211(e11, e22, e33, e12, e13, d23, S211, beta);
      flag220 = 1;
    }
  }
  if (!fE22 && !fE33 && !fE44 && !fE55) {
    if (det_lead(P3, 3) < h2min || det_lead(P5, 3) < h2min) {
      const double e11 = P3[0][0], e12 = P3[0][1], e13 = P3[0][2], e22 = P3[1][1], e23 = P3[1][2],
                   e33 = P3[2][2];
      const double d23 = S121 - e23, d22 = S220 - e22, d33 = S022 - e33;
      if (e11 > diagmin) { S220b = d22 + e13 * e13 / e11; S022b = d33 + e12 * e12 / e11; }
      S121 = realizability_S211(e11, e22, e33, e12, e13, d23, S121, beta);
      flag220 = 1;
    }
  }
  if (!fE22 && !fE33 && !fE55 && !fE66) {
    if (det_lead(P4, 3) < h2min || det_lead(P6, 3) < h2min) {
      const double e11 = P6[0][0], e12 = P6[0][1], e13 = P6[0][2], e22 = P6[1][1], e23 = P6[1][2],
                   e33 = P6[2][2];
      const double d23 = S112 - e23, d22 = S022 - e22, d33 = S202 - e33;
      if (e11 > diagmin) { S202c = d22 + e13 * e13 / e11; S022c = d33 + e12 * e12 / e11; }
      S112 = realizability_S211(e11, e22, e33, e12, e13, d23, S112, beta);
      flag220 = 1;
    }
  }
  const double S220_211 = jmax(S220a, S220b);
  const double S202_211 = jmax(S202a, S202c);
  const double S022_211 = jmax(S022b, S022c);

  // 4x4 minors (:455-608)
  auto minor4 = [&](const double E[6][6], double Sxx, double& smin, double& smax) {
    smin = Sxx;
    smax = Sxx;
    if (det_lead(E, 4) < 0) {
      flag220 = 1;
      const double ex = -E[1][1] + E[0][3], d22 = Sxx - E[1][1], Y = E[2][2];
      std::complex<double> R[3];
      rootsR_X_Y(Y, E[0][0], E[0][1], E[0][2], E[1][2], E[1][3], E[2][3], E[3][3], ex, R);
      double Rr[3];
      sort_real3(R, Rr);
      if (roots_real(R)) { smin = Rr[1] + d22; smax = Rr[2] + d22; }
      if (flags) *flags |= HYQ_FLAG_ROOTS4;
    }
  };
  double s220mina, s220maxa, s220minb, s220maxb, s202mina, s202maxa, s202minb, s202maxb;
  double s022mina, s022maxa, s022minb, s022maxb;
  delta2star3D(s, P1);
  delta2star3D(perm3(s), P3);
  minor4(P1, S220, s220mina, s220maxa);
  minor4(P3, S220, s220minb, s220maxb);
  delta2star3D(perm2(s), P2);
  delta2star3D(perm4(s), P4);
  minor4(P2, S202, s202mina, s202maxa);
  minor4(P4, S202, s202minb, s202maxb);
  delta2star3D(perm5(s), P5);
  delta2star3D(perm6(s), P6);
  minor4(P5, S022, s022mina, s022maxa);
  minor4(P6, S022, s022minb, s022maxb);

  auto max6 = [](double a, double b, double c, double d, double e, double f) {
    return jmax(jmax(jmax(jmax(jmax(a, b), c), d), e), f);
  };
  auto min4 = [](double a, double b, double c, double d) { return jmin(jmin(jmin(a, b), c), d); };
  S220 = max6(s220mina, s220minb, S220_diag, S220_310, S220_310r, S220_211);
  S220 = min4(S220, S220max, s220maxa, s220maxb);
  S202 = max6(s202mina, s202minb, S202_diag, S202_310, S202_310r, S202_211);
  S202 = min4(S202, S202max, s202maxa, s202maxb);
  S022 = max6(s022mina, s022minb, S022_diag, S022_310, S022_310r, S022_211);
  S022 = min4(S022, S022max, s022maxa, s022maxb);
  return flag220;
}

// edge_corner_correction (src/realizability/edge_corner_correction.jl:24-253).  The
// inputs carry the reference's "r" suffixes: 1/2/3 = outputs of check2D on the XY, XZ,
// YZ planes, bare r = values entering check2D.
struct EdgeIn {
  double R110, R101, R011, S110r, S101r, S011r;
  double S300r1, S030r1, S400r1, S040r1, S300r2, S003r2, S400r2, S004r2;
  double S030r3, S003r3, S040r3, S004r3;
  double S300r, S030r, S003r, S400r, S040r, S004r;
  double S210r, S201r, S120r, S111r, S102r, S021r, S012r;
  double S310r, S301r, S220r, S211r, S202r, S130r, S121r, S112r, S103r, S031r, S022r, S013r;
};
static void edge_corner_correction(const EdgeIn& q, S28& o) {
  double S110, S101, S011, S300, S030, S003, S400, S040, S004, dummy;
  double S210, S201, S120, S111, S102, S021, S012, S310, S301, S220, S211, S202, S130, S121,
      S112, S103, S031, S022, S013;
  if (q.R110 <= 0 && q.R101 > 0 && q.R011 > 0) {
    S110 = jsign(q.S110r);
    const double Sm = (q.S011r + q.S101r) / 2;
    S011 = jsign(q.S011r) * Sm;
    S101 = jsign(q.S101r) * Sm;
    realizability_S2(S110, S101, S011, dummy);
    S300 = q.S300r1; S030 = q.S030r1; S400 = q.S400r1; S040 = q.S040r1;
    S003 = q.S003r; S004 = q.S004r;
    // average_edge_variants_xy (:113-140); argument S300,S400 there are S300r,S400r
    const double a300 = q.S300r, a400 = q.S400r;
    S210 = (S030 + S110 * a300) / 2;
    S201 = (q.S021r + q.S201r) / 2;
    S120 = (S110 * S030 + a300) / 2;
    S111 = (S110 * q.S021r + S110 * q.S201r) / 2;
    S102 = (S110 * q.S012r + q.S102r) / 2;
    S021 = (q.S021r + q.S201r) / 2;
    S012 = (q.S012r + S110 * q.S102r) / 2;
    S310 = (S110 * S040 + S110 * a400) / 2;
    S301 = (S110 * q.S031r + q.S301r) / 2;
    S220 = (S040 + a400) / 2;
    S211 = (q.S031r + S110 * q.S301r) / 2;
    S202 = (q.S022r + q.S202r) / 2;
    S130 = (S110 * S040 + S110 * a400) / 2;
    S121 = (S110 * q.S031r + q.S301r) / 2;
    S112 = (S110 * q.S022r + S110 * q.S202r) / 2;
    S103 = (S110 * q.S013r + q.S103r) / 2;
    S031 = (q.S031r + S110 * q.S301r) / 2;
    S022 = (q.S022r + q.S202r) / 2;
    S013 = (q.S013r + S110 * q.S103r) / 2;
  } else if (q.R101 <= 0 && q.R110 > 0 && q.R011 > 0) {
    S101 = jsign(q.S101r);
    const double Sm = (q.S011r + q.S110r) / 2;
    S011 = jsign(q.S011r) * Sm;
    S110 = jsign(q.S110r) * Sm;
    realizability_S2(S110, S101, S011, dummy);
    S300 = q.S300r2; S003 = q.S003r2; S400 = q.S400r2; S004 = q.S004r2;
    S030 = q.S030r; S040 = q.S040r;
    // average_edge_variants_xz (:143-170)
    const double a300 = q.S300r, a400 = q.S400r;
    S210 = (q.S013r + q.S210r) / 2;
    S201 = (S003 + S101 * a300) / 2;
    S120 = (S101 * q.S021r + q.S120r) / 2;
    S111 = (S101 * q.S012r + S101 * q.S210r) / 2;
    S102 = (S101 * S003 + a300) / 2;
    S021 = (q.S021r + S101 * q.S120r) / 2;
    S012 = (q.S012r + q.S210r) / 2;
    S310 = (S101 * q.S013r + q.S310r) / 2;
    S301 = (S101 * S004 + S101 * a400) / 2;
    S220 = (q.S022r + q.S220r) / 2;
    S211 = (q.S013r + S101 * q.S310r) / 2;
    S202 = (S004 + a400) / 2;
    S130 = (S101 * q.S031r + q.S130r) / 2;
    S121 = (S101 * q.S022r + S101 * q.S220r) / 2;
    S112 = (S101 * q.S013r + q.S310r) / 2;
    S103 = (S101 * S004 + S101 * a400) / 2;
    S031 = (q.S031r + S101 * q.S130r) / 2;
    S022 = (q.S022r + q.S220r) / 2;
    S013 = (q.S013r + S101 * q.S310r) / 2;
  } else if (q.R011 <= 0 && q.R101 > 0 && q.R110 > 0) {
    S011 = jsign(q.S011r);
    const double Sm = (q.S101r + q.S110r) / 2;
    S101 = jsign(q.S101r) * Sm;
    S110 = jsign(q.S110r) * Sm;
    realizability_S2(S110, S101, S011, dummy);
    S030 = q.S030r3; S003 = q.S003r3; S040 = q.S040r3; S004 = q.S004r3;
    S300 = q.S300r; S400 = q.S400r;
    // average_edge_variants_yz (:173-200); S030,S040 there are S030r,S040r
    const double a030 = q.S030r, a040 = q.S040r;
    S210 = (S011 * q.S201r + q.S210r) / 2;
    S201 = (q.S201r + S011 * q.S210r) / 2;
    S120 = (q.S102r + q.S120r) / 2;
    S111 = (S011 * q.S102r + S011 * q.S120r) / 2;
    S102 = (q.S102r + q.S120r) / 2;
    S021 = (S003 + S011 * a030) / 2;
    S012 = (S011 * S003 + a030) / 2;
    S310 = (S011 * q.S301r + q.S310r) / 2;
    S301 = (q.S301r + S011 * q.S310r) / 2;
    S220 = (q.S202r + q.S220r) / 2;
    S211 = (S011 * q.S202r + S011 * q.S220r) / 2;
    S202 = (q.S202r + q.S220r) / 2;
    S130 = (S011 * q.S103r + q.S130r) / 2;
    S121 = (q.S103r + S011 * q.S130r) / 2;
    S112 = (S011 * q.S103r + q.S130r) / 2;
    S103 = (q.S103r + S011 * q.S130r) / 2;
    S031 = (S011 * S004 + S011 * a040) / 2;
    S022 = (S004 + a040) / 2;
    S013 = (S011 * S004 + S011 * a040) / 2;
  } else {
    S110 = jsign(q.S110r);
    S101 = jsign(q.S101r);
    S011 = jsign(q.S011r);
    realizability_S2(S110, S101, S011, dummy);
    S300 = (q.S030r1 + q.S030r + q.S300r) / 3;
    S030 = (q.S030r1 + q.S030r + q.S030r3) / 3;
    S003 = (q.S003r + q.S003r2 + q.S003r3) / 3;
    S400 = (q.S040r1 + q.S040r + q.S400r) / 3;
    S040 = (q.S040r1 + q.S040r + q.S040r3) / 3;
    S004 = (q.S004r + q.S004r2 + q.S004r3) / 3;
    // average_corner_variants (:203-253) with (S030a,S040a,S030b,S040b,S300c,S400c,S003c,S004c)
    // = (S030r1,S040r1,S030r,S040r,S300r,S400r,S003r3,S004r3)
    const double A3 = q.S030r1, A4 = q.S040r1, B3 = q.S030r, B4 = q.S040r, C3 = q.S003r3,
                 C4 = q.S004r3;
    const double x = S110, y = S101, z = S011;
    S210 = (x * A3 + B3 + z * C3) / 3;
    S201 = (y * A3 + z * B3 + C3) / 3;
    S120 = (A3 + x * B3 + y * C3) / 3;
    S111 = (x * y * A3 + x * z * B3 + y * z * C3) / 3;
    S102 = (A3 + x * B3 + y * C3) / 3;
    S021 = (y * A3 + z * B3 + C3) / 3;
    S012 = (x * A3 + B3 + z * C3) / 3;
    S310 = (x * A4 + x * B4 + y * z * C4) / 3;
    S301 = (y * A4 + x * z * B4 + y * C4) / 3;
    S220 = (A4 + B4 + C4) / 3;
    S211 = (x * y * A4 + z * B4 + z * C4) / 3;
    S202 = (A4 + B4 + C4) / 3;
    S130 = (x * A4 + x * B4 + y * z * C4) / 3;
    S121 = (y * A4 + x * z * B4 + y * C4) / 3;
    S112 = (x * A4 + x * B4 + y * z * C4) / 3;
    S103 = (y * A4 + x * z * B4 + y * C4) / 3;
    S031 = (x * y * A4 + z * B4 + z * C4) / 3;
    S022 = (A4 + B4 + C4) / 3;
    S013 = (x * y * A4 + z * B4 + z * C4) / 3;
  }
  o.S110 = S110; o.S101 = S101; o.S011 = S011; o.S300 = S300; o.S030 = S030; o.S003 = S003;
  o.S400 = S400; o.S040 = S040; o.S004 = S004; o.S210 = S210; o.S201 = S201; o.S120 = S120;
  o.S111 = S111; o.S102 = S102; o.S021 = S021; o.S012 = S012; o.S310 = S310; o.S301 = S301;
  o.S220 = S220; o.S211 = S211; o.S202 = S202; o.S130 = S130; o.S121 = S121; o.S112 = S112;
  o.S103 = S103; o.S031 = S031; o.S022 = S022; o.S013 = S013;
}

// hyqmom_3D (src/moments/hyqmom_3D.jl:10-53).  The closure has one formula per family,
// instantiated along each axis (pair): a 5th-order univariate family, a (4,1), a (3,2),
// a (3,1,1) and a (2,2,1) family.
static double f500(double s3, double s4) { return 0.5 * s3 * (5 * s4 - 3 * s3 * s3 - 1); }
static double f410(double a3, double a4, double ab11, double ab21, double ab31) {
  return (a3 - 2 * a3 * a4 + 9 * a3 * a3 * a3 / 4) * ab11 + (5 * a4 / 2 - 15 * a3 * a3 / 4 - 1.5) * ab21 +
         2 * a3 * ab31;
}
static double f320(double a3, double a4, double b3, double ab11, double ab21, double ab12,
                   double ab22, double ab31) {
  return (b3 * a3 * a3 / 2 - b3 + b3 * (a3 * a3 - a4 + 1)) * ab11 + (a4 - 3 * a3 * a3 / 2) * ab12 +
         (-(3 * b3 * a3) / 2) * ab21 + (3 * a3 / 2) * ab22 + b3 * ab31 - a3 / 2;
}
static double f311(double a3, double a4, double a211, double a111, double bc11) {
  return 3 * (a3 * a211) / 2 + (2 * a4 - 3 * a3 * a3) * a111 / 2 - (a3 * bc11) / 2;
}
static double f221(double a3, double b3, double b121, double a211, double ac11, double bc11,
                   double a201, double b021, double a111) {
  return a3 * (b121 - ac11) + b3 * (a211 - bc11) + a201 + b021 - a3 * b3 * a111;
}
struct S21 {
  double S500, S410, S320, S230, S140, S401, S302, S203, S104, S311, S221, S131, S212, S113,
      S122, S050, S041, S032, S023, S014, S005;
};
static S21 hyqmom_3D(const S28& s) {
  S21 o;
  o.S500 = f500(s.S300, s.S400);
  o.S050 = f500(s.S030, s.S040);
  o.S005 = f500(s.S003, s.S004);
  o.S320 = f320(s.S300, s.S400, s.S030, s.S110, s.S210, s.S120, s.S220, s.S310);
  o.S230 = f320(s.S030, s.S040, s.S300, s.S110, s.S120, s.S210, s.S220, s.S130);
  o.S302 = f320(s.S300, s.S400, s.S003, s.S101, s.S201, s.S102, s.S202, s.S301);
  o.S203 = f320(s.S003, s.S004, s.S300, s.S101, s.S102, s.S201, s.S202, s.S103);
  o.S032 = f320(s.S030, s.S040, s.S003, s.S011, s.S021, s.S012, s.S022, s.S031);
  o.S023 = f320(s.S003, s.S004, s.S030, s.S011, s.S012, s.S021, s.S022, s.S013);
  o.S410 = f410(s.S300, s.S400, s.S110, s.S210, s.S310);
  o.S140 = f410(s.S030, s.S040, s.S110, s.S120, s.S130);
  o.S401 = f410(s.S300, s.S400, s.S101, s.S201, s.S301);
  o.S104 = f410(s.S003, s.S004, s.S101, s.S102, s.S103);
  o.S041 = f410(s.S030, s.S040, s.S011, s.S021, s.S031);
  o.S014 = f410(s.S003, s.S004, s.S011, s.S012, s.S013);
  o.S311 = f311(s.S300, s.S400, s.S211, s.S111, s.S011);
  o.S131 = f311(s.S030, s.S040, s.S121, s.S111, s.S101);
  o.S113 = f311(s.S003, s.S004, s.S112, s.S111, s.S110);
  o.S221 = f221(s.S300, s.S030, s.S121, s.S211, s.S101, s.S011, s.S201, s.S021, s.S111);
  o.S212 = f221(s.S300, s.S003, s.S112, s.S211, s.S110, s.S011, s.S210, s.S012, s.S111);
  o.S122 = f221(s.S003, s.S030, s.S121, s.S112, s.S101, s.S110, s.S102, s.S120, s.S111);
  return o;
}

// Flux_closure35_and_realizable_3D (src/numerics/Flux_closure35_and_realizable_3D.jl:33-234)
static void flux_closure35(const double* M4, int flag2D, double Ma, double* Fx, double* Fy,
                           double* Fz, double* M4r, uint32_t* flags) {
  const double s3max = 4.0 + std::fabs(Ma) / 2.0;
  const double h2min = 1.0e-8;
  const int itrealmax = 6;
  const double M000 = M4[0];
  const double umean = M4[1] / M000, vmean = M4[5] / M000, wmean = M4[15] / M000;
  double C4[35], S4[35];
  m2cs4(M4, C4, S4);
  const double C200 = jmax(EPS, C4[2]), C020 = jmax(EPS, C4[9]), C002 = jmax(EPS, C4[19]);
  S28 s{S4[3],  S4[4],  S4[6],  S4[7],  S4[8],  S4[10], S4[11], S4[12], S4[13], S4[14],
        S4[16], S4[17], S4[18], S4[20], S4[21], S4[22], S4[23], S4[24], S4[25], S4[26],
        S4[27], S4[28], S4[29], S4[30], S4[31], S4[32], S4[33], S4[34]};
  double H200, H020, H002;
  const double o300 = s.S300, o400 = s.S400, o030 = s.S030, o040 = s.S040, o003 = s.S003,
               o004 = s.S004;
  enforce_univariate(s.S300, s.S400, H200, h2min, s3max);
  enforce_univariate(s.S030, s.S040, H020, h2min, s3max);
  enforce_univariate(s.S003, s.S004, H002, h2min, s3max);
  if (flags && (o300 != s.S300 || o400 != s.S400 || o030 != s.S030 || o040 != s.S040 ||
                o003 != s.S003 || o004 != s.S004))
    *flags |= HYQ_FLAG_UNIVARIATE;
  const double A220 = std::sqrt((H200 + s.S300 * s.S300) * (H020 + s.S030 * s.S030));
  s.S220 = realizablity_S220(s.S110, s.S220, A220);
  const double A202 = std::sqrt((H200 + s.S300 * s.S300) * (H002 + s.S003 * s.S003));
  s.S202 = realizablity_S220(s.S101, s.S202, A202);
  const double A022 = std::sqrt((H020 + s.S030 * s.S030) * (H002 + s.S003 * s.S003));
  s.S022 = realizablity_S220(s.S011, s.S022, A022);

  // check2D_all_planes (:83-88) always runs; its outputs are read only on the edge/corner
  // branch, so its branch flags are recorded only there.
  const double R110 = 1 - s.S110 * s.S110, R101 = 1 - s.S101 * s.S101, R011 = 1 - s.S011 * s.S011;
  uint32_t* f2d = (R110 <= 0 || R101 <= 0 || R011 <= 0) ? flags : nullptr;
  double p1[10], p2[10], p3[10];
  check2D(s.S300, s.S400, s.S110, s.S210, s.S310, s.S120, s.S220, s.S030, s.S130, s.S040, p1,
          f2d);
  check2D(s.S300, s.S400, s.S101, s.S201, s.S301, s.S102, s.S202, s.S003, s.S103, s.S004, p2,
          f2d);
  check2D(s.S030, s.S040, s.S011, s.S021, s.S031, s.S012, s.S022, s.S003, s.S013, s.S004, p3,
          f2d);
  if (R110 <= 0 || R101 <= 0 || R011 <= 0) {
    if (flags) *flags |= HYQ_FLAG_EDGE_CORNER;
    EdgeIn q;
    q.R110 = R110; q.R101 = R101; q.R011 = R011;
    q.S110r = p1[2]; q.S101r = p2[2]; q.S011r = p3[2];
    q.S300r1 = p1[0]; q.S030r1 = p1[7]; q.S400r1 = p1[1]; q.S040r1 = p1[9];
    q.S300r2 = p2[0]; q.S003r2 = p2[7]; q.S400r2 = p2[1]; q.S004r2 = p2[9];
    q.S030r3 = p3[0]; q.S003r3 = p3[7]; q.S040r3 = p3[1]; q.S004r3 = p3[9];
    q.S300r = s.S300; q.S030r = s.S030; q.S003r = s.S003; q.S400r = s.S400; q.S040r = s.S040;
    q.S004r = s.S004;
    q.S210r = p1[3]; q.S201r = p2[3]; q.S120r = p1[5]; q.S111r = s.S111; q.S102r = p2[5];
    q.S021r = p3[3]; q.S012r = p3[5];
    q.S310r = p1[4]; q.S301r = p2[4]; q.S220r = p1[6]; q.S211r = s.S211; q.S202r = p2[6];
    q.S130r = p1[8]; q.S121r = s.S121; q.S112r = s.S112; q.S103r = p2[8]; q.S031r = p3[4];
    q.S022r = p3[6]; q.S013r = p3[8];
    S28 o = s;
    edge_corner_correction(q, o);
    s = o;
    // :107-109.  realizable_2D returns (S21, S12, S31, S22, S13); the xz and yz calls
    // destructure that into (S201, S301, S102, S202, S103) and (S021, S031, S012, S022,
    // S013), so the corrected S102 lands in S301 and vice versa (likewise S012 / S031).
    realizable_2D(s.S300, s.S400, s.S110, s.S210, s.S310, s.S120, s.S220, s.S030, s.S130, s.S040,
                  flags);
    realizable_2D(s.S300, s.S400, s.S101, s.S201, s.S301, s.S102, s.S202, s.S003, s.S103, s.S004,
                  flags);
    std::swap(s.S301, s.S102);
    realizable_2D(s.S030, s.S040, s.S011, s.S021, s.S031, s.S012, s.S022, s.S003, s.S013, s.S004,
                  flags);
    std::swap(s.S031, s.S012);
  } else {
    int flag220 = realizable_3D(s, flags);
    int it = 0;
    if (flag220 == 1 && flags) *flags |= HYQ_FLAG_FLAG220;
    while (flag220 == 1 && it < itrealmax) {
      ++it;
      realizable_3D(s, flags);
    }
  }
  if (flag2D == 1) {
    s.S011 = 0.0; s.S101 = 0.0; s.S201 = 0.0; s.S021 = 0.0; s.S012 = 0.0; s.S102 = 0.0;
    s.S111 = 0.0; s.S202 = 1.0; s.S022 = 1.0; s.S112 = s.S110; s.S121 = 0.0; s.S211 = 0.0;
    s.S103 = 0.0; s.S013 = 0.0; s.S301 = 0.0; s.S031 = 0.0;
  }
  const S21 f = hyqmom_3D(s);
  const double sx = std::sqrt(jmax(EPS, C200)), sy = std::sqrt(jmax(EPS, C020)),
               sz = std::sqrt(jmax(EPS, C002));
  Cube Sc;
  s28_cube(s, Sc);
  Sc(5, 0, 0) = f.S500; Sc(4, 1, 0) = f.S410; Sc(3, 2, 0) = f.S320; Sc(2, 3, 0) = f.S230;
  Sc(1, 4, 0) = f.S140; Sc(4, 0, 1) = f.S401; Sc(3, 0, 2) = f.S302; Sc(2, 0, 3) = f.S203;
  Sc(1, 0, 4) = f.S104; Sc(3, 1, 1) = f.S311; Sc(2, 2, 1) = f.S221; Sc(1, 3, 1) = f.S131;
  Sc(2, 1, 2) = f.S212; Sc(1, 1, 3) = f.S113; Sc(1, 2, 2) = f.S122; Sc(0, 5, 0) = f.S050;
  Sc(0, 4, 1) = f.S041; Sc(0, 3, 2) = f.S032; Sc(0, 2, 3) = f.S023; Sc(0, 1, 4) = f.S014;
  Sc(0, 0, 5) = f.S005;
  // S_to_C_batch (src/utils/moment_array_utils.jl:137-204)
  Cube Cc;
  std::memset(&Cc, 0, sizeof(Cc));
  for (int i = 0; i <= 5; ++i)
    for (int j = 0; i + j <= 5; ++j)
      for (int k = 0; i + j + k <= 5; ++k)
        Cc(i, j, k) = Sc(i, j, k) * ipow(sx, i) * ipow(sy, j) * ipow(sz, k);
  Cc(2, 0, 0) = C200;
  Cc(0, 2, 0) = C020;
  Cc(0, 0, 2) = C002;
  Cube M5;
  central_to_raw(M000, umean, vmean, wmean, Cc, 5, M5);
  for (int n = 0; n < 35; ++n) {
    const int i = E35[n][0], j = E35[n][1], k = E35[n][2];
    if (Fx) Fx[n] = M5(i + 1, j, k);
    if (Fy) Fy[n] = M5(i, j + 1, k);
    if (Fz) Fz[n] = M5(i, j, k + 1);
    if (M4r) M4r[n] = M5(i, j, k);
  }
}

// Moments5_3D (src/moments/Moments5_3D.jl:16-138): the HyQMOM closure of the 5th-order
// moments without any realizability projection.  Outputs in the reference's order
// 500 410 320 230 140 401 302 203 104 311 221 131 212 113 122 050 041 032 023 014 005.
static const int E21[21][3] = {{5, 0, 0}, {4, 1, 0}, {3, 2, 0}, {2, 3, 0}, {1, 4, 0}, {4, 0, 1},
                               {3, 0, 2}, {2, 0, 3}, {1, 0, 4}, {3, 1, 1}, {2, 2, 1}, {1, 3, 1},
                               {2, 1, 2}, {1, 1, 3}, {1, 2, 2}, {0, 5, 0}, {0, 4, 1}, {0, 3, 2},
                               {0, 2, 3}, {0, 1, 4}, {0, 0, 5}};
static void moments5_3d(const double* M4, double* M5o, double* C5o, double* S5o) {
  double C4[35], S4[35];
  m2cs4(M4, C4, S4);
  const S28 s{S4[3],  S4[4],  S4[6],  S4[7],  S4[8],  S4[10], S4[11], S4[12], S4[13], S4[14],
              S4[16], S4[17], S4[18], S4[20], S4[21], S4[22], S4[23], S4[24], S4[25], S4[26],
              S4[27], S4[28], S4[29], S4[30], S4[31], S4[32], S4[33], S4[34]};
  const S21 f = hyqmom_3D(s);
  const double S5[21] = {f.S500, f.S410, f.S320, f.S230, f.S140, f.S401, f.S302,
                         f.S203, f.S104, f.S311, f.S221, f.S131, f.S212, f.S113,
                         f.S122, f.S050, f.S041, f.S032, f.S023, f.S014, f.S005};
  const double sx = std::sqrt(jmax(C4[2], EPS)), sy = std::sqrt(jmax(C4[9], EPS)),
               sz = std::sqrt(jmax(C4[19], EPS));
  Cube Cc;
  std::memset(&Cc, 0, sizeof(Cc));
  for (int n = 0; n < 35; ++n) Cc(E35[n][0], E35[n][1], E35[n][2]) = C4[n];
  for (int n = 0; n < 21; ++n) {
    const int i = E21[n][0], j = E21[n][1], k = E21[n][2];
    const double c = S5[n] * ipow(sx, i) * ipow(sy, j) * ipow(sz, k);
    Cc(i, j, k) = c;
    S5o[n] = S5[n];
    C5o[n] = c;
  }
  Cube M5;
  central_to_raw(M4[0], M4[1] / M4[0], M4[5] / M4[0], M4[15] / M4[0], Cc, 5, M5);
  for (int n = 0; n < 21; ++n) M5o[n] = M5(E21[n][0], E21[n][1], E21[n][2]);
}

// Eigenvalues of a real n x n matrix (n <= 6): balancing, Householder Hessenberg
// reduction, Francis double-shift QR with standardized 2x2 blocks (Golub & Van Loan,
// Matrix Computations, Algs. 7.4.2, 7.5.1-7.5.2; the LAPACK dgeev family).
static void balance(double* A, int n) {
  const double radix = 2.0, sqrdx = 4.0;
  bool done = false;
  while (!done) {
    done = true;
    for (int i = 0; i < n; ++i) {
      double r = 0, c = 0;
      for (int j = 0; j < n; ++j)
        if (j != i) {
          c += std::fabs(A[j * n + i]);
          r += std::fabs(A[i * n + j]);
        }
      if (c != 0.0 && r != 0.0) {
        double g = r / radix, f = 1.0;
        const double s = c + r;
        while (c < g) { f *= radix; c *= sqrdx; }
        g = r * radix;
        while (c > g) { f /= radix; c /= sqrdx; }
        if ((c + r) / f < 0.95 * s) {
          done = false;
          g = 1.0 / f;
          for (int j = 0; j < n; ++j) A[i * n + j] *= g;
          for (int j = 0; j < n; ++j) A[j * n + i] *= f;
        }
      }
    }
  }
}
static void hessenberg(double* A, int n) {
  for (int k = 0; k < n - 2; ++k) {
    double alpha = 0;
    for (int i = k + 1; i < n; ++i) alpha += A[i * n + k] * A[i * n + k];
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (A[(k + 1) * n + k] > 0) alpha = -alpha;
    double v[6] = {0};
    for (int i = k + 1; i < n; ++i) v[i] = A[i * n + k];
    v[k + 1] -= alpha;
    double vn = 0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    for (int j = 0; j < n; ++j) {  // A = (I - 2vv^T/vn) A
      double d = 0;
      for (int i = k + 1; i < n; ++i) d += v[i] * A[i * n + j];
      d = 2 * d / vn;
      for (int i = k + 1; i < n; ++i) A[i * n + j] -= d * v[i];
    }
    for (int i = 0; i < n; ++i) {  // A = A (I - 2vv^T/vn)
      double d = 0;
      for (int j = k + 1; j < n; ++j) d += A[i * n + j] * v[j];
      d = 2 * d / vn;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= d * v[j];
    }
    for (int i = k + 2; i < n; ++i) A[i * n + k] = 0.0;
  }
}
// Eigenvalues of the 2x2 block [[a,b],[c,d]]: a real pair (imaginary parts exactly 0) or
// a complex-conjugate pair.
static void eig2(double a, double b, double c, double d, double& r1, double& i1, double& r2,
                 double& i2) {
  const double p = 0.5 * (a - d);
  const double disc = p * p + b * c;
  if (disc >= 0) {
    const double sq = std::sqrt(disc);
    const double z = p + (p >= 0 ? sq : -sq);
    r1 = d + z;
    r2 = (z != 0.0) ? d - (b * c) / z : d;
    i1 = i2 = 0.0;
  } else {
    r1 = r2 = 0.5 * (a + d);
    i1 = std::sqrt(-disc);
    i2 = -i1;
  }
}
// Householder reflector for a 2- or 3-vector: P = I - beta v v^T with v[0] = 1.
static void house3(double x, double y, double z, int len, double v[3], double& beta) {
  const double sig = y * y + (len == 3 ? z * z : 0.0);
  v[0] = 1.0; v[1] = y; v[2] = (len == 3 ? z : 0.0);
  if (sig == 0.0) { beta = 0.0; return; }
  const double mu = std::sqrt(x * x + sig);
  const double v0 = (x <= 0) ? x - mu : -sig / (x + mu);
  beta = 2.0 * v0 * v0 / (sig + v0 * v0);
  v[1] = y / v0;
  v[2] = (len == 3 ? z / v0 : 0.0);
}
// Eigenvalues of an upper Hessenberg matrix (row-major n x n, n <= 6) by Francis
// double-shift QR steps on the active window, deflating when a subdiagonal entry falls
// below eps * (|h(k-1,k-1)| + |h(k,k)|).  Returns 0, or -1 if it did not converge.
static int hqr(double* H, int n, double* wr, double* wi) {
  auto h = [&](int i, int j) -> double& { return H[i * n + j]; };
  int hi = n - 1, its = 0;
  while (hi >= 0) {
    int lo = hi;
    while (lo > 0) {
      double s = std::fabs(h(lo - 1, lo - 1)) + std::fabs(h(lo, lo));
      if (s == 0.0) s = 1.0;
      if (std::fabs(h(lo, lo - 1)) <= EPS * s) { h(lo, lo - 1) = 0.0; break; }
      --lo;
    }
    if (lo == hi) {
      wr[hi] = h(hi, hi); wi[hi] = 0.0; --hi; its = 0; continue;
    }
    if (lo == hi - 1) {
      eig2(h(hi - 1, hi - 1), h(hi - 1, hi), h(hi, hi - 1), h(hi, hi), wr[hi - 1], wi[hi - 1],
           wr[hi], wi[hi]);
      hi -= 2; its = 0; continue;
    }
    if (its >= 100) return -1;
    ++its;
    double s, t;
    if (its % 10 == 0) {  // exceptional (ad hoc) shift
      const double w = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
      s = 1.5 * w;
      t = w * w;
    } else {
      s = h(hi - 1, hi - 1) + h(hi, hi);
      t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
    }
    double x = h(lo, lo) * h(lo, lo) + h(lo, lo + 1) * h(lo + 1, lo) - s * h(lo, lo) + t;
    double y = h(lo + 1, lo) * (h(lo, lo) + h(lo + 1, lo + 1) - s);
    double z = h(lo + 1, lo) * h(lo + 2, lo + 1);
    for (int k = lo - 1; k <= hi - 3; ++k) {
      double v[3], beta;
      house3(x, y, z, 3, v, beta);
      const int q = std::max(lo, k);
      for (int j = q; j <= hi; ++j) {
        const double d = beta * (h(k + 1, j) + v[1] * h(k + 2, j) + v[2] * h(k + 3, j));
        h(k + 1, j) -= d; h(k + 2, j) -= d * v[1]; h(k + 3, j) -= d * v[2];
      }
      const int r = std::min(k + 4, hi);
      for (int i = lo; i <= r; ++i) {
        const double d = beta * (h(i, k + 1) + v[1] * h(i, k + 2) + v[2] * h(i, k + 3));
        h(i, k + 1) -= d; h(i, k + 2) -= d * v[1]; h(i, k + 3) -= d * v[2];
      }
      x = h(k + 2, k + 1);
      y = h(k + 3, k + 1);
      if (k < hi - 3) z = h(k + 4, k + 1);
    }
    double v[3], beta;
    house3(x, y, 0.0, 2, v, beta);
    for (int j = hi - 2; j <= hi; ++j) {
      const double d = beta * (h(hi - 1, j) + v[1] * h(hi, j));
      h(hi - 1, j) -= d; h(hi, j) -= d * v[1];
    }
    for (int i = lo; i <= hi; ++i) {
      const double d = beta * (h(i, hi - 1) + v[1] * h(i, hi));
      h(i, hi - 1) -= d; h(i, hi) -= d * v[1];
    }
  }
  return 0;
}
// eigvals of a general real matrix (balance -> Hessenberg -> QR); NaN on failure.
static void eigvals(const double* Ain, int n, double* wr, double* wi) {
  double A[36];
  std::memcpy(A, Ain, sizeof(double) * n * n);
  balance(A, n);
  hessenberg(A, n);
  if (hqr(A, n, wr, wi) != 0)
    for (int i = 0; i < n; ++i) { wr[i] = NaN; wi[i] = NaN; }
}

// jacobian6 (src/autogen/jacobian6.jl:14-59) in central-moment form.  Arguments follow
// the reference's (m00,m01,...,m40) positional order, m_ij = E[f^i g^j] with f the flux
// direction.  The entries below are the reference Jacobian rewritten in the mean (a,b)
// and central moments c_ij of the plane; rows 1, 2 and 4 are shift rows.
static void jacobian6(const double* m, double J[6][6]) {
  const double m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3], m04 = m[4], m10 = m[5], m11 = m[6],
               m12 = m[7], m13 = m[8], m20 = m[9], m21 = m[10], m22 = m[11], m30 = m[12],
               m31 = m[13], m40 = m[14];
  double R[5][5] = {{0}};
  R[0][0] = 1; R[0][1] = m01 / m00; R[0][2] = m02 / m00; R[0][3] = m03 / m00; R[0][4] = m04 / m00;
  R[1][0] = m10 / m00; R[1][1] = m11 / m00; R[1][2] = m12 / m00; R[1][3] = m13 / m00;
  R[2][0] = m20 / m00; R[2][1] = m21 / m00; R[2][2] = m22 / m00; R[3][0] = m30 / m00;
  R[3][1] = m31 / m00; R[4][0] = m40 / m00;
  const double a = R[1][0], b = R[0][1];
  double c[5][5] = {{0}};
  for (int i = 0; i <= 4; ++i)
    for (int j = 0; i + j <= 4; ++j) {
      double sacc = 0;
      for (int p = 0; p <= i; ++p)
        for (int q = 0; q <= j; ++q)
          sacc += binom(i, p) * binom(j, q) * ipow(-a, i - p) * ipow(-b, j - q) * R[p][q];
      c[i][j] = sacc;
    }
  const double c20 = c[2][0], c02 = c[0][2], c11 = c[1][1], c30 = c[3][0], c03 = c[0][3],
               c21 = c[2][1], c12 = c[1][2], c40 = c[4][0], c04 = c[0][4], c31 = c[3][1],
               c13 = c[1][3], c22 = c[2][2];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) J[i][j] = 0.0;
  J[0][1] = 1.0;
  J[1][2] = 1.0;
  J[3][4] = 1.0;
  const double a2 = a * a, b2 = b * b;
  J[2][0] = (2 * a2 * a * c02 * c02 * c20 * c20 + 3 * a2 * c02 * c02 * c20 * c30 -
             2 * a * c02 * c02 * c20 * c40 + 3 * a * c02 * c02 * c30 * c30 +
             6 * b * c02 * c11 * c20 * c40 - 9 * b * c02 * c11 * c30 * c30 -
             6 * b * c02 * c20 * c20 * c31 + 9 * b * c02 * c20 * c21 * c30 -
             c02 * c02 * c20 * c20 * c30 + 2 * c03 * c11 * c20 * c40 - 3 * c03 * c11 * c30 * c30 -
             2 * c03 * c20 * c20 * c31 + 3 * c03 * c20 * c21 * c30) /
            (2 * c02 * c02 * c20 * c20);
  J[2][1] = -(6 * a2 * c20 * c20 + 6 * a * c20 * c30 - 2 * c20 * c40 + 3 * c30 * c30) / (2 * c20 * c20);
  J[2][2] = 3 * (2 * a * c20 + c30) / (2 * c20);
  J[2][3] = -(2 * c11 * c20 * c40 - 3 * c11 * c30 * c30 - 2 * c20 * c20 * c31 + 3 * c20 * c21 * c30) /
            (2 * c02 * c20 * c20);
  J[4][0] = (6 * a2 * b * c02 * c02 * c02 * c20 + 3 * a2 * c02 * c02 * c03 * c20 +
             6 * a * b * c02 * c02 * c02 * c30 + 3 * a * c02 * c02 * c03 * c30 -
             12 * b2 * c02 * c02 * c11 * c30 + 12 * b2 * c02 * c02 * c20 * c21 +
             3 * b * c02 * c02 * c02 * c20 * c20 + 9 * b * c02 * c02 * c12 * c30 -
             9 * b * c02 * c02 * c20 * c22 - 18 * b * c02 * c03 * c11 * c30 +
             18 * b * c02 * c03 * c20 * c21 + 3 * c02 * c03 * c12 * c30 - 3 * c02 * c03 * c20 * c22 +
             2 * c02 * c04 * c11 * c30 - 2 * c02 * c04 * c20 * c21 - 6 * c03 * c03 * c11 * c30 +
             6 * c03 * c03 * c20 * c21) /
            (2 * c02 * c02 * c02 * c20);
  J[4][1] = -3 * (2 * a * c20 + c30) * (2 * b * c02 + c03) / (2 * c02 * c20);
  J[4][2] = 3 * (2 * b * c02 + c03) / (2 * c02);
  J[4][3] = -(2 * a2 * c02 * c02 * c20 + 2 * a * c02 * c02 * c30 - 8 * b * c02 * c11 * c30 +
              8 * b * c02 * c20 * c21 + c02 * c02 * c20 * c20 + 3 * c02 * c12 * c30 -
              3 * c02 * c20 * c22 - 6 * c03 * c11 * c30 + 6 * c03 * c20 * c21) /
            (2 * c02 * c02 * c20);
  J[4][4] = (2 * a * c20 + c30) / c20;
  J[4][5] = -(c11 * c30 - c20 * c21) / (c02 * c20);
  const double c02_2 = c02 * c02, c02_3 = c02_2 * c02, c02_4 = c02_3 * c02;
  J[5][0] = (24 * a * b2 * c02_4 + 24 * a * b * c02_3 * c03 + 6 * a * c02_4 * c02 -
             10 * a * c02_3 * c04 + 15 * a * c02_2 * c03 * c03 + 60 * b2 * c02_3 * c12 -
             48 * b2 * c02_2 * c03 * c11 - 12 * b * c02_4 * c11 - 24 * b * c02_3 * c13 +
             90 * b * c02_2 * c03 * c12 + 24 * b * c02_2 * c04 * c11 - 81 * b * c02 * c03 * c03 * c11 -
             6 * c02_4 * c12 - 8 * c02_2 * c03 * c13 - 10 * c02_2 * c04 * c12 +
             30 * c02 * c03 * c03 * c12 + 16 * c02 * c03 * c04 * c11 - 27 * c03 * c03 * c03 * c11) /
            (4 * c02_4);
  J[5][1] = -(24 * b2 * c02_2 + 24 * b * c02 * c03 + 6 * c02_3 - 10 * c02 * c04 + 15 * c03 * c03) /
            (4 * c02_2);
  J[5][3] = -(16 * a * b * c02_3 + 8 * a * c02_2 * c03 + 40 * b * c02_2 * c12 -
              32 * b * c02 * c03 * c11 - 4 * c02_3 * c11 - 8 * c02_2 * c13 + 30 * c02 * c03 * c12 +
              8 * c02 * c04 * c11 - 27 * c03 * c03 * c11) /
            (4 * c02_3);
  J[5][4] = 2 * (2 * b * c02 + c03) / c02;
  J[5][5] = (2 * a * c02_2 + 5 * c02 * c12 - 4 * c03 * c11) / (2 * c02_2);
}

struct Eig6 {
  double wr[6], wi[6];
  bool finite;
};
static bool maxabsimag_gt(const Eig6& e, double thr) {  // Julia maximum(abs.(imag(lam))) > thr
  double m = 0;
  for (int i = 0; i < 6; ++i) {
    if (std::isnan(e.wi[i])) return false;  // maximum() is NaN; NaN > thr is false
    m = std::max(m, std::fabs(e.wi[i]));
  }
  return m > thr;
}
static void eig_of_slice(const double* m15, Eig6& e) {
  double J[6][6];
  jacobian6(m15, J);
  bool fin = true;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j)
      if (!std::isfinite(J[i][j])) fin = false;
  e.finite = fin;
  if (!fin) {
    for (int i = 0; i < 6; ++i) { e.wr[i] = NaN; e.wi[i] = NaN; }
    return;
  }
  eigvals(&J[0][0], 6, e.wr, e.wi);
}
static void minmax_real(const Eig6& e, double& mn, double& mx) {
  // sort(real(lam)) with NaN last; [1] and [6]
  double r[6];
  for (int i = 0; i < 6; ++i) r[i] = e.wr[i];
  std::sort(r, r + 6, [](double x, double y) {
    if (std::isnan(x)) return false;
    if (std::isnan(y)) return true;
    return x < y;
  });
  mn = r[0];
  mx = r[5];
}

// axis_moment_slice (src/utils/axis_moment_slice.jl:31-54); 1-based reference indices.
static const int SLICE[4][15] = {
    {1, 6, 10, 13, 15, 2, 7, 11, 14, 3, 8, 12, 4, 9, 5},     // UV (x)
    {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},     // VU (y)
    {1, 16, 20, 23, 25, 2, 17, 21, 24, 3, 18, 22, 4, 19, 5},  // UW (x)
    {1, 16, 20, 23, 25, 6, 26, 32, 34, 10, 29, 35, 13, 31, 15}};  // VW (y)
// z Jacobians, eigenvalues6z_hyperbolic_3D.jl:107 (WU) and :117-120 (WV).
static const int SLICE_WU[15] = {1, 2, 3, 4, 5, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25};
static void take15(const double* M, const int* idx, double* out) {
  for (int t = 0; t < 15; ++t) out[t] = M[idx[t] - 1];
}

// compute_jacobian_eigenvalues (src/numerics/compute_jacobian_eigenvalues.jl:18-50)
static void compute_eigs_axis(const double* M, int axis, double& vmin, double& vmax, Eig6& la,
                              Eig6& lb) {
  double sa[15], sb[15];
  take15(M, SLICE[axis == 1 ? 0 : 1], sa);
  take15(M, SLICE[axis == 1 ? 2 : 3], sb);
  eig_of_slice(sa, la);
  if (!la.finite) { vmin = NaN; vmax = NaN; }
  else minmax_real(la, vmin, vmax);
  eig_of_slice(sb, lb);
  if (lb.finite) {
    double mn, mx;
    minmax_real(lb, mn, mx);
    vmin = jmin(vmin, mn);
    vmax = jmax(vmax, mx);
  }
}

// C4toM4_3D on a 35-vector of central moments (35-order, C000 ignored).
static void c35_to_m35(double M000, double u, double v, double w, const double* C35, double* M35) {
  Cube C;
  std::memset(&C, 0, sizeof(C));
  for (int n = 0; n < 35; ++n) C(E35[n][0], E35[n][1], E35[n][2]) = C35[n];
  Cube R;
  central_to_raw(M000, u, v, w, C, 4, R);
  cube_to35(R, M35);
}
static int idx35(int i, int j, int k) {
  for (int n = 0; n < 35; ++n)
    if (E35[n][0] == i && E35[n][1] == j && E35[n][2] == k) return n;
  return -1;
}
// correct_{uv,uw,vw}_plane (eigenvalues6_hyperbolic_3D.jl:181-252 and the inlined copies in
// eigenvalues6z_hyperbolic_3D.jl:225-283): the 3rd/4th-order cross moments of plane (p,q)
// are rebuilt from the plane correlation, and the (2,2) moment is floored.
static void correct_plane(double* C35, const double* C4, const double* S4, int p, int q) {
  int e[3];
  auto id = [&](int ip, int iq) {
    e[0] = e[1] = e[2] = 0;
    e[p] += ip;
    e[q] += iq;
    return idx35(e[0], e[1], e[2]);
  };
  const double Spq = S4[id(1, 1)], Sp3 = S4[id(3, 0)], Sq3 = S4[id(0, 3)], Sp4 = S4[id(4, 0)],
               Sq4 = S4[id(0, 4)];
  double S22 = S4[id(2, 2)];
  const double Sqq_p = Spq * Sq3;  // S(1,2)
  const double Spp_q = Spq * Sp3;  // S(2,1)
  const double S13 = Spq * Sq4, S31 = Spq * Sp4;
  const double s22min = (2 + 5 * Sq3 * Spq * Sp3) / 6;
  if (S22 < s22min) S22 = s22min;
  const double sp = std::sqrt(jmax(EPS, C4[id(2, 0)])), sq = std::sqrt(jmax(EPS, C4[id(0, 2)]));
  C35[id(2, 1)] = Spp_q * ipow(sp, 2) * sq;
  C35[id(1, 2)] = Sqq_p * sp * ipow(sq, 2);
  C35[id(3, 1)] = S31 * ipow(sp, 3) * sq;
  C35[id(1, 3)] = S13 * sp * ipow(sq, 3);
  C35[id(2, 2)] = S22 * ipow(sp, 2) * ipow(sq, 2);
}

// eigenvalues6_hyperbolic_3D (src/numerics/eigenvalues6_hyperbolic_3D.jl:22-155)
static void eigenvalues6_hyperbolic(const double* M, int axis, int flag2D, double Ma, double& vmin,
                                    double& vmax, double* Mr, uint32_t* flags) {
  std::memcpy(Mr, M, sizeof(double) * 35);
  Eig6 la, lb;
  compute_eigs_axis(M, axis, vmin, vmax, la, lb);
  const double thr = 1000 * EPS;
  const bool ca = maxabsimag_gt(la, thr), cb = maxabsimag_gt(lb, thr);
  if (ca || cb) {
    if (flags) *flags |= (axis == 1 ? HYQ_FLAG_EIGCORR_X : HYQ_FLAG_EIGCORR_Y);
    const double M000 = M[0], um = M[1] / M000, vm = M[5] / M000, wm = M[15] / M000;
    double C4[35], S4[35];
    m2cs4(M, C4, S4);
    double Cn[35];
    std::memcpy(Cn, C4, sizeof(Cn));
    if (ca) correct_plane(Cn, C4, S4, 0, 1);  // always the UV plane
    if (cb) {
      if (axis == 1) correct_plane(Cn, C4, S4, 0, 2);  // UW
      else correct_plane(Cn, C4, S4, 1, 2);            // VW
    }
    double Mh[35];
    c35_to_m35(M000, um, vm, wm, Cn, Mh);
    flux_closure35(Mh, flag2D, Ma, nullptr, nullptr, nullptr, Mr, flags);
    Eig6 a2, b2;
    compute_eigs_axis(Mr, axis, vmin, vmax, a2, b2);
  }
}

// eigenvalues6z_hyperbolic_3D (src/numerics/eigenvalues6z_hyperbolic_3D.jl:23-365)
static void eigenvalues6z_hyperbolic(const double* M, int flag2D, double Ma, bool positivity,
                                     double& vmin, double& vmax, double* Mr, uint32_t* flags) {
  std::memcpy(Mr, M, sizeof(double) * 35);
  if (positivity && M[0] <= 0.0) {
    Mr[0] = jmax(M[0], 1e-14);
    if (flags) *flags |= HYQ_FLAG_RHOCLAMP;
  }
  double W[35];
  std::memcpy(W, M, sizeof(W));
  W[0] = Mr[0];
  if (positivity && W[0] > 1e-15) {
    const double m000 = W[0];
    const double vx = M[2] / m000 - (M[1] / m000) * (M[1] / m000);
    const double vy = M[9] / m000 - (M[5] / m000) * (M[5] / m000);
    const double vz = M[19] / m000 - (M[15] / m000) * (M[15] / m000);
    if (vx < 0.0 || vy < 0.0 || vz < 0.0) {
      if (flags) *flags |= HYQ_FLAG_ZNEGVAR;
      flux_closure35(M, flag2D, Ma, nullptr, nullptr, nullptr, Mr, flags);
      std::memcpy(W, Mr, sizeof(W));
    }
  }
  double sa[15], sb[15];
  take15(W, SLICE_WU, sa);
  take15(W, SLICE[3], sb);
  Eig6 la, lb;
  eig_of_slice(sa, la);
  eig_of_slice(sb, lb);
  if (!la.finite || !lb.finite) {
    // Julia's eigvals throws on a non-finite matrix here (no guard in the reference).
    if (flags) *flags |= HYQ_FLAG_NONFINITE;
  }
  double mn, mx;
  minmax_real(la, vmin, vmax);
  minmax_real(lb, mn, mx);
  vmin = jmin(vmin, mn);
  vmax = jmax(vmax, mx);
  const double thr = 1000 * EPS;
  const bool ca = maxabsimag_gt(la, thr), cb = maxabsimag_gt(lb, thr);
  if (ca || cb) {
    if (flags) *flags |= HYQ_FLAG_EIGCORR_Z;
    const double M000 = M[0], um = M[1] / M000, vm = M[5] / M000, wm = M[15] / M000;
    double C4[35], S4[35];
    m2cs4(M, C4, S4);
    double Cn[35];
    std::memcpy(Cn, C4, sizeof(Cn));
    if (ca) correct_plane(Cn, C4, S4, 0, 2);  // WU: the xz cross moments
    if (cb) correct_plane(Cn, C4, S4, 1, 2);  // WV: the yz cross moments
    double Mh[35];
    c35_to_m35(M000, um, vm, wm, Cn, Mh);
    flux_closure35(Mh, flag2D, Ma, nullptr, nullptr, nullptr, Mr, flags);
    take15(Mr, SLICE_WU, sa);
    take15(Mr, SLICE[3], sb);
    eig_of_slice(sa, la);
    eig_of_slice(sb, lb);
    minmax_real(la, vmin, vmax);
    minmax_real(lb, mn, mx);
    vmin = jmin(vmin, mn);
    vmax = jmax(vmax, mx);
  }
}

// closure_and_eigenvalues (src/numerics/closure_and_eigenvalues.jl:17-78), N = 2.
static void closure_and_eigenvalues(const double* mom, double& Mp, double& vpmin, double& vpmax) {
  const int N = 2;
  double sig[N + 3][2 * N + 4];
  double a[N + 2], b[N + 3];
  for (auto& r : sig)
    for (double& x : r) x = 0.0;
  for (double& x : a) x = 0.0;
  for (double& x : b) x = 0.0;
  // 1-based indices as in the reference
  for (int i = 2; i <= 2 * N + 2; ++i) sig[2][i] = mom[i - 2];
  a[1] = mom[1] / mom[0];
  for (int k = 3; k <= N + 1; ++k) {
    for (int l = k; l <= 2 * N - k + 4; ++l)
      sig[k][l] = sig[k - 1][l + 1] - a[k - 2] * sig[k - 1][l] - b[k - 2] * sig[k - 2][l];
    a[k - 1] = sig[k][k + 1] / sig[k][k] - sig[k - 1][k] / sig[k - 1][k - 1];
    b[k - 1] = sig[k][k] / sig[k - 1][k - 1];
  }
  int k = N + 2;
  sig[k][k] = sig[k - 1][k + 1] - a[k - 2] * sig[k - 1][k] - b[k - 2] * sig[k - 2][k];
  b[k - 1] = sig[k][k] / sig[k - 1][k - 1];
  double sa = 0;
  for (int i = 1; i <= N; ++i) sa += a[i];
  a[N + 1] = sa / N;
  sig[N + 2][N + 3] = sig[N + 2][N + 2] * (a[N + 1] + sig[N + 1][N + 2] / sig[N + 1][N + 1]);
  for (k = N + 2; k >= 3; --k) {
    const int l = 2 * N - k + 5;
    sig[k - 1][l + 1] = sig[k][l] + a[k - 2] * sig[k - 1][l] + b[k - 2] * sig[k - 2][l];
  }
  Mp = sig[2][2 * N + 3];
  b[N + 1] = b[N + 1] * (2 * N + 1) / N;
  // Jacobi matrix z: diag a[1..3], off-diagonal sqrt(b[2]), sqrt(b[3]) (complex if b < 0).
  // Same characteristic polynomial as the real matrix [[a1,b2,0],[1,a2,b3],[0,1,a3]].
  const double T[9] = {a[1], b[2], 0.0, 1.0, a[2], b[3], 0.0, 1.0, a[3]};
  for (double x : T)
    if (!std::isfinite(x)) { vpmin = NaN; vpmax = NaN; return; }
  double wr[3], wi[3];
  eigvals(T, 3, wr, wi);
  vpmin = jmin(jmin(wr[0], wr[1]), wr[2]);
  vpmax = jmax(jmax(wr[0], wr[1]), wr[2]);
}

// 3x3 symmetric eigen-decomposition (cyclic Jacobi) for sqrt(C2) in S4toC4_3D_r.
static void sym3_sqrt(const double C[3][3], double A[3][3]) {
  double a[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  std::memcpy(a, C, sizeof(a));
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  double sq[3];
  for (int i = 0; i < 3; ++i) sq[i] = std::sqrt(std::max(0.0, a[i][i]));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += V[i][k] * sq[k] * V[j][k];
      A[i][j] = s;
    }
}
// S4toC4_3D_r (src/autogen/S4toC4_3D_r.jl:19-80): central moments of x = A s, A = sqrt(C2),
// from the standardized moments of s (given as a cube, order <= 4).
static void s4_to_c4(const double C2[3][3], const Cube& S, Cube& C) {
  double A[3][3];
  sym3_sqrt(C2, A);
  std::memset(&C, 0, sizeof(C));
  for (int i = 0; i <= 4; ++i)
    for (int j = 0; i + j <= 4; ++j)
      for (int k = 0; i + j + k <= 4; ++k) {
        // expand prod_r (A_r . s)^{e_r} as a polynomial in s
        Cube P;
        std::memset(&P, 0, sizeof(P));
        P(0, 0, 0) = 1.0;
        const int ex[3] = {i, j, k};
        for (int r = 0; r < 3; ++r)
          for (int e = 0; e < ex[r]; ++e) {
            Cube Qc;
            std::memset(&Qc, 0, sizeof(Qc));
            for (int p = 0; p <= 4; ++p)
              for (int q = 0; p + q <= 4; ++q)
                for (int t = 0; p + q + t <= 4; ++t) {
                  const double v = P(p, q, t);
                  if (v == 0.0) continue;
                  if (p + q + t + 1 <= 5) {
                    Qc(p + 1, q, t) += v * A[r][0];
                    Qc(p, q + 1, t) += v * A[r][1];
                    Qc(p, q, t + 1) += v * A[r][2];
                  }
                }
            P = Qc;
          }
        double s = 0;
        for (int p = 0; p <= 4; ++p)
          for (int q = 0; p + q <= 4; ++q)
            for (int t = 0; p + q + t <= 4; ++t) s += P(p, q, t) * S(p, q, t);
        C(i, j, k) = s;
      }
}
// InitializeM4_35 (src/moments/InitializeM4_35.jl:14-76): Gaussian standardized moments.
static void initialize_m4_35(double M000, double u, double v, double w, double C200, double C110,
                             double C101, double C020, double C011, double C002, double* M35) {
  Cube S;
  std::memset(&S, 0, sizeof(S));
  S(0, 0, 0) = 1;
  S(2, 0, 0) = S(0, 2, 0) = S(0, 0, 2) = 1;
  S(4, 0, 0) = S(0, 4, 0) = S(0, 0, 4) = 3;
  S(2, 2, 0) = S(2, 0, 2) = S(0, 2, 2) = 1;
  const double C2[3][3] = {{C200, C110, C101}, {C110, C020, C011}, {C101, C011, C002}};
  Cube C;
  s4_to_c4(C2, S, C);
  Cube R;
  central_to_raw(M000, u, v, w, C, 4, R);
  cube_to35(R, M35);
}

// collision35 (src/numerics/collision35.jl:17-45)
static void collision35(const double* M, double dt, double Kn, bool positivity, double* Mout) {
  const double rho = M[0], um = M[1] / rho, vm = M[5] / rho, wm = M[15] / rho;
  const double C200 = M[2] / rho - um * um, C020 = M[9] / rho - vm * vm,
               C002 = M[19] / rho - wm * wm;
  double Th = (C200 + C020 + C002) / 3;
  if (positivity) Th = jmax(Th, 1e-14);
  const double tc = Kn / (rho * std::sqrt(Th) * 2);
  double MG[35];
  initialize_m4_35(rho, um, vm, wm, Th, 0.0, 0.0, Th, 0.0, Th, MG);
  const double e = std::exp(-dt / tc);
  for (int n = 0; n < 35; ++n) Mout[n] = MG[n] - e * (MG[n] - M[n]);
}

// pas_HLL + flux_HLL (src/numerics/pas_HLL.jl:24-84, flux_HLL.jl:20-32) on one line of Np
// cells, physical BC at both ends.  M, F: [Np][35] row-major.  Mp out.
static void pas_hll_line(const double* M, const double* Fin, int Np, double dt, double dx,
                         const double* vpmin, const double* vpmax, double* Mp) {
  const int K = 35;
  std::vector<double> W(Np * K, 0.0), F(Fin, Fin + Np * K), ll(Np, 0.0), lr(Np, 0.0);
  for (int j = 1; j <= Np - 2; ++j) {  // 0-based j = reference j-1, j in 2..Np-1
    ll[j] = jmin(vpmin[j], vpmin[j + 1]);
    lr[j] = jmax(vpmax[j], vpmax[j + 1]);
    if (std::fabs(ll[j] - lr[j]) > 1e-10) {
      for (int k = 0; k < K; ++k)
        W[j * K + k] = (ll[j] * M[j * K + k] - lr[j] * M[(j + 1) * K + k]) / (ll[j] - lr[j]) -
                       (F[j * K + k] - F[(j + 1) * K + k]) / (ll[j] - lr[j]);
    }
  }
  if (Np >= 2) {
    for (int k = 0; k < K; ++k) W[0 * K + k] = W[1 * K + k];
    for (int k = 0; k < K; ++k) W[(Np - 1) * K + k] = W[(Np - 2) * K + k];
    for (int k = 0; k < K; ++k) F[0 * K + k] = F[1 * K + k];
    for (int k = 0; k < K; ++k) F[(Np - 1) * K + k] = F[(Np - 2) * K + k];
  }
  std::vector<double> Fl((Np > 1 ? Np - 1 : 0) * K);
  for (int j = 0; j < Np - 1; ++j)
    for (int k = 0; k < K; ++k)
      Fl[j * K + k] = 0.5 * (F[j * K + k] + F[(j + 1) * K + k]) -
                      0.5 * (std::fabs(ll[j]) * (W[j * K + k] - M[j * K + k]) -
                             std::fabs(lr[j]) * (W[j * K + k] - M[(j + 1) * K + k]));
  for (int i = 0; i < Np * K; ++i) Mp[i] = M[i];
  for (int j = 1; j <= Np - 2; ++j)
    for (int k = 0; k < K; ++k)
      Mp[j * K + k] = M[j * K + k] - dt / dx * (Fl[j * K + k] - Fl[(j - 1) * K + k]);
  if (Np >= 2) {
    for (int k = 0; k < K; ++k) Mp[0 * K + k] = Mp[1 * K + k];
    for (int k = 0; k < K; ++k) Mp[(Np - 1) * K + k] = Mp[(Np - 2) * K + k];
  }
}

// One time step of simulation_runner (src/simulation_runner.jl:346-560) on a single
// device domain.  State layout: Julia column-major (Nx,Ny,Nz,35).
struct Run {
  int Nx, Ny, Nz, flag2D;
  double Kn, Ma, CFL, dtmax, dx, dy, dz;
  bool positivity;
  int nthreads;
};
static inline size_t at(const Run& r, int i, int j, int k, int m) {
  return (size_t)i + (size_t)r.Nx * ((size_t)j + (size_t)r.Ny * ((size_t)k + (size_t)r.Nz * m));
}
static int step(const Run& r, double* M, double tmax, double& t, double& dt_out,
                double* vp_out, uint32_t* flags_out) {
  const size_t N = (size_t)r.Nx * r.Ny * r.Nz;
  std::vector<double> Fx(N * 35), Fy(N * 35), Fz(N * 35), Mnp(N * 35), vp(6 * N);
  int err = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(r.nthreads) reduction(| : err)
  for (long long c = 0; c < (long long)N; ++c) {
    const int i = c % r.Nx, j = (c / r.Nx) % r.Ny, k = c / ((size_t)r.Nx * r.Ny);
    double MOM[35], Mr[35], Mr2[35], fx[35], fy[35], fz[35];
    uint32_t fl = 0;
    for (int m = 0; m < 35; ++m) MOM[m] = M[at(r, i, j, k, m)];
    double xmn, xmx, ymn, ymx, zmn, zmx;
    flux_closure35(MOM, r.flag2D, r.Ma, nullptr, nullptr, nullptr, Mr, &fl);
    eigenvalues6_hyperbolic(Mr, 1, r.flag2D, r.Ma, xmn, xmx, Mr2, &fl);
    eigenvalues6_hyperbolic(Mr2, 2, r.flag2D, r.Ma, ymn, ymx, Mr, &fl);
    eigenvalues6z_hyperbolic(Mr, r.flag2D, r.Ma, r.positivity, zmn, zmx, Mr2, &fl);
    flux_closure35(Mr2, r.flag2D, r.Ma, fx, fy, fz, Mr, &fl);
    double Mp, a5, b5;
    double v5[3][2];
    const int sl[3][5] = {{0, 1, 2, 3, 4}, {0, 5, 9, 12, 14}, {0, 15, 19, 22, 24}};
    for (int d = 0; d < 3; ++d) {
      double mom[5];
      for (int t2 = 0; t2 < 5; ++t2) mom[t2] = Mr[sl[d][t2]];
      closure_and_eigenvalues(mom, Mp, a5, b5);
      v5[d][0] = a5;
      v5[d][1] = b5;
    }
    vp[0 * N + c] = jmin(v5[0][0], xmn);
    vp[1 * N + c] = jmax(v5[0][1], xmx);
    vp[2 * N + c] = jmin(v5[1][0], ymn);
    vp[3 * N + c] = jmax(v5[1][1], ymx);
    vp[4 * N + c] = jmin(v5[2][0], zmn);
    vp[5 * N + c] = jmax(v5[2][1], zmx);
    for (int m = 0; m < 35; ++m) {
      Fx[at(r, i, j, k, m)] = fx[m];
      Fy[at(r, i, j, k, m)] = fy[m];
      Fz[at(r, i, j, k, m)] = fz[m];
      Mnp[at(r, i, j, k, m)] = Mr[m];
    }
    if (fl & HYQ_FLAG_NONFINITE) err = 1;
    if (flags_out) flags_out[c] = fl;
  }
  std::memcpy(M, Mnp.data(), sizeof(double) * N * 35);
  if (vp_out) std::memcpy(vp_out, vp.data(), sizeof(double) * 6 * N);
  // vmax over interior |vp| (NaN-propagating maximum), :402
  double vmax = -std::numeric_limits<double>::infinity();
  for (size_t q = 0; q < 6 * N; ++q) vmax = jmax(vmax, std::fabs(vp[q]));
  double dt = jmin(r.CFL * jmin(jmin(r.dx, r.dy), r.dz) / vmax, jmin(r.dtmax, r.Kn));
  dt = jmin(dt, tmax - t);
  t += dt;
  dt_out = dt;
  // directional HLL updates (:499-513) and additive combine (:516-520)
  std::vector<double> Mx(M, M + N * 35), My(M, M + N * 35), Mz(M, M + N * 35);
  auto sweep = [&](int axis, std::vector<double>& F, std::vector<double>& Mout, double ds) {
    const int n0 = axis == 0 ? r.Nx : (axis == 1 ? r.Ny : r.Nz);
    if (axis == 2 && r.Nz == 1) return;  // apply_flux_update_3d.jl:183-187
    const int na = axis == 0 ? r.Ny : r.Nx, nb = axis == 2 ? r.Ny : r.Nz;
#pragma omp parallel for collapse(2) num_threads(r.nthreads)
    for (int b = 0; b < nb; ++b)
      for (int a = 0; a < na; ++a) {
        std::vector<double> lm(n0 * 35), lf(n0 * 35), lo(n0 * 35), mn(n0), mx(n0);
        for (int s = 0; s < n0; ++s) {
          int i, j, k;
          if (axis == 0) { i = s; j = a; k = b; }
          else if (axis == 1) { i = a; j = s; k = b; }
          else { i = a; j = b; k = s; }
          const size_t c = (size_t)i + (size_t)r.Nx * ((size_t)j + (size_t)r.Ny * k);
          for (int m = 0; m < 35; ++m) {
            lm[s * 35 + m] = M[at(r, i, j, k, m)];
            lf[s * 35 + m] = F[at(r, i, j, k, m)];
          }
          mn[s] = vp[(2 * axis) * N + c];
          mx[s] = vp[(2 * axis + 1) * N + c];
        }
        pas_hll_line(lm.data(), lf.data(), n0, dt, ds, mn.data(), mx.data(), lo.data());
        for (int s = 0; s < n0; ++s) {
          int i, j, k;
          if (axis == 0) { i = s; j = a; k = b; }
          else if (axis == 1) { i = a; j = s; k = b; }
          else { i = a; j = b; k = s; }
          for (int m = 0; m < 35; ++m) Mout[at(r, i, j, k, m)] = lo[s * 35 + m];
        }
      }
  };
  sweep(0, Fx, Mx, r.dx);
  sweep(1, Fy, My, r.dy);
  sweep(2, Fz, Mz, r.dz);
  for (size_t q = 0; q < N * 35; ++q) M[q] = Mx[q] + My[q] + Mz[q] - 2.0 * M[q];
  // pass 2: realizability (:526-543) then BGK (:546-557)
#pragma omp parallel for schedule(dynamic, 16) num_threads(r.nthreads)
  for (long long c = 0; c < (long long)N; ++c) {
    const int i = c % r.Nx, j = (c / r.Nx) % r.Ny, k = c / ((size_t)r.Nx * r.Ny);
    double MOM[35], Mr[35], Mr2[35], Mo[35];
    uint32_t fl = 0;
    double d0, d1;
    for (int m = 0; m < 35; ++m) MOM[m] = M[at(r, i, j, k, m)];
    flux_closure35(MOM, r.flag2D, r.Ma, nullptr, nullptr, nullptr, Mr, &fl);
    eigenvalues6_hyperbolic(Mr, 1, r.flag2D, r.Ma, d0, d1, Mr2, &fl);
    eigenvalues6_hyperbolic(Mr2, 2, r.flag2D, r.Ma, d0, d1, Mr, &fl);
    flux_closure35(Mr, r.flag2D, r.Ma, nullptr, nullptr, nullptr, Mr2, &fl);
    collision35(Mr2, dt, r.Kn, r.positivity, Mo);
    for (int m = 0; m < 35; ++m) Mnp[at(r, i, j, k, m)] = Mo[m];
  }
  std::memcpy(M, Mnp.data(), sizeof(double) * N * 35);
  return err ? HYQ_ENONFINITE : HYQ_OK;
}

}  // namespace orc

// ---------------------------------------------------------------------------------------
// C entry points used by the tests (ctypes).  All arrays are host memory.
extern "C" {

void orc_m2cs4(const double* M, double* C4, double* S4) { orc::m2cs4(M, C4, S4); }

void orc_flux_closure35(const double* M, int flag2D, double Ma, double* Fx, double* Fy,
                        double* Fz, double* Mr, uint32_t* flags) {
  uint32_t f = 0;
  orc::flux_closure35(M, flag2D, Ma, Fx, Fy, Fz, Mr, &f);
  if (flags) *flags = f;
}

void orc_delta2star3D(const double* s28, double* E36) {
  orc::S28 s;
  std::memcpy(&s, s28, sizeof(s));
  double E[6][6];
  orc::delta2star3D(s, E);
  std::memcpy(E36, E, sizeof(E));
}
void orc_lower_bound_S220(const double* s28, double* out3) {
  orc::S28 s;
  std::memcpy(&s, s28, sizeof(s));
  orc::lower_bound_S220(s, out3);
}
void orc_bound_minor1(const double* s28, double* out12) {
  orc::S28 s;
  std::memcpy(&s, s28, sizeof(s));
  orc::bound_minor1(s, out12);
}
double orc_delta2starchol_L3(const double* a10) {
  return orc::delta2starchol_L3(a10[0], a10[1], a10[2], a10[3], a10[4], a10[5], a10[6], a10[7],
                                a10[8], a10[9]);
}
void orc_rootsR(const double* a10, double* re3, double* im3) {
  std::complex<double> R[3];
  orc::rootsR(a10[0], a10[1], a10[2], a10[3], a10[4], a10[5], a10[6], a10[7], a10[8], a10[9], R);
  for (int k = 0; k < 3; ++k) { re3[k] = R[k].real(); im3[k] = R[k].imag(); }
}
void orc_rootsR_X_Y(const double* a9, double* re3, double* im3) {
  std::complex<double> R[3];
  orc::rootsR_X_Y(a9[0], a9[1], a9[2], a9[3], a9[4], a9[5], a9[6], a9[7], a9[8], R);
  for (int k = 0; k < 3; ++k) { re3[k] = R[k].real(); im3[k] = R[k].imag(); }
}
void orc_jacobian6(const double* m15, double* J36) {
  double J[6][6];
  orc::jacobian6(m15, J);
  std::memcpy(J36, J, sizeof(J));
}
void orc_eigvals(const double* A, int n, double* wr, double* wi) { orc::eigvals(A, n, wr, wi); }
void orc_realizability_S2(const double* in3, double* out4) {
  double a = in3[0], b = in3[1], c = in3[2], s2;
  orc::realizability_S2(a, b, c, s2);
  out4[0] = a; out4[1] = b; out4[2] = c; out4[3] = s2;
}
void orc_check2D(const double* in10, double* out10) {
  orc::check2D(in10[0], in10[1], in10[2], in10[3], in10[4], in10[5], in10[6], in10[7], in10[8],
               in10[9], out10, nullptr);
}
void orc_realizable_3D(const double* s28, double* out28, int* flag220) {
  orc::S28 s;
  std::memcpy(&s, s28, sizeof(s));
  *flag220 = orc::realizable_3D(s, nullptr);
  std::memcpy(out28, &s, sizeof(s));
}
void orc_eigenvalues6(const double* M, int axis, int flag2D, double Ma, double* vmin, double* vmax,
                      double* Mr, uint32_t* flags) {
  uint32_t f = 0;
  if (axis == 3) orc::eigenvalues6z_hyperbolic(M, flag2D, Ma, true, *vmin, *vmax, Mr, &f);
  else orc::eigenvalues6_hyperbolic(M, axis, flag2D, Ma, *vmin, *vmax, Mr, &f);
  if (flags) *flags = f;
}
void orc_eigenvalues6z(const double* M, int flag2D, double Ma, int positivity, double* vmin,
                       double* vmax, double* Mr, uint32_t* flags) {
  uint32_t f = 0;
  orc::eigenvalues6z_hyperbolic(M, flag2D, Ma, positivity != 0, *vmin, *vmax, Mr, &f);
  if (flags) *flags = f;
}
void orc_realizable_3D_flags(const double* s28, double* out28, int* flag220, uint32_t* flags) {
  orc::S28 s;
  std::memcpy(&s, s28, sizeof(s));
  uint32_t f = 0;
  *flag220 = orc::realizable_3D(s, &f);
  std::memcpy(out28, &s, sizeof(s));
  *flags = f;
}
void orc_moments5_3d(const double* M4, double* M5, double* C5, double* S5) {
  orc::moments5_3d(M4, M5, C5, S5);
}
void orc_closure_and_eigenvalues(const double* mom5, double* out3) {
  orc::closure_and_eigenvalues(mom5, out3[0], out3[1], out3[2]);
}
void orc_initialize_m4_35(const double* in10, double* M35) {
  orc::initialize_m4_35(in10[0], in10[1], in10[2], in10[3], in10[4], in10[5], in10[6], in10[7],
                        in10[8], in10[9], M35);
}
void orc_s4toc4(const double* C2_9, const double* S35, double* C35) {
  double C2[3][3];
  std::memcpy(C2, C2_9, sizeof(C2));
  orc::Cube S, C;
  std::memset(&S, 0, sizeof(S));
  for (int n = 0; n < 35; ++n) S(orc::E35[n][0], orc::E35[n][1], orc::E35[n][2]) = S35[n];
  orc::s4_to_c4(C2, S, C);
  orc::cube_to35(C, C35);
}
void orc_central_to_raw5(double M000, double u, double v, double w, const double* C56_cube216,
                         double* M216) {
  orc::Cube C, M;
  std::memcpy(&C, C56_cube216, sizeof(C));
  orc::central_to_raw(M000, u, v, w, C, 5, M);
  std::memcpy(M216, &M, sizeof(M));
}
void orc_collision35(const double* M, double dt, double Kn, int positivity, double* Mout) {
  orc::collision35(M, dt, Kn, positivity != 0, Mout);
}
void orc_pas_hll(const double* M, const double* F, int Np, double dt, double dx,
                 const double* vpmin, const double* vpmax, double* Mp) {
  orc::pas_hll_line(M, F, Np, dt, dx, vpmin, vpmax, Mp);
}
// Pass-1 per-cell pipeline on a batch: outputs Mr, Fx, Fy, Fz ([n][35]), vp [n][6], flags.
void orc_pass1_batch(const double* M, long long n, int flag2D, double Ma, int positivity,
                     double* Mr_out, double* Fx, double* Fy, double* Fz, double* vp,
                     uint32_t* flags, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (long long c = 0; c < n; ++c) {
    double Mr[35], Mr2[35];
    uint32_t fl = 0;
    double xmn, xmx, ymn, ymx, zmn, zmx;
    orc::flux_closure35(M + 35 * c, flag2D, Ma, nullptr, nullptr, nullptr, Mr, &fl);
    orc::eigenvalues6_hyperbolic(Mr, 1, flag2D, Ma, xmn, xmx, Mr2, &fl);
    orc::eigenvalues6_hyperbolic(Mr2, 2, flag2D, Ma, ymn, ymx, Mr, &fl);
    orc::eigenvalues6z_hyperbolic(Mr, flag2D, Ma, positivity != 0, zmn, zmx, Mr2, &fl);
    orc::flux_closure35(Mr2, flag2D, Ma, Fx + 35 * c, Fy + 35 * c, Fz + 35 * c, Mr, &fl);
    const int sl[3][5] = {{0, 1, 2, 3, 4}, {0, 5, 9, 12, 14}, {0, 15, 19, 22, 24}};
    double v5[3][2];
    for (int d = 0; d < 3; ++d) {
      double mom[5], Mp;
      for (int t = 0; t < 5; ++t) mom[t] = Mr[sl[d][t]];
      orc::closure_and_eigenvalues(mom, Mp, v5[d][0], v5[d][1]);
    }
    vp[6 * c + 0] = orc::jmin(v5[0][0], xmn);
    vp[6 * c + 1] = orc::jmax(v5[0][1], xmx);
    vp[6 * c + 2] = orc::jmin(v5[1][0], ymn);
    vp[6 * c + 3] = orc::jmax(v5[1][1], ymx);
    vp[6 * c + 4] = orc::jmin(v5[2][0], zmn);
    vp[6 * c + 5] = orc::jmax(v5[2][1], zmx);
    std::memcpy(Mr_out + 35 * c, Mr, sizeof(Mr));
    flags[c] = fl;
  }
}
// Full step loop: while t < tmax && nn < nnmax.  M in Julia layout (Nx,Ny,Nz,35).
int orc_run(double* M, int Nx, int Ny, int Nz, double Kn, double Ma, int flag2D, double CFL,
            double dtmax, double dx, double dy, double dz, int positivity, double tmax,
            long long nnmax, double* t_inout, long long* nn_inout, int nthreads, double* dts) {
  orc::Run r{Nx, Ny, Nz, flag2D, Kn, Ma, CFL, dtmax, dx, dy, dz, positivity != 0, nthreads};
  double t = *t_inout;
  long long nn = *nn_inout;
  int rc = HYQ_OK;
  while (t < tmax && nn < nnmax) {
    ++nn;
    double dt;
    const int e = orc::step(r, M, tmax, t, dt, nullptr, nullptr);
    if (dts) dts[nn - 1 - *nn_inout] = dt;
    if (e != HYQ_OK) { rc = e; break; }
  }
  *t_inout = t;
  *nn_inout = nn;
  return rc;
}
}  // extern "C"
