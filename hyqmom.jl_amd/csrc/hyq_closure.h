// Device realizability projection and HyQMOM closure (FP64, one cell per lane).
//
// Follows Flux_closure35_and_realizable_3D (src/numerics/Flux_closure35_and_realizable_3D.jl)
// and realizable_3D (src/realizability/realizable_3D.jl) branch for branch.  The GPU form
// differs from the reference's arithmetic in three places (all exact rewrites):
//  * Delta2* is evaluated once per state as the Schur complement
//        E = Q - u u^T - W^T adj(B) W / det(B)
//    of the standardized Gram matrix over (1,x,y,z | x2,xy,xz,y2,yz,z2) (B = correlation
//    block, W = third-order block, Q = fourth-order block); the six axis permutations of
//    delta2star3D_permutation.jl are index permutations of the same matrix;
//  * leading-minor determinants use closed-form cofactor expansions;
//  * rootsR / rootsR_X_Y build the cubic's coefficients in closed form (hyq_cubics.h) and
//    solve it with Cardano's formula using a real square root (NaN for three real roots,
//    as the reference's @fastmath sqrt).
#pragma once
#include "hyq_cubics.h"
#include "hyq_moments.h"

namespace hyq {

// Flags (values of HYQ_FLAG_* in include/hyqmom.h)
constexpr uint32_t F_UNIV = 1u << 0, F_EDGE = 1u << 1, F_FACES = 1u << 2, F_220 = 1u << 3,
                   F_R2D = 1u << 4, F_R4 = 1u << 5, F_ECX = 1u << 6, F_ECY = 1u << 7,
                   F_ECZ = 1u << 8, F_ZNEG = 1u << 9, F_NONFIN = 1u << 10, F_RHO = 1u << 11,
                   F_FALLBACK = 1u << 12, F_NOCONV = 1u << 13;

struct S28 {
  double S300, S400, S110, S210, S310, S120, S220, S030, S130, S040, S101, S201, S301, S102,
      S202, S003, S103, S004, S011, S111, S211, S021, S121, S031, S012, S112, S013, S022;
};

// ------------------------------------------------------------------------------------
// Delta2* as a symmetric 6x6 over (x2, xy, xz, y2, yz, z2), 21 packed entries.
HD constexpr int s6(int a, int b) {
  return (a < b ? a : b) * 6 - (a < b ? a : b) * ((a < b ? a : b) - 1) / 2 +
         ((a < b ? b : a) - (a < b ? a : b));
}
struct Sym6 {
  double e[21];
};

HD void delta2star(const S28& s, Sym6& E) {
  const double a = s.S110, b = s.S101, c = s.S011;
  const double A00 = 1 - c * c, A01 = b * c - a, A02 = a * c - b, A11 = 1 - b * b,
               A12 = a * b - c, A22 = 1 - a * a;
  const double rd = 1.0 / (1 + 2 * a * b * c - (a * a + b * b + c * c));
  const double W[3][6] = {{s.S300, s.S210, s.S201, s.S120, s.S111, s.S102},
                          {s.S210, s.S120, s.S111, s.S030, s.S021, s.S012},
                          {s.S201, s.S111, s.S102, s.S021, s.S012, s.S003}};
  const double u[6] = {1.0, a, b, 1.0, c, 1.0};
  double T[3][6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    T[0][q] = A00 * W[0][q] + A01 * W[1][q] + A02 * W[2][q];
    T[1][q] = A01 * W[0][q] + A11 * W[1][q] + A12 * W[2][q];
    T[2][q] = A02 * W[0][q] + A12 * W[1][q] + A22 * W[2][q];
  }
  const double Q[21] = {s.S400, s.S310, s.S301, s.S220, s.S211, s.S202, s.S220,
                        s.S211, s.S130, s.S121, s.S112, s.S202, s.S121, s.S112,
                        s.S103, s.S040, s.S031, s.S022, s.S022, s.S013, s.S004};
#pragma unroll
  for (int p = 0; p < 6; ++p)
#pragma unroll
    for (int q = p; q < 6; ++q)
      E.e[s6(p, q)] = Q[s6(p, q)] - u[p] * u[q] -
                      (W[0][p] * T[0][q] + W[1][p] * T[1][q] + W[2][p] * T[2][q]) * rd;
}
// Z = (Q - E) entries used by lower_bound_S220.jl and bound_minor1.jl, from the same
// Schur complement (the value of a fourth-order moment that zeroes its Delta2* entry).
HD double zpart(const Sym6& E, const S28& s, int p, int q) {
  const double Q[21] = {s.S400, s.S310, s.S301, s.S220, s.S211, s.S202, s.S220,
                        s.S211, s.S130, s.S121, s.S112, s.S202, s.S121, s.S112,
                        s.S103, s.S040, s.S031, s.S022, s.S022, s.S013, s.S004};
  return Q[s6(p, q)] - E.e[s6(p, q)];
}

// Axis permutations of delta2star3D_permutation.jl:26-61 as basis index maps into E1.
struct PermT {
  int p[6][6];
};
HD constexpr PermT perms() {
  return PermT{{{0, 1, 2, 3, 4, 5},    // E1 ijk
                {0, 2, 1, 5, 4, 3},    // E2 ikj
                {3, 1, 4, 0, 2, 5},    // E3 jik
                {5, 2, 4, 0, 1, 3},    // E4 jki
                {3, 4, 1, 5, 2, 0},    // E5 kij
                {5, 4, 2, 3, 1, 0}}};  // E6 kji
}
struct Lead4 {  // leading 4x4 block of a permuted Delta2*
  double e11, e12, e13, e14, e22, e23, e24, e33, e34, e44;
};
template <int P>
HD Lead4 lead4(const Sym6& E) {
  constexpr int a = perms().p[P][0], b = perms().p[P][1], c = perms().p[P][2],
                d = perms().p[P][3];
  Lead4 L;
  L.e11 = E.e[s6(a, a)]; L.e12 = E.e[s6(a, b)]; L.e13 = E.e[s6(a, c)]; L.e14 = E.e[s6(a, d)];
  L.e22 = E.e[s6(b, b)]; L.e23 = E.e[s6(b, c)]; L.e24 = E.e[s6(b, d)];
  L.e33 = E.e[s6(c, c)]; L.e34 = E.e[s6(c, d)]; L.e44 = E.e[s6(d, d)];
  return L;
}
HD double det2(const Lead4& L) { return L.e11 * L.e22 - L.e12 * L.e12; }
HD double det3(const Lead4& L) {
  return L.e11 * (L.e22 * L.e33 - L.e23 * L.e23) - L.e12 * (L.e12 * L.e33 - L.e23 * L.e13) +
         L.e13 * (L.e12 * L.e23 - L.e22 * L.e13);
}
HD double det4(const Lead4& L) {
  // Laplace expansion over the 2x2 minors of rows (1,2) and rows (3,4)
  const double a12 = L.e11 * L.e22 - L.e12 * L.e12, a13 = L.e11 * L.e23 - L.e13 * L.e12,
               a14 = L.e11 * L.e24 - L.e14 * L.e12, a23 = L.e12 * L.e23 - L.e13 * L.e22,
               a24 = L.e12 * L.e24 - L.e14 * L.e22, a34 = L.e13 * L.e24 - L.e14 * L.e23;
  const double b34 = L.e33 * L.e44 - L.e34 * L.e34, b24 = L.e23 * L.e44 - L.e34 * L.e24,
               b23 = L.e23 * L.e34 - L.e33 * L.e24, b14 = L.e13 * L.e44 - L.e34 * L.e14,
               b13 = L.e13 * L.e34 - L.e33 * L.e14, b12 = L.e13 * L.e24 - L.e23 * L.e14;
  return a12 * b34 - a13 * b24 + a14 * b23 + a23 * b14 - a24 * b13 + a34 * b12;
}

// Cardano with a real square root of the discriminant.
struct Roots3 {
  double re[3], im[3];
};
HD Roots3 cardano(const double c4[4]) {
  const double a = c4[0], b = c4[1], c = c4[2], d = c4[3];
  const double p = (3.0 * a * c - b * b) / (3.0 * a * a);
  const double q = (2.0 * b * b * b - 9.0 * a * b * c + 27.0 * a * a * d) / (27.0 * a * a * a);
  const double disc = (q / 2.0) * (q / 2.0) + (p / 3.0) * (p / 3.0) * (p / 3.0);
  const double sq = disc < 0 ? NAN : sqrt(disc);
  const double u = cbrt(-q / 2.0 + sq);
  const double sh = -b / (3.0 * a);
  const double t = u - p / (3.0 * u);
  const double im = 0.8660254037844386 * (u + p / (3.0 * u));
  Roots3 R;
  R.re[0] = sh + t; R.im[0] = 0.0;
  R.re[1] = sh - 0.5 * t; R.im[1] = -im;
  R.re[2] = sh - 0.5 * t; R.im[2] = im;
  return R;
}
HD bool nlt(double p, double q) { return (q != q) ? (p == p) : (p < q); }  // NaN-last order
// Second and third entries of sort(real(R)).
HD void roots_mid_hi(const Roots3& R, double& r2, double& r3) {
  double x = R.re[0], y = R.re[1], z = R.re[2], t;
  if (nlt(y, x)) { t = x; x = y; y = t; }
  if (nlt(z, y)) { t = y; y = z; z = t; }
  if (nlt(y, x)) { t = x; x = y; y = t; }
  r2 = y;
  r3 = z;
}
// maximum(abs.(imag(R))) / maximum(abs.(R)): NaN if any root is NaN.
HD double roots_imag_ratio(const Roots3& R) {
  double mi = 0, ma = 0;
  bool nan = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double ai = fabs(R.im[k]), ar = hypot(R.re[k], R.im[k]);
    nan = nan || (ai != ai) || (ar != ar);
    mi = fmax(mi, ai);
    ma = fmax(ma, ar);
  }
  return nan ? NAN : mi / ma;
}

// ------------------------------------------------------------------------------------
// realizability sub-projections (src/realizability/*.jl)

// enforce_univariate (src/moments/enforce_univariate.jl:19-39)
HD void enforce_univariate(double& S3, double& S4, double& H, double h2min, double s3max) {
  H = S4 - S3 * S3 - 1;
  if (H <= h2min) { H = h2min; S4 = H + S3 * S3 + 1; }
  if (S3 < -s3max) { S3 = -s3max; S4 = H + S3 * S3 + 1; }
  else if (S3 > s3max) { S3 = s3max; S4 = H + S3 * S3 + 1; }
}
// realizablity_S220 (src/realizability/realizablity_S220.jl:19-31)
HD double clamp_S220(double S110, double S220, double A220) {
  const double lo = jmax(S110 * S110, 1 - A220), hi = 1 + A220;
  return S220 < lo ? lo : (S220 > hi ? hi : S220);
}
// realizability_S2 + find_zero_bisection (src/realizability/realizability_S2.jl:18-66)
HD double realizability_S2(double& S110, double& S101, double& S011) {
  const double q = S110 * S110 + S101 * S101 + S011 * S011, r = 2 * S110 * S101 * S011;
  const double S2 = 1 + r - q;
  double xr = 1.0;
  if (S2 < 0) {
    double a = 0.0, b = 1.0;
    const double fa0 = 1.0;  // Y(0)
    double fa = fa0;
    bool done = false;
    for (int it = 0; it < 100; ++it) {
      const double c = (a + b) / 2, fc = 1 + r * c * c * c - q * c * c;
      if (fabs(fc) < 1e-12 || (b - a) / 2 < 1e-12) { xr = c; done = true; break; }
      if (jsign(fc) == jsign(fa)) { a = c; fa = fc; }
      else { b = c; }
    }
    if (!done) xr = (a + b) / 2;
  }
  xr = 0.9999 * xr;
  S110 *= xr;
  S101 *= xr;
  S011 *= xr;
  return 1 + 2 * S110 * S101 * S011 - (S110 * S110 + S101 * S101 + S011 * S011);
}
// realizability_S210 (src/realizability/realizability_S210.jl:20-44).  With U = sqrt(D1),
// V = U X and Xr = U \ (beta xr V) = beta xr X exactly, and V.V = X^T D1 X.
HD void realizability_S210(double S110, double S101, double S011, double S300, double& S210,
                           double& S201, double H200, double beta) {
  const double X0 = S210 - S110 * S300, X1 = S201 - S101 * S300;
  const double d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  const double L = jmax(0.0, X0 * (d11 * X0 + d12 * X1) + X1 * (d12 * X0 + d22 * X1));
  const double dD1 = jmax(0.0, d11 * d22 - d12 * d12);
  const double R = H200 * dD1;
  double xr = 1.0;
  if (R <= 0 || X0 * X0 + X1 * X1 < 1000 * EPS) xr = 0.0;
  else if (L > R) xr = sqrt(R / L);
  S210 = beta * xr * X0 + S110 * S300;
  S201 = beta * xr * X1 + S101 * S300;
}
// realizability_S111 (src/realizability/realizability_S111.jl:18-36)
HD double realizability_S111(double S110, double S101, double S011, double S210, double S201,
                             double S120, double S021, double S102, double S012, double S111) {
  const double A110 = ((S101 - S011 * S110) * S210 + (S011 - S101 * S110) * S120) / (1 - S110 * S110);
  const double A101 = ((S110 - S011 * S101) * S201 + (S011 - S110 * S101) * S102) / (1 - S101 * S101);
  const double A011 = ((S110 - S101 * S011) * S021 + (S101 - S110 * S011) * S012) / (1 - S011 * S011);
  const double Rmin = jmin(jmin(A110, A101), A011), Rmax = jmax(jmax(A110, A101), A011);
  return S111 > Rmax ? Rmax : (S111 < Rmin ? Rmin : S111);
}
HD double quadD2(double S110, double S101, double S011, double v0, double v1, double v2) {
  const double d00 = 1 - S011 * S011, d01 = S011 * S101 - S110, d02 = S011 * S110 - S101,
               d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  return v0 * (d00 * v0 + d01 * v1 + d02 * v2) + v1 * (d01 * v0 + d11 * v1 + d12 * v2) +
         v2 * (d02 * v0 + d12 * v1 + d22 * v2);
}
// realizability_S310 (src/realizability/realizability_S310.jl:20-69); S220 passes through.
HD double realizability_S310(double S110, double S101, double S011, double S300, double S210,
                             double S201, double S120, double S111, double S310, double S220,
                             double H200, double beta) {
  const double b310 =
      S111 * ((1 - S110 * S110) * S201 + (S101 * S110 - S011) * S210 + (S011 * S110 - S101) * S300) +
      S120 * ((1 - S101 * S101) * S210 + (S011 * S101 - S110) * S300 + (S101 * S110 - S011) * S201) +
      S210 * ((1 - S011 * S011) * S300 + (S011 * S101 - S110) * S210 + (S011 * S110 - S101) * S201);
  const double d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  const double dD1 = d11 * d22 - d12 * d12;
  const double V0 = S210 - S110 * S300, V1 = S201 - S101 * S300;
  const double L1 = (V0 * (d11 * V0 + d12 * V1) + V1 * (d12 * V0 + d22 * V1)) / dD1;
  double G310b = H200 - L1;
  if (G310b < 0) G310b = 0.0;
  const double L2 = quadD2(S110, S101, S011, S210, S120, S111) / dD1;
  double G310a = S220 - S110 * S110 - L2 + 1000 * EPS;
  if (G310a < 0 || L2 < 0) G310a = 0.0;
  const double sG = beta * sqrt(G310a * G310b);
  const double lo = S110 + b310 / dD1 - sG, hi = S110 + b310 / dD1 + sG;
  return S310 <= lo ? lo : (S310 >= hi ? hi : S310);
}
// realizability_S310_220 (src/realizability/realizability_S310_220.jl:18-37)
HD double realizability_S310_220(double S110, double S101, double S011, double S210,
                                 double S120, double S111, double S220) {
  const double d11 = 1 - S101 * S101, d12 = S101 * S110 - S011, d22 = 1 - S110 * S110;
  const double L2 = quadD2(S110, S101, S011, S210, S120, S111) / (d11 * d22 - d12 * d12);
  const double mn = S110 * S110 + L2 + 1000 * EPS;
  return mn > S220 ? mn : S220;
}
// realizability_S211 (src/realizability/realizability_S211.jl:19-34)
HD double realizability_S211(double e11, double e22, double e33, double e12, double e13,
                             double d23, double S211, double beta) {
  const double b = e12 * e13;
  const double sG = beta * sqrt(jmax(0.0, (e11 * e22 - e13 * e13) * (e11 * e33 - e12 * e12)));
  const double lo = d23 + (b - sG) / e11, hi = d23 + (b + sG) / e11;
  return S211 <= lo ? lo : (S211 >= hi ? hi : S211);
}

// ------------------------------------------------------------------------------------
// 2D: realizable_2D (src/realizability/realizable_2D.jl:24-118) and check2D
// (src/utils/diagnostics.jl:17-42).  Only reached on the edge/corner branch: the interior
// branch of Flux_closure35 never reads check2D's outputs.
HD void schur2d(double s03, double s04, double s11, double s12, double s13, double s21,
                double s22, double s30, double s31, double s40, double e[6]) {
  // basis (x2, xy, y2) against (1, x, y); B = [[1,s11],[s11,1]]
  const double rd = 1.0 / (1 - s11 * s11);
  const double W[2][3] = {{s30, s21, s12}, {s21, s12, s03}};
  const double u[3] = {1.0, s11, 1.0};
  const double Q[6] = {s40, s31, s22, s22, s13, s04};  // (00,01,02,11,12,22)
  const int P[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const int p = P[t][0], q = P[t][1];
    const double tq0 = W[0][q] - s11 * W[1][q], tq1 = W[1][q] - s11 * W[0][q];
    e[t] = Q[t] - u[p] * u[q] - (W[0][p] * tq0 + W[1][p] * tq1) * rd;
  }
}
HD double det3s(const double e[6]) {
  return e[0] * (e[3] * e[5] - e[4] * e[4]) - e[1] * (e[1] * e[5] - e[4] * e[2]) +
         e[2] * (e[1] * e[4] - e[3] * e[2]);
}
HD void realizable_2D(double S30, double S40, double S11, double& S21, double& S31, double& S12,
                      double& S22, double S03, double& S13, double S04, uint32_t& fl) {
  const double Del1 = jmax(0.0, 1 - S11 * S11);
  const double H20 = jmax(EPS, S40 - S30 * S30 - 1), H02 = jmax(EPS, S04 - S03 * S03 - 1);
  double G1 = sqrt(Del1 * H02);
  double lo = S11 * S03 - G1, hi = S11 * S03 + G1;
  S12 = S12 <= lo ? lo : (S12 >= hi ? hi : S12);
  G1 = sqrt(Del1 * H20);
  lo = S11 * S30 - G1;
  hi = S11 * S30 + G1;
  S21 = S21 <= lo ? lo : (S21 >= hi ? hi : S21);
  const double G22 = sqrt((H20 + S30 * S30) * (H02 + S03 * S03));
  const double s22min = jmax(S11 * S11, 1 - G22), s22max = 1 + G22;
  S22 = S22 < s22min ? s22min : (S22 > s22max ? s22max : S22);
  const double base = Del1 * S22 - Del1 * S11 * S11 - S12 * S12 + 2 * S11 * S12 * S21 - S21 * S21;
  double G31 = base * (Del1 * H20 - (S21 - S11 * S30) * (S21 - S11 * S30));
  double G13 = base * (Del1 * H02 - (S12 - S11 * S03) * (S12 - S11 * S03));
  if (G31 < 0 || G13 < 0) { G31 = 0.0; G13 = 0.0; }
  else { G13 = sqrt(G13); G31 = sqrt(G31); }
  const double n31 = S12 * S21 + S21 * S30 - S11 * S21 * S21 - S11 * S12 * S30;
  lo = S11 + (n31 - G31) / (Del1 + EPS);
  hi = S11 + (n31 + G31) / (Del1 + EPS);
  S31 = S31 <= lo ? lo : (S31 >= hi ? hi : S31);
  const double n13 = S12 * S21 + S12 * S03 - S11 * S12 * S12 - S11 * S21 * S03;
  lo = S11 + (n13 - G13) / (Del1 + EPS);
  hi = S11 + (n13 + G13) / (Del1 + EPS);
  S13 = S13 <= lo ? lo : (S13 >= hi ? hi : S13);
  double e[6];
  schur2d(S03, S04, S11, S12, S13, S21, S22, S30, S31, S40, e);
  const double L3 = (1.0 - S11 * S11) * det3s(e);  // delta2starchol_L3.jl
  if (L3 < 0) {
    fl |= F_R2D;
    double e0[6], c[4];
    schur2d(S03, H02 + S03 * S03 + 1.0, S11, S12, S13, S21, 0.0, S30, S31, H20 + S30 * S30 + 1.0, e0);
    cubic_2d(e0[0], e0[1], e0[2], e0[3], e0[4], e0[5], c);
    const Roots3 R = cardano(c);
    double r2, r3;
    roots_mid_hi(R, r2, r3);
    const double mx = jmin(s22max, r3);
    double mn = jmax(s22min, r2);
    if (mx < mn) mn = mx;
    const double S22a = S22 > mx ? mx : (S22 < mn ? mn : S22);
    // "ratio > 1000 eps" keeps S22; a NaN ratio (three real roots) falls to the else branch
    const double ratio = roots_imag_ratio(R);
    if (!(ratio > 1000 * EPS)) S22 = S22a;
  }
}
struct P10 {
  double S30, S40, S11, S21, S31, S12, S22, S03, S13, S04;
};
HD P10 check2D(P10 p, uint32_t& fl) {
  const double h2min = 1000 * EPS;
  if (fabs(p.S11) >= 1) {
    p.S11 = jsign(p.S11);
    const double s3m = sqrt(fabs(p.S30 * p.S03));
    double s4m = sqrt(p.S40 * p.S04);
    const double H2m = jmax(h2min, s4m - s3m * s3m - 1);
    s4m = H2m + s3m * s3m + 1;
    p.S30 = jsign(p.S30) * s3m;
    p.S03 = p.S11 * p.S30;
    p.S40 = s4m;
    p.S04 = s4m;
    p.S12 = p.S11 * p.S03;
    p.S21 = p.S11 * p.S30;
    p.S13 = p.S11 * p.S04;
    p.S31 = p.S11 * p.S40;
    p.S22 = s4m;
  } else {
    realizable_2D(p.S30, p.S40, p.S11, p.S21, p.S31, p.S12, p.S22, p.S03, p.S13, p.S04, fl);
  }
  return p;
}


// edge_corner_correction (src/realizability/edge_corner_correction.jl:24-253).  p1, p2, p3
// are check2D's outputs on the XY, XZ and YZ planes; s holds the values that entered them.
HD void edge_corner(double R110, double R101, double R011, S28& s, const P10& p1, const P10& p2,
                    const P10& p3) {
  // names as in the reference call (Flux_closure35_and_realizable_3D.jl:112-120)
  const double S110r = p1.S11, S101r = p2.S11, S011r = p3.S11;
  const double S300r1 = p1.S30, S030r1 = p1.S03, S400r1 = p1.S40, S040r1 = p1.S04;
  const double S300r2 = p2.S30, S003r2 = p2.S03, S400r2 = p2.S40, S004r2 = p2.S04;
  const double S030r3 = p3.S30, S003r3 = p3.S03, S040r3 = p3.S40, S004r3 = p3.S04;
  const double S300r = s.S300, S030r = s.S030, S003r = s.S003, S400r = s.S400, S040r = s.S040,
               S004r = s.S004;
  const double S210r = p1.S21, S201r = p2.S21, S120r = p1.S12, S102r = p2.S12, S021r = p3.S21,
               S012r = p3.S12;
  const double S310r = p1.S31, S301r = p2.S31, S220r = p1.S22, S202r = p2.S22, S130r = p1.S13,
               S103r = p2.S13, S031r = p3.S31, S022r = p3.S22, S013r = p3.S13;
  double S110, S101, S011;
  if (R110 <= 0 && R101 > 0 && R011 > 0) {
    S110 = jsign(S110r);
    const double Sm = (S011r + S101r) / 2;
    S011 = jsign(S011r) * Sm;
    S101 = jsign(S101r) * Sm;
    realizability_S2(S110, S101, S011);
    s.S300 = S300r1; s.S030 = S030r1; s.S400 = S400r1; s.S040 = S040r1; s.S003 = S003r; s.S004 = S004r;
    const double x = S110, a300 = S300r, a400 = S400r, S030 = s.S030, S040 = s.S040;
    s.S210 = (S030 + x * a300) / 2;            s.S201 = (S021r + S201r) / 2;
    s.S120 = (x * S030 + a300) / 2;            s.S111 = (x * S021r + x * S201r) / 2;
    s.S102 = (x * S012r + S102r) / 2;          s.S021 = (S021r + S201r) / 2;
    s.S012 = (S012r + x * S102r) / 2;          s.S310 = (x * S040 + x * a400) / 2;
    s.S301 = (x * S031r + S301r) / 2;          s.S220 = (S040 + a400) / 2;
    s.S211 = (S031r + x * S301r) / 2;          s.S202 = (S022r + S202r) / 2;
    s.S130 = (x * S040 + x * a400) / 2;        s.S121 = (x * S031r + S301r) / 2;
    s.S112 = (x * S022r + x * S202r) / 2;      s.S103 = (x * S013r + S103r) / 2;
    s.S031 = (S031r + x * S301r) / 2;          s.S022 = (S022r + S202r) / 2;
    s.S013 = (S013r + x * S103r) / 2;
  } else if (R101 <= 0 && R110 > 0 && R011 > 0) {
    S101 = jsign(S101r);
    const double Sm = (S011r + S110r) / 2;
    S011 = jsign(S011r) * Sm;
    S110 = jsign(S110r) * Sm;
    realizability_S2(S110, S101, S011);
    s.S300 = S300r2; s.S003 = S003r2; s.S400 = S400r2; s.S004 = S004r2; s.S030 = S030r; s.S040 = S040r;
    const double y = S101, a300 = S300r, a400 = S400r, S003 = s.S003, S004 = s.S004;
    s.S210 = (S013r + S210r) / 2;              s.S201 = (S003 + y * a300) / 2;
    s.S120 = (y * S021r + S120r) / 2;          s.S111 = (y * S012r + y * S210r) / 2;
    s.S102 = (y * S003 + a300) / 2;            s.S021 = (S021r + y * S120r) / 2;
    s.S012 = (S012r + S210r) / 2;              s.S310 = (y * S013r + S310r) / 2;
    s.S301 = (y * S004 + y * a400) / 2;        s.S220 = (S022r + S220r) / 2;
    s.S211 = (S013r + y * S310r) / 2;          s.S202 = (S004 + a400) / 2;
    s.S130 = (y * S031r + S130r) / 2;          s.S121 = (y * S022r + y * S220r) / 2;
    s.S112 = (y * S013r + S310r) / 2;          s.S103 = (y * S004 + y * a400) / 2;
    s.S031 = (S031r + y * S130r) / 2;          s.S022 = (S022r + S220r) / 2;
    s.S013 = (S013r + y * S310r) / 2;
  } else if (R011 <= 0 && R101 > 0 && R110 > 0) {
    S011 = jsign(S011r);
    const double Sm = (S101r + S110r) / 2;
    S101 = jsign(S101r) * Sm;
    S110 = jsign(S110r) * Sm;
    realizability_S2(S110, S101, S011);
    s.S030 = S030r3; s.S003 = S003r3; s.S040 = S040r3; s.S004 = S004r3; s.S300 = S300r; s.S400 = S400r;
    const double z = S011, a030 = S030r, a040 = S040r, S003 = s.S003, S004 = s.S004;
    s.S210 = (z * S201r + S210r) / 2;          s.S201 = (S201r + z * S210r) / 2;
    s.S120 = (S102r + S120r) / 2;              s.S111 = (z * S102r + z * S120r) / 2;
    s.S102 = (S102r + S120r) / 2;              s.S021 = (S003 + z * a030) / 2;
    s.S012 = (z * S003 + a030) / 2;            s.S310 = (z * S301r + S310r) / 2;
    s.S301 = (S301r + z * S310r) / 2;          s.S220 = (S202r + S220r) / 2;
    s.S211 = (z * S202r + z * S220r) / 2;      s.S202 = (S202r + S220r) / 2;
    s.S130 = (z * S103r + S130r) / 2;          s.S121 = (S103r + z * S130r) / 2;
    s.S112 = (z * S103r + S130r) / 2;          s.S103 = (S103r + z * S130r) / 2;
    s.S031 = (z * S004 + z * a040) / 2;        s.S022 = (S004 + a040) / 2;
    s.S013 = (z * S004 + z * a040) / 2;
  } else {
    S110 = jsign(S110r);
    S101 = jsign(S101r);
    S011 = jsign(S011r);
    realizability_S2(S110, S101, S011);
    s.S300 = (S030r1 + S030r + S300r) / 3;
    s.S030 = (S030r1 + S030r + S030r3) / 3;
    s.S003 = (S003r + S003r2 + S003r3) / 3;
    s.S400 = (S040r1 + S040r + S400r) / 3;
    s.S040 = (S040r1 + S040r + S040r3) / 3;
    s.S004 = (S004r + S004r2 + S004r3) / 3;
    // average of the three corner variants (average_corner_variants, :203-253)
    const double A3 = S030r1, A4 = S040r1, B3 = S030r, B4 = S040r, C3 = S003r3, C4 = S004r3;
    const double x = S110, y = S101, z = S011;
    s.S210 = (x * A3 + B3 + z * C3) / 3;       s.S201 = (y * A3 + z * B3 + C3) / 3;
    s.S120 = (A3 + x * B3 + y * C3) / 3;       s.S111 = (x * y * A3 + x * z * B3 + y * z * C3) / 3;
    s.S102 = (A3 + x * B3 + y * C3) / 3;       s.S021 = (y * A3 + z * B3 + C3) / 3;
    s.S012 = (x * A3 + B3 + z * C3) / 3;       s.S310 = (x * A4 + x * B4 + y * z * C4) / 3;
    s.S301 = (y * A4 + x * z * B4 + y * C4) / 3; s.S220 = (A4 + B4 + C4) / 3;
    s.S211 = (x * y * A4 + z * B4 + z * C4) / 3; s.S202 = (A4 + B4 + C4) / 3;
    s.S130 = (x * A4 + x * B4 + y * z * C4) / 3; s.S121 = (y * A4 + x * z * B4 + y * C4) / 3;
    s.S112 = (x * A4 + x * B4 + y * z * C4) / 3; s.S103 = (y * A4 + x * z * B4 + y * C4) / 3;
    s.S031 = (x * y * A4 + z * B4 + z * C4) / 3; s.S022 = (A4 + B4 + C4) / 3;
    s.S013 = (x * y * A4 + z * B4 + z * C4) / 3;
  }
  s.S110 = S110;
  s.S101 = S101;
  s.S011 = S011;
}

// ------------------------------------------------------------------------------------
// realizable_3D (src/realizability/realizable_3D.jl:39-624).  Returns flag220.
HD int realizable_3D(S28& s, uint32_t& fl) {
  int flag220 = 0;
  const double diagmin = 1.0e-10, h2min = 1.0e-10, S2min = 1.0e-12;
  const double H200 = jmax(EPS, s.S400 - s.S300 * s.S300 - 1);
  const double H020 = jmax(EPS, s.S040 - s.S030 * s.S030 - 1);
  const double H002 = jmax(EPS, s.S004 - s.S003 * s.S003 - 1);
  const double S220max =
      clamp_S220(s.S110, s.S220, sqrt((H200 + s.S300 * s.S300) * (H020 + s.S030 * s.S030)));
  const double S202max =
      clamp_S220(s.S101, s.S202, sqrt((H200 + s.S300 * s.S300) * (H002 + s.S003 * s.S003)));
  const double S022max =
      clamp_S220(s.S011, s.S022, sqrt((H020 + s.S030 * s.S030) * (H002 + s.S003 * s.S003)));
  s.S220 = jmin(s.S220, S220max);
  s.S202 = jmin(s.S202, S202max);
  s.S022 = jmin(s.S022, S022max);
  const double S2 = realizability_S2(s.S110, s.S101, s.S011);
  const double R110 = sqrt(jmax(0.0, 1 - s.S110 * s.S110));
  const double R101 = sqrt(jmax(0.0, 1 - s.S101 * s.S101));
  const double R011 = sqrt(jmax(0.0, 1 - s.S011 * s.S011));

  if (S2 <= S2min) {  // faces (:79-142)
    fl |= F_FACES;
    const double Rmax = jmax(jmax(R110, R101), R011);
    double g1, g2;
    if (R110 == Rmax) {
      g1 = (s.S101 - s.S110 * s.S011) / (1 - s.S110 * s.S110);
      g2 = (s.S011 - s.S110 * s.S101) / (1 - s.S110 * s.S110);
      s.S021 = g2 * s.S030 + g1 * s.S120;
      s.S201 = g2 * s.S210 + g1 * s.S300;
      s.S031 = g2 * s.S040 + g1 * s.S130;
      s.S301 = g2 * s.S310 + g1 * s.S400;
      s.S111 = g2 * s.S120 + g1 * s.S210;
      s.S211 = g2 * s.S220 + g1 * s.S310;
      s.S121 = g2 * s.S130 + g1 * s.S220;
      s.S012 = g2 * g2 * s.S030 + 2 * g1 * g2 * s.S120 + g1 * g1 * s.S210;
      s.S102 = g2 * g2 * s.S120 + 2 * g1 * g2 * s.S210 + g1 * g1 * s.S300;
      s.S022 = g2 * g2 * s.S040 + 2 * g1 * g2 * s.S130 + g1 * g1 * s.S220;
      s.S202 = g2 * g2 * s.S220 + 2 * g1 * g2 * s.S310 + g1 * g1 * s.S400;
      s.S112 = g2 * g2 * s.S130 + 2 * g1 * g2 * s.S220 + g1 * g1 * s.S310;
      s.S003 = g2 * g2 * g2 * s.S030 + 3 * g1 * g2 * g2 * s.S120 + 3 * g1 * g1 * g2 * s.S210 + g1 * g1 * g1 * s.S300;
      s.S013 = g2 * g2 * g2 * s.S040 + 3 * g1 * g2 * g2 * s.S130 + 3 * g1 * g1 * g2 * s.S220 + g1 * g1 * g1 * s.S310;
      s.S103 = g2 * g2 * g2 * s.S130 + 3 * g1 * g2 * g2 * s.S220 + 3 * g1 * g1 * g2 * s.S310 + g1 * g1 * g1 * s.S400;
      s.S004 = g2 * g2 * g2 * g2 * s.S040 + 4 * g1 * g2 * g2 * g2 * s.S130 + 6 * g1 * g1 * g2 * g2 * s.S220 +
               4 * g1 * g1 * g1 * g2 * s.S310 + g1 * g1 * g1 * g1 * s.S400;
    } else if (R101 == Rmax) {
      g1 = (s.S110 - s.S101 * s.S011) / (1 - s.S101 * s.S101);
      g2 = (s.S011 - s.S110 * s.S101) / (1 - s.S101 * s.S101);
      s.S012 = g2 * s.S003 + g1 * s.S102;
      s.S210 = g2 * s.S201 + g1 * s.S300;
      s.S013 = g2 * s.S004 + g1 * s.S103;
      s.S310 = g2 * s.S301 + g1 * s.S400;
      s.S111 = g2 * s.S102 + g1 * s.S201;
      s.S211 = g2 * s.S202 + g1 * s.S301;
      s.S112 = g2 * s.S103 + g1 * s.S202;
      s.S021 = g2 * g2 * s.S003 + 2 * g1 * g2 * s.S102 + g1 * g1 * s.S201;
      s.S120 = g2 * g2 * s.S102 + 2 * g1 * g2 * s.S201 + g1 * g1 * s.S300;
      s.S022 = g2 * g2 * s.S004 + 2 * g1 * g2 * s.S103 + g1 * g1 * s.S202;
      s.S220 = g2 * g2 * s.S202 + 2 * g1 * g2 * s.S301 + g1 * g1 * s.S400;
      s.S121 = g2 * g2 * s.S103 + 2 * g1 * g2 * s.S202 + g1 * g1 * s.S301;
      s.S030 = g2 * g2 * g2 * s.S003 + 3 * g1 * g2 * g2 * s.S102 + 3 * g1 * g1 * g2 * s.S201 + g1 * g1 * g1 * s.S300;
      s.S031 = g2 * g2 * g2 * s.S004 + 3 * g1 * g2 * g2 * s.S103 + 3 * g1 * g1 * g2 * s.S202 + g1 * g1 * g1 * s.S301;
      s.S130 = g2 * g2 * g2 * s.S103 + 3 * g1 * g2 * g2 * s.S202 + 3 * g1 * g1 * g2 * s.S301 + g1 * g1 * g1 * s.S400;
      s.S040 = g2 * g2 * g2 * g2 * s.S004 + 4 * g1 * g2 * g2 * g2 * s.S103 + 6 * g1 * g1 * g2 * g2 * s.S202 +
               4 * g1 * g1 * g1 * g2 * s.S301 + g1 * g1 * g1 * g1 * s.S400;
    } else {
      g1 = (s.S101 - s.S110 * s.S011) / (1 - s.S011 * s.S011);
      g2 = (s.S110 - s.S101 * s.S011) / (1 - s.S011 * s.S011);
      s.S102 = g2 * s.S003 + g1 * s.S012;
      s.S120 = g2 * s.S021 + g1 * s.S030;
      s.S103 = g2 * s.S004 + g1 * s.S013;
      s.S130 = g2 * s.S031 + g1 * s.S040;
      s.S111 = g2 * s.S012 + g1 * s.S021;
      s.S121 = g2 * s.S022 + g1 * s.S031;
      s.S112 = g2 * s.S013 + g1 * s.S022;
      s.S201 = g2 * g2 * s.S003 + 2 * g1 * g2 * s.S012 + g1 * g1 * s.S021;
      s.S210 = g2 * g2 * s.S012 + 2 * g1 * g2 * s.S021 + g1 * g1 * s.S030;
      s.S202 = g2 * g2 * s.S004 + 2 * g1 * g2 * s.S013 + g1 * g1 * s.S022;
      s.S220 = g2 * g2 * s.S022 + 2 * g1 * g2 * s.S031 + g1 * g1 * s.S040;
      s.S211 = g2 * g2 * s.S013 + 2 * g1 * g2 * s.S022 + g1 * g1 * s.S031;
      s.S300 = g2 * g2 * g2 * s.S003 + 3 * g1 * g2 * g2 * s.S012 + 3 * g1 * g1 * g2 * s.S021 + g1 * g1 * g1 * s.S030;
      s.S301 = g2 * g2 * g2 * s.S004 + 3 * g1 * g2 * g2 * s.S013 + 3 * g1 * g1 * g2 * s.S022 + g1 * g1 * g1 * s.S031;
      s.S310 = g2 * g2 * g2 * s.S013 + 3 * g1 * g2 * g2 * s.S022 + 3 * g1 * g1 * g2 * s.S031 + g1 * g1 * g1 * s.S040;
      s.S400 = g2 * g2 * g2 * g2 * s.S004 + 4 * g1 * g2 * g2 * g2 * s.S013 + 6 * g1 * g1 * g2 * g2 * s.S022 +
               4 * g1 * g1 * g1 * g2 * s.S031 + g1 * g1 * g1 * g1 * s.S040;
    }
    return flag220;
  }

  // interior (:144-619)
  const double beta = 1.0;
  Sym6 E;
  delta2star(s, E);
  bool f11 = false, f22 = false, f33 = false, f44 = false, f55 = false, f66 = false;
  if (E.e[s6(0, 0)] < 0) { realizability_S210(s.S110, s.S101, s.S011, s.S300, s.S210, s.S201, H200, beta); f11 = true; }
  if (E.e[s6(3, 3)] < 0) { realizability_S210(s.S110, s.S011, s.S101, s.S030, s.S120, s.S021, H020, beta); f44 = true; }
  if (E.e[s6(5, 5)] < 0) { realizability_S210(s.S011, s.S101, s.S110, s.S003, s.S012, s.S102, H002, beta); f66 = true; }
  s.S111 = realizability_S111(s.S110, s.S101, s.S011, s.S210, s.S201, s.S120, s.S021, s.S102, s.S012, s.S111);
  const double S220_310 = realizability_S310_220(s.S110, s.S101, s.S011, s.S210, s.S120, s.S111, s.S220);
  const double S202_310 = realizability_S310_220(s.S101, s.S110, s.S011, s.S201, s.S102, s.S111, s.S202);
  const double S022_310 = realizability_S310_220(s.S011, s.S110, s.S101, s.S021, s.S012, s.S111, s.S022);
  double S220_diag = s.S220, S202_diag = s.S202, S022_diag = s.S022;
  delta2star(s, E);  // third-order moments may have changed: the lower bounds use them
  {
    const double lb0 = zpart(E, s, 1, 1), lb1 = zpart(E, s, 2, 2), lb2 = zpart(E, s, 4, 4);
    if (s.S220 < lb0) { S220_diag = lb0; f22 = true; flag220 = 1; }
    if (s.S202 < lb1) { S202_diag = lb1; f33 = true; flag220 = 1; }
    if (s.S022 < lb2) { S022_diag = lb2; f55 = true; flag220 = 1; }
  }
  if (E.e[s6(0, 0)] < diagmin) f11 = true;
  if (E.e[s6(1, 1)] < diagmin) f22 = true;
  if (E.e[s6(2, 2)] < diagmin) f33 = true;
  if (E.e[s6(3, 3)] < diagmin) f44 = true;
  if (E.e[s6(4, 4)] < diagmin) f55 = true;
  if (E.e[s6(5, 5)] < diagmin) f66 = true;
  if (f11 || f22 || f33 || f44 || f55 || f66) {
    // bound_minor1.jl: values that zero the off-diagonal entries
    const double S310b = zpart(E, s, 0, 1), S301b = zpart(E, s, 0, 2), S130b = zpart(E, s, 1, 3),
                 S103b = zpart(E, s, 2, 5), S031b = zpart(E, s, 3, 4), S013b = zpart(E, s, 4, 5),
                 S211b = zpart(E, s, 0, 4), S121b = zpart(E, s, 2, 3), S112b = zpart(E, s, 1, 5);
    if (f11) { s.S310 = S310b; s.S301 = S301b; s.S211 = S211b; }
    if (f22) { s.S310 = S310b; s.S130 = S130b; s.S211 = S211b; s.S121 = S121b; s.S112 = S112b; }
    if (f33) { s.S301 = S301b; s.S103 = S103b; s.S211 = S211b; s.S121 = S121b; s.S112 = S112b; }
    if (f44) { s.S130 = S130b; s.S031 = S031b; s.S121 = S121b; }
    if (f55) { s.S031 = S031b; s.S013 = S013b; s.S211 = S211b; s.S121 = S121b; s.S112 = S112b; }
    if (f66) { s.S103 = S103b; s.S013 = S013b; s.S112 = S112b; }
  }
  if (f11 && f22 && f33 && f44 && f55 && f66) return flag220;

  // 2x2 minors (:284-378)
  double S220a = s.S220, S220b = s.S220, S202a = s.S202, S202b = s.S202, S022a = s.S022,
         S022b = s.S022;
  delta2star(s, E);
  bool flag310 = false;
  if (!f11 && !f44 && det2(lead4<0>(E)) < h2min) {
    s.S310 = realizability_S310(s.S110, s.S101, s.S011, s.S300, s.S210, s.S120, s.S201, s.S111, s.S310, s.S220, H200, beta);
    S220a = s.S220; flag310 = true;
  }
  if (!f22 && !f44 && det2(lead4<2>(E)) < h2min) {
    s.S130 = realizability_S310(s.S110, s.S011, s.S101, s.S030, s.S120, s.S021, s.S210, s.S111, s.S130, s.S220, H020, beta);
    S220b = s.S220; flag310 = true;
  }
  if (!f11 && !f33 && det2(lead4<1>(E)) < h2min) {
    s.S301 = realizability_S310(s.S101, s.S110, s.S011, s.S300, s.S201, s.S210, s.S102, s.S111, s.S301, s.S202, H200, beta);
    S202a = s.S202; flag310 = true;
  }
  if (!f33 && !f66 && det2(lead4<3>(E)) < h2min) {
    s.S103 = realizability_S310(s.S101, s.S011, s.S110, s.S003, s.S102, s.S012, s.S201, s.S111, s.S103, s.S202, H002, beta);
    S202b = s.S202; flag310 = true;
  }
  if (!f44 && !f55 && det2(lead4<4>(E)) < h2min) {
    s.S031 = realizability_S310(s.S011, s.S110, s.S101, s.S030, s.S021, s.S120, s.S012, s.S111, s.S031, s.S022, H020, beta);
    S022a = s.S022; flag310 = true;
  }
  if (!f55 && !f66 && det2(lead4<5>(E)) < h2min) {
    s.S013 = realizability_S310(s.S011, s.S101, s.S110, s.S003, s.S012, s.S102, s.S021, s.S111, s.S013, s.S022, H002, beta);
    S022b = s.S022; flag310 = true;
  }
  double S220_310r = S220_310, S022_310r = S022_310, S202_310r = S202_310;
  {
    const double r = jmax(S220a, S220b);
    if (s.S220 < r) { S220_310r = r; flag220 = 1; }
  }
  {
    const double r = jmax(S022a, S022b);
    if (s.S022 < r) { S022_310r = r; flag220 = 1; }
  }
  {
    const double r = jmax(S202a, S202b);
    if (s.S202 < r) { S202_310r = r; flag220 = 1; }
  }
  if (flag310) {
    delta2star(s, E);
    double m = det2(lead4<0>(E));
    m = jmin(m, det2(lead4<1>(E)));
    m = jmin(m, det2(lead4<2>(E)));
    m = jmin(m, det2(lead4<3>(E)));
    m = jmin(m, det2(lead4<4>(E)));
    m = jmin(m, det2(lead4<5>(E)));
    if (m < 0) flag220 = 1;
  }

  // 3x3 minors (:380-453): E is current here (recomputed above when anything changed)
  S220a = s.S220; S220b = s.S220; S202a = s.S202;
  double S202c = s.S202;
  S022b = s.S022;
  double S022c = s.S022;
  if (!f11 && !f22 && !f33 && !f55) {
    const Lead4 A = lead4<0>(E);
    if (det3(A) <= h2min || det3(lead4<1>(E)) <= h2min) {
      const double d23 = s.S211 - A.e23, d22 = s.S220 - A.e22, d33 = s.S202 - A.e33;
      if (A.e11 > diagmin) { S220a = d22 + A.e13 * A.e13 / A.e11; S202a = d33 + A.e12 * A.e12 / A.e11; }
      s.S211 = realizability_S211(A.e11, A.e22, A.e33, A.e12, A.e13, d23, s.S211, beta);
      flag220 = 1;
    }
  }
  if (!f22 && !f33 && !f44 && !f55) {
    const Lead4 A = lead4<2>(E);
    if (det3(A) < h2min || det3(lead4<4>(E)) < h2min) {
      const double d23 = s.S121 - A.e23, d22 = s.S220 - A.e22, d33 = s.S022 - A.e33;
      if (A.e11 > diagmin) { S220b = d22 + A.e13 * A.e13 / A.e11; S022b = d33 + A.e12 * A.e12 / A.e11; }
      s.S121 = realizability_S211(A.e11, A.e22, A.e33, A.e12, A.e13, d23, s.S121, beta);
      flag220 = 1;
    }
  }
  if (!f22 && !f33 && !f55 && !f66) {
    const Lead4 A = lead4<5>(E);
    if (det3(lead4<3>(E)) < h2min || det3(A) < h2min) {
      const double d23 = s.S112 - A.e23, d22 = s.S022 - A.e22, d33 = s.S202 - A.e33;
      if (A.e11 > diagmin) { S202c = d22 + A.e13 * A.e13 / A.e11; S022c = d33 + A.e12 * A.e12 / A.e11; }
      s.S112 = realizability_S211(A.e11, A.e22, A.e33, A.e12, A.e13, d23, s.S112, beta);
      flag220 = 1;
    }
  }
  const double S220_211 = jmax(S220a, S220b), S202_211 = jmax(S202a, S202c),
               S022_211 = jmax(S022b, S022c);

  // 4x4 minors (:455-608)
  delta2star(s, E);
  auto minor4 = [&](const Lead4& L, double Sxx, double& smin, double& smax) {
    smin = Sxx;
    smax = Sxx;
    if (det4(L) < 0) {
      flag220 = 1;
      fl |= F_R4;
      double c[4];
      cubic_x_y(L.e33, L.e11, L.e12, L.e13, L.e23, L.e24, L.e34, L.e44, -L.e22 + L.e14, c);
      const Roots3 R = cardano(c);
      if (roots_imag_ratio(R) < 1000 * EPS) {
        double r2, r3;
        roots_mid_hi(R, r2, r3);
        const double d22 = Sxx - L.e22;
        smin = r2 + d22;
        smax = r3 + d22;
      }
    }
  };
  double s220mina, s220maxa, s220minb, s220maxb, s202mina, s202maxa, s202minb, s202maxb,
      s022mina, s022maxa, s022minb, s022maxb;
  minor4(lead4<0>(E), s.S220, s220mina, s220maxa);
  minor4(lead4<2>(E), s.S220, s220minb, s220maxb);
  minor4(lead4<1>(E), s.S202, s202mina, s202maxa);
  minor4(lead4<3>(E), s.S202, s202minb, s202maxb);
  minor4(lead4<4>(E), s.S022, s022mina, s022maxa);
  minor4(lead4<5>(E), s.S022, s022minb, s022maxb);
  double v;
  v = jmax(jmax(jmax(jmax(jmax(s220mina, s220minb), S220_diag), S220_310), S220_310r), S220_211);
  s.S220 = jmin(jmin(jmin(v, S220max), s220maxa), s220maxb);
  v = jmax(jmax(jmax(jmax(jmax(s202mina, s202minb), S202_diag), S202_310), S202_310r), S202_211);
  s.S202 = jmin(jmin(jmin(v, S202max), s202maxa), s202maxb);
  v = jmax(jmax(jmax(jmax(jmax(s022mina, s022minb), S022_diag), S022_310), S022_310r), S022_211);
  s.S022 = jmin(jmin(jmin(v, S022max), s022maxa), s022maxb);
  return flag220;
}

}  // namespace hyq
