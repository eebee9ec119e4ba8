// Per-cell pipeline pieces: the 35-moment closure, flux-Jacobian wave speeds, the 1D
// HyQMOM closure, BGK relaxation.  FP64, one cell per lane.
#pragma once
#include "hyq_closure.h"

namespace hyq {

// hyqmom_3D (src/moments/hyqmom_3D.jl:10-53): one formula per family, instantiated along
// each axis / axis pair.
HD double f500(double s3, double s4) { return 0.5 * s3 * (5 * s4 - 3 * s3 * s3 - 1); }
HD double f410(double a3, double a4, double ab11, double ab21, double ab31) {
  return (a3 - 2 * a3 * a4 + 9 * a3 * a3 * a3 / 4) * ab11 +
         (5 * a4 / 2 - 15 * a3 * a3 / 4 - 1.5) * ab21 + 2 * a3 * ab31;
}
HD double f320(double a3, double a4, double b3, double ab11, double ab21, double ab12,
               double ab22, double ab31) {
  return (b3 * a3 * a3 / 2 - b3 + b3 * (a3 * a3 - a4 + 1)) * ab11 + (a4 - 3 * a3 * a3 / 2) * ab12 +
         (-(3 * b3 * a3) / 2) * ab21 + (3 * a3 / 2) * ab22 + b3 * ab31 - a3 / 2;
}
HD double f311(double a3, double a4, double a211, double a111, double bc11) {
  return 3 * (a3 * a211) / 2 + (2 * a4 - 3 * a3 * a3) * a111 / 2 - (a3 * bc11) / 2;
}
HD double f221(double a3, double b3, double b121, double a211, double ac11, double bc11,
               double a201, double b021, double a111) {
  return a3 * (b121 - ac11) + b3 * (a211 - bc11) + a201 + b021 - a3 * b3 * a111;
}

// Standardized moments (order <= 4) in the packed layout, from the 28-set.
HD void s28_to_packed(const S28& s, double* c) {
  c[tpos(0, 0, 0)] = 1.0;
  c[tpos(1, 0, 0)] = 0.0; c[tpos(0, 1, 0)] = 0.0; c[tpos(0, 0, 1)] = 0.0;
  c[tpos(2, 0, 0)] = 1.0; c[tpos(0, 2, 0)] = 1.0; c[tpos(0, 0, 2)] = 1.0;
  c[tpos(1, 1, 0)] = s.S110; c[tpos(1, 0, 1)] = s.S101; c[tpos(0, 1, 1)] = s.S011;
  c[tpos(3, 0, 0)] = s.S300; c[tpos(2, 1, 0)] = s.S210; c[tpos(2, 0, 1)] = s.S201;
  c[tpos(1, 2, 0)] = s.S120; c[tpos(1, 1, 1)] = s.S111; c[tpos(1, 0, 2)] = s.S102;
  c[tpos(0, 3, 0)] = s.S030; c[tpos(0, 2, 1)] = s.S021; c[tpos(0, 1, 2)] = s.S012;
  c[tpos(0, 0, 3)] = s.S003;
  c[tpos(4, 0, 0)] = s.S400; c[tpos(3, 1, 0)] = s.S310; c[tpos(3, 0, 1)] = s.S301;
  c[tpos(2, 2, 0)] = s.S220; c[tpos(2, 1, 1)] = s.S211; c[tpos(2, 0, 2)] = s.S202;
  c[tpos(1, 3, 0)] = s.S130; c[tpos(1, 2, 1)] = s.S121; c[tpos(1, 1, 2)] = s.S112;
  c[tpos(1, 0, 3)] = s.S103; c[tpos(0, 4, 0)] = s.S040; c[tpos(0, 3, 1)] = s.S031;
  c[tpos(0, 2, 2)] = s.S022; c[tpos(0, 1, 3)] = s.S013; c[tpos(0, 0, 4)] = s.S004;
}

// The closure's state between its two halves: the realized standardized 4th-order set, the
// floored variances and the density and means.
struct ClosureState {
  S28 s;
  double C200, C020, C002, M000, um, vm, wm;
};

// M2CS4_35 (src/moments/M2CS4_35.jl:47-49) standardization of the central moments c
// (packed, order <= 4) with sx, sy, sz = sqrt(max(C, eps)) of the three variances: the 28
// standardized moments the closure uses.
HD S28 standardize28(const double* c, double sx, double sy, double sz) {
  S28 s;
#define SD(i, j, k) (c[tpos(i, j, k)] / (pw<i>(sx) * pw<j>(sy) * pw<k>(sz)))
  s.S300 = SD(3, 0, 0); s.S400 = SD(4, 0, 0); s.S110 = SD(1, 1, 0); s.S210 = SD(2, 1, 0);
  s.S310 = SD(3, 1, 0); s.S120 = SD(1, 2, 0); s.S220 = SD(2, 2, 0); s.S030 = SD(0, 3, 0);
  s.S130 = SD(1, 3, 0); s.S040 = SD(0, 4, 0); s.S101 = SD(1, 0, 1); s.S201 = SD(2, 0, 1);
  s.S301 = SD(3, 0, 1); s.S102 = SD(1, 0, 2); s.S202 = SD(2, 0, 2); s.S003 = SD(0, 0, 3);
  s.S103 = SD(1, 0, 3); s.S004 = SD(0, 0, 4); s.S011 = SD(0, 1, 1); s.S111 = SD(1, 1, 1);
  s.S211 = SD(2, 1, 1); s.S021 = SD(0, 2, 1); s.S121 = SD(1, 2, 1); s.S031 = SD(0, 3, 1);
  s.S012 = SD(0, 1, 2); s.S112 = SD(1, 1, 2); s.S013 = SD(0, 1, 3); s.S022 = SD(0, 2, 2);
#undef SD
  return s;
}

// Flux_closure35_and_realizable_3D (src/numerics/Flux_closure35_and_realizable_3D.jl:33-234),
// first half: standardization and the realizability projection (:33-176).
HD void closure_project(const double* M4, int flag2D, double Ma, ClosureState& q, uint32_t& fl) {
  const double s3max = 4.0 + fabs(Ma) / 2.0, h2min = 1.0e-8;
  double c[35];
  double M000, um, vm, wm;
  raw_to_central4(M4, c, M000, um, vm, wm);
  const double C200 = jmax(EPS, c[tpos(2, 0, 0)]), C020 = jmax(EPS, c[tpos(0, 2, 0)]),
               C002 = jmax(EPS, c[tpos(0, 0, 2)]);
  const double sx = sqrt(jmax(c[tpos(2, 0, 0)], EPS)), sy = sqrt(jmax(c[tpos(0, 2, 0)], EPS)),
               sz = sqrt(jmax(c[tpos(0, 0, 2)], EPS));
  S28 s = standardize28(c, sx, sy, sz);
  double H200, H020, H002;
  {
    const double o[6] = {s.S300, s.S400, s.S030, s.S040, s.S003, s.S004};
    enforce_univariate(s.S300, s.S400, H200, h2min, s3max);
    enforce_univariate(s.S030, s.S040, H020, h2min, s3max);
    enforce_univariate(s.S003, s.S004, H002, h2min, s3max);
    if (o[0] != s.S300 || o[1] != s.S400 || o[2] != s.S030 || o[3] != s.S040 ||
        o[4] != s.S003 || o[5] != s.S004)
      fl |= F_UNIV;
  }
  s.S220 = clamp_S220(s.S110, s.S220, sqrt((H200 + s.S300 * s.S300) * (H020 + s.S030 * s.S030)));
  s.S202 = clamp_S220(s.S101, s.S202, sqrt((H200 + s.S300 * s.S300) * (H002 + s.S003 * s.S003)));
  s.S022 = clamp_S220(s.S011, s.S022, sqrt((H020 + s.S030 * s.S030) * (H002 + s.S003 * s.S003)));

  const double R110 = 1 - s.S110 * s.S110, R101 = 1 - s.S101 * s.S101, R011 = 1 - s.S011 * s.S011;
  if (R110 <= 0 || R101 <= 0 || R011 <= 0) {
    // edge/corner branch (:107-124) with check2D of the three planes (:83-88)
    fl |= F_EDGE;
    const P10 p1 = check2D(P10{s.S300, s.S400, s.S110, s.S210, s.S310, s.S120, s.S220, s.S030, s.S130, s.S040}, fl);
    const P10 p2 = check2D(P10{s.S300, s.S400, s.S101, s.S201, s.S301, s.S102, s.S202, s.S003, s.S103, s.S004}, fl);
    const P10 p3 = check2D(P10{s.S030, s.S040, s.S011, s.S021, s.S031, s.S012, s.S022, s.S003, s.S013, s.S004}, fl);
    edge_corner(R110, R101, R011, s, p1, p2, p3);
    // :107-109: realizable_2D returns (S21, S12, S31, S22, S13), which the xz and yz calls
    // destructure into (S201, S301, S102, S202, S103) and (S021, S031, S012, S022, S013):
    // the corrected S102 lands in S301 and S301's in S102 (likewise S012 and S031)
    realizable_2D(s.S300, s.S400, s.S110, s.S210, s.S310, s.S120, s.S220, s.S030, s.S130, s.S040, fl);
    realizable_2D(s.S300, s.S400, s.S101, s.S201, s.S301, s.S102, s.S202, s.S003, s.S103, s.S004, fl);
    { const double t = s.S301; s.S301 = s.S102; s.S102 = t; }
    realizable_2D(s.S030, s.S040, s.S011, s.S021, s.S031, s.S012, s.S022, s.S003, s.S013, s.S004, fl);
    { const double t = s.S031; s.S031 = s.S012; s.S012 = t; }
  } else {
    const int f220 = realizable_3D(s, fl);
    if (f220) {
      fl |= F_220;
      for (int it = 0; it < 6; ++it) realizable_3D(s, fl);
    }
  }
  if (flag2D == 1) {
    s.S011 = 0.0; s.S101 = 0.0; s.S201 = 0.0; s.S021 = 0.0; s.S012 = 0.0; s.S102 = 0.0;
    s.S111 = 0.0; s.S202 = 1.0; s.S022 = 1.0; s.S112 = s.S110; s.S121 = 0.0; s.S211 = 0.0;
    s.S103 = 0.0; s.S013 = 0.0; s.S301 = 0.0; s.S031 = 0.0;
  }
  q.s = s;
  q.C200 = C200; q.C020 = C020; q.C002 = C002;
  q.M000 = M000; q.um = um; q.vm = vm; q.wm = wm;
}

// Second half: the 5th-order closure, S -> C (S_to_C_batch, moment_array_utils.jl:137-204),
// C -> M.  With FLUX, Fx/Fy/Fz (35 each) are written through the pointers with stride `st`.
// hyqmom_3D (src/autogen/hyqmom_3D.jl): the 21 standardized 5th-order moments of the
// closure from the 28 standardized moments s, into the packed array c.
HD void hyqmom5_std(const S28& s, double* c) {
  c[tpos(5, 0, 0)] = f500(s.S300, s.S400);
  c[tpos(0, 5, 0)] = f500(s.S030, s.S040);
  c[tpos(0, 0, 5)] = f500(s.S003, s.S004);
  c[tpos(3, 2, 0)] = f320(s.S300, s.S400, s.S030, s.S110, s.S210, s.S120, s.S220, s.S310);
  c[tpos(2, 3, 0)] = f320(s.S030, s.S040, s.S300, s.S110, s.S120, s.S210, s.S220, s.S130);
  c[tpos(3, 0, 2)] = f320(s.S300, s.S400, s.S003, s.S101, s.S201, s.S102, s.S202, s.S301);
  c[tpos(2, 0, 3)] = f320(s.S003, s.S004, s.S300, s.S101, s.S102, s.S201, s.S202, s.S103);
  c[tpos(0, 3, 2)] = f320(s.S030, s.S040, s.S003, s.S011, s.S021, s.S012, s.S022, s.S031);
  c[tpos(0, 2, 3)] = f320(s.S003, s.S004, s.S030, s.S011, s.S012, s.S021, s.S022, s.S013);
  c[tpos(4, 1, 0)] = f410(s.S300, s.S400, s.S110, s.S210, s.S310);
  c[tpos(1, 4, 0)] = f410(s.S030, s.S040, s.S110, s.S120, s.S130);
  c[tpos(4, 0, 1)] = f410(s.S300, s.S400, s.S101, s.S201, s.S301);
  c[tpos(1, 0, 4)] = f410(s.S003, s.S004, s.S101, s.S102, s.S103);
  c[tpos(0, 4, 1)] = f410(s.S030, s.S040, s.S011, s.S021, s.S031);
  c[tpos(0, 1, 4)] = f410(s.S003, s.S004, s.S011, s.S012, s.S013);
  c[tpos(3, 1, 1)] = f311(s.S300, s.S400, s.S211, s.S111, s.S011);
  c[tpos(1, 3, 1)] = f311(s.S030, s.S040, s.S121, s.S111, s.S101);
  c[tpos(1, 1, 3)] = f311(s.S003, s.S004, s.S112, s.S111, s.S110);
  c[tpos(2, 2, 1)] = f221(s.S300, s.S030, s.S121, s.S211, s.S101, s.S011, s.S201, s.S021, s.S111);
  c[tpos(2, 1, 2)] = f221(s.S300, s.S003, s.S112, s.S211, s.S110, s.S011, s.S210, s.S012, s.S111);
  c[tpos(1, 2, 2)] = f221(s.S003, s.S030, s.S121, s.S112, s.S101, s.S110, s.S102, s.S120, s.S111);
}

template <bool FLUX, typename FT = double>
HD void closure_emit(const ClosureState& q, double* Mr, FT* Fx, FT* Fy, FT* Fz, size_t st) {
  const S28& s = q.s;
  const double C200 = q.C200, C020 = q.C020, C002 = q.C002;
  double M000 = q.M000, um = q.um, vm = q.vm, wm = q.wm;
  double c[56];
  s28_to_packed(s, c);
  hyqmom5_std(s, c);
  const double qx = sqrt(jmax(EPS, C200)), qy = sqrt(jmax(EPS, C020)), qz = sqrt(jmax(EPS, C002));
  double px[6], py[6], pz[6];
  px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
  for (int t = 1; t < 6; ++t) { px[t] = px[t - 1] * qx; py[t] = py[t - 1] * qy; pz[t] = pz[t - 1] * qz; }
#pragma unroll
  for (int i = 0; i <= 5; ++i)
#pragma unroll
    for (int j = 0; i + j <= 5; ++j)
#pragma unroll
      for (int k = 0; i + j + k <= 5; ++k) c[tpos(i, j, k)] *= px[i] * py[j] * pz[k];
  c[tpos(2, 0, 0)] = C200;
  c[tpos(0, 2, 0)] = C020;
  c[tpos(0, 0, 2)] = C002;
  if (FLUX) {
    central_to_raw<5>(c, M000, um, vm, wm);
#pragma unroll
    for (int n = 0; n < 35; ++n) {
      constexpr E35T E = e35();
      const int i = E.e[n][0], j = E.e[n][1], k = E.e[n][2];
      Mr[n] = c[tpos(i, j, k)];
      HYQ_STORE_STREAM(&Fx[n * st], (FT)c[tpos(i + 1, j, k)]);
      HYQ_STORE_STREAM(&Fy[n * st], (FT)c[tpos(i, j + 1, k)]);
      HYQ_STORE_STREAM(&Fz[n * st], (FT)c[tpos(i, j, k + 1)]);
    }
  } else {
    central_to_raw<4>(c, M000, um, vm, wm);
#pragma unroll
    for (int n = 0; n < 35; ++n) Mr[n] = c[T35(n)];
  }
}

template <bool FLUX, typename FT = double>
HD void flux_closure35(const double* M4, int flag2D, double Ma, double* Mr, FT* Fx, FT* Fy,
                       FT* Fz, size_t st, uint32_t& fl) {
  ClosureState q;
  closure_project(M4, flag2D, Ma, q, fl);
  closure_emit<FLUX, FT>(q, Mr, Fx, Fy, Fz, st);
}

// Moments5_3D (src/moments/Moments5_3D.jl:16-138): the HyQMOM closure of the 5th-order
// moments of M4 without any realizability projection.  S5: hyqmom_3D of M4's standardized
// moments; C5 = S5 sx^i sy^j sz^k; M5: raw moments of order 5 from C5 and M4's own central
// moments (C5toM5_3D).  Outputs in the reference's order 500 410 320 230 140 401 302 203 104
// 311 221 131 212 113 122 050 041 032 023 014 005, each with stride st.
HD double ipow5(double x, int n) {  // x^n, 0 <= n <= 5
  double r = 1.0;
#pragma unroll
  for (int t = 0; t < 5; ++t)
    if (t < n) r *= x;
  return r;
}
HD void moments5_cell(const double* M4, double* M5, double* C5, double* S5, size_t st) {
  constexpr int E21[21][3] = {{5, 0, 0}, {4, 1, 0}, {3, 2, 0}, {2, 3, 0}, {1, 4, 0}, {4, 0, 1},
                              {3, 0, 2}, {2, 0, 3}, {1, 0, 4}, {3, 1, 1}, {2, 2, 1}, {1, 3, 1},
                              {2, 1, 2}, {1, 1, 3}, {1, 2, 2}, {0, 5, 0}, {0, 4, 1}, {0, 3, 2},
                              {0, 2, 3}, {0, 1, 4}, {0, 0, 5}};
  double c[56];
  double M000, um, vm, wm;
  raw_to_central4(M4, c, M000, um, vm, wm);
  const double sx = sqrt(jmax(c[tpos(2, 0, 0)], EPS)), sy = sqrt(jmax(c[tpos(0, 2, 0)], EPS)),
               sz = sqrt(jmax(c[tpos(0, 0, 2)], EPS));
  hyqmom5_std(standardize28(c, sx, sy, sz), c);
#pragma unroll
  for (int n = 0; n < 21; ++n) {
    const int i = E21[n][0], j = E21[n][1], k = E21[n][2];
    const double sv = c[tpos(i, j, k)];
    const double cv = sv * (ipow5(sx, i) * ipow5(sy, j) * ipow5(sz, k));
    S5[n * st] = sv;
    C5[n * st] = cv;
    c[tpos(i, j, k)] = cv;
  }
  central_to_raw<5>(c, M000, um, vm, wm);
#pragma unroll
  for (int n = 0; n < 21; ++n) M5[n * st] = c[tpos(E21[n][0], E21[n][1], E21[n][2])];
}

}  // namespace hyq
