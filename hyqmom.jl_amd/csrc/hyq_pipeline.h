// The per-cell passes of one time step (simulation_runner.jl:352-380 and :526-556).
#pragma once
#include "hyq_fasteig.h"

namespace hyq {

// correct_{uv,uw,vw}_plane (eigenvalues6_hyperbolic_3D.jl:181-252, and the inlined copies
// in eigenvalues6z_hyperbolic_3D.jl:225-283) on packed central moments c0 (kept) -> c.
// Plane (P, Q): cross moments of orders 3 and 4 rebuilt from the plane correlation.
template <int P, int Q>
HD void correct_plane(const double* c0, double* c) {
  auto id = [](int i, int j) {
    int e[3] = {0, 0, 0};
    e[P] += i;
    e[Q] += j;
    return tpos(e[0], e[1], e[2]);
  };
  const double sp = sqrt(jmax(EPS, c0[id(2, 0)])), sq = sqrt(jmax(EPS, c0[id(0, 2)]));
  // standardized values as M2CS4_35 computes them: sqrt(max(C, eps))
  const double Spq = c0[id(1, 1)] / (sp * sq), Sp3 = c0[id(3, 0)] / (sp * sp * sp),
               Sq3 = c0[id(0, 3)] / (sq * sq * sq), Sp4 = c0[id(4, 0)] / (sp * sp * sp * sp),
               Sq4 = c0[id(0, 4)] / (sq * sq * sq * sq);
  double S22 = c0[id(2, 2)] / (sp * sp * sq * sq);
  const double s22min = (2 + 5 * Sq3 * Spq * Sp3) / 6;
  if (S22 < s22min) S22 = s22min;
  c[id(2, 1)] = (Spq * Sp3) * sp * sp * sq;
  c[id(1, 2)] = (Spq * Sq3) * sp * sq * sq;
  c[id(3, 1)] = (Spq * Sp4) * sp * sp * sp * sq;
  c[id(1, 3)] = (Spq * Sq4) * sp * sq * sq * sq;
  c[id(2, 2)] = S22 * sp * sp * sq * sq;
}

HD void pack35(const double* c, double* M) {
#pragma unroll
  for (int n = 0; n < 35; ++n) M[n] = c[T35(n)];
}

// Largest deviation of a plane's standardized moments from a Gaussian's (S30 = S03 = S21 =
// S12 = 0, S40 = S04 = 3, S31 = S13 = 3 S11, S22 = 1 + 2 S11^2).  Below kGaussTol the
// Gaussian's extreme eigenvalues seed the fast solve (the seed is still verified by the
// Sturm count).
constexpr double kGaussTol = 1e-3;
// The learned mode hands a plane to the network only when its standardized moments are
// at least this far from a Gaussian's; nearer ones take the Gaussian's seeds (verified by
// the Sturm count like any seed): a seed from the network saves at most about one Laguerre
// iteration per end there, less than the network's own cost.
constexpr double kNetTol = 0.1;
HD double plane_gauss_dev(const Plane& p) {
  double d = fmax(fmax(fabs(p.S30), fabs(p.S03)), fmax(fabs(p.S21), fabs(p.S12)));
  d = fmax(d, fmax(fabs(p.S40 - 3.0), fabs(p.S04 - 3.0)));
  d = fmax(d, fmax(fabs(p.S31 - 3.0 * p.S11), fabs(p.S13 - 3.0 * p.S11)));
  return fmax(d, fabs(p.S22 - 1.0 - 2.0 * p.S11 * p.S11));
}
HD bool plane_near_gaussian(const Plane& p) { return p.ok && plane_gauss_dev(p) < kGaussTol; }

// Eigen solve of one plane in mode EM: 0 = full FP64 solve (QR); 1 = the fast path from
// warm starts (wmin/wmax in standardized units, NaN = none), falling back to the full solve
// when hyperbolicity is not proven or the iteration does not converge.
// One plane by the fast path: Gaussian seed (when no warm start is given and the plane is
// near-Gaussian), Sturm test (with the 2x2 block test of a nearly double pair), Laguerre
// extremes.  Returns kPlaneProven with vmin/vmax set, kPlaneSolve when the plane is not
// proven hyperbolic (the full solve of this plane decides), kPlaneComplex when a complex
// pair above the reference's test is proven (the eigenvalue correction), or kPlaneFull
// when the plane or its Jacobian is not finite (the full path carries the reference's NaN
// semantics).
constexpr int kPlaneProven = 0, kPlaneSolve = 1, kPlaneFull = 2, kPlaneComplex = 3;
template <bool PAIR = true>
HD int plane_fast_status(const Plane& p, float wmin, float wmax, bool need_vals, uint32_t& fl,
                         double& vmin, double& vmax) {
  if (!p.ok) return kPlaneFull;
  if (wmin != wmin && plane_near_gaussian(p)) {
    // a Gaussian plane's standardized extreme eigenvalues are -sqrt(3), +sqrt(3)
    wmin = -1.7320508075688772f;
    wmax = 1.7320508075688772f;
  }
  double J[36];
  jacobian_std(p, J);
  bool fin = true;
#pragma unroll
  for (int t = 0; t < 36; ++t) fin = fin && (fabs(J[t]) < 1e300);
  if (!fin) return kPlaneFull;
  const FastEig fe = fast_eigs_std<PAIR>(J, (double)wmin, (double)wmax, need_vals, p.sig);
  // root-polishing iterations, summed over the cell's plane solves, in flag bits 16..27
  const uint32_t it = (fl >> 16) + (uint32_t)fe.iters;
  fl = (fl & 0xFFFFu) | ((it < 0xFFFu ? it : 0xFFFu) << 16);
  if (fe.cplx) return kPlaneComplex;
  if (!fe.real) return kPlaneSolve;
  vmin = need_vals ? p.a + p.sig * fe.lmin : 0.0;
  vmax = need_vals ? p.a + p.sig * fe.lmax : 0.0;
  return kPlaneProven;
}
HD bool plane_fast(const Plane& p, float wmin, float wmax, bool need_vals, uint32_t& fl,
                   double& vmin, double& vmax) {
  return plane_fast_status(p, wmin, wmax, need_vals, fl, vmin, vmax) == kPlaneProven;
}

// CORRECTS: a complex pair leads to the eigenvalue correction, which replaces the speeds, so
// a proven complex pair needs no values; after the correction (CORRECTS = false) the
// reference takes the real parts of whatever the eigenvalues are, so the QR gives them.
template <int EM, bool CORRECTS = true>
HD EigRes plane_solve(const Plane& p, float wmin, float wmax, bool need_vals, uint32_t& fl) {
  if (EM == 0) return plane_eigs(p, fl);
  EigRes r;
  const int st = plane_fast_status<true>(p, wmin, wmax, need_vals, fl, r.vmin, r.vmax);
  if (st == kPlaneProven) {
    r.cplx = false;
    r.ok = true;
    return r;
  }
  if (CORRECTS && st == kPlaneComplex) {
    r.vmin = r.vmax = NAN;
    r.cplx = true;
    r.ok = true;
    return r;
  }
  fl |= F_FALLBACK;
  return plane_eigs(p, fl);
}
// Warm starts of the five distinct plane Jacobians of a cell (UV, UW, VU, VW, WU), each
// (min, max); NaN when absent.
struct Warm {
  float w[10];
};
HD Warm no_warm() {
  Warm w;
#pragma unroll
  for (int t = 0; t < 10; ++t) w.w[t] = NAN;
  return w;
}

// Minimum/maximum over two Jacobians with compute_jacobian_eigenvalues.jl's NaN guard: a
// non-finite first Jacobian makes both NaN; a non-finite second one is skipped.
HD void combine_guarded(const EigRes& A, const EigRes& B, double& vmin, double& vmax) {
  vmin = A.ok ? A.vmin : NAN;
  vmax = A.ok ? A.vmax : NAN;
  if (B.ok) {
    vmin = jmin(vmin, B.vmin);
    vmax = jmax(vmax, B.vmax);
  }
}

// The eigenvalue correction of one axis (eigenvalues6_hyperbolic_3D.jl:35-60): the planes
// with a complex pair (ca: UV or VU, cb: UW or VW) corrected in the central moments c of the
// state, which returns to raw moments and is projected again into Mr.
template <int AXIS>
HD void axis_correct(const double* c, double M000, double um, double vm, double wm, bool ca,
                     bool cb, int flag2D, double Ma, double* Mr, uint32_t& fl) {
  fl |= AXIS == 0 ? F_ECX : F_ECY;
  double cn[35];
#pragma unroll
  for (int t = 0; t < 35; ++t) cn[t] = c[t];
  if (ca) correct_plane<0, 1>(c, cn);
  if (cb) {
    if (AXIS == 0) correct_plane<0, 2>(c, cn);
    else correct_plane<1, 2>(c, cn);
  }
  central_to_raw<4>(cn, M000, um, vm, wm);
  double Mh[35];
  pack35(cn, Mh);
  flux_closure35<false, double>(Mh, flag2D, Ma, Mr, nullptr, nullptr, nullptr, 0, fl);
}

// eigenvalues6_hyperbolic_3D (src/numerics/eigenvalues6_hyperbolic_3D.jl:22-155).
// AXIS 0: UV and UW Jacobians; AXIS 1: VU and VW.  In: M.  Out: Mr (may alias nothing).
// need_vals = false (pass 2, whose speeds are not used): after a correction the speeds of
// the corrected state are not computed.  Returns true when the state was corrected.
template <int AXIS, int EM = 0>
HD bool eig6_axis(const double* M, int flag2D, double Ma, double& vmin, double& vmax,
                  double* Mr, uint32_t& fl, const Warm& wm_ = no_warm(), bool need_vals = true) {
  double c[35];
  double M000, um, vm, wm;
  raw_to_central4(M, c, M000, um, vm, wm);
  EigRes A, B;
  const int ia = AXIS == 0 ? 0 : 2, ib = AXIS == 0 ? 1 : 3;  // Warm slots: UV UW VU VW WU
  if (AXIS == 0) {
    A = plane_solve<EM>(plane<0, 1>(c, um, vm, wm), wm_.w[2 * ia], wm_.w[2 * ia + 1], need_vals, fl);
    B = plane_solve<EM>(plane<0, 2>(c, um, vm, wm), wm_.w[2 * ib], wm_.w[2 * ib + 1], need_vals, fl);
  } else {
    A = plane_solve<EM>(plane<1, 0>(c, um, vm, wm), wm_.w[2 * ia], wm_.w[2 * ia + 1], need_vals, fl);
    B = plane_solve<EM>(plane<1, 2>(c, um, vm, wm), wm_.w[2 * ib], wm_.w[2 * ib + 1], need_vals, fl);
  }
  combine_guarded(A, B, vmin, vmax);
  if (A.cplx || B.cplx) {
    axis_correct<AXIS>(c, M000, um, vm, wm, A.cplx, B.cplx, flag2D, Ma, Mr, fl);
    if (!need_vals) return true;
    raw_to_central4(Mr, c, M000, um, vm, wm);
    if (AXIS == 0) {
      A = plane_solve<EM, false>(plane<0, 1>(c, um, vm, wm), wm_.w[2 * ia], wm_.w[2 * ia + 1], need_vals, fl);
      B = plane_solve<EM, false>(plane<0, 2>(c, um, vm, wm), wm_.w[2 * ib], wm_.w[2 * ib + 1], need_vals, fl);
    } else {
      A = plane_solve<EM, false>(plane<1, 0>(c, um, vm, wm), wm_.w[2 * ia], wm_.w[2 * ia + 1], need_vals, fl);
      B = plane_solve<EM, false>(plane<1, 2>(c, um, vm, wm), wm_.w[2 * ib], wm_.w[2 * ib + 1], need_vals, fl);
    }
    combine_guarded(A, B, vmin, vmax);
    return true;
  }
#pragma unroll
  for (int n = 0; n < 35; ++n) Mr[n] = M[n];
  return false;
}

// eigenvalues6z_hyperbolic_3D (src/numerics/eigenvalues6z_hyperbolic_3D.jl:23-365): the WU
// Jacobian (flux along z, second variable x) and, as the reference maps it, the VW one.
template <int EM = 0>
HD void eig6_z(const double* M, int flag2D, double Ma, bool positivity, double& vmin,
               double& vmax, double* Mr, uint32_t& fl, const Warm& wm_ = no_warm()) {
#pragma unroll
  for (int n = 0; n < 35; ++n) Mr[n] = M[n];
  if (positivity && M[0] <= 0.0) {
    Mr[0] = jmax(M[0], 1e-14);
    fl |= F_RHO;
  }
  double W[35];
#pragma unroll
  for (int n = 0; n < 35; ++n) W[n] = M[n];
  W[0] = Mr[0];
  if (positivity && W[0] > 1e-15) {
    const double m0 = W[0];
    const double vx = M[2] / m0 - (M[1] / m0) * (M[1] / m0);
    const double vy = M[9] / m0 - (M[5] / m0) * (M[5] / m0);
    const double vz = M[19] / m0 - (M[15] / m0) * (M[15] / m0);
    if (vx < 0.0 || vy < 0.0 || vz < 0.0) {
      fl |= F_ZNEG;
      flux_closure35<false, double>(M, flag2D, Ma, Mr, nullptr, nullptr, nullptr, 0, fl);
#pragma unroll
      for (int n = 0; n < 35; ++n) W[n] = Mr[n];
    }
  }
  double c[35];
  double M000, um, vm, wm;
  raw_to_central4(W, c, M000, um, vm, wm);
  EigRes A = plane_solve<EM>(plane<2, 0>(c, um, vm, wm), wm_.w[8], wm_.w[9], true, fl);
  EigRes B = plane_solve<EM>(plane<1, 2>(c, um, vm, wm), wm_.w[6], wm_.w[7], true, fl);
  if (!A.ok || !B.ok) fl |= F_NONFIN;  // the reference's eigvals throws here
  vmin = jmin(A.vmin, B.vmin);
  vmax = jmax(A.vmax, B.vmax);
  if (A.cplx || B.cplx) {
    fl |= F_ECZ;
    raw_to_central4(M, c, M000, um, vm, wm);  // corrections start from the input M
    double cn[35];
#pragma unroll
    for (int t = 0; t < 35; ++t) cn[t] = c[t];
    if (A.cplx) correct_plane<0, 2>(c, cn);
    if (B.cplx) correct_plane<1, 2>(c, cn);
    central_to_raw<4>(cn, M000, um, vm, wm);
    double Mh[35];
    pack35(cn, Mh);
    flux_closure35<false, double>(Mh, flag2D, Ma, Mr, nullptr, nullptr, nullptr, 0, fl);
    raw_to_central4(Mr, c, M000, um, vm, wm);
    A = plane_solve<EM, false>(plane<2, 0>(c, um, vm, wm), wm_.w[8], wm_.w[9], true, fl);
    B = plane_solve<EM, false>(plane<1, 2>(c, um, vm, wm), wm_.w[6], wm_.w[7], true, fl);
    vmin = jmin(A.vmin, B.vmin);
    vmax = jmax(A.vmax, B.vmax);
  }
}

// Pass 1 of the step for one cell (simulation_runner.jl:356-379).  Outputs the projected
// state Mr, the fluxes (strided), and the six wave speeds.
// First half of pass 1: the projection, returned in A (the state the wave speeds see).
HD void pass1a_cell(const double* M, int flag2D, double Ma, double* A, uint32_t& fl) {
  flux_closure35<false, double>(M, flag2D, Ma, A, nullptr, nullptr, nullptr, 0, fl);
}
// Second half of pass 1 from the projected state A.
template <typename FT = double, int EM = 0>
HD void pass1b_cell(double* A, int flag2D, double Ma, bool positivity, double* Mr, FT* Fx,
                    FT* Fy, FT* Fz, size_t st, double vp[6], uint32_t& fl, const Warm& w) {
  double B[35];
  double xmn, xmx, ymn, ymx, zmn, zmx;
  eig6_axis<0, EM>(A, flag2D, Ma, xmn, xmx, B, fl, w);
  eig6_axis<1, EM>(B, flag2D, Ma, ymn, ymx, A, fl, w);
  eig6_z<EM>(A, flag2D, Ma, positivity, zmn, zmx, B, fl, w);
  flux_closure35<true, FT>(B, flag2D, Ma, Mr, Fx, Fy, Fz, st, fl);
  double a, b;
  closure1d_speeds(Mr[0], Mr[1], Mr[2], Mr[3], Mr[4], a, b);
  vp[0] = jmin(a, xmn);
  vp[1] = jmax(b, xmx);
  closure1d_speeds(Mr[0], Mr[5], Mr[9], Mr[12], Mr[14], a, b);
  vp[2] = jmin(a, ymn);
  vp[3] = jmax(b, ymx);
  closure1d_speeds(Mr[0], Mr[15], Mr[19], Mr[22], Mr[24], a, b);
  vp[4] = jmin(a, zmn);
  vp[5] = jmax(b, zmx);
}
template <typename FT = double, int EM = 0>
HD void pass1_cell(const double* M, int flag2D, double Ma, bool positivity, double* Mr, FT* Fx,
                   FT* Fy, FT* Fz, size_t st, double vp[6], uint32_t& fl, const Warm& w = no_warm()) {
  double A[35];
  pass1a_cell(M, flag2D, Ma, A, fl);
  pass1b_cell<FT, EM>(A, flag2D, Ma, positivity, Mr, Fx, Fy, Fz, st, vp, fl, w);
}

// collision35 (src/numerics/collision35.jl:17-45): relaxation toward the isotropic
// Maxwellian with the same density, velocity and trace temperature.
template <bool MIXED = false>
HD void collision_cell(const double* M, double dt, double Kn, bool positivity, double* Mo) {
  const double rho = M[0], r = 1.0 / rho;
  const double um = M[1] / rho, vm = M[5] / rho, wm = M[15] / rho;
  const double C200 = M[2] / rho - um * um, C020 = M[9] / rho - vm * vm,
               C002 = M[19] / rho - wm * wm;
  (void)r;
  double Th = (C200 + C020 + C002) / 3;
  if (positivity) Th = jmax(Th, 1e-14);
  const double tc = Kn / (rho * sqrt(Th) * 2);
  double g[35];
#pragma unroll
  for (int t = 0; t < 35; ++t) g[t] = 0.0;
  g[tpos(2, 0, 0)] = Th; g[tpos(0, 2, 0)] = Th; g[tpos(0, 0, 2)] = Th;
  g[tpos(4, 0, 0)] = 3 * Th * Th; g[tpos(0, 4, 0)] = 3 * Th * Th; g[tpos(0, 0, 4)] = 3 * Th * Th;
  g[tpos(2, 2, 0)] = Th * Th; g[tpos(2, 0, 2)] = Th * Th; g[tpos(0, 2, 2)] = Th * Th;
  central_to_raw<4>(g, rho, um, vm, wm);
  if (MIXED) {
    // relaxation in FP32: M + (1 - e) (MG - M), the difference formed in FP64
    const float a = 1.0f - expf((float)(-dt / tc));
#pragma unroll
    for (int n = 0; n < 35; ++n) Mo[n] = M[n] + (double)(a * (float)(g[T35(n)] - M[n]));
  } else {
    const double e = exp(-dt / tc);
#pragma unroll
    for (int n = 0; n < 35; ++n) {
      const double MG = g[T35(n)];
      Mo[n] = MG - e * (MG - M[n]);
    }
  }
}

// Pass 2 for one cell (simulation_runner.jl:526-556): realizability, x and y wave-speed
// corrections, realizability, then BGK.
template <bool MIXED = false, int EM = 0>
HD void pass2_cell(const double* M, int flag2D, double Ma, bool positivity, double dt,
                   double Kn, double* Mo, uint32_t& fl) {
  double A[35], B[35];
  double d0, d1;
  flux_closure35<false, double>(M, flag2D, Ma, A, nullptr, nullptr, nullptr, 0, fl);
  // the wave speeds of pass 2 are not used (only the complex-eigenvalue corrections are)
  eig6_axis<0, EM>(A, flag2D, Ma, d0, d1, B, fl, no_warm(), false);
  eig6_axis<1, EM>(B, flag2D, Ma, d0, d1, A, fl, no_warm(), false);
  flux_closure35<false, double>(A, flag2D, Ma, B, nullptr, nullptr, nullptr, 0, fl);
  collision_cell<MIXED>(B, dt, Kn, positivity, Mo);
}

// ---- the fast step with deferred cells (mixed_newton / mixed_learned modes) -----------------
// Most cells of a step are hyperbolic on every plane with no density clamp or
// re-projection on the z path: for them pass 1 and pass 2 reduce to closures and the five
// fast plane solves.  The rest (a QR fallback, a complex-eigenvalue correction, the z
// path's clamp or re-projection) are deferred to a second kernel that runs the full path
// on a compacted list, so the rare, register-hungry branches neither widen the common
// kernel nor hold up its waves.

// The five distinct plane Jacobians of a cell, in this order (the z direction's second plane
// is the y direction's VW, as eigenvalues6z_hyperbolic_3D.jl:117-120 maps it).
constexpr int kUV = 0, kUW = 1, kVU = 2, kVW = 3, kWU = 4;
HD Plane cell_plane(const double* c, double um, double vm, double wm, int q) {
  switch (q) {
    case kUV: return plane<0, 1>(c, um, vm, wm);
    case kUW: return plane<0, 2>(c, um, vm, wm);
    case kVU: return plane<1, 0>(c, um, vm, wm);
    case kVW: return plane<1, 2>(c, um, vm, wm);
    default: return plane<2, 0>(c, um, vm, wm);
  }
}

// Decisions of a cell's x/y planes kUV, kUW, kVU, kVW, 2 bits per plane q at bit 2q: the
// plane is hyperbolic, has a complex pair (the eigenvalue correction), or is not decided.
constexpr unsigned kKnownNone = 0u, kKnownReal = 1u, kKnownComplex = 2u;
HD unsigned known_of(unsigned known, int q) { return (known >> (2 * q)) & 3u; }

// The fast step's view of a cell's planes in pass 1 (need_vals, five planes) or pass 2 (the
// x/y checks, four planes).  Returns kPlaneFull when the cell must take the full path
// (z density clamp or re-projection, eigenvalues6z_hyperbolic_3D.jl:27-104, a complex pair
// or a non-finite plane); otherwise the planes the fast solve proved have their (vmin, vmax)
// in v[2q], v[2q + 1], and bit q of `pend` marks each plane that needs its own full solve.
// Pass 2 (VALS = false) decides all four planes and returns the decisions in `known`; pass 1
// stops at the first plane that sends the cell to the full path.
template <bool PAIR = true, bool VALS = true>
HD int fast_cell_planes(const double* A, bool positivity, const Warm& w, double v[10],
                        unsigned& pend, uint32_t& fl, unsigned* known = nullptr) {
  pend = 0;
  unsigned kn = 0u;
  if (VALS) {
    if (positivity && !(A[0] > 0.0)) return kPlaneFull;
    if (positivity && A[0] > 1e-15) {
      const double m0 = A[0];
      const double vx = A[2] / m0 - (A[1] / m0) * (A[1] / m0);
      const double vy = A[9] / m0 - (A[5] / m0) * (A[5] / m0);
      const double vz = A[19] / m0 - (A[15] / m0) * (A[15] / m0);
      if (vx < 0.0 || vy < 0.0 || vz < 0.0) return kPlaneFull;
    }
  }
  double c[35], M000, um, vm, wm;
  raw_to_central4(A, c, M000, um, vm, wm);
  const float nan = NAN;
  bool full = false;
#pragma unroll
  for (int q = 0; q < (VALS ? 5 : 4); ++q) {
    v[2 * q] = v[2 * q + 1] = 0.0;
    if (VALS && full) continue;
    const int st = plane_fast_status<PAIR>(cell_plane(c, um, vm, wm, q), VALS ? w.w[2 * q] : nan,
                                           VALS ? w.w[2 * q + 1] : nan, VALS, fl, v[2 * q],
                                           v[2 * q + 1]);
    if (st == kPlaneFull || st == kPlaneComplex) full = true;
    if (st == kPlaneSolve) pend |= 1u << q;
    kn |= (st == kPlaneProven ? kKnownReal : (st == kPlaneComplex ? kKnownComplex : kKnownNone)) << (2 * q);
  }
  if (known != nullptr) *known = kn;
  if (full) return kPlaneFull;
  return pend ? kPlaneSolve : kPlaneProven;
}

// x/y/z (min, max) from the five planes' extremes, combined as pass 1 of the full path does.
HD void combine_planes(const double v[10], double v6[6]) {
  v6[0] = jmin(v[2 * kUV], v[2 * kUW]);
  v6[1] = jmax(v[2 * kUV + 1], v[2 * kUW + 1]);
  v6[2] = jmin(v[2 * kVU], v[2 * kVW]);
  v6[3] = jmax(v[2 * kVU + 1], v[2 * kVW + 1]);
  v6[4] = jmin(v[2 * kWU], v[2 * kVW]);
  v6[5] = jmax(v[2 * kWU + 1], v[2 * kVW + 1]);
}

// x/y/z (min, max) of the 6x6 Jacobians of projected state A when the cell takes none of
// the rare branches.  Returns false when the cell needs the full path or a plane solve.
HD bool fast_cell_speeds(const double* A, bool positivity, const Warm& w, double v6[6],
                         uint32_t& fl) {
  double v[10];
  unsigned pend;
  if (fast_cell_planes<true, true>(A, positivity, w, v, pend, fl) != kPlaneProven) return false;
  combine_planes(v, v6);
  return true;
}

// The 1D closure speeds of the x, y and z slices of Mr (simulation_runner.jl:370-372):
// (min, max) per direction.
HD void closure1d_all(const double* Mr, double s[6]) {
  closure1d_speeds(Mr[0], Mr[1], Mr[2], Mr[3], Mr[4], s[0], s[1]);
  closure1d_speeds(Mr[0], Mr[5], Mr[9], Mr[12], Mr[14], s[2], s[3]);
  closure1d_speeds(Mr[0], Mr[15], Mr[19], Mr[22], Mr[24], s[4], s[5]);
}
// vp from the 1D speeds s and the 6x6 speeds v6 (simulation_runner.jl:374-379).
HD void combine_vp(const double s[6], const double v6[6], double vp[6]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vp[2 * d] = jmin(s[2 * d], v6[2 * d]);
    vp[2 * d + 1] = jmax(s[2 * d + 1], v6[2 * d + 1]);
  }
}

// The rest of pass 1 for a cell that took none of the rare branches: closure and fluxes of
// A and the 1D speeds, combined with the 6x6 speeds v6.
template <typename FT>
HD void pass1_finish(const double* A, int flag2D, double Ma, double* Mr, FT* Fx, FT* Fy, FT* Fz,
                     size_t st, const double v6[6], double vp[6], uint32_t& fl) {
  flux_closure35<true, FT>(A, flag2D, Ma, Mr, Fx, Fy, Fz, st, fl);
  double s[6];
  closure1d_all(Mr, s);
  combine_vp(s, v6, vp);
}

// Pass 2 of a cell when its x/y planes need no correction: closure, hyperbolicity check of
// UV, UW, VU, VW, closure, BGK.  Returns false (Mo untouched) when the cell needs the full
// pass 2.
// First half: projection of M to A and the hyperbolicity checks of UV, UW, VU, VW.
// Returns kPlaneProven, kPlaneSolve (bit q of pend: plane q needs its own full solve) or
// kPlaneFull (the full pass 2).
template <bool PAIR = true>
HD int pass2_check_planes(const double* M, int flag2D, double Ma, double* A, unsigned& pend,
                          uint32_t& fl, unsigned* known = nullptr) {
  flux_closure35<false, double>(M, flag2D, Ma, A, nullptr, nullptr, nullptr, 0, fl);
  double v[10];
  return fast_cell_planes<PAIR, false>(A, false, no_warm(), v, pend, fl, known);
}
// The full decision of plane q of projected state A (the fast solve with the pair test, else
// the QR): kKnownReal, with (v0, v1) its extremes when need_vals; kKnownComplex; or kKnownNone
// when the plane is not finite.  Either of the last two sends the cell to the full path.
HD unsigned solve_plane(const double* A, int q, bool need_vals, float wmin, float wmax,
                        double& v0, double& v1, uint32_t& fl) {
  double c[35], M000, um, vm, wm;
  raw_to_central4(A, c, M000, um, vm, wm);
  const float nan = NAN;
  const Plane p = cell_plane(c, um, vm, wm, q);
  const int st = plane_fast_status<true>(p, need_vals ? wmin : nan, need_vals ? wmax : nan,
                                         need_vals, fl, v0, v1);
  if (st == kPlaneFull) return kKnownNone;
  if (st == kPlaneComplex) return kKnownComplex;
  if (st == kPlaneSolve) {
    fl |= F_FALLBACK;
    const EigRes r = plane_eigs(p, fl);
    if (!r.ok) return kKnownNone;
    v0 = r.vmin;
    v1 = r.vmax;
    return r.cplx ? kKnownComplex : kKnownReal;
  }
  return kKnownReal;
}
// Pass 2's x or y check of a full-path cell whose two planes (UV, UW or VU, VW) were already
// decided on this state (known, as fast_cell_planes returns it): the correction of the planes
// with a complex pair, without solving the planes again, or a copy when there is none; when
// a decision is missing, eig6_axis.  The same result as eig6_axis with need_vals = false:
// the decisions came from the same solves on the same state.  Returns true when the state
// was corrected.
template <int AXIS>
HD bool eig6_axis_decided(const double* M, int flag2D, double Ma, double* Mr, uint32_t& fl,
                          unsigned known) {
  const unsigned sa = known_of(known, AXIS == 0 ? kUV : kVU);
  const unsigned sb = known_of(known, AXIS == 0 ? kUW : kVW);
  if (sa == kKnownNone || sb == kKnownNone) {
    double d0, d1;
    return eig6_axis<AXIS, 1>(M, flag2D, Ma, d0, d1, Mr, fl, no_warm(), false);
  }
  if (sa == kKnownReal && sb == kKnownReal) {
#pragma unroll
    for (int n = 0; n < 35; ++n) Mr[n] = M[n];
    return false;
  }
  double c[35], M000, um, vm, wm;
  raw_to_central4(M, c, M000, um, vm, wm);
  axis_correct<AXIS>(c, M000, um, vm, wm, sa == kKnownComplex, sb == kKnownComplex, flag2D, Ma, Mr, fl);
  return true;
}
HD bool pass2_check(const double* M, int flag2D, double Ma, double* A, uint32_t& fl) {
  unsigned pend;
  return pass2_check_planes<true>(M, flag2D, Ma, A, pend, fl) == kPlaneProven;
}
// Second half: projection of A and BGK.
template <bool MIXED>
HD void pass2_relax(const double* A, int flag2D, double Ma, bool positivity, double dt, double Kn,
                    double* Mo, uint32_t& fl) {
  double B[35];
  flux_closure35<false, double>(A, flag2D, Ma, B, nullptr, nullptr, nullptr, 0, fl);
  collision_cell<MIXED>(B, dt, Kn, positivity, Mo);
}
template <bool MIXED>
HD bool pass2_fast(const double* M, int flag2D, double Ma, bool positivity, double dt, double Kn,
                   double* Mo, uint32_t& fl) {
  double A[35];
  if (!pass2_check(M, flag2D, Ma, A, fl)) return false;
  pass2_relax<MIXED>(A, flag2D, Ma, positivity, dt, Kn, Mo, fl);
  return true;
}

}  // namespace hyq
