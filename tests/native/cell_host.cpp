// Host build of the device cell numerics (hyqmom.jl_amd/csrc/*.h) for CPU tests only:
// lets the CPU suite compare the GPU code paths with the oracle without a GPU.  Not part
// of the product; the product is libhyqmom_hip.so.
#include <cstring>
#include "../../hyqmom.jl_amd/csrc/hyq_pipeline.h"
#include "../../hyqmom.jl_amd/csrc/hyq_update.h"
#include "../../hyqmom.jl_amd/csrc/hyq_fasteig.h"
#include <vector>
#include <cmath>

using namespace hyq;
extern "C" {
void cell_flux_closure35(const double* M, int flag2D, double Ma, double* Fx, double* Fy,
                         double* Fz, double* Mr, uint32_t* flags) {
  uint32_t fl = 0;
  flux_closure35<true>(M, flag2D, Ma, Mr, Fx, Fy, Fz, 1, fl);
  *flags = fl;
}
void cell_eig6(const double* M, int axis, int flag2D, double Ma, double* vmin, double* vmax,
               double* Mr, uint32_t* flags) {
  uint32_t fl = 0;
  if (axis == 1) eig6_axis<0>(M, flag2D, Ma, *vmin, *vmax, Mr, fl);
  else if (axis == 2) eig6_axis<1>(M, flag2D, Ma, *vmin, *vmax, Mr, fl);
  else eig6_z(M, flag2D, Ma, true, *vmin, *vmax, Mr, fl);
  *flags = fl;
}
void cell_pass1(const double* M, int flag2D, double Ma, int positivity, double* Mr, double* Fx,
                double* Fy, double* Fz, double* vp, uint32_t* flags) {
  uint32_t fl = 0;
  pass1_cell(M, flag2D, Ma, positivity != 0, Mr, Fx, Fy, Fz, 1, vp, fl);
  *flags = fl;
}
void cell_pass2(const double* M, int flag2D, double Ma, int positivity, double dt, double Kn,
                double* Mo, uint32_t* flags) {
  uint32_t fl = 0;
  pass2_cell(M, flag2D, Ma, positivity != 0, dt, Kn, Mo, fl);
  *flags = fl;
}
void cell_collision(const double* M, double dt, double Kn, int positivity, double* Mo) {
  collision_cell(M, dt, Kn, positivity != 0, Mo);
}
void cell_closure1d(const double* m5, double* vmin, double* vmax) {
  closure1d_speeds(m5[0], m5[1], m5[2], m5[3], m5[4], *vmin, *vmax);
}
}
extern "C" int cell_realizable_3D(const double* in28, double* out28, uint32_t* flags) {
  S28 s;
  std::memcpy(&s, in28, sizeof(s));
  uint32_t fl = 0;
  const int f = realizable_3D(s, fl);
  std::memcpy(out28, &s, sizeof(s));
  *flags = fl;
  return f;
}
extern "C" void cell_delta2star(const double* in28, double* E21) {
  S28 s;
  std::memcpy(&s, in28, sizeof(s));
  Sym6 E;
  delta2star(s, E);
  std::memcpy(E21, E.e, sizeof(E.e));
}
// Host emulation of one device step (k_pass1, k_dt, k_update, k_pass2) on the SoA state
// M[35][N]; returns dt.  Same cell functions and line logic as the kernels.
static double host_step_impl(double* M, int Nx, int Ny, int Nz, double Kn, double Ma, int flag2D,
                             double CFL, double dtmax, double dx, double dy, double dz,
                             double tmax, double t, uint32_t* flags, bool skip_pass2) {
  const long long N = (long long)Nx * Ny * Nz;
  std::vector<double> Mr(35 * N), F(105 * N), vp(6 * N);
  double vmax = 0.0;
  bool nan = false;
  for (long long c = 0; c < N; ++c) {
    double m[35], mr[35], v[6];
    for (int n = 0; n < 35; ++n) m[n] = M[n * N + c];
    uint32_t fl = 0;
    pass1_cell<double>(m, flag2D, Ma, true, mr, &F[c], &F[35 * N + c], &F[70 * N + c], (size_t)N, v, fl);
    for (int n = 0; n < 35; ++n) Mr[n * N + c] = mr[n];
    for (int d = 0; d < 6; ++d) {
      vp[d * N + c] = v[d];
      if (std::isnan(v[d])) nan = true;
      vmax = std::fmax(vmax, std::fabs(v[d]));
    }
    if (flags) flags[c] = fl;
  }
  if (nan) vmax = NAN;
  double dt = jmin(CFL * jmin(jmin(dx, dy), dz) / vmax, jmin(dtmax, Kn));
  dt = jmin(dt, tmax - t);
  for (long long c = 0; c < N; ++c) {
    const int i = (int)(c % Nx), j = (int)((c / Nx) % Ny), k = (int)(c / ((long long)Nx * Ny));
    update_cell<double>(Mr.data(), F.data(), vp.data(), M, c, i, j, k, Nx, Ny, Nz, N, dt / dx,
                        dt / dy, dt / dz);
  }
  if (skip_pass2) return dt;
  for (long long c = 0; c < N; ++c) {
    double m[35], o[35];
    for (int n = 0; n < 35; ++n) m[n] = M[n * N + c];
    uint32_t fl = 0;
    pass2_cell<false>(m, flag2D, Ma, true, dt, Kn, o, fl);
    for (int n = 0; n < 35; ++n) M[n * N + c] = o[n];
  }
  return dt;
}
extern "C" double host_step(double* M, int Nx, int Ny, int Nz, double Kn, double Ma, int flag2D,
                            double CFL, double dtmax, double dx, double dy, double dz,
                            double tmax, double t, uint32_t* flags) {
  return host_step_impl(M, Nx, Ny, Nz, Kn, Ma, flag2D, CFL, dtmax, dx, dy, dz, tmax, t, flags, false);
}
// The same step without pass 2: the state after the update, as pass 2 receives it.
extern "C" double host_step_to_update(double* M, int Nx, int Ny, int Nz, double Kn, double Ma,
                                      int flag2D, double CFL, double dtmax, double dx, double dy,
                                      double dz, double tmax, double t) {
  return host_step_impl(M, Nx, Ny, Nz, Kn, Ma, flag2D, CFL, dtmax, dx, dy, dz, tmax, t, nullptr, true);
}
extern "C" void cell_pass2_flags(const double* M, int flag2D, double Ma, double dt, double Kn,
                                 double* Mo, uint32_t* fl_out) {
  uint32_t fl = 0;
  pass2_cell<false>(M, flag2D, Ma, true, dt, Kn, Mo, fl);
  *fl_out = fl;
}
// Plane solve both ways on the standardized Jacobian of plane (f,g) of a 35-moment state.
// out: [qr_min, qr_max, qr_cplx, fast_min, fast_max, fast_real, fast_iters, shift, scale,
// fast_cplx]
// (raw scale; standardized = (raw - shift) / scale)
extern "C" void cell_plane_compare(const double* M, int f, int g, double wmin, double wmax,
                                   double* out) {
  double c[35], M000, um, vm, wm;
  raw_to_central4(M, c, M000, um, vm, wm);
  Plane p;
  if (f == 0 && g == 1) p = plane<0, 1>(c, um, vm, wm);
  else if (f == 0 && g == 2) p = plane<0, 2>(c, um, vm, wm);
  else if (f == 1 && g == 0) p = plane<1, 0>(c, um, vm, wm);
  else if (f == 1 && g == 2) p = plane<1, 2>(c, um, vm, wm);
  else p = plane<2, 0>(c, um, vm, wm);
  uint32_t fl = 0;
  const EigRes q = plane_eigs(p, fl);
  double J[36];
  jacobian_std(p, J);
  const FastEig fe = fast_eigs_std(J, wmin, wmax, true, p.sig);
  out[0] = q.vmin; out[1] = q.vmax; out[2] = q.cplx ? 1 : 0;
  out[3] = p.a + p.sig * fe.lmin; out[4] = p.a + p.sig * fe.lmax; out[5] = fe.real ? 1 : 0;
  out[6] = fe.iters;
  out[7] = p.a; out[8] = p.sig; out[9] = fe.cplx ? 1 : 0;
}

// The QR iteration both ways on a 6x6 matrix (balanced and reduced to Hessenberg form first):
// hqr_core on an array accessor and the register form hqr6_reg.  out: for each, the smallest
// and largest real part, the largest |imag| and the convergence flag (8 values).
extern "C" void cell_qr_both(const double* J36, double* out) {
  double H[36];
  for (int t = 0; t < 36; ++t) H[t] = J36[t];
  balance_n<6>(H);
  hessenberg_n<6>(H);
  double G[36];
  for (int t = 0; t < 36; ++t) G[t] = H[t];
  ArrMat<6> m{H};
  ExtremeSink a{0.0, 0.0, 0.0, true}, b{0.0, 0.0, 0.0, true};
  const bool ca = hqr_core<6>(m, a);
  const bool cb = hqr6_reg(G, b);
  out[0] = a.mn; out[1] = a.mx; out[2] = a.mi; out[3] = ca ? 1 : 0;
  out[4] = b.mn; out[5] = b.mx; out[6] = b.mi; out[7] = cb ? 1 : 0;
}

// The standardized Jacobian of plane (f, g) of a 35-moment state (row-major 6x6).
extern "C" void cell_jacobian_std(const double* M, int f, int g, double* J) {
  double c[35], M000, um, vm, wm;
  raw_to_central4(M, c, M000, um, vm, wm);
  Plane p;
  if (f == 0 && g == 1) p = plane<0, 1>(c, um, vm, wm);
  else if (f == 0 && g == 2) p = plane<0, 2>(c, um, vm, wm);
  else if (f == 1 && g == 0) p = plane<1, 0>(c, um, vm, wm);
  else if (f == 1 && g == 2) p = plane<1, 2>(c, um, vm, wm);
  else p = plane<2, 0>(c, um, vm, wm);
  jacobian_std(p, J);
}


// The fast standardized solve on an arbitrary 6x6 matrix (row-major).  out: [real, cplx,
// lmin, lmax, iters].
extern "C" void cell_fast_eigs_J(const double* J36, double sig, int need_vals, double* out) {
  const FastEig fe = fast_eigs_std(J36, NAN, NAN, need_vals != 0, sig);
  out[0] = fe.real ? 1 : 0; out[1] = fe.cplx ? 1 : 0; out[2] = fe.lmin; out[3] = fe.lmax;
  out[4] = fe.iters;
}
