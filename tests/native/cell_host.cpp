// Host build of the device cell numerics (hyqmom.jl_amd/csrc/*.h) for CPU tests only:
// lets the CPU suite compare the GPU code paths with the oracle without a GPU.  Not part
// of the product; the product is libhyqmom_hip.so.
#include <cstring>
#include "../../hyqmom.jl_amd/csrc/hyq_pipeline.h"

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
