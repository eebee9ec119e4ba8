// HLL line update with the reference's physical-BC semantics (pas_HLL.jl:24-84,
// flux_HLL.jl:20-32, apply_flux_update_3d.jl:30-213) on the SoA state.
//
// Cell s of a line of n cells takes the value of cell j = clamp(s, 1, n-2) updated with the
// fluxes of faces j-1 and j (pas_HLL.jl:72-81: Mp[1] = Mp[2], Mp[Np] = Mp[Np-1]), so one
// cell reads cells j-1, j, j+1 of each of its three lines.  The HLL wave-speed bounds of
// those faces do not depend on the moment: they are formed once per cell (UpdDir) and the
// 35 moments then reuse them.  W* = (ll M_q - lr M_{q+1} - (F_q - F_{q+1})) / (ll - lr)
// where |ll - lr| > 1e-10, else 0; face 0 has zero wave speeds (lleft[1] = lright[1] = 0)
// and the flux of cell 1; the last face uses F_{n-2} twice (F[Np] = F[Np-1]).
#pragma once
#include "hyq_moments.h"

namespace hyq {

// HLL data of one face: bounds, their magnitudes and 1/(ll - lr) (sep: |ll - lr| > 1e-10).
template <typename T>
struct FaceHLL {
  T ll, lr, all, alr, inv;
  bool sep;
};
HD FaceHLL<double> face_hll(double vl0, double vl1, double vh0, double vh1) {
  FaceHLL<double> f;
  f.ll = jmin(vl0, vl1);  // Julia min/max: NaN propagates
  f.lr = jmax(vh0, vh1);
  f.all = fabs(f.ll);
  f.alr = fabs(f.lr);
  f.sep = fabs(f.ll - f.lr) > 1e-10;
  f.inv = f.sep ? 1.0 / (f.ll - f.lr) : 0.0;
  return f;
}
HD FaceHLL<float> face_hll(float vl0, float vl1, float vh0, float vh1) {
  FaceHLL<float> f;
  f.ll = fminf(vl0, vl1);
  f.lr = fmaxf(vh0, vh1);
  f.all = fabsf(f.ll);
  f.alr = fabsf(f.lr);
  f.sep = fabsf(f.ll - f.lr) > 1e-10f;
  f.inv = f.sep ? 1.0f / (f.ll - f.lr) : 0.0f;
  return f;
}

// One direction of one cell: offsets of cells j-1, j, j+1 (elements, relative to the cell)
// and the HLL data of faces j-1 (A) and j (B).  mode 0: line of one cell (unchanged);
// 1: two cells (both take cell 1); 2: general.
template <typename T>
struct UpdDir {
  int o0, o1, o2;  // a neighbour is at most one z plane away: |offset| <= Nx * Ny < 2^31
  int mode;
  bool first, last;
  FaceHLL<T> A, B;
};
// vl/vh: the line's lower/upper wave-speed planes at this cell's position.
template <typename T>
HD UpdDir<T> upd_dir(const T* vl, const T* vh, int s, int n, int st) {
  UpdDir<T> d;
  d.mode = n == 1 ? 0 : (n == 2 ? 1 : 2);
  d.first = d.last = false;
  if (d.mode != 2) {
    // the cell keeps its own value (one cell) or takes cell 1's (two cells); the faces are
    // not used, but upd_value evaluates them, so they hold zeros
    d.o0 = d.o1 = d.o2 = d.mode == 0 ? 0 : (1 - s) * st;
    d.A = d.B = FaceHLL<T>{T(0), T(0), T(0), T(0), T(0), false};
    return d;
  }
  const int j = s < 1 ? 1 : (s > n - 2 ? n - 2 : s);
  d.o0 = (j - 1 - s) * st;
  d.o1 = (j - s) * st;
  d.o2 = (j + 1 - s) * st;
  d.first = j == 1;
  d.last = j == n - 2;
  const T l0 = vl[d.o0], l1 = vl[d.o1], l2 = vl[d.o2];
  const T h0 = vh[d.o0], h1 = vh[d.o1], h2 = vh[d.o2];
  d.A = face_hll(l0, l1, h0, h1);
  d.B = face_hll(l1, l2, h1, h2);
  return d;
}

// Updated value of one moment along one direction (FP64 fluxes and arithmetic), general
// mode: m0, m1, m2 and f0, f1, f2 are the moment's state and flux at cells j-1, j, j+1
// (f0 is not read when d.first).
template <typename FT>
HD double upd_core(const UpdDir<double>& d, double m0, double m1, double m2, FT f0_, FT f1_, FT f2_,
                   double dtdx) {
  const double f1 = (double)f1_, f2 = (double)f2_;
  const double wsB = d.B.sep ? (d.B.ll * m1 - d.B.lr * m2) * d.B.inv - (f1 - f2) * d.B.inv : 0.0;
  const double fB = 0.5 * (f1 + (d.last ? f1 : f2)) - 0.5 * (d.B.all * (wsB - m1) - d.B.alr * (wsB - m2));
  // face 0 (d.first): zero wave speeds, W* of face 1, F[1] := F[2]; both forms are evaluated
  // and selected, so that no load waits behind a branch
  const double fA0 = 0.5 * (f1 + f1) - 0.5 * (0.0 * (wsB - m0) - 0.0 * (wsB - m1));
  const double f0 = (double)f0_;
  const double wsA = d.A.sep ? (d.A.ll * m0 - d.A.lr * m1) * d.A.inv - (f0 - f1) * d.A.inv : 0.0;
  const double fAg = 0.5 * (f0 + f1) - 0.5 * (d.A.all * (wsA - m0) - d.A.alr * (wsA - m1));
  const double fA = d.first ? fA0 : fAg;
  return m1 - dtdx * (fB - fA);
}

// Mixed precision: FP32 fluxes, wave speeds and HLL arithmetic in difference form (W* - M_q
// and W* - M_{q+1} from M_q - M_{q+1}), so FP32 rounding applies to differences and to the
// increment, never to the FP64 state.
HD double upd_core(const UpdDir<float>& d, double m0, double m1, double m2, float f0, float f1,
                   float f2, double dtdx) {
  const float dmB = (float)(m1 - m2);
  const float w0B = d.B.sep ? (d.B.lr * dmB - (f1 - f2)) * d.B.inv : -(float)m1;
  const float w1B = w0B + dmB;
  const float fB = 0.5f * (f1 + (d.last ? f1 : f2)) - 0.5f * (d.B.all * w0B - d.B.alr * w1B);
  const float dmA = (float)(m0 - m1);
  const float w0A = d.A.sep ? (d.A.lr * dmA - (f0 - f1)) * d.A.inv : -(float)m0;
  const float w1A = w0A + dmA;
  const float fAg = 0.5f * (f0 + f1) - 0.5f * (d.A.all * w0A - d.A.alr * w1A);
  const float fA = d.first ? f1 : fAg;  // zero wave speeds on face 0: the flux of cell 1
  const float inc = -(float)dtdx * (fB - fA);
  return m1 + (double)inc;
}

// Byte offsets, within a state plane (FP64) and a flux plane (FT), of one direction's cells
// j-1, j, j+1 (f0: cell j for face 0, whose flux is F[1] := F[2]).  32 bits: a plane holds at
// most kMaxUpdateCells cells.
constexpr long long kMaxUpdateCells = 1ll << 29;
struct UpdOff {
  unsigned m0, m1, m2, f0, f1, f2;
};
template <typename T, typename FT>
HD UpdOff upd_off(const UpdDir<T>& d, unsigned c) {
  UpdOff r;
  r.m0 = (c + (unsigned)d.o0) * 8u;
  r.m1 = (c + (unsigned)d.o1) * 8u;
  r.m2 = (c + (unsigned)d.o2) * 8u;
  r.f0 = (c + (unsigned)(d.first ? d.o1 : d.o0)) * (unsigned)sizeof(FT);
  r.f1 = (c + (unsigned)d.o1) * (unsigned)sizeof(FT);
  r.f2 = (c + (unsigned)d.o2) * (unsigned)sizeof(FT);
  return r;
}
template <typename T>
HD T at_byte(const T* plane, unsigned off) {
  return *(const T*)((const char*)plane + off);
}

// One moment's values along one direction: state and flux at cells j-1, j, j+1.
template <typename FT>
struct UpdVals {
  double m0, m1, m2;
  FT f0, f1, f2;
};
// The six loads are made in every mode, from offsets that stay inside the line (the mode
// selects the result in upd_value): no load sits behind a branch.
template <typename FT>
HD UpdVals<FT> upd_load(const UpdOff& o, const double* m, const FT* f) {
  UpdVals<FT> v;
  v.m0 = at_byte(m, o.m0);
  v.m1 = at_byte(m, o.m1);
  v.m2 = at_byte(m, o.m2);
  v.f0 = at_byte(f, o.f0);
  v.f1 = at_byte(f, o.f1);
  v.f2 = at_byte(f, o.f2);
  return v;
}
// Updated value of one moment along one direction.
template <typename T, typename FT>
HD double upd_value(const UpdDir<T>& d, const UpdVals<FT>& v, double dtdx) {
  const double r = upd_core(d, v.m0, v.m1, v.m2, v.f0, v.f1, v.f2, dtdx);
  return d.mode == 2 ? r : v.m1;
}

// Combined x/y/z update of cell c (i, j, k) for all 35 moments: Mx + My + Mz - 2M
// (simulation_runner.jl:516-520); when Nz == 1 the z sweep leaves the value unchanged
// (apply_flux_update_3d.jl:183), as a line of one cell does.  Mr: state planes [35][N]; F:
// flux planes [105][N] (x, y, z blocks); vp: wave-speed planes [6][N]; out: [35][N]; N is
// the plane stride (a grid's cells may be a contiguous part of longer planes; at most
// kMaxUpdateCells), c the cell's plane index.  Each moment's 19 loads are issued together,
// before any of them is used, so the moment waits for memory once; the results are stored
// non-temporally, which leaves the caches to the neighbours' values (324 against 333 us per
// launch at 120x120x60, mixed).
template <typename FT>
HD void update_cell(const double* Mr, const FT* F, const FT* vp, double* out, long long c, int i,
                    int j, int k, int Nx, int Ny, int Nz, long long N, double dtdx, double dtdy,
                    double dtdz) {
  const int sy = Nx, sz = Nx * Ny;
  const UpdDir<FT> dx = upd_dir(vp + c, vp + N + c, i, Nx, 1);
  const UpdDir<FT> dy = upd_dir(vp + 2 * N + c, vp + 3 * N + c, j, Ny, sy);
  const UpdDir<FT> dz = upd_dir(vp + 4 * N + c, vp + 5 * N + c, k, Nz, sz);
  const unsigned c32 = (unsigned)c;
  const UpdOff ox = upd_off<FT, FT>(dx, c32), oy = upd_off<FT, FT>(dy, c32), oz = upd_off<FT, FT>(dz, c32);
  for (int n = 0; n < 35; ++n) {
    const double* m = Mr + (size_t)n * N;
    const double m0 = at_byte(m, c32 * 8u);
    const UpdVals<FT> vx = upd_load(ox, m, F + (size_t)n * N);
    const UpdVals<FT> vy = upd_load(oy, m, F + (size_t)(35 + n) * N);
    const UpdVals<FT> vz = upd_load(oz, m, F + (size_t)(70 + n) * N);
    HYQ_SCHED_BARRIER();
    const double mx = upd_value(dx, vx, dtdx);
    const double my = upd_value(dy, vy, dtdy);
    const double mz = upd_value(dz, vz, dtdz);  // m0 when Nz == 1
    HYQ_STORE_STREAM(&out[(size_t)n * N + c], mx + my + mz - 2.0 * m0);
  }
}

}  // namespace hyq
