// HLL line update with the reference's physical-BC semantics (pas_HLL.jl:24-84,
// flux_HLL.jl:20-32, apply_flux_update_3d.jl:30-213), on strided lines of the SoA state.
#pragma once
#include "hyq_moments.h"

namespace hyq {
template <typename FT>
struct Line {
  const double* M;
  const FT* F;
  const FT* lo;
  const FT* hi;
  long long st;  // stride between consecutive cells of the line
  int n;
  HD double m(int q) const { return M[(long long)q * st]; }
  HD double f(int q) const { return (double)F[(long long)q * st]; }
  HD double vl(int q) const { return (double)lo[(long long)q * st]; }
  HD double vh(int q) const { return (double)hi[(long long)q * st]; }
  // intermediate HLL state of face q (between cells q and q+1), F unmodified
  HD double wstar(int q, double ll, double lr) const {
    if (fabs(ll - lr) > 1e-10)
      return (ll * m(q) - lr * m(q + 1)) / (ll - lr) - (f(q) - f(q + 1)) / (ll - lr);
    return 0.0;
  }
  // numerical flux of face q, 0 <= q <= n-2, with the physical-BC copies at both ends
  HD double flux(int q) const {
    if (q == 0) {
      // lleft[1] = lright[1] = 0; F[1] := F[2]; Wstar[1] := Wstar[2]
      double ws = 0.0;
      if (n >= 3) {
        const double ll = jmin(vl(1), vl(2)), lr = jmax(vh(1), vh(2));
        ws = wstar(1, ll, lr);
      }
      const double f1 = f(1);
      return 0.5 * (f1 + f1) - 0.5 * (0.0 * (ws - m(0)) - 0.0 * (ws - m(1)));
    }
    const double ll = jmin(vl(q), vl(q + 1)), lr = jmax(vh(q), vh(q + 1));
    const double ws = wstar(q, ll, lr);
    const double fa = f(q), fb = (q == n - 2) ? f(q) : f(q + 1);  // F[Np] := F[Np-1]
    return 0.5 * (fa + fb) - 0.5 * (fabs(ll) * (ws - m(q)) - fabs(lr) * (ws - m(q + 1)));
  }
  // Mp at cell s of the line
  HD double updated(int s, double dtdx) const {
    if (n == 1) return m(s);
    if (n == 2) return m(1);
    const int j = s < 1 ? 1 : (s > n - 2 ? n - 2 : s);
    return m(j) - dtdx * (flux(j) - flux(j - 1));
  }
};

// Mixed precision: the HLL fluxes and increments in FP32 on FP32 fluxes and wave speeds,
// written in difference form (W* - M_q and W* - M_{q+1} from M_q - M_{q+1}) so that FP32
// rounding applies to differences, not to the FP64 state; the state stays FP64.
struct LineF {
  const double* M;
  const float* F;
  const float* lo;
  const float* hi;
  long long st;
  int n;
  HD double m(int q) const { return M[(long long)q * st]; }
  HD float f(int q) const { return F[(long long)q * st]; }
  HD float vl(int q) const { return lo[(long long)q * st]; }
  HD float vh(int q) const { return hi[(long long)q * st]; }
  // (W* - M_q, W* - M_{q+1}) of face q
  HD void wdiff(int q, float ll, float lr, float& w0, float& w1) const {
    const float dm = (float)(m(q) - m(q + 1));
    if (fabsf(ll - lr) > 1e-10f) {
      w0 = (lr * dm - (f(q) - f(q + 1))) / (ll - lr);
    } else {
      w0 = -(float)m(q);
    }
    w1 = w0 + dm;
  }
  HD float flux(int q) const {
    if (q == 0) {
      const float f1 = f(1);
      return f1;  // zero wave speeds on the first face: the flux of cell 1
    }
    const float ll = fminf(vl(q), vl(q + 1)), lr = fmaxf(vh(q), vh(q + 1));
    float w0, w1;
    wdiff(q, ll, lr, w0, w1);
    const float fa = f(q), fb = (q == n - 2) ? f(q) : f(q + 1);
    return 0.5f * (fa + fb) - 0.5f * (fabsf(ll) * w0 - fabsf(lr) * w1);
  }
  // (updated value at cell s) - M[s]: the increment plus the boundary copy offset
  HD double updated(int s, double dtdx) const {
    if (n == 1) return m(s);
    if (n == 2) return m(1);
    const int j = s < 1 ? 1 : (s > n - 2 ? n - 2 : s);
    const float inc = -(float)dtdx * (flux(j) - flux(j - 1));
    return m(j) + (double)inc;
  }
};


// Combined x/y/z update of moment plane Mn at cell c (i,j,k): Mx + My + Mz - 2M.
template <typename L>
HD double update_cell(const L& lx, const L& ly, const L& lz, double m0, int i, int j, int k,
                      int Nz, double dtdx, double dtdy, double dtdz) {
  const double mx = lx.updated(i, dtdx);
  const double my = ly.updated(j, dtdy);
  const double mz = (Nz == 1) ? m0 : lz.updated(k, dtdz);  // apply_flux_update_3d.jl:183
  return mx + my + mz - 2.0 * m0;
}

}  // namespace hyq
