// Ensemble cell kernels, mixed_newton instantiation (see hyq_ensemble_kernels.h).
#include "hyq_ensemble_kernels.h"

namespace hyq_ens {
void launch_newton(int which, int b, unsigned nblk, const Blk* T, const EnsPtrs& p, hipStream_t s) {
  launch_cells<float, 1, true>(which, b, nblk, T, p, s);
}
}  // namespace hyq_ens
