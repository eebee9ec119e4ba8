// Ensemble cell kernels, FP64 instantiation (see hyq_ensemble_kernels.h).
#include "hyq_ensemble_kernels.h"

namespace hyq_ens {
void launch_fp64(int which, int b, unsigned nblk, const Blk* T, const EnsPtrs& p, hipStream_t s) {
  launch_cells<double, 0, false>(which, b, nblk, T, p, s);
}
}  // namespace hyq_ens
