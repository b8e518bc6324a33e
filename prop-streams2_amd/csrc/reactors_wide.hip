// reactors_wide.hip — the wide instances of the batched reactor kernel (compiled for more
// waves per SIMD, for batches larger than the GPU holds at one wave per SIMD).  A separate
// translation unit, so that the narrow instances in reactors.hip do not share device code,
// and with it a register budget, with these.
#include "reactors_dev.h"

namespace streams {

int reactor_dispatch_wide(int n_s, int indx_cp_l, int indx_cp_r, const ReactorArgs &A, int ncu, int wg_per_cu,
                          hipStream_t st) {
  return reactor_dispatch<true>(n_s, indx_cp_l, indx_cp_r, A, ncu, wg_per_cu, st);
}

}  // namespace streams
