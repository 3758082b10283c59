// Dense k-qubit vector kernels instantiated for double states (one translation
// unit per dtype so that the instantiations compile in parallel).
#include "dense_kernels.h"

namespace ua {
namespace dense_detail {

void launch_fast_k_f64(int k, int mode, bool lds, void* out, const void* in, const void* op,
                       const FastParams& p, int unr, hipStream_t s) {
  if (k >= 4) return launch_fast_k45<double>(k, mode, lds, out, in, op, p, unr, s);
  launch_fast_k123<double>(k, mode, lds, out, in, op, p, unr, s);
}

}  // namespace dense_detail
}  // namespace ua
