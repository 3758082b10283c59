// Dense 1..3-qubit vector kernels instantiated for double2 states; the 4/5-qubit
// ones are in dense_c128_k45.hip (separate translation units compile in
// parallel).
#include "dense_kernels.h"

namespace ua {
namespace dense_detail {

void launch_fast_k_c128(int k, int mode, bool lds, void* out, const void* in, const void* op,
                       const FastParams& p, int unr, hipStream_t s) {
  if (k >= 4) return launch_fast_k45_c128(k, mode, lds, out, in, op, p, unr, s);
  launch_fast_k123<double2>(k, mode, lds, out, in, op, p, unr, s);
}

}  // namespace dense_detail
}  // namespace ua
