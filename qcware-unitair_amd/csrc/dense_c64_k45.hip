// Dense 4/5-qubit vector kernels instantiated for float2 states.
#include "dense_kernels.h"

namespace ua {
namespace dense_detail {

void launch_fast_k45_c64(int k, int mode, bool lds, void* out, const void* in,
                         const void* op, const FastParams& p, int unr, hipStream_t s) {
  launch_fast_k45<float2>(k, mode, lds, out, in, op, p, unr, s);
}

}  // namespace dense_detail
}  // namespace ua
