// Shared by the fused-pass interpreter (fused.hip) and the compiled passes
// (fused_jit.hip): the structural validation of a ua_fused_pass.
#pragma once
#include "ua_common.h"

namespace ua {

// What the validation learns about a pass.
struct FusedPassInfo {
  bool heavy;        // has a DENSE2 op
  bool only_layers;  // only LAYER1Q heads and their DENSE1 sub-ops
  bool needs_index;  // a cube stage has an op that reads the global index
  bool reg_head;     // the first stage is diagonal with PHASE_FACTORS ops only
  bool reg_tail;     // the last stage is diagonal with PHASE_FACTORS ops only
};

bool is_diag_kind(int kind);

// Checks dtype, qubit count, tile bits (4..max_tile_bits), stages and every
// op of `pass` against host_ops; fills *info.  Returns UA_OK or
// UA_ERR_INVALID_ARGUMENT (with the message set).
int check_fused_pass(int dtype, int n, const ua_fused_pass* pass, const ua_fused_op* host_ops,
                     FusedPassInfo* info, int max_tile_bits);

// UA_TUNE_COMPILED_WAVES (fused_jit.hip).
int set_compiled_waves(int waves);

}  // namespace ua
