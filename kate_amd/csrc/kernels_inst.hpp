// kernels_inst.hpp -- entry points of the translation units that instantiate
// the pass and batch kernels (one unit per precision and kind, so that they
// compile in parallel).  Same arguments and return codes as launch_pass /
// launch_batch.
#pragma once

#include "kernels.hpp"

namespace kate {

int launch_pass_f64(int tile_bits, void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, void* stream);
int launch_pass_f32(int tile_bits, void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, void* stream);
int launch_batch_kernel_f64(const BatchArgs& a, void* stream);
int launch_batch_kernel_f32(const BatchArgs& a, void* stream);

}  // namespace kate
