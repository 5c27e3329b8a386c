// pass_f32.hip -- instantiates k_pass for complex64 states.
#include "kernels_dev.hpp"
#include "kernels_inst.hpp"

namespace kate {

int launch_pass_f32(int tile_bits, void* state, int nq_int, const PassArgs& pa, const DevPlan& dp, void* stream)
{
    return pass_r<float>(tile_bits, state, nq_int, pa, dp, (hipStream_t)stream);
}

}  // namespace kate
