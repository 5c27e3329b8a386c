// batch_f32.hip -- instantiates the persistent k_batch kernel for complex64 states.
#include "kernels_dev.hpp"
#include "kernels_inst.hpp"

namespace kate {

int launch_batch_kernel_f32(const BatchArgs& a, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    return a.wide ? batch_w<float, true>(a, s) : batch_w<float, false>(a, s);
}

}  // namespace kate
