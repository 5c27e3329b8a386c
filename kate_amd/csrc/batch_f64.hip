// batch_f64.hip -- instantiates the persistent k_batch kernel for complex128 states.
#include "kernels_dev.hpp"
#include "kernels_inst.hpp"

namespace kate {

int launch_batch_kernel_f64(const BatchArgs& a, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    return a.wide ? batch_w<double, true>(a, s) : batch_w<double, false>(a, s);
}

}  // namespace kate
