// The k = 3 instances of the cell apply (55 local flux dofs for a 10-gon: a
// fully unrolled cell body of up to 1540 packed operator entries), in a
// translation unit of their own so that they compile beside device.hip.
#include "apply.h"

namespace vemd {

template <typename T>
void launch_apply_k3(const ApplyArgs<T>& a, dim3 grid, hipStream_t stream) {
    hipLaunchKernelGGL((k_apply<T, 3>), grid, dim3(256), 0, stream, a);
}
template void launch_apply_k3<double>(const ApplyArgs<double>&, dim3, hipStream_t);
template void launch_apply_k3<float>(const ApplyArgs<float>&, dim3, hipStream_t);

void launch_adjoint_k3(const ApplyArgs<double>& a, int ntiles, hipStream_t stream, const double* lam, const double* x,
                       const int* cell_of_gpos, double* grad) {
    hipLaunchKernelGGL((k_adjoint<double, 3>), dim3(ntiles), dim3(64), 0, stream, a, lam, x, cell_of_gpos, grad);
}

}  // namespace vemd
