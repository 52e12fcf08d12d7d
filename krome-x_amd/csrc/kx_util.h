// Small device helpers of the batch entry points: the parameter check of
// device-resident calls (the reference's "krome_params_check" reduction,
// runtime/core/solve_kokkos.hpp:1711-1719) and the flag fill
// ("krome_fill_flags", solve_kokkos.hpp:1670-1673).
#pragma once
#include <hip/hip_runtime.h>

namespace kx {

// rows with an unset REQUIRED parameter get FLAG_PARAM_UNSET and count
template <class Net>
__global__ void params_check_kernel(int ncell, const double* par, int pst,
                                    int* flags, int* count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const bool m = Net::params_missing(par + (long)c * pst) >= 0;
  flags[c] = m ? -105 : 0;
  if (m) atomicAdd(count, 1);
}

__global__ inline void fill_flags_kernel(int ncell, int* flags, int v) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) flags[c] = v;
}

// number of rows with a REQUIRED parameter unset (< 0 on a HIP error);
// flags receive -105 or 0
template <class Net>
inline int params_check_device(int ncell, const double* par, int pst,
                               int* flags, int* cnt) {
  int h = 0;
  hipError_t e = hipMemcpy(cnt, &h, sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL((params_check_kernel<Net>), dim3((ncell + 255) / 256),
                       dim3(256), 0, nullptr, ncell, par, pst, flags, cnt);
    e = hipMemcpy(&h, cnt, sizeof(int), hipMemcpyDeviceToHost);
  }
  return e == hipSuccess ? h : -1;
}

inline void fill_flags(int ncell, int* flags, int v) {
  hipLaunchKernelGGL(fill_flags_kernel, dim3((ncell + 255) / 256), dim3(256),
                     0, nullptr, ncell, flags, v);
  (void)hipDeviceSynchronize();
}

}  // namespace kx
