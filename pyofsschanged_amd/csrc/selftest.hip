// Engine self-test of the workgroup reductions (ops.hpp wg_sum, wg_sum_n):
// M sums at once must equal M single sums bit for bit, in every thread, at
// each workgroup size the propagation kernels use.
#include <hip/hip_runtime.h>

#include <string>

#include "ctx.hpp"

namespace gn {

template <int T, int M>
__global__ void __launch_bounds__(T) k_sum_check(const double* __restrict__ in, double* __restrict__ out_n,
                                                 double* __restrict__ out_1) {
  constexpr int NW = T / 64;
  __shared__ double buf[M * NW];
  __shared__ double red[NW];
  const size_t base = (size_t)blockIdx.x * M * T + threadIdx.x;
  double v[M], w[M];
#pragma unroll
  for (int m = 0; m < M; ++m) w[m] = v[m] = in[base + (size_t)m * T];
  wg_sum_n<NW, M>(v, buf);
#pragma unroll
  for (int m = 0; m < M; ++m) w[m] = wg_sum<NW>(w[m], red);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    out_n[base + (size_t)m * T] = v[m];
    out_1[base + (size_t)m * T] = w[m];
  }
}

template <int T>
static hipError_t launch_sum_check(int m, unsigned blocks, const double* in, double* on, double* o1) {
  switch (m) {
    case 2: hipLaunchKernelGGL((k_sum_check<T, 2>), dim3(blocks), dim3(T), 0, 0, in, on, o1); break;
    case 3: hipLaunchKernelGGL((k_sum_check<T, 3>), dim3(blocks), dim3(T), 0, 0, in, on, o1); break;
    case 4: hipLaunchKernelGGL((k_sum_check<T, 4>), dim3(blocks), dim3(T), 0, 0, in, on, o1); break;
    case 9: hipLaunchKernelGGL((k_sum_check<T, 9>), dim3(blocks), dim3(T), 0, 0, in, on, o1); break;
    case 18: hipLaunchKernelGGL((k_sum_check<T, 18>), dim3(blocks), dim3(T), 0, 0, in, on, o1); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gn

extern "C" int gnlse_check_reductions(int device, int threads, int m, int64_t blocks, const double* in, double* out_n,
                                      double* out_1) {
  if (!in || !out_n || !out_1 || blocks < 1 || blocks > 4096) return fail(GNLSE_ERR_ARG, "invalid arguments");
  if (m != 2 && m != 3 && m != 4 && m != 9 && m != 18) return fail(GNLSE_ERR_ARG, "m must be 2, 3, 4, 9 or 18");
  if (threads != 64 && threads != 128 && threads != 256 && threads != 512 && threads != 1024)
    return fail(GNLSE_ERR_ARG, "threads must be 64, 128, 256, 512 or 1024");
  HIPCHK(hipSetDevice(device));
  const size_t bytes = (size_t)blocks * m * threads * sizeof(double);
  double *d_in = nullptr, *d_n = nullptr, *d_1 = nullptr;
  HIPCHK(hipMalloc(&d_in, bytes));
  if (hipMalloc(&d_n, bytes) != hipSuccess || hipMalloc(&d_1, bytes) != hipSuccess) {
    (void)hipFree(d_in);
    if (d_n) (void)hipFree(d_n);
    return fail(GNLSE_ERR_HIP, "hipMalloc failed");
  }
  hipError_t e = hipMemcpy(d_in, in, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const unsigned b = (unsigned)blocks;
    switch (threads) {
      case 64: e = launch_sum_check<64>(m, b, d_in, d_n, d_1); break;
      case 128: e = launch_sum_check<128>(m, b, d_in, d_n, d_1); break;
      case 256: e = launch_sum_check<256>(m, b, d_in, d_n, d_1); break;
      case 512: e = launch_sum_check<512>(m, b, d_in, d_n, d_1); break;
      default: e = launch_sum_check<1024>(m, b, d_in, d_n, d_1); break;
    }
  }
  if (e == hipSuccess) e = hipMemcpy(out_n, d_n, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(out_1, d_1, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d_in);
  (void)hipFree(d_n);
  (void)hipFree(d_1);
  if (e != hipSuccess) return fail(GNLSE_ERR_HIP, std::string("reduction check: ") + hipGetErrorString(e));
  return GNLSE_OK;
}
