// selftest.hip — exhaustive device check of the fast rsqrt of device.hpp.
//
// Every float bit pattern in [lo, hi] is evaluated with dev::rsq (hardware
// estimates + corrections) and with the IEEE expression 1.0f / sqrtf(x); the
// kernel counts the patterns whose results differ bit for bit (NaN results are
// compared as "both NaN") and reports the smallest differing input.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "internal.hpp"

namespace ot {
namespace {

__global__ void __launch_bounds__(256) k_rsq_check(uint32_t lo, uint32_t hi, unsigned long long* mismatches,
                                                   uint32_t* first_bad) {
  const uint64_t n = (uint64_t)hi - lo + 1;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long bad = 0;
  uint32_t first = 0xFFFFFFFFu;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t bits = lo + (uint32_t)i;
    const float x = __builtin_bit_cast(float, bits);
    const float a = dev::rsq(x), b = dev::rsq_ieee(x);
    const bool same = (a != a && b != b) || __builtin_bit_cast(uint32_t, a) == __builtin_bit_cast(uint32_t, b);
    if (!same) {
      ++bad;
      first = bits < first ? bits : first;
    }
  }
  if (bad) {
    atomicAdd(mismatches, bad);
    atomicMin(first_bad, first);
  }
}

}  // namespace
}  // namespace ot

extern "C" ot_status ot_selftest_rsq(uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches, uint32_t* first_bad) {
  if (!mismatches || !first_bad || hi_bits < lo_bits)
    return ot::set_error(OT_EINVAL, "ot_selftest_rsq: bad argument");
  unsigned long long* d_count = nullptr;
  uint32_t* d_first = nullptr;
  hipError_t e = hipMalloc(&d_count, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&d_first, sizeof(uint32_t));
  const unsigned long long zero = 0;
  const uint32_t none = 0xFFFFFFFFu;
  if (e == hipSuccess) e = hipMemcpy(d_count, &zero, sizeof(zero), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_first, &none, sizeof(none), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ot::k_rsq_check, dim3(8192), dim3(256), 0, 0, lo_bits, hi_bits, d_count, d_first);
    e = hipGetLastError();
  }
  unsigned long long count = 0;
  uint32_t first = none;
  if (e == hipSuccess) e = hipMemcpy(&count, d_count, sizeof(count), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost);
  if (d_count) (void)hipFree(d_count);
  if (d_first) (void)hipFree(d_first);
  if (e != hipSuccess) return ot::set_error(OT_EHIP, "ot_selftest_rsq: %s", hipGetErrorString(e));
  *mismatches = count;
  *first_bad = first;
  return OT_OK;
}
