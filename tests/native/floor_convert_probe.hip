// Test-only library: checks on the device the one-instruction floor-convert
// that the raymarcher's exit point uses (meso::floor_i32, v_cvt_flr_i32_f32).
// Over every finite binary32 bit pattern, both signs, it counts
//   - results that differ from v_floor_f32 followed by v_cvt_i32_f32, and
//   - results in the int32 range that differ from floor(x) computed exactly
//     in integers from the bit pattern (denormals included: -denormal -> -1).
// No memory is read; one 4 x u64 buffer of results is written.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "traverse.h"

namespace {

constexpr uint32_t kFinite = 0x7F800000u;  // finite magnitudes: bits [0, kFinite)

__device__ __forceinline__ int32_t floor_then_convert(float x) {
  float f;
  int32_t r;
  asm("v_floor_f32 %0, %1" : "=v"(f) : "v"(x));
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
}

// floor(x) from the bits of x when it lies in [-2^31, 2^31); in_range false
// otherwise.
__host__ __device__ __forceinline__ int32_t exact_floor(uint32_t b, bool &in_range) {
  const bool neg = (b >> 31) != 0u;
  const uint32_t mag = b & 0x7FFFFFFFu;
  const int e = (int)(mag >> 23) - 127;
  in_range = true;
  if (e < 0) return (neg && mag != 0u) ? -1 : 0;  // |x| < 1
  if (e > 31 || (e == 31 && !(neg && (mag & 0x7FFFFFu) == 0u))) {
    in_range = false;  // |x| >= 2^31, except x = -2^31
    return 0;
  }
  const uint64_t m = (uint64_t)((mag & 0x7FFFFFu) | 0x800000u);
  int64_t v;
  bool frac = false;
  if (e >= 23) {
    v = (int64_t)(m << (e - 23));
  } else {
    v = (int64_t)(m >> (23 - e));
    frac = (m & ((1ull << (23 - e)) - 1ull)) != 0ull;
  }
  if (neg) v = -v - (frac ? 1 : 0);
  return (int32_t)v;
}

// out[0]: differences from floor-then-convert, out[1]: differences from the
// exact floor, out[2]: patterns checked against the exact floor, out[3]: the
// lowest differing pattern index (UINT64_MAX when none).
__global__ void floor_convert_probe_kernel(unsigned long long *out) {
  const uint64_t total = 2ull * kFinite;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long exact_checked = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint32_t b = i < kFinite ? (uint32_t)i : ((uint32_t)(i - kFinite) | 0x80000000u);
    const float x = __uint_as_float(b);
    const int32_t r = meso::floor_i32(x);
    bool in_range;
    const int32_t e = exact_floor(b, in_range);
    const bool bad_pair = r != floor_then_convert(x);
    const bool bad_exact = in_range && r != e;
    exact_checked += in_range ? 1ull : 0ull;
    if (bad_pair) atomicAdd(&out[0], 1ull);
    if (bad_exact) atomicAdd(&out[1], 1ull);
    if (bad_pair || bad_exact) atomicMin(&out[3], (unsigned long long)i);
  }
  atomicAdd(&out[2], exact_checked);
}

}  // namespace

extern "C" int floor_convert_probe(unsigned long long *host_out) {
  unsigned long long init[4] = {0ull, 0ull, 0ull, ~0ull};
  unsigned long long *d = nullptr;
  if (hipMalloc(&d, sizeof(init)) != hipSuccess) return 1;
  int rc = 0;
  if (hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) rc = 2;
  if (rc == 0) {
    floor_convert_probe_kernel<<<2048, 256>>>(d);
    if (hipGetLastError() != hipSuccess) rc = 3;
  }
  if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = 4;
  if (rc == 0 && hipMemcpy(host_out, d, sizeof(init), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
  (void)hipFree(d);
  return rc;
}
