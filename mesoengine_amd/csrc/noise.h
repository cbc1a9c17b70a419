// Seeded procedural voxel functions shared by the voxelization kernels.
// Integer lattice hashing (lowbias32) plus float +, -, *, / and floor in a
// fixed order: a scalar IEEE implementation of the same steps reproduces every
// voxel bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace meso {

__host__ __device__ __forceinline__ uint32_t nhash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ float hash_unit(uint32_t h) {
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float fade(float t) {
  float q = t * 6.0f;
  q = q - 15.0f;
  q = q * t;
  q = q + 10.0f;
  float t3 = t * t;
  t3 = t3 * t;
  return t3 * q;
}
__device__ __forceinline__ float lerpf(float a, float b, float t) {
  float d = b - a;
  d = d * t;
  return a + d;
}

__device__ __forceinline__ float value_noise2(float x, float z, uint32_t seed) {
  float fx = __builtin_floorf(x), fz = __builtin_floorf(z);
  int32_t ix = (int32_t)fx, iz = (int32_t)fz;
  float ux = fade(x - fx), uz = fade(z - fz);
  uint32_t hz0 = nhash32((uint32_t)iz ^ seed), hz1 = nhash32((uint32_t)(iz + 1) ^ seed);
  float v00 = hash_unit(nhash32((uint32_t)ix ^ hz0));
  float v10 = hash_unit(nhash32((uint32_t)(ix + 1) ^ hz0));
  float v01 = hash_unit(nhash32((uint32_t)ix ^ hz1));
  float v11 = hash_unit(nhash32((uint32_t)(ix + 1) ^ hz1));
  float a = lerpf(v00, v10, ux);
  float b = lerpf(v01, v11, ux);
  return lerpf(a, b, uz);
}

__device__ __forceinline__ float value_noise3(float x, float y, float z, uint32_t seed) {
  float fx = __builtin_floorf(x), fy = __builtin_floorf(y), fz = __builtin_floorf(z);
  int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  float ux = fade(x - fx), uy = fade(y - fy), uz = fade(z - fz);
  // lattice hash h(x,y,z) = H(x ^ H(y ^ H(z ^ seed)))
  const uint32_t hz0 = nhash32((uint32_t)iz ^ seed), hz1 = nhash32((uint32_t)(iz + 1) ^ seed);
  const uint32_t h00 = nhash32((uint32_t)iy ^ hz0), h10 = nhash32((uint32_t)(iy + 1) ^ hz0);
  const uint32_t h01 = nhash32((uint32_t)iy ^ hz1), h11 = nhash32((uint32_t)(iy + 1) ^ hz1);
  const uint32_t x0 = (uint32_t)ix, x1 = (uint32_t)(ix + 1);
  float v0 = hash_unit(nhash32(x0 ^ h00)), v1 = hash_unit(nhash32(x1 ^ h00));
  float v2 = hash_unit(nhash32(x0 ^ h10)), v3 = hash_unit(nhash32(x1 ^ h10));
  float v4 = hash_unit(nhash32(x0 ^ h01)), v5 = hash_unit(nhash32(x1 ^ h01));
  float v6 = hash_unit(nhash32(x0 ^ h11)), v7 = hash_unit(nhash32(x1 ^ h11));
  float x00 = lerpf(v0, v1, ux);
  float x10 = lerpf(v2, v3, ux);
  float x01 = lerpf(v4, v5, ux);
  float x11 = lerpf(v6, v7, ux);
  float y0 = lerpf(x00, x10, uy);
  float y1 = lerpf(x01, x11, uy);
  return lerpf(y0, y1, uz);
}

constexpr uint32_t kOctaveSeedStep = 0x9E3779B9u;

__device__ __forceinline__ float fbm2(float x, float z, int octaves, float base_freq,
                                      uint32_t seed) {
  float sum = 0.0f, norm = 0.0f, amp = 1.0f, f = base_freq;
  for (int o = 0; o < octaves; o++) {
    float px = x * f, pz = z * f;
    float n = value_noise2(px, pz, seed + (uint32_t)o * kOctaveSeedStep);
    float an = amp * n;
    sum = sum + an;
    norm = norm + amp;
    amp = amp * 0.5f;
    f = f * 2.0f;
  }
  return sum / norm;
}

__device__ __forceinline__ float fbm3(float x, float y, float z, int octaves, float base_freq,
                                      uint32_t seed) {
  float sum = 0.0f, norm = 0.0f, amp = 1.0f, f = base_freq;
  for (int o = 0; o < octaves; o++) {
    float px = x * f, py = y * f, pz = z * f;
    float n = value_noise3(px, py, pz, seed + (uint32_t)o * kOctaveSeedStep);
    float an = amp * n;
    sum = sum + an;
    norm = norm + amp;
    amp = amp * 0.5f;
    f = f * 2.0f;
  }
  return sum / norm;
}

// heightfield: h = floor(n * (0.10 + 0.40 * (fbm2 - 1/2))), 4 octaves, period n/4
__device__ __forceinline__ int32_t heightfield_height(int32_t x, int32_t z, int32_t n,
                                                      uint32_t seed) {
  float f = fbm2((float)x, (float)z, 4, 4.0f / (float)n, seed);
  f = f - 0.5f;
  float h = 0.40f * f;
  h = 0.10f + h;
  h = (float)n * h;
  return (int32_t)__builtin_floorf(h);
}
__device__ __forceinline__ int32_t terrain_height(int32_t x, int32_t z, int32_t n, uint32_t seed) {
  float f = fbm2((float)x, (float)z, 4, 4.0f / (float)n, seed);
  float h = 0.40f * f;
  h = 0.05f + h;
  h = (float)n * h;
  return (int32_t)__builtin_floorf(h);
}
// material by height band of n/16
__device__ __forceinline__ uint16_t height_band(int32_t y, int32_t n) {
  int32_t b = (y * 16) / n;
  return (uint16_t)(1 + (b < 3 ? b : 3));
}
__device__ __forceinline__ uint16_t terrain_caves_voxel(int32_t x, int32_t y, int32_t z,
                                                        int32_t n, uint32_t seed, int32_t h) {
  if (y >= h) return 0;
  float c = fbm3((float)x, (float)y, (float)z, 3, 8.0f / (float)n, seed ^ 0x5bd1e995u);
  c = c - 0.5f;
  if (__builtin_fabsf(c) < 0.03f) return 0;
  return height_band(y, n);
}
__device__ __forceinline__ uint16_t caves_voxel(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  float c = fbm3((float)x, (float)y, (float)z, 4, 1.0f / 64.0f, seed);
  c = c - 0.5f;
  if (__builtin_fabsf(c) >= 0.0055f) return 0;
  return (uint16_t)(1 + ((uint32_t)y >> 7) % 4u);
}

}  // namespace meso
