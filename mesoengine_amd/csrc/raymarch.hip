// Raymarch kernels: one lane per pixel, wave64 packets of 8x8 pixels so the
// rays of a wave fetch the same bricks.  Tiled variants remap workgroups so
// that each XCD (own L2) renders a contiguous band of tiles; the persistent
// variant keeps every wave slot busy and pulls packets from per-band queues.
#include "traverse.h"

namespace meso {

// Hardware hands workgroup i to XCD i % 8.  Give XCD x a contiguous range of
// logical tiles so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
  const uint32_t per = nblocks / 8u, rem = nblocks % 8u;
  const uint32_t xcd = bid % 8u, local = bid / 8u;
  return xcd < rem ? xcd * (per + 1u) + local : rem * (per + 1u) + (xcd - rem) * per + local;
}

template <int KIND, bool COUNT, int TW, int TH>
__global__ void __launch_bounds__(TW *TH) raymarch_kernel(DevScene s, DevCamera cam, uint32_t flags,
                                                          DevShading sh, RayOut out,
                                                          uint32_t tiles_x, uint32_t n_tiles) {
  // tile of TW x TH pixels made of 8x8 wave packets
  const uint32_t tile = xcd_remap(blockIdx.x, n_tiles);
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  constexpr uint32_t waves_x = TW / 8;
  const int32_t px = (int32_t)(tx * TW + (wave % waves_x) * 8u + (lane & 7u));
  const int32_t py = (int32_t)(ty * TH + (wave / waves_x) * 8u + (lane >> 3));

  StepCounts cnt = {0, 0, 0, 0, 0, 0, 0};
  uint32_t n_shadow = 0, n_ao = 0;
  bool valid = px < cam.width && py < cam.height;
  PixelResult r{-1, kFaceMiss, 0.0f, 0.0f};
  if (valid) {
    r = shade_pixel<KIND, COUNT>(s, cam, flags, sh, px, py, cnt, n_shadow, n_ao);
    const int64_t pix = (int64_t)py * cam.width + px;
    if (out.index_bytes == 8)
      ((int64_t *)out.index)[pix] = r.index;
    else
      ((int32_t *)out.index)[pix] = (int32_t)r.index;
    if (out.face) out.face[pix] = (uint8_t)r.face;
    if (out.t) out.t[pix] = r.t;
    if (out.shade) out.shade[pix] = r.shade;
  }
  if constexpr (COUNT) {
    __shared__ unsigned long long acc[MESO_CNT_COUNT];
    if (threadIdx.x < MESO_CNT_COUNT) acc[threadIdx.x] = 0;
    __syncthreads();
    if (valid) {
      atomicAdd(&acc[MESO_CNT_N_VOX], (unsigned long long)cnt.n_vox);
      atomicAdd(&acc[MESO_CNT_N_TOP], (unsigned long long)cnt.n_top);
      atomicAdd(&acc[MESO_CNT_N_NODE], (unsigned long long)cnt.n_node);
      atomicAdd(&acc[MESO_CNT_N_BRICKSLOT], (unsigned long long)cnt.n_brickslot);
      atomicAdd(&acc[MESO_CNT_N_BRICKMASK], (unsigned long long)cnt.n_brickmask);
      atomicAdd(&acc[MESO_CNT_HITS], (unsigned long long)(r.index >= 0 ? 1 : 0));
      atomicAdd(&acc[MESO_CNT_RAYS], 1ull);
      atomicAdd(&acc[MESO_CNT_SHADOW_RAYS], (unsigned long long)n_shadow);
      atomicAdd(&acc[MESO_CNT_AO_RAYS], (unsigned long long)n_ao);
      atomicAdd(&acc[MESO_CNT_N_COARSE], (unsigned long long)cnt.n_coarse);
      atomicAdd(&acc[MESO_CNT_N_ITER], (unsigned long long)cnt.n_iter);
    }
    __syncthreads();
    if (threadIdx.x < MESO_CNT_COUNT && out.counters)
      atomicAdd(&out.counters[threadIdx.x], acc[threadIdx.x]);
  }
}

// End of a persistent wave: add its counters, then leave.  The band counters
// persist per stream: the last wave to leave (queue[128] counts leavers) sets
// them back to zero, after every wave's last counter access has returned,
// for the next launch.
template <bool COUNT>
__device__ __forceinline__ void persistent_leave(const RayOut &out, unsigned int *queue,
                                                 const StepCounts &cnt, uint32_t n_hits,
                                                 uint32_t n_rays, uint32_t n_shadow,
                                                 uint32_t n_ao) {
  const uint32_t lane = threadIdx.x;
  if constexpr (COUNT) {
    unsigned long long v[MESO_CNT_COUNT] = {cnt.n_vox, cnt.n_top, cnt.n_node, cnt.n_brickslot,
                                            cnt.n_brickmask, n_hits, n_rays, n_shadow, n_ao,
                                            cnt.n_coarse, cnt.n_iter};
#pragma unroll
    for (int c = 0; c < MESO_CNT_COUNT; c++) {
      unsigned long long x = v[c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
      if (lane == 0 && out.counters && x) atomicAdd(&out.counters[c], x);
    }
  }
  if (lane == 0 && atomicAdd(&queue[128], 1u) == gridDim.x - 1u) {
#pragma unroll
    for (uint32_t b = 0; b < 8u; b++)
      __hip_atomic_store(&queue[b * 16u], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&queue[128], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Persistent waves: one 64-lane workgroup per wave slot of the device.  The
// screen's PW x (64 / PW) packets (8x8, 16x4 or 32x2) are cut into 8 bands
// of rows; the waves of XCD x
// (workgroup i runs on XCD i % 8) take packets of band x from that band's
// counter, then help the other bands in turn.  A wave leaves once every
// counter has passed the end of its band, so the grid always drains; slots
// freed by short (sky) packets are refilled at once instead of waiting for
// the next workgroup dispatch.
template <int KIND, bool COUNT, int PW>
__global__ void __launch_bounds__(64) raymarch_persistent_kernel(
    DevScene s, DevCamera cam, uint32_t flags, DevShading sh, RayOut out, uint32_t packets_x,
    uint32_t n_packets, unsigned int *queue) {
  constexpr uint32_t PH = 64 / PW;  // packet of PW x PH pixels
  const uint32_t lane = threadIdx.x;
  const uint32_t home = blockIdx.x & 7u;
  StepCounts cnt = {0, 0, 0, 0, 0, 0, 0};
  uint32_t n_shadow = 0, n_ao = 0, n_hits = 0, n_rays = 0;
  for (uint32_t k = 0; k < 8u; k++) {
    const uint32_t band = (home + k) & 7u;
    const uint32_t lo = (uint32_t)(((uint64_t)n_packets * band) >> 3);
    const uint32_t hi = (uint32_t)(((uint64_t)n_packets * (band + 1u)) >> 3);
    for (;;) {
      uint32_t p = 0;
      if (lane == 0) p = atomicAdd(&queue[band * 16u], 1u);  // one counter per 64 B line
      p = lo + (uint32_t)__shfl((int)p, 0, 64);
      if (p >= hi) break;  // uniform over the wave
      const int32_t px = (int32_t)((p % packets_x) * (uint32_t)PW + lane % (uint32_t)PW);
      const int32_t py = (int32_t)((p / packets_x) * PH + lane / (uint32_t)PW);
      if (px < cam.width && py < cam.height) {
        const PixelResult r = shade_pixel<KIND, COUNT>(s, cam, flags, sh, px, py, cnt, n_shadow, n_ao);
        const int64_t pix = (int64_t)py * cam.width + px;
        if (out.index_bytes == 8)
          ((int64_t *)out.index)[pix] = r.index;
        else
          ((int32_t *)out.index)[pix] = (int32_t)r.index;
        if (out.face) out.face[pix] = (uint8_t)r.face;
        if (out.t) out.t[pix] = r.t;
        if (out.shade) out.shade[pix] = r.shade;
        n_hits += r.index >= 0 ? 1u : 0u;
        n_rays++;
      }
    }
  }
  persistent_leave<COUNT>(out, queue, cnt, n_hits, n_rays, n_shadow, n_ao);
}


// j-th (0-based) set bit of a 64-bit mask that has more than j set bits
__device__ __forceinline__ uint32_t select_bit64(uint64_t m, uint32_t j) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t w = 32; w >= 1; w >>= 1) {
    const uint32_t c = (uint32_t)__popcll(m & ((1ull << w) - 1ull));
    if (j >= c) {
      j -= c;
      m >>= w;
      pos += w;
    }
  }
  return pos;
}

// A pixel's primary ray: writes index, face and t; returns true (and fills
// sc and face) when the pixel needs secondary rays, else writes the shade.
template <int KIND, bool COUNT>
__device__ __forceinline__ bool primary_pixel(const DevScene &s, const DevCamera &cam,
                                              uint32_t flags, const DevShading &sh, int32_t px,
                                              int32_t py, const RayOut &out, StepCounts &cnt,
                                              uint32_t &n_hits, Secondary &sc, int32_t &face) {
  float o[3], d[3];
  primary_ray(cam, px, py, o, d);
  Hit h;
  const bool hit = trace<KIND, COUNT>(s, o, d, h, cnt);
  const int64_t pix = (int64_t)py * cam.width + px;
  if (out.index_bytes == 8)
    ((int64_t *)out.index)[pix] = h.index;
  else
    ((int32_t *)out.index)[pix] = (int32_t)h.index;
  if (out.face) out.face[pix] = (uint8_t)h.face;
  if (out.t) out.t[pix] = h.t;
  n_hits += h.index >= 0 ? 1u : 0u;
  face = h.face;
  if (hit && h.face < 6) {
    secondary_setup(sh, o, d, h.face, h.t, h.cell, sc);
    if (needs_secondary(flags, sh, h.face)) return true;
    if (out.shade) out.shade[pix] = shade_value(sc.lambert, 1.0f, false, 0, sh.ao_samples);
    return false;
  }
  if (out.shade) out.shade[pix] = hit ? 0.25f : 0.0f;
  return false;
}

// Persistent waves over pairs of packets (variants 8 and 9).  The wave traces
// the primary rays of two packets, then shares out the pixels of both that
// need secondary rays (shadow, AO) over its 64 lanes, so that the secondary
// phase runs with up to twice the busy lanes of one packet.  A pixel's
// secondary state (hit point, start cell, Lambert term, face, pixel) moves to
// the lane that traces it by lane shuffles.  Every ray computes what
// shade_pixel computes, so the outputs are the same bits.  It holds more
// registers than the one-packet waves (5 waves/SIMD against 7), and pays off
// when issue, not latency, bounds the frame (for example with two frames in
// flight on the 1080p heightfield).
template <int KIND, bool COUNT, int PW>
__global__ void __launch_bounds__(64) raymarch_pair_kernel(
    DevScene s, DevCamera cam, uint32_t flags, DevShading sh, RayOut out, uint32_t packets_x,
    uint32_t n_packets, unsigned int *queue) {
  constexpr uint32_t PH = 64 / PW;
  const uint32_t lane = threadIdx.x;
  const uint32_t home = blockIdx.x & 7u;
  StepCounts cnt = {0, 0, 0, 0, 0, 0, 0};
  uint32_t n_shadow = 0, n_ao = 0, n_hits = 0, n_rays = 0;
  for (uint32_t k = 0; k < 8u; k++) {
    const uint32_t band = (home + k) & 7u;
    const uint32_t lo = (uint32_t)(((uint64_t)n_packets * band) >> 3);
    const uint32_t hi = (uint32_t)(((uint64_t)n_packets * (band + 1u)) >> 3);
    for (;;) {
      uint32_t p = 0;
      if (lane == 0) p = atomicAdd(&queue[band * 16u], 2u);  // two packets at a time
      p = lo + (uint32_t)__shfl((int)p, 0, 64);
      if (p >= hi) break;  // uniform over the wave
      Secondary sa = {{0.0f, 0.0f, 0.0f}, {0, 0, 0}, 0.0f}, sb = sa;
      int32_t fa = 0, fb = 0;
      uint32_t pa = 0, pb = 0;
      bool na = false, nb = false;
      {
        const int32_t px = (int32_t)((p % packets_x) * (uint32_t)PW + lane % (uint32_t)PW);
        const int32_t py = (int32_t)((p / packets_x) * PH + lane / (uint32_t)PW);
        if (px < cam.width && py < cam.height) {
          n_rays++;
          pa = (uint32_t)((int64_t)py * cam.width + px);
          na = primary_pixel<KIND, COUNT>(s, cam, flags, sh, px, py, out, cnt, n_hits, sa, fa);
        }
      }
      if (p + 1u < hi) {  // uniform
        const uint32_t q = p + 1u;
        const int32_t px = (int32_t)((q % packets_x) * (uint32_t)PW + lane % (uint32_t)PW);
        const int32_t py = (int32_t)((q / packets_x) * PH + lane / (uint32_t)PW);
        if (px < cam.width && py < cam.height) {
          n_rays++;
          pb = (uint32_t)((int64_t)py * cam.width + px);
          nb = primary_pixel<KIND, COUNT>(s, cam, flags, sh, px, py, out, cnt, n_hits, sb, fb);
        }
      }
      const uint64_t ma = __ballot(na), mb = __ballot(nb);
      const uint32_t ca = (uint32_t)__popcll(ma), total = ca + (uint32_t)__popcll(mb);
      for (uint32_t base = 0; base < total; base += 64u) {  // uniform
        const uint32_t j = base + lane;
        const bool from_a = j < ca;
        const uint32_t src = j < total ? (from_a ? select_bit64(ma, j) : select_bit64(mb, j - ca))
                                       : lane;
        // every lane takes part in every shuffle; the lane keeps A's or B's copy
        Secondary sc;
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const float xa = __shfl(sa.p[i], (int)src, 64), xb = __shfl(sb.p[i], (int)src, 64);
          sc.p[i] = from_a ? xa : xb;
          const int32_t ia = __shfl(sa.start[i], (int)src, 64), ib = __shfl(sb.start[i], (int)src, 64);
          sc.start[i] = from_a ? ia : ib;
        }
        const float la = __shfl(sa.lambert, (int)src, 64), lb = __shfl(sb.lambert, (int)src, 64);
        sc.lambert = from_a ? la : lb;
        const int32_t ga = __shfl(fa, (int)src, 64), gb = __shfl(fb, (int)src, 64);
        const uint32_t qa = (uint32_t)__shfl((int)pa, (int)src, 64), qb = (uint32_t)__shfl((int)pb, (int)src, 64);
        const int32_t face = from_a ? ga : gb;
        const uint32_t pix = from_a ? qa : qb;
        if (j < total) {
          const float shade = shade_secondary<KIND, COUNT>(s, flags, sh, pix, face, sc, cnt,
                                                           n_shadow, n_ao);
          if (out.shade) out.shade[pix] = shade;
        }
      }
    }
  }
  persistent_leave<COUNT>(out, queue, cnt, n_hits, n_rays, n_shadow, n_ao);
}

template <int KIND, bool COUNT, int PW>
static meso_status launch_pair(meso_context *ctx, const meso_scene *scene, const DevCamera &cam,
                               uint32_t flags, const DevShading &sh, const RayOut &out,
                               hipStream_t stream, int grid) {
  constexpr int PH = 64 / PW;
  const uint32_t packets_x = (uint32_t)((cam.width + PW - 1) / PW);
  const uint32_t n_packets = packets_x * (uint32_t)((cam.height + PH - 1) / PH);
  if (grid <= 0) {
    int per_cu = 0;
    MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, raymarch_pair_kernel<KIND, COUNT, PW>, 64, 0));
    grid = (per_cu < 1 ? 1 : per_cu) * ctx->props.multiProcessorCount;
  }
  grid = (grid + 7) & ~7;
  unsigned int *queue = nullptr;
  const meso_status st = band_queue(ctx, stream, &queue);
  if (st != MESO_OK) return st;
  hipLaunchKernelGGL((raymarch_pair_kernel<KIND, COUNT, PW>), dim3((uint32_t)grid), dim3(64), 0,
                     stream, scene->dev(), cam, flags, sh, out, packets_x, n_packets, queue);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

template <int KIND, bool COUNT, int PW>
static meso_status launch_persistent(meso_context *ctx, const meso_scene *scene,
                                     const DevCamera &cam, uint32_t flags, const DevShading &sh,
                                     const RayOut &out, hipStream_t stream, int grid) {
  constexpr int PH = 64 / PW;
  const uint32_t packets_x = (uint32_t)((cam.width + PW - 1) / PW);
  const uint32_t n_packets = packets_x * (uint32_t)((cam.height + PH - 1) / PH);
  if (grid <= 0) {
    int per_cu = 0;
    MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, raymarch_persistent_kernel<KIND, COUNT, PW>, 64, 0));
    grid = (per_cu < 1 ? 1 : per_cu) * ctx->props.multiProcessorCount;
  }
  grid = (grid + 7) & ~7;  // every XCD gets the same number of waves
  // one counter set per stream: launches on one stream are ordered, and
  // launches on different streams never share counters
  unsigned int *queue = nullptr;
  const meso_status st = band_queue(ctx, stream, &queue);
  if (st != MESO_OK) return st;
  hipLaunchKernelGGL((raymarch_persistent_kernel<KIND, COUNT, PW>), dim3((uint32_t)grid), dim3(64), 0,
                     stream, scene->dev(), cam, flags, sh, out, packets_x, n_packets, queue);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

template <int KIND, bool COUNT, int TW, int TH>
static meso_status launch_shape(const meso_scene *scene, const DevCamera &cam, uint32_t flags,
                                const DevShading &sh, const RayOut &out, hipStream_t stream) {
  const uint32_t tiles_x = (uint32_t)((cam.width + TW - 1) / TW);
  const uint32_t tiles_y = (uint32_t)((cam.height + TH - 1) / TH);
  const uint32_t n_tiles = tiles_x * tiles_y;
  hipLaunchKernelGGL((raymarch_kernel<KIND, COUNT, TW, TH>), dim3(n_tiles), dim3(TW * TH), 0,
                     stream, scene->dev(), cam, flags, sh, out, tiles_x, n_tiles);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

template <int KIND, bool COUNT>
static meso_status launch_kind(meso_context *ctx, const meso_scene *scene, const DevCamera &cam,
                               uint32_t flags, const DevShading &sh, const RayOut &out,
                               hipStream_t stream, int variant, int grid) {
  switch (variant) {
    case 0: return launch_shape<KIND, COUNT, 16, 16>(scene, cam, flags, sh, out, stream);
    case 1: return launch_shape<KIND, COUNT, 32, 8>(scene, cam, flags, sh, out, stream);
    case 2: return launch_shape<KIND, COUNT, 8, 8>(scene, cam, flags, sh, out, stream);
    case 3: return launch_shape<KIND, COUNT, 16, 8>(scene, cam, flags, sh, out, stream);
    case 4: return launch_shape<KIND, COUNT, 32, 16>(scene, cam, flags, sh, out, stream);
    case 5: return launch_persistent<KIND, COUNT, 8>(ctx, scene, cam, flags, sh, out, stream, grid);
    case 6: return launch_persistent<KIND, COUNT, 16>(ctx, scene, cam, flags, sh, out, stream, grid);
    case 7: return launch_persistent<KIND, COUNT, 32>(ctx, scene, cam, flags, sh, out, stream, grid);
    case 8: return launch_pair<KIND, COUNT, 8>(ctx, scene, cam, flags, sh, out, stream, grid);
    case 9: return launch_pair<KIND, COUNT, 16>(ctx, scene, cam, flags, sh, out, stream, grid);
    default: return fail(MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown raymarch variant %d", variant);
  }
}

RayLaunch default_ray_launch(const meso_scene *) { return RayLaunch{0, 256, 16, 16, 0}; }

meso_status launch_raymarch(meso_context *ctx, const meso_scene *scene, const DevCamera &cam,
                            uint32_t flags, const DevShading &sh, const RayOut &out,
                            hipStream_t stream, const RayLaunch *shape) {
  const int variant = shape ? shape->variant : 0;
  const int grid = shape ? shape->grid : 0;
  const bool count = (flags & MESO_RAY_COUNT) != 0;
  switch (scene->kind) {
    case MESO_SCENE_DENSE_U8:
      return count ? launch_kind<1, true>(ctx, scene, cam, flags, sh, out, stream, variant, grid)
                   : launch_kind<1, false>(ctx, scene, cam, flags, sh, out, stream, variant, grid);
    case MESO_SCENE_BRICKMAP_U16:
      return count ? launch_kind<2, true>(ctx, scene, cam, flags, sh, out, stream, variant, grid)
                   : launch_kind<2, false>(ctx, scene, cam, flags, sh, out, stream, variant, grid);
    case MESO_SCENE_TWO_LEVEL_U16:
      return count ? launch_kind<3, true>(ctx, scene, cam, flags, sh, out, stream, variant, grid)
                   : launch_kind<3, false>(ctx, scene, cam, flags, sh, out, stream, variant, grid);
    default: return fail(MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown scene kind %d", scene->kind);
  }
}

}  // namespace meso
