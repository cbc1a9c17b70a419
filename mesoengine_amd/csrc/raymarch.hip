// Raymarch kernels: one lane per pixel, wave64 packets of 8x8 pixels so the
// rays of a wave fetch the same bricks; workgroups are remapped so that each
// XCD (own L2) renders a contiguous band of tiles.
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
static meso_status launch_kind(const meso_scene *scene, const DevCamera &cam, uint32_t flags,
                               const DevShading &sh, const RayOut &out, hipStream_t stream,
                               int variant) {
  switch (variant) {
    case 0: return launch_shape<KIND, COUNT, 16, 16>(scene, cam, flags, sh, out, stream);
    case 1: return launch_shape<KIND, COUNT, 32, 8>(scene, cam, flags, sh, out, stream);
    case 2: return launch_shape<KIND, COUNT, 8, 8>(scene, cam, flags, sh, out, stream);
    case 3: return launch_shape<KIND, COUNT, 16, 8>(scene, cam, flags, sh, out, stream);
    case 4: return launch_shape<KIND, COUNT, 32, 16>(scene, cam, flags, sh, out, stream);
    default: return fail(MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown raymarch variant %d", variant);
  }
}

RayLaunch default_ray_launch(const meso_scene *) { return RayLaunch{0, 256, 16, 16}; }

meso_status launch_raymarch(meso_context *, const meso_scene *scene, const DevCamera &cam,
                            uint32_t flags, const DevShading &sh, const RayOut &out,
                            hipStream_t stream, const RayLaunch *shape) {
  const int variant = shape ? shape->variant : 0;
  const bool count = (flags & MESO_RAY_COUNT) != 0;
  switch (scene->kind) {
    case MESO_SCENE_DENSE_U8:
      return count ? launch_kind<1, true>(scene, cam, flags, sh, out, stream, variant)
                   : launch_kind<1, false>(scene, cam, flags, sh, out, stream, variant);
    case MESO_SCENE_BRICKMAP_U16:
      return count ? launch_kind<2, true>(scene, cam, flags, sh, out, stream, variant)
                   : launch_kind<2, false>(scene, cam, flags, sh, out, stream, variant);
    case MESO_SCENE_TWO_LEVEL_U16:
      return count ? launch_kind<3, true>(scene, cam, flags, sh, out, stream, variant)
                   : launch_kind<3, false>(scene, cam, flags, sh, out, stream, variant);
    default: return fail(MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown scene kind %d", scene->kind);
  }
}

}  // namespace meso
