// Quad visibility: the raster consumer of the meshes.  For every pixel of a
// camera, the nearest front-facing quad of a chunk pool's resident slots
// along the pixel's primary ray, written in the raymarcher's output format.
// It takes the place of the reference's per-cell imposter draw
// (Samples/SimpleVoxel.cpp:352-398, a 3-face fan per solid cell), drawing the
// packed quads of the mesher instead.
//
// Work is split by quad size, on the device and without a host round trip:
//   raster_quads_kernel: one workgroup per slot, skipped when the chunk's
//     box projects outside the frame; a lane per quad culls back faces, clips the quad at 1e-3 in front of the eye and projects it to a
//     conservative pixel rectangle (1 pixel of margin).  A quad of at most
//     kSmall pixels is tested by its own lane.  A quad of at most kMedium
//     pixels is tested by its whole wave, 64 pixels a step, the wave taking
//     its lanes' medium quads in turn.  Larger quads are cut into segments
//     of kSegPixels pixels appended to a list, with their rectangle.
//   raster_segments_kernel: persistent waves, one segment at a time, a lane
//     per pixel.
//   raster_resolve_kernel: a lane per pixel turns the winning key into the
//     voxel index, face, t and shade.
// A hit is a 64-bit key (float bits of t << 32) | (slot * per_slot + k),
// combined with atomicMin, so the result does not depend on the order in
// which quads are tested.  The per-pixel test is written one FP32 operation
// at a time, as in the oracle, and uses the raymarcher's primary ray and its
// plane distance (p - o) * (1 / d), so t equals the DDA's entry t bit for bit
// when both find the same face.
#include <cmath>
#include <cstring>

#include "traverse.h"

namespace meso {

constexpr uint32_t kPoolNoSlot = 0xFFFFFFFFu;
constexpr int kSmall = 64;        // pixels tested by the quad's own lane
constexpr int kMedium = 4096;     // pixels tested by the quad's wave
constexpr int kSegPixels = 16384; // pixels per segment of a larger quad
constexpr uint32_t kMaxSegments = 1u << 22;

struct QuadGeom {
  int a, u, v;
  bool pos;
  float P;
  int32_t U0, V0, W, H, slice;
  int face;
};

__device__ __forceinline__ QuadGeom quad_decode(uint32_t q, uint32_t chunk, uint32_t nc) {
  QuadGeom g;
  const int32_t ox = (int32_t)(chunk % nc) * 32, oy = (int32_t)((chunk / nc) % nc) * 32,
                oz = (int32_t)(chunk / (nc * nc)) * 32;
  const int32_t x = ox + (int32_t)(q & 31u), y = oy + (int32_t)((q >> 5) & 31u),
                z = oz + (int32_t)((q >> 10) & 31u);
  g.W = (int32_t)((q >> 15) & 31u) + 1;
  g.H = (int32_t)((q >> 20) & 31u) + 1;
  g.face = (int)((q >> 25) & 7u);
  g.pos = (g.face & 1) == 0;
  if (g.face < 2) {
    g.a = 2; g.u = 0; g.v = 1; g.slice = z; g.U0 = x; g.V0 = y;
  } else if (g.face < 4) {
    g.a = 1; g.u = 0; g.v = 2; g.slice = y; g.U0 = x; g.V0 = z;
  } else {
    g.a = 0; g.u = 1; g.v = 2; g.slice = x; g.U0 = y; g.V0 = z;
  }
  g.P = (float)(g.slice + (g.pos ? 1 : 0));
  return g;
}

__device__ __forceinline__ float pick3(const float v[3], int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// t of the ray's hit on the quad, or -1
__device__ __forceinline__ float quad_hit(const QuadGeom &g, const float o[3], const float d[3]) {
  const float da = pick3(d, g.a);
  if (g.pos ? !(da < 0.0f) : !(da > 0.0f)) return -1.0f;
  const float inv = 1.0f / da;
  const float q = g.P - pick3(o, g.a);
  const float t = q * inv;
  if (!(t > 0.0f)) return -1.0f;
  const float du = pick3(d, g.u) * t, dv = pick3(d, g.v) * t;
  const float hu = pick3(o, g.u) + du, hv = pick3(o, g.v) + dv;
  if (!(hu >= (float)g.U0 && hu < (float)(g.U0 + g.W))) return -1.0f;
  if (!(hv >= (float)g.V0 && hv < (float)(g.V0 + g.H))) return -1.0f;
  return t;
}

// conservative pixel rectangle [r0, r1) x [r2, r3); false when empty.  The
// quad is clipped at z >= 1e-3 in front of the eye (one Sutherland-Hodgman
// pass) and the projected vertices of the clipped polygon are bounded as
// they are found, so no vertex array is kept.
__device__ __forceinline__ void bound_point(const DevCamera &c, double x, double y, double z,
                                            double b[4]) {
  const double px = (x / z / c.sx + 1.0) * 0.5 * c.width - 0.5;
  const double py = (1.0 - y / z / c.sy) * 0.5 * c.height - 0.5;
  b[0] = px < b[0] ? px : b[0];
  b[1] = px > b[1] ? px : b[1];
  b[2] = py < b[2] ? py : b[2];
  b[3] = py > b[3] ? py : b[3];
}

__device__ bool quad_rect(const DevCamera &c, const QuadGeom &g, int32_t r[4]) {
  double cx[4], cy[4], cz[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double pa = g.P, pu = g.U0 + ((k == 1 || k == 2) ? g.W : 0),
                 pv = g.V0 + (k >= 2 ? g.H : 0);
    const double p0 = g.a == 0 ? pa : (g.u == 0 ? pu : pv);
    const double p1 = g.a == 1 ? pa : (g.u == 1 ? pu : pv);
    const double p2 = g.a == 2 ? pa : (g.u == 2 ? pu : pv);
    const double e0 = p0 - c.eye[0], e1 = p1 - c.eye[1], e2 = p2 - c.eye[2];
    cx[k] = e0 * c.right[0] + e1 * c.right[1] + e2 * c.right[2];
    cy[k] = e0 * c.up[0] + e1 * c.up[1] + e2 * c.up[2];
    cz[k] = e0 * c.fwd[0] + e1 * c.fwd[1] + e2 * c.fwd[2];
  }
  const double zn = 1e-3;
  double b[4] = {1e30, -1e30, 1e30, -1e30};
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int l = (k + 1) & 3;
    const bool ia = cz[k] >= zn, ib = cz[l] >= zn;
    if (ia) {
      bound_point(c, cx[k], cy[k], cz[k], b);
      any = true;
    }
    if (ia != ib) {
      const double s = (zn - cz[k]) / (cz[l] - cz[k]);
      bound_point(c, cx[k] + (cx[l] - cx[k]) * s, cy[k] + (cy[l] - cy[k]) * s,
                  cz[k] + (cz[l] - cz[k]) * s, b);
      any = true;
    }
  }
  if (!any) return false;
  double fx0 = floor(b[0]) - 1.0, fx1 = ceil(b[1]) + 2.0, fy0 = floor(b[2]) - 1.0,
         fy1 = ceil(b[3]) + 2.0;
  fx0 = fx0 < 0.0 ? 0.0 : fx0;
  fy0 = fy0 < 0.0 ? 0.0 : fy0;
  fx1 = fx1 > c.width ? (double)c.width : fx1;
  fy1 = fy1 > c.height ? (double)c.height : fy1;
  if (!(fx0 < fx1 && fy0 < fy1)) return false;
  r[0] = (int32_t)fx0;
  r[1] = (int32_t)fx1;
  r[2] = (int32_t)fy0;
  r[3] = (int32_t)fy1;
  return true;
}

// false when no part of the 32^3 chunk at (ox, oy, oz) can reach a pixel:
// its 12 edges, clipped at z >= 1e-3, project outside the frame (or nothing
// is in front of the eye)
__device__ bool chunk_visible(const DevCamera &c, int32_t ox, int32_t oy, int32_t oz) {
  double cx[8], cy[8], cz[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const double e0 = ox + ((k & 1) ? 32 : 0) - (double)c.eye[0];
    const double e1 = oy + ((k & 2) ? 32 : 0) - (double)c.eye[1];
    const double e2 = oz + ((k & 4) ? 32 : 0) - (double)c.eye[2];
    cx[k] = e0 * c.right[0] + e1 * c.right[1] + e2 * c.right[2];
    cy[k] = e0 * c.up[0] + e1 * c.up[1] + e2 * c.up[2];
    cz[k] = e0 * c.fwd[0] + e1 * c.fwd[1] + e2 * c.fwd[2];
  }
  const double zn = 1e-3;
  double b[4] = {1e30, -1e30, 1e30, -1e30};
  bool any = false;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (cz[k] >= zn) {
      bound_point(c, cx[k], cy[k], cz[k], b);
      any = true;
    }
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1) {  // the edges k -> k | bit
      if (k & bit) continue;
      const int l = k | bit;
      if ((cz[k] >= zn) != (cz[l] >= zn)) {
        const double t = (zn - cz[k]) / (cz[l] - cz[k]);
        bound_point(c, cx[k] + (cx[l] - cx[k]) * t, cy[k] + (cy[l] - cy[k]) * t, zn, b);
        any = true;
      }
    }
  }
  if (!any) return false;
  return b[1] >= -2.0 && b[0] <= c.width + 1.0 && b[3] >= -2.0 && b[2] <= c.height + 1.0;
}

__device__ __forceinline__ void test_pixel(const DevCamera &cam, const QuadGeom &g, int32_t px,
                                           int32_t py, uint64_t id, unsigned long long *keys) {
  float o[3], d[3];
  primary_ray(cam, px, py, o, d);
  const float t = quad_hit(g, o, d);
  if (t < 0.0f) return;
  const unsigned long long key = ((unsigned long long)__float_as_uint(t) << 32) | id;
  atomicMin(&keys[(uint64_t)py * (uint32_t)cam.width + (uint32_t)px], key);
}

struct RasterScratch {
  unsigned long long *keys;  // width * height
  uint4 *segs;               // quad id, first pixel, x0 | y0 << 16, width | height << 16
  unsigned int *counters;    // [0] segments, [1] overflowed quads, [2] small, [3] medium + large
};

// pixels [j0, j1) of a quad's rectangle, row-major, a lane per pixel
__device__ __forceinline__ void wave_pixels(const DevCamera &cam, const QuadGeom &g, int32_t x0,
                                            int32_t y0, int32_t rw, uint32_t j0, uint32_t j1,
                                            uint64_t id, unsigned long long *keys) {
  for (uint32_t j = j0 + (threadIdx.x & 63u); j < j1; j += 64u)
    test_pixel(cam, g, x0 + (int32_t)(j % (uint32_t)rw), y0 + (int32_t)(j / (uint32_t)rw), id,
               keys);
}

__global__ void __launch_bounds__(256) raster_clear_kernel(RasterScratch s, uint64_t npx) {
  const uint64_t i = block_id() * 256u + threadIdx.x;
  if (i < npx) s.keys[i] = ~0ull;
  if (i < 4) s.counters[i] = 0u;
}

__global__ void __launch_bounds__(256) raster_quads_kernel(DevCamera cam, const uint32_t *slot_chunk,
                                                           const uint32_t *slot_count,
                                                           const uint32_t *quads, uint32_t per_slot,
                                                           uint32_t nc, RasterScratch s) {
  const uint32_t sl = blockIdx.x;
  const uint32_t chunk = slot_chunk[sl];
  if (chunk == kPoolNoSlot) return;
  // the whole chunk outside the frame: none of its quads can be hit
  if (!chunk_visible(cam, (int32_t)(chunk % nc) * 32, (int32_t)((chunk / nc) % nc) * 32,
                     (int32_t)(chunk / (nc * nc)) * 32))
    return;
  const uint32_t cnt = slot_count[sl] < per_slot ? slot_count[sl] : per_slot;
  const float eye[3] = {cam.eye[0], cam.eye[1], cam.eye[2]};
  // the trip count is the same for every lane, so the waves stay whole for
  // the wave-wide steps below
  for (uint32_t base = 0; base < cnt; base += 256u) {
    const uint32_t k = base + threadIdx.x;
    const uint64_t id = (uint64_t)sl * per_slot + k;
    int32_t r[4] = {0, 0, 0, 0};
    bool live = false;
    QuadGeom g;
    if (k < cnt) {
      g = quad_decode(quads[id], chunk, nc);
      const float e = pick3(eye, g.a);
      // back-facing quads cannot be hit from this eye
      live = (g.pos ? e > g.P : e < g.P) && quad_rect(cam, g, r);
    }
    const int32_t rw = r[1] - r[0], rh = r[3] - r[2];
    const int32_t area = live ? rw * rh : 0;
    if (live && area <= kSmall) {
      atomicAdd(&s.counters[2], 1u);
      for (int32_t py = r[2]; py < r[3]; py++)
        for (int32_t px = r[0]; px < r[1]; px++) test_pixel(cam, g, px, py, id, s.keys);
    }
    const bool medium = live && area > kSmall && area <= kMedium;
    if (live && area > kMedium) {
      const uint32_t nseg = ((uint32_t)area + kSegPixels - 1) / kSegPixels;
      atomicAdd(&s.counters[3], 1u);
      const uint32_t j = atomicAdd(&s.counters[0], nseg);
      if (j + nseg <= kMaxSegments) {
        for (uint32_t q = 0; q < nseg; q++)
          s.segs[j + q] = make_uint4((uint32_t)id, q * kSegPixels,
                                     (uint32_t)r[0] | ((uint32_t)r[2] << 16),
                                     (uint32_t)rw | ((uint32_t)rh << 16));
      } else {
        // the segment list is full: test the quad's pixels here (slow, never wrong)
        atomicAdd(&s.counters[1], 1u);
        for (int32_t py = r[2]; py < r[3]; py++)
          for (int32_t px = r[0]; px < r[1]; px++) test_pixel(cam, g, px, py, id, s.keys);
      }
    }
    // the wave takes its medium quads one at a time, 64 pixels a step
    uint64_t todo = __ballot(medium);
    if (medium) atomicAdd(&s.counters[3], 1u);
    while (todo) {
      const int L = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1ull;
      const uint32_t mk = (uint32_t)__shfl((int)k, L, 64);
      const int32_t mx0 = __shfl(r[0], L, 64), my0 = __shfl(r[2], L, 64);
      const int32_t mrw = __shfl(rw, L, 64), marea = __shfl(area, L, 64);
      const uint64_t mid = (uint64_t)sl * per_slot + mk;
      const QuadGeom mg = quad_decode(quads[mid], chunk, nc);
      wave_pixels(cam, mg, mx0, my0, mrw, 0u, (uint32_t)marea, mid, s.keys);
    }
  }
}

// persistent waves over the segments of the large quads
__global__ void __launch_bounds__(64) raster_segments_kernel(DevCamera cam, const uint32_t *slot_chunk,
                                                             const uint32_t *quads, uint32_t per_slot,
                                                             uint32_t nc, RasterScratch s) {
  const uint32_t n = s.counters[0] < kMaxSegments ? s.counters[0] : kMaxSegments;
  for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {
    const uint4 seg = s.segs[j];
    const uint32_t id = seg.x;
    const QuadGeom g = quad_decode(quads[id], slot_chunk[id / per_slot], nc);
    const int32_t x0 = (int32_t)(seg.z & 0xFFFFu), y0 = (int32_t)(seg.z >> 16);
    const int32_t rw = (int32_t)(seg.w & 0xFFFFu), rh = (int32_t)(seg.w >> 16);
    const uint32_t area = (uint32_t)rw * (uint32_t)rh;
    const uint32_t end = seg.y + kSegPixels < area ? seg.y + kSegPixels : area;
    wave_pixels(cam, g, x0, y0, rw, seg.y, end, id, s.keys);
  }
}

__global__ void __launch_bounds__(256) raster_resolve_kernel(DevCamera cam, DevShading sh,
                                                             const uint32_t *slot_chunk,
                                                             const uint32_t *quads,
                                                             uint32_t per_slot, uint32_t nc,
                                                             int32_t n, RayOut out,
                                                             const unsigned long long *keys) {
  const uint64_t i = block_id() * 256u + threadIdx.x;
  const uint64_t npx = (uint64_t)cam.width * (uint32_t)cam.height;
  if (i >= npx) return;
  const unsigned long long key = keys[i];
  int64_t index = -1;
  int32_t face = kFaceMiss;
  float t = __builtin_inff(), shade = 0.0f;
  if (key != ~0ull) {
    const uint32_t id = (uint32_t)key;
    const uint32_t q = quads[id];
    const QuadGeom g = quad_decode(q, slot_chunk[id / per_slot], nc);
    const int32_t px = (int32_t)(i % (uint32_t)cam.width), py = (int32_t)(i / (uint32_t)cam.width);
    float o[3], d[3];
    primary_ray(cam, px, py, o, d);
    t = __uint_as_float((uint32_t)(key >> 32));
    const float du = pick3(d, g.u) * t, dv = pick3(d, g.v) * t;
    const float hu = pick3(o, g.u) + du, hv = pick3(o, g.v) + dv;
    int32_t cu = (int32_t)__builtin_floorf(hu), cv = (int32_t)__builtin_floorf(hv);
    cu = cu < g.U0 ? g.U0 : (cu > g.U0 + g.W - 1 ? g.U0 + g.W - 1 : cu);
    cv = cv < g.V0 ? g.V0 : (cv > g.V0 + g.H - 1 ? g.V0 + g.H - 1 : cv);
    int32_t cell[3];
#pragma unroll
    for (int k = 0; k < 3; k++) cell[k] = k == g.a ? g.slice : (k == g.u ? cu : cv);
    index = lin3(cell[0], cell[1], cell[2], n);
    face = g.face;
    // the primary-only shade of the raymarcher: no shadow or AO rays
    const float lightc = pick3(sh.light, g.a);
    float lambert = g.pos ? lightc : -lightc;
    if (lambert < 0.0f) lambert = 0.0f;
    shade = shade_value(lambert, 1.0f, false, 0, 1);
  }
  if (out.index_bytes == 8)
    ((int64_t *)out.index)[i] = index;
  else
    ((int32_t *)out.index)[i] = (int32_t)index;
  if (out.face) out.face[i] = (uint8_t)face;
  if (out.t) out.t[i] = t;
  if (out.shade) out.shade[i] = shade;
}

}  // namespace meso

using namespace meso;

struct meso_raster {
  meso_context *ctx = nullptr;
  int32_t max_width = 0, max_height = 0;
  RasterScratch s{};
};

extern "C" {

meso_status meso_raster_create(meso_context *ctx, int32_t max_width, int32_t max_height,
                               meso_raster **out) {
  MESO_CHECK(ctx && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  *out = nullptr;
  MESO_CHECK(max_width >= 1 && max_height >= 1 && max_width <= 16384 && max_height <= 16384,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "frame %dx%d outside 1..16384", max_width,
             max_height);
  meso_raster *r = new meso_raster;
  r->ctx = ctx;
  r->max_width = max_width;
  r->max_height = max_height;
  const uint64_t npx = (uint64_t)max_width * (uint64_t)max_height;
  if (hipMalloc((void **)&r->s.keys, npx * 8) != hipSuccess ||
      hipMalloc((void **)&r->s.segs, (uint64_t)kMaxSegments * 16) != hipSuccess ||
      hipMalloc((void **)&r->s.counters, 16) != hipSuccess) {
    (void)hipGetLastError();
    if (r->s.keys) (void)hipFree(r->s.keys);
    if (r->s.segs) (void)hipFree(r->s.segs);
    if (r->s.counters) (void)hipFree(r->s.counters);
    delete r;
    return fail(MESO_ERR_OUT_OF_MEMORY, "raster: cannot allocate its scratch");
  }
  *out = r;
  return MESO_OK;
}

meso_status meso_raster_destroy(meso_raster *r) {
  if (!r) return MESO_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(r->s.keys);
  (void)hipFree(r->s.segs);
  (void)hipFree(r->s.counters);
  delete r;
  return MESO_OK;
}

meso_status meso_raster_pool(meso_raster *r, meso_chunk_pool *pool, const meso_camera *cam,
                             const meso_shading *shading, const meso_ray_outputs *out,
                             meso_stream stream) {
  MESO_CHECK(r && pool && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  meso_status st = check_camera(cam);
  if (st != MESO_OK) return st;
  MESO_CHECK(cam->width <= r->max_width && cam->height <= r->max_height,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "frame %dx%d larger than the raster's %dx%d",
             cam->width, cam->height, r->max_width, r->max_height);
  MESO_CHECK(out->size >= sizeof(meso_ray_outputs), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_ray_outputs.size too small");
  MESO_CHECK(out->hit_index, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "hit_index is NULL");
  MESO_CHECK(out->index_bytes == 4 || out->index_bytes == 8, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "index_bytes must be 4 or 8");
  const PoolView pv = pool_view(pool);
  MESO_CHECK(pv.scene->ctx == r->ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "pool of another context");
  const uint64_t nvox = (uint64_t)pv.scene->n * pv.scene->n * pv.scene->n;
  MESO_CHECK(out->index_bytes == 8 || nvox <= 0x7FFFFFFFull, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "scene of %llu voxels needs 8-byte hit indices", (unsigned long long)nvox);
  MESO_CHECK((uint64_t)pv.capacity * (uint64_t)pv.per_slot <= 0xFFFFFFFFull,
             MESO_ERR_NOT_SUPPORTED, "pool of more than 2^32 quad places");
  DevShading sh{};
  if (shading && shading->size >= sizeof(meso_shading) &&
      (shading->light[0] != 0.0f || shading->light[1] != 0.0f || shading->light[2] != 0.0f))
    derive_shading(shading, &sh);
  else
    sh.light[1] = 1.0f;
  DevCamera dc;
  derive_camera(cam, &dc);
  RayOut ro{out->hit_index, out->index_bytes, out->face, out->t, out->shade, nullptr};
  const hipStream_t s = stream_of(r->ctx, stream);
  const uint64_t npx = (uint64_t)cam->width * (uint64_t)cam->height;
  hipLaunchKernelGGL(raster_clear_kernel, grid_1d((npx + 255) / 256), dim3(256), 0, s, r->s, npx);
  MESO_HIP(hipGetLastError());
  hipLaunchKernelGGL(raster_quads_kernel, dim3((uint32_t)pv.capacity), dim3(256), 0, s, dc,
                     pv.slot_chunk, pv.slot_count, pv.quads, (uint32_t)pv.per_slot,
                     (uint32_t)pv.nc, r->s);
  MESO_HIP(hipGetLastError());
  const uint32_t waves = (uint32_t)r->ctx->props.multiProcessorCount * 16u;
  hipLaunchKernelGGL(raster_segments_kernel, dim3(waves), dim3(64), 0, s, dc, pv.slot_chunk,
                     pv.quads, (uint32_t)pv.per_slot, (uint32_t)pv.nc, r->s);
  MESO_HIP(hipGetLastError());
  hipLaunchKernelGGL(raster_resolve_kernel, grid_1d((npx + 255) / 256), dim3(256), 0, s, dc, sh,
                     pv.slot_chunk, pv.quads, (uint32_t)pv.per_slot, (uint32_t)pv.nc,
                     pv.scene->n, ro, r->s.keys);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

meso_status meso_raster_stats_get(meso_raster *r, meso_raster_stats *out, meso_stream stream) {
  MESO_CHECK(r && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(out->size >= sizeof(meso_raster_stats), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_raster_stats too small");
  const hipStream_t s = stream_of(r->ctx, stream);
  uint32_t c[4];
  MESO_HIP(hipMemcpyAsync(c, r->s.counters, 16, hipMemcpyDeviceToHost, s));
  MESO_HIP(hipStreamSynchronize(s));
  out->segments = c[0];
  out->overflow_quads = c[1];
  out->small_quads = c[2];
  out->large_quads = c[3];
  return MESO_OK;
}

}  // extern "C"
