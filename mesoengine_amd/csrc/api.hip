// C-ABI entry points: errors, context, buffers, streams, events, scenes,
// raymarch.  Kernels live in raymarch.hip / mesh.hip / voxelize.hip.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "meso_internal.h"
#include "traverse.h"

namespace meso {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

meso_status fail(meso_status code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  // the error is now this call's status: consume the runtime's per-thread
  // last error too, or the next launch check (hipGetLastError after a
  // kernel launch) would report a failure that was already returned
  (void)hipGetLastError();
  return code;
}

hipStream_t stream_of(meso_context *ctx, meso_stream s) {
  return s ? (hipStream_t)s : ctx->stream;
}

__global__ void zero_kernel(uint4 *p, uint64_t n16, uint8_t *tail, uint32_t ntail) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    p[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// Clear device memory with a kernel rather than hipMemsetAsync, so that
// callers may capture it into a graph.  ptr must be 16-byte aligned.
meso_status launch_zero(void *ptr, uint64_t bytes, hipStream_t stream) {
  if (!bytes) return MESO_OK;
  MESO_CHECK(((uintptr_t)ptr & 15u) == 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "zero-fill pointer must be 16-byte aligned");
  const uint64_t n16 = bytes / 16;
  const uint32_t ntail = (uint32_t)(bytes % 16);
  uint64_t blocks = (n16 + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(zero_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, (uint4 *)ptr, n16,
                     (uint8_t *)ptr + n16 * 16, ntail);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

meso_status band_queue(meso_context *ctx, hipStream_t stream, unsigned int **out) {
  std::lock_guard<std::mutex> lk(ctx->queue_mu);
  auto it = ctx->band_queues.find(stream);
  if (it != ctx->band_queues.end()) {
    *out = it->second;
    return MESO_OK;
  }
  unsigned int *q = nullptr;
  MESO_HIP(hipMalloc((void **)&q, 1024));
  // cleared on the stream that uses it: a hipMemset runs on the null stream,
  // which the context's non-blocking streams do not wait for, so the first
  // launch could read the counters before they are zero (and a clear landing
  // during that launch leaves them past the end for every later launch)
  const meso_status st = launch_zero(q, 1024, stream);
  if (st != MESO_OK) {
    hipFree(q);
    return st;
  }
  ctx->band_queues[stream] = q;
  *out = q;
  return MESO_OK;
}

// ------------------------------------------------------------------ empty-space levels
// Brick map: one bit per 32^3 and per 128^3 voxel block holding a brick.
__global__ void coarse_brickmap_kernel(const uint32_t *slots, uint32_t nb, uint32_t *occ32,
                                       uint32_t *occ128) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)nb * nb * nb || slots[i] == kEmptySlot) return;
  const uint32_t bx = (uint32_t)(i % nb), by = (uint32_t)((i / nb) % nb),
                 bz = (uint32_t)(i / ((uint64_t)nb * nb));
  const uint32_t n32 = (nb + 3) / 4, n128 = (nb + 15) / 16;
  const uint32_t i32 = (bx >> 2) + ((by >> 2) + (bz >> 2) * n32) * n32;
  const uint32_t i128 = (bx >> 4) + ((by >> 4) + (bz >> 4) * n128) * n128;
  atomicOr(&occ32[i32 >> 5], 1u << (i32 & 31u));
  atomicOr(&occ128[i128 >> 5], 1u << (i128 & 31u));
}

// Two-level: per node, one bit per 4^3-brick group holding a brick.
__global__ void coarse_groups_kernel(const uint32_t *node_slots, uint64_t n_nodes,
                                     unsigned long long *groups) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes * 4096u || node_slots[i] == kEmptySlot) return;
  const uint32_t l = (uint32_t)(i % 4096u);
  const uint32_t bx = l & 15u, by = (l >> 4) & 15u, bz = l >> 8;
  atomicOr(&groups[i / 4096u], 1ull << ((bx >> 2) + 4u * (by >> 2) + 16u * (bz >> 2)));
}

// Both brick kinds: per 32^3 block, one bit per allocated brick of its 4^3.
__global__ void bricks32_kernel(const uint32_t *top, const uint32_t *slots, uint32_t nb,
                                int32_t two_level, uint64_t cells, uint32_t n32,
                                unsigned long long *bricks32) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  uint32_t bx, by, bz;
  if (two_level) {
    const uint32_t nt = nb / 16u;
    const uint64_t t = i / 4096u;
    const uint32_t l = (uint32_t)(i % 4096u), node = top[t];
    if (node == kEmptySlot || slots[(uint64_t)node * 4096u + l] == kEmptySlot) return;
    bx = (uint32_t)(t % nt) * 16u + (l & 15u);
    by = (uint32_t)((t / nt) % nt) * 16u + ((l >> 4) & 15u);
    bz = (uint32_t)(t / ((uint64_t)nt * nt)) * 16u + (l >> 8);
  } else {
    if (slots[i] == kEmptySlot) return;
    bx = (uint32_t)(i % nb);
    by = (uint32_t)((i / nb) % nb);
    bz = (uint32_t)(i / ((uint64_t)nb * nb));
  }
  const uint64_t w = (bx >> 2) + ((uint64_t)(by >> 2) + (uint64_t)(bz >> 2) * n32) * n32;
  atomicOr(&bricks32[w], 1ull << ((bx & 3u) + 4u * (by & 3u) + 16u * (bz & 3u)));
}

// min / max brick coordinates of the allocated bricks: box[0..2] min,
// box[3..5] max (one atomic per wave)
// The box of the cells a thread saw, in registers.
struct BoxAcc {
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {-1, -1, -1};
  __device__ __forceinline__ void add(uint32_t bx, uint32_t by, uint32_t bz) {
    const int c[3] = {(int)bx, (int)by, (int)bz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      mn[a] = c[a] < mn[a] ? c[a] : mn[a];
      mx[a] = c[a] > mx[a] ? c[a] : mx[a];
    }
  }
};

// Merges the boxes of a wave's lanes, then into `box` with one atomic per
// bound (the kernels loop over their cells, so a wave merges once).
__device__ __forceinline__ void box_reduce(BoxAcc acc, int *box) {
  int *mn = acc.mn, *mx = acc.mx;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int o1 = __shfl_xor(mn[a], off, 64), o2 = __shfl_xor(mx[a], off, 64);
      mn[a] = o1 < mn[a] ? o1 : mn[a];
      mx[a] = o2 > mx[a] ? o2 : mx[a];
    }
  if ((threadIdx.x & 63u) == 0 && mx[0] >= 0)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      atomicMin(&box[a], mn[a]);
      atomicMax(&box[3 + a], mx[a]);
    }
}

__global__ void box_brickmap_kernel(const uint32_t *slots, uint32_t nb, int *box) {
  const uint64_t cells = (uint64_t)nb * nb * nb;
  BoxAcc acc;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (slots[i] != kEmptySlot)
      acc.add((uint32_t)(i % nb), (uint32_t)((i / nb) % nb), (uint32_t)(i / ((uint64_t)nb * nb)));
  box_reduce(acc, box);
}

__global__ void box_two_level_kernel(const uint32_t *top, uint32_t nt, const uint32_t *node_slots,
                                     int *box) {
  const uint64_t cells = (uint64_t)nt * nt * nt * 4096u;
  BoxAcc acc;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = i / 4096u;
    const uint32_t l = (uint32_t)(i % 4096u);
    const uint32_t node = top[t];
    if (node == kEmptySlot || node_slots[(uint64_t)node * 4096u + l] == kEmptySlot) continue;
    acc.add((uint32_t)(t % nt) * 16u + (l & 15u), (uint32_t)((t / nt) % nt) * 16u + ((l >> 4) & 15u),
            (uint32_t)(t / ((uint64_t)nt * nt)) * 16u + (l >> 8));
  }
  box_reduce(acc, box);
}

// the box of the allocated bricks, in voxels (empty when there are none)
static meso_status build_box(meso_scene *sc, hipStream_t stream) {
  int init[6] = {INT_MAX, INT_MAX, INT_MAX, -1, -1, -1}, res[6];
  int *d = nullptr;
  MESO_HIP(hipMalloc((void **)&d, sizeof(init)));
  hipError_t e = hipMemcpyAsync(d, init, sizeof(init), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) {
    if (sc->kind == MESO_SCENE_BRICKMAP_U16) {
      const uint32_t nb = (uint32_t)sc->n / 8;
      const uint64_t cells = (uint64_t)nb * nb * nb;
      hipLaunchKernelGGL(box_brickmap_kernel, dim3((uint32_t)std::min<uint64_t>((cells + 255) / 256, 1024)),
                         dim3(256), 0, stream, sc->slots, nb, d);
    } else {
      const uint32_t nt = (uint32_t)sc->n / 128;
      const uint64_t cells = (uint64_t)nt * nt * nt * 4096u;
      hipLaunchKernelGGL(box_two_level_kernel, dim3((uint32_t)std::min<uint64_t>((cells + 255) / 256, 1024)),
                         dim3(256), 0, stream, sc->top, nt, sc->slots, d);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(res, d, sizeof(res), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(MESO_ERR_RUNTIME, "scene box: %s", hipGetErrorString(e));
  const bool empty = res[3] < 0;
  for (int a = 0; a < 3; a++) {
    sc->box_lo[a] = empty ? 0 : res[a] * 8;
    sc->box_hi[a] = empty ? 0 : (res[3 + a] + 1) * 8;
  }
  return MESO_OK;
}

// Per brick cell (index i into slots): the 2^3-block bits of its brick, from
// the brick's mask (0 for an empty cell), and 1 + the brick's cell.
__global__ void brick_blocks_kernel(const uint32_t *slots, uint64_t cells, const uint64_t *masks,
                                    uint64_t *blocks, uint32_t *slot_cell) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const uint32_t slot = slots[i];
  blocks[i] = slot == kEmptySlot ? 0ull : brick_blocks2(masks + (uint64_t)slot * 8u);
  if (slot != kEmptySlot) slot_cell[slot] = (uint32_t)i + 1u;
}

meso_status build_coarse(meso_scene *sc, hipStream_t stream) {
  if (sc->kind == MESO_SCENE_BRICKMAP_U16 || sc->kind == MESO_SCENE_TWO_LEVEL_U16) {
    // cells = entries of slots (below 2^30: edges are at most 8192)
    const uint64_t nb = (uint64_t)sc->n / 8;
    const uint64_t cells = sc->kind == MESO_SCENE_BRICKMAP_U16 ? nb * nb * nb
                                                               : (uint64_t)sc->n_nodes * 4096u;
    const uint64_t nbr = (uint64_t)sc->n_bricks;
    MESO_HIP(hipMalloc((void **)&sc->blocks, (cells > 0 ? cells : 1) * 8));
    MESO_HIP(hipMalloc((void **)&sc->slot_cell, (nbr > 0 ? nbr : 1) * 4));
    sc->bytes += cells * 8 + nbr * 4;
    meso_status st = launch_zero(sc->slot_cell, (nbr > 0 ? nbr : 1) * 4, stream);
    if (st != MESO_OK) return st;
    if (cells > 0) {
      hipLaunchKernelGGL(brick_blocks_kernel, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0,
                         stream, sc->slots, cells, sc->masks, sc->blocks, sc->slot_cell);
      MESO_HIP(hipGetLastError());
    }
  }
  if (sc->kind == MESO_SCENE_BRICKMAP_U16) {
    const uint64_t nb = (uint64_t)sc->n / 8;
    const uint64_t n32 = (nb + 3) / 4, n128 = (nb + 15) / 16;
    const uint64_t w32 = (n32 * n32 * n32 + 31) / 32, w128 = (n128 * n128 * n128 + 31) / 32;
    MESO_HIP(hipMalloc((void **)&sc->occ32, w32 * 4));
    MESO_HIP(hipMalloc((void **)&sc->occ128, w128 * 4));
    sc->bytes += (w32 + w128) * 4;
    meso_status st = launch_zero(sc->occ32, w32 * 4, stream);
    if (st == MESO_OK) st = launch_zero(sc->occ128, w128 * 4, stream);
    if (st != MESO_OK) return st;
    const uint64_t cells = nb * nb * nb;
    hipLaunchKernelGGL(coarse_brickmap_kernel, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0,
                       stream, sc->slots, (uint32_t)nb, sc->occ32, sc->occ128);
    MESO_HIP(hipGetLastError());
  } else if (sc->kind == MESO_SCENE_TWO_LEVEL_U16) {
    const uint64_t nn = (uint64_t)(sc->n_nodes > 0 ? sc->n_nodes : 1);
    MESO_HIP(hipMalloc((void **)&sc->groups, nn * 8));
    sc->bytes += nn * 8;
    meso_status st = launch_zero(sc->groups, nn * 8, stream);
    if (st != MESO_OK) return st;
    if (sc->n_nodes > 0) {
      const uint64_t cells = (uint64_t)sc->n_nodes * 4096u;
      hipLaunchKernelGGL(coarse_groups_kernel, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0,
                         stream, sc->slots, (uint64_t)sc->n_nodes,
                         (unsigned long long *)sc->groups);
      MESO_HIP(hipGetLastError());
    }
  }
  if (sc->kind == MESO_SCENE_BRICKMAP_U16 || sc->kind == MESO_SCENE_TWO_LEVEL_U16) {
    const uint32_t nb = (uint32_t)sc->n / 8, n32 = (nb + 3) / 4;
    const uint64_t words = (uint64_t)n32 * n32 * n32;
    MESO_HIP(hipMalloc((void **)&sc->bricks32, words * 8));
    sc->bytes += words * 8;
    meso_status st = launch_zero(sc->bricks32, words * 8, stream);
    if (st != MESO_OK) return st;
    const bool two = sc->kind == MESO_SCENE_TWO_LEVEL_U16;
    const uint64_t cells =
        two ? (uint64_t)(nb / 16) * (nb / 16) * (nb / 16) * 4096u : (uint64_t)nb * nb * nb;
    hipLaunchKernelGGL(bricks32_kernel, dim3((uint32_t)((cells + 255) / 256)), dim3(256), 0, stream,
                       sc->top, sc->slots, nb, two ? 1 : 0, cells, n32,
                       (unsigned long long *)sc->bricks32);
    MESO_HIP(hipGetLastError());
  }
  MESO_HIP(hipStreamSynchronize(stream));
  if (sc->kind == MESO_SCENE_BRICKMAP_U16 || sc->kind == MESO_SCENE_TWO_LEVEL_U16)
    return build_box(sc, stream);
  return MESO_OK;
}

void free_scene_arrays(meso_scene *sc) {
  hipSetDevice(sc->ctx->device);
  for (void *p : {(void *)sc->dense, (void *)sc->slots, (void *)sc->masks, (void *)sc->blocks,
                  (void *)sc->slot_cell, (void *)sc->voxels, (void *)sc->top, (void *)sc->occ32, (void *)sc->occ128,
                  (void *)sc->groups, (void *)sc->bricks32})
    if (p) hipFree(p);
  sc->dense = nullptr;
  sc->slots = nullptr;
  sc->masks = nullptr;
  sc->blocks = nullptr;
  sc->slot_cell = nullptr;
  sc->voxels = nullptr;
  sc->top = nullptr;
  sc->occ32 = sc->occ128 = nullptr;
  sc->groups = nullptr;
  sc->bricks32 = nullptr;
}

// ------------------------------------------------------------------ RNG
uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint32_t Pcg32::next() {
  uint64_t old = state;
  state = old * 6364136223846793005ull + inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
void Pcg32::srandom(uint64_t initstate, uint64_t initseq) {
  state = 0;
  inc = (initseq << 1u) | 1u;
  next();
  state += initstate;
  next();
}
void Pcg32::seed(uint64_t s) {
  uint64_t sm = s;
  uint64_t a = splitmix64(&sm);
  uint64_t b = splitmix64(&sm);
  srandom(a, b);
}

// ------------------------------------------------------------------ camera
static void normalize3(float v[3]) {
  float xx = v[0] * v[0], yy = v[1] * v[1], zz = v[2] * v[2];
  float s = xx + yy;
  s = s + zz;
  float len = std::sqrt(s);
  v[0] = v[0] / len;
  v[1] = v[1] / len;
  v[2] = v[2] / len;
}
static void cross3(const float a[3], const float b[3], float out[3]) {
  float t0 = a[1] * b[2], t1 = a[2] * b[1];
  float t2 = a[2] * b[0], t3 = a[0] * b[2];
  float t4 = a[0] * b[1], t5 = a[1] * b[0];
  out[0] = t0 - t1;
  out[1] = t2 - t3;
  out[2] = t4 - t5;
}

// Vertical-FOV pinhole: FVoxelCamera::GetProjectionMatrix uses
// glm::perspective with the vertical fov (Runtimes/Instance/VoxelCamera.cpp:12-17);
// forward is eye->target (VoxelCamera.cpp:24-29 takes -row 2 of the view).
void derive_camera(const meso_camera *c, DevCamera *o) {
  for (int i = 0; i < 3; i++) {
    o->eye[i] = c->eye[i];
    o->fwd[i] = c->target[i] - c->eye[i];
  }
  normalize3(o->fwd);
  cross3(o->fwd, c->up, o->right);
  normalize3(o->right);
  cross3(o->right, o->fwd, o->up);
  float th = (float)std::tan((double)c->vfov_deg * 3.14159265358979323846 / 360.0);
  float aspect = (float)c->width / (float)c->height;
  o->sx = th * aspect;
  o->sy = th;
  o->inv_w2 = 2.0f / (float)c->width;
  o->inv_h2 = 2.0f / (float)c->height;
  o->width = c->width;
  o->height = c->height;
}

void derive_shading(const meso_shading *s, DevShading *o) {
  float l[3] = {s->light[0], s->light[1], s->light[2]};
  normalize3(l);
  for (int i = 0; i < 3; i++) o->light[i] = l[i];
  o->ao_seed = s->ao_seed;
  o->ao_samples = s->ao_samples;
  o->ao_radius = s->ao_radius;
}

meso_status check_camera(const meso_camera *cam) {
  MESO_CHECK(cam, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "camera is NULL");
  MESO_CHECK(cam->size >= sizeof(meso_camera), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_camera.size %u < %zu", cam->size, sizeof(meso_camera));
  MESO_CHECK(cam->width > 0 && cam->height > 0 && cam->width <= 16384 && cam->height <= 16384,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "image %dx%d outside 1..16384", cam->width,
             cam->height);
  MESO_CHECK(cam->vfov_deg > 0.0f && cam->vfov_deg < 180.0f, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "vfov %f outside (0,180)", (double)cam->vfov_deg);
  for (int k = 0; k < 3; k++)
    MESO_CHECK(std::isfinite(cam->eye[k]) && std::isfinite(cam->target[k]) &&
                   std::isfinite(cam->up[k]),
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "camera eye / target / up must be finite");
  float f[3] = {cam->target[0] - cam->eye[0], cam->target[1] - cam->eye[1],
                cam->target[2] - cam->eye[2]};
  float c[3];
  cross3(f, cam->up, c);
  MESO_CHECK(c[0] != 0.0f || c[1] != 0.0f || c[2] != 0.0f, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "camera up is parallel to the view direction (or eye == target)");
  return MESO_OK;
}

}  // namespace meso

using namespace meso;

// ------------------------------------------------------------------ ABI
extern "C" {

const char *meso_last_error_message(void) { return g_error.c_str(); }
uint32_t meso_abi_version(void) { return MESO_ABI_VERSION; }

meso_status meso_context_create(int32_t device, meso_context **out) {
  MESO_CHECK(out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "out is NULL");
  *out = nullptr;
  int count = 0;
  MESO_HIP(hipGetDeviceCount(&count));
  MESO_CHECK(device >= 0 && device < count, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "device %d not in [0,%d)", device, count);
  MESO_HIP(hipSetDevice(device));
  auto *ctx = new meso_context();
  ctx->device = device;
  hipError_t e = hipGetDeviceProperties(&ctx->props, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(MESO_ERR_RUNTIME, "context setup failed: %s", hipGetErrorString(e));
  }
  const meso_status st = init_mesh_scratch(ctx);
  if (st != MESO_OK) {
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return st;
  }
  *out = ctx;
  return MESO_OK;
}

meso_status meso_context_destroy(meso_context *ctx) {
  if (!ctx) return MESO_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) {
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
  }
  if (!ctx->band_queues.empty()) {
    hipDeviceSynchronize();
    for (auto &kv : ctx->band_queues) hipFree(kv.second);
  }
  delete ctx;
  return MESO_OK;
}


meso_status meso_device_info_get(meso_context *ctx, meso_device_info *out) {
  MESO_CHECK(ctx && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  out->size = sizeof(*out);
  std::snprintf(out->name, sizeof(out->name), "%s", ctx->props.gcnArchName);
  out->compute_units = ctx->props.multiProcessorCount;
  out->xcds = 8;
  out->lds_bytes_per_cu = (int32_t)ctx->props.maxSharedMemoryPerMultiProcessor;
  out->l2_bytes = ctx->props.l2CacheSize;
  out->hbm_bytes = ctx->props.totalGlobalMem;
  out->clock_khz = ctx->props.clockRate;
  out->wavefront = ctx->props.warpSize;
  return MESO_OK;
}

// ---------------- buffers
meso_status meso_buffer_create(meso_context *ctx, uint64_t size, const void *init,
                               meso_buffer **out) {
  MESO_CHECK(ctx && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  *out = nullptr;
  MESO_CHECK(size > 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "buffer size is 0");
  MESO_HIP(hipSetDevice(ctx->device));
  void *p = nullptr;
  MESO_HIP(hipMalloc(&p, size));
  if (init) {
    hipError_t e = hipMemcpyAsync(p, init, size, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      hipFree(p);
      return fail(MESO_ERR_RUNTIME, "initial upload failed: %s", hipGetErrorString(e));
    }
  }
  auto *b = new meso_buffer();
  b->ctx = ctx;
  b->ptr = p;
  b->size = size;
  *out = b;
  return MESO_OK;
}

meso_status meso_buffer_destroy(meso_buffer *buf) {
  if (!buf) return MESO_OK;
  hipSetDevice(buf->ctx->device);
  hipFree(buf->ptr);
  delete buf;
  return MESO_OK;
}

meso_status meso_buffer_upload(meso_buffer *buf, const void *src, uint64_t size, uint64_t offset) {
  MESO_CHECK(buf && src, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(offset <= buf->size && size <= buf->size - offset, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "upload [%llu,+%llu) outside buffer of %llu bytes", (unsigned long long)offset,
             (unsigned long long)size, (unsigned long long)buf->size);
  // on the context stream: after the work queued there that may still read
  // the buffer, and finished before the call returns
  MESO_HIP(hipMemcpyAsync((char *)buf->ptr + offset, src, size, hipMemcpyHostToDevice,
                          buf->ctx->stream));
  MESO_HIP(hipStreamSynchronize(buf->ctx->stream));
  return MESO_OK;
}

meso_status meso_buffer_download(meso_buffer *buf, void *dst, uint64_t size, uint64_t offset) {
  MESO_CHECK(buf && dst, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(offset <= buf->size && size <= buf->size - offset, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "download [%llu,+%llu) outside buffer of %llu bytes", (unsigned long long)offset,
             (unsigned long long)size, (unsigned long long)buf->size);
  // on the context stream: after the work queued there that writes the buffer
  MESO_HIP(hipMemcpyAsync(dst, (const char *)buf->ptr + offset, size, hipMemcpyDeviceToHost,
                          buf->ctx->stream));
  MESO_HIP(hipStreamSynchronize(buf->ctx->stream));
  return MESO_OK;
}

uint64_t meso_buffer_device_address(meso_buffer *buf, uint64_t offset) {
  if (!buf || offset >= buf->size || (offset & 7u)) {
    set_error("device address offset invalid or not 8-byte aligned");
    return 0;
  }
  return (uint64_t)(uintptr_t)buf->ptr + offset;
}

uint64_t meso_buffer_size(meso_buffer *buf) { return buf ? buf->size : 0; }

meso_status meso_buffer_fill_zero(meso_buffer *buf, meso_stream stream) {
  MESO_CHECK(buf, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL buffer");
  return launch_zero(buf->ptr, buf->size, stream_of(buf->ctx, stream));
}

// ---------------- streams / events
meso_status meso_stream_create(meso_context *ctx, meso_stream *out) {
  MESO_CHECK(ctx && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipStream_t s;
  MESO_HIP(hipSetDevice(ctx->device));
  MESO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = (meso_stream)s;
  return MESO_OK;
}
meso_status meso_stream_destroy(meso_context *ctx, meso_stream stream) {
  MESO_CHECK(ctx && stream, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  MESO_HIP(hipStreamSynchronize(s));
  {
    // the stream's raymarch counters go with it (a later stream may get the
    // same handle; it then starts from a fresh, zeroed set)
    std::lock_guard<std::mutex> lk(ctx->queue_mu);
    auto it = ctx->band_queues.find(s);
    if (it != ctx->band_queues.end()) {
      hipFree(it->second);
      ctx->band_queues.erase(it);
    }
  }
  MESO_HIP(hipStreamDestroy(s));
  return MESO_OK;
}
meso_status meso_stream_sync(meso_context *ctx, meso_stream stream) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL context");
  MESO_HIP(hipStreamSynchronize(stream_of(ctx, stream)));
  return MESO_OK;
}
meso_status meso_event_create(meso_context *ctx, meso_event **out) {
  MESO_CHECK(ctx && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipEvent_t e;
  MESO_HIP(hipEventCreate(&e));
  auto *ev = new meso_event();
  ev->ctx = ctx;
  ev->ev = e;
  *out = ev;
  return MESO_OK;
}
meso_status meso_event_destroy(meso_event *ev) {
  if (!ev) return MESO_OK;
  hipEventDestroy(ev->ev);
  delete ev;
  return MESO_OK;
}
meso_status meso_event_record(meso_event *ev, meso_stream stream) {
  MESO_CHECK(ev, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL event");
  MESO_HIP(hipEventRecord(ev->ev, stream_of(ev->ctx, stream)));
  return MESO_OK;
}
meso_status meso_stream_wait_event(meso_context *ctx, meso_stream stream, meso_event *ev) {
  MESO_CHECK(ctx && ev, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_HIP(hipStreamWaitEvent(stream_of(ctx, stream), ev->ev, 0));
  return MESO_OK;
}
meso_status meso_event_sync(meso_event *ev) {
  MESO_CHECK(ev, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL event");
  MESO_HIP(hipEventSynchronize(ev->ev));
  return MESO_OK;
}
meso_status meso_event_elapsed_ms(meso_event *a, meso_event *b, float *ms) {
  MESO_CHECK(a && b && ms, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_HIP(hipEventElapsedTime(ms, a->ev, b->ev));
  return MESO_OK;
}
meso_status meso_memcpy_d2h(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                            meso_stream stream) {
  MESO_CHECK(ctx && dst && src, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  MESO_HIP(hipStreamSynchronize(s));
  return MESO_OK;
}
meso_status meso_memcpy_h2d(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                            meso_stream stream) {
  MESO_CHECK(ctx && dst && src, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  MESO_HIP(hipStreamSynchronize(s));
  return MESO_OK;
}

meso_status meso_host_alloc(meso_context *ctx, uint64_t bytes, void **out) {
  MESO_CHECK(ctx && out && bytes, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument or size 0");
  *out = nullptr;
  MESO_HIP(hipSetDevice(ctx->device));
  MESO_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  return MESO_OK;
}

meso_status meso_host_free(meso_context *ctx, void *ptr) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  if (ptr) MESO_HIP(hipHostFree(ptr));
  return MESO_OK;
}

meso_status meso_memcpy_h2d_async(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                                  meso_stream stream) {
  MESO_CHECK(ctx && dst && src, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  return MESO_OK;
}

meso_status meso_memcpy_d2d(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                            meso_stream stream) {
  MESO_CHECK(ctx && dst && src, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
  return MESO_OK;
}

// ---------------- scenes
static meso_status scene_alloc(meso_context *ctx, void **p, uint64_t bytes, const void *src,
                               meso_scene *sc) {
  if (!bytes) {
    *p = nullptr;
    return MESO_OK;
  }
  MESO_HIP(hipMalloc(p, bytes));
  sc->bytes += bytes;
  if (src) MESO_HIP(hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice));
  return MESO_OK;
}

static void scene_free(meso_scene *sc) {
  if (!sc) return;
  free_scene_arrays(sc);
  delete sc;
}

#define SCENE_TRY(expr)        \
  do {                         \
    meso_status st_ = (expr);  \
    if (st_ != MESO_OK) {      \
      scene_free(sc);          \
      return st_;              \
    }                          \
  } while (0)

meso_status meso_scene_create_dense_u8(meso_context *ctx, int32_t n, const uint8_t *voxels,
                                       meso_scene **out) {
  MESO_CHECK(ctx && voxels && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(n >= 1 && n <= 2048, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "dense edge %d not in 1..2048", n);
  MESO_HIP(hipSetDevice(ctx->device));
  auto *sc = new meso_scene();
  sc->ctx = ctx;
  sc->kind = MESO_SCENE_DENSE_U8;
  sc->n = n;
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->dense, (uint64_t)n * n * n, voxels, sc));
  *out = sc;
  return MESO_OK;
}

meso_status meso_scene_create_brickmap_u16(meso_context *ctx, int32_t n, const uint32_t *slots,
                                           int64_t n_bricks, const uint64_t *masks,
                                           const uint16_t *voxels, meso_scene **out) {
  MESO_CHECK(ctx && slots && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(n >= 8 && n % 8 == 0 && n <= 8192, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "brick map edge %d must be a multiple of 8 in 8..8192", n);
  MESO_CHECK(n_bricks >= 0 && n_bricks < (1ll << 29), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_bricks %lld out of range (< 2^29)", (long long)n_bricks);
  MESO_CHECK(n_bricks == 0 || (masks && voxels), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "masks/voxels NULL with %lld bricks", (long long)n_bricks);
  const int64_t nb = n / 8, ncell = nb * nb * nb;
  // every slot must name an existing brick (the kernels index the pools with
  // it), and no brick may be held by two cells (a brick's derived words are
  // kept per cell)
  std::vector<bool> held((size_t)n_bricks);
  for (int64_t i = 0; i < ncell; i++) {
    MESO_CHECK(slots[i] == MESO_EMPTY_SLOT || (int64_t)slots[i] < n_bricks,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "slot[%lld] = %u >= n_bricks %lld", (long long)i,
               slots[i], (long long)n_bricks);
    if (slots[i] == MESO_EMPTY_SLOT) continue;
    MESO_CHECK(!held[slots[i]], MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "brick %u is held by two cells (slot[%lld])", slots[i], (long long)i);
    held[slots[i]] = true;
  }
  MESO_HIP(hipSetDevice(ctx->device));
  auto *sc = new meso_scene();
  sc->ctx = ctx;
  sc->kind = MESO_SCENE_BRICKMAP_U16;
  sc->n = n;
  sc->n_bricks = n_bricks;
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->slots, (uint64_t)ncell * 4, slots, sc));
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->masks, (uint64_t)n_bricks * 64, masks, sc));
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->voxels, (uint64_t)n_bricks * 1024, voxels, sc));
  SCENE_TRY(build_coarse(sc, ctx->stream));
  *out = sc;
  return MESO_OK;
}

meso_status meso_scene_create_two_level_u16(meso_context *ctx, int32_t n, const uint32_t *top,
                                            int64_t n_nodes, const uint32_t *node_slots,
                                            int64_t n_bricks, const uint64_t *masks,
                                            const uint16_t *voxels, meso_scene **out) {
  MESO_CHECK(ctx && top && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(n >= 128 && n % 128 == 0 && n <= 8192, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "two-level edge %d must be a multiple of 128 in 128..8192", n);
  MESO_CHECK(n_nodes >= 0 && n_bricks >= 0 && n_bricks < (1ll << 29),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "negative or oversized counts (bricks < 2^29)");
  MESO_CHECK(n_nodes == 0 || node_slots, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "node_slots NULL");
  MESO_CHECK(n_bricks == 0 || (masks && voxels), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "masks/voxels NULL");
  const int64_t nt = n / 128, ntop = nt * nt * nt;
  for (int64_t i = 0; i < ntop; i++)
    MESO_CHECK(top[i] == MESO_EMPTY_SLOT || (int64_t)top[i] < n_nodes,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "top[%lld] = %u >= n_nodes", (long long)i, top[i]);
  std::vector<bool> held((size_t)n_bricks);
  for (int64_t i = 0; i < n_nodes * 4096; i++) {
    MESO_CHECK(node_slots[i] == MESO_EMPTY_SLOT || (int64_t)node_slots[i] < n_bricks,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "node slot %lld = %u >= n_bricks", (long long)i,
               node_slots[i]);
    if (node_slots[i] == MESO_EMPTY_SLOT) continue;
    MESO_CHECK(!held[node_slots[i]], MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "brick %u is held by two node cells (%lld)", node_slots[i], (long long)i);
    held[node_slots[i]] = true;
  }
  MESO_HIP(hipSetDevice(ctx->device));
  auto *sc = new meso_scene();
  sc->ctx = ctx;
  sc->kind = MESO_SCENE_TWO_LEVEL_U16;
  sc->n = n;
  sc->n_nodes = n_nodes;
  sc->n_bricks = n_bricks;
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->top, (uint64_t)ntop * 4, top, sc));
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->slots, (uint64_t)n_nodes * 4096 * 4, node_slots, sc));
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->masks, (uint64_t)n_bricks * 64, masks, sc));
  SCENE_TRY(scene_alloc(ctx, (void **)&sc->voxels, (uint64_t)n_bricks * 1024, voxels, sc));
  SCENE_TRY(build_coarse(sc, ctx->stream));
  *out = sc;
  return MESO_OK;
}

meso_status meso_scene_get_info(meso_scene *sc, meso_scene_info *out) {
  MESO_CHECK(sc && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  out->size = sizeof(*out);
  out->kind = sc->kind;
  out->n = sc->n;
  out->n_bricks = sc->n_bricks;
  out->n_nodes = sc->n_nodes;
  out->device_bytes = sc->bytes;
  out->solid_voxels = sc->solid;
  return MESO_OK;
}

meso_status meso_scene_download(meso_scene *sc, int32_t part, void *dst, uint64_t bytes) {
  MESO_CHECK(sc && dst, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  const void *src = nullptr;
  uint64_t have = 0;
  const uint64_t n = (uint64_t)sc->n;
  switch (part) {
    case MESO_PART_DENSE:
      src = sc->dense;
      have = sc->dense ? n * n * n : 0;
      break;
    case MESO_PART_SLOTS:
      src = sc->slots;
      have = sc->kind == MESO_SCENE_BRICKMAP_U16 ? (n / 8) * (n / 8) * (n / 8) * 4
                                                 : (uint64_t)sc->n_nodes * 4096 * 4;
      break;
    case MESO_PART_MASKS:
      src = sc->masks;
      have = (uint64_t)sc->n_bricks * 64;
      break;
    case MESO_PART_VOXELS:
      src = sc->voxels;
      have = (uint64_t)sc->n_bricks * 1024;
      break;
    case MESO_PART_TOP:
      src = sc->top;
      have = sc->top ? (n / 128) * (n / 128) * (n / 128) * 4 : 0;
      break;
    default: return fail(MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown scene part %d", part);
  }
  MESO_CHECK(src && bytes <= have, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "part %d holds %llu bytes, %llu requested", part, (unsigned long long)have,
             (unsigned long long)bytes);
  MESO_HIP(hipSetDevice(sc->ctx->device));
  MESO_HIP(hipStreamSynchronize(sc->ctx->stream));
  MESO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return MESO_OK;
}

meso_status meso_scene_destroy(meso_scene *sc) {
  if (!sc) return MESO_OK;
  hipStreamSynchronize(sc->ctx->stream);
  scene_free(sc);
  return MESO_OK;
}

// ---------------- raymarch
meso_status meso_raymarch(meso_context *ctx, meso_scene *scene, const meso_camera *cam,
                          uint32_t flags, const meso_shading *shading, const meso_ray_outputs *out,
                          meso_stream stream) {
  MESO_CHECK(ctx && scene && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->ctx == ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "scene belongs to another context");
  meso_status st = check_camera(cam);
  if (st != MESO_OK) return st;
  MESO_CHECK(out->size >= sizeof(meso_ray_outputs), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_ray_outputs.size too small");
  MESO_CHECK(out->hit_index, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "hit_index is NULL");
  MESO_CHECK(out->index_bytes == 4 || out->index_bytes == 8, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "index_bytes must be 4 or 8");
  const uint64_t nvox = (uint64_t)scene->n * scene->n * scene->n;
  MESO_CHECK(out->index_bytes == 8 || nvox <= 0x7FFFFFFFull, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "scene of %llu voxels needs 8-byte hit indices", (unsigned long long)nvox);
  MESO_CHECK((flags & ~(MESO_RAY_SHADOW | MESO_RAY_AO | MESO_RAY_COUNT)) == 0,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown ray flags 0x%x", flags);
  MESO_CHECK(!(flags & MESO_RAY_COUNT) || out->counters, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "MESO_RAY_COUNT needs a counters buffer");
  DevShading sh{};
  if (flags & (MESO_RAY_SHADOW | MESO_RAY_AO)) {
    MESO_CHECK(shading && shading->size >= sizeof(meso_shading), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "shadow/AO rays need a meso_shading");
    MESO_CHECK(shading->light[0] != 0.0f || shading->light[1] != 0.0f || shading->light[2] != 0.0f,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "light direction is zero");
    if (flags & MESO_RAY_AO)
      MESO_CHECK(shading->ao_samples >= 1 && shading->ao_samples <= 8 && shading->ao_radius > 0.0f,
                 MESO_ERR_ARGUMENT_OUT_OF_RANGE, "AO needs 1..8 samples and a positive radius");
    derive_shading(shading, &sh);
  } else if (shading && shading->size >= sizeof(meso_shading) &&
             (shading->light[0] != 0.0f || shading->light[1] != 0.0f || shading->light[2] != 0.0f)) {
    derive_shading(shading, &sh);
  } else {
    sh.light[1] = 1.0f;
  }
  DevCamera dc;
  derive_camera(cam, &dc);
  RayOut ro{out->hit_index, out->index_bytes, out->face, out->t, out->shade,
            (unsigned long long *)out->counters};
  MESO_HIP(hipSetDevice(ctx->device));
  const RayLaunch *shape = nullptr;
  RayLaunch tuned;
  {
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    char key[128];
    std::snprintf(key, sizeof(key), "ray:%d:%d:%dx%d:%u", scene->kind, scene->n, cam->width,
                  cam->height, flags & ~MESO_RAY_COUNT);
    auto it = ctx->tune.find(key);
    if (it != ctx->tune.end()) {
      tuned = RayLaunch{it->second.variant, it->second.block, it->second.tile_w, it->second.tile_h,
                        it->second.grid};
      shape = &tuned;
    }
  }
  return launch_raymarch(ctx, scene, dc, flags, sh, ro, stream_of(ctx, stream), shape);
}

}  // extern "C"
