// Internal declarations of the gfx950 voxel library (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>

#include "../../include/meso_hip.h"

namespace meso {

constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
meso_status fail(meso_status code, const char *fmt, ...);

#define MESO_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ::meso::fail(e_ == hipErrorOutOfMemory ? MESO_ERR_OUT_OF_MEMORY        \
                                                    : MESO_ERR_RUNTIME,             \
                          "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                          __FILE__, __LINE__);                                      \
  } while (0)

#define MESO_CHECK(cond, code, ...)                                                 \
  do {                                                                              \
    if (!(cond)) return ::meso::fail(code, __VA_ARGS__);                            \
  } while (0)

// ---------------------------------------------------------------- device structs
// Pinhole camera derived on the host (same operation order as the oracle).
struct DevCamera {
  float eye[3], fwd[3], right[3], up[3];
  float sx, sy, inv_w2, inv_h2;
  int32_t width, height;
};

struct DevShading {
  float light[3];
  uint32_t ao_seed;
  int32_t ao_samples;
  float ao_radius;
};

struct DevScene {
  int32_t n;
  const uint8_t *dense;
  const uint32_t *slots;
  const uint64_t *masks;
  // per brick cell (indexed like slots): one bit per 2^3 block of the brick
  // holding a solid voxel (0 for an empty cell)
  const uint64_t *blocks;
  const uint16_t *voxels;
  const uint32_t *top;
  const uint32_t *occ32;   // brick map: one bit per 32^3 block holding a brick
  const uint32_t *occ128;  // brick map: one bit per 128^3 block
  const uint64_t *groups;  // two-level: per node, one bit per 4^3-brick group
  // both brick kinds: per 32^3 block of the grid, bit (bx&3) + 4 (by&3) +
  // 16 (bz&3) set when brick (bx,by,bz) of the block is allocated
  const uint64_t *bricks32;
  // the box [lo, hi) that holds every solid voxel: rays are clipped to it and
  // walks end when they step out of it
  int32_t lo[3], hi[3];
  // the same bounds as floats (exact: below 2^24), converted on the host so
  // that they stay scalar kernel arguments (no scalar float convert on gfx950)
  float flo[3], fhi[3];
};

struct RayOut {
  void *index;
  int32_t index_bytes;
  uint8_t *face;
  float *t;
  float *shade;
  unsigned long long *counters;
};

// raymarch launch shape (autotuner variants)
struct RayLaunch {
  int32_t variant;  // 0..4: 8x8 wave packets in 16x16, 32x8, 8x8, 16x8, 32x16 tiles;
                    // 5, 6, 7: persistent waves pulling 8x8, 16x4, 32x2 packets from
                    // per-XCD band queues
  int32_t block;
  int32_t tile_w, tile_h;
  int32_t grid;     // variants 5-7: workgroups (0 = one per wave slot of the device)
};

struct MeshLaunch {
  int32_t block;        // threads per workgroup
  int32_t grid;         // persistent workgroups (0 = derive from occupancy)
};

}  // namespace meso

// ---------------------------------------------------------------- handles
struct meso_context {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t props{};
  std::mutex tune_mu;
  std::map<std::string, meso_tune_result> tune;
  // per stream: the band counters of the persistent raymarch (zero between
  // launches: the last wave of each launch resets them)
  std::mutex queue_mu;
  std::map<hipStream_t, unsigned int *> band_queues;
  // this device's copy of the mesher's static scratch (meso_mesh_chunks_noalloc)
  void *mesh_scratch = nullptr;
  // polls of the batched mesher's look-back wait before it gives up
  uint32_t lookback_spin_limit = 1u << 24;
};

struct meso_buffer {
  meso_context *ctx = nullptr;
  void *ptr = nullptr;
  uint64_t size = 0;
};

struct meso_scene {
  meso_context *ctx = nullptr;
  int32_t kind = 0;
  int32_t n = 0;
  uint8_t *dense = nullptr;
  uint32_t *slots = nullptr;
  uint64_t *masks = nullptr;
  // per brick cell (indexed like slots): bit cx + 4 cy + 16 cz set when the
  // 2^3 block (cx,cy,cz) of the cell's brick holds a solid voxel; written
  // wherever a brick's mask is written.  slot_cell[slot]: 1 + the cell (index
  // into slots) that holds the brick, 0 when no cell does.
  uint64_t *blocks = nullptr;
  uint32_t *slot_cell = nullptr;
  uint16_t *voxels = nullptr;
  uint32_t *top = nullptr;
  uint32_t *occ32 = nullptr;
  uint32_t *occ128 = nullptr;
  uint64_t *groups = nullptr;
  uint64_t *bricks32 = nullptr;  // see DevScene::bricks32
  int64_t n_bricks = 0;
  int64_t n_nodes = 0;
  uint64_t bytes = 0;
  int64_t solid = -1;
  // voxels [box_lo, box_hi) hold every allocated brick (brick scenes; set by
  // build_coarse); dense grids use the whole grid
  int32_t box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};
  meso::DevScene dev() const {
    const bool dense_grid = kind == MESO_SCENE_DENSE_U8;
    meso::DevScene d{n,      dense,  slots,  masks, blocks, voxels, top, occ32,
                     occ128, groups, bricks32, {},  {},     {},     {}};
    for (int a = 0; a < 3; a++) {
      d.lo[a] = dense_grid ? 0 : box_lo[a];
      d.hi[a] = dense_grid ? n : box_hi[a];
      d.flo[a] = (float)d.lo[a];
      d.fhi[a] = (float)d.hi[a];
    }
    return d;
  }
};

struct meso_event {
  meso_context *ctx = nullptr;
  hipEvent_t ev = nullptr;
};

struct meso_graph {
  meso_context *ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace meso {
// A 1D launch of nblocks workgroups spread over a 2D grid, so that
// gridDim.x * blockDim.x stays below 2^32 work-items; kernels recover the
// linear block with block_id() and skip ids >= nblocks.
inline dim3 grid_1d(uint64_t nblocks) {
  const uint64_t gx = nblocks < 65536u ? (nblocks ? nblocks : 1u) : 65536u;
  return dim3((uint32_t)gx, (uint32_t)((nblocks + gx - 1) / gx));
}
__device__ __forceinline__ uint64_t block_id() {
  return (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
}

// host helpers shared across translation units
void derive_camera(const meso_camera *c, DevCamera *out);
void derive_shading(const meso_shading *s, DevShading *out);
// validates a caller's camera (size, image, fov, finite vectors, up not parallel)
meso_status check_camera(const meso_camera *cam);
// what the quad rasteriser reads of a chunk pool
struct PoolView {
  const meso_scene *scene;
  const uint32_t *slot_chunk, *slot_count, *quads;
  int32_t capacity, per_slot, nc;
};
PoolView pool_view(const meso_chunk_pool *pool);
hipStream_t stream_of(meso_context *ctx, meso_stream s);
meso_status launch_zero(void *ptr, uint64_t bytes, hipStream_t stream);
// the device address of the mesher's static scratch on ctx's (current) device
meso_status init_mesh_scratch(meso_context *ctx);
// the persistent raymarch's band counters for `stream` (1 KiB, zeroed once)
meso_status band_queue(meso_context *ctx, hipStream_t stream, unsigned int **out);
// builds the empty-space levels (occ32/occ128 or node groups) of a brick scene
meso_status build_coarse(meso_scene *sc, hipStream_t stream);
void free_scene_arrays(meso_scene *sc);

// kernels launched by other translation units
meso_status launch_raymarch(meso_context *ctx, const meso_scene *scene, const DevCamera &cam,
                            uint32_t flags, const DevShading &sh, const RayOut &out,
                            hipStream_t stream, const RayLaunch *shape);
RayLaunch default_ray_launch(const meso_scene *scene);
meso_status launch_mesh_batched(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                                uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                uint64_t capacity, uint64_t *d_total, void *d_work,
                                hipStream_t stream, const MeshLaunch *shape);
MeshLaunch default_mesh_launch(meso_context *ctx);
// meshes the first *n_dev chunks of a scene chunk list into fixed slots
// (quads at slots[c] * per_slot); ticket: 2 counters, zero before the first
// launch, left at zero by every launch
meso_status launch_mesh_slotted(meso_context *ctx, const meso_scene *scene, const int32_t *xyz,
                                const uint32_t *slots, const uint32_t *n_dev, int32_t max_chunks,
                                uint32_t mode, uint32_t *slot_count, uint32_t *quads,
                                uint32_t per_slot, unsigned int *ticket, hipStream_t stream);

// host PCG32 / SplitMix64 (same published algorithms as the generators use)
uint64_t splitmix64(uint64_t *s);
struct Pcg32 {
  uint64_t state, inc;
  void srandom(uint64_t initstate, uint64_t initseq);
  void seed(uint64_t s);
  uint32_t next();
  float u01() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};
}  // namespace meso
