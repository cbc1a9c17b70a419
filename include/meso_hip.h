/*
 * meso_hip.h -- C-ABI of the MI355X (gfx950) voxel hot path.
 *
 * This is the drop-in boundary for MesoEngine's voxel path.  In the reference
 * engine, the Instance hooks (Runtimes/Instance/VoxelWindowsInstance.h:137-162,
 * overridden in Samples/SimpleVoxel.cpp:273-280,352-398) drive the GPU through
 * lvk::IContext / lvk::ICommandBuffer
 * (ThirdParty/lightweightvk/lvk/LVK.h:720-845).  Here the same hooks call this
 * library instead.  Each entry point names the reference interface it
 * replaces.
 *
 * Conventions (shared with the engine's data model):
 *  - voxel linear index x + y*n + z*n*n, x fastest
 *    (Runtimes/Helper/VoxelMathHelper.h:73-76);
 *  - face ids 0:+Z 1:-Z 2:+Y 3:-Y 4:+X 5:-X, the vertex-group order of
 *    Runtimes/Shape/Cube.h:25-54; 6 = ray origin inside a solid voxel;
 *    0xFF = miss;
 *  - brick = 8^3 voxels (Runtimes/Voxel/VoxelSceneConfig.h:22); top node =
 *    16^3 bricks (VoxelSceneConfig.h:24); brick occupancy = 8 x u64 (word z,
 *    bit x+8y), the 64-byte FVolume mask of Runtimes/Voxel/VoxelStructure.h:30-39;
 *  - quad = x:5 | y:5 | z:5 | (w-1):5 | (h-1):5 | face:3 | 0:4, packed
 *    little-endian like FGPUBlock (Runtimes/Voxel/Block/Block.h:21-27).
 *
 * No exceptions cross this boundary.  Every call returns a meso_status.  On
 * failure, meso_last_error_message() holds a per-thread description.  A
 * context is externally synchronised (one issuing thread), as lvk's is; the
 * server created by meso_serve_start is thread-safe for submit/poll/wait.
 */
#ifndef MESO_HIP_H
#define MESO_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MESO_ABI_VERSION 1u

/* lvk::Result::Code (LVK.h:250-278): Ok, ArgumentOutOfRange, RuntimeError;
 * extended with allocation failure, admission refusal and unsupported. */
typedef int32_t meso_status;
#define MESO_OK 0
#define MESO_ERR_ARGUMENT_OUT_OF_RANGE 1
#define MESO_ERR_RUNTIME 2
#define MESO_ERR_OUT_OF_MEMORY 3
#define MESO_ERR_BUSY 4
#define MESO_ERR_NOT_SUPPORTED 5

typedef struct meso_context meso_context;
typedef struct meso_buffer meso_buffer;
typedef struct meso_scene meso_scene;
typedef struct meso_event meso_event;
typedef struct meso_graph meso_graph;
typedef struct meso_server meso_server;
typedef void *meso_stream; /* a hipStream_t; NULL = the context's stream */

/* Per-thread message of the last failed call on this thread
 * (replaces lvk::Result::message, LVK.h:258). */
const char *meso_last_error_message(void);
uint32_t meso_abi_version(void);

/* ------------------------------------------------------------------ */
/* context (replaces lvk::createVulkanContextWithSwapchain,              */
/* Runtimes/Instance/VoxelWindowsInstance.cpp:102-127)                   */
/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t size;
  char name[64];
  int32_t compute_units;
  int32_t xcds;
  int32_t lds_bytes_per_cu;
  int32_t l2_bytes;
  uint64_t hbm_bytes;
  int32_t clock_khz;
  int32_t wavefront;
} meso_device_info;

meso_status meso_context_create(int32_t device, meso_context **out);
meso_status meso_context_destroy(meso_context *ctx);
meso_status meso_device_info_get(meso_context *ctx, meso_device_info *out);

/* ------------------------------------------------------------------ */
/* buffers (IContext::createBuffer / upload / gpuAddress / destroy,     */
/* LVK.h:803,822,824,817; BufferDesc LVK.h:675-681)                      */
/* ------------------------------------------------------------------ */
/* size 0 -> MESO_ERR_ARGUMENT_OUT_OF_RANGE (lvk rejects usage==0 the same
 * way, VulkanClasses.cpp:3148). init may be NULL. */
meso_status meso_buffer_create(meso_context *ctx, uint64_t size, const void *init,
                               meso_buffer **out);
meso_status meso_buffer_destroy(meso_buffer *buf);
/* out-of-range copies fail with ARGUMENT_OUT_OF_RANGE (VulkanClasses.cpp:3709-3711); a copy
 * runs on the context stream after the work queued there and is finished on return */
meso_status meso_buffer_upload(meso_buffer *buf, const void *src, uint64_t size, uint64_t offset);
meso_status meso_buffer_download(meso_buffer *buf, void *dst, uint64_t size, uint64_t offset);
/* device address of buf+offset; offset must be 8-byte aligned and inside the
 * buffer, else 0 (IContext::gpuAddress, VulkanClasses.cpp:3726-3734). */
uint64_t meso_buffer_device_address(meso_buffer *buf, uint64_t offset);
uint64_t meso_buffer_size(meso_buffer *buf);
/* zero-fill with a kernel (graph-capture safe) */
meso_status meso_buffer_fill_zero(meso_buffer *buf, meso_stream stream);

/* ------------------------------------------------------------------ */
/* streams and events (acquireCommandBuffer / submit / wait,             */
/* LVK.h:798-801; SubmitHandle LVK.h:774-789)                            */
/* ------------------------------------------------------------------ */
meso_status meso_stream_create(meso_context *ctx, meso_stream *out);
meso_status meso_stream_destroy(meso_context *ctx, meso_stream stream);
meso_status meso_stream_sync(meso_context *ctx, meso_stream stream);
meso_status meso_event_create(meso_context *ctx, meso_event **out);
meso_status meso_event_destroy(meso_event *ev);
meso_status meso_event_record(meso_event *ev, meso_stream stream);
/* Work queued on `stream` after this call starts after `ev` completes: the
 * dependency a submit can name (ICommandBuffer::cmdDispatchThreadGroups'
 * Dependencies, LVK.h:730-732), between two of this library's streams. */
meso_status meso_stream_wait_event(meso_context *ctx, meso_stream stream, meso_event *ev);
meso_status meso_event_sync(meso_event *ev);
meso_status meso_event_elapsed_ms(meso_event *start, meso_event *end, float *ms);
/* copies between device pointers and host memory, on a stream */
meso_status meso_memcpy_d2h(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                            meso_stream stream);
meso_status meso_memcpy_h2d(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                            meso_stream stream);
/* Host-to-device copy queued on the stream without waiting: src must stay
 * valid and unchanged until the stream reaches the copy (pinned memory from
 * meso_host_alloc copies at full PCIe speed). */
meso_status meso_memcpy_h2d_async(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                                  meso_stream stream);
/* Pinned (page-locked) host memory for full-speed copies between host and
 * device (the host-visible storage behind IContext::getMappedPtr,
 * LVK.h:823).  Free with meso_host_free. */
meso_status meso_host_alloc(meso_context *ctx, uint64_t bytes, void **out);
meso_status meso_host_free(meso_context *ctx, void *ptr);
/* device-to-device copy, asynchronous on the stream.  lvk has no buffer to
 * buffer copy (only IContext::upload, LVK.h:822); this one moves results
 * between device buffers and is the known-size copy that calibrates the
 * HBM traffic counters. */
meso_status meso_memcpy_d2d(meso_context *ctx, void *dst, const void *src, uint64_t bytes,
                            meso_stream stream);

/* ------------------------------------------------------------------ */
/* scenes: the engine's voxel data model (FVoxelSceneConfig,             */
/* Runtimes/Voxel/VoxelSceneConfig.h:20-49; FVolume VoxelStructure.h:30-39) */
/* resident in HBM                                                       */
/* ------------------------------------------------------------------ */
#define MESO_SCENE_DENSE_U8 1
#define MESO_SCENE_BRICKMAP_U16 2
#define MESO_SCENE_TWO_LEVEL_U16 3
#define MESO_EMPTY_SLOT 0xFFFFFFFFu

typedef struct {
  uint32_t size;
  int32_t kind;
  int32_t n;          /* voxels per axis */
  int64_t n_bricks;   /* allocated bricks (brick map, two-level) */
  int64_t n_nodes;    /* allocated top nodes (two-level) */
  uint64_t device_bytes;
  int64_t solid_voxels; /* counted at generation; -1 if unknown */
} meso_scene_info;

/* host arrays are copied into library-owned device memory.  Brick scenes:
 * every non-empty cell names a brick below n_bricks, and no brick is named
 * by two cells (MESO_ERR_ARGUMENT_OUT_OF_RANGE otherwise). */
meso_status meso_scene_create_dense_u8(meso_context *ctx, int32_t n, const uint8_t *voxels,
                                       meso_scene **out);
meso_status meso_scene_create_brickmap_u16(meso_context *ctx, int32_t n, const uint32_t *slots,
                                           int64_t n_bricks, const uint64_t *masks,
                                           const uint16_t *voxels, meso_scene **out);
meso_status meso_scene_create_two_level_u16(meso_context *ctx, int32_t n, const uint32_t *top,
                                            int64_t n_nodes, const uint32_t *node_slots,
                                            int64_t n_bricks, const uint64_t *masks,
                                            const uint16_t *voxels, meso_scene **out);

/* GPU voxelization (replaces the CPU generator callback
 * FChunkManage::Initialize(..., Generator) of ChunkManager.h:61-66,90-102 and
 * FGeneratorHelper, Runtimes/Helper/GeneratorHelper.h:90-150). */
#define MESO_GEN_SPHERES 0       /* n_spheres seeded spheres (dense only) */
#define MESO_GEN_HEIGHTFIELD 1   /* 4-octave value-noise fBm heightfield */
#define MESO_GEN_TERRAIN_CAVES 2 /* heightfield terrain with fBm tunnels */
#define MESO_GEN_CAVES 3         /* thin sheets |fBm - 1/2| < tau */
typedef struct {
  uint32_t size;
  int32_t generator;  /* MESO_GEN_* */
  int32_t layout;     /* MESO_SCENE_* */
  int32_t n;          /* edge in voxels; multiple of 8 (brick map) or 128 (two-level) */
  uint64_t seed;
  int32_t n_spheres;
} meso_generate_desc;
meso_status meso_scene_generate(meso_context *ctx, const meso_generate_desc *desc,
                                meso_scene **out);
meso_status meso_scene_get_info(meso_scene *scene, meso_scene_info *out);
#define MESO_PART_DENSE 0
#define MESO_PART_SLOTS 1  /* brick map slots, or node brick-slot tables */
#define MESO_PART_MASKS 2
#define MESO_PART_VOXELS 3
#define MESO_PART_TOP 4
meso_status meso_scene_download(meso_scene *scene, int32_t part, void *dst, uint64_t bytes);
/* Overwrite the voxels of existing bricks (slot ids) from device memory,
 * rebuilding their masks.  Bricks keep their slot; a brick that becomes empty
 * keeps a zero mask.  d_voxels: n_edit * 512 u16. */
meso_status meso_scene_update_bricks(meso_scene *scene, const uint32_t *d_slots,
                                     const uint16_t *d_voxels, int32_t n_edit, meso_stream stream);
/* The same without waiting, for a pipelined upload of host edits (the role
 * of the double-buffered modify queues and staging upload of
 * FChunkPool, Runtimes/Voxel/Chunk/ChunkPool.h:137-168,641-705): slots >=
 * the scene's brick count are skipped and counted into *d_skipped (device
 * u32, may be NULL).  The slots of one call must be distinct. */
meso_status meso_scene_update_bricks_async(meso_scene *scene, const uint32_t *d_slots,
                                           const uint16_t *d_voxels, int32_t n_edit,
                                           uint32_t *d_skipped, meso_stream stream);
/* Seeded edits of 32^3 chunks (chunk coordinates d_chunk_xyz[3k], device):
 * inside a sphere drawn per chunk from lowbias32(seed ^ lowbias32(chunk
 * index)) every voxel of an allocated brick toggles (solid -> 0, empty -> 7)
 * and the brick masks follow.  Bricks are neither allocated nor freed.  The
 * chunk grid has ceil(n/32) chunks per axis (chunk index x + nc*(y + nc*z));
 * chunks outside it are skipped.  The chunks of one call must be distinct
 * (the edits of a repeated chunk race); a second call applies a repeat.
 * Replaces the per-chunk regeneration path FChunkManage::MultiThreadGenerator
 * (Runtimes/Voxel/Chunk/ChunkManager.h:160-210) for edited chunks. */
meso_status meso_scene_edit_chunks(meso_scene *scene, const int32_t *d_chunk_xyz, int32_t n_chunks,
                                   uint32_t seed, meso_stream stream);
meso_status meso_scene_destroy(meso_scene *scene);

/* ------------------------------------------------------------------ */
/* raymarch (replaces the per-frame draw in Render(),                    */
/* Samples/SimpleVoxel.cpp:352-398, and the camera UBO of                */
/* Runtimes/Shader/GPUStructures.h:36-56)                                */
/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t size;
  float eye[3];
  float target[3];
  float up[3];
  float vfov_deg;     /* vertical field of view, as glm::perspective */
  int32_t width, height;
} meso_camera;

typedef struct {
  uint32_t size;
  float light[3];     /* direction toward the light; normalised by the library */
  uint32_t ao_seed;
  int32_t ao_samples; /* 0..8 */
  float ao_radius;    /* voxels */
} meso_shading;

#define MESO_RAY_PRIMARY 0u
#define MESO_RAY_SHADOW 1u
#define MESO_RAY_AO 2u
#define MESO_RAY_COUNT 0x100u /* also accumulate step counters (slower) */

/* Step counters (accumulated with atomics when MESO_RAY_COUNT is set).
 * N_VOX: voxel reads of a dense grid, or voxel-level mask-word reads (8 B)
 * of a brick scene; N_BRICKMASK: non-empty bricks entered (their 8-byte
 * 2^3-block word read). */
#define MESO_CNT_N_VOX 0
#define MESO_CNT_N_TOP 1
#define MESO_CNT_N_NODE 2
#define MESO_CNT_N_BRICKSLOT 3
#define MESO_CNT_N_BRICKMASK 4
#define MESO_CNT_HITS 5
#define MESO_CNT_RAYS 6
#define MESO_CNT_SHADOW_RAYS 7
#define MESO_CNT_AO_RAYS 8
#define MESO_CNT_N_COARSE 9  /* 32^3 / 128^3 occupancy words read */
#define MESO_CNT_N_ITER 10   /* traversal steps at any level */
#define MESO_CNT_COUNT 11

typedef struct {
  uint32_t size;
  void *hit_index;      /* device: int32 (index_bytes 4) or int64 (8) per pixel; -1 = miss */
  int32_t index_bytes;
  uint8_t *face;        /* device, per pixel; may be NULL */
  float *t;             /* device, per pixel; +inf on miss; may be NULL */
  float *shade;         /* device, per pixel; may be NULL */
  uint64_t *counters;   /* device, MESO_CNT_COUNT u64, used with MESO_RAY_COUNT */
} meso_ray_outputs;

meso_status meso_raymarch(meso_context *ctx, meso_scene *scene, const meso_camera *cam,
                          uint32_t flags, const meso_shading *shading,
                          const meso_ray_outputs *out, meso_stream stream);

/* ------------------------------------------------------------------ */
/* meshing (replaces the per-cell imposter draw of                       */
/* Samples/SimpleVoxel.cpp:31-195,387 and the CPU interior cull of       */
/* Runtimes/Voxel/Chunk/ChunkPool.h:385-390)                             */
/* ------------------------------------------------------------------ */
#define MESO_MESH_CULLED 0u
#define MESO_MESH_GREEDY 1u
#define MESO_CHUNKS_PACKED 0      /* chunk c at voxels + c * 32^3 elements */
#define MESO_CHUNKS_DENSE_GRID 1  /* chunks tiled over a dense grid of edge grid_n */

typedef struct {
  uint32_t size;
  const void *voxels;   /* device */
  int32_t elem_bytes;   /* 1 (u8) or 2 (u16); 0 = empty */
  int32_t layout;       /* MESO_CHUNKS_* */
  int32_t grid_n;       /* dense grid edge (multiple of 32) for DENSE_GRID */
  int32_t n_chunks;
} meso_chunk_source;

/* Plain path: one launch meshes one 32^3 chunk.  Writes the quad count to
 * d_count and min(count, capacity) quads to d_quads in canonical order
 * (face, slice, v, u). */
meso_status meso_mesh_chunk(meso_context *ctx, const void *d_voxels, int32_t elem_bytes,
                            uint32_t mode, uint32_t *d_count, uint32_t *d_quads,
                            uint32_t capacity, meso_stream stream);

/* Batched path: one persistent launch over all chunks, with a device work
 * queue and an ordered decoupled look-back for the quad prefix.
 * d_counts[n_chunks], d_prefix[n_chunks] (exclusive), d_quads[capacity],
 * d_total[1] (u64 total quads; quads beyond capacity are not written and
 * d_total still reports the full count).  d_work: workspace of
 * meso_mesh_workspace_bytes(n_chunks) bytes. */
uint64_t meso_mesh_workspace_bytes(int32_t n_chunks);
meso_status meso_mesh_chunks(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                             uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                             uint64_t capacity, uint64_t *d_total, void *d_work,
                             meso_stream stream);

/* Whether the batched meshing that last used d_work (NULL: the static
 * scratch of meso_mesh_chunks_noalloc) completed its look-back: waits for
 * `stream`, then RUNTIME_ERROR if a chunk's bounded wait for the chunks
 * ticketed before it gave up (its prefix, quads and the total are then
 * wrong).  The wait is bounded so that every wave of the persistent launch
 * exits whatever happens to the others. */
meso_status meso_mesh_check(meso_context *ctx, const void *d_work, meso_stream stream);
/* Polls of that wait before it gives up (default 2^24; 0 gives up at the
 * first look, which tests use to force the failure from the host). */
meso_status meso_mesh_set_lookback_spin_limit(meso_context *ctx, uint32_t limit);

/* meso_mesh_chunks without a workspace argument and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The workspace is a static device array of the
 * library, one per GPU, sized for MESO_MESH_NOALLOC_MAX_CHUNKS chunks;
 * larger batches return ARGUMENT_OUT_OF_RANGE.  Every call on one GPU shares
 * that array (each launch clears it first): calls that overlap on one GPU
 * (different streams, or threads) race; calls ordered on one stream do not.
 * Outputs and results are those of meso_mesh_chunks. */
#define MESO_MESH_NOALLOC_MAX_CHUNKS 16384
meso_status meso_mesh_chunks_noalloc(meso_context *ctx, const meso_chunk_source *src,
                                     uint32_t mode, uint32_t *d_counts, uint32_t *d_prefix,
                                     uint32_t *d_quads, uint64_t capacity, uint64_t *d_total,
                                     meso_stream stream);

/* Batched meshing of 32^3 chunks read straight from a resident brick map or
 * two-level scene (chunk coordinates d_chunk_xyz[3k] on the device, on the
 * grid of ceil(n/32) chunks per axis; voxels past the scene edge and chunks
 * outside the grid mesh as empty).  Outputs and workspace as in
 * meso_mesh_chunks; quads are chunk-local. */
meso_status meso_mesh_scene_chunks(meso_context *ctx, meso_scene *scene, const int32_t *d_chunk_xyz,
                                   int32_t n_chunks, uint32_t mode, uint32_t *d_counts,
                                   uint32_t *d_prefix, uint32_t *d_quads, uint64_t capacity,
                                   uint64_t *d_total, void *d_work, meso_stream stream);

/* Captures meso_mesh_chunks (workspace reset + persistent launch) into a
 * hipGraph that replays with the same arguments. */
meso_status meso_mesh_graph_create(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                                   uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                   uint64_t capacity, uint64_t *d_total, void *d_work,
                                   meso_graph **out);
/* Captures n plain meso_mesh_chunk launches (chunk c -> d_counts[c],
 * d_quads + c*per_chunk_cap) into a hipGraph. */
meso_status meso_mesh_plain_graph_create(meso_context *ctx, const meso_chunk_source *src,
                                         uint32_t mode, uint32_t *d_counts, uint32_t *d_quads,
                                         uint32_t per_chunk_cap, meso_graph **out);
meso_status meso_graph_launch(meso_graph *g, meso_stream stream);
meso_status meso_graph_destroy(meso_graph *g);

/* Seeded chunk batch (Beta(1/2,1/2) fill, value noise, mixed spheres):
 * d_out[n_chunks * 32^3] u16. */
meso_status meso_generate_chunks(meso_context *ctx, uint64_t seed, int32_t n_chunks,
                                 uint16_t *d_out, meso_stream stream);

/* Reference voxelizer on the GPU: FGeneratorHelper::GenerateSphere
 * (GeneratorHelper.h:120-150) for n_chunks chunks of edge res at integer
 * chunk coordinates d_chunk_xyz[3*c], writing occupancy (res^3 u8 per chunk)
 * and the erosion mip chain of FChunk::CalculateOccupancyErodeMipmaps
 * (Chunk.h:73-94): depth * res^3 u8 per chunk, mip 0 = occupancy. */
meso_status meso_generate_sphere_chunks(meso_context *ctx, const int32_t *d_chunk_xyz,
                                        int32_t n_chunks, int32_t res, float block_size,
                                        int32_t depth, uint8_t *d_mips, meso_stream stream);

/* ------------------------------------------------------------------ */
/* autotuner: times tile / packet shapes on this device and keeps the     */
/* fastest per workload key (in memory and in an optional text cache)    */
/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t size;
  int32_t variant;     /* chosen variant id */
  int32_t block;       /* threads per workgroup */
  int32_t grid;        /* workgroups (persistent kernels) */
  int32_t tile_w, tile_h;
  float best_ms;
  float default_ms;
  int32_t tried;
} meso_tune_result;
meso_status meso_autotune_mesh(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                               uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                               uint64_t capacity, uint64_t *d_total, void *d_work,
                               meso_tune_result *out);
meso_status meso_autotune_raymarch(meso_context *ctx, meso_scene *scene, const meso_camera *cam,
                                   uint32_t flags, const meso_shading *shading,
                                   const meso_ray_outputs *out, meso_tune_result *result);
/* Pins a raymarch launch shape for the workload key (scene kind and edge,
 * frame size, flags without MESO_RAY_COUNT), replacing the tuner's choice:
 * shape->variant 0..4 (tiles 16x16, 32x8, 8x8, 16x8, 32x16), 5..7
 * (persistent waves on 8x8, 16x4, 32x2 packets) or 8..9 (persistent waves
 * over pairs of 8x8 or 16x4 packets, their secondary rays shared over the
 * wave); for 5..9 shape->grid is the number of workgroups of 64 lanes, 0 =
 * one per wave slot, at most 32 per CU; tile_w / tile_h must match the
 * variant's tile or packet.  A caller that shares the
 * device with latency-bound work pins a smaller persistent grid so that
 * the other work keeps wave slots.  No lvk counterpart. */
meso_status meso_autotune_pin_raymarch(meso_context *ctx, meso_scene *scene, const meso_camera *cam,
                                       uint32_t flags, const meso_tune_result *shape);
/* Load / save the tuning cache: text lines "key variant block grid tile_w
 * tile_h best_ms default_ms".  A missing file is not an error. */
meso_status meso_autotune_load(meso_context *ctx, const char *path);
meso_status meso_autotune_save(meso_context *ctx, const char *path);

/* ------------------------------------------------------------------ */
/* chunk pool: the chunks around a camera, ranked by importance and kept */
/* meshed in a fixed number of slots                                     */
/* ------------------------------------------------------------------ */
/* Replaces the reference's chunk streaming: FChunkManage::UpdateChunks /
 * UpdateLoadingQueue (Runtimes/Voxel/Chunk/ChunkManager.h:160-330), the
 * fixed-capacity FChunkPool with EChunkOverrideMode::FindMin replacement
 * (ChunkPool.h:446-560) and FChunkManageHelper's importance and desired set
 * (ChunkManagerHelper.h:21-148).  Chunks are the 32^3 chunks of a resident
 * brick map or two-level scene (index x + nc*(y + nc*z), nc = ceil(n/32));
 * a slot holds the packed quads of one chunk, meshed from the scene. */
typedef struct meso_chunk_pool meso_chunk_pool;
typedef struct meso_pool_desc {
  uint32_t size;           /* sizeof(meso_pool_desc) */
  int32_t capacity;        /* slots, 1..2^22 (FVoxelSceneConfig::MaxChunkCount: 16384) */
  int32_t quads_per_slot;  /* quad capacity of a slot, 1..98304; longer meshes are cut */
  int32_t forward_radius;  /* ViewForwardLoadChunkSize, 1..64 chunks (24) */
  int32_t backward_radius; /* ViewBackwardLoadChunkSize, 0..forward_radius (6) */
  float view_angle;        /* ViewChunkAngle, degrees in (0, 360] (120) */
  int32_t max_loads;       /* chunks loaded per update, 1..capacity
                              (MaxSyncedLoadChunkCount + MaxUnsyncedLoadChunkCount: 256) */
  uint32_t mesh_mode;      /* MESO_MESH_CULLED or MESO_MESH_GREEDY */
  int32_t bake_directions; /* BakeVisibilityViewNum: 0 = the exact forward vector; 2..65536 =
                              the nearest of that many Fibonacci-sphere directions
                              (GetFibonacciSphere, VoxelMathHelper.h:48-71; the reference's
                              default is 256), used for the whole update */
} meso_pool_desc;
typedef struct meso_pool_stats {
  uint32_t size;          /* sizeof(meso_pool_stats) */
  uint32_t stamp;         /* updates so far (the frame stamp; wraps at 2^32) */
  uint32_t n_candidates;  /* last update: desired non-empty chunks not resident */
  uint32_t n_loads;       /* last update: chunks loaded and meshed */
  uint32_t n_evicted;     /* last update: resident chunks they replaced */
  uint32_t n_resident;    /* slots holding a chunk */
  uint32_t n_truncated;   /* resident chunks with more quads than a slot holds */
  uint32_t pad;
  uint64_t n_quads;       /* quads of the resident chunks (full counts) */
  uint64_t device_bytes;  /* device memory of the pool */
} meso_pool_stats;
typedef struct meso_pool_arrays {
  uint32_t size;                /* sizeof(meso_pool_arrays) */
  int32_t chunks_per_axis;      /* nc */
  const uint32_t *lookup;       /* nc^3: slot of each chunk, or 0xFFFFFFFF */
  const uint32_t *slot_chunk;   /* capacity: chunk index held, or 0xFFFFFFFF (free) */
  const uint32_t *slot_stamp;   /* capacity: stamp of the update that loaded it */
  const uint32_t *slot_count;   /* capacity: the chunk's quad count (may exceed a slot) */
  const uint32_t *quads;        /* capacity * quads_per_slot packed quads, chunk-local */
  const int32_t *load_xyz;      /* the last update's loads: max_loads x (x, y, z) */
  const uint32_t *load_slot;    /* and their slots; n_loads of them are valid */
} meso_pool_arrays;
/* Creates an empty pool over a resident scene (the scene must outlive it). */
meso_status meso_chunk_pool_create(meso_context *ctx, meso_scene *scene, const meso_pool_desc *desc,
                                   meso_chunk_pool **out);
meso_status meso_chunk_pool_destroy(meso_chunk_pool *pool);
/* One update for a camera in chunk camera_chunk looking along forward (any
 * length; normalised here).  Candidates are the desired chunks (inside the
 * forward sphere; the 3^3 core, or nearer than a radius that blends from the
 * forward radius inside the view cone to the backward radius outside) that
 * are inside the scene, hold a non-empty brick and are not resident, ranked
 * by CalculateChunkImportance (descending), then chunk index.  Slots rank
 * free first, then by the importance of their chunk under this camera
 * (ascending).  Candidate i takes slot i while i < max_loads and the slot is
 * free or its chunk strictly less important; the candidates left wait for
 * later updates (the reference's re-queue).  The loaded chunks are meshed
 * into their slots.  Runs on `stream` without a host wait.  Calls on one
 * pool must be ordered (one stream, or synchronised). */
meso_status meso_chunk_pool_update(meso_chunk_pool *pool, const int32_t camera_chunk[3],
                                   const float forward[3], meso_stream stream);
/* Remeshes the resident chunks of a dirty list (device chunk coordinates,
 * e.g. after meso_scene_edit_chunks) into their slots; others are skipped. */
meso_status meso_chunk_pool_remesh(meso_chunk_pool *pool, const int32_t *d_chunk_xyz,
                                   int32_t n_chunks, meso_stream stream);
/* Counts over the pool; waits for `stream`. */
meso_status meso_chunk_pool_stats(meso_chunk_pool *pool, meso_pool_stats *out, meso_stream stream);
/* Device arrays of the pool, for consumers (valid until destroy). */
meso_status meso_chunk_pool_arrays(meso_chunk_pool *pool, meso_pool_arrays *out);

/* ------------------------------------------------------------------ */
/* quad visibility: the raster consumer of the meshes                    */
/* ------------------------------------------------------------------ */
/* Replaces the per-cell imposter draw of Samples/SimpleVoxel.cpp:352-398
 * (a 3-face fan per solid cell, FTriplePlanarCube, Shape/TriplePlanarCube.h:31-46)
 * with the packed quads of a chunk pool.  For every pixel, the nearest
 * front-facing quad along the pixel's primary ray (the ray meso_raymarch
 * casts): the ray must travel against the face normal, t = (plane - o) *
 * (1 / d) > 0, and the hit point must lie in the quad's [u0, u0 + w) x
 * [v0, v0 + h).  Ties go to the smaller slot * quads_per_slot + k.  The
 * outputs are meso_raymarch's with flags 0: the voxel behind the face
 * (-1 on a miss), the face, t (+inf on a miss) and the shade
 * 0.25 + 0.75 * lambert (0 on a miss; shading may be NULL for light +y);
 * counters are not written.  Where the pool holds every non-empty chunk the
 * view reaches, the result equals the raymarcher's primary hit (t to the
 * bit) except on rays through voxel edges. */
typedef struct meso_raster meso_raster;
typedef struct meso_raster_stats {
  uint32_t size;           /* sizeof(meso_raster_stats) */
  uint32_t small_quads;    /* last frame: quads of at most 64 pixels, tested by one lane */
  uint32_t large_quads;    /* last frame: larger quads, tested by a wave (up to 4096 pixels)
                              or cut into segments of 16384 pixels */
  uint32_t segments;       /* last frame: the segments */
  uint32_t overflow_quads; /* last frame: quads tested by one lane (segment list full) */
} meso_raster_stats;
/* Scratch for frames up to max_width x max_height (8 B per pixel plus a
 * 64 MiB segment list). */
meso_status meso_raster_create(meso_context *ctx, int32_t max_width, int32_t max_height,
                               meso_raster **out);
meso_status meso_raster_destroy(meso_raster *raster);
/* One frame of the pool's resident quads; runs on `stream` without a host
 * wait.  Calls on one raster object must be ordered. */
meso_status meso_raster_pool(meso_raster *raster, meso_chunk_pool *pool, const meso_camera *cam,
                             const meso_shading *shading, const meso_ray_outputs *out,
                             meso_stream stream);
/* Counts of the last frame; waits for `stream`. */
meso_status meso_raster_stats_get(meso_raster *raster, meso_raster_stats *out, meso_stream stream);

/* ------------------------------------------------------------------ */
/* serving: continuous batching of raymarch requests over resident       */
/* scenes, admitted by an HBM budget (the GPU counterpart of the         */
/* streaming loop FChunkManage::UpdateLoadingQueue, ChunkManager.h:211-400, */
/* with its back-pressure re-queue at :355-372)                          */
/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t size;
  uint64_t hbm_budget_bytes;  /* bytes of request outputs admitted at once */
  int32_t max_pending;        /* queued, not yet admitted; submit -> BUSY beyond */
  int32_t max_inflight;       /* admitted requests on the device at once: entries of the
                                 device ring the persistent kernel draws from (1..4096) */
  int32_t workgroups_per_cu;  /* persistent workgroups of 4 waves per CU, 0..8 (0: all that fit) */
  int32_t tile;               /* pixel edge of a work item: a wave's square packet (0 or 8) */
  int32_t idle_exit_us;       /* the persistent kernel returns after this long with nothing
                                 to claim, and is launched again for new work (0: 500) */
} meso_serve_config;

typedef struct {
  uint32_t size;
  uint64_t submitted, admitted, completed, refused_busy;
  uint64_t launches;          /* launches of the persistent kernel */
  uint64_t primary_rays;      /* pixels of the admitted requests */
  uint64_t bytes_in_use, peak_bytes_in_use;
  double busy_ms;             /* summed lifetimes of those launches */
  double queue_us_sum, queue_us_max;  /* submit -> admission */
  uint64_t queued_over_1ms;   /* requests that waited more than 1 ms for admission */
  uint64_t admission_blocked_passes;  /* scheduler passes that left requests waiting */
} meso_serve_stats;

typedef struct {
  uint32_t size;
  int32_t index_bytes;        /* 4 or 8 */
  void *hit_index;            /* host */
  uint8_t *face;
  float *t;
  float *shade;
} meso_host_outputs;

meso_status meso_serve_start(meso_context *ctx, const meso_serve_config *cfg, meso_server **out);
/* Queue a request; returns a ticket.  ARGUMENT_OUT_OF_RANGE if the request
 * alone exceeds the budget; BUSY if the pending queue is full. */
meso_status meso_serve_submit(meso_server *srv, meso_scene *scene, const meso_camera *cam,
                              uint32_t flags, const meso_shading *shading, uint64_t *ticket);
/* done: 0 pending/running, 1 complete, -1 failed */
meso_status meso_serve_poll(meso_server *srv, uint64_t ticket, int32_t *done);
/* Block until complete or timeout (microseconds; <0 waits forever).
 * Returns BUSY on timeout.  latency_us (may be NULL): submit -> result ready. */
meso_status meso_serve_wait(meso_server *srv, uint64_t ticket, int64_t timeout_us,
                            double *latency_us);
/* Copy a completed result to host memory and release its HBM budget. */
meso_status meso_serve_take(meso_server *srv, uint64_t ticket, const meso_host_outputs *dst);
meso_status meso_serve_stats_get(meso_server *srv, meso_serve_stats *out);
meso_status meso_serve_stop(meso_server *srv);

#ifdef __cplusplus
}
#endif
#endif
