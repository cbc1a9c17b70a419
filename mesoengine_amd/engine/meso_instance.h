// Headless voxel instance over libmeso_hip: the engine side of the C-ABI.
//
// The hooks keep the names and call order of VoxelWindowsInstance
// (Runtimes/Instance/VoxelWindowsInstance.h:137-162, loop at
// VoxelWindowsInstance.cpp:368-393: UpdateCamera, UpdatePhysics,
// RenderStart, Render, RenderEnd, UpdateFrameIndex) and of its SimpleVoxel
// overrides (Samples/SimpleVoxel.cpp:273-280,352-398).  Where those issue lvk
// buffer uploads and command buffers, these issue meso_* calls:
//   WhenCameraUpdate -> derives the pinhole camera of the next frame
//   UpdatePhysics    -> applies queued chunk edits and remeshes the dirty
//                       chunks straight from the resident scene
//   Render           -> one meso_raymarch of the frame (primary, shadow, AO)
// No window, swapchain or lvk object exists; results stay in device buffers.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "meso_hip.h"

namespace meso_engine {

struct FVec3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct FIVec3 {
  int32_t x = 0, y = 0, z = 0;
  bool operator==(const FIVec3 &o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const FIVec3 &o) const { return !(*this == o); }
};

// How a full pool picks a slot to override (Runtimes/Voxel/VoxelSceneConfig.h:12-17).
// The chunk pool on the GPU implements FindMin: a candidate replaces the
// least important resident chunk when it is more important.
enum class EChunkOverrideMode : uint8_t {
  FindLess = 1 << 0,
  FindMin = 1 << 1,
  OverrideMin = 1 << 2,
};

// The reference's scene configuration (Runtimes/Voxel/VoxelSceneConfig.h:20-49):
// every field, in the same order and with the same default, so the
// reference's initialisers (designated or member by member) compile
// unchanged.  GetChunkSize() keeps its meaning: ChunkResolution blocks of
// BlockSize.  Fields after the reference's are this engine's own:
// MeshChunkResolution is the edge, in voxels, of the chunks that edits,
// remeshing, streaming and the camera's chunk tracking work on (32: a mesh
// chunk's rows are 32-bit masks), and ChunkQuadCapacity the quads one pool
// slot holds.  The pool reads MaxChunkCount, the two load radii,
// ViewChunkAngle, the two load counts and BakeVisibilityViewNum; the other
// reference fields size CPU pools, thread batches and erosion depths that
// the GPU path does not have, and are kept for source compatibility.
struct FVoxelSceneConfig {
  unsigned char BlockResolution = 8;
  float BlockSize = 1.0f;
  unsigned char ChunkResolution = 16;
  uint32_t MaxBlockCount = 65536 * 16;
  uint32_t MaxVolumeCount = 65536 * 16;
  uint32_t MaxChunkCount = 8192 * 2;
  uint32_t MaxEmptyChunkCount = 8192 * 4;

  uint32_t MaxChunkCheckTimes = 128;
  uint32_t MaxEmptyChunkCheckTimes = 128;
  uint32_t MaxBlockCheckTimes = 16;

  uint32_t BakeVisibilityViewNum = 256;
  uint32_t ViewForwardLoadChunkSize = 24;
  uint32_t ViewBackwardLoadChunkSize = 6;
  uint32_t MaxSyncedLoadChunkCount = 0;
  uint32_t MaxUnsyncedLoadChunkCount = 256;
  uint32_t ChunkTaskPerCore = 8;
  EChunkOverrideMode ChunkOverrideMode = EChunkOverrideMode::FindMin;
  float ViewChunkAngle = 120.0f;

  uint32_t ChunkOccupancyDepth = 4;
  uint32_t ChunkInnerVoxelCullDepthThreshold = 1;
  float GetChunkSize() const { return (float)ChunkResolution * BlockSize; }

  // this engine's fields
  uint32_t MeshChunkResolution = 32;
  uint32_t ChunkQuadCapacity = 8192;
  float GetMeshChunkSize() const { return (float)MeshChunkResolution * BlockSize; }
};

// Camera with FVoxelCamera's floating origin (Runtimes/Instance/VoxelCamera.cpp:42-59):
// Position stays inside one chunk and CameraChunkLocation counts whole
// chunks.  The chunk callback fires when the camera crosses a chunk border,
// the update callback when it moved chunk or turned.
class FVoxelCamera {
 public:
  FVec3 Position;                 // inside the current chunk
  FVec3 Forward{1.0f, 0.0f, 0.0f};  // look direction (need not be unit length)
  FVec3 Up{0.0f, 0.0f, 1.0f};     // +Z up as in VoxelCamera.h:23
  float FovDegrees = 60.0f;       // vertical field of view
  FIVec3 CameraChunkLocation;
  FVec3 CameraForward;

  using Callback = std::function<void()>;
  void SetCameraChunkUpdateCallback(Callback cb) { CameraChunkUpdateCallback = std::move(cb); }
  void SetCameraUpdateCallback(Callback cb) { CameraUpdateCallback = std::move(cb); }

  // places the camera at a world position looking along `forward`
  void SetPose(const FVec3 &world_eye, const FVec3 &forward, const FVoxelSceneConfig &cfg);
  FVec3 GetForwardVector() const { return Forward; }
  FVec3 GetWorldPosition(const FVoxelSceneConfig &cfg) const;
  // moves the floating origin to the chunk holding Position and fires the
  // callbacks (FVoxelCamera::UpdateCamera)
  void UpdateCamera(const FVoxelSceneConfig &cfg);
  // the ray camera of a width x height frame
  meso_camera GetRayCamera(const FVoxelSceneConfig &cfg, int32_t width, int32_t height) const;

 private:
  Callback CameraChunkUpdateCallback, CameraUpdateCallback;
};

// Double-buffered upload of bricks edited on the host: the engine fills one
// pinned batch while the GPU copies and applies the others (the role of the
// reference's double-buffered modify queues and staging upload,
// Runtimes/Voxel/Chunk/ChunkPool.h:137-168,641-705, here as pinned host
// batches, async copies and an async brick update on one stream).
class FBrickUploadRing {
 public:
  FBrickUploadRing(meso_context *ctx, meso_scene *scene, int32_t capacity_bricks, int32_t depth = 2);
  ~FBrickUploadRing();
  FBrickUploadRing(const FBrickUploadRing &) = delete;
  FBrickUploadRing &operator=(const FBrickUploadRing &) = delete;

  meso_status Init();
  // makes the next batch writable, waiting until the GPU is done with it
  meso_status Begin();
  // queues one brick (slot id, 512 u16 voxels x + 8y + 64z); false when full
  bool Add(uint32_t slot, const uint16_t *voxels);
  // queues the batch's copies and the brick update on `stream`; no wait
  meso_status Submit(meso_stream stream = nullptr);
  // waits for every submitted batch
  meso_status Drain();
  int32_t Capacity() const { return capacity_; }
  // bricks skipped (slot out of range) over all finished batches
  uint64_t Skipped();

 private:
  struct Batch {
    uint32_t *h_slots = nullptr;
    uint16_t *h_voxels = nullptr;
    meso_buffer *d_slots = nullptr, *d_voxels = nullptr;
    meso_event *done = nullptr;
    int32_t count = 0;
    bool in_flight = false;
  };
  meso_context *ctx_;
  meso_scene *scene_;
  int32_t capacity_, depth_;
  std::vector<Batch> batches_;
  meso_buffer *d_skipped_ = nullptr;
  int32_t cur_ = -1;
};

struct FFrameStats {
  float physics_ms = 0.0f;  // edits + remesh
  float render_ms = 0.0f;
  uint64_t quads = 0;       // quads of the chunks remeshed this frame
  int32_t dirty_chunks = 0;
  float stream_ms = 0.0f;   // chunk streaming: remesh of resident dirty chunks + pool update
  uint32_t pool_loads = 0;  // chunks the pool loaded this frame
  uint32_t pool_resident = 0;
};

struct FInstanceDesc {
  int32_t generator = MESO_GEN_CAVES;
  int32_t layout = MESO_SCENE_TWO_LEVEL_U16;
  int32_t n = 256;
  uint64_t seed = 5;
  int32_t n_spheres = 8;  // dense sphere scenes
  int32_t width = 256, height = 256;
  uint32_t ray_flags = MESO_RAY_SHADOW;
  float light[3] = {0.4f, 1.0f, 0.3f};
  uint32_t ao_seed = 0;
  int32_t ao_samples = 4;
  float ao_radius = 16.0f;
  uint32_t mesh_mode = MESO_MESH_GREEDY;
  int32_t max_dirty_chunks = 4096;  // per frame
  uint64_t quad_capacity = 16u << 20;
  // keep the chunks around the camera meshed in a meso_chunk_pool sized by
  // VoxelSceneConfig (the reference's ChunkManager streaming); brick scenes
  bool chunk_streaming = false;
};

class FHipVoxelInstance {
 public:
  explicit FHipVoxelInstance(const FInstanceDesc &desc);
  virtual ~FHipVoxelInstance();
  FHipVoxelInstance(const FHipVoxelInstance &) = delete;
  FHipVoxelInstance &operator=(const FHipVoxelInstance &) = delete;

  // creates the context, generates the scene and the frame's buffers
  meso_status Initialize(int32_t device = 0);
  // runs `frames` iterations of the instance loop; stops at the first error
  meso_status RunInstance(uint32_t frames);

  // hooks, in loop order
  virtual void InitializeBegin();
  virtual void UpdateCamera();
  virtual void UpdatePhysics();
  virtual void RenderStart();
  virtual void Render();
  virtual void RenderEnd();
  virtual void UpdateFrameIndex();
  virtual void WhenCameraChunkUpdate() {}
  virtual void WhenCameraUpdate();
  // the queued chunk edits and the remesh of the dirty chunks (part of UpdatePhysics)
  void ApplyEdits();

  // chunk edits applied by the next UpdatePhysics: chunk coordinates
  // (distinct x, y, z triples) and the seed of their sphere toggles
  void QueueChunkEdits(const std::vector<int32_t> &xyz, uint32_t seed);

  meso_status Status() const { return status_; }
  const std::string &ErrorMessage() const { return error_; }
  const FFrameStats &LastFrame() const { return last_; }
  const meso_camera &FrameCamera() const { return frame_camera_; }
  uint32_t RenderGlobalFrameIndex() const { return frame_index_; }
  meso_context *Context() const { return ctx_; }
  meso_scene *Scene() const { return scene_; }
  meso_chunk_pool *ChunkPool() const { return pool_; }
  // the camera chunk and forward vector the last pool update used
  const int32_t *StreamChunk() const { return stream_chunk_; }
  const float *StreamForward() const { return stream_forward_; }

  // copies of the last frame's outputs and the last remesh (host memory)
  meso_status DownloadFrame(std::vector<uint8_t> &index, std::vector<uint8_t> &face,
                            std::vector<float> &t, std::vector<float> &shade) const;
  meso_status DownloadMesh(std::vector<uint32_t> &counts, std::vector<uint32_t> &quads) const;

  FVoxelSceneConfig VoxelSceneConfig;
  FVoxelCamera WindowsCamera;

 protected:
  bool Check(meso_status st);

  FInstanceDesc desc_;
  meso_context *ctx_ = nullptr;
  meso_scene *scene_ = nullptr;
  meso_buffer *index_ = nullptr, *face_ = nullptr, *t_ = nullptr, *shade_ = nullptr;
  meso_buffer *dirty_ = nullptr, *counts_ = nullptr, *prefix_ = nullptr, *quads_ = nullptr;
  meso_buffer *total_ = nullptr, *work_ = nullptr;
  meso_event *ev_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  meso_chunk_pool *pool_ = nullptr;
  int32_t stream_chunk_[3] = {0, 0, 0};
  float stream_forward_[3] = {1.0f, 0.0f, 0.0f};
  std::vector<int32_t> queued_xyz_;
  uint32_t queued_seed_ = 0;
  int32_t meshed_chunks_ = 0;
  bool physics_ran_ = false;
  meso_camera frame_camera_{};
  FFrameStats last_;
  uint32_t frame_index_ = 0;
  meso_status status_ = MESO_OK;
  std::string error_;
};

}  // namespace meso_engine
