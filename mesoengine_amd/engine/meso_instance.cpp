// Headless voxel instance over libmeso_hip (see meso_instance.h).
#include "meso_instance.h"

#include <cmath>
#include <cstring>

namespace meso_engine {

// ------------------------------------------------------------------ camera
void FVoxelCamera::SetPose(const FVec3 &world_eye, const FVec3 &forward,
                           const FVoxelSceneConfig &cfg) {
  const float cs = cfg.GetMeshChunkSize();
  // floating origin: whole chunks in CameraChunkLocation, the rest in Position
  // (x - floor(x / cs) * cs is exact in float for a power-of-two chunk size)
  CameraChunkLocation = {(int32_t)std::floor(world_eye.x / cs), (int32_t)std::floor(world_eye.y / cs),
                         (int32_t)std::floor(world_eye.z / cs)};
  Position = {world_eye.x - (float)CameraChunkLocation.x * cs,
              world_eye.y - (float)CameraChunkLocation.y * cs,
              world_eye.z - (float)CameraChunkLocation.z * cs};
  Forward = forward;
}

FVec3 FVoxelCamera::GetWorldPosition(const FVoxelSceneConfig &cfg) const {
  const float cs = cfg.GetMeshChunkSize();
  return {(float)CameraChunkLocation.x * cs + Position.x, (float)CameraChunkLocation.y * cs + Position.y,
          (float)CameraChunkLocation.z * cs + Position.z};
}

void FVoxelCamera::UpdateCamera(const FVoxelSceneConfig &cfg) {
  const float cs = cfg.GetMeshChunkSize();
  // ConvertToChunkLocation (Runtimes/Helper/VoxelMathHelper.h:17-22): whole
  // chunks of the in-chunk position move into CameraChunkLocation
  const FIVec3 off{(int32_t)std::floor(Position.x / cs), (int32_t)std::floor(Position.y / cs),
                   (int32_t)std::floor(Position.z / cs)};
  const bool moved = off != FIVec3{};
  if (moved) {
    if (CameraChunkUpdateCallback) CameraChunkUpdateCallback();
    CameraChunkLocation = {CameraChunkLocation.x + off.x, CameraChunkLocation.y + off.y,
                           CameraChunkLocation.z + off.z};
    Position = {Position.x - (float)off.x * cs, Position.y - (float)off.y * cs,
                Position.z - (float)off.z * cs};
  }
  const FVec3 f = GetForwardVector();
  const bool turned = f.x != CameraForward.x || f.y != CameraForward.y || f.z != CameraForward.z;
  if (turned || moved) {
    CameraForward = f;
    if (CameraUpdateCallback) CameraUpdateCallback();
  }
}

meso_camera FVoxelCamera::GetRayCamera(const FVoxelSceneConfig &cfg, int32_t width,
                                       int32_t height) const {
  meso_camera c{};
  c.size = sizeof(c);
  const FVec3 e = GetWorldPosition(cfg);
  c.eye[0] = e.x;
  c.eye[1] = e.y;
  c.eye[2] = e.z;
  c.target[0] = e.x + Forward.x;
  c.target[1] = e.y + Forward.y;
  c.target[2] = e.z + Forward.z;
  c.up[0] = Up.x;
  c.up[1] = Up.y;
  c.up[2] = Up.z;
  c.vfov_deg = FovDegrees;
  c.width = width;
  c.height = height;
  return c;
}

// ------------------------------------------------------------------ instance
FHipVoxelInstance::FHipVoxelInstance(const FInstanceDesc &desc) : desc_(desc) {
  WindowsCamera.SetCameraChunkUpdateCallback([this] { WhenCameraChunkUpdate(); });
  WindowsCamera.SetCameraUpdateCallback([this] { WhenCameraUpdate(); });
}

FHipVoxelInstance::~FHipVoxelInstance() {
  if (pool_) meso_chunk_pool_destroy(pool_);
  for (meso_event *e : ev_)
    if (e) meso_event_destroy(e);
  for (meso_buffer *b : {index_, face_, t_, shade_, dirty_, counts_, prefix_, quads_, total_, work_})
    if (b) meso_buffer_destroy(b);
  if (scene_) meso_scene_destroy(scene_);
  if (ctx_) meso_context_destroy(ctx_);
}

bool FHipVoxelInstance::Check(meso_status st) {
  if (st != MESO_OK && status_ == MESO_OK) {
    status_ = st;
    error_ = meso_last_error_message();
  }
  return st == MESO_OK;
}

meso_status FHipVoxelInstance::Initialize(int32_t device) {
  if (VoxelSceneConfig.MeshChunkResolution != 32 || !(VoxelSceneConfig.BlockSize > 0.0f)) {
    // edits, remeshing and the pool work on the library's 32^3 chunks
    status_ = MESO_ERR_ARGUMENT_OUT_OF_RANGE;
    error_ = "FVoxelSceneConfig: MeshChunkResolution must be 32 and BlockSize positive";
    return status_;
  }
  if (!Check(meso_context_create(device, &ctx_))) return status_;
  InitializeBegin();
  return status_;
}

void FHipVoxelInstance::InitializeBegin() {
  meso_generate_desc g{};
  g.size = sizeof(g);
  g.generator = desc_.generator;
  g.layout = desc_.layout;
  g.n = desc_.n;
  g.seed = desc_.seed;
  g.n_spheres = desc_.n_spheres;
  if (!Check(meso_scene_generate(ctx_, &g, &scene_))) return;
  const uint64_t px = (uint64_t)desc_.width * (uint64_t)desc_.height;
  const int32_t ib = desc_.n > 1024 ? 8 : 4;  // int64 index once n^3 passes int32
  const uint64_t nd = (uint64_t)desc_.max_dirty_chunks;
  if (!Check(meso_buffer_create(ctx_, px * ib, nullptr, &index_)) ||
      !Check(meso_buffer_create(ctx_, px, nullptr, &face_)) ||
      !Check(meso_buffer_create(ctx_, px * 4, nullptr, &t_)) ||
      !Check(meso_buffer_create(ctx_, px * 4, nullptr, &shade_)) ||
      !Check(meso_buffer_create(ctx_, nd * 12, nullptr, &dirty_)) ||
      !Check(meso_buffer_create(ctx_, nd * 4, nullptr, &counts_)) ||
      !Check(meso_buffer_create(ctx_, nd * 4, nullptr, &prefix_)) ||
      !Check(meso_buffer_create(ctx_, desc_.quad_capacity * 4, nullptr, &quads_)) ||
      !Check(meso_buffer_create(ctx_, 16, nullptr, &total_)) ||
      !Check(meso_buffer_create(ctx_, meso_mesh_workspace_bytes((int32_t)nd), nullptr, &work_)))
    return;
  for (meso_event *&e : ev_)
    if (!Check(meso_event_create(ctx_, &e))) return;
  if (desc_.chunk_streaming) {
    // the pool of ChunkPool.h:250-300, sized by the scene config
    const FVoxelSceneConfig &c = VoxelSceneConfig;
    meso_pool_desc pd{};
    pd.size = sizeof(pd);
    pd.capacity = (int32_t)c.MaxChunkCount;
    pd.quads_per_slot = (int32_t)c.ChunkQuadCapacity;
    pd.forward_radius = (int32_t)c.ViewForwardLoadChunkSize;
    pd.backward_radius = (int32_t)c.ViewBackwardLoadChunkSize;
    pd.view_angle = c.ViewChunkAngle;
    pd.max_loads = (int32_t)(c.MaxSyncedLoadChunkCount + c.MaxUnsyncedLoadChunkCount);
    pd.mesh_mode = desc_.mesh_mode;
    pd.bake_directions = (int32_t)c.BakeVisibilityViewNum;
    Check(meso_chunk_pool_create(ctx_, scene_, &pd, &pool_));
  }
}

void FHipVoxelInstance::QueueChunkEdits(const std::vector<int32_t> &xyz, uint32_t seed) {
  queued_xyz_ = xyz;
  queued_seed_ = seed;
}

void FHipVoxelInstance::UpdateCamera() { WindowsCamera.UpdateCamera(VoxelSceneConfig); }

void FHipVoxelInstance::WhenCameraUpdate() {
  frame_camera_ = WindowsCamera.GetRayCamera(VoxelSceneConfig, desc_.width, desc_.height);
}

void FHipVoxelInstance::UpdatePhysics() {
  last_ = FFrameStats{};
  physics_ran_ = false;
  const int32_t n_dirty = (int32_t)(queued_xyz_.size() / 3);
  ApplyEdits();
  if (status_ != MESO_OK || !pool_) return;
  // the loading queue (FChunkManage::UpdateLoadingQueue, ChunkManager.h:211-330):
  // remesh the resident chunks the edits touched, then rank and load
  const int32_t *d_xyz = (const int32_t *)meso_buffer_device_address(dirty_, 0);
  const FIVec3 &cc = WindowsCamera.CameraChunkLocation;
  const FVec3 f = WindowsCamera.GetForwardVector();
  stream_chunk_[0] = cc.x;
  stream_chunk_[1] = cc.y;
  stream_chunk_[2] = cc.z;
  stream_forward_[0] = f.x;
  stream_forward_[1] = f.y;
  stream_forward_[2] = f.z;
  (void)(Check(meso_event_record(ev_[4], nullptr)) &&
         (!physics_ran_ || Check(meso_chunk_pool_remesh(pool_, d_xyz, n_dirty, nullptr))) &&
         Check(meso_chunk_pool_update(pool_, stream_chunk_, stream_forward_, nullptr)) &&
         Check(meso_event_record(ev_[5], nullptr)));
}

void FHipVoxelInstance::ApplyEdits() {
  if (status_ != MESO_OK || queued_xyz_.empty()) return;
  const int32_t n = (int32_t)(queued_xyz_.size() / 3);
  if (n > desc_.max_dirty_chunks) {
    Check(MESO_ERR_ARGUMENT_OUT_OF_RANGE);
    error_ = "more dirty chunks than max_dirty_chunks";
    return;
  }
  if (!Check(meso_buffer_upload(dirty_, queued_xyz_.data(), queued_xyz_.size() * 4, 0))) return;
  const int32_t *d_xyz = (const int32_t *)meso_buffer_device_address(dirty_, 0);
  if (!Check(meso_event_record(ev_[0], nullptr)) ||
      !Check(meso_scene_edit_chunks(scene_, d_xyz, n, queued_seed_, nullptr)) ||
      !Check(meso_mesh_scene_chunks(
          ctx_, scene_, d_xyz, n, desc_.mesh_mode,
          (uint32_t *)meso_buffer_device_address(counts_, 0),
          (uint32_t *)meso_buffer_device_address(prefix_, 0),
          (uint32_t *)meso_buffer_device_address(quads_, 0), desc_.quad_capacity,
          (uint64_t *)meso_buffer_device_address(total_, 0),
          (void *)meso_buffer_device_address(work_, 0), nullptr)) ||
      !Check(meso_event_record(ev_[1], nullptr)))
    return;
  meshed_chunks_ = n;
  last_.dirty_chunks = n;
  physics_ran_ = true;
  queued_xyz_.clear();
}

void FHipVoxelInstance::RenderStart() {
  // the camera of this frame (the reference uploads its camera UBO here,
  // VoxelWindowsInstance.cpp:410-419); the ray camera is passed by value
  if (frame_camera_.size == 0) WhenCameraUpdate();
}

void FHipVoxelInstance::Render() {
  if (status_ != MESO_OK) return;
  meso_shading sh{};
  sh.size = sizeof(sh);
  std::memcpy(sh.light, desc_.light, sizeof(sh.light));
  sh.ao_seed = desc_.ao_seed;
  sh.ao_samples = desc_.ao_samples;
  sh.ao_radius = desc_.ao_radius;
  meso_ray_outputs out{};
  out.size = sizeof(out);
  out.hit_index = (void *)meso_buffer_device_address(index_, 0);
  out.index_bytes = desc_.n > 1024 ? 8 : 4;
  out.face = (uint8_t *)meso_buffer_device_address(face_, 0);
  out.t = (float *)meso_buffer_device_address(t_, 0);
  out.shade = (float *)meso_buffer_device_address(shade_, 0);
  if (!Check(meso_event_record(ev_[2], nullptr)) ||
      !Check(meso_raymarch(ctx_, scene_, &frame_camera_, desc_.ray_flags, &sh, &out, nullptr)) ||
      !Check(meso_event_record(ev_[3], nullptr)))
    return;
}

void FHipVoxelInstance::RenderEnd() {
  // the frame's submit + wait (lvk submit/wait, LVK.h:798-801)
  if (status_ != MESO_OK) return;
  if (!Check(meso_event_sync(ev_[3]))) return;
  float ms = 0.0f;
  if (!Check(meso_event_elapsed_ms(ev_[2], ev_[3], &ms))) return;
  last_.render_ms = ms;
  if (pool_) {
    if (!Check(meso_event_elapsed_ms(ev_[4], ev_[5], &ms))) return;
    last_.stream_ms = ms;
    meso_pool_stats ps{};
    ps.size = sizeof(ps);
    if (!Check(meso_chunk_pool_stats(pool_, &ps, nullptr))) return;
    last_.pool_loads = ps.n_loads;
    last_.pool_resident = ps.n_resident;
  }
  if (physics_ran_) {
    if (!Check(meso_event_elapsed_ms(ev_[0], ev_[1], &ms))) return;
    last_.physics_ms = ms;
    uint64_t total = 0;
    if (!Check(meso_buffer_download(total_, &total, 8, 0))) return;
    last_.quads = total;
  }
}

void FHipVoxelInstance::UpdateFrameIndex() { frame_index_++; }

meso_status FHipVoxelInstance::RunInstance(uint32_t frames) {
  for (uint32_t f = 0; f < frames && status_ == MESO_OK; f++) {
    UpdateCamera();
    UpdatePhysics();
    RenderStart();
    Render();
    RenderEnd();
    UpdateFrameIndex();
  }
  return status_;
}

meso_status FHipVoxelInstance::DownloadFrame(std::vector<uint8_t> &index, std::vector<uint8_t> &face,
                                             std::vector<float> &t, std::vector<float> &shade) const {
  const uint64_t px = (uint64_t)desc_.width * (uint64_t)desc_.height;
  const uint64_t ib = desc_.n > 1024 ? 8 : 4;
  index.resize(px * ib);
  face.resize(px);
  t.resize(px);
  shade.resize(px);
  meso_status st = meso_buffer_download(index_, index.data(), px * ib, 0);
  if (st == MESO_OK) st = meso_buffer_download(face_, face.data(), px, 0);
  if (st == MESO_OK) st = meso_buffer_download(t_, t.data(), px * 4, 0);
  if (st == MESO_OK) st = meso_buffer_download(shade_, shade.data(), px * 4, 0);
  return st;
}

meso_status FHipVoxelInstance::DownloadMesh(std::vector<uint32_t> &counts,
                                            std::vector<uint32_t> &quads) const {
  counts.resize((size_t)meshed_chunks_);
  uint64_t total = 0;
  meso_status st = meso_buffer_download(total_, &total, 8, 0);
  if (st != MESO_OK) return st;
  if (total > desc_.quad_capacity) total = desc_.quad_capacity;
  quads.resize((size_t)total);
  if (meshed_chunks_) st = meso_buffer_download(counts_, counts.data(), counts.size() * 4, 0);
  if (st == MESO_OK && total) st = meso_buffer_download(quads_, quads.data(), total * 4, 0);
  return st;
}

// ------------------------------------------------------------------ uploads
FBrickUploadRing::FBrickUploadRing(meso_context *ctx, meso_scene *scene, int32_t capacity_bricks,
                                   int32_t depth)
    : ctx_(ctx), scene_(scene), capacity_(capacity_bricks), depth_(depth < 2 ? 2 : depth) {}

FBrickUploadRing::~FBrickUploadRing() {
  Drain();
  for (Batch &b : batches_) {
    if (b.h_slots) meso_host_free(ctx_, b.h_slots);
    if (b.h_voxels) meso_host_free(ctx_, b.h_voxels);
    if (b.d_slots) meso_buffer_destroy(b.d_slots);
    if (b.d_voxels) meso_buffer_destroy(b.d_voxels);
    if (b.done) meso_event_destroy(b.done);
  }
  if (d_skipped_) meso_buffer_destroy(d_skipped_);
}

meso_status FBrickUploadRing::Init() {
  if (!ctx_ || !scene_ || capacity_ <= 0) return MESO_ERR_ARGUMENT_OUT_OF_RANGE;
  batches_.resize((size_t)depth_);
  const uint64_t n = (uint64_t)capacity_;
  for (Batch &b : batches_) {
    meso_status st = meso_host_alloc(ctx_, n * 4, (void **)&b.h_slots);
    if (st == MESO_OK) st = meso_host_alloc(ctx_, n * 1024, (void **)&b.h_voxels);
    if (st == MESO_OK) st = meso_buffer_create(ctx_, n * 4, nullptr, &b.d_slots);
    if (st == MESO_OK) st = meso_buffer_create(ctx_, n * 1024, nullptr, &b.d_voxels);
    if (st == MESO_OK) st = meso_event_create(ctx_, &b.done);
    if (st != MESO_OK) return st;
  }
  meso_status st = meso_buffer_create(ctx_, 16, nullptr, &d_skipped_);
  if (st == MESO_OK) st = meso_buffer_fill_zero(d_skipped_, nullptr);
  return st;
}

meso_status FBrickUploadRing::Begin() {
  cur_ = (cur_ + 1) % depth_;
  Batch &b = batches_[(size_t)cur_];
  if (b.in_flight) {  // the GPU may still read this batch's host memory
    const meso_status st = meso_event_sync(b.done);
    if (st != MESO_OK) return st;
    b.in_flight = false;
  }
  b.count = 0;
  return MESO_OK;
}

bool FBrickUploadRing::Add(uint32_t slot, const uint16_t *voxels) {
  if (cur_ < 0) return false;
  Batch &b = batches_[(size_t)cur_];
  if (b.count >= capacity_) return false;
  b.h_slots[b.count] = slot;
  std::memcpy(b.h_voxels + (size_t)b.count * 512u, voxels, 1024);
  b.count++;
  return true;
}

meso_status FBrickUploadRing::Submit(meso_stream stream) {
  if (cur_ < 0) return MESO_ERR_ARGUMENT_OUT_OF_RANGE;
  Batch &b = batches_[(size_t)cur_];
  if (b.count == 0) return MESO_OK;
  void *ds = (void *)meso_buffer_device_address(b.d_slots, 0);
  void *dv = (void *)meso_buffer_device_address(b.d_voxels, 0);
  meso_status st = meso_memcpy_h2d_async(ctx_, ds, b.h_slots, (uint64_t)b.count * 4, stream);
  if (st == MESO_OK)
    st = meso_memcpy_h2d_async(ctx_, dv, b.h_voxels, (uint64_t)b.count * 1024, stream);
  if (st == MESO_OK)
    st = meso_scene_update_bricks_async(
        scene_, (const uint32_t *)ds, (const uint16_t *)dv, b.count,
        (uint32_t *)meso_buffer_device_address(d_skipped_, 0), stream);
  if (st == MESO_OK) st = meso_event_record(b.done, stream);
  if (st == MESO_OK) b.in_flight = true;
  return st;
}

meso_status FBrickUploadRing::Drain() {
  for (Batch &b : batches_)
    if (b.in_flight) {
      const meso_status st = meso_event_sync(b.done);
      if (st != MESO_OK) return st;
      b.in_flight = false;
    }
  return MESO_OK;
}

uint64_t FBrickUploadRing::Skipped() {
  if (Drain() != MESO_OK || !d_skipped_) return 0;
  uint32_t v = 0;
  return meso_buffer_download(d_skipped_, &v, 4, 0) == MESO_OK ? v : 0;
}

}  // namespace meso_engine
