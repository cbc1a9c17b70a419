// meso_instance: runs the headless voxel instance for a number of frames.
//
// Each frame follows the instance loop (UpdateCamera, UpdatePhysics,
// RenderStart, Render, RenderEnd, UpdateFrameIndex).  The camera flies a
// fixed step along its forward vector every frame, and a seeded set of dirty
// chunks is edited and remeshed in UpdatePhysics.  One JSON line reports the
// per-frame timings (and, with --upload-bricks, the double-buffered upload of
// host-edited bricks, checked against the scene afterwards); --dump writes the
// last frame's outputs, the last remesh and every frame's dirty set to a
// little-endian binary file:
//   "MESOINST" u32 version=1
//   i32 width, height, index_bytes, frames
//   f32 eye[3], target[3], up[3], vfov      (the last frame's ray camera)
//   per frame: u32 edit_seed, i32 n_dirty, i32 xyz[3 n_dirty]
//   u32 n_chunks, u32 counts[n_chunks], u64 total, u32 quads[total]
//   index[width*height*index_bytes], u8 face[], f32 t[], f32 shade[]
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "meso_instance.h"

using namespace meso_engine;

namespace {

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// `count` distinct chunks of the nc^3 grid, drawn by rejection from SplitMix64
std::vector<int32_t> dirty_set(int32_t nc, int32_t count, uint64_t seed, uint32_t frame) {
  const uint64_t total = (uint64_t)nc * nc * nc;
  if ((uint64_t)count > total) count = (int32_t)total;
  uint64_t s = seed * 0x100000001B3ull + frame;
  std::unordered_set<uint64_t> seen;
  std::vector<int32_t> xyz;
  xyz.reserve(3 * (size_t)count);
  while ((int32_t)seen.size() < count) {
    const uint64_t c = splitmix64(s) % total;
    if (!seen.insert(c).second) continue;
    xyz.push_back((int32_t)(c % nc));
    xyz.push_back((int32_t)((c / nc) % nc));
    xyz.push_back((int32_t)(c / ((uint64_t)nc * nc)));
  }
  return xyz;
}

// Uploads `batches` batches of `per` host-edited bricks through the
// double-buffered ring, then checks the scene's voxels and masks of every
// touched brick against the last payload written to it.
bool run_upload(FHipVoxelInstance &inst, int32_t per, int32_t batches, uint64_t seed,
                std::string &out) {
  meso_scene_info info{};
  info.size = sizeof(info);
  if (meso_scene_get_info(inst.Scene(), &info) != MESO_OK || info.n_bricks <= 0) {
    out = "no bricks";
    return false;
  }
  const uint64_t nb = (uint64_t)info.n_bricks;
  if ((uint64_t)per > nb) per = (int32_t)nb;
  FBrickUploadRing ring(inst.Context(), inst.Scene(), per, 2);
  if (ring.Init() != MESO_OK) {
    out = meso_last_error_message();
    return false;
  }
  auto payload = [&](uint32_t slot, uint32_t batch, uint16_t *v) {
    uint64_t st = seed * 0x9E3779B97F4A7C15ull ^ ((uint64_t)batch << 40) ^ slot;
    for (int i = 0; i < 512; i += 4) {
      const uint64_t r = splitmix64(st);
      for (int k = 0; k < 4; k++) {
        const uint16_t x = (uint16_t)(r >> (16 * k));
        v[i + k] = (x & 1u) ? (uint16_t)(1u + (x >> 1) % 7u) : (uint16_t)0;
      }
    }
  };
  // the edits are made before the timed loop: it measures the upload path
  // (copy into the pinned batch, async copies, brick update), not the editor
  std::vector<int32_t> last(nb, -1);  // batch that last wrote each slot
  std::vector<uint16_t> vox(512);
  std::vector<uint32_t> slots((size_t)per * batches);
  std::vector<uint16_t> edits((size_t)per * batches * 512);
  uint64_t s = seed + 77u;
  for (int32_t b = 0; b < batches; b++) {
    std::unordered_set<uint32_t> used;
    size_t k = (size_t)b * per;
    while ((int32_t)used.size() < per) {
      const uint32_t slot = (uint32_t)(splitmix64(s) % nb);
      if (!used.insert(slot).second) continue;
      slots[k] = slot;
      payload(slot, (uint32_t)b, &edits[k * 512]);
      last[slot] = b;
      k++;
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  for (int32_t b = 0; b < batches; b++) {
    if (ring.Begin() != MESO_OK) {
      out = meso_last_error_message();
      return false;
    }
    for (size_t k = (size_t)b * per; k < (size_t)(b + 1) * per; k++) ring.Add(slots[k], &edits[k * 512]);
    if (ring.Submit() != MESO_OK) {
      out = meso_last_error_message();
      return false;
    }
  }
  if (ring.Drain() != MESO_OK) {
    out = meso_last_error_message();
    return false;
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // check every touched brick
  std::vector<uint16_t> all(nb * 512);
  std::vector<uint64_t> masks(nb * 8);
  if (meso_scene_download(inst.Scene(), MESO_PART_VOXELS, all.data(), all.size() * 2) != MESO_OK ||
      meso_scene_download(inst.Scene(), MESO_PART_MASKS, masks.data(), masks.size() * 8) != MESO_OK) {
    out = meso_last_error_message();
    return false;
  }
  bool ok = ring.Skipped() == 0;
  uint64_t touched = 0;
  for (uint64_t slot = 0; slot < nb && ok; slot++) {
    if (last[slot] < 0) continue;
    touched++;
    payload((uint32_t)slot, (uint32_t)last[slot], vox.data());
    for (int i = 0; i < 512 && ok; i++) {
      ok = all[slot * 512 + i] == vox[i];
      const bool bit = (masks[slot * 8 + i / 64] >> (i % 64)) & 1u;
      ok = ok && bit == (vox[i] != 0);
    }
  }
  const double bytes = (double)per * batches * (1024.0 + 4.0);
  char buf[256];
  std::snprintf(buf, sizeof(buf),
                "{\"bricks_per_batch\": %d, \"batches\": %d, \"ms\": %.3f, \"mbricks_s\": %.3f, "
                "\"gb_s\": %.2f, \"touched\": %llu, \"check\": %s}",
                per, batches, ms, per * (double)batches / ms / 1e3, bytes / ms / 1e6,
                (unsigned long long)touched, ok ? "true" : "false");
  out = buf;
  return ok;
}

bool parse3(const char *s, float v[3]) { return std::sscanf(s, "%f,%f,%f", &v[0], &v[1], &v[2]) == 3; }

int usage() {
  std::fprintf(stderr,
               "usage: meso_instance [--scene caves|terrain|terrain_caves|spheres] [--n N]\n"
               "  [--seed S] [--width W] [--height H] [--frames F] [--dirty D] [--ao K]\n"
               "  [--no-shadow] [--eye x,y,z] [--forward x,y,z] [--up x,y,z] [--step V]\n"
               "  [--mesh culled|greedy] [--dump FILE] [--device I]\n"
               "  [--upload-bricks N --upload-batches B]\n"
               "  [--stream [--pool-capacity C --pool-radius R --pool-loads L] [--dump-pool FILE]]\n");
  return 2;
}

}  // namespace

// The pool's state after the last frame and the per-frame inputs of its
// updates, for a parity replay: "MESOPOOL", version, frames, capacity,
// quads_per_slot, nc, then per frame the camera chunk (3 x i32) and forward
// (3 x f32), then lookup (nc^3 u32), slot_chunk, slot_stamp, slot_count
// (capacity u32 each) and every slot's quads (capacity x quads_per_slot u32).
static bool write_pool_dump(const FHipVoxelInstance &inst, const std::string &path,
                            const std::vector<int32_t> &chunks, const std::vector<float> &fwd) {
  meso_pool_arrays a{};
  a.size = sizeof(a);
  if (meso_chunk_pool_arrays(inst.ChunkPool(), &a) != MESO_OK) return false;
  const FVoxelSceneConfig &c = inst.VoxelSceneConfig;
  const uint64_t cap = c.MaxChunkCount, q = c.ChunkQuadCapacity;
  const uint64_t nc3 = (uint64_t)a.chunks_per_axis * a.chunks_per_axis * a.chunks_per_axis;
  std::vector<uint32_t> lookup(nc3), sc(cap), ss(cap), cnt(cap), quads(cap * q);
  meso_context *ctx = inst.Context();
  if (meso_memcpy_d2h(ctx, lookup.data(), a.lookup, nc3 * 4, nullptr) != MESO_OK ||
      meso_memcpy_d2h(ctx, sc.data(), a.slot_chunk, cap * 4, nullptr) != MESO_OK ||
      meso_memcpy_d2h(ctx, ss.data(), a.slot_stamp, cap * 4, nullptr) != MESO_OK ||
      meso_memcpy_d2h(ctx, cnt.data(), a.slot_count, cap * 4, nullptr) != MESO_OK ||
      meso_memcpy_d2h(ctx, quads.data(), a.quads, cap * q * 4, nullptr) != MESO_OK) {
    std::fprintf(stderr, "meso_instance: pool download: %s\n", meso_last_error_message());
    return false;
  }
  FILE *fo = std::fopen(path.c_str(), "wb");
  if (!fo) {
    std::perror(path.c_str());
    return false;
  }
  const uint32_t hdr[5] = {1u, (uint32_t)(chunks.size() / 3), (uint32_t)cap, (uint32_t)q,
                           (uint32_t)a.chunks_per_axis};
  std::fwrite("MESOPOOL", 1, 8, fo);
  std::fwrite(hdr, 4, 5, fo);
  for (size_t f = 0; f < chunks.size() / 3; f++) {
    std::fwrite(&chunks[3 * f], 4, 3, fo);
    std::fwrite(&fwd[3 * f], 4, 3, fo);
  }
  std::fwrite(lookup.data(), 4, nc3, fo);
  std::fwrite(sc.data(), 4, cap, fo);
  std::fwrite(ss.data(), 4, cap, fo);
  std::fwrite(cnt.data(), 4, cap, fo);
  std::fwrite(quads.data(), 4, cap * q, fo);
  std::fclose(fo);
  return true;
}

int main(int argc, char **argv) {
  FInstanceDesc d;
  std::string scene = "caves", dump, dump_pool;
  int pool_capacity = 0, pool_radius = 0, pool_loads = 0;
  uint32_t frames = 4;
  int32_t dirty = 64, device = 0;
  float eye[3] = {-1, -1, -1}, fwd[3] = {1.0f, -0.3f, 0.6f}, up[3] = {0.0f, 1.0f, 0.0f};
  float step = 1.5f;
  bool shadow = true;
  int32_t upload_bricks = 0, upload_batches = 0;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    const char *v = i + 1 < argc ? argv[i + 1] : nullptr;
    auto need = [&]() { return v != nullptr; };
    if (a == "--stream") {
      d.chunk_streaming = true;
      continue;
    }
    if (a == "--no-shadow") {
      shadow = false;
      continue;
    }
    if (!need()) return usage();
    i++;
    if (a == "--scene") scene = v;
    else if (a == "--n") d.n = std::atoi(v);
    else if (a == "--seed") d.seed = std::strtoull(v, nullptr, 10);
    else if (a == "--width") d.width = std::atoi(v);
    else if (a == "--height") d.height = std::atoi(v);
    else if (a == "--frames") frames = (uint32_t)std::atoi(v);
    else if (a == "--dirty") dirty = std::atoi(v);
    else if (a == "--ao") d.ao_samples = std::atoi(v);
    else if (a == "--step") step = (float)std::atof(v);
    else if (a == "--device") device = std::atoi(v);
    else if (a == "--dump") dump = v;
    else if (a == "--dump-pool") dump_pool = v;
    else if (a == "--pool-capacity") pool_capacity = std::atoi(v);
    else if (a == "--pool-radius") pool_radius = std::atoi(v);
    else if (a == "--pool-loads") pool_loads = std::atoi(v);
    else if (a == "--upload-bricks") upload_bricks = std::atoi(v);
    else if (a == "--upload-batches") upload_batches = std::atoi(v);
    else if (a == "--mesh") d.mesh_mode = std::string(v) == "culled" ? MESO_MESH_CULLED : MESO_MESH_GREEDY;
    else if (a == "--eye") { if (!parse3(v, eye)) return usage(); }
    else if (a == "--forward") { if (!parse3(v, fwd)) return usage(); }
    else if (a == "--up") { if (!parse3(v, up)) return usage(); }
    else return usage();
  }
  if (scene == "caves") {
    d.generator = MESO_GEN_CAVES;
    d.layout = MESO_SCENE_TWO_LEVEL_U16;
  } else if (scene == "terrain") {
    d.generator = MESO_GEN_HEIGHTFIELD;
    d.layout = MESO_SCENE_BRICKMAP_U16;
  } else if (scene == "terrain_caves") {
    d.generator = MESO_GEN_TERRAIN_CAVES;
    d.layout = MESO_SCENE_BRICKMAP_U16;
  } else if (scene == "spheres") {
    d.generator = MESO_GEN_SPHERES;
    d.layout = MESO_SCENE_DENSE_U8;
    dirty = 0;  // edits need a brick scene
  } else {
    return usage();
  }
  d.ray_flags = (shadow ? MESO_RAY_SHADOW : 0u) | (d.ao_samples > 0 ? MESO_RAY_AO : 0u);
  d.ao_seed = (uint32_t)d.seed;
  d.max_dirty_chunks = dirty > 0 ? dirty : 1;

  FHipVoxelInstance inst(d);
  if (pool_capacity > 0) inst.VoxelSceneConfig.MaxChunkCount = (uint32_t)pool_capacity;
  if (pool_radius > 0) inst.VoxelSceneConfig.ViewForwardLoadChunkSize = (uint32_t)pool_radius;
  if (pool_loads > 0) inst.VoxelSceneConfig.MaxUnsyncedLoadChunkCount = (uint32_t)pool_loads;
  if (inst.VoxelSceneConfig.ViewBackwardLoadChunkSize > inst.VoxelSceneConfig.ViewForwardLoadChunkSize)
    inst.VoxelSceneConfig.ViewBackwardLoadChunkSize = inst.VoxelSceneConfig.ViewForwardLoadChunkSize;
  if (inst.Initialize(device) != MESO_OK) {
    std::fprintf(stderr, "meso_instance: %s\n", inst.ErrorMessage().c_str());
    return 1;
  }
  int chunk_updates = 0;
  inst.WindowsCamera.SetCameraChunkUpdateCallback([&] { chunk_updates++; });
  if (eye[0] < 0) {
    eye[0] = 0.15f * (float)d.n;
    eye[1] = 0.6f * (float)d.n;
    eye[2] = 0.2f * (float)d.n;
  }
  inst.WindowsCamera.Up = {up[0], up[1], up[2]};
  inst.WindowsCamera.SetPose({eye[0], eye[1], eye[2]}, {fwd[0], fwd[1], fwd[2]},
                             inst.VoxelSceneConfig);
  const float fl = std::sqrt(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2]);
  const int32_t nc = (d.n + 31) / 32;

  std::vector<uint32_t> seeds;
  std::vector<std::vector<int32_t>> sets;
  std::vector<FFrameStats> stats;
  std::vector<int32_t> stream_chunks;  // per frame: the pool update's camera chunk
  std::vector<float> stream_fwd;       // and forward vector
  for (uint32_t f = 0; f < frames; f++) {
    if (dirty > 0) {
      sets.push_back(dirty_set(nc, dirty, d.seed, f));
      seeds.push_back((uint32_t)(d.seed * 7919u + f));
      inst.QueueChunkEdits(sets.back(), seeds.back());
    }
    if (inst.RunInstance(1) != MESO_OK) {
      std::fprintf(stderr, "meso_instance: frame %u: %s\n", f, inst.ErrorMessage().c_str());
      return 1;
    }
    stats.push_back(inst.LastFrame());
    stream_chunks.insert(stream_chunks.end(), inst.StreamChunk(), inst.StreamChunk() + 3);
    stream_fwd.insert(stream_fwd.end(), inst.StreamForward(), inst.StreamForward() + 3);
    // fly along the view direction; the next UpdateCamera moves the origin
    inst.WindowsCamera.Position.x += step * fwd[0] / fl;
    inst.WindowsCamera.Position.y += step * fwd[1] / fl;
    inst.WindowsCamera.Position.z += step * fwd[2] / fl;
  }

  std::printf("{\"scene\": \"%s\", \"n\": %d, \"width\": %d, \"height\": %d, \"frames\": %u, "
              "\"dirty_chunks\": %d, \"ao_samples\": %d, \"chunk_updates\": %d, \"render_ms\": [",
              scene.c_str(), d.n, d.width, d.height, frames, dirty, d.ao_samples, chunk_updates);
  for (size_t i = 0; i < stats.size(); i++) std::printf("%s%.4f", i ? ", " : "", stats[i].render_ms);
  std::printf("], \"physics_ms\": [");
  for (size_t i = 0; i < stats.size(); i++) std::printf("%s%.4f", i ? ", " : "", stats[i].physics_ms);
  std::printf("], \"quads\": [");
  for (size_t i = 0; i < stats.size(); i++)
    std::printf("%s%" PRIu64, i ? ", " : "", stats[i].quads);
  if (d.chunk_streaming) {
    std::printf("], \"stream_ms\": [");
    for (size_t i = 0; i < stats.size(); i++) std::printf("%s%.4f", i ? ", " : "", stats[i].stream_ms);
    std::printf("], \"pool_loads\": [");
    for (size_t i = 0; i < stats.size(); i++) std::printf("%s%u", i ? ", " : "", stats[i].pool_loads);
    std::printf("], \"pool_resident\": [");
    for (size_t i = 0; i < stats.size(); i++)
      std::printf("%s%u", i ? ", " : "", stats[i].pool_resident);
  }
  const meso_camera &c = inst.FrameCamera();
  std::printf("], \"eye\": [%.9g, %.9g, %.9g]", c.eye[0], c.eye[1], c.eye[2]);
  if (upload_bricks > 0 && upload_batches > 0 && dirty > 0) {
    std::string up;
    if (!run_upload(inst, upload_bricks, upload_batches, d.seed, up)) {
      std::fprintf(stderr, "meso_instance: upload: %s\n", up.c_str());
      return 1;
    }
    std::printf(", \"upload\": %s", up.c_str());
  }
  std::printf("}\n");
  std::fflush(stdout);

  if (!dump_pool.empty() && inst.ChunkPool()) {
    if (!write_pool_dump(inst, dump_pool, stream_chunks, stream_fwd)) return 1;
  }
  if (!dump.empty()) {
    std::vector<uint8_t> index, face;
    std::vector<float> t, shade;
    std::vector<uint32_t> counts, quads;
    if (inst.DownloadFrame(index, face, t, shade) != MESO_OK ||
        inst.DownloadMesh(counts, quads) != MESO_OK) {
      std::fprintf(stderr, "meso_instance: download: %s\n", meso_last_error_message());
      return 1;
    }
    FILE *fo = std::fopen(dump.c_str(), "wb");
    if (!fo) {
      std::perror(dump.c_str());
      return 1;
    }
    const uint32_t version = 1;
    const int32_t hdr[4] = {d.width, d.height, d.n > 1024 ? 8 : 4, (int32_t)frames};
    std::fwrite("MESOINST", 1, 8, fo);
    std::fwrite(&version, 4, 1, fo);
    std::fwrite(hdr, 4, 4, fo);
    std::fwrite(c.eye, 4, 3, fo);
    std::fwrite(c.target, 4, 3, fo);
    std::fwrite(c.up, 4, 3, fo);
    std::fwrite(&c.vfov_deg, 4, 1, fo);
    for (size_t f = 0; f < frames; f++) {
      const uint32_t s = f < seeds.size() ? seeds[f] : 0u;
      const int32_t nd = f < sets.size() ? (int32_t)(sets[f].size() / 3) : 0;
      std::fwrite(&s, 4, 1, fo);
      std::fwrite(&nd, 4, 1, fo);
      if (nd) std::fwrite(sets[f].data(), 4, sets[f].size(), fo);
    }
    const uint32_t nch = (uint32_t)counts.size();
    const uint64_t total = quads.size();
    std::fwrite(&nch, 4, 1, fo);
    if (nch) std::fwrite(counts.data(), 4, nch, fo);
    std::fwrite(&total, 8, 1, fo);
    if (total) std::fwrite(quads.data(), 4, total, fo);
    std::fwrite(index.data(), 1, index.size(), fo);
    std::fwrite(face.data(), 1, face.size(), fo);
    std::fwrite(t.data(), 4, t.size(), fo);
    std::fwrite(shade.data(), 4, shade.size(), fo);
    std::fclose(fo);
  }
  return 0;
}
