// Host code of the engine adapter (mesoengine_amd/engine/meso_instance.cpp)
// under AddressSanitizer + UndefinedBehaviorSanitizer, built by
// tests/test_native_sanitizers.py and linked against libmeso_hip.so.  Runs
// without a GPU: the floating-origin camera (FVoxelCamera, after
// Runtimes/Instance/VoxelCamera.cpp:42-59) over a long random walk, the ray
// camera it derives, and the instance's set-up error paths (a bad scene
// config; no device), with the instance destroyed after each failure.
#include <cmath>
#include <cstdio>
#include <random>

#include "meso_instance.h"

using namespace meso_engine;

static int failures = 0;
#define EXPECT(c)                                                                  \
  do {                                                                             \
    if (!(c)) {                                                                    \
      std::fprintf(stderr, "engine_sanitize: %s failed at line %d\n", #c, __LINE__); \
      failures++;                                                                  \
    }                                                                              \
  } while (0)

int main() {
  FVoxelSceneConfig cfg;
  FVoxelCamera cam;
  int chunk_moves = 0, updates = 0;
  cam.SetCameraChunkUpdateCallback([&] { chunk_moves++; });
  cam.SetCameraUpdateCallback([&] { updates++; });
  cam.SetPose({1000.5f, 300.25f, -77.0f}, {1.0f, 0.2f, 0.1f}, cfg);
  std::mt19937 rng(3);
  std::uniform_real_distribution<float> step(-20.0f, 20.0f);
  double wx = 1000.5, wy = 300.25, wz = -77.0;
  for (int i = 0; i < 20000; i++) {
    const float dx = step(rng), dy = step(rng), dz = step(rng);
    cam.Position.x += dx;
    cam.Position.y += dy;
    cam.Position.z += dz;
    wx += dx;
    wy += dy;
    wz += dz;
    if (i % 97 == 0) cam.Forward = {step(rng), step(rng), 1.0f};
    cam.UpdateCamera(cfg);
    const float cs = cfg.GetMeshChunkSize();
    EXPECT(cam.Position.x >= 0.0f && cam.Position.x < cs);
    EXPECT(cam.Position.y >= 0.0f && cam.Position.y < cs);
    EXPECT(cam.Position.z >= 0.0f && cam.Position.z < cs);
    const FVec3 w = cam.GetWorldPosition(cfg);
    EXPECT(std::fabs(w.x - wx) < 1e-2 && std::fabs(w.y - wy) < 1e-2 && std::fabs(w.z - wz) < 1e-2);
    const meso_camera rc = cam.GetRayCamera(cfg, 640, 360);
    EXPECT(rc.size == sizeof(meso_camera) && rc.width == 640 && rc.height == 360);
  }
  EXPECT(chunk_moves > 1000 && updates >= chunk_moves);

  // a scene config the library cannot run: refused before any device call
  {
    FInstanceDesc d;
    FHipVoxelInstance inst(d);
    inst.VoxelSceneConfig.MeshChunkResolution = 16;
    EXPECT(inst.Initialize(0) == MESO_ERR_ARGUMENT_OUT_OF_RANGE);
    EXPECT(!inst.ErrorMessage().empty());
    EXPECT(inst.RunInstance(2) != MESO_OK);
  }
  // no GPU here: the context cannot be created, the instance reports it and
  // its destructor releases what exists
  {
    FInstanceDesc d;
    d.chunk_streaming = true;
    FHipVoxelInstance inst(d);
    const meso_status st = inst.Initialize(0);
    EXPECT(st != MESO_OK);
    EXPECT(!inst.ErrorMessage().empty());
    EXPECT(inst.RunInstance(1) != MESO_OK);
  }
  if (failures) {
    std::fprintf(stderr, "engine_sanitize: %d failures\n", failures);
    return 1;
  }
  std::printf("engine_sanitize ok: %d chunk moves, %d camera updates\n", chunk_moves, updates);
  return 0;
}
