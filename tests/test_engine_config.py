"""The engine header's FVoxelSceneConfig takes the reference's initialisers.

Runtimes/Voxel/VoxelSceneConfig.h:20-49 defines the struct; the reference
value-initialises it (Runtimes/Voxel/Chunk/ChunkManager.h:131), keeps it as a
member (Runtimes/Instance/VoxelWindowsInstance.h:37,55) and reads its fields
(Samples/SimpleVoxel.cpp:261,329,415-423).  A translation unit written the way
reference code fills the struct -- every field by designated initialiser in
declaration order, member assignment, value-initialisation -- must compile
against mesoengine_amd/engine/meso_instance.h, and the defaults must be the
reference's.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SOURCE = r"""
#include "meso_instance.h"
using meso_engine::FVoxelSceneConfig;
using meso_engine::EChunkOverrideMode;

// defaults (VoxelSceneConfig.h:22-45)
constexpr FVoxelSceneConfig D{};
static_assert(D.BlockResolution == 8, "");
static_assert(D.BlockSize == 1.0f, "");
static_assert(D.ChunkResolution == 16, "");
static_assert(D.MaxBlockCount == 65536 * 16, "");
static_assert(D.MaxVolumeCount == 65536 * 16, "");
static_assert(D.MaxChunkCount == 8192 * 2, "");
static_assert(D.MaxEmptyChunkCount == 8192 * 4, "");
static_assert(D.MaxChunkCheckTimes == 128, "");
static_assert(D.MaxEmptyChunkCheckTimes == 128, "");
static_assert(D.MaxBlockCheckTimes == 16, "");
static_assert(D.BakeVisibilityViewNum == 256, "");
static_assert(D.ViewForwardLoadChunkSize == 24, "");
static_assert(D.ViewBackwardLoadChunkSize == 6, "");
static_assert(D.MaxSyncedLoadChunkCount == 0, "");
static_assert(D.MaxUnsyncedLoadChunkCount == 256, "");
static_assert(D.ChunkTaskPerCore == 8, "");
static_assert(D.ChunkOverrideMode == EChunkOverrideMode::FindMin, "");
static_assert(D.ViewChunkAngle == 120.0f, "");
static_assert(D.ChunkOccupancyDepth == 4, "");
static_assert(D.ChunkInnerVoxelCullDepthThreshold == 1, "");
static_assert(static_cast<uint8_t>(EChunkOverrideMode::FindLess) == 1, "");
static_assert(static_cast<uint8_t>(EChunkOverrideMode::OverrideMin) == 4, "");

// every reference field by designated initialiser, in declaration order
const FVoxelSceneConfig Full{
    .BlockResolution = 8,
    .BlockSize = 0.5f,
    .ChunkResolution = 16,
    .MaxBlockCount = 65536,
    .MaxVolumeCount = 65536,
    .MaxChunkCount = 4096,
    .MaxEmptyChunkCount = 8192,
    .MaxChunkCheckTimes = 64,
    .MaxEmptyChunkCheckTimes = 64,
    .MaxBlockCheckTimes = 8,
    .BakeVisibilityViewNum = 128,
    .ViewForwardLoadChunkSize = 16,
    .ViewBackwardLoadChunkSize = 4,
    .MaxSyncedLoadChunkCount = 2,
    .MaxUnsyncedLoadChunkCount = 128,
    .ChunkTaskPerCore = 4,
    .ChunkOverrideMode = EChunkOverrideMode::FindLess,
    .ViewChunkAngle = 90.0f,
    .ChunkOccupancyDepth = 4,
    .ChunkInnerVoxelCullDepthThreshold = 1,
};

// value-initialised member, member assignment and reads as the reference does
struct Holder {
  FVoxelSceneConfig VoxelSceneConfig = {};
};

int main() {
  Holder h;
  h.VoxelSceneConfig.MaxUnsyncedLoadChunkCount = 64;
  h.VoxelSceneConfig.ChunkOverrideMode = EChunkOverrideMode::OverrideMin;
  const unsigned threads = h.VoxelSceneConfig.MaxUnsyncedLoadChunkCount;
  const float chunk = Full.GetChunkSize();          // 16 blocks of 0.5
  const float mesh_chunk = Full.GetMeshChunkSize();  // 32 voxels of 0.5
  return (threads == 64 && chunk == 8.0f && mesh_chunk == 16.0f &&
          Full.MeshChunkResolution == 32 && Full.ChunkQuadCapacity == 8192) ? 0 : 1;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_reference_initialisers_compile_and_defaults_match(tmp_path):
    src = tmp_path / "cfg.cpp"
    src.write_text(SOURCE)
    exe = tmp_path / "cfg"
    cmd = ["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror=missing-field-initializers",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "mesoengine_amd", "engine"),
           str(src), "-o", str(exe)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    assert subprocess.run([str(exe)]).returncode == 0
