// Culled-face and greedy meshing of 32^3 chunks on gfx950.
//
// One workgroup meshes one chunk:
//  1. Row masks.  Wave64 ballots turn pairs of 32-voxel rows into 32-bit
//     masks: occupancy O[z][y] (bits over x) and material equality with the
//     +x, +y and +z neighbour (Ex, Ey, Ez).  A ballot transpose from LDS gives
//     the same masks with bits over y (OT, EyT, EzT) for the X-facing planes.
//  2. One lane per (face, slice) plane (6 x 32 = 192 lanes) derives its 32
//     face rows from those masks and merges them greedily with bit
//     operations: width from the run of set bits and of equal neighbours,
//     height by testing whole rows.  Quads are counted first, the workgroup
//     scans the 192 counts, and the lanes run again to write each quad at its
//     canonical position (face, slice, v, u).
//  3. Batched mode: chunks are taken from an atomic ticket, and each chunk
//     finds its output offset by an ordered decoupled look-back over the
//     chunks ticketed before it, so one persistent launch writes one compact
//     quad buffer plus per-chunk counts and exclusive prefixes.
#include "meso_internal.h"

namespace meso {

constexpr int kPlanes = 192;    // 6 faces x 32 slices
constexpr int kMeshBlock = 256;

struct ChunkAddr {
  const void *base;
  uint64_t off;   // element offset of voxel (0,0,0)
  uint64_t sy, sz;  // element strides of y and z
};

__device__ __forceinline__ ChunkAddr chunk_addr(const meso_chunk_source &src, uint32_t c) {
  ChunkAddr a;
  a.base = src.voxels;
  if (src.layout == MESO_CHUNKS_DENSE_GRID) {
    const uint32_t nc = (uint32_t)src.grid_n / 32u;
    const uint64_t N = (uint64_t)src.grid_n;
    const uint32_t cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
    a.off = (uint64_t)cx * 32u + (uint64_t)cy * 32u * N + (uint64_t)cz * 32u * N * N;
    a.sy = N;
    a.sz = N * N;
  } else {
    a.off = (uint64_t)c * 32768u;
    a.sy = 32;
    a.sz = 1024;
  }
  return a;
}

template <int ES>
__device__ __forceinline__ uint32_t vox(const ChunkAddr &a, int x, int y, int z) {
  const uint64_t i = a.off + (uint64_t)x + (uint64_t)y * a.sy + (uint64_t)z * a.sz;
  if constexpr (ES == 1)
    return ((const uint8_t *)a.base)[i];
  else
    return ((const uint16_t *)a.base)[i];
}

struct MeshLds {
  uint32_t O[32][32], Ex[32][32], Ey[32][32], Ez[32][32];  // [z][y], bits over x
  uint32_t OT[32][32], EyT[32][32], EzT[32][32];            // [z][x], bits over y
  uint32_t R[32][kPlanes];                                   // working face rows [v][plane]
  uint32_t scan[kMeshBlock / 64];
  uint32_t chunk;
  uint32_t base_lo, base_hi;
};

// Step 1: masks from global memory (each voxel and its +x/+y/+z neighbours).
template <int ES>
__device__ __forceinline__ void build_masks(MeshLds &L, const ChunkAddr &a) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  constexpr uint32_t nwaves = kMeshBlock / 64;
  const int x = (int)(lane & 31u);
  // 512 row pairs (y even, z)
  for (uint32_t p = wave; p < 512u; p += nwaves) {
    const int z = (int)(p >> 4);
    const int y = (int)((p & 15u) * 2u + (lane >> 5));
    const uint32_t c = vox<ES>(a, x, y, z);
    const uint32_t cx = x < 31 ? vox<ES>(a, x + 1, y, z) : 0u;
    const uint32_t cy = y < 31 ? vox<ES>(a, x, y + 1, z) : 0u;
    const uint32_t cz = z < 31 ? vox<ES>(a, x, y, z + 1) : 0u;
    const uint64_t bo = __ballot(c != 0u);
    const uint64_t bx = __ballot(x < 31 && c == cx);
    const uint64_t by = __ballot(y < 31 && c == cy);
    const uint64_t bz = __ballot(z < 31 && c == cz);
    if (lane == 0) {
      const int y0 = (int)((p & 15u) * 2u);
      L.O[z][y0] = (uint32_t)bo;
      L.O[z][y0 + 1] = (uint32_t)(bo >> 32);
      L.Ex[z][y0] = (uint32_t)bx;
      L.Ex[z][y0 + 1] = (uint32_t)(bx >> 32);
      L.Ey[z][y0] = (uint32_t)by;
      L.Ey[z][y0 + 1] = (uint32_t)(by >> 32);
      L.Ez[z][y0] = (uint32_t)bz;
      L.Ez[z][y0 + 1] = (uint32_t)(bz >> 32);
    }
  }
  __syncthreads();
  // ballot transpose: lane = y + 32 * (z & 1) holds row [z][y]; ballot of bit x
  // over the lanes is column x (bits over y) of planes z and z+1.
  for (uint32_t zp = wave; zp < 16u; zp += nwaves) {
    const int z = (int)(zp * 2u + (lane >> 5));
    const int y = (int)(lane & 31u);
    const uint32_t o = L.O[z][y], ey = L.Ey[z][y], ez = L.Ez[z][y];
    for (int xb = 0; xb < 32; xb++) {
      const uint64_t t0 = __ballot((o >> xb) & 1u);
      const uint64_t t1 = __ballot((ey >> xb) & 1u);
      const uint64_t t2 = __ballot((ez >> xb) & 1u);
      if (lane == 0) {
        const int z0 = (int)(zp * 2u);
        L.OT[z0][xb] = (uint32_t)t0;
        L.OT[z0 + 1][xb] = (uint32_t)(t0 >> 32);
        L.EyT[z0][xb] = (uint32_t)t1;
        L.EyT[z0 + 1][xb] = (uint32_t)(t1 >> 32);
        L.EzT[z0][xb] = (uint32_t)t2;
        L.EzT[z0 + 1][xb] = (uint32_t)(t2 >> 32);
      }
    }
  }
  __syncthreads();
}

// Face rows of plane (f, s): bit u of row v is set where the face is exposed.
// z faces: (u,v) = (x,y); y faces: (x,z); x faces: (y,z).
__device__ __forceinline__ uint32_t face_row(const MeshLds &L, int f, int s, int v) {
  switch (f) {
    case 0: return L.O[s][v] & ~(s < 31 ? L.O[s + 1][v] : 0u);
    case 1: return L.O[s][v] & ~(s > 0 ? L.O[s - 1][v] : 0u);
    case 2: return L.O[v][s] & ~(s < 31 ? L.O[v][s + 1] : 0u);
    case 3: return L.O[v][s] & ~(s > 0 ? L.O[v][s - 1] : 0u);
    case 4: return L.OT[v][s] & ~(s < 31 ? L.OT[v][s + 1] : 0u);
    default: return L.OT[v][s] & ~(s > 0 ? L.OT[v][s - 1] : 0u);
  }
}
// equality along u (bit u: cell u == cell u+1) and along v (cell (u,v) == (u,v+1))
__device__ __forceinline__ uint32_t eq_u(const MeshLds &L, int f, int s, int v) {
  return f < 2 ? L.Ex[s][v] : (f < 4 ? L.Ex[v][s] : L.EyT[v][s]);
}
__device__ __forceinline__ uint32_t eq_v(const MeshLds &L, int f, int s, int v) {
  return f < 2 ? L.Ey[s][v] : (f < 4 ? L.Ez[v][s] : L.EzT[v][s]);
}

__device__ __forceinline__ uint32_t pack_quad(int f, int s, int u, int v, int w, int h) {
  int x, y, z;
  if (f < 2) {
    x = u; y = v; z = s;
  } else if (f < 4) {
    x = u; y = s; z = v;
  } else {
    x = s; y = u; z = v;
  }
  return (uint32_t)x | ((uint32_t)y << 5) | ((uint32_t)z << 10) | ((uint32_t)(w - 1) << 15) |
         ((uint32_t)(h - 1) << 20) | ((uint32_t)f << 25);
}

// Step 2 for one plane.  Returns the quad count; with EMIT writes quad k to
// out[base + k] while base + k < cap.
template <bool GREEDY, bool EMIT>
__device__ __forceinline__ uint32_t mesh_plane(MeshLds &L, int t, uint32_t *out, uint64_t base,
                                               uint64_t cap) {
  const int f = t >> 5, s = t & 31;
  for (int v = 0; v < 32; v++) L.R[v][t] = face_row(L, f, s, v);
  uint32_t count = 0;
  for (int v = 0; v < 32; v++) {
    uint32_t row = L.R[v][t];
    if (!GREEDY && !EMIT) {
      count += (uint32_t)__popc(row);
      continue;
    }
    while (row) {
      const int u = __ffs(row) - 1;
      int w = 1, h = 1;
      uint32_t wmask = 1u << u;
      if (GREEDY) {
        // length of the run of set bits from u; a full row (u = 0, all 32
        // bits set) leaves no zero in ~row
        const uint32_t run = ~(row >> u);
        const int wr = run ? __ffs(run) - 1 : 32;
        const uint32_t eh = ~(eq_u(L, f, s, v) >> u);
        const int we = (eh ? __ffs(eh) - 1 : 32) + 1;
        w = wr < we ? wr : we;
        wmask = (w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1u)) << u;
        const uint32_t emask = w > 1 ? (((1u << (w - 1)) - 1u) << u) : 0u;
        while (v + h < 32) {
          const uint32_t r2 = L.R[v + h][t];
          if ((r2 & wmask) != wmask) break;
          if ((eq_u(L, f, s, v + h) & emask) != emask) break;
          if (!((eq_v(L, f, s, v + h - 1) >> u) & 1u)) break;
          h++;
        }
        for (int j = 1; j < h; j++) L.R[v + j][t] &= ~wmask;
      }
      row &= ~wmask;
      if (EMIT) {
        const uint64_t k = base + count;
        if (k < cap) out[k] = pack_quad(f, s, u, v, w, h);
      }
      count++;
    }
  }
  return count;
}

// exclusive scan of one value per thread over the workgroup; returns the total
__device__ __forceinline__ uint32_t block_exclusive_scan(MeshLds &L, uint32_t v, uint32_t &excl) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += o;
  }
  if (lane == 63) L.scan[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < kMeshBlock / 64; w++) {
    const uint32_t s = L.scan[w];
    if (w < wave) before += s;
    total += s;
  }
  excl = before + inc - v;
  __syncthreads();
  return total;
}

// look-back status word: flag in bits 62-63 (1 = aggregate, 2 = inclusive)
constexpr uint64_t kAgg = 1ull << 62, kInc = 2ull << 62, kValMask = (1ull << 62) - 1;

struct MeshWork {
  unsigned int ticket;
  unsigned int error;
  unsigned int pad[62];
  unsigned long long status[1];  // n_chunks entries
};

template <int ES, bool GREEDY>
__global__ void __launch_bounds__(kMeshBlock) mesh_batched_kernel(
    meso_chunk_source src, uint32_t *counts, uint32_t *prefix, uint32_t *quads, uint64_t capacity,
    unsigned long long *total, MeshWork *work) {
  __shared__ MeshLds L;
  const int t = (int)threadIdx.x;
  const uint32_t n_chunks = (uint32_t)src.n_chunks;
  for (;;) {
    if (t == 0) L.chunk = atomicAdd(&work->ticket, 1u);
    __syncthreads();
    const uint32_t c = L.chunk;
    if (c >= n_chunks) break;  // uniform: every wave leaves
    const ChunkAddr a = chunk_addr(src, c);
    build_masks<ES>(L, a);
    const uint32_t mine = t < kPlanes ? mesh_plane<GREEDY, false>(L, t, nullptr, 0, 0) : 0u;
    uint32_t excl = 0;
    const uint32_t chunk_total = block_exclusive_scan(L, mine, excl);
    if (t == 0) {
      // publish the aggregate, then add up predecessors until an inclusive one
      __hip_atomic_store(&work->status[c], kAgg | (uint64_t)chunk_total, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      uint64_t before = 0;
      int64_t j = (int64_t)c - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        const uint64_t st = __hip_atomic_load(&work->status[j], __ATOMIC_ACQUIRE,
                                              __HIP_MEMORY_SCOPE_AGENT);
        if (st & kInc) {
          before += st & kValMask;
          break;
        }
        if (st & kAgg) {
          before += st & kValMask;
          j--;
          continue;
        }
        // chunk j was ticketed earlier, so its workgroup is resident and will
        // publish; the bound only guarantees that every wave exits
        if (++spins > (1u << 24)) {
          atomicOr(&work->error, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(&work->status[c], kInc | (before + chunk_total), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      counts[c] = chunk_total;
      prefix[c] = (uint32_t)before;
      if (c == n_chunks - 1) *total = before + chunk_total;
      L.base_lo = (uint32_t)before;
      L.base_hi = (uint32_t)(before >> 32);
    }
    __syncthreads();
    const uint64_t base = ((uint64_t)L.base_hi << 32) | L.base_lo;
    if (t < kPlanes) mesh_plane<GREEDY, true>(L, t, quads, base + excl, capacity);
    __syncthreads();
  }
}

template <int ES, bool GREEDY>
__global__ void __launch_bounds__(kMeshBlock) mesh_plain_kernel(const void *voxels, uint32_t *count,
                                                                uint32_t *quads, uint32_t capacity) {
  __shared__ MeshLds L;
  const int t = (int)threadIdx.x;
  const ChunkAddr a{voxels, 0, 32, 1024};
  build_masks<ES>(L, a);
  const uint32_t mine = t < kPlanes ? mesh_plane<GREEDY, false>(L, t, nullptr, 0, 0) : 0u;
  uint32_t excl = 0;
  const uint32_t chunk_total = block_exclusive_scan(L, mine, excl);
  if (t == 0) *count = chunk_total;
  if (t < kPlanes) mesh_plane<GREEDY, true>(L, t, quads, excl, capacity);
}

template <int ES, bool GREEDY>
static meso_status launch_batched_t(meso_context *ctx, const meso_chunk_source *src,
                                    uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                    uint64_t capacity, uint64_t *d_total, void *d_work,
                                    hipStream_t stream, const MeshLaunch *shape) {
  int grid = shape && shape->grid > 0 ? shape->grid : 0;
  if (!grid) {
    int per_cu = 0;
    MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, mesh_batched_kernel<ES, GREEDY>, kMeshBlock, 0));
    if (per_cu < 1) per_cu = 1;
    grid = per_cu * ctx->props.multiProcessorCount;
  }
  if (grid > src->n_chunks) grid = src->n_chunks;
  meso_status st = launch_zero(d_work, meso_mesh_workspace_bytes(src->n_chunks), stream);
  if (st != MESO_OK) return st;
  hipLaunchKernelGGL((mesh_batched_kernel<ES, GREEDY>), dim3(grid), dim3(kMeshBlock), 0, stream,
                     *src, d_counts, d_prefix, d_quads, capacity, (unsigned long long *)d_total,
                     (MeshWork *)d_work);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

MeshLaunch default_mesh_launch(meso_context *) { return MeshLaunch{kMeshBlock, 0}; }

static meso_status check_source(const meso_chunk_source *src) {
  MESO_CHECK(src && src->size >= sizeof(meso_chunk_source), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_chunk_source missing or too small");
  MESO_CHECK(src->voxels, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "chunk voxels NULL");
  MESO_CHECK(src->elem_bytes == 1 || src->elem_bytes == 2, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "elem_bytes must be 1 or 2");
  MESO_CHECK(src->n_chunks >= 1 && src->n_chunks <= 40000, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d not in 1..40000 (quad offsets are u32)", src->n_chunks);
  if (src->layout == MESO_CHUNKS_DENSE_GRID) {
    MESO_CHECK(src->grid_n >= 32 && src->grid_n % 32 == 0 && src->grid_n <= 4096,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "grid_n %d must be a multiple of 32", src->grid_n);
    const int64_t nc = src->grid_n / 32;
    MESO_CHECK(src->n_chunks <= nc * nc * nc, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "n_chunks %d exceeds the %lld chunks of the grid", src->n_chunks,
               (long long)(nc * nc * nc));
  } else {
    MESO_CHECK(src->layout == MESO_CHUNKS_PACKED, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "unknown chunk layout %d", src->layout);
  }
  return MESO_OK;
}

meso_status launch_mesh_batched(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                                uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                uint64_t capacity, uint64_t *d_total, void *d_work,
                                hipStream_t stream, const MeshLaunch *shape) {
  const bool g = mode == MESO_MESH_GREEDY;
  if (src->elem_bytes == 1)
    return g ? launch_batched_t<1, true>(ctx, src, d_counts, d_prefix, d_quads, capacity, d_total,
                                         d_work, stream, shape)
             : launch_batched_t<1, false>(ctx, src, d_counts, d_prefix, d_quads, capacity,
                                          d_total, d_work, stream, shape);
  return g ? launch_batched_t<2, true>(ctx, src, d_counts, d_prefix, d_quads, capacity, d_total,
                                       d_work, stream, shape)
           : launch_batched_t<2, false>(ctx, src, d_counts, d_prefix, d_quads, capacity, d_total,
                                        d_work, stream, shape);
}

meso_status launch_mesh_plain(const void *voxels, int32_t es, uint32_t mode, uint32_t *d_count,
                              uint32_t *d_quads, uint32_t capacity, hipStream_t stream) {
  const bool g = mode == MESO_MESH_GREEDY;
  if (es == 1) {
    if (g)
      hipLaunchKernelGGL((mesh_plain_kernel<1, true>), dim3(1), dim3(kMeshBlock), 0, stream,
                         voxels, d_count, d_quads, capacity);
    else
      hipLaunchKernelGGL((mesh_plain_kernel<1, false>), dim3(1), dim3(kMeshBlock), 0, stream,
                         voxels, d_count, d_quads, capacity);
  } else {
    if (g)
      hipLaunchKernelGGL((mesh_plain_kernel<2, true>), dim3(1), dim3(kMeshBlock), 0, stream,
                         voxels, d_count, d_quads, capacity);
    else
      hipLaunchKernelGGL((mesh_plain_kernel<2, false>), dim3(1), dim3(kMeshBlock), 0, stream,
                         voxels, d_count, d_quads, capacity);
  }
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

}  // namespace meso

using namespace meso;

extern "C" {

uint64_t meso_mesh_workspace_bytes(int32_t n_chunks) {
  if (n_chunks < 1) n_chunks = 1;
  return 256u + 8u * (uint64_t)n_chunks;
}

meso_status meso_mesh_chunk(meso_context *ctx, const void *d_voxels, int32_t elem_bytes,
                            uint32_t mode, uint32_t *d_count, uint32_t *d_quads, uint32_t capacity,
                            meso_stream stream) {
  MESO_CHECK(ctx && d_voxels && d_count, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(capacity == 0 || d_quads, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "quads NULL");
  MESO_CHECK(elem_bytes == 1 || elem_bytes == 2, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "elem_bytes must be 1 or 2");
  MESO_CHECK(mode == MESO_MESH_CULLED || mode == MESO_MESH_GREEDY, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "unknown mesh mode %u", mode);
  return launch_mesh_plain(d_voxels, elem_bytes, mode, d_count, d_quads, capacity,
                           stream_of(ctx, stream));
}

meso_status meso_mesh_chunks(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                             uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                             uint64_t capacity, uint64_t *d_total, void *d_work,
                             meso_stream stream) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL context");
  meso_status st = check_source(src);
  if (st != MESO_OK) return st;
  MESO_CHECK(d_counts && d_prefix && d_total && d_work && (capacity == 0 || d_quads),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL output");
  MESO_CHECK(mode == MESO_MESH_CULLED || mode == MESO_MESH_GREEDY, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "unknown mesh mode %u", mode);
  const MeshLaunch *shape = nullptr;
  MeshLaunch tuned;
  {
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    char key[96];
    snprintf(key, sizeof(key), "mesh:%d:%d:%d:%u", src->elem_bytes, src->layout, src->n_chunks,
             mode);
    auto it = ctx->tune.find(key);
    if (it != ctx->tune.end()) {
      tuned = MeshLaunch{it->second.block, it->second.grid};
      shape = &tuned;
    }
  }
  return launch_mesh_batched(ctx, src, mode, d_counts, d_prefix, d_quads, capacity, d_total,
                             d_work, stream_of(ctx, stream), shape);
}

}  // extern "C"
