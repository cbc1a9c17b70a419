// GPU voxelization: seeded scene generators into dense / brick-map /
// two-level layouts, the chunk batch for meshing, the reference sphere
// generator with its erosion mip chain, and brick edits.
#include <algorithm>
#include <cstring>
#include <vector>

#include "meso_internal.h"
#include "noise.h"
#include "traverse.h"

namespace meso {

// ------------------------------------------------------------------ scan
// Exclusive prefix sum of u32 flags (1024 elements per workgroup, recursive
// over block sums).  Used to assign brick and node slots in linear order.
__global__ void __launch_bounds__(256) scan_block_kernel(const uint32_t *in, uint32_t *out,
                                                         uint64_t n, uint32_t *block_sums) {
  __shared__ uint32_t s[256];
  const uint64_t base = (uint64_t)blockIdx.x * 1024u + threadIdx.x * 4u;
  uint32_t v[4];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;  // exclusive
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) block_sums[blockIdx.x] = s[255];
}

__global__ void scan_add_kernel(uint32_t *out, uint64_t n, const uint32_t *offsets) {
  const uint64_t i = (uint64_t)blockIdx.x * 1024u + threadIdx.x * 4u;
  const uint32_t add = offsets[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (i + k < n) out[i + k] += add;
}

// out[i] = sum(in[0..i)); returns the total through *total_host.
static meso_status scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *total_host,
                            hipStream_t stream) {
  if (n == 0) {
    *total_host = 0;
    return MESO_OK;
  }
  const uint64_t nb = (n + 1023) / 1024;
  uint32_t *sums = nullptr, *sums_scan = nullptr;
  MESO_HIP(hipMallocAsync((void **)&sums, nb * 4, stream));
  MESO_HIP(hipMallocAsync((void **)&sums_scan, nb * 4, stream));
  hipLaunchKernelGGL(scan_block_kernel, dim3((uint32_t)nb), dim3(256), 0, stream, in, out, n, sums);
  MESO_HIP(hipGetLastError());
  uint64_t total = 0;
  if (nb > 1) {
    meso_status st = scan_u32(sums, sums_scan, nb, &total, stream);
    if (st != MESO_OK) return st;
    hipLaunchKernelGGL(scan_add_kernel, dim3((uint32_t)nb), dim3(256), 0, stream, out, n, sums_scan);
    MESO_HIP(hipGetLastError());
  } else {
    uint32_t t = 0;
    MESO_HIP(hipMemcpyAsync(&t, sums, 4, hipMemcpyDeviceToHost, stream));
    MESO_HIP(hipStreamSynchronize(stream));
    total = t;
  }
  MESO_HIP(hipFreeAsync(sums, stream));
  MESO_HIP(hipFreeAsync(sums_scan, stream));
  *total_host = total;
  return MESO_OK;
}

// ------------------------------------------------------------------ dense spheres
constexpr int kMaxSpheres = 64;
struct SphereSet {
  int32_t count;
  int32_t p[4 * kMaxSpheres];  // cx, cy, cz, r in 1/256 voxel
};

// voxel centre (x + 1/2) against a sphere, exact integer arithmetic
__device__ __forceinline__ bool in_sphere_fp(const int32_t *s, int32_t x, int32_t y, int32_t z) {
  const int64_t dx = (int64_t)x * 256 + 128 - s[0];
  const int64_t dy = (int64_t)y * 256 + 128 - s[1];
  const int64_t dz = (int64_t)z * 256 + 128 - s[2];
  const int64_t r = s[3];
  return dx * dx + dy * dy + dz * dz <= r * r;
}

__global__ void spheres_dense_kernel(uint8_t *grid, int32_t n, SphereSet set,
                                     unsigned long long *solid) {
  const uint64_t total = (uint64_t)n * n * n;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // grid-stride: every wave runs the same number of iterations
  const uint64_t rounds = (total + stride - 1) / stride;
  for (uint64_t it = 0; it < rounds; it++) {
    const uint64_t i = it * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint8_t m = 0;
    if (i < total) {
      const int32_t x = (int32_t)(i % n), y = (int32_t)((i / n) % n),
                    z = (int32_t)(i / ((uint64_t)n * n));
      for (int k = 0; k < set.count; k++)
        if (in_sphere_fp(&set.p[4 * k], x, y, z)) {
          m = (uint8_t)(1 + k);
          break;
        }
      grid[i] = m;
    }
    const unsigned long long b = __ballot(m != 0);
    if ((threadIdx.x & 63u) == 0 && b) atomicAdd(solid, (unsigned long long)__popcll(b));
  }
}

// ------------------------------------------------------------------ brick scenes
__global__ void heights_kernel(int32_t *h, int32_t n, uint32_t seed, int32_t gen) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)n * n) return;
  const int32_t x = (int32_t)(i % n), z = (int32_t)(i / n);
  h[i] = gen == MESO_GEN_HEIGHTFIELD ? heightfield_height(x, z, n, seed)
                                     : terrain_height(x, z, n, seed);
}

// One workgroup of 512 lanes per brick: wave w holds local z = w and lane
// l = x + 8y, so the wave's ballot is exactly the brick's mask word z.
__device__ __forceinline__ uint16_t gen_voxel(int32_t gen, int32_t x, int32_t y, int32_t z,
                                              int32_t n, uint32_t seed, const int32_t *heights) {
  if (gen == MESO_GEN_HEIGHTFIELD) {
    return y < heights[(int64_t)z * n + x] ? height_band(y, n) : (uint16_t)0;
  } else if (gen == MESO_GEN_TERRAIN_CAVES) {
    return terrain_caves_voxel(x, y, z, n, seed, heights[(int64_t)z * n + x]);
  }
  return caves_voxel(x, y, z, seed);
}

// brick order: linear (brick map) or node-major (two-level: node * 4096 + local)
__device__ __forceinline__ void brick_of(uint64_t ordered, int32_t nb, bool two_level, int32_t &bx,
                                         int32_t &by, int32_t &bz) {
  if (!two_level) {
    bx = (int32_t)(ordered % nb);
    by = (int32_t)((ordered / nb) % nb);
    bz = (int32_t)(ordered / ((uint64_t)nb * nb));
  } else {
    const int32_t nt = nb / 16;
    const uint64_t node = ordered / 4096u;
    const uint32_t local = (uint32_t)(ordered % 4096u);
    const int32_t tx = (int32_t)(node % nt), ty = (int32_t)((node / nt) % nt),
                  tz = (int32_t)(node / ((uint64_t)nt * nt));
    bx = tx * 16 + (int32_t)(local & 15u);
    by = ty * 16 + (int32_t)((local >> 4) & 15u);
    bz = tz * 16 + (int32_t)(local >> 8);
  }
}

__global__ void __launch_bounds__(512) brick_flags_kernel(int32_t gen, int32_t n, uint32_t seed,
                                                          const int32_t *heights, int32_t two_level,
                                                          uint32_t *flags,
                                                          unsigned long long *solid,
                                                          uint64_t n_bricks) {
  const uint64_t bid = block_id();
  if (bid >= n_bricks) return;  // uniform over the workgroup
  const int32_t nb = n / 8;
  int32_t bx, by, bz;
  brick_of(bid, nb, two_level != 0, bx, by, bz);
  const int32_t lx = threadIdx.x & 7, ly = (threadIdx.x >> 3) & 7, lz = threadIdx.x >> 6;
  const uint16_t v = gen_voxel(gen, bx * 8 + lx, by * 8 + ly, bz * 8 + lz, n, seed, heights);
  const unsigned long long b = __ballot(v != 0);
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  if ((threadIdx.x & 63u) == 0 && b) atomicAdd(&cnt, (uint32_t)__popcll(b));
  __syncthreads();
  if (threadIdx.x == 0) {
    flags[bid] = cnt ? 1u : 0u;
    if (cnt) atomicAdd(solid, (unsigned long long)cnt);
  }
}

__global__ void __launch_bounds__(512) brick_write_kernel(int32_t gen, int32_t n, uint32_t seed,
                                                          const int32_t *heights, int32_t two_level,
                                                          const uint32_t *flags,
                                                          const uint32_t *prefix,
                                                          const uint32_t *node_ids,
                                                          uint32_t *slots, uint64_t *masks,
                                                          uint16_t *voxels, uint64_t n_bricks) {
  const int32_t nb = n / 8;
  const uint64_t ordered = block_id();
  if (ordered >= n_bricks) return;  // uniform over the workgroup
  if (two_level) {
    // bricks of empty nodes have no slot table
    const uint32_t node = node_ids[ordered / 4096u];
    if (node == kEmptySlot) return;
    if (threadIdx.x == 0)
      slots[(uint64_t)node * 4096u + ordered % 4096u] = flags[ordered] ? prefix[ordered] : kEmptySlot;
  } else if (threadIdx.x == 0) {
    slots[ordered] = flags[ordered] ? prefix[ordered] : kEmptySlot;
  }
  if (!flags[ordered]) return;
  int32_t bx, by, bz;
  brick_of(ordered, nb, two_level != 0, bx, by, bz);
  const int32_t lx = threadIdx.x & 7, ly = (threadIdx.x >> 3) & 7, lz = threadIdx.x >> 6;
  const uint16_t v = gen_voxel(gen, bx * 8 + lx, by * 8 + ly, bz * 8 + lz, n, seed, heights);
  const unsigned long long b = __ballot(v != 0);
  const uint64_t slot = prefix[ordered];
  voxels[slot * 512u + threadIdx.x] = v;
  if ((threadIdx.x & 63u) == 0) masks[slot * 8u + lz] = (uint64_t)b;
}

__global__ void node_flags_kernel(const uint32_t *brick_flags, uint32_t *node_flags) {
  __shared__ uint32_t any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  uint32_t a = 0;
  for (uint32_t i = threadIdx.x; i < 4096u; i += blockDim.x)
    a |= brick_flags[(uint64_t)blockIdx.x * 4096u + i];
  if (a) atomicOr(&any, 1u);
  __syncthreads();
  if (threadIdx.x == 0) node_flags[blockIdx.x] = any;
}

__global__ void node_ids_kernel(const uint32_t *node_flags, const uint32_t *node_prefix,
                                uint32_t *top, uint64_t n_top) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_top) top[i] = node_flags[i] ? node_prefix[i] : kEmptySlot;
}

// ------------------------------------------------------------------ chunk batch
struct ChunkParamsDev {
  uint32_t noise_seed;
  uint32_t target;       // solid voxels wanted from the noise
  int32_t n_spheres;
  int32_t sph[4][4];
};

__device__ __forceinline__ int32_t chunk_noise_q(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  float v = fbm3((float)x, (float)y, (float)z, 2, 0.125f, seed);
  v = v * 256.0f;
  int32_t q = (int32_t)v;
  return q > 255 ? 255 : q;
}

// One workgroup per chunk: quantised noise into LDS with a 256-bin histogram,
// threshold at the fill quantile, spheres drawn over it.
__global__ void __launch_bounds__(1024) chunk_gen_kernel(const ChunkParamsDev *params,
                                                         uint16_t *out) {
  __shared__ uint8_t q[32768];
  __shared__ uint32_t hist[256];
  __shared__ uint32_t thr;
  const ChunkParamsDev p = params[blockIdx.x];
  if (threadIdx.x < 256) hist[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 32768u; i += 1024u) {
    const int32_t x = (int32_t)(i & 31u), y = (int32_t)((i >> 5) & 31u), z = (int32_t)(i >> 10);
    const int32_t v = chunk_noise_q(x, y, z, p.noise_seed);
    q[i] = (uint8_t)v;
    atomicAdd(&hist[v], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0, cum = 0;
    while (cum < p.target && t < 256u) {
      cum += hist[t];
      t++;
    }
    thr = t;
  }
  __syncthreads();
  const uint32_t T = thr;
  uint16_t *o = out + (uint64_t)blockIdx.x * 32768u;
  for (uint32_t i = threadIdx.x; i < 32768u; i += 1024u) {
    const int32_t x = (int32_t)(i & 31u), y = (int32_t)((i >> 5) & 31u), z = (int32_t)(i >> 10);
    uint16_t m = 0;
    for (int s = 0; s < p.n_spheres; s++)
      if (in_sphere_fp(p.sph[s], x, y, z)) {
        m = (uint16_t)(5 + s);
        break;
      }
    const uint32_t qi = q[i];
    if (!m && qi < T) m = (uint16_t)(1u + (qi * 4u) / T);
    o[i] = m;
  }
}

// ------------------------------------------------------------------ reference sphere + erosion
// FGeneratorHelper::GenerateSphere (Runtimes/Helper/GeneratorHelper.h:120-150)
// and FChunk::CalculateOccupancyErodeMipmaps (Runtimes/Voxel/Chunk/Chunk.h:73-94)
// with the 26-neighbour ErodeSingleVoxel rule (BinaryOccupancyVolume.h:76-99).
__global__ void __launch_bounds__(256) sphere_chunk_kernel(const int32_t *chunk_xyz, int32_t res,
                                                           double block_size, int32_t depth,
                                                           uint8_t *mips) {
  extern __shared__ uint8_t lds[];  // 2 * res^3
  const int32_t vol = res * res * res;
  uint8_t *cur = lds, *prev = lds + vol;
  const int32_t cx = chunk_xyz[3 * blockIdx.x], cy = chunk_xyz[3 * blockIdx.x + 1],
                cz = chunk_xyz[3 * blockIdx.x + 2];
  const double cr = (double)res;
  uint8_t *out = mips + (uint64_t)blockIdx.x * depth * vol;
  for (int32_t i = threadIdx.x; i < vol; i += blockDim.x) {
    const int32_t X = i % res, Y = (i / res) % res, Z = i / (res * res);
    double px = (double)cx * block_size * cr + (double)X * block_size;
    double py = (double)cy * block_size * cr + (double)Y * block_size;
    double pz = (double)cz * block_size * cr + (double)Z * block_size;
    double dx = px - 100.0, dy = py, dz = pz;
    double dd = dx * dx + dy * dy;
    dd = dd + dz * dz;
    double d = __builtin_sqrt(dd) - 50.0;
    const uint8_t s = d < 0.0 ? 1 : 0;
    prev[i] = s;
    out[i] = s;
  }
  __syncthreads();
  for (int32_t lvl = 1; lvl < depth; lvl++) {
    for (int32_t i = threadIdx.x; i < vol; i += blockDim.x) {
      const int32_t X = i % res, Y = (i / res) % res, Z = i / (res * res);
      uint8_t e = 0;
      if (out[i] && X >= 1 && Y >= 1 && Z >= 1 && X < res - 1 && Y < res - 1 && Z < res - 1) {
        e = 1;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              if (!dx && !dy && !dz) continue;
              e &= prev[(X + dx) + (Y + dy) * res + (Z + dz) * res * res];
            }
      }
      cur[i] = e;
      out[(uint64_t)lvl * vol + i] = e;
    }
    __syncthreads();
    uint8_t *t = cur;
    cur = prev;
    prev = t;
  }
}

// ------------------------------------------------------------------ brick edits
__global__ void __launch_bounds__(512) brick_update_kernel(const uint32_t *edit_slots,
                                                           const uint16_t *edit_vox,
                                                           int64_t n_bricks, uint64_t *masks,
                                                           uint64_t *blocks,
                                                           const uint32_t *slot_cell,
                                                           uint16_t *voxels,
                                                           uint32_t *bad, uint64_t n_edit) {
  const uint64_t e = block_id();
  if (e >= n_edit) return;
  const uint32_t slot = edit_slots[e];
  if ((int64_t)slot >= n_bricks) {  // every lane takes the same branch
    if (threadIdx.x == 0 && bad) atomicAdd(bad, 1u);
    return;
  }
  const uint16_t v = edit_vox[e * 512u + threadIdx.x];
  const unsigned long long b = __ballot(v != 0);
  voxels[(uint64_t)slot * 512u + threadIdx.x] = v;
  if ((threadIdx.x & 63u) == 0) masks[(uint64_t)slot * 8u + (threadIdx.x >> 6)] = (uint64_t)b;
  const uint64_t sub = brick_blocks2_from_planes((uint64_t)b);
  if (threadIdx.x == 0 && slot_cell[slot] != 0u) blocks[slot_cell[slot] - 1u] = sub;
}

// ------------------------------------------------------------------ chunk edits
// cell (index into slots) of brick (bx,by,bz) in a brick map or two-level
// scene, or kEmptySlot outside the grid or in an empty node
__device__ __forceinline__ uint32_t scene_brick_cell(const DevScene &s, int32_t kind, int32_t bx,
                                                     int32_t by, int32_t bz) {
  const int32_t nb = s.n >> 3;
  if (bx < 0 || by < 0 || bz < 0 || bx >= nb || by >= nb || bz >= nb) return kEmptySlot;
  if (kind == MESO_SCENE_BRICKMAP_U16)
    return (uint32_t)bx + ((uint32_t)by + (uint32_t)bz * (uint32_t)nb) * (uint32_t)nb;
  const uint32_t nt = (uint32_t)s.n >> 7;
  const uint32_t node = s.top[((uint32_t)bx >> 4) + (((uint32_t)by >> 4) + ((uint32_t)bz >> 4) * nt) * nt];
  if (node == kEmptySlot) return kEmptySlot;
  return node * 4096u + ((uint32_t)bx & 15u) + 16u * ((uint32_t)by & 15u) +
         256u * ((uint32_t)bz & 15u);
}

// One workgroup of 512 lanes per chunk (lane = voxel of a brick), walking the
// chunk's allocated bricks: a launch of one workgroup per brick spent its
// time dispatching workgroups, two thirds of which found no brick allocated.
// The first wave looks up the 64 bricks' cells and slots at once (one lane
// each) and lists the allocated ones; each brick's voxels are loaded while
// the one before is folded.
__global__ void __launch_bounds__(512) edit_chunks_kernel(DevScene s, int32_t kind,
                                                          const int32_t *xyz, uint32_t seed,
                                                          uint64_t *masks, uint64_t *blocks,
                                                          uint16_t *voxels, uint64_t n_chunks) {
  __shared__ uint32_t live_b[64], live_cell[64], live_slot[64];
  __shared__ uint32_t n_live;
  const uint64_t chunk = block_id();
  if (chunk >= n_chunks) return;
  const int32_t nc = (s.n + 31) >> 5;  // a partial last chunk counts
  const int32_t cx = xyz[3 * chunk], cy = xyz[3 * chunk + 1], cz = xyz[3 * chunk + 2];
  if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc) return;
  const uint32_t i = threadIdx.x;
  if (i < 64u) {
    const int32_t bx = cx * 4 + (int32_t)(i & 3u), by = cy * 4 + (int32_t)((i >> 2) & 3u),
                  bz = cz * 4 + (int32_t)(i >> 4);
    const uint32_t cell = scene_brick_cell(s, kind, bx, by, bz);
    const uint32_t slot = cell == kEmptySlot ? kEmptySlot : s.slots[cell];
    const unsigned long long live = __ballot(slot != kEmptySlot);
    if (slot != kEmptySlot) {
      const uint32_t k = (uint32_t)__popcll(live & ((1ull << i) - 1ull));  // in brick order
      live_b[k] = i;
      live_cell[k] = cell;
      live_slot[k] = slot;
    }
    if (i == 0) n_live = (uint32_t)__popcll(live);
  }
  __syncthreads();
  const uint32_t n = n_live;
  if (n == 0) return;  // uniform over the workgroup
  const uint32_t h0 = nhash32(seed ^ nhash32((uint32_t)cx + ((uint32_t)cy + (uint32_t)cz * (uint32_t)nc) * (uint32_t)nc));
  const uint32_t h1 = nhash32(h0), h2 = nhash32(h1), h3 = nhash32(h2);
  const int64_t sx = (int64_t)cx * 8192 + (h0 & 0x1FFFu), sy = (int64_t)cy * 8192 + (h1 & 0x1FFFu),
                sz = (int64_t)cz * 8192 + (h2 & 0x1FFFu);
  const int64_t r = 4 * 256 + (int64_t)(h3 % (8u * 256u));
  const int32_t lx = (int32_t)(i & 7u), ly = (int32_t)((i >> 3) & 7u), lz = (int32_t)(i >> 6);
  uint16_t next = voxels[(uint64_t)live_slot[0] * 512u + i];
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t b = live_b[k], cell = live_cell[k], slot = live_slot[k];
    uint16_t val = next;
    // the next brick's voxels are in flight meanwhile (bricks own their slots)
    if (k + 1 < n) next = voxels[(uint64_t)live_slot[k + 1] * 512u + i];
    const int32_t bx = cx * 4 + (int32_t)(b & 3u), by = cy * 4 + (int32_t)((b >> 2) & 3u),
                  bz = cz * 4 + (int32_t)(b >> 4);
    const int64_t dx = (int64_t)(bx * 8 + lx) * 256 + 128 - sx;
    const int64_t dy = (int64_t)(by * 8 + ly) * 256 + 128 - sy;
    const int64_t dz = (int64_t)(bz * 8 + lz) * 256 + 128 - sz;
    if (dx * dx + dy * dy + dz * dz <= r * r) {
      val = val ? (uint16_t)0 : (uint16_t)7;
      voxels[(uint64_t)slot * 512u + i] = val;
    }
    const unsigned long long m = __ballot(val != 0);
    if ((i & 63u) == 0) masks[(uint64_t)slot * 8u + (uint32_t)lz] = (uint64_t)m;
    const uint64_t sub = brick_blocks2_from_planes((uint64_t)m);
    if (i == 0) blocks[cell] = sub;
    __syncthreads();  // the fold's LDS planes are written again for the next brick
  }
}

}  // namespace meso

using namespace meso;

static meso_status gen_free(meso_scene *sc, meso_status st) {
  if (sc) {
    hipSetDevice(sc->ctx->device);
    hipStreamSynchronize(sc->ctx->stream);
    free_scene_arrays(sc);
    delete sc;
  }
  return st;
}

#define GEN_TRY(expr)                      \
  do {                                     \
    meso_status st_ = (expr);              \
    if (st_ != MESO_OK) return gen_free(sc, st_); \
  } while (0)
#define GEN_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return gen_free(sc, fail(e_ == hipErrorOutOfMemory ? MESO_ERR_OUT_OF_MEMORY      \
                                                         : MESO_ERR_RUNTIME,           \
                               "%s: %s", #call, hipGetErrorString(e_)));               \
  } while (0)

extern "C" {

meso_status meso_scene_generate(meso_context *ctx, const meso_generate_desc *d, meso_scene **out) {
  MESO_CHECK(ctx && d && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(d->size >= sizeof(meso_generate_desc), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_generate_desc.size too small");
  *out = nullptr;
  const int32_t n = d->n;
  const uint32_t seed = (uint32_t)(d->seed ^ (d->seed >> 32));
  MESO_HIP(hipSetDevice(ctx->device));
  hipStream_t stream = ctx->stream;
  meso_scene *sc = nullptr;
  if (d->generator == MESO_GEN_SPHERES) {
    MESO_CHECK(d->layout == MESO_SCENE_DENSE_U8, MESO_ERR_NOT_SUPPORTED,
               "sphere scenes are generated as dense u8 grids");
    MESO_CHECK(n >= 1 && n <= 2048, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "dense edge %d", n);
    MESO_CHECK(d->n_spheres >= 1 && d->n_spheres <= kMaxSpheres && d->n_spheres < 256,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "n_spheres %d not in 1..%d", d->n_spheres,
               kMaxSpheres);
    SphereSet set{};
    set.count = d->n_spheres;
    Pcg32 r;
    r.seed(d->seed);
    const uint32_t span = (uint32_t)n * 256u;
    for (int i = 0; i < d->n_spheres; i++) {
      set.p[4 * i + 0] = (int32_t)(r.next() % span);
      set.p[4 * i + 1] = (int32_t)(r.next() % span);
      set.p[4 * i + 2] = (int32_t)(r.next() % span);
      set.p[4 * i + 3] = (int32_t)(8 * 256 + r.next() % (24u * 256u + 1u));
    }
    sc = new meso_scene();
    sc->ctx = ctx;
    sc->kind = MESO_SCENE_DENSE_U8;
    sc->n = n;
    const uint64_t total = (uint64_t)n * n * n;
    GEN_HIP(hipMalloc((void **)&sc->dense, total));
    sc->bytes = total;
    unsigned long long *solid = nullptr;
    GEN_HIP(hipMallocAsync((void **)&solid, 8, stream));
    GEN_TRY(launch_zero(solid, 8, stream));
    hipLaunchKernelGGL(spheres_dense_kernel, dim3((uint32_t)std::min<uint64_t>((total + 255) / 256, 1u << 20)), dim3(256), 0,
                       stream, sc->dense, n, set, solid);
    GEN_HIP(hipGetLastError());
    unsigned long long s = 0;
    GEN_HIP(hipMemcpyAsync(&s, solid, 8, hipMemcpyDeviceToHost, stream));
    GEN_HIP(hipFreeAsync(solid, stream));
    GEN_HIP(hipStreamSynchronize(stream));
    sc->solid = (int64_t)s;
    *out = sc;
    return MESO_OK;
  }
  MESO_CHECK(d->generator == MESO_GEN_HEIGHTFIELD || d->generator == MESO_GEN_TERRAIN_CAVES ||
                 d->generator == MESO_GEN_CAVES,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown generator %d", d->generator);
  const bool two = d->layout == MESO_SCENE_TWO_LEVEL_U16;
  MESO_CHECK(d->layout == MESO_SCENE_BRICKMAP_U16 || two, MESO_ERR_NOT_SUPPORTED,
             "procedural scenes are built as brick maps or two-level maps");
  MESO_CHECK(two ? (n >= 128 && n % 128 == 0 && n <= 4096) : (n >= 8 && n % 8 == 0 && n <= 4096),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "edge %d invalid for layout %d", n, d->layout);
  const int32_t nb = n / 8;
  const uint64_t n_brick_cells = (uint64_t)nb * nb * nb;
  MESO_CHECK(n_brick_cells < 0xFFFFFFFFull, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "too many bricks");
  sc = new meso_scene();
  sc->ctx = ctx;
  sc->kind = two ? MESO_SCENE_TWO_LEVEL_U16 : MESO_SCENE_BRICKMAP_U16;
  sc->n = n;
  int32_t *heights = nullptr;
  uint32_t *flags = nullptr, *prefix = nullptr, *node_flags = nullptr, *node_prefix = nullptr;
  unsigned long long *solid = nullptr;
  if (d->generator != MESO_GEN_CAVES) {
    GEN_HIP(hipMallocAsync((void **)&heights, (uint64_t)n * n * 4, stream));
    hipLaunchKernelGGL(heights_kernel, dim3((uint32_t)(((uint64_t)n * n + 255) / 256)), dim3(256),
                       0, stream, heights, n, seed, d->generator);
    GEN_HIP(hipGetLastError());
  }
  GEN_HIP(hipMallocAsync((void **)&flags, n_brick_cells * 4, stream));
  GEN_HIP(hipMallocAsync((void **)&prefix, n_brick_cells * 4, stream));
  GEN_HIP(hipMallocAsync((void **)&solid, 8, stream));
  GEN_TRY(launch_zero(solid, 8, stream));
  hipLaunchKernelGGL(brick_flags_kernel, grid_1d(n_brick_cells), dim3(512), 0, stream,
                     d->generator, n, seed, heights, two ? 1 : 0, flags, solid, n_brick_cells);
  GEN_HIP(hipGetLastError());
  uint64_t n_bricks = 0;
  GEN_TRY(scan_u32(flags, prefix, n_brick_cells, &n_bricks, stream));
  uint64_t n_nodes = 0;
  const int32_t nt = n / 128;
  const uint64_t n_top = two ? (uint64_t)nt * nt * nt : 0;
  if (two) {
    GEN_HIP(hipMallocAsync((void **)&node_flags, n_top * 4, stream));
    GEN_HIP(hipMallocAsync((void **)&node_prefix, n_top * 4, stream));
    hipLaunchKernelGGL(node_flags_kernel, dim3((uint32_t)n_top), dim3(256), 0, stream, flags,
                       node_flags);
    GEN_HIP(hipGetLastError());
    GEN_TRY(scan_u32(node_flags, node_prefix, n_top, &n_nodes, stream));
    GEN_HIP(hipMalloc((void **)&sc->top, n_top * 4));
    hipLaunchKernelGGL(node_ids_kernel, dim3((uint32_t)((n_top + 255) / 256)), dim3(256), 0,
                       stream, node_flags, node_prefix, sc->top, n_top);
    GEN_HIP(hipGetLastError());
    GEN_HIP(hipMalloc((void **)&sc->slots, std::max<uint64_t>(n_nodes, 1) * 4096 * 4));
    sc->bytes += n_top * 4 + n_nodes * 4096 * 4;
  } else {
    GEN_HIP(hipMalloc((void **)&sc->slots, n_brick_cells * 4));
    sc->bytes += n_brick_cells * 4;
  }
  GEN_HIP(hipMalloc((void **)&sc->masks, std::max<uint64_t>(n_bricks, 1) * 64));
  GEN_HIP(hipMalloc((void **)&sc->voxels, std::max<uint64_t>(n_bricks, 1) * 1024));
  sc->bytes += n_bricks * (64 + 1024);
  hipLaunchKernelGGL(brick_write_kernel, grid_1d(n_brick_cells), dim3(512), 0, stream,
                     d->generator, n, seed, heights, two ? 1 : 0, flags, prefix, sc->top, sc->slots,
                     sc->masks, sc->voxels, n_brick_cells);
  GEN_HIP(hipGetLastError());
  unsigned long long s = 0;
  GEN_HIP(hipMemcpyAsync(&s, solid, 8, hipMemcpyDeviceToHost, stream));
  GEN_HIP(hipStreamSynchronize(stream));
  if (heights) hipFreeAsync(heights, stream);
  hipFreeAsync(flags, stream);
  hipFreeAsync(prefix, stream);
  hipFreeAsync(solid, stream);
  if (node_flags) hipFreeAsync(node_flags, stream);
  if (node_prefix) hipFreeAsync(node_prefix, stream);
  GEN_HIP(hipStreamSynchronize(stream));
  sc->n_bricks = (int64_t)n_bricks;
  sc->n_nodes = (int64_t)n_nodes;
  sc->solid = (int64_t)s;
  MESO_CHECK(n_bricks < (1ull << 29), MESO_ERR_OUT_OF_MEMORY, "%llu bricks exceed 2^29",
             (unsigned long long)n_bricks);
  GEN_TRY(build_coarse(sc, stream));
  *out = sc;
  return MESO_OK;
}

meso_status meso_generate_chunks(meso_context *ctx, uint64_t seed, int32_t n_chunks,
                                 uint16_t *d_out, meso_stream stream) {
  MESO_CHECK(ctx && d_out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(n_chunks >= 1 && n_chunks <= (1 << 20), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d not in 1..2^20", n_chunks);
  // Per-chunk parameters from PCG32: Beta(1/2,1/2) fill as cos^2 of a uniform
  // angle (a point uniform in the unit disk by rejection), noise seed, spheres.
  std::vector<ChunkParamsDev> host((size_t)n_chunks);
  Pcg32 r;
  r.seed(seed);
  for (int32_t c = 0; c < n_chunks; c++) {
    float a, b, r2;
    for (;;) {
      a = r.u01() * 2.0f;
      a = a - 1.0f;
      b = r.u01() * 2.0f;
      b = b - 1.0f;
      float aa = a * a, bb = b * b;
      r2 = aa + bb;
      if (r2 < 1.0f && r2 > 0.0f) break;
    }
    float aa = a * a;
    float fill = aa / r2;
    ChunkParamsDev &p = host[(size_t)c];
    p.noise_seed = r.next();
    p.n_spheres = 1 + (int32_t)(r.next() % 4u);
    for (int i = 0; i < 4; i++) {
      p.sph[i][0] = (int32_t)(r.next() % (32u * 256u));
      p.sph[i][1] = (int32_t)(r.next() % (32u * 256u));
      p.sph[i][2] = (int32_t)(r.next() % (32u * 256u));
      p.sph[i][3] = (int32_t)(3 * 256 + r.next() % (9u * 256u + 1u));
    }
    float tf = fill * 32768.0f;
    tf = tf + 0.5f;
    p.target = (uint32_t)tf;
  }
  hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipSetDevice(ctx->device));
  ChunkParamsDev *dp = nullptr;
  MESO_HIP(hipMallocAsync((void **)&dp, host.size() * sizeof(ChunkParamsDev), s));
  MESO_HIP(hipMemcpyAsync(dp, host.data(), host.size() * sizeof(ChunkParamsDev),
                          hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(chunk_gen_kernel, dim3((uint32_t)n_chunks), dim3(1024), 0, s, dp, d_out);
  MESO_HIP(hipGetLastError());
  MESO_HIP(hipFreeAsync(dp, s));
  MESO_HIP(hipStreamSynchronize(s));
  return MESO_OK;
}

meso_status meso_generate_sphere_chunks(meso_context *ctx, const int32_t *d_chunk_xyz,
                                        int32_t n_chunks, int32_t res, float block_size,
                                        int32_t depth, uint8_t *d_mips, meso_stream stream) {
  MESO_CHECK(ctx && d_chunk_xyz && d_mips, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(n_chunks >= 1 && n_chunks <= (1 << 20) && res >= 1 && res <= 32 && depth >= 1 &&
                 depth <= 8,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "n_chunks %d, res %d (1..32), depth %d (1..8)",
             n_chunks, res, depth);
  hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipSetDevice(ctx->device));
  const size_t lds = 2u * (size_t)res * res * res;
  hipLaunchKernelGGL(sphere_chunk_kernel, dim3((uint32_t)n_chunks), dim3(256), lds, s,
                     d_chunk_xyz, res, (double)block_size, depth, d_mips);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

meso_status meso_scene_edit_chunks(meso_scene *scene, const int32_t *d_chunk_xyz, int32_t n_chunks,
                                   uint32_t seed, meso_stream stream) {
  MESO_CHECK(scene && d_chunk_xyz, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->kind != MESO_SCENE_DENSE_U8, MESO_ERR_NOT_SUPPORTED,
             "chunk edits need a brick map or two-level scene");
  MESO_CHECK(n_chunks >= 0 && n_chunks <= (1 << 24), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d out of range", n_chunks);
  if (n_chunks == 0) return MESO_OK;
  hipStream_t s = stream_of(scene->ctx, stream);
  MESO_HIP(hipSetDevice(scene->ctx->device));
  hipLaunchKernelGGL(edit_chunks_kernel, grid_1d((uint64_t)n_chunks), dim3(512), 0, s,
                     scene->dev(), scene->kind, d_chunk_xyz, seed, scene->masks, scene->blocks,
                     scene->voxels, (uint64_t)n_chunks);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

meso_status meso_scene_update_bricks(meso_scene *scene, const uint32_t *d_slots,
                                     const uint16_t *d_voxels, int32_t n_edit, meso_stream stream) {
  MESO_CHECK(scene && d_slots && d_voxels, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->kind != MESO_SCENE_DENSE_U8, MESO_ERR_NOT_SUPPORTED,
             "brick edits need a brick map or two-level scene");
  MESO_CHECK(n_edit >= 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "negative edit count");
  if (n_edit == 0) return MESO_OK;
  hipStream_t s = stream_of(scene->ctx, stream);
  MESO_HIP(hipSetDevice(scene->ctx->device));
  uint32_t *bad = nullptr;
  MESO_HIP(hipMallocAsync((void **)&bad, 16, s));
  meso_status st = launch_zero(bad, 16, s);
  if (st != MESO_OK) return st;
  hipLaunchKernelGGL(brick_update_kernel, grid_1d((uint64_t)n_edit), dim3(512), 0, s, d_slots,
                     d_voxels, scene->n_bricks, scene->masks, scene->blocks, scene->slot_cell,
                     scene->voxels, bad,
                     (uint64_t)n_edit);
  MESO_HIP(hipGetLastError());
  uint32_t nbad = 0;
  MESO_HIP(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
  MESO_HIP(hipFreeAsync(bad, s));
  MESO_HIP(hipStreamSynchronize(s));
  MESO_CHECK(nbad == 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "%u edits named a slot >= n_bricks", nbad);
  return MESO_OK;
}

meso_status meso_scene_update_bricks_async(meso_scene *scene, const uint32_t *d_slots,
                                           const uint16_t *d_voxels, int32_t n_edit,
                                           uint32_t *d_skipped, meso_stream stream) {
  MESO_CHECK(scene && d_slots && d_voxels, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->kind != MESO_SCENE_DENSE_U8, MESO_ERR_NOT_SUPPORTED,
             "brick edits need a brick map or two-level scene");
  MESO_CHECK(n_edit >= 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "negative edit count");
  if (n_edit == 0) return MESO_OK;
  hipStream_t s = stream_of(scene->ctx, stream);
  MESO_HIP(hipSetDevice(scene->ctx->device));
  hipLaunchKernelGGL(brick_update_kernel, grid_1d((uint64_t)n_edit), dim3(512), 0, s, d_slots,
                     d_voxels, scene->n_bricks, scene->masks, scene->blocks, scene->slot_cell,
                     scene->voxels, d_skipped,
                     (uint64_t)n_edit);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

}  // extern "C"
