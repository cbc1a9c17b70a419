// Chunk pool: the 32^3 chunks around a camera, ranked by importance and kept
// meshed in a fixed number of slots.  This is the device-side counterpart of
// the reference's chunk streaming (FChunkManage::UpdateChunks /
// UpdateLoadingQueue, Runtimes/Voxel/Chunk/ChunkManager.h:160-330;
// FChunkPool::PushToPool with EChunkOverrideMode::FindMin, ChunkPool.h:446-560;
// importance and the desired set, ChunkManagerHelper.h:21-148).  Where the
// reference ranks on the CPU and hands chunks to worker threads one at a
// time, one update here ranks every desired chunk and every slot in
// parallel, pairs the best candidates with the least important slots, and
// meshes the loaded chunks from the resident scene straight into their
// slots.  The host only launches; nothing returns to it.
//
// Float expressions follow the same one-operation-at-a-time order as the
// oracle (-ffp-contract=off), so the ranking is bit-identical to it.
#include <cmath>
#include <cstring>  // before rocprim, whose texture iterator calls memset
#include <vector>

#include <rocprim/rocprim.hpp>

#include "meso_internal.h"

namespace meso {

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
// counters[]: the state one update leaves for the next kernels and stats
enum : int {
  kCntLoads = 0,       // loads of the last update (set to the limit, lowered by the pairing)
  kCntCandidates = 1,  // desired, non-empty, non-resident chunks of the last update
  kCntEvicted = 2,     // resident chunks replaced by the last update
  kCntStamp = 3,       // updates so far
  kCntTicket = 4,      // two counters of the slotted mesher
  kCntRemesh = 6,      // resident dirty chunks listed by the last remesh
  kCntResident = 8,    // stats: slots holding a chunk
  kCntTruncated = 9,   // stats: resident chunks longer than a slot
  kCntQuadsLo = 10,    // stats: quads of the resident chunks (u64 as two words)
  kCntWords = 16
};

struct PoolGeom {
  int32_t R, RB;       // forward / backward radius, chunks
  float thr;           // view-cone threshold
  int32_t nc;          // chunks per scene axis
  int32_t side;        // 2R + 1
  int32_t capacity;
};

__device__ __forceinline__ float offset_len(int32_t dx, int32_t dy, int32_t dz) {
  const float x = (float)dx, y = (float)dy, z = (float)dz;
  float ss = x * x + y * y;
  ss = ss + z * z;
  return __builtin_sqrtf(ss);
}

// dot of the offset's direction with the unit forward vector f
__device__ __forceinline__ float offset_dir_dot(int32_t dx, int32_t dy, int32_t dz, float len,
                                                float f0, float f1, float f2) {
  const float ux = (float)dx / len, uy = (float)dy / len, uz = (float)dz / len;
  float d = ux * f0 + uy * f1;
  d = d + uz * f2;
  return d;
}

// FImportanceComputeInfo::CalculateChunkImportance (ChunkManagerHelper.h:21-43)
__device__ __forceinline__ float chunk_importance(int32_t dx, int32_t dy, int32_t dz, float f0,
                                                  float f1, float f2) {
  if (dx >= -2 && dx <= 2 && dy >= -2 && dy <= 2 && dz >= -2 && dz <= 2) return 1.0e6f;
  const float len = offset_len(dx, dy, dz);
  float a = offset_dir_dot(dx, dy, dz, len, f0, f1, f2);
  a = a > 0.0f ? a : 0.0f;
  a = a - 0.5f;
  a = a * 2.0f;
  a = a > 0.75f ? a : 0.75f;
  float di = 64.0f - len;
  di = di > 0.25f ? di : 0.25f;
  return a * di;
}

// the acceptance of GetDesiredShowChunkLocationByView (ChunkManagerHelper.h:98-135):
// inside the forward sphere, and the 3^3 core or nearer than a radius that
// blends from the forward radius (inside the view cone) to the backward one
__device__ __forceinline__ bool chunk_desired(int32_t dx, int32_t dy, int32_t dz, float f0,
                                              float f1, float f2, const PoolGeom &g) {
  const float len = offset_len(dx, dy, dz);
  if ((double)len > (double)g.R + 1e-6) return false;
  if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) return true;
  float a = offset_dir_dot(dx, dy, dz, len, f0, f1, f2);
  a = a > 0.0f ? a : 0.0f;
  a = a > g.thr ? 1.0f : a / g.thr;
  a = a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
  const float t1 = a * (float)g.R;
  const float b = 1.0f - a;
  const float t2 = b * (float)g.RB;
  const float thr = t1 + t2;
  return len < thr;
}

// the scene's 32^3 level: a chunk with no non-empty brick needs no slot
template <int KIND>
__device__ __forceinline__ bool chunk_nonempty(const DevScene &s, int32_t nc, int32_t cx,
                                               int32_t cy, int32_t cz) {
  const uint32_t x = (uint32_t)cx, y = (uint32_t)cy, z = (uint32_t)cz;
  if constexpr (KIND == 2) {
    const uint32_t i = x + (uint32_t)nc * (y + (uint32_t)nc * z);
    return (s.occ32[i >> 5] >> (i & 31u)) & 1u;
  } else {
    const uint32_t n128 = (uint32_t)s.n >> 7;
    const uint32_t node = s.top[(x >> 2) + n128 * ((y >> 2) + n128 * (z >> 2))];
    if (node == kNoSlot) return false;
    return (s.groups[node] >> ((x & 3u) + 4u * (y & 3u) + 16u * (z & 3u))) & 1ull;
  }
}

__global__ void pool_begin_kernel(uint32_t *counters, uint32_t limit) {
  counters[kCntLoads] = limit;
  counters[kCntCandidates] = 0u;
  counters[kCntEvicted] = 0u;
  counters[kCntStamp] = counters[kCntStamp] + 1u;
}

// Candidate keys: importance descending (complemented float bits;
// importances are positive), then chunk index.  With `compact` the
// candidates are appended to `keys` in any order (their number is the
// candidate counter); otherwise there is one key per offset of the
// (2R+1)^3 cube and offsets that are no candidate get the largest key, so
// that a full sort puts them last.
template <int KIND>
__global__ void __launch_bounds__(256) pool_rank_kernel(DevScene s, PoolGeom g, int32_t cx0,
                                                        int32_t cy0, int32_t cz0, float f0,
                                                        float f1, float f2,
                                                        const uint32_t *lookup, uint64_t *keys,
                                                        uint64_t cube, uint32_t *counters,
                                                        bool compact) {
  const uint64_t k = block_id() * 256u + threadIdx.x;
  bool cand = false;
  uint64_t key = ~0ull;
  if (k < cube) {
    const uint32_t side = (uint32_t)g.side;
    const int32_t dx = (int32_t)(k % side) - g.R, dy = (int32_t)((k / side) % side) - g.R,
                  dz = (int32_t)(k / ((uint64_t)side * side)) - g.R;
    const int32_t cx = cx0 + dx, cy = cy0 + dy, cz = cz0 + dz;
    if (cx >= 0 && cy >= 0 && cz >= 0 && cx < g.nc && cy < g.nc && cz < g.nc) {
      const uint32_t idx = (uint32_t)cx + (uint32_t)g.nc * ((uint32_t)cy + (uint32_t)g.nc * (uint32_t)cz);
      if (lookup[idx] == kNoSlot && chunk_desired(dx, dy, dz, f0, f1, f2, g) &&
          chunk_nonempty<KIND>(s, g.nc, cx, cy, cz)) {
        const float imp = chunk_importance(dx, dy, dz, f0, f1, f2);
        key = ((uint64_t)(~__float_as_uint(imp)) << 32) | idx;
        cand = true;
      }
    }
    if (!compact) keys[k] = key;
  }
  const uint64_t b = __ballot(cand);
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t base = 0u;
  if (lane == 0u && b) base = atomicAdd(&counters[kCntCandidates], (uint32_t)__popcll(b));
  if (compact) {
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
    if (cand) keys[base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))] = key;
  }
}

// Victim keys, one per slot: free slots first (by slot), then resident
// chunks by importance under this camera (ascending), then slot.
__global__ void __launch_bounds__(256) pool_victim_kernel(PoolGeom g, int32_t cx0, int32_t cy0,
                                                          int32_t cz0, float f0, float f1,
                                                          float f2, const uint32_t *slot_chunk,
                                                          uint64_t *keys) {
  const uint32_t sl = (uint32_t)(block_id() * 256u + threadIdx.x);
  if (sl >= (uint32_t)g.capacity) return;
  const uint32_t c = slot_chunk[sl];
  if (c == kNoSlot) {
    keys[sl] = sl;
    return;
  }
  const uint32_t nc = (uint32_t)g.nc;
  const int32_t x = (int32_t)(c % nc), y = (int32_t)((c / nc) % nc), z = (int32_t)(c / (nc * nc));
  const float imp = chunk_importance(x - cx0, y - cy0, z - cz0, f0, f1, f2);
  keys[sl] = ((uint64_t)__float_as_uint(imp) << 32) | sl;
}

// The first `limit` keys of a list in ascending order, for limits up to
// kSelMax: one workgroup per list, block 0 the compact candidates (their
// number is the candidate counter), block 1 the victims (one per slot).
// An MSB-first radix select over 8-bit digits finds t, the limit-th
// smallest key: per digit a histogram in LDS of the keys that share the
// digits found so far, and a scan by one wave; once at most kSelSmall keys
// share them, one wave ranks those keys directly.  The keys below t are
// gathered into LDS, the remaining places up to the limit take t, and a
// bitonic sort orders the kpad (power of two >= limit) places.  That is
// the first `limit` places of a full sort of the list, repeated keys
// included.  A list no longer than the limit is taken whole and padded
// with the largest key, which is what a full sort of the cube's keys puts
// after the candidates.  The first kSelPer * kSelThreads keys stay in
// registers across the passes; the rest are re-read from memory.
constexpr uint32_t kSelThreads = 1024u, kSelMax = 4096u, kSelPer = 16u, kSelSmall = 64u;
// Longest list the one-workgroup select takes: keys past the first
// kSelPer * kSelThreads are re-read on every radix pass, so past eight times
// that the device-wide radix sorts are faster.
constexpr uint64_t kSelListMax = 8ull * kSelPer * kSelThreads;

__global__ void __launch_bounds__(kSelThreads)
    pool_select_kernel(const uint64_t *cand, const uint64_t *vict, uint32_t cap, uint32_t limit,
                       uint32_t kpad, const uint32_t *counters, uint64_t *cand_out,
                       uint64_t *vict_out) {
  __shared__ uint64_t buf[kSelMax];
  __shared__ uint32_t hist[256];
  __shared__ uint32_t spread_hist[16 * 257];  // 16 copies, by lane, for the top two digits
  __shared__ uint32_t s_bin, s_below, s_cnt, s_fill;
  __shared__ uint64_t s_key;
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const bool victims = blockIdx.x != 0u;
  const uint64_t *keys = victims ? vict : cand;
  uint64_t *out = victims ? vict_out : cand_out;
  const uint32_t n = victims ? cap : counters[kCntCandidates];
  constexpr uint32_t kHeld = kSelPer * kSelThreads;
  uint64_t r[kSelPer];
#pragma unroll
  for (uint32_t j = 0; j < kSelPer; j++) {
    const uint32_t i = tid + j * kSelThreads;
    r[j] = i < n ? keys[i] : 0ull;
  }
  // f(key, valid) for every key of the list; each call is made by all lanes
  // of the workgroup together, so f may use wave-wide operations
  auto each = [&](auto &&f) {
#pragma unroll
    for (uint32_t j = 0; j < kSelPer; j++) f(r[j], tid + j * kSelThreads < n);
    for (uint32_t b = kHeld; b < n; b += kSelThreads) {
      const uint32_t i = b + tid;
      f(i < n ? keys[i] : 0ull, i < n);
    }
  };
  uint64_t t = ~0ull;     // the limit-th smallest key
  uint32_t rem = limit;   // its rank among the keys sharing the digits found
  uint32_t less = 0u;     // keys sharing them that are below t
  if (n > limit) {
    uint64_t prefix = 0ull, hi = 0ull;  // digits found, and their bits
    int shift = 56;
    bool small = false;
    for (;; shift -= 8) {
      // The top digits are shared by most keys of a wave; lanes that add to
      // one LDS word serialize, so those passes count into 16 copies
      // (lane & 15, stride 257 words) that are summed afterwards.
      const bool spread = shift >= 48;
      if (tid < 256u) hist[tid] = 0u;
      if (spread)
        for (uint32_t i = tid; i < 16u * 257u; i += kSelThreads) spread_hist[i] = 0u;
      __syncthreads();
      each([&](uint64_t key, bool valid) {
        const bool m = valid && (key & hi) == prefix;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        if (m) atomicAdd(spread ? &spread_hist[(lane & 15u) * 257u + d] : &hist[d], 1u);
      });
      __syncthreads();
      if (spread) {
        if (tid < 256u) {
          uint32_t acc = 0u;
#pragma unroll
          for (uint32_t c = 0; c < 16u; c++) acc += spread_hist[c * 257u + tid];
          hist[tid] = acc;
        }
        __syncthreads();
      }
      if (tid < 64u) {
        const uint32_t h0 = hist[4u * tid], h1 = hist[4u * tid + 1u], h2 = hist[4u * tid + 2u],
                       h3 = hist[4u * tid + 3u];
        const uint32_t c = h0 + h1 + h2 + h3;
        uint32_t inc = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t v = __shfl_up(inc, o);
          if ((int)tid >= o) inc += v;
        }
        // the bins hold at least rem keys in all, so some lane reaches rem
        const uint64_t over = __ballot(inc >= rem);
        const uint32_t at = over ? (uint32_t)__ffsll((long long)over) - 1u : 63u;
        if (tid == at) {
          uint32_t cum = inc - c, bin = 4u * tid, hb = h0;
          if (cum + h0 < rem) {
            cum += h0, bin++, hb = h1;
            if (cum + h1 < rem) {
              cum += h1, bin++, hb = h2;
              if (cum + h2 < rem) cum += h2, bin++, hb = h3;
            }
          }
          s_bin = bin;
          s_below = cum;
          s_cnt = hb;
        }
      }
      __syncthreads();
      prefix |= (uint64_t)s_bin << shift;
      hi |= 255ull << shift;
      rem -= s_below;
      if (s_cnt <= kSelSmall) {
        small = true;
        break;
      }
      if (shift == 0) break;
    }
    if (small) {
      // few keys hold the digits found: gather them, and one wave ranks
      // them (ties by place); the key of rank rem - 1 is t
      if (tid == 0u) s_fill = 0u;
      __syncthreads();
      each([&](uint64_t key, bool valid) {
        if (valid && (key & hi) == prefix) {
          const uint32_t pos = atomicAdd(&s_fill, 1u);
          if (pos < kSelSmall) buf[pos] = key;
        }
      });
      __syncthreads();
      const uint32_t cnt = s_cnt;
      if (tid < cnt) {
        const uint64_t k = buf[tid];
        uint32_t rank = 0u, lt = 0u;
        for (uint32_t j = 0; j < cnt; j++) {
          const uint64_t o = buf[j];
          lt += o < k ? 1u : 0u;
          rank += (o < k || (o == k && j < tid)) ? 1u : 0u;
        }
        if (rank == rem - 1u) {
          s_key = k;
          s_fill = lt;
        }
      }
      __syncthreads();
      t = s_key;
      less = s_fill;
      __syncthreads();  // s_fill is reset for the gather below
    } else {
      t = prefix;  // every key holding all eight digits equals t
    }
  }
  // keys below t: n when the list is taken whole, else the keys under the
  // digits' bin plus those of the small set below t
  const uint32_t below = n > limit ? limit - rem + less : n;
  if (tid == 0u) s_fill = 0u;
  for (uint32_t i = below + tid; i < kpad; i += kSelThreads) buf[i] = i < limit ? t : ~0ull;
  __syncthreads();
  each([&](uint64_t key, bool valid) {
    const bool m = valid && key < t;
    const uint64_t mb = __ballot(m);
    if (!mb) return;
    uint32_t base = 0u;
    if (lane == 0u) base = atomicAdd(&s_fill, (uint32_t)__popcll(mb));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
    const uint32_t pos = base + (uint32_t)__popcll(mb & ((1ull << lane) - 1ull));
    if (m && pos < below) buf[pos] = key;
  });
  // bitonic sort, two layers (strides s and s/2) per barrier where the
  // stage has them: a thread orders the four places i, i + s/2, i + s,
  // i + 3s/2 in registers
  for (uint32_t size = 2u; size <= kpad; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0u;) {
      __syncthreads();
      if (stride >= 2u) {
        const uint32_t h = stride >> 1, p1 = (uint32_t)__ffs((int)h) - 1u;
        for (uint32_t q = tid; q < kpad / 4u; q += kSelThreads) {
          const uint32_t i = ((q >> p1) << (p1 + 2u)) | (q & (h - 1u));
          const bool up = (i & size) == 0u;
          uint64_t a0 = buf[i], a1 = buf[i + h], a2 = buf[i + stride], a3 = buf[i + stride + h];
          auto order = [up](uint64_t &x, uint64_t &y) {
            if ((x > y) == up) {
              const uint64_t z = x;
              x = y;
              y = z;
            }
          };
          order(a0, a2);
          order(a1, a3);
          order(a0, a1);
          order(a2, a3);
          buf[i] = a0;
          buf[i + h] = a1;
          buf[i + stride] = a2;
          buf[i + stride + h] = a3;
        }
        stride >>= 2;
      } else {
        for (uint32_t q = tid; q < kpad / 2u; q += kSelThreads) {
          const uint32_t i = 2u * q;
          const uint64_t a = buf[i], b = buf[i + 1u];
          if ((a > b) == ((i & size) == 0u)) {
            buf[i] = b;
            buf[i + 1u] = a;
          }
        }
        stride = 0u;
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < limit; i += kSelThreads) out[i] = buf[i];
}

// Candidate i takes victim i when the victim is free or strictly less
// important.  Candidates descend and victims ascend in importance, so the
// pairs that qualify are a prefix; its length is the first failure.
__global__ void __launch_bounds__(256) pool_pair_kernel(const uint64_t *cand, const uint64_t *vict,
                                                        uint32_t limit, uint32_t *counters) {
  const uint32_t i = (uint32_t)(block_id() * 256u + threadIdx.x);
  if (i >= limit) return;
  const uint64_t ck = cand[i], vk = vict[i];
  bool ok = ck != ~0ull;
  if (ok && (uint32_t)(vk >> 32) != 0u) {
    const float ci = __uint_as_float(~(uint32_t)(ck >> 32));
    const float vi = __uint_as_float((uint32_t)(vk >> 32));
    ok = vi < ci;
  }
  if (!ok) atomicMin(&counters[kCntLoads], i);
}

// Applies the pairs: the victim's chunk leaves the lookup table, the
// candidate takes its slot, and the load list for the mesher is written.
// Old chunks (resident) and new ones (not resident) are disjoint sets and
// every pair has its own slot, so the writes never collide.
__global__ void __launch_bounds__(256) pool_apply_kernel(const uint64_t *cand, const uint64_t *vict,
                                                         uint32_t limit, int32_t nc,
                                                         uint32_t *lookup, uint32_t *slot_chunk,
                                                         uint32_t *slot_stamp, int32_t *load_xyz,
                                                         uint32_t *load_slot, uint32_t *counters) {
  const uint32_t i = (uint32_t)(block_id() * 256u + threadIdx.x);
  if (i >= limit || i >= counters[kCntLoads]) return;
  const uint32_t sl = (uint32_t)vict[i], c = (uint32_t)cand[i];
  const uint32_t old = slot_chunk[sl];
  if (old != kNoSlot) {
    lookup[old] = kNoSlot;
    atomicAdd(&counters[kCntEvicted], 1u);
  }
  lookup[c] = sl;
  slot_chunk[sl] = c;
  slot_stamp[sl] = counters[kCntStamp];
  const uint32_t n = (uint32_t)nc;
  load_xyz[3 * i] = (int32_t)(c % n);
  load_xyz[3 * i + 1] = (int32_t)((c / n) % n);
  load_xyz[3 * i + 2] = (int32_t)(c / (n * n));
  load_slot[i] = sl;
}

__global__ void pool_remesh_begin_kernel(uint32_t *counters) { counters[kCntRemesh] = 0u; }

// the resident chunks of a dirty list, with their slots
__global__ void __launch_bounds__(256) pool_dirty_kernel(const int32_t *xyz, int32_t n, int32_t nc,
                                                         const uint32_t *lookup, uint32_t room,
                                                         int32_t *out_xyz, uint32_t *out_slot,
                                                         uint32_t *counters) {
  const int32_t k = (int32_t)(block_id() * 256u + threadIdx.x);
  if (k >= n) return;
  const int32_t x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
  if (x < 0 || y < 0 || z < 0 || x >= nc || y >= nc || z >= nc) return;
  const uint32_t sl = lookup[(uint32_t)x + (uint32_t)nc * ((uint32_t)y + (uint32_t)nc * (uint32_t)z)];
  if (sl == kNoSlot) return;
  const uint32_t j = atomicAdd(&counters[kCntRemesh], 1u);
  if (j >= room) return;  // only repeated entries can overflow a list of `capacity`
  out_xyz[3 * j] = x;
  out_xyz[3 * j + 1] = y;
  out_xyz[3 * j + 2] = z;
  out_slot[j] = sl;
}

// stats over the slots
__global__ void __launch_bounds__(256) pool_stats_kernel(const uint32_t *slot_chunk,
                                                         const uint32_t *slot_count,
                                                         int32_t capacity, uint32_t per_slot,
                                                         uint32_t *counters) {
  const int32_t sl = (int32_t)(block_id() * 256u + threadIdx.x);
  uint32_t res = 0, trunc = 0;
  uint64_t q = 0;
  if (sl < capacity && slot_chunk[sl] != kNoSlot) {
    res = 1;
    q = slot_count[sl];
    trunc = slot_count[sl] > per_slot ? 1u : 0u;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    res += (uint32_t)__shfl_xor((int)res, off, 64);
    trunc += (uint32_t)__shfl_xor((int)trunc, off, 64);
    q += (uint64_t)__shfl_xor((long long)q, off, 64);
  }
  if ((threadIdx.x & 63u) == 0 && (res || q)) {
    atomicAdd(&counters[kCntResident], res);
    atomicAdd(&counters[kCntTruncated], trunc);
    atomicAdd((unsigned long long *)&counters[kCntQuadsLo], (unsigned long long)q);
  }
}

__global__ void pool_stats_begin_kernel(uint32_t *counters) {
  counters[kCntResident] = 0u;
  counters[kCntTruncated] = 0u;
  counters[kCntQuadsLo] = 0u;
  counters[kCntQuadsLo + 1] = 0u;
}

}  // namespace meso

using namespace meso;

struct meso_chunk_pool {
  meso_context *ctx = nullptr;
  meso_scene *scene = nullptr;
  meso_pool_desc desc{};
  PoolGeom geom{};
  uint64_t cube = 0;
  uint32_t limit = 0;  // pairs considered per update: min(max_loads, capacity, cube)
  uint32_t *lookup = nullptr, *slot_chunk = nullptr, *slot_stamp = nullptr,
           *slot_count = nullptr, *quads = nullptr, *counters = nullptr;
  uint64_t *cand = nullptr, *cand_sorted = nullptr, *vict = nullptr, *vict_sorted = nullptr;
  int32_t *load_xyz = nullptr, *remesh_xyz = nullptr;
  uint32_t *load_slot = nullptr, *remesh_slot = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  uint64_t bytes = 0;
  std::vector<float> dirs;  // baked view directions, 3 per direction (empty: exact)
};

// GetFibonacciSphere<float> (Runtimes/Helper/VoxelMathHelper.h:47-68) as
// BakeVisibilityByView instantiates it (ChunkManagerHelper.h:205): Phi in
// double stored as float, then float arithmetic throughout (cosf, sinf,
// sqrtf), and glm::normalize's v * (1 / sqrt(dot(v, v))).  Host code, so the
// directions are the same bits as any other host evaluation on this libm.
static void bake_directions(int32_t n, std::vector<float> &dirs) {
  dirs.resize((size_t)n * 3);
  const float phi = (float)(3.14159265358979323846 * (std::sqrt(5.0) - 1.0));
  for (int32_t i = 0; i < n; i++) {
    float y = (float)i / (float)(n - 1);
    y = y * 2.0f;
    y = 1.0f - y;
    const float yy = y * y;
    const float r = sqrtf(1.0f - yy);
    const float th = phi * (float)i;
    const float x = cosf(th) * r;
    const float z = sinf(th) * r;
    float dd = x * x + yy;
    dd = dd + z * z;
    const float inv = 1.0f / sqrtf(dd);
    dirs[(size_t)i * 3 + 0] = x * inv;
    dirs[(size_t)i * 3 + 1] = y * inv;
    dirs[(size_t)i * 3 + 2] = z * inv;
  }
}

static void pool_free(meso_chunk_pool *p) {
  void *ptrs[] = {p->lookup, p->slot_chunk, p->slot_stamp, p->slot_count, p->quads, p->counters,
                  p->cand, p->cand_sorted, p->vict, p->vict_sorted, p->load_xyz,
                  p->remesh_xyz, p->load_slot, p->remesh_slot, p->sort_tmp};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
}

static meso_status pool_alloc(meso_chunk_pool *p, void **ptr, uint64_t bytes) {
  if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    return fail(MESO_ERR_OUT_OF_MEMORY, "chunk pool: cannot allocate %llu bytes",
                (unsigned long long)bytes);
  }
  p->bytes += bytes;
  return MESO_OK;
}

namespace meso {
PoolView pool_view(const meso_chunk_pool *p) {
  return PoolView{p->scene, p->slot_chunk, p->slot_count, p->quads, p->geom.capacity,
                  p->desc.quads_per_slot, p->geom.nc};
}
}  // namespace meso

extern "C" {

meso_status meso_chunk_pool_create(meso_context *ctx, meso_scene *scene, const meso_pool_desc *desc,
                                   meso_chunk_pool **out) {
  MESO_CHECK(ctx && scene && desc && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  *out = nullptr;
  MESO_CHECK(desc->size >= sizeof(meso_pool_desc), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_pool_desc too small");
  MESO_CHECK(scene->ctx == ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "scene of another context");
  MESO_CHECK(scene->kind == MESO_SCENE_BRICKMAP_U16 || scene->kind == MESO_SCENE_TWO_LEVEL_U16,
             MESO_ERR_NOT_SUPPORTED, "the chunk pool needs a brick map or two-level scene");
  const meso_pool_desc &d = *desc;
  MESO_CHECK(d.capacity >= 1 && d.capacity <= (1 << 22), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "capacity %d not in 1..%d", d.capacity, 1 << 22);
  MESO_CHECK(d.quads_per_slot >= 1 && d.quads_per_slot <= 98304, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "quads_per_slot %d not in 1..98304", d.quads_per_slot);
  MESO_CHECK(d.forward_radius >= 1 && d.forward_radius <= 64, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "forward_radius %d not in 1..64", d.forward_radius);
  MESO_CHECK(d.backward_radius >= 0 && d.backward_radius <= d.forward_radius,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "backward_radius %d not in 0..%d",
             d.backward_radius, d.forward_radius);
  MESO_CHECK(std::isfinite(d.view_angle) && d.view_angle > 0.0f && d.view_angle <= 360.0f,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "view_angle must be in (0, 360] degrees");
  MESO_CHECK(d.max_loads >= 1 && d.max_loads <= d.capacity, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "max_loads %d not in 1..capacity", d.max_loads);
  MESO_CHECK(d.mesh_mode == MESO_MESH_CULLED || d.mesh_mode == MESO_MESH_GREEDY,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown mesh mode %u", d.mesh_mode);
  MESO_CHECK(d.bake_directions == 0 || (d.bake_directions >= 2 && d.bake_directions <= 65536),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "bake_directions %d not 0 or 2..65536",
             d.bake_directions);
  meso_chunk_pool *p = new meso_chunk_pool;
  if (d.bake_directions) bake_directions(d.bake_directions, p->dirs);
  p->ctx = ctx;
  p->scene = scene;
  p->desc = d;
  p->geom.R = d.forward_radius;
  p->geom.RB = d.backward_radius;
  // max(cos(radians(angle) / 2), 0.01) as ChunkManagerHelper.h:97
  float r = d.view_angle * 0.017453292f;
  r = r * 0.5f;
  const float c = cosf(r);
  p->geom.thr = c > 0.01f ? c : 0.01f;
  p->geom.nc = (scene->n + 31) / 32;
  p->geom.side = 2 * d.forward_radius + 1;
  p->geom.capacity = d.capacity;
  p->cube = (uint64_t)p->geom.side * p->geom.side * p->geom.side;
  uint64_t lim = (uint64_t)d.max_loads;
  if (lim > p->cube) lim = p->cube;
  p->limit = (uint32_t)lim;
  const uint64_t nc3 = (uint64_t)p->geom.nc * p->geom.nc * p->geom.nc;
  const uint64_t cap = (uint64_t)d.capacity;
  meso_status st = MESO_OK;
#define POOL_ALLOC(field, bytes)                                   \
  if (st == MESO_OK) st = pool_alloc(p, (void **)&p->field, bytes)
  POOL_ALLOC(lookup, nc3 * 4);
  POOL_ALLOC(slot_chunk, cap * 4);
  POOL_ALLOC(slot_stamp, cap * 4);
  POOL_ALLOC(slot_count, cap * 4);
  POOL_ALLOC(quads, cap * (uint64_t)d.quads_per_slot * 4);
  POOL_ALLOC(counters, kCntWords * 4);
  POOL_ALLOC(cand, p->cube * 8);
  POOL_ALLOC(cand_sorted, p->cube * 8);
  POOL_ALLOC(vict, cap * 8);
  POOL_ALLOC(vict_sorted, cap * 8);
  POOL_ALLOC(load_xyz, (uint64_t)d.max_loads * 12);
  POOL_ALLOC(load_slot, (uint64_t)d.max_loads * 4);
  POOL_ALLOC(remesh_xyz, cap * 12);
  POOL_ALLOC(remesh_slot, cap * 4);
#undef POOL_ALLOC
  if (st != MESO_OK) {
    pool_free(p);
    return st;
  }
  size_t b1 = 0, b2 = 0;
  if (rocprim::radix_sort_keys(nullptr, b1, p->cand, p->cand_sorted, (size_t)p->cube, 0, 64,
                               ctx->stream) != hipSuccess ||
      rocprim::radix_sort_keys(nullptr, b2, p->vict, p->vict_sorted, (size_t)cap, 0, 64,
                               ctx->stream) != hipSuccess) {
    pool_free(p);
    return fail(MESO_ERR_RUNTIME, "chunk pool: sort set-up failed");
  }
  p->sort_bytes = b1 > b2 ? b1 : b2;
  st = pool_alloc(p, &p->sort_tmp, p->sort_bytes);
  if (st != MESO_OK) {
    pool_free(p);
    return st;
  }
  if (hipMemsetAsync(p->lookup, 0xFF, nc3 * 4, ctx->stream) != hipSuccess ||
      hipMemsetAsync(p->slot_chunk, 0xFF, cap * 4, ctx->stream) != hipSuccess ||
      hipMemsetAsync(p->slot_stamp, 0, cap * 4, ctx->stream) != hipSuccess ||
      hipMemsetAsync(p->slot_count, 0, cap * 4, ctx->stream) != hipSuccess ||
      hipMemsetAsync(p->counters, 0, kCntWords * 4, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    (void)hipGetLastError();
    pool_free(p);
    return fail(MESO_ERR_RUNTIME, "chunk pool: initialisation failed");
  }
  *out = p;
  return MESO_OK;
}

meso_status meso_chunk_pool_destroy(meso_chunk_pool *pool) {
  if (!pool) return MESO_OK;
  (void)hipDeviceSynchronize();
  pool_free(pool);
  return MESO_OK;
}

meso_status meso_chunk_pool_update(meso_chunk_pool *pool, const int32_t camera_chunk[3],
                                   const float forward[3], meso_stream stream) {
  MESO_CHECK(pool && camera_chunk && forward, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  for (int a = 0; a < 3; a++) {
    MESO_CHECK(std::isfinite(forward[a]), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "forward vector not finite");
    MESO_CHECK(camera_chunk[a] >= -(1 << 24) && camera_chunk[a] <= (1 << 24),
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "camera chunk coordinate %d out of range",
               camera_chunk[a]);
  }
  float ss = forward[0] * forward[0] + forward[1] * forward[1];
  ss = ss + forward[2] * forward[2];
  MESO_CHECK(ss > 0.0f && std::isfinite(ss), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "forward vector has no direction");
  const float len = sqrtf(ss);
  float f0 = forward[0] / len, f1 = forward[1] / len, f2 = forward[2] / len;
  if (!pool->dirs.empty()) {
    // the nearest baked direction (TNearestMap::Query, NearestMap.h:36-50;
    // the first one on ties)
    const size_t nd = pool->dirs.size() / 3;
    size_t best = 0;
    float bd = INFINITY;
    for (size_t i = 0; i < nd; i++) {
      const float dx = pool->dirs[3 * i] - f0, dy = pool->dirs[3 * i + 1] - f1,
                  dz = pool->dirs[3 * i + 2] - f2;
      float dd = dx * dx + dy * dy;
      dd = dd + dz * dz;
      if (dd < bd) {
        bd = dd;
        best = i;
      }
    }
    f0 = pool->dirs[3 * best];
    f1 = pool->dirs[3 * best + 1];
    f2 = pool->dirs[3 * best + 2];
  }
  const hipStream_t s = stream_of(pool->ctx, stream);
  const PoolGeom &g = pool->geom;
  const int32_t cx = camera_chunk[0], cy = camera_chunk[1], cz = camera_chunk[2];
  hipLaunchKernelGGL(pool_begin_kernel, dim3(1), dim3(1), 0, s, pool->counters, pool->limit);
  MESO_HIP(hipGetLastError());
  const DevScene ds = pool->scene->dev();
  // limits up to kSelMax select the first keys of lists up to kSelListMax
  // keys (candidates: at most one per cube offset; victims: one per slot);
  // larger limits or longer lists sort the lists
  const bool select = pool->limit <= kSelMax && (uint64_t)pool->cube <= kSelListMax &&
                      (uint64_t)pool->geom.capacity <= kSelListMax;
  if (pool->scene->kind == MESO_SCENE_BRICKMAP_U16)
    hipLaunchKernelGGL(pool_rank_kernel<2>, grid_1d((pool->cube + 255) / 256), dim3(256), 0, s,
                       ds, g, cx, cy, cz, f0, f1, f2, pool->lookup, pool->cand, pool->cube,
                       pool->counters, select);
  else
    hipLaunchKernelGGL(pool_rank_kernel<3>, grid_1d((pool->cube + 255) / 256), dim3(256), 0, s,
                       ds, g, cx, cy, cz, f0, f1, f2, pool->lookup, pool->cand, pool->cube,
                       pool->counters, select);
  MESO_HIP(hipGetLastError());
  const uint64_t cap = (uint64_t)g.capacity;
  hipLaunchKernelGGL(pool_victim_kernel, grid_1d((cap + 255) / 256), dim3(256), 0, s, g, cx, cy,
                     cz, f0, f1, f2, pool->slot_chunk, pool->vict);
  MESO_HIP(hipGetLastError());
  if (select) {
    uint32_t kpad = 1u;
    while (kpad < pool->limit) kpad <<= 1;
    hipLaunchKernelGGL(pool_select_kernel, dim3(2), dim3(kSelThreads), 0, s, pool->cand,
                       pool->vict, (uint32_t)cap, pool->limit, kpad, pool->counters,
                       pool->cand_sorted, pool->vict_sorted);
    MESO_HIP(hipGetLastError());
  } else {
    size_t b = pool->sort_bytes;
    MESO_HIP(rocprim::radix_sort_keys(pool->sort_tmp, b, pool->cand, pool->cand_sorted,
                                      (size_t)pool->cube, 0, 64, s));
    b = pool->sort_bytes;
    MESO_HIP(rocprim::radix_sort_keys(pool->sort_tmp, b, pool->vict, pool->vict_sorted,
                                      (size_t)cap, 0, 64, s));
  }
  const uint64_t lim = pool->limit;
  hipLaunchKernelGGL(pool_pair_kernel, grid_1d((lim + 255) / 256), dim3(256), 0, s,
                     pool->cand_sorted, pool->vict_sorted, pool->limit, pool->counters);
  MESO_HIP(hipGetLastError());
  hipLaunchKernelGGL(pool_apply_kernel, grid_1d((lim + 255) / 256), dim3(256), 0, s,
                     pool->cand_sorted, pool->vict_sorted, pool->limit, g.nc, pool->lookup,
                     pool->slot_chunk, pool->slot_stamp, pool->load_xyz, pool->load_slot,
                     pool->counters);
  MESO_HIP(hipGetLastError());
  return launch_mesh_slotted(pool->ctx, pool->scene, pool->load_xyz, pool->load_slot,
                             pool->counters + kCntLoads, (int32_t)pool->limit,
                             pool->desc.mesh_mode, pool->slot_count, pool->quads,
                             (uint32_t)pool->desc.quads_per_slot,
                             (unsigned int *)pool->counters + kCntTicket, s);
}

meso_status meso_chunk_pool_remesh(meso_chunk_pool *pool, const int32_t *d_chunk_xyz,
                                   int32_t n_chunks, meso_stream stream) {
  MESO_CHECK(pool && d_chunk_xyz, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(n_chunks >= 0 && n_chunks <= (1 << 24), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d not in 0..%d", n_chunks, 1 << 24);
  if (n_chunks == 0) return MESO_OK;
  const hipStream_t s = stream_of(pool->ctx, stream);
  hipLaunchKernelGGL(pool_remesh_begin_kernel, dim3(1), dim3(1), 0, s, pool->counters);
  MESO_HIP(hipGetLastError());
  hipLaunchKernelGGL(pool_dirty_kernel, grid_1d(((uint64_t)n_chunks + 255) / 256), dim3(256), 0,
                     s, d_chunk_xyz, n_chunks, pool->geom.nc, pool->lookup,
                     (uint32_t)pool->geom.capacity, pool->remesh_xyz, pool->remesh_slot,
                     pool->counters);
  MESO_HIP(hipGetLastError());
  return launch_mesh_slotted(pool->ctx, pool->scene, pool->remesh_xyz, pool->remesh_slot,
                             pool->counters + kCntRemesh, pool->geom.capacity,
                             pool->desc.mesh_mode, pool->slot_count, pool->quads,
                             (uint32_t)pool->desc.quads_per_slot,
                             (unsigned int *)pool->counters + kCntTicket, s);
}

meso_status meso_chunk_pool_stats(meso_chunk_pool *pool, meso_pool_stats *out, meso_stream stream) {
  MESO_CHECK(pool && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(out->size >= sizeof(meso_pool_stats), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_pool_stats too small");
  const hipStream_t s = stream_of(pool->ctx, stream);
  hipLaunchKernelGGL(pool_stats_begin_kernel, dim3(1), dim3(1), 0, s, pool->counters);
  MESO_HIP(hipGetLastError());
  const uint64_t cap = (uint64_t)pool->geom.capacity;
  hipLaunchKernelGGL(pool_stats_kernel, grid_1d((cap + 255) / 256), dim3(256), 0, s,
                     pool->slot_chunk, pool->slot_count, pool->geom.capacity,
                     (uint32_t)pool->desc.quads_per_slot, pool->counters);
  MESO_HIP(hipGetLastError());
  uint32_t c[kCntWords];
  MESO_HIP(hipMemcpyAsync(c, pool->counters, sizeof(c), hipMemcpyDeviceToHost, s));
  MESO_HIP(hipStreamSynchronize(s));
  const uint32_t sz = out->size;
  std::memset(out, 0, sizeof(meso_pool_stats));
  out->size = sz;
  out->stamp = c[kCntStamp];
  out->n_candidates = c[kCntCandidates];
  out->n_loads = c[kCntLoads];
  out->n_evicted = c[kCntEvicted];
  out->n_resident = c[kCntResident];
  out->n_truncated = c[kCntTruncated];
  out->n_quads = (uint64_t)c[kCntQuadsLo] | ((uint64_t)c[kCntQuadsLo + 1] << 32);
  out->device_bytes = pool->bytes;
  return MESO_OK;
}

meso_status meso_chunk_pool_arrays(meso_chunk_pool *pool, meso_pool_arrays *out) {
  MESO_CHECK(pool && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(out->size >= sizeof(meso_pool_arrays), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_pool_arrays too small");
  out->lookup = pool->lookup;
  out->slot_chunk = pool->slot_chunk;
  out->slot_stamp = pool->slot_stamp;
  out->slot_count = pool->slot_count;
  out->quads = pool->quads;
  out->load_xyz = pool->load_xyz;
  out->load_slot = pool->load_slot;
  out->chunks_per_axis = pool->geom.nc;
  return MESO_OK;
}

}  // extern "C"
