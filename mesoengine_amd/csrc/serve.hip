// Serving: continuous batching of raymarch requests over HBM-resident scenes
// in one persistent kernel fed from a device ring.
//
// Host side (serve_core.h): requests are admitted in arrival order while
// their outputs fit an HBM arena of hbm_budget_bytes and a ring entry is
// free; an admitted request is published as the next ring entry.
//
// Device side.  An entry is cut into 8x8-pixel packets (one wave pass each,
// row-major).  Device-scope atomics and loads are performed at the memory
// side, a few microseconds each under load, so the kernel schedules in two
// levels and touches device-wide words about once per kChunk packets:
//  - Workgroups of four waves claim chunks of up to kChunk consecutive
//    packets from eight heads, one per XCD.  Head x hands out the
//    contiguous band [u x / 8, u (x + 1) / 8) of an entry's u units (its
//    packets, or 8 when it has fewer: units past the packets hold no work),
//    so an XCD's workgroups sweep neighbouring packets.  A head packs the
//    entry it draws from (22 bits), its band's unit count (21 bits) and the
//    next unit (21 bits): an atomic add of kChunk claims.  A workgroup claims
//    from its own XCD's head with the most units left, then from the others.
//  - Inside a workgroup the waves take packets by tickets from an LDS
//    counter, with no barrier.  Chunks cover consecutive tickets; the wave
//    whose ticket is the first past the newest chunk claims the next one
//    (the others wait in LDS), so no ticket is lost.  Installed chunks live
//    in an LDS ring of kSlots; a slot is reused once its chunk is finished.
//  - The wave that finishes a chunk's last packet publishes it: the waves'
//    stores are ordered before it by a workgroup release/acquire through
//    the LDS count, one system release fence writes them back, and a relaxed
//    add moves the entry's counter (one line per ring slot).  The add that
//    completes an entry resets the counter and sets the slot's completion
//    flag in host memory, which the scheduler thread polls.
//  - Every entry has at least one unit on every head and completes only when
//    all units are finished, so every head passes an entry before it
//    completes and its ring slot can be reused: a head, and a workgroup
//    holding a chunk, only read descriptors of open entries.
//  - One publisher wave (wave 0 of workgroup 0) is the only reader of host
//    memory: it copies newly published descriptors into a device ring,
//    announces them in a line per XCD, moves exhausted heads on to the next
//    entry, passes on the host's stop request, and after idle_exit_us with
//    every head exhausted and nothing new published marks the launch as
//    ending; the scheduler launches the kernel again when work is
//    outstanding.  Workgroups that find nothing to claim poll only their
//    XCD's line.
// Every pixel is computed by the same per-pixel code whatever else shares the
// kernel, so a result is identical alone, in a mix and on the plain path.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>

#include "serve_core.h"
#include "traverse.h"

namespace meso {

constexpr int kServePacket = 8;          // a wave's 8x8 pixels
constexpr uint32_t kServeWaves = 4;      // waves of a workgroup
constexpr uint32_t kServeThreads = 64 * kServeWaves;
constexpr uint32_t kServeHeads = 8;      // one per XCD: workgroup i runs on XCD i % 8
constexpr uint32_t kChunk = 32;          // packets a workgroup claims at once
constexpr uint32_t kSlots = 4;           // installed chunks a workgroup holds
constexpr uint32_t kRescanUs = 200;      // an idle workgroup looks at the heads this often
constexpr uint32_t kCntStride = 64;      // completion counters one 256-byte line apart
// head word: entry (22 bits) | unit count (21 bits) | cursor (21 bits)
constexpr uint64_t kFieldMask = (1ull << 21) - 1;
constexpr uint64_t kEntryMask = (1ull << 22) - 1;
constexpr uint32_t kMaxRing = 4096;
// an entry's packets, so that a band's units (and a cursor past them by a
// claim from every workgroup) fit the 21-bit fields
constexpr uint32_t kMaxPackets = 1u << 20;

__host__ __device__ inline uint64_t head_word(uint64_t entry, uint64_t units) {
  return ((entry & kEntryMask) << 42) | ((units & kFieldMask) << 21);
}
// first unit of head x's band of an entry with `units` units
__device__ __forceinline__ uint32_t band_start(uint32_t units, uint32_t x) {
  return (uint32_t)(((uint64_t)units * x) / kServeHeads);
}
// units in head x's band: at least one, as units >= 8
__device__ __forceinline__ uint32_t band_units(uint32_t units, uint32_t x) {
  return band_start(units, x + 1) - band_start(units, x);
}

// One ring entry, in host-coherent memory (written by the host); the
// publisher copies it into the device ring, which the workgroups read.
struct alignas(8) ServeDesc {
  uint64_t seq;  // entry number (from 1)
  uint32_t packets, packets_x;
  uint32_t units;  // max(packets, 8)
  int32_t kind;
  uint32_t flags;
  int32_t index_bytes;
  void *index;
  uint8_t *face;
  float *t;
  float *shade;
  DevCamera cam;
  DevShading sh;
  DevScene scene;
};
static_assert(sizeof(ServeDesc) % 8 == 0, "descriptor words are copied 8 bytes at a time");
constexpr uint32_t kDescWords = sizeof(ServeDesc) / 8;
static_assert(kDescWords <= 64, "one lane per descriptor word");

// host-coherent control words
struct ServeCtl {
  uint64_t published;  // last entry published (0: none)
  uint32_t stop;       // the host asks every wave to return
  uint32_t pad;
};

// device-memory state (zeroed once when the server starts); the completion
// counters (kCntStride words apart) and the device descriptor ring follow.
// Each head and each XCD's line is 256 bytes of its own.
struct alignas(256) ServeHead {
  unsigned long long word;  // entry 0 is a placeholder with no units
};
struct alignas(256) ServeXcd {
  unsigned long long published;  // entries whose descriptors are in the device ring
  unsigned int exiting;          // launch number whose waves are returning
};
struct ServeState {
  ServeHead heads[kServeHeads];
  ServeXcd xcd[kServeHeads];
};
static_assert(sizeof(ServeState) % 256 == 0, "the counters start on a line of their own");

__device__ __forceinline__ uint64_t sys_load64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t sys_load32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ unsigned long long dev_load64(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t dev_load32(const unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// lane r's value, in every lane
__device__ __forceinline__ uint32_t lane_at(uint32_t v, uint32_t r) {
  return __builtin_amdgcn_readlane(v, r);
}
__device__ __forceinline__ uint64_t lane_at(uint64_t v, uint32_t r) {
  return (uint64_t)lane_at((uint32_t)v, r) | ((uint64_t)lane_at((uint32_t)(v >> 32), r) << 32);
}
__device__ __forceinline__ uint32_t lane0(uint32_t v) { return lane_at(v, 0); }
__device__ __forceinline__ uint64_t lane0(uint64_t v) { return lane_at(v, 0); }

// a copy of an LDS struct whose every 32-bit word is wave-uniform
template <class T>
__device__ __forceinline__ T uniform_copy(const T &lds) {
  static_assert(sizeof(T) % 4 == 0, "");
  T out;
  const uint32_t *s = (const uint32_t *)&lds;
  uint32_t *d = (uint32_t *)&out;
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; i++) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
  return out;
}

template <int KIND>
__device__ __forceinline__ void serve_pixel(const ServeDesc &d, const DevScene &scene,
                                            const DevCamera &cam, int32_t px, int32_t py) {
  StepCounts cnt = {0, 0, 0, 0, 0, 0, 0};
  uint32_t ns = 0, na = 0;
  const DevShading sh = uniform_copy(d.sh);
  const uint32_t flags = __builtin_amdgcn_readfirstlane(d.flags);
  const PixelResult res = shade_pixel<KIND, false>(scene, cam, flags, sh, px, py, cnt, ns, na);
  const int64_t pix = (int64_t)py * cam.width + px;
  if (d.index_bytes == 8)
    ((int64_t *)d.index)[pix] = res.index;
  else
    ((int32_t *)d.index)[pix] = (int32_t)res.index;
  d.face[pix] = (uint8_t)res.face;
  d.t[pix] = res.t;
  d.shade[pix] = res.shade;
}

// lane 0 performs f() (an atomic, or a load of a word every lane would read
// alike); every lane gets its value.  The region is one access with no loop,
// entered and left by the whole wave, so the scheduler's control flow stays
// wave-uniform: a loop around a lane-0 region lets the compiler iterate the
// other lanes past it on their own.
template <class F>
__device__ __forceinline__ uint64_t by_lane0_64(uint32_t lane, F f) {
  uint64_t v = 0;
  if (lane == 0) v = f();
  return lane0(v);
}
template <class F>
__device__ __forceinline__ uint32_t by_lane0_32(uint32_t lane, F f) {
  uint32_t v = 0;
  if (lane == 0) v = f();
  return lane0(v);
}

// a head word (read, or returned by an add): units left
__device__ __forceinline__ bool claimable(uint64_t c) {
  return (c & kFieldMask) < ((c >> 21) & kFieldMask);
}
__device__ __forceinline__ uint32_t units_left(uint64_t c) {
  return claimable(c) ? (uint32_t)(((c >> 21) & kFieldMask) - (c & kFieldMask)) : 0u;
}
// the largest v over the wave
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Move head x (read as h, exhausted) on to the entry after its own when that
// one is published (pub): it is open, as this head has not passed it, so its
// descriptor is stable.  Called by single lanes.
__device__ __forceinline__ void advance_head(ServeState *st, const ServeDesc *ddesc,
                                             uint32_t ring_mask, uint32_t x, uint64_t h,
                                             uint64_t pub) {
  const uint64_t ahead = (pub - (h >> 42)) & kEntryMask;
  if (ahead < 1 || ahead >= (1ull << 21)) return;
  const uint64_t n = ((h >> 42) + 1) & kEntryMask;
  const uint32_t units = __hip_atomic_load(&ddesc[n & ring_mask].units, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  atomicCAS(&st->heads[x].word, (unsigned long long)h,
            (unsigned long long)head_word(n, band_units(units, x)));
}

// An installed chunk: tickets [lo, hi) (packed in `span`, written last) are
// units [u0, u0 + hi - lo) of ring entry `entry`, whose descriptor is
// desc[slot]; `done` counts its finished tickets.
struct ChunkSlot {
  uint64_t span;  // lo | hi << 32
  uint64_t entry;
  uint32_t u0, done;
};

// The publisher wave (see the header).
__device__ void serve_publisher(const ServeDesc *desc, ServeDesc *ddesc, const ServeCtl *ctl,
                                ServeState *st, uint32_t ring_mask, uint32_t gen,
                                uint64_t idle_ticks) {
  const uint32_t lane = threadIdx.x & 63u;
  uint64_t copied = by_lane0_64(lane, [&] { return (uint64_t)dev_load64(&st->xcd[0].published); });
  uint64_t last_busy = wall_clock64();
  for (;;) {
    const uint32_t stop = by_lane0_32(lane, [&] { return sys_load32(&ctl->stop); });
    const uint64_t hp = by_lane0_64(lane, [&] { return sys_load64(&ctl->published); });
    if (hp > copied) {
      for (uint64_t e = copied + 1; e <= hp; e++) {
        // entry e's slot held an entry that has completed, which no wave reads
        const uint64_t *src = (const uint64_t *)&desc[e & ring_mask];
        uint64_t *dst = (uint64_t *)&ddesc[e & ring_mask];
        if (lane < kDescWords) dst[lane] = sys_load64(src + lane);
      }
      // the copies are written back before the counts that announce them move
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      copied = hp;
      if (lane < kServeHeads) atomicMax(&st->xcd[lane].published, (unsigned long long)hp);
      last_busy = wall_clock64();
    }
    // heads: move exhausted ones on; any with units left means work
    uint64_t h = 0;
    if (lane < kServeHeads) h = dev_load64(&st->heads[lane].word);
    if (lane < kServeHeads && !claimable(h)) advance_head(st, ddesc, ring_mask, lane, h, copied);
    const uint64_t now = wall_clock64();
    if (__ballot(lane < kServeHeads && claimable(h))) last_busy = now;
    if (stop || now - last_busy > idle_ticks) {
      if (lane < kServeHeads) atomicMax(&st->xcd[lane].exiting, stop ? 0xFFFFFFFFu : gen);
      return;
    }
    __builtin_amdgcn_s_sleep(16);
  }
}

__global__ void __launch_bounds__(kServeThreads)
    serve_persistent_kernel(const ServeDesc *desc, ServeDesc *ddesc, const ServeCtl *ctl,
                            uint64_t *done_flag, ServeState *st, uint32_t *done_cnt,
                            uint32_t ring_mask, uint32_t gen, uint64_t idle_ticks,
                            uint64_t rescan_ticks) {
  __shared__ ServeDesc cdesc[kSlots];  // descriptors of the installed chunks' entries
  __shared__ ChunkSlot ch[kSlots];
  __shared__ uint32_t next_ticket, installed, n_installed, leave;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t home = blockIdx.x % kServeHeads;
  ServeXcd *const mine = &st->xcd[home];
  if (threadIdx.x < kSlots) {
    ch[threadIdx.x].span = 0;
    ch[threadIdx.x].entry = ~0ull;
    ch[threadIdx.x].u0 = 0;
    ch[threadIdx.x].done = 0;
  }
  if (threadIdx.x == 0) {
    next_ticket = 0;
    installed = 0;
    n_installed = 0;
    leave = 0;
  }
  __syncthreads();  // the only barrier: every wave runs on its own from here
  if (blockIdx.x == 0 && wave == 0) {
    serve_publisher(desc, ddesc, ctl, st, ring_mask, gen, idle_ticks);
    return;
  }
  // Claim the next chunk from the heads for ticket t, the first past the
  // newest chunk, and install it; false when the launch ends.
  auto refill = [&](uint32_t t) -> bool {
    const uint32_t k = lane0(n_installed) % kSlots;
    // the slot's previous chunk must be finished (its tickets are all
    // taken, and their waves finish without waiting for anything)
    for (;;) {
      const uint64_t span = lane0(ch[k].span);
      const uint32_t n = (uint32_t)(span >> 32) - (uint32_t)span;
      if (lane0(ch[k].done) == n) break;
      __builtin_amdgcn_s_sleep(1);
    }
    uint64_t c = 0;
    uint32_t x = 0;
    uint32_t nap = 1;
    bool wide = false;
    for (;;) {
      // own head (lane 0), the XCD's ... then every head, one per lane
      uint64_t hl = 0;
      const uint32_t hx = (home + lane) % kServeHeads;
      if (lane < (wide ? kServeHeads : 1u)) hl = dev_load64(&st->heads[hx].word);
      const uint32_t left = units_left(hl);
      const uint32_t most = wave_max(left);
      if (most) {
        const uint64_t at = __ballot(left == most);
        x = lane_at(hx, (uint32_t)__ffsll((long long)at) - 1);
        c = by_lane0_64(lane, [&] {
          return (uint64_t)atomicAdd(&st->heads[x].word, (unsigned long long)kChunk);
        });
        if (claimable(c)) break;
        wide = false;
        continue;  // that head ran out meanwhile: look again
      }
      if (!wide) {
        wide = true;
        continue;
      }
      // every head exhausted: move on the first one behind the published
      // count (the publisher moves them on too)
      const uint64_t pub =
          by_lane0_64(lane, [&] { return (uint64_t)dev_load64(&mine->published); });
      const uint64_t ahead = (pub - (hl >> 42)) & kEntryMask;
      const uint64_t behind = __ballot(lane < kServeHeads && ahead >= 1 && ahead < (1ull << 21));
      if (behind) {
        const uint32_t r = (uint32_t)__ffsll((long long)behind) - 1;
        const uint32_t bx = lane_at(hx, r);
        const uint64_t bh = lane_at(hl, r);
        if (lane == 0) advance_head(st, ddesc, ring_mask, bx, bh, pub);
        wide = false;
        continue;
      }
      // nothing to claim: wait for the XCD's published count to move (or
      // kRescanUs), or for the launch to end
      const uint64_t since = wall_clock64();
      for (;;) {
        for (uint32_t j = 0; j < nap; j++) __builtin_amdgcn_s_sleep(32);
        nap = nap < 64 ? nap * 2 : 64;
        if (by_lane0_32(lane, [&] { return dev_load32(&mine->exiting); }) >= gen) {
          if (lane == 0) leave = 1;
          return false;
        }
        const uint64_t p =
            by_lane0_64(lane, [&] { return (uint64_t)dev_load64(&mine->published); });
        if (p != pub || wall_clock64() - since > rescan_ticks) break;
      }
      wide = false;
    }
    const uint64_t ent = c >> 42;
    const uint32_t cur = (uint32_t)(c & kFieldMask), count = (uint32_t)((c >> 21) & kFieldMask);
    const uint32_t n = count - cur < kChunk ? count - cur : kChunk;
    if (cur + n == count) {
      // this claim took the band's last units: move the head on at once
      const uint64_t pub =
          by_lane0_64(lane, [&] { return (uint64_t)dev_load64(&mine->published); });
      if (lane == 0) advance_head(st, ddesc, ring_mask, x, c + kChunk, pub);
    }
    if (lane0(ch[k].entry) != ent) {
      // the chunk holds the entry open, so its descriptor is stable
      const unsigned long long *src = (const unsigned long long *)&ddesc[ent & ring_mask];
      if (lane < kDescWords) ((uint64_t *)&cdesc[k])[lane] = dev_load64(src + lane);
    }
    const uint32_t units = __builtin_amdgcn_readfirstlane(cdesc[k].units);
    if (lane == 0) {
      ch[k].entry = ent;
      ch[k].u0 = band_start(units, x) + cur;
      ch[k].done = 0;
      ch[k].span = (uint64_t)t | ((uint64_t)(t + n) << 32);  // last: it validates the slot
      installed = t + n;
      n_installed = n_installed + 1;
    }
    __builtin_amdgcn_wave_barrier();
    return true;
  };
  for (;;) {
    const uint32_t t = by_lane0_32(lane, [&] { return atomicAdd(&next_ticket, 1u); });
    // the chunk holding ticket t; the slot search reads each span whole, so
    // a slot being refilled never matches a ticket of another chunk
    uint32_t k = kSlots;
    for (;;) {
      for (uint32_t j = 0; j < kSlots; j++) {
        const uint64_t span = lane0(ch[j].span);
        if (t - (uint32_t)span < (uint32_t)(span >> 32) - (uint32_t)span) {
          k = j;
          break;
        }
      }
      if (k < kSlots) break;
      if (lane0(leave)) return;
      if (t == lane0(installed)) {
        if (!refill(t)) return;
        continue;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    // this ticket's packet
    const ServeDesc &d = cdesc[k];
    const uint64_t span = lane0(ch[k].span);
    const uint32_t u = lane0(ch[k].u0) + (t - (uint32_t)span);
    const uint32_t n = (uint32_t)(span >> 32) - (uint32_t)span;
    // what publishing the chunk needs, read before this wave's count: once
    // the count is in, the slot may be refilled
    const uint64_t seq = lane0((uint64_t)d.seq);
    const uint32_t total = __builtin_amdgcn_readfirstlane(d.units);
    const uint32_t packets = __builtin_amdgcn_readfirstlane(d.packets);
    const uint32_t slot = (uint32_t)seq & ring_mask;
    if (u < packets) {
      const uint32_t packets_x = __builtin_amdgcn_readfirstlane(d.packets_x);
      const DevCamera cam = uniform_copy(d.cam);
      const DevScene scene = uniform_copy(d.scene);
      const int kind = (int)__builtin_amdgcn_readfirstlane(d.kind);
      const int32_t px = (int32_t)((u % packets_x) * kServePacket + (lane & 7u));
      const int32_t py = (int32_t)((u / packets_x) * kServePacket + (lane >> 3));
      if (px < cam.width && py < cam.height) {
        switch (kind) {
          case MESO_SCENE_DENSE_U8: serve_pixel<1>(d, scene, cam, px, py); break;
          case MESO_SCENE_BRICKMAP_U16: serve_pixel<2>(d, scene, cam, px, py); break;
          default: serve_pixel<3>(d, scene, cam, px, py); break;
        }
      }
    }
    // count the ticket: this wave's stores are ordered before the count
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    const uint32_t finished = by_lane0_32(lane, [&] { return atomicAdd(&ch[k].done, 1u); }) + 1;
    if (finished == n) {
      // the chunk's last ticket: every wave's stores for it are ordered
      // before this point; write them back, then move the entry's counter
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
      uint32_t *const cnt = &done_cnt[slot * kCntStride];
      const uint32_t old = by_lane0_32(lane, [&] {
        return __hip_atomic_fetch_add(cnt, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      });
      if (old + n == total && lane == 0) {
        __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&done_flag[slot], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

struct ReqPayload {
  meso_scene *scene;
  DevCamera cam;
  DevShading sh;
  uint32_t flags;
  int32_t index_bytes;
  uint64_t pixels;
  uint32_t packets_x;
};

static uint64_t request_bytes(uint64_t pixels, int32_t index_bytes) {
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  return al(pixels * (uint64_t)index_bytes) + al(pixels) + al(pixels * 4) + al(pixels * 4);
}

// The device side of the server over the persistent kernel.  publish() and
// tick() run on the scheduler thread only; launches() / busy_ms() are read
// from client threads, hence the atomics.
class HipServeDevice final : public meso_serve::ServeDevice<ReqPayload> {
 public:
  HipServeDevice(meso_context *ctx, uint32_t ring, int grid, uint64_t idle_ticks,
                 uint64_t rescan_ticks)
      : ctx_(ctx), ring_(ring), grid_(grid), idle_ticks_(idle_ticks), rescan_ticks_(rescan_ticks) {}
  ~HipServeDevice() override { release(); }

  meso_status init(uint64_t arena_bytes) {
    MESO_HIP(hipSetDevice(ctx_->device));
    hipError_t e = hipMalloc((void **)&arena_, arena_bytes);
    if (e != hipSuccess)
      return fail(MESO_ERR_OUT_OF_MEMORY, "arena of %llu bytes: %s",
                  (unsigned long long)arena_bytes, hipGetErrorString(e));
    // host-coherent ring, control words and completion flags
    MESO_HIP(hipHostMalloc((void **)&desc_, sizeof(ServeDesc) * ring_, hipHostMallocCoherent));
    MESO_HIP(hipHostMalloc((void **)&ctl_, sizeof(ServeCtl), hipHostMallocCoherent));
    MESO_HIP(hipHostMalloc((void **)&done_flag_, sizeof(uint64_t) * ring_, hipHostMallocCoherent));
    std::memset(desc_, 0, sizeof(ServeDesc) * ring_);
    std::memset(ctl_, 0, sizeof(ServeCtl));
    std::memset(done_flag_, 0, sizeof(uint64_t) * ring_);
    MESO_HIP(hipHostGetDevicePointer((void **)&d_desc_, desc_, 0));
    MESO_HIP(hipHostGetDevicePointer((void **)&d_ctl_, ctl_, 0));
    MESO_HIP(hipHostGetDevicePointer((void **)&d_done_flag_, done_flag_, 0));
    // device state: ServeState, the completion counters, the device ring
    const uint64_t state_bytes = ddesc_offset() + sizeof(ServeDesc) * ring_;
    MESO_HIP(hipMalloc((void **)&state_, state_bytes));
    MESO_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    MESO_HIP(hipEventCreate(&ev_start_));
    MESO_HIP(hipEventCreate(&ev_end_));
    const meso_status st = launch_zero(state_, state_bytes, stream_);
    if (st != MESO_OK) return st;
    MESO_HIP(hipStreamSynchronize(stream_));
    return MESO_OK;
  }

  uint8_t *arena() const { return arena_; }

  void publish(uint64_t seq, uint32_t slot, uint64_t arena_off, uint32_t packets,
               const ReqPayload &p) override {
    ServeDesc d{};
    d.seq = seq;
    d.packets = packets;
    d.packets_x = p.packets_x;
    d.units = std::max(packets, kServeHeads);  // at least one unit on every head
    d.kind = p.scene->kind;
    d.flags = p.flags;
    d.index_bytes = p.index_bytes;
    auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
    uint8_t *base = arena_ + arena_off;
    d.index = base;
    d.face = base + al(p.pixels * (uint64_t)p.index_bytes);
    d.t = (float *)(d.face + al(p.pixels));
    d.shade = (float *)((uint8_t *)d.t + al(p.pixels * 4));
    d.cam = p.cam;
    d.sh = p.sh;
    d.scene = p.scene->dev();
    // the slot's previous entry has completed, so no workgroup reads it; the
    // count is published after the words, in order (x86 stores are ordered
    // and the release keeps the compiler from moving them)
    uint64_t *dst = (uint64_t *)&desc_[slot];
    const uint64_t *src = (const uint64_t *)&d;
    for (uint32_t i = 0; i < kDescWords; i++) __atomic_store_n(dst + i, src[i], __ATOMIC_RELAXED);
    __atomic_store_n(&ctl_->published, seq, __ATOMIC_RELEASE);
  }

  bool completed(uint32_t slot, uint64_t seq) override {
    return __atomic_load_n(&done_flag_[slot], __ATOMIC_ACQUIRE) == seq;
  }

  bool tick(bool outstanding) override {
    if (running_) {
      const hipError_t q = hipEventQuery(ev_end_);
      if (q == hipSuccess) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ev_start_, ev_end_);
        busy_us_.fetch_add((uint64_t)(ms * 1000.0f));
        running_ = false;
      } else if (q != hipErrorNotReady) {
        set_error(std::string("serve kernel: ") + hipGetErrorString(q));
        return false;
      }
    }
    if (!outstanding || running_) return true;
    // a new launch: its number lets the previous launch's ending mark pass it
    gen_++;
    hipSetDevice(ctx_->device);
    hipEventRecord(ev_start_, stream_);
    hipLaunchKernelGGL(serve_persistent_kernel, dim3((uint32_t)grid_), dim3(kServeThreads), 0,
                       stream_, d_desc_, (ServeDesc *)((uint8_t *)state_ + ddesc_offset()),
                       d_ctl_, d_done_flag_, (ServeState *)state_,
                       (uint32_t *)((uint8_t *)state_ + sizeof(ServeState)), ring_ - 1, gen_,
                       idle_ticks_, rescan_ticks_);
    const hipError_t e = hipGetLastError();
    hipEventRecord(ev_end_, stream_);
    if (e != hipSuccess) {
      set_error(std::string("serve kernel launch: ") + hipGetErrorString(e));
      return false;
    }
    running_ = true;
    launches_.fetch_add(1);
    return true;
  }

  std::string error() override {
    std::lock_guard<std::mutex> lk(err_mu_);
    return error_;
  }
  uint64_t launches() override { return launches_.load(); }
  double busy_ms() override { return (double)busy_us_.load() / 1000.0; }

  void shutdown() override {
    if (!ctl_) return;
    __atomic_store_n(&ctl_->stop, 1u, __ATOMIC_RELEASE);
    if (stream_) hipStreamSynchronize(stream_);
    if (running_) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev_start_, ev_end_) == hipSuccess)
        busy_us_.fetch_add((uint64_t)(ms * 1000.0f));
      running_ = false;
    }
  }

 private:
  void set_error(const std::string &m) {
    std::lock_guard<std::mutex> lk(err_mu_);
    error_ = m;
  }
  uint64_t ddesc_offset() const {
    return sizeof(ServeState) + sizeof(uint32_t) * kCntStride * ring_;
  }
  void release() {
    shutdown();
    hipSetDevice(ctx_->device);
    if (ev_start_) hipEventDestroy(ev_start_);
    if (ev_end_) hipEventDestroy(ev_end_);
    if (stream_) hipStreamDestroy(stream_);
    if (state_) hipFree(state_);
    if (arena_) hipFree(arena_);
    if (desc_) hipHostFree(desc_);
    if (ctl_) hipHostFree(ctl_);
    if (done_flag_) hipHostFree(done_flag_);
    ev_start_ = ev_end_ = nullptr;
    stream_ = nullptr;
    state_ = nullptr;
    arena_ = nullptr;
    desc_ = nullptr;
    ctl_ = nullptr;
    done_flag_ = nullptr;
  }

  meso_context *ctx_;
  uint32_t ring_;
  int grid_;
  uint64_t idle_ticks_, rescan_ticks_;
  uint8_t *arena_ = nullptr;
  ServeDesc *desc_ = nullptr, *d_desc_ = nullptr;
  ServeCtl *ctl_ = nullptr, *d_ctl_ = nullptr;
  uint64_t *done_flag_ = nullptr, *d_done_flag_ = nullptr;
  void *state_ = nullptr;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_start_ = nullptr, ev_end_ = nullptr;
  bool running_ = false;
  uint32_t gen_ = 0;
  std::atomic<uint64_t> launches_{0}, busy_us_{0};
  std::mutex err_mu_;
  std::string error_;
};

}  // namespace meso

using namespace meso;

struct meso_server {
  meso_context *ctx = nullptr;
  meso_serve_config cfg{};
  std::unique_ptr<HipServeDevice> dev;
  std::unique_ptr<meso_serve::ServeCore<ReqPayload>> core;
};

extern "C" {

meso_status meso_serve_start(meso_context *ctx, const meso_serve_config *cfg, meso_server **out) {
  MESO_CHECK(ctx && cfg && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  *out = nullptr;
  MESO_CHECK(cfg->size >= sizeof(meso_serve_config), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_serve_config.size too small");
  MESO_CHECK(cfg->hbm_budget_bytes >= (1u << 20), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "budget below 1 MiB");
  MESO_CHECK(cfg->max_inflight >= 1 && cfg->max_inflight <= (int32_t)kMaxRing,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "max_inflight %d not in 1..%u", cfg->max_inflight,
             kMaxRing);
  MESO_CHECK(cfg->max_pending >= 1, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "max_pending must be >= 1");
  MESO_CHECK(cfg->workgroups_per_cu >= 0 && cfg->workgroups_per_cu <= 8,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "workgroups_per_cu %d not in 0..8",
             cfg->workgroups_per_cu);
  MESO_CHECK(cfg->idle_exit_us >= 0 && cfg->idle_exit_us <= 1000000, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "idle_exit_us %d not in 0..1000000", cfg->idle_exit_us);
  MESO_CHECK(cfg->tile == 0 || cfg->tile == kServePacket, MESO_ERR_NOT_SUPPORTED,
             "packet edge must be %d", kServePacket);
  MESO_HIP(hipSetDevice(ctx->device));
  int per_cu = 0;
  MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_persistent_kernel,
                                                         kServeThreads, 0));
  if (per_cu < 1) per_cu = 1;
  if (cfg->workgroups_per_cu > 0 && cfg->workgroups_per_cu < per_cu) per_cu = cfg->workgroups_per_cu;
  const int grid = std::max<int>(per_cu * ctx->props.multiProcessorCount, (int)kServeHeads);
  int clock_khz = 0;
  MESO_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
  const int idle_us = cfg->idle_exit_us ? cfg->idle_exit_us : 500;
  const uint64_t ticks_per_ms = (uint64_t)(clock_khz > 0 ? clock_khz : 100000);
  const uint64_t idle_ticks = (uint64_t)idle_us * ticks_per_ms / 1000;
  const uint64_t rescan_ticks = (uint64_t)kRescanUs * ticks_per_ms / 1000;
  uint32_t ring = 1;
  while (ring < (uint32_t)cfg->max_inflight) ring <<= 1;
  auto s = std::make_unique<meso_server>();
  s->ctx = ctx;
  s->cfg = *cfg;
  s->dev = std::make_unique<HipServeDevice>(ctx, ring, grid, idle_ticks, rescan_ticks);
  const meso_status st = s->dev->init(cfg->hbm_budget_bytes);
  if (st != MESO_OK) return st;
  meso_serve::CoreConfig cc;
  cc.budget_bytes = cfg->hbm_budget_bytes;
  cc.max_pending = cfg->max_pending;
  cc.ring_entries = (int32_t)ring;
  s->core = std::make_unique<meso_serve::ServeCore<ReqPayload>>(cc, s->dev.get());
  s->core->start();
  *out = s.release();
  return MESO_OK;
}

meso_status meso_serve_submit(meso_server *s, meso_scene *scene, const meso_camera *cam,
                              uint32_t flags, const meso_shading *shading, uint64_t *ticket) {
  MESO_CHECK(s && scene && cam && ticket, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->ctx == s->ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "scene of another context");
  {
    // finite eye / target / up, an up vector not parallel to the view, and
    // the size and fov bounds, as meso_raymarch checks them
    const meso_status cs = check_camera(cam);
    if (cs != MESO_OK) return cs;
  }
  MESO_CHECK((flags & ~(MESO_RAY_SHADOW | MESO_RAY_AO)) == 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "serving takes MESO_RAY_SHADOW / MESO_RAY_AO only");
  ReqPayload r{};
  r.scene = scene;
  derive_camera(cam, &r.cam);
  r.flags = flags;
  r.sh.light[1] = 1.0f;
  if (flags) {
    MESO_CHECK(shading && shading->size >= sizeof(meso_shading), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "shadow/AO rays need a meso_shading");
    MESO_CHECK(shading->light[0] != 0.0f || shading->light[1] != 0.0f || shading->light[2] != 0.0f,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "light direction is zero");
    if (flags & MESO_RAY_AO)
      MESO_CHECK(shading->ao_samples >= 1 && shading->ao_samples <= 8 && shading->ao_radius > 0.0f,
                 MESO_ERR_ARGUMENT_OUT_OF_RANGE, "AO needs 1..8 samples and a positive radius");
    derive_shading(shading, &r.sh);
  }
  const uint64_t nvox = (uint64_t)scene->n * scene->n * scene->n;
  r.index_bytes = nvox > 0x7FFFFFFFull ? 8 : 4;
  r.pixels = (uint64_t)cam->width * (uint64_t)cam->height;
  r.packets_x = (uint32_t)((cam->width + kServePacket - 1) / kServePacket);
  const uint64_t packets =
      (uint64_t)r.packets_x * (uint64_t)((cam->height + kServePacket - 1) / kServePacket);
  MESO_CHECK(packets < kMaxPackets, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "request of %dx%d pixels exceeds the %u 8x8 packets of one served request",
             cam->width, cam->height, kMaxPackets - 1);
  const uint64_t bytes = request_bytes(r.pixels, r.index_bytes);
  MESO_CHECK(bytes <= s->cfg.hbm_budget_bytes, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "request needs %llu bytes, budget is %llu", (unsigned long long)bytes,
             (unsigned long long)s->cfg.hbm_budget_bytes);
  const meso_status st = s->core->submit(r, bytes, (uint32_t)packets, r.pixels, ticket);
  if (st == MESO_ERR_BUSY)
    return fail(MESO_ERR_BUSY, "pending queue full (%d)", s->cfg.max_pending);
  if (st != MESO_OK) return fail(st, "server stopped or failed: %s", s->core->error().c_str());
  return MESO_OK;
}

meso_status meso_serve_poll(meso_server *s, uint64_t ticket, int32_t *done) {
  MESO_CHECK(s && done, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(s->core->poll(ticket, done), MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown ticket %llu",
             (unsigned long long)ticket);
  return MESO_OK;
}

meso_status meso_serve_wait(meso_server *s, uint64_t ticket, int64_t timeout_us,
                            double *latency_us) {
  MESO_CHECK(s, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL server");
  const meso_status st = s->core->wait(ticket, timeout_us, latency_us);
  switch (st) {
    case MESO_OK: return MESO_OK;
    case MESO_ERR_BUSY:
      return fail(MESO_ERR_BUSY, "timeout waiting for ticket %llu", (unsigned long long)ticket);
    case MESO_ERR_ARGUMENT_OUT_OF_RANGE:
      return fail(st, "unknown ticket %llu", (unsigned long long)ticket);
    default:
      return fail(st, "request failed or the server stopped: %s", s->core->error().c_str());
  }
}

meso_status meso_serve_take(meso_server *s, uint64_t ticket, const meso_host_outputs *dst) {
  MESO_CHECK(s && dst && dst->size >= sizeof(meso_host_outputs), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "NULL argument");
  meso_serve::Ticket<ReqPayload> r;
  const meso_status bt = s->core->begin_take(ticket, &r);
  if (bt == MESO_ERR_ARGUMENT_OUT_OF_RANGE) return fail(bt, "unknown ticket");
  if (bt != MESO_OK) return fail(bt, "ticket not complete");
  const ReqPayload &p = r.payload;
  if (dst->index_bytes != 0 && dst->index_bytes != p.index_bytes) {
    s->core->end_take(ticket, false);
    return fail(MESO_ERR_ARGUMENT_OUT_OF_RANGE, "result holds %d-byte indices", p.index_bytes);
  }
  hipSetDevice(s->ctx->device);
  const uint8_t *base = s->dev->arena() + r.arena_off;
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  const uint64_t o_face = al(p.pixels * (uint64_t)p.index_bytes);
  const uint64_t o_t = o_face + al(p.pixels);
  const uint64_t o_shade = o_t + al(p.pixels * 4);
  hipError_t e = hipSuccess;
  if (e == hipSuccess && dst->hit_index)
    e = hipMemcpy(dst->hit_index, base, p.pixels * p.index_bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dst->face)
    e = hipMemcpy(dst->face, base + o_face, p.pixels, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dst->t)
    e = hipMemcpy(dst->t, base + o_t, p.pixels * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dst->shade)
    e = hipMemcpy(dst->shade, base + o_shade, p.pixels * 4, hipMemcpyDeviceToHost);
  s->core->end_take(ticket, e == hipSuccess);
  if (e != hipSuccess) return fail(MESO_ERR_RUNTIME, "take: %s", hipGetErrorString(e));
  return MESO_OK;
}

meso_status meso_serve_stats_get(meso_server *s, meso_serve_stats *out) {
  MESO_CHECK(s && out && out->size >= sizeof(meso_serve_stats), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "NULL argument or meso_serve_stats.size too small");
  *out = s->core->stats();
  return MESO_OK;
}

meso_status meso_serve_stop(meso_server *s) {
  if (!s) return MESO_OK;
  s->core->stop();
  s->core.reset();
  s->dev.reset();
  delete s;
  return MESO_OK;
}

}  // extern "C"
