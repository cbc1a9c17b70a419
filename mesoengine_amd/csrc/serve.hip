// Serving: continuous batching of raymarch requests over HBM-resident scenes
// in one persistent kernel fed from a device ring.
//
// Host side (serve_core.h): requests are admitted in arrival order while
// their outputs fit an HBM arena of hbm_budget_bytes and a ring entry is
// free; an admitted request is published as the next ring entry.
//
// Device side: a persistent kernel of 256-lane workgroups pulls 16x16-pixel
// tiles (four 8x8 wave packets) of the published entries, oldest entry
// first, so a request that arrives while others run starts on tiles freed by
// them, inside the same launch.
//  - `head` (device memory, 64 bits) packs the entry being drawn from, its
//    tile count and the next tile: an atomic add claims a tile, and a claim
//    at or past the count means the entry is exhausted.  Exhaustion is thus
//    decided in device memory; a claimant reads the entry's descriptor (in
//    host-coherent memory) only for a valid claim, when the entry cannot
//    complete and its ring slot cannot be reused before that tile is done.
//  - An exhausted head moves to the next published entry by compare-and-swap.
//  - Every tile ends with a system-scope release add on its slot's counter
//    (its outputs are written back to memory first); the add that completes
//    the entry resets the counter and sets the slot's completion flag in host
//    memory, which the scheduler thread polls.
//  - Eight leader workgroups poll the host's published count and stop
//    request and keep them in device memory for the others, so the claim
//    loop of every other workgroup touches device memory only.  With nothing
//    to claim for idle_exit_us the leaders mark the launch as ending and
//    every workgroup returns; the scheduler launches the kernel again when
//    work is outstanding.
// Every pixel is computed by the same per-pixel code whatever else shares the
// kernel, so a result is identical alone, in a mix and on the plain path.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>

#include "serve_core.h"
#include "traverse.h"

namespace meso {

constexpr int kServeTile = 16;
constexpr uint32_t kServeThreads = 256;
constexpr uint32_t kServeLeaders = 8;  // one per XCD
// head word: entry (22 bits) | tile count (21 bits) | cursor (21 bits)
constexpr uint64_t kFieldMask = (1ull << 21) - 1;
constexpr uint64_t kEntryMask = (1ull << 22) - 1;
constexpr uint32_t kMaxRing = 4096;

__host__ __device__ inline uint64_t head_word(uint64_t entry, uint64_t tiles) {
  return ((entry & kEntryMask) << 42) | ((tiles & kFieldMask) << 21);
}

// One ring entry, in host-coherent memory (written by the host, read by the
// kernel with system-scope loads).
struct alignas(8) ServeDesc {
  uint64_t seq;  // entry number (from 1)
  uint32_t tiles_total, tiles_x;
  int32_t kind;
  uint32_t flags;
  int32_t index_bytes;
  int32_t pad;
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
static_assert(kDescWords <= kServeThreads, "one lane per descriptor word");

// host-coherent control words
struct ServeCtl {
  uint64_t published;  // last entry published (0: none)
  uint32_t stop;       // the host asks every workgroup to return
  uint32_t pad;
};

// device-memory state (zeroed once when the server starts); done_cnt[ring]
// follows it
struct ServeState {
  unsigned long long head;    // entry 0 is a placeholder with no tiles
  unsigned long long mirror;  // the leaders' copy of ServeCtl::published
  unsigned int exiting;       // launch number whose workgroups are returning
  unsigned int pad;
};

__device__ __forceinline__ uint64_t sys_load64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t sys_load32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ unsigned long long dev_load64(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

__global__ void __launch_bounds__(kServeThreads)
    serve_persistent_kernel(const ServeDesc *desc, const ServeCtl *ctl, uint64_t *done_flag,
                            ServeState *st, uint32_t *done_cnt, uint32_t ring_mask, uint32_t gen,
                            uint64_t idle_ticks) {
  __shared__ ServeDesc d;
  __shared__ uint32_t s_cmd, s_tile, s_load;
  __shared__ uint64_t s_entry;
  const bool leader = blockIdx.x < kServeLeaders;
  uint64_t have = 0;  // thread 0: entry + 1 whose descriptor d holds (0: none)
  uint64_t idle_since = wall_clock64();
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t cmd = 0, nap = 1;
      for (;;) {
        // only the leaders read host memory here: the stop request and the
        // published count, passed on to the others in device memory
        uint64_t pub = dev_load64(&st->mirror);
        if (leader) {
          if (sys_load32(&ctl->stop)) atomicMax(&st->exiting, 0xFFFFFFFFu);
          const uint64_t hp = sys_load64(&ctl->published);
          if (hp > pub) {
            atomicMax(&st->mirror, (unsigned long long)hp);
            pub = hp;
          }
        }
        if (__hip_atomic_load(&st->exiting, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= gen) {
          cmd = 2;
          break;
        }
        uint64_t h = dev_load64(&st->head);
        if ((h & kFieldMask) < ((h >> 21) & kFieldMask)) {
          // the entry has tiles left: claim one
          h = atomicAdd(&st->head, 1ull);
          const uint64_t tiles = (h >> 21) & kFieldMask, cur = h & kFieldMask;
          if (cur < tiles) {
            const uint64_t e = h >> 42;
            s_entry = e;
            s_tile = (uint32_t)cur;
            s_load = have != e + 1;
            have = e + 1;
            cmd = 1;
            break;
          }
          continue;  // exhausted meanwhile: look again
        }
        // the entry is exhausted: move the head to the next published entry
        const uint64_t e = h >> 42;
        const uint64_t ahead = (pub - e) & kEntryMask;
        if (ahead >= 1 && ahead < (1ull << 21)) {
          const uint64_t n = (e + 1) & kEntryMask;
          const uint32_t nt = sys_load32(&desc[n & ring_mask].tiles_total);
          atomicCAS(&st->head, (unsigned long long)h, (unsigned long long)head_word(n, nt));
          continue;  // whoever won, claim from the new head
        }
        // nothing to claim
        have = 0;
        if (wall_clock64() - idle_since > idle_ticks) {
          if (leader) atomicMax(&st->exiting, gen);
          cmd = 2;
          break;
        }
        for (uint32_t k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(32);
        nap = nap < 16 ? nap * 2 : 16;
      }
      s_cmd = cmd;
    }
    __syncthreads();
    if (s_cmd == 2) return;
    if (s_load) {
      // a valid claim holds the entry open, so its descriptor is stable
      const uint64_t *src = (const uint64_t *)&desc[s_entry & ring_mask];
      if (threadIdx.x < kDescWords) ((uint64_t *)&d)[threadIdx.x] = sys_load64(src + threadIdx.x);
      __syncthreads();
    }
    const uint32_t tiles_x = __builtin_amdgcn_readfirstlane(d.tiles_x);
    const uint32_t tx = s_tile % tiles_x, ty = s_tile / tiles_x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const DevCamera cam = uniform_copy(d.cam);
    const int32_t px = (int32_t)(tx * kServeTile + (wave & 1u) * 8u + (lane & 7u));
    const int32_t py = (int32_t)(ty * kServeTile + (wave >> 1) * 8u + (lane >> 3));
    if (px < cam.width && py < cam.height) {
      const DevScene scene = uniform_copy(d.scene);
      switch (__builtin_amdgcn_readfirstlane(d.kind)) {
        case MESO_SCENE_DENSE_U8: serve_pixel<1>(d, scene, cam, px, py); break;
        case MESO_SCENE_BRICKMAP_U16: serve_pixel<2>(d, scene, cam, px, py); break;
        default: serve_pixel<3>(d, scene, cam, px, py); break;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      // the tile's outputs are written back before the count moves
      const uint32_t slot = (uint32_t)s_entry & ring_mask;
      const uint32_t old =
          __hip_atomic_fetch_add(&done_cnt[slot], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      if (old + 1 == d.tiles_total) {
        __hip_atomic_store(&done_cnt[slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&done_flag[slot], (uint64_t)d.seq, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
      idle_since = wall_clock64();
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
  uint32_t tiles_x;
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
  HipServeDevice(meso_context *ctx, uint32_t ring, int grid, uint64_t idle_ticks)
      : ctx_(ctx), ring_(ring), grid_(grid), idle_ticks_(idle_ticks) {}
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
    const uint64_t state_bytes = sizeof(ServeState) + sizeof(uint32_t) * ring_;
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

  void publish(uint64_t seq, uint32_t slot, uint64_t arena_off, uint32_t tiles,
               const ReqPayload &p) override {
    ServeDesc d{};
    d.seq = seq;
    d.tiles_total = tiles;
    d.tiles_x = p.tiles_x;
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
                       stream_, d_desc_, d_ctl_, d_done_flag_, (ServeState *)state_,
                       (uint32_t *)((uint8_t *)state_ + sizeof(ServeState)), ring_ - 1, gen_,
                       idle_ticks_);
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
  uint64_t idle_ticks_;
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
  MESO_CHECK(cfg->tile == 0 || cfg->tile == kServeTile, MESO_ERR_NOT_SUPPORTED,
             "tile edge must be %d", kServeTile);
  MESO_HIP(hipSetDevice(ctx->device));
  int per_cu = 0;
  MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_persistent_kernel,
                                                         kServeThreads, 0));
  if (per_cu < 1) per_cu = 1;
  if (cfg->workgroups_per_cu > 0 && cfg->workgroups_per_cu < per_cu) per_cu = cfg->workgroups_per_cu;
  const int grid = std::max<int>(per_cu * ctx->props.multiProcessorCount, (int)kServeLeaders);
  int clock_khz = 0;
  MESO_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
  const int idle_us = cfg->idle_exit_us ? cfg->idle_exit_us : 500;
  const uint64_t idle_ticks = (uint64_t)idle_us * (uint64_t)(clock_khz > 0 ? clock_khz : 100000) / 1000;
  uint32_t ring = 1;
  while (ring < (uint32_t)cfg->max_inflight) ring <<= 1;
  auto s = std::make_unique<meso_server>();
  s->ctx = ctx;
  s->cfg = *cfg;
  s->dev = std::make_unique<HipServeDevice>(ctx, ring, grid, idle_ticks);
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
  r.tiles_x = (uint32_t)((cam->width + kServeTile - 1) / kServeTile);
  const uint32_t tiles = r.tiles_x * (uint32_t)((cam->height + kServeTile - 1) / kServeTile);
  const uint64_t bytes = request_bytes(r.pixels, r.index_bytes);
  MESO_CHECK(bytes <= s->cfg.hbm_budget_bytes, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "request needs %llu bytes, budget is %llu", (unsigned long long)bytes,
             (unsigned long long)s->cfg.hbm_budget_bytes);
  const meso_status st = s->core->submit(r, bytes, tiles, r.pixels, ticket);
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
