// Serving: continuous batching of raymarch requests over HBM-resident scenes.
//
// A scheduler thread
//  - admits queued requests in arrival order while their output bytes fit in
//    a preallocated HBM arena of hbm_budget_bytes (the budget binds: a request
//    waits until earlier results are taken and their bytes freed);
//  - packs 16x16-pixel tiles of every admitted request, oldest first, into one
//    shared launch of at most rays_per_launch rays, so a large request is
//    spread over several launches and never blocks small ones for long;
//  - keeps up to max_inflight launches queued on separate streams, so requests
//    admitted while others run start on the device as soon as it has room;
//  - completes a request when the launch holding its last tile has finished.
// Every pixel is computed by the same per-pixel code whatever else shares the
// launch, so a result is identical alone and in a mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <thread>
#include <unordered_map>
#include <vector>

#include "traverse.h"

namespace meso {

constexpr int kServeTile = 16;

struct ServeReqDev {
  DevScene scene;
  int32_t kind;
  uint32_t flags;
  DevCamera cam;
  DevShading sh;
  void *index;
  int32_t index_bytes;
  uint8_t *face;
  float *t;
  float *shade;
  uint32_t tiles_x;
};

struct WorkItem {
  uint32_t req;   // index into the launch's request table
  uint32_t tile;  // tile index within that request
};

template <int KIND>
__device__ __forceinline__ void serve_pixel(const ServeReqDev &r, int32_t px, int32_t py) {
  StepCounts cnt = {0, 0, 0, 0, 0, 0, 0};
  uint32_t ns = 0, na = 0;
  const PixelResult res = shade_pixel<KIND, false>(r.scene, r.cam, r.flags, r.sh, px, py, cnt, ns, na);
  const int64_t pix = (int64_t)py * r.cam.width + px;
  if (r.index_bytes == 8)
    ((int64_t *)r.index)[pix] = res.index;
  else
    ((int32_t *)r.index)[pix] = (int32_t)res.index;
  r.face[pix] = (uint8_t)res.face;
  r.t[pix] = res.t;
  r.shade[pix] = res.shade;
}

// one workgroup = one 16x16 tile of one request (four 8x8 wave packets)
__global__ void __launch_bounds__(256) serve_kernel(const ServeReqDev *reqs, const WorkItem *items) {
  const WorkItem it = items[blockIdx.x];
  const ServeReqDev &r = reqs[it.req];
  const uint32_t tx = it.tile % r.tiles_x, ty = it.tile / r.tiles_x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const int32_t px = (int32_t)(tx * kServeTile + (wave & 1u) * 8u + (lane & 7u));
  const int32_t py = (int32_t)(ty * kServeTile + (wave >> 1) * 8u + (lane >> 3));
  if (px >= r.cam.width || py >= r.cam.height) return;
  switch (r.kind) {
    case MESO_SCENE_DENSE_U8: serve_pixel<1>(r, px, py); break;
    case MESO_SCENE_BRICKMAP_U16: serve_pixel<2>(r, px, py); break;
    default: serve_pixel<3>(r, px, py); break;
  }
}

using Clock = std::chrono::steady_clock;

enum class ReqState { Pending, Admitted, Done, Failed };

struct Request {
  uint64_t ticket;
  meso_scene *scene;
  DevCamera cam;
  uint32_t flags;
  DevShading sh;
  int32_t index_bytes;
  uint64_t pixels;
  uint64_t bytes;
  uint64_t arena_off = 0;
  uint32_t tiles_x = 0, tiles_total = 0, tiles_issued = 0;
  int64_t last_launch = -1;
  ReqState state = ReqState::Pending;
  Clock::time_point t_submit, t_done;
};

// First-fit allocator over the arena (offsets, 256-byte granules).
struct Arena {
  uint64_t size = 0;
  std::list<std::pair<uint64_t, uint64_t>> free_list;  // (offset, bytes), sorted
  void init(uint64_t s) {
    size = s;
    free_list.clear();
    free_list.emplace_back(0, s);
  }
  bool alloc(uint64_t bytes, uint64_t &off) {
    bytes = (bytes + 255) & ~255ull;
    for (auto it = free_list.begin(); it != free_list.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        it->first += bytes;
        it->second -= bytes;
        if (!it->second) free_list.erase(it);
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = (bytes + 255) & ~255ull;
    auto it = free_list.begin();
    while (it != free_list.end() && it->first < off) ++it;
    it = free_list.insert(it, {off, bytes});
    auto nx = std::next(it);
    if (nx != free_list.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_list.erase(nx);
    }
    if (it != free_list.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_list.erase(it);
      }
    }
  }
};

struct Launch {
  int64_t id;
  int slot;  // stream slot
  hipEvent_t start, done;
  std::vector<uint64_t> finishing;  // tickets whose last tile is in this launch
};

static uint64_t request_bytes(uint64_t pixels, int32_t index_bytes) {
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  return al(pixels * (uint64_t)index_bytes) + al(pixels) + al(pixels * 4) + al(pixels * 4);
}

}  // namespace meso

using namespace meso;

struct meso_server {
  meso_context *ctx = nullptr;
  meso_serve_config cfg{};
  uint8_t *arena_mem = nullptr;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;            // tickets waiting for admission
  std::deque<uint64_t> active;             // admitted, tiles not all issued
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread thread;
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;
  std::deque<Launch> inflight;
  // pinned staging for request tables / work lists, one set per stream slot
  std::vector<ServeReqDev *> h_reqs, d_reqs;
  std::vector<WorkItem *> h_items, d_items;
  uint32_t max_items = 0, max_reqs = 0;
  int64_t launch_seq = 0;
  meso_serve_stats stats{};
  std::string error;
};

static void serve_loop(meso_server *s);

static void server_free(meso_server *s) {
  hipSetDevice(s->ctx->device);
  for (auto st : s->streams) {
    hipStreamSynchronize(st);
    hipStreamDestroy(st);
  }
  for (auto p : s->h_reqs) hipHostFree(p);
  for (auto p : s->h_items) hipHostFree(p);
  for (auto p : s->d_reqs) hipFree(p);
  for (auto p : s->d_items) hipFree(p);
  if (s->arena_mem) hipFree(s->arena_mem);
  delete s;
}

extern "C" {

meso_status meso_serve_start(meso_context *ctx, const meso_serve_config *cfg, meso_server **out) {
  MESO_CHECK(ctx && cfg && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(cfg->size >= sizeof(meso_serve_config), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_serve_config.size too small");
  MESO_CHECK(cfg->hbm_budget_bytes >= (1u << 20), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "budget below 1 MiB");
  MESO_CHECK(cfg->max_inflight >= 1 && cfg->max_inflight <= 4, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "max_inflight %d not in 1..4", cfg->max_inflight);
  MESO_CHECK(cfg->max_pending >= 1 && cfg->rays_per_launch >= 256, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "max_pending >= 1 and rays_per_launch >= 256 required");
  MESO_CHECK(cfg->tile == 0 || cfg->tile == kServeTile, MESO_ERR_NOT_SUPPORTED,
             "tile edge must be %d", kServeTile);
  MESO_HIP(hipSetDevice(ctx->device));
  auto *s = new meso_server();
  s->ctx = ctx;
  s->cfg = *cfg;
  hipError_t e = hipMalloc((void **)&s->arena_mem, cfg->hbm_budget_bytes);
  if (e != hipSuccess) {
    delete s;
    return fail(MESO_ERR_OUT_OF_MEMORY, "arena of %llu bytes: %s",
                (unsigned long long)cfg->hbm_budget_bytes, hipGetErrorString(e));
  }
  s->arena.init(cfg->hbm_budget_bytes);
  s->max_items = (uint32_t)((cfg->rays_per_launch + kServeTile * kServeTile - 1) /
                            (kServeTile * kServeTile)) + 1;
  s->max_reqs = (uint32_t)std::min<int64_t>(cfg->max_pending + 64, 4096);
  for (int i = 0; i < cfg->max_inflight && e == hipSuccess; i++) {
    hipStream_t st;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) break;
    s->streams.push_back(st);
    s->slot_busy.push_back(false);
    ServeReqDev *hr = nullptr, *dr = nullptr;
    WorkItem *hi = nullptr, *di = nullptr;
    if (e == hipSuccess) e = hipHostMalloc((void **)&hr, s->max_reqs * sizeof(ServeReqDev));
    if (e == hipSuccess) e = hipHostMalloc((void **)&hi, s->max_items * sizeof(WorkItem));
    if (e == hipSuccess) e = hipMalloc((void **)&dr, s->max_reqs * sizeof(ServeReqDev));
    if (e == hipSuccess) e = hipMalloc((void **)&di, s->max_items * sizeof(WorkItem));
    s->h_reqs.push_back(hr);
    s->h_items.push_back(hi);
    s->d_reqs.push_back(dr);
    s->d_items.push_back(di);
  }
  if (e != hipSuccess) {
    server_free(s);
    return fail(MESO_ERR_RUNTIME, "server setup: %s", hipGetErrorString(e));
  }
  s->stats.size = sizeof(meso_serve_stats);
  s->thread = std::thread(serve_loop, s);
  *out = s;
  return MESO_OK;
}

meso_status meso_serve_submit(meso_server *s, meso_scene *scene, const meso_camera *cam,
                              uint32_t flags, const meso_shading *shading, uint64_t *ticket) {
  MESO_CHECK(s && scene && cam && ticket, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->ctx == s->ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "scene of another context");
  MESO_CHECK(cam->size >= sizeof(meso_camera) && cam->width > 0 && cam->height > 0 &&
                 cam->width <= 16384 && cam->height <= 16384 && cam->vfov_deg > 0.0f &&
                 cam->vfov_deg < 180.0f,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "invalid camera");
  MESO_CHECK((flags & ~(MESO_RAY_SHADOW | MESO_RAY_AO)) == 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "serving takes MESO_RAY_SHADOW / MESO_RAY_AO only");
  Request r{};
  r.scene = scene;
  derive_camera(cam, &r.cam);
  r.flags = flags;
  std::memset(&r.sh, 0, sizeof(r.sh));
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
  r.bytes = request_bytes(r.pixels, r.index_bytes);
  MESO_CHECK(r.bytes <= s->cfg.hbm_budget_bytes, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "request needs %llu bytes, budget is %llu", (unsigned long long)r.bytes,
             (unsigned long long)s->cfg.hbm_budget_bytes);
  r.tiles_x = (uint32_t)((cam->width + kServeTile - 1) / kServeTile);
  r.tiles_total = r.tiles_x * (uint32_t)((cam->height + kServeTile - 1) / kServeTile);
  r.t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    MESO_CHECK(!s->stop, MESO_ERR_RUNTIME, "server stopped");
    if ((int64_t)s->pending.size() >= s->cfg.max_pending) {
      s->stats.refused_busy++;
      return fail(MESO_ERR_BUSY, "pending queue full (%d)", s->cfg.max_pending);
    }
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    s->pending.push_back(r.ticket);
    s->reqs.emplace(r.ticket, r);
    s->stats.submitted++;
  }
  s->cv_work.notify_one();
  return MESO_OK;
}

meso_status meso_serve_poll(meso_server *s, uint64_t ticket, int32_t *done) {
  MESO_CHECK(s && done, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  MESO_CHECK(it != s->reqs.end(), MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown ticket %llu",
             (unsigned long long)ticket);
  *done = it->second.state == ReqState::Done ? 1 : (it->second.state == ReqState::Failed ? -1 : 0);
  return MESO_OK;
}

meso_status meso_serve_wait(meso_server *s, uint64_t ticket, int64_t timeout_us,
                            double *latency_us) {
  MESO_CHECK(s, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto deadline = Clock::now() + std::chrono::microseconds(timeout_us < 0 ? 0 : timeout_us);
  for (;;) {
    auto it = s->reqs.find(ticket);
    MESO_CHECK(it != s->reqs.end(), MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown ticket %llu",
               (unsigned long long)ticket);
    if (it->second.state == ReqState::Done) {
      if (latency_us)
        *latency_us = std::chrono::duration<double, std::micro>(it->second.t_done -
                                                                it->second.t_submit).count();
      return MESO_OK;
    }
    if (it->second.state == ReqState::Failed)
      return fail(MESO_ERR_RUNTIME, "request failed: %s", s->error.c_str());
    if (timeout_us >= 0) {
      if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
        auto it2 = s->reqs.find(ticket);
        if (it2 != s->reqs.end() && it2->second.state == ReqState::Done) continue;
        return fail(MESO_ERR_BUSY, "timeout waiting for ticket %llu", (unsigned long long)ticket);
      }
    } else {
      s->cv_done.wait(lk);
    }
  }
}

meso_status meso_serve_take(meso_server *s, uint64_t ticket, const meso_host_outputs *dst) {
  MESO_CHECK(s && dst && dst->size >= sizeof(meso_host_outputs), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "NULL argument");
  Request r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    MESO_CHECK(it != s->reqs.end(), MESO_ERR_ARGUMENT_OUT_OF_RANGE, "unknown ticket");
    MESO_CHECK(it->second.state == ReqState::Done, MESO_ERR_BUSY, "ticket not complete");
    r = it->second;
  }
  MESO_CHECK(dst->index_bytes == 0 || dst->index_bytes == r.index_bytes,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "result holds %d-byte indices", r.index_bytes);
  MESO_HIP(hipSetDevice(s->ctx->device));
  const uint8_t *base = s->arena_mem + r.arena_off;
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  const uint64_t o_face = al(r.pixels * (uint64_t)r.index_bytes);
  const uint64_t o_t = o_face + al(r.pixels);
  const uint64_t o_shade = o_t + al(r.pixels * 4);
  if (dst->hit_index) MESO_HIP(hipMemcpy(dst->hit_index, base, r.pixels * r.index_bytes, hipMemcpyDeviceToHost));
  if (dst->face) MESO_HIP(hipMemcpy(dst->face, base + o_face, r.pixels, hipMemcpyDeviceToHost));
  if (dst->t) MESO_HIP(hipMemcpy(dst->t, base + o_t, r.pixels * 4, hipMemcpyDeviceToHost));
  if (dst->shade) MESO_HIP(hipMemcpy(dst->shade, base + o_shade, r.pixels * 4, hipMemcpyDeviceToHost));
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->arena.release(r.arena_off, r.bytes);
    s->stats.bytes_in_use -= r.bytes;
    s->reqs.erase(ticket);
  }
  s->cv_work.notify_one();
  return MESO_OK;
}

meso_status meso_serve_stats_get(meso_server *s, meso_serve_stats *out) {
  MESO_CHECK(s && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return MESO_OK;
}

meso_status meso_serve_stop(meso_server *s) {
  if (!s) return MESO_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->thread.joinable()) s->thread.join();
  for (auto &l : s->inflight) {
    hipEventSynchronize(l.done);
    hipEventDestroy(l.start);
    hipEventDestroy(l.done);
  }
  server_free(s);
  return MESO_OK;
}

}  // extern "C"

// Scheduler: admission, batching, launch, completion.
static void serve_loop(meso_server *s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // 1. completion of finished launches (in order)
    while (!s->inflight.empty()) {
      Launch &l = s->inflight.front();
      if (hipEventQuery(l.done) != hipSuccess) break;
      float ms = 0.0f;
      hipEventElapsedTime(&ms, l.start, l.done);
      s->stats.busy_ms += ms;
      const auto now = Clock::now();
      for (uint64_t tk : l.finishing) {
        auto it = s->reqs.find(tk);
        if (it != s->reqs.end()) {
          it->second.state = ReqState::Done;
          it->second.t_done = now;
          s->stats.completed++;
        }
      }
      s->slot_busy[(size_t)l.slot] = false;
      hipEventDestroy(l.start);
      hipEventDestroy(l.done);
      s->inflight.pop_front();
      s->cv_done.notify_all();
    }
    if (s->stop && s->inflight.empty()) break;
    // 2. admission in arrival order while the arena has room
    while (!s->stop && !s->pending.empty()) {
      Request &r = s->reqs[s->pending.front()];
      uint64_t off = 0;
      if (!s->arena.alloc(r.bytes, off)) break;  // budget binds: wait for frees
      r.arena_off = off;
      r.state = ReqState::Admitted;
      s->stats.admitted++;
      s->stats.bytes_in_use += r.bytes;
      s->stats.peak_bytes_in_use = std::max(s->stats.peak_bytes_in_use, s->stats.bytes_in_use);
      s->active.push_back(r.ticket);
      s->pending.pop_front();
    }
    // 3. one shared launch per free stream slot
    int slot = -1;
    for (size_t i = 0; i < s->slot_busy.size(); i++)
      if (!s->slot_busy[i]) {
        slot = (int)i;
        break;
      }
    if (!s->stop && slot >= 0 && !s->active.empty()) {
      ServeReqDev *hr = s->h_reqs[(size_t)slot];
      WorkItem *hi = s->h_items[(size_t)slot];
      uint32_t n_items = 0, n_reqs = 0;
      const uint32_t item_cap = s->max_items - 1;
      Launch l;
      l.id = s->launch_seq++;
      l.slot = slot;
      uint64_t rays = 0;
      for (auto it = s->active.begin(); it != s->active.end() && n_items < item_cap &&
                                        n_reqs < s->max_reqs;) {
        Request &r = s->reqs[*it];
        const uint32_t take = std::min(r.tiles_total - r.tiles_issued, item_cap - n_items);
        ServeReqDev &d = hr[n_reqs];
        d.scene = r.scene->dev();
        d.kind = r.scene->kind;
        d.flags = r.flags;
        d.cam = r.cam;
        d.sh = r.sh;
        uint8_t *base = s->arena_mem + r.arena_off;
        auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
        d.index = base;
        d.index_bytes = r.index_bytes;
        d.face = base + al(r.pixels * (uint64_t)r.index_bytes);
        d.t = (float *)(base + al(r.pixels * (uint64_t)r.index_bytes) + al(r.pixels));
        d.shade = (float *)((uint8_t *)d.t + al(r.pixels * 4));
        d.tiles_x = r.tiles_x;
        for (uint32_t k = 0; k < take; k++) hi[n_items++] = WorkItem{n_reqs, r.tiles_issued + k};
        rays += (uint64_t)take * kServeTile * kServeTile;
        r.tiles_issued += take;
        n_reqs++;
        if (r.tiles_issued == r.tiles_total) {
          r.last_launch = l.id;
          l.finishing.push_back(r.ticket);
          it = s->active.erase(it);
        } else {
          ++it;
        }
      }
      hipStream_t st = s->streams[(size_t)slot];
      hipEventCreate(&l.start);
      hipEventCreate(&l.done);
      hipEventRecord(l.start, st);
      hipMemcpyAsync(s->d_reqs[(size_t)slot], hr, n_reqs * sizeof(ServeReqDev),
                     hipMemcpyHostToDevice, st);
      hipMemcpyAsync(s->d_items[(size_t)slot], hi, n_items * sizeof(WorkItem),
                     hipMemcpyHostToDevice, st);
      hipLaunchKernelGGL(serve_kernel, dim3(n_items), dim3(256), 0, st, s->d_reqs[(size_t)slot],
                         s->d_items[(size_t)slot]);
      hipError_t e = hipGetLastError();
      hipEventRecord(l.done, st);
      if (e != hipSuccess) {
        s->error = hipGetErrorString(e);
        for (uint64_t tk : l.finishing) s->reqs[tk].state = ReqState::Failed;
      }
      s->slot_busy[(size_t)slot] = true;
      s->stats.launches++;
      s->stats.rays += rays;
      s->inflight.push_back(std::move(l));
      continue;  // try to fill the next slot right away
    }
    // 4. nothing to issue: wait briefly for a submission or a completion
    if (s->inflight.empty())
      s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
    else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    }
  }
}
