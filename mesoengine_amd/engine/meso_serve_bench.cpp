// meso_serve_bench: replays a raymarch request stream against libmeso_hip
// without an interpreter in the loop, on the plain path (one meso_raymarch
// per request, first come first served) and on the continuous-batching
// server (meso_serve_*), and prints one JSON line with p50 / p99 latency
// (arrival -> result ready) and Mrays/s per path.
//
// The request file (little-endian) is written by the Python workload code:
//   "MESOREQS" u32 version=1
//   i32 n_scenes, i32 n_requests
//   per scene:   i32 edge, i32 generator, i32 layout, u64 seed
//   per request: f64 arrival_s, i32 scene, i32 width, i32 height,
//                f32 eye[3], target[3], up[3], vfov
// Every request asks for primary + shadow rays toward light (0.4, 1, 0.3);
// the client takes the face ids of each result (1 byte per pixel).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "meso_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct SceneSpec {
  int32_t edge, generator, layout;
  uint64_t seed;
};

struct Request {
  double arrival;
  int32_t scene, width, height;
  float eye[3], target[3], up[3], vfov;
};

bool read_requests(const char *path, std::vector<SceneSpec> &scenes, std::vector<Request> &reqs) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  char magic[8];
  uint32_t version = 0;
  int32_t ns = 0, nr = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "MESOREQS", 8) == 0 &&
            std::fread(&version, 4, 1, f) == 1 && version == 1 && std::fread(&ns, 4, 1, f) == 1 &&
            std::fread(&nr, 4, 1, f) == 1 && ns > 0 && ns < 4096 && nr > 0 && nr < (1 << 24);
  for (int32_t i = 0; ok && i < ns; i++) {
    SceneSpec s{};
    ok = std::fread(&s.edge, 4, 1, f) == 1 && std::fread(&s.generator, 4, 1, f) == 1 &&
         std::fread(&s.layout, 4, 1, f) == 1 && std::fread(&s.seed, 8, 1, f) == 1;
    scenes.push_back(s);
  }
  for (int32_t i = 0; ok && i < nr; i++) {
    Request r{};
    ok = std::fread(&r.arrival, 8, 1, f) == 1 && std::fread(&r.scene, 4, 1, f) == 1 &&
         std::fread(&r.width, 4, 1, f) == 1 && std::fread(&r.height, 4, 1, f) == 1 &&
         std::fread(r.eye, 4, 3, f) == 3 && std::fread(r.target, 4, 3, f) == 3 &&
         std::fread(r.up, 4, 3, f) == 3 && std::fread(&r.vfov, 4, 1, f) == 1 && r.scene >= 0 &&
         r.scene < ns && r.width > 0 && r.height > 0 && r.width <= 8192 && r.height <= 8192;
    reqs.push_back(r);
  }
  std::fclose(f);
  return ok;
}

meso_camera camera_of(const Request &r) {
  meso_camera c{};
  c.size = sizeof(c);
  std::memcpy(c.eye, r.eye, sizeof(c.eye));
  std::memcpy(c.target, r.target, sizeof(c.target));
  std::memcpy(c.up, r.up, sizeof(c.up));
  c.vfov_deg = r.vfov;
  c.width = r.width;
  c.height = r.height;
  return c;
}

// sleep until shortly before t, then spin
void wait_until(Clock::time_point t) {
  for (;;) {
    const auto left = t - Clock::now();
    if (left <= Clock::duration::zero()) return;
    if (left > std::chrono::microseconds(300)) std::this_thread::sleep_for(left - std::chrono::microseconds(200));
  }
}

double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q / 100.0 * (double)(v.size() - 1);
  const size_t lo = (size_t)pos;
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

#define CHECK(call)                                                                   \
  do {                                                                                \
    const meso_status st_ = (call);                                                   \
    if (st_ != MESO_OK) {                                                             \
      std::fprintf(stderr, "meso_serve_bench: %s -> %d: %s\n", #call, (int)st_,       \
                   meso_last_error_message());                                        \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

struct Result {
  std::vector<double> lat_ms;
  double makespan_s = 0.0;
  uint64_t launches = 0;
  meso_serve_stats stats{};  // served path only
};

// plain path: one raymarch per request on the context stream, in arrival order
Result run_plain(meso_context *ctx, std::vector<meso_scene *> &scenes, const std::vector<Request> &reqs,
                 bool timed, const meso_shading &sh) {
  uint64_t maxpx = 0;
  for (const Request &r : reqs) maxpx = std::max<uint64_t>(maxpx, (uint64_t)r.width * r.height);
  meso_buffer *idx, *face, *t, *shade;
  CHECK(meso_buffer_create(ctx, maxpx * 4, nullptr, &idx));
  CHECK(meso_buffer_create(ctx, maxpx, nullptr, &face));
  CHECK(meso_buffer_create(ctx, maxpx * 4, nullptr, &t));
  CHECK(meso_buffer_create(ctx, maxpx * 4, nullptr, &shade));
  meso_ray_outputs out{sizeof(out), (void *)meso_buffer_device_address(idx, 0), 4,
                       (uint8_t *)meso_buffer_device_address(face, 0),
                       (float *)meso_buffer_device_address(t, 0),
                       (float *)meso_buffer_device_address(shade, 0), nullptr};
  std::vector<uint8_t> host(maxpx);
  Result res;
  const auto start = Clock::now();
  for (const Request &r : reqs) {
    const auto arrival = start + std::chrono::duration_cast<Clock::duration>(
                                     std::chrono::duration<double>(timed ? r.arrival : 0.0));
    wait_until(arrival);
    const meso_camera cam = camera_of(r);
    CHECK(meso_raymarch(ctx, scenes[r.scene], &cam, MESO_RAY_SHADOW, &sh, &out, nullptr));
    CHECK(meso_stream_sync(ctx, nullptr));
    res.lat_ms.push_back(std::chrono::duration<double, std::milli>(Clock::now() - arrival).count());
    CHECK(meso_buffer_download(face, host.data(), (uint64_t)r.width * r.height, 0));
    res.launches++;
  }
  res.makespan_s = std::chrono::duration<double>(Clock::now() - start).count();
  for (meso_buffer *b : {idx, face, t, shade}) meso_buffer_destroy(b);
  return res;
}

// served path: a submitter (this thread) and a consumer thread
Result run_served(meso_context *ctx, std::vector<meso_scene *> &scenes, const std::vector<Request> &reqs,
                  bool timed, const meso_shading &sh, const meso_serve_config &cfg) {
  meso_server *srv = nullptr;
  CHECK(meso_serve_start(ctx, &cfg, &srv));
  const size_t n = reqs.size();
  std::vector<uint64_t> tickets(n);
  std::vector<double> lag_ms(n), lat_ms(n);
  std::mutex mu;
  std::condition_variable cv;
  size_t submitted = 0;
  std::atomic<bool> failed{false};
  std::thread consumer([&] {
    uint64_t maxpx = 0;
    for (const Request &r : reqs) maxpx = std::max<uint64_t>(maxpx, (uint64_t)r.width * r.height);
    std::vector<uint8_t> face(maxpx);
    for (size_t i = 0; i < n; i++) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return submitted > i || failed.load(); });
        if (failed.load()) return;
      }
      double us = 0.0;
      if (meso_serve_wait(srv, tickets[i], 60000000, &us) != MESO_OK) {
        failed = true;
        return;
      }
      lat_ms[i] = us * 1e-3 + lag_ms[i];
      meso_host_outputs dst{sizeof(dst), 4, nullptr, face.data(), nullptr, nullptr};
      if (meso_serve_take(srv, tickets[i], &dst) != MESO_OK) {
        failed = true;
        return;
      }
    }
  });
  const auto start = Clock::now();
  for (size_t i = 0; i < n && !failed.load(); i++) {
    const Request &r = reqs[i];
    const auto arrival = start + std::chrono::duration_cast<Clock::duration>(
                                     std::chrono::duration<double>(timed ? r.arrival : 0.0));
    wait_until(arrival);
    const meso_camera cam = camera_of(r);
    uint64_t tk = 0;
    const meso_status st = meso_serve_submit(srv, scenes[r.scene], &cam, MESO_RAY_SHADOW, &sh, &tk);
    lag_ms[i] = std::chrono::duration<double, std::milli>(Clock::now() - arrival).count();
    if (st != MESO_OK) {
      std::fprintf(stderr, "meso_serve_bench: submit %zu -> %d: %s\n", i, (int)st,
                   meso_last_error_message());
      failed = true;
    } else {
      std::lock_guard<std::mutex> lk(mu);
      tickets[i] = tk;
      submitted = i + 1;
    }
    cv.notify_one();
  }
  consumer.join();
  Result res;
  res.makespan_s = std::chrono::duration<double>(Clock::now() - start).count();
  meso_serve_stats stats{};
  stats.size = sizeof(stats);
  CHECK(meso_serve_stats_get(srv, &stats));
  CHECK(meso_serve_stop(srv));
  if (failed.load()) {
    std::fprintf(stderr, "meso_serve_bench: served run failed: %s\n", meso_last_error_message());
    std::exit(1);
  }
  res.lat_ms = lat_ms;
  res.launches = stats.launches;
  res.stats = stats;
  return res;
}

int usage() {
  std::fprintf(stderr,
               "usage: meso_serve_bench REQUESTS [--budget-mb M] [--inflight K]\n"
               "  [--wg-per-cu W] [--capacity] [--device I]\n");
  return 2;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return usage();
  uint64_t budget_mb = 512;
  int32_t inflight = 64, wg_per_cu = 0, device = 0;
  bool capacity = false;
  for (int i = 2; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--capacity") {
      capacity = true;
      continue;
    }
    if (i + 1 >= argc) return usage();
    const char *v = argv[++i];
    if (a == "--budget-mb") budget_mb = std::strtoull(v, nullptr, 10);
    else if (a == "--inflight") inflight = std::atoi(v);
    else if (a == "--wg-per-cu") wg_per_cu = std::atoi(v);
    else if (a == "--device") device = std::atoi(v);
    else return usage();
  }
  std::vector<SceneSpec> specs;
  std::vector<Request> reqs;
  if (!read_requests(argv[1], specs, reqs)) {
    std::fprintf(stderr, "meso_serve_bench: cannot read %s\n", argv[1]);
    return 1;
  }
  meso_context *ctx = nullptr;
  CHECK(meso_context_create(device, &ctx));
  std::vector<meso_scene *> scenes;
  uint64_t resident = 0;
  for (const SceneSpec &s : specs) {
    meso_generate_desc g{sizeof(g), s.generator, s.layout, s.edge, s.seed, 0};
    meso_scene *sc = nullptr;
    CHECK(meso_scene_generate(ctx, &g, &sc));
    meso_scene_info info{};
    info.size = sizeof(info);
    CHECK(meso_scene_get_info(sc, &info));
    resident += info.device_bytes;
    scenes.push_back(sc);
  }
  meso_shading sh{sizeof(sh), {0.4f, 1.0f, 0.3f}, 0, 0, 16.0f};

  // rays per request (primary + shadow), counted once
  uint64_t rays = 0;
  {
    meso_buffer *cnt, *idx;
    uint64_t maxpx = 0;
    for (const Request &r : reqs) maxpx = std::max<uint64_t>(maxpx, (uint64_t)r.width * r.height);
    CHECK(meso_buffer_create(ctx, 8 * MESO_CNT_COUNT, nullptr, &cnt));
    CHECK(meso_buffer_create(ctx, maxpx * 4, nullptr, &idx));
    CHECK(meso_buffer_fill_zero(cnt, nullptr));
    meso_ray_outputs out{sizeof(out), (void *)meso_buffer_device_address(idx, 0), 4, nullptr,
                         nullptr, nullptr, (uint64_t *)meso_buffer_device_address(cnt, 0)};
    for (const Request &r : reqs) {
      const meso_camera cam = camera_of(r);
      CHECK(meso_raymarch(ctx, scenes[r.scene], &cam, MESO_RAY_SHADOW | MESO_RAY_COUNT, &sh, &out,
                          nullptr));
    }
    CHECK(meso_stream_sync(ctx, nullptr));
    uint64_t c[MESO_CNT_COUNT];
    CHECK(meso_buffer_download(cnt, c, sizeof(c), 0));
    rays = c[MESO_CNT_RAYS] + c[MESO_CNT_SHADOW_RAYS];
    meso_buffer_destroy(cnt);
    meso_buffer_destroy(idx);
  }

  meso_serve_config cfg{};
  cfg.size = sizeof(cfg);
  cfg.hbm_budget_bytes = budget_mb << 20;
  cfg.max_pending = (int32_t)reqs.size() + 8;
  cfg.max_inflight = inflight;
  cfg.workgroups_per_cu = wg_per_cu;
  cfg.tile = 8;
  const bool timed = !capacity;
  // warm both paths once on the first few requests
  {
    std::vector<Request> warm(reqs.begin(), reqs.begin() + std::min<size_t>(reqs.size(), 16));
    run_plain(ctx, scenes, warm, false, sh);
    run_served(ctx, scenes, warm, false, sh, cfg);
  }
  const Result plain = run_plain(ctx, scenes, reqs, timed, sh);
  const Result served = run_served(ctx, scenes, reqs, timed, sh, cfg);
  const double span = reqs.back().arrival;
  std::printf("{\"requests\": %zu, \"scenes\": %zu, \"resident_mb\": %.1f, \"rays\": %llu, "
              "\"mode\": \"%s\", \"offered_qps\": %.1f, \"budget_mb\": %llu, "
              "\"plain\": {\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"makespan_s\": %.4f, "
              "\"mrays\": %.1f}, "
              "\"served\": {\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"makespan_s\": %.4f, "
              "\"mrays\": %.1f, \"launches\": %llu, \"peak_budget_mb\": %.1f, "
              "\"queue_ms_mean\": %.3f, \"queue_ms_max\": %.3f, \"queued_over_1ms\": %llu}}\n",
              reqs.size(), scenes.size(), resident / 1e6, (unsigned long long)rays,
              capacity ? "capacity" : "timed", span > 0 ? (double)reqs.size() / span : 0.0,
              (unsigned long long)budget_mb, pct(plain.lat_ms, 50), pct(plain.lat_ms, 99),
              pct(plain.lat_ms, 100), plain.makespan_s, rays / plain.makespan_s / 1e6,
              pct(served.lat_ms, 50), pct(served.lat_ms, 99), pct(served.lat_ms, 100),
              served.makespan_s, rays / served.makespan_s / 1e6,
              (unsigned long long)served.launches, served.stats.peak_bytes_in_use / 1048576.0,
              served.stats.admitted ? served.stats.queue_us_sum / served.stats.admitted * 1e-3 : 0.0,
              served.stats.queue_us_max * 1e-3, (unsigned long long)served.stats.queued_over_1ms);
  for (meso_scene *s : scenes) meso_scene_destroy(s);
  meso_context_destroy(ctx);
  return 0;
}
