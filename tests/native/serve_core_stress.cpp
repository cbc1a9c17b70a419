// Stress of the server's host core (mesoengine_amd/csrc/serve_core.h) with a
// fake device, built under ThreadSanitizer and under AddressSanitizer +
// UndefinedBehaviorSanitizer by tests/test_native_sanitizers.py.
//
// The fake device runs on its own thread like the persistent kernel: it
// takes published entries, writes each entry's tag over its bytes of the
// arena, sleeps a little and sets the slot's completion word.  Client
// threads submit requests of random sizes (retrying on BUSY), wait, check
// the tag in the arena during take, and release.  Checked: every request
// completes exactly once, a ring slot is never republished while its entry is
// on the device, the bytes held never exceed the budget, results are the
// submitter's own, and stop() with requests still queued returns cleanly.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <random>
#include <thread>
#include <vector>

#include "serve_core.h"

namespace {

struct Payload {
  uint32_t tag = 0;
};

struct Entry {
  uint64_t seq;
  uint32_t slot;
  uint64_t off;
  uint64_t bytes;
  uint32_t tag;
};

std::atomic<int> g_errors{0};
#define EXPECT(c)                                                             \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::fprintf(stderr, "serve_core_stress: %s failed at line %d\n", #c, __LINE__); \
      g_errors++;                                                             \
    }                                                                         \
  } while (0)

class FakeDevice final : public meso_serve::ServeDevice<Payload> {
 public:
  FakeDevice(uint32_t ring, uint64_t arena_bytes)
      : done_(ring), busy_(ring), arena_(arena_bytes, 0) {
    for (auto &d : done_) d.store(0);
    for (auto &b : busy_) b.store(0);
  }
  ~FakeDevice() override { shutdown(); }

  // bytes of a request are carried in `items` for this fake (items * 64)
  void publish(uint64_t seq, uint32_t slot, uint64_t off, uint32_t items,
               const Payload &p) override {
    EXPECT(busy_[slot].exchange(seq) == 0);  // slot free
    EXPECT(seq == last_seq_ + 1);             // in order
    last_seq_ = seq;
    std::lock_guard<std::mutex> lk(mu_);
    queue_.push_back(Entry{seq, slot, off, (uint64_t)items * 64, p.tag});
    cv_.notify_one();
  }
  bool completed(uint32_t slot, uint64_t seq) override {
    return done_[slot].load(std::memory_order_acquire) == seq;
  }
  bool tick(bool outstanding) override {
    if (outstanding && !running_) {
      running_ = true;
      launches_++;
      worker_ = std::thread([this] { run(); });
    }
    return true;
  }
  std::string error() override { return ""; }
  uint64_t launches() override { return launches_.load(); }
  double busy_ms() override { return 0.0; }
  void shutdown() override {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
  }
  // what a take reads: the arena bytes of a completed entry
  bool check(uint64_t off, uint64_t bytes, uint32_t tag) const {
    for (uint64_t i = 0; i < bytes; i++)
      if (arena_[off + i] != (uint8_t)(tag & 0xFF)) return false;
    return true;
  }

 private:
  void run() {
    std::mt19937 rng(7);
    for (;;) {
      Entry e;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        if (stop_ && queue_.empty()) return;
        e = queue_.front();
        queue_.pop_front();
      }
      std::memset(&arena_[e.off], (int)(e.tag & 0xFF), e.bytes);
      std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
      busy_[e.slot].store(0);
      done_[e.slot].store(e.seq, std::memory_order_release);
    }
  }
  std::vector<std::atomic<uint64_t>> done_, busy_;
  std::vector<uint8_t> arena_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Entry> queue_;
  bool stop_ = false;
  bool running_ = false;  // scheduler thread only
  uint64_t last_seq_ = 0;  // scheduler thread only
  std::atomic<uint64_t> launches_{0};
  std::thread worker_;
};

// A device that stops working: after `fail_after` publications, tick()
// reports an error (as a failed kernel launch would).
class FailingDevice final : public meso_serve::ServeDevice<Payload> {
 public:
  explicit FailingDevice(int fail_after) : fail_after_(fail_after) {}
  void publish(uint64_t, uint32_t, uint64_t, uint32_t, const Payload &) override { published_++; }
  bool completed(uint32_t, uint64_t) override { return false; }  // never completes
  bool tick(bool) override { return published_ < fail_after_; }
  std::string error() override { return "injected device failure"; }
  uint64_t launches() override { return 0; }
  double busy_ms() override { return 0.0; }
  void shutdown() override {}

 private:
  int fail_after_;
  int published_ = 0;  // scheduler thread only
};

void device_failure_phase() {
  FailingDevice dev(3);
  meso_serve::CoreConfig cfg;
  cfg.budget_bytes = 1 << 20;
  cfg.max_pending = 64;
  cfg.ring_entries = 4;
  auto core = std::make_unique<meso_serve::ServeCore<Payload>>(cfg, &dev);
  core->start();
  std::vector<uint64_t> tickets(10);
  Payload p;
  for (auto &t : tickets) EXPECT(core->submit(p, 1024, 16, 16, &t) == MESO_OK);
  // every request fails: the three published and the ones still queued
  for (uint64_t t : tickets) EXPECT(core->wait(t, 5000000, nullptr) == MESO_ERR_RUNTIME);
  int32_t done = 0;
  EXPECT(core->poll(tickets[0], &done) && done == -1);
  uint64_t t2 = 0;
  EXPECT(core->submit(p, 1024, 16, 16, &t2) == MESO_ERR_RUNTIME);
  EXPECT(core->error() == "injected device failure");
  core->stop();
}

}  // namespace

int main() {
  device_failure_phase();
  const uint64_t budget = 1 << 20;
  const uint32_t ring = 8;
  FakeDevice dev(ring, budget);
  meso_serve::CoreConfig cfg;
  cfg.budget_bytes = budget;
  cfg.max_pending = 2;  // fewer than the clients: BUSY refusals happen
  cfg.ring_entries = (int32_t)ring;
  cfg.poll_us = 5;
  auto core = std::make_unique<meso_serve::ServeCore<Payload>>(cfg, &dev);
  core->start();

#ifndef STRESS_PER_CLIENT
#define STRESS_PER_CLIENT 150
#endif
  const int clients = 4, per_client = STRESS_PER_CLIENT;
  std::atomic<uint64_t> busy_refusals{0}, completed{0};
  std::vector<std::thread> th;
  for (int c = 0; c < clients; c++) {
    th.emplace_back([&, c] {
      std::mt19937 rng(100 + c);
      for (int i = 0; i < per_client; i++) {
        const uint32_t tiles = 1 + rng() % 4000;  // up to 256 KB of a 1 MB budget
        Payload p;
        p.tag = (uint32_t)(c * 1000 + i);
        uint64_t ticket = 0;
        meso_status st;
        while ((st = core->submit(p, (uint64_t)tiles * 64, tiles, tiles, &ticket)) == MESO_ERR_BUSY) {
          busy_refusals++;
          std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        EXPECT(st == MESO_OK);
        double lat = 0.0;
        EXPECT(core->wait(ticket, 20000000, &lat) == MESO_OK);
        EXPECT(lat >= 0.0);
        int32_t done = 0;
        EXPECT(core->poll(ticket, &done) && done == 1);
        meso_serve::Ticket<Payload> t;
        EXPECT(core->begin_take(ticket, &t) == MESO_OK);
        EXPECT(core->begin_take(ticket, &t) == MESO_ERR_BUSY);  // one take at a time
        EXPECT(t.payload.tag == p.tag);
        EXPECT(dev.check(t.arena_off, t.bytes, p.tag));
        core->end_take(ticket, true);
        EXPECT(core->begin_take(ticket, &t) == MESO_ERR_ARGUMENT_OUT_OF_RANGE);  // forgotten
        completed++;
      }
    });
  }
  // a reader of the statistics while the clients run
  std::atomic<bool> stop_reader{false};
  std::thread reader([&] {
    while (!stop_reader.load()) {
      const meso_serve_stats s = core->stats();
      EXPECT(s.peak_bytes_in_use <= budget);
      EXPECT(s.bytes_in_use <= budget);
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
  });
  for (auto &t : th) t.join();
  stop_reader = true;
  reader.join();
  const meso_serve_stats s = core->stats();
  EXPECT(completed.load() == (uint64_t)clients * per_client);
  EXPECT(s.completed == (uint64_t)clients * per_client);
  EXPECT(s.admitted == s.completed && s.bytes_in_use == 0);
  EXPECT(s.peak_bytes_in_use <= budget);
  EXPECT(s.refused_busy == busy_refusals.load());
  EXPECT(s.queue_us_max >= 0.0);

  // stop with requests that can never be admitted: larger than the free
  // budget while an earlier result is held
  uint64_t held = 0, starved = 0;
  Payload p;
  EXPECT(core->submit(p, budget - 4096, (uint32_t)((budget - 4096) / 64), 1, &held) == MESO_OK);
  EXPECT(core->wait(held, 20000000, nullptr) == MESO_OK);
  EXPECT(core->submit(p, 8192, 128, 1, &starved) == MESO_OK);
  EXPECT(core->wait(starved, 20000, nullptr) == MESO_ERR_BUSY);  // times out: budget binds
  core->stop();
  EXPECT(core->wait(starved, 1000, nullptr) == MESO_ERR_RUNTIME);
  EXPECT(core->submit(p, 64, 1, 1, &starved) == MESO_ERR_RUNTIME);
  core.reset();
  if (g_errors.load()) {
    std::fprintf(stderr, "serve_core_stress: %d failures\n", g_errors.load());
    return 1;
  }
  std::printf("serve_core_stress ok: %llu requests, %llu BUSY retries, peak %llu of %llu bytes\n",
              (unsigned long long)s.completed, (unsigned long long)busy_refusals.load(),
              (unsigned long long)s.peak_bytes_in_use, (unsigned long long)budget);
  return 0;
}
