// Host side of the raymarch server: tickets, admission by an HBM budget and
// by free entries of the device ring, publication in arrival order,
// completion from per-entry flags, take / release, statistics.
//
// This file holds no HIP: the device is reached through ServeDevice, which
// serve.hip implements over a persistent kernel and the CPU tests implement
// over a thread, so the whole scheduler (its locking included) is exercised
// under ThreadSanitizer and AddressSanitizer without a GPU.
//
// The role in the reference: FChunkManage::UpdateLoadingQueue admits chunk
// tasks in importance order while worker slots are free and re-queues what
// does not fit (Runtimes/Voxel/Chunk/ChunkManager.h:211-400, re-queue at
// :355-372).  Here requests are admitted in arrival order while their output
// bytes fit the budget and a ring entry is free; one that does not fit waits
// at the head of the queue until results are taken and their bytes freed.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "meso_hip.h"

namespace meso_serve {

using Clock = std::chrono::steady_clock;

// First-fit allocator over the output arena: byte offsets, 256-byte granules.
class Arena {
 public:
  void init(uint64_t size) {
    size_ = size;
    free_.clear();
    free_.emplace_back(0, size);
  }
  static uint64_t granule(uint64_t bytes) { return (bytes + 255) & ~255ull; }
  bool alloc(uint64_t bytes, uint64_t &off) {
    bytes = granule(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        it->first += bytes;
        it->second -= bytes;
        if (!it->second) free_.erase(it);
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = granule(bytes);
    auto it = free_.begin();
    while (it != free_.end() && it->first < off) ++it;
    it = free_.insert(it, {off, bytes});
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t free_bytes() const {
    uint64_t s = 0;
    for (const auto &f : free_) s += f.second;
    return s;
  }
  uint64_t size() const { return size_; }

 private:
  uint64_t size_ = 0;
  std::list<std::pair<uint64_t, uint64_t>> free_;  // (offset, bytes), sorted, coalesced
};

// The device side of the server.  Entries are numbered from 1 in publication
// order; entry `seq` lives in slot seq % ring_entries (a power of two), and
// the core reuses a slot only after its entry completed.
template <class Payload>
class ServeDevice {
 public:
  virtual ~ServeDevice() = default;
  // Make entry `seq` visible to the device: its payload, the arena offset of
  // its outputs and its work-item count.  Called in increasing seq order.
  virtual void publish(uint64_t seq, uint32_t slot, uint64_t arena_off, uint32_t items,
                       const Payload &p) = 0;
  // true once the device has completed entry `seq` (held in `slot`)
  virtual bool completed(uint32_t slot, uint64_t seq) = 0;
  // Called on every scheduler pass: keeps a worker running while entries
  // are outstanding.  false on a device error (see error()).
  virtual bool tick(bool outstanding) = 0;
  virtual std::string error() = 0;
  // device statistics: worker launches and their summed durations
  virtual uint64_t launches() = 0;
  virtual double busy_ms() = 0;
  // ends the worker and waits for it
  virtual void shutdown() = 0;
};

struct CoreConfig {
  uint64_t budget_bytes = 0;
  int32_t max_pending = 1;
  int32_t ring_entries = 1;  // a power of two
  // pause of the scheduler thread between passes while entries are on the device
  int32_t poll_us = 10;
};

enum class State { Pending, Admitted, Done, Failed, Taking };

template <class Payload>
struct Ticket {
  uint64_t id = 0;
  Payload payload{};
  uint64_t bytes = 0;
  uint32_t items = 0;  // work items of the device (the server's 8x8 packets)
  uint64_t pixels = 0;
  uint64_t seq = 0;
  uint32_t slot = 0;
  uint64_t arena_off = 0;
  State state = State::Pending;
  Clock::time_point t_submit, t_admit, t_done;
};

template <class Payload>
class ServeCore {
 public:
  ServeCore(const CoreConfig &cfg, ServeDevice<Payload> *dev) : cfg_(cfg), dev_(dev) {
    arena_.init(cfg.budget_bytes);
    slot_seq_.assign((size_t)cfg.ring_entries, 0);
    stats_.size = sizeof(meso_serve_stats);
  }
  ~ServeCore() { stop(); }
  ServeCore(const ServeCore &) = delete;
  ServeCore &operator=(const ServeCore &) = delete;

  void start() { thread_ = std::thread([this] { loop(); }); }

  // MESO_OK with *ticket set; MESO_ERR_BUSY when the pending queue is full;
  // MESO_ERR_RUNTIME after stop or a device error.
  meso_status submit(const Payload &p, uint64_t bytes, uint32_t items, uint64_t pixels,
                     uint64_t *ticket) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stop_ || failed_) return MESO_ERR_RUNTIME;
      if ((int64_t)pending_.size() >= cfg_.max_pending) {
        stats_.refused_busy++;
        return MESO_ERR_BUSY;
      }
      Ticket<Payload> t;
      t.id = next_ticket_++;
      t.payload = p;
      t.bytes = bytes;
      t.items = items;
      t.pixels = pixels;
      t.t_submit = Clock::now();
      *ticket = t.id;
      pending_.push_back(t.id);
      tickets_.emplace(t.id, t);
      stats_.submitted++;
    }
    cv_work_.notify_one();
    return MESO_OK;
  }

  // 1 done, -1 failed, 0 pending or running; false: unknown ticket
  bool poll(uint64_t ticket, int32_t *done) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = tickets_.find(ticket);
    if (it == tickets_.end()) return false;
    const State s = it->second.state;
    *done = (s == State::Done || s == State::Taking) ? 1 : (s == State::Failed ? -1 : 0);
    return true;
  }

  // MESO_OK when complete (latency_us: submit -> completion seen), BUSY on
  // timeout (timeout_us < 0 waits without limit), ARGUMENT_OUT_OF_RANGE for
  // an unknown ticket, RUNTIME_ERROR when the request failed or the server
  // stopped before completing it.
  meso_status wait(uint64_t ticket, int64_t timeout_us, double *latency_us) {
    std::unique_lock<std::mutex> lk(mu_);
    const auto deadline = Clock::now() + std::chrono::microseconds(timeout_us < 0 ? 0 : timeout_us);
    for (;;) {
      auto it = tickets_.find(ticket);
      if (it == tickets_.end()) return MESO_ERR_ARGUMENT_OUT_OF_RANGE;
      const Ticket<Payload> &t = it->second;
      if (t.state == State::Done || t.state == State::Taking) {
        if (latency_us)
          *latency_us = std::chrono::duration<double, std::micro>(t.t_done - t.t_submit).count();
        return MESO_OK;
      }
      if (t.state == State::Failed || loop_exited_) return MESO_ERR_RUNTIME;
      if (timeout_us >= 0) {
        if (cv_done_.wait_until(lk, deadline) == std::cv_status::timeout) {
          auto it2 = tickets_.find(ticket);
          if (it2 != tickets_.end() &&
              (it2->second.state == State::Done || it2->second.state == State::Taking))
            continue;
          return MESO_ERR_BUSY;
        }
      } else {
        cv_done_.wait(lk);
      }
    }
  }

  // First half of a take: a completed ticket is marked Taking and its record
  // copied out, so the caller can copy its outputs without the lock.
  // MESO_ERR_BUSY if not complete (or already being taken).
  meso_status begin_take(uint64_t ticket, Ticket<Payload> *out) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = tickets_.find(ticket);
    if (it == tickets_.end()) return MESO_ERR_ARGUMENT_OUT_OF_RANGE;
    if (it->second.state != State::Done) return MESO_ERR_BUSY;
    it->second.state = State::Taking;
    *out = it->second;
    return MESO_OK;
  }
  // Second half: frees the ticket's budget and forgets it (or puts it back
  // to Done when the copy failed, `ok` false).
  void end_take(uint64_t ticket, bool ok) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = tickets_.find(ticket);
      if (it == tickets_.end()) return;
      if (!ok) {
        it->second.state = State::Done;
        return;
      }
      arena_.release(it->second.arena_off, it->second.bytes);
      stats_.bytes_in_use -= Arena::granule(it->second.bytes);
      tickets_.erase(it);
    }
    cv_work_.notify_one();
  }

  meso_serve_stats stats() {
    std::lock_guard<std::mutex> lk(mu_);
    meso_serve_stats s = stats_;
    s.launches = dev_->launches();
    s.busy_ms = dev_->busy_ms();
    return s;
  }
  std::string error() {
    std::lock_guard<std::mutex> lk(mu_);
    return error_;
  }

  // Stops admission, lets the entries on the device complete, then ends the
  // scheduler thread and the device worker.  Requests still pending are
  // never run; waiting on them returns RUNTIME_ERROR.
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    if (thread_.joinable()) thread_.join();
    dev_->shutdown();
  }

 private:
  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      // 1. completions, in any order
      bool any = false;
      for (auto it = inflight_.begin(); it != inflight_.end();) {
        Ticket<Payload> &t = tickets_[*it];
        if (dev_->completed(t.slot, t.seq)) {
          t.state = State::Done;
          t.t_done = Clock::now();
          slot_seq_[t.slot] = 0;
          stats_.completed++;
          it = inflight_.erase(it);
          any = true;
        } else {
          ++it;
        }
      }
      if (any) cv_done_.notify_all();
      if (stop_ && (inflight_.empty() || failed_)) break;
      // 2. admission in arrival order: a free ring entry and room in the budget
      while (!stop_ && !failed_ && !pending_.empty()) {
        Ticket<Payload> &t = tickets_[pending_.front()];
        const uint32_t slot = (uint32_t)(next_seq_ & (uint64_t)(cfg_.ring_entries - 1));
        if (slot_seq_[slot] != 0) break;  // the entry using this slot is still on the device
        uint64_t off = 0;
        if (!arena_.alloc(t.bytes, off)) break;  // the budget binds: wait for takes
        t.arena_off = off;
        t.seq = next_seq_++;
        t.slot = slot;
        t.state = State::Admitted;
        t.t_admit = Clock::now();
        slot_seq_[slot] = t.seq + 1;
        const double q = std::chrono::duration<double, std::micro>(t.t_admit - t.t_submit).count();
        stats_.queue_us_sum += q;
        stats_.queue_us_max = std::max(stats_.queue_us_max, q);
        if (q > 1000.0) stats_.queued_over_1ms++;
        stats_.admitted++;
        stats_.bytes_in_use += Arena::granule(t.bytes);
        stats_.peak_bytes_in_use = std::max(stats_.peak_bytes_in_use, stats_.bytes_in_use);
        stats_.primary_rays += t.pixels;
        dev_->publish(t.seq, slot, off, t.items, t.payload);
        inflight_.push_back(t.id);
        pending_.pop_front();
      }
      if (!pending_.empty() && !stop_) stats_.admission_blocked_passes++;
      // 3. keep the device worker running while entries are outstanding
      if (!failed_ && !dev_->tick(!inflight_.empty())) {
        failed_ = true;
        error_ = dev_->error();
        for (uint64_t id : inflight_) tickets_[id].state = State::Failed;
        for (uint64_t id : pending_) tickets_[id].state = State::Failed;
        inflight_.clear();
        pending_.clear();
        cv_done_.notify_all();
      }
      // 4. wait for a submission, a take or a completion
      if (inflight_.empty()) {
        cv_work_.wait_for(lk, std::chrono::milliseconds(2));
      } else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(cfg_.poll_us));
        lk.lock();
      }
    }
    loop_exited_ = true;
    cv_done_.notify_all();
  }

  CoreConfig cfg_;
  ServeDevice<Payload> *dev_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  Arena arena_;
  std::deque<uint64_t> pending_;
  std::list<uint64_t> inflight_;
  std::unordered_map<uint64_t, Ticket<Payload>> tickets_;
  std::vector<uint64_t> slot_seq_;  // seq + 1 of the entry in each slot, 0 when free
  uint64_t next_ticket_ = 1, next_seq_ = 1;
  bool stop_ = false, failed_ = false, loop_exited_ = false;
  meso_serve_stats stats_{};
  std::string error_;
  std::thread thread_;
};

}  // namespace meso_serve
