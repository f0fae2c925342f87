// Host side of a batch on a pool of host threads: validation, relabelling
// and feed order of the requests before a launch, and decoding of the
// results after it.  No device code here.
//
// Feed order (the kernel's claim order within each feed shard): requests
// without slack first, then by move bound (deepest first), then by symbols
// (most first), stable in the caller's order; the sorted list is then dealt
// round-robin to the feed shards (init_feed's contiguous ranges), so every
// shard runs its own longest-expected-first sequence.  The order is a
// counting sort over (no slack, bound, symbols): 2 x 9 x 15 buckets.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/td16_abi.h"
#include "td16_internal.h"

namespace td16 {

namespace {

// ---------------------------------------------------------------------------
// Worker pool: a fixed set of threads that take parts of one parallel loop
// at a time.  The calling thread works too, so `parts` = 1 runs inline.
// ---------------------------------------------------------------------------

class Pool {
 public:
  static Pool &get() {
    static Pool *p = new Pool();  // never destroyed: workers outlive static teardown
    return *p;
  }

  void run(int parts, const std::function<void(int)> &fn) {
    if (parts <= 1) {
      fn(0);
      return;
    }
    std::lock_guard<std::mutex> one(runMu_);
    uint64_t gen;
    {
      std::lock_guard<std::mutex> lk(mu_);
      while ((int)workers_.size() < parts - 1 && (int)workers_.size() < kMaxWorkers)
        workers_.emplace_back([this] { loop(); });
      gen = ++gen_;
      fn_ = &fn;
      parts_ = parts;
      finished_ = 0;
      // the part counter carries the generation, so a worker still holding
      // an earlier loop's function can never claim a part of this one
      next_.store(gen << 32, std::memory_order_release);
    }
    cv_.notify_all();
    work(gen, &fn, parts);
    std::unique_lock<std::mutex> lk(mu_);
    cvDone_.wait(lk, [&] { return finished_ == parts; });
    fn_ = nullptr;
  }

 private:
  static constexpr int kMaxWorkers = 63;

  void work(uint64_t gen, const std::function<void(int)> *fn, int parts) {
    int done = 0;
    for (;;) {
      uint64_t v = next_.load(std::memory_order_acquire);
      bool claimed = false;
      while ((v >> 32) == gen && (int)(uint32_t)v < parts) {
        if (next_.compare_exchange_weak(v, v + 1, std::memory_order_acq_rel)) {
          claimed = true;
          break;
        }
      }
      if (!claimed) break;
      (*fn)((int)(uint32_t)v);
      ++done;
    }
    if (done) {
      std::lock_guard<std::mutex> lk(mu_);
      finished_ += done;
      if (finished_ == parts_) cvDone_.notify_all();
    }
  }

  void loop() {
    uint64_t seen = 0;
    for (;;) {
      uint64_t gen;
      const std::function<void(int)> *fn;
      int parts;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen = gen_;
        fn = fn_;
        parts = parts_;
      }
      if (fn != nullptr) work(gen, fn, parts);
    }
  }

  std::mutex runMu_, mu_;
  std::condition_variable cv_, cvDone_;
  std::vector<std::thread> workers_;
  const std::function<void(int)> *fn_ = nullptr;
  int parts_ = 0;
  int finished_ = 0;
  uint64_t gen_ = 0;
  std::atomic<uint64_t> next_{0};
};

constexpr int kBuckets = 2 * kMaxDepth * 15;  // (no slack) x (bound 1..9) x (N 3..17)
constexpr uint16_t kNoJob = 0xFFFF;

inline uint16_t bucket_of(const Job &j) {
  const int slack = (j.flags & kFlagTight) ? 0 : 1;
  return (uint16_t)(slack * (kMaxDepth * 15) + (kMaxDepth - j.maxDepth) * 15 + (kMaxN - j.N));
}

}  // namespace

void decode_result(const DevResult &r, td16_result *out) {
  std::memset(out, 0, sizeof(*out));
  out->status = TD16_OK;
  out->found = r.found;
  out->depth = r.found ? r.depth : 0;
  out->split = r.split != 0 ? 1 : 0;
  out->expanded = r.expanded;
  out->solutions = r.solutions;
  out->device_us = (r.t1 >= r.t0) ? (double)(r.t1 - r.t0) * 0.01 : 0.0;  // 100 MHz ticks
  unpack_moves(r.moves, out->depth, out->moves);
}

void trivial_result(int trivial, td16_result *out) {
  std::memset(out, 0, sizeof(*out));
  out->status = TD16_OK;
  if (trivial == 2) out->found = 1;  // IDDFS on a sorted input: distance 0
}

int host_threads(int want) {
  if (want > 0) return std::min(want, 64);
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
}

void parallel_parts(int parts, const std::function<void(int)> &fn) { Pool::get().run(parts, fn); }

uint32_t deal_position(uint32_t r, uint32_t nj, uint32_t per) {
  // shards 0..F-1 hold `per` requests, shard F (when rem > 0) holds rem:
  // round t < rem deals to F + 1 shards, later rounds to F
  const uint32_t F = nj / per, rem = nj % per;
  const uint32_t wide = F + (rem ? 1u : 0u);
  if (r < rem * wide) return (r % wide) * per + r / wide;
  const uint32_t r2 = r - rem * wide;
  return (r2 % F) * per + rem + r2 / F;
}

int prepare_batch_into(const td16_request *reqs, int count, td16_result *results, Job *jobsOut,
                       int *whereOut, uint32_t *njOut, PrepareScratch &ws, int threads,
                       char *msg, int cap) {
  *njOut = 0;
  if (count <= 0) return TD16_OK;
  const int parts = std::max(1, std::min(host_threads(threads), (count + 4095) / 4096));
  ws.jobs.resize((size_t)count);
  ws.keys.resize((size_t)count);
  ws.pos.resize((size_t)count);
  ws.counts.assign((size_t)parts * kBuckets, 0u);
  if ((int)ws.memos.size() < parts) ws.memos.resize((size_t)parts);
  std::atomic<int> firstBad{count};
  auto range = [&](int p, int &lo, int &hi) {
    lo = (int)((int64_t)count * p / parts);
    hi = (int)((int64_t)count * (p + 1) / parts);
  };
  // 1: every request to its label form, and its bucket
  parallel_parts(parts, [&](int p) {
    int lo, hi;
    range(p, lo, hi);
    uint32_t *cnt = &ws.counts[(size_t)p * kBuckets];
    RateMemo &memo = ws.memos[(size_t)p];
    for (int q = lo; q < hi; ++q) {
      int trivial = 0;
      Job &job = ws.jobs[(size_t)q];
      if (make_job(&reqs[q], &job, &trivial, memo, nullptr, 0) != TD16_OK) {
        int cur = firstBad.load(std::memory_order_relaxed);
        while (q < cur && !firstBad.compare_exchange_weak(cur, q)) {
        }
        ws.keys[(size_t)q] = kNoJob;
        continue;
      }
      if (trivial) {
        trivial_result(trivial, &results[q]);
        ws.keys[(size_t)q] = kNoJob;
        continue;
      }
      const uint16_t b = bucket_of(job);
      ws.keys[(size_t)q] = b;
      cnt[b]++;
    }
  });
  const int bad = firstBad.load();
  if (bad < count) {
    Job job;
    int trivial = 0;
    const int st = make_job_checked(&reqs[bad], &job, &trivial, msg, cap);
    std::memset(&results[bad], 0, sizeof(td16_result));
    results[bad].status = st;
    return st;
  }
  // 2: bucket starts per part (parts in request order keep the sort stable)
  uint32_t nj = 0;
  for (int b = 0; b < kBuckets; ++b)
    for (int p = 0; p < parts; ++p) {
      const uint32_t c = ws.counts[(size_t)p * kBuckets + b];
      ws.counts[(size_t)p * kBuckets + b] = nj;
      nj += c;
    }
  *njOut = nj;
  if (nj == 0) {
    std::fill(ws.pos.begin(), ws.pos.end(), kNotLaunched);
    return TD16_OK;
  }
  const uint32_t per = (nj + kFeedShards - 1) / kFeedShards;
  // 3: each request to its dealt position, then its record there (the
  // positions are scattered: the destination lines are fetched ahead)
  parallel_parts(parts, [&](int p) {
    int lo, hi;
    range(p, lo, hi);
    uint32_t *start = &ws.counts[(size_t)p * kBuckets];
    for (int q = lo; q < hi; ++q) {
      const uint16_t b = ws.keys[(size_t)q];
      ws.pos[(size_t)q] = b == kNoJob ? kNotLaunched : deal_position(start[b]++, nj, per);
    }
    constexpr int kAhead = 16;
    for (int q = lo; q < hi; ++q) {
      if (q + kAhead < hi && ws.pos[(size_t)(q + kAhead)] != kNotLaunched) {
        const char *ahead = reinterpret_cast<const char *>(&jobsOut[ws.pos[(size_t)(q + kAhead)]]);
        __builtin_prefetch(ahead, 1);
        __builtin_prefetch(ahead + 64, 1);
      }
      const uint32_t pos = ws.pos[(size_t)q];
      if (pos == kNotLaunched) continue;
      jobsOut[pos] = ws.jobs[(size_t)q];
      jobsOut[pos].tag = (uint32_t)q;
      whereOut[pos] = q;
    }
  });
  return TD16_OK;
}

int prepare_batch(const td16_request *reqs, int count, td16_result *results,
                  std::vector<Job> &jobs, std::vector<int> &where, char *msg, int cap) {
  PrepareScratch ws;
  jobs.resize((size_t)std::max(count, 0));
  where.resize((size_t)std::max(count, 0));
  uint32_t nj = 0;
  const int st = prepare_batch_into(reqs, count, results, jobs.data(), where.data(), &nj, ws, 0,
                                    msg, cap);
  jobs.resize(st == TD16_OK ? nj : 0);
  where.resize(st == TD16_OK ? nj : 0);
  return st;
}

void decode_batch(const DevResult *dev, const uint32_t *pos, uint32_t count,
                  td16_result *results, int threads, uint64_t *expanded, uint64_t *solutions,
                  uint64_t *donations) {
  const int parts = std::max(1, std::min(host_threads(threads), (int)((count + 8191) / 8192)));
  std::vector<uint64_t> sums((size_t)parts * 3, 0);
  parallel_parts(parts, [&](int p) {
    const uint32_t lo = (uint32_t)((uint64_t)count * p / parts);
    const uint32_t hi = (uint32_t)((uint64_t)count * (p + 1) / parts);
    uint64_t e = 0, s = 0, d = 0;
    // the records are read in the caller's order, scattered over the feed
    // positions: fetching a few ahead keeps several misses in flight
    constexpr uint32_t kAhead = 16;
    for (uint32_t q = lo; q < hi; ++q) {
      if (q + kAhead < hi && pos[q + kAhead] != kNotLaunched) {
        const char *ahead = reinterpret_cast<const char *>(&dev[pos[q + kAhead]]);
        __builtin_prefetch(ahead);
        __builtin_prefetch(ahead + sizeof(DevResult) - 1);
      }
      if (pos[q] == kNotLaunched) continue;
      const DevResult &r = dev[pos[q]];
      decode_result(r, &results[q]);
      e += r.expanded;
      s += r.solutions;
      d += (uint64_t)r.split;
    }
    sums[(size_t)p * 3] = e;
    sums[(size_t)p * 3 + 1] = s;
    sums[(size_t)p * 3 + 2] = d;
  });
  *expanded = *solutions = *donations = 0;
  for (int p = 0; p < parts; ++p) {
    *expanded += sums[(size_t)p * 3];
    *solutions += sums[(size_t)p * 3 + 1];
    *donations += sums[(size_t)p * 3 + 2];
  }
}

void positions_of(const int *where, uint32_t nj, uint32_t count, std::vector<uint32_t> &pos) {
  pos.assign(count, kNotLaunched);
  for (uint32_t q = 0; q < nj; ++q) pos[(size_t)where[q]] = q;
}

}  // namespace td16
