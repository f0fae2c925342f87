// Host runtime of libtd16: devices, slots, the blocking one-request-per-launch
// search, the batched persistent launch and the served loop.
//
// Reference counterparts (emitter/src/native/): slot_manager.cu (streams,
// events, device checks), sorting_search.cu:714-972 (slot buffers, queue
// sizing, performGpuSortingSearch), core/cuda_error.cuh (error classes),
// core/cuda_memory.cuh (owned device/pinned buffers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/td16_abi.h"
#include "td16_internal.h"

namespace td16 {

int set_msg(char *msg, int cap, int status, const char *fmt, ...);
int make_job_checked(const td16_request *r, Job *job, int *trivial, char *msg, int cap);

namespace {

#define TD16_HIP(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_msg(msg, cap,                                                              \
                     (e_ == hipErrorOutOfMemory) ? TD16_OUT_OF_MEMORY : TD16_DEVICE_ERROR, \
                     "%s:%d: HIP error %s (%s)", __FILE__, __LINE__, hipGetErrorName(e_),   \
                     hipGetErrorString(e_));                                                \
  } while (0)

// ---------------------------------------------------------------------------
// Devices
// ---------------------------------------------------------------------------

struct DeviceInfo {
  int device = -1;
  int cus = 0;
  int blocksPerCU = 0;      // resident blocks of the 5-waves-per-SIMD kernel
  int blocksPerCUWide = 0;  // of the 6-waves-per-SIMD build
  uint16_t *dLex = nullptr;
};

std::mutex gDevMu;
std::vector<DeviceInfo> gDevs;

int device_info(int device, DeviceInfo *out, char *msg, int cap) {
  std::lock_guard<std::mutex> lk(gDevMu);
  for (auto &d : gDevs)
    if (d.device == device) {
      *out = d;
      return TD16_OK;
    }
  int count = 0;
  TD16_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "devices contains a device ordinal outside the available range");
  TD16_HIP(hipSetDevice(device));
  DeviceInfo d;
  d.device = device;
  TD16_HIP(hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, device));
  TD16_HIP((hipError_t)search_kernel_occupancy(&d.blocksPerCU, &d.blocksPerCUWide));
  if (d.blocksPerCU < 1) d.blocksPerCU = 1;
  std::vector<uint16_t> lex(kLexEntries);
  build_lex_table(lex.data());
  TD16_HIP(hipMalloc(&d.dLex, sizeof(uint16_t) * kLexEntries));
  TD16_HIP(hipMemcpy(d.dLex, lex.data(), sizeof(uint16_t) * kLexEntries, hipMemcpyHostToDevice));
  gDevs.push_back(d);
  *out = d;
  return TD16_OK;
}

// ---------------------------------------------------------------------------
// Buffers of one launch context (owned; freed in release())
// ---------------------------------------------------------------------------

struct Buffers {
  int device = -1;
  uint32_t jobsCap = 0;
  Job *dJobs = nullptr;
  JobState *dStates = nullptr;
  Control *dCtrl = nullptr;
  QueueState *dQueue = nullptr;
  Item *dItems = nullptr;
  uint64_t *dSeq = nullptr;
  uint64_t qcap = 0;
  Control *hCtrl = nullptr;      // pinned
  Job *hJobs = nullptr;          // pinned
  // results live in coherent pinned host memory: the kernel writes each one
  // over the bus as the request finishes, so no copy follows the launch
  DevResult *hResults = nullptr;
  DevResult *dResults = nullptr; // device view of hResults

  void release() {
    if (device >= 0) hipSetDevice(device);
    hipFree(dJobs);
    hipFree(dStates);
    hipFree(dCtrl);
    hipFree(dQueue);
    hipFree(dItems);
    hipFree(dSeq);
    hipHostFree(hCtrl);
    hipHostFree(hJobs);
    hipHostFree(hResults);
    *this = Buffers();
  }
};

int alloc_buffers(Buffers &b, int device, uint32_t jobsCap, int64_t queueBytes, hipStream_t stream,
                  char *msg, int cap) {
  b.device = device;
  b.jobsCap = jobsCap;
  TD16_HIP(hipSetDevice(device));
  // the budget is split evenly over the ring shards
  uint64_t qcap = (uint64_t)queueBytes / (sizeof(Item) + sizeof(uint64_t)) / kQueueShards;
  if (qcap < 64) qcap = 64;
  b.qcap = qcap;
  const uint64_t slots = qcap * kQueueShards;
  TD16_HIP(hipMalloc(&b.dJobs, sizeof(Job) * jobsCap));
  TD16_HIP(hipMalloc(&b.dStates, sizeof(JobState) * jobsCap));
  TD16_HIP(hipMalloc(&b.dCtrl, sizeof(Control)));
  TD16_HIP(hipMalloc(&b.dQueue, sizeof(QueueState) * kQueueShards));
  TD16_HIP(hipMalloc(&b.dItems, sizeof(Item) * slots));
  TD16_HIP(hipMalloc(&b.dSeq, sizeof(uint64_t) * slots));
  TD16_HIP(hipHostMalloc(&b.hCtrl, sizeof(Control), hipHostMallocDefault));
  TD16_HIP(hipHostMalloc(&b.hJobs, sizeof(Job) * jobsCap, hipHostMallocDefault));
  TD16_HIP(hipHostMalloc(&b.hResults, sizeof(DevResult) * jobsCap,
                         hipHostMallocCoherent | hipHostMallocMapped));
  TD16_HIP(hipHostGetDevicePointer((void **)&b.dResults, b.hResults, 0));
  std::memset(b.hCtrl, 0, sizeof(Control));
  // queue counters start at zero, slot q expects producer position q
  static const QueueState zeroQueue[kQueueShards] = {};
  TD16_HIP(hipMemcpyAsync(b.dQueue, zeroQueue, sizeof(zeroQueue), hipMemcpyHostToDevice, stream));
  TD16_HIP((hipError_t)launch_queue_init(b.dSeq, qcap, stream));
  TD16_HIP(hipStreamSynchronize(stream));
  return TD16_OK;
}

}  // namespace

size_t init_feed(Control *c, uint32_t numJobs) {
  c->numJobs = numJobs;
  const uint32_t per = (numJobs + kFeedShards - 1) / kFeedShards;
  c->perShard = per;
  uint32_t used = 0;
  for (int s = 0; s < kFeedShards; ++s) {
    const uint64_t lo = std::min<uint64_t>((uint64_t)s * per, numJobs);
    c->feed[s].next = (uint32_t)lo;
    if (lo < std::min<uint64_t>((uint64_t)(s + 1) * per, numJobs)) ++used;
  }
  // nonempty shards are the leading ones
  return offsetof(Control, feed) + (size_t)used * sizeof(FeedShard);
}

void default_knobs(Params &P) {
  P.splitSlack = 5;
  // children per donation: the autotuner's pick on the batched mix (0.80
  // against 0.84 ms per launch at 64), 1 % faster on the 1M emitter stream,
  // neutral on the per-request and served paths
  P.donateMax = 16;
  P.served = 0;
  // idle sleeps of up to ~6.4 us: 1 % faster on the batched mix than ~13 us
  // (the launch ends sooner after its last request), neutral on long batches
  P.idleBackoff = 16;
  // a sharing check every 16 expansions: heavy requests spread twice as
  // fast as at 32 (the batched mix 0.77 against 0.82 ms per launch, its
  // heaviest request alone 0.15 against 0.25 ms, the 1M emitter stream
  // 48.6 against 51.5 ms); at 8 the checks cost the mix more than they gain
  P.checkEvery = 16;
}

void apply_batch_knobs(Params &P, const td16_batch_options &o) {
  P.countCandidates = o.count_candidates ? 1 : 0;
  if (o.split_depth > 0) P.splitSlack = o.split_depth;
  if (o.donate_max > 0) P.donateMax = std::min(o.donate_max, 64);
  if (o.check_every > 0) P.checkEvery = std::min(o.check_every, 1024);
}

void batch_grid(int wavesPerCU, uint32_t numJobs, int cus, int blocksPerCU, int blocksPerCUWide,
                int *blocks, bool *wide) {
  // 0 = 20 waves per CU for a batch of at least 8 requests per wavefront at
  // 16 waves per CU, where requests keep the fifth wave per SIMD busy (the
  // mix 0.58 -> 0.55 ms, the 1M emitter stream 37.0 -> 33.6 ms), else 16
  // (fewer, longer requests share their subtrees better with fewer waves:
  // 10,000 n = 10 distances 0.85 against 0.94 ms, the 2,000 n = 16 bound-8
  // requests 27 against 29.5 ms)
  const int wantWaves =
      wavesPerCU > 0 ? wavesPerCU : ((uint64_t)numJobs >= 8ULL * 16 * (uint64_t)cus ? 20 : 16);
  // above the main build's residency (20 waves per CU: its 96-VGPR budget
  // allows 5 per SIMD) the 6-waves-per-SIMD build runs
  const int wantBlocks = std::max(1, wantWaves / kWavesPerBlock);
  *wide = wantBlocks > blocksPerCU && blocksPerCUWide > blocksPerCU;
  const int perCU = std::min(wantBlocks, *wide ? blocksPerCUWide : blocksPerCU);
  *blocks = cus * std::max(1, perCU);
}

namespace {

Params make_params(const Buffers &b, const DeviceInfo &dev) {
  Params P;
  std::memset(&P, 0, sizeof(P));
  P.jobs = b.dJobs;
  P.states = b.dStates;
  P.results = b.dResults;
  P.ctrl = b.dCtrl;
  P.queue = b.dQueue;
  P.items = b.dItems;
  P.seq = b.dSeq;
  P.qcap = b.qcap;
  P.lex = dev.dLex;
  default_knobs(P);
  return P;
}

// ---------------------------------------------------------------------------
// Slots
// ---------------------------------------------------------------------------

struct Slot {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int blocks = 0;
  DeviceInfo dev;
  Buffers buf;
  std::mutex mu;
};

std::mutex gSlotsMu;
std::vector<std::unique_ptr<Slot>> gSlots;

void destroy_slots_locked() {
  for (auto &s : gSlots) {
    if (!s) continue;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    s->buf.release();
    if (s->done) hipEventDestroy(s->done);
    if (s->stream) hipStreamDestroy(s->stream);
  }
  gSlots.clear();
}

Slot *get_slot(int slot) {
  std::lock_guard<std::mutex> lk(gSlotsMu);
  if (slot < 0 || slot >= (int)gSlots.size()) return nullptr;
  return gSlots[slot].get();
}

int run_plain(Slot &s, const td16_request *req, td16_result *res, char *msg, int cap) {
  Job job;
  int trivial = 0;
  const int st = make_job_checked(req, &job, &trivial, msg, cap);
  if (st != TD16_OK) return st;
  if (trivial) {
    trivial_result(trivial, res);
    return TD16_OK;
  }
  std::lock_guard<std::mutex> lk(s.mu);
  TD16_HIP(hipSetDevice(s.device));
  // the request rides in the control block: one copy sets up the launch
  std::memset(s.buf.hCtrl, 0, sizeof(Control));
  s.buf.hCtrl->single = job;
  const size_t ctrlBytes = init_feed(s.buf.hCtrl, 1);
  TD16_HIP(hipMemcpyAsync(s.buf.dCtrl, s.buf.hCtrl, ctrlBytes, hipMemcpyHostToDevice,
                          s.stream));
  Params P = make_params(s.buf, s.dev);
  feed_params(P, *s.buf.hCtrl);
  P.jobs = &s.buf.dCtrl->single;
  // one request: idle wavefronts only wait for its donations and its end, so
  // they poll every ~1.6 us instead of the batched ~13 us (measured: -5 us
  // per call, +8 % single-slot and +14 % seven-slot request rate)
  P.idleBackoff = 4;
  // a request that may run long gets the slot's whole grid; the rest run on
  // 32 blocks (128 wavefronts), whose idle pollers disturb its one working
  // wavefront far less than the slot's ~1,000 blocks would (measured on the
  // batched mix's requests: single-slot calls +40 %, p50 54 -> 37 us; its
  // heaviest requests, up to 5,271 nodes, take as long on 16 blocks as on
  // the whole chip)
  const bool longSearch = (job.flags & kFlagTight) && job.maxDepth >= kDeepBound;
  const int blocks = longSearch ? s.blocks : std::min(s.blocks, 32);
  TD16_HIP((hipError_t)launch_search(P, blocks, false, s.stream));
  TD16_HIP(hipEventRecord(s.done, s.stream));
  TD16_HIP(hipEventSynchronize(s.done));
  decode_result(s.buf.hResults[0], res);
  return TD16_OK;
}

// ---------------------------------------------------------------------------
// Batch context (one per process, grown on demand)
// ---------------------------------------------------------------------------

struct BatchCtx {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t evUp = nullptr, ev0 = nullptr, ev1 = nullptr;
  PrepareScratch ws;
  std::vector<int> where;   // caller index of each prepared request
  Buffers buf;
  int64_t queueBytes = 0;
  hipGraphExec_t graph = nullptr;
  int graphBlocks = 0;
  uint32_t graphJobs = 0;
};

std::mutex gBatchMu;
BatchCtx gBatch;

// Streamed batches: two launch contexts used in turn, so that the host
// prepares chunk c + 1 into one while chunk c runs from the other; uploads
// run on their own stream, beside the search.
struct StreamCtx {
  int device = -1;
  hipStream_t run = nullptr, copy = nullptr;
  hipEvent_t up0[2] = {}, up1[2] = {}, k0[2] = {}, k1[2] = {};
  Buffers buf[2];
  PrepareScratch ws[2];
  std::vector<int> where[2];
  int64_t queueBytes = 0;

  void release() {
    if (device >= 0) hipSetDevice(device);
    if (run) hipStreamSynchronize(run);
    if (copy) hipStreamSynchronize(copy);
    for (int s = 0; s < 2; ++s) {
      buf[s].release();
      for (hipEvent_t e : {up0[s], up1[s], k0[s], k1[s]})
        if (e) hipEventDestroy(e);
    }
    if (run) hipStreamDestroy(run);
    if (copy) hipStreamDestroy(copy);
    *this = StreamCtx();
  }
};

std::mutex gStreamMu;
StreamCtx gStream;

void reset_batch_graph() {
  if (gBatch.graph) hipGraphExecDestroy(gBatch.graph);
  gBatch.graph = nullptr;
}

// ---------------------------------------------------------------------------
// Served loop
// ---------------------------------------------------------------------------

// One submitted request.  Its result is handed over through `state` (0
// pending, 1 ready, 2 failed with the device) and the ticket's own
// condition variable, so a completion wakes exactly its waiter.
struct Ticket {
  std::atomic<int> state{0};
  std::atomic<bool> blocked{false};  // the waiter sleeps on cv (else it polls `state`)
  td16_result res;
  std::chrono::steady_clock::time_point t0;      // arrival
  std::chrono::steady_clock::time_point tAdmit;  // published to the device
  std::mutex m;
  std::condition_variable cv;
};
using TicketPtr = std::shared_ptr<Ticket>;

// Submit-to-result latency, in log-spaced buckets 2 % wide from 0.05 us to
// about 30 s (the last bucket takes anything longer): constant memory and an
// O(1) update however long the loop serves.
struct LatencyHist {
  static constexpr int kBuckets = 1024;
  static constexpr double kLo = 0.05, kRatio = 1.02;
  std::atomic<uint64_t> n[kBuckets];
  std::atomic<uint64_t> count{0}, sumNs{0};

  LatencyHist() { reset(); }
  void reset() {
    for (auto &b : n) b.store(0, std::memory_order_relaxed);
    count.store(0, std::memory_order_relaxed);
    sumNs.store(0, std::memory_order_relaxed);
  }
  void add(double us) {
    int b = 0;
    if (us > kLo) b = (int)(std::log(us / kLo) / std::log(kRatio));
    b = std::min(std::max(b, 0), kBuckets - 1);
    n[b].fetch_add(1, std::memory_order_relaxed);
    count.fetch_add(1, std::memory_order_relaxed);
    sumNs.fetch_add((uint64_t)(us * 1000.0), std::memory_order_relaxed);
  }
  // geometric middle of the bucket that holds the p-quantile
  double quantile(double p) const {
    const uint64_t total = count.load(std::memory_order_relaxed);
    if (total == 0) return 0.0;
    const uint64_t rank = (uint64_t)(p * (double)(total - 1)) + 1;
    uint64_t seen = 0;
    for (int b = 0; b < kBuckets; ++b) {
      seen += n[b].load(std::memory_order_relaxed);
      if (seen >= rank) return b == 0 ? kLo : kLo * std::pow(kRatio, b + 0.5);
    }
    return kLo * std::pow(kRatio, kBuckets);
  }
  double mean() const {
    const uint64_t total = count.load(std::memory_order_relaxed);
    return total ? (double)sumNs.load(std::memory_order_relaxed) * 1e-3 / (double)total : 0.0;
  }
};

constexpr int kTicketShards = 64;
struct TicketShard {
  std::mutex mu;
  std::unordered_map<uint64_t, TicketPtr> map;
};

struct Server {
  int device = -1;
  DeviceInfo dev;
  hipStream_t stream = nullptr;
  Buffers buf;
  uint32_t slots = 0, ringCap = 0, queueCap = 0;
  int deadlineMs = 0;
  ServeHost *hHost = nullptr;
  Job *hRing = nullptr;
  DevResult *hRes = nullptr;
  uint32_t *hComp = nullptr;
  ServeHost *dHost = nullptr;
  Job *dRing = nullptr;
  DevResult *dRes = nullptr;
  uint32_t *dComp = nullptr;

  // admission: requests waiting for a slot and the slots' tickets
  std::mutex admitMu;
  std::condition_variable cvSpace, cvIdle, cvWork;
  std::deque<std::pair<TicketPtr, Job>> waiting;
  std::vector<uint32_t> freeSlots;
  std::vector<TicketPtr> slotTicket;
  uint64_t submittedSeq = 0;
  bool stopping = false;  // written and read under admitMu
  std::atomic<uint32_t> inFlight{0};

  std::atomic<uint64_t> nextTicket{1};
  TicketShard shards[kTicketShards];
  std::thread poller;
  std::atomic<bool> pollerExit{false};
  // set once the served kernel is seen to have failed or ended unasked
  std::atomic<bool> failed{false};
  char failure[160] = {0};

  std::atomic<uint64_t> nSubmitted{0}, nCompleted{0}, nRejected{0};
  LatencyHist hist;

  TicketShard &shard(uint64_t t) { return shards[t % kTicketShards]; }
};

// gServeMu orders start and stop; gServerPtrMu guards only the pointer, so a
// stop that drains the loop never holds up the waits it is draining.
std::mutex gServeMu, gServerPtrMu;
std::shared_ptr<Server> gServer;

std::shared_ptr<Server> current_server() {
  std::lock_guard<std::mutex> lk(gServerPtrMu);
  return gServer;
}

// Publishes waiting requests while slots are free (caller holds admitMu).
// Called by submit and by the completion poller, so no thread hop sits
// between a free slot and the next request.
void admit_locked(Server *S) {
  bool any = false;
  while (!S->waiting.empty() && !S->freeSlots.empty()) {
    auto item = std::move(S->waiting.front());
    S->waiting.pop_front();
    const uint32_t slot = S->freeSlots.back();
    S->freeSlots.pop_back();
    item.first->tAdmit = std::chrono::steady_clock::now();
    S->slotTicket[slot] = std::move(item.first);
    S->inFlight.fetch_add(1, std::memory_order_relaxed);
    const uint64_t r = S->submittedSeq++;
    Job job = item.second;
    job.tag = slot;
    std::memcpy(&S->hRing[r & (S->ringCap - 1)], &job, sizeof(Job));
    any = true;
  }
  if (any) {
    std::atomic_thread_fence(std::memory_order_release);
    __atomic_store_n(&S->hHost->submitted, S->submittedSeq, __ATOMIC_RELEASE);
    S->cvSpace.notify_all();
    S->cvWork.notify_one();  // the poller may be parked on an idle loop
  }
}

// Fails every request still in the loop once the device is lost: their
// waiters return TD16_DEVICE_ERROR instead of waiting forever.
void fail_all(Server *S, const char *why) {
  bool expected = false;
  if (!S->failed.compare_exchange_strong(expected, true)) return;
  std::snprintf(S->failure, sizeof(S->failure), "%s", why);
  std::vector<TicketPtr> lost;
  {
    std::lock_guard<std::mutex> lk(S->admitMu);
    for (auto &t : S->slotTicket)
      if (t) lost.push_back(std::move(t));
    for (auto &w : S->waiting) lost.push_back(std::move(w.first));
    S->waiting.clear();
    S->inFlight.store(0);
    S->cvSpace.notify_all();
    S->cvIdle.notify_all();
  }
  for (auto &t : lost) {
    t->res.status = TD16_DEVICE_ERROR;
    t->state.store(2, std::memory_order_seq_cst);
    std::lock_guard<std::mutex> l(t->m);
    t->cv.notify_all();
  }
}

// Completion poller: takes every completion record the device has written,
// frees their slots and admits waiting requests under one short lock, then
// hands each result to its own waiter.  It spins while requests are in
// flight (a dedicated core buys microseconds of latency) and parks on cvWork
// otherwise; about every 50 ms it checks that the served kernel still runs.
void poll_loop(Server *S) {
  uint64_t c = 0;
  struct Done {
    uint32_t slot;
    DevResult r;
  };
  std::vector<Done> done;
  done.reserve(256);
  std::vector<TicketPtr> tickets;
  tickets.reserve(256);
  auto lastCheck = std::chrono::steady_clock::now();
  while (!S->pollerExit.load(std::memory_order_acquire)) {
    done.clear();
    while (done.size() < 256) {
      const uint32_t e = (uint32_t)(c & (S->ringCap - 1));
      const uint32_t seq = __atomic_load_n(&S->hComp[2 * e + 1], __ATOMIC_ACQUIRE);
      if (seq != (uint32_t)(c + 1)) break;
      Done d;
      d.slot = __atomic_load_n(&S->hComp[2 * e], __ATOMIC_ACQUIRE);
      // copied before the slot is freed: its next request reuses the record
      std::memcpy(&d.r, &S->hRes[d.slot], sizeof(DevResult));
      done.push_back(d);
      ++c;
    }
    const auto now = std::chrono::steady_clock::now();
    if (done.empty()) {
      if (now - lastCheck > std::chrono::milliseconds(50)) {
        lastCheck = now;
        const hipError_t q = hipStreamQuery(S->stream);
        if (q != hipErrorNotReady) {
          char why[128];
          std::snprintf(why, sizeof why, "the served kernel ended unexpectedly (%s)",
                        hipGetErrorName(q));
          fail_all(S, why);
          return;
        }
      }
      if (S->inFlight.load(std::memory_order_relaxed) == 0) {
        std::unique_lock<std::mutex> lk(S->admitMu);
        S->cvWork.wait_for(lk, std::chrono::milliseconds(1), [&] {
          return S->inFlight.load() > 0 || S->pollerExit.load(std::memory_order_acquire);
        });
      } else {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#endif
      }
      continue;
    }
    tickets.clear();
    {
      std::lock_guard<std::mutex> lk(S->admitMu);
      for (const Done &d : done) {
        tickets.push_back(std::move(S->slotTicket[d.slot]));
        S->freeSlots.push_back(d.slot);
      }
      S->inFlight.fetch_sub((uint32_t)done.size(), std::memory_order_relaxed);
      admit_locked(S);
      if (S->waiting.empty() && S->inFlight.load() == 0) S->cvIdle.notify_all();
    }
    for (size_t q = 0; q < done.size(); ++q) {
      Ticket *t = tickets[q].get();
      if (t == nullptr) continue;
      td16_result res;
      decode_result(done[q].r, &res);
      const double us = std::chrono::duration<double, std::micro>(now - t->t0).count();
      res.latency_us = us;
      res.queue_us = std::chrono::duration<double, std::micro>(t->tAdmit - t->t0).count();
      S->hist.add(us);
      t->res = res;
      // a waiter that polls sees the state; only one that sleeps needs a
      // wake-up (state and flag are sequentially consistent on both sides)
      t->state.store(1, std::memory_order_seq_cst);
      if (t->blocked.load(std::memory_order_seq_cst)) {
        std::lock_guard<std::mutex> l(t->m);
        t->cv.notify_all();
      }
    }
    S->nCompleted.fetch_add(done.size(), std::memory_order_relaxed);
  }
}

}  // namespace
}  // namespace td16

using namespace td16;

extern "C" {

int td16_init_slots(int totalSlots, const int32_t *devices, int ndevices, int64_t budget,
                    char *msg, int cap) {
  // slot_manager.cu:44-128
  if (devices == nullptr || ndevices <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "devices must contain at least one device ordinal");
  if (totalSlots <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "totalSlots must be positive");
  if (budget <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "queueBytesBudget must be positive");
  int count = 0;
  TD16_HIP(hipGetDeviceCount(&count));
  std::vector<int> seen;
  for (int q = 0; q < ndevices; ++q) {
    const int d = devices[q];
    if (d < 0 || d >= count)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "devices contains a device ordinal outside the available range");
    if (std::find(seen.begin(), seen.end(), d) != seen.end())
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "devices must contain distinct device ordinals");
    seen.push_back(d);
  }
  std::lock_guard<std::mutex> lk(gSlotsMu);
  destroy_slots_locked();
  std::vector<int> perDevice(count, 0), ordinal(count, 0);
  for (int s = 0; s < totalSlots; ++s) perDevice[devices[s % ndevices]]++;
  for (int q = 0; q < ndevices; ++q) {
    const int d = devices[q];
    TD16_HIP(hipSetDevice(d));
    size_t freeB = 0, totalB = 0;
    TD16_HIP(hipMemGetInfo(&freeB, &totalB));
    const unsigned long long need = (unsigned long long)budget * (unsigned long long)perDevice[d];
    if (need > freeB)
      return set_msg(msg, cap, TD16_OUT_OF_MEMORY,
                     "device %d has %zu free bytes, but its search queues require at least %llu bytes",
                     d, freeB, need);
  }
  for (int s = 0; s < totalSlots; ++s) {
    const int d = devices[s % ndevices];
    auto slot = std::make_unique<Slot>();
    slot->device = d;
    int st = device_info(d, &slot->dev, msg, cap);
    if (st != TD16_OK) {
      destroy_slots_locked();
      return st;
    }
    const int resident = slot->dev.cus * slot->dev.blocksPerCU;
    if (perDevice[d] > resident) {
      destroy_slots_locked();
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "slots per device exceeds the persistent kernel's resident-block capacity");
    }
    // sorting_search.cu:740-752: the device's resident blocks split evenly
    const int base = std::max(1, resident / perDevice[d]);
    const int rem = resident % perDevice[d];
    slot->blocks = base + (ordinal[d] < rem ? 1 : 0);
    ordinal[d]++;
    hipError_t e = hipSetDevice(d);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&slot->done, hipEventBlockingSync | hipEventDisableTiming);
    if (e != hipSuccess) {
      destroy_slots_locked();
      return set_msg(msg, cap, TD16_DEVICE_ERROR, "HIP error %s", hipGetErrorName(e));
    }
    st = alloc_buffers(slot->buf, d, 1, budget, slot->stream, msg, cap);
    gSlots.push_back(std::move(slot));
    if (st != TD16_OK) {
      destroy_slots_locked();
      return st;
    }
  }
  return TD16_OK;
}

int td16_shutdown(char *msg, int cap) {
  td16_serve_stop(msg, cap);
  {
    std::lock_guard<std::mutex> lk(gSlotsMu);
    destroy_slots_locked();
  }
  {
    std::lock_guard<std::mutex> lk(gBatchMu);
    reset_batch_graph();
    if (gBatch.stream) {
      hipSetDevice(gBatch.device);
      hipStreamSynchronize(gBatch.stream);
      hipEventDestroy(gBatch.evUp);
      hipEventDestroy(gBatch.ev0);
      hipEventDestroy(gBatch.ev1);
      hipStreamDestroy(gBatch.stream);
    }
    gBatch.buf.release();
    gBatch = BatchCtx();
  }
  {
    std::lock_guard<std::mutex> lk(gStreamMu);
    gStream.release();
  }
  return TD16_OK;
}

int td16_slot_count(void) {
  std::lock_guard<std::mutex> lk(gSlotsMu);
  return (int)gSlots.size();
}

int td16_slot_device(int slot) {
  Slot *s = get_slot(slot);
  return s ? s->device : -1;
}

int td16_search_ex(int slot, const td16_request *req, td16_result *res, char *msg, int cap) {
  if (req == nullptr || res == nullptr)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "request and result must not be null");
  Slot *s = get_slot(slot);
  if (s == nullptr) {
    if (td16_slot_count() == 0)
      return set_msg(msg, cap, TD16_NOT_INITIALIZED, "GPU sorting slots are not initialized");
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid GPU sorting slot");
  }
  const int st = run_plain(*s, req, res, msg, cap);
  res->status = st;
  return st;
}

int td16_search(int slot, const uint8_t *beta, const uint8_t *omega, int n, int initialEven,
                float minRate, int maxMoves, int fullSorting, int32_t *outPositions,
                int32_t *outLen, char *msg, int cap) {
  // jni_sorter.cu:49-117
  if (beta == nullptr || omega == nullptr || outPositions == nullptr || outLen == nullptr)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "beta and omega must not be null");
  if (n < 0 || n > TD16_MAX_N)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "performGpuSortingSearch: n exceeds MAX_N");
  td16_request req;
  std::memset(&req, 0, sizeof(req));
  std::memcpy(req.beta, beta, (size_t)n);
  std::memcpy(req.omega, omega, (size_t)n);
  req.n = n;
  req.initial_even = initialEven;
  req.min_rate = minRate;
  req.max_moves = maxMoves;
  req.full_sorting = fullSorting ? 1 : 0;
  req.moveset = TD16_MOVES_REF;
  req.mode = TD16_FIRST;
  td16_result res;
  const int st = td16_search_ex(slot, &req, &res, msg, cap);
  if (st != TD16_OK) return st;
  *outLen = res.found ? res.depth : 0;
  for (int d = 0; d < *outLen; ++d)
    for (int q = 0; q < 3; ++q) outPositions[3 * d + q] = res.moves[d][q];
  return TD16_OK;
}

}  // extern "C"

namespace td16 {
namespace {

td16_batch_options normalize_batch_options(const td16_batch_options *opt) {
  td16_batch_options o;
  std::memset(&o, 0, sizeof(o));
  if (opt) o = *opt;
  if (o.repeats < 1) o.repeats = 1;
  if (o.queue_bytes <= 0) o.queue_bytes = 256LL << 20;
  return o;
}

}  // namespace

namespace {

// Makes the process's batch context fit `capJobsNeeded` requests on device
// o.device (the caller holds gBatchMu).
int ensure_batch_ctx(const td16_batch_options &o, uint32_t capJobsNeeded, DeviceInfo &dev,
                     char *msg, int cap) {
  int st = device_info(o.device, &dev, msg, cap);
  if (st != TD16_OK) return st;
  TD16_HIP(hipSetDevice(o.device));
  if (gBatch.device != o.device || gBatch.buf.jobsCap < std::max<uint32_t>(capJobsNeeded, 1) ||
      gBatch.queueBytes != o.queue_bytes) {
    reset_batch_graph();
    if (gBatch.stream) {
      hipSetDevice(gBatch.device);
      hipStreamSynchronize(gBatch.stream);
      hipEventDestroy(gBatch.evUp);
      hipEventDestroy(gBatch.ev0);
      hipEventDestroy(gBatch.ev1);
      hipStreamDestroy(gBatch.stream);
      gBatch.stream = nullptr;
    }
    gBatch.buf.release();
    TD16_HIP(hipSetDevice(o.device));
    gBatch.device = o.device;
    gBatch.queueBytes = o.queue_bytes;
    TD16_HIP(hipStreamCreateWithFlags(&gBatch.stream, hipStreamNonBlocking));
    TD16_HIP(hipEventCreate(&gBatch.evUp));
    TD16_HIP(hipEventCreate(&gBatch.ev0));
    TD16_HIP(hipEventCreate(&gBatch.ev1));
    uint32_t capJobs = 1024;
    while (capJobs < capJobsNeeded) capJobs *= 2;
    st = alloc_buffers(gBatch.buf, o.device, capJobs, o.queue_bytes, gBatch.stream, msg, cap);
    if (st != TD16_OK) return st;
  }
  return TD16_OK;
}

// Device side of a batch: uploads the nj requests prepared in the pinned
// table, runs `o.repeats` persistent launches over them and decodes the last
// one's results into results[q] for the launched requests among the
// caller's `count` (request q at feed position pos[q]; the caller holds
// gBatchMu).
int launch_prepared(uint32_t nj, const uint32_t *pos, uint32_t count,
                    const td16_batch_options &o, const DeviceInfo &dev, td16_result *results,
                    td16_batch_stats &localStats, char *msg, int cap) {
  Buffers &b = gBatch.buf;
  int blocks = 0;
  bool wide = false;
  batch_grid(o.waves_per_cu, nj, dev.cus, dev.blocksPerCU, dev.blocksPerCUWide, &blocks, &wide);
  localStats.blocks = blocks;
  localStats.waves = blocks * kWavesPerBlock;
  if (nj == 0) return TD16_OK;
  TD16_HIP(hipEventRecord(gBatch.evUp, gBatch.stream));
  TD16_HIP(hipMemcpyAsync(b.dJobs, b.hJobs, sizeof(Job) * nj, hipMemcpyHostToDevice, gBatch.stream));
  std::memset(b.hCtrl, 0, sizeof(Control));
  const size_t ctrlBytes = init_feed(b.hCtrl, nj);
  Params P = make_params(b, dev);
  feed_params(P, *b.hCtrl);
  apply_batch_knobs(P, o);

  if (o.use_graph) {
    reset_batch_graph();
    hipGraph_t g;
    TD16_HIP(hipStreamBeginCapture(gBatch.stream, hipStreamCaptureModeThreadLocal));
    // the graph resets the control block on the device and runs the search,
    // so a replay is one launch (no host-to-device copy); the request upload
    // and the events that time the launch stay outside it
    launch_control_reset(b.dCtrl, P.numJobs, P.perShard, gBatch.stream);
    launch_search(P, blocks, wide, gBatch.stream);
    TD16_HIP(hipStreamEndCapture(gBatch.stream, &g));
    hipError_t e = hipGraphInstantiate(&gBatch.graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    TD16_HIP(e);
    gBatch.graphBlocks = blocks;
    gBatch.graphJobs = nj;
  }

  double kernelMs = 0.0, wallMs = 0.0;
  for (int rep = 0; rep < o.repeats; ++rep) {
    const auto w0 = std::chrono::steady_clock::now();
    if (!o.use_graph)
      TD16_HIP(hipMemcpyAsync(b.dCtrl, b.hCtrl, ctrlBytes, hipMemcpyHostToDevice,
                              gBatch.stream));
    TD16_HIP(hipEventRecord(gBatch.ev0, gBatch.stream));
    if (o.use_graph)
      TD16_HIP(hipGraphLaunch(gBatch.graph, gBatch.stream));
    else
      TD16_HIP((hipError_t)launch_search(P, blocks, wide, gBatch.stream));
    TD16_HIP(hipEventRecord(gBatch.ev1, gBatch.stream));
    TD16_HIP(hipStreamSynchronize(gBatch.stream));
    const auto w1 = std::chrono::steady_clock::now();
    float ms = 0.f;
    TD16_HIP(hipEventElapsedTime(&ms, gBatch.ev0, gBatch.ev1));
    kernelMs += ms;
    wallMs += std::chrono::duration<double, std::milli>(w1 - w0).count();
    if (rep == 0) {
      float up = 0.f;  // the table upload (and, without a graph, the first control copy)
      TD16_HIP(hipEventElapsedTime(&up, gBatch.evUp, gBatch.ev0));
      localStats.upload_ms = up;
    }
  }
  localStats.kernel_ms = kernelMs / o.repeats;
  localStats.wall_ms = wallMs / o.repeats;
  TD16_HIP(hipMemcpy(&localStats.candidates, &b.dCtrl->candidates, sizeof(uint64_t),
                     hipMemcpyDeviceToHost));
  const auto d0 = std::chrono::steady_clock::now();
  decode_batch(b.hResults, pos, count, results, o.host_threads, &localStats.expanded,
               &localStats.solutions, &localStats.donations);
  localStats.decode_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d0).count();
  return TD16_OK;
}

// `o.repeats` persistent launches over requests prepared once (the
// autotuner's inner loop); results of the last one land in results[where[q]].
int run_batch(const std::vector<Job> &jobs, const std::vector<int> &where, int count,
              const td16_batch_options &o, td16_result *results, td16_batch_stats *stats,
              char *msg, int cap) {
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t nj = (uint32_t)jobs.size();
  td16_batch_stats localStats;
  std::memset(&localStats, 0, sizeof(localStats));
  std::lock_guard<std::mutex> lk(gBatchMu);
  DeviceInfo dev;
  int st = ensure_batch_ctx(o, nj, dev, msg, cap);
  if (st != TD16_OK) return st;
  if (nj) std::memcpy(gBatch.buf.hJobs, jobs.data(), sizeof(Job) * nj);
  std::vector<uint32_t> pos;
  positions_of(where.data(), nj, (uint32_t)count, pos);
  st = launch_prepared(nj, pos.data(), (uint32_t)count, o, dev, results, localStats, msg, cap);
  if (st != TD16_OK) return st;
  localStats.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = localStats;
  return TD16_OK;
}

}  // namespace
}  // namespace td16

extern "C" {

int td16_batch(const td16_request *reqs, int count, td16_result *results,
               const td16_batch_options *opt, td16_batch_stats *stats, char *msg, int cap) {
  if (count < 0 || (count > 0 && (reqs == nullptr || results == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid batch arguments");
  const auto t0 = std::chrono::steady_clock::now();
  const td16_batch_options o = normalize_batch_options(opt);
  td16_batch_stats localStats;
  std::memset(&localStats, 0, sizeof(localStats));
  std::lock_guard<std::mutex> lk(gBatchMu);
  DeviceInfo dev;
  int st = ensure_batch_ctx(o, (uint32_t)count, dev, msg, cap);
  if (st != TD16_OK) return st;
  // requests go straight to their feed positions in the pinned upload table
  gBatch.where.resize((size_t)std::max(count, 1));
  uint32_t nj = 0;
  st = prepare_batch_into(reqs, count, results, gBatch.buf.hJobs, gBatch.where.data(), &nj,
                          gBatch.ws, o.host_threads, msg, cap);
  if (st != TD16_OK) return st;
  localStats.prepare_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  st = launch_prepared(nj, gBatch.ws.pos.data(), (uint32_t)count, o, dev, results, localStats,
                       msg, cap);
  if (st != TD16_OK) return st;
  localStats.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = localStats;
  return TD16_OK;
}

int td16_batch_stream(const td16_request *reqs, int count, int chunk, td16_result *results,
                      const td16_batch_options *opt, td16_batch_stats *stats, char *msg,
                      int cap) {
  if (count < 0 || (count > 0 && (reqs == nullptr || results == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid batch arguments");
  if (chunk < 1)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "chunk must be at least 1 request");
  const auto t0 = std::chrono::steady_clock::now();
  const td16_batch_options o = normalize_batch_options(opt);
  td16_batch_stats total;
  std::memset(&total, 0, sizeof(total));
  if (count == 0) {
    if (stats) *stats = total;
    return TD16_OK;
  }
  chunk = std::min(chunk, count);
  std::lock_guard<std::mutex> lk(gStreamMu);
  DeviceInfo dev;
  int st = device_info(o.device, &dev, msg, cap);
  if (st != TD16_OK) return st;
  StreamCtx &X = gStream;
  TD16_HIP(hipSetDevice(o.device));
  if (X.device != o.device || X.buf[0].jobsCap < (uint32_t)chunk ||
      X.queueBytes != o.queue_bytes) {
    X.release();
    X.device = o.device;
    X.queueBytes = o.queue_bytes;
    TD16_HIP(hipStreamCreateWithFlags(&X.run, hipStreamNonBlocking));
    TD16_HIP(hipStreamCreateWithFlags(&X.copy, hipStreamNonBlocking));
    uint32_t capJobs = 1024;
    while (capJobs < (uint32_t)chunk) capJobs *= 2;
    for (int s = 0; s < 2; ++s) {
      for (hipEvent_t *e : {&X.up0[s], &X.up1[s], &X.k0[s], &X.k1[s]})
        TD16_HIP(hipEventCreate(e));
      st = alloc_buffers(X.buf[s], o.device, capJobs, o.queue_bytes, X.run, msg, cap);
      if (st != TD16_OK) return st;
    }
  }
  int blocks = 0;
  bool wide = false;
  batch_grid(o.waves_per_cu, (uint32_t)std::min(count, chunk), dev.cus, dev.blocksPerCU,
             dev.blocksPerCUWide, &blocks, &wide);
  total.blocks = blocks;
  total.waves = blocks * kWavesPerBlock;
  const int chunks = (count + chunk - 1) / chunk;
  uint32_t nj[2] = {0, 0};
  int base[2] = {0, 0}, cnt[2] = {0, 0};

  // Chunk c: prepared on the host into context c % 2, uploaded on the copy
  // stream, searched on the run stream once its upload is done.
  auto enqueue = [&](int s) -> int {
    Buffers &b = X.buf[s];
    std::memset(b.hCtrl, 0, sizeof(Control));
    const size_t ctrlBytes = init_feed(b.hCtrl, nj[s]);
    TD16_HIP(hipEventRecord(X.up0[s], X.copy));
    TD16_HIP(hipMemcpyAsync(b.dJobs, b.hJobs, sizeof(Job) * nj[s], hipMemcpyHostToDevice, X.copy));
    TD16_HIP(hipMemcpyAsync(b.dCtrl, b.hCtrl, ctrlBytes, hipMemcpyHostToDevice, X.copy));
    TD16_HIP(hipEventRecord(X.up1[s], X.copy));
    TD16_HIP(hipStreamWaitEvent(X.run, X.up1[s], 0));
    Params P = make_params(b, dev);
    feed_params(P, *b.hCtrl);
    apply_batch_knobs(P, o);
    TD16_HIP(hipEventRecord(X.k0[s], X.run));
    TD16_HIP((hipError_t)launch_search(P, blocks, wide, X.run));
    TD16_HIP(hipEventRecord(X.k1[s], X.run));
    return TD16_OK;
  };
  auto finish = [&](int s) -> int {
    if (nj[s] == 0) return TD16_OK;
    const auto w0 = std::chrono::steady_clock::now();
    TD16_HIP(hipEventSynchronize(X.k1[s]));
    total.wall_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    float ms = 0.f, up = 0.f;
    TD16_HIP(hipEventElapsedTime(&ms, X.k0[s], X.k1[s]));
    TD16_HIP(hipEventElapsedTime(&up, X.up0[s], X.up1[s]));
    total.kernel_ms += ms;
    total.upload_ms += up;
    uint64_t cands = 0;
    TD16_HIP(hipMemcpy(&cands, &X.buf[s].dCtrl->candidates, sizeof(cands), hipMemcpyDeviceToHost));
    total.candidates += cands;
    const auto d0 = std::chrono::steady_clock::now();
    uint64_t e = 0, so = 0, dn = 0;
    decode_batch(X.buf[s].hResults, X.ws[s].pos.data(), (uint32_t)cnt[s], results + base[s],
                 o.host_threads, &e, &so, &dn);
    total.decode_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d0).count();
    total.expanded += e;
    total.solutions += so;
    total.donations += dn;
    return TD16_OK;
  };

  for (int c = 0; c <= chunks; ++c) {
    const int s = c & 1;
    if (c < chunks) {
      // context s last served chunk c - 2, finished below in step c - 1
      base[s] = c * chunk;
      const int n = std::min(chunk, count - base[s]);
      cnt[s] = n;
      X.where[s].resize((size_t)n);
      const auto p0 = std::chrono::steady_clock::now();
      st = prepare_batch_into(reqs + base[s], n, results + base[s], X.buf[s].hJobs,
                              X.where[s].data(), &nj[s], X.ws[s], o.host_threads, msg, cap);
      total.prepare_ms +=
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p0)
              .count();
      if (st != TD16_OK) {
        hipStreamSynchronize(X.run);
        return st;
      }
      if (nj[s]) {
        st = enqueue(s);
        if (st != TD16_OK) return st;
      }
    }
    if (c >= 1) {
      st = finish(s ^ 1);  // chunk c - 1, searched while chunk c was prepared
      if (st != TD16_OK) return st;
    }
  }
  total.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = total;
  return TD16_OK;
}

int td16_autotune(const td16_request *reqs, int count, td16_result *results,
                  const td16_batch_options *base, int repeats, td16_batch_options *best,
                  td16_tune_point *points, int pointsCap, int *npoints, char *msg, int cap) {
  if (count <= 0 || reqs == nullptr || results == nullptr || best == nullptr)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "autotune needs a non-empty batch, its results and an output for the choice");
  if (repeats < 1 || repeats > 100)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "autotune repeats must be in 1..100");
  if (pointsCap < 0 || (pointsCap > 0 && points == nullptr))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid tuning-point buffer");
  if (npoints) *npoints = 0;
  td16_batch_options o = normalize_batch_options(base);
  o.repeats = 1;
  std::vector<Job> jobs;
  std::vector<int> where;
  int st = prepare_batch(reqs, count, results, jobs, where, msg, cap);
  if (st != TD16_OK) return st;

  // The knobs change how the work is shared, never the answer: whether a
  // request has a witness, its exact distance under iterative deepening and
  // the counts of an exhaustive search are the same at every point, and any
  // difference is reported as a device error.
  // trivial requests are answered by prepare_batch; run_batch rewrites the rest
  std::vector<td16_result> first, cur(results, results + count);
  auto same_answers = [&](const std::vector<td16_result> &a, const std::vector<td16_result> &b) {
    for (int q = 0; q < count; ++q) {
      if (a[q].found != b[q].found) return false;
      if (reqs[q].iddfs && a[q].depth != b[q].depth) return false;
      if (reqs[q].mode == TD16_COUNT_ALL &&
          (a[q].expanded != b[q].expanded || a[q].solutions != b[q].solutions))
        return false;
    }
    return true;
  };

  struct Point {
    int w, s, d, k;
    double ms;
    int waves;
  };
  std::vector<Point> seen;
  // one untimed launch, then `repeats` timed ones; the median device time
  auto measure = [&](int w, int sd, int dm, int ck, double *ms) -> int {
    for (const Point &p : seen)
      if (p.w == w && p.s == sd && p.d == dm && p.k == ck) {
        *ms = p.ms;
        return TD16_OK;
      }
    o.waves_per_cu = w;
    o.split_depth = sd;
    o.donate_max = dm;
    o.check_every = ck;
    td16_batch_stats bs;
    int rc = run_batch(jobs, where, count, o, cur.data(), &bs, msg, cap);
    if (rc != TD16_OK) return rc;
    std::vector<double> t;
    for (int r = 0; r < repeats; ++r) {
      rc = run_batch(jobs, where, count, o, cur.data(), &bs, msg, cap);
      if (rc != TD16_OK) return rc;
      t.push_back(bs.kernel_ms);
    }
    if (first.empty()) {
      first = cur;
    } else if (!same_answers(first, cur)) {
      return set_msg(msg, cap, TD16_DEVICE_ERROR,
                     "autotune: waves_per_cu=%d split_depth=%d donate_max=%d check_every=%d "
                     "changed a result",
                     w, sd, dm, ck);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    seen.push_back(Point{w, sd, dm, ck, *ms, bs.waves});
    return TD16_OK;
  };

  static const int kWaves[] = {4, 8, 12, 16, 20, 24};
  static const int kSplit[] = {2, 3, 4, 5, 6, 7};
  static const int kDonate[] = {4, 8, 16, 32, 64};
  static const int kCheck[] = {8, 16, 32};
  // the search starts at the batch's default grid (batch_grid's rule)
  int bw = o.waves_per_cu;
  if (bw <= 0) {
    DeviceInfo dev;
    st = device_info(o.device, &dev, msg, cap);
    if (st != TD16_OK) return st;
    bw = (uint64_t)jobs.size() >= 8ULL * 16 * (uint64_t)dev.cus ? 20 : 16;
  }
  int bsd = o.split_depth > 0 ? o.split_depth : 5;
  int bdm = o.donate_max > 0 ? std::min(o.donate_max, 64) : 16;
  int bck = o.check_every > 0 ? std::min(o.check_every, 1024) : 16;
  double bms = 0.0;
  st = measure(bw, bsd, bdm, bck, &bms);
  if (st != TD16_OK) return st;
  // coordinate descent, one knob at a time, until a full pass changes nothing
  for (int pass = 0; pass < 3; ++pass) {
    bool moved = false;
    for (int knob = 0; knob < 4; ++knob) {
      const int *vals = knob == 0 ? kWaves : knob == 1 ? kSplit : knob == 2 ? kDonate : kCheck;
      const int nv = knob == 0 ? (int)(sizeof(kWaves) / sizeof(int))
                     : knob == 1 ? (int)(sizeof(kSplit) / sizeof(int))
                     : knob == 2 ? (int)(sizeof(kDonate) / sizeof(int))
                                 : (int)(sizeof(kCheck) / sizeof(int));
      for (int v = 0; v < nv; ++v) {
        const int w = knob == 0 ? vals[v] : bw;
        const int sd = knob == 1 ? vals[v] : bsd;
        const int dm = knob == 2 ? vals[v] : bdm;
        const int ck = knob == 3 ? vals[v] : bck;
        double ms = 0.0;
        st = measure(w, sd, dm, ck, &ms);
        if (st != TD16_OK) return st;
        // a new point must win by more than 1 % to displace the current one
        if (ms < bms * 0.99) {
          bw = w;
          bsd = sd;
          bdm = dm;
          bck = ck;
          bms = ms;
          moved = true;
        }
      }
    }
    if (!moved) break;
  }
  *best = normalize_batch_options(base);
  best->waves_per_cu = bw;
  best->split_depth = bsd;
  best->donate_max = bdm;
  best->check_every = bck;
  for (int q = 0; q < count; ++q) results[q] = first[q];
  const int np = (int)seen.size();
  for (int q = 0; q < np && q < pointsCap; ++q) {
    points[q].waves_per_cu = seen[q].w;
    points[q].split_depth = seen[q].s;
    points[q].donate_max = seen[q].d;
    points[q].check_every = seen[q].k;
    points[q].reserved = 0;
    points[q].waves = seen[q].waves;
    points[q].kernel_ms = seen[q].ms;
  }
  if (npoints) *npoints = np;
  return TD16_OK;
}

int td16_serve_start(const td16_serve_options *opt, char *msg, int cap) {
  if (opt == nullptr) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "options must not be null");
  if (opt->slots < 1 || opt->slots > 65536)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "slots must be in 1..65536");
  if (opt->queue_capacity < 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "queue capacity must not be negative");
  if (opt->slot_bytes < (int64_t)(sizeof(Item) + sizeof(uint64_t)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "the slot budget must hold at least one work item (%zu bytes)",
                   sizeof(Item) + sizeof(uint64_t));
  if (opt->deadline_ms < 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "deadline_ms must not be negative");
  std::lock_guard<std::mutex> lk(gServeMu);
  if (current_server())
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "the served loop is already running");
  DeviceInfo dev;
  int st = device_info(opt->device, &dev, msg, cap);
  if (st != TD16_OK) return st;
  TD16_HIP(hipSetDevice(opt->device));
  size_t freeB = 0, totalB = 0;
  TD16_HIP(hipMemGetInfo(&freeB, &totalB));
  // every slot holds up to slot_bytes of HBM for its request's queued
  // subtrees (sorting_search.cu:755-793 sizes each slot's queues from its
  // budget; slot_manager.cu:87-106 checks budget x slots against free memory)
  const unsigned long long queueBytes =
      (unsigned long long)opt->slot_bytes * (unsigned long long)opt->slots;
  if (queueBytes > freeB)
    return set_msg(msg, cap, TD16_OUT_OF_MEMORY,
                   "device %d has %zu free bytes, but the served loop requires %llu bytes",
                   opt->device, freeB, queueBytes);
  auto S = std::make_shared<Server>();
  S->device = opt->device;
  S->dev = dev;
  S->slots = (uint32_t)opt->slots;
  S->queueCap = (uint32_t)opt->queue_capacity;
  S->deadlineMs = opt->deadline_ms;
  uint32_t ring = 64;
  while (ring < 2 * S->slots) ring *= 2;
  S->ringCap = ring;
  TD16_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  st = alloc_buffers(S->buf, opt->device, S->slots, (int64_t)queueBytes, S->stream, msg, cap);
  if (st != TD16_OK) {
    S->buf.release();
    hipStreamDestroy(S->stream);
    return st;
  }
  const unsigned flags = hipHostMallocCoherent | hipHostMallocMapped;
  TD16_HIP(hipHostMalloc(&S->hHost, sizeof(ServeHost), flags));
  TD16_HIP(hipHostMalloc(&S->hRing, sizeof(Job) * ring, flags));
  TD16_HIP(hipHostMalloc(&S->hRes, sizeof(DevResult) * S->slots, flags));
  TD16_HIP(hipHostMalloc(&S->hComp, sizeof(uint32_t) * 2 * ring, flags));
  std::memset(S->hHost, 0, sizeof(ServeHost));
  std::memset(S->hRing, 0, sizeof(Job) * ring);
  std::memset(S->hRes, 0, sizeof(DevResult) * S->slots);
  std::memset(S->hComp, 0, sizeof(uint32_t) * 2 * ring);
  TD16_HIP(hipHostGetDevicePointer((void **)&S->dHost, S->hHost, 0));
  TD16_HIP(hipHostGetDevicePointer((void **)&S->dRing, S->hRing, 0));
  TD16_HIP(hipHostGetDevicePointer((void **)&S->dRes, S->hRes, 0));
  TD16_HIP(hipHostGetDevicePointer((void **)&S->dComp, S->hComp, 0));
  S->freeSlots.resize(S->slots);
  for (uint32_t q = 0; q < S->slots; ++q) S->freeSlots[q] = S->slots - 1 - q;
  S->slotTicket.assign(S->slots, nullptr);

  std::memset(S->buf.hCtrl, 0, sizeof(Control));
  TD16_HIP(hipMemcpyAsync(S->buf.dCtrl, S->buf.hCtrl, sizeof(Control), hipMemcpyHostToDevice,
                          S->stream));
  Params P = make_params(S->buf, dev);
  P.served = 1;
  // idle wavefronts sleep up to ~0.8 ms: among 1,024 of them one wakes
  // about every microsecond, so a lone request is still picked up at once
  // (20k/s: p50 18.5 us, p99 42 us, the same as with ~51 us sleeps), while
  // the polling traffic that slows every request under load drops 16-fold
  // (capacity at 16 / 64 slots: 284k / 114k requests/s with ~51 us sleeps,
  // 736k / 597k with ~0.8 ms; against 92 / 141 us p50 / p99 at ~3.2 us)
  P.idleBackoff = 2048;
  // served requests are mostly light and arrive while others run: checks
  // every 32 expansions keep the p99 lower (20k/s: 141-143 against 164-171
  // us at 16; the median is the same)
  P.checkEvery = 32;
  // a request's queued subtrees stay within its slot's budget
  const uint64_t perSlot = (uint64_t)opt->slot_bytes / (sizeof(Item) + sizeof(uint64_t));
  P.itemCap = (int32_t)std::min<uint64_t>(perSlot, (uint64_t)INT32_MAX);
  P.ringCap = ring;
  P.host = S->dHost;
  P.hostJobs = S->dRing;
  P.hostResults = S->dRes;
  P.hostCompletionIdx = S->dComp;
  // 0 = the measured best for latency: 4 waves per CU (one per SIMD)
  const int wantWaves = opt->waves_per_cu > 0 ? opt->waves_per_cu : 4;
  const int perCU = std::max(1, std::min(dev.blocksPerCU, wantWaves / kWavesPerBlock));
  TD16_HIP((hipError_t)launch_search(P, dev.cus * perCU, false, S->stream));
  S->poller = std::thread(poll_loop, S.get());
  std::lock_guard<std::mutex> pk(gServerPtrMu);
  gServer = S;
  return TD16_OK;
}

namespace td16 {
namespace {

int serve_submit_at(const std::shared_ptr<Server> &S, const td16_request *req, int nonblocking,
                    uint64_t *ticket, std::chrono::steady_clock::time_point arrival, char *msg,
                    int cap) {
  if (req == nullptr || ticket == nullptr)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "request and ticket must not be null");
  if (S->failed.load())
    return set_msg(msg, cap, TD16_DEVICE_ERROR, "served loop: %s", S->failure);
  Job job;
  int trivial = 0;
  const int st = make_job_checked(req, &job, &trivial, msg, cap);
  if (st != TD16_OK) return st;
  auto tk = std::make_shared<Ticket>();
  tk->t0 = arrival;
  const uint64_t t = S->nextTicket.fetch_add(1, std::memory_order_relaxed);
  if (trivial) {
    trivial_result(trivial, &tk->res);
    tk->state.store(1, std::memory_order_release);
    {
      TicketShard &sh = S->shard(t);
      std::lock_guard<std::mutex> l(sh.mu);
      sh.map.emplace(t, tk);
    }
    *ticket = t;
    S->nSubmitted.fetch_add(1, std::memory_order_relaxed);
    S->nCompleted.fetch_add(1, std::memory_order_relaxed);
    S->hist.add(0.0);
    return TD16_OK;
  }
  {
    TicketShard &sh = S->shard(t);
    std::lock_guard<std::mutex> l(sh.mu);
    sh.map.emplace(t, tk);
  }
  auto drop = [&](int status, const char *why) {
    TicketShard &sh = S->shard(t);
    std::lock_guard<std::mutex> l(sh.mu);
    sh.map.erase(t);
    return set_msg(msg, cap, status, "%s", why);
  };
  {
    std::unique_lock<std::mutex> lk(S->admitMu);
    // GPUSortingSearch.java:46-52: pending capacity = slots + queue capacity
    const size_t limit = (size_t)S->queueCap + S->slots;
    if (S->stopping) {
      lk.unlock();
      return drop(TD16_NOT_INITIALIZED, "the served loop is stopping");
    }
    if (S->waiting.size() >= limit) {
      if (nonblocking) {
        S->nRejected.fetch_add(1, std::memory_order_relaxed);
        lk.unlock();
        return drop(TD16_QUEUE_FULL, "the served loop's pending queue is full");
      }
      S->cvSpace.wait(lk, [&] {
        return S->waiting.size() < limit || S->stopping || S->failed.load();
      });
      if (S->stopping || S->failed.load()) {
        lk.unlock();
        return drop(S->stopping ? TD16_NOT_INITIALIZED : TD16_DEVICE_ERROR,
                    S->stopping ? "the served loop is stopping" : "the served loop failed");
      }
    }
    S->waiting.emplace_back(tk, job);
    admit_locked(S.get());
  }
  *ticket = t;
  S->nSubmitted.fetch_add(1, std::memory_order_relaxed);
  return TD16_OK;
}

int serve_wait(const std::shared_ptr<Server> &S, uint64_t ticket, td16_result *res, char *msg,
               int cap) {
  TicketPtr tk;
  {
    TicketShard &sh = S->shard(ticket);
    std::lock_guard<std::mutex> l(sh.mu);
    auto it = sh.map.find(ticket);
    if (it == sh.map.end()) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "unknown ticket");
    tk = it->second;
  }
  // Most results arrive within tens of microseconds.  The waiter polls its
  // ticket -- spinning briefly, then yielding the core, so that with more
  // waiting threads than cores the completion poller still runs -- and
  // sleeps on the ticket only after about 2 ms.
  const auto start = std::chrono::steady_clock::now();
  auto late = [&] {
    return S->deadlineMs > 0 &&
           std::chrono::steady_clock::now() - start > std::chrono::milliseconds(S->deadlineMs);
  };
  for (int q = 0; q < 256 && tk->state.load(std::memory_order_acquire) == 0; ++q) {
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
  while (tk->state.load(std::memory_order_acquire) == 0 &&
         std::chrono::steady_clock::now() - start < std::chrono::milliseconds(2)) {
    if (late())
      // the ticket stays valid: a later wait may still collect the result
      return set_msg(msg, cap, TD16_DEVICE_ERROR,
                     "served loop: no result within the %d ms deadline", S->deadlineMs);
    std::this_thread::yield();
  }
  if (tk->state.load(std::memory_order_acquire) == 0) {
    std::unique_lock<std::mutex> l(tk->m);
    tk->blocked.store(true, std::memory_order_seq_cst);
    const int slice = S->deadlineMs > 0 ? std::min(S->deadlineMs, 20) : 20;
    while (tk->state.load(std::memory_order_seq_cst) == 0) {
      tk->cv.wait_for(l, std::chrono::milliseconds(slice));
      if (tk->state.load(std::memory_order_acquire) != 0) break;
      if (S->failed.load())
        return set_msg(msg, cap, TD16_DEVICE_ERROR, "served loop: %s", S->failure);
      if (late())
        return set_msg(msg, cap, TD16_DEVICE_ERROR,
                       "served loop: no result within the %d ms deadline", S->deadlineMs);
    }
  }
  if (tk->state.load(std::memory_order_acquire) == 2) {
    TicketShard &sh = S->shard(ticket);
    std::lock_guard<std::mutex> l(sh.mu);
    sh.map.erase(ticket);
    return set_msg(msg, cap, TD16_DEVICE_ERROR, "served loop: %s", S->failure);
  }
  *res = tk->res;
  TicketShard &sh = S->shard(ticket);
  std::lock_guard<std::mutex> l(sh.mu);
  sh.map.erase(ticket);
  return TD16_OK;
}

}  // namespace
}  // namespace td16

int td16_serve_submit(const td16_request *req, int nonblocking, uint64_t *ticket, char *msg,
                      int cap) {
  auto S = current_server();
  if (!S) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  return serve_submit_at(S, req, nonblocking, ticket, std::chrono::steady_clock::now(), msg, cap);
}

int td16_serve_wait(uint64_t ticket, td16_result *res, char *msg, int cap) {
  if (res == nullptr) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "result must not be null");
  auto S = current_server();
  if (!S) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  return serve_wait(S, ticket, res, msg, cap);
}

int td16_serve_search(const td16_request *req, td16_result *res, char *msg, int cap) {
  auto S = current_server();
  if (!S) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  uint64_t t = 0;
  const int st = serve_submit_at(S, req, 0, &t, std::chrono::steady_clock::now(), msg, cap);
  if (st != TD16_OK) return st;
  return serve_wait(S, t, res, msg, cap);
}

int td16_serve_stats_get(td16_serve_stats *out, int resetLatency, char *msg, int cap) {
  if (out == nullptr) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "stats must not be null");
  auto S = current_server();
  if (!S) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  std::memset(out, 0, sizeof(*out));
  out->submitted = S->nSubmitted.load();
  out->completed = S->nCompleted.load();
  out->rejected = S->nRejected.load();
  out->in_flight = (int32_t)S->inFlight.load();
  {
    std::lock_guard<std::mutex> lk(S->admitMu);
    out->waiting = (int32_t)S->waiting.size();
  }
  out->p50_us = S->hist.quantile(0.50);
  out->p99_us = S->hist.quantile(0.99);
  out->mean_us = S->hist.mean();
  if (resetLatency) S->hist.reset();
  return TD16_OK;
}

int td16_serve_replay(const td16_request *reqs, const double *arrivalUs, int count,
                      td16_result *results, double *elapsedUs, char *msg, int cap) {
  auto S = current_server();
  if (!S) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  if (count < 0 || (count > 0 && (reqs == nullptr || arrivalUs == nullptr || results == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid replay arguments");
  S->hist.reset();
  std::vector<uint64_t> tickets((size_t)count);
  const auto start = std::chrono::steady_clock::now();
  for (int q = 0; q < count; ++q) {
    const auto due = start + std::chrono::nanoseconds((int64_t)(arrivalUs[q] * 1000.0));
    while (std::chrono::steady_clock::now() < due) {
#if defined(__x86_64__)
      __builtin_ia32_pause();
#endif
    }
    const int st = serve_submit_at(S, &reqs[q], 0, &tickets[(size_t)q], due, msg, cap);
    if (st != TD16_OK) return st;
  }
  for (int q = 0; q < count; ++q) {
    const int st = serve_wait(S, tickets[(size_t)q], &results[q], msg, cap);
    if (st != TD16_OK) return st;
  }
  if (elapsedUs)
    *elapsedUs = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - start)
                     .count();
  return TD16_OK;
}

int td16_serve_closed_loop(const td16_request *reqs, int count, int producers, int window,
                           td16_result *results, double *elapsedUs, char *msg, int cap) {
  auto S = current_server();
  if (!S) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  if (count < 0 || (count > 0 && (reqs == nullptr || results == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid closed-loop arguments");
  if (producers < 1 || producers > 1024)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "producers must be in 1..1024");
  if (window < 1 || window > 4096)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "window must be in 1..4096");
  S->hist.reset();
  std::atomic<int> firstErr{TD16_OK};
  std::mutex errMu;
  std::string err;
  const auto start = std::chrono::steady_clock::now();
  std::vector<std::thread> ths;
  for (int p = 0; p < producers; ++p)
    ths.emplace_back([&, p] {
      char m[256];
      // up to `window` requests of this producer in flight, collected in
      // submission order
      std::deque<std::pair<int, uint64_t>> open;
      auto fail = [&](int st) {
        int expected = TD16_OK;
        if (firstErr.compare_exchange_strong(expected, st)) {
          std::lock_guard<std::mutex> l(errMu);
          err = m;
        }
      };
      for (int q = p; q < count || !open.empty();) {
        if (firstErr.load(std::memory_order_relaxed) != TD16_OK) return;
        if (q < count && (int)open.size() < window) {
          uint64_t t = 0;
          const int st = serve_submit_at(S, &reqs[q], 0, &t, std::chrono::steady_clock::now(),
                                         m, sizeof m);
          if (st != TD16_OK) return fail(st);
          open.emplace_back(q, t);
          q += producers;
          continue;
        }
        const int st = serve_wait(S, open.front().second, &results[open.front().first], m,
                                  sizeof m);
        if (st != TD16_OK) return fail(st);
        open.pop_front();
      }
    });
  for (auto &t : ths) t.join();
  if (elapsedUs)
    *elapsedUs = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - start)
                     .count();
  if (firstErr.load() != TD16_OK) return set_msg(msg, cap, firstErr.load(), "%s", err.c_str());
  return TD16_OK;
}

int td16_serve_stop(char *msg, int cap) {
  std::lock_guard<std::mutex> lk(gServeMu);
  auto S = current_server();
  if (!S) return TD16_OK;
  auto since = [](std::chrono::steady_clock::time_point t0, int ms) {
    return ms > 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(ms);
  };
  // no new request is admitted from here on and blocked submitters give up;
  // admitted and waiting requests finish (within the deadline, if any)
  const auto t0 = std::chrono::steady_clock::now();
  bool drained = true;
  {
    std::unique_lock<std::mutex> l2(S->admitMu);
    S->stopping = true;
    S->cvSpace.notify_all();
    while (!(S->waiting.empty() && S->inFlight.load() == 0)) {
      if (S->failed.load() || since(t0, S->deadlineMs)) {
        drained = false;
        break;
      }
      S->cvIdle.wait_for(l2, std::chrono::milliseconds(20));
    }
    // the kernel exits once every request it took has finished, so the
    // stop request is safe to send in any case
    __atomic_store_n(&S->hHost->submitted, S->submittedSeq | kServeStop, __ATOMIC_RELEASE);
  }
  // requests that can no longer be delivered end with an error, not a hang
  if (!drained) fail_all(S.get(), "the served loop stopped before the request finished");
  hipSetDevice(S->device);
  // the kernel exits within one idle sleep of seeing the stop request; with
  // a deadline, wait at least a second for that
  const auto t1 = std::chrono::steady_clock::now();
  const int exitMs = S->deadlineMs > 0 ? std::max(S->deadlineMs, 1000) : 0;
  hipError_t e;
  while ((e = hipStreamQuery(S->stream)) == hipErrorNotReady && !since(t1, exitMs))
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  S->pollerExit.store(true, std::memory_order_release);
  S->cvWork.notify_all();
  if (S->poller.joinable()) S->poller.join();
  {
    std::lock_guard<std::mutex> pk(gServerPtrMu);
    gServer.reset();
  }
  if (!drained || e != hipSuccess) {
    // a kernel that may still run keeps its memory (it could write it yet)
    if (e == hipErrorNotReady)
      return set_msg(msg, cap, TD16_DEVICE_ERROR,
                     "served loop: the kernel did not stop within %d ms", exitMs);
    if (e != hipSuccess)
      return set_msg(msg, cap, TD16_DEVICE_ERROR, "served kernel failed: %s", hipGetErrorName(e));
    if (S->failed.load() && std::strcmp(S->failure, "the served loop stopped before the request "
                                                    "finished") != 0)
      return set_msg(msg, cap, TD16_DEVICE_ERROR, "served loop: %s", S->failure);
    return set_msg(msg, cap, TD16_DEVICE_ERROR,
                   "served loop: requests still running at the %d ms deadline were failed",
                   S->deadlineMs);
  }
  S->buf.release();
  hipHostFree(S->hHost);
  hipHostFree(S->hRing);
  hipHostFree(S->hRes);
  hipHostFree(S->hComp);
  hipStreamDestroy(S->stream);
  S->hHost = nullptr;
  S->hRing = nullptr;
  S->hRes = nullptr;
  S->hComp = nullptr;
  return TD16_OK;
}

}  // extern "C"

