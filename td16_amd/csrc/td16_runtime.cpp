// Host runtime of libtd16: devices, slots, the blocking one-request-per-launch
// search, the batched persistent launch and the served loop.
//
// Reference counterparts (emitter/src/native/): slot_manager.cu (streams,
// events, device checks), sorting_search.cu:714-972 (slot buffers, queue
// sizing, performGpuSortingSearch), core/cuda_error.cuh (error classes),
// core/cuda_memory.cuh (owned device/pinned buffers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
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
  int blocksPerCU = 0;      // resident blocks of the 4-waves-per-SIMD kernel
  int blocksPerCUWide = 0;  // of the 5-waves-per-SIMD build
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
  if (o.split_depth > 0) P.splitSlack = o.split_depth;
  if (o.donate_max > 0) P.donateMax = std::min(o.donate_max, 64);
  if (o.check_every > 0) P.checkEvery = std::min(o.check_every, 1024);
}

void batch_grid(int wavesPerCU, int cus, int blocksPerCU, int blocksPerCUWide, int *blocks,
                bool *wide) {
  // 0 = the measured best for the batched mix: 16 waves per CU (4 per SIMD,
  // the most the 128-VGPR build allows)
  const int wantWaves = wavesPerCU > 0 ? wavesPerCU : 16;
  // above the 4-waves-per-SIMD kernel's residency (16 waves per CU) the
  // 5-waves-per-SIMD build runs
  const int wantBlocks = std::max(1, wantWaves / kWavesPerBlock);
  *wide = wantBlocks > blocksPerCU && blocksPerCUWide > blocksPerCU;
  const int perCU = std::min(wantBlocks, *wide ? blocksPerCUWide : blocksPerCU);
  *blocks = cus * std::max(1, perCU);
}

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
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Buffers buf;
  int64_t queueBytes = 0;
  hipGraphExec_t graph = nullptr;
  int graphBlocks = 0;
  uint32_t graphJobs = 0;
};

std::mutex gBatchMu;
BatchCtx gBatch;

void reset_batch_graph() {
  if (gBatch.graph) hipGraphExecDestroy(gBatch.graph);
  gBatch.graph = nullptr;
}

// ---------------------------------------------------------------------------
// Served loop
// ---------------------------------------------------------------------------

struct Ticket {
  bool ready = false;
  td16_result res;
  std::chrono::steady_clock::time_point t0;      // arrival
  std::chrono::steady_clock::time_point tAdmit;  // published to the device
};

struct Server {
  bool running = false;
  int device = -1;
  DeviceInfo dev;
  hipStream_t stream = nullptr;
  Buffers buf;
  uint32_t slots = 0, ringCap = 0, queueCap = 0;
  ServeHost *hHost = nullptr;
  Job *hRing = nullptr;
  DevResult *hRes = nullptr;
  uint32_t *hComp = nullptr;
  ServeHost *dHost = nullptr;
  Job *dRing = nullptr;
  DevResult *dRes = nullptr;
  uint32_t *dComp = nullptr;

  std::mutex mu;
  std::condition_variable cvDone, cvSpace, cvWork;
  std::deque<std::pair<uint64_t, Job>> waiting;
  std::vector<uint32_t> freeSlots;
  std::vector<uint64_t> slotTicket;
  std::unordered_map<uint64_t, Ticket> tickets;
  uint64_t nextTicket = 1;
  uint64_t submittedSeq = 0;
  uint32_t inFlight = 0;
  bool stopping = false;
  std::thread poller;
  std::atomic<bool> pollerExit{false};

  uint64_t nSubmitted = 0, nCompleted = 0, nRejected = 0;
  std::vector<double> latUs;
};

std::mutex gServeMu;  // start/stop
Server *gServer = nullptr;

// Publishes waiting requests while slots are free (caller holds S->mu).
// Called by submit and by the completion poller, so no thread hop sits
// between a free slot and the next request.
void admit_locked(Server *S) {
  bool any = false;
  while (!S->waiting.empty() && S->inFlight < S->slots) {
    auto item = S->waiting.front();
    S->waiting.pop_front();
    const uint32_t slot = S->freeSlots.back();
    S->freeSlots.pop_back();
    S->slotTicket[slot] = item.first;
    auto tk = S->tickets.find(item.first);
    if (tk != S->tickets.end()) tk->second.tAdmit = std::chrono::steady_clock::now();
    __atomic_fetch_add(&S->inFlight, 1u, __ATOMIC_RELAXED);
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

// Completion poller: spins on the next completion record while requests are
// in flight (a dedicated core buys microseconds of latency); with nothing in
// flight it parks on cvWork, which admission signals.
void poll_loop(Server *S) {
  uint64_t c = 0;
  while (!S->pollerExit.load(std::memory_order_acquire)) {
    const uint32_t e = (uint32_t)(c & (S->ringCap - 1));
    const uint32_t seq = __atomic_load_n(&S->hComp[2 * e + 1], __ATOMIC_ACQUIRE);
    if (seq != (uint32_t)(c + 1)) {
      if (__atomic_load_n(&S->inFlight, __ATOMIC_RELAXED) == 0) {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cvWork.wait_for(lk, std::chrono::milliseconds(1), [&] {
          return S->inFlight > 0 || S->pollerExit.load(std::memory_order_acquire);
        });
      } else {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#endif
      }
      continue;
    }
    const uint32_t slot = __atomic_load_n(&S->hComp[2 * e], __ATOMIC_ACQUIRE);
    DevResult r;
    std::memcpy(&r, &S->hRes[slot], sizeof(DevResult));
    const auto now = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      const uint64_t t = S->slotTicket[slot];
      auto it = S->tickets.find(t);
      if (it != S->tickets.end()) {
        decode_result(r, &it->second.res);
        it->second.ready = true;
        const double us = std::chrono::duration<double, std::micro>(now - it->second.t0).count();
        it->second.res.latency_us = us;
        it->second.res.queue_us =
            std::chrono::duration<double, std::micro>(it->second.tAdmit - it->second.t0).count();
        S->latUs.push_back(us);
      }
      S->freeSlots.push_back(slot);
      __atomic_fetch_sub(&S->inFlight, 1u, __ATOMIC_RELAXED);
      S->nCompleted++;
      admit_locked(S);
    }
    S->cvDone.notify_all();
    ++c;
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
      hipEventDestroy(gBatch.ev0);
      hipEventDestroy(gBatch.ev1);
      hipStreamDestroy(gBatch.stream);
    }
    gBatch.buf.release();
    gBatch = BatchCtx();
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

// Host side of a batch: validates and relabels every request, answers the
// trivial ones in `results`, and returns the rest in feed order (`where[q]`
// is the caller's index of jobs[q]).
int prepare_batch(const td16_request *reqs, int count, td16_result *results,
                  std::vector<Job> &jobs, std::vector<int> &where, char *msg, int cap) {
  jobs.clear();
  where.clear();
  jobs.reserve((size_t)count);
  for (int q = 0; q < count; ++q) {
    Job job;
    int trivial = 0;
    const int st = make_job_checked(&reqs[q], &job, &trivial, msg, cap);
    if (st != TD16_OK) {
      std::memset(&results[q], 0, sizeof(td16_result));
      results[q].status = st;
      return st;
    }
    if (trivial) {
      trivial_result(trivial, &results[q]);
      continue;
    }
    job.tag = (uint32_t)q;
    jobs.push_back(job);
    where.push_back(q);
  }
  const uint32_t nj = (uint32_t)jobs.size();
  // Feed order: longest expected search first.  Heavy requests then start
  // with the launch and are split once the feed drains, instead of starting
  // late and becoming the launch's tail.  Long searches come from requests
  // without slack: the first depth at which a witness is possible at all
  // (every move gaining two odd cycles) is the move bound itself, so every
  // path must be perfect; among those the deeper bound is the longer search
  // (on the batched mix every request that needed 1,024 nodes or more had no
  // slack, averaging 27,201 nodes at bound 8 against 341 at 7 and 22 at 6,
  // and about 5 with slack).  Those go first, deepest first; the rest follow
  // by move bound, then symbols.  The sorted list is dealt round-robin to the
  // feed shards (contiguous ranges of the table, see init_feed), so every
  // shard runs its own longest-first sequence.
  {
    std::vector<uint32_t> order(nj);
    for (uint32_t q = 0; q < nj; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      const Job &x = jobs[a], &y = jobs[b];
      const bool tx = (x.flags & kFlagTight) != 0, ty = (y.flags & kFlagTight) != 0;
      if (tx != ty) return tx;
      return x.maxDepth != y.maxDepth ? x.maxDepth > y.maxDepth : x.N > y.N;
    });
    const uint32_t per = (nj + kFeedShards - 1) / kFeedShards;
    std::vector<Job> dealtJobs(nj);
    std::vector<int> dealtWhere(nj);
    uint32_t next = 0;
    for (uint32_t t = 0; next < nj; ++t) {
      for (uint32_t s = 0; s < (uint32_t)kFeedShards && next < nj; ++s) {
        const uint64_t pos = (uint64_t)s * per + t;
        const uint64_t end = std::min<uint64_t>((uint64_t)(s + 1) * per, nj);
        if (pos >= end) continue;
        dealtJobs[pos] = jobs[order[next]];
        dealtWhere[pos] = where[order[next]];
        ++next;
      }
    }
    jobs.swap(dealtJobs);
    where.swap(dealtWhere);
  }
  return TD16_OK;
}

namespace {

// Device side of a batch: `o.repeats` persistent launches over the prepared
// requests; results of the last one land in results[where[q]].
int run_batch(const std::vector<Job> &jobs, const std::vector<int> &where,
              const td16_batch_options &o, td16_result *results, td16_batch_stats *stats,
              char *msg, int cap) {
  const uint32_t nj = (uint32_t)jobs.size();
  td16_batch_stats localStats;
  std::memset(&localStats, 0, sizeof(localStats));
  std::lock_guard<std::mutex> lk(gBatchMu);
  DeviceInfo dev;
  int st = device_info(o.device, &dev, msg, cap);
  if (st != TD16_OK) return st;
  TD16_HIP(hipSetDevice(o.device));
  if (gBatch.device != o.device || gBatch.buf.jobsCap < std::max<uint32_t>(nj, 1) ||
      gBatch.queueBytes != o.queue_bytes) {
    reset_batch_graph();
    if (gBatch.stream) {
      hipSetDevice(gBatch.device);
      hipStreamSynchronize(gBatch.stream);
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
    TD16_HIP(hipEventCreate(&gBatch.ev0));
    TD16_HIP(hipEventCreate(&gBatch.ev1));
    uint32_t capJobs = 1024;
    while (capJobs < nj) capJobs *= 2;
    st = alloc_buffers(gBatch.buf, o.device, capJobs, o.queue_bytes, gBatch.stream, msg, cap);
    if (st != TD16_OK) return st;
  }
  Buffers &b = gBatch.buf;
  int blocks = 0;
  bool wide = false;
  batch_grid(o.waves_per_cu, dev.cus, dev.blocksPerCU, dev.blocksPerCUWide, &blocks, &wide);
  localStats.blocks = blocks;
  localStats.waves = blocks * kWavesPerBlock;
  if (nj == 0) {
    if (stats) *stats = localStats;
    return TD16_OK;
  }
  std::memcpy(b.hJobs, jobs.data(), sizeof(Job) * nj);
  TD16_HIP(hipMemcpyAsync(b.dJobs, b.hJobs, sizeof(Job) * nj, hipMemcpyHostToDevice, gBatch.stream));
  std::memset(b.hCtrl, 0, sizeof(Control));
  const size_t ctrlBytes = init_feed(b.hCtrl, nj);
  Params P = make_params(b, dev);
  apply_batch_knobs(P, o);

  if (o.use_graph) {
    reset_batch_graph();
    hipGraph_t g;
    TD16_HIP(hipStreamBeginCapture(gBatch.stream, hipStreamCaptureModeThreadLocal));
    // the graph holds the kernel alone; the copies and the events that time
    // the kernel stay outside it
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
  }
  localStats.kernel_ms = kernelMs / o.repeats;
  localStats.wall_ms = wallMs / o.repeats;
  for (uint32_t q = 0; q < nj; ++q) {
    const DevResult &r = b.hResults[q];
    td16_result &out = results[where[q]];
    decode_result(r, &out);
    localStats.expanded += r.expanded;
    localStats.solutions += r.solutions;
    localStats.donations += (uint64_t)r.split;
  }
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
  const td16_batch_options o = normalize_batch_options(opt);
  std::vector<Job> jobs;
  std::vector<int> where;
  const int st = prepare_batch(reqs, count, results, jobs, where, msg, cap);
  if (st != TD16_OK) return st;
  return run_batch(jobs, where, o, results, stats, msg, cap);
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
    int rc = run_batch(jobs, where, o, cur.data(), &bs, msg, cap);
    if (rc != TD16_OK) return rc;
    std::vector<double> t;
    for (int r = 0; r < repeats; ++r) {
      rc = run_batch(jobs, where, o, cur.data(), &bs, msg, cap);
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

  static const int kWaves[] = {4, 8, 12, 16, 20};
  static const int kSplit[] = {2, 3, 4, 5, 6, 7};
  static const int kDonate[] = {4, 8, 16, 32, 64};
  static const int kCheck[] = {8, 16, 32};
  int bw = o.waves_per_cu > 0 ? o.waves_per_cu : 16;
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
      const int nv = knob == 0 ? 5 : knob == 1 ? 6 : knob == 2 ? 5 : 3;
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
  if (opt->slot_bytes <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "slot budget must be positive");
  std::lock_guard<std::mutex> lk(gServeMu);
  if (gServer) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "the served loop is already running");
  DeviceInfo dev;
  int st = device_info(opt->device, &dev, msg, cap);
  if (st != TD16_OK) return st;
  TD16_HIP(hipSetDevice(opt->device));
  size_t freeB = 0, totalB = 0;
  TD16_HIP(hipMemGetInfo(&freeB, &totalB));
  // admission budget: every slot may hold slot_bytes of HBM for its work items
  const unsigned long long need = (unsigned long long)opt->slot_bytes * (unsigned long long)opt->slots;
  const unsigned long long queueBytes = std::min<unsigned long long>(need, 4ULL << 30);
  if (queueBytes > freeB)
    return set_msg(msg, cap, TD16_OUT_OF_MEMORY,
                   "device %d has %zu free bytes, but the served loop requires %llu bytes",
                   opt->device, freeB, queueBytes);
  auto S = std::make_unique<Server>();
  S->device = opt->device;
  S->dev = dev;
  S->slots = (uint32_t)opt->slots;
  S->queueCap = (uint32_t)opt->queue_capacity;
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
  S->slotTicket.assign(S->slots, 0);

  std::memset(S->buf.hCtrl, 0, sizeof(Control));
  TD16_HIP(hipMemcpyAsync(S->buf.dCtrl, S->buf.hCtrl, sizeof(Control), hipMemcpyHostToDevice,
                          S->stream));
  Params P = make_params(S->buf, dev);
  P.served = 1;
  // idle wavefronts sleep up to ~51 us: with a request arriving every few
  // microseconds some wavefront is always awake to claim it, and the quiet
  // memory system makes every step of a request faster (20k/s: p50 18 us,
  // p99 42-51 us, against 92 / 141 us at ~3.2 us sleeps)
  P.idleBackoff = 128;
  // served requests are mostly light and arrive while others run: checks
  // every 32 expansions keep the p99 lower (20k/s: 141-143 against 164-171
  // us at 16; the median is the same)
  P.checkEvery = 32;
  P.ringCap = ring;
  P.host = S->dHost;
  P.hostJobs = S->dRing;
  P.hostResults = S->dRes;
  P.hostCompletionIdx = S->dComp;
  // 0 = the measured best for latency: 4 waves per CU (one per SIMD)
  const int wantWaves = opt->waves_per_cu > 0 ? opt->waves_per_cu : 4;
  const int perCU = std::max(1, std::min(dev.blocksPerCU, wantWaves / kWavesPerBlock));
  TD16_HIP((hipError_t)launch_search(P, dev.cus * perCU, false, S->stream));
  S->running = true;
  Server *raw = S.release();
  raw->poller = std::thread(poll_loop, raw);
  gServer = raw;
  return TD16_OK;
}

namespace td16 {
namespace {
int serve_submit_at(const td16_request *req, int nonblocking, uint64_t *ticket,
                    std::chrono::steady_clock::time_point arrival, char *msg, int cap);
}
}  // namespace td16

int td16_serve_submit(const td16_request *req, int nonblocking, uint64_t *ticket, char *msg,
                      int cap) {
  return td16::serve_submit_at(req, nonblocking, ticket, std::chrono::steady_clock::now(), msg,
                               cap);
}

namespace td16 {
namespace {
int serve_submit_at(const td16_request *req, int nonblocking, uint64_t *ticket,
                    std::chrono::steady_clock::time_point arrival, char *msg, int cap) {
  Server *S = gServer;
  if (S == nullptr) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  if (req == nullptr || ticket == nullptr)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "request and ticket must not be null");
  Job job;
  int trivial = 0;
  const int st = make_job_checked(req, &job, &trivial, msg, cap);
  if (st != TD16_OK) return st;
  std::unique_lock<std::mutex> lk(S->mu);
  const uint64_t t = S->nextTicket++;
  Ticket &tk = S->tickets[t];
  tk.t0 = arrival;
  *ticket = t;
  S->nSubmitted++;
  if (trivial) {
    trivial_result(trivial, &tk.res);
    tk.ready = true;
    S->nCompleted++;
    S->latUs.push_back(0.0);
    return TD16_OK;
  }
  // GPUSortingSearch.java:46-52: pending capacity = slots + queue capacity
  const size_t limit = (size_t)S->queueCap + S->slots;
  if (S->waiting.size() >= limit) {
    if (nonblocking) {
      S->tickets.erase(t);
      S->nSubmitted--;
      S->nRejected++;
      return set_msg(msg, cap, TD16_QUEUE_FULL, "the served loop's pending queue is full");
    }
    S->cvSpace.wait(lk, [&] { return S->waiting.size() < limit || S->stopping; });
  }
  S->waiting.emplace_back(t, job);
  admit_locked(S);
  return TD16_OK;
}
}  // namespace
}  // namespace td16

int td16_serve_wait(uint64_t ticket, td16_result *res, char *msg, int cap) {
  Server *S = gServer;
  if (S == nullptr) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "unknown ticket");
  S->cvDone.wait(lk, [&] { return it->second.ready; });
  *res = it->second.res;
  S->tickets.erase(it);
  return TD16_OK;
}

int td16_serve_search(const td16_request *req, td16_result *res, char *msg, int cap) {
  uint64_t t = 0;
  const int st = td16_serve_submit(req, 0, &t, msg, cap);
  if (st != TD16_OK) return st;
  return td16_serve_wait(t, res, msg, cap);
}

int td16_serve_stats_get(td16_serve_stats *out, int resetLatency, char *msg, int cap) {
  Server *S = gServer;
  if (S == nullptr) return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  std::lock_guard<std::mutex> lk(S->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = S->nSubmitted;
  out->completed = S->nCompleted;
  out->rejected = S->nRejected;
  out->in_flight = (int32_t)S->inFlight;
  out->waiting = (int32_t)S->waiting.size();
  if (!S->latUs.empty()) {
    std::vector<double> v = S->latUs;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      size_t i = (size_t)(p * (double)(v.size() - 1) + 0.5);
      return v[std::min(i, v.size() - 1)];
    };
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    double s = 0;
    for (double x : v) s += x;
    out->mean_us = s / (double)v.size();
  }
  if (resetLatency) S->latUs.clear();
  return TD16_OK;
}

int td16_serve_replay(const td16_request *reqs, const double *arrivalUs, int count,
                      td16_result *results, double *elapsedUs, char *msg, int cap) {
  if (gServer == nullptr)
    return set_msg(msg, cap, TD16_NOT_INITIALIZED, "the served loop is not running");
  if (count < 0 || (count > 0 && (reqs == nullptr || arrivalUs == nullptr || results == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid replay arguments");
  td16_serve_stats tmp;
  td16_serve_stats_get(&tmp, 1, nullptr, 0);  // reset the latency window
  std::vector<uint64_t> tickets((size_t)count);
  const auto start = std::chrono::steady_clock::now();
  for (int q = 0; q < count; ++q) {
    const auto due = start + std::chrono::nanoseconds((int64_t)(arrivalUs[q] * 1000.0));
    while (std::chrono::steady_clock::now() < due) {
#if defined(__x86_64__)
      __builtin_ia32_pause();
#endif
    }
    const int st = td16::serve_submit_at(&reqs[q], 0, &tickets[(size_t)q], due, msg, cap);
    if (st != TD16_OK) return st;
  }
  for (int q = 0; q < count; ++q) {
    const int st = td16_serve_wait(tickets[(size_t)q], &results[q], msg, cap);
    if (st != TD16_OK) return st;
  }
  if (elapsedUs)
    *elapsedUs = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - start)
                     .count();
  return TD16_OK;
}

int td16_serve_stop(char *msg, int cap) {
  std::lock_guard<std::mutex> lk(gServeMu);
  Server *S = gServer;
  if (S == nullptr) return TD16_OK;
  {
    std::unique_lock<std::mutex> l2(S->mu);
    S->stopping = true;
    S->cvSpace.notify_all();
    S->cvDone.wait(l2, [&] { return S->waiting.empty() && S->inFlight == 0; });
  }
  __atomic_store_n(&S->hHost->submitted, S->submittedSeq | kServeStop, __ATOMIC_RELEASE);
  hipSetDevice(S->device);
  hipError_t e = hipStreamSynchronize(S->stream);
  S->pollerExit.store(true, std::memory_order_release);
  S->poller.join();
  S->buf.release();
  hipHostFree(S->hHost);
  hipHostFree(S->hRing);
  hipHostFree(S->hRes);
  hipHostFree(S->hComp);
  hipStreamDestroy(S->stream);
  gServer = nullptr;
  delete S;
  if (e != hipSuccess)
    return set_msg(msg, cap, TD16_DEVICE_ERROR, "served kernel failed: %s", hipGetErrorName(e));
  return TD16_OK;
}

}  // extern "C"
