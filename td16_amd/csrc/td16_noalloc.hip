// td16_batch_noalloc: the batched persistent search without any device
// allocation.  Every device buffer a batch needs -- control block, ring
// counters and slots, request table, request state, result staging and the
// triple tables -- is a fixed region of one file-scope device array,
// `scratch`.  Each GPU holds its own copy of that array for the life of the
// process; its address on a device comes from hipGetSymbolAddress, taken
// once per device with that device current.  The array's name is a
// host-side stand-in and is never passed to a kernel or a copy.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/td16_abi.h"
#include "td16_internal.h"

namespace td16 {
namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ring slots per shard: a full ring only stops donation (the children stay
// with the wavefront that holds them), so this bounds memory, not results
constexpr uint64_t kRingPerShard = TD16_NOALLOC_QUEUE_ITEMS / kQueueShards;
static_assert(kRingPerShard * kQueueShards == TD16_NOALLOC_QUEUE_ITEMS, "ring split");
constexpr size_t kMaxReq = TD16_NOALLOC_MAX_REQUESTS;

constexpr size_t kOffCtrl = 0;
constexpr size_t kOffQueue = align256(kOffCtrl + sizeof(Control));
constexpr size_t kOffLex = align256(kOffQueue + sizeof(QueueState) * kQueueShards);
constexpr size_t kOffSeq = align256(kOffLex + sizeof(uint16_t) * kLexEntries);
constexpr size_t kOffItems = align256(kOffSeq + sizeof(uint64_t) * TD16_NOALLOC_QUEUE_ITEMS);
constexpr size_t kOffJobs = align256(kOffItems + sizeof(Item) * TD16_NOALLOC_QUEUE_ITEMS);
constexpr size_t kOffStates = align256(kOffJobs + sizeof(Job) * kMaxReq);
constexpr size_t kOffResults = align256(kOffStates + sizeof(JobState) * kMaxReq);
constexpr size_t kScratchBytes = align256(kOffResults + sizeof(DevResult) * kMaxReq);
static_assert(kScratchBytes <= TD16_NOALLOC_SCRATCH_BYTES, "scratch size as declared");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

struct DeviceScratch {
  unsigned char *ptr = nullptr;  // this device's copy of `scratch`
  int cus = 0;
  int blocksPerCU = 0;
  int blocksPerCUWide = 0;
};

// Calls are serialized per process: they share each GPU's array, and two
// batches in it at once would overwrite each other's state.
std::mutex gMu;
DeviceScratch gDev[kMaxDevices];

// Host staging reused by every call (the lock above covers them).
Control gHostCtrl;
const QueueState kZeroQueue[kQueueShards] = {};

#define NA_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_msg(msg, cap, TD16_DEVICE_ERROR, "%s:%d: HIP error %s (%s)", __FILE__,     \
                     __LINE__, hipGetErrorName(e_), hipGetErrorString(e_));                 \
  } while (0)

// Once per device, with that device current: the array's address on it, the
// grid shape, and the triple tables copied into their region.
int setup_device(int device, DeviceScratch &d, char *msg, int cap) {
  if (d.ptr != nullptr) return TD16_OK;
  void *ptr = nullptr;
  NA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (ptr == nullptr)
    return set_msg(msg, cap, TD16_DEVICE_ERROR, "td16_batch_noalloc: no scratch address");
  int cus = 0, bpc = 0, bpcWide = 0;
  NA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  NA_HIP((hipError_t)search_kernel_occupancy(&bpc, &bpcWide));
  uint16_t lex[kLexEntries];
  build_lex_table(lex);
  NA_HIP(hipMemcpy((unsigned char *)ptr + kOffLex, lex, sizeof(lex), hipMemcpyHostToDevice));
  d.cus = cus;
  d.blocksPerCU = bpc < 1 ? 1 : bpc;
  d.blocksPerCUWide = bpcWide;
  d.ptr = (unsigned char *)ptr;
  return TD16_OK;
}

}  // namespace
}  // namespace td16

using namespace td16;

extern "C" int td16_batch_noalloc(const td16_request *reqs, int count, td16_result *results,
                                  const td16_batch_options *opt, td16_batch_stats *stats,
                                  char *msg, int cap) {
  if (count < 0 || (count > 0 && (reqs == nullptr || results == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid batch arguments");
  if (count > TD16_NOALLOC_MAX_REQUESTS)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "td16_batch_noalloc: %d requests exceed its fixed maximum of %d per call",
                   count, TD16_NOALLOC_MAX_REQUESTS);
  td16_batch_options o;
  std::memset(&o, 0, sizeof(o));
  if (opt) o = *opt;
  if (o.use_graph)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "td16_batch_noalloc: use_graph must be 0 (instantiating a graph allocates)");
  if (o.repeats < 1) o.repeats = 1;
  if (o.device < 0 || o.device >= kMaxDevices)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "td16_batch_noalloc: invalid device %d",
                   o.device);

  std::vector<Job> jobs;
  std::vector<int> where;
  int st = prepare_batch(reqs, count, results, jobs, where, msg, cap);
  if (st != TD16_OK) return st;
  const uint32_t nj = (uint32_t)jobs.size();

  std::lock_guard<std::mutex> lk(gMu);
  int ndev = 0;
  NA_HIP(hipGetDeviceCount(&ndev));
  if (o.device >= ndev)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "td16_batch_noalloc: invalid device %d",
                   o.device);
  NA_HIP(hipSetDevice(o.device));
  DeviceScratch &d = gDev[o.device];
  st = setup_device(o.device, d, msg, cap);
  if (st != TD16_OK) return st;

  td16_batch_stats local;
  std::memset(&local, 0, sizeof(local));
  int blocks = 0;
  bool wide = false;
  batch_grid(o.waves_per_cu, (uint32_t)count, d.cus, d.blocksPerCU, d.blocksPerCUWide, &blocks,
             &wide);
  local.blocks = blocks;
  local.waves = blocks * kWavesPerBlock;
  if (nj == 0) {
    if (stats) *stats = local;
    return TD16_OK;
  }

  unsigned char *base = d.ptr;  // never the array's own name
  Params P;
  std::memset(&P, 0, sizeof(P));
  P.jobs = reinterpret_cast<const Job *>(base + kOffJobs);
  P.states = reinterpret_cast<JobState *>(base + kOffStates);
  P.results = reinterpret_cast<DevResult *>(base + kOffResults);
  P.ctrl = reinterpret_cast<Control *>(base + kOffCtrl);
  P.queue = reinterpret_cast<QueueState *>(base + kOffQueue);
  P.items = reinterpret_cast<Item *>(base + kOffItems);
  P.seq = reinterpret_cast<uint64_t *>(base + kOffSeq);
  P.qcap = kRingPerShard;
  P.lex = reinterpret_cast<const uint16_t *>(base + kOffLex);
  default_knobs(P);
  apply_batch_knobs(P, o);

  hipStream_t stream = nullptr;  // the null stream: nothing to create
  // The array keeps the previous call's contents.  What a launch reads
  // before writing is reset here: the ring counters and slot sequence
  // numbers (a copy and a fill kernel, no memset) and the control block's
  // leading part (feed cursors, demand and completion counters).  Request
  // state is written before it is read; every result word is written.
  NA_HIP(hipMemcpyAsync((void *)P.jobs, jobs.data(), sizeof(Job) * nj, hipMemcpyHostToDevice,
                        stream));
  NA_HIP(hipMemcpyAsync(P.queue, kZeroQueue, sizeof(kZeroQueue), hipMemcpyHostToDevice, stream));
  NA_HIP((hipError_t)launch_queue_init(P.seq, kRingPerShard, stream));
  std::memset(&gHostCtrl, 0, sizeof(gHostCtrl));
  const size_t ctrlBytes = init_feed(&gHostCtrl, nj);
  feed_params(P, gHostCtrl);

  double wallMs = 0.0;
  for (int rep = 0; rep < o.repeats; ++rep) {
    const auto w0 = std::chrono::steady_clock::now();
    NA_HIP(hipMemcpyAsync(P.ctrl, &gHostCtrl, ctrlBytes, hipMemcpyHostToDevice, stream));
    NA_HIP((hipError_t)launch_search(P, blocks, wide, stream));
    NA_HIP(hipStreamSynchronize(stream));
    wallMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0)
                  .count();
  }
  std::vector<DevResult> out(nj);
  NA_HIP(hipMemcpy(out.data(), P.results, sizeof(DevResult) * nj, hipMemcpyDeviceToHost));
  NA_HIP(hipMemcpy(&local.candidates, &P.ctrl->candidates, sizeof(uint64_t),
                   hipMemcpyDeviceToHost));
  local.kernel_ms = wallMs / o.repeats;  // host-timed: no events are created
  local.wall_ms = wallMs / o.repeats;
  std::vector<uint32_t> pos;
  positions_of(where.data(), nj, (uint32_t)count, pos);
  decode_batch(out.data(), pos.data(), (uint32_t)count, results, o.host_threads, &local.expanded,
               &local.solutions, &local.donations);
  if (stats) *stats = local;
  return TD16_OK;
}
