// Kernels of the persistent search (the body is in td16_search.h) at up to
// five waves per SIMD, the queue initialisation kernel, and the launch
// wrappers the host runtime calls.  The six-waves-per-SIMD builds are
// compiled in td16_device_w6.hip.
#include "td16_search.h"

namespace td16 {

// Register budget for five waves per SIMD (96 VGPRs; the body needs 90 and
// nothing spills), so the same build runs 4 to 20 waves per CU.  The batched
// mix (about 6 nodes per request) runs fastest at 20.
__global__ void __launch_bounds__(kBlockThreads, 5)
td16_search_kernel(Params P) {
  __shared__ BlockLds lds;
  search_blocks<false>(P, lds);
}


// Counted build: the same search that also sums the candidate triples it
// evaluates (Control::candidates); for measurement launches.
__global__ void __launch_bounds__(kBlockThreads, 5)
td16_search_kernel_counted(Params P) {
  __shared__ BlockLds lds;
  search_blocks<true>(P, lds);
}

// The six-waves-per-SIMD builds (td16_device_w6.hip).
__global__ void td16_search_kernel_w6(Params P);
__global__ void td16_search_kernel_w6_counted(Params P);


// Every shard's slot q starts out free for producer position q.
__global__ void td16_queue_init_kernel(uint64_t *seq, uint64_t capPerShard, uint64_t total) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) seq[i] = i % capPerShard;
}

// Look-ahead claims (kLookAheadClaim requests held per atomic) save a claim
// round trip per request, which pays when every wavefront runs many
// requests; with few requests per wavefront a held request mostly leaves
// another wavefront without work (all 10,000 n = 10 distances, 2.4 requests
// per wavefront: 1.10 ms per launch without look-ahead against 1.20 ms with
// it), so a batch of fewer than kLookAheadMinPerWave requests per wavefront
// has none.  Otherwise the last 96 requests of each feed shard are claimed
// one at a time: a request held there could wait behind a long search while
// the feed drains.
constexpr int kLookAheadMinPerWave = 8;
int look_margin(const Params &p, int blocks) {
  const uint64_t waves = (uint64_t)blocks * kWavesPerBlock;
  return (uint64_t)p.numJobs >= kLookAheadMinPerWave * waves ? 96 : kNoLookAhead;
}

int launch_search(const Params &p0, int blocks, bool wide, void *stream) {
  Params p = p0;
  p.lookMargin = look_margin(p0, blocks);
  if (p.countCandidates && wide)
    hipLaunchKernelGGL(td16_search_kernel_w6_counted, dim3(blocks), dim3(kBlockThreads), 0,
                       (hipStream_t)stream, p);
  else if (p.countCandidates)
    hipLaunchKernelGGL(td16_search_kernel_counted, dim3(blocks), dim3(kBlockThreads), 0,
                       (hipStream_t)stream, p);
  else if (wide)
    hipLaunchKernelGGL(td16_search_kernel_w6, dim3(blocks), dim3(kBlockThreads), 0,
                       (hipStream_t)stream, p);
  else
    hipLaunchKernelGGL(td16_search_kernel, dim3(blocks), dim3(kBlockThreads), 0,
                       (hipStream_t)stream, p);
  return (int)hipGetLastError();
}

// Puts a batched launch's control block back in its starting state (what
// init_feed and a zeroed block give): the feed cursors at their shards'
// first requests and every counter at zero.  Captured ahead of the search in
// a batch's hipGraph, so that a replay needs no host-to-device copy.
__global__ void td16_control_reset_kernel(Control *c, uint32_t numJobs, uint32_t perShard) {
  const int t = threadIdx.x;
  if (t < kFeedShards) {
    const uint64_t lo = (uint64_t)t * perShard;
    c->feed[t].next = lo < numJobs ? (uint32_t)lo : numJobs;
  }
  if (t < kBalanceShards) {
    c->balance[t].v = 0;
    c->published[t].v = 0;
  }
  if (t == 0) {
    c->numJobs = numJobs;
    c->perShard = perShard;
    c->drained = 0;
    c->jobsDone = 0;
    c->consumed = 0;
    c->completions = 0;
    c->lastPoll = 0;
    c->submittedMirror = 0;
    c->stopMirror = 0;
    c->candidates = 0;
  }
}

int launch_control_reset(Control *c, uint32_t numJobs, uint32_t perShard, void *stream) {
  hipLaunchKernelGGL(td16_control_reset_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, c,
                     numJobs, perShard);
  return (int)hipGetLastError();
}

int launch_queue_init(uint64_t *seq, uint64_t capPerShard, void *stream) {
  const uint64_t total = capPerShard * kQueueShards;
  const uint64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(td16_queue_init_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, seq, capPerShard, total);
  return (int)hipGetLastError();
}

int search_kernel_occupancy(int *blocksPerCU, int *blocksPerCUWide) {
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocksPerCU, td16_search_kernel, kBlockThreads, 0);
  if (e != hipSuccess) return (int)e;
  return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCUWide,
                                                           td16_search_kernel_w6,
                                                           kBlockThreads, 0);
}

}  // namespace td16
