// Internal layout shared by the HIP kernels and the host runtime of libtd16.
//
// Search state
//   A search is carried in "label" form: the beta sequence of the request is
//   relabelled so that label(beta0[p]) = p.  The start sequence is then the
//   identity 0..n-1 (n = N-1) with the sentinel label n fixed at position n
//   (moves t(i,j,k) have k <= n, so position n never moves).  Positions
//   0..n-1 are nibble-packed into one uint64 (position p in bits 4p..4p+3;
//   n <= 16).  The request's target rho = omega0 o beta0 becomes a map R on
//   labels, and omega at any node is omega(y) = R(pred(y)), pred being the
//   cyclic predecessor in the sequence.  The reference keeps beta's function
//   form implicit the same way through rho^-1
//   (emitter/src/native/backtracking/sorting_search.cu:542-561, 899-913).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/td16_abi.h"

namespace td16 {

constexpr int kMaxN = 17;
constexpr int kMaxDepth = 9;
constexpr int kMaxChunks = 11;      // ceil(C(17,3) / 64)
static_assert(kMaxChunks <= 32, "expand() keeps a node's masks in lanes ch and 32 + ch");
constexpr int kLexEntries = 3060;   // sum over m = 3..17 of C(m,3) = C(18,4)
constexpr int kWavesPerBlock = 4;
constexpr int kBlockThreads = 64 * kWavesPerBlock;

enum JobFlags : uint8_t {
  kFlagAllMoves = 1,   // every position triple (else the reference filter)
  kFlagCountAll = 2,   // no early exit
  kFlagIddfs = 4,      // iterative deepening on the depth limit
  kFlagLexOrder = 8,   // FIRST search keeps the lexicographic order throughout
  kFlagTight = 16,     // FIRST search without slack: the move bound is the first
                       // depth at which a witness is possible (a long search)
};

// Requests without slack and with a move bound of at least kDeepBound are
// the long searches (tens of thousands of nodes on average on the batched
// mix): they are split at the first check, and a per-request launch gives
// them the slot's whole grid.
constexpr int kDeepBound = 8;

// One request as the kernel reads it (host-written, 128 B).
struct alignas(16) Job {
  uint64_t pi0;          // start sequence (labels), nibble-packed
  uint8_t R[20];         // omega(y) = R[pred(y)]; R[0..N-1] used
  uint8_t N;             // symbols = n + 1 (3..17)
  uint8_t maxDepth;      // move bound (IDDFS: last limit tried)
  uint8_t flags;         // JobFlags
  uint8_t firstLimit;    // IDDFS: first limit
  uint32_t sat[10];      // bit e of sat[d]: a depth-d child with e odd cycles is a witness
  uint32_t reach[10];    // bit e of reach[d]: such a child may still lead to one
  uint32_t pad[3];
  uint32_t tag;          // served loop: low bits of the ticket
};
static_assert(sizeof(Job) == 128, "Job layout");

// Mutable per-request state on the device (64 B).
struct alignas(64) JobState {
  int32_t pending;       // work items of this request not yet finished
  int32_t found;         // 0/1, set once by the wavefront that records the witness
  int32_t limit;         // current depth limit
  int32_t split;         // subtrees donated from this request
  uint64_t expanded;
  uint64_t solutions;
  int32_t depth;         // witness length
  uint16_t moves[10];    // witness, packed i | j<<5 | k<<10
  uint8_t pad[4];
  uint64_t t0;           // device real-time counter (100 MHz) when the request started
  uint8_t pad2[56];
};
static_assert(sizeof(JobState) == 128, "JobState layout");

// What the host reads back per request (80 B).
struct alignas(16) DevResult {
  int32_t found;
  int32_t depth;
  int32_t split;
  int32_t limit;
  uint64_t expanded;
  uint64_t solutions;
  uint16_t moves[10];
  uint32_t tag;
  uint32_t seq;          // served loop: completion sequence + 1 once written
  uint8_t pad[4];
  uint64_t t0, t1;       // device real-time counter at start and at publication
};
static_assert(sizeof(DevResult) == 80, "DevResult layout");

// A donated subtree: a node still to be expanded (48 B).
struct alignas(16) Item {
  uint64_t pi;
  uint32_t job;
  uint32_t depth;        // moves already applied
  uint16_t moves[10];    // the moves that led here, packed
  uint32_t pad[3];
};
static_assert(sizeof(Item) == 48, "Item layout");

// Shared work queue: kQueueShards bounded lock-free rings (per-slot sequence
// numbers, D. Vyukov's bounded MPMC scheme), so that idle wavefronts poll and
// claim on different cache lines.  Shard s owns ring slots
// [s*cap, (s+1)*cap).  head/tail only grow and persist across launches;
// every launch ends with every shard empty.
constexpr int kQueueShards = 64;
struct alignas(128) QueueState {
  uint64_t head;
  uint8_t pad0[120];
  uint64_t tail;
  uint8_t pad1[120];
};

// Device-scope atomics on one address serialize at the memory side (about
// 12 ns each on MI355X) while atomics on different 256-byte lines proceed in
// parallel, so every counter that many wavefronts update is either sharded
// over such lines or updated rarely.

// Batched feed: the request table is cut into kFeedShards contiguous feed
// shards of perShard requests; shard s owns [s*perShard, min(numJobs,
// (s+1)*perShard)).  A wavefront claims from its home shard and scans the
// others only once that one is exhausted.
constexpr int kFeedShards = 64;
static_assert(kFeedShards == kQueueShards, "a wave's home shard indexes both");
struct alignas(256) FeedShard {
  uint32_t next;         // next unclaimed request
  uint32_t pad0[63];
};

// Donation demand (queued items minus idle wavefronts), kept as the sum of
// kBalanceShards counters; wavefront w updates shard w % kBalanceShards.
constexpr int kBalanceShards = 8;
struct alignas(256) Balance {
  int64_t v;
  int64_t pad[31];
};

// Per-launch control block (reset by the host before each launch).  Only
// the feed shards that own requests are copied; the kernel never reads the
// others' counters.
struct alignas(256) Control {
  uint32_t numJobs;
  uint32_t perShard;
  uint32_t pad0[30];
  // per-request launch: its one request travels in the same copy as the
  // counters (the kernel then reads Params::jobs pointing here)
  Job single;
  uint32_t drained;      // batched: 1 once some wavefront saw every feed shard exhausted
  uint32_t pad1[63];
  Balance balance[kBalanceShards];
  // batched: requests published, summed over the shards; each wavefront
  // counts its own in a register and adds them when it runs out of work
  Balance published[kBalanceShards];
  uint32_t jobsDone;     // served: requests published
  uint32_t pad4[63];
  // served loop
  uint64_t consumed;     // submit-ring entries taken by the device
  uint64_t completions;  // completion-ring entries written
  uint64_t lastPoll;     // device clock of the last host-ring poll
  uint64_t submittedMirror;
  uint64_t stopMirror;   // device copy of ServeHost::submitted's stop bit
  uint64_t candidates;   // candidate triples evaluated, summed at each wavefront's exit
  uint64_t pad5[26];
  FeedShard feed[kFeedShards];
};
static_assert(offsetof(Control, single) == 128 && offsetof(Control, drained) == 256,
              "Control layout");

// Fills the batched feed fields for `numJobs` requests (host side) and
// returns how many leading bytes of the block a launch must copy.
size_t init_feed(Control *c, uint32_t numJobs);

// Host-visible rings of the served loop (fine-grained pinned host memory).
struct alignas(128) ServeHost {
  // entries published by the host; bit 63 (kServeStop), set only once every
  // submitted request has completed, asks the kernel to exit
  uint64_t submitted;
  uint64_t pad0[31];
};
constexpr uint64_t kServeStop = 1ULL << 63;

// Params::lookMargin of a batch with too few requests per wavefront for
// look-ahead claims (no request is ever within it of a shard's end).
constexpr int32_t kNoLookAhead = 0x3FFFFFFF;

struct Params {
  const Job *jobs;
  JobState *states;
  DevResult *results;
  Control *ctrl;
  QueueState *queue;     // kQueueShards shards
  Item *items;
  uint64_t *seq;
  uint64_t qcap;         // slots per shard
  const uint16_t *lex;   // lexicographic triple tables, kLexEntries entries
  int32_t splitSlack;    // donate only frames with depth <= limit - splitSlack
  int32_t donateMax;
  int32_t served;        // 0 batched, 1 served loop
  int32_t idleBackoff;   // longest idle sleep, in units of s_sleep(16)
  int32_t checkEvery;    // expansions between an item's sharing checks
  int32_t itemCap;       // served: most work items one request may have pending (0 = no cap)
  uint32_t ringCap;      // served: power of two
  ServeHost *host;       // served: device view of the host block
  const Job *hostJobs;   // served: device view of the host submit ring
  DevResult *hostResults;// served: device view of the host completion ring
  uint32_t *hostCompletionIdx; // served: ring index per completion sequence
  int32_t countCandidates; // 1: run the counted build (sums Control::candidates)
  // batched feed: requests and requests per feed shard, the values init_feed
  // writes into the control block, kept here too so that wavefronts compute
  // shard bounds from kernel arguments instead of loading them
  uint32_t numJobs;
  uint32_t perShard;
  // batched feed: a request within lookMargin of its feed shard's end is
  // not a look-ahead point (set by launch_search from the batch and grid;
  // kNoLookAhead for a batch of few requests per wavefront)
  int32_t lookMargin;
};

// Copies the feed sizes init_feed wrote into `c` into the kernel arguments.
inline void feed_params(Params &P, const Control &c) {
  P.numJobs = c.numJobs;
  P.perShard = c.perShard;
}

// Launch wrappers (td16_device.hip).
// wide: the six-waves-per-SIMD build of the kernel (above 20 waves per CU)
int launch_search(const Params &p, int blocks, bool wide, void *stream);
int launch_queue_init(uint64_t *seq, uint64_t capPerShard, void *stream);
// a batched launch's control block back to its starting state, on the device
int launch_control_reset(Control *c, uint32_t numJobs, uint32_t perShard, void *stream);
int search_kernel_occupancy(int *blocksPerCU, int *blocksPerCUWide);

// Host helpers (td16_host.cpp).
int set_msg(char *msg, int cap, int status, const char *fmt, ...);
void build_lex_table(uint16_t *out);   // kLexEntries entries
int rate_tables(int N, int initialEven, float minRate, int maxDepth, int fullSorting,
                uint32_t sat[10], uint32_t reach[10]);
void unpack_moves(const uint16_t *packed, int depth, int32_t (*moves)[3]);

// Launch-path helpers shared by the batched entry points (td16_runtime.cpp).
// The library's default sharing knobs, and a batch's overrides of them.
void default_knobs(Params &P);
void apply_batch_knobs(Params &P, const td16_batch_options &o);
// Persistent grid for `wavesPerCU` (0 = default for a batch of `numJobs`
// requests) on a device with `cus` CUs: blocks, and whether the
// six-waves-per-SIMD build runs.
void batch_grid(int wavesPerCU, uint32_t numJobs, int cus, int blocksPerCU, int blocksPerCUWide,
                int *blocks, bool *wide);
void decode_result(const DevResult &r, td16_result *out);
void trivial_result(int trivial, td16_result *out);

// The witness / prune tables and the no-slack flag depend only on (N,
// initial_even, min_rate, max_moves, full_sorting) -- and on whether the
// request is an iterative-deepening one -- and a batch has few distinct such
// keys (the batched mix: about 300 among 100,000 requests), so the tables
// last built are kept in a small direct-mapped memo, one per preparing
// thread.
struct RateMemo {
  static constexpr int kEntries = 512;
  struct Entry {
    uint64_t k1 = ~0ULL, k2 = ~0ULL;
    uint32_t sat[10], reach[10];
    bool tight;  // FIRST search without slack (kFlagTight)
  } e[kEntries];
};
// Validates one request and writes its label form (td16_host.cpp); *trivial
// is set when no launch is needed.  make_job_checked uses a per-thread memo.
int make_job(const td16_request *r, Job *job, int *trivial, RateMemo &memo, char *msg, int cap);
int make_job_checked(const td16_request *r, Job *job, int *trivial, char *msg, int cap);

// Host preparation and decoding of a batch (td16_prepare.cpp), on a pool of
// host threads.  host_threads(0) is the default thread count.
int host_threads(int want);
// Position of the r-th request of the feed order once the sorted list is
// dealt round-robin to the feed shards of `per` requests (init_feed).
uint32_t deal_position(uint32_t r, uint32_t nj, uint32_t per);
struct PrepareScratch {
  std::vector<Job> jobs;
  std::vector<uint16_t> keys;    // per request: its feed bucket
  std::vector<uint32_t> pos;     // per request: its feed position, or kNotLaunched
  std::vector<uint32_t> counts;
  std::vector<RateMemo> memos;  // one per part
};
// Validates and relabels every request, answers the trivial ones in
// `results` and writes the rest in feed order to jobsOut[0 .. *njOut)
// (`whereOut[q]` is the caller's index of jobsOut[q]); both arrays hold
// `count` entries.  An invalid request stops the batch: its result gets the
// status and `msg` the reason.
int prepare_batch_into(const td16_request *reqs, int count, td16_result *results, Job *jobsOut,
                       int *whereOut, uint32_t *njOut, PrepareScratch &ws, int threads,
                       char *msg, int cap);
// The same into vectors sized to the requests that need a launch.
int prepare_batch(const td16_request *reqs, int count, td16_result *results,
                  std::vector<Job> &jobs, std::vector<int> &where, char *msg, int cap);
// Batched results sit at their requests' feed positions; pos[q] is request
// q's position (kNotLaunched for a request answered on the host).
// decode_batch fills results[q] from dev[pos[q]] for every launched q <
// count, walking the caller's order, and sums the counters.
constexpr uint32_t kNotLaunched = 0xFFFFFFFFu;
void decode_batch(const DevResult *dev, const uint32_t *pos, uint32_t count,
                  td16_result *results, int threads, uint64_t *expanded, uint64_t *solutions,
                  uint64_t *donations);
// pos for decode_batch from the caller indices of the launched requests.
void positions_of(const int *where, uint32_t nj, uint32_t count, std::vector<uint32_t> &pos);

}  // namespace td16
