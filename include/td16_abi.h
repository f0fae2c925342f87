/*
 * td16 witness search -- C ABI of libtd16.so (MI355X / gfx950).
 *
 * Drop-in boundary for the reference's JNI search library `td16_jni`
 * (reference: emitter/src/native/jni/jni_sorter.cu).  Every entry point takes
 * plain pointers and sizes; no HIP or torch types cross this boundary.  A JNI
 * shim that exports the reference's exact `Java_br_unb_cic_tdp_base_
 * GPUSortingSearch_*` names on top of these functions is in
 * td16_amd/csrc/jni_sorter.cpp.
 *
 * Conventions
 *  - Symbols are 0..n-1 with n <= 17 (the reference's MAX_N,
 *    emitter/src/native/CMakeLists.txt:42); move bounds are 1..9
 *    (MAX_DEPTH, :43).
 *  - A move is a position triple (i, j, k), 0 <= i < j < k < n, in the beta
 *    sequence current before the move; it swaps the blocks [i,j) and [j,k)
 *    of beta and maps omega to omega o (beta[i] beta[j] beta[k])^-1
 *    (emitter/src/native/backtracking/sorting_search.cu:228-230, 443-445).
 *  - Every function returns a td16_status.  On failure a NUL-terminated
 *    message is written to `msg` (when msg != NULL and msg_cap > 0); the JNI
 *    shim turns every non-OK status into java.lang.RuntimeException(msg), as
 *    jni_sorter.cu:15-22,111-116 does for every std::exception.
 *  - Functions are thread-safe.  Calls on distinct slots run concurrently;
 *    the reference also drives each slot from a single thread
 *    (emitter/src/main/java/br/unb/cic/tdp/base/GPUSortingSearch.java:123-145).
 */
#ifndef TD16_ABI_H
#define TD16_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TD16_MAX_N 17
#define TD16_MAX_DEPTH 9

typedef enum {
    TD16_OK = 0,
    TD16_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    TD16_OUT_OF_MEMORY = 2,    /* GpuMemoryExhaustedException */
    TD16_DEVICE_ERROR = 3,     /* CudaException */
    TD16_NOT_INITIALIZED = 4,  /* search before initSlots */
    TD16_QUEUE_FULL = 5        /* served loop: pending queue at capacity (try later) */
} td16_status;

/* Move sets.  REF is the reference filter (sorting_search.cu:375-380,
 * 439-441, 607-610): no cut at an omega fixed point, not three distinct
 * cycles, odd-cycle delta 0 or +2.  ALL takes every position triple. */
enum { TD16_MOVES_REF = 0, TD16_MOVES_ALL = 1 };
/* FIRST stops at the first witness (the reference's behaviour);
 * COUNT_ALL never stops early and counts every node and witness. */
enum { TD16_FIRST = 0, TD16_COUNT_ALL = 1 };

/* One search request in the reference's form.
 *   beta:  the cyclic sequence, beta[0..n-1] (a permutation of 0..n-1)
 *   omega: one-line permutation, omega[x] = image of x
 *   initial_even, min_rate, max_moves, full_sorting: as passed to
 *     GPUSortingSearch.searchForSortingSeq (GPUSortingSearch.java:117)
 *   moveset, mode: TD16_MOVES_*, TD16_*; the reference is REF + FIRST
 *   iddfs: 1 = exact distance by iterative deepening on limits
 *     max(ceil(b/3), odd-cycle bound), +1, ... up to max_moves (min_rate,
 *     initial_even and full_sorting are then ignored: the witness condition
 *     is "sorted").
 *   lex_order: visiting order of a FIRST search's expandable children.
 *     0 (default) = greedy: the children that gain two odd cycles first, then
 *     the others, each class in the reference's lexicographic triple order;
 *     1 = lexicographic, the reference's order.  A node's children are always
 *     evaluated in lexicographic order, so the witness move found at a node
 *     is the lexicographically first one in both.  COUNT_ALL ignores it. */
typedef struct {
    uint8_t beta[TD16_MAX_N];
    uint8_t omega[TD16_MAX_N];
    uint8_t lex_order;
    uint8_t pad0;
    int32_t n;
    int32_t initial_even;
    float min_rate;
    int32_t max_moves;
    int32_t full_sorting;
    int32_t moveset;
    int32_t mode;
    int32_t iddfs;
} td16_request;

/* Result of one request.
 *   status: td16_status of this request
 *   found: 1 when a witness was found
 *   depth: witness length (0 when none); in IDDFS mode the exact distance
 *   moves: positions (i, j, k) of the witness, moves[0..depth-1]
 *   expanded: nodes whose C(n,3) children were evaluated (root included)
 *   solutions: witnesses met (COUNT_ALL: all of them)
 *   split: 1 when the search was shared by several wavefronts.  Its FIRST
 *     witness and node count may then differ from a sequential DFS's in the
 *     same order; the witness is still valid and found / not-found never
 *     differs (a complete search expands the same node set in any order)
 *   device_us: device time from the request's start on a wavefront to its
 *     result's publication (the chip's 100 MHz real-time counter)
 *   latency_us: served loop only, submit (or scheduled arrival) to the
 *     result's arrival on the host
 *   queue_us: served loop only, submit (or scheduled arrival) to the
 *     request's publication to the device (time spent waiting for a slot) */
typedef struct {
    int32_t status;
    int32_t found;
    int32_t depth;
    int32_t split;
    int32_t moves[TD16_MAX_DEPTH][3];
    uint64_t expanded;
    uint64_t solutions;
    double device_us;
    double latency_us;
    double queue_us;
} td16_result;

/* ---------------------------------------------------------------------- */
/* Slots: the reference's per-slot lifecycle                              */
/* ---------------------------------------------------------------------- */

/* Replaces Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots
 * (jni_sorter.cu:24-47) -> initGpuSortSlots (backtracking/slot_manager.cu:
 * 44-128) -> allocateGpuSortingSearchSlot (sorting_search.cu:755-793).
 * Slot s runs on devices[s % ndevices].  Each slot gets its own stream and
 * `queue_bytes_budget` bytes of HBM for its work queue.  Fails with
 * TD16_INVALID_ARGUMENT on an empty, duplicate or out-of-range device list,
 * total_slots <= 0 or budget <= 0, and with TD16_OUT_OF_MEMORY when
 * budget x slots exceeds a device's free memory.  Re-initialising replaces
 * the previous slots. */
int td16_init_slots(int total_slots, const int32_t *devices, int ndevices,
                    int64_t queue_bytes_budget, char *msg, int msg_cap);

/* Releases every slot, stream and buffer (the reference never frees its
 * slots; this exists for tests and orderly shutdown). */
int td16_shutdown(char *msg, int msg_cap);

/* Number of slots and the device of one slot (-1 when invalid).
 * Replaces getGpuSortSlotDevice (slot_manager.cu:145-148). */
int td16_slot_count(void);
int td16_slot_device(int slot);

/* ---------------------------------------------------------------------- */
/* The drop-in search                                                      */
/* ---------------------------------------------------------------------- */

/* Replaces Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq
 * (jni_sorter.cu:49-117) -> performGpuSortingSearch (sorting_search.cu:
 * 811-972): one blocking search on `slot`, REF moves, FIRST mode.
 * out_positions receives up to 9 position triples (row-major [9][3]);
 * *out_len the witness length, 0 when there is none (including n < 3).
 * Errors: TD16_INVALID_ARGUMENT when beta/omega are not permutations of
 * 0..n-1, n > 17, max_moves outside 1..9 or the slot is invalid
 * (sorting_search.cu:820-825, 140-145). */
int td16_search(int slot, const uint8_t *beta, const uint8_t *omega, int n,
                int initial_even, float min_rate, int max_moves,
                int full_sorting, int32_t *out_positions, int32_t *out_len,
                char *msg, int msg_cap);

/* The same blocking one-instance-per-launch search with every option of
 * td16_request and the counters of td16_result. */
int td16_search_ex(int slot, const td16_request *req, td16_result *res,
                   char *msg, int msg_cap);

/* ---------------------------------------------------------------------- */
/* Batched search: one persistent launch over many requests                */
/* ---------------------------------------------------------------------- */

typedef struct {
    int32_t device;          /* HIP device ordinal */
    int32_t waves_per_cu;    /* resident wavefronts per CU, in steps of 4 up to 24;
                                0 = 20 for a batch of at least 128 requests per CU,
                                else 16; 24 runs an 80-VGPR build of the kernel */
    int32_t split_depth;     /* frames deeper than max_moves - split_depth are
                                never donated; 0 = default (5) */
    int32_t donate_max;      /* children donated per donation (at most 64); 0 = default (16) */
    int32_t use_graph;       /* 1 = launch through a captured hipGraph */
    int32_t repeats;         /* run the same batch this many times (>= 1) */
    int64_t queue_bytes;     /* HBM for the shared work queue; 0 = 256 MiB */
    int32_t check_every;     /* node expansions between an item's sharing
                                checks (at most 1024); 0 = default (16) */
    int32_t host_threads;    /* host threads that prepare the requests and decode
                                the results; 0 = default (all hardware threads,
                                at most 16) */
    int32_t count_candidates;/* 1 = run the counted build of the kernel, which
                                also sums the candidate triples it evaluates
                                (td16_batch_stats.candidates); for measurement:
                                the default build does not count */
    int32_t reserved;        /* 0 */
} td16_batch_options;

typedef struct {
    double kernel_ms;        /* mean device time of one persistent launch */
    double wall_ms;          /* mean host wall time of one repeat: the control-block
                                reset (a host copy, or inside the graph), the launch
                                and the wait for its end (the requests' preparation
                                and upload are not in it) */
    uint64_t expanded;       /* sum over requests (last repeat) */
    uint64_t solutions;
    uint64_t donations;      /* work items pushed to the shared queue */
    int32_t waves;           /* wavefronts in the persistent grid */
    int32_t blocks;
    double prepare_ms;       /* host: validation, relabelling and feed order of
                                the requests (td16_batch; 0 for re-runs of
                                requests prepared once) */
    double upload_ms;        /* device: the request table's copy to HBM */
    double decode_ms;        /* host: results of the last repeat to td16_result */
    double total_ms;         /* host wall time of the whole call */
    uint64_t candidates;     /* count_candidates launches: candidate triples the last
                                launch evaluated, C(m,3) per expanded node, m the
                                positions it enumerates (breakpoints under the
                                reference filter); 0 otherwise */
} td16_batch_stats;

/* Runs `count` requests in one persistent, work-sharing launch per repeat.
 * results[q] is request q's result (from the last repeat). */
int td16_batch(const td16_request *reqs, int count, td16_result *results,
               const td16_batch_options *opt, td16_batch_stats *stats,
               char *msg, int msg_cap);

/* A stream of requests as consecutive persistent launches of at most
 * `chunk` requests each, double-buffered: while chunk c is searched, the
 * host prepares chunk c + 1 (validation, relabelling, feed order) and
 * decodes chunk c - 1's results, and chunk c + 1's upload runs beside the
 * search on a second stream.  This is how a producer that hands over fresh
 * batches keeps the device busy; td16_batch pays the same host work before
 * and after its one launch.  Results and options as td16_batch (use_graph
 * and repeats do not apply); in stats the times are sums over the chunks
 * (wall_ms: the host's waits for searches to end) and total_ms the wall
 * time of the whole call.  The reference searches one instance per call
 * (sorting_search.cu:811-972); the Java producers stay the source of the
 * stream. */
int td16_batch_stream(const td16_request *reqs, int count, int chunk, td16_result *results,
                      const td16_batch_options *opt, td16_batch_stats *stats,
                      char *msg, int msg_cap);

/* Allocation-free variant of td16_batch.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * running while a HIP graph is captured, tight serving loops, harnesses that
 * forbid hipMalloc.  It takes the same requests, results, options and
 * statistics as td16_batch and runs the same kernels, but calls no
 * hipMalloc, hipFree, hipHostMalloc or other allocating HIP routine: every
 * device buffer is a region of one static device array per GPU of at most
 * TD16_NOALLOC_SCRATCH_BYTES, which that GPU holds until the process ends.
 * Host staging uses ordinary (pageable) process memory.  Differences:
 *  - count must be at most TD16_NOALLOC_MAX_REQUESTS, else
 *    TD16_INVALID_ARGUMENT;
 *  - the shared work queue holds TD16_NOALLOC_QUEUE_ITEMS donated subtrees
 *    (queue_bytes is ignored; a full queue only stops donation, so results
 *    do not depend on it);
 *  - use_graph must be 0; launches go to the null stream;
 *  - kernel_ms is host-timed (launch to completion, no events).
 * Every call on one GPU uses that GPU's array, which keeps the last call's
 * contents (each call resets what a launch reads before writing).  Calls
 * that overlapped on one GPU would race on it, so the library serializes
 * them per process with a lock; each process has its own arrays. */
#define TD16_NOALLOC_MAX_REQUESTS 32768
#define TD16_NOALLOC_QUEUE_ITEMS 32768
#define TD16_NOALLOC_SCRATCH_BYTES (13 << 20)
int td16_batch_noalloc(const td16_request *reqs, int count, td16_result *results,
                       const td16_batch_options *opt, td16_batch_stats *stats,
                       char *msg, int msg_cap);

/* One configuration the autotuner timed. */
typedef struct {
    int32_t waves_per_cu;
    int32_t split_depth;
    int32_t donate_max;
    int32_t check_every;
    int32_t waves;           /* wavefronts in the grid it gave */
    int32_t reserved;
    double kernel_ms;        /* median device time of one launch */
} td16_tune_point;

/* Autotuner for the batched launch: picks waves_per_cu (wavefronts per CU),
 * split_depth (how far above the move bound a DFS frame may still be
 * donated: the stack depth at which work is shared), donate_max (the steal
 * granularity) and check_every (how often a search looks for idle
 * wavefronts to share with) for this batch by timing it.  Coordinate
 * descent from `base` over waves_per_cu {4, 8, 12, 16, 20, 24}, split_depth
 * {2..7}, donate_max {4, 8, 16, 32, 64} and check_every {8, 16, 32}; each
 * point is one untimed launch and
 * `repeats` timed ones (median), and a point must beat the current choice by
 * more than 1 % to replace it.  The knobs only change how work is shared:
 * every point's answers (found; the distance of iterative-deepening
 * requests; the counts of COUNT_ALL requests) must equal the first point's,
 * else TD16_DEVICE_ERROR.  On success `best` is `base` with the four knobs
 * set, results[q] holds request q's result from the first point and
 * points[0 .. min(*npoints, points_cap)) the timed configurations in the
 * order tried.  The reference picks none of these (its queue count comes
 * from the occupancy API, sorting_search.cu:714-752). */
int td16_autotune(const td16_request *reqs, int count, td16_result *results,
                  const td16_batch_options *base, int repeats,
                  td16_batch_options *best, td16_tune_point *points,
                  int points_cap, int *npoints, char *msg, int msg_cap);

/* ---------------------------------------------------------------------- */
/* Served loop: continuous batching into a running persistent kernel       */
/* ---------------------------------------------------------------------- */

typedef struct {
    int32_t device;
    int32_t slots;           /* requests admitted on the device at once (1..65536) */
    int32_t queue_capacity;  /* requests waiting for a slot (the reference's
                                tdp.gpuQueueCapacity, GPUSortingSearch.java:82-89) */
    int32_t waves_per_cu;    /* resident wavefronts per CU; 0 = 4 (tuned for latency) */
    int64_t slot_bytes;      /* HBM budget per slot (tdp.gpuBudgetMB): the loop
                                allocates slots x slot_bytes for queued subtrees
                                (TD16_OUT_OF_MEMORY beyond the device's free
                                memory) and an admitted request keeps at most
                                slot_bytes of them queued (48-byte subtree plus
                                8-byte ring word each; a full budget only stops
                                its donations, the work stays with the
                                wavefronts that hold it) */
    int32_t deadline_ms;     /* 0 = none; else td16_serve_wait and td16_serve_stop
                                give up after this long with TD16_DEVICE_ERROR */
    int32_t reserved;        /* 0 */
} td16_serve_options;

typedef struct {
    uint64_t submitted, completed, rejected;
    double p50_us, p99_us, mean_us; /* submit -> result visible on the host; p50 and
                                       p99 to 2 % (log-spaced buckets), mean exact */
    int32_t in_flight, waiting;
} td16_serve_stats;

/* Starts the served loop on one device: a persistent kernel that admits
 * requests as slots free up, shares their subtrees across wavefronts and
 * reports each completion to the host. */
int td16_serve_start(const td16_serve_options *opt, char *msg, int msg_cap);
/* Queues a request; blocks while `queue_capacity` requests are waiting
 * (GPUSortingSearch.java:216 blocks the same way) unless nonblocking, in
 * which case TD16_QUEUE_FULL is returned.  *ticket identifies the request.
 * Once td16_serve_stop has begun, new submits fail with
 * TD16_NOT_INITIALIZED (a submit blocked on a full queue too). */
int td16_serve_submit(const td16_request *req, int nonblocking, uint64_t *ticket,
                      char *msg, int msg_cap);
/* Blocks until the ticket's result is ready and returns it.  Returns
 * TD16_DEVICE_ERROR when the served kernel has failed or ended, or when the
 * deadline passes (the ticket then stays valid for a later wait). */
int td16_serve_wait(uint64_t ticket, td16_result *res, char *msg, int msg_cap);
/* Submit + wait. */
int td16_serve_search(const td16_request *req, td16_result *res, char *msg, int msg_cap);
int td16_serve_stats_get(td16_serve_stats *st, int reset_latency, char *msg, int msg_cap);
/* Open-loop load: submits reqs[q] at arrival_us[q] microseconds after the
 * call starts (spinning until then), waits for every result, and resets the
 * latency window first so that td16_serve_stats_get reports this replay's
 * p50/p99, measured from each scheduled arrival to the result's arrival on
 * the host.  *elapsed_us receives the replay's wall time. */
int td16_serve_replay(const td16_request *reqs, const double *arrival_us, int count,
                      td16_result *results, double *elapsed_us, char *msg, int msg_cap);
/* Closed-loop load: `producers` host threads (1..1024), each keeping up to
 * `window` (1..4096) of its requests in flight -- it submits while it has
 * fewer, otherwise waits for its oldest -- with no pacing (producer p takes
 * reqs[p], reqs[p + producers], ...).  window = 1 is a thread that calls
 * search and blocks, as GPUSortingSearch's slot threads do
 * (GPUSortingSearch.java:134-162, 203-240).  Measures the loop's capacity;
 * resets the latency window first. */
int td16_serve_closed_loop(const td16_request *reqs, int count, int producers, int window,
                           td16_result *results, double *elapsed_us, char *msg, int msg_cap);
/* Stops the kernel after the admitted requests finish.  With a deadline, a
 * loop that does not drain and stop in time (or whose kernel failed) returns
 * TD16_DEVICE_ERROR and leaves its device memory allocated. */
int td16_serve_stop(char *msg, int msg_cap);

/* ---------------------------------------------------------------------- */
/* Caller-side replay and the certificate format                           */
/* ---------------------------------------------------------------------- */

/* Positions -> symbols by replaying the block swaps on a copy of beta
 * (GPUSortingSearch.java:229-239,258-264). */
int td16_positions_to_symbols(const uint8_t *beta, int n, const int32_t *positions,
                              int len, int32_t *symbols, char *msg, int msg_cap);

/* Replays symbol triples (CertificateValidator.java:14-50): length in
 * 1..max_moves, each move an ordered triple of distinct symbols in the
 * current beta, omega o beta preserved, omega sorted at the end.  Returns
 * TD16_OK or TD16_INVALID_ARGUMENT with the reason. */
int td16_validate(const uint8_t *beta, const uint8_t *omega, int n,
                  const int32_t *symbols, int len, int max_moves,
                  char *msg, int msg_cap);

/* Certificate codec (CertificateCodec.java:22-118): the key is omega's
 * one-line notation at 5 bits per image, MSB first; the value is 15 bits per
 * move.  *out_len receives the byte count.  Decoders reject a wrong length
 * or nonzero padding. */
int td16_codec_key(const uint8_t *omega, int n, uint8_t *out, int cap, int32_t *out_len,
                   char *msg, int msg_cap);
int td16_codec_value(const int32_t *symbols, int len, uint8_t *out, int cap,
                     int32_t *out_len, char *msg, int msg_cap);
int td16_codec_decode_key(const uint8_t *in, int nbytes, int n, uint8_t *omega,
                          char *msg, int msg_cap);
int td16_codec_decode_value(const uint8_t *in, int nbytes, int32_t *symbols, int cap,
                            int32_t *out_len, char *msg, int msg_cap);

/* Batched certificate work on host threads (threads 0 = every hardware
 * thread).  Fixed-stride records: key q at keys + q*TD16_CERT_KEY_BYTES,
 * value q at values + q*TD16_CERT_VALUE_BYTES, with value_len[q] bytes used.
 *
 * td16_certify_batch: the emitter's step after a batch of searches
 * (GPUSortingSearch.java:229-239 positions -> symbols,
 * CertificateValidator.java:14-50 replay, CertificateCodec.java:22-70 key and
 * value).  status[q]: 0 certified (key, value written); 1 nothing to certify
 * (result not OK or no witness; value_len 0); -1..-6 the replay rejected the
 * witness (codes below).
 *
 * td16_check_batch: the checker's lookup of a batch of proof obligations
 * after the multi-get (checker/.../CertificateStore.java:97-122): each stored
 * value is decoded (CertificateCodec.java:75-97) and replayed from the
 * obligation's pair (beta q, omega q at stride TD16_MAX_N, ns[q] symbols)
 * under max_moves[q].  status[q]: 0 valid; 1 no stored value (value_len 0,
 * the reference's empty list); -1 length outside 1..max_moves, -2 beta or
 * omega not a permutation, -3 a move without three distinct symbols, -4 not
 * an ordered triple, -5 cyclic target not preserved, -6 not sorted, -7
 * malformed value (length or padding).  The reference throws on the first
 * invalid certificate; here every obligation gets its own status. */
#define TD16_CERT_KEY_BYTES 11   /* (5 * TD16_MAX_N + 7) / 8 */
#define TD16_CERT_VALUE_BYTES 17 /* (15 * TD16_MAX_DEPTH + 7) / 8 */
int td16_certify_batch(const td16_request *reqs, const td16_result *results, int count,
                       int threads, uint8_t *keys, uint8_t *values, int32_t *value_len,
                       int32_t *status, char *msg, int msg_cap);
int td16_check_batch(const uint8_t *betas, const uint8_t *omegas, const int32_t *ns,
                     const int32_t *max_moves, int count, const uint8_t *values,
                     const int32_t *value_len, int threads, int32_t *status,
                     char *msg, int msg_cap);

/* ---------------------------------------------------------------------- */
/* Synthetic instances                                                     */
/* ---------------------------------------------------------------------- */

/* Linear permutations, nibble-packed (position p in bits 4p..4p+3, unused
 * nibbles 0xF), from SplitMix64 with uniform draws by rejection.
 *   uniform: Durstenfeld Fisher-Yates on 0..n-1, redrawn while identity
 *   products: n ~ U{nmin..nmax}, k ~ U{kmin..kmax}, k uniform block
 *     transpositions t(i,j,k), 0 <= i < j < k <= n, applied to the
 *     identity; redrawn while identity.  ks receives k.
 *   arrivals: unit-rate exponential gaps, -ln(1 - u), u = top 53 bits. */
int td16_gen_uniform(uint64_t seed, int n, int count, uint64_t *perms);
int td16_gen_products(uint64_t seed, int nmin, int nmax, int kmin, int kmax, int count,
                      uint8_t *ns, uint8_t *ks, uint64_t *perms);
int td16_gen_exponential(uint64_t seed, int count, double *gaps);

/* A linear permutation of 0..n-1 as a request in the reference's form:
 * beta = [pi_0 .. pi_{n-1}, n], omega(y) = pred(y) + 1 mod (n+1),
 * initial_even = odd cycles of omega, min_rate = 2 / (k / (float)norm)
 * (the emitter's convention, UnorientedCycles.java:34), max_moves = k. */
int td16_linear_request(uint64_t perm, int n, int k, int moveset, int mode, int iddfs,
                        td16_request *req);

/* ---------------------------------------------------------------------- */
/* Producer side: cycle-configuration representatives                      */
/* ---------------------------------------------------------------------- */

#define TD16_BRACELET_PAIRS 0
#define TD16_BRACELET_WORDS 1

/* One representative per realizable extended-toric class of an annotated
 * partition -- omega's cycle lengths parts[0..m-1], oriented[q] != 0 for an
 * oriented cycle -- the set DirectTwistedBraceletGenerator emits
 * (DirectTwistedBraceletGenerator.java:78-104 generateRepresentatives and
 * countRepresentatives; validation :234-255).
 *   mode TD16_BRACELET_PAIRS: a row is omega's one-line notation (n bytes,
 *     n = sum of parts) of the pair (omega, beta = (0 1 .. n-1)), whose
 *     target omega o beta is one n-cycle (the reference's CyclicTargetPair
 *     from decodeWord, :1475-1505);
 *   mode TD16_BRACELET_WORDS: a row is a raw orbit-minimal word before the
 *     target filter (the reference's generateWords test hook, :225-232):
 *     symbols number the cycles longest first (unoriented before oriented of
 *     one length), one symbol per unoriented cycle and one per rank of an
 *     oriented cycle.
 * out: cap_rows rows of n bytes, or NULL to count only; *total receives the
 * number of rows.  When it exceeds cap_rows, TD16_INVALID_ARGUMENT is
 * returned with the count in the message and out holds cap_rows of the rows
 * (which ones is unspecified); at most cap_rows rows are ever held in host
 * memory.  Otherwise rows come in increasing word order for any `threads`
 * (0 = all hardware threads).  n <= 32; a searchable pair needs
 * n <= TD16_MAX_N.  The time grows with the number of classes (the whole
 * unoriented stage of 17 symbols, 3.86 M classes, takes about 0.5 s on 16
 * threads); families of many short cycles on 30 or more symbols have far
 * more. */
int td16_bracelets(const int32_t *parts, const uint8_t *oriented, int m, int mode,
                   int threads, uint8_t *out, int64_t cap_rows, int64_t *total,
                   char *msg, int msg_cap);

/* The class of a fixed-point-free pair (TwistedBraceletKey.of,
 * TwistedBraceletKey.java:22-29; canonicalRepresentative :96-98): beta (the
 * cyclic sequence) and omega, permutations of 0..n-1, 2 <= n <= 32.  Writes
 * the annotated type (*m cycles; parts[c], oriented[c] in the generator's
 * order, longest first, unoriented before oriented; room for 16 entries),
 * the class word (n symbols of the generator's alphabet; NULL to skip) and
 * the class representative (omega's one-line notation for beta =
 * (0 1 .. n-1), the row td16_bracelets emits for the class when its target
 * is one n-cycle; NULL to skip).  A cycle is unoriented when omega maps each
 * of its symbols to the cycle's previous symbol along beta.
 * TD16_INVALID_ARGUMENT for non-permutations and for omega with a fixed
 * point. */
int td16_bracelet_canonical(const uint8_t *beta, const uint8_t *omega, int n,
                            int32_t *parts, uint8_t *oriented, int32_t *m,
                            uint8_t *word, uint8_t *rep_omega, char *msg, int msg_cap);

/* Sustained int32 VALU rate of the device, in 10^12 lane-ops per second:
 * a grid of resident wavefronts runs independent dependent chains of
 * v_add_u32 / v_xor_b32 for `iterations` rounds.  Measurement aid for the
 * search kernel's integer-op roofline (the search issues no MFMA). */
int td16_probe_valu_peak(int device, int iterations, double *tops, char *msg, int msg_cap);

/* Copies `bytes` bytes between two device buffers with a plain grid-stride
 * kernel (launched twice; the second launch is timed) and reports read+write
 * GB/s: a kernel of known HBM traffic for calibrating byte counters. */
int td16_probe_copy(int device, int64_t bytes, double *gbps, char *msg, int msg_cap);

/* Library version string. */
const char *td16_version(void);

#ifdef __cplusplus
}
#endif
#endif
