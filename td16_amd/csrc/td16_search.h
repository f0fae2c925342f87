// Persistent work-sharing DFS for block-transposition witnesses on CDNA4
// (gfx950, wave64).
//
// One wavefront is one DFS worker.  A node is expanded by the whole wave:
//   1. set-up: lanes 0..N-1 own one sequence position each; the cycle
//      structure of omega (cycle, length, distance to the cycle's smallest
//      position) is found by 4 rounds of pointer doubling over ds_bpermute,
//      no LDS tables and no serial walks;
//   2. candidates: the C(m,3) position triples (m = breakpoints under the
//      reference filter, all N positions otherwise) are evaluated 64 lanes at
//      a time in lexicographic order; each lane gets the odd-cycle delta of
//      its triple in O(1) from three bpermuted cycle records, then a ballot
//      splits witnesses from expandable children.
// The wave keeps one LDS frame per depth (sequence + expandable-children
// bitmasks), so the DFS stack never spills.  Idle wavefronts take donated
// subtrees from a lock-free ring in HBM, or the next request.
//
// Reference semantics followed (emitter/src/native/backtracking/
// sorting_search.cu): move filter :375-380,439-441,607-610; delta cases
// :240-332; witness and prune predicates :336-373 (host-tabulated); positions
// returned exactly as :956-971 decodes them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "td16_internal.h"

namespace td16 {

namespace {

// --------------------------------------------------------------------------
// Wave primitives
// --------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// (the builtin takes the predicate as a lane mask; HIP's __ballot widens it
// to an int first, which costs a select and a compare per call)
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ int bperm(int srcLane, int v) {
  return __builtin_amdgcn_ds_bpermute(srcLane << 2, v);
}

__device__ __forceinline__ int fperm(int dstLane, int v) {
  return __builtin_amdgcn_ds_permute(dstLane << 2, v);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Orders this wave's LDS writes before later reads by other lanes of the
// same wave (LDS requests of one wave complete in order; this stops the
// compiler from moving accesses across the point).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// v = value in lane `lane` of v, the other lanes unchanged (v_writelane_b32,
// lane select in M0; this compiler has no builtin for it).  `value` and
// `lane` are wave-uniform.
__device__ __forceinline__ void write_lane(int &v, uint32_t value, int lane) {
  asm volatile("v_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(value), "{m0}"(lane));
}

__device__ __forceinline__ int rank_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// 64-bit value of lane q (wave-uniform).
__device__ __forceinline__ int64_t lane_value64(int64_t v, int q) {
  return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), q) << 32) |
                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, q));
}

// Sum of the values lanes 0..kBalanceShards-1 hold (wave-uniform).
__device__ __forceinline__ int64_t shard_lane_sum(int64_t v) {
  int64_t s = 0;
#pragma unroll
  for (int q = 0; q < kBalanceShards; ++q) s += lane_value64(v, q);
  return s;
}

template <class T>
__device__ __forceinline__ T ld_rlx(T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T ld_acq(T *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_rel(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_rlx(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T add_rlx(T *p, T v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T add_ar(T *p, T v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ bool cas(T *p, T expected, T desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ bool cas_rlx(T *p, T expected, T desired) {
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Agent-scope fences.  On gfx950 each XCD has its own L2, so these cost an
// L2 write-back (release) or invalidate (acquire): they are placed only where
// one wavefront hands data to another (a donated item, a finished request),
// never in polling loops.
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
// Waits until this wave's earlier memory operations have completed and keeps
// the compiler from moving memory operations across the wait.  Between
// agent-scope atomics (performed at the memory side) and agent-scope atomic
// stores (written through to memory) this orders them for every other wave
// without the L2 write-back of fence_rel; plain stores still need fence_rel.
// (A workgroup-scope release fence emits no wait at all outside
// thread-group-split mode, so the wait is spelled out.)
__device__ __forceinline__ void complete_mem() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}


// Lane 0 runs f, every lane gets its int result.
template <class F>
__device__ __forceinline__ int lane0(F f) {
  int r = 0;
  if (lane_id() == 0) r = f();
  return uni(r);
}

// Block transposition t(i,j,k) on a nibble-packed sequence: blocks [i,j) and
// [j,k) swap places (reference: emitter/src/native/backtracking/
// transposition.cuh:5-11, E/base/GPUSortingSearch.java:258-264).
__device__ __forceinline__ uint64_t apply_move(uint64_t pi, int i, int j, int k) {
  const uint32_t lo = 4u * i, mid = 4u * j, hi = 4u * k;
  const uint64_t mLo = (1ULL << lo) - 1ULL;
  const uint64_t mMid = (1ULL << mid) - 1ULL;
  const uint64_t mHi = (hi >= 64u) ? ~0ULL : ((1ULL << hi) - 1ULL);
  const uint64_t A = pi & (mMid ^ mLo);
  const uint64_t B = pi & (mHi ^ mMid);
  const uint64_t rest = pi & ~(mHi ^ mLo);
  return rest | (B >> (mid - lo)) | (A << (hi - mid));
}

// --------------------------------------------------------------------------
// Per-wave LDS state
// --------------------------------------------------------------------------

// A frame lists its expandable children as bitmasks over the lexicographic
// candidate index t (bit t%64 of chunk t/64).  In a reordered frame the
// chunks come twice: first the children that gain two odd cycles, then the
// others, so a visit in chunk order tries the greedy moves first.
struct Frame {
  uint64_t pi;
  uint64_t mask[2 * kMaxChunks];  // expand() builds them in lanes ch and 32 + ch
  uint32_t hdr;               // frame header, see hdr_* below
  uint8_t pos[20];            // element -> position
};

// Frame header word: lexicographic table offset of the enumerated set
// (C(m,4) <= 2380), candidate chunk count, "element q is position q" flag
// and "reordered" flag; written once when the node is expanded.  A taken
// child is cleared from its mask, so the frame's next child is always the
// lowest bit of its first nonzero mask chunk (no cursor to keep: storing one
// cost an LDS write per descent step).
__device__ __forceinline__ uint32_t hdr_make(int off, int nchunks, bool full, bool reordered) {
  return (uint32_t)off | ((uint32_t)nchunks << 12) | (full ? 1u << 16 : 0u) |
         (reordered ? 1u << 17 : 0u);
}
__device__ __forceinline__ int hdr_off(uint32_t h) { return (int)(h & 0xFFFu); }
__device__ __forceinline__ int hdr_nchunks(uint32_t h) { return (int)((h >> 12) & 15u); }
__device__ __forceinline__ bool hdr_full(uint32_t h) { return (h >> 16) & 1u; }
__device__ __forceinline__ bool hdr_reordered(uint32_t h) { return (h >> 17) & 1u; }
// mask chunks to scan, and the candidate chunk a mask chunk stands for
__device__ __forceinline__ int hdr_mask_chunks(uint32_t h) {
  return hdr_reordered(h) ? 2 * hdr_nchunks(h) : hdr_nchunks(h);
}
__device__ __forceinline__ int hdr_cand_chunk(uint32_t h, int ch) {
  const int n = hdr_nchunks(h);
  return ch >= n ? ch - n : ch;
}

struct WaveLds {
  Frame frames[kMaxDepth];
  uint32_t satTab[kMaxDepth + 1];
  uint32_t expTab[kMaxDepth + 1];
  uint16_t path[kMaxDepth + 1];  // path[d]: move applied at depth d
  uint16_t wit[kMaxDepth + 1];   // witness of a request no other wavefront touched
  uint16_t pad[4];
};

// The odd-cycle delta of a block transposition on positions a < b < c from
// 12 bits of their cycle records (D: the position's distance from its
// cycle's leader along the cycle, L: the cycle's length):
//   bits 0-5   L_a&1, D_a&1, L_b&1, D_b&1, L_c&1, D_c&1
//   bits 6-8   a,b on one cycle; b,c on one cycle; a,c on one cycle
//   bits 9-11  D_a < D_b, D_b < D_c, D_c < D_a (the arc between two
//              positions of one cycle wraps past the leader)
// Three positions on one cycle in cyclic order (one wrap) split it into
// three arcs; in the other order the cycle stays one.  Two on one cycle: the
// arc between them closes on itself and the rest joins the third position's
// cycle.  Three cycles merge into one.  Parities of the arcs: the distance
// difference's, flipped when the arc wraps a cycle of odd length.  Code:
// delta + 2 (0, 2 or 4) in the low three bits for the reference filter, 6
// when it rejects the move (a three-cycle merge or a negative delta), and
// delta + 2 in the next three for the full move set (never rejected); the
// code is a bit position in the per-node class masks of expand().
constexpr int kDeltaCodes = 4096;
__device__ __forceinline__ uint32_t delta_code(int idx) {
  const int la = idx & 1, qa = (idx >> 1) & 1, lb = (idx >> 2) & 1, qb = (idx >> 3) & 1;
  const int lc = (idx >> 4) & 1, qc = (idx >> 5) & 1;
  const int sab = (idx >> 6) & 1, sbc = (idx >> 7) & 1, sac = (idx >> 8) & 1;
  const int wab = (idx >> 9) & 1, wbc = (idx >> 10) & 1, wca = (idx >> 11) & 1;
  const int pab = qa ^ qb ^ (wab & la), pbc = qb ^ qc ^ (wbc & lb), pca = qc ^ qa ^ (wca & lc);
  int delta;
  if (sab && sbc) {
    delta = (wab ^ wbc ^ wca) ? pab + pbc + pca - la : 0;
  } else if (sab || sbc || sac) {
    const int pd = sab ? pab : (sbc ? pbc : pca);
    const int l1 = sbc ? lb : la;
    const int l2 = la ^ lb ^ lc;
    delta = pd + (l1 ^ l2 ^ pd) - l1 - l2;
  } else {
    delta = (la ^ lb ^ lc) - la - lb - lc;
  }
  const uint32_t cls = (uint32_t)(delta + 2);
  const bool refOk = (sab | sbc | sac) && delta >= 0;
  return (refOk ? cls : 6u) | (cls << 3);
}

struct BlockLds {
  // odd-cycle delta codes of a triple, by its 12 record bits (delta_code);
  // entry kDeltaCodes (and the padding) is code 7 (rejected) in both move
  // sets, read by lanes past a node's candidate count
  uint8_t dtab[kDeltaCodes + 8];
  uint16_t lex[kLexEntries];
  // the candidate loop reads one chunk ahead, up to entry C(17,4) +
  // 12 * 64 - 1 = 3147 for m = 17 (lanes past the candidate count discard
  // what they read)
  uint16_t lexPad[96];
  WaveLds wave[kWavesPerBlock];
  // counted builds: candidate triples each wavefront evaluated (after the
  // waves' state, so the layout the timed builds use does not move)
  uint64_t cands[kWavesPerBlock];
};

// Scalar (wave-uniform) description of the item being searched.
struct Ctx {
  int N, n;
  int flags;
  int limit;       // effective depth limit
  uint32_t job;
  int Rreg;        // lane y holds R[y]
  int depth0;
  bool shared;     // other wavefronts may work on this request
  uint64_t expanded;
  uint64_t solutions;
  int sinceCheck;
  int home;        // this wave's queue shard and feed shard
  uint32_t rng;    // LCG state for random shard probes
  bool drained;    // batched: this wave saw every feed shard exhausted
  bool homeDone;   // batched: this wave saw its home feed shard exhausted
  uint32_t published;  // batched: requests published and not yet counted globally
  // batched look-ahead, taken while a request's record loads: the next
  // requests claimed from the home shard (preN of them from preG on; lane 0
  // holds the claim's result, read when needed)
  int preG;
  bool preValid;
  int preN;
  bool reorder;    // greedy order: frames visit the children gaining odd cycles first
  // request-level state (the request this wave started, or the item's)
  uint32_t jw;       // lane q < 32: word q of the request's 128-byte record
  bool everShared;   // some item of the request was donated
  uint64_t t0;       // device clock when the request started
  uint64_t totExpanded, totSolutions;  // finished items, while never shared
  int found, wdepth; // witness kept in LDS (w.wit), while never shared
  int wmove, wlen;   // the witness an expansion just met: last move and length
};

enum { kExpandOk = 0, kExpandFound = 1, kExpandLeaf = 2 };

// Nodes after which an item counts as heavy: while unstarted requests
// remain, it keeps up to kEarlyStock of its subtrees queued.  Requests
// without slack (kFlagTight) are where long searches come from and qualify
// sooner; with a move bound of 8 or more (tens of thousands of nodes on
// average) at the first check.  Any item past kPriorityNodes gets the SIMD's
// issue priority.
constexpr int kPriorityNodes = 64;
constexpr int kHeavyNodes = 64;
constexpr int kHeavyNodesDeep = 0;
constexpr int kHeavyNodesSlack = 512;
constexpr int kEarlyStock = 1024;

// Packs the witness into the job state when this wave is the first finder.
__device__ void record_witness(const Params &P, Ctx &c, WaveLds &w, int depth, int lastMove) {
  JobState *st = &P.states[c.job];
  // the witness is published by the release on `pending` at the item's end;
  // a request no other wavefront has touched needs no atomic at all
  const int lane = lane_id();
  if (!c.everShared) {
    // no other wavefront ever held this request: the witness stays in LDS
    // and the result is published from there
    if (lane < depth) w.wit[lane] = (lane == depth - 1) ? (uint16_t)lastMove : w.path[lane];
    c.found = 1;
    c.wdepth = depth;
    wave_sync();
    return;
  }
  int won = 1;
  if (c.shared) won = lane0([&] { return cas_rlx(&st->found, 0, 1) ? 1 : 0; });
  else if (lane == 0) st_rlx(&st->found, 1);
  if (!won) return;
  // agent-scope stores, two moves per word: the item's release of `pending`
  // (release_item) then needs only to wait for them
  const int q0 = 2 * lane, q1 = 2 * lane + 1;
  const uint32_t m0 = q0 < depth ? (q0 == depth - 1 ? (uint32_t)lastMove : (uint32_t)w.path[q0]) : 0u;
  const uint32_t m1 = q1 < depth ? (q1 == depth - 1 ? (uint32_t)lastMove : (uint32_t)w.path[q1]) : 0u;
  if (q0 < depth) st_rlx(reinterpret_cast<uint32_t *>(st->moves) + lane, m0 | (m1 << 16));
  if (lane == 0) st_rlx(&st->depth, depth);
}

// Expands the node `pi` at depth d into frame d.  Returns kExpandFound when
// a witness was recorded (FIRST mode) and the item must stop, kExpandLeaf
// when the node has no expandable child (frame d is then left unwritten).
// Each candidate's odd-cycle delta is one lookup in dtab (delta_code).
// kCount: also add the node's candidate count to *cands (the counted builds).
template <bool kCount>
__device__ int expand(const Params &P, Ctx &c, WaveLds &w, const uint16_t *lex,
                      const uint8_t *dtab, uint64_t *cands, int d, uint64_t pi) {
  const int lane = lane_id();
  const int N = c.N, n = c.n;
  Frame &F = w.frames[d];
  c.expanded++;

  // ---- set-up: lane p owns position p ------------------------------------
  const int p = lane;
  const bool inN = p < N;
  const int sh = (4 * p) & 63;
  const int shPrev = (4 * (p - 1)) & 63;
  const int sPrev = (p == 0) ? n : (int)((pi >> shPrev) & 15u);
  const int om = bperm(sPrev & 31, c.Rreg);  // omega(s_p) = R(s_{p-1})
  // position of symbol om in pi (the nibble equal to om), or n for om = n
  const uint64_t x = pi ^ ((uint64_t)(om & 15) * 0x1111111111111111ULL);
  const uint64_t z = ((x - 0x1111111111111111ULL) & ~x & 0x8888888888888888ULL) |
                     0x8000000000000000ULL;
  int nx = om == n ? n : (__builtin_ctzll(z) >> 2);
  nx = inN ? nx : p;  // lanes past N: harmless fixed points
  (void)sh;

  // Cycle of position p under nx by pointer doubling.  Before round r the
  // window is nx^i(p) for 0 <= i <= 2^r: M holds its members, J = nx^(2^r)(p)
  // and D the first i at which the window's smallest member (the cycle's
  // leader once the window covers the cycle) occurs.  One bpermute per round
  // carries J | M << 5 | D << 22; four rounds cover cycles of up to 17, and
  // the rounds stop once no window is still growing (before round r a window
  // of fewer than 2^r + 1 members has closed on its cycle).
  int J = nx & 31;
  uint32_t M = (1u << (p & 31)) | (1u << J);
  int D = J < p ? 1 : 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (r >= 1 && ballot(__builtin_popcount(M) > (1u << r)) == 0) break;
    const uint32_t got =
        (uint32_t)bperm(J, (int)((uint32_t)J | ((M & 0x1FFFFu) << 5) | ((uint32_t)D << 22)));
    const uint32_t Mj = (got >> 5) & 0x1FFFFu;
    if (__builtin_ctz(Mj) < __builtin_ctz(M)) D = (1 << r) + (int)(got >> 22);
    M |= Mj;
    J = (int)(got & 31u);
  }
  const int leader = __builtin_ctz(M);
  const int L = __builtin_popcount(M);
  const bool isLeader = (p == leader);
  const int ec = __builtin_popcountll(ballot(inN && isLeader && (L & 1)));
  const uint64_t bpMask = ballot(inN && nx != p);
  // the triple's table bits come from three copies of the record whose
  // parity bits sit at different places (a: bits 8-9, b: 10-11, c: 12-13)
  const int info = (D << 24) | (leader << 16) | (((L & 1) | ((D & 1) << 1)) << 8);

  // ---- enumeration set --------------------------------------------------
  const int cd = d + 1;
  const uint32_t satMask = (uint32_t)uni((int)w.satTab[cd]);
  const uint32_t expMask = (uint32_t)uni((int)w.expTab[cd]);
  const bool allMoves = (c.flags & kFlagAllMoves) != 0;
  bool full = false;
  if (allMoves) {
    // triples through an omega fixed point change the odd-cycle count by
    // 0 or -2; enumerate them only when such a child can still matter
    uint32_t low = 1u << ec;
    if (ec >= 2) low |= 1u << (ec - 2);
    full = ((satMask | expMask) & low) != 0;
  }
  int m, einfo;
  if (full) {
    m = N;
    einfo = info;
  } else {
    m = __builtin_popcountll(bpMask);
    const bool isBp = inN && nx != p;
    const int r = rank_below(bpMask);
    einfo = fperm(isBp ? r : 63, info);
    if (isBp) F.pos[r] = (uint8_t)p;
  }
  const int einfoB = einfo + ((einfo & 0x300) * 3);
  const int einfoC = einfo + ((einfo & 0x300) * 15);
  m = uni(m);
  // C(m,3) and C(m,4) in unsigned arithmetic (both vanish below 3 and 4)
  const int ncand = (int)((uint32_t)(m * (m - 1) * (m - 2)) / 6u);
  if (kCount && lane == 0)
    __hip_atomic_fetch_add(cands, (uint64_t)ncand, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WAVEFRONT);
  const int nchunks = (ncand + 63) >> 6;
  const int off = (int)((uint32_t)(ncand * (m - 3)) >> 2);
  const int mode = c.flags & kFlagCountAll;
  // code -> witness / expandable child / greedy child: bit delta + 2 of a
  // class mask stands for a child with ec + delta odd cycles, i.e. bits
  // ec - 2, ec and ec + 2 of the depth's masks shifted down to 0, 2 and 4
  // (codes 6 and 7, rejected, hit no bit)
  const int codeShift = allMoves ? 3 : 0;
  const uint32_t satC = (uint32_t)(((uint64_t)satMask << 2) >> ec) & 0x15u;
  const uint32_t expC = (uint32_t)(((uint64_t)expMask << 2) >> ec) & ~satC & 0x15u;
  const uint32_t grdC = c.reorder ? (expC & 0x10u) : 0u;
  const uint64_t plainMask = c.reorder ? 0ULL : ~0ULL;
  // the frame's masks collect in two registers, lane ch holding mask ch and
  // lane 32 + ch mask nchunks + ch, and reach LDS in one store after the
  // loop (a lane-0 store per chunk cost scalar instructions)
  int mlo = 0, mhi = 0;

  // ---- candidates ------------------------------------------------------
  // software-pipelined: the next chunk's triple-table entries are read while
  // this chunk's cycle records are fetched
  int eNext = (int)lex[off + lane];
  uint64_t anyExp = 0;
  // one exit: a witness ends the loop through its bound
  int nend = nchunks;
  bool wfound = false;
  for (int ch = 0; ch < nend; ++ch) {
    const int t = (ch << 6) + lane;
    const int e = eNext;
    const int A = bperm(e & 31, einfo);
    const int B = bperm((e >> 5) & 31, einfoB);
    const int C = bperm((e >> 10) & 31, einfoC);
    eNext = (int)lex[off + t + 64];
    // the table index: parity bits, same-cycle bits, wrap bits
    const uint32_t ua = (uint32_t)A, ub = (uint32_t)B, uc = (uint32_t)C;
    int idx = (int)(((ua | ub | uc) >> 8) & 63u);
    idx |= ((ua >> 16) & 0xFFu) == ((ub >> 16) & 0xFFu) ? 64 : 0;
    idx |= ((ub >> 16) & 0xFFu) == ((uc >> 16) & 0xFFu) ? 128 : 0;
    idx |= ((ua >> 16) & 0xFFu) == ((uc >> 16) & 0xFFu) ? 256 : 0;
    idx |= (ua >> 24) < (ub >> 24) ? 512 : 0;
    idx |= (ub >> 24) < (uc >> 24) ? 1024 : 0;
    idx |= (uc >> 24) < (ua >> 24) ? 2048 : 0;
    // lanes past the candidate count read code 7 and drop out of the ballots
    // (a vector select: masking the ballots cost scalar instructions)
    if (lane >= ncand - (ch << 6)) idx = kDeltaCodes;
    const uint32_t code = ((uint32_t)dtab[idx] >> codeShift) & 7u;
    const uint64_t satb = ballot(((satC >> code) & 1u) != 0);
    const uint64_t expb = ballot(((expC >> code) & 1u) != 0);
    if (satb != 0) {
      if (mode != 0) {
        c.solutions += (uint64_t)__builtin_popcountll(satb);
      } else {
        wave_sync();
        const int wl = __builtin_ctzll(satb);
        const int we = bperm(wl, e);
        const int wi = we & 31, wj = (we >> 5) & 31, wk = (we >> 10) & 31;
        int pi_ = wi, pj = wj, pk = wk;
        if (!full) {
          pi_ = F.pos[wi];
          pj = F.pos[wj];
          pk = F.pos[wk];
        }
        c.solutions += 1;
        // recorded by run_item once its DFS loop has ended: lane-dependent
        // code on a loop's exit path makes the compiler treat the loop (and
        // every value it carries) as divergent
        c.wmove = uni(pi_ | (pj << 5) | (pk << 10));
        c.wlen = cd;
        wfound = true;
        nend = 0;
      }
    }
    anyExp |= expb;
    // reordered frame: the greedy children (two more odd cycles) first; a
    // frame in candidate order keeps them all in the first nchunks masks
    // (the second half is then never scanned).  No branch on the order.
    const uint64_t greedy = ballot(((grdC >> code) & 1u) != 0);
    const uint64_t lo = greedy | (expb & plainMask), hi = expb & ~greedy;
    write_lane(mlo, (uint32_t)lo, ch);
    write_lane(mhi, (uint32_t)(lo >> 32), ch);
    write_lane(mlo, (uint32_t)hi, ch + 32);
    write_lane(mhi, (uint32_t)(hi >> 32), ch + 32);
  }
  if (wfound) return kExpandFound;
  // a node without expandable children leaves its frame unread: the search
  // resumes at the parent's
  if (anyExp == 0) return kExpandLeaf;
  if ((lane & 31) < nchunks)
    F.mask[lane < 32 ? lane : nchunks + lane - 32] = ((uint64_t)(uint32_t)mhi << 32) | (uint32_t)mlo;
  if (lane == 0) {
    F.pi = pi;
    F.hdr = hdr_make(off, nchunks, full, c.reorder);
  }
  wave_sync();
  return kExpandOk;
}

// Decodes bit `bit` of chunk `ch` of frame F into a packed move; every lane
// decodes its own bit (no cross-lane step).
__device__ __forceinline__ int frame_move_lane(const Frame &F, uint32_t hdr, const uint16_t *lex,
                                               int ch, int bit) {
  const int e = lex[hdr_off(hdr) + (hdr_cand_chunk(hdr, ch) << 6) + bit];
  int i = e & 31, j = (e >> 5) & 31, k = (e >> 10) & 31;
  if (!hdr_full(hdr)) {
    i = F.pos[i];
    j = F.pos[j];
    k = F.pos[k];
  }
  return i | (j << 5) | (k << 10);
}

// The frame as 8-byte words, as run_item reads it (one word per lane).
constexpr int kFrameWords = (int)(sizeof(Frame) / 8);
constexpr int kHdrWord = (int)(offsetof(Frame, hdr) / 8);
static_assert(offsetof(Frame, hdr) % 8 == 0 && offsetof(Frame, pos) == offsetof(Frame, hdr) + 4 &&
                  sizeof(Frame) % 8 == 0 && kFrameWords <= 64,
              "run_item reads the frame by 8-byte word");


// ---------------------------------------------------------------------------
// Shared queue
// ---------------------------------------------------------------------------

// Tries to donate up to `want` still-unvisited children of the shallowest
// frame that has any (within the split depth).  Each donating lane pushes its
// own child into its own ring shard (home shard + its rank), so one donation
// lands on several shards and no tail pointer sees more than one CAS from
// it.  Children whose push fails stay with this wave.  Returns the number
// donated.
// `through`: the item records go out as agent-scope stores (written through
// to memory) and a wait for their completion publishes them, instead of
// plain stores and one L2 write-back.  Batches of few requests per
// wavefront, whose launches are mostly subtree hand-offs, gain from it
// (10,000 n = 10 distances 0.89 -> 0.79 ms, 2,000 n = 16 bound-8 requests
// 28.4 -> 23.8 ms); on the 100k mix it cost 6 % even when kept to the
// drained phase, so large batches keep the write-back.
__device__ int donate(const Params &P, Ctx &c, WaveLds &w, const uint16_t *lex, int d, int want,
                      bool through) {
  const int lane = lane_id();
  const int maxFd = c.limit - P.splitSlack;
  int fd = -1, ch = -1;
  for (int q = c.depth0; q <= d && q <= maxFd; ++q) {
    const uint32_t h = (uint32_t)uni((int)w.frames[q].hdr);
    const int nch = hdr_mask_chunks(h);
    for (int x = 0; x < nch; ++x) {
      if (uni64(w.frames[q].mask[x]) != 0) {
        fd = q;
        ch = x;
        break;
      }
    }
    if (fd >= 0) break;
  }
  if (fd < 0) return 0;
  if (P.served && P.itemCap > 0) {
    // a served request keeps at most its slot's budget of items pending
    const int pend = c.everShared ? lane0([&] { return ld_rlx(&P.states[c.job].pending); }) : 1;
    if (P.itemCap - pend < want) want = P.itemCap - pend;
    if (want < 1) return 0;
  }
  Frame &F = w.frames[fd];
  const uint32_t fhdr = (uint32_t)uni((int)F.hdr);
  uint64_t bits = uni64(F.mask[ch]);
  // keep at most `want` of them (the lowest bits go)
  if (want < 1) want = 1;
  if (__builtin_popcountll(bits) > want) {
    uint64_t keep = bits;
    for (int q = 0; q < want; ++q) keep &= keep - 1;
    bits ^= keep;
  }
  const uint64_t cap = P.qcap;
  const bool mine = (bits >> lane) & 1ULL;
  int shard = 0;
  uint64_t pos = 0;
  bool ok = false;
  if (mine) {
    shard = (c.home + rank_below(bits)) % kQueueShards;
    QueueState *q = &P.queue[shard];
    uint64_t *sq = P.seq + (uint64_t)shard * cap;
    for (int tries = 0; tries < 2 && !ok; ++tries) {
      const uint64_t t = ld_rlx(&q->tail);
      const uint64_t s = ld_rlx(&sq[t % cap]);
      if (s == t) {
        if (cas_rlx(&q->tail, t, t + 1)) {
          ok = true;
          pos = t;
        }
      } else if ((int64_t)(s - t) < 0) {
        break;  // shard full
      }
    }
  }
  const uint64_t okBits = ballot(ok);
  const int cnt = __builtin_popcountll(okBits);
  if (cnt == 0) return 0;
  // account before publishing so the request cannot complete underneath;
  // the release fence below orders these before any consumer's claim
  if (lane == 0) {
    JobState *st = &P.states[c.job];
    if (!c.everShared) {
      // first donation: the request's shared state is written only now,
      // from registers (a FIRST search still running has no witness yet;
      // the counts of finished IDDFS rounds join here) with agent-scope
      // stores, and the fence below publishes the items
      st_rlx(&st->pending, 1 + cnt);
      st_rlx(&st->found, 0);
      st_rlx(&st->limit, c.limit);
      st_rlx(&st->split, cnt);
      st_rlx(&st->expanded, c.totExpanded);
      st_rlx(&st->solutions, c.totSolutions);
      st_rlx(&st->depth, 0);
      st_rlx(&st->t0, c.t0);
    } else {
      add_rlx(&st->pending, cnt);
      add_rlx(&st->split, cnt);
    }
    add_rlx(&P.ctrl->balance[c.home % kBalanceShards].v, (int64_t)cnt);
  }
  c.everShared = true;
  const uint64_t slot = (uint64_t)shard * cap + pos % cap;
  if (ok) {
    const int mv = frame_move_lane(F, fhdr, lex, ch, lane);
    Item it;
    it.pi = apply_move(F.pi, mv & 31, (mv >> 5) & 31, (mv >> 10) & 31);
    it.job = c.job;
    it.depth = (uint32_t)(fd + 1);
#pragma unroll
    for (int q = 0; q < 10; ++q)
      it.moves[q] = (q < fd) ? w.path[q] : (q == fd ? (uint16_t)mv : (uint16_t)0);
    it.pad[0] = it.pad[1] = it.pad[2] = 0;
    if (through) {
      const uint64_t *iw = reinterpret_cast<const uint64_t *>(&it);
      uint64_t *dst = reinterpret_cast<uint64_t *>(&P.items[slot]);
#pragma unroll
      for (int q = 0; q < 6; ++q) st_rlx(dst + q, iw[q]);
    } else {
      P.items[slot] = it;
    }
  }
  if (through) complete_mem();
  else fence_rel();  // one L2 write-back publishes every item of this donation
  if (ok) st_rlx(&P.seq[slot], pos + 1);
  if (lane == 0) F.mask[ch] &= ~okBits;
  wave_sync();
  c.shared = true;
  return cnt;
}

// Claims one item from the home shard or one pseudo-random shard: lanes 0
// and 1 probe one shard each at the same time (head, then the head slot's
// sequence number), and lane 0 claims the first published slot with a CAS.
// While the batched feed lasts (`feeding`) only the random shard is probed:
// the early stock of a heavy request spreads over the shards after its
// donor's, so the home shard is no better a guess, and every probe is two
// device-scope reads.  Returns 1 and fills `it` (and the wave's path prefix)
// when it got one.
__device__ int pop_item(const Params &P, Ctx &c, Item &it, WaveLds &w, bool feeding) {
  const int lane = lane_id();
  const uint64_t cap = P.qcap;
  c.rng = c.rng * 1664525u + 1013904223u;
  const int other = (int)((c.rng >> 8) % kQueueShards);
  const int myShard = (lane == 0 && !feeding) ? c.home : other;
  bool ready = false;
  uint64_t hl = 0;
  if (lane < (feeding ? 1 : 2)) {
    // relaxed polling (served: the acquire fence is paid only once a slot is
    // claimed)
    hl = ld_rlx(&P.queue[myShard].head);
    ready = ld_rlx(&P.seq[(uint64_t)myShard * cap + hl % cap]) == hl + 1;
  }
  const uint64_t rdy = ballot(ready);
  if (rdy == 0) return 0;
  const int src = __builtin_ctzll(rdy);
  const int shard = __builtin_amdgcn_readlane(myShard, src);
  const uint64_t h = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(hl >> 32), src) << 32) |
                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)hl, src);
  if (lane0([&] { return cas_rlx(&P.queue[shard].head, h, h + 1) ? 1 : 0; }) == 0) return 0;
  // The record is read with agent-scope loads (lane q < 9: word q — pi, job,
  // depth, moves), which the producer's release made current, so the
  // batched path needs no acquire: an agent-scope acquire invalidates the
  // XCD's L2, and with many hand-offs in flight those invalidations (and the
  // producers' write-backs) stalled every wave's memory traffic (a probe
  // took ~16 us on configuration 3).  Served: the receiver reads the
  // request's job-table copy with plain loads, so it still acquires.
  if (P.served) fence_acq();
  const uint64_t slot = (uint64_t)shard * cap + h % cap;
  uint32_t *srcw = reinterpret_cast<uint32_t *>(&P.items[slot]);
  const uint32_t word = lane < 9 ? ld_rlx(srcw + lane) : 0u;
  it.pi = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)word, 1) << 32) |
          (uint32_t)__builtin_amdgcn_readlane((int)word, 0);
  it.job = (uint32_t)__builtin_amdgcn_readlane((int)word, 2);
  it.depth = (uint32_t)__builtin_amdgcn_readlane((int)word, 3);
  const uint32_t mw = (uint32_t)bperm(4 + ((lane & 15) >> 1), (int)word);
  if (lane < 10) w.path[lane] = (uint16_t)(mw >> (16 * (lane & 1)));
  // the item is in registers once every load has returned: free the slot
  __builtin_amdgcn_s_waitcnt(0);
  if (lane == 0) st_rlx(&P.seq[slot], h + cap);
  wave_sync();
  return 1;
}

// ---------------------------------------------------------------------------
// Requests
// ---------------------------------------------------------------------------

// End of feed shard s's request range (host-written size fields).
__device__ __forceinline__ uint32_t feed_end(const Params &P, int s) {
  const uint32_t per = P.perShard, nj = P.numJobs;
  const uint64_t e = (uint64_t)(s + 1) * per;
  return e < nj ? (uint32_t)e : nj;
}

// Requests one look-ahead claim takes (one atomic on the shard's cursor).
// The last Params::lookMargin requests of a feed shard are not claimed ahead
// (set at launch from the batch and grid sizes).
constexpr int kLookAheadClaim = 2;

// Batched feed: claims the next unstarted request, from the home feed shard
// first, then from whichever other shard still has some (every lane inspects
// one shard's cursor).  Returns -1 once every shard is exhausted.
__device__ int claim_job(const Params &P, Ctx &c) {
  if (c.drained) return -1;
  Control *ctl = P.ctrl;
  const int lane = lane_id();
  const uint32_t per = P.perShard;
  const uint32_t nj = P.numJobs;
  auto shard_end = [&](uint32_t s) -> uint32_t {
    const uint64_t e = (uint64_t)(s + 1) * per;
    return e < nj ? (uint32_t)e : nj;
  };
  // one atomic claims a request; the counter of a shard without requests is
  // never touched (the host does not reset it)
  auto take = [&](int s) {
    return lane0([&] {
      if ((uint64_t)s * per >= nj) return -1;
      const uint32_t g = add_rlx(&ctl->feed[s].next, 1u);
      return g < shard_end((uint32_t)s) ? (int)g : -1;
    });
  };
  int got = -1;
  if (!c.homeDone) {
    got = take(c.home);
    if (got >= 0) return got;
    c.homeDone = true;
  }
  if (lane0([&] { return (int)ld_rlx(&ctl->drained); }) != 0) {
    c.drained = true;
    return -1;
  }
  uint64_t tried = 1ULL << c.home;
  const int r = (c.home + 1) & (kFeedShards - 1);
  for (;;) {
    // shards without requests are never read: the host does not reset them
    const bool has = (uint64_t)lane * per < nj && ld_rlx(&ctl->feed[lane].next) < shard_end((uint32_t)lane);
    const uint64_t avail = ballot(has) & ~tried;
    if (avail == 0) break;
    // first available shard after the home shard, cyclically
    const uint64_t rot = r ? ((avail >> r) | (avail << (64 - r))) : avail;
    const int pick = (__builtin_ctzll(rot) + r) & (kFeedShards - 1);
    got = take(pick);
    if (got >= 0) return got;
    tried |= 1ULL << pick;
  }
  c.drained = true;
  if (lane == 0 && ld_rlx(&ctl->drained) == 0u) st_rlx(&ctl->drained, 1u);
  return -1;
}

// Lane q < 32 reads word q of request j's 128-byte record: one coalesced
// load brings everything the search needs into the wave.
__device__ __forceinline__ uint32_t job_word(const Params &P, uint32_t j) {
  const int lane = lane_id();
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&P.jobs[j]);
  return lane < 32 ? src[lane] : 0u;
}
static_assert(offsetof(Job, R) == 8 && offsetof(Job, N) == 28 && offsetof(Job, maxDepth) == 29 &&
                  offsetof(Job, flags) == 30 && offsetof(Job, firstLimit) == 31 &&
                  offsetof(Job, sat) == 32 && offsetof(Job, reach) == 72 &&
                  offsetof(Job, tag) == 124,
              "the record is read by word");
// word 7: N | maxDepth << 8 | flags << 16 | firstLimit << 24
__device__ __forceinline__ uint32_t job_w7(const Ctx &c) {
  return (uint32_t)__builtin_amdgcn_readlane((int)c.jw, 7);
}
__device__ __forceinline__ uint64_t job_pi0(const Ctx &c) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)c.jw, 1) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)c.jw, 0);
}
__device__ __forceinline__ int job_max_depth(const Ctx &c) { return (int)((job_w7(c) >> 8) & 0xFFu); }

// Loads the request's parameters (from c.jw) into the wave context and LDS
// tables.
__device__ void load_job(const Params &P, Ctx &c, WaveLds &w, uint32_t j, int limit) {
  const int lane = lane_id();
  c.job = j;
  const uint32_t w7 = job_w7(c);
  c.N = (int)(w7 & 0xFFu);
  c.n = c.N - 1;
  c.flags = (int)((w7 >> 16) & 0xFFu);
  // FIRST searches take the greedy order unless the request asks for the
  // lexicographic one; COUNT_ALL visits everything, in lexicographic order
  c.reorder = (c.flags & (kFlagCountAll | kFlagLexOrder)) == 0;
  c.limit = limit;
  // R[y] is byte 8 + y of the record, sat[d] word 8 + d, reach[d] word 18 + d
  const uint32_t rword = (uint32_t)bperm(lane < 20 ? 2 + (lane >> 2) : 2, (int)c.jw);
  c.Rreg = lane < 20 ? (int)((rword >> (8 * (lane & 3))) & 0xFFu) : 0;
  const int tl = lane <= kMaxDepth ? lane : 0;
  const uint32_t satw = (uint32_t)bperm(8 + tl, (int)c.jw);
  const uint32_t reachw = (uint32_t)bperm(18 + tl, (int)c.jw);
  if (lane <= kMaxDepth) {
    uint32_t s, r;
    if (c.flags & kFlagIddfs) {
      s = lane >= 1 ? (1u << c.N) : 0u;
      const int left = limit - lane;
      const int lo = c.N - 2 * left;
      r = left < 0 ? 0u : (lo <= 0 ? ((2u << c.N) - 1u) : (((2u << c.N) - 1u) & ~((1u << lo) - 1u)));
    } else {
      s = satw;
      r = reachw;
    }
    w.satTab[lane] = s;
    w.expTab[lane] = (lane < limit) ? r : 0u;
  }
  wave_sync();
}

__device__ bool release_item(const Params &P, Ctx &c);

// Runs one item (a node to expand and its whole subtree) to completion.
// Returns true when this wave released the request's last pending item.
template <bool kCount>
__device__ bool run_item(const Params &P, Ctx &c, WaveLds &w, const uint16_t *lex,
                         const uint8_t *dtab, uint64_t *cands, uint64_t pi, int depth0) {
  const int lane = lane_id();
  depth0 = uni(depth0);
  c.depth0 = depth0;
  c.expanded = 0;
  c.solutions = 0;
  c.sinceCheck = 0;
  const bool countAll = (c.flags & kFlagCountAll) != 0;
  JobState *st = &P.states[c.job];

  bool stop = false;
  bool met = false;  // this item met a witness (FIRST)
  if (c.shared && !countAll) stop = lane0([&] { return ld_rlx(&st->found); }) != 0;
  if (!stop) {
    const int r0 = uni(expand<kCount>(P, c, w, lex, dtab, cands, depth0, pi));
    stop = r0 != kExpandOk;
    met = r0 == kExpandFound;
  }
  int d = depth0;
  // the whole top frame in one LDS round trip: lane q < kFrameWords holds
  // 8-byte word q (word 0 the sequence, words 1..22 the child masks, word 23
  // the header and positions 0..3, words 24 and 25 positions 4..19)
  const int fq = lane < kFrameWords ? lane : 0;
  while (!stop) {
    d = uni(d);
    Frame &F = w.frames[d];
    const uint64_t fword = reinterpret_cast<const uint64_t *>(&F)[fq];
    const uint32_t hdr = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)fword, kHdrWord);
    // next expandable child: the first nonzero mask chunk (chunk q sits in
    // lane q + 1)
    const int q = lane - 1;
    const uint64_t nz =
        ballot((int)((uint32_t)q < (uint32_t)hdr_mask_chunks(hdr)) & (int)(fword != 0));
    if (nz == 0) {
      if (d == depth0) break;
      --d;
      continue;
    }
    const int chl = __builtin_ctzll(nz);
    const int ch = chl - 1;
    const uint64_t bits = (uint64_t)lane_value64((int64_t)fword, chl);
    const int bit = __builtin_ctzll(bits);
    // the child's move: lanes 0..2 each look up one element's position in
    // the frame's position map (one LDS read instead of a scalar chain of
    // lane reads, shifts and masks per element)
    int mv;
    {
      const int e = (int)lex[hdr_off(hdr) + (hdr_cand_chunk(hdr, ch) << 6) + bit];
      const int sh = lane < 3 ? 5 * lane : 0;
      const int el = (e >> sh) & 31;
      const int part = (hdr_full(hdr) ? el : (int)F.pos[el]) << sh;
      mv = __builtin_amdgcn_readlane(part, 0) | __builtin_amdgcn_readlane(part, 1) |
           __builtin_amdgcn_readlane(part, 2);
    }
    const uint64_t child = apply_move((uint64_t)lane_value64((int64_t)fword, 0), mv & 31,
                                      (mv >> 5) & 31, (mv >> 10) & 31);
    if (lane == 0) {
      F.mask[ch] = bits & (bits - 1);
      w.path[d] = (uint16_t)mv;
    }
    wave_sync();
    // sharing and cancellation checks, every P.checkEvery expansions (one
    // device-scope round trip each; fewer expansions between them spread a
    // heavy request sooner).  Subtrees are donated only when wavefronts sit
    // idle, no unstarted request is waiting for them and this item has
    // proved heavy: while requests remain, request-level parallelism keeps
    // every wavefront busy for free.
    if (++c.sinceCheck >= P.checkEvery) {
      c.sinceCheck = 0;
      // an item that has run long gets the SIMD's issue priority over the
      // light requests sharing it, so it does not become the launch's tail
      // (a request with slack can run a few hundred nodes without ever
      // counting as heavy); light work fills the remaining issue slots
      const uint64_t heavyAt =
          !(c.flags & kFlagTight) ? kHeavyNodesSlack
          : job_max_depth(c) >= kDeepBound ? kHeavyNodesDeep : kHeavyNodes;
      if (c.expanded >= kPriorityNodes) __builtin_amdgcn_s_setprio(2);
      // balance = queued items - idle wavefronts (lanes 0..7, one shard
      // each) and whether requests are still waiting (lanes 8 and 9), in
      // one round trip
      int64_t v = 0;
      if (lane < kBalanceShards) v = ld_rlx(&P.ctrl->balance[lane].v);
      else if (lane == kBalanceShards)
        v = P.served ? (int64_t)ld_rlx(&P.ctrl->consumed) : (int64_t)ld_rlx(&P.ctrl->drained);
      else if (lane == kBalanceShards + 1 && P.served)
        v = (int64_t)ld_rlx(&P.ctrl->submittedMirror);
      const int64_t bal = shard_lane_sum(v);
      const int64_t f0 = lane_value64(v, kBalanceShards);
      const bool drained = P.served ? (uint64_t)f0 >= (uint64_t)lane_value64(v, kBalanceShards + 1)
                                    : f0 != 0;
      // requests still waiting: a heavy item keeps a small stock of its
      // subtrees queued, which wavefronts take before new requests, so heavy
      // requests spread early instead of crawling on one wavefront that
      // shares its SIMD with light ones
      const int64_t room = drained ? -bal : (c.expanded >= heavyAt ? kEarlyStock - bal : 0);
      const int want = room <= 0 ? 0 : (room < (int64_t)P.donateMax ? (int)room : P.donateMax);
      if (want > 0) donate(P, c, w, lex, d, want, !P.served && P.lookMargin == kNoLookAhead);
      if (c.shared && !countAll) {
        if (lane0([&] { return ld_rlx(&st->found); }) != 0) break;
      }
    }
    const int r = uni(expand<kCount>(P, c, w, lex, dtab, cands, d + 1, child));
    if (r == kExpandFound) {
      met = true;
      break;
    }
    if (r == kExpandOk) ++d;
  }
  if (met) record_witness(P, c, w, c.wlen, c.wmove);
  __builtin_amdgcn_s_setprio(0);
  return uni(release_item(P, c) ? 1 : 0) != 0;
}

// Flushes the item's counters and releases it (the release publishes the
// counters and any witness to the wavefront that finalizes the request).
// Returns true when this wave released the request's last pending item.
__device__ bool release_item(const Params &P, Ctx &c) {
  const int lane = lane_id();
  JobState *st = &P.states[c.job];
  const uint64_t expanded = c.expanded;
  if (!c.shared) {
    // a request no other wavefront ever held keeps its counts in registers
    if (!c.everShared) {
      c.totExpanded += expanded;
      c.totSolutions += c.solutions;
      return true;
    }
    // root item of a later IDDFS round of a once-shared request: this wave
    // is the only one holding it, so it finalizes with plain stores
    if (lane == 0) {
      st->expanded += expanded;
      st->solutions += c.solutions;
      st->pending = 0;
    }
    wave_sync();
    return true;
  }
  int last = 0;
  if (lane == 0) {
    if (expanded) add_rlx(&st->expanded, expanded);
    if (c.solutions) add_rlx(&st->solutions, c.solutions);
    // every write an item makes to its request's state is an agent-scope
    // atomic or atomic store (counts, witness, limit): waiting for them to
    // complete orders them before the count below, no L2 write-back needed
    complete_mem();
    last = __hip_atomic_fetch_add(&st->pending, -1, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) == 1 ? 1 : 0;
  }
  last = uni(last);
  if (last) fence_acq();
  return last != 0;
}

// Word q of the 80-byte DevResult record, written by lane q (q < 20).  A
// request no other wavefront ever held is described by the wave's own
// registers and LDS; otherwise by its state in HBM (the caller has already
// acquired what other wavefronts wrote).
__device__ __forceinline__ uint32_t result_word(const Params &P, const Ctx &c, const WaveLds &w,
                                                uint32_t j, int q) {
  if (q == 13) return (uint32_t)__builtin_amdgcn_readlane((int)c.jw, 31);  // tag
  if (q == 18 || q == 19) {
    const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
    return q == 18 ? (uint32_t)t1 : (uint32_t)(t1 >> 32);
  }
  if (!c.everShared) {
    switch (q) {
      case 0: return (uint32_t)c.found;
      case 1: return c.found ? (uint32_t)c.wdepth : 0u;
      case 3: return (uint32_t)c.limit;
      case 4: return (uint32_t)c.totExpanded;
      case 5: return (uint32_t)(c.totExpanded >> 32);
      case 6: return (uint32_t)c.totSolutions;
      case 7: return (uint32_t)(c.totSolutions >> 32);
      case 16: return (uint32_t)c.t0;
      case 17: return (uint32_t)(c.t0 >> 32);
      default: break;
    }
    if (q >= 8 && q < 13) return reinterpret_cast<const uint32_t *>(w.wit)[q - 8];
    return 0u;  // split, seq, pad
  }
  const uint32_t *st = reinterpret_cast<const uint32_t *>(&P.states[j]);
  // found depth split limit | expanded solutions | moves[10] | tag seq pad | t0 t1
  int src = -1;
  if (q < 4) src = (int)((0x02030801u >> (8 * q)) & 0xffu);
  else if (q < 8) src = q;                    // expanded, solutions
  else if (q < 13) src = q + 1;               // moves
  else if (q == 16 || q == 17) src = q;       // t0
  if (src >= 0) return st[src];
  return 0u;  // seq, pad
}
static_assert(offsetof(WaveLds, wit) % 4 == 0, "the witness is read by word");
static_assert(offsetof(JobState, found) == 4 && offsetof(JobState, limit) == 8 &&
                  offsetof(JobState, split) == 12 && offsetof(JobState, expanded) == 16 &&
                  offsetof(JobState, depth) == 32 && offsetof(JobState, moves) == 36 &&
                  offsetof(JobState, t0) == 64,
              "result_word reads JobState by word");
static_assert(offsetof(DevResult, moves) == 32 && offsetof(DevResult, tag) == 52 &&
                  offsetof(DevResult, t0) == 64,
              "result_word writes DevResult by word");

__device__ void publish_result(const Params &P, Ctx &c, const WaveLds &w, uint32_t j) {
  const int lane = lane_id();
  if (!P.served) {
    // the host reads batched results after the kernel has ended
    if (lane < 20) reinterpret_cast<uint32_t *>(&P.results[j])[lane] = result_word(P, c, w, j, lane);
    c.published++;  // added to the shared count once this wave runs out of work
  } else {
    // record into the host ring slot of job slot j and its index into the
    // completion ring, then the sequence word the host polls: the host reads
    // the index and the record only after it has acquired that word, so one
    // system-scope release orders both before it
    if (lane < 20)
      reinterpret_cast<uint32_t *>(&P.hostResults[j])[lane] = result_word(P, c, w, j, lane);
    uint32_t cslot = 0, cnext = 0;
    if (lane == 0) {
      const uint64_t cseq = add_rlx(&P.ctrl->completions, (uint64_t)1);
      cslot = (uint32_t)(cseq & (uint64_t)(P.ringCap - 1));
      cnext = (uint32_t)(cseq + 1);
      P.hostCompletionIdx[2 * cslot] = j;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (lane == 0) {
      __hip_atomic_store(&P.hostCompletionIdx[2 * cslot + 1], cnext, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
      add_rlx(&P.ctrl->jobsDone, 1u);
    }
  }
  wave_sync();
}

// Starts request j (root item) in this wave; returns its first depth limit.
__device__ int start_job(Ctx &c) {
  const uint32_t w7 = job_w7(c);
  const int limit = (int)((((w7 >> 16) & kFlagIddfs) ? (w7 >> 24) : (w7 >> 8)) & 0xFFu);
  c.t0 = __builtin_amdgcn_s_memrealtime();
  c.everShared = false;
  c.totExpanded = 0;
  c.totSolutions = 0;
  c.found = 0;
  c.wdepth = 0;
  // the request's state in HBM is written at its first donation, if any:
  // until then no other wavefront reads it, and most requests never donate
  return limit;
}

// Served loop: takes the next submitted request (one claim attempt).  The submit ring and its
// count live in host memory and are read only with atomic loads, which reach
// host memory on every access (a probe that warmed the line with a plain
// load still saw the host's next write within one ~1.2 us poll), so no cache
// maintenance sits on this path: an agent-scope release or acquire here
// would walk an XCD's L2, tens of microseconds under load, to protect data
// that is never read through the cache.  One idle wavefront at a time (at
// most every ~2 us by the chip's 100 MHz counter) copies the host's count
// into a device-memory mirror; a claimer takes entry r with a CAS on
// `consumed`, reads it over the bus, keeps it in jwOut (lane q < 32 holds word
// q, as job_word returns it) and copies it to the job table for wavefronts
// that may later receive subtrees of it (a donation's release orders those
// stores before they are read).
__device__ int take_submitted(const Params &P, uint32_t &jOut, uint32_t &jwOut) {
  const int lane = lane_id();
  int got = 0;
  uint64_t r = 0;
  if (lane == 0) {
    uint64_t avail = ld_rlx(&P.ctrl->submittedMirror);
    r = ld_rlx(&P.ctrl->consumed);
    if (r >= avail) {
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      const uint64_t last = ld_rlx(&P.ctrl->lastPoll);
      if (now - last > 200ULL && cas_rlx(&P.ctrl->lastPoll, last, now)) {
        // one read gives the count and the stop bit together
        const uint64_t h =
            __hip_atomic_load(&P.host->submitted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (h & kServeStop) st_rlx(&P.ctrl->stopMirror, (uint64_t)1);
        const uint64_t cnt = h & ~kServeStop;
        uint64_t cur = ld_rlx(&P.ctrl->submittedMirror);
        while (cur < cnt && !cas_rlx(&P.ctrl->submittedMirror, cur, cnt))
          cur = ld_rlx(&P.ctrl->submittedMirror);
      }
      avail = ld_rlx(&P.ctrl->submittedMirror);
    }
    // one attempt: a wavefront that loses the race goes back to its loop.
    // Retrying at once turns a backlog into a storm of same-address CASes
    // (each ~12 ns at the memory side, every loser reloading and trying
    // again), which capped the loop near 100k requests/s with 128 or more
    // requests admitted; with one attempt it serves 0.7-1.5M/s
    if (r < avail && cas_rlx(&P.ctrl->consumed, r, r + 1)) got = 1;
  }
  got = uni(got);
  if (!got) return 0;
  r = uni64(r);
  // one read of the 128-byte record over the bus, 16 bytes per lane; ring
  // entries are transport only and the record's tag names its job slot
  const uint32_t e = (uint32_t)(r & (uint64_t)(P.ringCap - 1));
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&P.hostJobs[e]);
  uint32_t v[4] = {0, 0, 0, 0};
  if (lane < 8) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      v[q] = __hip_atomic_load(src + 4 * lane + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  const uint32_t j = (uint32_t)__builtin_amdgcn_readlane((int)v[3], 7);  // word 31: tag
  if (lane < 8) {
    uint32_t *dst = reinterpret_cast<uint32_t *>(const_cast<Job *>(&P.jobs[j]));
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[4 * lane + q] = v[q];
  }
  // lane q < 32 takes word q: element q % 4 of lane q / 4
  const int from = (lane & 31) >> 2, sel = lane & 3;
  const uint32_t w0 = (uint32_t)bperm(from, (int)v[0]);
  const uint32_t w1 = (uint32_t)bperm(from, (int)v[1]);
  const uint32_t w2 = (uint32_t)bperm(from, (int)v[2]);
  const uint32_t w3 = (uint32_t)bperm(from, (int)v[3]);
  const uint32_t wq = sel == 0 ? w0 : sel == 1 ? w1 : sel == 2 ? w2 : w3;
  jwOut = lane < 32 ? wq : 0u;
  jOut = j;
  return 1;
}

// The persistent search loop of one block; the kernels below instantiate it
// under two register budgets.
template <bool kCount>
__device__ __forceinline__ void search_blocks(Params P, BlockLds &lds) {
  for (int q = threadIdx.x; q < kLexEntries; q += blockDim.x) lds.lex[q] = P.lex[q];
  for (int q = threadIdx.x; q < kDeltaCodes + 8; q += blockDim.x)
    lds.dtab[q] = q < kDeltaCodes ? (uint8_t)delta_code(q) : (uint8_t)0xFFu;
  if (threadIdx.x < 96) lds.lexPad[threadIdx.x] = 0;
  __syncthreads();

  // the wave's index in the block: the same on every lane, which the
  // compiler cannot see in threadIdx.x, so it is made wave-uniform here
  const int wid = uni((int)(threadIdx.x >> 6));
  uint64_t *cands = &lds.cands[wid];
  if (kCount && (threadIdx.x & 63) == 0) *cands = 0;
  WaveLds &w = lds.wave[wid];
  const uint16_t *lex = lds.lex;
  const int lane = lane_id();

  Ctx c;
  c.shared = false;
  c.home = (int)((blockIdx.x * kWavesPerBlock + wid) % kQueueShards);
  c.rng = 0x9E3779B9u * (blockIdx.x * kWavesPerBlock + wid + 1);
  c.drained = false;
  c.homeDone = false;
  c.published = 0;
  c.preValid = false;
  c.preN = 0;
  bool idle = false;
  int backoff = 1;
  for (;;) {
    // The wave's loop-carried state is the same on every lane.  Said so
    // here, the compiler keeps it in scalar registers and compiles every
    // loop of the search as a uniform one: without this it cannot rule out
    // that a lane-dependent branch (a lane-0 store, a per-lane probe) on some
    // exit path left lanes disagreeing, and falls back to exec-mask loops,
    // divergent phis and 114 VGPRs instead of 90.
    c.preValid = uni(c.preValid ? 1 : 0) != 0;
    c.preN = uni(c.preN);
    c.drained = uni(c.drained ? 1 : 0) != 0;
    c.homeDone = uni(c.homeDone ? 1 : 0) != 0;
    idle = uni(idle ? 1 : 0) != 0;
    backoff = uni(backoff);
    c.shared = uni(c.shared ? 1 : 0) != 0;
    c.everShared = uni(c.everShared ? 1 : 0) != 0;
    c.published = (uint32_t)uni((int)c.published);
    Item it;
    uint32_t j = 0;
    uint32_t servedJw = 0;
    // While the batched feed lasts, queued items are the early stock of
    // heavy requests, taken before new requests: a wavefront probes one ring
    // shard at every turn (probing only every 4th turn after a miss saved
    // nothing measurable on the batched mix and made the emitter-shaped
    // stream 5 % slower, its heavy requests' stock waiting longer).
    const bool feeding = !P.served && !c.drained;
    bool haveItem = uni(pop_item(P, c, it, w, feeding)) != 0;
    bool haveJob = false;
    if (!haveItem) {
      if (!P.served) {
        int jj = -1;
        if (c.preValid) {
          const int g = uni(c.preG);
          if (--c.preN == 0) c.preValid = false;
          else c.preG = g + 1;
          if (g < (int)feed_end(P, c.home)) {
            jj = g;
          } else {
            c.homeDone = true;
            c.preValid = false;
          }
        }
        if (jj < 0) jj = uni(claim_job(P, c));
        if (jj >= 0) {
          j = (uint32_t)jj;
          haveJob = true;
        }
      } else {
        haveJob = uni(take_submitted(P, j, servedJw)) != 0;
      }
    }
    if (!haveItem && !haveJob) {
      if (!idle) {
        if (lane == 0) add_rlx(&P.ctrl->balance[c.home % kBalanceShards].v, (int64_t)-1);
        idle = true;
      }
      int done = 0;
      if (!P.served) {
        // this wave's count joins its shard (the add returns the shard's
        // value) while the other lanes read the other shards
        const int h = c.home % kBalanceShards;
        int64_t v = 0;
        if (lane < kBalanceShards) {
          if (lane == h && c.published)
            v = add_rlx(&P.ctrl->published[lane].v, (int64_t)c.published) + (int64_t)c.published;
          else
            v = ld_rlx(&P.ctrl->published[lane].v);
        }
        done = shard_lane_sum(v) >= (int64_t)P.numJobs ? 1 : 0;
        c.published = 0;
      } else {
        // device-side mirrors only: host memory is read by the rate-limited
        // poll in take_submitted, never by every idle wavefront
        done = lane0([&] {
          if (ld_rlx(&P.ctrl->stopMirror) == 0) return 0;
          const uint64_t cons = ld_rlx(&P.ctrl->consumed);
          return (cons >= ld_rlx(&P.ctrl->submittedMirror) &&
                  ld_rlx(&P.ctrl->jobsDone) >= (uint32_t)cons) ? 1 : 0;
        });
      }
      if (done) {
        if (kCount && lane == 0) add_rlx(&P.ctrl->candidates, *cands);
        break;
      }
      // idle polling touches a few words of HBM and contends with the
      // working wavefronts' atomics and cache write-backs; back off (about
      // 0.4 us per unit, up to P.idleBackoff units)
      for (int q = 0; q < backoff; ++q) __builtin_amdgcn_s_sleep(16);
      const int maxBackoff = P.idleBackoff;
      backoff = backoff < maxBackoff ? backoff * 2 : maxBackoff;
      continue;
    }
    // balance bookkeeping: an idle wave that takes an item leaves it
    // unchanged (one item fewer queued, one wave fewer idle)
    if (haveItem != idle && lane == 0)
      add_rlx(&P.ctrl->balance[c.home % kBalanceShards].v, haveItem ? (int64_t)-1 : (int64_t)1);
    idle = false;
    backoff = 1;

    uint64_t pi;
    int depth0, limit;
    if (haveJob) {
      c.jw = P.served ? servedJw : job_word(P, j);
      limit = start_job(c);
      // look ahead: once the last claim is used up, claim the home shard's
      // next kLookAheadClaim requests with one atomic, whose result is read
      // at the next request's start.  Not at a request without slack (a
      // held request would wait behind what may be a long search) and not
      // within lookMargin of the shard's end
      if (!P.served && !c.drained && !c.preValid && !c.homeDone &&
          !((job_w7(c) >> 16) & kFlagTight) &&
          (int64_t)j + P.lookMargin < (int64_t)feed_end(P, c.home) &&
          j >= (uint32_t)c.home * P.perShard) {
        if (lane == 0) c.preG = (int)add_rlx(&P.ctrl->feed[c.home].next, (uint32_t)kLookAheadClaim);
        c.preValid = true;
        c.preN = kLookAheadClaim;
      }
      pi = job_pi0(c);
      depth0 = 0;
      c.shared = false;
    } else {
      j = (uint32_t)uni((int)it.job);
      c.job = j;
      c.shared = true;
      c.everShared = true;
      c.jw = job_word(P, j);
      // a request that already has its witness: release the item unopened
      // (COUNT_ALL requests never set found)
      if (lane0([&] { return (int)ld_rlx(&P.states[j].found); }) != 0) {
        c.expanded = 0;
        c.solutions = 0;
        if (release_item(P, c)) publish_result(P, c, w, j);
        continue;
      }
      pi = uni64(it.pi);
      depth0 = uni((int)it.depth);
      limit = uni(ld_rlx(&P.states[j].limit));  // ordered by pop_item's acquire
    }
    load_job(P, c, w, j, limit);
    bool publish = false;
    for (;;) {
      if (!run_item<kCount>(P, c, w, lex, lds.dtab, cands, pi, depth0)) break;
      // last item of request j finished: next IDDFS limit or publish
      JobState *st = &P.states[j];
      int next = 0;
      if (!c.everShared) {
        if ((c.flags & kFlagIddfs) && !c.found && c.limit < job_max_depth(c)) {
          next = c.limit + 1;
          if (lane == 0) {
            st_rlx(&st->limit, next);
            st_rlx(&st->pending, 1);
          }
        }
      } else {
        next = lane0([&] {
          if ((c.flags & kFlagIddfs) && st->found == 0 && st->limit < job_max_depth(c)) {
            const int nl = st->limit + 1;
            st_rlx(&st->limit, nl);
            st_rlx(&st->pending, 1);
            return nl;
          }
          return 0;
        });
      }
      wave_sync();
      next = uni(next);
      if (next == 0) {
        publish = true;
        break;
      }
      load_job(P, c, w, j, next);
      pi = job_pi0(c);
      depth0 = 0;
      c.shared = false;
    }
    if (publish) publish_result(P, c, w, j);
  }
}

}  // namespace
}  // namespace td16
