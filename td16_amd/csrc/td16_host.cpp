// Host-side algorithms of libtd16: request validation and relabelling, the
// witness/prune tables, caller-side replay, the certificate codec and the
// synthetic instance generators.  No device code here.
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/td16_abi.h"
#include "td16_internal.h"

namespace td16 {

int set_msg(char *msg, int cap, int status, const char *fmt, ...) {
  if (msg && cap > 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, (size_t)cap, fmt, ap);
    va_end(ap);
  }
  return status;
}

void build_lex_table(uint16_t *out) {
  // block m (m = 3..17) starts at C(m,4) and lists the C(m,3) triples
  // i < j < k < m in lexicographic order, packed i | j<<5 | k<<10
  int q = 0;
  for (int m = 3; m <= kMaxN; ++m)
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j)
        for (int k = j + 1; k < m; ++k) out[q++] = (uint16_t)(i | (j << 5) | (k << 10));
}

namespace {

// float -> int with the device conversion's semantics (truncation,
// saturation, NaN -> 0), as used by sorting_search.cu:361-362.
int f2i_device(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT_MAX;
  if (x <= -2147483648.0f) return INT_MIN;
  return (int)x;
}

// emitter/src/native/backtracking/sorting_search.cu:336-346
bool satisfies(int N, int ie, float minRate, bool full, int ec, int depth) {
  if (depth <= 0) return false;
  if (full && ec != N) return false;
  const int gained = ec - ie;
  const float achieved = static_cast<float>(gained) / static_cast<float>(depth);
  return gained > 0 && achieved >= minRate;
}

// emitter/src/native/backtracking/sorting_search.cu:348-373
bool reachable(int N, int ie, float minRate, int maxDepth, bool full, int ec, int depth) {
  const int maxPossible = N - ec;
  if (full) {
    const int lb = (maxPossible + 1) / 2;
    return depth + lb <= maxDepth;
  }
  const int gained = ec - ie;
  const int optimistic = gained + 2 * (maxDepth - depth);
  const int required = f2i_device(std::floor(minRate * static_cast<float>(maxDepth)));
  if (optimistic < required) return false;
  const int lb = (maxPossible + 1) / 2;
  const int best = depth + lb;
  if (best > 0) {
    const float ratio = static_cast<float>(gained + maxPossible) / static_cast<float>(best);
    if (ratio < minRate) return false;
  }
  return true;
}

int odd_cycles(const uint8_t *omega, int N) {
  uint32_t seen = 0;
  int odd = 0;
  for (int s = 0; s < N; ++s) {
    if (seen & (1u << s)) continue;
    int len = 0;
    for (int x = s; !(seen & (1u << x)); x = omega[x]) {
      seen |= 1u << x;
      ++len;
    }
    odd += len & 1;
  }
  return odd;
}

bool is_perm(const uint8_t *a, int N) {
  uint32_t seen = 0;
  for (int i = 0; i < N; ++i) {
    if (a[i] >= N || (seen & (1u << a[i]))) return false;
    seen |= 1u << a[i];
  }
  return true;
}

void swap_blocks(uint8_t *seq, int i, int j, int k) {
  uint8_t tmp[32];
  std::memcpy(tmp, seq + i, (size_t)(k - i));
  std::memcpy(seq + i, tmp + (j - i), (size_t)(k - j));
  std::memcpy(seq + i + (k - j), tmp, (size_t)(j - i));
}

}  // namespace

int rate_tables(int N, int initialEven, float minRate, int maxDepth, int fullSorting,
                uint32_t sat[10], uint32_t reach[10]) {
  for (int d = 0; d <= kMaxDepth; ++d) {
    sat[d] = reach[d] = 0;
    for (int e = 0; e <= N; ++e) {
      if (satisfies(N, initialEven, minRate, fullSorting != 0, e, d)) sat[d] |= 1u << e;
      if (reachable(N, initialEven, minRate, maxDepth, fullSorting != 0, e, d))
        reach[d] |= 1u << e;
    }
  }
  return 0;
}

namespace {

// Callers without a memo of their own share one per thread.
thread_local RateMemo tMemo;

const RateMemo::Entry &rate_entry(RateMemo &memo, int N, int initialEven, float minRate,
                                  int maxMoves, int full, bool iddfs) {
  uint32_t rb;
  std::memcpy(&rb, &minRate, sizeof(rb));
  const uint64_t k1 = (uint64_t)rb | ((uint64_t)(uint8_t)N << 32) |
                      ((uint64_t)(uint8_t)maxMoves << 40) | ((uint64_t)(full != 0) << 48) |
                      ((uint64_t)iddfs << 49);
  const uint64_t k2 = (uint64_t)(uint32_t)initialEven;
  uint64_t h = (k1 ^ (k2 * 0x9E3779B97F4A7C15ULL)) * 0xBF58476D1CE4E5B9ULL;
  RateMemo::Entry &e = memo.e[(h >> 40) % RateMemo::kEntries];
  if (e.k1 == k1 && e.k2 == k2) return e;
  rate_tables(N, initialEven, minRate, maxMoves, full, e.sat, e.reach);
  // No slack: the odd-cycle count rises by at most two per move, so depth
  // d reaches initial_even .. initial_even + 2d; when the first depth with a
  // witness among those is the move bound itself, every move must be
  // perfect.  Long FIRST searches come from such requests.
  e.tight = false;
  if (!iddfs) {
    for (int d = 0; d <= maxMoves && d <= kMaxDepth; ++d) {
      uint32_t odds = 0;  // odd-cycle counts depth d can reach
      for (int i = 0; i <= d; ++i) {
        const int ev = initialEven + 2 * i;
        if (ev >= 0 && ev < 32) odds |= 1u << ev;
      }
      if (e.sat[d] & odds) {
        e.tight = d == maxMoves;
        break;
      }
    }
  }
  e.k1 = k1;
  e.k2 = k2;
  return e;
}

}  // namespace

// Validates a request (emitter/src/native/backtracking/sorting_search.cu:
// 820-825 and jni_sorter.cu:61-67) and turns it into the kernel's label form.
// Returns TD16_OK, or TD16_INVALID_ARGUMENT with a message.  *trivial is set
// when no launch is needed: n < 3 (no triple: sorting_search.cu:826-827) or
// an IDDFS request whose first limit exceeds max_moves or that is sorted.
int make_job(const td16_request *r, Job *job, int *trivial, RateMemo &memo, char *msg,
             int cap) {
  *trivial = 0;
  const int N = r->n;
  if (N < 0 || N > kMaxN)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "performGpuSortingSearch: n exceeds MAX_N");
  if (r->max_moves <= 0 || r->max_moves > kMaxDepth)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "performGpuSortingSearch: invalid maximumDepth");
  if (r->moveset != TD16_MOVES_REF && r->moveset != TD16_MOVES_ALL)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "unknown move set %d", r->moveset);
  if (r->mode != TD16_FIRST && r->mode != TD16_COUNT_ALL)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "unknown search mode %d", r->mode);
  // both are permutations of 0..N-1 iff every value is below N and the N
  // values cover all N bits
  uint32_t sb = 0, so = 0, over = 0;
  for (int p = 0; p < N; ++p) {
    const uint32_t b = r->beta[p], o = r->omega[p];
    over |= (uint32_t)(b >= (uint32_t)N) | (uint32_t)(o >= (uint32_t)N);
    sb |= 1u << (b & 31);
    so |= 1u << (o & 31);
  }
  const uint32_t full = (1u << N) - 1u;
  if (over || sb != full || so != full)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "beta and omega must be permutations of 0..n-1");
  std::memset(job, 0, sizeof(*job));
  if (N < 3) {
    *trivial = 1;
    return TD16_OK;
  }
  uint8_t lab[32];
  for (int p = 0; p < N; ++p) lab[r->beta[p]] = (uint8_t)p;
  // rho = omega o beta (sorting_search.cu:899-913): R[q] = label(rho(beta[q]))
  // and beta's function form maps beta[q] to beta[q + 1]
  for (int q = 0; q + 1 < N; ++q) job->R[q] = lab[r->omega[r->beta[q + 1]]];
  job->R[N - 1] = lab[r->omega[r->beta[0]]];
  // labels 0..N-2 at positions 0..N-2, unused nibbles 0xF
  const int used = 4 * (N - 1);  // 8..64 bits
  job->pi0 = used >= 64 ? 0xFEDCBA9876543210ULL
                        : (0xFEDCBA9876543210ULL & ((1ULL << used) - 1ULL)) | (~0ULL << used);
  job->N = (uint8_t)N;
  job->maxDepth = (uint8_t)r->max_moves;
  job->flags = (uint8_t)((r->moveset == TD16_MOVES_ALL ? kFlagAllMoves : 0) |
                         (r->mode == TD16_COUNT_ALL ? kFlagCountAll : 0) |
                         (r->iddfs ? kFlagIddfs : 0) | (r->lex_order ? kFlagLexOrder : 0));
  if (r->iddfs) {
    const int odd = odd_cycles(r->omega, N);
    int b = 0;
    for (int s = 0; s < N; ++s) b += r->omega[s] != s;
    if (odd == N) {
      *trivial = 2;  // already sorted: distance 0
      return TD16_OK;
    }
    int first = (N - odd) / 2;           // odd-cycle bound
    if ((b + 2) / 3 > first) first = (b + 2) / 3;  // breakpoint bound ceil(b/3)
    if (first < 1) first = 1;
    if (first > r->max_moves) {
      *trivial = 3;  // no witness within max_moves
      return TD16_OK;
    }
    job->firstLimit = (uint8_t)first;
    const RateMemo::Entry &e = rate_entry(memo, N, odd, 0.0f, r->max_moves, 1, true);
    std::memcpy(job->sat, e.sat, sizeof(job->sat));
    std::memcpy(job->reach, e.reach, sizeof(job->reach));
  } else {
    const RateMemo::Entry &e = rate_entry(memo, N, r->initial_even, r->min_rate, r->max_moves,
                                          r->full_sorting, false);
    std::memcpy(job->sat, e.sat, sizeof(job->sat));
    std::memcpy(job->reach, e.reach, sizeof(job->reach));
    if (r->mode == TD16_FIRST && e.tight) job->flags |= kFlagTight;
  }
  return TD16_OK;
}

int make_job_checked(const td16_request *r, Job *job, int *trivial, char *msg, int cap) {
  return make_job(r, job, trivial, tMemo, msg, cap);
}

void unpack_moves(const uint16_t *packed, int depth, int32_t (*moves)[3]) {
  for (int d = 0; d < depth && d < kMaxDepth; ++d) {
    moves[d][0] = packed[d] & 31;
    moves[d][1] = (packed[d] >> 5) & 31;
    moves[d][2] = (packed[d] >> 10) & 31;
  }
}

}  // namespace td16

using td16::set_msg;

extern "C" {

int td16_positions_to_symbols(const uint8_t *beta, int n, const int32_t *positions, int len,
                              int32_t *symbols, char *msg, int cap) {
  if (n < 0 || n > td16::kMaxN || !td16::is_perm(beta, n))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "beta must be a permutation of 0..n-1");
  uint8_t cur[td16::kMaxN];
  std::memcpy(cur, beta, (size_t)n);
  for (int m = 0; m < len; ++m) {
    const int i = positions[3 * m], j = positions[3 * m + 1], k = positions[3 * m + 2];
    if (!(0 <= i && i < j && j < k && k < n))
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "move %d is not an ordered position triple",
                     m);
    symbols[3 * m] = cur[i];
    symbols[3 * m + 1] = cur[j];
    symbols[3 * m + 2] = cur[k];
    td16::swap_blocks(cur, i, j, k);
  }
  return TD16_OK;
}

}  // extern "C"

namespace td16 {

// Replay of symbol triples (emitter/src/main/java/br/unb/cic/tdp/base/
// CertificateValidator.java:14-50).  Returns 0 when valid, else the reason:
// 1 length outside 1..maxMoves, 2 beta or omega not a permutation, 3 a move
// without three distinct symbols, 4 not an ordered triple in the current
// cyclic order, 5 the cyclic target not preserved, 6 omega not sorted at the
// end.  *bad receives the offending move for 3..5.
int replay(const uint8_t *beta0, const uint8_t *omega0, int n, const int32_t *sym, int len,
           int maxMoves, int *bad) {
  if (len < 1 || len > maxMoves) return 1;
  if (n < 1 || n > kMaxN || !is_perm(beta0, n) || !is_perm(omega0, n)) return 2;
  uint8_t beta[kMaxN], omega[kMaxN], rho[kMaxN], nxt[kMaxN];
  std::memcpy(beta, beta0, (size_t)n);
  std::memcpy(omega, omega0, (size_t)n);
  for (int p = 0; p < n; ++p) nxt[beta[p]] = beta[(p + 1) % n];
  for (int x = 0; x < n; ++x) rho[x] = omega[nxt[x]];
  for (int m = 0; m < len; ++m) {
    *bad = m;
    const int a = sym[3 * m], b = sym[3 * m + 1], c = sym[3 * m + 2];
    if (a == b || b == c || a == c || a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n)
      return 3;
    int pa = -1, pb = -1, pc = -1;
    for (int p = 0; p < n; ++p) {
      if (beta[p] == a) pa = p;
      if (beta[p] == b) pb = p;
      if (beta[p] == c) pc = p;
    }
    if (!(pa < pb && pb < pc)) return 4;
    // omega <- omega o (a b c)^-1
    const uint8_t oa = omega[a], ob = omega[b], oc = omega[c];
    omega[a] = oc;
    omega[b] = oa;
    omega[c] = ob;
    swap_blocks(beta, pa, pb, pc);
    for (int p = 0; p < n; ++p) nxt[beta[p]] = beta[(p + 1) % n];
    for (int x = 0; x < n; ++x)
      if (omega[nxt[x]] != rho[x]) return 5;
  }
  for (int x = 0; x < n; ++x)
    if (omega[x] != x) return 6;
  return 0;
}

}  // namespace td16

extern "C" {

int td16_validate(const uint8_t *beta0, const uint8_t *omega0, int n, const int32_t *sym, int len,
                  int maxMoves, char *msg, int cap) {
  int bad = 0;
  const int r = td16::replay(beta0, omega0, n, sym, len, maxMoves, &bad);
  const int32_t *mv = sym + 3 * bad;
  switch (r) {
    case 0: return TD16_OK;
    case 1:
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "Certificate has %d moves; the required range is 1..%d", len, maxMoves);
    case 2: return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "beta and omega must be permutations");
    case 3:
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "A move must contain three distinct symbols");
    case 4:
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "Move is not an ordered triple in the current cyclic order: (%d %d %d)",
                     mv[0], mv[1], mv[2]);
    case 5:
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "Certificate move does not preserve the cyclic target: (%d %d %d)", mv[0],
                     mv[1], mv[2]);
    default:
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "Certificate does not sort the algebraic permutation");
  }
}

// emitter/src/main/java/br/unb/cic/tdp/base/CertificateCodec.java:22-118
static int write5(uint8_t *buf, int bit, int v) {
  for (int q = 0; q < 5; ++q) {
    const int vb = (v >> (4 - q)) & 1;
    const int ob = bit + q;
    buf[ob >> 3] |= (uint8_t)(vb << (7 - (ob & 7)));
  }
  return bit + 5;
}

static int read5(const uint8_t *buf, int bit) {
  int v = 0;
  for (int q = 0; q < 5; ++q) {
    const int ib = bit + q;
    v = (v << 1) | ((buf[ib >> 3] >> (7 - (ib & 7))) & 1);
  }
  return v;
}

static bool payload_ok(const uint8_t *buf, int nbytes, long bits) {
  if (buf == nullptr || nbytes != (int)((bits + 7) / 8)) return false;
  const int pad = (int)(8L * nbytes - bits);
  return pad == 0 || (buf[nbytes - 1] & ((1 << pad) - 1)) == 0;
}

int td16_codec_key(const uint8_t *omega, int n, uint8_t *out, int outCap, int32_t *outLen,
                   char *msg, int cap) {
  if (n <= 0 || n > 32)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "The compact certificate format supports domains of size 1..32");
  const int nb = (5 * n + 7) / 8;
  if (nb > outCap) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "output buffer too small");
  std::memset(out, 0, (size_t)nb);
  int bit = 0;
  for (int s = 0; s < n; ++s) {
    if (omega[s] >= 32)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "Certificate symbols must lie in 0..31");
    bit = write5(out, bit, omega[s]);
  }
  *outLen = nb;
  return TD16_OK;
}

int td16_codec_value(const int32_t *sym, int len, uint8_t *out, int outCap, int32_t *outLen,
                     char *msg, int cap) {
  if (len <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "A certificate must contain at least one move");
  const int nb = (15 * len + 7) / 8;
  if (nb > outCap) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "output buffer too small");
  std::memset(out, 0, (size_t)nb);
  int bit = 0;
  for (int q = 0; q < 3 * len; ++q) {
    if (sym[q] < 0 || sym[q] >= 32)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "Certificate symbols must lie in 0..31");
    bit = write5(out, bit, sym[q]);
  }
  *outLen = nb;
  return TD16_OK;
}

int td16_codec_decode_key(const uint8_t *in, int nbytes, int n, uint8_t *omega, char *msg,
                          int cap) {
  if (n <= 0 || n > 32)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "The compact certificate format supports domains of size 1..32");
  if (!payload_ok(in, nbytes, 5L * n))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "Malformed certificate key length or padding");
  for (int s = 0; s < n; ++s) omega[s] = (uint8_t)read5(in, 5 * s);
  return TD16_OK;
}

int td16_codec_decode_value(const uint8_t *in, int nbytes, int32_t *sym, int symCap,
                            int32_t *outLen, char *msg, int cap) {
  if (in == nullptr || nbytes <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "Truncated certificate value");
  const int moves = (int)((8L * nbytes) / 15);
  if (moves == 0) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "Truncated certificate value");
  if (!payload_ok(in, nbytes, 15L * moves))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                   "Malformed certificate value length or padding");
  if (moves > symCap) return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "output buffer too small");
  for (int q = 0; q < 3 * moves; ++q) sym[q] = read5(in, 5 * q);
  *outLen = moves;
  return TD16_OK;
}

}  // extern "C"

namespace td16 {
namespace {

// Runs body(lo, hi) over [0, count) on `threads` host threads (0: all).
template <class Body>
void parallel_ranges(int count, int threads, Body body) {
  int t = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
  if (t < 1) t = 1;
  if (t > 256) t = 256;
  if (t > count) t = count > 0 ? count : 1;
  if (t == 1) {
    body(0, count);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve((size_t)t);
  for (int q = 0; q < t; ++q) {
    const int lo = (int)((int64_t)count * q / t), hi = (int)((int64_t)count * (q + 1) / t);
    pool.emplace_back([=] { body(lo, hi); });
  }
  for (auto &th : pool) th.join();
}

void encode5(const uint8_t *v, int count, uint8_t *out) {
  std::memset(out, 0, (size_t)((5 * count + 7) / 8));
  int bit = 0;
  for (int q = 0; q < count; ++q) bit = write5(out, bit, v[q]);
}

}  // namespace
}  // namespace td16

extern "C" {

int td16_certify_batch(const td16_request *reqs, const td16_result *results, int count,
                       int threads, uint8_t *keys, uint8_t *values, int32_t *value_len,
                       int32_t *status, char *msg, int cap) {
  if (count < 0 || (count > 0 && (reqs == nullptr || results == nullptr || keys == nullptr ||
                                   values == nullptr || value_len == nullptr ||
                                   status == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "null buffer or negative count");
  td16::parallel_ranges(count, threads, [&](int lo, int hi) {
    for (int q = lo; q < hi; ++q) {
      const td16_request &r = reqs[q];
      const td16_result &res = results[q];
      uint8_t *key = keys + (size_t)q * TD16_CERT_KEY_BYTES;
      uint8_t *val = values + (size_t)q * TD16_CERT_VALUE_BYTES;
      std::memset(key, 0, TD16_CERT_KEY_BYTES);
      value_len[q] = 0;
      if (res.status != TD16_OK || !res.found || res.depth < 1) {
        status[q] = 1;  // nothing to certify
        continue;
      }
      const int n = r.n;
      if (n < 1 || n > td16::kMaxN || !td16::is_perm(r.beta, n) || res.depth > TD16_MAX_DEPTH) {
        status[q] = -2;
        continue;
      }
      // positions -> symbols on a copy of beta (GPUSortingSearch.java:229-239)
      int32_t sym[3 * TD16_MAX_DEPTH];
      uint8_t cur[td16::kMaxN];
      std::memcpy(cur, r.beta, (size_t)n);
      bool ok = true;
      for (int m = 0; m < res.depth && ok; ++m) {
        const int i = res.moves[m][0], j = res.moves[m][1], k = res.moves[m][2];
        if (!(0 <= i && i < j && j < k && k < n)) {
          ok = false;
          break;
        }
        sym[3 * m] = cur[i];
        sym[3 * m + 1] = cur[j];
        sym[3 * m + 2] = cur[k];
        td16::swap_blocks(cur, i, j, k);
      }
      int bad = 0;
      const int rr = ok ? td16::replay(r.beta, r.omega, n, sym, res.depth, r.max_moves, &bad) : 4;
      if (rr != 0) {
        status[q] = -rr;
        continue;
      }
      td16::encode5(r.omega, n, key);
      uint8_t sym8[3 * TD16_MAX_DEPTH];
      for (int s = 0; s < 3 * res.depth; ++s) sym8[s] = (uint8_t)sym[s];
      td16::encode5(sym8, 3 * res.depth, val);
      value_len[q] = (15 * res.depth + 7) / 8;
      status[q] = 0;
    }
  });
  return TD16_OK;
}

int td16_check_batch(const uint8_t *betas, const uint8_t *omegas, const int32_t *ns,
                     const int32_t *max_moves, int count, const uint8_t *values,
                     const int32_t *value_len, int threads, int32_t *status, char *msg, int cap) {
  if (count < 0 || (count > 0 && (betas == nullptr || omegas == nullptr || ns == nullptr ||
                                   max_moves == nullptr || values == nullptr ||
                                   value_len == nullptr || status == nullptr)))
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "null buffer or negative count");
  td16::parallel_ranges(count, threads, [&](int lo, int hi) {
    for (int q = lo; q < hi; ++q) {
      const int nb = value_len[q];
      if (nb == 0) {
        status[q] = 1;  // no stored certificate: the obligation stays open
        continue;
      }
      const uint8_t *in = values + (size_t)q * TD16_CERT_VALUE_BYTES;
      const int moves = nb > 0 ? (8 * nb) / 15 : 0;
      if (nb < 0 || nb > TD16_CERT_VALUE_BYTES || moves == 0 || moves > TD16_MAX_DEPTH ||
          !payload_ok(in, nb, 15L * moves)) {
        status[q] = -7;  // malformed value (CertificateCodec.decodeValue rejects it)
        continue;
      }
      int32_t sym[3 * TD16_MAX_DEPTH];
      for (int s = 0; s < 3 * moves; ++s) sym[s] = read5(in, 5 * s);
      int bad = 0;
      const int n = ns[q];
      const int rr = (n < 1 || n > td16::kMaxN)
                         ? 2
                         : td16::replay(betas + (size_t)q * TD16_MAX_N,
                                        omegas + (size_t)q * TD16_MAX_N, n, sym, moves,
                                        max_moves[q], &bad);
      status[q] = -rr;
    }
  });
  return TD16_OK;
}

// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014)
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in [0, m): reject the 2^64 mod m smallest draws
static uint64_t uniform(uint64_t *s, uint64_t m) {
  const uint64_t threshold = (0 - m) % m;
  for (;;) {
    const uint64_t r = splitmix64(s);
    if (r >= threshold) return r % m;
  }
}

static uint64_t pack(const uint8_t *p, int n) {
  uint64_t v = ~0ULL;
  for (int q = 0; q < n; ++q) {
    v &= ~(0xFULL << (4 * q));
    v |= (uint64_t)p[q] << (4 * q);
  }
  return v;
}

static bool identity(const uint8_t *p, int n) {
  for (int q = 0; q < n; ++q)
    if (p[q] != q) return false;
  return true;
}

int td16_gen_uniform(uint64_t seed, int n, int count, uint64_t *perms) {
  if (n < 2 || n > 16 || count < 0) return TD16_INVALID_ARGUMENT;
  uint64_t st = seed;
  for (int c = 0; c < count; ++c) {
    uint8_t p[16];
    do {
      for (int q = 0; q < n; ++q) p[q] = (uint8_t)q;
      for (int i = n - 1; i >= 1; --i) {
        const int j = (int)uniform(&st, (uint64_t)i + 1);
        const uint8_t t = p[i];
        p[i] = p[j];
        p[j] = t;
      }
    } while (identity(p, n));
    perms[c] = pack(p, n);
  }
  return TD16_OK;
}

int td16_gen_products(uint64_t seed, int nmin, int nmax, int kmin, int kmax, int count,
                      uint8_t *ns, uint8_t *ks, uint64_t *perms) {
  if (nmin < 2 || nmax > 16 || nmin > nmax || kmin < 1 || kmin > kmax || count < 0)
    return TD16_INVALID_ARGUMENT;
  uint64_t st = seed;
  for (int c = 0; c < count; ++c) {
    uint8_t p[17];
    int n, k;
    do {
      n = nmin + (int)uniform(&st, (uint64_t)(nmax - nmin + 1));
      k = kmin + (int)uniform(&st, (uint64_t)(kmax - kmin + 1));
      for (int q = 0; q < n; ++q) p[q] = (uint8_t)q;
      const int N = n + 1;
      const uint64_t T = (uint64_t)N * (N - 1) * (N - 2) / 6;
      for (int mv = 0; mv < k; ++mv) {
        // lexicographic unranking over 0 <= i < j < kk <= n
        uint64_t t = uniform(&st, T);
        int i = 0, j = 1, kk = 2;
        bool done = false;
        for (int a = 0; a < N - 2 && !done; ++a)
          for (int b = a + 1; b < N - 1 && !done; ++b) {
            const uint64_t cnt = (uint64_t)(N - 1 - b);
            if (t < cnt) {
              i = a;
              j = b;
              kk = b + 1 + (int)t;
              done = true;
            } else {
              t -= cnt;
            }
          }
        td16::swap_blocks(p, i, j, kk);
      }
    } while (identity(p, n));
    ns[c] = (uint8_t)n;
    ks[c] = (uint8_t)k;
    perms[c] = pack(p, n);
  }
  return TD16_OK;
}

int td16_gen_exponential(uint64_t seed, int count, double *gaps) {
  uint64_t st = seed;
  for (int c = 0; c < count; ++c) {
    const double u = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
    gaps[c] = -std::log1p(-u);
  }
  return TD16_OK;
}

int td16_linear_request(uint64_t perm, int n, int k, int moveset, int mode, int iddfs,
                        td16_request *req) {
  if (n < 2 || n > 16 || k < 1 || k > td16::kMaxDepth) return TD16_INVALID_ARGUMENT;
  std::memset(req, 0, sizeof(*req));
  const int N = n + 1;
  for (int q = 0; q < n; ++q) req->beta[q] = (uint8_t)((perm >> (4 * q)) & 0xF);
  req->beta[n] = (uint8_t)n;
  if (!td16::is_perm(req->beta, N)) return TD16_INVALID_ARGUMENT;
  for (int p = 0; p < N; ++p) req->omega[req->beta[p]] = (uint8_t)((req->beta[(p + N - 1) % N] + 1) % N);
  const int odd = td16::odd_cycles(req->omega, N);
  const int norm = (N - odd) / 2;
  req->n = N;
  req->initial_even = odd;
  // emitter/src/main/java/br/unb/cic/tdp/UnorientedCycles.java:34
  req->min_rate = norm > 0 ? 2.0f / ((float)k / (float)norm) : 0.0f;
  req->max_moves = k;
  req->full_sorting = 1;
  req->moveset = moveset;
  req->mode = mode;
  req->iddfs = iddfs;
  return TD16_OK;
}

const char *td16_version(void) { return "td16-mi355x 0.1 (gfx950)"; }

}  // extern "C"
