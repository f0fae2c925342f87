// Representatives of the emitter's cycle configurations: for an annotated
// partition (omega's cycle lengths, each cycle oriented or not) one
// cyclic-target pair per realizable extended-toric class, the set that
// emitter/src/main/java/br/unb/cic/tdp/DirectTwistedBraceletGenerator.java
// emits.  Host code, no device work.
//
// A configuration is a word of n beads, one per position 0..n-1 of
// beta = (0 1 .. n-1).  A bead names its omega cycle, and a bead of an
// oriented cycle also its rank along that cycle.  The alphabet lists the
// cycles by length (longest first, unoriented before oriented of one length,
// DirectTwistedBraceletGenerator.java:257-270): an unoriented cycle is one
// symbol used `length` times, an oriented cycle `length` symbols (its ranks)
// used once each.  Classes are the orbits of words under rotation and
// reflection of the bead circle combined with the auxiliary relabelling:
// along a traversal, each cycle met is renamed to the least unused cycle of
// its annotated type and an oriented cycle's ranks are shifted so that the
// first one met is 0 (negated along a reflected traversal).  The
// representative of an orbit is its least image.  Decoding a word
// (DirectTwistedBraceletGenerator.java:1475-1505): an unoriented cycle met at
// positions p1 < .. < pk maps p(i) -> p(i-1) and p1 -> pk; an oriented cycle
// maps the position of rank r to that of rank r+1.  An oriented cycle whose
// ranks read r, r-1, r-2, .. along the word has the unoriented pattern and is
// not an oriented configuration (DirectTwistedBraceletGenerator.java:1290-1303).
// A representative is realizable when rho = omega o beta, rho(p-1) =
// omega(p), is one n-cycle.
//
// Method (orderly generation, not the reference's bracelet recursion):
// words grow left to right in increasing symbol order, and a prefix is cut as
// soon as
//  - it is not its own standard form (a cycle first met is not the least
//    unmet cycle of its type, or an oriented cycle's first rank is not 0),
//  - an oriented cycle is complete in the unoriented pattern,
//  - (pairs only) the partial target closes a cycle shorter than n, or
//  - a rotation or reflection whose image is determined on a prefix is
//    smaller there than the word.
// Images still tied with the word are settled once it is complete.  Rows come
// out in increasing word order; a parallel run splits the tree at a shallow
// depth and concatenates the subtrees in order, so the output does not depend
// on the thread count.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/td16_abi.h"
#include "td16_internal.h"

namespace td16 {
namespace {

constexpr int kMaxBeads = 32;
constexpr int kMaxParts = 16;

struct Alphabet {
  int n = 0, m = 0, S = 0;
  int size[kMaxParts];
  bool oriented[kMaxParts];
  int groupFirst[kMaxParts], groupEnd[kMaxParts];  // cycles of c's annotated type
  int firstSym[kMaxParts];
  int symCycle[kMaxBeads], symRank[kMaxBeads], mult[kMaxBeads];
};

int mod(int a, int b) {
  const int r = a % b;
  return r < 0 ? r + b : r;
}

// Relabelling state of one traversal (rotation s, direction dir).
struct Standardizer {
  int8_t target[kMaxParts];
  int8_t origin[kMaxParts];
  uint32_t used;

  void reset(int m) {
    for (int c = 0; c < m; ++c) target[c] = -1;
    used = 0;
  }

  int next(const Alphabet &A, int sym, int dir) {
    const int c = A.symCycle[sym];
    if (target[c] < 0) {
      int t = A.groupFirst[c];
      while ((used >> t) & 1u) ++t;
      target[c] = (int8_t)t;
      used |= 1u << t;
      origin[c] = (int8_t)A.symRank[sym];
    }
    const int t = target[c];
    if (!A.oriented[c]) return A.firstSym[t];
    return A.firstSym[t] + mod(dir * (A.symRank[sym] - origin[c]), A.size[c]);
  }
};

struct Undo {
  int placed, firstPos, lastPos, lastRank;
  bool reverse;
  int nArcs;
  int h[2], t[2], u[2], v[2];
  uint32_t fwd, bwd;
};

class Generator {
 public:
  Generator(const Alphabet &A, bool pairs) : A_(A), pairs_(pairs) {
    for (int s = 0; s < A.S; ++s) remaining_[s] = A.mult[s];
    for (int c = 0; c < A.m; ++c) {
      placed_[c] = 0;
      firstPos_[c] = lastPos_[c] = -1;
      lastRank_[c] = 0;
      reverse_[c] = true;
    }
    for (int s = 0; s < A.S; ++s) posOfSym_[s] = -1;
    for (int p = 0; p < A.n; ++p) other_[p] = p;
    arcs_ = 0;
    met_ = 0;
  }

  // Places symbol s at position t; false (with everything undone) when the
  // prefix is cut.  fwd / bwd: traversals still tied with the word.
  bool place(int t, int s, uint32_t &fwd, uint32_t &bwd, Undo &U) {
    if (remaining_[s] == 0) return false;
    const int c = A_.symCycle[s];
    const int r = A_.symRank[s];
    if (placed_[c] == 0) {
      if (A_.oriented[c] && r != 0) return false;
      int least = A_.groupFirst[c];
      while ((met_ >> least) & 1u) ++least;
      if (least != c) return false;
    }
    bool rev = reverse_[c];
    if (A_.oriented[c] && placed_[c] > 0) rev = rev && r == mod(lastRank_[c] - 1, A_.size[c]);
    if (A_.oriented[c] && placed_[c] + 1 == A_.size[c] && rev) return false;

    U.placed = placed_[c];
    U.firstPos = firstPos_[c];
    U.lastPos = lastPos_[c];
    U.lastRank = lastRank_[c];
    U.reverse = reverse_[c];
    U.nArcs = 0;
    U.fwd = fwd;
    U.bwd = bwd;
    w_[t] = s;
    --remaining_[s];
    if (placed_[c] == 0) {
      met_ |= 1u << c;
      firstPos_[c] = t;
    }
    ++placed_[c];
    reverse_[c] = rev;
    lastRank_[c] = r;
    const int prevPos = lastPos_[c];
    lastPos_[c] = t;
    posOfSym_[s] = t;

    bool ok = true;
    if (pairs_) {
      const int n = A_.n;
      if (!A_.oriented[c]) {
        if (placed_[c] >= 2) ok = add_arc(t - 1, prevPos, U);
        if (ok && placed_[c] == A_.size[c]) ok = add_arc(mod(firstPos_[c] - 1, n), t, U);
      } else {
        const int base = A_.firstSym[c], sz = A_.size[c];
        const int qPred = posOfSym_[base + mod(r - 1, sz)];
        const int qSucc = posOfSym_[base + mod(r + 1, sz)];
        if (qPred >= 0) ok = add_arc(mod(qPred - 1, n), t, U);
        if (ok && qSucc >= 0) ok = add_arc(mod(t - 1, n), qSucc, U);
      }
    }
    if (ok) ok = prefix_images(t, fwd, bwd);
    if (!ok) {
      unplace(t, s, U);
      fwd = U.fwd;
      bwd = U.bwd;
    }
    return ok;
  }

  void unplace(int t, int s, const Undo &U) {
    for (int q = U.nArcs - 1; q >= 0; --q) {
      other_[U.h[q]] = U.u[q];
      other_[U.t[q]] = U.v[q];
      --arcs_;
    }
    const int c = A_.symCycle[s];
    posOfSym_[s] = -1;
    placed_[c] = U.placed;
    firstPos_[c] = U.firstPos;
    lastPos_[c] = U.lastPos;
    lastRank_[c] = U.lastRank;
    reverse_[c] = U.reverse;
    if (U.placed == 0) met_ &= ~(1u << c);
    ++remaining_[s];
    (void)t;
  }

  // Depth-first generation from position t; rows go to `out` (n bytes each)
  // unless it is null.
  void run(int t, uint32_t fwd, uint32_t bwd, std::vector<uint8_t> *out, int64_t *count) {
    if (t == A_.n) {
      if (!complete_images(fwd, bwd)) return;
      ++*count;
      if (out && (stored_ == nullptr || stored_->fetch_add(1) < storeCap_)) emit(*out);
      return;
    }
    for (int s = 0; s < A_.S; ++s) {
      uint32_t f = fwd, b = bwd;
      Undo U;
      if (!place(t, s, f, b, U)) continue;
      run(t + 1, f, b, out, count);
      unplace(t, s, U);
    }
  }

  // Rows beyond `cap` over all generators sharing `stored` are counted but
  // not kept, so a too-small output cannot make the host memory grow.
  void limit_storage(std::atomic<int64_t> *stored, int64_t cap) {
    stored_ = stored;
    storeCap_ = cap;
  }

  // Prefixes of length `depth` that survive every cut, in increasing order.
  void frontier(int t, int depth, uint32_t fwd, uint32_t bwd, std::vector<int> &prefixes,
                int *count) {
    if (t == depth) {
      prefixes.insert(prefixes.end(), w_, w_ + depth);
      ++*count;
      return;
    }
    for (int s = 0; s < A_.S; ++s) {
      uint32_t f = fwd, b = bwd;
      Undo U;
      if (!place(t, s, f, b, U)) continue;
      frontier(t + 1, depth, f, b, prefixes, count);
      unplace(t, s, U);
    }
  }

 private:
  bool add_arc(int u, int v, Undo &U) {
    const int h = other_[u];  // head of the path u ends
    if (h == v && arcs_ + 1 < A_.n) return false;
    const int tl = other_[v];  // tail of the path v starts
    const int q = U.nArcs++;
    U.h[q] = h;
    U.t[q] = tl;
    U.u[q] = u;
    U.v[q] = v;
    other_[h] = tl;
    other_[tl] = h;
    ++arcs_;
    return true;
  }

  // Compares the images determined by the prefix w[0..t] with the word.
  bool prefix_images(int t, uint32_t &fwd, uint32_t &bwd) {
    Standardizer st;
    // rotations starting at s = 1..t, read forwards: image[k] = std(w[s+k])
    uint32_t cand = fwd;
    if (t >= 1) cand |= 1u << t;
    for (uint32_t bits = cand; bits; bits &= bits - 1) {
      const int s = __builtin_ctz(bits);
      st.reset(A_.m);
      int x = 0;
      for (int k = s; k <= t; ++k) x = st.next(A_, w_[k], +1);
      const int y = w_[t - s];
      if (x < y) return false;
      if (x > y) cand &= ~(1u << s);
    }
    fwd = cand;
    // the reflection through position t, read backwards: image[k] = std(w[t-k])
    st.reset(A_.m);
    int k = 0;
    for (; k <= t; ++k) {
      const int x = st.next(A_, w_[t - k], -1);
      if (x < w_[k]) return false;
      if (x > w_[k]) break;
    }
    if (k > t) bwd |= 1u << t;
    return true;
  }

  bool complete_images(uint32_t fwd, uint32_t bwd) {
    const int n = A_.n;
    Standardizer st;
    for (int dir = +1; dir >= -1; dir -= 2) {
      for (uint32_t bits = dir > 0 ? fwd : bwd; bits; bits &= bits - 1) {
        const int s = __builtin_ctz(bits);
        st.reset(A_.m);
        for (int k = 0; k < n; ++k) {
          const int x = st.next(A_, w_[mod(s + dir * k, n)], dir);
          if (x < w_[k]) return false;
          if (x > w_[k]) break;
        }
      }
    }
    return true;
  }

  void emit(std::vector<uint8_t> &out) {
    const int n = A_.n;
    const size_t at = out.size();
    out.resize(at + (size_t)n);
    uint8_t *row = out.data() + at;
    if (!pairs_) {
      for (int p = 0; p < n; ++p) row[p] = (uint8_t)w_[p];
      return;
    }
    decode_word(A_, w_, row);
  }

 public:
  // omega of a word over positions 0..n-1 (DirectTwistedBraceletGenerator.java:
  // 1475-1505): an unoriented cycle met at p1 < .. < pk maps p(i) to p(i-1)
  // and p1 to pk; an oriented cycle maps its rank r's position to rank r+1's.
  static void decode_word(const Alphabet &A, const int *w, uint8_t *row) {
    const int n = A.n;
    int prev[kMaxParts], first[kMaxParts];
    int posByRank[kMaxBeads];
    for (int c = 0; c < A.m; ++c) prev[c] = first[c] = -1;
    for (int p = 0; p < n; ++p) {
      const int s = w[p], c = A.symCycle[s];
      if (A.oriented[c]) {
        posByRank[s] = p;
        continue;
      }
      if (prev[c] >= 0) row[p] = (uint8_t)prev[c];
      else first[c] = p;
      prev[c] = p;
    }
    for (int c = 0; c < A.m; ++c) {
      if (A.oriented[c]) {
        const int base = A.firstSym[c], sz = A.size[c];
        for (int r = 0; r < sz; ++r) row[posByRank[base + r]] = (uint8_t)posByRank[base + (r + 1) % sz];
      } else {
        row[first[c]] = (uint8_t)prev[c];
      }
    }
  }

 private:

  const Alphabet &A_;
  const bool pairs_;
  int w_[kMaxBeads];
  int remaining_[kMaxBeads];
  int placed_[kMaxParts], firstPos_[kMaxParts], lastPos_[kMaxParts], lastRank_[kMaxParts];
  bool reverse_[kMaxParts];
  int posOfSym_[kMaxBeads];
  int other_[kMaxBeads];
  int arcs_;
  uint32_t met_;
  std::atomic<int64_t> *stored_ = nullptr;
  int64_t storeCap_ = 0;
};

int make_alphabet(const int32_t *parts, const uint8_t *oriented, int m, Alphabet *A, char *msg,
                  int cap) {
  // DirectTwistedBraceletGenerator.java:234-255 (the reference's messages)
  if (m <= 0 || parts == nullptr || oriented == nullptr)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "The annotated partition must be nonempty");
  if (m > kMaxParts)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "at most %d cycles are supported", kMaxParts);
  int n = 0;
  for (int q = 0; q < m; ++q) {
    if (parts[q] < 2)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "Every cycle must have length at least two");
    if (oriented[q] && parts[q] < 3)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "Only cycles of length at least three can be oriented");
    n += parts[q];
    if (n > kMaxBeads)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "at most %d beads are supported",
                     kMaxBeads);
  }
  std::vector<std::pair<int, int>> spec((size_t)m);  // (size, oriented)
  for (int q = 0; q < m; ++q) spec[(size_t)q] = {parts[q], oriented[q] ? 1 : 0};
  std::stable_sort(spec.begin(), spec.end(), [](const std::pair<int, int> &a,
                                                const std::pair<int, int> &b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  A->n = n;
  A->m = m;
  int S = 0;
  for (int c = 0; c < m; ++c) {
    A->size[c] = spec[(size_t)c].first;
    A->oriented[c] = spec[(size_t)c].second != 0;
    A->firstSym[c] = S;
    const int k = A->oriented[c] ? A->size[c] : 1;
    for (int r = 0; r < k; ++r) {
      A->symCycle[S] = c;
      A->symRank[S] = r;
      A->mult[S] = A->oriented[c] ? 1 : A->size[c];
      ++S;
    }
  }
  A->S = S;
  for (int c = 0; c < m;) {
    int e = c + 1;
    while (e < m && spec[(size_t)e] == spec[(size_t)c]) ++e;
    for (int q = c; q < e; ++q) {
      A->groupFirst[q] = c;
      A->groupEnd[q] = e;
    }
    c = e;
  }
  return TD16_OK;
}

}  // namespace
}  // namespace td16

extern "C" int td16_bracelets(const int32_t *parts, const uint8_t *oriented, int m, int mode,
                              int threads, uint8_t *out, int64_t cap_rows, int64_t *total,
                              char *msg, int cap) {
  using namespace td16;
  if (total == nullptr || (mode != TD16_BRACELET_PAIRS && mode != TD16_BRACELET_WORDS) ||
      cap_rows < 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid bracelet arguments");
  *total = 0;
  Alphabet A;
  const int st = make_alphabet(parts, oriented, m, &A, msg, cap);
  if (st != TD16_OK) return st;
  const bool pairs = mode == TD16_BRACELET_PAIRS;
  const bool keep = out != nullptr;

  int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 256));
  // split the tree where it has enough subtrees to share
  std::vector<int> prefixes;
  int nPrefix = 0, depth = 0;
  if (nt > 1) {
    for (depth = 1; depth < A.n; ++depth) {
      prefixes.clear();
      nPrefix = 0;
      Generator g(A, pairs);
      g.frontier(0, depth, 0, 0, prefixes, &nPrefix);
      if (nPrefix >= 16 * nt || nPrefix == 0) break;
    }
  }
  std::vector<std::vector<uint8_t>> rows;
  std::vector<int64_t> counts;
  std::atomic<int64_t> stored{0};
  if (nt == 1 || depth >= A.n || nPrefix == 0) {
    rows.resize(1);
    counts.assign(1, 0);
    Generator g(A, pairs);
    g.limit_storage(&stored, cap_rows);
    g.run(0, 0, 0, keep ? &rows[0] : nullptr, &counts[0]);
  } else {
    rows.resize((size_t)nPrefix);
    counts.assign((size_t)nPrefix, 0);
    std::atomic<int> next{0};
    auto work = [&] {
      for (;;) {
        const int q = next.fetch_add(1);
        if (q >= nPrefix) return;
        Generator g(A, pairs);
        g.limit_storage(&stored, cap_rows);
        uint32_t f = 0, b = 0;
        Undo U[kMaxBeads];
        for (int t = 0; t < depth; ++t) g.place(t, prefixes[(size_t)q * depth + t], f, b, U[t]);
        g.run(depth, f, b, keep ? &rows[(size_t)q] : nullptr, &counts[(size_t)q]);
      }
    };
    std::vector<std::thread> pool;
    for (int q = 0; q < nt; ++q) pool.emplace_back(work);
    for (auto &th : pool) th.join();
  }
  int64_t n = 0;
  for (int64_t c : counts) n += c;
  *total = n;
  if (keep) {
    int64_t written = 0;
    for (const auto &r : rows) {
      if (written == cap_rows) break;
      const int64_t k = std::min<int64_t>((int64_t)r.size() / A.n, cap_rows - written);
      if (k > 0) std::memcpy(out + written * A.n, r.data(), (size_t)(k * A.n));
      written += k;
    }
    if (n > cap_rows)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "output holds %lld rows; %lld are needed", (long long)cap_rows,
                     (long long)n);
  }
  return TD16_OK;
}

// The class of any fixed-point-free pair (TwistedBraceletKey.of, emitter/src/
// main/java/br/unb/cic/tdp/base/TwistedBraceletKey.java:22-29): read omega
// along beta's cyclic order, type its cycles (unoriented when omega maps each
// position of the cycle to the cycle's previous position along beta), write
// the word in the generator's alphabet and take the least standardized image
// over all rotations and reflections.  The representative is that word
// decoded, the pair the generator emits for the class when it is realizable.
extern "C" int td16_bracelet_canonical(const uint8_t *beta, const uint8_t *omega, int n,
                                       int32_t *parts, uint8_t *oriented, int32_t *m,
                                       uint8_t *word, uint8_t *rep_omega, char *msg, int cap) {
  using namespace td16;
  if (beta == nullptr || omega == nullptr || m == nullptr || parts == nullptr ||
      oriented == nullptr || n < 2 || n > kMaxBeads)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid pair arguments (2 <= n <= %d)",
                   kMaxBeads);
  int pos[kMaxBeads];
  bool seenB[kMaxBeads] = {}, seenO[kMaxBeads] = {};
  for (int p = 0; p < n; ++p) {
    if (beta[p] >= n || seenB[beta[p]] || omega[p] >= n || seenO[omega[p]])
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "beta and omega must be permutations");
    seenB[beta[p]] = seenO[omega[p]] = true;
    pos[beta[p]] = p;
  }
  int om[kMaxBeads];  // omega on beta's positions
  for (int p = 0; p < n; ++p) om[p] = pos[omega[beta[p]]];
  int cycleOf[kMaxBeads], base[kMaxParts], nC = 0;
  int32_t sz[kMaxParts];
  uint8_t ori[kMaxParts];
  for (int p = 0; p < n; ++p) cycleOf[p] = -1;
  for (int p = 0; p < n; ++p) {
    if (cycleOf[p] >= 0) continue;
    if (om[p] == p)
      return set_msg(msg, cap, TD16_INVALID_ARGUMENT,
                     "omega must be fixed-point free (every cycle of length at least two)");
    int len = 0;
    for (int x = p; cycleOf[x] < 0; x = om[x]) {
      cycleOf[x] = nC;
      ++len;
    }
    // the cycle's positions in increasing order: unoriented when each maps
    // to the one before it, the first to the last
    int prevPos = -1, firstPos = -1, lastPos = -1;
    bool unori = true;
    for (int x = 0; x < n; ++x) {
      if (cycleOf[x] != nC) continue;
      if (firstPos < 0) firstPos = x;
      else if (om[x] != prevPos) unori = false;
      prevPos = lastPos = x;
    }
    if (om[firstPos] != lastPos) unori = false;
    base[nC] = p;
    sz[nC] = len;
    ori[nC] = unori ? 0 : 1;
    ++nC;
  }
  Alphabet A;
  const int st = make_alphabet(sz, ori, nC, &A, msg, cap);
  if (st != TD16_OK) return st;
  // each cycle takes a free slot of its annotated type; ranks count omega
  // steps from the cycle's first position
  int slotOf[kMaxParts];
  bool taken[kMaxParts] = {};
  for (int c = 0; c < nC; ++c) {
    int s = 0;
    while (taken[s] || A.size[s] != sz[c] || A.oriented[s] != (ori[c] != 0)) ++s;
    taken[s] = true;
    slotOf[c] = s;
  }
  int w0[kMaxBeads];
  for (int c = 0; c < nC; ++c) {
    int x = base[c];
    for (int r = 0; r < sz[c]; ++r, x = om[x])
      w0[x] = A.firstSym[slotOf[c]] + (ori[c] ? r : 0);
  }
  int best[kMaxBeads], cand[kMaxBeads];
  bool have = false;
  for (int s = 0; s < n; ++s) {
    for (int dir = +1; dir >= -1; dir -= 2) {
      Standardizer sd;
      sd.reset(A.m);
      bool smaller = !have, larger = false;
      for (int k = 0; k < n && !larger; ++k) {
        cand[k] = sd.next(A, w0[mod(s + dir * k, n)], dir);
        if (!smaller) {
          if (cand[k] < best[k]) smaller = true;
          else if (cand[k] > best[k]) larger = true;
        }
      }
      if (smaller) {
        std::memcpy(best, cand, sizeof(int) * (size_t)n);
        have = true;
      }
    }
  }
  *m = A.m;
  for (int c = 0; c < A.m; ++c) {
    parts[c] = A.size[c];
    oriented[c] = A.oriented[c] ? 1 : 0;
  }
  if (word)
    for (int p = 0; p < n; ++p) word[p] = (uint8_t)best[p];
  if (rep_omega) Generator::decode_word(A, best, rep_omega);
  return TD16_OK;
}
