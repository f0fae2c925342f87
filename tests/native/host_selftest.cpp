// CPU self-test of libtd16's host core -- td16_host.cpp (validation,
// relabelling, tables, replay, codec, generators), td16_prepare.cpp (the
// threaded batch preparation and decoding) and td16_bracelet.cpp (the
// threaded class generator) -- with no HIP runtime.  tests/test_native_host.py
// builds it with g++ under AddressSanitizer + UndefinedBehaviorSanitizer and,
// separately, under ThreadSanitizer, and runs it.  Prints "host selftest ok".
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/td16_abi.h"
#include "../../td16_amd/csrc/td16_internal.h"

namespace td16 {
int make_job_checked(const td16_request *r, Job *job, int *trivial, char *msg, int cap);
void parallel_parts(int parts, const std::function<void(int)> &fn);
}  // namespace td16

using namespace td16;

static int g_fail = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);   \
      ++g_fail;                                                           \
    }                                                                     \
  } while (0)

// The feed order written out directly: a stable sort on (no slack first,
// deeper bound first, more symbols first), then the sorted list dealt to the
// kFeedShards contiguous shards one request per shard per round.
static void reference_prepare(const std::vector<td16_request> &reqs, std::vector<Job> &jobs,
                              std::vector<int> &where) {
  std::vector<Job> all;
  std::vector<int> idx;
  for (int q = 0; q < (int)reqs.size(); ++q) {
    Job j;
    int trivial = 0;
    if (make_job_checked(&reqs[q], &j, &trivial, nullptr, 0) != TD16_OK || trivial) continue;
    j.tag = (uint32_t)q;
    all.push_back(j);
    idx.push_back(q);
  }
  const uint32_t nj = (uint32_t)all.size();
  std::vector<uint32_t> order(nj);
  for (uint32_t q = 0; q < nj; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    const Job &x = all[a], &y = all[b];
    const bool tx = (x.flags & kFlagTight) != 0, ty = (y.flags & kFlagTight) != 0;
    if (tx != ty) return tx;
    return x.maxDepth != y.maxDepth ? x.maxDepth > y.maxDepth : x.N > y.N;
  });
  jobs.assign(nj, Job());
  where.assign(nj, -1);
  const uint32_t per = nj ? (nj + kFeedShards - 1) / kFeedShards : 1;
  uint32_t next = 0;
  for (uint32_t t = 0; next < nj; ++t)
    for (uint32_t s = 0; s < (uint32_t)kFeedShards && next < nj; ++s) {
      const uint64_t pos = (uint64_t)s * per + t;
      if (pos >= std::min<uint64_t>((uint64_t)(s + 1) * per, nj)) continue;
      jobs[pos] = all[order[next]];
      where[pos] = idx[order[next]];
      ++next;
    }
}

// A mixed batch: products of 1..9 block transpositions on 2..16 elements in
// every mode, sorted inputs, and requests too small to search.
static std::vector<td16_request> mixed(uint64_t seed, int count) {
  std::vector<uint8_t> ns(count), ks(count);
  std::vector<uint64_t> perms(count);
  td16_gen_products(seed, 2, 16, 1, 9, count, ns.data(), ks.data(), perms.data());
  std::vector<td16_request> reqs(count);
  for (int q = 0; q < count; ++q) {
    const int moveset = q % 5 == 0 ? TD16_MOVES_ALL : TD16_MOVES_REF;
    const int mode = q % 7 == 0 ? TD16_COUNT_ALL : TD16_FIRST;
    const int iddfs = q % 11 == 0;
    td16_linear_request(perms[q], ns[q], ks[q], moveset, mode, iddfs, &reqs[q]);
    reqs[q].lex_order = (uint8_t)(q % 3 == 0);
    if (q % 97 == 0) {  // a sorted input: trivial for iterative deepening
      td16_request &r = reqs[q];
      for (int x = 0; x < r.n; ++x) r.omega[x] = (uint8_t)x;
      r.iddfs = 1;
    }
    if (q % 89 == 0) {  // two symbols: no triple
      td16_request &r = reqs[q];
      r.n = 2;
      r.beta[0] = 0, r.beta[1] = 1, r.omega[0] = 1, r.omega[1] = 0;
    }
  }
  return reqs;
}

static void test_prepare() {
  const int sizes[] = {1, 2, 63, 64, 65, 100, 1000, 4097, 20000, 100000};
  for (int count : sizes) {
    std::vector<td16_request> reqs = mixed(1000 + count, count);
    std::vector<Job> want;
    std::vector<int> wantWhere;
    reference_prepare(reqs, want, wantWhere);
    for (int threads : {1, 3, 16}) {
      std::vector<td16_result> res(count);
      std::vector<Job> got(count);
      std::vector<int> gotWhere(count);
      PrepareScratch ws;
      uint32_t nj = 0;
      char msg[256];
      CHECK(prepare_batch_into(reqs.data(), count, res.data(), got.data(), gotWhere.data(), &nj,
                               ws, threads, msg, sizeof msg) == TD16_OK);
      CHECK(nj == want.size());
      if (nj != want.size()) continue;
      if (nj) CHECK(std::memcmp(got.data(), want.data(), sizeof(Job) * nj) == 0);
      CHECK(std::equal(wantWhere.begin(), wantWhere.end(), gotWhere.begin()));
      // each launched request knows its feed position, the rest are marked
      for (uint32_t p = 0; p < nj; ++p) CHECK(ws.pos[(size_t)gotWhere[p]] == p);
      size_t marked = 0;
      for (int q = 0; q < count; ++q) marked += ws.pos[(size_t)q] == kNotLaunched;
      CHECK(marked == (size_t)count - nj);
      // trivial requests are answered in place
      for (int q = 0; q < count; ++q)
        if (q % 89 == 0 && count > q) CHECK(res[q].status == TD16_OK && res[q].found == 0);
    }
  }
  // an invalid request stops the batch with its own status and a reason
  std::vector<td16_request> reqs = mixed(7, 30000);
  reqs[12345].omega[0] = reqs[12345].omega[1];
  reqs[22222].max_moves = 10;
  std::vector<td16_result> res(reqs.size());
  std::vector<Job> got(reqs.size());
  std::vector<int> where(reqs.size());
  PrepareScratch ws;
  uint32_t nj = 0;
  char msg[256] = {0};
  CHECK(prepare_batch_into(reqs.data(), (int)reqs.size(), res.data(), got.data(), where.data(),
                           &nj, ws, 16, msg, sizeof msg) == TD16_INVALID_ARGUMENT);
  CHECK(res[12345].status == TD16_INVALID_ARGUMENT);
  CHECK(std::strstr(msg, "permutations") != nullptr);
}

static void test_deal() {
  for (uint32_t nj = 1; nj <= 3000; ++nj) {
    const uint32_t per = (nj + kFeedShards - 1) / kFeedShards;
    std::vector<uint32_t> pos;
    for (uint32_t t = 0; pos.size() < nj; ++t)
      for (uint32_t s = 0; s < (uint32_t)kFeedShards && pos.size() < nj; ++s) {
        const uint64_t p = (uint64_t)s * per + t;
        if (p < std::min<uint64_t>((uint64_t)(s + 1) * per, nj)) pos.push_back((uint32_t)p);
      }
    for (uint32_t r = 0; r < nj; ++r) CHECK(deal_position(r, nj, per) == pos[r]);
  }
}

static void test_pool() {
  for (int rep = 0; rep < 200; ++rep) {
    const int parts = 1 + rep % 24;
    std::vector<uint64_t> part(parts, 0);
    parallel_parts(parts, [&](int p) {
      uint64_t s = 0;
      for (int x = 0; x < 1000; ++x) s += (uint64_t)(p * 1000 + x);
      part[p] = s;
    });
    uint64_t total = 0;
    for (uint64_t v : part) total += v;
    const uint64_t n = (uint64_t)parts * 1000;
    CHECK(total == n * (n - 1) / 2);
  }
}

static void test_decode() {
  const uint32_t nj = 50000;
  std::vector<DevResult> dev(nj);
  std::vector<int> where(nj);
  for (uint32_t q = 0; q < nj; ++q) {
    std::memset(&dev[q], 0, sizeof(DevResult));
    dev[q].found = (int)(q & 1);
    dev[q].depth = dev[q].found ? 1 : 0;
    dev[q].moves[0] = (uint16_t)(0 | (1 << 5) | (2 << 10));
    dev[q].expanded = q + 1;
    dev[q].solutions = q & 3;
    dev[q].split = (int)(q % 5 == 0);
    where[q] = (int)(nj - 1 - q);
  }
  std::vector<td16_result> res(nj);
  std::vector<uint32_t> pos;
  positions_of(where.data(), nj, nj, pos);
  uint64_t e = 0, s = 0, d = 0;
  decode_batch(dev.data(), pos.data(), nj, res.data(), 16, &e, &s, &d);
  uint64_t we = 0, wsol = 0, wd = 0;
  for (uint32_t q = 0; q < nj; ++q) {
    we += q + 1, wsol += q & 3, wd += q % 5 == 0;
    const td16_result &r = res[nj - 1 - q];
    CHECK(r.expanded == q + 1 && r.found == (int)(q & 1));
    if (r.found) CHECK(r.moves[0][0] == 0 && r.moves[0][1] == 1 && r.moves[0][2] == 2);
  }
  CHECK(e == we && s == wsol && d == wd);
}

static void test_certify_and_check_threads() {
  // the reference's smoke vector many times over, every third witness broken
  const int n = 60000;
  std::vector<td16_request> reqs(n);
  std::vector<td16_result> res(n);
  for (int q = 0; q < n; ++q) {
    td16_request &r = reqs[q];
    std::memset(&r, 0, sizeof r);
    r.n = 3;
    r.beta[0] = 0, r.beta[1] = 1, r.beta[2] = 2;
    r.omega[0] = 1, r.omega[1] = 2, r.omega[2] = 0;
    r.initial_even = 1, r.min_rate = 2.0f, r.max_moves = 1, r.full_sorting = 1;
    std::memset(&res[q], 0, sizeof(td16_result));
    res[q].found = 1, res[q].depth = 1;
    res[q].moves[0][0] = 0, res[q].moves[0][1] = q % 3 == 0 ? 2 : 1, res[q].moves[0][2] = 2;
  }
  std::vector<uint8_t> keys((size_t)n * TD16_CERT_KEY_BYTES), vals((size_t)n * TD16_CERT_VALUE_BYTES);
  std::vector<int32_t> vlen(n), st(n);
  char msg[256];
  const int rc = td16_certify_batch(reqs.data(), res.data(), n, 8, keys.data(), vals.data(),
                                    vlen.data(), st.data(), msg, sizeof msg);
  CHECK(rc == TD16_OK);
  int ok = 0;
  for (int q = 0; q < n; ++q) ok += st[q] == 0;
  CHECK(ok == n - (n + 2) / 3);
  std::vector<uint8_t> betas((size_t)n * TD16_MAX_N), omegas((size_t)n * TD16_MAX_N);
  std::vector<int32_t> ns(n, 3), mm(n, 1), cst(n);
  for (int q = 0; q < n; ++q) {
    std::memcpy(&betas[(size_t)q * TD16_MAX_N], reqs[q].beta, 3);
    std::memcpy(&omegas[(size_t)q * TD16_MAX_N], reqs[q].omega, 3);
  }
  CHECK(td16_check_batch(betas.data(), omegas.data(), ns.data(), mm.data(), n, vals.data(),
                         vlen.data(), 8, cst.data(), msg, sizeof msg) == TD16_OK);
  for (int q = 0; q < n; ++q) CHECK(cst[q] == (st[q] == 0 ? 0 : 1));
}

static void test_bracelets_threads() {
  // one of the unoriented stage's partitions of 17 symbols: rows are the
  // same, in the same order, for any thread count
  const int32_t parts[] = {7, 4, 2, 2, 2};
  const uint8_t oriented[] = {0, 0, 0, 0, 0};
  int64_t total1 = 0, total8 = 0;
  char msg[256];
  CHECK(td16_bracelets(parts, oriented, 5, TD16_BRACELET_PAIRS, 1, nullptr, 0, &total1, msg,
                       sizeof msg) == TD16_OK);
  CHECK(total1 == 4800);  // emitter.log:30
  std::vector<uint8_t> a((size_t)total1 * 17), b((size_t)total1 * 17);
  CHECK(td16_bracelets(parts, oriented, 5, TD16_BRACELET_PAIRS, 1, a.data(), total1, &total1,
                       msg, sizeof msg) == TD16_OK);
  CHECK(td16_bracelets(parts, oriented, 5, TD16_BRACELET_PAIRS, 8, b.data(), total1, &total8,
                       msg, sizeof msg) == TD16_OK);
  CHECK(total8 == total1 && a == b);
}

int main() {
  test_deal();
  test_pool();
  test_prepare();
  test_decode();
  test_certify_and_check_threads();
  test_bracelets_threads();
  if (g_fail) {
    std::printf("host selftest: %d failures\n", g_fail);
    return 1;
  }
  std::printf("host selftest ok\n");
  return 0;
}
