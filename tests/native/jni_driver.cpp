// Drives the JNI shim (td16_amd/csrc/jni_sorter.cpp) the way the Java
// dispatcher does (GPUSortingSearch.java:115,117,123-162), through a fake
// JNIEnv whose "Java heap" is a few C++ objects, and checks the marshalling
// contract of the reference's native methods (jni_sorter.cu:24-117):
// int[len][3] positions, int[0][] without a witness, null plus a
// java.lang.RuntimeException(message) on every error, bytes read as
// unsigned, every critical section released with JNI_ABORT.
//
// Built twice by the tests: against the stub C ABI in td16_stub.cpp (no GPU;
// the stub also records what the shim passed down) and against libtd16.so
// (a GPU test: the smoke vector and a real instance end to end).
#include <jni.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/td16_abi.h"

extern "C" {
void Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(JNIEnv *, jclass, jint, jintArray,
                                                         jlong);
jobjectArray Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq(
    JNIEnv *, jclass, jbyteArray, jbyteArray, jint, jfloat, jint, jboolean, jint);
}

// ---------------------------------------------------------------------------
// Fake Java heap
// ---------------------------------------------------------------------------

struct FakeObject {
  enum Kind { kClass, kBytes, kInts, kObjects } kind;
  std::string name;  // class name, or the element class of an object array
  std::vector<jbyte> bytes;
  std::vector<jint> ints;
  std::vector<FakeObject *> elems;
};

namespace {

std::vector<std::unique_ptr<FakeObject>> gHeap;
std::string gPendingClass, gPendingMessage;
int gOpenCritical = 0, gAbortReleases = 0, gOtherReleases = 0, gDeleted = 0;

FakeObject *alloc(FakeObject::Kind k) {
  gHeap.emplace_back(new FakeObject{k, "", {}, {}, {}});
  return gHeap.back().get();
}

jboolean exception_check(JNIEnv *) { return gPendingClass.empty() ? 0 : 1; }
void exception_clear(JNIEnv *) { gPendingClass.clear(), gPendingMessage.clear(); }
jclass find_class(JNIEnv *, const char *name) {
  FakeObject *c = alloc(FakeObject::kClass);
  c->name = name;
  return c;
}
jint throw_new(JNIEnv *, jclass c, const char *m) {
  gPendingClass = c->name;
  gPendingMessage = m ? m : "";
  return 0;
}
void delete_local_ref(JNIEnv *, jobject) { ++gDeleted; }
void *get_critical(JNIEnv *, jarray a, jboolean *copy) {
  if (copy) *copy = 0;
  ++gOpenCritical;
  return a->kind == FakeObject::kBytes ? (void *)a->bytes.data() : (void *)a->ints.data();
}
void release_critical(JNIEnv *, jarray, void *, jint mode) {
  --gOpenCritical;
  if (mode == JNI_ABORT) ++gAbortReleases;
  else ++gOtherReleases;
}
jsize get_array_length(JNIEnv *, jarray a) {
  switch (a->kind) {
    case FakeObject::kBytes: return (jsize)a->bytes.size();
    case FakeObject::kInts: return (jsize)a->ints.size();
    case FakeObject::kObjects: return (jsize)a->elems.size();
    default: return 0;
  }
}
void get_int_region(JNIEnv *, jintArray a, jsize s, jsize n, jint *b) {
  for (jsize i = 0; i < n; ++i) b[i] = a->ints[(size_t)(s + i)];
}
jobjectArray new_object_array(JNIEnv *, jsize n, jclass c, jobject init) {
  FakeObject *o = alloc(FakeObject::kObjects);
  o->name = c ? c->name : "";
  o->elems.assign((size_t)n, init);
  return o;
}
jintArray new_int_array(JNIEnv *, jsize n) {
  FakeObject *o = alloc(FakeObject::kInts);
  o->ints.assign((size_t)n, 0);
  return o;
}
void set_int_region(JNIEnv *, jintArray a, jsize s, jsize n, const jint *b) {
  for (jsize i = 0; i < n; ++i) a->ints[(size_t)(s + i)] = b[i];
}
void set_object_element(JNIEnv *, jobjectArray a, jsize i, jobject v) {
  a->elems[(size_t)i] = v;
}

const JNIEnvCalls kCalls = {exception_check, exception_clear, find_class, throw_new,
                            delete_local_ref, get_critical, release_critical,
                            get_array_length, get_int_region, new_object_array,
                            new_int_array, set_int_region, set_object_element};

jbyteArray bytes_of(const std::vector<int> &v) {
  FakeObject *o = alloc(FakeObject::kBytes);
  for (int x : v) o->bytes.push_back((jbyte)x);  // Java bytes are signed
  return o;
}
jintArray ints_of(const std::vector<int> &v) {
  FakeObject *o = alloc(FakeObject::kInts);
  for (int x : v) o->ints.push_back(x);
  return o;
}

int gFail = 0;
#define CHECK(c)                                                                  \
  do {                                                                            \
    if (!(c)) {                                                                   \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);           \
      ++gFail;                                                                    \
    }                                                                             \
  } while (0)

JNIEnv gEnv{&kCalls};

jobjectArray search(const std::vector<int> &beta, const std::vector<int> &omega, int even,
                    float rate, int maxMoves, int slot) {
  exception_clear(&gEnv);
  return Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq(
      &gEnv, nullptr, bytes_of(beta), bytes_of(omega), even, rate, maxMoves, 1, slot);
}

// The shim's answer as position triples; `ok` false when it returned null.
std::vector<std::vector<int>> triples(jobjectArray r, bool *ok) {
  std::vector<std::vector<int>> out;
  *ok = r != nullptr;
  if (!r) return out;
  CHECK(r->kind == FakeObject::kObjects && r->name == "[I");
  for (FakeObject *e : r->elems) {
    CHECK(e != nullptr && e->kind == FakeObject::kInts && e->ints.size() == 3);
    if (e && e->ints.size() == 3) out.push_back({e->ints[0], e->ints[1], e->ints[2]});
  }
  return out;
}

void expect_runtime_exception(const char *fragment) {
  CHECK(gPendingClass == "java/lang/RuntimeException");
  CHECK(gPendingMessage.find(fragment) != std::string::npos);
  if (gPendingMessage.find(fragment) == std::string::npos)
    std::fprintf(stderr, "  message was: %s\n", gPendingMessage.c_str());
}

void common_argument_checks() {
  bool ok;
  // null arrays, unequal lengths, n beyond MAX_N: null plus an exception
  exception_clear(&gEnv);
  CHECK(Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq(
            &gEnv, nullptr, nullptr, bytes_of({1, 2, 0}), 1, 2.0f, 1, 1, 0) == nullptr);
  expect_runtime_exception("must not be null");
  triples(search({0, 1, 2}, {1, 2, 0, 3}, 1, 2.0f, 1, 0), &ok);
  CHECK(!ok);
  expect_runtime_exception("equal lengths");
  std::vector<int> big(18);
  for (int i = 0; i < 18; ++i) big[(size_t)i] = i;
  triples(search(big, big, 18, 1.0f, 9, 0), &ok);
  CHECK(!ok);
  expect_runtime_exception("MAX_N");
  // initSlots argument checks come before the native call
  exception_clear(&gEnv);
  Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(&gEnv, nullptr, 2, nullptr, 1 << 20);
  expect_runtime_exception("devices must not be null");
  exception_clear(&gEnv);
  Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(&gEnv, nullptr, 2, ints_of({0}), 0);
  expect_runtime_exception("queueBytesBudget must be positive");
}

}  // namespace

#ifdef TD16_JNI_STUB_ABI
// Recorded by the stub C ABI (td16_stub.cpp).
extern int gStubSlots, gStubDevices[8], gStubNdev;
extern long long gStubBudget;
extern uint8_t gStubBeta[32];
extern int gStubN, gStubServeQueue, gStubServeSlots;

int run() {
  bool ok;
  common_argument_checks();
  // initSlots hands the slot count, device list and budget down unchanged
  exception_clear(&gEnv);
  Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(&gEnv, nullptr, 7, ints_of({3, 1}),
                                                      850LL << 20);
  CHECK(!exception_check(&gEnv));
  CHECK(gStubSlots == 7 && gStubNdev == 2 && gStubDevices[0] == 3 && gStubDevices[1] == 1);
  CHECK(gStubBudget == (850LL << 20));
  // a witness: int[1][3]
  auto t = triples(search({0, 1, 2}, {1, 2, 0}, 1, 2.0f, 1, 0), &ok);
  CHECK(ok && t.size() == 1 && t[0] == std::vector<int>({0, 1, 2}));
  // no witness: int[0][]
  t = triples(search({0, 1, 2, 3}, {0, 1, 2, 3}, 4, 1.0f, 2, 1), &ok);
  CHECK(ok && t.empty() && !exception_check(&gEnv));
  // a native error: null and RuntimeException(message)
  t = triples(search({0, 1, 2, 3, 4}, {1, 2, 3, 4, 0}, 1, 1.0f, 3, 0), &ok);
  CHECK(!ok);
  expect_runtime_exception("stub rejects five symbols");
  // bytes are unsigned: -56 as a Java byte reaches the native side as 200
  search({-56, 1, 2}, {1, 2, 0}, 1, 2.0f, 1, 0);
  CHECK(gStubN == 3 && gStubBeta[0] == 200);
  // the served mode: the same calls become served-loop tickets
  setenv("TD16_SERVE", "1", 1);
  setenv("TD16_SERVE_QUEUE", "77", 1);
  exception_clear(&gEnv);
  Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(&gEnv, nullptr, 16, ints_of({0}),
                                                      1LL << 30);
  CHECK(!exception_check(&gEnv) && gStubServeSlots == 16 && gStubServeQueue == 77);
  t = triples(search({0, 1, 2}, {1, 2, 0}, 1, 2.0f, 1, 0), &ok);
  CHECK(ok && t.size() == 1 && t[0] == std::vector<int>({0, 1, 2}));
  unsetenv("TD16_SERVE");
  return 0;
}
#else
int run() {
  bool ok;
  common_argument_checks();
  // the reference's GPU smoke vector (GPUSortingSearchSmokeTest.java:26-35)
  exception_clear(&gEnv);
  Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(&gEnv, nullptr, 2, ints_of({0}),
                                                      32LL << 20);
  CHECK(!exception_check(&gEnv));
  auto t = triples(search({0, 1, 2}, {1, 2, 0}, 1, 2.0f, 1, 1), &ok);
  CHECK(ok && t.size() == 1 && t[0] == std::vector<int>({0, 1, 2}));
  // a product of 6 block transpositions on 12 symbols: a witness of at most
  // 6 moves that replays
  uint64_t perm = 0;
  uint8_t ns = 0, ks = 0;
  td16_gen_products(7, 12, 12, 6, 6, 1, &ns, &ks, &perm);
  td16_request req;
  td16_linear_request(perm, 12, 6, TD16_MOVES_REF, TD16_FIRST, 0, &req);
  std::vector<int> beta(req.beta, req.beta + req.n), omega(req.omega, req.omega + req.n);
  t = triples(search(beta, omega, req.initial_even, req.min_rate, 6, 0), &ok);
  CHECK(ok && !t.empty() && t.size() <= 6);
  std::vector<int32_t> pos, sym(t.size() * 3);
  for (auto &m : t) pos.insert(pos.end(), m.begin(), m.end());
  char msg[256];
  CHECK(td16_positions_to_symbols(req.beta, req.n, pos.data(), (int)t.size(), sym.data(), msg,
                                  sizeof msg) == TD16_OK);
  CHECK(td16_validate(req.beta, req.omega, req.n, sym.data(), (int)t.size(), 6, msg,
                      sizeof msg) == TD16_OK);
  // a native error from the library: RuntimeException with its message
  t = triples(search({0, 0, 2}, {1, 2, 0}, 1, 2.0f, 1, 0), &ok);
  CHECK(!ok);
  expect_runtime_exception("permutations");
  td16_shutdown(msg, sizeof msg);
  return 0;
}
#endif

int main() {
  run();
  CHECK(gOpenCritical == 0);
  CHECK(gAbortReleases > 0 && gOtherReleases == 0);
  if (gFail) {
    std::printf("jni driver: %d failures\n", gFail);
    return 1;
  }
  std::printf("jni driver ok\n");
  return 0;
}
