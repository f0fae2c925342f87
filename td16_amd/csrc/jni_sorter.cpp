// JNI shim: the exact native methods of br.unb.cic.tdp.base.GPUSortingSearch
// (reference: emitter/src/main/java/br/unb/cic/tdp/base/GPUSortingSearch.java:
// 115,117), exported from libtd16_jni.so on top of the C ABI of libtd16.so.
//
// Build (needs a JDK for jni.h; this repository's container has none):
//   make -C td16_amd/csrc jni JAVA_HOME=/path/to/jdk
//
// Contract kept from the reference (emitter/src/native/jni/jni_sorter.cu):
//  - initSlots(totalSlots, devices[], queueBytesBudget); null devices or a
//    non-positive budget are argument errors (:28-33);
//  - searchForSortingSeq(beta[], omega[], initialEvenCycles, minRate,
//    maxMoves, fullSorting, slot) returns int[len][3] of positions, int[0][]
//    when there is no witness;
//  - bytes are read as unsigned (`b & 0xFF`) under GetPrimitiveArrayCritical
//    and released with JNI_ABORT (:71-83);
//  - every failure becomes java.lang.RuntimeException(message) and the
//    method returns null (:15-22, 111-116).
//
// With TD16_SERVE=1 in the environment, initSlots starts the served loop on
// the first device instead (slots = totalSlots, queue capacity from
// TD16_SERVE_QUEUE, default 512) and every searchForSortingSeq becomes a
// ticket in it: the Java side is unchanged, its slot threads become
// producers for one resident kernel.
#include <jni.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/td16_abi.h"

namespace {

bool served_mode() {
  const char *v = std::getenv("TD16_SERVE");
  return v != nullptr && std::strcmp(v, "1") == 0;
}

void throw_runtime(JNIEnv *env, const char *message) {
  if (env->ExceptionCheck()) env->ExceptionClear();
  jclass cls = env->FindClass("java/lang/RuntimeException");
  if (cls != nullptr) {
    env->ThrowNew(cls, message);
    env->DeleteLocalRef(cls);
  }
}

bool copy_unsigned(JNIEnv *env, jbyteArray arr, jsize n, uint8_t *out) {
  auto *p = static_cast<jbyte *>(env->GetPrimitiveArrayCritical(arr, nullptr));
  if (p == nullptr) return false;
  for (jsize i = 0; i < n; ++i) out[i] = static_cast<uint8_t>(p[i] & 0xFF);
  env->ReleasePrimitiveArrayCritical(arr, p, JNI_ABORT);
  return true;
}

}  // namespace

extern "C" JNIEXPORT void JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(
    JNIEnv *env, jclass, jint totalSlots, jintArray javaDevices, jlong queueBytesBudget) {
  char msg[512] = {0};
  if (javaDevices == nullptr) {
    throw_runtime(env, "devices must not be null");
    return;
  }
  if (queueBytesBudget <= 0) {
    throw_runtime(env, "queueBytesBudget must be positive");
    return;
  }
  const jsize nd = env->GetArrayLength(javaDevices);
  std::vector<jint> tmp(static_cast<size_t>(nd > 0 ? nd : 1));
  if (nd > 0) env->GetIntArrayRegion(javaDevices, 0, nd, tmp.data());
  std::vector<int32_t> devices(tmp.begin(), tmp.begin() + nd);
  int st;
  if (served_mode()) {
    if (nd < 1) {
      throw_runtime(env, "devices must contain at least one device ordinal");
      return;
    }
    td16_serve_options o;
    std::memset(&o, 0, sizeof(o));
    o.device = devices[0];
    o.slots = static_cast<int32_t>(totalSlots);
    const char *q = std::getenv("TD16_SERVE_QUEUE");
    o.queue_capacity = q ? std::atoi(q) : 512;
    o.slot_bytes = static_cast<int64_t>(queueBytesBudget);
    st = td16_serve_start(&o, msg, static_cast<int>(sizeof(msg)));
  } else {
    st = td16_init_slots(static_cast<int>(totalSlots), devices.data(), static_cast<int>(nd),
                         static_cast<int64_t>(queueBytesBudget), msg,
                         static_cast<int>(sizeof(msg)));
  }
  if (st != TD16_OK) throw_runtime(env, msg);
}

extern "C" JNIEXPORT jobjectArray JNICALL
Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq(
    JNIEnv *env, jclass, jbyteArray javaBeta, jbyteArray javaOmega, jint initialEvenCycles,
    jfloat minRate, jint maxMoves, jboolean fullSorting, jint slot) {
  char msg[512] = {0};
  if (javaBeta == nullptr || javaOmega == nullptr) {
    throw_runtime(env, "beta and omega must not be null");
    return nullptr;
  }
  const jsize n = env->GetArrayLength(javaBeta);
  if (env->GetArrayLength(javaOmega) != n) {
    throw_runtime(env, "beta and omega must have equal lengths");
    return nullptr;
  }
  if (n > TD16_MAX_N) {
    throw_runtime(env, "performGpuSortingSearch: n exceeds MAX_N");
    return nullptr;
  }
  uint8_t beta[TD16_MAX_N] = {0}, omega[TD16_MAX_N] = {0};
  if (!copy_unsigned(env, javaBeta, n, beta)) {
    throw_runtime(env, "failed to access beta array");
    return nullptr;
  }
  if (!copy_unsigned(env, javaOmega, n, omega)) {
    throw_runtime(env, "failed to access omega array");
    return nullptr;
  }
  int32_t positions[TD16_MAX_DEPTH * 3];
  int32_t len = 0;
  int st;
  if (served_mode()) {
    td16_request req;
    std::memset(&req, 0, sizeof(req));
    std::memcpy(req.beta, beta, static_cast<size_t>(n));
    std::memcpy(req.omega, omega, static_cast<size_t>(n));
    req.n = static_cast<int32_t>(n);
    req.initial_even = static_cast<int32_t>(initialEvenCycles);
    req.min_rate = static_cast<float>(minRate);
    req.max_moves = static_cast<int32_t>(maxMoves);
    req.full_sorting = fullSorting ? 1 : 0;
    req.moveset = TD16_MOVES_REF;
    req.mode = TD16_FIRST;
    td16_result res;
    st = td16_serve_search(&req, &res, msg, static_cast<int>(sizeof(msg)));
    if (st == TD16_OK) {
      len = res.found ? res.depth : 0;
      for (int d = 0; d < len; ++d)
        for (int q = 0; q < 3; ++q) positions[3 * d + q] = res.moves[d][q];
    }
  } else {
    st = td16_search(static_cast<int>(slot), beta, omega, static_cast<int>(n),
                     static_cast<int>(initialEvenCycles), static_cast<float>(minRate),
                     static_cast<int>(maxMoves), fullSorting ? 1 : 0, positions, &len, msg,
                     static_cast<int>(sizeof(msg)));
  }
  if (st != TD16_OK) {
    throw_runtime(env, msg);
    return nullptr;
  }
  jclass intArray = env->FindClass("[I");
  jobjectArray outer = env->NewObjectArray(static_cast<jsize>(len), intArray, nullptr);
  for (jsize d = 0; d < static_cast<jsize>(len); ++d) {
    const jint move[3] = {positions[3 * d], positions[3 * d + 1], positions[3 * d + 2]};
    jintArray inner = env->NewIntArray(3);
    env->SetIntArrayRegion(inner, 0, 3, move);
    env->SetObjectArrayElement(outer, d, inner);
    env->DeleteLocalRef(inner);
  }
  env->DeleteLocalRef(intArray);
  return outer;
}
