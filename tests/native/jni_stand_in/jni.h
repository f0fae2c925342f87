/* Test stand-in for <jni.h>: only the types, macros and JNIEnv calls that
 * td16_amd/csrc/jni_sorter.cpp uses, dispatched through a table that the
 * test driver fills with a fake Java heap.  Its layout is not a JVM's, so a
 * shim compiled against it is for tests only and is never shipped. */
#ifndef TD16_TEST_JNI_STAND_IN_H
#define TD16_TEST_JNI_STAND_IN_H

#include <stdint.h>

typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef uint8_t jboolean;
typedef float jfloat;
typedef jint jsize;

struct FakeObject;
typedef FakeObject *jobject;
typedef jobject jclass;
typedef jobject jarray;
typedef jobject jbyteArray;
typedef jobject jintArray;
typedef jobject jobjectArray;

#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL
#define JNI_ABORT 2

struct JNIEnv;
struct JNIEnvCalls {
  jboolean (*exception_check)(JNIEnv *);
  void (*exception_clear)(JNIEnv *);
  jclass (*find_class)(JNIEnv *, const char *);
  jint (*throw_new)(JNIEnv *, jclass, const char *);
  void (*delete_local_ref)(JNIEnv *, jobject);
  void *(*get_critical)(JNIEnv *, jarray, jboolean *);
  void (*release_critical)(JNIEnv *, jarray, void *, jint);
  jsize (*get_array_length)(JNIEnv *, jarray);
  void (*get_int_region)(JNIEnv *, jintArray, jsize, jsize, jint *);
  jobjectArray (*new_object_array)(JNIEnv *, jsize, jclass, jobject);
  jintArray (*new_int_array)(JNIEnv *, jsize);
  void (*set_int_region)(JNIEnv *, jintArray, jsize, jsize, const jint *);
  void (*set_object_element)(JNIEnv *, jobjectArray, jsize, jobject);
};

struct JNIEnv {
  const JNIEnvCalls *calls;
  jboolean ExceptionCheck() { return calls->exception_check(this); }
  void ExceptionClear() { calls->exception_clear(this); }
  jclass FindClass(const char *name) { return calls->find_class(this, name); }
  jint ThrowNew(jclass c, const char *m) { return calls->throw_new(this, c, m); }
  void DeleteLocalRef(jobject o) { calls->delete_local_ref(this, o); }
  void *GetPrimitiveArrayCritical(jarray a, jboolean *copy) {
    return calls->get_critical(this, a, copy);
  }
  void ReleasePrimitiveArrayCritical(jarray a, void *p, jint mode) {
    calls->release_critical(this, a, p, mode);
  }
  jsize GetArrayLength(jarray a) { return calls->get_array_length(this, a); }
  void GetIntArrayRegion(jintArray a, jsize s, jsize n, jint *b) {
    calls->get_int_region(this, a, s, n, b);
  }
  jobjectArray NewObjectArray(jsize n, jclass c, jobject init) {
    return calls->new_object_array(this, n, c, init);
  }
  jintArray NewIntArray(jsize n) { return calls->new_int_array(this, n); }
  void SetIntArrayRegion(jintArray a, jsize s, jsize n, const jint *b) {
    calls->set_int_region(this, a, s, n, b);
  }
  void SetObjectArrayElement(jobjectArray a, jsize i, jobject v) {
    calls->set_object_element(this, a, i, v);
  }
};

#endif
