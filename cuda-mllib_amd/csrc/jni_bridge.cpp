// jni_bridge.cpp -- Spark JNI bridge of the ALS half-sweep.
//
// Reference: als/src/CuMFJNIInterface.cpp:25-138.  Same marshalling and
// ordering:
//   * the source factors arrive as float[blocks][factors][f] and are
//     concatenated in order, so csrCol indexes that concatenation
//     (CuMFJNIInterface.cpp:60-76);
//   * the CSR arrays are pinned with Get/ReleasePrimitiveArrayCritical,
//     uploaded to mallocBest memory (:78-99) and handed to doALSWithCSR
//     with X_BATCH = 1, which consumes them (:109);
//   * the result is returned as float[m][f] (:111-118);
//   * the device is chosen round-robin with get_gpu() (:31-32).
// Differences: argument lengths are checked (the reference's checks are
// commented out, :35-42); every critical section is released on every
// path; failures raise a Java exception and return NULL instead of
// printing and falling off the end of the function (:129-131).
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/als.h"
#include "../../include/org_apache_spark_ml_recommendation_CuMFJNIInterface.h"
#include "../../include/utilities.h"

namespace {

void throw_java(JNIEnv* env, const char* cls, const std::string& msg) {
  if (env->ExceptionCheck()) return;  // keep the first exception
  jclass c = env->FindClass(cls);
  if (c) env->ThrowNew(c, msg.c_str());
}

// Pinned host buffer (hipHostMalloc), freed on scope exit.
struct Pinned {
  void* p = nullptr;
  explicit Pinned(size_t bytes) {
    if (bytes && hipHostMalloc(&p, bytes, 0) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

// Critical section over a primitive array, released on scope exit.
struct Critical {
  JNIEnv* env;
  jarray arr;
  void* p = nullptr;
  Critical(JNIEnv* e, jarray a) : env(e), arr(a) { p = a ? env->GetPrimitiveArrayCritical(a, nullptr) : nullptr; }
  void release() {
    if (p) env->ReleasePrimitiveArrayCritical(arr, p, JNI_ABORT);
    p = nullptr;
  }
  ~Critical() { release(); }
};

}  // namespace

extern "C" {

JNIEXPORT int cumf_jni_flatten_factors(JNIEnv* env, jobjectArray sortedSrcFactors, jint f, jint n,
                                       float* out) {
  if (!sortedSrcFactors) {
    throw_java(env, "java/lang/IllegalArgumentException", "sortedSrcFactors is null");
    return -1;
  }
  const jsize nblocks = env->GetArrayLength(sortedSrcFactors);
  jint index = 0;
  for (jsize b = 0; b < nblocks; ++b) {
    jobjectArray block = (jobjectArray)env->GetObjectArrayElement(sortedSrcFactors, b);
    if (!block) {
      throw_java(env, "java/lang/IllegalArgumentException", "null factor block");
      return -1;
    }
    const jsize nf = env->GetArrayLength(block);
    for (jsize j = 0; j < nf; ++j) {
      jfloatArray factor = (jfloatArray)env->GetObjectArrayElement(block, j);
      if (!factor || env->GetArrayLength(factor) != f || index >= n) {
        if (factor) env->DeleteLocalRef(factor);
        env->DeleteLocalRef(block);
        throw_java(env, "java/lang/IllegalArgumentException",
                   index >= n ? "more source factors than n" : "source factor of wrong length");
        return -1;
      }
      {
        Critical c(env, factor);
        if (!c.p) {
          env->DeleteLocalRef(factor);
          env->DeleteLocalRef(block);
          throw_java(env, "java/lang/OutOfMemoryError", "GetPrimitiveArrayCritical failed");
          return -1;
        }
        std::memcpy(out + (size_t)index * f, c.p, sizeof(float) * (size_t)f);
      }
      ++index;
      env->DeleteLocalRef(factor);
    }
    env->DeleteLocalRef(block);
  }
  if (index != n) {
    throw_java(env, "java/lang/IllegalArgumentException",
               "source factor count " + std::to_string(index) + " != n " + std::to_string(n));
    return -1;
  }
  return 0;
}

JNIEXPORT jobjectArray cumf_jni_rows_to_array(JNIEnv* env, jint m, jint f, const float* X) {
  jclass floatArrayClass = env->FindClass("[F");
  if (!floatArrayClass) return nullptr;
  jobjectArray output = env->NewObjectArray(m, floatArrayClass, nullptr);
  if (!output) return nullptr;
  for (jint i = 0; i < m; ++i) {
    jfloatArray row = env->NewFloatArray(f);
    if (!row) return nullptr;  // OutOfMemoryError pending
    env->SetFloatArrayRegion(row, 0, f, X + (size_t)i * f);
    env->SetObjectArrayElement(output, i, row);
    env->DeleteLocalRef(row);
    if (env->ExceptionCheck()) return nullptr;
  }
  return output;
}

JNIEXPORT jobjectArray JNICALL Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_doALSWithCSR(
    JNIEnv* env, jobject obj, jint m, jint n, jint f, jint nnz, jdouble lambda,
    jobjectArray sortedSrcFactors, jintArray csrRow, jintArray csrCol, jfloatArray csrVal) {
  (void)obj;
  if (m <= 0 || n <= 0 || f <= 0 || nnz < 0 || !csrRow || !csrCol || !csrVal) {
    throw_java(env, "java/lang/IllegalArgumentException", "doALSWithCSR: bad arguments");
    return nullptr;
  }
  if (env->GetArrayLength(csrRow) != m + 1 || env->GetArrayLength(csrCol) != nnz ||
      env->GetArrayLength(csrVal) != nnz) {
    throw_java(env, "java/lang/IllegalArgumentException",
               "doALSWithCSR: csrRow must have m+1 entries, csrCol and csrVal nnz");
    return nullptr;
  }
  const int dev = get_gpu();
  if (dev < 0 || hipSetDevice(dev) != hipSuccess) {
    (void)hipGetLastError();
    throw_java(env, "java/lang/RuntimeException", "doALSWithCSR: no GPU available");
    return nullptr;
  }
  Pinned theta(sizeof(float) * (size_t)n * f), X(sizeof(float) * (size_t)m * f);
  if (!theta.p || !X.p) {
    throw_java(env, "java/lang/OutOfMemoryError", "doALSWithCSR: pinned host allocation failed");
    return nullptr;
  }
  if (cumf_jni_flatten_factors(env, sortedSrcFactors, f, n, (float*)theta.p) != 0) return nullptr;

  int *dRow = nullptr, *dCol = nullptr;
  float* dVal = nullptr;
  std::string err;
  {
    Critical row(env, csrRow), col(env, csrCol), val(env, csrVal);
    if (!row.p || !col.p || !val.p) err = "GetPrimitiveArrayCritical failed";
    if (err.empty() && (mallocBest((void**)&dRow, sizeof(int) * ((size_t)m + 1)) != hipSuccess ||
                        mallocBest((void**)&dCol, sizeof(int) * (size_t)nnz) != hipSuccess ||
                        mallocBest((void**)&dVal, sizeof(float) * (size_t)nnz) != hipSuccess))
      err = "mallocBest failed";
    if (err.empty() && (hipMemcpy(dRow, row.p, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(dCol, col.p, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(dVal, val.p, sizeof(float) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess))
      err = std::string("upload failed: ") + hipGetErrorString(hipGetLastError());
  }  // critical sections released here, before any further JNI call
  if (!err.empty()) {
    (void)freeBest(dRow);
    (void)freeBest(dCol);
    (void)freeBest(dVal);
    throw_java(env, "java/lang/RuntimeException", "doALSWithCSR: " + err);
    return nullptr;
  }
  hipStream_t stream = nullptr;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    (void)freeBest(dRow);
    (void)freeBest(dCol);
    (void)freeBest(dVal);
    throw_java(env, "java/lang/RuntimeException", "doALSWithCSR: hipStreamCreate failed");
    return nullptr;
  }
  doALSWithCSR(stream, dRow, dCol, dVal, (float*)theta.p, (float*)X.p, m, n, f, nnz, (float)lambda, 1);
  err = cumf_last_error();
  (void)hipStreamDestroy(stream);
  if (!err.empty()) {
    throw_java(env, "java/lang/RuntimeException", "doALSWithCSR: " + err);
    return nullptr;
  }
  return cumf_jni_rows_to_array(env, m, f, (const float*)X.p);
}

JNIEXPORT void JNICALL Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_testjni(JNIEnv* env,
                                                                                         jobject obj) {
  (void)env;
  (void)obj;
  std::printf("*******in native code of testjni ...\n");
}

}  // extern "C"
