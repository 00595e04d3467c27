/*
 * jni_min.h -- the subset of the Java Native Interface that the ALS bridge
 * uses, written from the JNI specification ("JNI Functions" chapter: each
 * function's LINKAGE entry gives its index in the JNIEnv function table).
 * No JDK is installed in the build image, so <jni.h> is not available.
 *
 * The function table is laid out by index: unused slots are opaque
 * pointers, named slots sit at their specified index (checked below with
 * static_assert).  Only the C++ form of JNIEnv (a struct whose first
 * member points at the table) is provided; it is binary-identical to the
 * C form.  When a real <jni.h> is available, define CUMF_USE_SYSTEM_JNI
 * and it is used instead.
 */
#ifndef CUMF_JNI_MIN_H_
#define CUMF_JNI_MIN_H_

#ifdef CUMF_USE_SYSTEM_JNI
#include <jni.h>
#else

#include <stddef.h>
#include <stdint.h>

#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL
#define JNI_OK 0
#define JNI_FALSE 0
#define JNI_TRUE 1
#define JNI_ABORT 2

typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef uint8_t jboolean;
typedef uint16_t jchar;
typedef int16_t jshort;
typedef float jfloat;
typedef double jdouble;
typedef jint jsize;

struct _jobject;
typedef struct _jobject* jobject;
typedef jobject jclass;
typedef jobject jthrowable;
typedef jobject jstring;
typedef jobject jarray;
typedef jarray jobjectArray;
typedef jarray jintArray;
typedef jarray jfloatArray;

struct JNIEnv_;
typedef struct JNIEnv_ JNIEnv;

struct JNINativeInterface_ {
  void* slots_0_5[6];  /* reserved0..3, GetVersion, DefineClass */
  jclass(JNICALL* FindClass)(JNIEnv*, const char*);                       /* 6 */
  void* slots_7_13[7];
  jint(JNICALL* ThrowNew)(JNIEnv*, jclass, const char*);                  /* 14 */
  jthrowable(JNICALL* ExceptionOccurred)(JNIEnv*);                        /* 15 */
  void* slot_16;
  void(JNICALL* ExceptionClear)(JNIEnv*);                                 /* 17 */
  void* slots_18_22[5];
  void(JNICALL* DeleteLocalRef)(JNIEnv*, jobject);                        /* 23 */
  void* slots_24_170[147];
  jsize(JNICALL* GetArrayLength)(JNIEnv*, jarray);                        /* 171 */
  jobjectArray(JNICALL* NewObjectArray)(JNIEnv*, jsize, jclass, jobject); /* 172 */
  jobject(JNICALL* GetObjectArrayElement)(JNIEnv*, jobjectArray, jsize);  /* 173 */
  void(JNICALL* SetObjectArrayElement)(JNIEnv*, jobjectArray, jsize, jobject); /* 174 */
  void* slots_175_180[6];
  jfloatArray(JNICALL* NewFloatArray)(JNIEnv*, jsize);                    /* 181 */
  void* slots_182_212[31];
  void(JNICALL* SetFloatArrayRegion)(JNIEnv*, jfloatArray, jsize, jsize, const jfloat*); /* 213 */
  void* slots_214_221[8];
  void*(JNICALL* GetPrimitiveArrayCritical)(JNIEnv*, jarray, jboolean*);  /* 222 */
  void(JNICALL* ReleasePrimitiveArrayCritical)(JNIEnv*, jarray, void*, jint); /* 223 */
  void* slots_224_227[4];
  jboolean(JNICALL* ExceptionCheck)(JNIEnv*);                             /* 228 */
  void* slots_229_233[5];
};

#define CUMF_JNI_SLOT(name, idx) \
  static_assert(offsetof(JNINativeInterface_, name) == (idx) * sizeof(void*), #name " index")
CUMF_JNI_SLOT(FindClass, 6);
CUMF_JNI_SLOT(ThrowNew, 14);
CUMF_JNI_SLOT(ExceptionOccurred, 15);
CUMF_JNI_SLOT(ExceptionClear, 17);
CUMF_JNI_SLOT(DeleteLocalRef, 23);
CUMF_JNI_SLOT(GetArrayLength, 171);
CUMF_JNI_SLOT(NewObjectArray, 172);
CUMF_JNI_SLOT(GetObjectArrayElement, 173);
CUMF_JNI_SLOT(SetObjectArrayElement, 174);
CUMF_JNI_SLOT(NewFloatArray, 181);
CUMF_JNI_SLOT(SetFloatArrayRegion, 213);
CUMF_JNI_SLOT(GetPrimitiveArrayCritical, 222);
CUMF_JNI_SLOT(ReleasePrimitiveArrayCritical, 223);
CUMF_JNI_SLOT(ExceptionCheck, 228);
#undef CUMF_JNI_SLOT

struct JNIEnv_ {
  const struct JNINativeInterface_* functions;
  jclass FindClass(const char* name) { return functions->FindClass(this, name); }
  jint ThrowNew(jclass c, const char* msg) { return functions->ThrowNew(this, c, msg); }
  jboolean ExceptionCheck() { return functions->ExceptionCheck(this); }
  void DeleteLocalRef(jobject o) { functions->DeleteLocalRef(this, o); }
  jsize GetArrayLength(jarray a) { return functions->GetArrayLength(this, a); }
  jobjectArray NewObjectArray(jsize len, jclass c, jobject init) {
    return functions->NewObjectArray(this, len, c, init);
  }
  jobject GetObjectArrayElement(jobjectArray a, jsize i) {
    return functions->GetObjectArrayElement(this, a, i);
  }
  void SetObjectArrayElement(jobjectArray a, jsize i, jobject v) {
    functions->SetObjectArrayElement(this, a, i, v);
  }
  jfloatArray NewFloatArray(jsize len) { return functions->NewFloatArray(this, len); }
  void SetFloatArrayRegion(jfloatArray a, jsize s, jsize l, const jfloat* buf) {
    functions->SetFloatArrayRegion(this, a, s, l, buf);
  }
  void* GetPrimitiveArrayCritical(jarray a, jboolean* isCopy) {
    return functions->GetPrimitiveArrayCritical(this, a, isCopy);
  }
  void ReleasePrimitiveArrayCritical(jarray a, void* p, jint mode) {
    functions->ReleasePrimitiveArrayCritical(this, a, p, mode);
  }
};

#endif /* CUMF_USE_SYSTEM_JNI */
#endif /* CUMF_JNI_MIN_H_ */
