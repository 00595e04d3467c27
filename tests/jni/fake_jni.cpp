// fake_jni.cpp -- a JNIEnv with no JVM behind it, for testing the Spark
// bridge of libGPUALS.so.  Implements the JNI functions the bridge uses over
// a small in-process object heap; every other function-table slot points
// at a trap that aborts, so a bridge calling an unexpected JNI function
// fails loudly.  Exposed to Python through the fj_* C functions.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/jni_min.h"

namespace {

struct Obj {
  enum Kind { FLOATS, INTS, OBJECTS, CLASS } kind;
  std::vector<float> f;
  std::vector<int32_t> i;
  std::vector<Obj*> o;
  std::string name;
};

struct FakeEnv {
  JNIEnv_ env;  // must be first: JNIEnv* == FakeEnv*
  JNINativeInterface_ table;
  std::vector<std::unique_ptr<Obj>> heap;
  bool pending = false;
  std::string exc_class, exc_msg;
  int critical_depth = 0;
  int critical_calls = 0;
  int fail_critical_at = -1;  // this call index returns NULL
  int deleted_refs = 0;
};

FakeEnv* F(JNIEnv* e) { return reinterpret_cast<FakeEnv*>(e); }
Obj* O(jobject o) { return reinterpret_cast<Obj*>(o); }
jobject J(Obj* o) { return reinterpret_cast<jobject>(o); }

Obj* alloc(FakeEnv* fe, Obj::Kind k) {
  fe->heap.emplace_back(new Obj{k, {}, {}, {}, {}});
  return fe->heap.back().get();
}

void trap() {
  std::fprintf(stderr, "fake_jni: unexpected JNI function called\n");
  std::abort();
}

jclass FindClass(JNIEnv* e, const char* name) {
  Obj* c = alloc(F(e), Obj::CLASS);
  c->name = name;
  return J(c);
}
jint ThrowNew(JNIEnv* e, jclass c, const char* msg) {
  FakeEnv* fe = F(e);
  fe->pending = true;
  fe->exc_class = O(c)->name;
  fe->exc_msg = msg ? msg : "";
  return 0;
}
jthrowable ExceptionOccurred(JNIEnv* e) { return F(e)->pending ? J(alloc(F(e), Obj::CLASS)) : nullptr; }
void ExceptionClear(JNIEnv* e) { F(e)->pending = false; }
jboolean ExceptionCheck(JNIEnv* e) { return F(e)->pending ? JNI_TRUE : JNI_FALSE; }
void DeleteLocalRef(JNIEnv* e, jobject) { F(e)->deleted_refs++; }
jsize GetArrayLength(JNIEnv*, jarray a) {
  Obj* o = O(a);
  return (jsize)(o->kind == Obj::FLOATS ? o->f.size() : o->kind == Obj::INTS ? o->i.size() : o->o.size());
}
jobjectArray NewObjectArray(JNIEnv* e, jsize len, jclass, jobject init) {
  Obj* a = alloc(F(e), Obj::OBJECTS);
  a->o.assign((size_t)len, O(init));
  return J(a);
}
jobject GetObjectArrayElement(JNIEnv*, jobjectArray a, jsize i) { return J(O(a)->o.at((size_t)i)); }
void SetObjectArrayElement(JNIEnv*, jobjectArray a, jsize i, jobject v) { O(a)->o.at((size_t)i) = O(v); }
jfloatArray NewFloatArray(JNIEnv* e, jsize len) {
  Obj* a = alloc(F(e), Obj::FLOATS);
  a->f.assign((size_t)len, 0.f);
  return J(a);
}
void SetFloatArrayRegion(JNIEnv*, jfloatArray a, jsize s, jsize l, const jfloat* buf) {
  std::memcpy(O(a)->f.data() + s, buf, sizeof(float) * (size_t)l);
}
void* GetPrimitiveArrayCritical(JNIEnv* e, jarray a, jboolean* isCopy) {
  FakeEnv* fe = F(e);
  if (isCopy) *isCopy = JNI_FALSE;
  if (fe->critical_calls++ == fe->fail_critical_at) return nullptr;
  fe->critical_depth++;
  Obj* o = O(a);
  return o->kind == Obj::FLOATS ? (void*)o->f.data() : (void*)o->i.data();
}
void ReleasePrimitiveArrayCritical(JNIEnv* e, jarray, void*, jint) { F(e)->critical_depth--; }

}  // namespace

extern "C" {

void* fj_create() {
  FakeEnv* fe = new FakeEnv();
  void** slots = reinterpret_cast<void**>(&fe->table);
  for (size_t k = 0; k < sizeof(JNINativeInterface_) / sizeof(void*); ++k) slots[k] = (void*)&trap;
  fe->table.FindClass = FindClass;
  fe->table.ThrowNew = ThrowNew;
  fe->table.ExceptionOccurred = ExceptionOccurred;
  fe->table.ExceptionClear = ExceptionClear;
  fe->table.ExceptionCheck = ExceptionCheck;
  fe->table.DeleteLocalRef = DeleteLocalRef;
  fe->table.GetArrayLength = GetArrayLength;
  fe->table.NewObjectArray = NewObjectArray;
  fe->table.GetObjectArrayElement = GetObjectArrayElement;
  fe->table.SetObjectArrayElement = SetObjectArrayElement;
  fe->table.NewFloatArray = NewFloatArray;
  fe->table.SetFloatArrayRegion = SetFloatArrayRegion;
  fe->table.GetPrimitiveArrayCritical = GetPrimitiveArrayCritical;
  fe->table.ReleasePrimitiveArrayCritical = ReleasePrimitiveArrayCritical;
  fe->env.functions = &fe->table;
  return fe;
}
void fj_destroy(void* h) { delete static_cast<FakeEnv*>(h); }
void* fj_env(void* h) { return &static_cast<FakeEnv*>(h)->env; }
void* fj_new_floats(void* h, const float* v, int len) {
  Obj* a = alloc(static_cast<FakeEnv*>(h), Obj::FLOATS);
  a->f.assign(v, v + len);
  return a;
}
void* fj_new_ints(void* h, const int32_t* v, int len) {
  Obj* a = alloc(static_cast<FakeEnv*>(h), Obj::INTS);
  a->i.assign(v, v + len);
  return a;
}
void* fj_new_objects(void* h, int len) {
  Obj* a = alloc(static_cast<FakeEnv*>(h), Obj::OBJECTS);
  a->o.assign((size_t)len, nullptr);
  return a;
}
void fj_set(void* arr, int i, void* v) { static_cast<Obj*>(arr)->o.at((size_t)i) = static_cast<Obj*>(v); }
void* fj_get(void* arr, int i) { return static_cast<Obj*>(arr)->o.at((size_t)i); }
int fj_len(void* arr) {
  Obj* o = static_cast<Obj*>(arr);
  return (int)(o->kind == Obj::FLOATS ? o->f.size() : o->kind == Obj::INTS ? o->i.size() : o->o.size());
}
int fj_floats(void* arr, float* out, int cap) {
  Obj* o = static_cast<Obj*>(arr);
  int n = (int)o->f.size() < cap ? (int)o->f.size() : cap;
  std::memcpy(out, o->f.data(), sizeof(float) * (size_t)n);
  return (int)o->f.size();
}
int fj_exception(void* h, char* cls, int ccap, char* msg, int mcap) {
  FakeEnv* fe = static_cast<FakeEnv*>(h);
  if (!fe->pending) return 0;
  std::snprintf(cls, (size_t)ccap, "%s", fe->exc_class.c_str());
  std::snprintf(msg, (size_t)mcap, "%s", fe->exc_msg.c_str());
  return 1;
}
void fj_clear_exception(void* h) { static_cast<FakeEnv*>(h)->pending = false; }
int fj_critical_depth(void* h) { return static_cast<FakeEnv*>(h)->critical_depth; }
void fj_fail_critical_at(void* h, int call_index) {
  FakeEnv* fe = static_cast<FakeEnv*>(h);
  fe->critical_calls = 0;
  fe->fail_critical_at = call_index;
}

}  // extern "C"
