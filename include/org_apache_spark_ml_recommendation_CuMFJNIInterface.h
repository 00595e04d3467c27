/*
 * JNI entry points of libGPUALS.so for the Spark class
 * org.apache.spark.ml.recommendation.CuMFJNIInterface.
 *
 * Same symbols and signatures as the reference's javah-generated header
 * (IBMSparkGPU/CUDA-MLlib
 * als/src/org_apache_spark_ml_recommendation_CuMFJNIInterface.h:26-40):
 *   doALSWithCSR  (IIIID[[[F[I[I[F)[[F
 *   testjni       ()V
 * plus two marshalling helpers (no counterpart in the reference) that the
 * bridge is built from, exported so a test harness can drive them with a
 * fake JNIEnv and no GPU.
 */
#ifndef CUMF_JNI_INTERFACE_H_
#define CUMF_JNI_INTERFACE_H_

#include "jni_min.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Returns float[m][f] with the updated factors of the m rows of the CSR,
 * or NULL with a pending java.lang.RuntimeException /
 * IllegalArgumentException. */
JNIEXPORT jobjectArray JNICALL Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_doALSWithCSR(
    JNIEnv* env, jobject obj, jint m, jint n, jint f, jint nnz, jdouble lambda,
    jobjectArray sortedSrcFactors, jintArray csrRow, jintArray csrCol, jfloatArray csrVal);

JNIEXPORT void JNICALL Java_org_apache_spark_ml_recommendation_CuMFJNIInterface_testjni(JNIEnv* env,
                                                                                         jobject obj);

/* Concatenates the float[blocks][factors][f] source factors in order into
 * out (n x f).  Returns 0, or -1 with a pending exception. */
JNIEXPORT int cumf_jni_flatten_factors(JNIEnv* env, jobjectArray sortedSrcFactors, jint f, jint n,
                                       float* out);
/* Builds float[m][f] from X (m x f).  NULL with a pending exception. */
JNIEXPORT jobjectArray cumf_jni_rows_to_array(JNIEnv* env, jint m, jint f, const float* X);

#ifdef __cplusplus
}
#endif
#endif
