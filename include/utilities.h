/*
 * utilities.h -- device choice and allocation contract of libGPUALS.so.
 *
 * Replaces the reference's shared GPU runtime helpers
 * (IBMSparkGPU/CUDA-MLlib utilities.h:36-41, utilities.cu:23-94):
 *   get_gpu     C linkage, round-robin device choice        utilities.cu:56-94
 *   mallocBest  C++ linkage, device memory with a fallback
 *               to mapped pinned host memory                utilities.cu:31-40
 *   freeBest    C++ linkage, frees either kind              utilities.cu:43-53
 * The C++ declarations keep the reference's linkage so the mangled
 * symbols (_Z10mallocBestPPvm, _Z8freeBestPv) resolve for binaries built
 * against the reference header; the enum return type (hipError_t) has the
 * size and success value (0) of cudaError_t.  C callers use the
 * cumf_malloc_best / cumf_free_best aliases.
 *
 * Differences: get_gpu never resets devices (the reference resets every
 * device on first use, utilities.cu:86-89) and returns the device it
 * claimed, not a re-read of the shared counter (utilities.cu:93);
 * nothing here calls exit().
 */
#ifndef CUMF_UTILITIES_H_
#define CUMF_UTILITIES_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif
/* Next device in round-robin order (thread-safe); -1 if there is none. */
int get_gpu(void);
/* Returns 0 on success (hipSuccess). */
int cumf_malloc_best(void** devPtr, size_t size);
int cumf_free_best(void* devPtr);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
#include <hip/hip_runtime_api.h>
hipError_t mallocBest(void** devPtr, size_t size);
hipError_t freeBest(void* devPtr);
#endif

#endif /* CUMF_UTILITIES_H_ */
