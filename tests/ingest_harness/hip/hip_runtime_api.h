/*
 * Host-only stand-in for the few HIP runtime calls the rawspec program's
 * ingest makes, so that its threading and buffer sequencing can be tested
 * on a machine without a GPU: "device" memory is host memory, copies are
 * memcpy (optionally slowed down), events and streams are no-ops.
 */
#ifndef INGEST_HARNESS_HIP_H_
#define INGEST_HARNESS_HIP_H_

#include <stddef.h>

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum { hipMemcpyHostToDevice = 1 };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
enum { hipHostMallocDefault = 0, hipHostMallocWriteCombined = 4 };

hipError_t hipSetDevice(int dev);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, int kind, hipStream_t s);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
const char *hipGetErrorString(hipError_t e);

#endif
