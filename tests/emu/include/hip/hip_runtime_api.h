// Host API of the CPU wave emulator: the subset of the HIP runtime API that
// the sieve library calls, executed synchronously on host memory.
//
// Test infrastructure only.  The sieve's sources compile against this header
// (found before ROCm's through the include path) into a CPU library with the
// same C ABI, which the tests drive to check the kernels' control flow,
// barriers and cross-lane operations on the CPU.  Nothing in the product loads
// it.
#ifndef RPS_EMU_HIP_RUNTIME_API_H
#define RPS_EMU_HIP_RUNTIME_API_H

#include <cstddef>
#include <cstdint>

enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorInvalidDevice = 101,
    hipErrorLaunchFailure = 719,
};

enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
};

struct EmuStream;
struct EmuEvent;
typedef EmuStream *hipStream_t;
typedef EmuEvent *hipEvent_t;

constexpr unsigned hipStreamDefault = 0;
constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipEventDefault = 0;
constexpr unsigned hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;

struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t totalGlobalMem;
    int major;
    int minor;
};

#define HIP_SYMBOL(x) (reinterpret_cast<const void *>(&(x)))

hipError_t hipGetDeviceCount(int *count);
hipError_t hipSetDevice(int device);
hipError_t hipGetDevice(int *device);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *prop, int device);
hipError_t hipDeviceGetPCIBusId(char *bus, int len, int device);
const char *hipGetErrorString(hipError_t error);
hipError_t hipGetLastError();

hipError_t hipMalloc(void **ptr, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T **ptr, size_t bytes) {
    return hipMalloc(reinterpret_cast<void **>(ptr), bytes);
}
hipError_t hipFree(void *ptr);
hipError_t hipHostMalloc(void **ptr, size_t bytes, unsigned flags = 0);
template <class T>
inline hipError_t hipHostMalloc(T **ptr, size_t bytes, unsigned flags = 0) {
    return hipHostMalloc(reinterpret_cast<void **>(ptr), bytes, flags);
}
hipError_t hipHostFree(void *ptr);
hipError_t hipGetSymbolAddress(void **ptr, const void *symbol);

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t stream = nullptr);
hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t stream = nullptr);
hipError_t hipMemset(void *dst, int value, size_t bytes);

hipError_t hipStreamCreate(hipStream_t *stream);
hipError_t hipStreamCreateWithFlags(hipStream_t *stream, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t stream);
hipError_t hipStreamSynchronize(hipStream_t stream);
hipError_t hipStreamWaitEvent(hipStream_t stream, hipEvent_t event, unsigned flags = 0);
hipError_t hipDeviceSynchronize();

hipError_t hipEventCreate(hipEvent_t *event);
hipError_t hipEventCreateWithFlags(hipEvent_t *event, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t event);
hipError_t hipEventRecord(hipEvent_t event, hipStream_t stream = nullptr);
hipError_t hipEventSynchronize(hipEvent_t event);
hipError_t hipEventQuery(hipEvent_t event);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t start, hipEvent_t stop);

#endif
