// Internal to libtpgp_hip: the device context and the host-side helpers shared by its
// translation units (the resident engine in tpgp_hip.hip, the generator scan in
// tpgp_generator.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/tpgp_hip.h"

// One device, its stream and its persistent scratch; calls on a context are serialised by the
// mutex (as the reference's CUDA context, cuda_rev.cu:22).  It is recursive because a batch
// that records a step graph holds it from the start of the capture to its end, across the
// calls that enqueue the step.
struct tpgp_hip_context {
    int device;
    int cus;
    hipStream_t stream;
    void *storage;
    size_t capacity;
    std::recursive_mutex mutex;
    char name[256];
    // launch-shape choices (tpgp_hip_set_tuning): per key width (1 or 2 words), parameter and
    // size bucket; 0 = the default
    unsigned char tune[2][TPGP_HIP_TUNE_PARAMS][TPGP_HIP_TUNE_BUCKETS] = {};
};

namespace {
// printf into a caller's error buffer (NULL or zero capacity: nothing)
inline void msg(char *e, size_t n, const char *fmt, ...) {
    if (!e || !n) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e, n, fmt, ap);
    va_end(ap);
}
inline tpgp_hip_status hip_fail(hipError_t code, const char *what, char *e, size_t n) {
    msg(e, n, "%s: %s", what, hipGetErrorString(code));
    (void)hipGetLastError();  // reported now: a later launch check must not report it again
    return code == hipErrorOutOfMemory ? TPGP_HIP_ENOMEM : TPGP_HIP_EDEVICE;
}
// A failure after copies were queued on `st`: wait for the stream first, so that no queued copy
// outlives the host buffer it reads or writes, then report.
inline tpgp_hip_status drain_fail(hipStream_t st, hipError_t code, const char *what, char *e,
                                  size_t n) {
    (void)hipStreamSynchronize(st);
    return hip_fail(code, what, e, n);
}
struct DeviceGuard {  // save/restore the caller's current device
    int previous = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&previous) == hipSuccess && hipSetDevice(dev) == hipSuccess) ok = true;
    }
    ~DeviceGuard() {
        if (previous >= 0) (void)hipSetDevice(previous);
    }
};
}  // namespace
