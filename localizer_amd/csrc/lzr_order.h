/*
 * lzr_order.h -- stream ordering of an object's device work.
 *
 * A context, finder, network or renderer keeps scratch buffers in device memory that
 * every call reuses.  Calls may come on different streams; the scratch makes them
 * dependent, so a call on another stream than the previous call's first waits for that
 * call's kernels (an event recorded at the end of each call).  Inside a graph capture
 * nothing is recorded or waited for: the caller orders the replays.
 */
#ifndef LZR_ORDER_H
#define LZR_ORDER_H

#include <hip/hip_runtime.h>

struct LzrStreamOrder {
    hipEvent_t done = nullptr;
    bool recorded = false;
    hipStream_t last = nullptr;
    bool capturing = false;

    /* Before the call's first kernel. */
    hipError_t begin(hipStream_t s) {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        hipError_t e = hipStreamIsCapturing(s, &cap);
        if (e != hipSuccess) return e;
        capturing = cap != hipStreamCaptureStatusNone;
        if (capturing) return hipSuccess;
        if (!done && (e = hipEventCreateWithFlags(&done, hipEventDisableTiming)) != hipSuccess) return e;
        if (recorded && last != s) return hipStreamWaitEvent(s, done, 0);
        return hipSuccess;
    }
    /* After the call's last kernel. */
    hipError_t end(hipStream_t s) {
        if (capturing) return hipSuccess;
        hipError_t e = hipEventRecord(done, s);
        if (e != hipSuccess) return e;
        recorded = true;
        last = s;
        return hipSuccess;
    }
    void destroy() {
        if (done) (void)hipEventDestroy(done);
        done = nullptr;
    }
};

#endif /* LZR_ORDER_H */
