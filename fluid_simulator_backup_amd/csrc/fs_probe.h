// fs_probe.h -- in-loop duration samples of one kernel (measurement only).
//
// begin()/end() bracket one launch with events on the launching stream; the
// pair is read back at the next begin() or at harvest(), by which point the
// stream has long passed it, so the host never waits on work it has not
// already queued behind it.  `launches` counts every launch of the kernel the
// caller attributes to the probe, sampled or not, so that the kernel's share
// of a frame is avg * launches.
#pragma once
#include <hip/hip_runtime.h>

struct FsLaunchProbe {
    hipEvent_t a = nullptr, b = nullptr;
    bool pending = false;
    double sum_ms = 0;
    long samples = 0, launches = 0;

    void create() {
        if (!a) (void)hipEventCreate(&a);
        if (!b) (void)hipEventCreate(&b);
    }
    void destroy() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        a = b = nullptr;
    }
    void reset() {
        pending = false;
        sum_ms = 0;
        samples = launches = 0;
    }
    void harvest() {
        if (!pending) return;
        pending = false;
        float ms = 0;
        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
            sum_ms += ms;
            samples++;
        }
    }
    void begin(hipStream_t st) {
        harvest();
        (void)hipEventRecord(a, st);
    }
    void end(hipStream_t st) {
        (void)hipEventRecord(b, st);
        pending = true;
    }
};
