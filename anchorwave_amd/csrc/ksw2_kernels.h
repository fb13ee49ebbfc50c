// ksw2_kernels.h -- launch interface between the host library and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ksw2_common.h"

namespace kswhip {

struct StripeArgs {
    const uint8_t *seq;
    uint8_t *dir;
    int32_t *bnd;
    const KswJob *jobs;
    KswOut *out;
    const StripeTask *tasks;   // this launch's queue
    int ntasks;
    int32_t *queue_head;       // next task index (zeroed before the launch)
    int32_t *progress;         // per stripe: bottom-row columns published
    StripeRec *recs;           // per stripe: mqe partials
    KswScore sc;
};

struct BacktrackArgs {
    const uint8_t *dir;
    uint32_t *cig;
    const KswJob *jobs;
    KswOut *out;
    const StripeRec *recs;
    int njobs;
    KswScore sc;
};

// Persistent stripe kernel: `workers` waves drain the task queue.
hipError_t launch_stripe(const StripeArgs &a, int R, bool right, bool score_only, int workers, hipStream_t st);
hipError_t launch_backtrack(const BacktrackArgs &a, hipStream_t st);
// Zeroes n int32 (queue heads / progress counters) with a kernel, so the
// sequence stays capturable in a HIP graph.
hipError_t launch_zero(int32_t *p, size_t n, hipStream_t st);
// Waves of the stripe kernel that fit on the device at once.
int stripe_resident_waves(int R, bool right, bool score_only);

}  // namespace kswhip
