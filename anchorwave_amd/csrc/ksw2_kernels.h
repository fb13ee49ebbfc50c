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
    int njobs;
    KswScore sc;
};

struct BacktrackArgs {
    const uint8_t *dir;
    uint32_t *cig;
    const KswJob *jobs;
    KswOut *out;
    int njobs;
    int R;
};

hipError_t launch_stripe(const StripeArgs &a, bool right, bool score_only, hipStream_t st);
hipError_t launch_backtrack(const BacktrackArgs &a, hipStream_t st);

}  // namespace kswhip
