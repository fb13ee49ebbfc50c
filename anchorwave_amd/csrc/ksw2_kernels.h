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
    const int32_t *order;      // this launch's jobs, one workgroup each
    int njobs;
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

// Stripe kernel: one workgroup of `team` waves (1..kMaxTeam) per job.
hipError_t launch_stripe(const StripeArgs &a, int R, bool right, bool score_only, int team, hipStream_t st);
hipError_t launch_backtrack(const BacktrackArgs &a, hipStream_t st);
// Waves of the stripe kernel that fit on the device at once.
int stripe_resident_waves(int R, bool right, bool score_only);

}  // namespace kswhip
