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
    uint32_t *cig;
    const KswJob *jobs;
    KswOut *out;
    const int32_t *order;      // this launch's jobs, one workgroup each
    int njobs;
    StripeRec *recs;           // per stripe: mqe partials
    // optional host mirror (pinned, mapped): each pair's result, CIGAR and a
    // completion flag raised once both are there; null when not wanted
    KswOut *host_out;
    uint32_t *host_cig;
    int32_t *host_flag;
    KswScore sc;
    // wide launches only (launch_stripe_wide): the first workgroup of each
    // job of this launch (njobs + 1 ascending entries, the last one past the
    // launch), the jobs' control blocks (kWideCtl dwords each, in order
    // position) and the plan's redo count (pairs whose team gave up a wait
    // and were recomputed by their epilogue workgroup)
    const int32_t *wfirst;
    int32_t *wctl;
    int32_t *werr;
    int force_late;            // tests: every wide pair is recomputed as if a wait had given up
    // host-raised stop flag (pinned, mapped) sampled every 256 columns, or
    // null: a stopped pass publishes a count past every wait and ends, and
    // the epilogue skips the walk (the host reports such pairs stopped)
    const int32_t *stop;
};

struct GenArgs {
    const uint8_t *seq;
    uint8_t *state;            // state blocks (gen_layout)
    uint8_t *dir;              // direction rows
    int32_t *offs;             // off[] / off_end[] per pair
    const GenJob *jobs;
    const int32_t *order;      // this launch's jobs, one workgroup each
    int njobs;
    KswOut *out;
    const int32_t *stop;       // host-raised stop flag (pinned), or null
    GenScore sc;
    int ring;                  // > 0: state in LDS rings of this many slots (gen_ring_slots)
    const int32_t *only_if;    // non-null: run job j only when only_if[j] != 0
};

struct GenBacktrackArgs {
    const uint8_t *dir;
    const int32_t *offs;
    uint32_t *cig;
    const GenJob *jobs;
    KswOut *out;
    int njobs;
    int tiled_min;             // pairs with qlen + tlen >= this walk one per wave (staged rows)
};

// Band kernel (ksw2_band.hip): one workgroup of `team` waves per job.
struct BandArgs {
    const uint8_t *seq;
    uint8_t *dir;              // direction rows (reference layout)
    int32_t *offs;             // off[] / off_end[] per pair
    int32_t *bnd;              // per pair: kBandCtl control dwords, then three stripe hand-off
                               // buffers of 2 * (qlen+tlen-1) dwords
    const GenJob *jobs;
    const int32_t *order;      // this launch's jobs, `wgs` workgroups each
    int njobs;
    int wgs;                   // workgroups per job (set by launch_band)
    KswOut *out;
    int32_t *redo;             // per job: 1 when H > 0 was seen (the general kernel must run it)
    GenScore sc;
    int force_redo;            // tests: flag every job for the general kernel
    // host-raised stop flag (pinned, mapped) sampled every 256 anti-diagonals
    // of a stripe, or null: a stopped wave publishes a count past every wait
    // and ends (the host reports such pairs stopped)
    const int32_t *stop;
};
constexpr int kBandTeam = 16;       // waves per pair, at most
constexpr int kBandGroup = 4;       // waves per workgroup: one per SIMD of a CU
constexpr int kBandRowsPerLane = 4; // positions per lane; a stripe is 64 lanes of them
// Control block of a band job: progress counts (one per team wave, 16
// dwords apart), the finish ticket, per-wave result records.
constexpr int kBandCtlStride = 16;
constexpr int kBandCtlTicket = kBandTeam * kBandCtlStride;
constexpr int kBandCtlRec = kBandCtlTicket + 16;
constexpr int kBandCtl = 512;
static_assert(kBandCtlRec + 8 * kBandTeam <= kBandCtl, "band control block");
// Band kernel for a team of `team` waves per job: kBandGroup-wave
// workgroups on different CUs that hand stripes on through device-scope
// counters.  Launches go in chunks of at most half the device's workgroups;
// with at most two such waiting launches at once (the caller's admission),
// no spinning workgroup waits on one that cannot be placed.
hipError_t launch_band(const BandArgs &a, int team, bool right, bool score_only, hipStream_t st);

// Singletrack walk of general pairs: one wavefront per job (all general
// jobs of the plan; those without KSW_EZ_SINGLETRACK return at once).
struct GenWalkArgs {
    const uint8_t *seq;
    const uint8_t *dir;
    const int32_t *offs;
    uint32_t *cig;
    const GenJob *jobs;
    KswOut *out;
    GenScore sc;
    int njobs;
};

// General kernel: one workgroup of `threads` (64..1024) per job.
hipError_t launch_general(const GenArgs &a, int threads, hipStream_t st);
hipError_t launch_general_backtrack(const GenBacktrackArgs &a, hipStream_t st);
// One move of a plan's CIGAR runs into a dense region: n words from word
// offset src of the CIGAR arena to word offset dst of the region.
struct CigarMove {
    uint64_t src;
    uint32_t dst, n;
};
// Runs the moves (one workgroup each), so the host copies only the runs.
hipError_t launch_cigar_compact(const uint32_t *arena, uint32_t *dense, const CigarMove *moves, int n_moves,
                                hipStream_t st);
hipError_t launch_general_walk(const GenWalkArgs &a, hipStream_t st);

// Stripe-kernel variants ("groups" of a plan): direction bytes with left or
// right gap alignment, score only, or the two Singletrack tracks.
enum StripeGroup { kGroupLeft = 0, kGroupRight = 1, kGroupScoreOnly = 2, kGroupTracks = 3, kGroups = 4 };

// Stripe kernel: one workgroup of `team` waves (1..kMaxTeam) per job, R
// rows per lane (8 or 4; the jobs' layouts must use the same R).
hipError_t launch_stripe(const StripeArgs &a, int group, int team, int R, hipStream_t st);
// Persistent one-wave workers (as many as fit on the device, at most one per
// job) claiming the launch's jobs in list order from a device queue head
// (*qhead, zero at launch: clear_queue_heads); every job runs as a one-wave
// team.  A drop-in for launch_stripe with team 1.
hipError_t launch_stripe_queue(const StripeArgs &a, int group, int R, int32_t *qhead, hipStream_t st);
hipError_t clear_queue_heads(int32_t *qhead, int n, hipStream_t st);
int stripe_queue_workers(int group, int R);

// Wide stripe kernel: each job's passes spread over its own number of
// workgroups of kWideGroup waves on different CUs (a team of up to kMaxWide
// waves), handed on through device-scope counters in the job's control
// block.  h_wfirst is the host copy of a.wfirst.  Launches go in chunks of
// at most `max_wgs` workgroups, which the caller keeps within its share of
// the device (every workgroup of a chunk must be able to be resident at
// once beside the other waiting launches that may run with it).
// launch_stripe_wide_epilogue then runs each job's epilogue (mqe, start
// cell, backtrack), one workgroup per job.
constexpr int kMaxWide = 64;        // waves per wide team, at most
constexpr int kWideGroup = 4;       // waves per workgroup: one per SIMD of a CU
constexpr int kWideCtlStride = 16;  // dwords between two waves' progress counts
constexpr int kWideCtlLate = kMaxWide * kWideCtlStride;  // raised when a wait gave up
constexpr int kWideCtl = kWideCtlLate + 64;              // dwords per job
hipError_t launch_stripe_wide(const StripeArgs &a, int group, int R, const int32_t *h_wfirst, int max_wgs,
                              hipStream_t st);
// Workgroups of the wide stripe kernel the device holds at once.
int stripe_wide_capacity(int group, int R);
hipError_t launch_stripe_wide_epilogue(const StripeArgs &a, int group, int R, hipStream_t st);
// Zeroes a plan's wide redo count (kWideCtl dwords at werr).
hipError_t clear_wide_err(int32_t *werr, hipStream_t st);
// Waves of the stripe kernel that fit on the device at once.
int stripe_resident_waves(int group, int R = kRowsPerLane);

// Zero-fills `bytes` bytes at `p` (16-byte aligned) with a kernel on `st`
// (a fill kernel, not a memset, so that a captured graph replays it).
hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t st);

}  // namespace kswhip
