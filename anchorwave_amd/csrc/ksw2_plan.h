// ksw2_plan.h -- host-side batch machinery shared by the C ABI entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_common.h"

namespace kswhip {

struct ScoreClass {
    bool generic_mm;  // with GENERIC_SC: s = (t == q) ? mat[0] : mat[1]
    bool fast_ok;     // scoring for which the fast kernel is exact
    int m;
    int sc_N;         // wildcard score of the match/mismatch mode
};

bool normalise_scoring(int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, KswScore *sc,
                       ScoreClass *cls);
int device_init(int device);
// GenScore of a normalised scoring (matrix copied; m <= kMaxM checked by caller).
void make_gen_score(int8_t m, const int8_t *mat, const KswScore &sc, const ScoreClass &cls, GenScore *g);
// The calling thread's progress callback (ksw2_hip_set_progress_callback):
// polled on the host while a plan runs; a zero return raises the plan's stop
// flag.  cb == nullptr: none.
struct Progress {
    ksw2_hip_progress_cb cb = nullptr;
    void *args = nullptr;
    int interval = 256;
};
int run_batch_host(int n, const ksw2_pair_t *pairs, int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2,
                   int8_t e2, ksw_extz_t *ez, void *km, const Progress *prog = nullptr);
// Frees the device block the calling thread's single calls keep between
// calls (run_batch_host).
void release_call_context();
int deliver_cigar(ksw_extz_t *ez, const uint32_t *src, int n, void *km);
int drop_cigar(ksw_extz_t *ez, void *km);
// A pair whose CIGAR comes from the Singletrack walk over two stored tracks.
inline bool track_pair(int flag) { return (flag & KSW_EZ_SINGLETRACK) && !(flag & KSW_EZ_SCORE_ONLY); }
inline bool track_pair(const ksw2_pair_t &p) { return track_pair(p.flag); }
void reset_ez(ksw_extz_t *ez);

// Why a pair cannot take the fast (unbanded stripe) kernel; 0 when it can.
enum FastReject : int {
    kFastOK = 0,
    kRejBand = 1,
    kRejZdrop = 2,
    kRejFlags = 4,
    kRejScoring = 8,
    kRejWildcard = 16,
};
int fast_path_reject(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls);
// True when the pair runs on the general kernel (not fast-path eligible, or
// KSW2_HIP_FORCE_GENERAL set).
bool takes_general(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls);
// Orientation of a stripe-kernel pair: the reference's query is put in the
// stripes and its target swept (transposed) when that gives the pair more
// passes, so a long query no longer sets a long sweep; kq and kt are the
// swept and striped lengths.  KSW2_HIP_TRANSPOSE=0 / 1 forces either
// orientation (tests).
struct StripeGeom {
    bool tr;
    int kq, kt;
};
StripeGeom stripe_geom(const ksw2_pair_t &p);
// Device bytes a pair needs while its plan is alive.
uint64_t pair_footprint(const ksw2_pair_t &p, bool general);
// Threads of the general kernel's workgroup for one pair (64..1024).
int gen_threads(const ksw2_pair_t &p);
// LDS ring slots of the general kernel for one pair (0: state in HBM).
int gen_ring(const ksw2_pair_t &p);
// True when a general-kernel pair may run on the band kernel: its only
// departure from the stripe kernel's domain is a band narrower than the
// matrix, and no Singletrack.
bool band_domain(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls);
// Bands at least this wide in the band kernel's domain always take it;
// narrower ones take it when a plan's narrow pairs are predicted to finish
// sooner there than on the general kernel.
constexpr int kBandWide = 512;
// True when a pair in the band kernel's domain has a band at least
// KSW2_HIP_BAND_MIN (default kBandWide) wide.
bool takes_band(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls);
// Anti-diagonals the reference completes before the band leaves the
// matrix (qlen + tlen - 1 when it never does).
int band_r_end(int qlen, int tlen, int w);

// Device bytes an engine may hold: 85 % of the free device memory, with
// the cached engine block counted as free when with_cache.
uint64_t mem_budget(bool with_cache);

struct DevicePool;
// A pool over the process's cached device block (the batch engine's, kept
// between calls), of `bytes` when it must be allocated; null when the block
// is busy or holds less than `need` and cannot be replaced.  The block
// returns to the cache when the pool is destroyed.
std::unique_ptr<DevicePool> cached_block_pool(uint64_t bytes, uint64_t need);

// Device memory pool of a serving engine: one hipMalloc carved first-fit.
struct DevicePool {
    uint8_t *base = nullptr;
    uint64_t bytes = 0;
    virtual bool alloc(uint64_t n, uint64_t *off) = 0;
    virtual void free(uint64_t off) = 0;
    // gives back part of an allocated block before the block is freed
    virtual void release(uint64_t off, uint64_t n) { (void)off, (void)n; }
    virtual ~DevicePool() {}
};

// Where a plan gets its memory and stream: defaults are its own hipMalloc'd
// block and its own streams (one per kernel-group lane).
struct PlanEnv {
    DevicePool *pool = nullptr;
    hipStream_t stream = nullptr;
    uint8_t *staging = nullptr;      // pinned upload buffer (>= staging_bytes), reused
    uint64_t staging_bytes = 0;
    int32_t *h_stop = nullptr;       // shared stop flag (pinned, mapped) and its
    int32_t *d_stop = nullptr;       // device address
    // host mirror for per-pair completion (pinned, mapped, coherent): result
    // slots, CIGAR words (cig arena layout) and completion flags; host and
    // device addresses
    KswOut *mirror_out = nullptr, *mirror_out_dev = nullptr;
    uint32_t *mirror_cig = nullptr, *mirror_cig_dev = nullptr;
    int32_t *mirror_flag = nullptr, *mirror_flag_dev = nullptr;
    int team = 0;                    // > 0: every stripe-kernel pair gets this team size
    int rows = 0;                    // 4 or 8: the stripe kernel's rows per lane (0: the plan picks)
    bool spread = false;             // 5-16 wave teams may run spread (as in a plan of its own)
    // The plan's four timing events, created by the caller (the plan then
    // creates none and destroys none); nullptr: the plan creates its own.
    hipEvent_t *events = nullptr;
    // Sequences go straight from the caller's buffers to the device (no
    // pinned staging); the seq arena is zero-filled by a kernel first.
    bool direct_upload = false;
    // The stripe and band kernels sample h_stop even with no progress
    // callback (its owner raises it: an engine stopping its flights).
    bool poll_stop = false;
    // Team sizing by a caller that sees more than this plan (the engine):
    // team_T > 0 is the per-wave step budget in place of the plan's own, and
    // wave_bytes > 0 the direction bytes one wave may hold, so a pair gets
    // at least bytes / wave_bytes waves (a long pair then holds its memory
    // for a shorter time).
    double team_T = 0;
    uint64_t wave_bytes = 0;
    // Sequences resident on the device: pair k's query and target start at
    // seq_dev + seq_offs[2k] and seq_dev + seq_offs[2k + 1] (k indexing the
    // pairs array of plan_create; 16-byte aligned, zero-padded to 16 bytes
    // past the end); the plan then uploads none.
    uint8_t *seq_dev = nullptr;
    const uint64_t *seq_offs = nullptr;
};

// Queue heads a plan keeps for its persistent one-wave launches (one per
// launch of one-wave teams; further such launches use one workgroup per pair).
constexpr int kMaxQueueHeads = 16;

// A device-resident batch: sequences uploaded, arenas sized, jobs built.
// Created from host pairs, run any number of times, results read back.
struct Plan {
    int device = 0;
    hipStream_t stream = nullptr;
    KswScore sc{};
    GenScore gsc{};
    int n = 0;                       // pairs in the plan
    int n_tracks_gen = 0;            // general pairs with the Singletrack walk
    int n_fast = 0, n_gen = 0;       // slots [0, n_fast) stripe kernel, then general kernel
    int R = kRowsPerLane;            // rows per lane of the stripe kernel
    std::vector<KswJob> jobs;        // host copy, stripe-kernel slots
    std::vector<GenJob> gjobs;       // host copy, general-kernel slots
    std::vector<int> order;          // plan slot -> caller index
    std::vector<uint64_t> cig_off;   // plan slot -> uint32 offset in the cig arena
    // Per job group (0 left-aligned traceback, 1 right-aligned, 2 score-only,
    // 3 Singletrack tracks): one wide launch and one launch of one-workgroup
    // teams (`team` = the largest; each pair's own size is in its KswJob).
    // `off`/`count` index d_order.
    // wide: a wide launch (each pair on its own number of kWideGroup-wave
    // workgroups), whose control blocks start at block `woff` of d_wctl and
    // whose first-workgroup table starts at entry `wf` of h_wfirst/d_wfirst.
    struct Launch {
        int grp, team, off, count, lane, wide, woff, wf;
    };
    std::vector<Launch> launches;
    struct GLaunch {
        int threads, ring, off, count, lane;
    };
    std::vector<GLaunch> glaunches;  // general kernel, per block size and ring size
    // band kernel, per (group, team); `off`/`count` index d_gorder
    struct BLaunch {
        int grp, team, off, count, lane;
    };
    std::vector<BLaunch> blaunches;
    int n_band = 0, band_off = 0;    // band jobs: d_gorder[band_off .. band_off + n_band)
    int32_t *d_redo = nullptr;       // per general job: the band kernel asks the general kernel to rerun it
    int nstripes_total = 0;
    uint64_t seq_bytes = 0, dir_bytes = 0, bnd_dwords = 0, cig_words = 0;
    uint8_t *block = nullptr;        // own device block (no pool)
    DevicePool *pool = nullptr;      // or a block of this pool
    uint64_t pool_off = 0, block_bytes = 0;
    bool own_streams = false;
    uint8_t *h_seq = nullptr;        // pinned staging kept while an upload may be in flight
    bool own_seq = true, own_stop = true;
    bool poll_stop = false;          // PlanEnv::poll_stop
    int progress_calls = 0;
    KswOut *h_mout = nullptr, *d_mout = nullptr;   // host mirror (not owned)
    uint32_t *h_mcig = nullptr, *d_mcig = nullptr;
    int32_t *h_mflag = nullptr, *d_mflag = nullptr;
    uint8_t *d_seq = nullptr, *d_dir = nullptr;
    int32_t *d_bnd = nullptr;
    uint32_t *d_cig = nullptr;
    KswJob *d_jobs = nullptr;
    KswOut *d_out = nullptr;
    int32_t *d_order = nullptr;      // job indices, grouped by launch, longest first
    int32_t *d_qhead = nullptr;      // queue heads of the persistent one-wave launches (kMaxQueueHeads)
    uint64_t state_bytes = 0, offs_ints = 0;
    uint8_t *d_state = nullptr;      // general kernel state blocks
    int32_t *d_offs = nullptr;       // general kernel off[]/off_end[]
    GenJob *d_gjobs = nullptr;
    int32_t *d_gorder = nullptr;
    int32_t *h_stop = nullptr;       // pinned stop flag the general kernel polls
    int32_t *d_stop = nullptr;
    StripeRec *d_recs = nullptr;
    int n_wide = 0;                  // pairs of wide launches
    bool optional_wide = false;      // every wide pair is a spread pair (may run as a one-workgroup team)
    int32_t *d_wctl = nullptr;       // their control blocks (kWideCtl dwords each)
    int32_t *d_werr = nullptr;       // count of wide pairs recomputed by their epilogue (a wait gave up)
    std::vector<int32_t> h_wfirst;   // per wide launch: each pair's first workgroup, then the end
    int32_t *d_wfirst = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool own_events = true;          // false: ev[] belong to the caller (PlanEnv::events)
    // launches of different kernel groups overlap on their own streams
    static constexpr int kLanes = 4;
    hipStream_t lane_stream[kLanes] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t lane_done[kLanes] = {nullptr, nullptr, nullptr, nullptr};
    float last_dp_ms = 0.f, last_bt_ms = 0.f;
    // the plan's launches captured as one graph (plan_launch, from the
    // second run of an eligible plan), the capture's fork event, and counts
    hipGraphExec_t graph = nullptr;
    hipEvent_t fork_ev = nullptr;
    bool graph_failed = false;
    int runs = 0, graph_replays = 0;
    bool stop_requested = false;     // the progress callback stopped the last run
    bool started = false;            // the last run launched (not stopped at r = 0)
    int max_diags = 0;               // largest qlen + tlen - 1 in the plan
};

int plan_create(Plan **out, const ksw2_pair_t *pairs, const int *idx, int nidx, int8_t m, const ScoreClass &cls,
                const GenScore &gsc, const PlanEnv *env = nullptr);
// The device the library runs on (device_init), -1 before it.
int library_device();
// Creates the process-wide events of the waiting-launch slots now, so that
// later launches create none.  Returns false when a HIP call fails.
bool wait_slots_prepare();
// Called before a stream that may have run waiting launches is destroyed:
// the slots whose events it recorded last wait for them and take fresh
// events, so no slot's event refers to a destroyed stream.
void wait_slots_forget(hipStream_t st);
// Set-up of the no-allocation entry point on the library's device (address
// of its scratch array, its stream and events); done by ksw2_hip_init and
// otherwise by the entry point's first call.  Returns a status code.
int noalloc_prepare();
// One pair through the no-allocation path (the C entry point has checked
// the arguments); the CIGAR goes into ez->cigar.  Returns a status code.
// Host only: whether a pair's plan fits the no-allocation scratch.
bool noalloc_fits(const ksw2_pair_t &pair, int8_t m, const ScoreClass &cls);
int noalloc_run(const ksw2_pair_t &pair, int8_t m, const ScoreClass &cls, const GenScore &gsc, ksw_extz_t *ez);
// launch (asynchronous) / completed? / wait (polls the progress callback) / both
int plan_launch(Plan *p, const Progress *prog = nullptr);
bool plan_done(Plan *p);
int plan_wait(Plan *p, const Progress *prog = nullptr);
int plan_run(Plan *p, const Progress *prog = nullptr);
bool progress_continue(const Progress *prog, int &calls, int total);
// Results of the last run into ez[order[k]] for every slot k (those with
// skip[k] set are left alone).
int plan_fetch(Plan *p, ksw_extz_t *ez, void *km, const std::vector<char> *skip = nullptr);
int deliver_slot(const Plan *p, int k, const KswOut &o, const uint32_t *cig_arena, ksw_extz_t *z, void *km);
// The pool range [*off, *off + *len) of stripe slot k's direction bytes (a
// plan carved from a pool); false when the pair stores none.  Once the slot's
// completion flag is set no kernel of the plan reads or writes that range.
bool plan_dir_span(const Plan *p, int k, uint64_t *off, uint64_t *len);
void plan_destroy(Plan *p);

}  // namespace kswhip
