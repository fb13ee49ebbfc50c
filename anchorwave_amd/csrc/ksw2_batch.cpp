// ksw2_batch.cpp -- batches of pairs on one device: eligibility, staging,
// launch, read-back.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_common.h"
#include "ksw2_engine.h"
#include "ksw2_kernels.h"
#include "ksw2_plan.h"

namespace kswhip {

namespace {

std::mutex g_init_mu;
int g_device = -1;

// A failed HIP call names itself on stderr (the caller sees KSW2_HIP_EDEVICE)
#define KSW_HIP_TRY(x)                                                                               \
    do {                                                                                             \
        const hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                      \
            fprintf(stderr, "[ksw2] %s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return KSW2_HIP_EDEVICE;                                                                 \
        }                                                                                            \
    } while (0)

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

}  // namespace

// 85 % of the free device memory; with_cache: at least the cached engine
// block (the library's own, which an engine takes).  KSW2_HIP_MEM_BUDGET
// overrides (tests).
uint64_t mem_budget(bool with_cache)
{
    const char *env = getenv("KSW2_HIP_MEM_BUDGET");
    if (env) return strtoull(env, nullptr, 10);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 8ull << 30;
    const uint64_t own = (uint64_t)(fr * 0.85);
    return with_cache ? std::max<uint64_t>(own, Engine::cached_block_bytes()) : own;
}

namespace {

}  // namespace

namespace {
// Launches whose workgroups wait on one another (band kernel, wide stripe
// teams) need every workgroup of a chunk resident at once, beside the
// chunks of the other such launches that run at the same time; kernels that
// do not wait finish on their own.  The device is split into four quarters,
// each a chain of completion events.  A launch sequence takes two quarters
// (free ones first, else those taken longest ago), waits for them on its own
// stream, keeps every chunk within half the device and records the quarters'
// events again, all under one lock.  So at most two such sequences run at
// once and their chunks together fit the device.  Events rather than
// streams of their own: a stream would hold one of the process's few
// hardware queues for good.
constexpr int kQuarters = 4;
std::mutex g_wait_mu;
hipEvent_t g_quarter[kQuarters] = {};
hipStream_t g_quarter_stream[kQuarters] = {};  // the stream each event was last recorded on
uint64_t g_quarter_used[kQuarters] = {};
uint64_t g_quarter_seq = 0;

// Whether quarter k's last sequence has finished (caller holds g_wait_mu).
// A query error other than "not ready" counts as busy and is cleared from
// the thread's last error, which the launches that follow read.
bool quarter_done(int k)
{
    if (!g_quarter[k]) return true;
    const hipError_t e = hipEventQuery(g_quarter[k]);
    if (e == hipSuccess) return true;
    if (e != hipErrorNotReady) {
        (void)hipGetLastError();
        fprintf(stderr, "[ksw2] wait slot %d: hipEventQuery: %s\n", k, hipGetErrorString(e));
    }
    return false;
}

// Takes two quarters for a sequence about to be enqueued on `st` (caller
// holds g_wait_mu); their indices go to q[0..1].
bool wait_slots_acquire(hipStream_t st, int q[2])
{
    int n = 0;
    bool taken[kQuarters] = {};
    for (int k = 0; k < kQuarters && n < 2; ++k)
        if (quarter_done(k)) q[n++] = k, taken[k] = true;
    while (n < 2) {
        int best = -1;
        for (int k = 0; k < kQuarters; ++k)
            if (!taken[k] && (best < 0 || g_quarter_used[k] < g_quarter_used[best])) best = k;
        q[n++] = best, taken[best] = true;
    }
    for (int i = 0; i < 2; ++i) {
        hipEvent_t &ev = g_quarter[q[i]];
        if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            fprintf(stderr, "[ksw2] wait slot %d: hipEventCreateWithFlags failed\n", q[i]);
            ev = nullptr;
            return false;
        }
        const hipError_t e = hipStreamWaitEvent(st, ev, 0);
        if (e != hipSuccess) {
            fprintf(stderr, "[ksw2] wait slot %d: hipStreamWaitEvent: %s\n", q[i], hipGetErrorString(e));
            return false;
        }
        g_quarter_used[q[i]] = ++g_quarter_seq;
    }
    return true;
}

// Quarters whose last sequence has finished (caller holds g_wait_mu).
int wait_slots_free()
{
    int n = 0;
    for (int k = 0; k < kQuarters; ++k)
        if (quarter_done(k)) ++n;
    return n;
}

bool wait_slots_release(hipStream_t st, const int q[2])
{
    for (int i = 0; i < 2; ++i) {
        const hipError_t e = hipEventRecord(g_quarter[q[i]], st);
        if (e != hipSuccess) {
            fprintf(stderr, "[ksw2] wait slot %d: hipEventRecord: %s\n", q[i], hipGetErrorString(e));
            return false;
        }
        g_quarter_stream[q[i]] = st;
    }
    return true;
}
}  // namespace

void wait_slots_forget(hipStream_t st)
{
    if (!st) return;
    std::lock_guard<std::mutex> lk(g_wait_mu);
    for (int k = 0; k < kQuarters; ++k)
        if (g_quarter[k] && g_quarter_stream[k] == st) {
            // a fresh event in its place (never recorded: the slot is free),
            // so a caller that allocates nothing finds every slot's event
            (void)hipEventSynchronize(g_quarter[k]);
            hipEvent_t fresh = nullptr;
            if (hipEventCreateWithFlags(&fresh, hipEventDisableTiming) != hipSuccess) fresh = nullptr;
            (void)hipEventDestroy(g_quarter[k]);
            g_quarter[k] = fresh, g_quarter_stream[k] = nullptr;
        }
}

bool wait_slots_prepare()
{
    std::lock_guard<std::mutex> lk(g_wait_mu);
    for (hipEvent_t &ev : g_quarter)
        if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
            return false;
        }
    return true;
}

int library_device()
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    return g_device;
}

int device_init(int device)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_device >= 0 && (device < 0 || device == g_device)) return KSW2_HIP_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return KSW2_HIP_ENODEV;
    if (device < 0) device = 0;
    if (device >= count) return KSW2_HIP_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return KSW2_HIP_ENODEV;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return KSW2_HIP_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return KSW2_HIP_ENODEV;
    g_device = device;
    return KSW2_HIP_OK;
}

int fast_path_reject(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls)
{
    int why = 0;
    const int mx = std::max(p.qlen, p.tlen);
    if (!(p.w < 0 || p.w >= mx)) why |= kRejBand;
    if (p.zdrop >= 0) why |= kRejZdrop;
    if (p.flag & KSW_EZ_APPROX_MAX) why |= kRejFlags;
    if (!cls.fast_ok) why |= kRejScoring;
    if (p.flag & KSW_EZ_GENERIC_SC) {
        if (!cls.generic_mm) why |= kRejScoring;
    } else {
        // match/mismatch mode: the wildcard code m-1 scores sc_N; the fast
        // kernel has no wildcard lane, so such pairs go elsewhere
        const uint8_t w = (uint8_t)(m - 1);
        if (memchr(p.query, w, (size_t)p.qlen) || memchr(p.target, w, (size_t)p.tlen)) why |= kRejWildcard;
    }
    return why;
}

// Team size for one pair.  With T the batch's step count per resident wave
// (the time the batch needs when every wave is busy), a pair whose passes
// take P * (qlen + 127) steps on one wave gets the smallest power-of-two team
// that brings it under 2T, capped by its pass count and kMaxWide: small
// pairs run on one wave with no synchronisation, long pairs do not outlast
// the batch.  With a per-wave byte budget (an engine streaming a batch larger
// than device memory), the team also grows until each wave holds at most that
// many direction bytes: a pair holds its bytes from its first pass to its
// walk, so memory held times the time it is held shrinks with the team, and
// it is that product, not the waves, that limits such a batch.  Teams above
// kMaxTeam waves are wide (workgroups of kWideGroup
// waves on different CUs, rounded up to whole workgroups).  KSW2_HIP_TEAM
// forces a size; KSW2_HIP_WIDE=k runs every pair wide on k workgroups
// (tuning and tests).
static double pair_steps(const KswJob &j) { return (double)pass_count(j.tlen, j.R) * (double)pass_steps(j.qlen); }

static int choose_team(const KswJob &j, double T, int forced, uint64_t wave_bytes = 0)
{
    const int P = pass_count(j.tlen, j.R);
    int team = 1;
    if (forced > 0) {
        team = std::max(1, std::min(std::min(forced, kMaxWide), P));
    } else {
        const double steps = pair_steps(j);
        // memory: at most wave_bytes of direction bytes per wave
        const double bytes = (j.flag & KSW_EZ_SCORE_ONLY) ? 0.0
                             : (double)(track_pair(j.flag) ? 2 : 1) * (double)dir_bytes(j.qlen, j.tlen, j.R);
        while (team < kMaxWide && team < P &&
               (steps / team > 2 * T || (wave_bytes > 0 && bytes > (double)wave_bytes * team)))
            team <<= 1;
        team = std::min(team, P);
    }
    return team > kMaxTeam ? (team + kWideGroup - 1) / kWideGroup * kWideGroup : team;
}


// Lone-pair layouts measured by anchorwave_amd/autotune.py on the GPU: per
// bin of swept and striped lengths, the rows per lane, team size and layout
// (one workgroup, or spread over CUs) that ran a pair of that size fastest.
// Read once from KSW2_HIP_TUNING (a path; "0": none) or from
// tuning/gfx950.txt beside the library's package; used for plans that leave
// most of the device idle (latency_plan).
struct TuneEntry {
    int q_lo, q_hi, t_lo, t_hi, rows, team, passes, spread;
    // the team for a pair of P passes at these rows: the measured team scaled
    // to the pair's passes, at most one wave per pass
    int team_for(int P) const
    {
        const int t = team >= passes ? P : (int)(((int64_t)team * P + passes / 2) / passes);
        return std::max(1, std::min(t, P));
    }
};

static const std::vector<TuneEntry> &tune_table()
{
    static std::vector<TuneEntry> table;
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path;
        if (const char *ev = getenv("KSW2_HIP_TUNING")) {
            if (ev[0] == '0' && ev[1] == 0) return;
            path = ev;
        } else {
            Dl_info info;
            if (!dladdr((void *)&tune_table, &info) || !info.dli_fname) return;
            path = info.dli_fname;  // .../anchorwave_amd/lib/libksw2hip.so
            const size_t cut = path.rfind('/');
            path = (cut == std::string::npos ? std::string(".") : path.substr(0, cut)) + "/../tuning/gfx950.txt";
        }
        FILE *f = fopen(path.c_str(), "r");
        if (!f) return;
        char line[512];
        while (fgets(line, sizeof line, f)) {
            TuneEntry e;
            if (sscanf(line, "bin %d %d %d %d %d %d %d %d", &e.q_lo, &e.q_hi, &e.t_lo, &e.t_hi, &e.rows, &e.team,
                       &e.passes, &e.spread) == 8 &&
                (e.rows == 4 || e.rows == 8) && e.team >= 1 && e.team <= kMaxWide && e.passes >= 1)
                table.push_back(e);
        }
        fclose(f);
    });
    return table;
}

// The measured layout for a pair of swept length kq and striped length kt, or null.
static const TuneEntry *tune_lookup(int kq, int kt)
{
    for (const TuneEntry &e : tune_table())
        if (kq >= e.q_lo && kq < e.q_hi && kt >= e.t_lo && kt < e.t_hi) return &e;
    return nullptr;
}

// A plan of its own (or one whose owner allows spread teams) that leaves
// most of the device idle and forces nothing: its time is its pairs' chains
// of passes, so the measured lone-pair layouts apply.
static bool latency_plan(const ksw2_pair_t *pairs, const std::vector<int> &fast, const PlanEnv *env)
{
    if (fast.empty() || (env && (!env->spread || env->team != 0 || env->rows != 0))) return false;
    if (getenv("KSW2_HIP_ROWS") || getenv("KSW2_HIP_TEAM") || getenv("KSW2_HIP_WIDE")) return false;
    int64_t waves = 0;
    for (int id : fast) waves += pass_count(stripe_geom(pairs[id]).kt, 8);
    return 8 * waves <= stripe_resident_waves(kGroupLeft, 4);
}

// The stripe kernel's byte-table score (one v_perm gives two cells' score
// candidates) needs every code below 4 and both candidates, 8|mis| plus a
// tag of at most 7, in a byte.  KSW2_HIP_LUT=0 turns it off (tests).
static bool lut_ok(const ksw2_pair_t &p, const KswScore &sc)
{
    const char *ev = getenv("KSW2_HIP_LUT");
    if ((ev && atoi(ev) == 0) || sc.mch != 0 || -8 * sc.mis + 7 > 255) return false;
    uint8_t acc = 0;
    for (int i = 0; i < p.qlen; ++i) acc |= p.query[i];
    for (int i = 0; i < p.tlen; ++i) acc |= p.target[i];
    return (acc & 0xfc) == 0;
}

bool takes_general(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls)
{
    const char *ev = getenv("KSW2_HIP_FORCE_GENERAL");
    return (ev && atoi(ev) > 0) || fast_path_reject(p, m, cls) != kFastOK;
}

StripeGeom stripe_geom(const ksw2_pair_t &p)
{
    bool tr = pass_count(p.qlen) > pass_count(p.tlen);
    if (const char *ev = getenv("KSW2_HIP_TRANSPOSE")) {
        if (ev[0] == '0') tr = false;
        else if (ev[0] == '1') tr = true;
    }
    return StripeGeom{tr, tr ? p.tlen : p.qlen, tr ? p.qlen : p.tlen};
}

static int norm_w(const ksw2_pair_t &p) { return p.w < 0 ? std::max(p.qlen, p.tlen) : p.w; }

// Rows per lane of a plan's stripe kernel.  Eight rows per lane take the
// fewest VALU operations per cell; four rows give a pair twice the passes,
// so twice the waves can share its sweep.  A plan whose pairs would leave
// most of the device idle at eight rows (a lone pair, a few short ones: the
// sum over pairs of passes, each a wave, is below an eighth of the
// device's resident waves) takes four, which shortens the pair's time when
// its team was capped by its pass count.  A caller that knows the device's
// load (the engine) may choose instead (PlanEnv::rows); KSW2_HIP_ROWS=4 or
// 8 forces it.
static int choose_rows(const ksw2_pair_t *pairs, const std::vector<int> &fast, const PlanEnv *env)
{
    if (const char *ev = getenv("KSW2_HIP_ROWS")) {
        const int r = atoi(ev);
        if (r == 4 || r == 8) return r;
    }
    if (env && (env->rows == 4 || env->rows == 8)) return env->rows;
    if (fast.empty()) return kRowsPerLane;
    if (latency_plan(pairs, fast, env)) {
        // the measured layout of the plan's largest pair
        int big = fast[0];
        for (int id : fast)
            if ((double)pairs[id].qlen * pairs[id].tlen > (double)pairs[big].qlen * pairs[big].tlen) big = id;
        const StripeGeom g = stripe_geom(pairs[big]);
        if (const TuneEntry *e = tune_lookup(g.kq, g.kt)) return e->rows;
    }
    int64_t waves = 0;
    for (int id : fast) waves += pass_count(stripe_geom(pairs[id]).kt, 8);
    return 8 * waves <= stripe_resident_waves(kGroupLeft, 4) ? 4 : 8;
}

uint64_t pair_footprint(const ksw2_pair_t &p, bool general)
{
    // CIGAR slot, twice: the arena and the dense region plan_fetch copies from
    const uint64_t cig = 2 * (4ull * ((uint64_t)p.qlen + p.tlen + 2) + 64), seq = (uint64_t)p.qlen + p.tlen + 32;
    const bool cigar = !(p.flag & KSW_EZ_SCORE_ONLY);
    const uint64_t tracks = track_pair(p) ? 2 : 1;  // Singletrack: two bytes per cell
    if (!general) {
        const StripeGeom g = stripe_geom(p);
        // the larger of both rows-per-lane layouts: direction bytes at eight
        // rows (rows padded to 1024), stripe partials at four (twice the stripes)
        return (cigar ? tracks * dir_bytes(g.kq, g.kt, 8) : 0) + 8ull * g.kq + cig + seq +
               16ull * stripe_count(g.kt, 4) + 4ull * kWideCtl + 512;
    }
    const int tlenB = (p.tlen + 63) / 64 * 64, qlenB = (p.qlen + 63) / 64 * 64;
    const uint64_t diags = (uint64_t)p.qlen + p.tlen - 1;
    const uint64_t rows = cigar ? tracks * diags * (uint64_t)gen_ncolB(p.qlen, p.tlen, norm_w(p)) + 256 : 0;
    // (+ the band kernel's control block, three stripe buffers and redo flag, whichever kernel runs it)
    return gen_layout(tlenB, qlenB).bytes + rows + (cigar ? 8 * diags : 0) + cig + seq + 24 * diags + 4 * kBandCtl + 516;
}

void make_gen_score(int8_t m, const int8_t *mat, const KswScore &sc, const ScoreClass &cls, GenScore *g)
{
    memset(g, 0, sizeof(*g));
    g->k = sc;
    g->m = m;
    g->sc_N = cls.sc_N;
    for (int k = 0; k < m * m && k < kMaxM * kMaxM; ++k) g->mat[k] = mat[k];
}

// The band kernel's domain, whatever the band's width: the pair's only
// departure from the stripe kernel's is a band narrower than the matrix, no
// Singletrack (KSW2_HIP_BAND=0 or KSW2_HIP_FORCE_GENERAL: none).
bool band_domain(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls)
{
    const char *ev = getenv("KSW2_HIP_BAND");
    if (ev && atoi(ev) == 0) return false;
    const char *fg = getenv("KSW2_HIP_FORCE_GENERAL");
    if (fg && atoi(fg) > 0) return false;
    if (p.flag & KSW_EZ_SINGLETRACK) return false;
    // progress counts stripe * (qlen + tlen + 1) + anti-diagonal must fit int32
    const int64_t nstripes = ((int64_t)p.tlen + 63) / 64 * 64 / (kWave * kBandRowsPerLane) + 1;
    if ((nstripes + 1) * ((int64_t)p.qlen + p.tlen + 1) >= INT32_MAX) return false;
    return fast_path_reject(p, m, cls) == kRejBand;
}

bool takes_band(const ksw2_pair_t &p, int8_t m, const ScoreClass &cls)
{
    const char *mn = getenv("KSW2_HIP_BAND_MIN");
    const int min_width = mn ? atoi(mn) : kBandWide;
    return band_domain(p, m, cls) && norm_w(p) + 1 >= min_width;
}

// Narrow bands (w + 1 < kBandWide) of one plan: the band kernel computes
// whole 256-position stripes over 2w + 512 anti-diagonals each, the general
// kernel only the band's cells, one workgroup per pair with a barrier per
// anti-diagonal.  Measured on MI355X (2 kbp pairs, w = 256, and single
// pairs of 2-50 kbp): the band kernel takes about 1.1 us per anti-diagonal
// of a pair's chain and sustains 6.4e10 stripe positions per second, the
// general kernel about 5.8 us per anti-diagonal and 4.4e10 in-band cells per
// second.  So a few pairs finish several times sooner on the band kernel and
// thousands sooner on the general kernel; the plan's narrow pairs take the
// kernel whose predicted time is the shorter.
static bool narrow_bands_to_band(const std::vector<const ksw2_pair_t *> &narrow)
{
    if (narrow.empty()) return false;
    double nr_max = 0, stripe_work = 0, cells = 0;
    for (const ksw2_pair_t *p : narrow) {
        const double nr = (double)p->qlen + p->tlen - 1, w = norm_w(*p);
        nr_max = std::max(nr_max, nr);
        stripe_work += (double)((p->tlen + 255) / 256 * 256) * (2 * w + 512);
        cells += std::min(2 * w + 1, (double)std::min(p->qlen, p->tlen)) * nr / 2;
    }
    const double t_band = std::max(1.1e-6 * nr_max, stripe_work / 6.4e10);
    const double t_gen = std::max(5.8e-6 * nr_max, cells / 4.4e10);
    return t_band < t_gen;
}

int band_r_end(int qlen, int tlen, int w)
{
    const int nr = qlen + tlen - 1;
    for (int r = 0; r < nr; ++r) {
        int st = 0, en = tlen - 1;
        if (st < r - qlen + 1) st = r - qlen + 1;
        if (en > r) en = r;
        if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
        if (en > (r + w) >> 1) en = (r + w) >> 1;
        if (st > en) return r;  // minimap2/ksw2_extd2_sse.c:152-155
    }
    return nr;
}

// LDS ring slots of the general kernel for one pair (0: state in HBM).
// APPROX_MAX pairs keep their state in HBM: their path tracker may read a
// position far behind the band.  KSW2_HIP_GEN_RING=0 turns the rings off
// (tests of both forms).
int gen_ring(const ksw2_pair_t &p)
{
    const char *ev = getenv("KSW2_HIP_GEN_RING");
    if ((ev && atoi(ev) == 0) || (p.flag & KSW_EZ_APPROX_MAX)) return 0;
    return gen_ring_slots(std::min(std::min(p.qlen, p.tlen), norm_w(p) + 1) + 128);
}

// Threads of the general kernel's workgroup for one pair: about four cells
// of the widest anti-diagonal per thread, a power of two in [64, 1024].
// KSW2_HIP_GEN_THREADS forces a size (tuning and tests).
int gen_threads(const ksw2_pair_t &p)
{
    const char *ev = getenv("KSW2_HIP_GEN_THREADS");
    const int forced = ev ? atoi(ev) : 0;
    if (forced >= 64 && forced <= 1024 && forced % 64 == 0) return forced;
    const int window = std::min(std::min(p.qlen, p.tlen), norm_w(p) + 1) + 128;
    int t = 64;
    while (t < 1024 && 2 * t <= window / 4) t <<= 1;
    return t;
}

int plan_create(Plan **out, const ksw2_pair_t *pairs, const int *idx, int nidx, int8_t m, const ScoreClass &cls,
                const GenScore &gsc, const PlanEnv *env)
{
    int st = device_init(-1);
    if (st != KSW2_HIP_OK) return st;
    Plan *p = new Plan();
    p->device = g_device;
    p->sc = gsc.k;
    p->gsc = gsc;
    // slots: stripe-kernel pairs by group (left, right, score-only, tracks),
    // then general-kernel pairs
    std::vector<int> g[kGroups], gen, fast;
    for (int k = 0; k < nidx; ++k) {
        const ksw2_pair_t &pp = pairs[idx[k]];
        if (takes_general(pp, m, cls)) {
            gen.push_back(idx[k]);
            continue;
        }
        fast.push_back(idx[k]);
        const int grp = (pp.flag & KSW_EZ_SCORE_ONLY) ? kGroupScoreOnly
                        : track_pair(pp)               ? kGroupTracks
                        : (pp.flag & KSW_EZ_RIGHT)     ? kGroupRight
                                                       : kGroupLeft;
        g[grp].push_back(idx[k]);
    }
    p->R = choose_rows(pairs, fast, env);
    const int R = p->R, SR = kWave * R;
    std::vector<int> grp_of;
    for (int grp = 0; grp < kGroups; ++grp)
        for (int id : g[grp]) p->order.push_back(id), grp_of.push_back(grp);
    p->n_fast = (int)p->order.size();
    for (int id : gen) p->order.push_back(id);
    p->n_gen = (int)gen.size();
    p->n = (int)p->order.size();
    for (int k = 0; k < p->n; ++k)
        p->max_diags = std::max(p->max_diags, pairs[p->order[k]].qlen + pairs[p->order[k]].tlen - 1);
    p->jobs.resize(p->n_fast);
    p->gjobs.resize(p->n_gen);
    p->cig_off.resize(p->n);
    uint64_t so = 0, dof = 0, bo = 0, co = 0, sto = 0, oo = 0;
    int stripes = 0;
    // sequences already on the device (an engine's resident arena): the
    // jobs address them there and the plan uploads none
    const bool resident = env && env->seq_dev && env->seq_offs;
    for (int k = 0; k < p->n_fast; ++k) {
        const ksw2_pair_t &pp = pairs[p->order[k]];
        KswJob &j = p->jobs[k];
        // the kernel's lengths: swept (qlen) and striped (tlen)
        const StripeGeom g = stripe_geom(pp);
        j.transposed = g.tr ? 1 : 0;
        j.qlen = g.kq, j.tlen = g.kt, j.flag = pp.flag, j.end_bonus = pp.end_bonus;
        j.R = R;
        j.lut = lut_ok(pp, gsc.k) ? 1 : 0;
        if (resident) {
            const uint64_t oq = env->seq_offs[2 * p->order[k]], ot = env->seq_offs[2 * p->order[k] + 1];
            j.q_off = g.tr ? ot : oq, j.t_off = g.tr ? oq : ot;
        } else {
            j.q_off = so, so = round_up(so + g.kq, 16);
            j.t_off = so, so = round_up(so + g.kt, 16);
        }
        j.dir_off = dof;
        if (!(pp.flag & KSW_EZ_SCORE_ONLY))
            dof = round_up(dof + (track_pair(pp) ? 2 : 1) * dir_bytes(g.kq, g.kt, R), 256);
        j.bnd_off = bo, bo += round_up(2ull * g.kq, 64);
        j.cig_off = co, co += round_up((uint64_t)pp.qlen + pp.tlen + 2, 16);
        j.stripe_off = stripes, stripes += (g.kt + SR - 1) / SR;
        p->cig_off[k] = j.cig_off;
    }
    p->nstripes_total = stripes;
    std::vector<char> band(p->n_gen, 0);
    {
        // bands at least kBandWide wide always take the band kernel; the
        // narrower ones together, when it is predicted to finish them sooner
        // (KSW2_HIP_BAND_MIN: a fixed width rule instead)
        const bool fixed_rule = getenv("KSW2_HIP_BAND_MIN") != nullptr;
        std::vector<const ksw2_pair_t *> narrow;
        for (int gk = 0; gk < p->n_gen; ++gk) {
            const ksw2_pair_t &pp = pairs[p->order[p->n_fast + gk]];
            band[gk] = takes_band(pp, m, cls);
            if (!band[gk] && !fixed_rule && band_domain(pp, m, cls)) band[gk] = 2, narrow.push_back(&pp);
        }
        const bool to_band = narrow_bands_to_band(narrow);
        for (char &b : band)
            if (b == 2) b = to_band;
    }
    for (int gk = 0; gk < p->n_gen; ++gk) {
        const int slot = p->n_fast + gk;
        const ksw2_pair_t &pp = pairs[p->order[slot]];
        if (track_pair(pp)) ++p->n_tracks_gen;
        GenJob &j = p->gjobs[gk];
        j.qlen = pp.qlen, j.tlen = pp.tlen, j.w = norm_w(pp), j.zdrop = pp.zdrop;
        j.end_bonus = pp.end_bonus, j.flag = pp.flag;
        j.tlenB = (pp.tlen + 63) / 64 * 64, j.qlenB = (pp.qlen + 63) / 64 * 64;
        j.ncolB = gen_ncolB(pp.qlen, pp.tlen, j.w);
        j.out = slot;
        j.r_end = 0, j.bnd_off = 0;
        if (band[gk]) {
            j.r_end = band_r_end(pp.qlen, pp.tlen, j.w);
            j.bnd_off = bo, bo += kBandCtl + round_up(6ull * ((uint64_t)pp.qlen + pp.tlen - 1), 64);
        }
        if (resident) {
            j.q_off = env->seq_offs[2 * p->order[slot]], j.t_off = env->seq_offs[2 * p->order[slot] + 1];
        } else {
            j.q_off = so, so = round_up(so + pp.qlen, 16);
            j.t_off = so, so = round_up(so + pp.tlen, 16);
        }
        j.st_off = sto, sto += gen_layout(j.tlenB, j.qlenB).bytes;
        const uint64_t diags = (uint64_t)pp.qlen + pp.tlen - 1;
        j.dir_off = dof;
        j.off_off = oo;
        if (!(pp.flag & KSW_EZ_SCORE_ONLY)) {
            dof = round_up(dof + (track_pair(pp) ? 2 : 1) * diags * (uint64_t)j.ncolB, 256);
            oo += round_up(2 * diags, 16);
        }
        j.cig_off = co, co += round_up((uint64_t)pp.qlen + pp.tlen + 2, 16);
        p->cig_off[slot] = j.cig_off;
    }
    // stripe-kernel launches: per group, a wide launch and a launch of
    // one-workgroup teams, longest pairs first
    std::vector<int32_t> order;
    if (p->n_fast > 0) {
        double total_steps = 0;
        for (const KswJob &j : p->jobs) total_steps += pair_steps(j);
        const double T = env && env->team_T > 0
                             ? env->team_T
                             : std::max(1.0, total_steps / std::max(1, stripe_resident_waves(kGroupLeft, R)));
        const uint64_t wave_bytes = env ? env->wave_bytes : 0;
        int forced = env ? env->team : 0;
        if (const char *ev = getenv("KSW2_HIP_TEAM")) forced = atoi(ev);
        int wide_forced = 0;
        if (const char *ev = getenv("KSW2_HIP_WIDE")) wide_forced = std::max(0, std::min(kMaxWide / kWideGroup, atoi(ev)));
        std::vector<int> team(p->n_fast);
        for (int k = 0; k < p->n_fast; ++k) {
            // KSW2_HIP_TEAM=-1 (tests): team sizes 1..16 in turn, one launch
            const int f = forced < 0 ? 1 + k % kMaxTeam : forced;
            team[k] = wide_forced ? kWideGroup * wide_forced : choose_team(p->jobs[k], T, f, wave_bytes);
            p->jobs[k].team = team[k];
        }
        // a plan of its own or a serving flight (PlanEnv::spread) that leaves
        // most of the device idle: teams of 5-16 waves go wide too ("spread"), so that
        // each wave has a SIMD to itself (one workgroup would put two on a
        // SIMD); a lone 5 kbp pair runs 22 % faster.  A spread pair keeps its
        // team size in its KswJob: when no waiting-launch slots are free at
        // launch (other threads' calls hold them), it runs as a one-workgroup
        // team instead of waiting.
        std::vector<char> spread(p->n_fast, 0);
        std::vector<char> tuned(p->n_fast, 0);
        if (!wide_forced && forced == 0 && latency_plan(pairs, fast, env)) {
            // measured lone-pair layouts (a pair of another bin than the one
            // that set the plan's rows keeps its team size, within its passes)
            for (int k = 0; k < p->n_fast; ++k) {
                const KswJob &j = p->jobs[k];
                const TuneEntry *e = tune_lookup(j.qlen, j.tlen);
                if (!e) continue;
                int t = e->team_for(pass_count(j.tlen, j.R));
                if (t > kMaxTeam) t = std::min(kMaxWide, (t + kWideGroup - 1) / kWideGroup * kWideGroup);
                team[k] = t, p->jobs[k].team = t, tuned[k] = 1;
                spread[k] = e->spread && t > kWideGroup && t <= kMaxTeam;
            }
        }
        if (!wide_forced && ((!env && forced == 0) || (env && env->spread && forced >= 0))) {
            int64_t waves = 0;
            for (int k = 0; k < p->n_fast; ++k) waves += team[k];
            if (4 * waves <= stripe_resident_waves(kGroupLeft, R))
                for (int k = 0; k < p->n_fast; ++k)
                    if (!tuned[k] && team[k] > kWideGroup && team[k] <= kMaxTeam) spread[k] = 1;
        }
        auto wide = [&](int k) { return wide_forced > 0 || team[k] > kMaxTeam || spread[k]; };
        p->optional_wide = true;
        for (int k = 0; k < p->n_fast; ++k)
            if (wide(k) && !spread[k]) p->optional_wide = false;
        auto longest_first = [&](std::vector<int> &m) {
            std::stable_sort(m.begin(), m.end(), [&](int x, int y) { return pair_steps(p->jobs[x]) > pair_steps(p->jobs[y]); });
        };
        for (int grp = 0; grp < kGroups; ++grp) {
            // every wide pair of the group in one launch, each on its own
            // number of workgroups
            std::vector<int> members;
            for (int k = 0; k < p->n_fast; ++k)
                if (grp_of[k] == grp && wide(k)) members.push_back(k);
            if (!members.empty()) {
                longest_first(members);
                const int wf = (int)p->h_wfirst.size();
                int blocks = 0, tmax = 0;
                for (int k : members) {
                    p->h_wfirst.push_back(blocks);
                    blocks += (team[k] + kWideGroup - 1) / kWideGroup;
                    tmax = std::max(tmax, team[k]);
                }
                p->h_wfirst.push_back(blocks);
                // team: the largest, for a one-workgroup launch of spread pairs
                p->launches.push_back({grp, tmax, (int)order.size(), (int)members.size(), 0, 1, p->n_wide, wf});
                p->n_wide += (int)members.size();
                order.insert(order.end(), members.begin(), members.end());
            }
            // one-workgroup teams: one launch, workgroups of the largest
            // team (a smaller team's extra waves end at once)
            members.clear();
            int tmax = 0;
            for (int k = 0; k < p->n_fast; ++k)
                if (grp_of[k] == grp && !wide(k)) members.push_back(k), tmax = std::max(tmax, team[k]);
            if (!members.empty()) {
                longest_first(members);
                p->launches.push_back({grp, tmax, (int)order.size(), (int)members.size(), grp % Plan::kLanes, 0, 0, 0});
                order.insert(order.end(), members.begin(), members.end());
            }
        }
    }
    // general-kernel launches: per block size and LDS ring size, most
    // anti-diagonal cells first
    std::vector<int32_t> gorder;
    {
        std::vector<int> thr(p->n_gen), ring(p->n_gen);
        for (int gk = 0; gk < p->n_gen; ++gk) {
            const ksw2_pair_t &pp = pairs[p->order[p->n_fast + gk]];
            thr[gk] = gen_threads(pp);
            ring[gk] = gen_ring(pp);
        }
        for (int t = 1024; t >= 64; t >>= 1)
            for (int rs = kGenRingMax; rs >= 0; rs = rs > 256 ? rs >> 1 : rs - 256) {
                std::vector<int> members;
                for (int gk = 0; gk < p->n_gen; ++gk)
                    if (!band[gk] && thr[gk] == t && ring[gk] == rs) members.push_back(gk);
                if (members.empty()) continue;
                std::stable_sort(members.begin(), members.end(), [&](int x, int y) {
                    const GenJob &a = p->gjobs[x], &b = p->gjobs[y];
                    return (double)(a.qlen + a.tlen) * a.ncolB > (double)(b.qlen + b.tlen) * b.ncolB;
                });
                p->glaunches.push_back({t, rs, (int)gorder.size(), (int)members.size(),
                                        t >= 512 ? 0 : t >= 256 ? 1 : 2});
                gorder.insert(gorder.end(), members.begin(), members.end());
            }
        // band-kernel launches: per group and team (one wave per 512-position
        // stripe, at most kMaxTeam), longest first
        p->band_off = (int)gorder.size();
        const char *bt = getenv("KSW2_HIP_BAND_TEAM");  // forces a team size (tests)
        const int band_team = bt ? std::max(1, std::min(kBandTeam, atoi(bt))) : 0;
        for (int grp = 0; grp < 3; ++grp)
            for (int team = kBandTeam; team >= 1; --team) {
                std::vector<int> members;
                for (int gk = 0; gk < p->n_gen; ++gk) {
                    if (!band[gk]) continue;
                    const GenJob &j = p->gjobs[gk];
                    const int g = (j.flag & KSW_EZ_SCORE_ONLY) ? 2 : (j.flag & KSW_EZ_RIGHT) ? 1 : 0;
                    const int brows = kWave * kBandRowsPerLane;
                    const int tm = band_team ? band_team : std::min(kBandTeam, (j.tlenB + brows - 1) / brows);
                    if (g == grp && tm == team) members.push_back(gk);
                }
                if (members.empty()) continue;
                std::stable_sort(members.begin(), members.end(), [&](int x, int y) {
                    const GenJob &a = p->gjobs[x], &b = p->gjobs[y];
                    return (double)(a.qlen + a.tlen) * a.ncolB > (double)(b.qlen + b.tlen) * b.ncolB;
                });
                p->blaunches.push_back({grp, team, (int)gorder.size(), (int)members.size(), 0});
                gorder.insert(gorder.end(), members.begin(), members.end());
            }
        p->n_band = (int)gorder.size() - p->band_off;
    }
    p->seq_bytes = std::max<uint64_t>(so, 16), p->dir_bytes = std::max<uint64_t>(dof, 256);
    p->bnd_dwords = std::max<uint64_t>(bo, 64), p->cig_words = std::max<uint64_t>(co, 16);
    p->state_bytes = std::max<uint64_t>(sto, 256), p->offs_ints = std::max<uint64_t>(oo, 16);
    auto fail = [&](int code) {
        plan_destroy(p);
        return code;
    };
    if (env && env->stream) {
        p->stream = env->stream;  // caller's stream: every launch goes there
    } else {
        p->own_streams = true;  // lane streams are created on first use (plan_launch)
        if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    }
    if (env && env->events) {
        p->own_events = false;
        for (int k = 0; k < 4; ++k) p->ev[k] = env->events[k];
    } else {
        for (auto &e : p->ev)
            if (hipEventCreate(&e) != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    }
    // one device block holds every arena
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = round_up(at + std::max<uint64_t>(bytes, 1), 256);
        return o;
    };
    const uint64_t o_seq = take(p->seq_bytes), o_dir = take(p->dir_bytes), o_bnd = take(p->bnd_dwords * 4);
    const uint64_t o_cig = take(p->cig_words * 8), o_jobs = take(sizeof(KswJob) * std::max(p->n_fast, 1));
    const uint64_t o_out = take(sizeof(KswOut) * std::max(p->n, 1));
    const uint64_t o_order = take(sizeof(int32_t) * std::max<size_t>(order.size(), 1));
    const uint64_t o_recs = take(sizeof(StripeRec) * std::max(stripes, 1));
    uint64_t o_state = 0, o_offs = 0, o_gjobs = 0, o_gorder = 0;
    if (p->n_gen > 0) {
        o_state = take(p->state_bytes), o_offs = take(p->offs_ints * 4);
        o_gjobs = take(sizeof(GenJob) * p->n_gen), o_gorder = take(sizeof(int32_t) * gorder.size());
    }
    const uint64_t o_redo = p->n_band > 0 ? take(sizeof(int32_t) * p->n_gen) : 0;
    const uint64_t o_wctl = p->n_wide > 0 ? take(4ull * kWideCtl * (p->n_wide + 1)) : 0;  // failure word, blocks
    const uint64_t o_wfirst = p->n_wide > 0 ? take(4ull * p->h_wfirst.size()) : 0;
    const uint64_t o_qhead = take(4ull * kMaxQueueHeads);
    p->block_bytes = at;
    uint8_t *base = nullptr;
    if (env && env->pool) {
        if (!env->pool->alloc(p->block_bytes, &p->pool_off)) return fail(KSW2_HIP_ENOMEM);
        p->pool = env->pool;
        base = env->pool->base + p->pool_off;
    } else {
        if (hipMalloc(&p->block, p->block_bytes) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
        base = p->block;
    }
    p->d_seq = resident ? env->seq_dev : base + o_seq;
    p->d_dir = base + o_dir, p->d_bnd = (int32_t *)(base + o_bnd);
    p->d_cig = (uint32_t *)(base + o_cig), p->d_jobs = (KswJob *)(base + o_jobs), p->d_out = (KswOut *)(base + o_out);
    p->d_order = (int32_t *)(base + o_order), p->d_recs = (StripeRec *)(base + o_recs);
    p->d_qhead = (int32_t *)(base + o_qhead);
    if (p->n_wide > 0) {
        p->d_werr = (int32_t *)(base + o_wctl), p->d_wctl = p->d_werr + kWideCtl;
        p->d_wfirst = (int32_t *)(base + o_wfirst);
    }
    if (p->n_gen > 0) {
        p->d_state = base + o_state, p->d_offs = (int32_t *)(base + o_offs);
        p->d_gjobs = (GenJob *)(base + o_gjobs), p->d_gorder = (int32_t *)(base + o_gorder);
        if (p->n_band > 0) p->d_redo = (int32_t *)(base + o_redo);
    }
    // the stop flag the kernels sample (the general kernel always; the
    // stripe and band kernels when a progress callback runs or the owner of
    // a shared flag asks for it)
    if (env && env->h_stop) {
        p->h_stop = env->h_stop, p->d_stop = env->d_stop, p->own_stop = false;
        p->poll_stop = env->poll_stop;
    } else if (p->n > 0) {
        if (hipHostMalloc((void **)&p->h_stop, sizeof(int32_t), hipHostMallocMapped) != hipSuccess)
            return fail(KSW2_HIP_ENOMEM);
        *p->h_stop = 0;
        if (hipHostGetDevicePointer((void **)&p->d_stop, p->h_stop, 0) != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    }
    // host mirror of the stripe-kernel results (engine flights)
    if (env && env->mirror_flag && p->n_fast > 0) {
        p->h_mout = env->mirror_out, p->h_mcig = env->mirror_cig, p->h_mflag = env->mirror_flag;
        p->d_mout = env->mirror_out_dev, p->d_mcig = env->mirror_cig_dev, p->d_mflag = env->mirror_flag_dev;
        memset(p->h_mflag, 0, sizeof(int32_t) * p->n_fast);
    }
    hipError_t e[6] = {hipSuccess, hipSuccess, hipSuccess, hipSuccess, hipSuccess, hipSuccess};
    if (resident) {
        p->own_seq = false;  // nothing to upload
    } else if (env && env->direct_upload) {
        // straight from the caller's buffers; the arena's padding is zero
        // as with the staged copy
        p->own_seq = false;
        e[0] = launch_fill_zero(p->d_seq, p->seq_bytes, p->stream);
        for (int k = 0; k < p->n_fast && e[0] == hipSuccess; ++k) {
            const ksw2_pair_t &pp = pairs[p->order[k]];
            const KswJob &j = p->jobs[k];
            e[0] = hipMemcpyAsync(p->d_seq + j.q_off, j.transposed ? pp.target : pp.query, (size_t)j.qlen,
                                  hipMemcpyHostToDevice, p->stream);
            if (e[0] == hipSuccess)
                e[0] = hipMemcpyAsync(p->d_seq + j.t_off, j.transposed ? pp.query : pp.target, (size_t)j.tlen,
                                      hipMemcpyHostToDevice, p->stream);
        }
        for (int gk = 0; gk < p->n_gen && e[0] == hipSuccess; ++gk) {
            const ksw2_pair_t &pp = pairs[p->order[p->n_fast + gk]];
            e[0] = hipMemcpyAsync(p->d_seq + p->gjobs[gk].q_off, pp.query, (size_t)pp.qlen, hipMemcpyHostToDevice,
                                  p->stream);
            if (e[0] == hipSuccess)
                e[0] = hipMemcpyAsync(p->d_seq + p->gjobs[gk].t_off, pp.target, (size_t)pp.tlen,
                                      hipMemcpyHostToDevice, p->stream);
        }
    } else {
        // stage sequences through pinned memory (kept until the plan is
        // destroyed when the copy is left in flight on the caller's stream)
        uint8_t *h_seq = nullptr;
        if (env && env->staging && env->staging_bytes >= p->seq_bytes) {
            h_seq = env->staging, p->own_seq = false;
        } else if (hipHostMalloc((void **)&h_seq, p->seq_bytes, hipHostMallocDefault) != hipSuccess) {
            return fail(KSW2_HIP_ENOMEM);
        }
        p->h_seq = h_seq;
        memset(h_seq, 0, p->seq_bytes);
        for (int k = 0; k < p->n_fast; ++k) {
            const ksw2_pair_t &pp = pairs[p->order[k]];
            const KswJob &j = p->jobs[k];
            memcpy(h_seq + j.q_off, j.transposed ? pp.target : pp.query, (size_t)j.qlen);
            memcpy(h_seq + j.t_off, j.transposed ? pp.query : pp.target, (size_t)j.tlen);
        }
        for (int gk = 0; gk < p->n_gen; ++gk) {
            const ksw2_pair_t &pp = pairs[p->order[p->n_fast + gk]];
            memcpy(h_seq + p->gjobs[gk].q_off, pp.query, (size_t)pp.qlen);
            memcpy(h_seq + p->gjobs[gk].t_off, pp.target, (size_t)pp.tlen);
        }
        e[0] = hipMemcpyAsync(p->d_seq, h_seq, p->seq_bytes, hipMemcpyHostToDevice, p->stream);
    }
    if (p->n_fast > 0)
        e[1] = hipMemcpyAsync(p->d_jobs, p->jobs.data(), sizeof(KswJob) * p->n_fast, hipMemcpyHostToDevice, p->stream);
    if (!order.empty())
        e[2] = hipMemcpyAsync(p->d_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice,
                              p->stream);
    if (p->n_gen > 0) {
        e[3] = hipMemcpyAsync(p->d_gjobs, p->gjobs.data(), sizeof(GenJob) * p->n_gen, hipMemcpyHostToDevice,
                              p->stream);
        e[4] = hipMemcpyAsync(p->d_gorder, gorder.data(), sizeof(int32_t) * gorder.size(), hipMemcpyHostToDevice,
                              p->stream);
    }
    if (p->n_wide > 0 && e[2] == hipSuccess)
        e[2] = hipMemcpyAsync(p->d_wfirst, p->h_wfirst.data(), sizeof(int32_t) * p->h_wfirst.size(),
                              hipMemcpyHostToDevice, p->stream);
    if (p->own_streams) {
        e[5] = hipStreamSynchronize(p->stream);
        if (p->own_seq && p->h_seq) (void)hipHostFree(p->h_seq);
        p->h_seq = nullptr;
    }
    for (hipError_t x : e)
        if (x != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    *out = p;
    return KSW2_HIP_OK;
}

// Calls the progress callback the way ksw_progress_continue does
// (minimap2/ksw2_progress.c:30-38): at anti-diagonals 0, interval, 2*interval,
// ... of `total`; here the k-th call is made on the k-th host poll.
bool progress_continue(const Progress *prog, int &calls, int total)
{
    if (!prog || !prog->cb) return true;
    const int64_t r = std::min<int64_t>((int64_t)calls * prog->interval, std::max(total - 1, 0));
    ++calls;
    return prog->cb(prog->args, (int)r, total) != 0;
}

// Plans run again and again (a caller's resident batch, ksw2_hip_plan_run)
// replay their launches as one captured graph from the second run on.
// Eligible: stripe-kernel pairs only (general, band and wide launches are
// left out: the waiting-launch slots are events shared between plans, which
// a graph cannot wait on), no progress callback and no host mirror (both
// change per run).  KSW2_HIP_GRAPH=0 turns it off.
static bool graph_eligible(const Plan *p, const Progress *prog)
{
    const char *ev = getenv("KSW2_HIP_GRAPH");
    if (ev && atoi(ev) == 0) return false;
    return !(prog && prog->cb) && p->n_gen == 0 && p->n_wide == 0 && p->blaunches.empty() && !p->h_mflag &&
           p->n_fast > 0;
}

static void stripe_args(const Plan *p, const int32_t *poll, StripeArgs &sa)
{
    sa.stop = poll;
    sa.seq = p->d_seq, sa.dir = p->d_dir, sa.bnd = p->d_bnd, sa.cig = p->d_cig, sa.sc = p->sc;
    sa.jobs = p->d_jobs, sa.out = p->d_out, sa.recs = p->d_recs;
    sa.host_out = p->d_mout, sa.host_cig = p->d_mcig, sa.host_flag = p->d_mflag;
    sa.wfirst = nullptr, sa.wctl = p->d_wctl, sa.werr = p->d_werr;
    const char *wr = getenv("KSW2_HIP_WIDE_REDO");  // tests: every wide pair takes the redo path
    sa.force_late = wr && atoi(wr) > 0 ? 1 : 0;
}

// Captures an eligible plan's stripe launches on its stream (forking to the
// lane streams the first run created, joining back) and instantiates them.
// A failed capture leaves the plan on direct launches for good.
// A launch of one-wave teams can run as persistent queue workers
// (KSW2_HIP_QUEUE=1; =2: only when the launch has more pairs than the device
// holds workers).  By default it is a grid of one workgroup per pair: the
// hardware dispatcher hands out the same longest-first list with less work
// per pair (graph replays of the fixed batch, 4,096 pairs: 12.83 against
// 12.97 ms; 16,384 pairs of 300-800 bp: 6.93 against 6.99 ms).
static int queue_mode()
{
    const char *qv = getenv("KSW2_HIP_QUEUE");
    return qv ? std::max(0, std::min(2, atoi(qv))) : 0;
}
static bool use_queue(int mode, const Plan *p, const Plan::Launch &l)
{
    if (mode == 0 || l.team != 1) return false;
    return mode == 1 || l.count > stripe_queue_workers(l.grp, p->R);
}

static void plan_capture(Plan *p)
{
    hipStream_t s = p->stream;
    bool ok = p->fork_ev || hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming) == hipSuccess;
    if (ok) ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (!ok) {
        p->graph_failed = true;
        (void)hipGetLastError();
        return;
    }
    StripeArgs sa;
    stripe_args(p, p->poll_stop ? p->d_stop : nullptr, sa);
    ok = hipEventRecord(p->fork_ev, s) == hipSuccess;
    bool used[Plan::kLanes] = {false, false, false, false};
    // the same launches as a direct run (plan_launch): one-wave teams as
    // persistent queue workers, their heads cleared by a fill kernel node
    const int qmode = queue_mode();
    int nq = 0;
    for (const Plan::Launch &l : p->launches) {
        hipStream_t st = s;
        if (p->own_streams) {
            st = p->lane_stream[l.lane];
            if (!st) ok = false;
            if (ok && !used[l.lane]) ok = hipStreamWaitEvent(st, p->fork_ev, 0) == hipSuccess, used[l.lane] = true;
        }
        sa.order = p->d_order + l.off, sa.njobs = l.count;
        if (ok && use_queue(qmode, p, l) && nq < kMaxQueueHeads) {
            ok = clear_queue_heads(p->d_qhead + nq, 1, st) == hipSuccess &&
                 launch_stripe_queue(sa, l.grp, p->R, p->d_qhead + nq, st) == hipSuccess;
            ++nq;
        } else if (ok) {
            ok = launch_stripe(sa, l.grp, l.team, p->R, st) == hipSuccess;
        }
    }
    for (int k = 0; k < Plan::kLanes; ++k)
        if (used[k] && ok)
            ok = hipEventRecord(p->lane_done[k], p->lane_stream[k]) == hipSuccess &&
                 hipStreamWaitEvent(s, p->lane_done[k], 0) == hipSuccess;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    ok = ok && e == hipSuccess && g;
    if (ok) ok = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    if (!ok) {
        p->graph = nullptr, p->graph_failed = true;
        (void)hipGetLastError();
    }
}

int plan_launch(Plan *p, const Progress *prog)
{
    p->stop_requested = false;
    p->started = false;
    p->progress_calls = 0;
    if (p->n == 0) return KSW2_HIP_OK;
    if (!progress_continue(prog, p->progress_calls, p->max_diags)) {
        // stopped before anti-diagonal 0: nothing runs (ksw2_extd2_avx.c:249-252)
        p->stop_requested = true;
        p->last_dp_ms = p->last_bt_ms = 0.f;
        return KSW2_HIP_OK;
    }
    if (p->h_stop && p->own_stop) *(volatile int32_t *)p->h_stop = 0;
    p->started = true;
    const bool graphed = graph_eligible(p, prog);
    if (graphed && !p->graph && !p->graph_failed && p->runs >= 1) plan_capture(p);
    ++p->runs;
    KSW_HIP_TRY(hipEventRecord(p->ev[0], p->stream));
    if (graphed && p->graph) {
        if (getenv("KSW2_HIP_GRAPH_CHECK")) {
            // tests: clear the results first (fill kernels, outside the
            // graph), so a replay that wrote nothing cannot pass for one
            KSW_HIP_TRY(launch_fill_zero(p->d_out, sizeof(KswOut) * (uint64_t)p->n, p->stream));
            KSW_HIP_TRY(launch_fill_zero(p->d_cig, 4 * p->cig_words, p->stream));
        }
        KSW_HIP_TRY(hipGraphLaunch(p->graph, p->stream));
        ++p->graph_replays;
        KSW_HIP_TRY(hipEventRecord(p->ev[1], p->stream));
        KSW_HIP_TRY(hipEventRecord(p->ev[2], p->stream));
        return KSW2_HIP_OK;
    }
    bool used[Plan::kLanes] = {false, false, false, false};
    auto lane = [&](int l) -> hipStream_t {
        if (!p->own_streams) return p->stream;
        if (!p->lane_stream[l]) {
            // a lane is its stream and its completion event, created together
            hipStream_t s = nullptr;
            hipEvent_t ev = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
                (void)hipStreamDestroy(s);
                return nullptr;
            }
            p->lane_stream[l] = s, p->lane_done[l] = ev;
        }
        if (!used[l]) {
            if (hipStreamWaitEvent(p->lane_stream[l], p->ev[0], 0) != hipSuccess) return nullptr;
            used[l] = true;
        }
        return p->lane_stream[l];
    };
    // general pairs first: they are the long, banded ones
    GenArgs ga;
    ga.seq = p->d_seq, ga.state = p->d_state, ga.dir = p->d_dir, ga.offs = p->d_offs, ga.jobs = p->d_gjobs;
    ga.out = p->d_out, ga.stop = p->d_stop, ga.sc = p->gsc;
    for (const Plan::GLaunch &l : p->glaunches) {
        hipStream_t st = lane(l.lane);
        if (!st) return KSW2_HIP_EDEVICE;
        ga.order = p->d_gorder + l.off, ga.njobs = l.count, ga.ring = l.ring, ga.only_if = nullptr;
        KSW_HIP_TRY(launch_general(ga, l.threads, st));
    }
    if (!p->blaunches.empty()) {
        BandArgs ba;
        ba.stop = (p->poll_stop || (prog && prog->cb)) ? p->d_stop : nullptr;
        ba.seq = p->d_seq, ba.dir = p->d_dir, ba.offs = p->d_offs, ba.bnd = p->d_bnd, ba.jobs = p->d_gjobs;
        ba.out = p->d_out, ba.redo = p->d_redo, ba.sc = p->gsc;
        const char *fr = getenv("KSW2_HIP_BAND_REDO");
        ba.force_redo = fr && atoi(fr) > 0 ? 1 : 0;
        hipStream_t st = lane(0);
        if (!st) return KSW2_HIP_EDEVICE;
        {
            std::lock_guard<std::mutex> lk(g_wait_mu);
            int q[2];
            if (!wait_slots_acquire(st, q)) return KSW2_HIP_EDEVICE;
            for (const Plan::BLaunch &l : p->blaunches) {
                ba.order = p->d_gorder + l.off, ba.njobs = l.count;
                KSW_HIP_TRY(launch_band(ba, l.team, l.grp == 1, l.grp == 2, st));  // chunks of half the device
            }
            if (!wait_slots_release(st, q)) return KSW2_HIP_EDEVICE;
        }
        // pairs whose H went above 0 (the band kernel does not track the
        // maximum) run again on the general kernel, on the same stream
        ga.order = p->d_gorder + p->band_off, ga.njobs = p->n_band, ga.ring = 0, ga.only_if = p->d_redo;
        KSW_HIP_TRY(launch_general(ga, 1024, st));
    }
    // the stripe and band kernels sample the stop flag only when it may be
    // raised: a progress callback runs, or the flag's owner raises it
    const int32_t *poll = (p->poll_stop || (prog && prog->cb)) ? p->d_stop : nullptr;
    StripeArgs sa;
    stripe_args(p, poll, sa);
    if (p->n_wide > 0) {
        // wide launches (the longest pairs) first, within two quarters of
        // the device, then their epilogues, which do not wait
        hipStream_t st = lane(0);
        if (!st) return KSW2_HIP_EDEVICE;
        KSW_HIP_TRY(clear_wide_err(p->d_werr, st));
        bool wide_now = true;
        {
            std::lock_guard<std::mutex> lk(g_wait_mu);
            // spread pairs only: run wide only if two quarters are free now
            const bool no_spread = getenv("KSW2_HIP_SPREAD_FALLBACK") != nullptr;  // tests
            wide_now = !p->optional_wide || (!no_spread && wait_slots_free() >= 2);
            if (wide_now) {
                int q[2];
                if (!wait_slots_acquire(st, q)) return KSW2_HIP_EDEVICE;
                for (const Plan::Launch &l : p->launches) {
                    if (!l.wide) continue;
                    sa.order = p->d_order + l.off, sa.njobs = l.count;
                    sa.wctl = p->d_wctl + (int64_t)l.woff * kWideCtl, sa.wfirst = p->d_wfirst + l.wf;
                    const int half = std::max(1, stripe_wide_capacity(l.grp, p->R) / 2);
                    KSW_HIP_TRY(launch_stripe_wide(sa, l.grp, p->R, p->h_wfirst.data() + l.wf, half, st));
                }
                if (!wait_slots_release(st, q)) return KSW2_HIP_EDEVICE;
            }
        }
        for (const Plan::Launch &l : p->launches) {
            if (!l.wide) continue;
            sa.order = p->d_order + l.off, sa.njobs = l.count, sa.wctl = p->d_wctl + (int64_t)l.woff * kWideCtl;
            if (wide_now) KSW_HIP_TRY(launch_stripe_wide_epilogue(sa, l.grp, p->R, st));
            else KSW_HIP_TRY(launch_stripe(sa, l.grp, l.team, p->R, st));  // one-workgroup teams
        }
    }
    // one-wave teams run as persistent workers over the launch's job list (a
    // device queue, longest first)
    const int qmode = queue_mode();
    int nq = 0;
    for (size_t k = 0; k < p->launches.size(); ++k) {
        const Plan::Launch &l = p->launches[k];
        if (l.wide) continue;
        hipStream_t st = lane(l.lane);
        if (!st) return KSW2_HIP_EDEVICE;
        sa.order = p->d_order + l.off, sa.njobs = l.count;
        if (use_queue(qmode, p, l) && nq < kMaxQueueHeads) {
            // its head zeroed on the launch's own stream, before the workers
            KSW_HIP_TRY(clear_queue_heads(p->d_qhead + nq, 1, st));
            KSW_HIP_TRY(launch_stripe_queue(sa, l.grp, p->R, p->d_qhead + nq, st));
            ++nq;
        } else {
            KSW_HIP_TRY(launch_stripe(sa, l.grp, l.team, p->R, st));
        }
    }
    for (int k = 0; k < Plan::kLanes; ++k)
        if (used[k]) {
            KSW_HIP_TRY(hipEventRecord(p->lane_done[k], p->lane_stream[k]));
            KSW_HIP_TRY(hipStreamWaitEvent(p->stream, p->lane_done[k], 0));
        }
    KSW_HIP_TRY(hipEventRecord(p->ev[1], p->stream));
    if (p->n_gen > 0) {
        GenBacktrackArgs gb;
        gb.dir = p->d_dir, gb.offs = p->d_offs, gb.cig = p->d_cig, gb.jobs = p->d_gjobs, gb.out = p->d_out;
        gb.njobs = p->n_gen;
        // pairs of qlen + tlen >= 512 walk one per wave (KSW2_HIP_TILED_MIN: tests)
        const char *tm = getenv("KSW2_HIP_TILED_MIN");
        gb.tiled_min = tm ? atoi(tm) : 512;
        KSW_HIP_TRY(launch_general_backtrack(gb, p->stream));
        if (p->n_tracks_gen > 0) {
            GenWalkArgs gw;
            gw.seq = p->d_seq, gw.dir = p->d_dir, gw.offs = p->d_offs, gw.cig = p->d_cig;
            gw.jobs = p->d_gjobs, gw.out = p->d_out, gw.sc = p->gsc, gw.njobs = p->n_gen;
            KSW_HIP_TRY(launch_general_walk(gw, p->stream));
        }
    }
    KSW_HIP_TRY(hipEventRecord(p->ev[2], p->stream));
    return KSW2_HIP_OK;
}

bool plan_done(Plan *p)
{
    if (!p->started) return true;
    return hipEventQuery(p->ev[2]) != hipErrorNotReady;
}

int plan_wait(Plan *p, const Progress *prog)
{
    if (!p->started) return KSW2_HIP_OK;
    if (prog && prog->cb) {
        // poll the callback while the device works; a stop raises the flag
        // the general kernel samples every 256 anti-diagonals
        hipError_t q;
        while ((q = hipEventQuery(p->ev[2])) == hipErrorNotReady) {
            if (!p->stop_requested && !progress_continue(prog, p->progress_calls, p->max_diags)) {
                p->stop_requested = true;
                if (p->h_stop) __atomic_store_n(p->h_stop, 1, __ATOMIC_SEQ_CST);
            }
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        if (q != hipSuccess) return KSW2_HIP_EDEVICE;
    }
    KSW_HIP_TRY(hipEventSynchronize(p->ev[2]));
    KSW_HIP_TRY(hipEventElapsedTime(&p->last_dp_ms, p->ev[0], p->ev[1]));
    KSW_HIP_TRY(hipEventElapsedTime(&p->last_bt_ms, p->ev[1], p->ev[2]));
    return KSW2_HIP_OK;
}

int plan_run(Plan *p, const Progress *prog)
{
    const int st = plan_launch(p, prog);
    if (st != KSW2_HIP_OK) return st;
    return plan_wait(p, prog);
}

// One slot's result into the caller's ez (CIGAR grown with krealloc/realloc).
bool plan_dir_span(const Plan *p, int k, uint64_t *off, uint64_t *len)
{
    if (!p->pool || k < 0 || k >= p->n_fast) return false;
    const KswJob &j = p->jobs[k];
    if (j.flag & KSW_EZ_SCORE_ONLY) return false;
    *off = (uint64_t)(p->d_dir - p->pool->base) + j.dir_off;
    *len = (track_pair(j.flag) ? 2 : 1) * dir_bytes(j.qlen, j.tlen, j.R);
    return true;
}

// A fetch whose arena range is larger than this copies only the CIGAR runs
// (a compaction kernel, one more round trip); a smaller range, a single
// call's for instance, is copied as it is.
constexpr uint64_t kDenseFetchMin = 1ull << 20;

// Delivers slot k's result with its CIGAR runs at `runs`.
static int deliver_runs(const Plan *p, int k, const KswOut &o, const uint32_t *runs, ksw_extz_t *z, void *km)
{
    z->max = (uint32_t)o.max, z->zdropped = o.zdropped;
    z->max_q = o.max_q, z->max_t = o.max_t;
    z->mqe = o.mqe, z->mqe_t = o.mqe_t, z->mte = o.mte, z->mte_q = o.mte_q;
    z->score = o.score, z->reach_end = o.reach_end, z->stopped = o.stopped;
    if (p->stop_requested && (k < p->n_fast || p->gjobs[k - p->n_fast].r_end > 0)) {
        // a stripe- or band-kernel pair whose deadline passed while it ran
        // (the kernels end at their next sample of the flag, at most 256
        // columns or anti-diagonals later) is reported stopped with no path;
        // callers discard stopped results (alignSlidingWindow.cpp:957-963)
        z->stopped = 1, z->score = KSW_NEG_INF;
        z->n_cigar = 0;
        return KSW2_HIP_OK;
    }
    if (o.n_cigar < 0) return drop_cigar(z, km);  // failed Singletrack walk
    return deliver_cigar(z, runs, o.n_cigar, km);
}

int deliver_slot(const Plan *p, int k, const KswOut &o, const uint32_t *cig_arena, ksw_extz_t *z, void *km)
{
    return deliver_runs(p, k, o, cig_arena + p->cig_off[k], z, km);
}

int plan_fetch(Plan *p, ksw_extz_t *ez, void *km, const std::vector<char> *skip)
{
    if (p->n == 0) return KSW2_HIP_OK;
    if (p->stop_requested && !p->started) {
        for (int k = 0; k < p->n; ++k)
            if (!skip || !(*skip)[k]) ez[p->order[k]].stopped = 1;  // stopped at r = 0
        return KSW2_HIP_OK;
    }
    int todo = 0;
    for (int k = 0; k < p->n; ++k) todo += !skip || !(*skip)[k];
    if (todo == 0) return KSW2_HIP_OK;  // every pair taken from the host mirror
    auto slot_words = [&](int k) -> uint64_t {
        return (uint64_t)(k < p->n_fast ? p->jobs[k].qlen + p->jobs[k].tlen
                                        : p->gjobs[k - p->n_fast].qlen + p->gjobs[k - p->n_fast].tlen) + 2;
    };
    // the arena's words of the slots still to deliver (pairs already taken
    // from the host mirror are skipped)
    uint64_t lo = p->cig_words, hi = 0;
    for (int k = 0; k < p->n; ++k)
        if (!skip || !(*skip)[k])
            lo = std::min(lo, p->cig_off[k]),
            hi = std::max(hi, std::min<uint64_t>(p->cig_off[k] + slot_words(k), p->cig_words));
    std::vector<KswOut> outs(p->n);
    KSW_HIP_TRY(hipMemcpyAsync(outs.data(), p->d_out, sizeof(KswOut) * p->n, hipMemcpyDeviceToHost, p->stream));
    const uint64_t dense_min = getenv("KSW2_HIP_DENSE_FETCH") ? 0 : kDenseFetchMin;  // tests: always dense
    if (hi > lo && (hi - lo) * 4 > dense_min) {
        // A large range: the runs of the slots to deliver are moved on the
        // device into the dense region behind the arena (the move list at
        // its head), so the host copies only the runs, a few hundred KB for
        // a batch whose arena holds tens of MB.  When the runs and the list
        // would not fit the region, the range is copied as it is.
        KSW_HIP_TRY(hipStreamSynchronize(p->stream));
        std::vector<CigarMove> moves;
        std::vector<uint64_t> at(p->n, 0);
        uint64_t total = 0;
        bool in_slots = true;  // every run count within its slot (the walks never write more)
        for (int k = 0; k < p->n; ++k) {
            if ((skip && (*skip)[k]) || outs[k].n_cigar <= 0) continue;
            const uint64_t n = (uint64_t)outs[k].n_cigar;
            if (n > slot_words(k) || p->cig_off[k] + n > p->cig_words) in_slots = false;
            at[k] = total;
            moves.push_back({p->cig_off[k], (uint32_t)total, (uint32_t)n});
            total += n;
        }
        const uint64_t list_words = 4 * (uint64_t)moves.size();
        if (in_slots && total < (1ull << 32) && list_words + total <= p->cig_words) {
            std::unique_ptr<uint32_t[]> dense(new uint32_t[std::max<uint64_t>(total, 1)]);
            if (total > 0) {
                uint32_t *region = p->d_cig + p->cig_words;
                CigarMove *d_moves = (CigarMove *)region;
                uint32_t *d_dense = region + list_words;
                KSW_HIP_TRY(hipMemcpyAsync(d_moves, moves.data(), sizeof(CigarMove) * moves.size(),
                                           hipMemcpyHostToDevice, p->stream));
                KSW_HIP_TRY(launch_cigar_compact(p->d_cig, d_dense, d_moves, (int)moves.size(), p->stream));
                KSW_HIP_TRY(hipMemcpyAsync(dense.get(), d_dense, total * 4, hipMemcpyDeviceToHost, p->stream));
                KSW_HIP_TRY(hipStreamSynchronize(p->stream));
            }
            for (int k = 0; k < p->n; ++k) {
                if (skip && (*skip)[k]) continue;
                const int st = deliver_runs(p, k, outs[k], dense.get() + at[k], ez + p->order[k], km);
                if (st != KSW2_HIP_OK) return st;
            }
            return KSW2_HIP_OK;
        }
    }
    std::unique_ptr<uint32_t[]> cig(new uint32_t[p->cig_words]);  // only [lo, hi) is filled and read
    if (hi > lo)
        KSW_HIP_TRY(hipMemcpyAsync(cig.get() + lo, p->d_cig + lo, (hi - lo) * 4, hipMemcpyDeviceToHost, p->stream));
    KSW_HIP_TRY(hipStreamSynchronize(p->stream));
    for (int k = 0; k < p->n; ++k) {
        if (skip && (*skip)[k]) continue;
        const int st = deliver_slot(p, k, outs[k], cig.get(), ez + p->order[k], km);
        if (st != KSW2_HIP_OK) return st;
    }
    return KSW2_HIP_OK;
}

void plan_destroy(Plan *p)
{
    if (!p) return;
    // Work a failed plan_launch left queued may still write into the plan's
    // memory and its host mirror: wait for every stream it used before the
    // region goes back to a pool (free once a run has been waited for).
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (int l = 0; l < Plan::kLanes; ++l)
        if (p->lane_stream[l]) (void)hipStreamSynchronize(p->lane_stream[l]);
    if (p->pool) p->pool->free(p->pool_off);
    if (p->block) (void)hipFree(p->block);
    if (p->h_stop && p->own_stop) (void)hipHostFree(p->h_stop);
    if (p->h_seq && p->own_seq) (void)hipHostFree(p->h_seq);
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    if (p->own_events)
        for (auto &e : p->ev)
            if (e) (void)hipEventDestroy(e);
    if (p->own_streams) {
        for (int l = 0; l < Plan::kLanes; ++l) {
            if (p->lane_done[l]) (void)hipEventDestroy(p->lane_done[l]);
            if (p->lane_stream[l]) wait_slots_forget(p->lane_stream[l]), (void)hipStreamDestroy(p->lane_stream[l]);
        }
        if (p->stream) wait_slots_forget(p->stream), (void)hipStreamDestroy(p->stream);
    }
    delete p;
}

namespace {

// What a host thread's single calls reuse from one call to the next: a
// stream, the plan's timing events, pinned staging for the sequences, a
// pinned stop flag and a device block grown to the largest call so far (up
// to kCallBlockMax).  A call then allocates and frees nothing, which matters
// most when many executor threads call at once: hipFree and hipHostFree
// wait for the device.  Freed when the thread ends, or by
// ksw2_hip_release_cached_memory on the thread.
constexpr uint64_t kCallBlockMax = 1ull << 30;

struct GrowPool : DevicePool {
    // one plan at a time; a larger one replaces the block (the previous
    // call's work has finished: every call waits for its plan)
    bool alloc(uint64_t n, uint64_t *off) override
    {
        if (n > bytes) {
            if (base) (void)hipFree(base);
            base = nullptr, bytes = 0;
            if (hipMalloc(&base, n) != hipSuccess) {
                base = nullptr;
                return false;
            }
            bytes = n;
        }
        *off = 0;
        return true;
    }
    void free(uint64_t) override {}
    void drop()
    {
        if (base) (void)hipFree(base);
        base = nullptr, bytes = 0;
    }
};

struct CallContext {
    bool ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    uint8_t *staging = nullptr;
    uint64_t staging_bytes = 0;
    int32_t *h_stop = nullptr, *d_stop = nullptr;
    GrowPool pool;

    bool prepare(uint64_t seq_bytes)
    {
        if (!ready) {
            if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return false;
            for (auto &e : ev)
                if (hipEventCreate(&e) != hipSuccess) return false;
            if (hipHostMalloc((void **)&h_stop, sizeof(int32_t), hipHostMallocMapped) != hipSuccess) return false;
            if (hipHostGetDevicePointer((void **)&d_stop, h_stop, 0) != hipSuccess) return false;
            ready = true;
        }
        if (seq_bytes > staging_bytes) {
            if (staging) (void)hipHostFree(staging);
            staging = nullptr, staging_bytes = 0;
            const uint64_t want = std::max<uint64_t>(seq_bytes + seq_bytes / 4, 1u << 16);
            if (hipHostMalloc((void **)&staging, want, hipHostMallocDefault) != hipSuccess) {
                staging = nullptr;
                return false;
            }
            staging_bytes = want;
        }
        *(volatile int32_t *)h_stop = 0;
        return true;
    }
    ~CallContext()
    {
        pool.drop();
        if (staging) (void)hipHostFree(staging);
        if (h_stop) (void)hipHostFree(h_stop);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) wait_slots_forget(stream), (void)hipStreamDestroy(stream);
    }
};

thread_local CallContext t_call;

}  // namespace

void release_call_context() { t_call.pool.drop(); }

int run_batch_host(int n, const ksw2_pair_t *pairs, int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2,
                   int8_t e2, ksw_extz_t *ez, void *km, const Progress *prog)
{
    for (int k = 0; k < n; ++k) reset_ez(ez + k);
    if (n == 0 || m <= 1) return KSW2_HIP_OK;
    // the length overflow test comes before the piece swap and the
    // -min_sc > 2(q+e) early return (minimap2/ksw2_extd2_avx.c:136-140)
    for (int k = 0; k < n; ++k)
        if (pairs[k].qlen > 0 && pairs[k].tlen > 0 && pairs[k].qlen > 0x7fffffff - pairs[k].tlen + 1) ez[k].zdropped = 1;
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return KSW2_HIP_OK;  // reference returns early
    if (m > kMaxM) {
        for (int k = 0; k < n; ++k) ez[k].zdropped = 1;
        return KSW2_HIP_EUNSUPPORTED;
    }
    GenScore gsc;
    make_gen_score(m, mat, sc, cls, &gsc);
    int status = KSW2_HIP_OK;
    std::vector<int> todo;
    for (int k = 0; k < n; ++k) {
        const ksw2_pair_t &p = pairs[k];
        if (p.qlen <= 0 || p.tlen <= 0) continue;            // reference returns early
        if (!p.query || !p.target) return KSW2_HIP_EINVAL;
        if (p.qlen > 0x7fffffff - p.tlen + 1) {               // ksw2_extd2_avx.c:138-141
            ez[k].zdropped = 1;
            continue;
        }
        todo.push_back(k);
    }
    // a batch that fits the free memory runs as one plan (its own block)
    uint64_t total = Engine::kPlanSlack;
    for (int k : todo) total += pair_footprint(pairs[k], takes_general(pairs[k], m, cls));
    // (a batch of at most 1 GiB, a single call's usual size, skips the
    // device memory query)
    // A plan larger than a single call's block is carved from the process's
    // cached device block when that holds it (allocating and freeing
    // hundreds of GB per call costs seconds: the driver clears freed memory
    // before it hands it out again).
    const bool small = todo.size() <= 1 || (total <= (1ull << 30) && !getenv("KSW2_HIP_MEM_BUDGET"));
    std::unique_ptr<DevicePool> big;
    if (!small && total > kCallBlockMax && total <= mem_budget(true)) big = cached_block_pool(mem_budget(true), total);
    if (small || big || total <= mem_budget(false)) {
        // one plan holds the whole batch; a single pair up to kCallBlockMax
        // runs in the thread's cached context
        if (todo.empty()) return status;
        Plan *plan = nullptr;
        PlanEnv env;
        const PlanEnv *use = nullptr;
        if (big) {
            env.pool = big.get();
            env.spread = true;  // a plan of its own
            use = &env;
        }
        if (todo.size() == 1 && total <= kCallBlockMax) {
            const ksw2_pair_t &pp = pairs[todo[0]];
            if (t_call.prepare((uint64_t)pp.qlen + pp.tlen + 64)) {
                env.pool = &t_call.pool;
                env.stream = t_call.stream;
                env.events = t_call.ev;
                env.staging = t_call.staging, env.staging_bytes = t_call.staging_bytes;
                env.h_stop = t_call.h_stop, env.d_stop = t_call.d_stop;
                env.spread = true;  // a plan of its own: 5-16 wave teams may spread over CUs
                use = &env;
            }
        }
        int st = plan_create(&plan, pairs, todo.data(), (int)todo.size(), m, cls, gsc, use);
        if (st == KSW2_HIP_OK) st = plan_run(plan, prog);
        if (st == KSW2_HIP_OK) st = plan_fetch(plan, ez, km);
        plan_destroy(plan);
        if (st != KSW2_HIP_OK) {
            for (int k : todo) ez[k].zdropped = 1;  // the device path did not run
            return st;
        }
        return status;
    }
    // larger than the budget: stream it through the engine, longest first
    // (CIGARs grown with realloc, as with km == NULL)
    Engine eng(m, mat, sc, cls, mem_budget(true));
    int st = eng.start(true);  // the cached block when it is free and large enough
    if (st != KSW2_HIP_OK) {
        for (int k : todo) ez[k].zdropped = 1;
        return st;
    }
    eng.hold(true);
    for (int k : todo) eng.submit(pairs[k], ez + k, false, nullptr);
    eng.order_longest_first();
    eng.hold(false);
    st = eng.drain(prog);
    if (st != KSW2_HIP_OK) status = st;
    return status;
}

}  // namespace kswhip
