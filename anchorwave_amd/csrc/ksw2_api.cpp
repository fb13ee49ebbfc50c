// ksw2_api.cpp -- host side of the C ABI declared in include/ksw2_hip.h.
//
// Responsibilities: normalise the scoring exactly as the reference does
// (minimap2/ksw2_extd2_sse.c:80-110), reproduce its early returns, decide per
// pair whether the unbanded fast kernel applies, stage sequences into HBM,
// launch the DP and backtrack kernels, and hand the CIGAR back through
// realloc/krealloc with the reference's m_cigar growth (minimap2/ksw2.h:144-154).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_common.h"
#include "ksw2_engine.h"
#include "ksw2_kernels.h"
#include "ksw2_plan.h"

using namespace kswhip;

// kalloc's krealloc when the host program links it (AnchorWave does); with
// km == NULL kalloc itself falls back to realloc (minimap2/kalloc.c:167-175).
extern "C" void *krealloc(void *km, void *ptr, size_t size) __attribute__((weak));
extern "C" void kfree(void *km, void *ptr) __attribute__((weak));

namespace {

thread_local int t_last_error = KSW2_HIP_OK;
thread_local ksw2_hip_progress_cb t_cb = nullptr;
thread_local void *t_cb_args = nullptr;
thread_local int t_cb_interval = 256;

kswhip::Progress current_progress()
{
    kswhip::Progress p;
    p.cb = t_cb, p.args = t_cb_args, p.interval = t_cb_interval;
    return p;
}

int set_err(int s)
{
    t_last_error = s;
    return s;
}

void reset_extz(ksw_extz_t *ez)
{
    ez->max_q = ez->max_t = ez->mqe_t = ez->mte_q = -1;
    ez->max = 0, ez->score = ez->mqe = ez->mte = KSW_NEG_INF;
    ez->n_cigar = 0, ez->zdropped = 0, ez->reach_end = 0, ez->stopped = 0;
}

}  // namespace

namespace kswhip {

// Normalised scoring + the reference's early-return test.  Returns false when
// the reference returns right after ksw_reset_extz (ksw2_extd2_sse.c:81,105).
bool normalise_scoring(int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                       KswScore *sc, ScoreClass *cls)
{
    const int qe_orig = q + e;
    if (q2 + e2 < q + e) std::swap(q, q2), std::swap(e, e2);
    int max_sc = mat[0], min_sc = mat[1];
    for (int t = 1; t < m * m; ++t) max_sc = std::max<int>(max_sc, mat[t]), min_sc = std::min<int>(min_sc, mat[t]);
    if (-min_sc > 2 * (q + e)) return false;
    int long_thres = e != e2 ? (q2 - q) / (e - e2) - 1 : 0;
    if (q2 + e2 + long_thres * e2 > q + e + long_thres * e) ++long_thres;
    const int long_diff = long_thres * (e - e2) - (q2 - q) - e2;
    sc->q = q, sc->e = e, sc->q2 = q2, sc->e2 = e2;
    sc->G1 = q + e, sc->G2 = q2 + e2;
    sc->qe_orig = qe_orig;
    sc->long_thres = long_thres, sc->long_diff = long_diff;
    sc->mch = mat[0], sc->mis = mat[1];

    // Score class.  generic_mm: with KSW_EZ_GENERIC_SC the matrix is constant
    // on the diagonal and off it, so s = (t == q) ? mat[0] : mat[1].
    bool diag_const = true;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            if (mat[a * m + b] != (a == b ? mat[0] : mat[1])) diag_const = false;
    cls->generic_mm = diag_const;
    cls->m = m;
    cls->sc_N = mat[m * m - 1] == 0 ? -e2 : mat[m * m - 1];
    cls->sc_N = (int)(int8_t)(uint8_t)(cls->sc_N & 0xff);

    // Exactness of the fast kernel's int32 arithmetic.  With e > e2 the
    // boundary values (edge_diff) are exactly the two-piece gap cost
    // min(q+l*e, q2+l*e2), so the recurrence is the exact DP; with e <= e2 the
    // reference's boundary switches pieces at the wrong length and z can
    // exceed mat[0].  Every in-band intermediate of the int8 recurrence then
    // lies in [-128, 127] when the match score is 0, mismatch <= 0,
    // penalties are non-negative and
    //   2*G1 <= 128, G1+G2 <= 128, G1+e2 <= 128, q2-mis <= 128, 2*G1-q <= 128;
    // the bounds follow from u,v in [-G1, G1], x in [-G1,-e], x2 in [-G2,-e2],
    // z in [mis, 0] (tests/test_oracle.py checks them against the
    // instrumented restatement).
    const int G1 = sc->G1, G2 = sc->G2, mis = mat[1];
    bool ok = q >= 0 && e > e2 && q2 >= 0 && e2 >= 0 && mat[0] == 0 && mis <= 0;
    ok = ok && 2 * G1 <= 128 && G1 + G2 <= 128 && G1 + e2 <= 128 && q2 - mis <= 128 && 2 * G1 - q <= 128;
    ok = ok && G2 <= 127 && long_diff >= -128 && long_diff <= 127;
    cls->fast_ok = ok;
    return true;
}

}  // namespace kswhip

// ------------------------------------------------------------------ C ABI
extern "C" {

int ksw2_hip_last_error(void) { return t_last_error; }

const char *ksw2_hip_error_string(int status)
{
    switch (status) {
    case KSW2_HIP_OK: return "ok";
    case KSW2_HIP_EINVAL: return "invalid argument";
    case KSW2_HIP_EUNSUPPORTED: return "parameters not supported by the device path";
    case KSW2_HIP_ENOMEM: return "out of memory";
    case KSW2_HIP_EDEVICE: return "HIP runtime error";
    case KSW2_HIP_ENODEV: return "no usable gfx950 device";
    case KSW2_HIP_EAGAIN: return "deferred: device memory committed";
    case KSW2_HIP_ETOOBIG: return "larger than the no-allocation scratch (KSW2_HIP_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown status";
    }
}

int ksw2_hip_init(int device)
{
    int st = kswhip::device_init(device);
    if (st == KSW2_HIP_OK) st = kswhip::noalloc_prepare();  // the no-allocation entry point's set-up
    return set_err(st);
}

int ksw2_hip_route(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                   int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int flag)
{
    if (m <= 1 || qlen <= 0 || tlen <= 0 || !mat) return -1;
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return -1;
    ksw2_pair_t p;
    p.query = query, p.target = target, p.qlen = qlen, p.tlen = tlen;
    p.w = w, p.zdrop = zdrop, p.end_bonus = 0, p.flag = flag;
    if (!query || !target) return -1;
    return kswhip::fast_path_reject(p, m, cls);
}

void ksw2_hip_set_progress_callback(ksw2_hip_progress_cb callback, void *arguments, int antidiagonal_interval)
{
    t_cb = callback, t_cb_args = arguments;
    t_cb_interval = antidiagonal_interval > 0 ? antidiagonal_interval : 1;
}

void ksw2_hip_clear_progress_callback(void)
{
    t_cb = nullptr, t_cb_args = nullptr, t_cb_interval = 256;
}

int ksw2_hip_release_cached_memory(void)
{
    kswhip::Engine::release_cached_block();
    kswhip::release_call_context();
    return KSW2_HIP_OK;
}

int ksw_extd2_hip_batch(int n, const ksw2_pair_t *pairs, int8_t m, const int8_t *mat, int8_t q, int8_t e,
                        int8_t q2, int8_t e2, ksw_extz_t *ez)
{
    if (n < 0 || (n > 0 && (!pairs || !mat || !ez))) return set_err(KSW2_HIP_EINVAL);
    const kswhip::Progress prog = current_progress();
    int st = kswhip::run_batch_host(n, pairs, m, mat, q, e, q2, e2, ez, nullptr, &prog);
    return set_err(st);
}

void ksw_extd2_hip(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                   const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus,
                   int flag, ksw_extz_t *ez)
{
    if (!ez) { set_err(KSW2_HIP_EINVAL); return; }
    ksw2_pair_t p;
    p.query = query, p.target = target, p.qlen = qlen, p.tlen = tlen;
    p.w = w, p.zdrop = zdrop, p.end_bonus = end_bonus, p.flag = flag;
    const kswhip::Progress prog = current_progress();
    int st = kswhip::run_batch_host(1, &p, m, mat, q, e, q2, e2, ez, km, &prog);
    set_err(st);
}

int ksw_extd2_hip_noalloc(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                          const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop,
                          int end_bonus, int flag, ksw_extz_t *ez)
{
    if (!ez) return set_err(KSW2_HIP_EINVAL);
    reset_extz(ez);
    // the early returns and the length test in run_batch_host's order
    if (m <= 1) return set_err(KSW2_HIP_OK);
    const bool overflow = qlen > 0 && tlen > 0 && qlen > 0x7fffffff - tlen + 1;
    if (overflow) ez->zdropped = 1;
    if (!mat) return set_err(KSW2_HIP_EINVAL);
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return set_err(KSW2_HIP_OK);
    if (m > kMaxM) {
        ez->zdropped = 1;
        return set_err(KSW2_HIP_EUNSUPPORTED);
    }
    if (qlen <= 0 || tlen <= 0 || overflow) return set_err(KSW2_HIP_OK);
    if (!query || !target) return set_err(KSW2_HIP_EINVAL);
    if (!(flag & KSW_EZ_SCORE_ONLY) && (!ez->cigar || (int64_t)ez->m_cigar < (int64_t)qlen + tlen))
        return set_err(KSW2_HIP_EINVAL);  // the caller's CIGAR buffer is never grown
    ksw2_pair_t p;
    p.query = query, p.target = target, p.qlen = qlen, p.tlen = tlen;
    p.w = w, p.zdrop = zdrop, p.end_bonus = end_bonus, p.flag = flag;
    if (!kswhip::noalloc_fits(p, m, cls)) {  // rejected before any device work
        ez->zdropped = 1;
        return set_err(KSW2_HIP_ETOOBIG);
    }
    int st = kswhip::device_init(-1);
    if (st != KSW2_HIP_OK) {
        ez->zdropped = 1;
        return set_err(st);
    }
    GenScore gsc;
    kswhip::make_gen_score(m, mat, sc, cls, &gsc);
    st = kswhip::noalloc_run(p, m, cls, gsc, ez);
    if (st != KSW2_HIP_OK) {
        reset_extz(ez);
        ez->zdropped = 1;  // the device path did not run
    }
    return set_err(st);
}

void ksw_extd2_singletrack_hip(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                               int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w,
                               int zdrop, int end_bonus, int flag, ksw_extz_t *ez)
{
    ksw_extd2_hip(km, qlen, query, tlen, target, m, mat, q, e, q2, e2, w, zdrop, end_bonus,
                  flag | KSW_EZ_SINGLETRACK, ez);
}

struct ksw2_hip_plan {
    kswhip::Plan *p;
    double cells;
};

int ksw2_hip_plan_create(ksw2_hip_plan **plan, int n, const ksw2_pair_t *pairs, int8_t m, const int8_t *mat,
                         int8_t q, int8_t e, int8_t q2, int8_t e2)
{
    if (!plan || n < 0 || (n > 0 && (!pairs || !mat))) return set_err(KSW2_HIP_EINVAL);
    *plan = nullptr;
    if (m <= 1 || m > kMaxM) return set_err(KSW2_HIP_EUNSUPPORTED);
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return set_err(KSW2_HIP_EUNSUPPORTED);
    GenScore gsc;
    kswhip::make_gen_score(m, mat, sc, cls, &gsc);
    std::vector<int> idx(n);
    double cells = 0;
    for (int k = 0; k < n; ++k) {
        if (pairs[k].qlen <= 0 || pairs[k].tlen <= 0 || !pairs[k].query || !pairs[k].target)
            return set_err(KSW2_HIP_EINVAL);
        if (pairs[k].qlen > 0x7fffffff - pairs[k].tlen + 1) return set_err(KSW2_HIP_EINVAL);

        idx[k] = k;
        cells += (double)pairs[k].qlen * pairs[k].tlen;
    }
    kswhip::Plan *p = nullptr;
    int st = kswhip::plan_create(&p, pairs, idx.data(), n, m, cls, gsc);
    if (st != KSW2_HIP_OK) return set_err(st);
    *plan = new ksw2_hip_plan{p, cells};
    return set_err(KSW2_HIP_OK);
}

int ksw2_hip_plan_run(ksw2_hip_plan *plan)
{
    if (!plan) return set_err(KSW2_HIP_EINVAL);
    return set_err(kswhip::plan_run(plan->p));
}

int ksw2_hip_plan_fetch(ksw2_hip_plan *plan, ksw_extz_t *ez)
{
    if (!plan || (!ez && plan->p->n > 0)) return set_err(KSW2_HIP_EINVAL);
    for (int k = 0; k < plan->p->n; ++k) kswhip::reset_ez(ez + k);
    return set_err(kswhip::plan_fetch(plan->p, ez, nullptr));
}

int ksw2_hip_plan_timing(const ksw2_hip_plan *plan, float *dp_ms, float *backtrack_ms, double *cells)
{
    if (!plan) return set_err(KSW2_HIP_EINVAL);
    if (dp_ms) *dp_ms = plan->p->last_dp_ms;
    if (backtrack_ms) *backtrack_ms = plan->p->last_bt_ms;
    if (cells) *cells = plan->cells;
    return set_err(KSW2_HIP_OK);
}

int ksw2_hip_plan_graph(const ksw2_hip_plan *plan, int *runs, int *replays)
{
    if (!plan) return set_err(KSW2_HIP_EINVAL);
    if (runs) *runs = plan->p->runs;
    if (replays) *replays = plan->p->graph_replays;
    return set_err(KSW2_HIP_OK);
}

int ksw2_hip_plan_dirs(const ksw2_hip_plan *plan, int index, uint8_t *out, size_t bytes, int *R, int *transposed)
{
    if (!plan || index < 0 || index >= plan->p->n) return set_err(KSW2_HIP_EINVAL);
    const kswhip::Plan *p = plan->p;
    int k = -1;
    for (int x = 0; x < p->n; ++x)
        if (p->order[x] == index) k = x;
    if (k < 0 || k >= p->n_fast) return set_err(KSW2_HIP_EINVAL);  // general-kernel pair: other layout
    const kswhip::KswJob &j = p->jobs[k];
    if (R) *R = j.R;
    if (transposed) *transposed = j.transposed;
    if (!out) return set_err(KSW2_HIP_OK);  // the geometry only
    const uint64_t need = (track_pair(j.flag) ? 2 : 1) * kswhip::dir_bytes(j.qlen, j.tlen, j.R);
    if (bytes < need || (j.flag & KSW_EZ_SCORE_ONLY)) return set_err(KSW2_HIP_EINVAL);
    if (hipMemcpy(out, p->d_dir + j.dir_off, need, hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(KSW2_HIP_EDEVICE);
    return set_err(KSW2_HIP_OK);
}

void ksw2_hip_plan_destroy(ksw2_hip_plan *plan)
{
    if (!plan) return;
    kswhip::plan_destroy(plan->p);
    delete plan;
}

struct ksw2_hip_server {
    kswhip::Engine *e;
};

int ksw2_hip_server_create(ksw2_hip_server **server, int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2,
                           int8_t e2, size_t device_bytes)
{
    if (!server || !mat) return set_err(KSW2_HIP_EINVAL);
    *server = nullptr;
    if (m <= 1 || m > kMaxM) return set_err(KSW2_HIP_EUNSUPPORTED);
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return set_err(KSW2_HIP_EUNSUPPORTED);
    int st = kswhip::device_init(-1);
    if (st != KSW2_HIP_OK) return set_err(st);
    // device_bytes == 0: 85 % of the free memory, or the library's cached
    // engine block when that is larger (the server then holds it)
    const uint64_t budget = device_bytes ? (uint64_t)device_bytes : kswhip::mem_budget(true);
    auto *eng = new kswhip::Engine(m, mat, sc, cls, budget);
    st = eng->start(device_bytes == 0, true);
    if (st != KSW2_HIP_OK) {
        delete eng;
        return set_err(st);
    }
    *server = new ksw2_hip_server{eng};
    return set_err(KSW2_HIP_OK);
}

int ksw2_hip_server_submit(ksw2_hip_server *server, const ksw2_pair_t *pair, ksw_extz_t *ez, uint64_t *ticket)
{
    if (!server || !pair || !ez) return set_err(KSW2_HIP_EINVAL);
    return set_err(server->e->submit(*pair, ez, true, ticket));
}

int ksw2_hip_server_try_submit(ksw2_hip_server *server, const ksw2_pair_t *pair, ksw_extz_t *ez, uint64_t *ticket,
                               int *retry_ms)
{
    if (!server || !pair || !ez || !retry_ms) return set_err(KSW2_HIP_EINVAL);
    *retry_ms = 0;
    return set_err(server->e->submit(*pair, ez, true, ticket, retry_ms));
}

int ksw2_hip_server_load(ksw2_hip_server *server, ksw2_hip_load_t *load)
{
    if (!server || !load) return set_err(KSW2_HIP_EINVAL);
    const kswhip::Engine::Load l = server->e->load();
    load->budget_bytes = l.budget, load->held_bytes = l.held_bytes, load->queued_bytes = l.queued_bytes;
    load->pending = l.pending, load->pending_cells = l.pending_cells, load->cells_per_s = l.cells_per_s;
    load->seq_arena_bytes = l.seq_arena_bytes, load->resident_tasks = l.resident_tasks;
    load->staged_flights = l.staged_flights;
    return set_err(KSW2_HIP_OK);
}

uint64_t ksw2_hip_footprint(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                            const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int flag)
{
    if (m <= 1 || qlen <= 0 || tlen <= 0 || !mat) return 0;
    if (qlen > 0x7fffffff - tlen + 1) return UINT64_MAX;  // tested before the scoring (ksw2_extd2_avx.c:136-140)
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return 0;  // the reference returns early
    ksw2_pair_t p;
    p.query = query, p.target = target, p.qlen = qlen, p.tlen = tlen;
    p.w = w, p.zdrop = zdrop, p.end_bonus = 0, p.flag = flag;
    if (!query || !target) return 0;
    return kswhip::pair_footprint(p, kswhip::takes_general(p, m, cls));
}

int ksw2_hip_server_poll(ksw2_hip_server *server, uint64_t *tickets, double *latency_s, int max_tickets,
                         int wait_ms)
{
    if (!server || max_tickets < 0 || (max_tickets > 0 && !tickets)) return set_err(KSW2_HIP_EINVAL);
    const int n = server->e->poll(tickets, latency_s, max_tickets, wait_ms);
    set_err(server->e->status());
    return n;
}

int ksw2_hip_server_poll_status(ksw2_hip_server *server, uint64_t *tickets, double *latency_s, int32_t *status,
                                int max_tickets, int wait_ms)
{
    if (!server || max_tickets < 0 || (max_tickets > 0 && (!tickets || !status))) return set_err(KSW2_HIP_EINVAL);
    const int n = server->e->poll(tickets, latency_s, max_tickets, wait_ms, status);
    set_err(server->e->status());
    return n;
}

void ksw2_hip_server_destroy(ksw2_hip_server *server)
{
    if (!server) return;
    delete server->e;
    delete server;
}

}  // extern "C"

namespace kswhip {

// Grows ez->cigar to hold n entries with the reference's policy (capacity
// 4, then doubling from the caller's m_cigar) and copies them in.
int deliver_cigar(ksw_extz_t *ez, const uint32_t *src, int n, void *km)
{
    int m = ez->m_cigar;
    if (n > m) {
        while (m < n) m = m ? m << 1 : 4;
        void *p;
        if (km && krealloc) p = krealloc(km, ez->cigar, (size_t)m << 2);
        else if (km) return KSW2_HIP_EUNSUPPORTED;  // arena without kalloc linked
        else p = realloc(ez->cigar, (size_t)m << 2);
        if (!p) return KSW2_HIP_ENOMEM;
        ez->cigar = (uint32_t *)p;
        ez->m_cigar = m;
    }
    if (n > 0) memcpy(ez->cigar, src, (size_t)n * sizeof(uint32_t));
    ez->n_cigar = n;
    return KSW2_HIP_OK;
}

// A failed Singletrack walk: the reference frees the CIGAR buffer and leaves
// none (minimap2/ksw2_singletrack.cpp:160-165); the caller sees zdropped.
int drop_cigar(ksw_extz_t *ez, void *km)
{
    if (ez->cigar) {
        if (km && kfree) kfree(km, ez->cigar);
        else if (km) return KSW2_HIP_EUNSUPPORTED;  // arena without kalloc linked
        else free(ez->cigar);
    }
    ez->cigar = nullptr, ez->m_cigar = 0, ez->n_cigar = 0;
    ez->zdropped = 1;
    return KSW2_HIP_OK;
}

void reset_ez(ksw_extz_t *ez) { reset_extz(ez); }

}  // namespace kswhip
