// ksw2_batch.cpp -- batches of pairs on one device: eligibility, staging,
// launch, read-back.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_common.h"
#include "ksw2_kernels.h"
#include "ksw2_plan.h"

namespace kswhip {

namespace {

std::mutex g_init_mu;
int g_device = -1;

#define KSW_HIP_TRY(x)                                  \
    do {                                                \
        if ((x) != hipSuccess) return KSW2_HIP_EDEVICE; \
    } while (0)

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

uint64_t mem_budget()
{
    const char *env = getenv("KSW2_HIP_MEM_BUDGET");
    if (env) return strtoull(env, nullptr, 10);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 8ull << 30;
    return (uint64_t)(fr * 0.85);
}

}  // namespace

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
    if (p.flag & (KSW_EZ_APPROX_MAX | KSW_EZ_SINGLETRACK)) why |= kRejFlags;
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
// (the time the batch needs when every wave is busy), a pair whose stripes
// take S * (qlen + 63) steps on one wave gets the smallest power-of-two team
// that brings it under T, capped by its stripe count and kMaxTeam: small
// pairs run on one wave with no synchronisation, long pairs do not outlast
// the batch.  KSW2_HIP_TEAM forces a size (tuning and tests).
static int choose_team(const KswJob &j, int R, double T, int forced)
{
    const int SR = kWave * R;
    const int S = (j.tlen + SR - 1) / SR;
    if (forced > 0) return std::max(1, std::min(std::min(forced, kMaxTeam), S));
    const double steps = (double)S * (j.qlen + kWave - 1);
    int team = 1;
    while (team < kMaxTeam && team < S && steps / team > T) team <<= 1;
    return std::min(team, S);
}

static int lane_of_team(int team) { return team >= 8 ? 0 : team >= 4 ? 1 : team >= 2 ? 2 : 3; }

int plan_create(Plan **out, int n, const ksw2_pair_t *pairs, const int *idx, int nidx, const KswScore &sc)
{
    (void)n;
    int st = device_init(-1);
    if (st != KSW2_HIP_OK) return st;
    Plan *p = new Plan();
    p->device = g_device;
    p->sc = sc;
    p->R = kRowsPerLane;
    const int R = p->R, SR = kWave * R;
    // job groups: left-aligned traceback, right-aligned, score-only
    std::vector<int> g[3];
    for (int k = 0; k < nidx; ++k) {
        const ksw2_pair_t &pp = pairs[idx[k]];
        const int grp = (pp.flag & KSW_EZ_SCORE_ONLY) ? 2 : (pp.flag & KSW_EZ_RIGHT) ? 1 : 0;
        g[grp].push_back(idx[k]);
    }
    std::vector<int> grp_of;
    for (int grp = 0; grp < 3; ++grp)
        for (int id : g[grp]) p->order.push_back(id), grp_of.push_back(grp);
    p->n = (int)p->order.size();
    p->jobs.resize(p->n);
    uint64_t so = 0, dof = 0, bo = 0, co = 0;
    int stripes = 0;
    for (int k = 0; k < p->n; ++k) {
        const ksw2_pair_t &pp = pairs[p->order[k]];
        KswJob &j = p->jobs[k];
        j.qlen = pp.qlen, j.tlen = pp.tlen, j.flag = pp.flag, j.end_bonus = pp.end_bonus;
        j.R = R;
        j.q_off = so, so = round_up(so + pp.qlen, 16);
        j.t_off = so, so = round_up(so + pp.tlen, 16);
        j.dir_off = dof;
        if (!(pp.flag & KSW_EZ_SCORE_ONLY)) dof = round_up(dof + dir_bytes(pp.qlen, pp.tlen, R), 256);
        j.bnd_off = bo, bo += round_up(2ull * pp.qlen, 64);
        j.cig_off = co, co += round_up((uint64_t)pp.qlen + pp.tlen + 2, 16);
        j.stripe_off = stripes, stripes += (pp.tlen + SR - 1) / SR;
    }
    p->nstripes_total = stripes;
    // launches: per group, per team size, longest pairs first
    std::vector<int32_t> order;
    {
        double total_steps = 0;
        for (const KswJob &j : p->jobs) total_steps += (double)((j.tlen + SR - 1) / SR) * (j.qlen + kWave - 1);
        const double T = std::max(1.0, total_steps / std::max(1, stripe_resident_waves(R, false, false)));
        int forced = 0;
        if (const char *ev = getenv("KSW2_HIP_TEAM")) forced = atoi(ev);
        std::vector<int> team(p->n);
        for (int k = 0; k < p->n; ++k) team[k] = choose_team(p->jobs[k], R, T, forced);
        for (int grp = 0; grp < 3; ++grp)
            for (int t = kMaxTeam; t >= 1; --t) {
                std::vector<int> members;
                for (int k = 0; k < p->n; ++k)
                    if (grp_of[k] == grp && team[k] == t) members.push_back(k);
                if (members.empty()) continue;
                std::stable_sort(members.begin(), members.end(), [&](int x, int y) {
                    const KswJob &a = p->jobs[x], &b = p->jobs[y];
                    return (double)a.tlen * (a.qlen + kWave - 1) > (double)b.tlen * (b.qlen + kWave - 1);
                });
                p->launches.push_back({grp, t, (int)order.size(), (int)members.size(), lane_of_team(t)});
                order.insert(order.end(), members.begin(), members.end());
            }
    }
    p->seq_bytes = std::max<uint64_t>(so, 16), p->dir_bytes = std::max<uint64_t>(dof, 256);
    p->bnd_dwords = std::max<uint64_t>(bo, 64), p->cig_words = std::max<uint64_t>(co, 16);
    auto fail = [&](int code) {
        plan_destroy(p);
        return code;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    for (auto &e : p->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    for (int l = 0; l < Plan::kLanes; ++l) {
        if (hipStreamCreateWithFlags(&p->lane_stream[l], hipStreamNonBlocking) != hipSuccess)
            return fail(KSW2_HIP_EDEVICE);
        if (hipEventCreateWithFlags(&p->lane_done[l], hipEventDisableTiming) != hipSuccess)
            return fail(KSW2_HIP_EDEVICE);
    }
    if (hipMalloc(&p->d_seq, p->seq_bytes) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_dir, p->dir_bytes) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_bnd, p->bnd_dwords * 4) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_cig, p->cig_words * 4) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_jobs, sizeof(KswJob) * std::max(p->n, 1)) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_out, sizeof(KswOut) * std::max(p->n, 1)) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_order, sizeof(int32_t) * std::max<size_t>(order.size(), 1)) != hipSuccess)
        return fail(KSW2_HIP_ENOMEM);
    if (hipMalloc(&p->d_recs, sizeof(StripeRec) * std::max(stripes, 1)) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    // stage sequences through pinned memory
    uint8_t *h_seq = nullptr;
    if (hipHostMalloc((void **)&h_seq, p->seq_bytes, hipHostMallocDefault) != hipSuccess) return fail(KSW2_HIP_ENOMEM);
    memset(h_seq, 0, p->seq_bytes);
    for (int k = 0; k < p->n; ++k) {
        const ksw2_pair_t &pp = pairs[p->order[k]];
        memcpy(h_seq + p->jobs[k].q_off, pp.query, (size_t)pp.qlen);
        memcpy(h_seq + p->jobs[k].t_off, pp.target, (size_t)pp.tlen);
    }
    hipError_t e1 = hipMemcpyAsync(p->d_seq, h_seq, p->seq_bytes, hipMemcpyHostToDevice, p->stream);
    hipError_t e2 = hipMemcpyAsync(p->d_jobs, p->jobs.data(), sizeof(KswJob) * p->n, hipMemcpyHostToDevice, p->stream);
    hipError_t e4 = order.empty() ? hipSuccess
                                  : hipMemcpyAsync(p->d_order, order.data(), sizeof(int32_t) * order.size(),
                                                   hipMemcpyHostToDevice, p->stream);
    hipError_t e3 = hipStreamSynchronize(p->stream);
    (void)hipHostFree(h_seq);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) return fail(KSW2_HIP_EDEVICE);
    *out = p;
    return KSW2_HIP_OK;
}

int plan_run(Plan *p)
{
    if (p->n == 0) return KSW2_HIP_OK;
    KSW_HIP_TRY(hipEventRecord(p->ev[0], p->stream));
    StripeArgs sa;
    sa.seq = p->d_seq, sa.dir = p->d_dir, sa.bnd = p->d_bnd, sa.sc = p->sc;
    sa.jobs = p->d_jobs, sa.out = p->d_out, sa.recs = p->d_recs;
    bool used[Plan::kLanes] = {false, false, false, false};
    for (const Plan::Launch &l : p->launches) {
        hipStream_t st = p->lane_stream[l.lane];
        if (!used[l.lane]) {
            KSW_HIP_TRY(hipStreamWaitEvent(st, p->ev[0], 0));
            used[l.lane] = true;
        }
        sa.order = p->d_order + l.off, sa.njobs = l.count;
        KSW_HIP_TRY(launch_stripe(sa, p->R, l.grp == 1, l.grp == 2, l.team, st));
    }
    for (int k = 0; k < Plan::kLanes; ++k)
        if (used[k]) {
            KSW_HIP_TRY(hipEventRecord(p->lane_done[k], p->lane_stream[k]));
            KSW_HIP_TRY(hipStreamWaitEvent(p->stream, p->lane_done[k], 0));
        }
    KSW_HIP_TRY(hipEventRecord(p->ev[1], p->stream));
    BacktrackArgs ba;
    ba.dir = p->d_dir, ba.cig = p->d_cig, ba.jobs = p->d_jobs, ba.out = p->d_out, ba.recs = p->d_recs;
    ba.njobs = p->n, ba.sc = p->sc;
    KSW_HIP_TRY(launch_backtrack(ba, p->stream));
    KSW_HIP_TRY(hipEventRecord(p->ev[2], p->stream));
    KSW_HIP_TRY(hipStreamSynchronize(p->stream));
    KSW_HIP_TRY(hipEventElapsedTime(&p->last_dp_ms, p->ev[0], p->ev[1]));
    KSW_HIP_TRY(hipEventElapsedTime(&p->last_bt_ms, p->ev[1], p->ev[2]));
    return KSW2_HIP_OK;
}

int plan_fetch(Plan *p, ksw_extz_t *ez, void *km)
{
    if (p->n == 0) return KSW2_HIP_OK;
    std::vector<KswOut> outs(p->n);
    std::vector<uint32_t> cig(p->cig_words);
    KSW_HIP_TRY(hipMemcpy(outs.data(), p->d_out, sizeof(KswOut) * p->n, hipMemcpyDeviceToHost));
    KSW_HIP_TRY(hipMemcpy(cig.data(), p->d_cig, p->cig_words * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < p->n; ++k) {
        ksw_extz_t *z = ez + p->order[k];
        const KswOut &o = outs[k];
        z->max = (uint32_t)o.max, z->zdropped = o.zdropped;
        z->max_q = o.max_q, z->max_t = o.max_t;
        z->mqe = o.mqe, z->mqe_t = o.mqe_t, z->mte = o.mte, z->mte_q = o.mte_q;
        z->score = o.score, z->reach_end = o.reach_end, z->stopped = o.stopped;
        const int st = deliver_cigar(z, cig.data() + p->jobs[k].cig_off, o.n_cigar, km);
        if (st != KSW2_HIP_OK) return st;
    }
    return KSW2_HIP_OK;
}

void plan_destroy(Plan *p)
{
    if (!p) return;
    for (void *ptr : {(void *)p->d_seq, (void *)p->d_dir, (void *)p->d_bnd, (void *)p->d_cig, (void *)p->d_jobs,
                      (void *)p->d_out, (void *)p->d_order, (void *)p->d_recs})
        if (ptr) (void)hipFree(ptr);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    for (int l = 0; l < Plan::kLanes; ++l) {
        if (p->lane_done[l]) (void)hipEventDestroy(p->lane_done[l]);
        if (p->lane_stream[l]) (void)hipStreamDestroy(p->lane_stream[l]);
    }
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

int run_batch_host(int n, const ksw2_pair_t *pairs, int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2,
                   int8_t e2, ksw_extz_t *ez, void *km)
{
    for (int k = 0; k < n; ++k) reset_ez(ez + k);
    if (n == 0 || m <= 1) return KSW2_HIP_OK;
    KswScore sc;
    ScoreClass cls;
    if (!normalise_scoring(m, mat, q, e, q2, e2, &sc, &cls)) return KSW2_HIP_OK;  // reference returns early
    std::vector<int> fast;
    int status = KSW2_HIP_OK;
    for (int k = 0; k < n; ++k) {
        const ksw2_pair_t &p = pairs[k];
        if (p.qlen <= 0 || p.tlen <= 0) continue;            // reference returns early
        if (!p.query || !p.target) return KSW2_HIP_EINVAL;
        if (p.qlen > 0x7fffffff - p.tlen + 1) {               // ksw2_extd2_avx.c:138-141
            ez[k].zdropped = 1;
            continue;
        }
        if (fast_path_reject(p, m, cls) == kFastOK) {
            fast.push_back(k);
        } else {
            ez[k].zdropped = 1;  // not yet served by a device kernel
            status = KSW2_HIP_EUNSUPPORTED;
        }
    }
    // chunk by the device memory budget
    const uint64_t budget = mem_budget();
    size_t pos = 0;
    while (pos < fast.size()) {
        uint64_t used = 0;
        size_t end = pos;
        while (end < fast.size()) {
            const ksw2_pair_t &p = pairs[fast[end]];
            const uint64_t need = dir_bytes(p.qlen, p.tlen, kRowsPerLane) + 8ull * p.qlen + 4ull * (p.qlen + p.tlen) +
                                  (uint64_t)p.qlen + p.tlen + 512;
            if (end > pos && used + need > budget) break;
            used += need;
            ++end;
        }
        Plan *plan = nullptr;
        int st = plan_create(&plan, n, pairs, fast.data() + pos, (int)(end - pos), sc);
        if (st != KSW2_HIP_OK) return st;
        st = plan_run(plan);
        if (st == KSW2_HIP_OK) st = plan_fetch(plan, ez, km);
        plan_destroy(plan);
        if (st != KSW2_HIP_OK) return st;
        pos = end;
    }
    return status;
}

}  // namespace kswhip
