// ksw2_certify.cpp -- score-certified banded alignment over batched calls:
// alignScoreCertifiedKsw2 (src/myImportandFunction/alignSlidingWindow.cpp:
// 1197-1400) for many intervals, with every task's next ksw_extd2 calls in
// one batched launch per round: the score-only calls, then the chained
// traceback's gap alignments of every task with a selection plan
// (tryChainedKsw2Traceback, :1056-1195), then the widening bands.
#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <limits>
#include <string>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_plan.h"
#include "ksw2_rows.h"

namespace {

using kswhip::code_of;
using kswhip::Progress;

constexpr int32_t kNegInf = -0x40000000;  // KSW_NEG_INF (minimap2/ksw2.h:18)

struct Task {
    int index = 0;
    std::vector<uint8_t> q, d;   // codes
    int flags = 0;               // KSW_EZ_GENERIC_SC when a wildcard occurs
    uint64_t len_diff = 0, max_band = 0, requested = 0, doubled = 0, band = 0;
    bool active = false;         // in the widening loop
    bool fallback = false;       // owes the attempt at the skipped first band
    bool have_best = false;
    std::vector<uint32_t> best;
};

// The rows alignSlidingWindow_minimap2 builds from a CIGAR (:852-870), the
// score scoreGappedTwoPieceAffineAlignment gives them (:677-721) and whether
// both reconstruct their sequences (gappedAlignmentReconstructs, :662-675).
struct Rescored {
    int64_t score = 0;
    bool valid = false;
};

Rescored rescore(const char *dq, int64_t ql, const char *dd, int64_t tl, const uint32_t *cigar, int n, int32_t mis,
                 int32_t o1, int32_t e1, int32_t o2, int32_t e2)
{
    std::string aq, ad;
    aq.reserve((size_t)(ql + tl)), ad.reserve((size_t)(ql + tl));
    int64_t qp = 0, tp = 0;
    kswhip::append_rows(aq, ad, dq, ql, qp, dd, tl, tp, cigar, n);
    Rescored r;
    bool ok = false;
    r.score = kswhip::rescore_rows(aq, ad, mis, o1, e1, o2, e2, &ok);
    r.valid = ok && kswhip::reconstructs(aq, dq, ql) && kswhip::reconstructs(ad, dd, tl);
    return r;
}

// Runs of aligned (M), query-only (I) and reference-only (D) columns of two
// gapped rows.
std::vector<uint32_t> cigar_of_rows(const std::string &aq, const std::string &ad)
{
    std::vector<uint32_t> out;
    for (size_t c = 0; c < aq.size(); ++c) {
        const uint32_t op = ad[c] == '-' ? 1u : aq[c] == '-' ? 2u : 0u;
        if (!out.empty() && (out.back() & 0xf) == op) out.back() += 1u << 4;
        else out.push_back(1u << 4 | op);
    }
    return out;
}

// One piece of a chained alignment (tryChainedKsw2Traceback): an anchor's
// matched columns, a gap with one side empty, or a gap aligned by a banded
// ksw_extd2 call (alignSlidingWindow_minimap2 at band min(max(lengths),
// maximum band), :1098-1121).
struct Piece {
    int kind;                    // 0 anchor, 1 one-sided gap, 2 aligned gap
    int64_t qs, qe, rs, re;
    int call = -1;               // aligned gap: index of its call in the round
};

uint32_t *copy_cigar(const uint32_t *src, int n)
{
    uint32_t *c = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)std::max(n, 1));
    if (c && n > 0) memcpy(c, src, sizeof(uint32_t) * (size_t)n);
    return c;
}

struct Deadline {
    bool on = false;
    std::chrono::steady_clock::time_point at;
    bool passed() const { return on && std::chrono::steady_clock::now() >= at; }
};

int deadline_probe(void *args, int, int) { return static_cast<Deadline *>(args)->passed() ? 0 : 1; }

}  // namespace

extern "C" void ksw2_hip_certify_free(int n, ksw2_certify_result_t *results)
{
    if (!results) return;
    for (int k = 0; k < n; ++k) {
        free(results[k].cigar), free(results[k].best_effort_cigar);
        results[k].cigar = results[k].best_effort_cigar = nullptr;
        results[k].n_cigar = results[k].n_best_effort = 0;
    }
}

extern "C" int ksw2_hip_certify_batch(int n, const ksw2_certify_task_t *tasks, int32_t mis, int32_t o1, int32_t e1,
                                      int32_t o2, int32_t e2, double max_seconds, ksw2_certify_result_t *res)
{
    if (n < 0 || (n > 0 && (!tasks || !res))) return KSW2_HIP_EINVAL;
    for (int k = 0; k < n; ++k) {
        ksw2_certify_result_t &r = res[k];
        memset(&r, 0, sizeof(r));
        r.status = KSW2_CERT_FAILED, r.final_band_width = -1;
    }
    // the matrix and gap arguments of ksw2ScoreOnly (:745-758) and
    // alignSlidingWindow_minimap2 (:804-808, :841)
    const int8_t a = 0, b = (int8_t)(mis < 0 ? mis : -mis);
    const int8_t mat[25] = {a, b, b, b, b, b, a, b, b, b, b, b, a, b, b, b, b, b, a, b, b, b, b, b, a};
    const int8_t q = (int8_t)(-o1), e = (int8_t)(-e1), q2 = (int8_t)(-o2), e2q = (int8_t)(-e2);
    Deadline dl;
    if (max_seconds > 0.0 && isfinite(max_seconds)) {
        dl.on = true;
        dl.at = std::chrono::steady_clock::now() +
                std::chrono::duration_cast<std::chrono::steady_clock::duration>(std::chrono::duration<double>(max_seconds));
    }
    Progress prog;
    prog.cb = dl.on ? deadline_probe : nullptr, prog.args = &dl, prog.interval = 256;

    std::vector<Task> work;
    std::vector<int> need_score;  // indices into work
    for (int k = 0; k < n; ++k) {
        const ksw2_certify_task_t &t = tasks[k];
        ksw2_certify_result_t &r = res[k];
        const int64_t ql = std::max<int64_t>(t.qlen, 0), tl = std::max<int64_t>(t.tlen, 0);
        if ((ql > 0 && !t.query) || (tl > 0 && !t.target)) return KSW2_HIP_EINVAL;
        const uint64_t minlen = (uint64_t)std::min(ql, tl);
        const uint64_t diff = (uint64_t)(ql > tl ? ql - tl : tl - ql);
        if (minlen == 0) {
            // one side empty (:1216-1232): all gap, scored as ksw2ScoreOnly's
            // closed form (:731-739)
            const int64_t len = std::max(ql, tl);
            r.score = std::max<int64_t>((int64_t)o1 + (int64_t)e1 * len, (int64_t)o2 + (int64_t)e2 * len);
            r.optimal_score = r.score, r.final_band_width = 0, r.traceback_attempts = 1;
            r.status = KSW2_CERT_COMPLETED;
            if (len > 0) {
                const uint32_t op = ql > 0 ? 1u : 2u;  // only query: I; only target: D
                const uint32_t c = (uint32_t)len << 4 | op;
                r.cigar = copy_cigar(&c, 1), r.n_cigar = 1;
            }
            continue;
        }
        if (t.initial_band < 0 || t.maximum_band < 0 || (uint64_t)t.maximum_band < diff) {
            r.status = KSW2_CERT_NOT_CERTIFIED;  // :1233-1238
            continue;
        }
        Task w;
        w.index = k;
        w.len_diff = diff;
        if (ql > 0x7fffffff || tl > 0x7fffffff) {
            r.optimal_score = kNegInf;  // ksw2ScoreOnly returns KSW_NEG_INF (:740-744)
            r.status = KSW2_CERT_FAILED;
            continue;
        }
        w.q.resize((size_t)ql), w.d.resize((size_t)tl);
        for (int64_t i = 0; i < ql; ++i)
            if ((w.q[(size_t)i] = code_of(t.query[i])) == 4) w.flags |= KSW_EZ_GENERIC_SC;
        for (int64_t i = 0; i < tl; ++i)
            if ((w.d[(size_t)i] = code_of(t.target[i])) == 4) w.flags |= KSW_EZ_GENERIC_SC;
        if (t.precomputed_optimal_score) r.optimal_score = *t.precomputed_optimal_score;
        else need_score.push_back((int)work.size());
        work.push_back(std::move(w));
    }

    // one batched call over `idx`, band b(task) (-1: score-only)
    std::vector<ksw2_pair_t> pairs;
    std::vector<ksw_extz_t> ez;
    auto run_round = [&](const std::vector<int> &idx, bool score_only, const std::vector<uint64_t> &bands) -> int {
        pairs.assign(idx.size(), ksw2_pair_t{});
        for (auto &z : ez) free(z.cigar);
        ez.assign(idx.size(), ksw_extz_t{});
        for (size_t i = 0; i < idx.size(); ++i) {
            const Task &w = work[(size_t)idx[i]];
            ksw2_pair_t &p = pairs[i];
            p.query = w.q.data(), p.target = w.d.data();
            p.qlen = (int32_t)w.q.size(), p.tlen = (int32_t)w.d.size();
            p.w = score_only ? -1 : (int32_t)bands[i];
            p.zdrop = -1, p.end_bonus = 0;
            p.flag = w.flags | (score_only ? KSW_EZ_SCORE_ONLY : 0);
        }
        if (idx.empty()) return KSW2_HIP_OK;
        return kswhip::run_batch_host((int)idx.size(), pairs.data(), 5, mat, q, e, q2, e2q, ez.data(), nullptr,
                                      dl.on ? &prog : nullptr);
    };
    auto release = [&]() {
        for (auto &z : ez) free(z.cigar);
        ez.clear();
    };

    // the optimal scores (:1240-1256)
    {
        const int st = run_round(need_score, true, {});
        if (st != KSW2_HIP_OK) {
            release();
            return st;
        }
        const bool late = dl.passed();
        for (size_t i = 0; i < need_score.size(); ++i) {
            Task &w = work[(size_t)need_score[i]];
            ksw2_certify_result_t &r = res[w.index];
            r.optimal_score = ez[i].score;
            r.status = ez[i].stopped || late ? KSW2_CERT_TIME_LIMIT : KSW2_CERT_FAILED;
        }
    }
    std::vector<int> loop;
    for (size_t i = 0; i < work.size(); ++i) {
        Task &w = work[i];
        ksw2_certify_result_t &r = res[w.index];
        if (r.status == KSW2_CERT_TIME_LIMIT) continue;
        if (r.optimal_score == kNegInf) {
            r.status = KSW2_CERT_FAILED;
            continue;
        }
        const ksw2_certify_task_t &t = tasks[w.index];
        // the widening schedule (:1285-1300)
        w.max_band = std::min<uint64_t>(std::max(w.q.size(), w.d.size()), (uint64_t)t.maximum_band);
        w.requested = std::max<uint64_t>(w.len_diff, (uint64_t)std::max<int64_t>(1, t.initial_band));
        w.doubled = w.requested > std::numeric_limits<uint64_t>::max() / 2 ? w.max_band : w.requested * 2;
        w.band = std::min(w.max_band, std::max(w.requested, w.doubled));
        w.active = w.band >= w.len_diff && w.band <= w.max_band;
        w.fallback = !w.active;
        r.status = KSW2_CERT_NOT_CERTIFIED;
        loop.push_back((int)i);
    }

    // the chained traceback through the selection plan's anchors
    // (:1260-1289, tryChainedKsw2Traceback :1056-1195): every task's aligned
    // gaps in one batched call, then each chain assembled in order
    {
        std::vector<std::vector<Piece>> chains(work.size());
        std::vector<char> has_chain(work.size(), 0);
        std::vector<ksw2_pair_t> cp;  // the aligned gaps (codes point into the tasks)
        for (int i : loop) {
            Task &w = work[(size_t)i];
            const ksw2_certify_task_t &t = tasks[w.index];
            if (!t.anchors) continue;
            ++res[w.index].traceback_attempts;
            if (t.n_anchors < 2 || w.max_band == 0) continue;  // :1071-1074
            const int64_t ql = (int64_t)w.q.size(), tl = (int64_t)w.d.size();
            std::vector<Piece> pieces;
            bool ok = true;
            auto gap = [&](int64_t qs, int64_t qe, int64_t rs, int64_t re) {
                const int64_t a = qe - qs, b = re - rs;
                if (a == 0 || b == 0) {
                    if (a || b) pieces.push_back(Piece{1, qs, qe, rs, re});
                    return true;
                }
                const uint64_t d = (uint64_t)(a > b ? a - b : b - a);
                if (d > w.max_band || a > 0x7fffffff || b > 0x7fffffff) return false;
                Piece pc{2, qs, qe, rs, re};
                pc.call = (int)cp.size();
                ksw2_pair_t pr;
                pr.query = w.q.data() + qs, pr.target = w.d.data() + rs;
                pr.qlen = (int32_t)a, pr.tlen = (int32_t)b;
                pr.w = (int32_t)std::min<uint64_t>((uint64_t)std::max(a, b), w.max_band);
                pr.zdrop = -1, pr.end_bonus = 0, pr.flag = 0;
                for (int64_t k = 0; k < a && !pr.flag; ++k)
                    if (pr.query[k] == 4) pr.flag = KSW_EZ_GENERIC_SC;
                for (int64_t k = 0; k < b && !pr.flag; ++k)
                    if (pr.target[k] == 4) pr.flag = KSW_EZ_GENERIC_SC;
                cp.push_back(pr);
                pieces.push_back(pc);
                return true;
            };
            int64_t qpos = 0, rpos = 0, used = 0;
            for (int64_t k = 0; k < t.n_anchors && ok; ++k) {  // :1133-1170
                const ksw2_chain_anchor_t &an = t.anchors[k];
                const int64_t aq = an.query_position, ar = an.reference_position, al = an.length;
                if (al == 0 || aq < qpos || ar < rpos || aq + al > ql || ar + al > tl) continue;
                bool match = true;
                for (int64_t o = 0; o < al && match; ++o)
                    match = toupper((unsigned char)t.query[aq + o]) == toupper((unsigned char)t.target[ar + o]);
                if (!match) continue;
                if (!gap(qpos, aq, rpos, ar)) ok = false;
                else pieces.push_back(Piece{0, aq, aq + al, ar, ar + al}), qpos = aq + al, rpos = ar + al, ++used;
            }
            if (ok && used >= 2 && gap(qpos, ql, rpos, tl)) {  // :1171-1179
                chains[(size_t)i] = std::move(pieces);
                has_chain[(size_t)i] = 1;
            } else {
                // unavailable: none of its calls run (they are the last ones)
                cp.resize(cp.size() - std::count_if(pieces.begin(), pieces.end(),
                                                     [](const Piece &pc) { return pc.kind == 2; }));
            }
        }
        std::vector<ksw_extz_t> cez(cp.size(), ksw_extz_t{});
        if (!cp.empty()) {
            const int st = kswhip::run_batch_host((int)cp.size(), cp.data(), 5, mat, q, e, q2, e2q, cez.data(), nullptr,
                                                  dl.on ? &prog : nullptr);
            if (st != KSW2_HIP_OK) {
                for (auto &z : cez) free(z.cigar);
                release();
                return st;
            }
        }
        const bool late = dl.passed();
        for (int i : loop) {
            if (!has_chain[(size_t)i]) continue;
            Task &w = work[(size_t)i];
            const ksw2_certify_task_t &t = tasks[w.index];
            ksw2_certify_result_t &r = res[w.index];
            std::string aq, ad;
            bool ok = true, stopped = false;
            for (const Piece &pc : chains[(size_t)i]) {
                if (pc.kind == 0 || pc.kind == 1) {
                    if (pc.qe > pc.qs) aq.append(t.query + pc.qs, (size_t)(pc.qe - pc.qs));
                    else aq.append((size_t)(pc.re - pc.rs), '-');
                    if (pc.re > pc.rs) ad.append(t.target + pc.rs, (size_t)(pc.re - pc.rs));
                    else ad.append((size_t)(pc.qe - pc.qs), '-');
                    continue;
                }
                const ksw_extz_t &z = cez[(size_t)pc.call];
                if (z.stopped || late) {  // :1122-1127 (the deadline)
                    stopped = true;
                    break;
                }
                std::string sq, sd;
                int64_t qp = 0, tp = 0;
                const int64_t a = pc.qe - pc.qs, b = pc.re - pc.rs;
                kswhip::append_rows(sq, sd, t.query + pc.qs, a, qp, t.target + pc.rs, b, tp, z.cigar, z.n_cigar);
                if (sq.size() != sd.size() || !kswhip::reconstructs(sq, t.query + pc.qs, a) ||
                    !kswhip::reconstructs(sd, t.target + pc.rs, b)) {
                    ok = false;
                    break;
                }
                aq += sq, ad += sd;
            }
            if (stopped) {
                r.status = KSW2_CERT_TIME_LIMIT;
                w.active = w.fallback = false;
                continue;
            }
            if (!ok || !kswhip::reconstructs(aq, t.query, (int64_t)w.q.size()) ||
                !kswhip::reconstructs(ad, t.target, (int64_t)w.d.size()))
                continue;  // unavailable (:1180-1185)
            bool rok = false;
            const int64_t score = kswhip::rescore_rows(aq, ad, mis, o1, e1, o2, e2, &rok);
            if (!rok) continue;
            const std::vector<uint32_t> cig = cigar_of_rows(aq, ad);
            r.score = score;
            if (score == r.optimal_score) {
                r.final_band_width = -2;
                r.cigar = copy_cigar(cig.data(), (int)cig.size()), r.n_cigar = (int)cig.size();
                r.status = KSW2_CERT_COMPLETED;
                w.active = w.fallback = false;
            } else {
                w.best = cig;
                w.have_best = true;
            }
        }
        for (auto &z : cez) free(z.cigar);
    }

    // one attempt for each task in `idx` at its band; returns the status
    auto attempt_round = [&](const std::vector<int> &idx, bool at_requested) -> int {
        std::vector<uint64_t> bands;
        for (int i : idx) bands.push_back(at_requested ? work[(size_t)i].requested : work[(size_t)i].band);
        const int st = run_round(idx, false, bands);
        if (st != KSW2_HIP_OK) return st;
        const bool late = dl.passed();
        for (size_t j = 0; j < idx.size(); ++j) {
            Task &w = work[(size_t)idx[j]];
            const ksw2_certify_task_t &t = tasks[w.index];
            ksw2_certify_result_t &r = res[w.index];
            ++r.traceback_attempts;
            if (ez[j].stopped || late) {  // :1313-1316
                r.status = KSW2_CERT_TIME_LIMIT;
                w.active = w.fallback = false;
                continue;
            }
            const Rescored rs = rescore(t.query, (int64_t)w.q.size(), t.target, (int64_t)w.d.size(), ez[j].cigar,
                                        ez[j].n_cigar, mis, o1, e1, o2, e2);
            if (rs.score == r.optimal_score && rs.valid) {
                r.score = rs.score;
                r.final_band_width = (int64_t)bands[j];
                r.cigar = copy_cigar(ez[j].cigar, ez[j].n_cigar), r.n_cigar = ez[j].n_cigar;
                r.status = KSW2_CERT_COMPLETED;
                w.active = w.fallback = false;
                continue;
            }
            if (rs.valid && (!w.have_best || rs.score > r.score)) {
                r.score = rs.score;
                w.best.assign(ez[j].cigar, ez[j].cigar + ez[j].n_cigar);
                w.have_best = true;
            }
            if (at_requested) {
                w.fallback = false;
                continue;
            }
            if (w.band == w.max_band) {
                w.active = false, w.fallback = true;
                continue;
            }
            const uint64_t dbl = w.band > std::numeric_limits<uint64_t>::max() / 2 ? w.max_band : w.band * 2;
            const uint64_t next = std::min(w.max_band, std::max(w.band + 1, dbl));
            if (next <= w.band) {
                w.active = false, w.fallback = true;
                continue;
            }
            w.band = next;
        }
        return KSW2_HIP_OK;
    };

    for (;;) {
        std::vector<int> idx;
        for (int i : loop)
            if (work[(size_t)i].active) idx.push_back(i);
        if (idx.empty()) break;
        const int st = attempt_round(idx, false);
        if (st != KSW2_HIP_OK) {
            release();
            return st;
        }
    }
    // the skipped first band, tried only on failure (:1347-1394)
    {
        std::vector<int> idx;
        for (int i : loop) {
            Task &w = work[(size_t)i];
            if (w.fallback && res[w.index].status == KSW2_CERT_NOT_CERTIFIED &&
                w.requested < std::min(w.max_band, w.doubled))
                idx.push_back(i);
        }
        const int st = attempt_round(idx, true);
        if (st != KSW2_HIP_OK) {
            release();
            return st;
        }
    }
    for (int i : loop) {
        Task &w = work[(size_t)i];
        ksw2_certify_result_t &r = res[w.index];
        if (r.status == KSW2_CERT_NOT_CERTIFIED) {
            r.final_band_width = (int64_t)w.band;  // :1395
            if (w.have_best) r.best_effort_cigar = copy_cigar(w.best.data(), (int)w.best.size()),
                             r.n_best_effort = (int)w.best.size();
        }
    }
    release();
    return KSW2_HIP_OK;
}
