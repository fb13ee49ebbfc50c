// ksw2_window.cpp -- AnchorWave's sliding-window alignment over batched
// calls: alignSlidingWindow (src/myImportandFunction/alignSlidingWindow.cpp:
// 457-590) for many intervals, window k of every interval in one launch.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_plan.h"
#include "ksw2_rows.h"

namespace {

struct Scoring {
    int8_t mat[25];
    int8_t q, e, q2, e2;
};

// The matrix and gap arguments an adapter passes for AnchorWave penalties
// (alignSlidingWindow.cpp:246-256, :339-344, :386): match 0, mismatch
// -|mismatch|, gaps negated (narrowed to int8 as the call does).
Scoring scoring_of(int32_t mis, int32_t o1, int32_t e1, int32_t o2, int32_t e2)
{
    Scoring s;
    const int8_t a = 0, b = (int8_t)(mis < 0 ? mis : -mis);
    for (int i = 0; i < 25; ++i) s.mat[i] = i % 6 == 0 ? a : b;
    s.q = (int8_t)(-o1), s.e = (int8_t)(-e1), s.q2 = (int8_t)(-o2), s.e2 = (int8_t)(-e2);
    return s;
}

struct Task {
    const char *dq = nullptr, *dd = nullptr;
    int64_t ql = 0, tl = 0;
    int64_t qstart = 1, dstart = 1;  // 1-based, as the reference's queryStart / databaseStart
    bool active = false;
    std::string aq, ad;              // the rows so far
    std::vector<uint32_t> ops;       // the same alignment as merged M/I/D runs
    std::vector<uint8_t> wq, wd;     // the current window's codes
    int flags = 0;
};

void push_op(std::vector<uint32_t> &ops, uint32_t op, int64_t len)
{
    if (len <= 0) return;
    if (!ops.empty() && (ops.back() & 0xf) == op) ops.back() += (uint32_t)len << 4;
    else ops.push_back((uint32_t)len << 4 | op);
}

void encode(const char *s, int64_t from, int64_t len, std::vector<uint8_t> &out, int &flags)
{
    out.resize((size_t)len);
    for (int64_t i = 0; i < len; ++i)
        if ((out[(size_t)i] = kswhip::code_of(s[from + i])) == 4) flags |= KSW_EZ_GENERIC_SC;
}

}  // namespace

extern "C" void ksw2_hip_window_free(int n, ksw2_window_result_t *results)
{
    if (!results) return;
    for (int k = 0; k < n; ++k) {
        free(results[k].cigar);
        results[k].cigar = nullptr, results[k].n_cigar = 0;
    }
}

extern "C" int ksw2_hip_window_batch(int n, const ksw2_window_task_t *tasks, int64_t window, int32_t ms, int32_t mis,
                                     int32_t o1, int32_t e1, int32_t o2, int32_t e2, ksw2_window_result_t *res)
{
    if (n < 0 || (n > 0 && (!tasks || !res))) return KSW2_HIP_EINVAL;
    if (window < 0) return KSW2_HIP_EINVAL;
    for (int k = 0; k < n; ++k) memset(&res[k], 0, sizeof(res[k]));
    // penalties of the whole-interval alignment, and of the windows (:535-539)
    const Scoring whole = scoring_of(mis, o1, e1, o2, e2);
    const Scoring win = e2 + ms < 0 ? scoring_of(mis + ms, o1 + ms, e1 + ms, o2 + ms, e2 + ms)
                                    : scoring_of(mis + ms - 1, o1 + ms - 1, e1 + ms - 1, o2 + ms - 1, -1);
    std::vector<Task> work((size_t)n);
    std::vector<int> global_idx;
    for (int k = 0; k < n; ++k) {
        const ksw2_window_task_t &t = tasks[k];
        Task &w = work[(size_t)k];
        w.ql = std::max<int64_t>(t.qlen, 0), w.tl = std::max<int64_t>(t.tlen, 0);
        if ((w.ql > 0 && !t.query) || (w.tl > 0 && !t.target)) return KSW2_HIP_EINVAL;
        if (w.ql > 0x7fffffff || w.tl > 0x7fffffff) return KSW2_HIP_EUNSUPPORTED;
        w.dq = t.query, w.dd = t.target;
        w.aq.reserve((size_t)(w.ql + w.tl)), w.ad.reserve((size_t)(w.ql + w.tl));
        // one window covers the interval (:489): the whole-interval alignment
        if (((double)w.tl / (double)window) * ((double)w.ql / (double)window) <= 1) {
            global_idx.push_back(k);
        } else {
            w.active = w.dstart <= w.tl && w.qstart <= w.ql;
            if (w.active && window > 1073741824) res[k].status = 2, w.active = false;  // the reference exits (:529-531)
        }
    }
    std::vector<ksw2_pair_t> pairs;
    std::vector<ksw_extz_t> ez;
    auto release = [&]() {
        for (auto &z : ez) free(z.cigar);
        ez.clear();
    };
    auto run = [&](const std::vector<int> &idx, const Scoring &sc, int extra_flag) -> int {
        for (auto &z : ez) free(z.cigar);
        ez.assign(idx.size(), ksw_extz_t{});
        pairs.assign(idx.size(), ksw2_pair_t{});
        for (size_t i = 0; i < idx.size(); ++i) {
            const Task &w = work[(size_t)idx[i]];
            ksw2_pair_t &p = pairs[i];
            p.query = w.wq.data(), p.target = w.wd.data();
            p.qlen = (int32_t)w.wq.size(), p.tlen = (int32_t)w.wd.size();
            p.w = -1, p.zdrop = -1, p.end_bonus = 0, p.flag = w.flags | extra_flag;
        }
        if (idx.empty()) return KSW2_HIP_OK;
        return kswhip::run_batch_host((int)idx.size(), pairs.data(), 5, sc.mat, sc.q, sc.e, sc.q2, sc.e2, ez.data(),
                                      nullptr);
    };

    // whole-interval alignments (minimap2AlignmentWithArena, :241-332)
    for (int k : global_idx) {
        Task &w = work[(size_t)k];
        w.flags = 0;
        encode(w.dq, 0, w.ql, w.wq, w.flags);
        encode(w.dd, 0, w.tl, w.wd, w.flags);
    }
    {
        std::vector<int> run_idx;
        for (int k : global_idx)
            if (work[(size_t)k].ql > 0 && work[(size_t)k].tl > 0) run_idx.push_back(k);
        int st = run(run_idx, whole, 0);
        if (st != KSW2_HIP_OK) {
            release();
            return st;
        }
        for (size_t i = 0; i < run_idx.size(); ++i) {
            Task &w = work[(size_t)run_idx[i]];
            int64_t qp = 0, tp = 0;
            kswhip::append_rows(w.aq, w.ad, w.dq, w.ql, qp, w.dd, w.tl, tp, ez[i].cigar, ez[i].n_cigar);
            for (int c = 0; c < ez[i].n_cigar; ++c) push_op(w.ops, ez[i].cigar[c] & 0xf, ez[i].cigar[c] >> 4);
        }
        for (int k : global_idx) {
            Task &w = work[(size_t)k];
            w.dstart = w.tl + 1, w.qstart = w.ql + 1;  // :494-495
            res[k].windows = 1;  // one call (it returns at once when a side is empty)
        }
    }

    // windows, one round per window index (:497-557)
    for (;;) {
        std::vector<int> idx;
        for (int k = 0; k < n; ++k) {
            Task &w = work[(size_t)k];
            if (!w.active || res[k].status != 0) continue;
            // getSubsequence(seq, start, start + W) clipped to the sequence (1-based, inclusive)
            const int64_t qend = std::min(w.qstart + window, w.ql), dend = std::min(w.dstart + window, w.tl);
            w.flags = 0;
            encode(w.dq, w.qstart - 1, qend - w.qstart + 1, w.wq, w.flags);
            encode(w.dd, w.dstart - 1, dend - w.dstart + 1, w.wd, w.flags);
            idx.push_back(k);
        }
        if (idx.empty()) break;
        const int st = run(idx, win, KSW_EZ_SEMIGLOBAL_END);
        if (st != KSW2_HIP_OK) {
            release();
            return st;
        }
        for (size_t i = 0; i < idx.size(); ++i) {
            Task &w = work[(size_t)idx[i]];
            ksw2_window_result_t &r = res[idx[i]];
            ++r.windows;
            const ksw_extz_t &z = ez[i];
            const int64_t wql = (int64_t)w.wq.size(), wtl = (int64_t)w.wd.size();
            // the better end (:392-398)
            int64_t endq, endt;
            if (z.mqe > z.mte) endt = z.mqe_t + 1, endq = wql;
            else endt = wtl, endq = z.mte_q + 1;
            int64_t qp = 0, tp = 0;
            kswhip::append_rows(w.aq, w.ad, w.dq + (w.qstart - 1), wql, qp, w.dd + (w.dstart - 1), wtl, tp, z.cigar,
                                z.n_cigar);
            if (qp != endq || tp != endt) {  // :547-553
                r.status = 1;
                w.active = false;
                continue;
            }
            for (int c = 0; c < z.n_cigar; ++c) push_op(w.ops, z.cigar[c] & 0xf, z.cigar[c] >> 4);
            w.qstart += endq, w.dstart += endt;  // :555-556
            w.active = w.dstart <= w.tl && w.qstart <= w.ql;
        }
    }
    release();

    // the tails (:559-580) and the rescored alignment (:587-589)
    for (int k = 0; k < n; ++k) {
        Task &w = work[(size_t)k];
        ksw2_window_result_t &r = res[k];
        if (r.status != 0) continue;
        const int64_t c1 = w.tl - w.dstart;
        if (c1 >= 0) {
            w.aq.append((size_t)(c1 + 1), '-');
            w.ad.append(w.dd + (w.dstart - 1), (size_t)(c1 + 1));
            push_op(w.ops, 2, c1 + 1);
        }
        const int64_t c2 = w.ql - w.qstart;
        if (c2 >= 0) {
            w.aq.append(w.dq + (w.qstart - 1), (size_t)(c2 + 1));
            w.ad.append((size_t)(c2 + 1), '-');
            push_op(w.ops, 1, c2 + 1);
        }
        bool ok = false;
        r.score = kswhip::rescore_rows(w.aq, w.ad, mis, o1, e1, o2, e2, &ok);
        if (!ok) r.status = 1;
        r.n_cigar = (int32_t)w.ops.size();
        r.cigar = (uint32_t *)malloc(sizeof(uint32_t) * std::max<size_t>(w.ops.size(), 1));
        if (r.cigar && !w.ops.empty()) memcpy(r.cigar, w.ops.data(), sizeof(uint32_t) * w.ops.size());
    }
    return KSW2_HIP_OK;
}
