// ksw2_walk.h -- the Singletrack traceback on the device, one wavefront per
// pair.
//
// The walk is that of minimap2/ksw2_singletrack.cpp:63-176.  It reads two
// stored difference tracks, p(i,j) = H(i,j) - H(i,j-1) and
// p2(i,j) = H(i,j) - H(i-1,j), and runs from (tlen-1, qlen-1):
//  * While in the match state, it takes a diagonal step while
//    p(i,j) + p2(i,j-1) equals the substitution score of (i,j).  Row and
//    column -1 read the boundary's gap-cost difference.
//  * Otherwise it searches the gap length l = 1, 2, ...  The accumulated
//    column difference H(i,j) - H(i-l,j) (a deletion, tested first) or row
//    difference H(i,j) - H(i,j-l) (an insertion) must equal -(q + l*e) or
//    -(q2 + l*e2).
//  * With a band, a cell left of its anti-diagonal's stored range forces an
//    insertion and one right of it a deletion.  A cell out of range reads
//    KSW_NEG_INF, which poisons any sum it enters.
//  * When no gap length up to max(i,j)+1 fits, the walk fails.  The caller
//    then reports no CIGAR and zdropped.
//
// Here both searches run 64 steps at a time, one step per lane:
//  * Diagonal run: lane k tests cell (i-k, j-k).  The first failing lane
//    ends the run, found with one ballot.
//  * Gap search: lane k tests length l0+1+k.  The accumulated differences
//    are a wave prefix sum of one track value per lane, plus the carry of
//    earlier chunks.  The first lane that fits, or that passes max(i,j)+1,
//    decides.
// Each lane tests the same conditions, in the same order, as the scalar
// loop. So the walk takes the same path. Only the number of dependent
// memory round trips changes: one per 64 steps instead of one per step.
// Control flow is uniform, and lane 0 writes the run-length encoded CIGAR.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ksw2_common.h"

namespace kswhip {

// Gap costs as the walk uses them: negated (go = -q, ge = -e, ...).
struct WalkGaps {
    int32_t go, ge, go2, ge2;
};

__device__ __forceinline__ int32_t walk_gap(int32_t len, const WalkGaps &g)
{
    if (len <= 0) return 0;
    const int64_t a = (int64_t)g.go + (int64_t)len * g.ge, b = (int64_t)g.go2 + (int64_t)len * g.ge2;
    return (int32_t)(a > b ? a : b);
}

// Difference across the boundary row or column at coordinate c >= 0.
__device__ __forceinline__ int32_t walk_edge(int32_t c, const WalkGaps &g)
{
    return walk_gap(c + 1, g) - walk_gap(c, g);
}

__device__ __forceinline__ bool walk_fin(int32_t v)
{
    return v > KSW_NEG_INF_I / 2 && v < -(KSW_NEG_INF_I / 2);
}

__device__ __forceinline__ int walk_first(bool pred)
{
    const uint64_t m = __ballot(pred);
    return m ? __builtin_ctzll(m) : kWave;
}

__device__ __forceinline__ int32_t walk_scan(int32_t v, int lane)
{
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int32_t o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    return v;
}

// Runs the walk with the whole calling wavefront (all 64 lanes, uniform
// control flow).  The accessor supplies
//   int32_t get(int track, int i, int j)  track 0 = p, 1 = p2; i, j >= 0
//   int forced(int i, int j)              1 insertion, 2 deletion, 0 none
//   int32_t cost(int i, int j)            substitution score of (i, j)
// Writes the CIGAR (always reversed into forward order, as
// ksw2_singletrack.cpp:167-171 does whatever KSW_EZ_REV_CIGAR says) and
// returns its length, or -1 when the walk fails.
template <class Acc>
__device__ int track_walk(const Acc &A, int qlen, int tlen, const WalkGaps &g, uint32_t *cig)
{
    const int lane = threadIdx.x % kWave;
    int i = tlen - 1, j = qlen - 1, n = 0, run_op = -1, run_len = 0;
    auto push = [&](int op, int len) {
        if (op == run_op) {
            run_len += len;
        } else {
            if (run_len > 0) {
                if (lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                ++n;
            }
            run_op = op, run_len = len;
        }
    };
    bool ok = true;
    while (i >= 0 && j >= 0) {
        // ---- match state: a diagonal run, 64 cells per round
        {
            const int ci = i - lane, cj = j - lane;
            bool ev = true;
            if (ci >= 0 && cj >= 0 && A.forced(ci, cj) == 0) {
                const int32_t h = A.get(0, ci, cj);
                const int32_t v = cj > 0 ? A.get(1, ci, cj - 1) : walk_edge(ci, g);
                ev = !(walk_fin(h) && walk_fin(v) && (int64_t)h + v == (int64_t)A.cost(ci, cj));
            }
            const int k = walk_first(ev);
            if (k > 0) push(0, k), i -= k, j -= k;
            if (k == kWave) continue;
        }
        if (i < 0 || j < 0) break;
        const int f = A.forced(i, j);
        if (f == 1) {
            push(1, 1), --j;
            continue;
        }
        if (f == 2) {
            push(2, 1), --i;
            continue;
        }
        // ---- gap search from (i, j): lengths l0+1 .. l0+64 per round
        int32_t carry_i = 0, carry_d = 0;
        bool poison_i = false, poison_d = false;
        for (int l0 = 0;; l0 += kWave) {
            const int L = l0 + 1 + lane;  // this lane's gap length
            const int mt = L - 1;         // the track term it contributes
            int32_t ti = 0, td = 0;
            bool pi = false, pd = false;
            if (j - mt >= 0) {
                ti = A.get(0, i, j - mt);
                if (!walk_fin(ti)) pi = true, ti = 0;
            }
            if (i - mt >= 0) {
                td = A.get(1, i - mt, j);
                if (!walk_fin(td)) pd = true, td = 0;
            }
            const int32_t si = carry_i + walk_scan(ti, lane), sd = carry_d + walk_scan(td, lane);
            const uint64_t upto = (2ull << lane) - 1;  // lanes 0..lane
            const uint64_t bpi = __ballot(pi), bpd = __ballot(pd);
            const bool Pi = poison_i || (bpi & upto) != 0, Pd = poison_d || (bpd & upto) != 0;
            const int32_t a1 = g.go + L * g.ge, a2 = g.go2 + L * g.ge2;
            const bool bad = L > i + 1 && L > j + 1;
            const bool del = i - L >= -1 && !Pd && (sd == a1 || sd == a2);
            const bool ins = j - L >= -1 && !Pi && (si == a1 || si == a2);
            const int k = walk_first(bad || del || ins);
            if (k < kWave) {
                const int Lk = l0 + 1 + k;
                if (Lk > i + 1 && Lk > j + 1) {
                    ok = false;
                } else if (__shfl(del ? 1 : 0, k, kWave)) {
                    push(2, Lk), i -= Lk;
                } else {
                    push(1, Lk), j -= Lk;
                }
                break;
            }
            carry_i = __shfl(si, kWave - 1, kWave), carry_d = __shfl(sd, kWave - 1, kWave);
            poison_i = poison_i || bpi != 0, poison_d = poison_d || bpd != 0;
        }
        if (!ok) break;
    }
    if (!ok) return -1;
    if (i >= 0) push(2, i + 1);
    if (j >= 0) push(1, j + 1);
    push(-2, 0);  // flush the last run
    __threadfence_block();
    for (int x = lane; x < (n >> 1); x += kWave) {
        const uint32_t t = cig[x];
        cig[x] = cig[n - 1 - x];
        cig[n - 1 - x] = t;
    }
    return n;
}

}  // namespace kswhip
