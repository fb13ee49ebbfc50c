// ksw2_general.hip -- gfx950 kernel for every ksw_extd2 call the stripe
// kernel does not take: bands, z-drop, KSW_EZ_APPROX_MAX / APPROX_DROP,
// generic matrices, the wildcard code, e <= e2, positive match scores, and
// inputs whose int8 intermediates wrap.
//
// It evaluates the reference's anti-diagonal recurrence exactly as the
// AVX-512 kernel does (minimap2/ksw2_extd2_avx.c:248-745, SSE equivalent
// minimap2/ksw2_extd2_sse.c:136-396): int8 arithmetic with two's-complement
// wrap, cells computed over the 64-byte-block range [stB, en16] of every
// anti-diagonal r (including padding cells outside the band, whose values
// feed in-band cells and whose direction bytes a banded backtrack can read),
// scores written in 16-byte groups, the boundary injected at the 16-rounded
// start, and the AVX-512 lane wrap at a non-64-aligned start.
//
// Parallel form: one workgroup per pair; thread k owns target positions
// t = k (mod blockDim.x) for the whole run, so the arrays a cell reads and
// writes at its own t (u, y, y2, s, H) never cross threads.  x, v and x2,
// which a cell reads at t-1, are double-buffered by the parity of r.  The
// scalar bookkeeping of anti-diagonal r (H at en0, the maximum with the
// reference's 4-lane tie order, mte/mqe, z-drop, APPROX_MAX's path) is
// finished by every thread from LDS right after the barrier of r, at the
// start of r+1: one barrier per anti-diagonal.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ksw2_common.h"
#include "ksw2_kernels.h"
#include "ksw2_walk.h"

namespace kswhip {

namespace {

constexpr int kMaxWaves = 16;
constexpr int kFlagScoreOnly = 0x01, kFlagRight = 0x02, kFlagGeneric = 0x04, kFlagApproxMax = 0x08,
              kFlagApproxDrop = 0x10, kFlagExtzOnly = 0x40, kFlagSemiglobalEnd = 0x800, kFlagSingletrack = 0x1000;

__device__ __forceinline__ int w8(int v) { return (int)(int8_t)(uint8_t)(v & 0xff); }

// Per-anti-diagonal hand-off through LDS (double-buffered by parity of r).
struct Fin {
    int best_h[kMaxWaves];
    uint32_t best_rank[kMaxWaves];
    int h_old_en0m1;  // H[en0-1] before this anti-diagonal's update
    int h_old_en0;    // H[en0] before it (used when en0 == 0)
    int u_en0, v_en0; // new u, v at en0
    int h_st0;        // new H at st0 (st0 < en0)
    int v_h0, u_h0;   // APPROX_MAX: new v at last_H0_t, new u at last_H0_t + 1
    int stop;         // host stop flag as read by thread 0 (r % 256 == 0)
};

// ksw_apply_zdrop with is_rot = 1 (minimap2/ksw2.h:201-217); ez->max is a
// 31-bit field compared as int.
struct Track {
    int max, max_t, max_q, mqe, mqe_t, mte, mte_q, score, zdropped, stopped;
};
__device__ __forceinline__ bool apply_zdrop(Track &z, int H, int r, int t, int zdrop, int e2)
{
    if (H > z.max) {
        z.max = H, z.max_t = t, z.max_q = r - t;
    } else if (t >= z.max_t && r - t >= z.max_q) {
        const int tl = t - z.max_t, ql = (r - t) - z.max_q;
        const int l = tl > ql ? tl - ql : ql - tl;
        if (zdrop >= 0 && z.max - H > zdrop + l * e2) {
            z.zdropped = 1;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ int edge_i8(const KswScore &sc, int r) { return edge_diff(sc, r); }

}  // namespace

// RING: the per-position state lives in LDS, in rings of a.ring slots
// indexed by t mod a.ring, instead of the pair's state block in HBM.  The
// band [stB, en16] moves right monotonically, so a ring a few hundred
// slots wider than the band holds every position a cell reads (t-1, st16-1,
// stB+15, en0-1); a slot is reset to the reference's initial values
// (ksw2_extd2_avx.c:226-245) just before its position first enters.  The
// cells compute exactly what they compute from HBM, at LDS latency.
template <bool RING>
__global__ void __launch_bounds__(kWave * kMaxWaves) ksw_general_kernel(GenArgs a)
{
    __shared__ Fin fin[2];
    extern __shared__ __align__(16) uint8_t gen_ring[];
    const int job = __builtin_amdgcn_readfirstlane(a.order[blockIdx.x]);
    if (a.only_if && __builtin_amdgcn_readfirstlane(a.only_if[job]) == 0) return;  // uniform
    const GenJob jb = a.jobs[job];
    const GenScore &gs = a.sc;
    const KswScore &sc = gs.k;
    const int nthr = blockDim.x, tid = threadIdx.x, lane = tid % kWave, wv = tid / kWave;
    const int nw = nthr / kWave;
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    const int w = __builtin_amdgcn_readfirstlane(jb.w), flag = __builtin_amdgcn_readfirstlane(jb.flag);
    const int tlenB = jb.tlenB, qlenB = jb.qlenB;
    const bool with_cigar = !(flag & kFlagScoreOnly), right = flag & kFlagRight, generic = flag & kFlagGeneric;
    const bool approx = flag & kFlagApproxMax;
    // Singletrack: the new v and u of every cell instead of its direction
    // byte (minimap2/ksw2_extd2_avx.c:432-436), in two row arrays
    const bool tracks = with_cigar && (flag & kFlagSingletrack);
    const GenLayout L = gen_layout(tlenB, qlenB);
    uint8_t *base = a.state + jb.st_off;
    // state arrays: ring slots in LDS (index & mk) or the state block (mk = -1)
    const int RS = RING ? a.ring : 0;
    const int mk = RING ? RS - 1 : -1;
    uint8_t *sb = RING ? gen_ring : base;
    int8_t *u = (int8_t *)(sb + (RING ? 0 * RS : L.u)), *y = (int8_t *)(sb + (RING ? 1 * RS : L.y));
    int8_t *y2 = (int8_t *)(sb + (RING ? 2 * RS : L.y2));
    int8_t *s = (int8_t *)(sb + (RING ? 3 * RS : L.s));
    int8_t *const X0 = (int8_t *)(sb + (RING ? 4 * RS : L.x0)), *const X1 = (int8_t *)(sb + (RING ? 5 * RS : L.x1));
    int8_t *const V0 = (int8_t *)(sb + (RING ? 6 * RS : L.v0)), *const V1 = (int8_t *)(sb + (RING ? 7 * RS : L.v1));
    int8_t *const XX0 = (int8_t *)(sb + (RING ? 8 * RS : L.xx0));
    int8_t *const XX1 = (int8_t *)(sb + (RING ? 9 * RS : L.xx1));
    int32_t *H = (int32_t *)(sb + (RING ? 10 * RS : L.H));
    uint8_t *sf = base + L.sf, *qr = base + L.qr;
    uint8_t *p = a.dir + jb.dir_off;
    int32_t *off = a.offs + jb.off_off, *off_end = off + (qlen + tlen - 1);
    const int ncolB = jb.ncolB;
    uint8_t *p2 = p + (uint64_t)(qlen + tlen - 1) * (uint64_t)ncolB;

    const int q = sc.q, e = sc.e, q2 = sc.q2, e2 = sc.e2;
    const int qe8 = w8(q + e), qe28 = w8(q2 + e2);
    const int sc_mch = gs.mat[0], sc_mis = gs.mat[1], sc_N = gs.sc_N, m1 = w8(gs.m - 1), m = gs.m;

    // state set-up (ksw2_extd2_avx.c:226-245): u v x y = -(q+e), x2 y2 =
    // -(q2+e2), s = 0, H = NEG_INF, sf = target, qr = reversed query
    {
        const uint8_t *tq = a.seq + jb.t_off, *qq = a.seq + jb.q_off;
        for (int t = tid; t < (RING ? RS : tlenB); t += nthr) {
            u[t] = y[t] = X0[t] = X1[t] = V0[t] = V1[t] = (int8_t)w8(-q - e);
            y2[t] = XX0[t] = XX1[t] = (int8_t)w8(-q2 - e2);
            s[t] = 0;
            H[t] = KSW_NEG_INF_I;
        }
        for (int t = tid; t < tlenB; t += nthr) sf[t] = t < tlen ? tq[t] : 0;
        for (int t = tid; t < qlenB + 128; t += nthr) qr[t] = t < qlen ? qq[qlen - 1 - t] : 0;
    }
    __syncthreads();

    Track z;
    z.max = 0, z.max_t = -1, z.max_q = -1, z.mqe = KSW_NEG_INF_I, z.mqe_t = -1;
    z.mte = KSW_NEG_INF_I, z.mte_q = -1, z.score = KSW_NEG_INF_I, z.zdropped = 0, z.stopped = 0;
    int last_st = -1, last_en = -1;
    int H0 = 0, last_H0_t = 0;
    int init_end = RS - 1;  // RING: positions < init_end + 1 have their slots set
    // band of the previous anti-diagonal, for its deferred bookkeeping
    int p_st0 = 0, p_en0 = 0, p_en1 = 0;
    const int nr = qlen + tlen - 1;

    for (int r = 0; r <= nr; ++r) {
        // ---- finish anti-diagonal r-1 (every thread, identical values)
        if (r > 0) {
            const int rr = r - 1;
            const Fin &F = fin[rr & 1];
            bool stop = false;
            if (!approx) {
                const int st0 = p_st0, en0 = p_en0;
                int h_en0;
                if (rr == 0) h_en0 = F.v_en0 - sc.qe_orig;
                else h_en0 = en0 > 0 ? F.h_old_en0m1 + F.u_en0 : F.h_old_en0 + F.v_en0;
                int max_H = h_en0, max_t = en0;
                if (rr > 0) {
                    // ksw2_extd2_avx.c:643-708: four interleaved lanes over
                    // [st0, en1) keep their first strict maximum and are
                    // folded lane 0..3 with strict '<', starting from H[en0];
                    // then a strict-'>' scalar tail.  Encoded as: larger H
                    // wins, equal H goes to the smaller rank (en0 = 0).
                    int bh = h_en0;
                    uint32_t br = 0;
                    for (int k = 0; k < nw; ++k)
                        if (F.best_h[k] > bh || (F.best_h[k] == bh && F.best_rank[k] < br))
                            bh = F.best_h[k], br = F.best_rank[k];
                    max_H = bh;
                    if (br == 0) max_t = en0;
                    else if (br - 1 < (4u << 29)) max_t = st0 + 4 * (int)((br - 1) & ((1u << 29) - 1)) + (int)((br - 1) >> 29);
                    else max_t = p_en1 + (int)(br - 1 - (4u << 29));
                }
                const int h_st0 = st0 == en0 ? h_en0 : F.h_st0;
                if (en0 == tlen - 1 && h_en0 >= z.mte) z.mte = h_en0, z.mte_q = rr - en0;
                if (rr - st0 == qlen - 1 && h_st0 >= z.mqe) z.mqe = h_st0, z.mqe_t = st0;
                if (en0 % nthr == tid) H[en0 & mk] = h_en0;  // owner keeps H[en0]
                if (apply_zdrop(z, max_H, rr, max_t, jb.zdrop, e2)) stop = true;
                else if (rr == qlen + tlen - 2 && en0 == tlen - 1) z.score = h_en0;
            } else {
                // approximate max along one path (ksw2_extd2_sse.c:376-392)
                const int st0 = p_st0, en0 = p_en0;
                if (rr > 0) {
                    const bool in0 = last_H0_t >= st0 && last_H0_t <= en0;
                    const bool in1 = last_H0_t + 1 >= st0 && last_H0_t + 1 <= en0;
                    if (in0 && in1) {
                        if (F.v_h0 > F.u_h0) H0 += F.v_h0;
                        else H0 += F.u_h0, ++last_H0_t;
                    } else if (in0) {
                        H0 += F.v_h0;
                    } else {
                        ++last_H0_t, H0 += F.u_h0;
                    }
                } else {
                    H0 = F.v_h0 - sc.qe_orig, last_H0_t = 0;
                }
                if ((flag & kFlagApproxDrop) && apply_zdrop(z, H0, rr, last_H0_t, jb.zdrop, e2)) stop = true;
                else if (rr == qlen + tlen - 2 && en0 == tlen - 1) z.score = H0;
            }
            if (stop) break;
            // progress hook: thread 0 sampled the host's stop flag at
            // anti-diagonal rr (every 256, as the reference polls its
            // callback); every thread sees the same sample after the barrier
            if (a.stop && (rr & 255) == 0 && F.stop) {
                z.stopped = 1;
                break;
            }
        }
        if (r == nr) break;

        // ---- band of anti-diagonal r (ksw2_extd2_sse.c:147-156)
        int st = 0, en = tlen - 1;
        if (st < r - qlen + 1) st = r - qlen + 1;
        if (en > r) en = r;
        if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
        if (en > (r + w) >> 1) en = (r + w) >> 1;
        if (st > en) {
            z.zdropped = 1;
            break;
        }
        const int st0 = st, en0 = en;
        const int st16 = st0 / 16 * 16, en16 = (en0 + 16) / 16 * 16 - 1, stB = st0 / 64 * 64;
        const int en1 = st0 + (en0 - st0) / 4 * 4;
        const int cr = r & 1;
        // parity buffers: anti-diagonal r-1's x, v, x2 (read), r's (written)
        const int8_t *Xp = cr ? X0 : X1, *Vp = cr ? V0 : V1, *X2p = cr ? XX0 : XX1;
        int8_t *Xc = cr ? X1 : X0, *Vc = cr ? V1 : V0, *X2c = cr ? XX1 : XX0;
        // boundary at lane st16 (ksw2_extd2_sse.c:158-168)
        int x1, x21, v1;
        if (st16 > 0) {
            if (st16 - 1 >= last_st && st16 - 1 <= last_en)
                x1 = Xp[(st16 - 1) & mk], x21 = X2p[(st16 - 1) & mk], v1 = Vp[(st16 - 1) & mk];
            else
                x1 = w8(-q - e), x21 = w8(-q2 - e2), v1 = w8(-q - e);
        } else {
            x1 = w8(-q - e), x21 = w8(-q2 - e2), v1 = edge_i8(sc, r);
        }
        // AVX-512 lane wrap at a non-64-aligned 16-rounded start
        // (ksw2_extd2_avx.c:72-85): lane stB reads position stB+15
        const bool wrap = st16 != stB;
        int wx = 0, wvv = 0, wx2 = 0;
        if (wrap) wx = Xp[(stB + 15) & mk], wvv = Vp[(stB + 15) & mk], wx2 = X2p[(stB + 15) & mk];
        const bool inject = en16 >= r;  // y[r], y2[r], u[r] re-initialised
        const int sc_end = generic ? en0 : min(st0 + (en0 - st0 + 16) / 16 * 16 - 1, tlenB - 1);
        const int t_end = max(en16, sc_end);
        const int qoff = qlen - 1 - r;
        if (with_cigar && tid == 0) off[r] = stB, off_end[r] = en16;

        int bh = INT32_MIN;
        uint32_t br = 0xffffffffu;
        Fin &F = fin[cr];
        if (a.stop && (r & 255) == 0 && tid == 0)
            F.stop = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        int t = stB + ((tid - stB) % nthr + nthr) % nthr;
        for (; t <= t_end; t += nthr) {
            int sct = 0;
            const bool in_sc = t >= st0 && t <= sc_end;
            if (in_sc) {
                const int sq = sf[t], sk = qr[qoff + t];
                if (!generic) {
                    sct = sq == sk ? sc_mch : sc_mis;
                    if (w8(sq) == m1 || w8(sk) == m1) sct = sc_N;
                } else {
                    sct = gs.mat[(sq * m + sk) & (kMaxM * kMaxM - 1)];
                }
                s[t & mk] = (int8_t)sct;
            }
            if (t > en16) continue;
            // ---- one cell (ksw2_extd2_sse.c:43-71, 237-331)
            int xl, vl, x2l;
            if (t == st16) xl = x1, vl = v1, x2l = x21;
            else if (wrap && t == stB) xl = wx, vl = wvv, x2l = wx2;
            else xl = Xp[(t - 1) & mk], vl = Vp[(t - 1) & mk], x2l = X2p[(t - 1) & mk];
            int ut, yt, y2t;
            if (inject && t == r) ut = edge_i8(sc, r), yt = w8(-q - e), y2t = w8(-q2 - e2);
            else ut = u[t & mk], yt = y[t & mk], y2t = y2[t & mk];
            int zz = in_sc ? sct : (int)s[t & mk];
            int A = w8(xl + vl), B = w8(yt + ut), A2 = w8(x2l + vl), B2 = w8(y2t + ut);
            int d = 0;
            if (!with_cigar) {
                zz = max(max(max(max(zz, A), B), A2), B2);
            } else if (!right) {
                d = A > zz ? 1 : 0;  zz = max(zz, A);
                d = B > zz ? 2 : d;  zz = max(zz, B);
                d = A2 > zz ? 3 : d; zz = max(zz, A2);
                d = B2 > zz ? 4 : d; zz = max(zz, B2);
            } else {
                d = zz > A ? 0 : 1;  zz = max(zz, A);
                d = zz > B ? d : 2;  zz = max(zz, B);
                d = zz > A2 ? d : 3; zz = max(zz, A2);
                d = zz > B2 ? d : 4; zz = max(zz, B2);
            }
            zz = min(zz, sc_mch);
            const int un = w8(zz - vl), vn = w8(zz - ut);
            int tmp = w8(zz - q);
            A = w8(A - tmp), B = w8(B - tmp);
            tmp = w8(zz - q2);
            A2 = w8(A2 - tmp), B2 = w8(B2 - tmp);
            int xn, yn, x2n, y2n;
            if (!right) {
                xn = w8((A > 0 ? A : 0) - qe8);    d |= A > 0 ? 0x08 : 0;
                yn = w8((B > 0 ? B : 0) - qe8);    d |= B > 0 ? 0x10 : 0;
                x2n = w8((A2 > 0 ? A2 : 0) - qe28); d |= A2 > 0 ? 0x20 : 0;
                y2n = w8((B2 > 0 ? B2 : 0) - qe28); d |= B2 > 0 ? 0x40 : 0;
            } else {
                xn = w8((A < 0 ? 0 : A) - qe8);    d |= A < 0 ? 0 : 0x08;
                yn = w8((B < 0 ? 0 : B) - qe8);    d |= B < 0 ? 0 : 0x10;
                x2n = w8((A2 < 0 ? 0 : A2) - qe28); d |= A2 < 0 ? 0 : 0x20;
                y2n = w8((B2 < 0 ? 0 : B2) - qe28); d |= B2 < 0 ? 0 : 0x40;
            }
            const int tm = t & mk;
            u[tm] = (int8_t)un, Vc[tm] = (int8_t)vn, Xc[tm] = (int8_t)xn;
            y[tm] = (int8_t)yn, X2c[tm] = (int8_t)x2n, y2[tm] = (int8_t)y2n;
            if (tracks) {
                p[(uint64_t)r * ncolB + (uint64_t)(t - stB)] = (uint8_t)vn;
                p2[(uint64_t)r * ncolB + (uint64_t)(t - stB)] = (uint8_t)un;
            } else if (with_cigar) {
                p[(uint64_t)r * ncolB + (uint64_t)(t - stB)] = (uint8_t)d;
            }
            // ---- H bookkeeping (ksw2_extd2_sse.c:332-375)
            if (!approx) {
                if (t == en0) F.u_en0 = un, F.v_en0 = vn, F.h_old_en0 = H[tm];
                if (r > 0) {
                    if (t == en0 - 1) F.h_old_en0m1 = H[tm];
                    if (t >= st0 && t < en0) {
                        const int hn = H[tm] + vn;
                        H[tm] = hn;
                        if (t == st0) F.h_st0 = hn;
                        const int k = t - st0;
                        const uint32_t rank = t < en1 ? 1u + ((uint32_t)(k & 3) << 29) + (uint32_t)(k >> 2)
                                                      : 1u + (4u << 29) + (uint32_t)(t - en1);
                        if (hn > bh || (hn == bh && rank < br)) bh = hn, br = rank;
                    }
                }
            } else {
                if (t == last_H0_t) F.v_h0 = vn;
                if (t == last_H0_t + 1) F.u_h0 = un;
            }
        }
        // values the bookkeeping needs at positions outside [stB, en16]
        if (!approx && r > 0 && en0 >= 1 && en0 - 1 < stB && (en0 - 1) % nthr == tid)
            F.h_old_en0m1 = H[(en0 - 1) & mk];
        if (approx && r > 0) {
            const int t1 = last_H0_t + 1;
            if ((t1 < stB || t1 > en16) && t1 < tlenB && t1 % nthr == tid) F.u_h0 = u[t1 & mk];
        }
        if (!approx) {
            // wave reduction of (H, rank): larger H, then smaller rank
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int oh = __shfl_xor(bh, o, kWave);
                const uint32_t orank = (uint32_t)__shfl_xor((int)br, o, kWave);
                if (oh > bh || (oh == bh && orank < br)) bh = oh, br = orank;
            }
            if (lane == 0) F.best_h[wv] = bh, F.best_rank[wv] = br;
        }
        p_st0 = st0, p_en0 = en0, p_en1 = en1;
        last_st = st16, last_en = en16;
        if (RING) {
            // reset the slots of the positions anti-diagonal r+1 reaches first
            // (its band's right end moves by at most 17 past this one's)
            int en_n = tlen - 1;
            if (en_n > r + 1) en_n = r + 1;
            if (en_n > (r + 1 + w) >> 1) en_n = (r + 1 + w) >> 1;
            const int reach = min(max((en_n + 16) / 16 * 16 - 1, en_n / 16 * 16 + 15) + 16, tlenB - 1);
            for (int t = init_end + 1 + tid; t <= reach; t += nthr) {
                const int tm = t & mk;
                u[tm] = y[tm] = X0[tm] = X1[tm] = V0[tm] = V1[tm] = (int8_t)w8(-q - e);
                y2[tm] = XX0[tm] = XX1[tm] = (int8_t)w8(-q2 - e2);
                s[tm] = 0;
                H[tm] = KSW_NEG_INF_I;
            }
            init_end = max(init_end, reach);
        }
        __syncthreads();
    }
    if (tid == 0) {
        KswOut o;
        o.score = z.score, o.mqe = z.mqe, o.mqe_t = z.mqe_t, o.mte = z.mte, o.mte_q = z.mte_q;
        o.max = z.max, o.max_t = z.max_t, o.max_q = z.max_q;
        o.zdropped = z.zdropped, o.stopped = z.stopped, o.reach_end = 0, o.n_cigar = 0;
        a.out[jb.out] = o;
    }
}

// Backtrack over the direction rows of the general kernel: ksw_backtrack with
// is_rot = 1 (minimap2/ksw2.h:160-192), including the forcing of cells
// outside [off[r], off_end[r]], from the start cell of
// minimap2/ksw2_extd2_sse.c:398-418.  One thread per pair.
// Where the walk of a general pair starts (minimap2/ksw2_extd2_sse.c:
// the end-of-DP cases before ksw_backtrack): adjusts o's score and
// reach_end as the reference does; (-1, -1) when there is no CIGAR.
__device__ __forceinline__ void backtrack_start(const GenJob &jb, KswOut &o, int &i0, int &j0)
{
    const int flag = jb.flag, qlen = jb.qlen, tlen = jb.tlen;
    i0 = -1, j0 = -1;
    if (flag & kFlagScoreOnly) return;
    if (!o.stopped && !o.zdropped && (flag & 0x800)) {          // SEMIGLOBAL_END
        if (o.mqe > o.mte) o.score = o.mqe, i0 = o.mqe_t, j0 = qlen - 1;
        else o.score = o.mte, i0 = tlen - 1, j0 = o.mte_q;
    } else if (!o.stopped && !o.zdropped && !(flag & 0x40)) {   // global
        i0 = tlen - 1, j0 = qlen - 1;
    } else if (!o.zdropped && (flag & 0x40) && o.mqe + jb.end_bonus > o.max) {  // EXTZ_ONLY
        o.reach_end = 1, i0 = o.mqe_t, j0 = qlen - 1;
    } else if (o.max_t >= 0 && o.max_q >= 0) {
        i0 = o.max_t, j0 = o.max_q;
    }
}

// One step of ksw_backtrack's state machine (minimap2/ksw2.h:142-170):
// `tmp` is the direction byte (0 when forced or outside the row), `force`
// 1 or 2 outside [off, off_end], else -1.  Returns the CIGAR op: 0 (i and j
// step back), 2 (i only), 1 (j only).
__device__ __forceinline__ int backtrack_op(int tmp, int force, int &state)
{
    // a state continues when its bit (tmp >> (state + 2)) is set, else the
    // byte's own state (tmp & 7) takes over
    const bool cont = state != 0 && (tmp >> (state + 2) & 1);
    state = force >= 0 ? force : cont ? state : tmp & 7;
    return state == 0 ? 0 : state == 1 || state == 3 ? 2 : 1;
}

__global__ void __launch_bounds__(64) ksw_general_backtrack_kernel(GenBacktrackArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.njobs) return;
    const GenJob jb = a.jobs[k];
    const int flag = jb.flag, qlen = jb.qlen, tlen = jb.tlen;
    if ((flag & kFlagSingletrack) && !(flag & kFlagScoreOnly)) return;  // ksw_general_walk_kernel
    if (qlen + tlen >= a.tiled_min) return;                             // ksw_general_backtrack_tiled_kernel
    KswOut o = a.out[jb.out];
    int i0, j0;
    backtrack_start(jb, o, i0, j0);
    int n = 0;
    if (i0 >= 0 || j0 >= 0) {
        const uint8_t *p = a.dir + jb.dir_off;
        const int32_t *off = a.offs + jb.off_off, *off_end = off + (qlen + tlen - 1);
        uint32_t *cig = a.cig + jb.cig_off;
        int i = i0, j = j0, state = 0;
        uint32_t last = 0xffffffffu;
        auto push = [&](uint32_t op, int len) {
            if (n > 0 && op == last) cig[n - 1] += (uint32_t)len << 4;
            else cig[n++] = (uint32_t)len << 4 | op, last = op;
        };
        while (i >= 0 && j >= 0) {
            const int r = i + j;
            int force = -1;
            if (i < off[r]) force = 2;
            if (i > off_end[r]) force = 1;
            // (rows a path visits were all written; the width test only
            // keeps a damaged row from reading outside its own bytes)
            const bool inrow = (uint32_t)(i - off[r]) < (uint32_t)jb.ncolB;
            const int tmp = force < 0 && inrow ? p[(uint64_t)r * jb.ncolB + (uint64_t)(i - off[r])] : 0;
            const int op = backtrack_op(tmp, force, state);
            i -= op != 1, j -= op != 2;
            push(op, 1);
        }
        if (i >= 0) push(2, i + 1);
        if (j >= 0) push(1, j + 1);
        if (!(flag & 0x80))
            for (int x = 0; x < (n >> 1); ++x) {
                const uint32_t t = cig[x];
                cig[x] = cig[n - 1 - x];
                cig[n - 1 - x] = t;
            }
    }
    o.n_cigar = n;
    a.out[jb.out] = o;
}

// The same walk for long pairs, one wave per pair.  A step moves to
// anti-diagonal r-1 or r-2 and to target position i or i-1, so the next 128
// anti-diagonals stay within positions [i-127, i] of the current (r, i):
// the wave stages those 128 rows (132 bytes each, from the 4-byte boundary
// at or below i-127, zero outside the row; lane L stages rows r-L and
// r-64-L) and their off/off_end in LDS, and the walk reads LDS until it
// leaves the tile.  The walk is uniform (scalar) but for runs of diagonal
// moves and runs of gap moves, each taken at once by a ballot over the
// lanes; the CIGAR runs go out one store per run, and the wave reverses
// them at the end.
constexpr int kTileRows = 128, kTileDw = 33;

__global__ void __launch_bounds__(64) ksw_general_backtrack_tiled_kernel(GenBacktrackArgs a)
{
    __shared__ uint32_t tile[kTileRows * kTileDw];
    // per staged row: off[], off_end[] and the position of its first byte
    __shared__ int t_ro[kTileRows], t_re[kTileRows], t_base[kTileRows];
    const int lane = threadIdx.x;
    const GenJob jb = a.jobs[blockIdx.x];
    const int flag = __builtin_amdgcn_readfirstlane(jb.flag);
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    if ((flag & kFlagSingletrack) && !(flag & kFlagScoreOnly)) return;
    if (qlen + tlen < a.tiled_min) return;
    KswOut o = a.out[jb.out];
    int i0, j0;
    backtrack_start(jb, o, i0, j0);
    int n = 0;
    if (i0 >= 0 || j0 >= 0) {
        const int ncolB = __builtin_amdgcn_readfirstlane(jb.ncolB);
        const uint32_t *p32 = (const uint32_t *)(a.dir + jb.dir_off);
        const int32_t *off = a.offs + jb.off_off, *off_end = off + (qlen + tlen - 1);
        uint32_t *cig = a.cig + jb.cig_off;
        const uint8_t *tb = (const uint8_t *)tile;
        int i = __builtin_amdgcn_readfirstlane(i0), j = __builtin_amdgcn_readfirstlane(j0), state = 0;
        int run_op = -1, run_len = 0;
        int r0 = -1, ws = 0;  // tile: anti-diagonals (r0 - 128, r0], positions from ws
        // lane L: off[], off_end[] of rows r0 - L (ro[0] ...) and r0 - 64 - L
        // (ro[1] ...) and the position of their first staged bytes
        int ro[2] = {0, 0}, re[2] = {-1, -1}, base[2] = {0, 0};
        const int nd = ncolB >> 2;
        while (i >= 0 && j >= 0) {
            const int r = i + j;
            if (r0 < 0 || r <= r0 - kTileRows || i < ws) {
                r0 = r, ws = i - 127;
                __syncthreads();  // the walk's last reads of the old tile are done
                // all loads of both rows in flight at once (clamped in-row
                // addresses), then the bytes outside the row zeroed
                uint32_t v[2][kTileDw];
                int d0[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int row = r0 - kWave * h - lane;
                    ro[h] = 0, re[h] = -1, base[h] = 0, d0[h] = 0;
                    if (row >= 0) {
                        ro[h] = off[row], re[h] = off_end[row];
                        d0[h] = (ws - ro[h]) >> 2;  // floor
                        base[h] = ro[h] + 4 * d0[h];
                        const uint32_t *src = p32 + (uint64_t)row * (uint32_t)nd;
#pragma unroll
                        for (int q = 0; q < kTileDw; ++q) v[h][q] = src[min(max(d0[h] + q, 0), nd - 1)];
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int Lr = kWave * h + lane;
                    t_ro[Lr] = ro[h], t_re[Lr] = re[h], t_base[Lr] = base[h];
                    if (r0 - Lr >= 0) {
#pragma unroll
                        for (int q = 0; q < kTileDw; ++q) {
                            const int d = d0[h] + q;
                            tile[Lr * kTileDw + q] = d >= 0 && d < nd ? v[h][q] : 0u;
                        }
                    }
                }
                __syncthreads();
            }
            const int L = r0 - r;
            if (state == 0) {
                // A run of diagonal moves at once: from state 0 the walk
                // stays in state 0 and moves to (i-1, j-1), anti-diagonal
                // r-2, while a cell is not forced and its argmax (low three
                // bits) is 0.  Lane k tests cell (i-k, j-k) of the tile (row
                // L+2k); the run is the lanes before the first that fails.
                const int Lk = L + 2 * lane, ik = i - lane;
                const int Lc = min(Lk, kTileRows - 1);
                const int o_k = t_ro[Lc], e_k = t_re[Lc], b_k = t_base[Lc];
                int fail = 1;
                if (Lk < kTileRows && ik >= 0 && lane <= j && ik >= ws && ik >= o_k && ik <= e_k) {
                    const int byte = tb[Lk * kTileDw * 4 + (ik - b_k)];
                    const int tmp = (uint32_t)(ik - o_k) < (uint32_t)ncolB ? byte : 0;
                    fail = (tmp & 7) != 0;
                }
                const uint64_t fb = __ballot(fail);
                const int run = fb ? (int)__builtin_ctzll(fb) : kWave;
                if (run > 0) {
                    i -= run, j -= run;
                    if (run_op == 0) {
                        run_len += run;
                    } else {
                        if (run_len > 0 && lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                        if (run_len > 0) ++n;
                        run_op = 0, run_len = run;
                    }
                    continue;
                }
            } else {
                // A run of gap moves at once: in state s (1-4) the walk
                // keeps s while a cell's continuation bit (s + 2) is set, or
                // while the cell is forced into s itself, and moves to
                // (i-1, j) for s = 1, 3 or (i, j-1) for s = 2, 4, anti-
                // diagonal r-1.  Lane k tests the k-th cell of the run (row
                // L+k).
                const int s = state;
                const bool di = s == 1 || s == 3;
                const int Lk = L + lane, ik = di ? i - lane : i;
                const int Lc = min(Lk, kTileRows - 1);
                const int o_k = t_ro[Lc], e_k = t_re[Lc], b_k = t_base[Lc];
                int fail = 1;
                if (Lk < kTileRows && lane <= (di ? i : j)) {
                    const int force = ik > e_k ? 1 : ik < o_k ? 2 : -1;
                    const int byte = tb[Lk * kTileDw * 4 + (ik - b_k)];
                    const int tmp = force < 0 && (uint32_t)(ik - o_k) < (uint32_t)ncolB ? byte : 0;
                    fail = force >= 0 ? force != s : !(tmp >> (s + 2) & 1);
                }
                const uint64_t fb = __ballot(fail);
                const int run = fb ? (int)__builtin_ctzll(fb) : kWave;
                if (run > 0) {
                    const int op = di ? 2 : 1;
                    if (di) i -= run;
                    else j -= run;
                    if (run_op == op) {
                        run_len += run;
                    } else {
                        if (run_len > 0 && lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                        if (run_len > 0) ++n;
                        run_op = op, run_len = run;
                    }
                    continue;
                }
            }
            const int h = L >> 6, Lh = L & (kWave - 1);
            const int o_r = __builtin_amdgcn_readlane(h ? ro[1] : ro[0], Lh);
            const int e_r = __builtin_amdgcn_readlane(h ? re[1] : re[0], Lh);
            const int b_r = __builtin_amdgcn_readlane(h ? base[1] : base[0], Lh);
            // (the row's byte is read whether or not it is used: the tile
            // holds zeros outside the row, and i - b_r is within the tile)
            const int byte = __builtin_amdgcn_readfirstlane((int)tb[L * kTileDw * 4 + (i - b_r)]);
            const int force = i > e_r ? 1 : i < o_r ? 2 : -1;
            const int tmp = force < 0 && (uint32_t)(i - o_r) < (uint32_t)ncolB ? byte : 0;
            const int op = backtrack_op(tmp, force, state);
            i -= op != 1, j -= op != 2;
            if (op == run_op) {
                ++run_len;
            } else {
                if (run_len > 0 && lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                if (run_len > 0) ++n;
                run_op = op, run_len = 1;
            }
        }
        // the tails merge with the last run (push of ksw_backtrack)
        if (i >= 0) {
            if (run_op == 2) run_len += i + 1;
            else {
                if (run_len > 0 && lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                if (run_len > 0) ++n;
                run_op = 2, run_len = i + 1;
            }
        }
        if (j >= 0) {
            if (run_op == 1) run_len += j + 1;
            else {
                if (run_len > 0 && lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                if (run_len > 0) ++n;
                run_op = 1, run_len = j + 1;
            }
        }
        if (run_len > 0 && lane == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
        if (run_len > 0) ++n;
        if (!(flag & 0x80)) {
            // lane 0's stores complete before the wave reads them back with
            // device-scope loads (past the CU's L1)
            __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
            __syncthreads();
            for (int x = lane; x < (n >> 1); x += kWave) {
                const uint32_t lo = __hip_atomic_load(cig + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t hi = __hip_atomic_load(cig + n - 1 - x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                cig[x] = hi;
                cig[n - 1 - x] = lo;
            }
        }
    }
    if (lane == 0) {
        o.n_cigar = n;
        a.out[jb.out] = o;
    }
}

// Singletrack tracks of one general pair: reference layout, two row arrays
// of n_col*64 bytes per anti-diagonal, off[]/off_end[] per anti-diagonal.
struct GenTracks {
    const uint8_t *p, *p2;
    const int32_t *off, *off_end;
    const uint8_t *qs, *ts;
    const int8_t *mat;
    int32_t ncolB, m;
    __device__ int32_t get(int track, int i, int j) const
    {
        const int d = i + j;
        const int o = off[d];
        // (a stored range always fits its row; the width test only keeps a
        // damaged row from reading outside its own bytes)
        if (i < o || i > off_end[d] || i - o >= ncolB) return KSW_NEG_INF_I;
        return (int32_t)(int8_t)(track ? p2 : p)[(uint64_t)d * ncolB + (uint64_t)(i - o)];
    }
    __device__ int forced(int i, int j) const
    {
        const int d = i + j;
        return i < off[d] ? 1 : i > off_end[d] ? 2 : 0;
    }
    __device__ int32_t cost(int i, int j) const { return mat[(ts[i] * m + qs[j]) & (kMaxM * kMaxM - 1)]; }
};

// One wavefront per general pair with KSW_EZ_SINGLETRACK: the dispatch of
// minimap2/ksw2_extd2_avx.c:752-760 and the Singletrack walk (track_walk);
// other pairs return at once.
__global__ void __launch_bounds__(64) ksw_general_walk_kernel(GenWalkArgs a)
{
    const GenJob jb = a.jobs[blockIdx.x];
    const int flag = __builtin_amdgcn_readfirstlane(jb.flag);
    if (!(flag & kFlagSingletrack) || (flag & kFlagScoreOnly)) return;
    KswOut o = a.out[jb.out];
    const int qlen = jb.qlen, tlen = jb.tlen;
    int n = 0;
    if (!o.stopped && !o.zdropped && !(flag & (kFlagExtzOnly | kFlagSemiglobalEnd))) {
        GenTracks A;
        A.p = a.dir + jb.dir_off;
        A.p2 = A.p + (uint64_t)(qlen + tlen - 1) * (uint64_t)jb.ncolB;
        A.off = a.offs + jb.off_off, A.off_end = A.off + (qlen + tlen - 1);
        A.qs = a.seq + jb.q_off, A.ts = a.seq + jb.t_off;
        A.mat = a.sc.mat, A.ncolB = jb.ncolB, A.m = a.sc.m;
        const WalkGaps g = {-a.sc.k.q, -a.sc.k.e, -a.sc.k.q2, -a.sc.k.e2};
        n = track_walk(A, qlen, tlen, g, a.cig + jb.cig_off);
        if (n < 0) o.zdropped = 1;
    } else if (!o.stopped && !o.zdropped) {
        o.zdropped = 1;
    }
    o.n_cigar = n;
    if (threadIdx.x == 0) a.out[jb.out] = o;
}

hipError_t launch_general(const GenArgs &a, int threads, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (threads < kWave || threads > kWave * kMaxWaves || threads % kWave) return hipErrorInvalidValue;
    if (a.ring > 0) {
        if (a.ring < 256 || a.ring > kGenRingMax || (a.ring & (a.ring - 1))) return hipErrorInvalidValue;
        static bool attr_set = false;  // allow the largest ring's dynamic LDS (once)
        if (!attr_set) {
            const hipError_t e = hipFuncSetAttribute((const void *)ksw_general_kernel<true>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)gen_ring_bytes(kGenRingMax));
            if (e != hipSuccess) return e;
            attr_set = true;
        }
        hipLaunchKernelGGL(ksw_general_kernel<true>, dim3(a.njobs), dim3(threads), gen_ring_bytes(a.ring), st, a);
    } else {
        hipLaunchKernelGGL(ksw_general_kernel<false>, dim3(a.njobs), dim3(threads), 0, st, a);
    }
    return hipGetLastError();
}

__global__ void __launch_bounds__(64) ksw_cigar_compact_kernel(const uint32_t *arena, uint32_t *dense,
                                                               const CigarMove *moves)
{
    const CigarMove m = moves[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < m.n; i += 64) dense[m.dst + i] = arena[m.src + i];
}

hipError_t launch_cigar_compact(const uint32_t *arena, uint32_t *dense, const CigarMove *moves, int n_moves,
                                hipStream_t st)
{
    if (n_moves <= 0) return hipSuccess;
    hipLaunchKernelGGL(ksw_cigar_compact_kernel, dim3(n_moves), dim3(64), 0, st, arena, dense, moves);
    return hipGetLastError();
}

hipError_t launch_general_backtrack(const GenBacktrackArgs &a, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(ksw_general_backtrack_kernel, dim3((a.njobs + 63) / 64), dim3(64), 0, st, a);
    hipLaunchKernelGGL(ksw_general_backtrack_tiled_kernel, dim3(a.njobs), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_general_walk(const GenWalkArgs &a, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(ksw_general_walk_kernel, dim3(a.njobs), dim3(kWave), 0, st, a);
    return hipGetLastError();
}

}  // namespace kswhip
