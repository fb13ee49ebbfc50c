// ksw2_band.hip -- gfx950 kernel for banded ksw_extd2 calls inside the
// stripe kernel's exactness domain (match 0, mismatch <= 0, the scoring
// bounds that keep every int8 intermediate in range, no z-drop, no
// wildcard), typically long pairs under a wide band (w = 4096).
//
// The reference sweeps anti-diagonals r; on each it updates the positions
// t of the 16/64-rounded range [stB, en16] in place
// (minimap2/ksw2_extd2_avx.c:248-745): a cell reads x, v, x2 of position t-1
// and u, y, y2, s of position t as they were left by earlier anti-diagonals
// (stale outside the band), with the boundary rules of
// minimap2/ksw2_extd2_sse.c:158-172 (injection at st16, the AVX-512 lane
// wrap at a non-64-aligned start, row r re-initialised at column 0).  This
// kernel evaluates the same cells in (t, r) coordinates:
//
//  * Stripes.  The positions are cut into stripes of 512; lane k of a wave
//    holds positions t0 + 8k .. t0 + 8k + 7 of its stripe in registers (the
//    in-place state u, v, x, y, x2, y2, s and H of each) and steps through
//    the anti-diagonals that reach the stripe.  Within a lane the positions
//    are updated from the highest down, so each still reads its upper
//    neighbour's value from the previous anti-diagonal; the lane's top
//    position takes the lane above's bottom values through a DPP shift, and
//    lane 0 takes the previous stripe's bottom position from a per-pair
//    buffer indexed by r.  A position outside [stB, en16] is not updated:
//    its stale value is simply the register it keeps.
//  * Because stB is a multiple of 64 and en16 + 1 of 16, a lane's eight
//    positions are all inside [stB, en16] or all outside: the commit and the
//    direction-byte store are per lane (eight bytes, aligned).
//  * Teams.  One workgroup of W waves per pair; wave w takes stripes w,
//    w + W, ...; stripe s trails stripe s-1 by at least 128 anti-diagonals
//    through workgroup-scope progress counters in LDS (as the stripe kernel).
//    The buffer is double-buffered by stripe parity; that lead keeps a
//    producer two stripes ahead from overwriting an entry before it is read.
//  * H bookkeeping (minimap2/ksw2_extd2_sse.c:332-375): H[en0] from the
//    position above plus u, H[t] += v on [st0, en0), mte on position tlen-1,
//    mqe where column qlen-1 is the band's start, the score at the last
//    anti-diagonal.  The maximum is not tracked: in this domain H never
//    exceeds ez->max's initial 0 (match score 0), so ksw_apply_zdrop never
//    changes it; every position checks that and a pair that ever sees H > 0
//    is flagged for the general kernel (redo), which tracks the maximum.
//  * Band exit (st > en) is pure geometry: the host passes the number of
//    anti-diagonals the reference completes.
//
// Outputs: direction bytes in the reference's row layout
// (p[r * n_col*64 + t - off[r]], off[r] = stB, off_end[r] = en16), so the
// general backtrack kernel walks them; the KswOut fields; the redo flag.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ksw2_common.h"
#include "ksw2_kernels.h"

namespace kswhip {

namespace {

constexpr int R = kRowsPerLane;  // positions per lane
constexpr int SR = kWave * R;    // positions per stripe
constexpr int kBlock = 64;       // anti-diagonals between progress publications
constexpr int kLead = 128;       // lead a stripe waits for on its predecessor
constexpr int kDppWaveShr1 = 0x138;

struct Band {
    int st0, en0, st16, en16, stB, sc_end;
};

// Band of anti-diagonal r (minimap2/ksw2_extd2_sse.c:147-156) and the
// ranges derived from it: the 16-rounded start/end, the 64-block start and
// the last position whose score is written (16-byte groups from st0, or
// exactly en0 with KSW_EZ_GENERIC_SC, ksw2_extd2_sse.c:173-193).
__device__ __forceinline__ Band band_of(int r, int qlen, int tlen, int w, int tlenB, bool generic)
{
    int st = 0, en = tlen - 1;
    if (st < r - qlen + 1) st = r - qlen + 1;
    if (en > r) en = r;
    if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
    if (en > (r + w) >> 1) en = (r + w) >> 1;
    Band b;
    b.st0 = st, b.en0 = en;
    b.st16 = st / 16 * 16, b.en16 = (en + 16) / 16 * 16 - 1, b.stB = st / 64 * 64;
    b.sc_end = generic ? en : min(st + (en - st + 16) / 16 * 16 - 1, tlenB - 1);
    return b;
}

// Anti-diagonals that can touch positions [t0, t0 + SR): a position t is
// computed only when stB <= t <= en16 and scored only when st0 <= t <=
// sc_end, which bound r to [2(t-15) - w - 1, 2(t+63) + w].
__device__ __forceinline__ void stripe_range(int t0, int w, int r_end, int &lo, int &hi)
{
    lo = max(0, 2 * t0 - w - 40);
    hi = min(r_end - 1, 2 * (t0 + SR - 1) + w + 130);
}

__device__ __forceinline__ int shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, kDppWaveShr1, 0xF, 0xF, false); }

__device__ __forceinline__ void publish(int32_t *slot, int value)
{
    __hip_atomic_store(slot, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Spins until the producer's published count reaches `need` (uniform exit
// test on lane 0's value).  A producer always ends a stripe by publishing a
// count past every need of its consumer, so the wait ends.
__device__ __forceinline__ int wait_count(const int32_t *slot, int need, int known)
{
    if (known >= need) return known;
    int seen = __builtin_amdgcn_readfirstlane(__hip_atomic_load(slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    while (seen < need) {
        __builtin_amdgcn_s_sleep(1);
        seen = __builtin_amdgcn_readfirstlane(__hip_atomic_load(slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    }
    return seen;
}

__device__ __forceinline__ int pack3(int x, int v, int x2)
{
    return (x & 0xff) | (v & 0xff) << 8 | (x2 & 0xff) << 16;
}
__device__ __forceinline__ int byte_s(int w, int k) { return (int)(int8_t)(uint8_t)((unsigned)w >> (8 * k)); }

}  // namespace

// Up to kBandTeam waves per pair (two per SIMD), so each wave may hold up to
// 256 VGPRs: the eight positions' state and both step forms fit unspilled.
template <bool RIGHT, bool SCORE_ONLY>
__global__ void __launch_bounds__(kWave * kBandTeam) ksw_band_kernel(BandArgs a)
{
    __shared__ int32_t prog[kBandTeam];
    __shared__ int32_t red[kBandTeam][8];
    const int nthr = blockDim.x, tid = threadIdx.x, lane = tid % kWave;
    const int W = __builtin_amdgcn_readfirstlane(nthr / kWave), wv = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int job = __builtin_amdgcn_readfirstlane(a.order[blockIdx.x]);
    const GenJob jb = a.jobs[job];
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    const int bw = __builtin_amdgcn_readfirstlane(jb.w), flag = __builtin_amdgcn_readfirstlane(jb.flag);
    const int tlenB = __builtin_amdgcn_readfirstlane(jb.tlenB), r_end = __builtin_amdgcn_readfirstlane(jb.r_end);
    const int ncolB = jb.ncolB;
    const bool generic = flag & 0x04;
    const int nr = qlen + tlen - 1;
    const KswScore &sc = a.sc.k;
    const int G1 = sc.G1, G2 = sc.G2, q = sc.q, q2 = sc.q2, mch = sc.mch, mis = sc.mis;
    const uint8_t *qs = a.seq + jb.q_off, *ts = a.seq + jb.t_off;
    uint8_t *p = a.dir + jb.dir_off;
    int32_t *off = a.offs + jb.off_off, *off_end = off + nr;
    int32_t *buf = a.bnd + jb.bnd_off;  // two buffers of 2 * nr dwords

    if (!SCORE_ONLY)
        for (int r = tid; r < r_end; r += nthr) {
            const Band b = band_of(r, qlen, tlen, bw, tlenB, generic);
            off[r] = b.stB, off_end[r] = b.en16;
        }
    if (lane == 0) prog[wv] = 0;
    __syncthreads();

    // per-lane results over this wave's stripes
    int mqe_h = KSW_NEG_INF_I, mqe_t = -1, mte_h = KSW_NEG_INF_I, mte_q = -1, score = KSW_NEG_INF_I, pos = 0;
    const int nstripes = (tlenB + SR - 1) / SR;
    int known = 0;
    for (int s = wv, kk = 0; s < nstripes; s += W, ++kk) {
        const int t0 = s * SR, tl = t0 + lane * R;
        int r_lo, r_hi, p_lo = 0, p_hi = -1;
        stripe_range(t0, bw, r_end, r_lo, r_hi);
        if (s > 0) stripe_range(t0 - SR, bw, r_end, p_lo, p_hi);
        const int pw = (s + W - 1) % W, pbase = (s - 1) / W * (nr + 2);  // producer wave, its count base
        int32_t *out_buf = buf + (uint64_t)(s & 1) * 2 * nr;
        const int32_t *in_buf = buf + (uint64_t)((s + 1) & 1) * 2 * nr;
        // per position: the in-place state and H; the score s as one bit
        // (set: mismatch; s is mch or mis in this domain, 0 at first), the
        // target codes packed four to a register (zero past tlen, as the
        // reference's zero-padded target copy)
        int U[R], Y[R], Y2[R], X[R], V[R], X2[R], H[R];
        uint32_t smask = 0, tcw[2] = {0u, 0u};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            U[k] = Y[k] = X[k] = V[k] = -G1;
            Y2[k] = X2[k] = -G2;
            H[k] = KSW_NEG_INF_I;
            const uint32_t c = tl + k < tlen ? (uint32_t)ts[tl + k] : 0u;
            tcw[k >> 2] |= c << (8 * (k & 3));
        }
        // the predecessor's bottom position once it is final (read once)
        int fx = -G1, fv = -G1, fx2 = -G2, fh = KSW_NEG_INF_I;
        bool have_final = false;
        // query window: byte k of (qwin[1]:qwin[0]) is the query code of
        // position tl + k's column r - (tl + k) (0 outside the query, as the
        // reference's zero-tailed reversed query); one new byte per step
        uint32_t qwin[2] = {0u, 0u};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int j = r_lo - 1 - tl - k;
            const uint32_t c = j >= 0 && j < qlen ? (uint32_t)qs[j] : 0u;
            qwin[k >> 2] |= c << (8 * (k & 3));
        }
        int lst16 = -1, len16 = -1;  // st16, en16 of anti-diagonal r-1
        if (r_lo > 0) {
            const Band bp = band_of(r_lo - 1, qlen, tlen, bw, tlenB, generic);
            lst16 = bp.st16, len16 = bp.en16;
        }
        int hmax = KSW_NEG_INF_I;  // largest H on interior steps (checked against 0 at the end)
        for (int r = r_lo; r <= r_hi; ++r) {
            if (s > 0 && (r - r_lo) % kBlock == 0)
                known = wait_count(prog + pw, pbase + min(r + kLead, p_hi) + 1, known);
            const Band b = band_of(r, qlen, tlen, bw, tlenB, generic);
            {
                const int j = r - tl;
                const uint32_t c = j >= 0 && j < qlen ? (uint32_t)qs[j] : 0u;
                qwin[1] = qwin[1] << 8 | qwin[0] >> 24;
                qwin[0] = qwin[0] << 8 | c;
            }
            // position t0-1 as anti-diagonal r-1 left it (lane 0's upper neighbour)
            int cx = -G1, cv = -G1, cx2 = -G2, ch = KSW_NEG_INF_I;
            if (s > 0 && p_hi >= p_lo && r - 1 >= p_lo) {
                if (r - 1 >= p_hi) {
                    if (!have_final) {
                        const int e = p_hi;
                        const int pk = in_buf[2 * e], hh = in_buf[2 * e + 1];
                        fx = byte_s(pk, 0), fv = byte_s(pk, 1), fx2 = byte_s(pk, 2), fh = hh;
                        have_final = true;
                    }
                    cx = fx, cv = fv, cx2 = fx2, ch = fh;
                } else if (lane == 0) {
                    const int pk = in_buf[2 * (r - 1)], hh = in_buf[2 * (r - 1) + 1];
                    cx = byte_s(pk, 0), cv = byte_s(pk, 1), cx2 = byte_s(pk, 2), ch = hh;
                }
            }
            {
                const int sx = shr1(X[R - 1]), sv = shr1(V[R - 1]), sx2 = shr1(X2[R - 1]), sh = shr1(H[R - 1]);
                if (lane != 0) cx = sx, cv = sv, cx2 = sx2, ch = sh;
            }
            // the AVX-512 lane wrap: position stB reads position stB+15
            // (row 7 of the next lane) from anti-diagonal r-1
            const bool wrap = b.st16 != b.stB;
            int wx = 0, wvv = 0, wx2 = 0;
            if (wrap && b.stB >= t0 && b.stB < t0 + SR) {
                const int src = (b.stB - t0) / R + 1;
                wx = __builtin_amdgcn_readlane(X[R - 1], src);
                wvv = __builtin_amdgcn_readlane(V[R - 1], src);
                wx2 = __builtin_amdgcn_readlane(X2[R - 1], src);
            }
            // interior step: every position of the stripe strictly inside
            // (st0, en0) -- all computed and scored, no boundary rule, no H
            // entry, no mqe/mte/score position -- the lean form below
            if (b.st0 < t0 && t0 + SR - 1 < b.en0) {
                uint32_t dw[2] = {0u, 0u}, sm = 0;
#pragma unroll
                for (int k = R - 1; k >= 0; --k) {
                    const int xin = k > 0 ? X[k - 1] : cx, vin = k > 0 ? V[k - 1] : cv;
                    const int x2in = k > 0 ? X2[k - 1] : cx2;
                    const uint32_t tc = (tcw[k >> 2] >> (8 * (k & 3))) & 0xffu;
                    const uint32_t qc = (qwin[k >> 2] >> (8 * (k & 3))) & 0xffu;
                    const bool mm = tc != qc;
                    sm |= mm ? 1u << k : 0u;
                    int z = mm ? mis : mch, d = 0;
                    int aa = xin + vin, bb = Y[k] + U[k], a2 = x2in + vin, b2 = Y2[k] + U[k];
                    if (SCORE_ONLY) {
                        z = max(max(max(max(z, aa), bb), a2), b2);
                    } else if (!RIGHT) {
                        d = aa > z ? 1 : 0;  z = max(z, aa);
                        d = bb > z ? 2 : d;  z = max(z, bb);
                        d = a2 > z ? 3 : d;  z = max(z, a2);
                        d = b2 > z ? 4 : d;  z = max(z, b2);
                    } else {
                        d = z > aa ? 0 : 1;  z = max(z, aa);
                        d = z > bb ? d : 2;  z = max(z, bb);
                        d = z > a2 ? d : 3;  z = max(z, a2);
                        d = z > b2 ? d : 4;  z = max(z, b2);
                    }
                    z = min(z, mch);
                    const int un = z - vin, vn = z - U[k];
                    int tmp = z - q;
                    aa -= tmp, bb -= tmp;
                    tmp = z - q2;
                    a2 -= tmp, b2 -= tmp;
                    if (!RIGHT) {
                        d |= aa > 0 ? 0x08 : 0;
                        d |= bb > 0 ? 0x10 : 0;
                        d |= a2 > 0 ? 0x20 : 0;
                        d |= b2 > 0 ? 0x40 : 0;
                    } else {
                        d |= aa >= 0 ? 0x08 : 0;
                        d |= bb >= 0 ? 0x10 : 0;
                        d |= a2 >= 0 ? 0x20 : 0;
                        d |= b2 >= 0 ? 0x40 : 0;
                    }
                    dw[k >> 2] |= (uint32_t)d << (8 * (k & 3));
                    asm volatile("" : "+v"(dw[k >> 2]));
                    X[k] = max(aa, 0) - G1, Y[k] = max(bb, 0) - G1;
                    X2[k] = max(a2, 0) - G2, Y2[k] = max(b2, 0) - G2;
                    U[k] = un, V[k] = vn;
                    H[k] += vn;
                    hmax = max(hmax, H[k]);
                }
                smask = sm;
                if (!SCORE_ONLY) {
                    uint2 w8;
                    w8.x = dw[0], w8.y = dw[1];
                    *(uint2 *)(p + (uint64_t)r * ncolB + (uint64_t)(tl - b.stB)) = w8;
                }
            } else {
            const bool keep_prev = b.st16 > 0 && b.st16 - 1 >= lst16 && b.st16 - 1 <= len16;
            const int edge = edge_diff(sc, r);
            const bool inject = b.en16 >= r;
            const bool la = tl >= b.stB && tl + R - 1 <= b.en16;
            uint32_t dw[2] = {0u, 0u};
#pragma unroll
            for (int k = R - 1; k >= 0; --k) {
                const int t = tl + k;
                int xin = k > 0 ? X[k - 1] : cx, vin = k > 0 ? V[k - 1] : cv, x2in = k > 0 ? X2[k - 1] : cx2;
                if (t == b.st16) {
                    if (b.st16 == 0) xin = -G1, vin = edge, x2in = -G2;
                    else if (!keep_prev) xin = -G1, vin = -G1, x2in = -G2;
                } else if (wrap && t == b.stB) {
                    xin = wx, vin = wvv, x2in = wx2;
                }
                int uin = U[k], yin = Y[k], y2in = Y2[k];
                if (inject && t == r) uin = edge, yin = -G1, y2in = -G2;
                if (t >= b.st0 && t <= b.sc_end) {
                    const uint32_t qc = (qwin[k >> 2] >> (8 * (k & 3))) & 0xffu;
                    const uint32_t tc = (tcw[k >> 2] >> (8 * (k & 3))) & 0xffu;
                    smask = (smask & ~(1u << k)) | (tc != qc ? 1u << k : 0u);
                }
                int z = (smask >> k & 1u) ? mis : mch, d = 0;
                int aa = xin + vin, bb = yin + uin, a2 = x2in + vin, b2 = y2in + uin;
                if (SCORE_ONLY) {
                    z = max(max(max(max(z, aa), bb), a2), b2);
                } else if (!RIGHT) {
                    d = aa > z ? 1 : 0;  z = max(z, aa);
                    d = bb > z ? 2 : d;  z = max(z, bb);
                    d = a2 > z ? 3 : d;  z = max(z, a2);
                    d = b2 > z ? 4 : d;  z = max(z, b2);
                } else {
                    d = z > aa ? 0 : 1;  z = max(z, aa);
                    d = z > bb ? d : 2;  z = max(z, bb);
                    d = z > a2 ? d : 3;  z = max(z, a2);
                    d = z > b2 ? d : 4;  z = max(z, b2);
                }
                z = min(z, mch);
                const int un = z - vin, vn = z - uin;
                int tmp = z - q;
                aa -= tmp, bb -= tmp;
                tmp = z - q2;
                a2 -= tmp, b2 -= tmp;
                int xn, yn, x2n, y2n;
                if (!RIGHT) {
                    xn = max(aa, 0) - G1;   d |= aa > 0 ? 0x08 : 0;
                    yn = max(bb, 0) - G1;   d |= bb > 0 ? 0x10 : 0;
                    x2n = max(a2, 0) - G2;  d |= a2 > 0 ? 0x20 : 0;
                    y2n = max(b2, 0) - G2;  d |= b2 > 0 ? 0x40 : 0;
                } else {
                    xn = max(aa, 0) - G1;   d |= aa >= 0 ? 0x08 : 0;
                    yn = max(bb, 0) - G1;   d |= bb >= 0 ? 0x10 : 0;
                    x2n = max(a2, 0) - G2;  d |= a2 >= 0 ? 0x20 : 0;
                    y2n = max(b2, 0) - G2;  d |= b2 >= 0 ? 0x40 : 0;
                }
                dw[k >> 2] |= (uint32_t)d << (8 * (k & 3));
                // materialise the byte now (the compiler would otherwise sink
                // the direction logic of all eight positions to the store and
                // keep their operands alive)
                asm volatile("" : "+v"(dw[k >> 2]));
                if (la) {
                    // H bookkeeping on [st0, en0]; the position above still
                    // holds its H of anti-diagonal r-1
                    if (t == b.en0) {
                        const int hab = k > 0 ? H[k - 1] : ch;
                        const int h = r == 0 ? vn - sc.qe_orig : b.en0 > 0 ? hab + un : H[k] + vn;
                        H[k] = h;
                        if (b.en0 == tlen - 1 && h >= mte_h) mte_h = h, mte_q = r - b.en0;
                        if (r == nr - 1 && b.en0 == tlen - 1) score = h;
                    } else if (t >= b.st0 && t < b.en0) {
                        H[k] += vn;
                    }
                    if (t >= b.st0 && t <= b.en0) {
                        if (H[k] > 0) pos = 1;
                        if (t == b.st0 && r - b.st0 == qlen - 1 && H[k] >= mqe_h) mqe_h = H[k], mqe_t = t;
                    }
                    U[k] = un, V[k] = vn, X[k] = xn, Y[k] = yn, X2[k] = x2n, Y2[k] = y2n;
                }
            }
            if (!SCORE_ONLY && la) {
                uint2 w8;
                w8.x = dw[0], w8.y = dw[1];
                *(uint2 *)(p + (uint64_t)r * ncolB + (uint64_t)(tl - b.stB)) = w8;
            }
            }
            lst16 = b.st16, len16 = b.en16;
            // this stripe's bottom position for the next stripe
            if (lane == kWave - 1) {
                out_buf[2 * r] = pack3(X[R - 1], V[R - 1], X2[R - 1]);
                out_buf[2 * r + 1] = H[R - 1];
            }
            if ((r - r_lo) % kBlock == kBlock - 1) publish(prog + wv, kk * (nr + 2) + r + 1);
        }
        publish(prog + wv, kk * (nr + 2) + nr + 1);  // past every need of the next stripe
        if (hmax > 0) pos = 1;
    }

    // ---- results: mte and the score come from one position; mqe keeps the
    // largest H, ties to the later anti-diagonal (the larger t)
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int oh = __shfl_xor(mqe_h, o, kWave), ot = __shfl_xor(mqe_t, o, kWave);
        if (oh > mqe_h || (oh == mqe_h && ot > mqe_t)) mqe_h = oh, mqe_t = ot;
        const int th = __shfl_xor(mte_h, o, kWave), tq = __shfl_xor(mte_q, o, kWave);
        if (tq >= 0 && (mte_q < 0 || th > mte_h)) mte_h = th, mte_q = tq;
        score = max(score, __shfl_xor(score, o, kWave));
        pos |= __shfl_xor(pos, o, kWave);
    }
    if (lane == 0) {
        red[wv][0] = mqe_h, red[wv][1] = mqe_t, red[wv][2] = mte_h, red[wv][3] = mte_q;
        red[wv][4] = score, red[wv][5] = pos;
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < W; ++k) {
            if (red[k][0] > mqe_h || (red[k][0] == mqe_h && red[k][1] > mqe_t)) mqe_h = red[k][0], mqe_t = red[k][1];
            if (red[k][3] >= 0 && (mte_q < 0 || red[k][2] > mte_h)) mte_h = red[k][2], mte_q = red[k][3];
            score = max(score, red[k][4]);
            pos |= red[k][5];
        }
        KswOut o;
        o.score = score, o.mqe = mqe_h, o.mqe_t = mqe_t, o.mte = mte_h, o.mte_q = mte_q;
        o.max = 0, o.max_t = -1, o.max_q = -1;
        o.zdropped = r_end < nr ? 1 : 0, o.stopped = 0, o.reach_end = 0, o.n_cigar = 0;
        a.out[jb.out] = o;
        a.redo[job] = pos | a.force_redo;
    }
}

hipError_t launch_band(const BandArgs &a, int team, bool right, bool score_only, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (team < 1 || team > kBandTeam) return hipErrorInvalidValue;
    if (score_only)
        hipLaunchKernelGGL((ksw_band_kernel<false, true>), dim3(a.njobs), dim3(kWave * team), 0, st, a);
    else if (right)
        hipLaunchKernelGGL((ksw_band_kernel<true, false>), dim3(a.njobs), dim3(kWave * team), 0, st, a);
    else
        hipLaunchKernelGGL((ksw_band_kernel<false, false>), dim3(a.njobs), dim3(kWave * team), 0, st, a);
    return hipGetLastError();
}

}  // namespace kswhip
