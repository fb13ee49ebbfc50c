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
//  * Stripes.  The positions are cut into stripes of 256; lane k of a wave
//    holds positions t0 + 4k .. t0 + 4k + 3 of its stripe in registers (the
//    in-place state u, v, x, y, x2, y2, s and H of each) and steps through
//    the anti-diagonals that reach the stripe.  Within a lane the positions
//    are updated from the highest down, so each still reads its upper
//    neighbour's value from the previous anti-diagonal; the lane's top
//    position takes the lane above's bottom values through a DPP shift, and
//    lane 0 takes the previous stripe's bottom position from a per-pair
//    buffer indexed by r.  A position outside [stB, en16] is not updated:
//    its stale value is simply the register it keeps.
//  * Because stB is a multiple of 64 and en16 + 1 of 16, a lane's four
//    positions are all inside [stB, en16] or all outside: the commit and the
//    direction-byte store are per lane (four bytes, aligned).
//  * Teams.  W waves per pair (W = 16 for w = 4096: a stripe's
//    anti-diagonals start 2 SR after its predecessor's, so about 2w / 2 SR
//    stripes are in flight); wave w takes stripes w, w + W, ...  The waves
//    sit in workgroups of four (one wave per SIMD) that the dispatcher
//    spreads over the CUs, so a pair's waves do not share SIMDs.  Stripe s
//    trails stripe s-1 by at least 64 anti-diagonals through progress
//    counts in device memory; counts and hand-off entries use device-scope
//    (sc1) accesses, coherent across the XCDs' L2s, and a count is stored
//    only after the entries it covers are complete.  Three buffers by
//    stripe mod 3: a producer three stripes ahead reuses a buffer, and the
//    leads between the stripes in between keep it from overwriting an entry
//    before it is read.
//  * A wait that outlasts kWaitTicks (a workgroup the dispatcher could not
//    place) gives up: the pair is flagged for the general kernel, which
//    recomputes it, and every wave still runs to its end.
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

constexpr int R = kBandRowsPerLane;  // positions per lane (4: one step of a stripe is short)
constexpr int SR = kWave * R;        // positions per stripe
static_assert(R == 4, "a lane's direction bytes and query window are one dword");
// Hand-off granularity, measured on the banded long pairs (64 x 100-300
// kbp, w = 4096): blocks, lead and publications of 64/64/32 anti-diagonals
// 0.625 s per call, 32/32/16 0.596 s, 16/16/8 0.594 s, 16/16/16 0.582 s,
// 8/8/8 0.620 s.
constexpr int kBlock = 16;       // anti-diagonals per block of hand-off loads
constexpr int kPublish = 16;     // anti-diagonals between progress publications
constexpr int kLead = 16;        // lead a stripe waits for on its predecessor
constexpr int kBufs = 3;         // hand-off buffers, by stripe mod 3
static_assert(kBlock % kPublish == 0, "publications fall on block offsets");
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
constexpr uint64_t kWaitTicks = 1000000000ull;  // 10 s of the 100 MHz real-time clock

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
    // (0 <= st <= en: the 16- and 64-roundings are masks)
    Band b;
    b.st0 = st, b.en0 = en;
    b.st16 = st & ~15, b.en16 = en | 15, b.stB = st & ~63;
    b.sc_end = generic ? en : min(st + ((en - st + 16) & ~15) - 1, tlenB - 1);
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

// lane k takes lane k-1's v, lane 0 keeps `old`
__device__ __forceinline__ int dpp_shr1(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, kDppWaveShr1, 0xF, 0xF, false);
}
// lane k takes lane k+1's v (lane 63 keeps its own)
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(v, v, kDppWaveShl1, 0xF, 0xF, false); }

// Progress count: stored after the wave's earlier hand-off stores are
// complete (s_waitcnt vmcnt(0)); both are device-scope stores that bypass
// the XCD-private caches.
__device__ __forceinline__ void publish(int32_t *slot, int value)
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
    asm volatile("" ::: "memory");
    if (__lane_id() == 0) __hip_atomic_store(slot, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int load_count(const int32_t *slot)
{
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Spins until the producer's published count reaches `need` (uniform exit
// test on lane 0's value).  A producer always ends a stripe by publishing a
// count past every need of its consumer, so the wait ends once the
// producer runs; past kWaitTicks it gives up, sets `late` and returns a
// count that satisfies every later wait of the wave.
__device__ __forceinline__ int wait_count(const int32_t *slot, int need, int known, int &late)
{
    if (known >= need) return known;
    int seen = load_count(slot);
    if (seen < need) {
        const uint64_t t_start = wall_clock64();
        while (seen < need) {
            __builtin_amdgcn_s_sleep(2);
            seen = load_count(slot);
            if (seen < need && wall_clock64() - t_start > kWaitTicks) {
                late = 1;
                seen = INT32_MAX;
            }
        }
    }
    asm volatile("" ::: "memory");  // the entry loads stay behind the wait
    return seen;
}

__device__ __forceinline__ int pack3(int x, int v, int x2)
{
    return (x & 0xff) | (v & 0xff) << 8 | (x2 & 0xff) << 16;
}
__device__ __forceinline__ int byte_s(int w, int k) { return (int)(int8_t)(uint8_t)((unsigned)w >> (8 * k)); }

// A lane's R direction bytes (one aligned store).
__device__ __forceinline__ void store_dirs(uint8_t *dst, const uint32_t (&dw)[2])
{
    if (R == 8) {
        uint2 w8;
        w8.x = dw[0], w8.y = dw[1];
        *(uint2 *)dst = w8;
    } else {
        *(uint32_t *)dst = dw[0];
    }
}

}  // namespace

// Scaled, tagged state (the stripe kernel's representation): every
// difference w is kept as -8w (u, v) or -8(w + G) - BIAS (x, y with G1; x2,
// y2 with G2), BIAS = 0 for left and 8 for right alignment, so that one
// chain of mins over s, a, b, a2, b2 (each -8*value + a 3-bit tag) gives
// both the reference's maximum and the index its strict '>' (left) or '>='
// (right) chain picks; T = (x + v) + 8e - 8z is -8a' - BIAS, whose sign is
// the continuation flag and min(T, -BIAS) the new x.
struct BandConsts {
    int CA, CB, CA2, CB2, E1, E2, LIM, SM, SMM, CAP;
};

template <bool RIGHT>
__device__ __forceinline__ BandConsts band_consts(const KswScore &sc)
{
    const int bias = RIGHT ? 8 : 0;
    const int tS = RIGHT ? 7 : 0, tA = RIGHT ? 6 : 1, tB = RIGHT ? 5 : 2, tA2 = RIGHT ? 4 : 3, tB2 = RIGHT ? 3 : 4;
    BandConsts k;
    k.CA = 8 * sc.G1 + tA + bias, k.CB = 8 * sc.G1 + tB + bias;
    k.CA2 = 8 * sc.G2 + tA2 + bias, k.CB2 = 8 * sc.G2 + tB2 + bias;
    k.E1 = 8 * sc.e, k.E2 = 8 * sc.e2, k.LIM = -bias;
    k.SM = -8 * sc.mch + tS, k.SMM = -8 * sc.mis + tS;
    k.CAP = -8 * sc.mch;  // z = min(z, mat[0]) in the -8 domain
    return k;
}

// One cell in the scaled form.  In: x, v, x2 of the position above, u, y,
// y2 of this position (scaled), mm = mismatch.  Out: the new state; returns
// the direction byte (0 when SCORE_ONLY).
template <bool RIGHT, bool SCORE_ONLY>
__device__ __forceinline__ int band_cell(const BandConsts &k, int X, int V, int X2, int U, int Y, int Y2, bool mm,
                                         int &Un, int &Vn, int &Xn, int &Yn, int &X2n, int &Y2n)
{
    const int S = mm ? k.SMM : k.SM;
    const int A = X + V + k.CA, B = Y + U + k.CB, A2 = X2 + V + k.CA2, B2 = Y2 + U + k.CB2;
    const int Z = min(min(min(min(S, A), B), A2), B2);
    const int Z8 = max(Z & ~7, k.CAP);
    const int M1 = k.E1 - Z8, M2 = k.E2 - Z8;
    const int Ta = X + V + M1, Tb = Y + U + M1, Ta2 = X2 + V + M2, Tb2 = Y2 + U + M2;
    Un = Z8 - V, Vn = Z8 - U;
    Xn = min(Ta, k.LIM), Yn = min(Tb, k.LIM), X2n = min(Ta2, k.LIM), Y2n = min(Tb2, k.LIM);
    if (SCORE_ONLY) return 0;
    int d = RIGHT ? (~Z & 7) : (Z & 7);
    d |= (int)(((uint32_t)Ta >> 31) << 3) | (int)(((uint32_t)Tb >> 31) << 4);
    d |= (int)(((uint32_t)Ta2 >> 31) << 5) | (int)(((uint32_t)Tb2 >> 31) << 6);
    return d;
}

// Two positions per register: position 2m of a lane in the low half of
// register m, 2m+1 in the high half.  The scaled values and every sum the
// cell forms fit in 16 bits in this domain (the stripe kernel's argument:
// int8 differences times 8 plus constants below 2^11), so the packed cell
// gives the 32-bit cell's values.
typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned short v2u __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int bits(v2s v) { return __builtin_bit_cast(int, v); }
__device__ __forceinline__ int bits(v2u v) { return __builtin_bit_cast(int, v); }
__device__ __forceinline__ v2s vec(int x) { return __builtin_bit_cast(v2s, x); }
__device__ __forceinline__ v2u uvec(int x) { return __builtin_bit_cast(v2u, x); }
__device__ __forceinline__ v2s splat(int x) { return (v2s){(short)x, (short)x}; }
__device__ __forceinline__ v2s vmin(v2s a, v2s b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ v2s vmax(v2s a, v2s b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ int lo16(int x) { return (int)(int16_t)(x & 0xffff); }
__device__ __forceinline__ int hi16(int x) { return x >> 16; }
// (x >> 3) of each half, sign-extended to 32 bits
__device__ __forceinline__ int lo16_div8(int x) { return __builtin_amdgcn_sbfe(x, 3, 13); }
__device__ __forceinline__ int hi16_div8(int x) { return x >> 19; }
// the high half of `hi` and the low half of `lo`: positions (2m-1, 2m)
__device__ __forceinline__ int above_pair(int lo_reg, int hi_reg)
{
    return (int)__builtin_amdgcn_alignbit((unsigned)hi_reg, (unsigned)lo_reg, 16);
}

struct BandConsts2 {
    v2s CA, CB, CA2, CB2, E1, E2, LIM, SM, SD, CAP;
};

__device__ __forceinline__ BandConsts2 pack_consts(const BandConsts &k)
{
    BandConsts2 p;
    p.CA = splat(k.CA), p.CB = splat(k.CB), p.CA2 = splat(k.CA2), p.CB2 = splat(k.CB2);
    p.E1 = splat(k.E1), p.E2 = splat(k.E2), p.LIM = splat(k.LIM);
    p.SM = splat(k.SM), p.SD = splat(k.SMM - k.SM), p.CAP = splat(k.CAP);
    return p;
}

// band_cell on two positions.  X/V/X2: the positions above; U/Y/Y2 this
// pair's, replaced with the new values, as are X/V/X2 (the new values of
// this pair); mm 0/1 per half.  Returns the two direction bytes in the low
// byte of each half.
template <bool RIGHT, bool SCORE_ONLY>
__device__ __forceinline__ int band_cell2(const BandConsts2 &k, v2u mm, v2s &X, v2s &V, v2s &X2, v2s &U, v2s &Y,
                                          v2s &Y2)
{
    v2u m = mm;
    asm volatile("" : "+v"(m));  // keep the multiply-add form
    const v2s S = k.SM + __builtin_bit_cast(v2s, m) * k.SD;
    const v2s XV = X + V, X2V = X2 + V, YU = Y + U, Y2U = Y2 + U;
    const v2s A = XV + k.CA, B = YU + k.CB, A2 = X2V + k.CA2, B2 = Y2U + k.CB2;
    const v2s Z = vmin(vmin(vmin(vmin(S, A), B), A2), B2);
    const v2s Z8 = vmax(vec(bits(Z) & (int)0xFFF8FFF8u), k.CAP);
    const v2s M1 = k.E1 - Z8, M2 = k.E2 - Z8;
    const v2s Ta = XV + M1, Tb = YU + M1, Ta2 = X2V + M2, Tb2 = Y2U + M2;
    const v2s Un = Z8 - V, Vn = Z8 - U;
    X = vmin(Ta, k.LIM), Y = vmin(Tb, k.LIM), X2 = vmin(Ta2, k.LIM), Y2 = vmin(Tb2, k.LIM);
    U = Un, V = Vn;
    if (SCORE_ONLY) return 0;
    int d = RIGHT ? (~bits(Z) & 0x00070007) : (bits(Z) & 0x00070007);
    d |= bits(__builtin_bit_cast(v2u, Ta) >> (v2u){12, 12}) & 0x00080008;
    d |= bits(__builtin_bit_cast(v2u, Tb) >> (v2u){11, 11}) & 0x00100010;
    d |= bits(__builtin_bit_cast(v2u, Ta2) >> (v2u){10, 10}) & 0x00200020;
    d |= bits(__builtin_bit_cast(v2u, Tb2) >> (v2u){9, 9}) & 0x00400040;
    return d;
}

// Per-stripe state of a lane: four positions in two packed registers per
// quantity, the score bit of each position as 0/1 halves, H in 32 bits.
struct Lane4 {
    v2s U[2], V[2], X[2], Y[2], X2[2], Y2[2];
    v2u MM[2];
    int H[4];
};

// The inputs from the positions above for the lane's two pairs: (p-1, p0)
// and (p1, p2), with position -1 from `top` (high half).
struct Above {
    v2s X[2], V[2], X2[2];
};

__device__ __forceinline__ Above above_of(const Lane4 &st, int tx, int tv, int tx2)
{
    Above a;
    a.X[0] = vec(above_pair(tx, bits(st.X[0]))), a.X[1] = vec(above_pair(bits(st.X[0]), bits(st.X[1])));
    a.V[0] = vec(above_pair(tv, bits(st.V[0]))), a.V[1] = vec(above_pair(bits(st.V[0]), bits(st.V[1])));
    a.X2[0] = vec(above_pair(tx2, bits(st.X2[0]))), a.X2[1] = vec(above_pair(bits(st.X2[0]), bits(st.X2[1])));
    return a;
}

// The four direction bytes of a lane, positions in order.
__device__ __forceinline__ uint32_t pack_dirs(int d0, int d1)
{
    return __builtin_amdgcn_perm((unsigned)d1, (unsigned)d0, 0x06040200u);
}

// Query codes of the lane's pairs from its window: (q0, q1), (q2, q3).
__device__ __forceinline__ void query_pairs(uint32_t qwin, v2u &q0, v2u &q1)
{
    q0 = uvec((int)__builtin_amdgcn_perm(0u, qwin, 0x0c010c00u));
    q1 = uvec((int)__builtin_amdgcn_perm(0u, qwin, 0x0c030c02u));
}

__device__ __forceinline__ v2u mismatch(v2u tb, v2u qb) { return __builtin_elementwise_min(tb ^ qb, (v2u){1, 1}); }

// A band-edge step in a block clear of the rare boundary rules: the
// interior form plus what the band's edges change.  LEAD: the band starts
// above the stripe (st0 < t0), so no start rule applies; TRAIL: it ends
// below the stripe (en0 > t0 + SR - 1), so no end rule applies.
//  * Inputs of the start (minimap2/ksw2_extd2_sse.c:158-172): position stB
//    reads position stB+15 of anti-diagonal r-1 when st16 != stB (the
//    AVX-512 lane wrap), position st16 the initial x, v, x2 unless position
//    st16-1 was computed on r-1.  Both become the inputs of that position
//    (the above-values of a position are read by the one below only).
//  * Score bits: positions [st0, sc_end] take this anti-diagonal's mismatch.
//  * H: set at en0 from the position above, += v on [st0, en0); kept below
//    st0 (a band one position wide reads it), free to drift above en0 (a
//    position enters at en0 from the position above).
//  * Commit: a lane above en16 keeps its state (it enters the band later
//    with it); below stB a lane's positions are never read again.
template <bool RIGHT, bool SCORE_ONLY, bool LEAD, bool TRAIL>
__device__ __forceinline__ void edge_step(const BandConsts2 &kc, const Band &b, int t0, int tl, int lane, Lane4 &st,
                                          const v2u (&TB)[2], uint32_t qwin, int tx, int tv, int tx2, int ch,
                                          bool keep_prev, int U0, int L0, int &hmax, uint8_t *prow)
{
    Above a = above_of(st, tx, tv, tx2);
    if (!LEAD) {
        if (b.st16 != b.stB && b.stB >= t0 && b.stB < t0 + SR) {
            // position stB (low half of pair 0 of its lane) takes position
            // stB+15 (high half of pair 1 of lane (stB+15-t0)/R)
            const int src = (b.stB + 15 - t0) / R;
            const int wx = __builtin_amdgcn_readlane(bits(st.X[1]), src);
            const int wv = __builtin_amdgcn_readlane(bits(st.V[1]), src);
            const int wx2 = __builtin_amdgcn_readlane(bits(st.X2[1]), src);
            if (lane == (b.stB - t0) / R) {
                a.X[0] = vec((bits(a.X[0]) & (int)0xffff0000u) | ((unsigned)wx >> 16));
                a.V[0] = vec((bits(a.V[0]) & (int)0xffff0000u) | ((unsigned)wv >> 16));
                a.X2[0] = vec((bits(a.X2[0]) & (int)0xffff0000u) | ((unsigned)wx2 >> 16));
            }
        }
        if (!keep_prev && b.st16 >= t0 && b.st16 < t0 + SR) {
            const int k16 = (b.st16 - t0) % R, m = k16 >> 1;
            const int hm = (k16 & 1) ? (int)0xffff0000u : 0xffff;
            if (lane == (b.st16 - t0) / R) {
                const int l0 = bits(splat(L0)) & hm, u0 = bits(splat(U0)) & hm;
                if (m == 0) {
                    a.X[0] = vec((bits(a.X[0]) & ~hm) | l0), a.V[0] = vec((bits(a.V[0]) & ~hm) | u0);
                    a.X2[0] = vec((bits(a.X2[0]) & ~hm) | l0);
                } else {
                    a.X[1] = vec((bits(a.X[1]) & ~hm) | l0), a.V[1] = vec((bits(a.V[1]) & ~hm) | u0);
                    a.X2[1] = vec((bits(a.X2[1]) & ~hm) | l0);
                }
            }
        }
    }
    const bool la = (LEAD || tl >= b.stB) && (TRAIL || tl + R - 1 <= b.en16);
    const int lam = la ? -1 : 0;
    // score bits: positions [st0, sc_end] take this anti-diagonal's mismatch
    {
        v2u q0, q1;
        query_pairs(qwin, q0, q1);
        const int lo = LEAD ? 0 : min(max(b.st0 - tl, 0), R);
        const int hi = TRAIL ? R : min(max(b.sc_end - tl + 1, 0), R);
        const uint32_t upd = ((1u << hi) - 1u) & ~((1u << lo) - 1u);
        const int um0 = (int)((upd & 1u) * 0xffffu | (upd & 2u) * 0x7fff8000u);
        const int um1 = (int)((upd >> 2 & 1u) * 0xffffu | (upd >> 2 & 2u) * 0x7fff8000u);
        st.MM[0] = uvec((bits(mismatch(TB[0], q0)) & um0) | (bits(st.MM[0]) & ~um0));
        st.MM[1] = uvec((bits(mismatch(TB[1], q1)) & um1) | (bits(st.MM[1]) & ~um1));
    }
    Lane4 nw;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        nw.X[m] = a.X[m], nw.V[m] = a.V[m], nw.X2[m] = a.X2[m];
        nw.U[m] = st.U[m], nw.Y[m] = st.Y[m], nw.Y2[m] = st.Y2[m];
    }
    const int d0 = band_cell2<RIGHT, SCORE_ONLY>(kc, st.MM[0], nw.X[0], nw.V[0], nw.X2[0], nw.U[0], nw.Y[0], nw.Y2[0]);
    const int d1 = band_cell2<RIGHT, SCORE_ONLY>(kc, st.MM[1], nw.X[1], nw.V[1], nw.X2[1], nw.U[1], nw.Y[1], nw.Y2[1]);
    // H: set at en0 from the position above, += v on [st0, en0); kept below
    // st0, free to drift above en0
    const int un3[4] = {lo16_div8(bits(nw.U[0])), hi16_div8(bits(nw.U[0])), lo16_div8(bits(nw.U[1])),
                        hi16_div8(bits(nw.U[1]))};
    const int vn3[4] = {lo16_div8(bits(nw.V[0])), hi16_div8(bits(nw.V[0])), lo16_div8(bits(nw.V[1])),
                        hi16_div8(bits(nw.V[1]))};
#pragma unroll
    for (int k = R - 1; k >= 0; --k) {
        const int t = tl + k;
        const int hab = k > 0 ? st.H[k - 1] : ch;
        const int vdec = (LEAD || t >= b.st0) ? vn3[k] : 0;
        const int hn = !TRAIL && t == b.en0 ? hab - un3[k] : st.H[k] - vdec;
        st.H[k] = hn;
        const bool inb = (LEAD || t >= b.st0) && (TRAIL || t <= b.en0);
        hmax = max(hmax, inb ? hn : KSW_NEG_INF_I);
    }
    // commit: a lane above en16 keeps its state (it enters the band later
    // with it); below stB a lane's positions are never read again
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        if (TRAIL) {
            st.U[m] = nw.U[m], st.V[m] = nw.V[m], st.X[m] = nw.X[m];
            st.Y[m] = nw.Y[m], st.X2[m] = nw.X2[m], st.Y2[m] = nw.Y2[m];
        } else {
            st.U[m] = vec((bits(nw.U[m]) & lam) | (bits(st.U[m]) & ~lam));
            st.V[m] = vec((bits(nw.V[m]) & lam) | (bits(st.V[m]) & ~lam));
            st.X[m] = vec((bits(nw.X[m]) & lam) | (bits(st.X[m]) & ~lam));
            st.Y[m] = vec((bits(nw.Y[m]) & lam) | (bits(st.Y[m]) & ~lam));
            st.X2[m] = vec((bits(nw.X2[m]) & lam) | (bits(st.X2[m]) & ~lam));
            st.Y2[m] = vec((bits(nw.Y2[m]) & lam) | (bits(st.Y2[m]) & ~lam));
        }
    }
    if (!SCORE_ONLY && la) *(uint32_t *)(prow + (tl - b.stB)) = pack_dirs(d0, d1);
}

// Up to kBandTeam waves per pair in workgroups of up to kBandGroup.  A
// pair needs about (2w + 2 SR + 300) / (2 SR + 190) stripes in flight (a
// stripe's anti-diagonals start 2 SR after its predecessor's, plus the
// lead), 13 for w = 4096 at SR = 256.
template <bool RIGHT, bool SCORE_ONLY>
__global__ void __launch_bounds__(kWave * kBandGroup) ksw_band_kernel(BandArgs a)
{
    const int nthr = blockDim.x, tid = threadIdx.x, lane = tid % kWave;
    const int wgs = a.wgs, g = blockIdx.x % wgs;
    const int W = __builtin_amdgcn_readfirstlane(wgs * (nthr / kWave));
    const int wv = __builtin_amdgcn_readfirstlane(g * (nthr / kWave) + tid / kWave);
    const int job = __builtin_amdgcn_readfirstlane(a.order[blockIdx.x / wgs]);
    const GenJob jb = a.jobs[job];
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    const int bw = __builtin_amdgcn_readfirstlane(jb.w), flag = __builtin_amdgcn_readfirstlane(jb.flag);
    const int tlenB = __builtin_amdgcn_readfirstlane(jb.tlenB), r_end = __builtin_amdgcn_readfirstlane(jb.r_end);
    const int ncolB = jb.ncolB;
    const bool generic = flag & 0x04;
    const int nr = qlen + tlen - 1;
    const KswScore &sc = a.sc.k;
    const BandConsts kc = band_consts<RIGHT>(sc);
    const BandConsts2 kc2 = pack_consts(kc);
    const int G1 = sc.G1, G2 = sc.G2, bias = RIGHT ? 8 : 0;
    // scaled constants: u, v = -G1 -> 8 G1; x, y, x2, y2 at their initial
    // values (-G1, -G2) -> -BIAS
    const int U0 = 8 * G1, L0 = -bias;
    const uint8_t *qs = a.seq + jb.q_off, *ts = a.seq + jb.t_off;
    uint8_t *p = a.dir + jb.dir_off;
    int32_t *off = a.offs + jb.off_off, *off_end = off + nr;
    int32_t *ctl = a.bnd + jb.bnd_off;      // control block (zeroed before the launch)
    int32_t *buf = ctl + kBandCtl;          // kBufs buffers of 2 * nr dwords
    int32_t *prog = ctl;                    // count of wave w at prog[w * kBandCtlStride]

    if (!SCORE_ONLY)
        for (int r = g * nthr + tid; r < r_end; r += wgs * nthr) {
            const Band b = band_of(r, qlen, tlen, bw, tlenB, generic);
            off[r] = b.stB, off_end[r] = b.en16;
        }

    // per-lane results over this wave's stripes
    int mqe_h = KSW_NEG_INF_I, mqe_t = -1, mte_h = KSW_NEG_INF_I, mte_q = -1, score = KSW_NEG_INF_I, pos = 0;
    const int nstripes = (tlenB + SR - 1) / SR;
    int known = 0, late = 0, stop = 0;
    for (int s = wv, kk = 0; s < nstripes && !stop; s += W, ++kk) {
        const int t0 = s * SR, tl = t0 + lane * R;
        int r_lo, r_hi, p_lo = 0, p_hi = -1;
        stripe_range(t0, bw, r_end, r_lo, r_hi);
        if (s > 0) stripe_range(t0 - SR, bw, r_end, p_lo, p_hi);
        const int pw = (s + W - 1) % W, pbase = (s - 1) / W * (nr + 2);  // producer wave, its count base
        int32_t *out_buf = buf + (uint64_t)(s % kBufs) * 2 * nr;
        const int32_t *in_buf = buf + (uint64_t)((s + kBufs - 1) % kBufs) * 2 * nr;
        // per position: the in-place state (scaled, two positions per
        // register) and H; the score s as a 0/1 half (1: mismatch; s is mch
        // or mis in this domain, mch = 0 at first); the target codes as
        // 16-bit halves (zero past tlen, as the reference's zero-padded
        // target copy)
        Lane4 st;
        v2u TB[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            st.U[m] = st.V[m] = splat(U0);
            st.X[m] = st.Y[m] = st.X2[m] = st.Y2[m] = splat(L0);
            st.MM[m] = (v2u){0, 0};
            const int ta = tl + 2 * m, tb = ta + 1;
            TB[m] = (v2u){(unsigned short)(ta < tlen ? ts[ta] : 0), (unsigned short)(tb < tlen ? ts[tb] : 0)};
        }
#pragma unroll
        for (int k = 0; k < R; ++k) st.H[k] = KSW_NEG_INF_I;
        // query window: byte k of qwin[0] is the query code of position
        // tl + k's column r - (tl + k) (0 outside the query, as the
        // reference's zero-tailed reversed query).  Each step lane k's new
        // byte is the one lane k-1 drops (its byte 3); lane 0 takes Q(r - t0)
        // from the block registers.
        uint32_t qwin[1] = {0u};
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int j = r_lo - 1 - tl - k;
            const uint32_t c = j >= 0 && j < qlen ? (uint32_t)qs[j] : 0u;
            qwin[k >> 2] |= c << (8 * (k & 3));
        }
        // per block of kBlock steps starting at rb, lane k < kBlock holds in
        // registers what lane 0 needs at step rb + k: the query code
        // Q(rb + k - t0) and the predecessor's entry rb + k - 1 (x, v, x2
        // scaled, and H; initial values before its first anti-diagonal, its
        // last entry after its last).  Loaded once per block, after the wait
        // that guarantees the entries are written, so no step waits on
        // memory; every step moves them down one lane, so lane 0 holds the
        // current step's.
        int bq = 0, bx = 0, bv = 0, bx2 = 0, bh = KSW_NEG_INF_I;
        int lst16 = -1, len16 = -1;  // st16, en16 of anti-diagonal r-1
        if (r_lo > 0) {
            const Band bp = band_of(r_lo - 1, qlen, tlen, bw, tlenB, generic);
            lst16 = bp.st16, len16 = bp.en16;
        }
        int hmax = KSW_NEG_INF_I;  // largest in-band H (checked against 0 at the end)
        uint8_t *prow = p + (uint64_t)r_lo * ncolB;  // direction row of anti-diagonal r
        for (int rb = r_lo; rb <= r_hi; rb += kBlock) {
            if (a.stop && ((rb - r_lo) & 255) < kBlock) {
                // the host's stop flag, every 256 anti-diagonals: end the
                // wave (nothing of this pair is delivered)
                stop = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
                if (stop) break;
            }
            if (s > 0) known = wait_count(prog + pw * kBandCtlStride, pbase + min(rb + kLead, p_hi) + 1, known, late);
            {
                const int j = rb + lane - t0;
                bq = j >= 0 && j < qlen ? (int)qs[j] : 0;
                const int e = min(rb + lane - 1, p_hi);
                int pk = pack3(-G1, -G1, -G2), h = KSW_NEG_INF_I;
                if (lane < kBlock && s > 0 && p_hi >= p_lo && e >= p_lo) {
                    const uint64_t en = __hip_atomic_load((const uint64_t *)(in_buf + 2 * e), __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
                    pk = (int)(uint32_t)en, h = (int)(uint32_t)(en >> 32);
                }
                // take the loads here: the steps of the block then read the
                // registers without waiting on memory (the step's own stores
                // would otherwise be waited for on every step)
                asm volatile("" : "+v"(bq), "+v"(pk), "+v"(h));
                // scaled, in the high half (the position above pair 0)
                bx = (-8 * (byte_s(pk, 0) + G1) - bias) << 16, bv = (-8 * byte_s(pk, 1)) << 16;
                bx2 = (-8 * (byte_s(pk, 2) + G2) - bias) << 16, bh = h;
            }
            const int rbe = min(rb + kBlock - 1, r_hi);
            // whether the block is clear of the rare boundary rules (the
            // injection at t == r, a start below 16, the last position or
            // anti-diagonal, the mqe column): r - en0, st0, en0 and r - st0
            // never decrease with r, so the block's ends decide
            bool plain;
            {
                const Band b0 = band_of(rb, qlen, tlen, bw, tlenB, generic);
                const Band b1 = band_of(rbe, qlen, tlen, bw, tlenB, generic);
                plain = rb - b0.en0 > 15 && b0.st0 >= 16 && b0.en0 > 0 && b1.en0 < tlen - 1 && rbe < nr - 1 &&
                        rbe - b1.st0 < qlen - 1;
            }
            for (int r = rb; r <= rbe; ++r, prow += ncolB) {
                const Band b = band_of(r, qlen, tlen, bw, tlenB, generic);
                // query window: each lane's new byte is the one the lane
                // above drops, lane 0's the block's current query code
                qwin[0] = qwin[0] << 8 | ((uint32_t)dpp_shr1(bq << 24, (int)qwin[0]) >> 24);
                // position t0-1 as anti-diagonal r-1 left it (high halves): the
                // lane above's bottom position, lane 0's from the block
                // registers
                const int tx = dpp_shr1(bx, bits(st.X[1])), tv = dpp_shr1(bv, bits(st.V[1]));
                const int tx2 = dpp_shr1(bx2, bits(st.X2[1])), ch = dpp_shr1(bh, st.H[R - 1]);
                bq = dpp_shl1(bq), bx = dpp_shl1(bx), bv = dpp_shl1(bv), bx2 = dpp_shl1(bx2), bh = dpp_shl1(bh);
                // interior step: every position of the stripe strictly inside
                // (st0, en0) -- all computed and scored, no boundary rule, no H
                // entry, no mqe/mte/score position -- the lean form
                if (b.st0 < t0 && t0 + SR - 1 < b.en0) {
                    Above ab = above_of(st, tx, tv, tx2);
                    v2u q0, q1;
                    query_pairs(qwin[0], q0, q1);
                    st.MM[0] = mismatch(TB[0], q0), st.MM[1] = mismatch(TB[1], q1);
                    st.X[0] = ab.X[0], st.V[0] = ab.V[0], st.X2[0] = ab.X2[0];
                    st.X[1] = ab.X[1], st.V[1] = ab.V[1], st.X2[1] = ab.X2[1];
                    const int d0 = band_cell2<RIGHT, SCORE_ONLY>(kc2, st.MM[0], st.X[0], st.V[0], st.X2[0], st.U[0],
                                                                 st.Y[0], st.Y2[0]);
                    const int d1 = band_cell2<RIGHT, SCORE_ONLY>(kc2, st.MM[1], st.X[1], st.V[1], st.X2[1], st.U[1],
                                                                 st.Y[1], st.Y2[1]);
                    // v = -(V / 8)
                    st.H[0] -= lo16_div8(bits(st.V[0])), st.H[1] -= hi16_div8(bits(st.V[0]));
                    st.H[2] -= lo16_div8(bits(st.V[1])), st.H[3] -= hi16_div8(bits(st.V[1]));
                    hmax = max(hmax, max(max(st.H[0], st.H[1]), max(st.H[2], st.H[3])));
                    if (!SCORE_ONLY) *(uint32_t *)(prow + (tl - b.stB)) = pack_dirs(d0, d1);
                } else if (plain) {
                    const bool keep_prev = b.st16 - 1 >= lst16 && b.st16 - 1 <= len16;
                    if (b.st0 < t0)
                        edge_step<RIGHT, SCORE_ONLY, true, false>(kc2, b, t0, tl, lane, st, TB, qwin[0], tx, tv, tx2,
                                                                  ch, keep_prev, U0, L0, hmax, prow);
                    else if (t0 + SR - 1 < b.en0)
                        edge_step<RIGHT, SCORE_ONLY, false, true>(kc2, b, t0, tl, lane, st, TB, qwin[0], tx, tv, tx2,
                                                                  ch, keep_prev, U0, L0, hmax, prow);
                    else
                        edge_step<RIGHT, SCORE_ONLY, false, false>(kc2, b, t0, tl, lane, st, TB, qwin[0], tx, tv, tx2,
                                                                   ch, keep_prev, U0, L0, hmax, prow);
                } else {
                    // every rule, on the 32-bit form of the lane's state
                    int U[R], Y[R], Y2[R], X[R], V[R], X2[R];
                    int *H = st.H;
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        U[2 * m] = lo16(bits(st.U[m])), U[2 * m + 1] = hi16(bits(st.U[m]));
                        V[2 * m] = lo16(bits(st.V[m])), V[2 * m + 1] = hi16(bits(st.V[m]));
                        X[2 * m] = lo16(bits(st.X[m])), X[2 * m + 1] = hi16(bits(st.X[m]));
                        Y[2 * m] = lo16(bits(st.Y[m])), Y[2 * m + 1] = hi16(bits(st.Y[m]));
                        X2[2 * m] = lo16(bits(st.X2[m])), X2[2 * m + 1] = hi16(bits(st.X2[m]));
                        Y2[2 * m] = lo16(bits(st.Y2[m])), Y2[2 * m + 1] = hi16(bits(st.Y2[m]));
                    }
                    uint32_t smask = (bits(st.MM[0]) & 1) | (bits(st.MM[0]) >> 15 & 2) |
                                     (bits(st.MM[1]) & 1) << 2 | (bits(st.MM[1]) >> 13 & 8);
                    const uint32_t tcw[1] = {((uint32_t)bits(TB[0]) & 0xffu) | ((uint32_t)bits(TB[0]) >> 8 & 0xff00u) |
                                             ((uint32_t)bits(TB[1]) & 0xffu) << 16 | ((uint32_t)bits(TB[1]) & 0xff0000u) << 8};
                    const int cx = hi16(tx), cv = hi16(tv), cx2 = hi16(tx2);
                    // the AVX-512 lane wrap: position stB reads position stB+15
                    // (row 7 of the next lane) from anti-diagonal r-1
                    const bool wrap = b.st16 != b.stB;
                    int wx = 0, wvv = 0, wx2 = 0;
                    if (wrap && b.stB >= t0 && b.stB < t0 + SR) {
                        const int src = (b.stB + 15 - t0) / R;  // position stB+15 is row R-1 of this lane
                        wx = __builtin_amdgcn_readlane(X[R - 1], src);
                        wvv = __builtin_amdgcn_readlane(V[R - 1], src);
                        wx2 = __builtin_amdgcn_readlane(X2[R - 1], src);
                    }
                    const bool keep_prev = b.st16 > 0 && b.st16 - 1 >= lst16 && b.st16 - 1 <= len16;
                    const int edgeV = -8 * edge_diff(sc, r);
                    const bool inject = b.en16 >= r;
                    const bool la = tl >= b.stB && tl + R - 1 <= b.en16;
                    uint32_t dw[2] = {0u, 0u};
#pragma unroll
                    for (int k = R - 1; k >= 0; --k) {
                        const int t = tl + k;
                        int xin = k > 0 ? X[k - 1] : cx, vin = k > 0 ? V[k - 1] : cv, x2in = k > 0 ? X2[k - 1] : cx2;
                        if (t == b.st16) {
                            if (b.st16 == 0) xin = L0, vin = edgeV, x2in = L0;
                            else if (!keep_prev) xin = L0, vin = U0, x2in = L0;
                        } else if (wrap && t == b.stB) {
                            xin = wx, vin = wvv, x2in = wx2;
                        }
                        int uin = U[k], yin = Y[k], y2in = Y2[k];
                        if (inject && t == r) uin = edgeV, yin = L0, y2in = L0;
                        if (t >= b.st0 && t <= b.sc_end) {
                            const uint32_t qc = (qwin[k >> 2] >> (8 * (k & 3))) & 0xffu;
                            const uint32_t tc = (tcw[k >> 2] >> (8 * (k & 3))) & 0xffu;
                            smask = (smask & ~(1u << k)) | (tc != qc ? 1u << k : 0u);
                        }
                        int un, vn, xn, yn, x2n, y2n;
                        const int d = band_cell<RIGHT, SCORE_ONLY>(kc, xin, vin, x2in, uin, yin, y2in,
                                                                   (smask >> k) & 1u, un, vn, xn, yn, x2n, y2n);
                        if (!SCORE_ONLY) {
                            dw[k >> 2] |= (uint32_t)d << (8 * (k & 3));
                            asm volatile("" : "+v"(dw[k >> 2]));
                        }
                        if (la) {
                            // H bookkeeping on [st0, en0]; the position above still
                            // holds its H of anti-diagonal r-1
                            const int ut = -(un >> 3), vt = -(vn >> 3);
                            if (t == b.en0) {
                                const int hab = k > 0 ? H[k - 1] : ch;
                                const int h = r == 0 ? vt - sc.qe_orig : b.en0 > 0 ? hab + ut : H[k] + vt;
                                H[k] = h;
                                if (b.en0 == tlen - 1 && h >= mte_h) mte_h = h, mte_q = r - b.en0;
                                if (r == nr - 1 && b.en0 == tlen - 1) score = h;
                            } else if (t >= b.st0 && t < b.en0) {
                                H[k] += vt;
                            }
                            if (t >= b.st0 && t <= b.en0) {
                                if (H[k] > 0) pos = 1;
                                if (t == b.st0 && r - b.st0 == qlen - 1 && H[k] >= mqe_h) mqe_h = H[k], mqe_t = t;
                            }
                            U[k] = un, V[k] = vn, X[k] = xn, Y[k] = yn, X2[k] = x2n, Y2[k] = y2n;
                        }
                    }
                    if (!SCORE_ONLY && la) store_dirs(prow + (tl - b.stB), dw);
                    // back to the packed form
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        st.U[m] = vec((int)__builtin_amdgcn_perm((unsigned)U[2 * m + 1], (unsigned)U[2 * m], 0x05040100u));
                        st.V[m] = vec((int)__builtin_amdgcn_perm((unsigned)V[2 * m + 1], (unsigned)V[2 * m], 0x05040100u));
                        st.X[m] = vec((int)__builtin_amdgcn_perm((unsigned)X[2 * m + 1], (unsigned)X[2 * m], 0x05040100u));
                        st.Y[m] = vec((int)__builtin_amdgcn_perm((unsigned)Y[2 * m + 1], (unsigned)Y[2 * m], 0x05040100u));
                        st.X2[m] =
                            vec((int)__builtin_amdgcn_perm((unsigned)X2[2 * m + 1], (unsigned)X2[2 * m], 0x05040100u));
                        st.Y2[m] =
                            vec((int)__builtin_amdgcn_perm((unsigned)Y2[2 * m + 1], (unsigned)Y2[2 * m], 0x05040100u));
                        st.MM[m] = uvec((int)((smask >> (2 * m) & 1u) | (smask >> (2 * m + 1) & 1u) << 16));
                    }
                }
                lst16 = b.st16, len16 = b.en16;
                // this stripe's bottom position for the next stripe, unscaled
                if (lane == kWave - 1) {
                    const int x = -((hi16(bits(st.X[1])) + bias) >> 3) - G1, v = -(hi16(bits(st.V[1])) >> 3);
                    const int x2 = -((hi16(bits(st.X2[1])) + bias) >> 3) - G2;
                    const uint64_t en = (uint64_t)(uint32_t)pack3(x, v, x2) | (uint64_t)(uint32_t)st.H[R - 1] << 32;
                    __hip_atomic_store((uint64_t *)(out_buf + 2 * r), en, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (((r - rb) & (kPublish - 1)) == kPublish - 1)
                    publish(prog + wv * kBandCtlStride, kk * (nr + 2) + r + 1);
            }
        }
        // past every need of the next stripe (of every later one when stopped)
        publish(prog + wv * kBandCtlStride, stop ? 0x3fffffff : kk * (nr + 2) + nr + 1);
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
    // each wave leaves its record; the last to take a ticket combines them
    // in wave order and writes the pair's results
    int32_t *rec = ctl + kBandCtlRec;
    pos |= late;
    int last = 0;
    if (lane == 0) {
        int32_t *m = rec + 8 * wv;
        m[0] = mqe_h, m[1] = mqe_t, m[2] = mte_h, m[3] = mte_q, m[4] = score, m[5] = pos;
        last = __hip_atomic_fetch_add(ctl + kBandCtlTicket, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == W - 1;
    }
    if (lane == 0 && last) {
        mqe_h = KSW_NEG_INF_I, mqe_t = -1, mte_h = KSW_NEG_INF_I, mte_q = -1, score = KSW_NEG_INF_I, pos = 0;
        for (int k = 0; k < W; ++k) {
            const int32_t *m = rec + 8 * k;
            if (m[0] > mqe_h || (m[0] == mqe_h && m[1] > mqe_t)) mqe_h = m[0], mqe_t = m[1];
            if (m[3] >= 0 && (mte_q < 0 || m[2] > mte_h)) mte_h = m[2], mte_q = m[3];
            score = max(score, m[4]);
            pos |= m[5];
        }
        KswOut o;
        o.score = score, o.mqe = mqe_h, o.mqe_t = mqe_t, o.mte = mte_h, o.mte_q = mte_q;
        o.max = 0, o.max_t = -1, o.max_q = -1;
        o.zdropped = r_end < nr ? 1 : 0, o.stopped = 0, o.reach_end = 0, o.n_cigar = 0;
        a.out[jb.out] = o;
        a.redo[job] = (pos ? 1 : 0) | a.force_redo;
    }
}

// Zeroes the control blocks of a launch's jobs.
__global__ void __launch_bounds__(256) ksw_band_ctl_kernel(int32_t *bnd, const GenJob *jobs, const int32_t *order)
{
    int32_t *ctl = bnd + jobs[order[blockIdx.x]].bnd_off;
    for (int k = threadIdx.x; k < kBandCtl; k += blockDim.x) ctl[k] = 0;
}

namespace {

// Workgroups of one band-kernel variant that the device holds at once.
int band_capacity(const void *fn, int threads)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess) return 0;
    return cus * per_cu;
}

}  // namespace

hipError_t launch_band(const BandArgs &a0, int team, bool right, bool score_only, hipStream_t st)
{
    if (a0.njobs <= 0) return hipSuccess;
    if (team < 1 || team > kBandTeam) return hipErrorInvalidValue;
    BandArgs a = a0;
    const int wpg = team < kBandGroup ? team : kBandGroup;
    a.wgs = (team + wpg - 1) / wpg;
    const int threads = kWave * wpg;
    void (*fn)(BandArgs) = score_only ? ksw_band_kernel<false, true>
                           : right    ? ksw_band_kernel<true, false>
                                      : ksw_band_kernel<false, false>;
    // half the device's capacity per launch: the workgroups of a launch
    // wait on one another, so all of them must be resident together, beside
    // the kernels of other streams (which finish on their own) and one other
    // waiting launch of at most half the device (the caller's admission)
    const int cap = band_capacity((const void *)fn, threads);
    if (cap < 2 * a.wgs) return hipErrorInvalidConfiguration;
    const int chunk = cap / 2 / a.wgs;
    for (int off = 0; off < a0.njobs; off += chunk) {
        a.order = a0.order + off;
        a.njobs = a0.njobs - off < chunk ? a0.njobs - off : chunk;
        hipLaunchKernelGGL(ksw_band_ctl_kernel, dim3(a.njobs), dim3(256), 0, st, a.bnd, a.jobs, a.order);
        hipLaunchKernelGGL(fn, dim3(a.njobs * a.wgs), dim3(threads), 0, st, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace kswhip
