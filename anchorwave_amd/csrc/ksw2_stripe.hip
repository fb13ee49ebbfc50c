// ksw2_stripe.hip -- gfx950 kernels for unbanded ksw_extd2 (the fast path).
//
// Recurrence: the two-piece-affine difference recurrence of
// minimap2/ksw2_extd2_sse.c:43-71 and :237-331 (left/right gap alignment,
// score-only), evaluated for every cell (i = target, j = query).  This file
// evaluates it in a different order and representation than the reference:
//
//  * Stripes and passes.  The target is cut into stripes of 512 rows.  Lane
//    k of a wavefront owns 8 consecutive rows of a stripe and sweeps the
//    query; at step c it is at column j = c - k, so the cell above (i-1, j)
//    was produced by lane k-1 one step earlier and arrives through a DPP
//    wave rotate.  A wave computes two stripes at once (a "pass"): stripe 2p
//    in the low 16 bits of every register and stripe 2p+1 in the high 16
//    bits, 64 steps behind, so every VALU instruction of the cell is a
//    packed 16-bit operation on two cells.  The high stripe's lane 0 takes
//    the low stripe's lane-63 output of the previous step (the rotate wraps
//    it around); the low stripe's lane 0 takes the previous pass's bottom row
//    from a per-pair column buffer; the left neighbour (i, j-1) stays in
//    registers.
//  * Teams.  One workgroup of W waves computes one pair; wave w takes passes
//    w, w+W, ...  Pass p trails pass p-1 by at least two 64-column blocks:
//    the producer writes its bottom-row columns to the column buffer and then
//    publishes a column count in LDS (workgroup-scope release); the consumer
//    polls that count (workgroup-scope acquire) before it reads.  All waves
//    of a team share one CU, so no cache maintenance beyond the CU is
//    involved.  W = 1 (no waiting at all) for pairs that are small against
//    the batch; long pairs get up to 16 waves.  A pair's team may be smaller
//    than the launch's workgroups: the extra waves end at once.
//  * Wide teams (ksw_stripe_wide_kernel).  Above 16 waves, or to give each
//    wave a SIMD of its own, a pair's passes go to several 4-wave
//    workgroups on different CUs; the counts and column-buffer entries are
//    then device-scope accesses, and the epilogue runs in a launch of its
//    own.
//  * Scaled, negated state.  Every difference value w is kept as -8*w plus a
//    3-bit tag, so a chain of v_pk_min_i16 yields both the maximum of the
//    five candidates (s, a, b, a2, b2) and the index the reference's strict
//    '>' (or '>=' for right alignment) chain would pick: ties resolve to the
//    smallest (left) or largest (right) index through the tag.  With
//    T = (x+v) + 8e - 8z (tag-free), the new x is min(T, 0) and the
//    continuation flag is the sign of T (left); right alignment keeps x
//    biased by -8 so that the sign of T is its '>= 0' flag.
//  * Exactness.  The reference computes on int8 lanes.  The host only routes
//    a call here when every in-band intermediate provably stays inside int8
//    (then 16-bit arithmetic gives the same values), the match score is 0
//    and no mismatch scores above it (then z = min(z, mat[0]) never binds and
//    H never exceeds 0, so ez->max stays 0), there is no band and no z-drop.
//
// Outputs per pair: the direction byte of every cell (bits as in
// minimap2/ksw2.h:156-159, layout dir_index()), or, for KSW_EZ_SINGLETRACK,
// the two difference tracks of every cell (the new v and u,
// minimap2/ksw2_extd2_avx.c:432-436, layout track_index()); and the H
// bookkeeping of minimap2/ksw2_extd2_sse.c:369-375 (score, mqe/mqe_t,
// mte/mte_q).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "ksw2_common.h"
#include "ksw2_kernels.h"
#include "ksw2_walk.h"

namespace kswhip {

namespace {

constexpr int kDppWaveShl1 = 0x130;
constexpr int kDppWaveRor1 = 0x13C;
constexpr int kLag = kWave;      // steps the high stripe trails the low one
constexpr int kAhead = 16;       // steps before a block its boundary entries are loaded

// What a pass stores per cell.
enum Mode { kDirs = 0, kScoreOnly = 1, kTracks = 2 };

typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned short v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int bits(v2s v) { return __builtin_bit_cast(int, v); }
__device__ __forceinline__ v2s vec(int x) { return __builtin_bit_cast(v2s, x); }
__device__ __forceinline__ v2s splat(int x) { return (v2s){(short)x, (short)x}; }
__device__ __forceinline__ v2s vmin(v2s a, v2s b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ int lo16(int x) { return (int)(int16_t)(x & 0xffff); }
__device__ __forceinline__ int hi16(int x) { return x >> 16; }

// (a & m) | (b & ~m): one v_bfi_b32
__device__ __forceinline__ int bfi(int m, int a, int b) { return (a & m) | (b & ~m); }

// (a & m) | b as one v_and_or_b32
__device__ __forceinline__ int and_or(int a, int m, int b)
{
    int r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(m), "v"(b));
    return r;
}

__device__ __forceinline__ int dpp_shl1(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, kDppWaveShl1, 0xF, 0xF, false);
}

// The value from the cell above for every lane: lane k takes lane k-1's
// (both halves); lane 0 takes lane 63's low half of the previous step into
// its high half, and the low stripe's boundary value (the high half of b_hi)
// into its low half.
__device__ __forceinline__ int carry(int y, int b_hi, bool lane0)
{
    const int rot = __builtin_amdgcn_update_dpp(0, y, kDppWaveRor1, 0xF, 0xF, false);
    const int fix = (int)__builtin_amdgcn_alignbit((unsigned)rot, (unsigned)b_hi, 16);
    return lane0 ? fix : rot;
}

// Tags: state k = 0 (diagonal s), 1 (a), 2 (b), 3 (a2), 4 (b2).  Under
// v_min the smallest tag wins a tie, so left alignment (first strict maximum)
// tags k, right alignment (last maximum, '>=' chain) tags 7-k.
template <bool RIGHT> struct Tags {
    static constexpr int S = RIGHT ? 7 : 0, A = RIGHT ? 6 : 1, B = RIGHT ? 5 : 2;
    static constexpr int A2 = RIGHT ? 4 : 3, B2 = RIGHT ? 3 : 4;
};

// Constants of the scaled representation (both halves):
//   stored  U = -8u, V = -8v, X = -8(x+G1) - BIAS, X2 = -8(x2+G2) - BIAS,
//           Y, Y2 likewise; BIAS = 0 (left) or 8 (right)
//   cands   A = (X+V) + CA = -8a + tagA, ..., S = -8s + tagS
//   update  T = (X+V) + 8e - (Z & ~7) = -8(a-z+q) - BIAS, X' = min(T, -BIAS)
// A transposed pair (query in the stripes, target swept) has the cell above
// and the cell to the left exchanged: its A and A2 candidates are the
// reference's b and b2 (insertions), its B and B2 the reference's a and a2.
// They take those candidates' tags and direction-byte flag bits, so the min
// chain keeps the reference's tie order and the byte is the reference's byte.
struct Consts {
    v2s CA, CB, CA2, CB2, E1, E2, LIM, SM, SD;
    v2u shA, shB, shA2, shB2;  // logical shift taking a T's sign (bit 15) to its flag bit
    int mA, mB, mA2, mB2;      // that flag bit in both halves
};

template <bool RIGHT>
__device__ __forceinline__ Consts make_consts(const KswScore &sc, bool transposed)
{
    using T = Tags<RIGHT>;
    const int bias = RIGHT ? 8 : 0;
    const int tA = transposed ? T::B : T::A, tB = transposed ? T::A : T::B;
    const int tA2 = transposed ? T::B2 : T::A2, tB2 = transposed ? T::A2 : T::B2;
    // flag bits of the reference's byte: a 3, b 4, a2 5, b2 6
    const int fA = transposed ? 4 : 3, fB = transposed ? 3 : 4, fA2 = transposed ? 6 : 5, fB2 = transposed ? 5 : 6;
    Consts k;
    k.CA = splat(8 * sc.G1 + tA + bias);
    k.CB = splat(8 * sc.G1 + tB + bias);
    k.CA2 = splat(8 * sc.G2 + tA2 + bias);
    k.CB2 = splat(8 * sc.G2 + tB2 + bias);
    k.E1 = splat(8 * sc.e);
    k.E2 = splat(8 * sc.e2);
    k.LIM = splat(-bias);
    k.SM = splat(-8 * sc.mch + T::S);
    k.SD = splat(-8 * (sc.mis - sc.mch));
    k.shA = (v2u){(unsigned short)(15 - fA), (unsigned short)(15 - fA)};
    k.shB = (v2u){(unsigned short)(15 - fB), (unsigned short)(15 - fB)};
    k.shA2 = (v2u){(unsigned short)(15 - fA2), (unsigned short)(15 - fA2)};
    k.shB2 = (v2u){(unsigned short)(15 - fB2), (unsigned short)(15 - fB2)};
    k.mA = 0x10001 << fA, k.mB = 0x10001 << fB, k.mA2 = 0x10001 << fA2, k.mB2 = 0x10001 << fB2;
    return k;
}

// Two cells (low and high half).  In: X/X2/V from the cells above (replaced
// by these cells' outputs for the cells below), U/Y/Y2 from the cells to the
// left (replaced for the cells to the right), tb/qb the target and query
// codes.  Returns the two direction bytes in the low byte of each half (0
// when SCORE_ONLY).
template <bool RIGHT, bool SCORE_ONLY>
__device__ __forceinline__ int cell2(const Consts &k, v2s S, v2s &X, v2s &X2, v2s &V, v2s &U, v2s &Y, v2s &Y2)
{
    const v2s XV = X + V, X2V = X2 + V, YU = Y + U, Y2U = Y2 + U;
    const v2s A = XV + k.CA, B = YU + k.CB, A2 = X2V + k.CA2, B2 = Y2U + k.CB2;
    const v2s Z = vmin(vmin(vmin(vmin(S, A), B), A2), B2);
    const v2s Z8 = vec(bits(Z) & (int)0xFFF8FFF8u);
    const v2s M1 = k.E1 - Z8, M2 = k.E2 - Z8;
    const v2s Ta = XV + M1, Tb = YU + M1, Ta2 = X2V + M2, Tb2 = Y2U + M2;
    const v2s Un = Z8 - V, Vn = Z8 - U;
    X = vmin(Ta, k.LIM), Y = vmin(Tb, k.LIM), X2 = vmin(Ta2, k.LIM), Y2 = vmin(Tb2, k.LIM);
    U = Un, V = Vn;
    if (SCORE_ONLY) return 0;
    // state index (tags) and the four continuation flags (sign bits of T,
    // each shifted to its bit of the byte)
    int d = RIGHT ? (~bits(Z) & 0x00070007) : (bits(Z) & 0x00070007);
    d = and_or(bits(__builtin_bit_cast(v2s, __builtin_bit_cast(v2u, Ta) >> k.shA)), k.mA, d);
    d = and_or(bits(__builtin_bit_cast(v2s, __builtin_bit_cast(v2u, Tb) >> k.shB)), k.mB, d);
    d = and_or(bits(__builtin_bit_cast(v2s, __builtin_bit_cast(v2u, Ta2) >> k.shA2)), k.mA2, d);
    d = and_or(bits(__builtin_bit_cast(v2s, __builtin_bit_cast(v2u, Tb2) >> k.shB2)), k.mB2, d);
    return d;
}

// Score candidate S of two cells (match/mismatch) from their codes: with
// ONE = (1, 1) opaque to the optimiser, min(tb ^ qb, ONE) stays one
// v_pk_min_u16 (0 match, 1 mismatch) and the score one v_pk_mad_u16.
__device__ __forceinline__ v2s score_mm(const Consts &k, v2u tb, v2u qb, int one)
{
    const v2u mm = __builtin_elementwise_min(tb ^ qb, __builtin_bit_cast(v2u, one));
    return k.SM + __builtin_bit_cast(v2s, mm) * k.SD;
}

// Sets one half of a packed register.
__device__ __forceinline__ v2s set_half(v2s v, int value, bool hi)
{
    const int b = bits(v);
    return vec(hi ? (b & 0xffff) | (value << 16) : (int)((unsigned)b & 0xffff0000u) | (value & 0xffff));
}

// Team progress: wave w publishes, for its k-th pass, k*(qlen+1) + the
// number of bottom-row columns written so far (monotonic per wave).
__device__ __forceinline__ void publish(int32_t *slot, int value)
{
    // every lane stores the same value (no divergent branch); the release
    // orders the preceding column-buffer stores before the count
    __hip_atomic_store(slot, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Spins until the producer's published count reaches `need`.  The exit test
// uses lane 0's value, so the loop is uniform.  A producer always ends its
// pass at k*(qlen+1) + qlen, so `need` <= that bound terminates.
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

// Wide teams (waves on several CUs): counts and column-buffer entries are
// device-scope accesses, coherent across the XCDs' L2s; a count is stored
// only after the entries it covers are complete (s_waitcnt vmcnt(0)).
constexpr uint64_t kWaitTicks = 1000000000ull;  // 10 s of the 100 MHz real-time clock
constexpr int kStopCount = 0x3fffffff;          // published by a stopped wave: past every need

__device__ __forceinline__ void publish_wide(int32_t *slot, int value)
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
    asm volatile("" ::: "memory");
    if (__lane_id() == 0) __hip_atomic_store(slot, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int load_count_wide(const int32_t *slot)
{
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// As wait_count; past kWaitTicks (a workgroup the dispatcher could not
// place) it gives up, sets `late` and returns a count that satisfies every
// later wait, so every wave still runs to its end.
__device__ __forceinline__ int wait_count_wide(const int32_t *slot, int need, int known, int &late)
{
    if (known >= need) return known;
    int seen = load_count_wide(slot);
    if (seen < need) {
        const uint64_t t_start = wall_clock64();
        while (seen < need) {
            __builtin_amdgcn_s_sleep(1);
            seen = load_count_wide(slot);
            if (seen < need && wall_clock64() - t_start > kWaitTicks) {
                late = 1;
                seen = INT32_MAX;
            }
        }
    }
    asm volatile("" ::: "memory");  // the entry loads stay behind the wait
    return seen;
}

// Two column-buffer dwords (one entry): plain accesses within a workgroup,
// device-scope ones across CUs.
template <bool WIDE>
__device__ __forceinline__ void load_entry(const int32_t *src, int &a, int &b)
{
    if (WIDE) {
        const uint64_t v = __hip_atomic_load((const uint64_t *)src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a = (int)(uint32_t)v, b = (int)(uint32_t)(v >> 32);
    } else {
        a = src[0], b = src[1];
    }
}

template <bool WIDE>
__device__ __forceinline__ void store_entry(int32_t *dst, int a, int b)
{
    if (WIDE) {
        const uint64_t v = (uint64_t)(uint32_t)a | (uint64_t)(uint32_t)b << 32;
        __hip_atomic_store((uint64_t *)dst, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        dst[0] = a, dst[1] = b;
    }
}

// mqe partial of one stripe from the snapshot of U at column qlen-1 (one
// half of Us): H(i,qlen-1) = H(-1,qlen-1) + sum_{i'<=i} u(i'), kept in the
// -8 domain; the smallest running sum (largest H) with ties to the larger
// row, and the stripe total.
template <int R>
__device__ __forceinline__ void stripe_rec(const v2s (&Us)[R], bool hi, int i0, int tlen, int lane, StripeRec *out)
{
    int u[R], lsum = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = hi ? hi16(bits(Us[r])) : lo16(bits(Us[r])), lsum += u[r];
    int incl = lsum;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int o = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += o;
    }
    int run = incl - lsum;
    int best = 0x7fffffff, best_i = -1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        run += u[r];
        const int i = i0 + r;
        if (i < tlen && run <= best) best = run, best_i = i;
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int ob = __shfl_xor(best, off, kWave);
        const int oi = __shfl_xor(best_i, off, kWave);
        if (ob < best || (ob == best && oi > best_i)) best = ob, best_i = oi;
    }
    const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
    if (lane == 0) {
        StripeRec rec;
        rec.total = total, rec.best = best, rec.best_i = best_i, rec.pad = 0;
        *out = rec;
    }
}

// One pass (stripes 2p and 2p+1) of one pair, computed by the calling wave.
// Roles (uniform): the first pass takes row -1 as the low stripe's upper
// boundary instead of the column buffer; the last pass holds row tlen-1
// (score / mte) and publishes nothing; every other pass publishes the high
// stripe's bottom row.  Lanes compute every step; each half's left state is
// (re)initialised at its first column and U is snapshot at its last one, so
// the cells a half computes outside its columns never reach a result.
// WIDE: the team spans several CUs (progress counts in device memory,
// kWideCtlStride dwords apart; a wait that gives up sets `late`).
template <bool RIGHT, int MODE, bool WIDE, bool LUT, int R>
__device__ __forceinline__ void pass_body(const StripeArgs &a, const Consts &kc, int job_id, int p, const KswJob &jb,
                                          int qlen, int tlen, int32_t *prog, int W, int &late, int &stop)
{
    constexpr int SR = kWave * R;  // rows per stripe
    const KswScore &sc = a.sc;
    const int nstripes = (tlen + SR - 1) / SR, npass = (nstripes + 1) / 2;
    const bool first = p == 0, last = p == npass - 1, has_hi = 2 * p + 1 < nstripes;
    const int lane = threadIdx.x % kWave;
    const bool lane0 = lane == 0;
    const uint8_t *qs = a.seq + jb.q_off;
    const uint8_t *ts = a.seq + jb.t_off;
    uint8_t *dir = a.dir + jb.dir_off;
    int32_t *bnd = a.bnd + jb.bnd_off;
    const int nsteps = qlen + kWave - 1 + kLag;
    // row tlen-1 (score, mte): stripe s_last, half h_last of the last pass
    const int h_last = ((tlen - 1) / SR) & 1;
    const int rho_last = last ? (tlen - 1) % R : -1;
    const int lane_last = ((tlen - 1) % SR) / R;
    // progress slots: pass p-1 is wave (p-1) % W's ((p-1)/W)-th pass
    constexpr int PS = WIDE ? kWideCtlStride : 1;
    const int32_t *prog_prev = prog + (p + W - 1) % W * PS;
    const int base_prev = first ? 0 : (p - 1) / W * (qlen + 1);
    int32_t *prog_mine = prog + p % W * PS;
    const int base_mine = p / W * (qlen + 1);

    const int ilo = 2 * p * SR + lane * R, ihi = ilo + SR;  // this lane's first rows
    v2u TB[R];
    v2s U[R], Y[R], Y2[R], Us[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int tl = ilo + r < tlen ? (int)ts[ilo + r] : 0, th = ihi + r < tlen ? (int)ts[ihi + r] : 0;
        // LUT: the v_perm selector of row r's two score bytes (table byte
        // tl of the low half's table, byte th of the high half's, zero high
        // bytes); else the two target codes
        TB[r] = LUT ? __builtin_bit_cast(v2u, (uint32_t)tl | 12u << 8 | (uint32_t)(th + 4) << 16 | 12u << 24)
                    : (v2u){(unsigned short)tl, (unsigned short)th};
        U[r] = (v2s){(short)(-8 * edge_diff(sc, ilo + r)), (short)(-8 * edge_diff(sc, ihi + r))};
        Y[r] = kc.LIM, Y2[r] = kc.LIM, Us[r] = U[r];
    }
    int row_acc = 0, mte_best = 0x7fffffff, mte_j = -1;  // row tlen-1 in the -8 domain
    int Xo = bits(kc.LIM), X2o = bits(kc.LIM), Vo = 0, Qo = 0;  // this lane's bottom-row outputs
    int bin0 = 0, bin1 = 0, bnx0 = 0, bnx1 = 0;  // lane-0 boundary queue (rotating), next block
    int out0 = 0, out1 = 0;                      // lane-63 output queue (rotating)
    int known = 0;                               // columns of pass p-1 known published
    auto wait_prev = [&](int need) {
        if (WIDE) known = wait_count_wide(prog_prev, need, known, late);
        else known = wait_count(prog_prev, need, known);
    };
    if (!first) {
        wait_prev(base_prev + min(kWave, qlen));
        if (lane < qlen) load_entry<WIDE>(bnd + 2 * lane, bin0, bin1);
    }
    const int lim_hi = bits(kc.LIM) << 16;
    // LUT: per step, byte k of a half's table is the score candidate of
    // target code k against that half's query code (SM when equal, SX
    // otherwise); the query codes travel as 8 * code
    const uint32_t sx_rep = (uint32_t)(lo16(bits(kc.SM + kc.SD)) & 0xff) * 0x01010101u;
    const uint32_t s_flip = (uint32_t)((lo16(bits(kc.SM)) ^ lo16(bits(kc.SM + kc.SD))) & 0xff);
    int one = 0x00010001;
    asm volatile("" : "+s"(one));
    uint32_t qword = 0;
    for (int c = 0; c < nsteps; ++c) {
        if (a.stop && (c & 255) == 0) {
            // the host's stop flag: publish a count past every wait of the
            // team and end (nothing of this pair is delivered)
            stop = __builtin_amdgcn_readfirstlane(__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
            if (stop) {
                if (WIDE) publish_wide(prog_mine, kStopCount);
                else publish(prog_mine, kStopCount);
                return;
            }
        }
        // lane 0's low-stripe inputs for column c, each in a high half
        int bxh, bx2h, bvh;
        if (first) {
            bxh = lim_hi, bx2h = lim_hi, bvh = (-8 * edge_diff(sc, c)) << 16;
        } else {
            bxh = bin0, bx2h = bin0 << 16, bvh = bin1;
        }
        if ((c & 3) == 0) qword = c < qlen ? *(const uint32_t *)(qs + c) : 0u;
        const int qh = (int)((qword >> (8 * (c & 3))) & 0xffu) << (LUT ? 19 : 16);
        v2s X = vec(carry(Xo, bxh, lane0)), X2 = vec(carry(X2o, bx2h, lane0)), V = vec(carry(Vo, bvh, lane0));
        const int Qw = carry(Qo, qh, lane0);
        if (!first) {
            // rotate the queue so lane 0 sees column c+1 next step; the
            // next block's entries are waited for and loaded kAhead steps
            // before the block starts (late enough that the producer need
            // only be that far ahead, early enough to hide the load)
            if ((c & (kWave - 1)) == kWave - 1 - kAhead) {
                const int nb = c + 1 + kAhead;  // the next block's first column
                wait_prev(base_prev + min(nb + kWave, qlen));
                bnx0 = 0, bnx1 = 0;
                if (nb + lane < qlen) load_entry<WIDE>(bnd + 2 * (nb + lane), bnx0, bnx1);
            }
            if ((c & (kWave - 1)) == kWave - 1) {
                bin0 = bnx0, bin1 = bnx1;
            } else {
                bin0 = dpp_shl1(0, bin0), bin1 = dpp_shl1(0, bin1);
            }
        }
        // a half starts at its first column: the lanes there take the
        // half's initial left state (U from Us, which holds it until the
        // half's last column; only the first 128 steps can hold a start, the
        // guard is wave-uniform)
        const int jl = c - lane;
        if (c < kWave + kLag) {
            const int m = jl == 0 ? 0x0000ffff : jl == kLag ? (int)0xffff0000u : 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                U[r] = vec(bfi(m, bits(Us[r]), bits(U[r])));
                Y[r] = vec(bfi(m, bits(kc.LIM), bits(Y[r]))), Y2[r] = vec(bfi(m, bits(kc.LIM), bits(Y2[r])));
            }
        }
        const v2u qb = __builtin_bit_cast(v2u, Qw);
        uint32_t tab_lo = 0, tab_hi = 0;
        if (LUT) {
            tab_lo = sx_rep ^ (s_flip << (Qw & 31));
            tab_hi = sx_rep ^ (s_flip << ((Qw >> 16) & 31));
        }
        auto score = [&](int r) -> v2s {
            return LUT ? vec((int)__builtin_amdgcn_perm(tab_hi, tab_lo, bits(__builtin_bit_cast(v2s, TB[r]))))
                       : score_mm(kc, TB[r], qb, one);
        };
        int d[R];
        if (!last) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                d[r] = cell2<RIGHT, MODE != kDirs>(kc, score(r), X, X2, V, U[r], Y[r], Y2[r]);
                if (MODE == kTracks) d[r] = bits(V);
            }
        } else {
            // the last pass also follows V along row tlen-1
            v2s vl = V;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                d[r] = cell2<RIGHT, MODE != kDirs>(kc, score(r), X, X2, V, U[r], Y[r], Y2[r]);
                if (MODE == kTracks) d[r] = bits(V);
                if (r == rho_last) vl = V;
            }
            const int jh = jl - kLag * h_last;
            const int v = h_last ? hi16(bits(vl)) : lo16(bits(vl));
            if (lane == lane_last && jh >= 0 && jh < qlen) {
                row_acc += v;
                if (row_acc <= mte_best) mte_best = row_acc, mte_j = jh;
            }
        }
        // a half ends at column qlen-1: keep U for the mqe partials
        if (c >= qlen - 1 && c < qlen - 1 + kWave + kLag) {
            const int je = jl - (qlen - 1);
            const int m = je == 0 ? 0x0000ffff : je == kLag ? (int)0xffff0000u : 0;
#pragma unroll
            for (int r = 0; r < R; ++r) Us[r] = vec(bfi(m, bits(U[r]), bits(Us[r])));
        }
        Xo = bits(X), X2o = bits(X2), Vo = bits(V), Qo = Qw;
        if (MODE == kDirs) {
            // 2R bytes per lane: byte 4g + 2h + (rho & 1) is row 2g + (rho & 1)
            // of half h
            uint8_t *dst = dir + (((uint64_t)p * nsteps + c) * kWave + lane) * (2 * R);
            if (R == 8) {
                uint4 w;
                w.x = __builtin_amdgcn_perm((unsigned)d[1], (unsigned)d[0], 0x06020400u);
                w.y = __builtin_amdgcn_perm((unsigned)d[3], (unsigned)d[2], 0x06020400u);
                w.z = __builtin_amdgcn_perm((unsigned)d[5 % R], (unsigned)d[4 % R], 0x06020400u);
                w.w = __builtin_amdgcn_perm((unsigned)d[7 % R], (unsigned)d[6 % R], 0x06020400u);
                *(uint4 *)dst = w;
            } else {
                uint2 w;
                w.x = __builtin_amdgcn_perm((unsigned)d[1], (unsigned)d[0], 0x06020400u);
                w.y = __builtin_amdgcn_perm((unsigned)d[3], (unsigned)d[2], 0x06020400u);
                *(uint2 *)dst = w;
            }
        } else if (MODE == kTracks) {
            // 4R bytes per lane: the p track (new V) then the p2 track (new
            // U), each in the byte order of the direction bytes; a value is
            // stored as -8w >> 3 = -w (exact: both are multiples of 8)
            int vt[R], ut[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                vt[r] = bits(vec(d[r]) >> (v2s){3, 3});
                ut[r] = bits(U[r] >> (v2s){3, 3});
            }
            uint8_t *dst = dir + (((uint64_t)p * nsteps + c) * kWave + lane) * (4 * R);
            if (R == 8) {
                uint4 wv, wu;
                wv.x = __builtin_amdgcn_perm((unsigned)vt[1], (unsigned)vt[0], 0x06020400u);
                wv.y = __builtin_amdgcn_perm((unsigned)vt[3], (unsigned)vt[2], 0x06020400u);
                wv.z = __builtin_amdgcn_perm((unsigned)vt[5 % R], (unsigned)vt[4 % R], 0x06020400u);
                wv.w = __builtin_amdgcn_perm((unsigned)vt[7 % R], (unsigned)vt[6 % R], 0x06020400u);
                wu.x = __builtin_amdgcn_perm((unsigned)ut[1], (unsigned)ut[0], 0x06020400u);
                wu.y = __builtin_amdgcn_perm((unsigned)ut[3], (unsigned)ut[2], 0x06020400u);
                wu.z = __builtin_amdgcn_perm((unsigned)ut[5 % R], (unsigned)ut[4 % R], 0x06020400u);
                wu.w = __builtin_amdgcn_perm((unsigned)ut[7 % R], (unsigned)ut[6 % R], 0x06020400u);
                ((uint4 *)dst)[0] = wv, ((uint4 *)dst)[1] = wu;
            } else {
                uint2 wv, wu;
                wv.x = __builtin_amdgcn_perm((unsigned)vt[1], (unsigned)vt[0], 0x06020400u);
                wv.y = __builtin_amdgcn_perm((unsigned)vt[3], (unsigned)vt[2], 0x06020400u);
                wu.x = __builtin_amdgcn_perm((unsigned)ut[1], (unsigned)ut[0], 0x06020400u);
                wu.y = __builtin_amdgcn_perm((unsigned)ut[3], (unsigned)ut[2], 0x06020400u);
                ((uint2 *)dst)[0] = wv, ((uint2 *)dst)[1] = wu;
            }
        }
        if (!last) {
            // lane 63 appends the high stripe's bottom-row output (column
            // c-127: x in the high half, x2 in the low half; v in the high
            // half) to the queue; every 64 steps the queue is written out
            const int w0 = (int)__builtin_amdgcn_perm((unsigned)Xo, (unsigned)X2o, 0x07060302u);
            out0 = dpp_shl1(w0, out0), out1 = dpp_shl1(Vo, out1);
            if ((c & (kWave - 1)) == kWave - 1 || c == nsteps - 1) {
                const int col = c - (kWave - 1) - kLag - (kWave - 1) + lane;
                if (col >= 0 && col < qlen) store_entry<WIDE>(bnd + 2 * col, out0, out1);
                const int count = base_mine + max(0, min(c - (kWave - 1) - kLag + 1, qlen));
                if (WIDE) publish_wide(prog_mine, count);
                else publish(prog_mine, count);
            }
        }
    }
    stripe_rec<R>(Us, false, ilo, tlen, lane, a.recs + jb.stripe_off + 2 * p);
    if (has_hi) stripe_rec<R>(Us, true, ihi, tlen, lane, a.recs + jb.stripe_off + 2 * p + 1);
    if (last) {
        // score / mte from the lane that owns row tlen-1
        const int sc_row = __builtin_amdgcn_readlane(row_acc, lane_last);
        const int mte_b = __builtin_amdgcn_readlane(mte_best, lane_last);
        const int mte_jj = __builtin_amdgcn_readlane(mte_j, lane_last);
        if (lane == 0) {
            const int64_t h_anchor = (int64_t)sc.G1 - sc.qe_orig;
            const int h_left_last_row = (int)(h_anchor + edge_prefix(sc, tlen - 1));  // H(tlen-1, -1)
            KswOut o;
            o.score = h_left_last_row - (sc_row >> 3);
            o.mte = h_left_last_row - (mte_b >> 3);
            o.mte_q = mte_jj;
            o.mqe = KSW_NEG_INF_I, o.mqe_t = -1;  // finalised by the backtrack kernel
            o.max = 0, o.max_t = -1, o.max_q = -1;
            o.zdropped = 0, o.stopped = 0, o.reach_end = 0, o.n_cigar = 0;
            a.out[job_id] = o;
        }
    }
}

}  // namespace

// Epilogue of one pair, run by the whole workgroup once every pass is done:
// folds the per-stripe mqe partials in row order (ties to the later row, the
// '>=' of minimap2/ksw2_extd2_sse.c:371), picks the start cell as in
// minimap2/ksw2_extd2_sse.c:398-418 and walks the direction bytes with the
// state machine of ksw_backtrack (minimap2/ksw2.h:160-192) for an unbanded
// matrix (no cell of the path can leave the band there, so the off/off_end
// forcing never triggers).  The walk reads a tile of 8 lanes x 64 columns (up
// to 64 rows x 64 columns of one stripe) staged in LDS by the workgroup; when
// the path leaves the tile the next one is staged, so HBM latency is paid
// once per tile rather than once per step.  Every thread runs the same walk
// (uniform control flow, LDS broadcast reads) while other workgroups keep the
// SIMDs busy; thread 0 writes the run-length encoded CIGAR, which is reversed
// at the end unless KSW_EZ_REV_CIGAR.  With a host mirror, the result and the
// CIGAR are then copied to pinned host memory and the pair's completion flag
// is raised (system-scope release), so a caller can take this pair's result
// while the launch still runs.
//
// With the two Singletrack tracks instead of direction bytes, the first wave
// runs the Singletrack walk (track_walk) for a global alignment; with
// KSW_EZ_EXTZ_ONLY or KSW_EZ_SEMIGLOBAL_END the result is zdropped with no
// path (minimap2/ksw2_extd2_avx.c:752-760), and a failed walk reports
// n_cigar = -1 (no CIGAR, zdropped).

// Singletrack tracks of one stripe-kernel pair (layout of the tracks mode),
// read in the reference's coordinates (i target, j query).  A transposed
// pair stores cell (i, j) at (j, i) and its horizontal and vertical
// differences exchanged.
template <int R>
struct StripeTracks {
    static constexpr int SR = kWave * R;
    const uint8_t *dir;
    const uint8_t *qs, *ts;   // the reference's query and target
    uint64_t nsteps;
    int32_t mch, mis;
    bool tr;
    __device__ int32_t get(int track, int i, int j) const
    {
        if (tr) {
            const int t = i;
            i = j, j = t, track = 1 - track;
        }
        const int s = i / SR, pp = s >> 1, h = s & 1, k = (i % SR) / R, rho = i % R;
        const uint64_t c = (uint64_t)(j + k + kLag * h);
        const uint8_t b =
            dir[((pp * nsteps + c) * kWave + k) * (4 * R) + track * (2 * R) + 4 * (rho >> 1) + 2 * h + (rho & 1)];
        return -(int32_t)(int8_t)b;
    }
    __device__ int forced(int, int) const { return 0; }
    __device__ int32_t cost(int i, int j) const { return ts[i] == qs[j] ? mch : mis; }
};

template <int R>
__device__ __forceinline__ void pair_epilogue(const StripeArgs &a, int job_id, const KswJob &jb,
                                              uint64_t (*tile)[64], bool tracks, int *shared_n, int nthr)
{
    constexpr int SR = kWave * R;  // rows per stripe
    const int tid = threadIdx.x;
    // the host's stop flag, read once for the whole workgroup: a stopped
    // pair's passes may have ended early, so it takes no walk
    int stopped = 0;
    if (a.stop) {
        if (tid == 0) *shared_n = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();
        stopped = *shared_n;
        __syncthreads();
    }
    KswOut o = a.out[job_id];
    if (stopped) o.stopped = 1;
    const int flag = jb.flag;
    // kq, kt: the swept and striped lengths; qlen, tlen: the reference's
    const bool tr = jb.transposed != 0;
    const int kq = jb.qlen, kt = jb.tlen;
    const int qlen = tr ? kt : kq, tlen = tr ? kq : kt;
    {
        // the stripes' partials at the last swept column, in row order
        const int nstripes = (kt + SR - 1) / SR;
        int carry_ = 0, best = 0x7fffffff, best_i = -1;
        for (int s = 0; s < nstripes; ++s) {
            const StripeRec rec = a.recs[jb.stripe_off + s];
            if (rec.best_i >= 0 && carry_ + rec.best <= best) best = carry_ + rec.best, best_i = rec.best_i;
            carry_ += rec.total;
        }
        const int64_t h_anchor = (int64_t)a.sc.G1 - a.sc.qe_orig;
        const int h_top_last_col = (int)(h_anchor + edge_prefix(a.sc, kq - 1));  // H(-1, kq-1)
        const int col_max = h_top_last_col - (best >> 3);
        if (!tr) {
            o.mqe = col_max, o.mqe_t = best_i;
        } else {
            // the last swept column is the reference's last row (mte) and the
            // last striped row, tracked by the last pass, its last column (mqe)
            o.mqe = o.mte, o.mqe_t = o.mte_q;
            o.mte = col_max, o.mte_q = best_i;
        }
    }
    int i0 = -1, j0 = -1;
    int n = 0;
    uint32_t *cig = a.cig + jb.cig_off;
    if (stopped) {
        // no walk: the passes, the stripe partials and this pair's result
        // slot may be incomplete
    } else if (tracks) {
        if (!o.stopped && !o.zdropped && !(flag & (0x40 | 0x800))) {
            if (tid < kWave) {
                StripeTracks<R> A;
                A.dir = a.dir + jb.dir_off;
                A.qs = a.seq + (tr ? jb.t_off : jb.q_off), A.ts = a.seq + (tr ? jb.q_off : jb.t_off);
                A.nsteps = (uint64_t)kq + kWave - 1 + kLag;
                A.mch = a.sc.mch, A.mis = a.sc.mis;
                A.tr = tr;
                const WalkGaps g = {-a.sc.q, -a.sc.e, -a.sc.q2, -a.sc.e2};
                const int nn = track_walk(A, qlen, tlen, g, cig);
                if (tid == 0) *shared_n = nn;
            }
            __syncthreads();
            n = *shared_n;
            if (n < 0) o.zdropped = 1;
        } else if (!o.stopped && !o.zdropped) {
            o.zdropped = 1;
        }
    } else if (!(flag & 0x01)) {  // KSW_EZ_SCORE_ONLY: no path
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
    if (i0 >= 0 || j0 >= 0) {
        const uint8_t *dir = a.dir + jb.dir_off;
        const uint64_t nsteps = (uint64_t)kq + kWave - 1 + kLag;
        int i = i0, j = j0, state = 0;
        // the staged tile, in stored coordinates (striped row, swept column)
        int t_r_lo = 1, t_r_hi = 0, t_c_lo = 1, t_c_hi = 0;
        int run_op = -1, run_len = 0;
        auto push = [&](int op, int len) {
            if (op == run_op) {
                run_len += len;
            } else {
                if (run_len > 0 && tid == 0) cig[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
                if (run_len > 0) ++n;
                run_op = op, run_len = len;
            }
        };
        while (i >= 0 && j >= 0) {
            const int r = tr ? j : i, c = tr ? i : j;  // stored coordinates of cell (i, j)
            if (r < t_r_lo || r > t_r_hi || c < t_c_lo || c > t_c_hi) {
                // stage lanes kb..k of stripe s (rows up to r) x columns c-63..c
                const int s = r / SR, k = (r - s * SR) / R, kb = k >= 7 ? k - 7 : 0;
                const int pp = s >> 1, h = s & 1;
                t_r_lo = s * SR + kb * R, t_r_hi = s * SR + k * R + R - 1;
                t_c_hi = c, t_c_lo = c >= 63 ? c - 63 : 0;
                const int c_lo = t_c_lo + kb + kLag * h, c_hi = t_c_hi + k + kLag * h;
                __syncthreads();  // the previous tile is no longer read
                for (int cc = c_lo + tid; cc <= c_hi; cc += nthr) {
                    const uint8_t *src = dir + (((uint64_t)pp * nsteps + (uint64_t)cc) * kWave + kb) * (2 * R);
                    for (int kr = 0; kr <= k - kb; ++kr) {
                        const int jj = cc - kLag * h - (kb + kr);
                        if (jj >= t_c_lo && jj <= t_c_hi) {
                            // the R row bytes of half h: bytes 2h, 2h+1 of each dword
                            const unsigned sel = h ? 0x07060302u : 0x05040100u;
                            if (R == 8) {
                                const uint4 w = ((const uint4 *)src)[kr];
                                const uint32_t lo = __builtin_amdgcn_perm(w.y, w.x, sel);
                                const uint32_t hi = __builtin_amdgcn_perm(w.w, w.z, sel);
                                tile[kr][jj - t_c_lo] = (uint64_t)hi << 32 | lo;
                            } else {
                                const uint2 w = ((const uint2 *)src)[kr];
                                tile[kr][jj - t_c_lo] = __builtin_amdgcn_perm(w.y, w.x, sel);
                            }
                        }
                    }
                }
                __syncthreads();
            }
            const int rel = r - t_r_lo;
            if (state == 0) {
                // A run of diagonal moves at once: from state 0 the walk
                // stays in state 0 and moves diagonally while each cell's
                // argmax (its low three bits) is 0.  Lane k tests cell
                // (r - k, c - k) of the staged tile (within the tile and the
                // matrix); the run is the lanes before the first that fails.
                const int lim = min(min(rel, c - t_c_lo), min(i, j));
                const int lk = tid & (kWave - 1);
                int fail = 1;
                if (lk <= lim) {
                    const int rr = rel - lk;
                    fail = ((int)(tile[rr / R][c - lk - t_c_lo] >> (8 * (rr % R))) & 7) != 0;
                }
                const uint64_t fb = __ballot(fail);
                const int run = fb ? (int)__builtin_ctzll(fb) : kWave;
                if (run > 0) {
                    push(0, run);
                    i -= run, j -= run;
                    continue;
                }
            } else {
                // A run of gap moves at once: state s (1-4) stays while each
                // cell's continuation bit (s + 2) is set, moving i (s = 1, 3)
                // or j (s = 2, 4) back by one, which is one stored row or one
                // column.  Lane k tests the k-th cell of the run.
                const int s = state;
                const bool di = s == 1 || s == 3, rows = di != tr;
                const int lim = min(rows ? rel : c - t_c_lo, di ? i : j);
                const int lk = tid & (kWave - 1);
                int fail = 1;
                if (lk <= lim) {
                    const int rr = rows ? rel - lk : rel, cc = rows ? c : c - lk;
                    fail = !((int)(tile[rr / R][cc - t_c_lo] >> (8 * (rr % R) + s + 2)) & 1);
                }
                const uint64_t fb = __ballot(fail);
                const int run = fb ? (int)__builtin_ctzll(fb) : kWave;
                if (run > 0) {
                    push(di ? 2 : 1, run);
                    if (di) i -= run;
                    else j -= run;
                    continue;
                }
            }
            const int tmp = (int)(tile[rel / R][c - t_c_lo] >> (8 * (rel % R))) & 0xff;
            if (state == 0) state = tmp & 7;
            else if (!(tmp >> (state + 2) & 1)) state = 0;
            if (state == 0) state = tmp & 7;
            if (state == 0) push(0, 1), --i, --j;
            else if (state == 1 || state == 3) push(2, 1), --i;
            else push(1, 1), --j;
        }
        if (i >= 0) push(2, i + 1);
        if (j >= 0) push(1, j + 1);
        push(-2, 0);  // flush the last run
        __syncthreads();
        if (!(flag & 0x80))
            for (int x = tid; x < (n >> 1); x += nthr) {
                const uint32_t t = cig[x];
                cig[x] = cig[n - 1 - x];
                cig[n - 1 - x] = t;
            }
    }
    o.n_cigar = n;
    if (tid == 0) a.out[job_id] = o;
    if (a.host_flag) {
        __syncthreads();
        uint32_t *hc = a.host_cig + jb.cig_off;
        if (n < 0) n = 0;
        for (int x = tid; x < n; x += nthr) hc[x] = cig[x];
        if (tid == 0) a.host_out[job_id] = o;
        __threadfence_system();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(a.host_flag + job_id, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// One workgroup per pair of this launch's job list.  A pair's team is its
// own W waves (jb.team, at most blockDim.x / 64): the waves beyond it end at
// once (s_barrier then waits on the surviving waves only), so pairs of
// different team sizes share one launch.  Wave w computes passes w, w+W,
// ...; then the team runs the pair's epilogue (mqe, start cell, backtrack).
template <bool RIGHT, int MODE, int R>
__global__ void __launch_bounds__(kWave * kMaxTeam) ksw_stripe_kernel(StripeArgs a)
{
    constexpr int SR = kWave * R;
    __shared__ int32_t prog[kMaxTeam];
    __shared__ uint64_t tile[8][64];
    __shared__ int shared_n;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int job = __builtin_amdgcn_readfirstlane(a.order[blockIdx.x]);
    const KswJob jb = a.jobs[job];
    const int W = __builtin_amdgcn_readfirstlane(max(1, min(jb.team, (int)blockDim.x / kWave)));
    if (w >= W) return;
    prog[w] = 0;  // every lane of wave w writes the same zero
    __syncthreads();
    const Consts kc = make_consts<RIGHT>(a.sc, jb.transposed != 0);
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    const int npass = ((tlen + SR - 1) / SR + 1) / 2;
    int late = 0, stop = 0;  // (no bounded waits within a workgroup)
    if (jb.lut)
        for (int p = w; p < npass && !stop; p += W)
            pass_body<RIGHT, MODE, false, true, R>(a, kc, job, p, jb, qlen, tlen, prog, W, late, stop);
    else
        for (int p = w; p < npass && !stop; p += W)
            pass_body<RIGHT, MODE, false, false, R>(a, kc, job, p, jb, qlen, tlen, prog, W, late, stop);
    __syncthreads();  // every pass written: stripe partials, score / mte, direction bytes
    pair_epilogue<R>(a, job, jb, tile, MODE == kTracks, &shared_n, W * kWave);
}

// Persistent one-wave workers over a device work queue.  The launch's job
// list is the queue, ordered longest first by the host (length-binned: the
// longest pairs are claimed first, short ones fill the tail); each wave
// claims the next entry with one device-scope atomic add on qhead, runs the
// pair's passes and its epilogue, and claims again until the list is
// exhausted, then ends.  No wave waits on another: a one-wave pair hands
// each pass to the next through the wave's own LDS count.  qhead is zero at
// launch (ksw_stripe_qhead_kernel, a fill kernel rather than a memset so a
// captured graph replays it).
template <bool RIGHT, int MODE, int R>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4)))
ksw_stripe_queue_kernel(StripeArgs a, int32_t *qhead)
{
    constexpr int SR = kWave * R;
    __shared__ int32_t prog[1];
    __shared__ uint64_t tile[8][64];
    __shared__ int shared_n;
    for (;;) {
        int idx = 0;
        if (threadIdx.x == 0) idx = __hip_atomic_fetch_add(qhead, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        idx = __builtin_amdgcn_readlane(idx, 0);
        if (idx >= a.njobs) return;
        const int job = __builtin_amdgcn_readfirstlane(a.order[idx]);
        const KswJob jb = a.jobs[job];
        prog[0] = 0;
        __syncthreads();
        const Consts kc = make_consts<RIGHT>(a.sc, jb.transposed != 0);
        const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
        const int npass = ((tlen + SR - 1) / SR + 1) / 2;
        int late = 0, stop = 0;
        if (jb.lut)
            for (int p = 0; p < npass && !stop; ++p)
                pass_body<RIGHT, MODE, false, true, R>(a, kc, job, p, jb, qlen, tlen, prog, 1, late, stop);
        else
            for (int p = 0; p < npass && !stop; ++p)
                pass_body<RIGHT, MODE, false, false, R>(a, kc, job, p, jb, qlen, tlen, prog, 1, late, stop);
        __syncthreads();
        pair_epilogue<R>(a, job, jb, tile, MODE == kTracks, &shared_n, kWave);
        __syncthreads();  // tile and prog are the next pair's
    }
}

// Zeroes the queue heads of a plan's queue launches.
__global__ void __launch_bounds__(64) ksw_stripe_qhead_kernel(int32_t *qhead, int n)
{
    if ((int)threadIdx.x < n) qhead[threadIdx.x] = 0;
}

// Wide teams: each pair has its own run of consecutive blocks (a.wfirst),
// which the dispatcher spreads over the XCDs' CUs; wave v of a team of W
// waves computes passes v, v+W, ...  A pair's control block holds one
// progress count per wave; a wave whose wait gave up raises its late word.
template <bool RIGHT, int MODE, int R>
__global__ void __launch_bounds__(kWave * kMaxTeam) ksw_stripe_wide_kernel(StripeArgs a)
{
    constexpr int SR = kWave * R;
    // the pair of this block: wfirst[slot] <= block < wfirst[slot + 1]
    const int blk = a.wfirst[0] + (int)blockIdx.x;
    int lo = 0, hi = a.njobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.wfirst[mid] <= blk) lo = mid;
        else hi = mid;
    }
    const int slot = __builtin_amdgcn_readfirstlane(lo);
    const int first = __builtin_amdgcn_readfirstlane(a.wfirst[slot]);
    const int W = __builtin_amdgcn_readfirstlane(a.wfirst[slot + 1] - first) * kWideGroup;
    const int v = __builtin_amdgcn_readfirstlane((blk - first) * kWideGroup + (int)threadIdx.x / kWave);
    const int job = __builtin_amdgcn_readfirstlane(a.order[slot]);
    int32_t *ctl = a.wctl + (int64_t)slot * kWideCtl;
    const KswJob jb = a.jobs[job];
    const Consts kc = make_consts<RIGHT>(a.sc, jb.transposed != 0);
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    const int npass = ((tlen + SR - 1) / SR + 1) / 2;
    int late = 0, stop = 0;
    if (jb.lut)
        for (int p = v; p < npass && !stop; p += W)
            pass_body<RIGHT, MODE, true, true, R>(a, kc, job, p, jb, qlen, tlen, ctl, W, late, stop);
    else
        for (int p = v; p < npass && !stop; p += W)
            pass_body<RIGHT, MODE, true, false, R>(a, kc, job, p, jb, qlen, tlen, ctl, W, late, stop);
    if ((late || a.force_late) && threadIdx.x % kWave == 0)
        __hip_atomic_store(ctl + kWideCtlLate, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Zeroes the control blocks of a wide launch's pairs (a fill kernel rather
// than a memset, so a captured graph replays it correctly).
__global__ void __launch_bounds__(256) ksw_stripe_wctl_kernel(int32_t *wctl)
{
    int32_t *ctl = wctl + (int64_t)blockIdx.x * kWideCtl;
    for (int k = threadIdx.x; k < kWideCtl; k += blockDim.x) ctl[k] = 0;
}

// The epilogue of every pair of a wide launch, one workgroup per pair, after
// all of the launch's passes.  A pair whose team gave up a wait (its late
// word is raised) is first recomputed by this workgroup as a one-workgroup
// team of kWideGroup waves (LDS progress counts, no cross-workgroup waits),
// into the same direction bytes, column buffer and stripe partials; the
// plan's redo word counts such pairs.
template <bool RIGHT, int MODE, int R>
__global__ void __launch_bounds__(kWave * kWideGroup) ksw_stripe_wide_epilogue_kernel(StripeArgs a)
{
    constexpr int SR = kWave * R;
    __shared__ int32_t prog[kWideGroup];
    __shared__ uint64_t tile[8][64];
    __shared__ int shared_n;
    const int slot = blockIdx.x;
    const int job = __builtin_amdgcn_readfirstlane(a.order[slot]);
    const KswJob jb = a.jobs[job];
    const int late = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(a.wctl + (int64_t)slot * kWideCtl + kWideCtlLate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (late) {
        const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
        prog[w] = 0;
        __syncthreads();
        const Consts kc = make_consts<RIGHT>(a.sc, jb.transposed != 0);
        const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
        const int npass = ((tlen + SR - 1) / SR + 1) / 2;
        int late2 = 0, stop = 0;
        for (int p = w; p < npass && !stop; p += kWideGroup) {
            if (jb.lut) pass_body<RIGHT, MODE, false, true, R>(a, kc, job, p, jb, qlen, tlen, prog, kWideGroup, late2, stop);
            else pass_body<RIGHT, MODE, false, false, R>(a, kc, job, p, jb, qlen, tlen, prog, kWideGroup, late2, stop);
        }
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(a.werr, 1);
    }
    pair_epilogue<R>(a, job, jb, tile, MODE == kTracks, &shared_n, (int)blockDim.x);
}

// Dispatch over the kernel variants: group (left, right, score-only,
// tracks) and rows per lane (8 or 4).
template <template <bool, int, int> class F, typename... Args>
static auto by_variant(int group, int R, Args &&...args)
{
    const bool r8 = R != 4;
    switch (group) {
    case kGroupRight:
        return r8 ? F<true, kDirs, 8>::run(args...) : F<true, kDirs, 4>::run(args...);
    case kGroupScoreOnly:
        return r8 ? F<false, kScoreOnly, 8>::run(args...) : F<false, kScoreOnly, 4>::run(args...);
    case kGroupTracks:
        return r8 ? F<false, kTracks, 8>::run(args...) : F<false, kTracks, 4>::run(args...);
    default:
        return r8 ? F<false, kDirs, 8>::run(args...) : F<false, kDirs, 4>::run(args...);
    }
}

template <bool RIGHT, int MODE, int R>
struct LaunchStripe {
    static hipError_t run(const StripeArgs &a, int team, hipStream_t st)
    {
        hipLaunchKernelGGL((ksw_stripe_kernel<RIGHT, MODE, R>), dim3(a.njobs), dim3(kWave * team), 0, st, a);
        return hipGetLastError();
    }
};

template <bool RIGHT, int MODE, int R>
struct LaunchQueue {
    static hipError_t run(const StripeArgs &a, int32_t *qhead, int workers, hipStream_t st)
    {
        hipLaunchKernelGGL((ksw_stripe_queue_kernel<RIGHT, MODE, R>), dim3(workers), dim3(kWave), 0, st, a, qhead);
        return hipGetLastError();
    }
};

template <bool RIGHT, int MODE, int R>
struct QueueWorkers {
    static int run()
    {
        int per_cu = 0, dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1024;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ksw_stripe_queue_kernel<RIGHT, MODE, R>, kWave, 0) !=
                hipSuccess ||
            per_cu <= 0)
            return prop.multiProcessorCount * 4;
        return per_cu * prop.multiProcessorCount;
    }
};

hipError_t launch_stripe_queue(const StripeArgs &a, int group, int R, int32_t *qhead, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (!qhead || (R != 4 && R != 8) || group < 0 || group >= kGroups) return hipErrorInvalidValue;
    return by_variant<LaunchQueue>(group, R, a, qhead, std::min(a.njobs, stripe_queue_workers(group, R)), st);
}

hipError_t clear_queue_heads(int32_t *qhead, int n, hipStream_t st)
{
    if (n < 1 || n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ksw_stripe_qhead_kernel, dim3(1), dim3(64), 0, st, qhead, n);
    return hipGetLastError();
}

// Cached per variant: the occupancy query reads the device properties.
int stripe_queue_workers(int group, int R)
{
    static int cache[kGroups][2] = {};
    if (group < 0 || group >= kGroups) group = kGroupLeft;
    const int ri = R == 4 ? 1 : 0;
    if (cache[group][ri] > 0) return cache[group][ri];
    const int n = by_variant<QueueWorkers>(group, R);
    __atomic_store_n(&cache[group][ri], n, __ATOMIC_RELAXED);
    return n;
}

hipError_t launch_stripe(const StripeArgs &a, int group, int team, int R, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (team < 1 || team > kMaxTeam || (R != 4 && R != 8) || group < 0 || group >= kGroups) return hipErrorInvalidValue;
    return by_variant<LaunchStripe>(group, R, a, team, st);
}

namespace {

template <bool RIGHT, int MODE, int R>
struct LaunchWide {
    static hipError_t run(const StripeArgs &a0, const int32_t *h_wfirst, int max_wgs, hipStream_t st)
    {
        void (*fn)(StripeArgs) = ksw_stripe_wide_kernel<RIGHT, MODE, R>;
        StripeArgs a = a0;
        int off = 0;
        while (off < a0.njobs) {
            // as many pairs as fit in max_wgs workgroups (at least one)
            int end = off + 1;
            while (end < a0.njobs && h_wfirst[end + 1] - h_wfirst[off] <= max_wgs) ++end;
            const int blocks = h_wfirst[end] - h_wfirst[off];
            if (blocks > max_wgs) return hipErrorInvalidConfiguration;
            a.order = a0.order + off;
            a.wfirst = a0.wfirst + off;
            a.wctl = a0.wctl + (int64_t)off * kWideCtl;
            a.njobs = end - off;
            hipLaunchKernelGGL(ksw_stripe_wctl_kernel, dim3(a.njobs), dim3(256), 0, st, a.wctl);
            hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWave * kWideGroup), 0, st, a);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            off = end;
        }
        return hipSuccess;
    }
};

template <bool RIGHT, int MODE, int R>
struct WideCapacity {
    static int run()
    {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)ksw_stripe_wide_kernel<RIGHT, MODE, R>,
                                                         kWave * kWideGroup, 0) != hipSuccess)
            return 0;
        return cus * per_cu;
    }
};

template <bool RIGHT, int MODE, int R>
struct LaunchWideEpilogue {
    static hipError_t run(const StripeArgs &a, hipStream_t st)
    {
        hipLaunchKernelGGL((ksw_stripe_wide_epilogue_kernel<RIGHT, MODE, R>), dim3(a.njobs), dim3(kWave * kWideGroup),
                           0, st, a);
        return hipGetLastError();
    }
};

template <bool RIGHT, int MODE, int R>
struct Resident {
    static int run()
    {
        int per_cu = 0, dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1024;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ksw_stripe_kernel<RIGHT, MODE, R>, kWave, 0) !=
                hipSuccess ||
            per_cu <= 0)
            return prop.multiProcessorCount * 4;
        return per_cu * prop.multiProcessorCount;
    }
};

}  // namespace

hipError_t launch_stripe_wide(const StripeArgs &a, int group, int R, const int32_t *h_wfirst, int max_wgs,
                              hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (!a.wfirst || !h_wfirst || !a.wctl || !a.werr || max_wgs < 1) return hipErrorInvalidValue;
    if ((R != 4 && R != 8) || group < 0 || group >= kGroups) return hipErrorInvalidValue;
    for (int k = 0; k < a.njobs; ++k) {
        const int n = h_wfirst[k + 1] - h_wfirst[k];
        if (n < 1 || n * kWideGroup > kMaxWide) return hipErrorInvalidValue;
    }
    return by_variant<LaunchWide>(group, R, a, h_wfirst, max_wgs, st);
}

// Cached per variant: the occupancy query reads the device properties.
int stripe_wide_capacity(int group, int R)
{
    static int cache[kGroups][2] = {};
    if (group < 0 || group >= kGroups) group = kGroupLeft;
    const int ri = R == 4 ? 1 : 0;
    if (cache[group][ri] > 0) return cache[group][ri];
    const int n = by_variant<WideCapacity>(group, R);
    __atomic_store_n(&cache[group][ri], n, __ATOMIC_RELAXED);
    return n;
}

hipError_t clear_wide_err(int32_t *werr, hipStream_t st)
{
    hipLaunchKernelGGL(ksw_stripe_wctl_kernel, dim3(1), dim3(256), 0, st, werr);
    return hipGetLastError();
}

hipError_t launch_stripe_wide_epilogue(const StripeArgs &a, int group, int R, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (!a.wctl || !a.werr) return hipErrorInvalidValue;
    if ((R != 4 && R != 8) || group < 0 || group >= kGroups) return hipErrorInvalidValue;
    return by_variant<LaunchWideEpilogue>(group, R, a, st);
}

// Cached per variant: the occupancy query reads the device properties.
int stripe_resident_waves(int group, int R)
{
    static int cache[kGroups][2] = {};
    if (group < 0 || group >= kGroups) group = kGroupLeft;
    const int ri = R == 4 ? 1 : 0;
    if (cache[group][ri] > 0) return cache[group][ri];
    const int n = by_variant<Resident>(group, R);
    __atomic_store_n(&cache[group][ri], n, __ATOMIC_RELAXED);
    return n;
}

}  // namespace kswhip
