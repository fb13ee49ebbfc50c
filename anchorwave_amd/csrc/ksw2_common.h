// ksw2_common.h -- structures shared by the host library and the gfx950 kernels.
//
// Data layout in HBM (one batch):
//   seq arena   query and target codes of every pair, each padded to 16 bytes
//   dir arena   direction bytes of every pair in stripe layout (see dir_index())
//   bnd arena   per pair, 2 dwords per query column: the bottom row of the
//               previous stripe (x, x2 packed as int16 pair; v)
//   cig arena   per pair, qlen+tlen+2 uint32 CIGAR slots
//   jobs / outs one KswJob and one KswOut per pair
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define KSW_HD
#else
#define KSW_HD __host__ __device__
#endif

namespace kswhip {

constexpr int kWave = 64;
constexpr int kMaxTeam = 16;   // waves cooperating on one pair (one workgroup)
constexpr int32_t KSW_NEG_INF_I = -0x40000000;

// Scoring shared by a batch, already normalised the way the reference
// normalises it (minimap2/ksw2_extd2_sse.c:83-110): pieces swapped so that
// q+e <= q2+e2, long-gap threshold and boundary difference precomputed.
struct KswScore {
    int32_t q, e, q2, e2;      // after the swap
    int32_t G1, G2;            // q+e, q2+e2
    int32_t qe_orig;           // q+e before the swap (H at r == 0 uses it)
    int32_t long_thres, long_diff;
    int32_t mch, mis;          // class-1 scores: equal codes / unequal codes
};

struct KswJob {
    uint64_t q_off, t_off;     // byte offsets in the seq arena
    uint64_t dir_off;          // byte offset in the dir arena
    uint64_t bnd_off;          // dword offset in the bnd arena
    uint64_t cig_off;          // uint32 offset in the cig arena
    int32_t qlen, tlen;
    int32_t flag, end_bonus;
    int32_t stripe_off;        // first slot of this pair in the per-stripe arrays
    int32_t R;                 // rows per lane of the stripe kernel that ran it
    int32_t team;              // waves of this pair's team (a launch's workgroups may be larger)
    int32_t lut;               // 1: codes below 4, score candidates from a per-step byte table
    // 1: the reference's query is in the stripes and its target is swept
    // (q_off/qlen then describe the reference's target, t_off/tlen its query)
    int32_t transposed;
    int32_t pad;
};

// Per-stripe partial of the mqe bookkeeping (column qlen-1): the sum of
// -8u over the stripe's rows, and the stripe-local minimum of the running
// sum with the row that attains it (ties to the larger row).
struct StripeRec {
    int32_t total, best, best_i, pad;
};

// Scoring of the general kernel: the normalised pieces plus the full matrix
// (KSW_EZ_GENERIC_SC) and the wildcard score of match/mismatch mode.
constexpr int kMaxM = 16;
struct GenScore {
    KswScore k;
    int32_t m, sc_N;
    int8_t mat[kMaxM * kMaxM];
};

// One pair of the general (anti-diagonal) kernel.  Its state block in the
// state arena holds the reference's per-target-position int8 arrays, the
// target copy and the reversed query (laid out so that the reference's
// over-reads see the same bytes), and H; see gen_layout().
struct GenJob {
    uint64_t q_off, t_off;     // byte offsets in the seq arena
    uint64_t st_off;           // byte offset of the state block
    uint64_t dir_off;          // byte offset of row 0 of the direction rows
    uint64_t off_off;          // int32 offset of off[] (off_end[] follows)
    uint64_t cig_off;          // uint32 offset in the cig arena
    int32_t qlen, tlen, w, zdrop, end_bonus, flag;
    int32_t ncolB;             // bytes per direction row: n_col_ * 64
    int32_t tlenB, qlenB;      // lengths rounded up to 64
    int32_t out;               // slot in the output array
    int32_t r_end;             // band kernel: anti-diagonals the reference completes
    uint64_t bnd_off;          // band kernel: dword offset of its two stripe buffers
};

// State block of one general pair: ten int8 arrays of tlenB (u, y, y2, s and
// two parity copies of x, v, x2), the target copy sf (zero-padded to tlenB),
// the reversed query qr followed by a zero tail, then int32 H[tlenB].
struct GenLayout {
    uint64_t u, y, y2, s, x0, x1, v0, v1, xx0, xx1, sf, qr, H, bytes;
};
KSW_HD inline GenLayout gen_layout(int tlenB, int qlenB)
{
    GenLayout L;
    const uint64_t T = (uint64_t)tlenB;
    L.u = 0, L.y = T, L.y2 = 2 * T, L.s = 3 * T, L.x0 = 4 * T, L.x1 = 5 * T, L.v0 = 6 * T, L.v1 = 7 * T;
    L.xx0 = 8 * T, L.xx1 = 9 * T, L.sf = 10 * T, L.qr = 11 * T;
    L.H = (L.qr + (uint64_t)qlenB + 128 + 255) / 256 * 256;
    L.bytes = (L.H + 4 * T + 255) / 256 * 256;
    return L;
}

// LDS rings of the general kernel: ten int8 arrays and int32 H per slot.
constexpr int kGenRingMax = 8192;
KSW_HD inline uint32_t gen_ring_bytes(int ring) { return (uint32_t)ring * 14u; }
// Ring slots for a pair whose band (plus the 64/16 rounding) spans at most
// `window` positions: a power of two at least window + 384, or 0 when that
// exceeds kGenRingMax (the state then stays in HBM).
KSW_HD inline int gen_ring_slots(int window)
{
    int r = 256;
    while (r < window + 384) r <<= 1;
    return r <= kGenRingMax ? r : 0;
}

// The reference's direction-row width for 64-byte blocks
// (minimap2/ksw2_extd2_avx.c:198-206): n_col_ = ceil(min(qlen, tlen, w+1)/64) + 1.
KSW_HD inline int gen_ncolB(int qlen, int tlen, int w)
{
    int n = qlen < tlen ? qlen : tlen;
    n = n < w + 1 ? n : w + 1;
    return ((n + 63) / 64 + 1) * 64;
}

struct KswOut {
    int32_t score, mqe, mqe_t, mte, mte_q;
    int32_t max, max_t, max_q;
    int32_t zdropped, stopped, reach_end, n_cigar;
};

// Boundary difference injected on the first row/column at anti-diagonal r
// (minimap2/ksw2_extd2_sse.c:167,171), as the int8 the reference stores.
KSW_HD inline int32_t edge_diff(const KswScore &sc, int r)
{
    int v = r == 0 ? -sc.G1 : r < sc.long_thres ? -sc.e : r == sc.long_thres ? sc.long_diff : -sc.e2;
    return (int32_t)(int8_t)(uint8_t)(v & 0xff);
}

// Sum of edge_diff(0..n): with G1 - qe_orig added this is H on the
// boundary row/column, H(-1, n) = H(n, -1).
KSW_HD inline int64_t edge_prefix(const KswScore &sc, int n)
{
    if (n < 0) return 0;
    int64_t s = edge_diff(sc, 0);
    if (n == 0) return s;
    // k in [1, n]: -e below long_thres, long_diff at it, -e2 above
    int lt = sc.long_thres;
    int64_t lo_end = (lt - 1 < n ? lt - 1 : n);          // last k with k < lt
    if (lo_end >= 1) s += (int64_t)(lo_end) * edge_diff(sc, 1);
    if (lt >= 1 && lt <= n) s += edge_diff(sc, lt);
    int64_t hi_start = (lt + 1 > 1 ? lt + 1 : 1);        // first k with k > lt
    if (hi_start <= n) s += (int64_t)(n - hi_start + 1) * edge_diff(sc, (int)hi_start);
    return s;
}

// Rows (target positions) handled by one lane of the stripe kernel: R = 8
// (the throughput form) or 4 (half the rows per pass, for pairs whose
// latency matters more than the device's throughput).  A stripe is 64 lanes
// x R rows, a pass is two stripes computed by one wave (the second in the
// high halves, kWave steps behind).
constexpr int kRowsPerLane = 8;
constexpr int kStripeRows = kWave * kRowsPerLane;

KSW_HD inline int stripe_count(int tlen, int R = kRowsPerLane) { return (tlen + kWave * R - 1) / (kWave * R); }
KSW_HD inline int pass_count(int tlen, int R = kRowsPerLane) { return (stripe_count(tlen, R) + 1) / 2; }
KSW_HD inline uint64_t pass_steps(int qlen) { return (uint64_t)qlen + 2 * kWave - 1; }

// Direction-byte address of cell (i = striped, j = swept) in the pass
// layout: stripe s = i / 64R is half h = s & 1 of pass p = s >> 1; lane k,
// row rho; the lane visits column j of half h at step c = j + k + 64h; each
// step the wave writes 64 lanes x 2R bytes, byte 4*(rho/2) + 2h + (rho&1)
// of a lane's 2R being row rho of half h.
KSW_HD inline uint64_t dir_index(int i, int j, int qlen, int R = kRowsPerLane)
{
    const int SR = kWave * R;
    const int s = i / SR, rem = i - s * SR, k = rem / R, rho = rem % R;
    const int p = s >> 1, h = s & 1;
    const uint64_t c = (uint64_t)j + k + (uint64_t)kWave * h;
    return (((uint64_t)p * pass_steps(qlen) + c) * kWave + (uint64_t)k) * (2 * R) + 4 * (rho >> 1) + 2 * h + (rho & 1);
}

KSW_HD inline uint64_t dir_bytes(int qlen, int tlen, int R = kRowsPerLane)
{
    return (uint64_t)pass_count(tlen, R) * pass_steps(qlen) * kWave * (2 * R);
}

// Singletrack track byte of cell (i, j): 4R bytes per lane per step, the p
// track (H(i,j) - H(i,j-1)) in the first 2R and the p2 track
// (H(i,j) - H(i-1,j)) in the next 2R, each in the direction-byte order;
// the stored byte is the negated value.  Twice dir_bytes() per pair.
KSW_HD inline uint64_t track_index(int track, int i, int j, int qlen, int R = kRowsPerLane)
{
    const uint64_t d = dir_index(i, j, qlen, R);
    return (d / (2 * R)) * (4 * R) + (uint64_t)track * (2 * R) + (d % (2 * R));
}

}  // namespace kswhip
