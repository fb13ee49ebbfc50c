// ksw2_stripe.hip -- gfx950 kernels for unbanded ksw_extd2 (the fast path).
//
// Recurrence: the two-piece-affine difference recurrence of
// minimap2/ksw2_extd2_sse.c:43-71 and :237-331 (left/right gap alignment,
// score-only), evaluated for every cell (i = target, j = query).  This file
// evaluates it in a different order and representation than the reference:
//
//  * Stripes.  The target is cut into stripes of 64*R rows.  Lane k of a
//    wavefront owns R consecutive rows and sweeps the query; at step c it is
//    at column j = c - k, so the cell above (i-1, j) was produced by lane k-1
//    one step earlier and arrives through a DPP wave_shr:1.  Lane 0 takes the
//    bottom row of the previous stripe from a per-pair column buffer; the
//    left neighbour (i, j-1) stays in the lane's registers.
//  * Teams.  One workgroup of W waves computes one pair; wave w takes
//    stripes w, w+W, w+2W, ...  Stripe s trails stripe s-1 by at least two
//    64-column blocks: the producer publishes its bottom-row columns to the
//    pair's column buffer and then a column count in LDS (workgroup-scope
//    release), the consumer polls that count (workgroup-scope acquire) before
//    it reads.  All waves of a team share one CU, so no cache maintenance
//    beyond the CU is involved.  W = 1 (no waiting at all) for pairs that
//    are small against the batch; long pairs get up to 16 waves.
//  * Scaled, negated state.  Every difference value w is kept as -8*w plus a
//    3-bit tag, so one v_min3 pair yields both the maximum of the five
//    candidates (s, a, b, a2, b2) and the index the reference's strict '>'
//    (or '>=' for right alignment) chain would pick: ties resolve to the
//    smallest (left) or largest (right) index through the tag.
//  * Exactness.  The reference computes on int8 lanes.  The host only routes
//    a call here when every in-band intermediate provably stays inside int8
//    (then int32 arithmetic gives the same values), the match score is 0 and
//    no mismatch scores above it (then z = min(z, mat[0]) never binds and H
//    never exceeds 0, so ez->max stays 0), there is no band and no z-drop.
//
// Outputs per pair: the direction byte of every cell (bits as in
// minimap2/ksw2.h:156-159), and the H bookkeeping of
// minimap2/ksw2_extd2_sse.c:369-375 (score, mqe/mqe_t, mte/mte_q).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ksw2_common.h"
#include "ksw2_kernels.h"

namespace kswhip {

namespace {

constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

__device__ __forceinline__ int dpp_shr1(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, kDppWaveShr1, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_shl1(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, kDppWaveShl1, 0xF, 0xF, false);
}
__device__ __forceinline__ int min3i(int a, int b, int c) { return min(min(a, b), c); }
__device__ __forceinline__ int med3i(int a, int lo, int hi) { return max(min(a, hi), lo); }

// Tags: state k = 0 (diagonal s), 1 (a), 2 (b), 3 (a2), 4 (b2).  Under
// v_min the smallest tag wins a tie, so left alignment (first strict maximum)
// tags k, right alignment (last maximum, '>=' chain) tags 7-k.
template <bool RIGHT> struct Tags {
    static constexpr int S = RIGHT ? 7 : 0, A = RIGHT ? 6 : 1, B = RIGHT ? 5 : 2;
    static constexpr int A2 = RIGHT ? 4 : 3, B2 = RIGHT ? 3 : 4;
};

// Per-kernel constants of the scaled representation:
//   stored  U = -8u, V = -8v, Xh = -8(x+G1), X2h = -8(x2+G2), Yh, Y2h likewise
//   cands   A = Xh+V+CA = -8a+tagA, ... , S = -8s+tagS
//   after   A' = A+nZ+KA = -8(a-z+q), Xh' = min(A',0)
struct Consts {
    int CA, CB, CA2, CB2, KA, KB, KA2, KB2, SM, SX;
};

template <bool RIGHT>
__device__ __forceinline__ Consts make_consts(const KswScore &sc)
{
    using T = Tags<RIGHT>;
    Consts k;
    k.CA = 8 * sc.G1 + T::A;
    k.CB = 8 * sc.G1 + T::B;
    k.CA2 = 8 * sc.G2 + T::A2;
    k.CB2 = 8 * sc.G2 + T::B2;
    k.KA = -(8 * sc.q + T::A);
    k.KB = -(8 * sc.q + T::B);
    k.KA2 = -(8 * sc.q2 + T::A2);
    k.KB2 = -(8 * sc.q2 + T::B2);
    k.SM = -8 * sc.mch + T::S;
    k.SX = -8 * sc.mis + T::S;
    return k;
}

// One cell.  In: Xa/X2a/Va from the cell above (replaced by this cell's
// outputs for the cell below), U/Yh/Y2h from the cell to the left (replaced
// by this cell's outputs for the cell to the right).  Returns the direction
// byte (0 when SCORE_ONLY).
template <bool RIGHT, bool SCORE_ONLY>
__device__ __forceinline__ int cell(const Consts &k, int tb, int qb, int &Xa, int &X2a, int &Va,
                                    int &U, int &Yh, int &Y2h)
{
    const int S = tb == qb ? k.SM : k.SX;
    const int A = Xa + Va + k.CA;
    const int B = Yh + U + k.CB;
    const int A2 = X2a + Va + k.CA2;
    const int B2 = Y2h + U + k.CB2;
    const int Z = min3i(min3i(S, A, B), A2, B2);
    const int Z8 = Z & ~7;
    const int nZ = -Z8;
    const int Un = Z8 - Va;
    const int Vn = Z8 - U;
    const int Ap = A + nZ + k.KA;
    const int Bp = B + nZ + k.KB;
    const int A2p = A2 + nZ + k.KA2;
    const int B2p = B2 + nZ + k.KB2;
    const int Xn = min(Ap, 0), Yn = min(Bp, 0), X2n = min(A2p, 0), Y2n = min(B2p, 0);
    int d = 0;
    if (!SCORE_ONLY) {
        if (!RIGHT) {
            // continuation flags: a' > 0  <=>  A' < 0  <=>  max(Xh', -8) == -8
            d = Z & 7;
            d |= max(Xn, -8) & 0x08;
            d |= max(Yn, -8) & 0x10;
            d |= max(X2n, -8) & 0x20;
            d |= max(Y2n, -8) & 0x40;
        } else {
            // a' >= 0  <=>  A' <= 0: med3(A',0,8) is 0 when set, 8 when clear
            int t = med3i(Ap, 0, 8) | (med3i(Bp, 0, 8) << 1) | (med3i(A2p, 0, 8) << 2) |
                    (med3i(B2p, 0, 8) << 3);
            d = ((Z & 7) | t) ^ 0x7F;
        }
    }
    Xa = Xn, X2a = X2n, Va = Vn;
    U = Un, Yh = Yn, Y2h = Y2n;
    return d;
}

template <int R>
__device__ __forceinline__ void store_dirs(uint8_t *p, const int (&d)[R])
{
    uint32_t w[R / 4];
#pragma unroll
    for (int g = 0; g < R / 4; ++g)
        w[g] = (uint32_t)d[4 * g] | ((uint32_t)d[4 * g + 1] << 8) | ((uint32_t)d[4 * g + 2] << 16) |
               ((uint32_t)d[4 * g + 3] << 24);
    if constexpr (R == 4) {
        *(uint32_t *)p = w[0];
    } else if constexpr (R == 8) {
        *(uint2 *)p = make_uint2(w[0], w[1]);
    } else {
        *(uint4 *)p = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Team progress: wave w publishes, for its k-th stripe, k*(qlen+1) + the
// number of bottom-row columns written so far (monotonic per wave).
__device__ __forceinline__ void publish(int32_t *slot, int value)
{
    // every lane stores the same value (no divergent branch); the release
    // orders the preceding column-buffer stores before the count
    __hip_atomic_store(slot, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Spins until the producer's published count reaches `need`.  The exit test
// uses lane 0's value, so the loop is uniform.  A producer always ends its
// stripe at k*(qlen+1) + qlen, so `need` <= that bound terminates.
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

// One stripe of one pair, computed by the calling wave.  Roles (uniform
// flags): the first stripe (s == 0) takes row -1 as its upper boundary
// instead of the column buffer; the last stripe owns row tlen-1 (score / mte)
// and publishes nothing; every other stripe publishes its bottom row.
template <int R, bool RIGHT, bool SCORE_ONLY>
__device__ __forceinline__ void stripe_body(const StripeArgs &a, const Consts &kc, int job_id, int s,
                                            const KswJob &jb, int qlen, int tlen, int32_t *prog, int W)
{
    constexpr int SR_ = kWave * R;
    const bool first = s == 0;
    const bool last = s == (tlen + SR_ - 1) / SR_ - 1;
    const KswScore &sc = a.sc;
    const int lane = threadIdx.x % kWave;
    const uint8_t *qs = a.seq + jb.q_off;
    const uint8_t *ts = a.seq + jb.t_off;
    uint8_t *dir = a.dir + jb.dir_off;
    int32_t *bnd = a.bnd + jb.bnd_off;
    constexpr int SR = kWave * R;
    const int nsteps = qlen + kWave - 1;
    const int rho_last = last ? (tlen - 1) % R : -1;
    // progress slots: stripe s-1 is wave (s-1) % W's ((s-1)/W)-th stripe
    const int32_t *prog_prev = prog + (s + W - 1) % W;
    const int base_prev = first ? 0 : (s - 1) / W * (qlen + 1);
    int32_t *prog_mine = prog + s % W;
    const int base_mine = s / W * (qlen + 1);

    int U[R], Yh[R], Y2h[R], TB[R];
    const int i0 = s * SR + lane * R;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + r;
        TB[r] = i < tlen ? (int)ts[i] : 0;
        U[r] = -8 * edge_diff(sc, i);
        Yh[r] = 0;
        Y2h[r] = 0;
    }
    int row_acc = 0, mte_best = 0x7fffffff, mte_j = -1;  // row tlen-1 in the -8 domain
    int Xo = 0, X2o = 0, Vo = 0;        // this lane's bottom-row outputs
    int bin_xx = 0, bin_v = 0;          // lane-0 boundary queue (rotating)
    int bnx_xx = 0, bnx_v = 0;          // next block of the queue
    int out_xx = 0, out_v = 0;          // lane-63 output queue (rotating)
    int known = 0;                      // columns of stripe s-1 known published
    if (!first) {
        known = wait_count(prog_prev, base_prev + min(2 * kWave, qlen), known);
        const int c0 = lane, c1 = kWave + lane;
        bin_xx = c0 < qlen ? bnd[2 * c0] : 0;
        bin_v = c0 < qlen ? bnd[2 * c0 + 1] : 0;
        bnx_xx = c1 < qlen ? bnd[2 * c1] : 0;
        bnx_v = c1 < qlen ? bnd[2 * c1 + 1] : 0;
    }
    int qb_next = lane == 0 ? (int)qs[0] : 0;
    for (int c = 0; c < nsteps; ++c) {
        const int j = c - lane;
        // value for lane 0: the cell above the stripe in column c
        int bx, bx2, bv;
        if (first) {
            bx = 0, bx2 = 0, bv = -8 * edge_diff(sc, c);
        } else {
            bx = (int)(int16_t)(bin_xx & 0xffff);
            bx2 = bin_xx >> 16;
            bv = bin_v;
        }
        int Xa = dpp_shr1(bx, Xo), X2a = dpp_shr1(bx2, X2o), Va = dpp_shr1(bv, Vo);
        if (!first) {
            // rotate the queue so lane 0 sees column c+1 next step
            if ((c & (kWave - 1)) == kWave - 1) {
                bin_xx = bnx_xx, bin_v = bnx_v;
                known = wait_count(prog_prev, base_prev + min(c + 1 + 2 * kWave, qlen), known);
                const int cn = c + 1 + kWave + lane;
                bnx_xx = cn < qlen ? bnd[2 * cn] : 0;
                bnx_v = cn < qlen ? bnd[2 * cn + 1] : 0;
            } else {
                bin_xx = dpp_shl1(0, bin_xx);
                bin_v = dpp_shl1(0, bin_v);
            }
        }
        const int qb = qb_next;
        qb_next = (j + 1 >= 0 && j + 1 < qlen) ? (int)qs[j + 1] : 0;
        if (j >= 0 && j < qlen) {
            int d[R];
            int vl = 0;  // V of the cell in row tlen-1 (last stripe)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                d[r] = cell<RIGHT, SCORE_ONLY>(kc, TB[r], qb, Xa, X2a, Va, U[r], Yh[r], Y2h[r]);
                vl = r == rho_last ? Va : vl;
            }
            if (last) {
                row_acc += vl;
                if (row_acc <= mte_best) mte_best = row_acc, mte_j = j;
            }
            Xo = Xa, X2o = X2a, Vo = Va;
            if (!SCORE_ONLY)
                store_dirs<R>(dir + (((uint64_t)s * nsteps + c) * kWave + lane) * R, d);
        }
        if (!last) {
            // lane 63 appends its bottom-row output (column c-63) to the
            // queue; every 64 steps the queue is written out and published
            const int packed = (Xo & 0xffff) | (X2o << 16);
            out_xx = dpp_shl1(packed, out_xx);
            out_v = dpp_shl1(Vo, out_v);
            if ((c & (kWave - 1)) == kWave - 1 || c == nsteps - 1) {
                const int col = c - 2 * (kWave - 1) + lane;
                if (col >= 0 && col < qlen) {
                    bnd[2 * col] = out_xx;
                    bnd[2 * col + 1] = out_v;
                }
                publish(prog_mine, base_mine + min(c - (kWave - 1) + 1, qlen));
            }
        }
    }
    // mqe partials: every lane holds U = -8u(i, qlen-1) of its rows (frozen
    // after its last active step); H(i,qlen-1) = H(-1,qlen-1) + sum_{i'<=i} u.
    int lsum = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) lsum += U[r];
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
        run += U[r];
        const int i = i0 + r;
        if (i < tlen && run <= best) best = run, best_i = i;
    }
    // wave reduction: smallest value (largest H), ties to the larger row
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
        a.recs[jb.stripe_off + s] = rec;
    }
    if (last) {
        // publish score / mte from the lane that owns row tlen-1
        const int lane_last = ((tlen - 1) % SR) / R;
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

// One workgroup (a team of blockDim.x / 64 waves) per pair of this launch's
// job list; wave w computes stripes w, w+W, ...
template <int R, bool RIGHT, bool SCORE_ONLY>
__global__ void __launch_bounds__(kWave * kMaxTeam) ksw_stripe_kernel(StripeArgs a)
{
    __shared__ int32_t prog[kMaxTeam];
    const int W = __builtin_amdgcn_readfirstlane(blockDim.x / kWave);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int job = __builtin_amdgcn_readfirstlane(a.order[blockIdx.x]);
    prog[w] = 0;  // every lane of wave w writes the same zero
    __syncthreads();
    const Consts kc = make_consts<RIGHT>(a.sc);
    const KswJob jb = a.jobs[job];
    const int qlen = __builtin_amdgcn_readfirstlane(jb.qlen), tlen = __builtin_amdgcn_readfirstlane(jb.tlen);
    constexpr int SR = kWave * R;
    const int nstripes = (tlen + SR - 1) / SR;
    for (int s = w; s < nstripes; s += W) stripe_body<R, RIGHT, SCORE_ONLY>(a, kc, job, s, jb, qlen, tlen, prog, W);
}

// Device backtrack: one thread per pair.  First folds the per-stripe mqe
// partials in row order (ties to the later row, the '>=' of
// minimap2/ksw2_extd2_sse.c:371), then walks the direction bytes from the
// start cell chosen as in minimap2/ksw2_extd2_sse.c:398-418, running the
// state machine of ksw_backtrack (minimap2/ksw2.h:160-192) for an unbanded
// matrix (no cell of the path can leave the band there, so the off/off_end
// forcing never triggers).  CIGAR ops are pushed run-length encoded and
// reversed at the end unless KSW_EZ_REV_CIGAR.
__global__ void __launch_bounds__(64) ksw_backtrack_kernel(BacktrackArgs a)
{
    const int job_id = blockIdx.x * blockDim.x + threadIdx.x;
    if (job_id >= a.njobs) return;
    const KswJob jb = a.jobs[job_id];
    KswOut o = a.out[job_id];
    const int flag = jb.flag;
    const int qlen = jb.qlen, tlen = jb.tlen;
    {
        const int SR = kWave * jb.R;
        const int nstripes = (tlen + SR - 1) / SR;
        int carry = 0, best = 0x7fffffff, best_i = -1;
        for (int s = 0; s < nstripes; ++s) {
            const StripeRec rec = a.recs[jb.stripe_off + s];
            if (rec.best_i >= 0 && carry + rec.best <= best) best = carry + rec.best, best_i = rec.best_i;
            carry += rec.total;
        }
        const int64_t h_anchor = (int64_t)a.sc.G1 - a.sc.qe_orig;
        const int h_top_last_col = (int)(h_anchor + edge_prefix(a.sc, qlen - 1));  // H(-1, qlen-1)
        o.mqe = h_top_last_col - (best >> 3);
        o.mqe_t = best_i;
    }
    int i0 = -1, j0 = -1;
    if (!(flag & 0x01)) {  // KSW_EZ_SCORE_ONLY: no path
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
    int n = 0;
    if (i0 >= 0 || j0 >= 0) {
        const uint8_t *dir = a.dir + jb.dir_off;
        uint32_t *cig = a.cig + jb.cig_off;
        int i = i0, j = j0, state = 0;
        uint32_t last = 0xffffffffu;
        auto push = [&](uint32_t op, int len) {
            if (n > 0 && op == last) cig[n - 1] += (uint32_t)len << 4;
            else cig[n++] = (uint32_t)len << 4 | op, last = op;
        };
        while (i >= 0 && j >= 0) {
            const int tmp = dir[dir_index(i, j, qlen, jb.R)];
            if (state == 0) state = tmp & 7;
            else if (!(tmp >> (state + 2) & 1)) state = 0;
            if (state == 0) state = tmp & 7;
            if (state == 0) push(0, 1), --i, --j;
            else if (state == 1 || state == 3) push(2, 1), --i;
            else push(1, 1), --j;
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
    a.out[job_id] = o;
}

template <int R, bool RIGHT, bool SCORE_ONLY>
static hipError_t launch_stripe_t(const StripeArgs &a, int team, hipStream_t st)
{
    hipLaunchKernelGGL((ksw_stripe_kernel<R, RIGHT, SCORE_ONLY>), dim3(a.njobs), dim3(kWave * team), 0, st, a);
    return hipGetLastError();
}

template <int R>
static hipError_t launch_stripe_r(const StripeArgs &a, bool right, bool score_only, int team, hipStream_t st)
{
    if (score_only) return launch_stripe_t<R, false, true>(a, team, st);
    if (right) return launch_stripe_t<R, true, false>(a, team, st);
    return launch_stripe_t<R, false, false>(a, team, st);
}

hipError_t launch_stripe(const StripeArgs &a, int R, bool right, bool score_only, int team, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    if (team < 1 || team > kMaxTeam) return hipErrorInvalidValue;
    switch (R) {
    case 4: return launch_stripe_r<4>(a, right, score_only, team, st);
    case 8: return launch_stripe_r<8>(a, right, score_only, team, st);
    case 16: return launch_stripe_r<16>(a, right, score_only, team, st);
    default: return hipErrorInvalidValue;
    }
}

template <int R, bool RIGHT, bool SCORE_ONLY>
static int resident_t()
{
    int per_cu = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ksw_stripe_kernel<R, RIGHT, SCORE_ONLY>, kWave, 0) !=
            hipSuccess ||
        per_cu <= 0)
        return prop.multiProcessorCount * 4;
    return per_cu * prop.multiProcessorCount;
}

template <int R>
static int resident_r(bool right, bool score_only)
{
    if (score_only) return resident_t<R, false, true>();
    if (right) return resident_t<R, true, false>();
    return resident_t<R, false, false>();
}

int stripe_resident_waves(int R, bool right, bool score_only)
{
    switch (R) {
    case 4: return resident_r<4>(right, score_only);
    case 8: return resident_r<8>(right, score_only);
    case 16: return resident_r<16>(right, score_only);
    default: return 1024;
    }
}

hipError_t launch_backtrack(const BacktrackArgs &a, hipStream_t st)
{
    if (a.njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(ksw_backtrack_kernel, dim3((a.njobs + 63) / 64), dim3(64), 0, st, a);
    return hipGetLastError();
}

}  // namespace kswhip
