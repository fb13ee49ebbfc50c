/*
 * ksw2_oracle.c -- TEST INFRASTRUCTURE ONLY (see ksw2_oracle.h).
 *
 * A scalar restatement of the reference's two-piece-affine KSW2 aligner,
 * written cell by cell from the behaviour of
 *   minimap2/ksw2_extd2_sse.c:39-419     (SSE, 16-byte blocks)
 *   minimap2/ksw2_extd2_avx.c:51-789     (AVX-512, 64-byte blocks; the pinned reference)
 *   minimap2/ksw2.h:144-217              (CIGAR push, backtrack, reset, z-drop)
 *
 * Everything the reference computes on int8 lanes is computed here on int8
 * with two's-complement wrap.  Cells outside the biological band but inside
 * the vector blocks ("padding") are computed too, from the same stale memory
 * the vector code would see, because a banded backtrack can read them.
 * The state arrays therefore live in one buffer laid out like the
 * reference's kcalloc'd block (ksw2_extd2_sse.c:118-121,
 * ksw2_extd2_avx.c:226-233): u v x y x2 y2 s sf qr.
 */
#include <stdlib.h>
#include <string.h>
#include "ksw2_oracle.h"

/* ------------------------------------------------------------------ range probe */
static int g_range_on = 0;
static oracle_range_t g_range;

void oracle_range_reset(void)
{
    int k;
    for (k = 0; k < 16; ++k) g_range.lo[k] = 0x7fffffff, g_range.hi[k] = -0x7fffffff;
    g_range.wraps = 0;
    g_range.cap_binds = 0;
}
void oracle_range_enable(int on) { g_range_on = on; if (on) oracle_range_reset(); }
void oracle_range_get(oracle_range_t *out) { *out = g_range; }

static int32_t g_cell[16];
static inline void rng(int k, int32_t v, int *wrapped)
{
    g_cell[k] = v;
    if (v < -128 || v > 127) *wrapped = 1;
}
static inline void rng_commit(void)
{
    int k;
    for (k = 0; k < 16; ++k) {
        if (g_cell[k] < g_range.lo[k]) g_range.lo[k] = g_cell[k];
        if (g_cell[k] > g_range.hi[k]) g_range.hi[k] = g_cell[k];
    }
}

/* ------------------------------------------------------------------ helpers */
static inline int8_t w8(int v) { return (int8_t)(uint8_t)(v & 0xff); } /* int8 wrap */
static inline int8_t max8(int8_t a, int8_t b) { return a > b ? a : b; }
static inline int8_t min8(int8_t a, int8_t b) { return a < b ? a : b; }

/* ksw_push_cigar, minimap2/ksw2.h:144-154 (km == NULL: realloc). */
static uint32_t *push_cigar(int *n_cigar, int *m_cigar, uint32_t *cigar, uint32_t op, int len)
{
    if (*n_cigar == 0 || op != (cigar[*n_cigar - 1] & 0xf)) {
        if (*n_cigar == *m_cigar) {
            *m_cigar = *m_cigar ? (*m_cigar) << 1 : 4;
            cigar = (uint32_t *)realloc(cigar, (size_t)(*m_cigar) << 2);
        }
        cigar[(*n_cigar)++] = (uint32_t)len << 4 | op;
    } else
        cigar[*n_cigar - 1] += (uint32_t)len << 4;
    return cigar;
}

/* ksw_backtrack with is_rot = 1 and min_intron_len = 0, minimap2/ksw2.h:160-192.
 * p row r holds the cells [off[r], off_end[r]] of anti-diagonal r. */
static void backtrack(int is_rev, const uint8_t *p, const int *off, const int *off_end,
                      size_t n_col, int i0, int j0, int *m_cigar_, int *n_cigar_, uint32_t **cigar_)
{
    int n_cigar = 0, m_cigar = *m_cigar_, i = i0, j = j0, state = 0;
    uint32_t *cigar = *cigar_, tmp;
    while (i >= 0 && j >= 0) {
        int r = i + j, force_state = -1;
        if (i < off[r]) force_state = 2;
        if (off_end && i > off_end[r]) force_state = 1;
        tmp = force_state < 0 ? p[(size_t)r * n_col + (size_t)(i - off[r])] : 0;
        if (state == 0) state = tmp & 7;
        else if (!(tmp >> (state + 2) & 1)) state = 0;
        if (state == 0) state = tmp & 7;
        if (force_state >= 0) state = force_state;
        if (state == 0) cigar = push_cigar(&n_cigar, &m_cigar, cigar, 0, 1), --i, --j;
        else if (state == 1 || state == 3) cigar = push_cigar(&n_cigar, &m_cigar, cigar, 2, 1), --i;
        else cigar = push_cigar(&n_cigar, &m_cigar, cigar, 1, 1), --j;
    }
    if (i >= 0) cigar = push_cigar(&n_cigar, &m_cigar, cigar, 2, i + 1);
    if (j >= 0) cigar = push_cigar(&n_cigar, &m_cigar, cigar, 1, j + 1);
    if (!is_rev)
        for (i = 0; i < n_cigar >> 1; ++i)
            tmp = cigar[i], cigar[i] = cigar[n_cigar - 1 - i], cigar[n_cigar - 1 - i] = tmp;
    *m_cigar_ = m_cigar, *n_cigar_ = n_cigar, *cigar_ = cigar;
}

/* ksw_reset_extz, minimap2/ksw2.h:194-199 */
/* Singletrack traceback (minimap2/ksw2_singletrack.cpp:29-176): walks the two
 * stored difference tracks of the same anti-diagonal layout as the direction
 * bytes, p = H(i,j) - H(i,j-1) and p2 = H(i,j) - H(i-1,j), from (tlen-1,
 * qlen-1).  Cells outside [off, off_end] of their anti-diagonal read as
 * KSW_NEG_INF; row/column -1 read the gap-cost difference of the boundary.
 * A diagonal step is taken while p(i,j) + p2(i,j-1) equals the pair's
 * substitution score; otherwise a gap length l = 1, 2, ... is searched whose
 * accumulated difference equals -(q + l*e) or -(q2 + l*e2), deletion first.
 * Returns 0 when no gap length is found (the reference then frees the CIGAR
 * and reports zdropped). */
static int32_t st_gap(int32_t len, int32_t go, int32_t ge, int32_t go2, int32_t ge2)
{
    int64_t a, b;
    if (len <= 0) return 0;
    a = (int64_t)go + (int64_t)len * ge, b = (int64_t)go2 + (int64_t)len * ge2;
    return (int32_t)(a > b ? a : b);
}

static int32_t st_get(const int8_t *tr, int32_t i, int32_t j, size_t n_col, const int *off, const int *off_end,
                      int32_t go, int32_t ge, int32_t go2, int32_t ge2)
{
    int64_t d;
    if (i < 0 || j < 0) {
        int32_t c = i < 0 ? j : i;
        return st_gap(c + 1, go, ge, go2, ge2) - st_gap(c, go, ge, go2, ge2);
    }
    d = (int64_t)i + j;
    if (i < off[d] || i > off_end[d]) return KSW_NEG_INF;
    return tr[(size_t)d * n_col + (size_t)(i - off[d])];
}

static int st_fin(int32_t v) { return v > KSW_NEG_INF / 2 && v < -KSW_NEG_INF / 2; }

static int st_backtrace(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                        const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, size_t n_col,
                        const int *off, const int *off_end, const int8_t *p, const int8_t *p2, ksw_extz_t *ez)
{
    const int32_t go = -q, ge = -e, go2 = -q2, ge2 = -e2;
    int in_m = 1, n_cigar = 0, m_cigar = ez->m_cigar, k;
    uint32_t *cigar = ez->cigar;
    int32_t i = tlen - 1, j = qlen - 1, l = 0, dd = 0, di = 0;
    while (i >= 0 && j >= 0) {
        const int32_t diag = i + j;
        if (i < off[diag]) { cigar = push_cigar(&n_cigar, &m_cigar, cigar, 1, 1); --j; in_m = 1; continue; }
        if (i > off_end[diag]) { cigar = push_cigar(&n_cigar, &m_cigar, cigar, 2, 1); --i; in_m = 1; continue; }
        if (in_m) {
            int32_t h = st_get(p, i, j, n_col, off, off_end, go, ge, go2, ge2);
            int32_t v = st_get(p2, i, j - 1, n_col, off, off_end, go, ge, go2, ge2);
            int32_t cost = mat[target[i] * m + query[j]];
            if (st_fin(h) && st_fin(v) && (int64_t)h + v == cost) {
                cigar = push_cigar(&n_cigar, &m_cigar, cigar, 0, 1);
                --i, --j;
            } else {
                in_m = 0, l = 0;
                di = st_get(p, i, j, n_col, off, off_end, go, ge, go2, ge2);
                dd = st_get(p2, i, j, n_col, off, off_end, go, ge, go2, ge2);
            }
        } else {
            int32_t a1, a2, ji, id;
            ++l;
            if (l > i + 1 && l > j + 1) { free(cigar); ez->cigar = 0; ez->m_cigar = ez->n_cigar = 0; return 0; }
            a1 = go + l * ge, a2 = go2 + l * ge2, ji = j - l, id = i - l;
            if (id >= -1 && (dd == a1 || dd == a2)) {
                cigar = push_cigar(&n_cigar, &m_cigar, cigar, 2, l);
                i = id, in_m = 1;
            } else if (ji >= -1 && (di == a1 || di == a2)) {
                cigar = push_cigar(&n_cigar, &m_cigar, cigar, 1, l);
                j = ji, in_m = 1;
            } else {
                if (ji >= 0) {
                    int32_t hh = st_get(p, i, ji, n_col, off, off_end, go, ge, go2, ge2);
                    di = st_fin(di) && st_fin(hh) ? di + hh : KSW_NEG_INF;
                }
                if (id >= 0) {
                    int32_t vv = st_get(p2, id, j, n_col, off, off_end, go, ge, go2, ge2);
                    dd = st_fin(dd) && st_fin(vv) ? dd + vv : KSW_NEG_INF;
                }
            }
        }
    }
    if (i >= 0) cigar = push_cigar(&n_cigar, &m_cigar, cigar, 2, i + 1);
    if (j >= 0) cigar = push_cigar(&n_cigar, &m_cigar, cigar, 1, j + 1);
    for (k = 0; k < n_cigar >> 1; ++k) {
        uint32_t tmp = cigar[k];
        cigar[k] = cigar[n_cigar - 1 - k], cigar[n_cigar - 1 - k] = tmp;
    }
    ez->cigar = cigar, ez->m_cigar = m_cigar, ez->n_cigar = n_cigar;
    return 1;
}

static void reset_extz(ksw_extz_t *ez)
{
    ez->max_q = ez->max_t = ez->mqe_t = ez->mte_q = -1;
    ez->max = 0, ez->score = ez->mqe = ez->mte = KSW_NEG_INF;
    ez->n_cigar = 0, ez->zdropped = 0, ez->reach_end = 0, ez->stopped = 0;
}

/* ksw_apply_zdrop with is_rot = 1, minimap2/ksw2.h:201-217.  ez->max is a
 * 31-bit unsigned field and takes part in int arithmetic. */
static int apply_zdrop(ksw_extz_t *ez, int32_t H, int r, int t, int zdrop, int8_t e)
{
    if (H > (int32_t)ez->max) {
        ez->max = (uint32_t)H, ez->max_t = t, ez->max_q = r - t;
    } else if (t >= ez->max_t && r - t >= ez->max_q) {
        int tl = t - ez->max_t, ql = (r - t) - ez->max_q, l;
        l = tl > ql ? tl - ql : ql - tl;
        if (zdrop >= 0 && (int)ez->max - H > zdrop + l * e) {
            ez->zdropped = 1;
            return 1;
        }
    }
    return 0;
}

/* Gap boundary difference on the first row/column: value injected for
 * anti-diagonal r (ksw2_extd2_sse.c:167,171). */
static inline int8_t edge_diff(int r, int q, int e, int e2, int long_thres, int long_diff)
{
    return w8(r == 0 ? -q - e : r < long_thres ? -e : r == long_thres ? long_diff : -e2);
}

/* ------------------------------------------------------------------ main */
void oracle_ksw_extd2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                      int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                      int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez,
                      int isa, uint8_t *dir_ij, int stop_r)
{
    const int B = isa == ORACLE_ISA_AVX512 ? 64 : 16;
    /* qe is fixed before the piece swap (ksw2_extd2_sse.c:73 vs :83) and is
     * only used for H at r == 0. */
    const int qe_orig = q + e;
    const int with_cigar = !(flag & KSW_EZ_SCORE_ONLY), approx_max = !!(flag & KSW_EZ_APPROX_MAX);
    const int right = !!(flag & KSW_EZ_RIGHT), generic = !!(flag & KSW_EZ_GENERIC_SC);
    /* two-track traceback: the AVX2/AVX-512 kernels only (ksw2_extd2_avx.c:128);
     * the SSE kernel ignores the flag */
    const int singletrack = !!(flag & KSW_EZ_SINGLETRACK) && isa == ORACLE_ISA_AVX512;
    int r, t, tlen_, qlen_, n_col_, max_sc, min_sc, long_thres, long_diff, last_st, last_en;
    int32_t *H = 0, H0 = 0, last_H0_t = 0;
    int8_t qe8, qe28, sc_mch, sc_mis, sc_N, m1;
    uint8_t *mem, *p = 0, *p2 = 0;
    int8_t *u, *v, *x, *y, *x2, *y2, *s;
    uint8_t *sf, *qr;
    int *off = 0, *off_end = 0;

    reset_extz(ez);
    if (m <= 1 || qlen <= 0 || tlen <= 0) return;
    if (q2 + e2 < q + e) { int8_t tq = q; q = q2; q2 = tq; tq = e; e = e2; e2 = tq; }

    qe8 = w8(q + e), qe28 = w8(q2 + e2);
    sc_mch = mat[0], sc_mis = mat[1];
    sc_N = mat[m * m - 1] == 0 ? w8(-e2) : mat[m * m - 1];
    m1 = w8(m - 1);

    if (w < 0) w = tlen > qlen ? tlen : qlen;
    tlen_ = (tlen + B - 1) / B;
    qlen_ = (qlen + B - 1) / B;
    n_col_ = qlen < tlen ? qlen : tlen;
    n_col_ = ((n_col_ < w + 1 ? n_col_ : w + 1) + B - 1) / B + 1;
    for (t = 1, max_sc = mat[0], min_sc = mat[1]; t < m * m; ++t) {
        max_sc = max_sc > mat[t] ? max_sc : mat[t];
        min_sc = min_sc < mat[t] ? min_sc : mat[t];
    }
    (void)max_sc;
    if (-min_sc > 2 * (q + e)) return;

    long_thres = e != e2 ? (q2 - q) / (e - e2) - 1 : 0;
    if (q2 + e2 + long_thres * e2 > q + e + long_thres * e) ++long_thres;
    long_diff = long_thres * (e - e2) - (q2 - q) - e2;

    if (with_cigar) {
        p = (uint8_t *)malloc(((size_t)(qlen + tlen - 1) * n_col_ + 1) * B);
        if (singletrack) p2 = (uint8_t *)malloc(((size_t)(qlen + tlen - 1) * n_col_ + 1) * B);
        off = (int *)malloc((size_t)(qlen + tlen - 1) * sizeof(int) * 2);
        off_end = off + qlen + tlen - 1;
    }
    /* u v x y x2 y2 s | sf | qr, zero-filled like kcalloc; generous tail so
     * that over-reads past qr see zeros as in the reference allocation. */
    mem = (uint8_t *)calloc((size_t)tlen_ * 8 + qlen_ + 2 + 64, B);
    u = (int8_t *)mem;
    v = u + (size_t)tlen_ * B, x = v + (size_t)tlen_ * B, y = x + (size_t)tlen_ * B;
    x2 = y + (size_t)tlen_ * B, y2 = x2 + (size_t)tlen_ * B, s = y2 + (size_t)tlen_ * B;
    sf = (uint8_t *)(s + (size_t)tlen_ * B), qr = sf + (size_t)tlen_ * B;
    memset(u, (uint8_t)w8(-q - e), (size_t)tlen_ * B);
    memset(v, (uint8_t)w8(-q - e), (size_t)tlen_ * B);
    memset(x, (uint8_t)w8(-q - e), (size_t)tlen_ * B);
    memset(y, (uint8_t)w8(-q - e), (size_t)tlen_ * B);
    memset(x2, (uint8_t)w8(-q2 - e2), (size_t)tlen_ * B);
    memset(y2, (uint8_t)w8(-q2 - e2), (size_t)tlen_ * B);
    if (!approx_max) {
        H = (int32_t *)malloc((size_t)tlen_ * B * 4);
        for (t = 0; t < tlen_ * B; ++t) H[t] = KSW_NEG_INF;
    }
    for (t = 0; t < qlen; ++t) qr[t] = query[qlen - 1 - t];
    memcpy(sf, target, (size_t)tlen);

    for (r = 0, last_st = last_en = -1; r < qlen + tlen - 1; ++r) {
        int st = 0, en = tlen - 1, st0, en0, st16, en16, stB, wrap = 0;
        int8_t x1, x21, v1, wx = 0, wv = 0, wx2 = 0;
        const uint8_t *qrr = qr + (qlen - 1 - r);
        if (stop_r >= 0 && r == stop_r) { ez->stopped = 1; break; }
        /* band of anti-diagonal r (ksw2_extd2_sse.c:147-156) */
        if (st < r - qlen + 1) st = r - qlen + 1;
        if (en > r) en = r;
        if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
        if (en > (r + w) >> 1) en = (r + w) >> 1;
        if (st > en) { ez->zdropped = 1; break; }
        st0 = st, en0 = en;
        st16 = st0 / 16 * 16, en16 = (en0 + 16) / 16 * 16 - 1;
        stB = st0 / B * B;
        /* boundary injected at lane st16 (ksw2_extd2_sse.c:158-168; the AVX-512
         * kernel uses the 16-rounded start too, ksw2_extd2_avx.c:272-286) */
        if (st16 > 0) {
            if (st16 - 1 >= last_st && st16 - 1 <= last_en)
                x1 = x[st16 - 1], x21 = x2[st16 - 1], v1 = v[st16 - 1];
            else
                x1 = w8(-q - e), x21 = w8(-q2 - e2), v1 = w8(-q - e);
        } else {
            x1 = w8(-q - e), x21 = w8(-q2 - e2);
            v1 = edge_diff(r, q, e, e2, long_thres, long_diff);
        }
        if (en16 >= r) {
            y[r] = w8(-q - e), y2[r] = w8(-q2 - e2);
            u[r] = edge_diff(r, q, e, e2, long_thres, long_diff);
        }
        /* scores (ksw2_extd2_sse.c:174-193): the fast path writes whole
         * 16-byte groups from st0, the generic path exactly [st0, en0] */
        if (!generic) {
            int n16 = (en0 - st0 + 16) / 16 * 16;
            for (t = st0; t < st0 + n16; ++t) {
                uint8_t sq = sf[t], sk = qrr[t];
                int8_t sc = sq == sk ? sc_mch : sc_mis;
                if ((int8_t)sq == m1 || (int8_t)sk == m1) sc = sc_N;
                s[t] = sc;
            }
        } else {
            for (t = st0; t <= en0; ++t) s[t] = mat[sf[t] * m + qrr[t]];
        }
        /* AVX-512 lane 0 of the first block when the 16-rounded start is not
         * 64-aligned: its byte shuffle wraps within the 128-bit lane and the
         * cell reads position stB+15 of the previous anti-diagonal
         * (ksw2_extd2_avx.c:72-85, mask mskc_ar[(st0-t*64)/16]). */
        if (B == 64 && st16 != stB) {
            wrap = 1;
            wx = x[stB + 15], wv = v[stB + 15], wx2 = x2[stB + 15];
        }
        if (with_cigar) off[r] = stB, off_end[r] = en16;
        /* cells, descending so that t-1 still holds anti-diagonal r-1 */
        for (t = en16; t >= stB; --t) {
            int8_t xl, vl, x2l, ut = u[t], z, a, b, a2, b2, tmp, d = 0;
            int wrapped = 0;
            if (t == st16) xl = x1, vl = v1, x2l = x21;
            else if (wrap && t == stB) xl = wx, vl = wv, x2l = wx2;
            else xl = x[t - 1], vl = v[t - 1], x2l = x2[t - 1];
            z = s[t];
            a = w8(xl + vl), b = w8(y[t] + ut), a2 = w8(x2l + vl), b2 = w8(y2[t] + ut);
            if (g_range_on) {
                rng(0, xl + vl, &wrapped), rng(1, y[t] + ut, &wrapped);
                rng(2, x2l + vl, &wrapped), rng(3, y2[t] + ut, &wrapped);
            }
            if (!with_cigar) {
                z = max8(z, a), z = max8(z, b), z = max8(z, a2), z = max8(z, b2);
            } else if (!right) {
                d = a > z ? 1 : 0;  z = max8(z, a);
                d = b > z ? 2 : d;  z = max8(z, b);
                d = a2 > z ? 3 : d; z = max8(z, a2);
                d = b2 > z ? 4 : d; z = max8(z, b2);
            } else {
                d = z > a ? 0 : 1;  z = max8(z, a);
                d = z > b ? d : 2;  z = max8(z, b);
                d = z > a2 ? d : 3; z = max8(z, a2);
                d = z > b2 ? d : 4; z = max8(z, b2);
            }
            if (g_range_on && z > sc_mch && t >= st0 && t <= en0) ++g_range.cap_binds;
            z = min8(z, sc_mch);
            if (g_range_on) {
                rng(4, z - vl, &wrapped), rng(5, z - ut, &wrapped);
                rng(6, z - q, &wrapped), rng(7, z - q2, &wrapped);
            }
            u[t] = w8(z - vl);
            v[t] = w8(z - ut);
            tmp = w8(z - q);
            if (g_range_on) rng(8, a - tmp, &wrapped), rng(9, b - tmp, &wrapped);
            a = w8(a - tmp), b = w8(b - tmp);
            tmp = w8(z - q2);
            if (g_range_on) rng(10, a2 - tmp, &wrapped), rng(11, b2 - tmp, &wrapped);
            a2 = w8(a2 - tmp), b2 = w8(b2 - tmp);
            if (!right) {
                if (g_range_on) {
                    rng(12, (a > 0 ? a : 0) - qe8, &wrapped), rng(13, (b > 0 ? b : 0) - qe8, &wrapped);
                    rng(14, (a2 > 0 ? a2 : 0) - qe28, &wrapped), rng(15, (b2 > 0 ? b2 : 0) - qe28, &wrapped);
                }
                x[t] = w8((a > 0 ? a : 0) - qe8);    d |= a > 0 ? 0x08 : 0;
                y[t] = w8((b > 0 ? b : 0) - qe8);    d |= b > 0 ? 0x10 : 0;
                x2[t] = w8((a2 > 0 ? a2 : 0) - qe28); d |= a2 > 0 ? 0x20 : 0;
                y2[t] = w8((b2 > 0 ? b2 : 0) - qe28); d |= b2 > 0 ? 0x40 : 0;
            } else {
                if (g_range_on) {
                    rng(12, (a < 0 ? 0 : a) - qe8, &wrapped), rng(13, (b < 0 ? 0 : b) - qe8, &wrapped);
                    rng(14, (a2 < 0 ? 0 : a2) - qe28, &wrapped), rng(15, (b2 < 0 ? 0 : b2) - qe28, &wrapped);
                }
                x[t] = w8((a < 0 ? 0 : a) - qe8);    d |= a < 0 ? 0 : 0x08;
                y[t] = w8((b < 0 ? 0 : b) - qe8);    d |= b < 0 ? 0 : 0x10;
                x2[t] = w8((a2 < 0 ? 0 : a2) - qe28); d |= a2 < 0 ? 0 : 0x20;
                y2[t] = w8((b2 < 0 ? 0 : b2) - qe28); d |= b2 < 0 ? 0 : 0x40;
            }
            if (g_range_on && t >= st0 && t <= en0) {
                rng_commit();
                if (wrapped) ++g_range.wraps;
            }
            if (with_cigar && singletrack) {
                /* p = z - u(t) (ksw2_extd2_avx.c:432-436): the new v; p2 = the new u */
                p[(size_t)r * n_col_ * B + (size_t)(t - stB)] = (uint8_t)v[t];
                p2[(size_t)r * n_col_ * B + (size_t)(t - stB)] = (uint8_t)u[t];
                if (dir_ij && t >= st0 && t <= en0) {
                    dir_ij[(size_t)t * qlen + (size_t)(r - t)] = (uint8_t)v[t];
                    dir_ij[(size_t)tlen * qlen + (size_t)t * qlen + (size_t)(r - t)] = (uint8_t)u[t];
                }
            } else if (with_cigar) {
                p[(size_t)r * n_col_ * B + (size_t)(t - stB)] = (uint8_t)d;
                if (dir_ij && t >= st0 && t <= en0) dir_ij[(size_t)t * qlen + (size_t)(r - t)] = (uint8_t)d;
            }
        }
        if (!approx_max) {
            /* exact H and its maximum on the anti-diagonal
             * (ksw2_extd2_sse.c:332-375, ksw2_extd2_avx.c:639-724): four
             * interleaved lanes over [st0, st0+(en0-st0)/4*4), each keeping
             * its first strict maximum, folded lane 0..3 with strict '<',
             * then a scalar tail; H[en0] is the starting candidate. */
            int32_t max_H, max_t;
            if (r > 0) {
                int32_t HH[4], tt[4];
                int en1 = st0 + (en0 - st0) / 4 * 4, i;
                max_H = H[en0] = en0 > 0 ? H[en0 - 1] + u[en0] : H[en0] + v[en0];
                max_t = en0;
                for (i = 0; i < 4; ++i) HH[i] = max_H, tt[i] = max_t;
                for (t = st0; t < en1; t += 4)
                    for (i = 0; i < 4; ++i) {
                        H[t + i] += v[t + i];
                        if (H[t + i] > HH[i]) HH[i] = H[t + i], tt[i] = t;
                    }
                for (i = 0; i < 4; ++i)
                    if (max_H < HH[i]) max_H = HH[i], max_t = tt[i] + i;
                for (; t < en0; ++t) {
                    H[t] += v[t];
                    if (H[t] > max_H) max_H = H[t], max_t = t;
                }
            } else {
                H[0] = v[0] - qe_orig, max_H = H[0], max_t = 0;
            }
            if (en0 == tlen - 1 && H[en0] >= ez->mte) ez->mte = H[en0], ez->mte_q = r - en0;
            if (r - st0 == qlen - 1 && H[st0] >= ez->mqe) ez->mqe = H[st0], ez->mqe_t = st0;
            if (apply_zdrop(ez, max_H, r, max_t, zdrop, e2)) break;
            if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H[tlen - 1];
        } else {
            /* approximate max along one path (ksw2_extd2_sse.c:376-392) */
            if (r > 0) {
                if (last_H0_t >= st0 && last_H0_t <= en0 && last_H0_t + 1 >= st0 && last_H0_t + 1 <= en0) {
                    int32_t d0 = v[last_H0_t], d1 = u[last_H0_t + 1];
                    if (d0 > d1) H0 += d0;
                    else H0 += d1, ++last_H0_t;
                } else if (last_H0_t >= st0 && last_H0_t <= en0) {
                    H0 += v[last_H0_t];
                } else {
                    ++last_H0_t, H0 += u[last_H0_t];
                }
            } else {
                H0 = v[0] - qe_orig, last_H0_t = 0;
            }
            if ((flag & KSW_EZ_APPROX_DROP) && apply_zdrop(ez, H0, r, last_H0_t, zdrop, e2)) break;
            if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H0;
        }
        last_st = st16, last_en = en16;
    }
    free(mem);
    free(H);
    if (with_cigar && singletrack) { /* ksw2_extd2_avx.c:752-762 */
        if (!ez->stopped && !ez->zdropped && !(flag & (KSW_EZ_EXTZ_ONLY | KSW_EZ_SEMIGLOBAL_END))) {
            st_backtrace(qlen, query, tlen, target, m, mat, q, e, q2, e2, (size_t)n_col_ * B, off, off_end,
                         (const int8_t *)p, (const int8_t *)p2, ez);
            if (ez->n_cigar == 0) ez->zdropped = 1;
        } else if (!ez->stopped && !ez->zdropped) {
            ez->zdropped = 1;
        }
        free(p);
        free(p2);
        free(off);
    } else if (with_cigar) { /* ksw2_extd2_sse.c:398-418 */
        int rev = !!(flag & KSW_EZ_REV_CIGAR);
        size_t nc = (size_t)n_col_ * B;
        if (!ez->stopped && !ez->zdropped && (flag & KSW_EZ_SEMIGLOBAL_END)) {
            if (ez->mqe > ez->mte) {
                ez->score = ez->mqe;
                backtrack(rev, p, off, off_end, nc, ez->mqe_t, qlen - 1, &ez->m_cigar, &ez->n_cigar, &ez->cigar);
            } else {
                ez->score = ez->mte;
                backtrack(rev, p, off, off_end, nc, tlen - 1, ez->mte_q, &ez->m_cigar, &ez->n_cigar, &ez->cigar);
            }
        } else if (!ez->stopped && !ez->zdropped && !(flag & KSW_EZ_EXTZ_ONLY)) {
            backtrack(rev, p, off, off_end, nc, tlen - 1, qlen - 1, &ez->m_cigar, &ez->n_cigar, &ez->cigar);
        } else if (!ez->zdropped && (flag & KSW_EZ_EXTZ_ONLY) && ez->mqe + end_bonus > (int)ez->max) {
            ez->reach_end = 1;
            backtrack(rev, p, off, off_end, nc, ez->mqe_t, qlen - 1, &ez->m_cigar, &ez->n_cigar, &ez->cigar);
        } else if (ez->max_t >= 0 && ez->max_q >= 0) {
            backtrack(rev, p, off, off_end, nc, ez->max_t, ez->max_q, &ez->m_cigar, &ez->n_cigar, &ez->cigar);
        }
        free(p);
        free(off);
    }
}

/* ------------------------------------------------------------------ checks */
static int eff_score(int tc, int qc, int8_t m, const int8_t *mat, int generic, int8_t e2)
{
    int sc_N = mat[m * m - 1] == 0 ? -e2 : mat[m * m - 1];
    if (generic) return mat[tc * m + qc];
    if (tc == m - 1 || qc == m - 1) return sc_N;
    return tc == qc ? mat[0] : mat[1];
}

int oracle_cigar_score(int n_cigar, const uint32_t *cigar, int qlen, const uint8_t *query,
                       int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                       int8_t q, int8_t e, int8_t q2, int8_t e2)
{
    int k, i = 0, j = 0, score = 0;
    (void)qlen, (void)tlen;
    for (k = 0; k < n_cigar; ++k) {
        int op = cigar[k] & 0xf, len = (int)(cigar[k] >> 4), l;
        if (op == 0) {
            for (l = 0; l < len; ++l, ++i, ++j) score += eff_score(target[i], query[j], m, mat, 1, e2);
        } else {
            int g1 = q + len * e, g2 = q2 + len * e2;
            score -= g1 < g2 ? g1 : g2;
            if (op == 1) j += len; else i += len;
        }
    }
    return score;
}

int oracle_gotoh2_global(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                         int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2)
{
    /* H(i,j): best score of target[0..i) vs query[0..j); E/F: ending in a
     * deletion/insertion of piece 1, E2/F2 of piece 2.  O(tlen*qlen) ints. */
    const int NEG = -0x3fffffff / 2;
    int i, j, res;
    size_t W = (size_t)qlen + 1;
    int *H = (int *)malloc(sizeof(int) * W * 2), *E = (int *)malloc(sizeof(int) * W * 2);
    int *E2 = (int *)malloc(sizeof(int) * W * 2);
    int *Hp = H, *Hc = H + W, *Ep = E, *Ec = E + W, *E2p = E2, *E2c = E2 + W;
    for (j = 0; j <= qlen; ++j) {
        int g1 = q + j * e, g2 = q2 + j * e2;
        Hp[j] = j == 0 ? 0 : -(g1 < g2 ? g1 : g2);
        Ep[j] = E2p[j] = NEG;
    }
    for (i = 1; i <= tlen; ++i) {
        int F = NEG, F2 = NEG, *sw;
        int g1 = q + i * e, g2 = q2 + i * e2;
        Hc[0] = -(g1 < g2 ? g1 : g2);
        Ec[0] = E2c[0] = NEG;
        for (j = 1; j <= qlen; ++j) {
            int h, ee, ee2;
            /* gap along the target (deletion): from row i-1, same column */
            ee = Ep[j] - e > Hp[j] - q - e ? Ep[j] - e : Hp[j] - q - e;
            ee2 = E2p[j] - e2 > Hp[j] - q2 - e2 ? E2p[j] - e2 : Hp[j] - q2 - e2;
            /* gap along the query (insertion): same row, column j-1 */
            F = F - e > Hc[j - 1] - q - e ? F - e : Hc[j - 1] - q - e;
            F2 = F2 - e2 > Hc[j - 1] - q2 - e2 ? F2 - e2 : Hc[j - 1] - q2 - e2;
            h = Hp[j - 1] + eff_score(target[i - 1], query[j - 1], m, mat, 1, e2);
            if (ee > h) h = ee;
            if (ee2 > h) h = ee2;
            if (F > h) h = F;
            if (F2 > h) h = F2;
            Hc[j] = h, Ec[j] = ee, E2c[j] = ee2;
        }
        sw = Hp, Hp = Hc, Hc = sw;
        sw = Ep, Ep = Ec, Ec = sw;
        sw = E2p, E2p = E2c, E2c = sw;
    }
    res = Hp[qlen];
    free(H), free(E), free(E2);
    return res;
}
