/*
 * ksw2_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar CPU restatement of the reference ksw_extd2 (minimap2/ksw2_extd2_sse.c
 * and minimap2/ksw2_extd2_avx.c).  It is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links or calls it.
 */
#ifndef KSW2_ORACLE_H
#define KSW2_ORACLE_H
#include <stdint.h>
#include "../include/ksw2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Vector width whose block semantics are emulated for the band edges:
 * 16 = ksw_extd2_sse, 64 = ksw_extd2_avx512 (the pinned reference). */
#define ORACLE_ISA_SSE    16
#define ORACLE_ISA_AVX512 64

/* Same arguments and results as ksw_extd2_* (minimap2/ksw2.h:91-107), with
 * km == NULL semantics (cigar grown with realloc).
 *   isa      ORACLE_ISA_SSE or ORACLE_ISA_AVX512
 *   dir_ij   optional (may be NULL): tlen*qlen bytes, receives the direction
 *            byte of every in-band cell (i = target, j = query) at
 *            dir_ij[i*qlen + j]; other bytes are left untouched.  With
 *            KSW_EZ_SINGLETRACK (ISA 64) it holds 2*tlen*qlen bytes instead:
 *            the two tracks, p = H(i,j)-H(i,j-1) at [i*qlen + j] and
 *            p2 = H(i,j)-H(i-1,j) at [tlen*qlen + i*qlen + j] (int8)
 *   stop_r   anti-diagonal at which a progress callback would return 0
 *            (-1: never), emulating minimap2/ksw2_progress.c:30-38 */
void oracle_ksw_extd2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                      int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                      int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez,
                      int isa, uint8_t *dir_ij, int stop_r);

/* Range instrumentation: when enabled, every int8 intermediate of the cell
 * recurrence is also computed in int32 and the extremes over in-band cells
 * are recorded, so a test can show where (if anywhere) the reference's int8
 * lanes wrap and whether the match-score cap ever changes z. */
typedef struct {
    int32_t lo[16], hi[16];   /* per intermediate, see ksw2_oracle.c */
    int64_t wraps;            /* cells where some intermediate left [-128,127] */
    int64_t cap_binds;        /* cells where z exceeded mat[0] before z = min(z, mat[0]) */
} oracle_range_t;
void oracle_range_enable(int on);
void oracle_range_get(oracle_range_t *out);
void oracle_range_reset(void);

/* Reference-free consistency helpers (used by the pinning tests). */
/* Score of a CIGAR under two-piece affine gaps, restating
 * scoreGappedTwoPieceAffineAlignment (alignSlidingWindow.cpp:677-721) on codes. */
int oracle_cigar_score(int n_cigar, const uint32_t *cigar, int qlen, const uint8_t *query,
                       int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
                       int8_t q, int8_t e, int8_t q2, int8_t e2);
/* Textbook full-matrix Gotoh DP with two gap pieces (global, score only),
 * gap of length l costs min(q+l*e, q2+l*e2).  Independent of the
 * difference recurrence. */
int oracle_gotoh2_global(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                         int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2);

#ifdef __cplusplus
}
#endif
#endif
