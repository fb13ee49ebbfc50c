/* The drop-in under AnchorWave's arena contract: ksw_extd2_hip called with a
 * kalloc arena (km != NULL) grows ez->cigar with the arena's krealloc, as the
 * reference's ksw_push_cigar does (minimap2/ksw2.h:144-154), and the caller
 * releases it with kfree (alignSlidingWindow.cpp:223-239, 328, 427).
 *
 * Links the reference's kalloc and ksw_extd2_avx512 (oracle/_ref, built from
 * the reference sources) and libksw2hip.so.  For each pair and flag set, one
 * arena per task as ScopedKswScratchArena makes it: both implementations run
 * into their own zeroed ksw_extz_t from the same arena, every field
 * (m_cigar included) and the CIGAR must agree, then a second call reuses each
 * ez (its CIGAR buffer grown from the first call's size), then kfree and
 * km_destroy.  Prints one line per failure and a summary; exits 1 on any
 * difference. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ksw2_hip.h"

/* the reference's kalloc and kernels (minimap2/kalloc.h, minimap2/ksw2.h) */
void *km_init(void);
void km_destroy(void *km);
void kfree(void *km, void *ptr);
void ksw_extd2_avx512(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m,
                      const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus,
                      int flag, ksw_extz_t *ez);
void ksw_extd2_singletrack_avx512(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                                  int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2, int w,
                                  int zdrop, int end_bonus, int flag, ksw_extz_t *ez);

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint32_t rnd(void)
{
    rng_state ^= rng_state << 13, rng_state ^= rng_state >> 7, rng_state ^= rng_state << 17;
    return (uint32_t)(rng_state >> 16);
}

/* target: uniform ACGT; query: 8% substitutions, 2% short indels */
static int make_pair(int tlen, uint8_t *t, uint8_t *q, int qcap)
{
    for (int i = 0; i < tlen; ++i) t[i] = rnd() & 3;
    int n = 0;
    for (int i = 0; i < tlen && n < qcap; ++i) {
        const uint32_t r = rnd() % 1000;
        if (r < 10) continue;                      /* deletion */
        if (r < 20 && n < qcap) q[n++] = rnd() & 3; /* insertion */
        q[n++] = r < 100 ? (uint8_t)((t[i] + 1 + rnd() % 3) & 3) : t[i];
    }
    return n;
}

static int same(const ksw_extz_t *a, const ksw_extz_t *b)
{
    if (a->max != b->max || a->zdropped != b->zdropped || a->max_q != b->max_q || a->max_t != b->max_t ||
        a->mqe != b->mqe || a->mqe_t != b->mqe_t || a->mte != b->mte || a->mte_q != b->mte_q ||
        a->score != b->score || a->m_cigar != b->m_cigar || a->n_cigar != b->n_cigar ||
        a->reach_end != b->reach_end || a->stopped != b->stopped)
        return 0;
    if (a->n_cigar > 0 && memcmp(a->cigar, b->cigar, sizeof(uint32_t) * (size_t)a->n_cigar)) return 0;
    return 1;
}

static void show(const char *what, const ksw_extz_t *a)
{
    fprintf(stderr, "  %s: score %d max %u zdropped %u mqe %d/%d mte %d/%d m_cigar %d n_cigar %d\n", what, a->score,
            a->max, a->zdropped, a->mqe, a->mqe_t, a->mte, a->mte_q, a->m_cigar, a->n_cigar);
}

int main(void)
{
    const int8_t m = 5;
    int8_t mat[25];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) mat[i * 5 + j] = i == j ? 0 : -6;
    struct {
        int w, zdrop, flag;
    } cfg[] = {{-1, -1, 0}, {-1, -1, KSW_EZ_RIGHT}, {-1, -1, KSW_EZ_SEMIGLOBAL_END}, {300, -1, 0},
               {-1, 200, KSW_EZ_EXTZ_ONLY}, {-1, -1, KSW_EZ_REV_CIGAR}, {-1, -1, KSW_EZ_SINGLETRACK}};
    const int ncfg = (int)(sizeof(cfg) / sizeof(cfg[0]));
    const int lens[] = {1, 7, 120, 700, 2000, 4500};
    int checked = 0, bad = 0;
    static uint8_t t[5000], q[6000];
    for (int li = 0; li < (int)(sizeof(lens) / sizeof(lens[0])); ++li)
        for (int c = 0; c < ncfg; ++c) {
            const int tlen = lens[li];
            int qlen = make_pair(tlen, t, q, (int)sizeof(q));
            if (qlen == 0) q[qlen++] = 0;
            const int st = cfg[c].flag & KSW_EZ_SINGLETRACK;
            void *km = km_init();
            ksw_extz_t a, b;
            memset(&a, 0, sizeof(a));
            memset(&b, 0, sizeof(b));
            for (int call = 0; call < 2; ++call) {
                /* call 1 reuses each ez and its arena-grown CIGAR buffer */
                if (st) {
                    ksw_extd2_singletrack_avx512(km, qlen, q, tlen, t, m, mat, 8, 2, 75, 1, cfg[c].w, cfg[c].zdrop, 0,
                                                 cfg[c].flag & ~KSW_EZ_SINGLETRACK, &a);
                    ksw_extd2_singletrack_hip(km, qlen, q, tlen, t, m, mat, 8, 2, 75, 1, cfg[c].w, cfg[c].zdrop, 0,
                                              cfg[c].flag & ~KSW_EZ_SINGLETRACK, &b);
                } else {
                    ksw_extd2_avx512(km, qlen, q, tlen, t, m, mat, 8, 2, 75, 1, cfg[c].w, cfg[c].zdrop, 0,
                                     cfg[c].flag, &a);
                    ksw_extd2_hip(km, qlen, q, tlen, t, m, mat, 8, 2, 75, 1, cfg[c].w, cfg[c].zdrop, 0, cfg[c].flag,
                                  &b);
                }
                const int err = ksw2_hip_last_error();
                ++checked;
                if (err != KSW2_HIP_OK || !same(&a, &b)) {
                    ++bad;
                    fprintf(stderr, "differ: qlen %d tlen %d w %d zdrop %d flag 0x%x call %d status %d\n", qlen,
                            tlen, cfg[c].w, cfg[c].zdrop, cfg[c].flag, call, err);
                    show("reference", &a);
                    show("hip", &b);
                }
            }
            kfree(km, a.cigar);
            kfree(km, b.cigar);
            km_destroy(km);
        }
    printf("kalloc arena: %d calls checked, %d differ\n", checked, bad);
    return bad ? 1 : 0;
}
