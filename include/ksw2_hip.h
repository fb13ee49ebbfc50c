/*
 * ksw2_hip.h -- C ABI of the MI355X (gfx950) implementation of KSW2's
 * two-piece-affine difference-recurrence aligner, ksw_extd2.
 *
 * Drop-in boundary: ksw_extd2_hip() has exactly the signature, argument
 * meaning, output struct and error conventions of
 *     ksw_extd2_sse / ksw_extd2_avx2 / ksw_extd2_avx512
 * declared in the reference at minimap2/ksw2.h:91-107 (called from
 * src/myImportandFunction/alignSlidingWindow.cpp:290-296, 382-388, 763-783,
 * 841-845).  The batch and serving entry points below have no counterpart in
 * the reference; they carry many such calls per launch.
 *
 * Plain C types only: no HIP or torch types cross this boundary.
 */
#ifndef KSW2_HIP_H
#define KSW2_HIP_H

#include <stddef.h>
#include <stdint.h>

/* ---- Types and flags shared with the reference header (minimap2/ksw2.h:18-50).
 * When the reference ksw2.h was included first (it defines KSW2_H_) its
 * definitions are used and these are skipped, so both headers can coexist. */
#ifndef KSW2_H_
#define KSW_NEG_INF -0x40000000

#define KSW_EZ_SCORE_ONLY     0x01   /* no path/CIGAR */
#define KSW_EZ_RIGHT          0x02   /* right-align gaps */
#define KSW_EZ_GENERIC_SC     0x04   /* full m*m matrix; else match/mismatch + wildcard m-1 */
#define KSW_EZ_APPROX_MAX     0x08   /* approximate max */
#define KSW_EZ_APPROX_DROP    0x10   /* approximate z-drop */
#define KSW_EZ_EXTZ_ONLY      0x40   /* extension only */
#define KSW_EZ_REV_CIGAR      0x80   /* reversed CIGAR */
#define KSW_EZ_SPLICE_FOR     0x100
#define KSW_EZ_SPLICE_REV     0x200
#define KSW_EZ_SPLICE_FLANK   0x400
#define KSW_EZ_SEMIGLOBAL_END 0x800  /* traceback from the better sequence end */
#define KSW_EZ_SINGLETRACK    0x1000 /* two-track (Singletrack) traceback */

#ifdef __cplusplus
extern "C" {
#endif
/* Layout identical to minimap2/ksw2.h:40-50 (56 bytes on x86-64). */
typedef struct {
    uint32_t max:31, zdropped:1;
    int max_q, max_t;   /* cell of the best score seen (extension) */
    int mqe, mqe_t;     /* best score reaching the end of the query */
    int mte, mte_q;     /* best score reaching the end of the target */
    int score;          /* score reaching both ends; KSW_NEG_INF if not reached */
    int m_cigar, n_cigar;
    int reach_end;
    int stopped;        /* the progress callback asked to stop */
    uint32_t *cigar;    /* len<<4 | op, op 0=M 1=I 2=D */
} ksw_extz_t;
#ifdef __cplusplus
}
#endif
#endif /* KSW2_H_ */

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the entry points that return one. */
#define KSW2_HIP_OK               0
#define KSW2_HIP_EINVAL          -1  /* bad argument (NULL pointer, size out of range) */
#define KSW2_HIP_EUNSUPPORTED    -2  /* parameters outside what the device path implements */
#define KSW2_HIP_ENOMEM          -3  /* device or host allocation failed */
#define KSW2_HIP_EDEVICE         -4  /* a HIP runtime call failed */
#define KSW2_HIP_ENODEV          -5  /* no usable gfx950 device */
#define KSW2_HIP_EAGAIN          -6  /* deferred: the server's device memory is committed, retry later */
#define KSW2_HIP_ETOOBIG         -7  /* larger than the no-allocation entry point's fixed scratch */

/*
 * One alignment, same contract as ksw_extd2_sse (minimap2/ksw2.h:91-107,
 * body minimap2/ksw2_extd2_sse.c:39-419):
 *   - ez is caller-owned; every field except cigar/m_cigar is reset
 *     (minimap2/ksw2.h:194-199);
 *   - ez->cigar is grown with krealloc(km, ...) when km != NULL and the
 *     process links kalloc (as AnchorWave does), else with realloc(); the
 *     caller frees it with kfree(km, ...) / free();
 *   - no return value: ez->score == KSW_NEG_INF when the end is not reached,
 *     ez->zdropped = 1 on band exit / z-drop and also when the device path
 *     cannot run the call (the AVX kernels' allocation-failure convention,
 *     minimap2/ksw2_extd2_avx.c:138-141); ksw2_hip_last_error() then tells why.
 */
void ksw_extd2_hip(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                   int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                   int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez);

/*
 * Singletrack variant, same contract as ksw_extd2_singletrack_avx512
 * (minimap2/ksw2_singletrack.hpp:41-46, body ksw2_singletrack.cpp:191-201):
 * ksw_extd2_hip with KSW_EZ_SINGLETRACK added to flag.  The DP stores the two
 * score-difference tracks instead of direction bytes and the CIGAR comes from
 * the Singletrack walk (ksw2_singletrack.cpp:63-176), always in forward
 * order.  Only a global alignment is walked: with KSW_EZ_EXTZ_ONLY or
 * KSW_EZ_SEMIGLOBAL_END the result is zdropped with no CIGAR, and a walk that
 * finds no path frees ez->cigar, leaves it NULL and sets zdropped
 * (minimap2/ksw2_extd2_avx.c:752-760).
 */
void ksw_extd2_singletrack_hip(void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                               int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                               int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez);

/*
 * No-allocation variant of ksw_extd2_hip.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and tie-breaking as ksw_extd2_hip(NULL, ...), except:
 *   - every device buffer of the call (direction bytes, column buffers,
 *     state, CIGAR slots, job tables) lives in a fixed scratch array of the
 *     library, KSW2_HIP_NOALLOC_SCRATCH_BYTES on each device it runs on,
 *     held until the process ends; the call allocates no device, pinned or
 *     managed memory and creates no stream or event (its set-up, done by
 *     ksw2_hip_init or else by its first call on a device, creates one
 *     stream and four events);
 *   - a call that needs more than the scratch is rejected before any device
 *     work: KSW2_HIP_ETOOBIG, ez->zdropped = 1;
 *   - the CIGAR is written to the caller's ez->cigar, which must hold at
 *     least qlen + tlen words (ez->m_cigar), else KSW2_HIP_EINVAL; it is
 *     never grown, freed or replaced (a Singletrack walk that finds no path
 *     leaves n_cigar = 0 with zdropped = 1);
 *   - no progress callback is polled;
 *   - calls on one device share its scratch, which keeps the last call's
 *     contents (each call clears what it needs zeroed): the library
 *     serialises them with a lock, since calls that overlapped on the array
 *     would race.
 * Returns the status (also in ksw2_hip_last_error()).
 */
#define KSW2_HIP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int ksw_extd2_hip_noalloc(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                          int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                          int w, int zdrop, int end_bonus, int flag, ksw_extz_t *ez);

/* Status of the last entry point called on this host thread, and its text. */
int ksw2_hip_last_error(void);
const char *ksw2_hip_error_string(int status);

/* Library / device set-up.  Optional: every entry point initialises lazily.
 * Returns KSW2_HIP_OK or KSW2_HIP_ENODEV. */
int ksw2_hip_init(int device);

/* Host-only (no GPU needed): which device kernel a ksw_extd2_hip call with
 * these arguments takes.  0 = the unbanded fast kernel; otherwise a bit mask
 * of why not: 1 band narrower than the matrix (exactly 1: the band kernel
 * when the band is at least 512 wide and KSW_EZ_SINGLETRACK is off, else
 * the general kernel; any other mask: the general kernel), 2 z-drop on, 4 APPROX_MAX
 * flag, 8 scoring outside the fast kernel's exactness domain,
 * 16 wildcard code m-1 present in match/mismatch mode.  -1 when the
 * reference returns before aligning (m <= 1, empty sequence, or
 * -min(mat) > 2(q+e), minimap2/ksw2_extd2_sse.c:81,105). */
int ksw2_hip_route(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                   int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                   int w, int zdrop, int flag);

/* ---- Progress / deadline hook: same meaning as minimap2/ksw2.h:52-63 and
 * minimap2/ksw2_progress.c:15-38.  The callback is polled on the host while
 * a call runs; returning 0 raises a device stop flag, the call ends with
 * ez->stopped = 1 and no CIGAR (callers discard stopped results:
 * alignSlidingWindow.cpp:1890-1893).  Thread-local, like the reference. */
typedef int (*ksw2_hip_progress_cb)(void *arguments, int antidiagonal, int total_antidiagonals);
void ksw2_hip_set_progress_callback(ksw2_hip_progress_cb callback, void *arguments, int antidiagonal_interval);
void ksw2_hip_clear_progress_callback(void);

/* ---- Batch API (no reference counterpart: replaces many per-interval
 * ksw_extd2 calls, AnchorTaskExecutor.cpp:406-470, with one launch).
 * Scoring (m, mat, q, e, q2, e2) is shared by the batch; band, z-drop,
 * end bonus and flags are per pair. */
typedef struct {
    const uint8_t *query;   /* host pointer, codes 0..m-1 */
    const uint8_t *target;  /* host pointer, codes 0..m-1 */
    int32_t qlen, tlen;
    int32_t w, zdrop, end_bonus, flag;
} ksw2_pair_t;

/* Aligns n pairs; ez[i] receives pair i's result exactly as n calls to
 * ksw_extd2_hip(NULL, ...) would (cigar grown with realloc).  Returns a
 * status code; on error no ez[] is touched beyond ksw_reset_extz. */
int ksw_extd2_hip_batch(int n, const ksw2_pair_t *pairs,
                        int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                        ksw_extz_t *ez);
/* A batch larger than the device memory budget streams through a device
 * block of the whole budget, which the library keeps for the next such call
 * (allocating and releasing hundreds of GB costs seconds per call).  Single
 * calls (ksw_extd2_hip) likewise keep, per host thread, a stream, pinned
 * staging and a device block of up to 1 GiB grown to the thread's largest
 * call, released when the thread ends.  This releases the batch block and
 * the calling thread's device block; the next such call allocates again.
 * No reference counterpart.  Returns KSW2_HIP_OK (also when nothing is
 * held). */
int ksw2_hip_release_cached_memory(void);

/* ---- Device-resident batch ("plan"; no reference counterpart).  The
 * pairs' codes are uploaded to HBM once; each run aligns all of them (DP
 * sweep + device backtrack) with results left in HBM until fetched.  This
 * is how a caller that owns many intervals keeps them resident and times
 * the device work alone.  All pairs must be eligible for a device kernel,
 * else KSW2_HIP_EUNSUPPORTED. */
typedef struct ksw2_hip_plan ksw2_hip_plan;
int ksw2_hip_plan_create(ksw2_hip_plan **plan, int n, const ksw2_pair_t *pairs,
                         int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2);
int ksw2_hip_plan_run(ksw2_hip_plan *plan);
/* results of the last run into ez[0..n) (cigar grown with realloc) */
int ksw2_hip_plan_fetch(ksw2_hip_plan *plan, ksw_extz_t *ez);
/* device time of the last run: DP sweep kernels and backtrack kernel, from
 * HIP events on the plan's stream; also the sum of qlen*tlen */
int ksw2_hip_plan_timing(const ksw2_hip_plan *plan, float *dp_ms, float *backtrack_ms, double *cells);
/* runs of the plan so far, and how many of them replayed the plan's
 * launches as one captured hipGraph (from the second run of a plan of
 * stripe-kernel pairs only, no progress callback; KSW2_HIP_GRAPH=0 turns
 * capture off) */
int ksw2_hip_plan_graph(const ksw2_hip_plan *plan, int *runs, int *replays);
void ksw2_hip_plan_destroy(ksw2_hip_plan *plan);
/* ---- Serving (no reference counterpart: the request loop that keeps the
 * device busy while interval tasks arrive, replacing AnchorWave's
 * per-thread ksw_extd2 calls under AnchorTaskExecutor).  A server owns a
 * device memory budget and a worker thread.  Submitted pairs are admitted
 * by their device footprint (direction bytes + state + CIGAR slots) as
 * memory frees up, grouped with other waiting pairs of similar size into
 * launches that overlap on the device (continuous batching), and complete
 * independently.  Each result is exactly what ksw_extd2_hip(NULL, ...)
 * returns for the same pair, alone or in any mix. */
typedef struct ksw2_hip_server ksw2_hip_server;
/* device_bytes == 0: 85% of the device's free memory, or the library's
 * cached engine block when it is free and larger (ksw_extd2_hip_batch
 * below); the server then takes that block and hands it back to the cache
 * when destroyed. */
int ksw2_hip_server_create(ksw2_hip_server **server, int8_t m, const int8_t *mat, int8_t q, int8_t e,
                           int8_t q2, int8_t e2, size_t device_bytes);
/* Queues one pair (its sequences are copied).  *ez must stay valid until the
 * ticket is returned by ksw2_hip_server_poll; ez->cigar is grown with realloc
 * (release it with free()).  Does not wait for the device. */
int ksw2_hip_server_submit(ksw2_hip_server *server, const ksw2_pair_t *pair, ksw_extz_t *ez, uint64_t *ticket);
/* Writes up to max_tickets completed tickets, waiting up to wait_ms
 * milliseconds for the first (-1: until one completes; 0: no wait).  When
 * latency_s is not NULL it receives each ticket's time from submission to
 * completion, in seconds.  Returns how many were written, or KSW2_HIP_EINVAL.
 * ksw2_hip_last_error() then reports KSW2_HIP_OK, or the first failure the
 * server has met (a pair the device path could not run completes with
 * zdropped = 1 and no CIGAR, the AVX kernels' allocation-failure convention;
 * ksw2_hip_server_poll_status tells such pairs from real z-drops). */
int ksw2_hip_server_poll(ksw2_hip_server *server, uint64_t *tickets, double *latency_s, int max_tickets,
                         int wait_ms);
/* ksw2_hip_server_poll, and status[i] receives ticket i's status: KSW2_HIP_OK,
 * or the error that kept the device path from running that pair. */
int ksw2_hip_server_poll_status(ksw2_hip_server *server, uint64_t *tickets, double *latency_s, int32_t *status,
                                int max_tickets, int wait_ms);
/* Finishes every submitted pair, then releases the server. */
void ksw2_hip_server_destroy(ksw2_hip_server *server);

/* ---- Admission for a caller's scheduler.  These replace the CPU memory
 * model of AnchorWave's algorithm selector (ksw2TracebackMemoryBytes,
 * src/myImportandFunction/AlignmentAlgorithmSelector.cpp:307-358) and its
 * memory back-pressure (AlignmentTaskDeferred,
 * src/service/AnchorTaskExecutor.h:19-32) for the GPU path. */
/* Device bytes a ksw_extd2_hip call with these arguments holds while it
 * runs (direction bytes or tracks, state, off/off_end, CIGAR slots,
 * sequences), on whichever kernel ksw2_hip_route picks; the admission
 * footprint of the batch and serving paths.  0 when the reference returns
 * before aligning; UINT64_MAX when qlen + tlen overflows.  Host only. */
uint64_t ksw2_hip_footprint(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                            int8_t m, const int8_t *mat, int8_t q, int8_t e, int8_t q2, int8_t e2,
                            int w, int zdrop, int flag);
/* ksw2_hip_server_submit, except that a pair whose footprint does not fit
 * beside the memory the server already holds and has queued is not queued:
 * the call returns KSW2_HIP_EAGAIN and *retry_ms receives a retry delay
 * (half the predicted time for the pending work, 250 ms to 5 s, as
 * alignSlidingWindow.cpp:2500-2511 sizes AlignmentTaskDeferred).  A pair
 * larger than the whole budget is submitted and fails as with
 * ksw2_hip_server_submit. */
int ksw2_hip_server_try_submit(ksw2_hip_server *server, const ksw2_pair_t *pair, ksw_extz_t *ez,
                               uint64_t *ticket, int *retry_ms);
typedef struct {
    uint64_t budget_bytes;   /* device memory the server admits against */
    uint64_t held_bytes;     /* footprints of the pairs on the device */
    uint64_t queued_bytes;   /* footprints of the pairs waiting for admission */
    int64_t pending;         /* pairs submitted and not yet completed */
    double pending_cells;    /* their qlen * tlen */
    double cells_per_s;      /* recent completion rate under load (0: none yet) */
    uint64_t seq_arena_bytes;  /* device arena holding submitted sequences */
    int64_t resident_tasks;    /* pairs whose sequences went to the arena at submission */
    int64_t staged_flights;    /* flights that uploaded their own sequences (no arena, or it was full) */
} ksw2_hip_load_t;
int ksw2_hip_server_load(ksw2_hip_server *server, ksw2_hip_load_t *load);

/* ---- Score-certified banded alignment, batched.  Replaces
 * alignScoreCertifiedKsw2 (src/myImportandFunction/alignSlidingWindow.cpp:
 * 1197-1400) for many intervals at once.  Per task, as the reference:
 * the optimal score from a full score-only ksw_extd2 call (ksw2ScoreOnly,
 * :723-796) unless precomputed; with a selection plan (anchors != NULL), the
 * chained traceback (tryChainedKsw2Traceback, :1056-1195): the alignment
 * through the plan's usable anchors, each gap between them aligned by a
 * banded ksw_extd2 call, which certifies the task with final band -2 when
 * it rescores to the optimum; then banded ksw_extd2 tracebacks
 * (alignSlidingWindow_minimap2, :800-890) at bands starting from twice
 * max(|qlen - tlen|, initial_band, 1) and doubling up to
 * min(max(qlen, tlen), maximum_band), until the rescored alignment
 * (scoreGappedTwoPieceAffineAlignment, :677-721) equals the optimal score
 * and both rows reconstruct their sequences; on failure the skipped first
 * band is tried last.  Every task's next call goes into one batched launch
 * per round, so a batch of n tasks costs one score-only launch, one launch
 * of every chain's gaps and about log2(maximum/initial) + 2 banded
 * launches.
 * Sequences are characters (A/C/G/T in either case; anything else is the
 * wildcard code 4 with KSW_EZ_GENERIC_SC, :644-660); the scoring is
 * AnchorWave's penalties (mismatch and the two gap pieces, negative), used
 * as the reference uses them.  The deadline is per batch: max_seconds > 0
 * stops the batch's device work once that much time has passed, and every
 * task not yet certified ends with status TimeLimit. */
enum {
    KSW2_CERT_COMPLETED = 0,
    KSW2_CERT_TIME_LIMIT = 1,
    KSW2_CERT_NOT_CERTIFIED = 2,
    KSW2_CERT_FAILED = 3
};
/* One certification anchor of AnchorWave's selection plan: an exact match
 * of `length` bases at these positions (AlignmentChainAnchor,
 * src/myImportandFunction/AlignmentAlgorithmSelector.h:129-133). */
typedef struct {
    uint32_t query_position;
    uint32_t reference_position;
    uint16_t length;
} ksw2_chain_anchor_t;
typedef struct {
    const char *query;      /* dna_q */
    const char *target;     /* dna_d */
    int64_t qlen, tlen;
    int64_t initial_band, maximum_band;
    const int64_t *precomputed_optimal_score;  /* NULL: computed by a score-only call */
    /* the selection plan's certification anchors
     * (selectionPlan->certificationAnchors); NULL: no selection plan */
    const ksw2_chain_anchor_t *anchors;
    int64_t n_anchors;
} ksw2_certify_task_t;
typedef struct {
    int32_t status;             /* KSW2_CERT_* (Ksw2CertifiedStatus order) */
    int64_t score, optimal_score, final_band_width;
    uint64_t traceback_attempts;
    /* COMPLETED: the certified alignment, len<<4 | op as in ksw_extz_t.  A
     * task whose certification failed keeps its best-scoring valid banded
     * alignment in best_effort_cigar.  Both are malloc'd (released by
     * ksw2_hip_certify_free). */
    uint32_t *cigar;
    int32_t n_cigar;
    uint32_t *best_effort_cigar;
    int32_t n_best_effort;
} ksw2_certify_result_t;
int ksw2_hip_certify_batch(int n, const ksw2_certify_task_t *tasks, int32_t mismatching_penalty,
                           int32_t open_gap1, int32_t extend_gap1, int32_t open_gap2, int32_t extend_gap2,
                           double max_seconds, ksw2_certify_result_t *results);
void ksw2_hip_certify_free(int n, ksw2_certify_result_t *results);

/* ---- Sliding-window alignment, batched.  Replaces alignSlidingWindow
 * (src/myImportandFunction/alignSlidingWindow.cpp:457-590) for many
 * intervals at once.  Per task, as the reference: when
 * (tlen/W)(qlen/W) <= 1 one global ksw_extd2 alignment
 * (minimap2AlignmentWithArena, :241-332); otherwise windows of W+1 bases
 * from the current start of each sequence, each aligned with
 * KSW_EZ_SEMIGLOBAL_END under the shifted penalties of :535-539
 * (alignmentMinimap2WithArena, :334-431), the next window starting where
 * the better end (mqe > mte: the query's end, else the target's) left
 * off; then the unaligned tails as one target gap and one query gap
 * (:559-580).  The score is the rescored end-to-end alignment under the
 * given penalties (:587-589).  Window k of every task runs in one batched
 * launch (the windows of one task depend on each other).
 * Sequences are characters as in ksw2_hip_certify_batch; window_size >= 0. */
typedef struct {
    const char *query;      /* dna_q */
    const char *target;     /* dna_d */
    int64_t qlen, tlen;
} ksw2_window_task_t;
typedef struct {
    int32_t status;         /* 0 done; 1 a window's CIGAR did not reach its
                             * reported end (the reference throws); 2 a
                             * window above 2^30 (the reference exits) */
    int64_t score;
    int32_t windows;        /* ksw_extd2 calls made */
    uint32_t *cigar;        /* the end-to-end alignment (M/I/D runs), malloc'd */
    int32_t n_cigar;
} ksw2_window_result_t;
int ksw2_hip_window_batch(int n, const ksw2_window_task_t *tasks, int64_t window_size, int32_t matching_score,
                          int32_t mismatching_penalty, int32_t open_gap1, int32_t extend_gap1, int32_t open_gap2,
                          int32_t extend_gap2, ksw2_window_result_t *results);
void ksw2_hip_window_free(int n, ksw2_window_result_t *results);

/* Diagnostics: copies pair `index`'s raw direction bytes (stripe-kernel
 * pairs only) into `out` and reports the rows per lane (8) and whether the
 * pair is stored transposed; with out == NULL it reports those two only.
 * Stored layout, with i the striped index and j the swept one (target and
 * query, or query and target when *transposed is 1, which the library picks
 * when the query gives more 1024-row passes than the target): stripe
 * s = i/512 is half h = s&1 of pass p = s>>1, lane k = (i%512)/8, row
 * rho = i%8, and cell (i, j) is byte
 * ((p*(swept+127) + j+k+64h)*64 + k)*16 + 4*(rho/2) + 2h + rho%2;
 * `bytes` must be at least ceil(ceil(striped/512)/2)*(swept+127)*1024.  The
 * bytes are the reference's direction bytes in either orientation.  A
 * KSW_EZ_SINGLETRACK pair has twice that: per lane and step 32 bytes, the
 * negated horizontal difference (along the swept sequence) in the first 16
 * and the negated vertical one in the next 16, each in the byte order
 * above.  Returns a status code. */
int ksw2_hip_plan_dirs(const ksw2_hip_plan *plan, int index, uint8_t *out, size_t bytes, int *R,
                       int *transposed);

#ifdef __cplusplus
}
#endif
#endif /* KSW2_HIP_H */
