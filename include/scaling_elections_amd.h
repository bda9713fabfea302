/*
 * scaling_elections_amd.h -- C ABI of the MI355X (gfx950) Schulze engine.
 *
 * Plain C: pointers, sizes and status codes; no C++ exceptions, no torch or numpy types.
 * Every entry point returns SE_OK (0) or a negative SE_ERR_* code; the message of the
 * last failure on the calling thread is available from se_last_error().  Messages that
 * the reference raises as Python RuntimeError keep the reference's exact text.
 *
 * Reference interfaces replaced (paths relative to the reference repository root):
 *   se_closure / se_closure_cpu   compute_strongest_paths(preferences, *, allow_tma,
 *                                 allow_gpu, tile_size)      scaling_elections.cu:863-989
 *   se_count / se_count_rankings  build_pairwise_preferences + populate_preferences_from_ranking
 *                                                            scaling_elections.py:37-88
 *   se_rank                       get_winner_and_ranking     scaling_elections.py:274-296
 *   se_device_count / se_device_info   log_gpus()            scaling_elections.cu:996-1016
 *   se_reduce_f32                 reduce(float32[])          scaling_elections.cu:1019-1032
 *
 * Threading: the library keeps one context (stream, lock, grow-only device scratch) per
 * device, and one more per caller stream passed to a *_device entry point.  Calls sharing
 * a context are serialised; calls on different caller streams run concurrently and their
 * work overlaps on the device.  The *_device variants take caller-owned device pointers
 * and an optional hipStream_t (passed as void*; NULL = the library's stream).
 * se_stream_release() frees the scratch the library keeps for a caller stream.
 */
#ifndef SCALING_ELECTIONS_AMD_H
#define SCALING_ELECTIONS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define SE_OK 0
#define SE_ERR_INVALID (-1)      /* bad argument (shape, size, pointer, flag)            */
#define SE_ERR_TILE (-2)         /* "Unsupported tile size" and the CPU tile checks      */
#define SE_ERR_ALLOC (-3)        /* "Failed to allocate memory on device"                */
#define SE_ERR_DEVICE (-4)       /* kernel launch / execution error                      */
#define SE_ERR_COPY (-5)         /* host<->device copy failed                            */
#define SE_ERR_NO_DEVICE (-6)    /* no usable HIP device                                 */
#define SE_ERR_INTERRUPTED (-7)  /* "Stopped by signal"                                  */
#define SE_ERR_RANGE (-8)        /* value does not fit the requested width (u16)         */

/* ---- closure flags ------------------------------------------------------------------ */
#define SE_FLAG_AUTO_WIDTH 0u       /* u16 packed when max(p0) <= 65535, else u32         */
#define SE_FLAG_FORCE_U32 1u        /* always run the 32-bit closure                      */
#define SE_FLAG_FORCE_U16 2u        /* require the 16-bit closure (SE_ERR_RANGE if not)   */
#define SE_FLAG_TIME_KERNELS 4u     /* record hipEvents around every kernel (se_timing)   */
#define SE_FLAG_ALLOW_TMA 8u        /* accepted and ignored: AMD GPUs have no TMA         */
                                    /* (scaling_elections.cu:555-558)                     */
#define SE_FLAG_PIVOT32 16u         /* force pivot depth 32 or 64 (default: the tuning     */
#define SE_FLAG_PIVOT64 32u         /* table); the result never depends on it              */
#define SE_FLAG_NO_F16 128u         /* 16-bit path: keep integer min/max even when the     */
                                    /* values fit the f16-ordered range                   */
#define SE_FLAG_NO_LOOKAHEAD 256u   /* force off / on the overlap of the next round's      */
#define SE_FLAG_LOOKAHEAD 512u      /* pivot and panels with the current independent tiles */
                                    /* (default: the tuning table); LOOKAHEAD runs the    */
                                    /* next pivot's strips in a launch of their own       */
#define SE_FLAG_LOOKAHEAD_FUSED 1024u /* look-ahead in one launch per round: the panel     */
                                    /* workgroups update the next pivot's strips (16-bit  */
                                    /* paths with pivot 64; otherwise as LOOKAHEAD)       */

/* Per-call measurements, filled when a non-NULL pointer is passed. */
typedef struct se_timing {
    double total_ms;          /* whole device pipeline: p0 init + rounds + output      */
    double init_ms;           /* p0 init kernels (SE_FLAG_TIME_KERNELS)                */
    double diag_ms;           /* sum over rounds: pivot-tile kernel                    */
    double panel_ms;          /* sum over rounds: row/column panel kernel              */
    double indep_ms;          /* sum over rounds: independent-tile kernel              */
    double output_ms;         /* widening + diagonal zeroing + unpadding               */
    uint32_t rounds;          /* pivot rounds run                                      */
    uint32_t pivot;           /* pivot depth used                                      */
    uint32_t width_bits;      /* 16 or 32                                              */
    uint32_t padded_n;        /* matrix side after zero padding                        */
    uint64_t indep_cells;     /* cell updates issued by the independent kernel (all    */
                              /* rounds) = sum over rounds of (N - T)^2 * T            */
    uint32_t f16_ordered;     /* 1: 16-bit values offset into positive f16 normals and  */
                              /* closed with f16 minimum/maximum (exact, see closure)  */
    uint32_t reserved;
} se_timing;

/* ---- diagnostics -------------------------------------------------------------------- */
const char *se_last_error(void);
int se_version(void);
int se_device_count(int *count);
/* name: caller buffer of name_len bytes (NUL-terminated on return). */
int se_device_info(int device, char *name, size_t name_len, int *compute_units,
                   size_t *global_mem_bytes, int *major, int *minor, char *arch, size_t arch_len);

/* ---- strongest paths (the closure) ---------------------------------------------------- */
/* Host buffers.  prefs: n x n uint32, row stride `row_stride` elements (>= n).  paths:
 * n x n uint32, contiguous, written entirely (p0 rule, max-min closure, zero diagonal).
 * tile_size must be 4, 8, 16 or 32 (scaling_elections.cu:895-901; 0 is rejected as in
 * the reference); the result does not depend on it.  Any n >= 0 is accepted: the GPU
 * matrix is zero padded, which is exact. */
int se_closure(const uint32_t *prefs, uint32_t n, size_t row_stride, uint32_t *paths,
               uint32_t tile_size, uint32_t flags, se_timing *timing);

/* Device buffers (caller-owned), same semantics; paths has row stride paths_stride. */
int se_closure_device(const uint32_t *prefs_dev, uint32_t n, size_t row_stride,
                      uint32_t *paths_dev, size_t paths_stride, uint32_t flags, void *stream,
                      se_timing *timing);

/* CPU branch of compute_strongest_paths (allow_gpu=False), scaling_elections.cu:933-988:
 * tile_size 0 picks 4 when n >= 4, else "Number of candidates should be at least 4,
 * ideally divisible by 4"; tile_size > n gives "Tile size should be less than or equal
 * to the number of candidates"; tiles other than 4..128 powers of two give "Unsupported
 * tile size".  threads <= 0 uses every hardware thread. */
int se_closure_cpu(const uint32_t *prefs, uint32_t n, size_t row_stride, uint32_t *paths,
                   uint32_t tile_size, int threads);

/* Serial CPU closure (the CLI's validation baseline), scaling_elections.py:91-125. */
int se_closure_cpu_serial(const uint32_t *prefs, uint32_t n, size_t row_stride, uint32_t *paths);

/* ---- pairwise preference counter ---------------------------------------------------- */
/* positions: m x n int32 (row v = voter v), positions[v][c] = place of candidate c in
 * voter v's ballot.  Every row must be a permutation of 0..n-1 (a strict order), which the
 * builder guarantees by filling incomplete ballots.  prefs: n x n uint32 (row stride
 * prefs_stride) receives d[i][j] = #{v : positions[v][i] < positions[v][j]}.
 * se_count returns SE_ERR_INVALID when a position lies outside the range the counter packs
 * (below 128 for n <= 128, below 32768 for n <= 32768); se_count_device does not check. */
int se_count(const int32_t *positions, uint32_t m, uint32_t n, uint32_t *prefs,
             size_t prefs_stride);
int se_count_device(const int32_t *positions_dev, uint32_t m, uint32_t n, uint32_t *prefs_dev,
                    size_t prefs_stride, void *stream, double *kernel_ms);

/* General ballots, exactly populate_preferences_from_ranking applied to each ballot
 * (scaling_elections.py:46-48), duplicates included: ids is the concatenation of m
 * (already filled) ballots, offsets has m+1 entries.  Every id must lie in [0, n). */
int se_count_rankings(const int32_t *ids, const int64_t *offsets, uint32_t m, uint32_t n,
                      uint32_t *prefs, size_t prefs_stride);

/* ---- winner and ranking ------------------------------------------------------------- */
/* wins[i] = #{j != i : p[i][j] > p[j][i]}; ranking = indices by descending wins, ties in
 * ascending index (Python's stable sort), scaling_elections.py:287-292. */
int se_rank(const uint32_t *paths, uint32_t n, size_t row_stride, uint32_t *wins,
            uint32_t *ranking);
int se_rank_device(const uint32_t *paths_dev, uint32_t n, size_t row_stride, uint32_t *wins_host,
                   uint32_t *ranking_host, void *stream);

/* ---- batches of independent elections -------------------------------------------------- */
/* No single reference counterpart: the reference handles one election per call
 * (build_pairwise_preferences + compute_strongest_paths + get_winner_and_ranking,
 * scaling_elections.py:51-296); this entry point runs the same three steps for many
 * elections at once and returns, per election, exactly what those calls return.
 * candidates/voters: host arrays of `elections` entries (C_e <= 1024; m_e voters).
 * positions_dev: concatenation over e of m_e x C_e position arrays (strict orders), int32,
 * or uint16 with SE_BATCH_POS16.
 * paths_dev: concatenation of C_e x C_e uint32 strongest-path matrices (zero diagonal).
 * ranking_dev: concatenation of C_e candidate indices by descending wins, ties ascending.
 * Default strategy: one counting launch over work items for the whole batch, then one
 * workgroup per election for p0, closure, wins and ranking; on the 16-bit paths the
 * elections above 128 candidates share one pair of launches per round instead (panels,
 * then tiles).  SE_BATCH_PER_LAUNCH runs every election through the single-election path
 * instead (for comparison). */
#define SE_BATCH_PER_LAUNCH 64u
/* Close every batched election in one workgroup, also those above 128 candidates, which by
 * default share per-round launches (panels, then tiles) with the batch's other large
 * elections.  Results never change; kept for comparison. */
#define SE_BATCH_ONE_GROUP 1024u
/* positions_dev holds uint16 positions instead of int32 (every position of a batched
 * election is below 1024): half the bytes the counter reads.  Results never change. */
#define SE_BATCH_POS16 2048u

typedef struct se_batch_timing {
    double total_ms;          /* counting + closure                                    */
    double count_ms;          /* counting launch (batched) / all per-election work     */
    double closure_ms;        /* closure launch (batched)                              */
    uint32_t elections;
    uint32_t reserved;
    uint64_t cells;           /* sum of C^3                                            */
    uint64_t pair_votes;      /* sum of m * C * (C - 1) / 2                            */
} se_batch_timing;

/* Device-side conversion of int32 positions (count values) to the uint16 form
 * SE_BATCH_POS16 takes; synchronous.  SE_ERR_INVALID when a value lies outside 0..65535
 * (dst then holds its low 16 bits).  No reference counterpart (an input format). */
int se_positions16_device(const int32_t *src_dev, size_t count, uint16_t *dst_dev, void *stream);

int se_batch_device(uint32_t elections, const uint32_t *candidates, const uint32_t *voters,
                    const void *positions_dev, uint32_t *paths_dev, uint32_t *ranking_dev,
                    uint32_t flags, void *stream, se_batch_timing *timing);

/* ---- launch-shape tuning (no reference counterpart) ------------------------------------ */
/* The closure's pivot depth and look-ahead, and the batch's split between elections closed
 * by one workgroup and elections closed by shared per-round launches, are picked from a
 * process-wide table.  The built-in table holds the choices measured fastest on MI355X;
 * scalingelections_amd.autotune re-measures them on the current device and installs its own
 * table.  Results never depend on the table: every choice computes the same closure.
 * SE_FLAG_PIVOT32/64 and SE_FLAG_(NO_)LOOKAHEAD on a call override it. */
#define SE_TUNE_F16 0u   /* 16-bit closure, f16-ordered (nonzero p0 values span <= 30720)   */
#define SE_TUNE_U16 1u   /* 16-bit closure, integer min/max                                */
#define SE_TUNE_U32 2u   /* 32-bit closure                                                 */
#define SE_TUNE_MAX_ENTRIES 64u
typedef struct se_tune_entry {
    uint32_t width;      /* SE_TUNE_F16, SE_TUNE_U16 or SE_TUNE_U32                        */
    uint32_t max_side;   /* applies to zero-padded sides N <= max_side (N: n rounded up to */
                         /* 128); the smallest max_side >= N of the width is used          */
    uint32_t pivot;      /* pivot depth: 32 or 64                                          */
    uint32_t lookahead;  /* 0: off; 1: the next pivot's panel overlaps the current round,   */
                         /* its strips in a launch of their own; 2: the same in one launch */
                         /* per round (16-bit widths with pivot 64 only)                   */
} se_tune_entry;
/* Installs a table (entries need not be sorted); count 0 restores the built-in table.
 * Widths without an entry covering N fall back to the built-in table. */
int se_tune_closure(const se_tune_entry *entries, uint32_t count);
/* Copies the active table (at most capacity entries); *count receives its size. */
int se_tune_closure_get(se_tune_entry *entries, uint32_t capacity, uint32_t *count);
/* Batched elections with at least rounds_min_candidates candidates share per-round launches
 * (16-bit paths); smaller ones are closed by one workgroup each.  Accepted: 129..1025
 * (1025: none share launches); 0 restores the built-in value. */
int se_tune_batch(uint32_t rounds_min_candidates);
int se_tune_batch_get(uint32_t *rounds_min_candidates);

/* ---- workload generation (no reference counterpart) ------------------------------------ */
/* Mallows ballots by repeated insertion (item k of the reference order inserted at
 * position j <= k with probability proportional to phi^(k - j)), seeded per (seed,
 * election, voter); writes positions in the se_batch_device layout.  reference: the
 * concatenated reference orders (C_e entries each).  threads <= 0: up to 16. */
int se_synth_mallows(uint32_t elections, const uint32_t *candidates, const uint32_t *voters,
                     const double *phi, const int32_t *reference, uint64_t seed, int32_t *positions,
                     int threads);

/* ---- misc ---------------------------------------------------------------------------- */
/* Sequential float sum, the reference's non-CUDA reduce() (std::accumulate order). */
float se_reduce_f32(const float *data, size_t count);

/* Interruption: se_interrupt() sets a flag that running closures check between pivot
 * rounds ("Stopped by signal", scaling_elections.cu:758-759); se_install_sigint() makes
 * SIGINT set it (returns the previous handler state for se_restore_sigint). */
int se_stream_release(void *stream);
void se_interrupt(void);
void se_clear_interrupt(void);
int se_install_sigint(void);
void se_restore_sigint(void);

#ifdef __cplusplus
}
#endif

#endif /* SCALING_ELECTIONS_AMD_H */
