/*
 * oracle.c -- CPU restatement of the reference Schulze path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load the library built from it.  The product (scalingelections_amd) never
 * links, loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic is the reference's own: uint32
 * counts, uint32 min/max, strict '>' comparisons.
 *
 * Pinning (each function is checked by tests/test_oracle.py against):
 *   oracle_p0, oracle_closure_serial, oracle_closure_tiled:
 *       - the validate_paradox.py 10/11-voter fixture (hand-derived d, p, ranking),
 *       - the 45-voter textbook Schulze example (d, p, ranking),
 *       - brute force by per-threshold reachability (a different algorithm),
 *       - oracle_widest_row (modified Dijkstra, a third algorithm),
 *       - the reference's own OpenMP CPU path compiled from source (oracle/_ref),
 *       - closure properties (idempotence, monotone-remap and relabelling
 *         equivariance, tile-size independence, zero padding neutrality).
 *   oracle_populate / oracle_build_prefs:
 *       - the same fixtures, brute-force recount, d + d^T = m (1 - I) for
 *         complete ballots, the row-sum identity, incomplete-ballot fill rule.
 *   oracle_wins / oracle_ranking:
 *       - fixture rankings, Python's stable sorted() on random ties.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* p0 rule: scaling_elections.py:103-112 (serial), :179-188 (tiled);
 * same rule in scaling_elections.cu:495-503 and :742-752.
 * p[i][j] = d[i][j] if d[i][j] > d[j][i] else 0; p[i][i] = 0. */
void oracle_p0(const uint32_t *d, int64_t n, int64_t stride, uint32_t *p, int64_t p_stride) {
    /* 64 x 64 blocks so the transposed operand d[j][i] stays in cache */
#pragma omp parallel for schedule(dynamic)
    for (int64_t bi = 0; bi < n; bi += 64)
        for (int64_t bj = 0; bj < n; bj += 64)
            for (int64_t i = bi; i < bi + 64 && i < n; i++)
                for (int64_t j = bj; j < bj + 64 && j < n; j++) {
                    uint32_t v = 0;
                    if (i != j && d[i * stride + j] > d[j * stride + i]) v = d[i * stride + j];
                    p[i * p_stride + j] = v;
                }
}

/* Serial strongest paths: scaling_elections.py:91-125.
 * For pivot i, row j != i, column k not in {i, j}:
 *     p[j][k] = max(p[j][k], min(p[j][i], p[i][k])). */
void oracle_closure_serial(const uint32_t *d, int64_t n, int64_t stride, uint32_t *p) {
    oracle_p0(d, n, stride, p, n);
    for (int64_t i = 0; i < n; i++) {
        const uint32_t *pi = p + i * n;
        for (int64_t j = 0; j < n; j++) {
            if (i == j) continue;
            uint32_t *pj = p + j * n;
            uint32_t pji = pj[i];
            for (int64_t k = 0; k < n; k++) {
                if (i == k || j == k) continue;
                uint32_t m = pji < pi[k] ? pji : pi[k];
                if (m > pj[k]) pj[k] = m;
            }
        }
    }
}

/* Tile micro-kernel: scaling_elections.py:128-159 (compute_strongest_paths_tile_numba).
 * c, a and b are views into one matrix of row stride `ld` and may alias, exactly as in
 * the reference.  For k, i, j < T: when the global c, a and b cells are all off the
 * diagonal and min(a[i][k], b[k][j]) > c[i][j], c[i][j] takes that value. */
static void tile_masked(uint32_t *m, int64_t ld, int64_t T,
                        int64_t c_row, int64_t c_col, int64_t a_row, int64_t a_col,
                        int64_t b_row, int64_t b_col) {
    for (int64_t k = 0; k < T; k++)
        for (int64_t i = 0; i < T; i++) {
            uint32_t *c = m + (c_row + i) * ld + c_col;
            const uint32_t av = m[(a_row + i) * ld + a_col + k];
            const uint32_t *b = m + (b_row + k) * ld + b_col;
            const int a_ok = (a_row + i) != (a_col + k);
            for (int64_t j = 0; j < T; j++) {
                if ((c_row + i) != (c_col + j) && a_ok && (b_row + k) != (b_col + j)) {
                    uint32_t r = av < b[j] ? av : b[j];
                    if (r > c[j]) c[j] = r;
                }
            }
        }
}

/* Same micro-kernel when no c, a or b cell can lie on the diagonal (all three tiles are
 * off-diagonal tiles of an aligned tiling): every mask above is true, so the body reduces
 * to the unmasked update.  Kept separate only so the compiler can vectorise it. */
static void tile_plain(uint32_t *restrict c, const uint32_t *restrict a, const uint32_t *restrict b,
                       int64_t ld, int64_t T) {
    for (int64_t k = 0; k < T; k++)
        for (int64_t i = 0; i < T; i++) {
            uint32_t *ci = c + i * ld;
            const uint32_t av = a[i * ld + k];
            const uint32_t *bk = b + k * ld;
            for (int64_t j = 0; j < T; j++) {
                uint32_t r = av < bk[j] ? av : bk[j];
                ci[j] = r > ci[j] ? r : ci[j];
            }
        }
}

static void tile_call(uint32_t *m, int64_t ld, int64_t T,
                      int64_t c_row, int64_t c_col, int64_t a_row, int64_t a_col,
                      int64_t b_row, int64_t b_col) {
    if (c_row != c_col && a_row != a_col && b_row != b_col && !(c_row == a_row && c_col == a_col) &&
        !(c_row == b_row && c_col == b_col))
        tile_plain(m + c_row * ld + c_col, m + a_row * ld + a_col, m + b_row * ld + b_col, ld, T);
    else
        tile_masked(m, ld, T, c_row, c_col, a_row, a_col, b_row, b_col);
}

/* Tiled strongest paths: scaling_elections.py:162-271 (compute_strongest_paths_numba_parallel).
 * Three phases per pivot tile k: diagonal tile (:197-208), column panel (:211-227),
 * row panel (:230-246), independent tiles (:249-269).  The reference is only defined
 * for n % T == 0; here the matrix is zero padded to a multiple of T, which is exact:
 * padded rows and columns hold 0 and min(x, 0) = 0 never raises a cell.
 * `max_rounds` < 0 runs every round; otherwise only the first max_rounds pivot tiles
 * (a bounded sample of the same work, for timing).  `threads` <= 0 keeps the OpenMP
 * default.  Returns 0, or -1 on allocation failure. */
int oracle_closure_tiled(const uint32_t *d, int64_t n, int64_t stride, uint32_t *p, int64_t T,
                         int threads, int64_t max_rounds) {
    if (T <= 0) return -1;
    int64_t tiles = (n + T - 1) / T;
    int64_t N = tiles * T;
    uint32_t *m = p;
    if (N != n) {
        m = (uint32_t *)calloc((size_t)(N * N), sizeof(uint32_t));
        if (!m) return -1;
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    oracle_p0(d, n, stride, m, N);
    int64_t rounds = tiles;
    if (max_rounds >= 0 && max_rounds < rounds) rounds = max_rounds;
    for (int64_t k = 0; k < rounds; k++) {
        int64_t ks = k * T;
        tile_call(m, N, T, ks, ks, ks, ks, ks, ks);
#pragma omp parallel for schedule(dynamic)
        for (int64_t i = 0; i < tiles; i++) {
            if (i == k) continue;
            tile_call(m, N, T, i * T, ks, i * T, ks, ks, ks);
        }
#pragma omp parallel for schedule(dynamic)
        for (int64_t j = 0; j < tiles; j++) {
            if (j == k) continue;
            tile_call(m, N, T, ks, j * T, ks, ks, ks, j * T);
        }
#pragma omp parallel for schedule(dynamic)
        for (int64_t i = 0; i < tiles; i++) {
            if (i == k) continue;
            for (int64_t j = 0; j < tiles; j++) {
                if (j == k) continue;
                tile_call(m, N, T, i * T, j * T, i * T, ks, ks, j * T);
            }
        }
    }
    if (m != p) {
        for (int64_t i = 0; i < n; i++) memcpy(p + i * n, m + i * N, (size_t)n * sizeof(uint32_t));
        free(m);
    }
    return 0;
}

/* Per-source widest path by a modified Dijkstra -- an algorithm independent of
 * Floyd-Warshall, used to spot-check single rows of large closures.
 * out[j] = max over paths src -> j of the minimum p0 edge on the path; out[src] = 0.
 * p0 is the strongest-path initialisation (oracle_p0 output), row stride `ld`.
 * Returns 0, or -1 on allocation failure. */
int oracle_widest_row(const uint32_t *p0, int64_t n, int64_t ld, int64_t src, uint32_t *out) {
    unsigned char *done = (unsigned char *)calloc((size_t)n, 1);
    if (!done) return -1;
    for (int64_t j = 0; j < n; j++) out[j] = p0[src * ld + j];
    out[src] = 0;
    done[src] = 1;
    for (int64_t it = 1; it < n; it++) {
        int64_t best = -1;
        uint32_t bv = 0;
        for (int64_t j = 0; j < n; j++)
            if (!done[j] && (best < 0 || out[j] > bv)) { best = j; bv = out[j]; }
        if (best < 0 || bv == 0) break; /* everything left is unreachable (width 0) */
        done[best] = 1;
        const uint32_t *row = p0 + best * ld;
        for (int64_t j = 0; j < n; j++) {
            if (done[j]) continue;
            uint32_t w = bv < row[j] ? bv : row[j];
            if (w > out[j]) out[j] = w;
        }
    }
    out[src] = 0;
    free(done);
    return 0;
}

/* populate_preferences_from_ranking: scaling_elections.py:37-48.
 * prefs[ranking[i]][ranking[j]] += 1 for every i < j (uint32, wraps mod 2^32).
 * Ids must lie in [0, n); duplicates behave as in the reference (they touch the
 * diagonal and the reverse cells). */
void oracle_populate(uint32_t *prefs, int64_t n, const int64_t *ranking, int64_t len) {
    for (int64_t i = 0; i < len; i++) {
        uint32_t *row = prefs + ranking[i] * n;
        for (int64_t j = i + 1; j < len; j++) row[ranking[j]] += 1u;
    }
}

/* build_pairwise_preferences: scaling_elections.py:51-88, for ballots given as one flat
 * id array plus m+1 offsets.  n must be max(1, max id + 1) (:64-66; computed by the
 * caller).  A ballot shorter than n gets the ids it lacks appended in ascending order
 * (:75-83), then every ballot is counted (:86).  Returns 0, or -1 on allocation failure. */
int oracle_build_prefs(const int64_t *flat, const int64_t *offsets, int64_t m, int64_t n,
                       uint32_t *prefs) {
    memset(prefs, 0, (size_t)(n * n) * sizeof(uint32_t));
    unsigned char *present = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    if (!present) return -1;
    for (int64_t v = 0; v < m; v++) {
        const int64_t *r = flat + offsets[v];
        int64_t len = offsets[v + 1] - offsets[v];
        if (len != n) {
            /* len may exceed n only with duplicates; the mask is still over 0..n-1 */
            int64_t *buf = (int64_t *)malloc((size_t)(len + n + 1) * sizeof(int64_t));
            if (!buf) { free(present); return -1; }
            memset(present, 0, (size_t)n);
            for (int64_t i = 0; i < len; i++) { buf[i] = r[i]; present[r[i]] = 1; }
            int64_t L = len;
            for (int64_t c = 0; c < n; c++)
                if (!present[c]) buf[L++] = c;
            oracle_populate(prefs, n, buf, L);
            free(buf);
        } else {
            oracle_populate(prefs, n, r, len);
        }
    }
    free(present);
    return 0;
}

/* Wins count: scaling_elections.py:287-290.  wins[i] = #{j != i : p[i][j] > p[j][i]}. */
void oracle_wins(const uint32_t *p, int64_t n, int64_t *wins) {
    for (int64_t i = 0; i < n; i++) {
        int64_t w = 0;
        for (int64_t j = 0; j < n; j++)
            if (i != j && p[i * n + j] > p[j * n + i]) w++;
        wins[i] = w;
    }
}

/* Ranking: scaling_elections.py:292 -- sorted(range(n), key=wins, reverse=True).
 * Python's sort is stable and reverse=True keeps equal keys in their original order,
 * so ties stay in ascending index.  Written as a stable insertion sort (O(n^2), test
 * sizes only). */
void oracle_ranking(const int64_t *wins, int64_t n, int64_t *ranking) {
    for (int64_t i = 0; i < n; i++) {
        int64_t x = i;
        int64_t pos = i;
        while (pos > 0 && wins[ranking[pos - 1]] < wins[x]) {
            ranking[pos] = ranking[pos - 1];
            pos--;
        }
        ranking[pos] = x;
    }
}
