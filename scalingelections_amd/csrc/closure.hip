// closure.hip -- strongest-path (max-min Floyd-Warshall) closure for gfx950.
//
// Semantics follow the reference's serial kernel (scaling_elections.py:91-125):
//   p0[i][j] = d[i][j] if d[i][j] > d[j][i] else 0, p0[i][i] = 0   (:103-112)
//   p[j][k] = max(p[j][k], min(p[j][i], p[i][k]))  for every pivot i (:115-123)
// The result is the max-min closure with a zero diagonal.  It is computed here as a
// blocked Floyd-Warshall over pivot blocks K of depth T, three phases per block, like the
// reference's tiled kernels (scaling_elections.py:162-271, scaling_elections.cu:224-343),
// but laid out for CDNA:
//   * pivot + panels -- one fused kernel: every workgroup closes the T x T tile (K,K) in
//                registers, then its chunk of rows K / columns K is one max-min product
//                with the closed pivot tile (no sequential dependence once it is closed);
//   * independent -- every cell outside rows/columns K gets max(c, A (x) B) with A the
//                column panel and B the row panel: a depth-T max-min "GEMM" on 128 x 128
//                macro tiles, 8 x 8 register micro tiles per lane, strips staged in LDS.
// The reference's diagonal masks are dropped: an update of p[i][j] through pivot i or j
// is min(p[i][i], p[i][j]) or min(p[i][j], p[j][j]), never above p[i][j], so diagonal
// values cannot change any off-diagonal result; the output kernel writes 0 there.
// The matrix is zero padded to a multiple of 128: padded rows/columns stay 0 and
// min(x, 0) = 0 never raises a cell.  When max(p0) <= 65535 the whole closure runs on
// 16-bit values with packed v_pk_min_u16 / v_pk_max_u16 (closure values are a subset of
// the p0 values and 0, so the narrowing is exact); otherwise on 32-bit values, where two
// pivot steps fold into one v_max3_u32.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "closure_tiles.h"
#include "se_internal.h"

namespace se {

// ---------------------------------------------------------------------------------------
// p0 init over 64 x 64 tile pairs {(I, J), (J, I)} with I <= J: each pair reads both blocks
// of d once (row-major, coalesced; the (J, I) block goes through LDS transposed) and
// handles both orientations, so d is read once per kernel.  Pairs are numbered row-major
// over the upper triangle of the nt x nt tile grid and visited grid-stride.

constexpr uint32_t kInitTile = 64;
// The p0 range (max, min of the nonzero p0 values) is reduced into kRangeSlots pairs,
// kRangePitch words apart, that the host combines.
constexpr uint32_t kRangeSlots = 32, kRangePitch = 16;

// Pair t -> (I, J), I <= J, of the row-major upper triangle of an nt x nt grid.
__device__ __forceinline__ void upper_pair(uint32_t t, uint32_t nt, uint32_t* I, uint32_t* J) {
    // rows before I hold I nt - I (I - 1) / 2 pairs; solve for I, then correct by +-1
    const double b = 2.0 * nt + 1.0;
    int64_t i = (int64_t)((b - sqrt(b * b - 8.0 * (double)t)) / 2.0);
    auto first = [&](int64_t r) { return (uint64_t)(r * (int64_t)nt - r * (r - 1) / 2); };
    if (i < 0) i = 0;
    while (i > 0 && first(i) > t) --i;
    while (i + 1 < (int64_t)nt && first(i + 1) <= t) ++i;
    *I = (uint32_t)i;
    *J = (uint32_t)(i + (t - first(i)));
}

// Loads d block (I, J) into registers (row a = ty + 4 q) and block (J, I) transposed into
// tr (tr[b][a] = d[J0 + b][I0 + a]); zeros outside n x n.
template <typename D>
__device__ __forceinline__ void load_pair(const D* __restrict__ d, uint32_t n, size_t ld, uint32_t I0,
                                          uint32_t J0, uint32_t (*tr)[kInitTile + 1], uint32_t* v) {
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t q = 0; q < kInitTile / 4; ++q) {
        const uint32_t a = ty + 4 * q;
        const uint32_t i = I0 + a, j = J0 + tx;
        v[q] = (i < n && j < n) ? (uint32_t)d[(size_t)i * ld + j] : 0u;
        const uint32_t r = J0 + a, c = I0 + tx;
        tr[tx][a] = (r < n && c < n) ? (uint32_t)d[(size_t)r * ld + c] : 0u;  // tr[c - I0][r - J0]
    }
}

// Range of p0 over the n x n domain: out[0] = max (atomicMax), out[1] = smallest nonzero
// value (atomicMin, stays 0xFFFFFFFF when p0 is all zero).  Decides the 16-bit eligibility
// and whether the values fit the f16-ordered range.  A pair (i, j), i != j, contributes
// max(d[i][j], d[j][i]) when the two differ (that larger count is a nonzero p0 entry).
// D: element type of d (uint32_t, or uint16_t when the host entry narrowed it in transit).
template <typename D>
__global__ __launch_bounds__(256) void p0_max_kernel(const D* __restrict__ d, uint32_t n,
                                                     size_t ld, uint32_t nt, uint32_t* __restrict__ out) {
    __shared__ uint32_t tr[kInitTile][kInitTile + 1];
    __shared__ uint32_t red[kThreads / 64], redmin[kThreads / 64];
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const uint32_t pairs = nt * (nt + 1) / 2;
    uint32_t m = 0, mn = 0xFFFFFFFFu;
    for (uint32_t t = blockIdx.x; t < pairs; t += gridDim.x) {
        uint32_t I, J;
        upper_pair(t, nt, &I, &J);
        const uint32_t I0 = I * kInitTile, J0 = J * kInitTile;
        uint32_t v[kInitTile / 4];
        __syncthreads();
        load_pair(d, n, ld, I0, J0, tr, v);
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kInitTile / 4; ++q) {
            const uint32_t a = ty + 4 * q;
            const uint32_t x = v[q], y = tr[a][tx];  // d[i][j], d[j][i] with i = I0 + a, j = J0 + tx
            if (I0 + a != J0 + tx && x != y) {
                const uint32_t w = max(x, y);
                m = max(m, w);
                mn = min(mn, w);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = m;
        redmin[threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = max(max(red[0], red[1]), max(red[2], red[3]));
        uint32_t rn = min(min(redmin[0], redmin[1]), min(redmin[2], redmin[3]));
        // one of kRangeSlots (max, min) pairs, a 64-byte line each: same-address atomics
        // from thousands of workgroups serialize (4096 candidates: 60 us with one pair)
        uint32_t* slot = out + (blockIdx.x % kRangeSlots) * kRangePitch;
        if (r) atomicMax(slot, r);
        if (rn != 0xFFFFFFFFu) atomicMin(slot + 1, rn);
    }
}

// Writes p0 into the padded work matrix W (N x N, row pitch ldw, element E), both
// orientations of each tile pair; nonzero values are stored as v + bias (bias = 0, or the
// f16-ordered offset 0x0400 - min).  Block (J, I) is assembled transposed in LDS, then
// written row-major.
template <typename E, typename D>
__global__ __launch_bounds__(256) void p0_write_kernel(const D* __restrict__ d, uint32_t n,
                                                       size_t ld, E* __restrict__ W, size_t ldw, uint32_t nt,
                                                       uint32_t bias) {
    __shared__ uint32_t tr[kInitTile][kInitTile + 1];
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const uint32_t pairs = nt * (nt + 1) / 2;
    for (uint32_t t = blockIdx.x; t < pairs; t += gridDim.x) {
        uint32_t I, J;
        upper_pair(t, nt, &I, &J);
        const uint32_t I0 = I * kInitTile, J0 = J * kInitTile;
        uint32_t v[kInitTile / 4];
        __syncthreads();
        load_pair(d, n, ld, I0, J0, tr, v);
        __syncthreads();
        // (I, J) orientation straight out; (J, I) values go back into tr transposed
#pragma unroll
        for (uint32_t q = 0; q < kInitTile / 4; ++q) {
            const uint32_t a = ty + 4 * q;
            const uint32_t i = I0 + a, j = J0 + tx;
            const uint32_t x = v[q], y = tr[a][tx];
            const bool off = i != j && i < n && j < n;
            W[(size_t)i * ldw + j] = (E)((off && x > y) ? x + bias : 0u);
            v[q] = (off && y > x) ? y + bias : 0u;  // p0[j][i]
        }
        if (I == J) continue;  // the diagonal pair wrote both orientations above
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kInitTile / 4; ++q) tr[tx][ty + 4 * q] = v[q];  // tr[j - J0][i - I0]
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kInitTile / 4; ++q) {
            const uint32_t b = ty + 4 * q;  // row J0 + b, column I0 + tx
            W[(size_t)(J0 + b) * ldw + I0 + tx] = (E)tr[b][tx];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Pivot + panels, fused: every workgroup closes the pivot tile P = W[K][K] itself
// (close_pivot), then computes its chunk of the row panel (blocks [0, nb)) or the column
// panel (blocks [nb, 2 nb)).  Block 0 stores the closed P into pivot_out; the next round's
// block 0 (or commit_pivot_kernel after the last round) copies it into W.  Nothing reads
// W[K][K] before then: the independent kernel of this round only uses it for pivot-row and
// pivot-column cells, which it does not store.

// One workgroup of the 32-bit panel of the pivot block at kb: index < nb is row chunk
// `index`, the others column chunk index - nb.  Workgroup 0 commits the previous closed
// pivot (when given) and publishes this one.  LDS: pivot_panel_lds_words<T>() words.
template <uint32_t T>
constexpr uint32_t pivot_panel_lds_words() { return T * (T + 1) + kMacro * (T + 1) + 4 * T; }

template <typename E, uint32_t T>
__device__ __forceinline__ void pivot_panel(E* __restrict__ W, size_t ldw, uint32_t kb, uint32_t nb,
                                            uint32_t index, uint32_t* __restrict__ pivot_out,
                                            const uint32_t* __restrict__ pivot_prev, uint32_t kb_prev,
                                            uint32_t* lds) {
    uint32_t(*P)[T + 1] = reinterpret_cast<uint32_t(*)[T + 1]>(lds);
    uint32_t* S = lds + T * (T + 1);
    uint32_t(*rowbuf)[T] = reinterpret_cast<uint32_t(*)[T]>(S + kMacro * (T + 1));
    uint32_t(*colbuf)[T] = rowbuf + 2;
    const uint32_t t = threadIdx.x;
    const bool row_panel = index < nb;
    const uint32_t c0 = (row_panel ? index : index - nb) * kMacro;

    if (index == 0 && pivot_prev != nullptr) {
        for (uint32_t q = t; q < T * T; q += kThreads)
            W[(size_t)(kb_prev + q / T) * ldw + kb_prev + q % T] = (E)pivot_prev[q];
    }
    // stage the chunk's old values while the pivot closes
    if (row_panel)
        load_row_chunk<E, T>(W, ldw, kb, c0, S);
    else
        load_col_chunk<E, T>(W, ldw, kb, c0, S);
    close_pivot<E, T>(W, ldw, kb, T, P, rowbuf, colbuf);
    if (index == 0)
        for (uint32_t q = t; q < T * T; q += kThreads) pivot_out[q] = P[q / T][q % T];
    if (row_panel)
        row_chunk_product<E, T>(W, ldw, kb, c0, P, S);
    else
        col_chunk_product<E, T>(W, ldw, kb, c0, P, S);
}

template <typename E, uint32_t T>
__global__ __launch_bounds__(256) void pivot_panel_kernel(E* __restrict__ W, size_t ldw, uint32_t kb,
                                                          uint32_t nb, uint32_t* __restrict__ pivot_out,
                                                          const uint32_t* __restrict__ pivot_prev,
                                                          uint32_t kb_prev) {
    __shared__ uint32_t lds[pivot_panel_lds_words<T>()];
    pivot_panel<E, T>(W, ldw, kb, nb, blockIdx.x, pivot_out, pivot_prev, kb_prev, lds);
}

// Copies the last round's closed pivot tile into W.
template <typename E, uint32_t T>
__global__ __launch_bounds__(256) void commit_pivot_kernel(E* __restrict__ W, size_t ldw, uint32_t kb,
                                                           const uint32_t* __restrict__ pivot) {
    for (uint32_t q = threadIdx.x; q < T * T; q += kThreads)
        W[(size_t)(kb + q / T) * ldw + kb + q % T] = (E)pivot[q];
}

// ---------------------------------------------------------------------------------------
// Independent tiles: one 128 x 128 macro tile per workgroup (indep_tile_u16 / _u32).

// Tile selection: which = 0 -> every macro tile (grid nb x nb); which = 1 -> the tiles of
// macro row/column `mnext`, i.e. the strips of the next pivot block (grid 2 nb - 1);
// which = 2 -> every other tile (grid nb x nb, strip tiles exit at once).
__device__ __forceinline__ bool select_tile(uint32_t which, uint32_t mnext, uint32_t nb, uint32_t* bi,
                                            uint32_t* bj) {
    if (which == 1) {
        const uint32_t t = blockIdx.x;
        if (t < nb) {
            *bi = mnext;
            *bj = t;
        } else {
            const uint32_t r = t - nb;
            *bi = r < mnext ? r : r + 1;
            *bj = mnext;
        }
        return true;
    }
    *bi = blockIdx.y;
    *bj = blockIdx.x;
    return !(which == 2 && (*bi == mnext || *bj == mnext));
}

template <uint32_t T, bool F16>
__global__ __launch_bounds__(256) void indep_u16_kernel(uint16_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                        uint32_t which, uint32_t mnext, uint32_t nb) {
    __shared__ __align__(16) uint16_t lds[indep_u16_lds_elems<T>()];
    uint32_t bi, bj;
    if (!select_tile(which, mnext, nb, &bi, &bj)) return;
    indep_tile_u16<T, F16>(W, ldw, kb, bi * kMacro, bj * kMacro, lds);
}

template <uint32_t T>
__global__ __launch_bounds__(256) void indep_u32_kernel(uint32_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                        uint32_t which, uint32_t mnext, uint32_t nb) {
    __shared__ __align__(16) uint32_t lds[indep_u32_lds_elems<T>()];
    uint32_t bi, bj;
    if (!select_tile(which, mnext, nb, &bi, &bj)) return;
    indep_tile_u32<T>(W, ldw, kb, bi * kMacro, bj * kMacro, lds);
}

// 32-bit look-ahead launch, as panel_indep_u16_kernel: workgroups [0, npanel) are the panel
// of the pivot block at kb_next (2 nb chunks; npanel is 2 nb rounded up to a multiple of
// 8, the extra workgroups exit), the others the independent tiles of the pivot block at
// kb outside macro row/column mnext (the strip launch did those).  The panel reads and
// writes only macro row/column mnext and, in workgroup 0, commits the previous closed
// pivot into the block at kb, which the tiles neither read nor store.
template <uint32_t T>
__global__ __launch_bounds__(256) void panel_indep_u32_kernel(uint32_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                              uint32_t mnext, uint32_t nb, uint32_t npanel,
                                                              uint32_t kb_next, uint32_t* __restrict__ pivot_out,
                                                              const uint32_t* __restrict__ pivot_prev) {
    constexpr uint32_t kWords = indep_u32_lds_elems<T>() > pivot_panel_lds_words<T>() ? indep_u32_lds_elems<T>()
                                                                                      : pivot_panel_lds_words<T>();
    __shared__ __align__(16) uint32_t lds[kWords];
    if (blockIdx.x < npanel) {
        if (blockIdx.x >= 2 * nb) return;
        __builtin_amdgcn_s_setprio(3);  // the round's critical path, as in the 16-bit launch
        pivot_panel<uint32_t, T>(W, ldw, kb_next, nb, blockIdx.x, pivot_out, pivot_prev, kb, lds);
        return;
    }
    const uint32_t t = blockIdx.x - npanel;
    const uint32_t bi = t / nb, bj = t % nb;
    if (bi == mnext || bj == mnext) return;  // strip tiles ran in the strip launch
    indep_tile_u32<T>(W, ldw, kb, bi * kMacro, bj * kMacro, lds);
}

// Strip launch of the 16-bit look-ahead schedule in 64 x 64 quarter tiles: the 2 nb - 1
// macro tiles of macro row/column mnext, four quarters each (workgroup 4 t + q: quarter
// q = 2 qi + qj of strip tile t, numbered as select_tile(1, ...) numbers them).
template <uint32_t T, bool F16>
__global__ __launch_bounds__(256) void strip64_u16_kernel(uint16_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                          uint32_t mnext, uint32_t nb) {
    __shared__ __align__(16) uint16_t lds[indep64_u16_lds_elems<T>()];
    const uint32_t t = blockIdx.x >> 2, q = blockIdx.x & 3;
    uint32_t bi, bj;
    if (t < nb) {
        bi = mnext;
        bj = t;
    } else {
        const uint32_t r = t - nb;
        bi = r < mnext ? r : r + 1;
        bj = mnext;
    }
    indep_tile64_u16<T, F16>(W, ldw, kb, bi * kMacro + 64 * (q >> 1), bj * kMacro + 64 * (q & 1), lds);
}

// ---------------------------------------------------------------------------------------
// 16-bit look-ahead launch: workgroups [0, npanel) are the panel of the pivot block at
// kb_next (one row chunk + one column chunk each, panel_u16), the others the independent
// tiles of the pivot block at kb outside macro row/column mnext (the strip launch did
// those).  The two groups touch disjoint cells: the panel reads and writes only macro
// row/column mnext (final after the strip launch) and, in workgroup 0, commits the
// previous closed pivot into the pivot block at kb, which the independent tiles neither
// read nor store.  Panel workgroups come first, so they are dispatched first; npanel is
// rounded up to a multiple of 8 (the extra workgroups exit) and the tiles are numbered
// row-major over the full nb x nb grid (strip tiles exit), so tile (bi, bj) lands on XCD
// bj mod 8 when nb is a multiple of 8: each XCD's L2 serves the B strips of its columns.
// With a grid of npanel and nothing else the kernel is a plain panel launch.
//
// strips_here = 1 (T = 64): no strip launch ran.  The panel workgroups first give their
// row and column chunks (the next pivot's T rows and columns) this round's update (see
// panel_u16), and the tiles of macro row/column mnext run here too, with the next pivot's
// rows and columns masked (kx): every cell still has exactly one writer per launch.
template <uint32_t T, bool F16>
__global__ __launch_bounds__(256) void panel_indep_u16_kernel(uint16_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                              uint32_t mnext, uint32_t nb, uint32_t npanel,
                                                              uint32_t kb_next, uint32_t* __restrict__ pivot_out,
                                                              const uint32_t* __restrict__ pivot_prev,
                                                              uint32_t kb_prev, uint32_t strips_here) {
    static_assert(panel_u16_lds_elems<T>() <= indep_u16_lds_elems<T>(), "panel LDS must fit the tile LDS");
    __shared__ __align__(16) uint16_t lds[indep_u16_lds_elems<T>()];
    if (blockIdx.x < npanel) {
        if (blockIdx.x >= nb) return;
        // the panel is the round's critical path: its waves issue ahead of the tiles' waves
        // sharing their SIMDs (the tiles fill the issue slots the panel's barriers leave)
        __builtin_amdgcn_s_setprio(3);
        panel_u16<T, F16>(W, ldw, kb_next, blockIdx.x, pivot_out, pivot_prev, kb_prev, lds,
                          strips_here ? kb : kNoRange);
        return;
    }
    const uint32_t t = blockIdx.x - npanel;
    const uint32_t bi = t / nb, bj = t % nb;
    if (bi == mnext || bj == mnext) {
        if (!strips_here) return;  // strip tiles ran in the strip launch
        indep_tile_u16<T, F16>(W, ldw, kb, bi * kMacro, bj * kMacro, lds, kb_next);
        return;
    }
    indep_tile_u16<T, F16>(W, ldw, kb, bi * kMacro, bj * kMacro, lds);
}

// ---------------------------------------------------------------------------------------
// Small elections (n <= 128): the whole pipeline in one workgroup and one launch.  p0
// (scaling_elections.py:103-112) straight into registers, 32-bit (so no width decision and
// no host round trip), then Floyd-Warshall over the n pivots (:115-123), then the output
// with its zero diagonal.  Lane (ty, tx) holds rows ty + 16 r and columns tx + 16 c
// (r, c < 8).  Pivot k's row and column are published through LDS, double-buffered, so one
// barrier per pivot: neither changes during step k (min(p[i][k], p[k][k]) <= p[i][k]).
// The diagonal may pick up values meanwhile; like the reference's skipped diagonal it
// never feeds an off-diagonal cell, and the output zeroes it.
constexpr uint32_t kSmallN = 128;

template <typename D, typename O>
__global__ __launch_bounds__(256) void small_closure_kernel(const D* __restrict__ d, uint32_t n, size_t ld,
                                                            O* __restrict__ out, size_t ldo) {
    __shared__ uint32_t rowbuf[2][kSmallN], colbuf[2][kSmallN];
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    uint32_t cell[8][8];
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t i = ty + 16 * r, j = tx + 16 * c;
            uint32_t v = 0;
            if (i < n && j < n && i != j) {
                const uint32_t x = d[(size_t)i * ld + j];
                if (x > (uint32_t)d[(size_t)j * ld + i]) v = x;
            }
            cell[r][c] = v;
        }
    // pivot k = 16 q + kk: its row lives at r = q in the lanes with ty == kk, its column at
    // c = q in the lanes with tx == kk (register indices stay compile-time constants)
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) {
        for (uint32_t kk = 0; kk < 16; ++kk) {
            const uint32_t k = 16 * q + kk;
            if (k >= n) break;
            const uint32_t buf = k & 1;
            if (ty == kk) {
#pragma unroll
                for (uint32_t c = 0; c < 8; ++c) rowbuf[buf][tx + 16 * c] = cell[q][c];
            }
            if (tx == kk) {
#pragma unroll
                for (uint32_t r = 0; r < 8; ++r) colbuf[buf][ty + 16 * r] = cell[r][q];
            }
            __syncthreads();
            uint32_t a[8], b[8];
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r) a[r] = colbuf[buf][ty + 16 * r];
#pragma unroll
            for (uint32_t c = 0; c < 8; ++c) b[c] = rowbuf[buf][tx + 16 * c];
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
                for (uint32_t c = 0; c < 8; ++c) cell[r][c] = max(cell[r][c], min(a[r], b[c]));
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t i = ty + 16 * r, j = tx + 16 * c;
            if (i < n && j < n) out[(size_t)i * ldo + j] = (O)(i == j ? 0u : cell[r][c]);
        }
}

// The same for a 16-bit input (every value <= 65535, the host entry's 16-bit transit):
// each lane keeps its 8 x 8 cells as 32 packed pairs (columns tx + 32 h and tx + 32 h + 16)
// and updates two cells per v_pk_min_u16 / v_pk_max_u16, half the VALU work per pivot.  The
// pivot column is published with its value in both halves, the pivot row as the pairs the
// lanes hold, so every lane reads one word per register.
template <typename O>
__global__ __launch_bounds__(256) void small_closure16_kernel(const uint16_t* __restrict__ d, uint32_t n, size_t ld,
                                                              O* __restrict__ out, size_t ldo) {
    __shared__ uint32_t rowbuf[2][kSmallN / 2], colbuf[2][kSmallN];
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    u16x2 cell[8][4];
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
        for (uint32_t h = 0; h < 4; ++h) {
            const uint32_t i = ty + 16 * r;
            uint32_t v[2];
#pragma unroll
            for (uint32_t e = 0; e < 2; ++e) {
                const uint32_t j = tx + 32 * h + 16 * e;
                v[e] = 0;
                if (i < n && j < n && i != j) {
                    const uint32_t x = d[(size_t)i * ld + j];
                    if (x > (uint32_t)d[(size_t)j * ld + i]) v[e] = x;
                }
            }
            cell[r][h] = as_u16x2(v[0] | (v[1] << 16));
        }
    // pivot k = 16 q + kk: its row lives at r = q in the lanes with ty == kk, its column in
    // pair q / 2, half q % 2, in the lanes with tx == kk
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) {
        for (uint32_t kk = 0; kk < 16; ++kk) {
            const uint32_t k = 16 * q + kk;
            if (k >= n) break;
            const uint32_t buf = k & 1;
            if (ty == kk) {
#pragma unroll
                for (uint32_t h = 0; h < 4; ++h) rowbuf[buf][16 * h + tx] = as_u32(cell[q][h]);
            }
            if (tx == kk) {
#pragma unroll
                for (uint32_t r = 0; r < 8; ++r) {
                    const uint32_t v = (as_u32(cell[r][q >> 1]) >> (16 * (q & 1))) & 0xFFFFu;
                    colbuf[buf][ty + 16 * r] = v | (v << 16);
                }
            }
            __syncthreads();
            u16x2 a[8], b[4];
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r) a[r] = as_u16x2(colbuf[buf][ty + 16 * r]);
#pragma unroll
            for (uint32_t h = 0; h < 4; ++h) b[h] = as_u16x2(rowbuf[buf][16 * h + tx]);
#pragma unroll
            for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
                for (uint32_t h = 0; h < 4; ++h) cell[r][h] = pk_max(cell[r][h], pk_min(a[r], b[h]));
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
        for (uint32_t h = 0; h < 4; ++h)
#pragma unroll
            for (uint32_t e = 0; e < 2; ++e) {
                const uint32_t i = ty + 16 * r, j = tx + 32 * h + 16 * e;
                const uint32_t v = (as_u32(cell[r][h]) >> (16 * e)) & 0xFFFFu;
                if (i < n && j < n) out[(size_t)i * ldo + j] = (O)(i == j ? 0u : v);
            }
}

// ---------------------------------------------------------------------------------------
// Output: out[i][j] = W[i][j] widened to u32, 0 on the diagonal, padding dropped.

template <typename E, typename O>
__global__ __launch_bounds__(256) void output_kernel(const E* __restrict__ W, size_t ldw, uint32_t n,
                                                     O* __restrict__ out, size_t ldo, uint32_t bias) {
    const uint32_t i = blockIdx.x;
    const E* src = W + (size_t)i * ldw;
    O* dst = out + (size_t)i * ldo;
    for (uint32_t j = threadIdx.x; j < n; j += kThreads) {
        const uint32_t w = (uint32_t)src[j];
        dst[j] = (O)((i == j || w == 0) ? 0u : w - bias);
    }
}

__global__ void init_range_kernel(uint32_t* p) {
    if (threadIdx.x < kRangeSlots) {
        p[threadIdx.x * kRangePitch] = 0u;
        p[threadIdx.x * kRangePitch + 1] = 0xFFFFFFFFu;
    }
}

// ---------------------------------------------------------------------------------------
// Host driver.

namespace {

struct EventTimer {
    bool on;
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    EventTimer(bool enabled, hipStream_t stream) : on(enabled && stream != nullptr), s(stream) {}
    ~EventTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    // marks the start (even index) or end (odd index) of a kernel of `k`
    void mark(int k) {
        if (!on) return;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
        (void)hipEventRecord(e, s);
        ev.push_back(e);
        kind.push_back(k);
    }
    // sums elapsed times per kind, pairing consecutive marks
    void collect(double* sums) {
        if (!on) return;
        for (size_t i = 0; i + 1 < ev.size(); i += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) sums[kind[i]] += ms;
        }
    }
};

enum { kKindInit = 0, kKindDiag = 1, kKindPanel = 2, kKindIndep = 3, kKindOut = 4 };

// Workgroups of the p0 passes: one per tile pair up to 4096 (16 per CU), grid-stride beyond.
uint32_t init_grid(uint32_t nt) {
    const uint64_t pairs = (uint64_t)nt * (nt + 1) / 2;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(pairs, 4096));
}

template <typename E, uint32_t T, bool F16>
void launch_indep(E* W, size_t ldw, uint32_t kb, uint32_t which, uint32_t mnext, uint32_t nb, hipStream_t s) {
    const dim3 grid = which == 1 ? dim3(2 * nb - 1) : dim3(nb, nb);
    if constexpr (sizeof(E) == 2)
        indep_u16_kernel<T, F16><<<grid, kThreads, 0, s>>>(W, ldw, kb, which, mnext, nb);
    else
        indep_u32_kernel<T><<<grid, kThreads, 0, s>>>(W, ldw, kb, which, mnext, nb);
}

// Rounds of the 32-bit closure, all on stream s.  With look-ahead (kLaStrips, >= 2
// rounds) round r is two launches: the tiles of the next pivot's strips (macro row/column
// mnext), then panel_indep_u32_kernel holding the next round's panel and the rest of round
// r's tiles, so the panel overlaps the bulk of the round.  Without it (kLaOff): all
// independent tiles, then the next panel.
template <uint32_t T>
int run_rounds32(uint32_t* W, size_t ldw, uint32_t n, uint32_t N, uint32_t* pivot_buf, hipStream_t s,
                 EventTimer& tm, uint32_t la_mode, uint32_t* rounds_out) {
    const uint32_t nb = N / kMacro;
    const uint32_t rounds = (n + T - 1) / T;
    auto pivot_of = [&](uint32_t r) { return pivot_buf + (r & 1) * T * T; };
    if (rounds == 0) {
        *rounds_out = 0;
        return SE_OK;
    }
    const uint32_t npanel = (2 * nb + 7) / 8 * 8;
    tm.mark(kKindPanel);
    pivot_panel_kernel<uint32_t, T><<<2 * nb, kThreads, 0, s>>>(W, ldw, 0, nb, pivot_of(0), nullptr, 0);
    tm.mark(kKindPanel);
    for (uint32_t r = 0; r < rounds; ++r) {
        if (g_interrupted.load(std::memory_order_relaxed)) {
            set_error("Stopped by signal");
            return SE_ERR_INTERRUPTED;
        }
        const uint32_t kb = r * T;
        const bool next = r + 1 < rounds;
        if (!next || la_mode == kLaOff) {
            tm.mark(kKindIndep);
            launch_indep<uint32_t, T, false>(W, ldw, kb, 0, 0, nb, s);
            tm.mark(kKindIndep);
            if (next) {
                tm.mark(kKindPanel);
                pivot_panel_kernel<uint32_t, T>
                    <<<2 * nb, kThreads, 0, s>>>(W, ldw, kb + T, nb, pivot_of(r + 1), pivot_of(r), kb);
                tm.mark(kKindPanel);
            }
        } else {
            const uint32_t mnext = (kb + T) / kMacro;
            tm.mark(kKindIndep);
            launch_indep<uint32_t, T, false>(W, ldw, kb, 1, mnext, nb, s);
            panel_indep_u32_kernel<T><<<npanel + nb * nb, kThreads, 0, s>>>(W, ldw, kb, mnext, nb, npanel, kb + T,
                                                                             pivot_of(r + 1), pivot_of(r));
            tm.mark(kKindIndep);
        }
        SE_CHECK_LAUNCH("closure round launch failed");
    }
    tm.mark(kKindPanel);
    commit_pivot_kernel<uint32_t, T><<<1, kThreads, 0, s>>>(W, ldw, (rounds - 1) * T, pivot_of(rounds - 1));
    tm.mark(kKindPanel);
    SE_CHECK_LAUNCH("closure round launch failed");
    *rounds_out = rounds;
    return SE_OK;
}

// Rounds of the 16-bit closure, by look-ahead mode:
//   kLaStrips: round r is two launches, the strip tiles of the next pivot (macro
//     row/column mnext), then one launch holding the next round's panel and the rest of
//     round r's independent tiles (panel_indep_u16_kernel);
//   kLaFused (T = 64): one launch per round, the panel workgroups giving the next pivot's
//     strips round r's update themselves (no strip launch);
//   kLaOff: all independent tiles, then the next panel.
template <uint32_t T, bool F16>
int run_rounds16(uint16_t* W, size_t ldw, uint32_t n, uint32_t N, uint32_t* pivot_buf, hipStream_t s,
                 EventTimer& tm, uint32_t la_mode, uint32_t* rounds_out) {
    const uint32_t nb = N / kMacro;
    const uint32_t rounds = (n + T - 1) / T;
    auto pivot_of = [&](uint32_t r) { return pivot_buf + (r & 1) * T * T; };
    if (rounds == 0) {
        *rounds_out = 0;
        return SE_OK;
    }
    const uint32_t npanel = (nb + 7) / 8 * 8;
    tm.mark(kKindPanel);
    panel_indep_u16_kernel<T, F16><<<npanel, kThreads, 0, s>>>(W, ldw, 0, 0, nb, npanel, 0, pivot_of(0), nullptr, 0,
                                                                0u);
    tm.mark(kKindPanel);
    for (uint32_t r = 0; r < rounds; ++r) {
        if (g_interrupted.load(std::memory_order_relaxed)) {
            set_error("Stopped by signal");
            return SE_ERR_INTERRUPTED;
        }
        const uint32_t kb = r * T;
        const bool next = r + 1 < rounds;
        if (!next) {
            tm.mark(kKindIndep);
            launch_indep<uint16_t, T, F16>(W, ldw, kb, 0, 0, nb, s);
            tm.mark(kKindIndep);
        } else if (la_mode == kLaOff) {
            tm.mark(kKindIndep);
            launch_indep<uint16_t, T, F16>(W, ldw, kb, 0, 0, nb, s);
            tm.mark(kKindIndep);
            tm.mark(kKindPanel);
            panel_indep_u16_kernel<T, F16><<<npanel, kThreads, 0, s>>>(W, ldw, kb, 0, nb, npanel, kb + T,
                                                                        pivot_of(r + 1), pivot_of(r), kb, 0u);
            tm.mark(kKindPanel);
        } else if (T == 64 && la_mode == kLaFused) {
            const uint32_t mnext = (kb + T) / kMacro;
            tm.mark(kKindIndep);
            panel_indep_u16_kernel<T, F16><<<npanel + nb * nb, kThreads, 0, s>>>(
                W, ldw, kb, mnext, nb, npanel, kb + T, pivot_of(r + 1), pivot_of(r), kb, 1u);
            tm.mark(kKindIndep);
        } else {
            const uint32_t mnext = (kb + T) / kMacro;
            tm.mark(kKindIndep);
            strip64_u16_kernel<T, F16><<<4 * (2 * nb - 1), kThreads, 0, s>>>(W, ldw, kb, mnext, nb);
            panel_indep_u16_kernel<T, F16><<<npanel + nb * nb, kThreads, 0, s>>>(
                W, ldw, kb, mnext, nb, npanel, kb + T, pivot_of(r + 1), pivot_of(r), kb, 0u);
            tm.mark(kKindIndep);
        }
        SE_CHECK_LAUNCH("closure round launch failed");
    }
    tm.mark(kKindPanel);
    commit_pivot_kernel<uint16_t, T><<<1, kThreads, 0, s>>>(W, ldw, (rounds - 1) * T, pivot_of(rounds - 1));
    tm.mark(kKindPanel);
    SE_CHECK_LAUNCH("closure round launch failed");
    *rounds_out = rounds;
    return SE_OK;
}

}  // namespace

// Device pipeline: d (device, n x n, row stride ld, u32 or, with d16, u16) -> out
// (device, n x n, ldo).  out is u32, except that with out16 != nullptr and a 16-bit
// closure it receives u16 values (every strongest path is then at most 65535) and
// *out16 is set; the host entry points use both to halve their PCIe traffic.
int closure_device_ex(DeviceContext* ctx, const void* d_any, bool d16, uint32_t n, size_t ld, void* out_any,
                      size_t ldo, bool* out16, uint32_t flags, hipStream_t s, se_timing* timing) {
    if (out16) *out16 = false;
    if (n == 0) {
        if (timing) *timing = se_timing{};
        return SE_OK;
    }
    const uint32_t* d32 = static_cast<const uint32_t*>(d_any);
    const uint16_t* d16p = static_cast<const uint16_t*>(d_any);
    const uint32_t N = (uint32_t)round_up(n, kMacro);
    const size_t ldw = N;
    const bool want_timing = timing != nullptr;
    EventTimer tm((flags & SE_FLAG_TIME_KERNELS) != 0, s);

    // up to 128 candidates: one launch (small_closure_kernel), unless a flag asks for a
    // particular width, pivot depth or schedule of the blocked path
    constexpr uint32_t kShapeFlags = SE_FLAG_FORCE_U32 | SE_FLAG_FORCE_U16 | SE_FLAG_PIVOT32 | SE_FLAG_PIVOT64 |
                                     SE_FLAG_NO_F16 | SE_FLAG_NO_LOOKAHEAD | SE_FLAG_LOOKAHEAD |
                                     SE_FLAG_LOOKAHEAD_FUSED;
    if (n <= kSmallN && !(flags & kShapeFlags)) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (want_timing && (!hip_ok(hipEventCreate(&e0), "Failed to create events") ||
                            !hip_ok(hipEventCreate(&e1), "Failed to create events")))
            return SE_ERR_DEVICE;
        if (want_timing) (void)hipEventRecord(e0, s);
        // values of a 16-bit input fit the 16-bit output (p takes values of d only)
        if (d16 && out16) {
            small_closure16_kernel<uint16_t>
                <<<1, kThreads, 0, s>>>(d16p, n, ld, static_cast<uint16_t*>(out_any), ldo);
            *out16 = true;
        } else if (d16) {
            small_closure16_kernel<uint32_t>
                <<<1, kThreads, 0, s>>>(d16p, n, ld, static_cast<uint32_t*>(out_any), ldo);
        } else {
            small_closure_kernel<uint32_t, uint32_t>
                <<<1, kThreads, 0, s>>>(d32, n, ld, static_cast<uint32_t*>(out_any), ldo);
        }
        SE_CHECK_LAUNCH("closure launch failed");
        if (want_timing) {
            (void)hipEventRecord(e1, s);
            if (!hip_ok(hipEventSynchronize(e1), "CUDA operations did not complete successfully"))
                return SE_ERR_DEVICE;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            se_timing t{};
            t.total_ms = ms;
            t.indep_ms = ms;
            t.rounds = n;  // one pivot per step, no blocking
            t.pivot = 1;
            t.width_bits = 32;
            t.padded_n = n;
            t.indep_cells = (uint64_t)n * n * n;
            *timing = t;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return SE_OK;
    }

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (want_timing) {
        if (!hip_ok(hipEventCreate(&e0), "Failed to create events") ||
            !hip_ok(hipEventCreate(&e1), "Failed to create events"))
            return SE_ERR_DEVICE;
        (void)hipEventRecord(e0, s);
    }

    // width decision: exact criterion max(p0) <= 65535; the 16-bit path runs f16-ordered
    // when the nonzero p0 values span at most 0x7C00 - 0x0400 (offset into positive normals)
    bool use16, f16 = false;
    uint32_t bias = 0;
    if (flags & SE_FLAG_FORCE_U32) {
        use16 = false;
    } else {
        uint32_t* dmax = static_cast<uint32_t*>(scratch(ctx, kSlotSmall, kRangeSlots * kRangePitch * 4));
        if (!dmax) return SE_ERR_ALLOC;
        // read back through pinned memory: a pageable destination goes through the
        // runtime's own staging and takes tens of microseconds longer
        static_assert(kRangeSlots * kRangePitch * sizeof(uint32_t) <= kSmallPinnedBytes, "range slots");
        uint32_t* slots = small_pinned(ctx);
        if (!slots) return SE_ERR_ALLOC;
        tm.mark(kKindInit);
        init_range_kernel<<<1, 64, 0, s>>>(dmax);
        const uint32_t ntn = (n + kInitTile - 1) / kInitTile;
        if (d16)
            p0_max_kernel<uint16_t><<<init_grid(ntn), kThreads, 0, s>>>(d16p, n, ld, ntn, dmax);
        else
            p0_max_kernel<uint32_t><<<init_grid(ntn), kThreads, 0, s>>>(d32, n, ld, ntn, dmax);
        tm.mark(kKindInit);
        SE_CHECK_LAUNCH("p0 init launch failed");
        if (!hip_ok(hipMemcpyAsync(slots, dmax, kRangeSlots * kRangePitch * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                    "Failed to copy data from device to host") ||
            !hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully"))
            return SE_ERR_COPY;
        uint32_t range[2] = {0u, 0xFFFFFFFFu};
        for (uint32_t q = 0; q < kRangeSlots; ++q) {
            range[0] = std::max(range[0], slots[q * kRangePitch]);
            range[1] = std::min(range[1], slots[q * kRangePitch + 1]);
        }
        use16 = range[0] <= 0xFFFFu;
        if ((flags & SE_FLAG_FORCE_U16) && !use16) {
            set_error("strongest-path values exceed 65535; the 16-bit closure is not exact here");
            return SE_ERR_RANGE;
        }
        if (use16 && !(flags & SE_FLAG_NO_F16) && range[1] != 0xFFFFFFFFu &&
            range[0] - range[1] <= 0x7C00u - 0x0400u) {
            f16 = true;
            bias = 0x0400u - range[1];  // modulo 2^32: v + bias = v - min + 0x0400
        }
    }

    const size_t esz = use16 ? 2 : 4;
    void* W = scratch(ctx, kSlotWork, (size_t)N * ldw * esz);
    if (!W) return SE_ERR_ALLOC;

    const uint32_t ntN = N / kInitTile;
    tm.mark(kKindInit);
    if (use16 && d16)
        p0_write_kernel<uint16_t, uint16_t>
            <<<init_grid(ntN), kThreads, 0, s>>>(d16p, n, ld, static_cast<uint16_t*>(W), ldw, ntN, bias);
    else if (use16)
        p0_write_kernel<uint16_t, uint32_t>
            <<<init_grid(ntN), kThreads, 0, s>>>(d32, n, ld, static_cast<uint16_t*>(W), ldw, ntN, bias);
    else if (d16)
        p0_write_kernel<uint32_t, uint16_t>
            <<<init_grid(ntN), kThreads, 0, s>>>(d16p, n, ld, static_cast<uint32_t*>(W), ldw, ntN, 0u);
    else
        p0_write_kernel<uint32_t, uint32_t>
            <<<init_grid(ntN), kThreads, 0, s>>>(d32, n, ld, static_cast<uint32_t*>(W), ldw, ntN, 0u);
    tm.mark(kKindInit);
    SE_CHECK_LAUNCH("p0 init launch failed");

    // pivot depth and look-ahead from the tuning table: deeper pivots halve the HBM traffic
    // per cell update of the independent kernel (2 s / T bytes per update) and the number
    // of rounds, shallower ones shorten each round's dependent chain (pivot closure and
    // panels), which wins while rounds are short.  Flags override the table.
    uint32_t T = 64, la_mode = kLaStrips;
    tuned_closure_shape(N, use16 ? (f16 ? SE_TUNE_F16 : SE_TUNE_U16) : SE_TUNE_U32, &T, &la_mode);
    if (flags & SE_FLAG_PIVOT32) T = 32;
    if (flags & SE_FLAG_PIVOT64) T = 64;
    uint32_t* pivot_buf = static_cast<uint32_t*>(scratch(ctx, kSlotPivot, 2 * 64 * 64 * sizeof(uint32_t)));
    if (!pivot_buf) return SE_ERR_ALLOC;
    uint32_t rounds = 0;
    int rc;
    uint16_t* W16 = static_cast<uint16_t*>(W);
    uint32_t* W32 = static_cast<uint32_t*>(W);
    // look-ahead (the table's mode, or a flag): the next panel runs inside the launch of
    // the current round's remaining tiles, on the one stream s (16-bit mode 2 also drops
    // the strip launch).
    if (flags & SE_FLAG_NO_LOOKAHEAD)
        la_mode = kLaOff;
    else if (flags & SE_FLAG_LOOKAHEAD_FUSED)
        la_mode = kLaFused;
    else if (flags & SE_FLAG_LOOKAHEAD)
        la_mode = kLaStrips;
    if (la_mode == kLaFused && (!use16 || T != 64)) la_mode = kLaStrips;  // 16-bit, pivot 64 only
    if (use16) {
        if (f16)
            rc = T == 64 ? run_rounds16<64, true>(W16, ldw, n, N, pivot_buf, s, tm, la_mode, &rounds)
                         : run_rounds16<32, true>(W16, ldw, n, N, pivot_buf, s, tm, la_mode, &rounds);
        else
            rc = T == 64 ? run_rounds16<64, false>(W16, ldw, n, N, pivot_buf, s, tm, la_mode, &rounds)
                         : run_rounds16<32, false>(W16, ldw, n, N, pivot_buf, s, tm, la_mode, &rounds);
    } else {
        rc = T == 64 ? run_rounds32<64>(W32, ldw, n, N, pivot_buf, s, tm, la_mode, &rounds)
                     : run_rounds32<32>(W32, ldw, n, N, pivot_buf, s, tm, la_mode, &rounds);
    }
    if (rc != SE_OK) return rc;

    dim3 go(n);
    tm.mark(kKindOut);
    if (use16 && out16) {
        output_kernel<uint16_t, uint16_t>
            <<<go, kThreads, 0, s>>>(static_cast<uint16_t*>(W), ldw, n, static_cast<uint16_t*>(out_any), ldo, bias);
        *out16 = true;
    } else if (use16) {
        output_kernel<uint16_t, uint32_t>
            <<<go, kThreads, 0, s>>>(static_cast<uint16_t*>(W), ldw, n, static_cast<uint32_t*>(out_any), ldo, bias);
    } else {
        output_kernel<uint32_t, uint32_t>
            <<<go, kThreads, 0, s>>>(static_cast<uint32_t*>(W), ldw, n, static_cast<uint32_t*>(out_any), ldo, 0u);
    }
    tm.mark(kKindOut);
    SE_CHECK_LAUNCH("output launch failed");

    if (want_timing) {
        (void)hipEventRecord(e1, s);
        if (!hip_ok(hipEventSynchronize(e1), "CUDA operations did not complete successfully")) return SE_ERR_DEVICE;
        se_timing t{};
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t.total_ms = ms;
        double sums[5] = {0, 0, 0, 0, 0};
        tm.collect(sums);
        t.init_ms = sums[kKindInit];
        t.diag_ms = sums[kKindDiag];
        t.panel_ms = sums[kKindPanel];
        t.indep_ms = sums[kKindIndep];
        t.output_ms = sums[kKindOut];
        t.rounds = rounds;
        t.pivot = T;
        t.width_bits = use16 ? 16 : 32;
        t.f16_ordered = f16 ? 1u : 0u;
        t.padded_n = N;
        t.indep_cells = (uint64_t)rounds * (uint64_t)(N - T) * (uint64_t)(N - T) * T;
        *timing = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return SE_OK;
}

int closure_device_impl(DeviceContext* ctx, const uint32_t* d, uint32_t n, size_t ld, uint32_t* out,
                        size_t ldo, uint32_t flags, hipStream_t s, se_timing* timing) {
    return closure_device_ex(ctx, d, false, n, ld, out, ldo, nullptr, flags, s, timing);
}

}  // namespace se
