// closure_tiles.h -- workgroup-level building blocks of the blocked max-min closure,
// shared by the single-election kernels (closure.hip) and the batched kernels (batch.hip).
// Every function is called by all 256 lanes of a workgroup with uniform arguments.
//
// Semantics (reference: scaling_elections.py:115-123): a cell update is
//     c = max(c, min(a, b)),
// evaluated without the reference's diagonal masks, which is exact: an update of p[i][j]
// through pivot i or j is min(p[i][i], p[i][j]) or min(p[i][j], p[j][j]) and never raises
// p[i][j]; the diagonal is zeroed when results are written out.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace se {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kMacro = 128;  // macro tile side of the independent kernel
constexpr uint32_t kThreads = 256;

__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u32(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ u16x2 pk_min(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ h16x2 as_h16x2(uint32_t x) { return __builtin_bit_cast(h16x2, x); }
__device__ __forceinline__ uint32_t h_as_u32(h16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ u16x2 pk_max(u16x2 a, u16x2 b) { return __builtin_elementwise_max(a, b); }

__device__ __forceinline__ uint32_t word(const uint4& v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}


// ---------------------------------------------------------------------------------------
// Pivot tile closure, in registers.  Lane t owns an R x R block (R = T / 16) at rows
// R (t / 16), columns R (t % 16).  At step k the owners of row k and column k publish them
// through LDS (double-buffered by k parity, so one barrier per step suffices) and every lane
// updates its block.  The cells read in step k cannot change in step k
// (min(p[i][k], p[k][k]) <= p[i][k]), so the step is race-free.  Only the first `steps`
// pivots run (pivots past the last real candidate are no-ops on zero padding).  The closed
// tile lands in P (LDS); the function ends with a barrier.
template <typename E, uint32_t T>
__device__ __forceinline__ void close_pivot(const E* __restrict__ W, size_t ldw, uint32_t kb, uint32_t steps,
                                            uint32_t (*P)[T + 1], uint32_t (*rowbuf)[T],
                                            uint32_t (*colbuf)[T]) {
    constexpr uint32_t R = T / 16;
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    const uint32_t pr = R * ty, pc = R * tx;
    uint32_t cell[R][R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) cell[r][c] = W[(size_t)(kb + pr + r) * ldw + kb + pc + c];
#pragma unroll
    for (uint32_t k = 0; k < T; ++k) {
        if (k >= steps) break;
        const uint32_t kr = k % R, buf = k & 1;
        if (ty == k / R) {
#pragma unroll
            for (uint32_t c = 0; c < R; ++c) rowbuf[buf][pc + c] = cell[kr][c];
        }
        if (tx == k / R) {
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) colbuf[buf][pr + r] = cell[r][kr];
        }
        __syncthreads();
        uint32_t a[R], b[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) a[r] = colbuf[buf][pr + r];
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) b[c] = rowbuf[buf][pc + c];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r)
#pragma unroll
            for (uint32_t c = 0; c < R; ++c) cell[r][c] = max(cell[r][c], min(a[r], b[c]));
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) P[pr + r][pc + c] = cell[r][c];
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Panels.  With the pivot tile P closed, the row panel (K, J) becomes, in one shot (split
// any path at its last vertex in K),
//     new[i][j] = max(old[i][j], max_k min(P[i][k], old[k][j])),
// and the column panel (I, K)
//     new[i][j] = max(old[i][j], max_k min(old[i][k], P[k][j])).
// A chunk is 128 columns of the row panel or 128 rows of the column panel; cells inside
// (K, K) are never written.  S holds the chunk's old values (T x 128, or 128 x (T + 1)).

template <typename E, uint32_t T>
__device__ __forceinline__ void load_row_chunk(const E* __restrict__ W, size_t ldw, uint32_t kb, uint32_t c0,
                                               uint32_t* S) {
    for (uint32_t q = threadIdx.x; q < T * kMacro; q += kThreads) {
        uint32_t i = q / kMacro, j = q % kMacro;
        S[i * kMacro + j] = W[(size_t)(kb + i) * ldw + c0 + j];
    }
}

template <typename E, uint32_t T>
__device__ __forceinline__ void load_col_chunk(const E* __restrict__ W, size_t ldw, uint32_t kb, uint32_t c0,
                                               uint32_t* S) {
    for (uint32_t q = threadIdx.x; q < kMacro * T; q += kThreads) {
        uint32_t i = q / T, j = q % T;
        S[i * (T + 1) + j] = W[(size_t)(c0 + i) * ldw + kb + j];
    }
}

// lane: rows (T/16) ty .. of the chunk, columns 8 tx .. 8 tx + 7
template <typename E, uint32_t T>
__device__ __forceinline__ void row_chunk_product(E* __restrict__ W, size_t ldw, uint32_t kb, uint32_t c0,
                                                  const uint32_t (*P)[T + 1], const uint32_t* S) {
    constexpr uint32_t RP = T / 16;
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    uint32_t acc[RP][8];
#pragma unroll
    for (uint32_t r = 0; r < RP; ++r)
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) acc[r][c] = S[(RP * ty + r) * kMacro + 8 * tx + c];
#pragma unroll 4
    for (uint32_t k = 0; k < T; ++k) {
        uint32_t a[RP], b[8];
#pragma unroll
        for (uint32_t r = 0; r < RP; ++r) a[r] = P[RP * ty + r][k];
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) b[c] = S[k * kMacro + 8 * tx + c];
#pragma unroll
        for (uint32_t r = 0; r < RP; ++r)
#pragma unroll
            for (uint32_t c = 0; c < 8; ++c) acc[r][c] = max(acc[r][c], min(a[r], b[c]));
    }
    if (c0 + 8 * tx - kb >= T) {  // columns of the pivot block stay untouched
#pragma unroll
        for (uint32_t r = 0; r < RP; ++r)
#pragma unroll
            for (uint32_t c = 0; c < 8; ++c) W[(size_t)(kb + RP * ty + r) * ldw + c0 + 8 * tx + c] = (E)acc[r][c];
    }
}

// lane: rows 8 ty .. 8 ty + 7 of the chunk, columns (T/16) tx ..
template <typename E, uint32_t T>
__device__ __forceinline__ void col_chunk_product(E* __restrict__ W, size_t ldw, uint32_t kb, uint32_t c0,
                                                  const uint32_t (*P)[T + 1], const uint32_t* S) {
    constexpr uint32_t RP = T / 16;
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    uint32_t acc[8][RP];
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
        for (uint32_t c = 0; c < RP; ++c) acc[r][c] = S[(8 * ty + r) * (T + 1) + RP * tx + c];
#pragma unroll 4
    for (uint32_t k = 0; k < T; ++k) {
        uint32_t a[8], b[RP];
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) a[r] = S[(8 * ty + r) * (T + 1) + k];
#pragma unroll
        for (uint32_t c = 0; c < RP; ++c) b[c] = P[k][RP * tx + c];
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
            for (uint32_t c = 0; c < RP; ++c) acc[r][c] = max(acc[r][c], min(a[r], b[c]));
    }
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r) {
        const uint32_t i = c0 + 8 * ty + r;
        if (i - kb < T) continue;  // rows of the pivot block stay untouched
#pragma unroll
        for (uint32_t c = 0; c < RP; ++c) W[(size_t)i * ldw + kb + RP * tx + c] = (E)acc[r][c];
    }
}

// ---------------------------------------------------------------------------------------
// Independent tiles, 16-bit.  Workgroup = 256 lanes = 128 x 128 macro tile; lane (ty, tx)
// owns rows ty + 16 r (r < 8) and columns 8 tx .. 8 tx + 7, i.e. 8 x 4 packed u16 pairs.
// Rows are interleaved by 16 so the four row addresses a wave reads from the A strip fall
// in different LDS banks; the B strip row a wave reads is 256 contiguous bytes.
// Per pivot step: 32 v_pk_min_u16 + 32 v_pk_max_u16 for 64 cells, a[i][k] broadcast to
// both halves by op_sel, and two ds_read_b128 per lane per step.

template <uint32_t T>
constexpr uint32_t indep_u16_lds_elems() { return kMacro * (T + 8) + T * kMacro; }

// C(I0.., J0..) = max(C, A (x) B) on one 128 x 128 macro tile; As/Bs are LDS buffers of
// indep_u16_lds_elems<T>() u16 in total.  Starts with a barrier (LDS reuse between calls).
// F16 = true: the 16-bit values are bit patterns of non-negative f16 numbers (0 or within
// [0x0400, 0x7C00] after an order-preserving offset), whose float order equals their
// unsigned order; IEEE-754-2019 minimum/maximum return one operand unchanged, so two pivot
// steps fold into 2 v_pk_minimum3_f16 + 1 v_pk_maximum3_f16 (0.75 VALU per update).
// MASKED = false: the tile holds no pivot row or column (almost every tile), so loads and
// stores need no per-row checks.  kx (kNoRange: none) names a second block of T rows and
// columns of C that is neither read nor stored: the next pivot's rows and columns when the
// panel workgroups of the same launch update them (the strips are still read in full).
constexpr uint32_t kNoRange = 0xFFFFFFFFu;

template <uint32_t T, bool F16, bool MASKED>
__device__ __forceinline__ void indep_tile_u16_impl(uint16_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                    uint32_t I0, uint32_t J0, uint16_t* __restrict__ lds,
                                                    uint32_t kx = kNoRange) {
    constexpr uint32_t AP = T + 8;  // A strip pitch (u16), keeps 16-byte alignment
    uint16_t* As = lds;
    uint16_t* Bs = lds + kMacro * AP;
    const uint32_t t = threadIdx.x, tx = t & 15, ty = t >> 4;
    __syncthreads();

    // Loads: the strips first, then the C tile.  Memory loads retire in order, so the
    // strips' LDS stores wait for the strip loads only, and the C loads complete behind the
    // product A (x) B, which starts from zero (the least value) and is merged into C at the
    // end -- exact, max being associative.  Cells of the pivot block (K, K) only feed
    // pivot-row/column cells, which are not stored, so they are not read (zeros instead):
    // the next round's panel kernel may be committing that block concurrently; pivot
    // rows/columns of C are neither read nor stored either.
    constexpr uint32_t kAq = kMacro * (T / 8) / kThreads, kBq = T * 16 / kThreads;
    static_assert(kMacro * (T / 8) % kThreads == 0 && T * 16 % kThreads == 0, "strip chunks per lane");
    uint4 sa[kAq], sb[kBq];
#pragma unroll
    for (uint32_t i = 0; i < kAq; ++i) {
        const uint32_t q = t + i * kThreads, row = q / (T / 8), ch = q % (T / 8);
        sa[i] = make_uint4(0, 0, 0, 0);
        if (!MASKED || I0 + row - kb >= T)
            sa[i] = *reinterpret_cast<const uint4*>(W + (size_t)(I0 + row) * ldw + kb + ch * 8);
    }
#pragma unroll
    for (uint32_t i = 0; i < kBq; ++i) {
        const uint32_t q = t + i * kThreads, row = q >> 4, ch = q & 15;
        sb[i] = make_uint4(0, 0, 0, 0);
        if (!MASKED || J0 + ch * 8 - kb >= T)
            sb[i] = *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + J0 + ch * 8);
    }
    // C rows/columns that are neither read nor stored: this round's pivot block and kx
    auto skipped = [&](uint32_t x) { return x - kb < T || (kx != kNoRange && x - kx < T); };
    uint4 c[8];
    const uint16_t* cp = W + (size_t)(I0 + ty) * ldw + J0 + tx * 8;
    const bool col_in_pivot = MASKED && skipped(J0 + tx * 8);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if constexpr (MASKED) {
            c[r] = make_uint4(0, 0, 0, 0);
            if (!col_in_pivot && !skipped(I0 + ty + 16 * r))
                c[r] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw);
        } else {
            c[r] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw);
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kAq; ++i) {
        const uint32_t q = t + i * kThreads;
        *reinterpret_cast<uint4*>(As + (q / (T / 8)) * AP + (q % (T / 8)) * 8) = sa[i];
    }
#pragma unroll
    for (uint32_t i = 0; i < kBq; ++i) {
        const uint32_t q = t + i * kThreads;
        *reinterpret_cast<uint4*>(Bs + (q >> 4) * kMacro + (q & 15) * 8) = sb[i];
    }
    __syncthreads();

    u16x2 acc[8][4];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][q] = as_u16x2(0u);

    // A fragments 4 pivot steps at a time (8 B per row): 16 registers, which keeps the
    // kernel within 128 VGPRs (4 waves per SIMD) while C is held for the final merge
#pragma unroll 1
    for (uint32_t k0 = 0; k0 < T; k0 += 4) {
        uint2 a4[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            a4[r] = *reinterpret_cast<const uint2*>(As + (ty + 16 * r) * AP + k0);
        if constexpr (F16) {
#pragma unroll
            for (int kk = 0; kk < 4; kk += 2) {
                const uint4 b4 = *reinterpret_cast<const uint4*>(Bs + (k0 + kk) * kMacro + tx * 8);
                const uint4 b5 = *reinterpret_cast<const uint4*>(Bs + (k0 + kk + 1) * kMacro + tx * 8);
                const h16x2 b0[4] = {as_h16x2(b4.x), as_h16x2(b4.y), as_h16x2(b4.z), as_h16x2(b4.w)};
                const h16x2 b1[4] = {as_h16x2(b5.x), as_h16x2(b5.y), as_h16x2(b5.z), as_h16x2(b5.w)};
                // all minimums of two rows first, then their maximums: every maximum3 reads
                // results produced several instructions earlier (no dependency stalls)
#pragma unroll
                for (int r = 0; r < 8; r += 2) {
                    h16x2 m0[2][4], m1[2][4];
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr) {
                        const h16x2 a2 = as_h16x2((kk >> 1) ? a4[r + rr].y : a4[r + rr].x);
                        const h16x2 a0 = a2.xx, a1 = a2.yy;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            m0[rr][q] = __builtin_elementwise_minimum(a0, b0[q]);
                            m1[rr][q] = __builtin_elementwise_minimum(a1, b1[q]);
                        }
                    }
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const h16x2 cur = as_h16x2(as_u32(acc[r + rr][q]));
                            acc[r + rr][q] = as_u16x2(h_as_u32(__builtin_elementwise_maximum(
                                __builtin_elementwise_maximum(cur, m0[rr][q]), m1[rr][q])));
                        }
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const uint4 b4 = *reinterpret_cast<const uint4*>(Bs + (k0 + kk) * kMacro + tx * 8);
                const u16x2 b[4] = {as_u16x2(b4.x), as_u16x2(b4.y), as_u16x2(b4.z), as_u16x2(b4.w)};
#pragma unroll
                for (int r = 0; r < 8; r += 2) {
                    u16x2 mn[2][4];
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr) {
                        const u16x2 a2 = as_u16x2((kk >> 1) ? a4[r + rr].y : a4[r + rr].x);
                        const u16x2 av = (kk & 1) ? a2.yy : a2.xx;
#pragma unroll
                        for (int q = 0; q < 4; ++q) mn[rr][q] = pk_min(av, b[q]);
                    }
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[r + rr][q] = pk_max(acc[r + rr][q], mn[rr][q]);
                }
            }
        }
    }

    // merge the product into C
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if constexpr (F16)
                acc[r][q] = as_u16x2(h_as_u32(
                    __builtin_elementwise_maximum(as_h16x2(as_u32(acc[r][q])), as_h16x2(word(c[r], q)))));
            else
                acc[r][q] = pk_max(acc[r][q], as_u16x2(word(c[r], q)));
        }

    if (col_in_pivot) return;  // these columns belong to a pivot block
    uint16_t* op = W + (size_t)(I0 + ty) * ldw + J0 + tx * 8;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (MASKED && skipped(I0 + ty + 16 * r)) continue;  // pivot rows: final, or the panel's
        uint4 v;
        v.x = as_u32(acc[r][0]);
        v.y = as_u32(acc[r][1]);
        v.z = as_u32(acc[r][2]);
        v.w = as_u32(acc[r][3]);
        *reinterpret_cast<uint4*>(op + (size_t)(16 * r) * ldw) = v;
    }
}

template <uint32_t T, bool F16 = false>
__device__ __forceinline__ void indep_tile_u16(uint16_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t I0,
                                               uint32_t J0, uint16_t* __restrict__ lds, uint32_t kx = kNoRange) {
    auto apart = [&](uint32_t k) {
        return (I0 + kMacro <= k || I0 >= k + T) && (J0 + kMacro <= k || J0 >= k + T);
    };
    if (apart(kb) && (kx == kNoRange || apart(kx)))
        indep_tile_u16_impl<T, F16, false>(W, ldw, kb, I0, J0, lds);
    else
        indep_tile_u16_impl<T, F16, true>(W, ldw, kb, I0, J0, lds, kx);
}

// Quarter tiles (64 x 64) for the strip launch, which has only 2 nb - 1 macro tiles: four
// times as many workgroups keep four times as many CUs busy on that short launch.  Lane
// (ty, tx) owns rows ty + 16 r (r < 4) and columns 4 tx .. 4 tx + 3 (two packed pairs).
// LDS: A strip 64 x (T + 8) and B strip T x 64 u16 (indep64_u16_lds_elems<T>()).
template <uint32_t T>
constexpr uint32_t indep64_u16_lds_elems() { return 64 * (T + 8) + T * 64; }

// out_lds != nullptr: the updated quarter goes to LDS (pitch 64, every cell) instead of W.
template <uint32_t T, bool F16, bool MASKED>
__device__ __forceinline__ void indep_tile64_u16_impl(uint16_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                      uint32_t I0, uint32_t J0, uint16_t* __restrict__ lds,
                                                      uint16_t* __restrict__ out_lds) {
    constexpr uint32_t AP = T + 8;
    uint16_t* As = lds;
    uint16_t* Bs = lds + 64 * AP;
    const uint32_t t = threadIdx.x, tx = t & 15, ty = t >> 4;
    __syncthreads();

    // strips first, then C (merged after the product, as in indep_tile_u16); pivot
    // rows/columns neither read nor stored
    constexpr uint32_t kAq = 64 * (T / 8) / kThreads, kBq = T * 8 / kThreads;
    static_assert(64 * (T / 8) % kThreads == 0 && T * 8 % kThreads == 0, "strip chunks per lane");
    uint4 sa[kAq], sb[kBq];
#pragma unroll
    for (uint32_t i = 0; i < kAq; ++i) {
        const uint32_t q = t + i * kThreads, row = q / (T / 8), ch = q % (T / 8);
        sa[i] = make_uint4(0, 0, 0, 0);
        if (!MASKED || I0 + row - kb >= T)
            sa[i] = *reinterpret_cast<const uint4*>(W + (size_t)(I0 + row) * ldw + kb + ch * 8);
    }
#pragma unroll
    for (uint32_t i = 0; i < kBq; ++i) {
        const uint32_t q = t + i * kThreads, row = q >> 3, ch = q & 7;
        sb[i] = make_uint4(0, 0, 0, 0);
        if (!MASKED || J0 + ch * 8 - kb >= T)
            sb[i] = *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + J0 + ch * 8);
    }
    uint2 c[4];
    const uint16_t* cp = W + (size_t)(I0 + ty) * ldw + J0 + tx * 4;
    const bool col_in_pivot = MASKED && J0 + tx * 4 - kb < T;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if constexpr (MASKED) {
            c[r] = make_uint2(0, 0);
            if (!col_in_pivot && I0 + ty + 16 * r - kb >= T)
                c[r] = *reinterpret_cast<const uint2*>(cp + (size_t)(16 * r) * ldw);
        } else {
            c[r] = *reinterpret_cast<const uint2*>(cp + (size_t)(16 * r) * ldw);
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kAq; ++i) {
        const uint32_t q = t + i * kThreads;
        *reinterpret_cast<uint4*>(As + (q / (T / 8)) * AP + (q % (T / 8)) * 8) = sa[i];
    }
#pragma unroll
    for (uint32_t i = 0; i < kBq; ++i) {
        const uint32_t q = t + i * kThreads;
        *reinterpret_cast<uint4*>(Bs + (q >> 3) * 64 + (q & 7) * 8) = sb[i];
    }
    __syncthreads();

    u16x2 acc[4][2];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r][0] = acc[r][1] = as_u16x2(0u);
#pragma unroll 1
    for (uint32_t k0 = 0; k0 < T; k0 += 8) {
        uint4 a4[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) a4[r] = *reinterpret_cast<const uint4*>(As + (ty + 16 * r) * AP + k0);
#pragma unroll
        for (int kk = 0; kk < 8; kk += 2) {
            const uint2 b4 = *reinterpret_cast<const uint2*>(Bs + (k0 + kk) * 64 + tx * 4);
            const uint2 b5 = *reinterpret_cast<const uint2*>(Bs + (k0 + kk + 1) * 64 + tx * 4);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t a2w = word(a4[r], kk >> 1);
                if constexpr (F16) {
                    const h16x2 a2 = as_h16x2(a2w);
                    const h16x2 a0 = a2.xx, a1 = a2.yy;
                    const h16x2 m00 = __builtin_elementwise_minimum(a0, as_h16x2(b4.x));
                    const h16x2 m01 = __builtin_elementwise_minimum(a0, as_h16x2(b4.y));
                    const h16x2 m10 = __builtin_elementwise_minimum(a1, as_h16x2(b5.x));
                    const h16x2 m11 = __builtin_elementwise_minimum(a1, as_h16x2(b5.y));
                    acc[r][0] = as_u16x2(h_as_u32(__builtin_elementwise_maximum(
                        __builtin_elementwise_maximum(as_h16x2(as_u32(acc[r][0])), m00), m10)));
                    acc[r][1] = as_u16x2(h_as_u32(__builtin_elementwise_maximum(
                        __builtin_elementwise_maximum(as_h16x2(as_u32(acc[r][1])), m01), m11)));
                } else {
                    const u16x2 a2 = as_u16x2(a2w);
                    const u16x2 a0 = a2.xx, a1 = a2.yy;
                    acc[r][0] = pk_max(pk_max(acc[r][0], pk_min(a0, as_u16x2(b4.x))), pk_min(a1, as_u16x2(b5.x)));
                    acc[r][1] = pk_max(pk_max(acc[r][1], pk_min(a0, as_u16x2(b4.y))), pk_min(a1, as_u16x2(b5.y)));
                }
            }
        }
    }

#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if constexpr (F16) {
            acc[r][0] = as_u16x2(h_as_u32(__builtin_elementwise_maximum(as_h16x2(as_u32(acc[r][0])), as_h16x2(c[r].x))));
            acc[r][1] = as_u16x2(h_as_u32(__builtin_elementwise_maximum(as_h16x2(as_u32(acc[r][1])), as_h16x2(c[r].y))));
        } else {
            acc[r][0] = pk_max(acc[r][0], as_u16x2(c[r].x));
            acc[r][1] = pk_max(acc[r][1], as_u16x2(c[r].y));
        }
    }

    if (out_lds != nullptr) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            *reinterpret_cast<uint2*>(out_lds + (ty + 16 * r) * 64 + tx * 4) =
                make_uint2(as_u32(acc[r][0]), as_u32(acc[r][1]));
        return;
    }
    if (MASKED && J0 + tx * 4 - kb < T) return;
    uint16_t* op = W + (size_t)(I0 + ty) * ldw + J0 + tx * 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (MASKED && I0 + ty + 16 * r - kb < T) continue;
        *reinterpret_cast<uint2*>(op + (size_t)(16 * r) * ldw) = make_uint2(as_u32(acc[r][0]), as_u32(acc[r][1]));
    }
}

template <uint32_t T, bool F16>
__device__ __forceinline__ void indep_tile64_u16(uint16_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t I0,
                                                 uint32_t J0, uint16_t* __restrict__ lds,
                                                 uint16_t* __restrict__ out_lds = nullptr) {
    const bool clean = (I0 + 64 <= kb || I0 >= kb + T) && (J0 + 64 <= kb || J0 >= kb + T);
    if (clean)
        indep_tile64_u16_impl<T, F16, false>(W, ldw, kb, I0, J0, lds, out_lds);
    else
        indep_tile64_u16_impl<T, F16, true>(W, ldw, kb, I0, J0, lds, out_lds);
}

// Independent tiles, 32-bit.  Lane (ty, tx) owns rows ty + 16 r (r < 8) and columns
// 4 tx .. 4 tx + 3 and 64 + 4 tx .. 64 + 4 tx + 3 (each half is 256 contiguous bytes per
// row across 16 lanes, so B-strip reads and global accesses stay conflict-free and
// coalesced).  Two pivot steps fold into one v_max3_u32: 1.5 VALU per cell update.

template <uint32_t T>
constexpr uint32_t indep_u32_lds_elems() { return kMacro * (T + 4) + T * kMacro; }

// LATE (pivot 64, where LDS already limits the kernel to 2 waves per SIMD and 256 VGPRs):
// strips loaded first, C loaded behind them and merged after the product, as in
// indep_tile_u16_impl; with pivot 32 (4 waves, 128 VGPRs) C seeds the accumulators.
// MASKED = false: no pivot row or column in the tile, so no per-row checks (pivot 64).
template <uint32_t T, bool MASKED>
__device__ __forceinline__ void indep_tile_u32_impl(uint32_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                    uint32_t I0, uint32_t J0, uint32_t* __restrict__ lds) {
    constexpr bool LATE = T == 64;
    constexpr uint32_t AP = T + 4;  // A strip pitch (u32)
    uint32_t* As = lds;
    uint32_t* Bs = lds + kMacro * AP;
    const uint32_t t = threadIdx.x, tx = t & 15, ty = t >> 4;
    __syncthreads();

    // C tile without its pivot rows/columns (see indep_tile_u16_impl)
    uint4 c[8][2];
    const uint32_t* cp = W + (size_t)(I0 + ty) * ldw + J0 + tx * 4;
    const bool lo_in_pivot = MASKED && J0 + tx * 4 - kb < T, hi_in_pivot = MASKED && J0 + 64 + tx * 4 - kb < T;
    auto load_c = [&]() {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const bool row_ok = !MASKED || I0 + ty + 16 * r - kb >= T;
            c[r][0] = make_uint4(0, 0, 0, 0);
            c[r][1] = make_uint4(0, 0, 0, 0);
            if (row_ok && !lo_in_pivot) c[r][0] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw);
            if (row_ok && !hi_in_pivot) c[r][1] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw + 64);
        }
    };
    if constexpr (LATE) {
        // strips through registers, pivot block as zeros; C behind them
        constexpr uint32_t kAq = kMacro * (T / 4) / kThreads, kBq = T * 32 / kThreads;
        static_assert(kMacro * (T / 4) % kThreads == 0 && T * 32 % kThreads == 0, "strip chunks per lane");
        uint4 sa[kAq], sb[kBq];
#pragma unroll
        for (uint32_t i = 0; i < kAq; ++i) {
            const uint32_t q = t + i * kThreads, row = q / (T / 4), ch = q % (T / 4);
            sa[i] = make_uint4(0, 0, 0, 0);
            if (!MASKED || I0 + row - kb >= T)
                sa[i] = *reinterpret_cast<const uint4*>(W + (size_t)(I0 + row) * ldw + kb + ch * 4);
        }
#pragma unroll
        for (uint32_t i = 0; i < kBq; ++i) {
            const uint32_t q = t + i * kThreads, row = q >> 5, ch = q & 31;
            sb[i] = make_uint4(0, 0, 0, 0);
            if (!MASKED || J0 + ch * 4 - kb >= T)
                sb[i] = *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + J0 + ch * 4);
        }
        load_c();
#pragma unroll
        for (uint32_t i = 0; i < kAq; ++i) {
            const uint32_t q = t + i * kThreads;
            *reinterpret_cast<uint4*>(As + (q / (T / 4)) * AP + (q % (T / 4)) * 4) = sa[i];
        }
#pragma unroll
        for (uint32_t i = 0; i < kBq; ++i) {
            const uint32_t q = t + i * kThreads;
            *reinterpret_cast<uint4*>(Bs + (q >> 5) * kMacro + (q & 31) * 4) = sb[i];
        }
    } else {
        // C first, then the strips (pivot block as zeros) straight into LDS
        load_c();
        for (uint32_t q = t; q < kMacro * (T / 4); q += kThreads) {
            const uint32_t row = q / (T / 4), ch = q % (T / 4);
            uint4 v = make_uint4(0, 0, 0, 0);
            if (!MASKED || I0 + row - kb >= T)
                v = *reinterpret_cast<const uint4*>(W + (size_t)(I0 + row) * ldw + kb + ch * 4);
            *reinterpret_cast<uint4*>(As + row * AP + ch * 4) = v;
        }
        for (uint32_t q = t; q < T * 32; q += kThreads) {
            const uint32_t row = q >> 5, ch = q & 31;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (!MASKED || J0 + ch * 4 - kb >= T)
                v = *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + J0 + ch * 4);
            *reinterpret_cast<uint4*>(Bs + row * kMacro + ch * 4) = v;
        }
    }
    __syncthreads();

    uint32_t acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[r][q] = LATE ? 0u : word(c[r][0], q);
            acc[r][4 + q] = LATE ? 0u : word(c[r][1], q);
        }

#pragma unroll 1
    for (uint32_t k0 = 0; k0 < T; k0 += 4) {
        uint4 a4[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            a4[r] = *reinterpret_cast<const uint4*>(As + (ty + 16 * r) * AP + k0);
#pragma unroll
        for (int kk = 0; kk < 4; kk += 2) {
            const uint32_t* b0p = Bs + (k0 + kk) * kMacro + tx * 4;
            const uint32_t* b1p = b0p + kMacro;
            const uint4 b00 = *reinterpret_cast<const uint4*>(b0p);
            const uint4 b01 = *reinterpret_cast<const uint4*>(b0p + 64);
            const uint4 b10 = *reinterpret_cast<const uint4*>(b1p);
            const uint4 b11 = *reinterpret_cast<const uint4*>(b1p + 64);
            const uint32_t b0[8] = {b00.x, b00.y, b00.z, b00.w, b01.x, b01.y, b01.z, b01.w};
            const uint32_t b1[8] = {b10.x, b10.y, b10.z, b10.w, b11.x, b11.y, b11.z, b11.w};
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const uint32_t a0 = word(a4[r], kk), a1 = word(a4[r], kk + 1);
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    acc[r][q] = max(max(acc[r][q], min(a0, b0[q])), min(a1, b1[q]));
            }
        }
    }
    if constexpr (LATE) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[r][q] = max(acc[r][q], word(c[r][0], q));
                acc[r][4 + q] = max(acc[r][4 + q], word(c[r][1], q));
            }
    }

    uint32_t* op = W + (size_t)(I0 + ty) * ldw + J0 + tx * 4;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (MASKED && I0 + ty + 16 * r - kb < T) continue;
        if (!lo_in_pivot)
            *reinterpret_cast<uint4*>(op + (size_t)(16 * r) * ldw) =
                make_uint4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
        if (!hi_in_pivot)
            *reinterpret_cast<uint4*>(op + (size_t)(16 * r) * ldw + 64) =
                make_uint4(acc[r][4], acc[r][5], acc[r][6], acc[r][7]);
    }
}

template <uint32_t T>
__device__ __forceinline__ void indep_tile_u32(uint32_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t I0,
                                               uint32_t J0, uint32_t* __restrict__ lds) {
    // pivot 32 keeps the checks everywhere: unconditional strip loads would be hoisted
    // together, past the 128 VGPRs of 4 waves per SIMD
    if (T == 64 && (I0 + kMacro <= kb || I0 >= kb + T) && (J0 + kMacro <= kb || J0 >= kb + T))
        indep_tile_u32_impl<T, false>(W, ldw, kb, I0, J0, lds);
    else
        indep_tile_u32_impl<T, true>(W, ldw, kb, I0, J0, lds);
}

// ---------------------------------------------------------------------------------------
// 16-bit panels (one workgroup: the row chunk and the column chunk `chunk` of the pivot
// block starting at kb).  Every workgroup closes the pivot tile in registers, keeps it as
// u16 in LDS (pitch T + 2) and runs both chunk products on packed u16 pairs
// (v_pk_min_u16 / v_pk_max_u16: exact on the integer values and on the f16-ordered bit
// patterns alike, whose unsigned order is their value order).  LDS use (about 27 KB at
// T = 64) stays below the independent kernel's, so panel workgroups and independent tiles
// can share one launch.
//   P16: T x (T + 2) u16                  closed pivot
//   S:   T x 128 u16 (row chunk, pitch 128) or T x 136 u16 (column chunk, transposed:
//        S[k][i] = W[c0 + i][kb + k], pitch 136 spreads consecutive k over the banks)
//   rowbuf, colbuf: 2 x T u32 each         pivot-closure broadcast
template <uint32_t T>
constexpr uint32_t panel_u16_lds_elems() {
    return T * (T + 2) + T * 136 + 8 * T;
}

template <uint32_t T>
__device__ __forceinline__ void close_pivot_u16(const uint16_t* __restrict__ W, size_t ldw, uint32_t kb,
                                                uint16_t* P16, uint32_t* rowbuf, uint32_t* colbuf) {
    constexpr uint32_t R = T / 16, PP = T + 2;
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    const uint32_t pr = R * ty, pc = R * tx;
    uint32_t cell[R][R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) cell[r][c] = W[(size_t)(kb + pr + r) * ldw + kb + pc + c];
#pragma unroll
    for (uint32_t k = 0; k < T; ++k) {
        const uint32_t kr = k % R, buf = (k & 1) * T;
        if (ty == k / R) {
#pragma unroll
            for (uint32_t c = 0; c < R; ++c) rowbuf[buf + pc + c] = cell[kr][c];
        }
        if (tx == k / R) {
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) colbuf[buf + pr + r] = cell[r][kr];
        }
        __syncthreads();
        uint32_t a[R], b[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) a[r] = colbuf[buf + pr + r];
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) b[c] = rowbuf[buf + pc + c];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r)
#pragma unroll
            for (uint32_t c = 0; c < R; ++c) cell[r][c] = max(cell[r][c], min(a[r], b[c]));
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) P16[(pr + r) * PP + pc + c] = (uint16_t)cell[r][c];
    __syncthreads();
}

// f16-ordered packed max(acc, min(a0, b0), min(a1, b1)) in 3 instructions (F16), or
// the integer form in 4.
template <bool F16>
__device__ __forceinline__ u16x2 maxmin2(u16x2 acc, u16x2 a0, u16x2 b0, u16x2 a1, u16x2 b1) {
    if constexpr (F16) {
        const h16x2 m0 = __builtin_elementwise_minimum(as_h16x2(as_u32(a0)), as_h16x2(as_u32(b0)));
        const h16x2 m1 = __builtin_elementwise_minimum(as_h16x2(as_u32(a1)), as_h16x2(as_u32(b1)));
        return as_u16x2(h_as_u32(
            __builtin_elementwise_maximum(__builtin_elementwise_maximum(as_h16x2(as_u32(acc)), m0), m1)));
    } else {
        return pk_max(pk_max(acc, pk_min(a0, b0)), pk_min(a1, b1));
    }
}

// Row chunk (K, [c0, c0 + 128)): new[i][j] = max(old[i][j], max_k min(P[i][k], old[k][j])).
// Lane (ty, tx): pivot rows R ty + r (r < R), columns c0 + 8 tx .. 8 tx + 7 (4 packed pairs).
// S holds the chunk's old values (staged by the caller, followed by a barrier).
template <uint32_t T, bool F16>
__device__ __forceinline__ void panel_rows_u16(uint16_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t c0,
                                               const uint16_t* P16, const uint16_t* S) {
    constexpr uint32_t R = T / 16, PP = T + 2;
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    u16x2 acc[R][4];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint4 v = *reinterpret_cast<const uint4*>(S + (R * ty + r) * kMacro + 8 * tx);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][q] = as_u16x2(word(v, q));
    }
#pragma unroll 2
    for (uint32_t k = 0; k < T; k += 2) {
        const uint4 b4 = *reinterpret_cast<const uint4*>(S + k * kMacro + 8 * tx);
        const uint4 b5 = *reinterpret_cast<const uint4*>(S + (k + 1) * kMacro + 8 * tx);
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const u16x2 a2 = as_u16x2(*reinterpret_cast<const uint32_t*>(P16 + (R * ty + r) * PP + k));
            const u16x2 a0 = a2.xx, a1 = a2.yy;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[r][q] = maxmin2<F16>(acc[r][q], a0, as_u16x2(word(b4, q)), a1, as_u16x2(word(b5, q)));
        }
    }
    if (c0 + 8 * tx - kb < T) return;  // columns of the pivot block stay untouched
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
        *reinterpret_cast<uint4*>(W + (size_t)(kb + R * ty + r) * ldw + c0 + 8 * tx) =
            make_uint4(as_u32(acc[r][0]), as_u32(acc[r][1]), as_u32(acc[r][2]), as_u32(acc[r][3]));
}

// Column chunk ([c0, c0 + 128), K): new[i][j] = max(old[i][j], max_k min(old[i][k], P[k][j])).
// Lane (ty, tx): rows c0 + 8 ty + r (r < 8), pivot columns R tx .. R tx + R - 1 (R / 2 pairs).
// S holds the chunk transposed (S[k][i] = old[c0 + i][kb + k], pitch 136), staged by the
// caller and followed by a barrier.
template <uint32_t T, bool F16>
__device__ __forceinline__ void panel_cols_u16(uint16_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t c0,
                                               const uint16_t* P16, const uint16_t* S) {
    constexpr uint32_t R = T / 16, PP = T + 2, SP = 136;
    static_assert(R % 2 == 0, "column chunk pairs need an even share of pivot columns");
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    u16x2 acc[8][R / 2];
#pragma unroll
    for (uint32_t c = 0; c < R; c += 2) {
        const uint4 v0 = *reinterpret_cast<const uint4*>(S + (R * tx + c) * SP + 8 * ty);
        const uint4 v1 = *reinterpret_cast<const uint4*>(S + (R * tx + c + 1) * SP + 8 * ty);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t x0 = (word(v0, r >> 1) >> (16 * (r & 1))) & 0xFFFFu;
            const uint32_t x1 = (word(v1, r >> 1) >> (16 * (r & 1))) & 0xFFFFu;
            acc[r][c / 2] = as_u16x2(x0 | (x1 << 16));
        }
    }
#pragma unroll 2
    for (uint32_t k = 0; k < T; k += 2) {
        const uint4 a4 = *reinterpret_cast<const uint4*>(S + k * SP + 8 * ty);
        const uint4 a5 = *reinterpret_cast<const uint4*>(S + (k + 1) * SP + 8 * ty);
        u16x2 b0[R / 2], b1[R / 2];
#pragma unroll
        for (uint32_t c = 0; c < R / 2; ++c) {
            b0[c] = as_u16x2(*reinterpret_cast<const uint32_t*>(P16 + k * PP + R * tx + 2 * c));
            b1[c] = as_u16x2(*reinterpret_cast<const uint32_t*>(P16 + (k + 1) * PP + R * tx + 2 * c));
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const u16x2 x2 = as_u16x2(word(a4, r >> 1)), y2 = as_u16x2(word(a5, r >> 1));
            const u16x2 av = (r & 1) ? x2.yy : x2.xx, aw = (r & 1) ? y2.yy : y2.xx;
#pragma unroll
            for (uint32_t c = 0; c < R / 2; ++c) acc[r][c] = maxmin2<F16>(acc[r][c], av, b0[c], aw, b1[c]);
        }
    }
    if (c0 + 8 * ty - kb < T) return;  // rows of the pivot block stay untouched
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        uint16_t* dst = W + (size_t)(c0 + 8 * ty + r) * ldw + kb + R * tx;
        if constexpr (R == 4) {
            *reinterpret_cast<uint2*>(dst) = make_uint2(as_u32(acc[r][0]), as_u32(acc[r][1]));
        } else {
#pragma unroll
            for (uint32_t c = 0; c < R / 2; ++c) reinterpret_cast<uint32_t*>(dst)[c] = as_u32(acc[r][c]);
        }
    }
}

// The whole panel workgroup: commits the previous round's closed pivot (workgroup 0, when
// given), closes this pivot, publishes it (workgroup 0), then row chunk and column chunk
// `chunk`.  Both chunks are read from W before the pivot closes (neither product writes a
// cell the other reads: they meet only in the pivot block, which neither stores), so their
// loads overlap the pivot closure.  Starts with a barrier.
//
// kr != kNoRange (T = 64 only): this panel's strips have not yet received the update of
// the round whose pivot block starts at kr (strips final).  The workgroup applies it
// first, with the 64 x 64 quarter-tile code: to the two quarters of its row chunk and the
// two of its column chunk (written to W; the pivot block quarter is skipped) and to the
// pivot block itself, into LDS only -- every panel workgroup needs it, and the block in W
// is replaced by the closed pivot one launch later.  Cells in kr's rows or columns are
// final and stay untouched (the quarter code masks them), so the block at kr, which
// workgroup 0 may be committing meanwhile, is never read.
template <uint32_t T, bool F16>
__device__ __forceinline__ void panel_u16(uint16_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t chunk,
                                          uint32_t* __restrict__ pivot_out, const uint32_t* __restrict__ pivot_prev,
                                          uint32_t kb_prev, uint16_t* lds, uint32_t kr = kNoRange) {
    constexpr uint32_t NQ = T * 16 / kThreads, SP = 136;
    uint16_t* P16 = lds;
    uint16_t* S = lds + T * (T + 2);
    uint32_t* rowbuf = reinterpret_cast<uint32_t*>(S + T * SP);
    uint32_t* colbuf = rowbuf + 2 * T;
    const uint32_t t = threadIdx.x, c0 = chunk * kMacro;
    __syncthreads();
    if (chunk == 0 && pivot_prev != nullptr) {
        for (uint32_t q = t; q < T * T; q += kThreads)
            W[(size_t)(kb_prev + q / T) * ldw + kb_prev + q % T] = (uint16_t)pivot_prev[q];
    }
    // the pivot block to close: W, or its updated copy in LDS
    const uint16_t* pv = W;
    size_t pv_ld = ldw;
    uint32_t pv_k = kb;
    if constexpr (T == 64) {
        if (kr != kNoRange) {
            uint16_t* pn = lds + indep64_u16_lds_elems<T>();  // beyond the quarter code's LDS
#pragma unroll 1
            for (uint32_t q = 0; q < 2; ++q) {
                const uint32_t j0 = c0 + 64 * q;
                if (j0 == kb) continue;  // the pivot block: below, into LDS
                indep_tile64_u16<T, F16>(W, ldw, kr, kb, j0, lds);  // row chunk quarter
                indep_tile64_u16<T, F16>(W, ldw, kr, j0, kb, lds);  // column chunk quarter
            }
            indep_tile64_u16<T, F16>(W, ldw, kr, kb, kb, lds, pn);
            __syncthreads();  // the quarters' stores and pn are visible to the whole workgroup
            pv = pn;
            pv_ld = 64;
            pv_k = 0;
        }
    }
    // both chunks into registers; thread q of the column gather takes (k = q % T, rows
    // c0 + 8 g .. 8 g + 7 with g = q / T) and packs them for the transposed layout
    static_assert(NQ == 2 || NQ == 4, "chunk share per thread");
    auto gather = [&](uint32_t i) {  // 8 rows of one pivot column, packed in pairs
        const uint32_t q = t + i * kThreads, k = q % T, g = q / T;
        uint32_t w[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const uint32_t lo = W[(size_t)(c0 + 8 * g + 2 * h) * ldw + kb + k];
            const uint32_t hi = W[(size_t)(c0 + 8 * g + 2 * h + 1) * ldw + kb + k];
            w[h] = lo | (hi << 16);
        }
        return make_uint4(w[0], w[1], w[2], w[3]);
    };
    auto put_col = [&](uint32_t i, const uint4& v) {
        const uint32_t q = t + i * kThreads, k = q % T, g = q / T;
        *reinterpret_cast<uint4*>(S + k * SP + 8 * g) = v;
    };
    auto load_row = [&](uint32_t i) {
        const uint32_t q = t + i * kThreads, row = q >> 4, ch = q & 15;
        return *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + c0 + ch * 8);
    };
    auto put_row = [&](uint32_t i, const uint4& v) {
        const uint32_t q = t + i * kThreads, row = q >> 4, ch = q & 15;
        *reinterpret_cast<uint4*>(S + row * kMacro + ch * 8) = v;
    };
    // named registers (not arrays, which the compiler keeps on the stack here): the values
    // live across the fully unrolled pivot closure
    const uint4 row0 = load_row(0), row1 = load_row(1);
    uint4 row2 = make_uint4(0, 0, 0, 0), row3 = make_uint4(0, 0, 0, 0);
    const uint4 col0 = gather(0), col1 = gather(1);
    uint4 col2 = make_uint4(0, 0, 0, 0), col3 = make_uint4(0, 0, 0, 0);
    if constexpr (NQ == 4) {
        row2 = load_row(2);
        row3 = load_row(3);
        col2 = gather(2);
        col3 = gather(3);
    }
    close_pivot_u16<T>(pv, pv_ld, pv_k, P16, rowbuf, colbuf);  // ends with a barrier
    if (chunk == 0)
        for (uint32_t q = t; q < T * T; q += kThreads) pivot_out[q] = P16[(q / T) * (T + 2) + q % T];
    put_row(0, row0);
    put_row(1, row1);
    if constexpr (NQ == 4) {
        put_row(2, row2);
        put_row(3, row3);
    }
    __syncthreads();
    panel_rows_u16<T, F16>(W, ldw, kb, c0, P16, S);
    __syncthreads();  // S held the row chunk
    put_col(0, col0);
    put_col(1, col1);
    if constexpr (NQ == 4) {
        put_col(2, col2);
        put_col(3, col3);
    }
    __syncthreads();
    panel_cols_u16<T, F16>(W, ldw, kb, c0, P16, S);
}

// The panel of one pivot block in one workgroup: the pivot is closed once, then every
// row chunk and column chunk of the nb = ldw / 128 chunks in turn (for small matrices,
// where one workgroup per chunk would close the same pivot nb times).  Same LDS layout
// and hand-off as panel_u16.
template <uint32_t T, bool F16>
__device__ __forceinline__ void panel_all_u16(uint16_t* __restrict__ W, size_t ldw, uint32_t kb, uint32_t nb,
                                              uint32_t* __restrict__ pivot_out,
                                              const uint32_t* __restrict__ pivot_prev, uint32_t kb_prev,
                                              uint16_t* lds) {
    constexpr uint32_t NQ = T * 16 / kThreads, SP = 136;
    uint16_t* P16 = lds;
    uint16_t* S = lds + T * (T + 2);
    uint32_t* rowbuf = reinterpret_cast<uint32_t*>(S + T * SP);
    uint32_t* colbuf = rowbuf + 2 * T;
    const uint32_t t = threadIdx.x;
    __syncthreads();
    if (pivot_prev != nullptr) {
        for (uint32_t q = t; q < T * T; q += kThreads)
            W[(size_t)(kb_prev + q / T) * ldw + kb_prev + q % T] = (uint16_t)pivot_prev[q];
    }
    close_pivot_u16<T>(W, ldw, kb, P16, rowbuf, colbuf);  // ends with a barrier
    for (uint32_t q = t; q < T * T; q += kThreads) pivot_out[q] = P16[(q / T) * (T + 2) + q % T];
    for (uint32_t chunk = 0; chunk < nb; ++chunk) {
        const uint32_t c0 = chunk * kMacro;
#pragma unroll
        for (uint32_t i = 0; i < NQ; ++i) {
            const uint32_t q = t + i * kThreads, row = q >> 4, ch = q & 15;
            *reinterpret_cast<uint4*>(S + row * kMacro + ch * 8) =
                *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + c0 + ch * 8);
        }
        __syncthreads();
        panel_rows_u16<T, F16>(W, ldw, kb, c0, P16, S);
        __syncthreads();
        for (uint32_t q = t; q < T * 16; q += kThreads) {
            const uint32_t k = q % T, g = q / T;
            uint32_t w[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const uint32_t lo = W[(size_t)(c0 + 8 * g + 2 * h) * ldw + kb + k];
                const uint32_t hi = W[(size_t)(c0 + 8 * g + 2 * h + 1) * ldw + kb + k];
                w[h] = lo | (hi << 16);
            }
            *reinterpret_cast<uint4*>(S + k * SP + 8 * g) = make_uint4(w[0], w[1], w[2], w[3]);
        }
        __syncthreads();
        panel_cols_u16<T, F16>(W, ldw, kb, c0, P16, S);
        __syncthreads();
    }
}

}  // namespace se
