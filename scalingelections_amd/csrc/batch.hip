// batch.hip -- thousands of independent elections of uneven size in a few launches.
//
// Per election e (C_e candidates, m_e voters, positions m_e x C_e int32) the batch
// produces exactly what the single-election path produces: the strongest-path matrix
// (scaling_elections.py:91-125 semantics, zero diagonal) and the ranking by descending
// wins with ties in ascending index (:274-296).
//
// Counting, batch_count_kernel: one workgroup per work item, items sorted by decreasing
//   cost so the longest start first.  Elections with C <= 128 are one item each, counted
//   voter-parallel: the C x C counter matrix is cut into 8 x 8 blocks (<= 256) and the
//   256 lanes are split into G = 256 / blocks voter groups; partial counts are summed
//   through LDS.  Larger elections are cut into 128 x 128 tile pairs (I <= J), each an
//   item, counted like the single-election kernel (count_tile), d[j][i] = m - d[i][j].
//   Items far above a workgroup's fair share are cut into voter slices that add into a
//   zeroed d atomically.
// Closure, batch_closure_kernel: one workgroup per election, largest first.  p0 into a
//   zero-padded work matrix (global memory, L2-resident at these sizes), then the blocked
//   closure run by the workgroup alone (close_pivot, panel chunk products, independent
//   macro tiles -- the same building blocks as the single-election kernels, with barriers
//   in place of kernel boundaries), then output, win counts and the stable ranking, all
//   on the device.  Elections with C <= 64 are closed entirely in registers/LDS.  On the
//   16-bit paths the elections above a threshold (launch-shape table, 129 by default) are
//   instead closed together by one pair of launches per round (batch_panel_kernel, then
//   batch_tile_kernel over all their macro tiles), followed by a tail phase.
//
// The per-launch strategy (SE_BATCH_PER_LAUNCH) runs every election through the plain
// single-election entry points instead, for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <type_traits>
#include <vector>

#include "closure_tiles.h"
#include "count_tiles.h"
#include "se_internal.h"

namespace se {

int count_device_impl(DeviceContext* ctx, const void* pos, bool pos16, uint32_t m, uint32_t n, uint32_t* d, size_t ldd,
                      hipStream_t s, double* kernel_ms, uint32_t* flag);
int closure_device_impl(DeviceContext* ctx, const uint32_t* d, uint32_t n, size_t ld, uint32_t* out,
                        size_t ldo, uint32_t flags, hipStream_t s, se_timing* timing);
int wins_device_impl(DeviceContext* ctx, const uint32_t* p, uint32_t n, size_t ld, uint32_t* wins,
                     hipStream_t s);

namespace {

constexpr uint32_t kMaxBatchC = 1024;  // larger elections belong on the single-election path

struct ElectionMeta {
    uint64_t pos_off;  // element offset of the positions (m x C int32)
    uint64_t d_off;    // element offset of d and of p (C x C uint32)
    uint64_t w_off;    // element offset of the work matrix (Cp x Cp)
    uint64_t r_off;    // element offset of the ranking (C uint32)
    uint32_t C, m, Cp, K;  // K: counter packing width (voters per word)
    uint32_t piv;          // index among the round-closed elections (kNoRounds: closed alone)
    uint32_t pad[3];
};

constexpr uint32_t kNoRounds = 0xFFFFFFFFu;

// ------------------------------------------------------------------------- counting
// One work item: rows [I0, I0 + ri) x columns [J0, J0 + rj) of election e's d, over voters
// [vb, ve).  shape = mode | ri << 8 | rj << 16; mode 1: full 128 x 128 tile, mode 0:
// voter-parallel sub-block (small elections and the partial tiles at the edge).  Items
// whose work far exceeds a workgroup's fair share are cut into voter slices, which add
// their counts atomically into the zeroed d (put = kPutAtomic); the others store.
struct CountItem {
    uint32_t e, I0, J0, shape;
    uint32_t vb, ve, put, pad;
};

template <uint32_t K, typename P>
__device__ __forceinline__ void batch_count_item(const P* pos, uint32_t* d, uint32_t C, const CountItem& it,
                                                 uint32_t (*PI)[kCT], uint32_t (*PJ)[kCT]) {
    if ((it.shape & 0xFFu) == 1) {
        count_tile<K, P>(pos, it.vb, it.ve, C, C, d, C, it.I0 / kCT, it.J0 / kCT, PI, PJ, &PI[0][0], it.put);
        return;
    }
    const uint32_t ri = (it.shape >> 8) & 0xFFu, rj = (it.shape >> 16) & 0xFFu;
    count_block<K, P>(pos, it.vb, it.ve, C, it.I0, it.J0, ri, rj, d, C, PI, PJ, &PI[0][0], it.put);
}

// Positions P: int32, or uint16 (SE_BATCH_POS16: half the bytes the staging reads).
template <typename P>
__global__ __launch_bounds__(256) void batch_count_kernel(const P* __restrict__ pos_all,
                                                          uint32_t* __restrict__ d_all,
                                                          const ElectionMeta* __restrict__ meta,
                                                          const CountItem* __restrict__ items) {
    __shared__ __align__(16) uint32_t PI[kVW][kCT];
    __shared__ __align__(16) uint32_t PJ[kVW][kCT];
    const CountItem it = items[blockIdx.x];
    const ElectionMeta em = meta[it.e];
    const P* pos = pos_all + em.pos_off;
    uint32_t* d = d_all + em.d_off;
    if (em.K == 4)
        batch_count_item<4, P>(pos, d, em.C, it, PI, PJ);
    else if (em.K == 2)
        batch_count_item<2, P>(pos, d, em.C, it, PI, PJ);
    else
        batch_count_item<1, P>(pos, d, em.C, it, PI, PJ);
}

// ------------------------------------------------------------------------- closure
template <typename E, uint32_t T>
struct BatchLds {
    uint32_t P[T][T + 1];
    uint32_t rowbuf[2][T], colbuf[2][T];
    uint32_t wins[kMaxBatchC];
    union alignas(16) {
        uint32_t S[kMacro * (T + 1) > 2 * 64 * 65 ? kMacro * (T + 1) : 2 * 64 * 65];
        E indep[sizeof(E) == 2 ? indep_u16_lds_elems<T>() : indep_u32_lds_elems<T>()];
    } u;
};

template <typename E, uint32_t T, bool F16>
__device__ __forceinline__ void indep_tile(E* W, size_t ldw, uint32_t kb, uint32_t I0, uint32_t J0, E* lds) {
    if constexpr (sizeof(E) == 2)
        indep_tile_u16<T, F16>(W, ldw, kb, I0, J0, lds);
    else
        indep_tile_u32<T>(W, ldw, kb, I0, J0, lds);
}

// Per-election square passes over 64 x 64 tile pairs (I, J), I <= J: the tile and its
// transposed partner are staged in LDS (u32, pitch 65: conflict-free column reads), so both
// orientations are read and written by whole rows (coalesced, no division per cell).  Wave
// w of the workgroup handles rows w, w + 4, ... of a tile, lane c column c.
constexpr uint32_t kPT = 64, kPP = kPT + 1;

// p0 (scaling_elections.py:103-112) of tile pair (I, J), I <= J, of one election (d: C x C)
// into its padded work matrix (Cp x Cp): v = d[i][j] + bias when d[i][j] > d[j][i], else 0
// (diagonal and padding 0).  A, B: 2 x 64 x 65 u32 of LDS.  Starts with a barrier.
// U: rows staged per unrolled step (loads in flight per thread: 2 U).
template <typename E, uint32_t U>
__device__ __forceinline__ void p0_pair(const uint32_t* __restrict__ d, uint32_t C, E* __restrict__ W, uint32_t Cp,
                                        uint32_t bias, uint32_t I, uint32_t J, uint32_t* A, uint32_t* B) {
    const uint32_t t = threadIdx.x, c = t & (kPT - 1), r0 = t / kPT;
    __syncthreads();
    // unconditional loads at clamped indices (cells outside C x C count as 0), so the loads
    // of U rows issue together instead of one round trip each
#pragma unroll U
    for (uint32_t r = r0; r < kPT; r += kThreads / kPT) {
        const uint32_t i = I * kPT + r, j = J * kPT + c, i2 = J * kPT + r, j2 = I * kPT + c;
        const uint32_t x = d[(size_t)min(i, C - 1) * C + min(j, C - 1)];
        const uint32_t y = d[(size_t)min(i2, C - 1) * C + min(j2, C - 1)];
        A[r * kPP + c] = (i < C && j < C) ? x : 0u;
        B[r * kPP + c] = (i2 < C && j2 < C) ? y : 0u;
    }
    __syncthreads();
    for (uint32_t r = r0; r < kPT; r += kThreads / kPT) {
        const uint32_t i = I * kPT + r, j = J * kPT + c;
        if (i < Cp && j < Cp) {
            const uint32_t x = A[r * kPP + c], y = B[c * kPP + r];
            W[(size_t)i * Cp + j] = (E)((i != j && x > y) ? x + bias : 0u);
        }
        const uint32_t i2 = J * kPT + r, j2 = I * kPT + c;
        if (I != J && i2 < Cp && j2 < Cp) {
            const uint32_t x = B[r * kPP + c], y = A[c * kPP + r];
            W[(size_t)i2 * Cp + j2] = (E)(x > y ? x + bias : 0u);
        }
    }
}

// p0 of a whole election, pair by pair.
template <typename E, uint32_t U>
__device__ void batch_p0(const uint32_t* __restrict__ d, uint32_t C, E* __restrict__ W, uint32_t Cp,
                         uint32_t bias, uint32_t* A, uint32_t* B) {
    const uint32_t nt = (Cp + kPT - 1) / kPT;
    for (uint32_t I = 0; I < nt; ++I)
        for (uint32_t J = I; J < nt; ++J) p0_pair<E, U>(d, C, W, Cp, bias, I, J, A, B);
    __syncthreads();
}

// Tail of tile pair (I, J), I <= J: output p (C x C, zero diagonal, bias removed) of both
// tiles, and wins[i] += #{j in the pair's columns, j != i : W[i][j] > W[j][i]}
// (scaling_elections.py:287-290; the offset keeps the order) for the rows of both tiles,
// added atomically (wins: LDS or global).  A, B as in p0_pair.  Starts with a barrier.
template <typename E, uint32_t U>
__device__ __forceinline__ void tail_pair(const E* __restrict__ W, uint32_t Cp, uint32_t C, uint32_t bias,
                                          uint32_t* __restrict__ p, uint32_t I, uint32_t J, uint32_t* wins,
                                          uint32_t* A, uint32_t* B) {
    const uint32_t t = threadIdx.x, c = t & (kPT - 1), r0 = t / kPT;
    __syncthreads();
    // unconditional loads at clamped indices (cells outside C x C count as 0), so the loads
    // of U rows issue together
#pragma unroll U
    for (uint32_t r = r0; r < kPT; r += kThreads / kPT) {
        const uint32_t i = I * kPT + r, j = J * kPT + c, i2 = J * kPT + r, j2 = I * kPT + c;
        const uint32_t x = W[(size_t)min(i, C - 1) * Cp + min(j, C - 1)];
        const uint32_t y = W[(size_t)min(i2, C - 1) * Cp + min(j2, C - 1)];
        A[r * kPP + c] = (i < C && j < C) ? x : 0u;
        B[r * kPP + c] = (i2 < C && j2 < C) ? y : 0u;
    }
    __syncthreads();
    uint32_t beats = 0;  // column j = J * 64 + c over this tile's rows (I != J)
    for (uint32_t r = r0; r < kPT; r += kThreads / kPT) {
        const uint32_t i = I * kPT + r, j = J * kPT + c;
        const bool in = i < C && j < C;
        const uint32_t x = A[r * kPP + c], y = B[c * kPP + r];
        if (in) p[(size_t)i * C + j] = (i == j || x == 0) ? 0u : x - bias;
        // row i is this wave's alone within the pair: count its wins with one ballot
        const unsigned long long won = __ballot(in && i != j && x > y);
        if (c == 0 && won) atomicAdd(&wins[i], (uint32_t)__popcll(won));
        if (I != J) {
            beats += (in && y > x) ? 1u : 0u;
            const uint32_t i2 = J * kPT + r, j2 = I * kPT + c;
            const uint32_t x2 = B[r * kPP + c];
            if (i2 < C && j2 < C) p[(size_t)i2 * C + j2] = x2 == 0 ? 0u : x2 - bias;
        }
    }
    if (beats) atomicAdd(&wins[J * kPT + c], beats);
}

// The stable ranking by descending wins, ties in ascending index (:292):
// rank(i) = #{j : w_j > w_i or (w_j = w_i and j < i)}.  wins: C words of LDS.
__device__ __forceinline__ void rank_from_wins(const uint32_t* wins, uint32_t C, uint32_t* __restrict__ rank_out) {
    for (uint32_t i = threadIdx.x; i < C; i += kThreads) {
        const uint32_t wi = wins[i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < C; ++j) {
            const uint32_t wj = wins[j];
            rank += (wj > wi || (wj == wi && j < i)) ? 1u : 0u;
        }
        rank_out[rank] = i;
    }
}

// Output, wins and ranking of a whole election, pair by pair.  wins: C words of LDS; A, B
// as in p0_pair.
template <typename E, uint32_t U>
__device__ void batch_tail(const E* __restrict__ W, uint32_t Cp, uint32_t C, uint32_t bias, uint32_t* __restrict__ p,
                           uint32_t* __restrict__ rank_out, uint32_t* wins, uint32_t* A, uint32_t* B) {
    for (uint32_t i = threadIdx.x; i < C; i += kThreads) wins[i] = 0;
    const uint32_t nt = (C + kPT - 1) / kPT;
    for (uint32_t I = 0; I < nt; ++I)
        for (uint32_t J = I; J < nt; ++J) tail_pair<E, U>(W, Cp, C, bias, p, I, J, wins, A, B);
    __syncthreads();
    rank_from_wins(wins, C, rank_out);
}

// Offset of the f16-ordered batch closure: nonzero p0 values v <= m <= 30720 are stored as
// v + 0x3FF, positive f16 normals in [0x0400, 0x7BFF] whose unsigned order is their value
// order; the pivot and panel code compares them as integers, the tiles as f16.
constexpr uint32_t kBatchF16Bias = 0x3FFu;
constexpr uint32_t kBatchF16MaxVoters = 0x7C00u - 0x0400u;

// The elections closed alone (piv == kNoRounds), one workgroup each: p0 into the work
// matrix, the closure, then the tail (output, wins, ranking).  The round-closed elections
// exit at once (batch_p0_pairs_kernel, batch_tail_pairs_kernel and batch_rank_kernel
// handle their p0 and tail).
template <typename E, uint32_t T, bool F16>
__device__ __forceinline__ void batch_closure_body(const uint32_t* __restrict__ d_all, E* __restrict__ W_all,
                                                   uint32_t* __restrict__ p_all, uint32_t* __restrict__ rank_all,
                                                   const ElectionMeta* __restrict__ meta,
                                                   const uint32_t* __restrict__ order) {
    __shared__ __align__(16) BatchLds<E, T> L;
    const uint32_t t = threadIdx.x;
    const ElectionMeta em = meta[order[blockIdx.x]];
    const uint32_t C = em.C, Cp = em.Cp;
    const uint32_t* d = d_all + em.d_off;
    uint32_t* p = p_all + em.d_off;
    E* W = W_all + em.w_off;
    if (C == 0 || em.piv != kNoRounds) return;
    constexpr uint32_t bias = F16 ? kBatchF16Bias : 0u;

    batch_p0<E, 4>(d, C, W, Cp, bias, L.u.S, L.u.S + kPT * kPP);

    for (uint32_t r = 0, rounds = (C + T - 1) / T; r < rounds; ++r) {
        const uint32_t kb = r * T;
        close_pivot<E, T>(W, Cp, kb, min(T, C - kb), L.P, L.rowbuf, L.colbuf);
        if (Cp > T) {
            for (uint32_t c0 = 0; c0 < Cp; c0 += kMacro) {
                load_row_chunk<E, T>(W, Cp, kb, c0, L.u.S);
                __syncthreads();
                row_chunk_product<E, T>(W, Cp, kb, c0, L.P, L.u.S);
                __syncthreads();
                load_col_chunk<E, T>(W, Cp, kb, c0, L.u.S);
                __syncthreads();
                col_chunk_product<E, T>(W, Cp, kb, c0, L.P, L.u.S);
                __syncthreads();
            }
        }
        for (uint32_t q = t; q < T * T; q += kThreads) W[(size_t)(kb + q / T) * Cp + kb + q % T] = (E)L.P[q / T][q % T];
        __syncthreads();
        if (Cp > T) {
            for (uint32_t I0 = 0; I0 < Cp; I0 += kMacro)
                for (uint32_t J0 = 0; J0 < Cp; J0 += kMacro) indep_tile<E, T, F16>(W, Cp, kb, I0, J0, L.u.indep);
            __syncthreads();
        }
    }

    // output with zero diagonal, wins (:287-290) and the stable ranking (:292)
    __syncthreads();  // the closure's last LDS use (the union) ends here
    batch_tail<E, 4>(W, Cp, C, bias, p, rank_all + em.r_off, L.wins, L.u.S, L.u.S + kPT * kPP);
}

template <typename E, uint32_t T, bool F16>
__global__ __launch_bounds__(256) void batch_closure_kernel(const uint32_t* __restrict__ d_all,
                                                            E* __restrict__ W_all, uint32_t* __restrict__ p_all,
                                                            uint32_t* __restrict__ rank_all,
                                                            const ElectionMeta* __restrict__ meta,
                                                            const uint32_t* __restrict__ order) {
    batch_closure_body<E, T, F16>(d_all, W_all, p_all, rank_all, meta, order);
}

// The 32-bit form: its LDS (43.6 KB) admits three workgroups per CU; the bounds keep its
// registers within three waves per SIMD.
__global__ __launch_bounds__(256, 3) void batch_closure32_kernel(const uint32_t* __restrict__ d_all,
                                                                 uint32_t* __restrict__ W_all,
                                                                 uint32_t* __restrict__ p_all,
                                                                 uint32_t* __restrict__ rank_all,
                                                                 const ElectionMeta* __restrict__ meta,
                                                                 const uint32_t* __restrict__ order) {
    batch_closure_body<uint32_t, 32, false>(d_all, W_all, p_all, rank_all, meta, order);
}

// p0 of the round-closed elections, one workgroup per tile pair (items: election, I, J,
// over the padded side).  The (I, I) workgroups also zero the election's win counts of rows
// I, which batch_tail_pairs_kernel adds to after the rounds.
template <bool F16>
__global__ __launch_bounds__(256) void batch_p0_pairs_kernel(const uint32_t* __restrict__ d_all,
                                                             uint16_t* __restrict__ W_all,
                                                             const ElectionMeta* __restrict__ meta,
                                                             const uint4* __restrict__ items,
                                                             uint32_t* __restrict__ wins_all) {
    __shared__ __align__(16) uint32_t AB[2 * kPT * kPP];
    const uint4 it = items[blockIdx.x];
    const ElectionMeta em = meta[it.x];
    constexpr uint32_t bias = F16 ? kBatchF16Bias : 0u;
    p0_pair<uint16_t, 16>(d_all + em.d_off, em.C, W_all + em.w_off, em.Cp, bias, it.y, it.z, AB, AB + kPT * kPP);
    if (it.y == it.z && threadIdx.x < kPT && it.y * kPT + threadIdx.x < em.C)
        wins_all[em.r_off + it.y * kPT + threadIdx.x] = 0u;
}

// Output and win counts of the round-closed elections, one workgroup per tile pair (items
// over the C x C domain), after the last pivot is committed.
template <bool F16>
__global__ __launch_bounds__(256) void batch_tail_pairs_kernel(const uint16_t* __restrict__ W_all,
                                                               uint32_t* __restrict__ p_all,
                                                               const ElectionMeta* __restrict__ meta,
                                                               const uint4* __restrict__ items,
                                                               uint32_t* __restrict__ wins_all) {
    __shared__ __align__(16) uint32_t AB[2 * kPT * kPP];
    const uint4 it = items[blockIdx.x];
    const ElectionMeta em = meta[it.x];
    constexpr uint32_t bias = F16 ? kBatchF16Bias : 0u;
    tail_pair<uint16_t, 16>(W_all + em.w_off, em.Cp, em.C, bias, p_all + em.d_off, it.y, it.z,
                            wins_all + em.r_off, AB, AB + kPT * kPP);
}

// Rankings of the round-closed elections from their win counts, one workgroup each.
__global__ __launch_bounds__(256) void batch_rank_kernel(const ElectionMeta* __restrict__ meta,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ wins_all,
                                                         uint32_t* __restrict__ rank_all) {
    __shared__ uint32_t wins[kMaxBatchC];
    const ElectionMeta em = meta[list[blockIdx.x]];
    for (uint32_t i = threadIdx.x; i < em.C; i += kThreads) wins[i] = wins_all[em.r_off + i];
    __syncthreads();
    rank_from_wins(wins, em.C, rank_all + em.r_off);
}

// Round r of the round-closed elections, 16-bit: one workgroup per election for the
// round's panel (panel_all_u16: the pivot closed once in registers, then every row chunk
// and column chunk), then one workgroup per (election, macro tile) of the round's
// independent tiles.  Each election keeps its own double-buffered pivot hand-off
// (pivots + piv * 2 T^2); the panel of round r commits the pivot of round r - 1.
template <uint32_t T, bool F16>
__global__ __launch_bounds__(256) void batch_panel_kernel(uint16_t* __restrict__ W_all,
                                                          const ElectionMeta* __restrict__ meta,
                                                          const uint2* __restrict__ items, uint32_t r,
                                                          uint32_t* __restrict__ pivots) {
    __shared__ __align__(16) uint16_t lds[panel_u16_lds_elems<T>()];
    const uint2 it = items[blockIdx.x];  // (election, unused)
    const ElectionMeta em = meta[it.x];
    uint32_t* pv = pivots + (size_t)em.piv * 2 * T * T;
    const uint32_t kb = r * T;
    panel_all_u16<T, F16>(W_all + em.w_off, em.Cp, kb, em.Cp / kMacro, pv + (r & 1) * T * T,
                          r ? pv + ((r - 1) & 1) * T * T : nullptr, r ? kb - T : 0u, lds);
}

template <uint32_t T, bool F16>
__global__ __launch_bounds__(256) void batch_tile_kernel(uint16_t* __restrict__ W_all,
                                                         const ElectionMeta* __restrict__ meta,
                                                         const uint4* __restrict__ items, uint32_t r) {
    __shared__ __align__(16) uint16_t lds[indep_u16_lds_elems<T>()];
    const uint4 it = items[blockIdx.x];  // (election, I0, J0)
    const ElectionMeta em = meta[it.x];
    indep_tile_u16<T, F16>(W_all + em.w_off, em.Cp, r * T, it.y, it.z, lds);
}

// After the last round: every round-closed election commits its last closed pivot.
template <uint32_t T>
__global__ __launch_bounds__(256) void batch_commit_kernel(uint16_t* __restrict__ W_all,
                                                           const ElectionMeta* __restrict__ meta,
                                                           const uint32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ pivots) {
    const ElectionMeta em = meta[list[blockIdx.x]];
    const uint32_t last = (em.C + T - 1) / T - 1, kb = last * T;
    const uint32_t* pv = pivots + (size_t)em.piv * 2 * T * T + (last & 1) * T * T;
    uint16_t* W = W_all + em.w_off;
    for (uint32_t q = threadIdx.x; q < T * T; q += kThreads)
        W[(size_t)(kb + q / T) * em.Cp + kb + q % T] = (uint16_t)pv[q];
}

inline uint32_t padded_side(uint32_t C, uint32_t T) { return C <= T ? T : (uint32_t)round_up(C, kMacro); }

}  // namespace

int batch_device_impl(DeviceContext* ctx, uint32_t E, const uint32_t* Cs, const uint32_t* ms, const void* pos_any,
                      uint32_t* paths, uint32_t* ranking, uint32_t flags, hipStream_t s, se_batch_timing* timing) {
    uint64_t pos_total = 0, d_total = 0, r_total = 0;
    uint32_t max_m = 0;
    const bool pos16 = (flags & SE_BATCH_POS16) != 0;
    for (uint32_t e = 0; e < E; ++e) {
        if (Cs[e] > kMaxBatchC) {
            set_error("batched elections are limited to 1024 candidates; use se_closure_device for larger ones");
            return SE_ERR_INVALID;
        }
        pos_total += (uint64_t)ms[e] * Cs[e];
        d_total += (uint64_t)Cs[e] * Cs[e];
        r_total += Cs[e];
        max_m = std::max(max_m, ms[e]);
    }
    hipEvent_t ev[3] = {};
    if (timing) {
        for (auto& x : ev)
            if (!hip_ok(hipEventCreate(&x), "Failed to create events")) return SE_ERR_DEVICE;
        (void)hipEventRecord(ev[0], s);
    }

    if (flags & SE_BATCH_PER_LAUNCH) {
        // the plain form: one count launch, one closure pipeline and one ranking per election
        uint32_t maxC = 0;
        for (uint32_t e = 0; e < E; ++e) maxC = std::max(maxC, Cs[e]);
        uint32_t* dbuf = static_cast<uint32_t*>(scratch(ctx, kSlotBatchD, (size_t)maxC * maxC * 4 + 256));
        uint32_t* wins = static_cast<uint32_t*>(scratch(ctx, kSlotBatchMeta, (size_t)maxC * 4 + 256));
        if (!dbuf || !wins) return SE_ERR_ALLOC;
        std::vector<uint32_t> hw(maxC), hr(maxC);
        uint64_t po = 0, doff = 0, ro = 0;
        for (uint32_t e = 0; e < E; ++e) {
            const uint32_t C = Cs[e], m = ms[e];
            if (C) {
                const void* pe = pos16 ? static_cast<const void*>(static_cast<const uint16_t*>(pos_any) + po)
                                       : static_cast<const void*>(static_cast<const int32_t*>(pos_any) + po);
                int rc = count_device_impl(ctx, pe, pos16, m, C, dbuf, C, s, nullptr, nullptr);
                if (rc == SE_OK) rc = closure_device_impl(ctx, dbuf, C, C, paths + doff, C, 0, s, nullptr);
                if (rc == SE_OK) rc = wins_device_impl(ctx, paths + doff, C, C, wins, s);
                if (rc != SE_OK) return rc;
                if (!hip_ok(hipMemcpyAsync(hw.data(), wins, C * 4, hipMemcpyDeviceToHost, s), "Failed to copy data from device to host") ||
                    !hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully"))
                    return SE_ERR_COPY;
                std::iota(hr.begin(), hr.begin() + C, 0u);
                std::stable_sort(hr.begin(), hr.begin() + C, [&](uint32_t a, uint32_t b) { return hw[a] > hw[b]; });
                if (!hip_ok(hipMemcpyAsync(ranking + ro, hr.data(), C * 4, hipMemcpyHostToDevice, s),
                            "Failed to copy data from host to device") ||
                    !hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully"))
                    return SE_ERR_COPY;
            }
            po += (uint64_t)m * C;
            doff += (uint64_t)C * C;
            ro += C;
        }
        if (timing) {
            (void)hipEventRecord(ev[1], s);
            (void)hipEventRecord(ev[2], s);
        }
    } else {
        const bool use16 = max_m <= 0xFFFFu;  // d <= m, so every p0 value fits 16 bits
        const uint32_t T = use16 ? 64 : 32;
        std::vector<ElectionMeta> meta(E);
        uint64_t po = 0, doff = 0, wo = 0, ro = 0;
        for (uint32_t e = 0; e < E; ++e) {
            ElectionMeta& em = meta[e];
            em.C = Cs[e];
            em.m = ms[e];
            em.Cp = em.C ? padded_side(em.C, T) : 0;
            em.K = count_pack(em.C, em.m);
            em.piv = kNoRounds;
            em.pad[0] = em.pad[1] = em.pad[2] = 0;
            em.pos_off = po;
            em.d_off = doff;
            em.w_off = wo;
            em.r_off = ro;
            po += (uint64_t)em.m * em.C;
            doff += (uint64_t)em.C * em.C;
            wo += (uint64_t)em.Cp * em.Cp;
            ro += em.C;
        }
        // work items, most expensive first
        std::vector<CountItem> items;
        std::vector<double> cost;
        for (uint32_t e = 0; e < E; ++e) {
            const uint32_t C = Cs[e];
            if (!C) continue;
            // items: full 128 x 128 tiles (mode 1) and voter-parallel sub-blocks (mode 0) for
            // small elections and the partial tiles at the edge.  (Cutting full diagonal
            // tiles into two 64-wide upper triangles and a 64 x 64 square, to skip their
            // lower half, measured slower: the voter-parallel pieces cost more than the half
            // tile they save.)
            const uint32_t K = meta[e].K;
            const double words = (double)((ms[e] + K - 1) / K);
            auto add = [&](uint32_t I0, uint32_t J0, uint32_t ri, uint32_t rj, bool full) {
                items.push_back(CountItem{e, I0, J0, (full ? 1u : 0u) | (ri << 8) | (rj << 16), 0u, ms[e], kPutStore, 0u});
                // per-lane instructions: staging (K loads + packing per word) and 128 per
                // word shared by the G word groups of a voter-parallel item
                const uint32_t nbi = (ri + 7) / 8, nbj = (rj + 7) / 8;
                const uint32_t blocks = I0 == J0 ? nbi * (nbi + 1) / 2 : nbi * nbj;
                const double G = full ? 1.0 : (double)(kCThreads / blocks);
                cost.push_back(words * (K + 4) + words * 128.0 / G + 64.0);
            };
            const uint32_t nb = (C + kCT - 1) / kCT;
            for (uint32_t I = 0; I < nb; ++I)
                for (uint32_t J = I; J < nb; ++J) {
                    const uint32_t I0 = I * kCT, J0 = J * kCT;
                    const uint32_t ri = std::min(kCT, C - I0), rj = std::min(kCT, C - J0);
                    add(I0, J0, ri, rj, ri == kCT && rj == kCT);
                }
        }
        // Items far above a workgroup's fair share (total / kSliceShare) would set the
        // launch's tail, running at a quarter speed beside three other workgroups on their
        // CU: cut them into voter slices that add into a zeroed d atomically.
        constexpr double kSliceShare = 2048.0;
        constexpr uint32_t kMaxSlices = 32;
        double total_cost = 0.0;
        for (double c : cost) total_cost += c;
        const double fair = total_cost / kSliceShare;
        bool any_sliced = false;
        for (size_t i = 0, n0 = items.size(); i < n0; ++i) {
            const uint32_t m = items[i].ve;
            const uint32_t S = (uint32_t)std::min<double>({std::ceil(cost[i] / fair), (double)kMaxSlices, (double)m});
            if (S < 2) continue;
            any_sliced = true;
            const uint32_t step = (m + S - 1) / S;
            const double part = cost[i] / S;
            CountItem base = items[i];
            base.put = kPutAtomic;
            for (uint32_t v0 = 0, k = 0; v0 < m; v0 += step, ++k) {
                CountItem it = base;
                it.vb = v0;
                it.ve = std::min(m, v0 + step);
                if (k == 0) {
                    items[i] = it;
                    cost[i] = part;
                } else {
                    items.push_back(it);
                    cost.push_back(part);
                }
            }
        }
        std::vector<uint32_t> idx(items.size());
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
        std::vector<CountItem> sorted_items(items.size());
        for (size_t i = 0; i < idx.size(); ++i) sorted_items[i] = items[idx[i]];
        std::vector<uint32_t> order(E);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return Cs[a] > Cs[b]; });

        // Elections with at least rounds_min candidates (tuning table, >= 129) are closed by
        // launches shared with the other large elections of the batch, one pair per round
        // (panels, then tiles), on the 16-bit paths; the others by one workgroup each.
        const bool rounds_mode = use16 && !(flags & SE_BATCH_ONE_GROUP);
        const uint32_t rounds_min = tuned_batch_rounds_min_c();
        std::vector<uint32_t> large;
        if (rounds_mode)
            for (uint32_t e = 0; e < E; ++e)
                if (Cs[e] >= rounds_min) {
                    meta[e].piv = (uint32_t)large.size();
                    large.push_back(e);
                }
        std::sort(large.begin(), large.end(), [&](uint32_t a, uint32_t b) { return Cs[a] > Cs[b]; });
        const size_t meta_bytes = round_up(E * sizeof(ElectionMeta), 256);
        const size_t item_bytes = round_up(std::max<size_t>(1, sorted_items.size()) * sizeof(CountItem), 256);
        const size_t order_bytes = round_up((size_t)E * 4, 256);
        char* mbuf = static_cast<char*>(scratch(ctx, kSlotBatchMeta, meta_bytes + item_bytes + order_bytes));
        uint32_t* dbuf = static_cast<uint32_t*>(scratch(ctx, kSlotBatchD, d_total * 4 + 256));
        void* wbuf = scratch(ctx, kSlotWork, wo * (use16 ? 2 : 4) + 256);
        if (!mbuf || !dbuf || !wbuf) return SE_ERR_ALLOC;
        ElectionMeta* dmeta = reinterpret_cast<ElectionMeta*>(mbuf);
        CountItem* ditems = reinterpret_cast<CountItem*>(mbuf + meta_bytes);
        uint32_t* dorder = reinterpret_cast<uint32_t*>(mbuf + meta_bytes + item_bytes);
        if (!hip_ok(hipMemcpyAsync(dmeta, meta.data(), E * sizeof(ElectionMeta), hipMemcpyHostToDevice, s),
                    "Failed to copy data from host to device") ||
            (!sorted_items.empty() &&
             !hip_ok(hipMemcpyAsync(ditems, sorted_items.data(), sorted_items.size() * sizeof(CountItem),
                                    hipMemcpyHostToDevice, s),
                     "Failed to copy data from host to device")) ||
            !hip_ok(hipMemcpyAsync(dorder, order.data(), (size_t)E * 4, hipMemcpyHostToDevice, s),
                    "Failed to copy data from host to device"))
            return SE_ERR_COPY;
        if (any_sliced && !hip_ok(hipMemsetAsync(dbuf, 0, d_total * 4, s), "Failed to clear device memory"))
            return SE_ERR_DEVICE;
        if (!sorted_items.empty() && pos16)
            batch_count_kernel<uint16_t><<<(uint32_t)sorted_items.size(), kCThreads, 0, s>>>(
                static_cast<const uint16_t*>(pos_any), dbuf, dmeta, ditems);
        else if (!sorted_items.empty())
            batch_count_kernel<int32_t><<<(uint32_t)sorted_items.size(), kCThreads, 0, s>>>(
                static_cast<const int32_t*>(pos_any), dbuf, dmeta, ditems);
        SE_CHECK_LAUNCH("batch counter launch failed");
        // work lists of the round-closed elections, built while the counter runs
        uint32_t max_rounds = 0;
        for (uint32_t e : large) max_rounds = std::max(max_rounds, (Cs[e] + T - 1) / T);
        std::vector<uint2> panel_items;
        std::vector<uint4> tile_items;
        std::vector<size_t> panel_off(max_rounds + 1), tile_off(max_rounds + 1);
        for (uint32_t r = 0; r < max_rounds; ++r) {
            panel_off[r] = panel_items.size();
            tile_off[r] = tile_items.size();
            for (uint32_t e : large) {
                if ((Cs[e] + T - 1) / T <= r) continue;
                const uint32_t nbe = meta[e].Cp / kMacro;
                panel_items.push_back(make_uint2(e, 0));
                for (uint32_t I = 0; I < nbe; ++I)
                    for (uint32_t J = 0; J < nbe; ++J) tile_items.push_back(make_uint4(e, I * kMacro, J * kMacro, 0));
            }
        }
        panel_off[max_rounds] = panel_items.size();
        tile_off[max_rounds] = tile_items.size();
        // p0 and tail tile pairs (I <= J) of the round-closed elections: p0 over the padded
        // side, the tail over C x C
        std::vector<uint4> p0_pairs, tail_pairs;
        for (uint32_t e : large) {
            const uint32_t np = meta[e].Cp / kPT, nc = (Cs[e] + kPT - 1) / kPT;
            for (uint32_t I = 0; I < np; ++I)
                for (uint32_t J = I; J < np; ++J) {
                    p0_pairs.push_back(make_uint4(e, I, J, 0));
                    if (J < nc) tail_pairs.push_back(make_uint4(e, I, J, 0));
                }
        }

        uint32_t* pivots = nullptr;
        uint2* dpanels = nullptr;
        uint4* dtiles = nullptr;
        uint32_t* dlarge = nullptr;
        uint4 *dp0_pairs = nullptr, *dtail_pairs = nullptr;
        uint32_t* dwins = nullptr;
        if (!large.empty()) {
            const size_t tb = round_up(tile_items.size() * sizeof(uint4), 256);
            const size_t pb = round_up(panel_items.size() * sizeof(uint2), 256);
            const size_t lb = round_up(large.size() * 4, 256);
            const size_t qb = round_up(p0_pairs.size() * sizeof(uint4), 256);
            const size_t ub = round_up(tail_pairs.size() * sizeof(uint4), 256);
            const size_t wb = round_up(r_total * 4, 256);
            pivots = static_cast<uint32_t*>(scratch(ctx, kSlotBatchPivot, large.size() * 2 * T * T * 4));
            char* rbuf = static_cast<char*>(scratch(ctx, kSlotBatchRounds, tb + pb + lb + qb + ub + wb));
            if (!pivots || !rbuf) return SE_ERR_ALLOC;
            dtiles = reinterpret_cast<uint4*>(rbuf);
            dpanels = reinterpret_cast<uint2*>(rbuf + tb);
            dlarge = reinterpret_cast<uint32_t*>(rbuf + tb + pb);
            dp0_pairs = reinterpret_cast<uint4*>(rbuf + tb + pb + lb);
            dtail_pairs = reinterpret_cast<uint4*>(rbuf + tb + pb + lb + qb);
            dwins = reinterpret_cast<uint32_t*>(rbuf + tb + pb + lb + qb + ub);
            if (!hip_ok(hipMemcpyAsync(dtiles, tile_items.data(), tile_items.size() * sizeof(uint4),
                                       hipMemcpyHostToDevice, s),
                        "Failed to copy data from host to device") ||
                !hip_ok(hipMemcpyAsync(dpanels, panel_items.data(), panel_items.size() * sizeof(uint2),
                                       hipMemcpyHostToDevice, s),
                        "Failed to copy data from host to device") ||
                !hip_ok(hipMemcpyAsync(dlarge, large.data(), large.size() * 4, hipMemcpyHostToDevice, s),
                        "Failed to copy data from host to device") ||
                !hip_ok(hipMemcpyAsync(dp0_pairs, p0_pairs.data(), p0_pairs.size() * sizeof(uint4),
                                       hipMemcpyHostToDevice, s),
                        "Failed to copy data from host to device") ||
                !hip_ok(hipMemcpyAsync(dtail_pairs, tail_pairs.data(), tail_pairs.size() * sizeof(uint4),
                                       hipMemcpyHostToDevice, s),
                        "Failed to copy data from host to device"))
                return SE_ERR_COPY;
        }
        if (timing) (void)hipEventRecord(ev[1], s);
        if (E) {
            uint16_t* W16 = static_cast<uint16_t*>(wbuf);
            auto rounds16 = [&](auto f16_tag) {
                constexpr bool F16 = decltype(f16_tag)::value;
                const uint32_t nl = (uint32_t)large.size();
                if (nl)
                    batch_p0_pairs_kernel<F16><<<(uint32_t)p0_pairs.size(), kThreads, 0, s>>>(dbuf, W16, dmeta,
                                                                                             dp0_pairs, dwins);
                batch_closure_kernel<uint16_t, 64, F16><<<E, kThreads, 0, s>>>(dbuf, W16, paths, ranking, dmeta, dorder);
                if (!nl) return;
                for (uint32_t r = 0; r < max_rounds; ++r) {
                    const uint32_t np = (uint32_t)(panel_off[r + 1] - panel_off[r]);
                    const uint32_t nt = (uint32_t)(tile_off[r + 1] - tile_off[r]);
                    batch_panel_kernel<64, F16><<<np, kThreads, 0, s>>>(W16, dmeta, dpanels + panel_off[r], r, pivots);
                    batch_tile_kernel<64, F16><<<nt, kThreads, 0, s>>>(W16, dmeta, dtiles + tile_off[r], r);
                }
                batch_commit_kernel<64><<<nl, kThreads, 0, s>>>(W16, dmeta, dlarge, pivots);
                batch_tail_pairs_kernel<F16>
                    <<<(uint32_t)tail_pairs.size(), kThreads, 0, s>>>(W16, paths, dmeta, dtail_pairs, dwins);
                batch_rank_kernel<<<nl, kThreads, 0, s>>>(dmeta, dlarge, dwins, ranking);
            };
            if (use16 && max_m <= kBatchF16MaxVoters)
                rounds16(std::true_type{});
            else if (use16)
                rounds16(std::false_type{});
            else
                batch_closure32_kernel<<<E, kThreads, 0, s>>>(dbuf, static_cast<uint32_t*>(wbuf), paths, ranking,
                                                              dmeta, dorder);
        }
        SE_CHECK_LAUNCH("batch closure launch failed");
        if (timing) (void)hipEventRecord(ev[2], s);
        // the host metadata vectors die here; the copies above must have left them first
        if (!hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully")) return SE_ERR_DEVICE;
    }
    if (timing) {
        if (!hip_ok(hipEventSynchronize(ev[2]), "CUDA operations did not complete successfully")) return SE_ERR_DEVICE;
        float a = 0.f, b = 0.f;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        timing->count_ms = a;
        timing->closure_ms = b;
        timing->total_ms = a + b;
        timing->elections = E;
        timing->cells = 0;
        for (uint32_t e = 0; e < E; ++e) timing->cells += (uint64_t)Cs[e] * Cs[e] * Cs[e];
        timing->pair_votes = 0;
        for (uint32_t e = 0; e < E; ++e) timing->pair_votes += (uint64_t)ms[e] * Cs[e] * (Cs[e] ? Cs[e] - 1 : 0) / 2;
        for (auto x : ev) (void)hipEventDestroy(x);
    }
    return SE_OK;
}

}  // namespace se
