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

#include "se_internal.h"

namespace se {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kMacro = 128;  // macro tile side of the independent kernel
constexpr uint32_t kThreads = 256;

__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u32(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ u16x2 pk_min(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ u16x2 pk_max(u16x2 a, u16x2 b) { return __builtin_elementwise_max(a, b); }

__device__ __forceinline__ uint32_t word(const uint4& v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// ---------------------------------------------------------------------------------------
// p0 init.  64 x 64 tiles; the transposed operand d[j][i] goes through LDS so both reads
// are coalesced.

constexpr uint32_t kInitTile = 64;

// max over p0 of the n x n domain -> atomicMax(*out).  Decides the 16-bit eligibility.
__global__ __launch_bounds__(256) void p0_max_kernel(const uint32_t* __restrict__ d, uint32_t n,
                                                     size_t ld, uint32_t* __restrict__ out) {
    __shared__ uint32_t tr[kInitTile][kInitTile + 1];
    __shared__ uint32_t red[kThreads / 64];
    const uint32_t I0 = blockIdx.y * kInitTile, J0 = blockIdx.x * kInitTile;
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // tr[b][a] = d[J0 + b][I0 + a]
    for (uint32_t b = ty; b < kInitTile; b += 4) {
        uint32_t r = J0 + b, c = I0 + tx;
        tr[b][tx] = (r < n && c < n) ? d[(size_t)r * ld + c] : 0u;
    }
    __syncthreads();
    uint32_t m = 0;
    for (uint32_t a = ty; a < kInitTile; a += 4) {
        uint32_t i = I0 + a, j = J0 + tx;
        if (i < n && j < n && i != j) {
            uint32_t v = d[(size_t)i * ld + j];
            if (v > tr[tx][a]) m = max(m, v);
        }
    }
    // wave reduction then block reduction
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = max(max(red[0], red[1]), max(red[2], red[3]));
        if (r) atomicMax(out, r);
    }
}

// Writes p0 into the padded work matrix W (N x N, row pitch ldw, element E).
template <typename E>
__global__ __launch_bounds__(256) void p0_write_kernel(const uint32_t* __restrict__ d, uint32_t n,
                                                       size_t ld, E* __restrict__ W, uint32_t N,
                                                       size_t ldw) {
    __shared__ uint32_t tr[kInitTile][kInitTile + 1];
    const uint32_t I0 = blockIdx.y * kInitTile, J0 = blockIdx.x * kInitTile;
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (uint32_t b = ty; b < kInitTile; b += 4) {
        uint32_t r = J0 + b, c = I0 + tx;
        tr[b][tx] = (r < n && c < n) ? d[(size_t)r * ld + c] : 0u;
    }
    __syncthreads();
    for (uint32_t a = ty; a < kInitTile; a += 4) {
        uint32_t i = I0 + a, j = J0 + tx;
        if (i >= N || j >= N) continue;
        uint32_t v = 0;
        if (i < n && j < n && i != j) {
            uint32_t x = d[(size_t)i * ld + j];
            if (x > tr[tx][a]) v = x;
        }
        W[(size_t)i * ldw + j] = (E)v;
    }
}

// ---------------------------------------------------------------------------------------
// Pivot + panels, fused.  Every workgroup closes the pivot tile P = W[K][K] (T x T) itself,
// in registers: lane t owns an R x R block (R = T / 16) at rows R (t / 16), columns
// R (t % 16).  At step k the owners of row k and column k publish them through LDS
// (double-buffered by k parity, so one barrier per step suffices) and every lane updates
// its block.  The cells read in step k cannot change in step k (min(p[i][k], p[k][k]) <=
// p[i][k]), so the step is race-free.
// With P closed, the row panel (K, J) becomes, in one shot (split any path at its last
// vertex in K),
//     new[i][j] = max(old[i][j], max_k min(P[i][k], old[k][j])),
// and the column panel (I, K)
//     new[i][j] = max(old[i][j], max_k min(old[i][k], P[k][j])).
// Blocks [0, nb) take 128-column chunks of the row panel, [nb, 2 nb) 128-row chunks of
// the column panel; cells of (K, K) are not written here.  Block 0 stores the closed P
// into pivot_out; the next round's block 0 (or commit_pivot_kernel after the last round)
// copies it into W -- nothing reads W[K][K] before then: the independent kernel of this
// round only uses it for pivot-row/column cells, which it does not store.

template <typename E, uint32_t T>
__global__ __launch_bounds__(256) void pivot_panel_kernel(E* __restrict__ W, size_t ldw, uint32_t kb,
                                                          uint32_t nb, uint32_t* __restrict__ pivot_out,
                                                          const uint32_t* __restrict__ pivot_prev,
                                                          uint32_t kb_prev) {
    constexpr uint32_t R = T / 16;  // pivot rows/columns per lane
    constexpr uint32_t RP = T / 16; // panel: row-panel rows per lane / column-panel columns per lane
    __shared__ uint32_t P[T][T + 1];
    __shared__ uint32_t S[kMacro * (T + 1)];  // the chunk: T x 128 (row) or 128 x (T + 1) (column)
    __shared__ uint32_t rowbuf[2][T], colbuf[2][T];
    const uint32_t t = threadIdx.x, ty = t >> 4, tx = t & 15;
    const bool row_panel = blockIdx.x < nb;
    const uint32_t c0 = (row_panel ? blockIdx.x : blockIdx.x - nb) * kMacro;

    // hand the previous round's closed pivot to W
    if (blockIdx.x == 0 && pivot_prev != nullptr) {
        for (uint32_t q = t; q < T * T; q += kThreads)
            W[(size_t)(kb_prev + q / T) * ldw + kb_prev + q % T] = (E)pivot_prev[q];
    }

    // stage the panel chunk (old values) while the pivot closes
    if (row_panel) {
        for (uint32_t q = t; q < T * kMacro; q += kThreads) {
            uint32_t i = q / kMacro, j = q % kMacro;
            S[i * kMacro + j] = W[(size_t)(kb + i) * ldw + c0 + j];
        }
    } else {
        for (uint32_t q = t; q < kMacro * T; q += kThreads) {
            uint32_t i = q / T, j = q % T;
            S[i * (T + 1) + j] = W[(size_t)(c0 + i) * ldw + kb + j];
        }
    }

    // pivot closure in registers
    const uint32_t pr = R * ty, pc = R * tx;
    uint32_t cell[R][R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r)
#pragma unroll
        for (uint32_t c = 0; c < R; ++c) cell[r][c] = W[(size_t)(kb + pr + r) * ldw + kb + pc + c];
#pragma unroll
    for (uint32_t k = 0; k < T; ++k) {
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
        for (uint32_t c = 0; c < R; ++c) {
            P[pr + r][pc + c] = cell[r][c];
            if (blockIdx.x == 0) pivot_out[(pr + r) * T + pc + c] = cell[r][c];
        }
    __syncthreads();

    if (row_panel) {
        // lane: rows RP ty .. +RP-1 of the chunk, columns 8 tx .. 8 tx + 7
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
                for (uint32_t c = 0; c < 8; ++c)
                    W[(size_t)(kb + RP * ty + r) * ldw + c0 + 8 * tx + c] = (E)acc[r][c];
        }
    } else {
        // lane: rows 8 ty .. 8 ty + 7 of the chunk, columns RP tx .. +RP-1
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
}

// Copies the last round's closed pivot tile into W.
template <typename E, uint32_t T>
__global__ __launch_bounds__(256) void commit_pivot_kernel(E* __restrict__ W, size_t ldw, uint32_t kb,
                                                           const uint32_t* __restrict__ pivot) {
    for (uint32_t q = threadIdx.x; q < T * T; q += kThreads)
        W[(size_t)(kb + q / T) * ldw + kb + q % T] = (E)pivot[q];
}

// ---------------------------------------------------------------------------------------
// Independent tiles, 16-bit.  Workgroup = 256 lanes = 128 x 128 macro tile; lane (ty, tx)
// owns rows ty + 16 r (r < 8) and columns 8 tx .. 8 tx + 7, i.e. 8 x 4 packed u16 pairs.
// Rows are interleaved by 16 so the four row addresses a wave reads from the A strip fall
// in different LDS banks; the B strip row a wave reads is 256 contiguous bytes.
// Per pivot step: 32 v_pk_min_u16 + 32 v_pk_max_u16 for 64 cells, a[i][k] broadcast to
// both halves by op_sel, and two ds_read_b128 per lane per step.

template <uint32_t T>
__global__ __launch_bounds__(256) void indep_u16_kernel(uint16_t* __restrict__ W, size_t ldw,
                                                        uint32_t kb, uint32_t nb) {
    constexpr uint32_t AP = T + 8;  // A strip pitch (u16), keeps 16-byte alignment
    __shared__ __align__(16) uint16_t As[kMacro * AP];
    __shared__ __align__(16) uint16_t Bs[T * kMacro];
    const uint32_t t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const uint32_t I0 = blockIdx.y * kMacro, J0 = blockIdx.x * kMacro;
    (void)nb;

    uint4 c[8];
    const uint16_t* cp = W + (size_t)(I0 + ty) * ldw + J0 + tx * 8;
#pragma unroll
    for (int r = 0; r < 8; ++r) c[r] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw);

    for (uint32_t q = t; q < kMacro * (T / 8); q += kThreads) {
        uint32_t row = q / (T / 8), ch = q % (T / 8);
        *reinterpret_cast<uint4*>(As + row * AP + ch * 8) =
            *reinterpret_cast<const uint4*>(W + (size_t)(I0 + row) * ldw + kb + ch * 8);
    }
    for (uint32_t q = t; q < T * 16; q += kThreads) {
        uint32_t row = q >> 4, ch = q & 15;
        *reinterpret_cast<uint4*>(Bs + row * kMacro + ch * 8) =
            *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + J0 + ch * 8);
    }
    __syncthreads();

    u16x2 acc[8][4];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][q] = as_u16x2(word(c[r], q));

#pragma unroll 1
    for (uint32_t k0 = 0; k0 < T; k0 += 8) {
        uint4 a4[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            a4[r] = *reinterpret_cast<const uint4*>(As + (ty + 16 * r) * AP + k0);
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const uint4 b4 = *reinterpret_cast<const uint4*>(Bs + (k0 + kk) * kMacro + tx * 8);
            const u16x2 b[4] = {as_u16x2(b4.x), as_u16x2(b4.y), as_u16x2(b4.z), as_u16x2(b4.w)};
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const u16x2 a2 = as_u16x2(word(a4[r], kk >> 1));
                const u16x2 av = (kk & 1) ? a2.yy : a2.xx;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[r][q] = pk_max(acc[r][q], pk_min(av, b[q]));
            }
        }
    }

    if (J0 + tx * 8 - kb < T) return;  // these columns belong to the pivot block
    uint16_t* op = W + (size_t)(I0 + ty) * ldw + J0 + tx * 8;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (I0 + ty + 16 * r - kb < T) continue;  // pivot rows are final already
        uint4 v;
        v.x = as_u32(acc[r][0]);
        v.y = as_u32(acc[r][1]);
        v.z = as_u32(acc[r][2]);
        v.w = as_u32(acc[r][3]);
        *reinterpret_cast<uint4*>(op + (size_t)(16 * r) * ldw) = v;
    }
}

// Independent tiles, 32-bit.  Lane (ty, tx) owns rows ty + 16 r (r < 8) and columns
// 4 tx .. 4 tx + 3 and 64 + 4 tx .. 64 + 4 tx + 3 (each half is 256 contiguous bytes per
// row across 16 lanes, so B-strip reads and global accesses stay conflict-free and
// coalesced).  Two pivot steps fold into one v_max3_u32: 1.5 VALU per cell update.

template <uint32_t T>
__global__ __launch_bounds__(256) void indep_u32_kernel(uint32_t* __restrict__ W, size_t ldw,
                                                        uint32_t kb, uint32_t nb) {
    constexpr uint32_t AP = T + 4;  // A strip pitch (u32)
    __shared__ __align__(16) uint32_t As[kMacro * AP];
    __shared__ __align__(16) uint32_t Bs[T * kMacro];
    const uint32_t t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const uint32_t I0 = blockIdx.y * kMacro, J0 = blockIdx.x * kMacro;
    (void)nb;

    uint4 c[8][2];
    const uint32_t* cp = W + (size_t)(I0 + ty) * ldw + J0 + tx * 4;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        c[r][0] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw);
        c[r][1] = *reinterpret_cast<const uint4*>(cp + (size_t)(16 * r) * ldw + 64);
    }
    for (uint32_t q = t; q < kMacro * (T / 4); q += kThreads) {
        uint32_t row = q / (T / 4), ch = q % (T / 4);
        *reinterpret_cast<uint4*>(As + row * AP + ch * 4) =
            *reinterpret_cast<const uint4*>(W + (size_t)(I0 + row) * ldw + kb + ch * 4);
    }
    for (uint32_t q = t; q < T * 32; q += kThreads) {
        uint32_t row = q >> 5, ch = q & 31;
        *reinterpret_cast<uint4*>(Bs + row * kMacro + ch * 4) =
            *reinterpret_cast<const uint4*>(W + (size_t)(kb + row) * ldw + J0 + ch * 4);
    }
    __syncthreads();

    uint32_t acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[r][q] = word(c[r][0], q);
            acc[r][4 + q] = word(c[r][1], q);
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

    const bool lo_pivot = J0 + tx * 4 - kb < T, hi_pivot = J0 + 64 + tx * 4 - kb < T;
    uint32_t* op = W + (size_t)(I0 + ty) * ldw + J0 + tx * 4;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (I0 + ty + 16 * r - kb < T) continue;
        if (!lo_pivot)
            *reinterpret_cast<uint4*>(op + (size_t)(16 * r) * ldw) =
                make_uint4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
        if (!hi_pivot)
            *reinterpret_cast<uint4*>(op + (size_t)(16 * r) * ldw + 64) =
                make_uint4(acc[r][4], acc[r][5], acc[r][6], acc[r][7]);
    }
}

// ---------------------------------------------------------------------------------------
// Output: out[i][j] = W[i][j] widened to u32, 0 on the diagonal, padding dropped.

template <typename E>
__global__ __launch_bounds__(256) void output_kernel(const E* __restrict__ W, size_t ldw, uint32_t n,
                                                     uint32_t* __restrict__ out, size_t ldo) {
    const uint32_t i = blockIdx.x;
    const E* src = W + (size_t)i * ldw;
    uint32_t* dst = out + (size_t)i * ldo;
    for (uint32_t j = threadIdx.x; j < n; j += kThreads) dst[j] = (i == j) ? 0u : (uint32_t)src[j];
}

// Zeroes a few words with a kernel (keeps the pipeline free of memset nodes).
__global__ void zero_words_kernel(uint32_t* p, uint32_t count) {
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0u;
}

// ---------------------------------------------------------------------------------------
// Host driver.

namespace {

struct EventTimer {
    bool on;
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    EventTimer(bool enabled, hipStream_t stream) : on(enabled), s(stream) {}
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

template <typename E, uint32_t T>
int run_rounds(E* W, size_t ldw, uint32_t n, uint32_t N, uint32_t* pivot_buf, hipStream_t s, EventTimer& tm,
               uint32_t* rounds_out) {
    const uint32_t nb = N / kMacro;
    const uint32_t rounds = (n + T - 1) / T;
    for (uint32_t r = 0; r < rounds; ++r) {
        if (g_interrupted.load(std::memory_order_relaxed)) {
            set_error("Stopped by signal");
            return SE_ERR_INTERRUPTED;
        }
        const uint32_t kb = r * T;
        uint32_t* out = pivot_buf + (r & 1) * T * T;
        const uint32_t* prev = r ? pivot_buf + ((r - 1) & 1) * T * T : nullptr;
        tm.mark(kKindPanel);
        pivot_panel_kernel<E, T><<<2 * nb, kThreads, 0, s>>>(W, ldw, kb, nb, out, prev, kb - T);
        tm.mark(kKindPanel);
        tm.mark(kKindIndep);
        if constexpr (sizeof(E) == 2)
            indep_u16_kernel<T><<<dim3(nb, nb), kThreads, 0, s>>>(W, ldw, kb, nb);
        else
            indep_u32_kernel<T><<<dim3(nb, nb), kThreads, 0, s>>>(W, ldw, kb, nb);
        tm.mark(kKindIndep);
        SE_CHECK_LAUNCH("closure round launch failed");
    }
    if (rounds) {
        tm.mark(kKindPanel);
        commit_pivot_kernel<E, T><<<1, kThreads, 0, s>>>(W, ldw, (rounds - 1) * T,
                                                          pivot_buf + ((rounds - 1) & 1) * T * T);
        tm.mark(kKindPanel);
        SE_CHECK_LAUNCH("closure round launch failed");
    }
    *rounds_out = rounds;
    return SE_OK;
}

}  // namespace

// Device pipeline: d (device, n x n u32, row stride ld) -> out (device, n x n u32, ldo).
int closure_device_impl(DeviceContext* ctx, const uint32_t* d, uint32_t n, size_t ld, uint32_t* out,
                        size_t ldo, uint32_t flags, hipStream_t s, se_timing* timing) {
    if (n == 0) {
        if (timing) *timing = se_timing{};
        return SE_OK;
    }
    const uint32_t N = (uint32_t)round_up(n, kMacro);
    const size_t ldw = N;
    const bool want_timing = timing != nullptr;
    EventTimer tm((flags & SE_FLAG_TIME_KERNELS) != 0, s);

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (want_timing) {
        if (!hip_ok(hipEventCreate(&e0), "Failed to create events") ||
            !hip_ok(hipEventCreate(&e1), "Failed to create events"))
            return SE_ERR_DEVICE;
        (void)hipEventRecord(e0, s);
    }

    // width decision: exact criterion max(p0) <= 65535
    bool use16;
    if (flags & SE_FLAG_FORCE_U32) {
        use16 = false;
    } else {
        uint32_t* dmax = static_cast<uint32_t*>(scratch(ctx, kSlotSmall, 256));
        if (!dmax) return SE_ERR_ALLOC;
        uint32_t hmax = 0;
        tm.mark(kKindInit);
        zero_words_kernel<<<1, 64, 0, s>>>(dmax, 1);
        dim3 g((n + kInitTile - 1) / kInitTile, (n + kInitTile - 1) / kInitTile);
        p0_max_kernel<<<g, kThreads, 0, s>>>(d, n, ld, dmax);
        tm.mark(kKindInit);
        SE_CHECK_LAUNCH("p0 init launch failed");
        if (!hip_ok(hipMemcpyAsync(&hmax, dmax, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                    "Failed to copy data from device to host") ||
            !hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully"))
            return SE_ERR_COPY;
        use16 = hmax <= 0xFFFFu;
        if ((flags & SE_FLAG_FORCE_U16) && !use16) {
            set_error("strongest-path values exceed 65535; the 16-bit closure is not exact here");
            return SE_ERR_RANGE;
        }
    }

    const size_t esz = use16 ? 2 : 4;
    void* W = scratch(ctx, kSlotWork, (size_t)N * ldw * esz);
    if (!W) return SE_ERR_ALLOC;

    dim3 gi(N / kInitTile, N / kInitTile);
    tm.mark(kKindInit);
    if (use16)
        p0_write_kernel<uint16_t><<<gi, kThreads, 0, s>>>(d, n, ld, static_cast<uint16_t*>(W), N, ldw);
    else
        p0_write_kernel<uint32_t><<<gi, kThreads, 0, s>>>(d, n, ld, static_cast<uint32_t*>(W), N, ldw);
    tm.mark(kKindInit);
    SE_CHECK_LAUNCH("p0 init launch failed");

    // pivot depth: deeper pivots halve the HBM traffic per cell update of the independent
    // kernel (2 s / T bytes per update); 64 fits the 16-bit strips in 34 KB of LDS
    uint32_t T = use16 ? 64 : 32;
    if (flags & SE_FLAG_PIVOT32) T = 32;
    if (flags & SE_FLAG_PIVOT64) T = 64;
    uint32_t* pivot_buf = static_cast<uint32_t*>(scratch(ctx, kSlotPivot, 2 * 64 * 64 * sizeof(uint32_t)));
    if (!pivot_buf) return SE_ERR_ALLOC;
    uint32_t rounds = 0;
    int rc;
    if (use16)
        rc = T == 64 ? run_rounds<uint16_t, 64>(static_cast<uint16_t*>(W), ldw, n, N, pivot_buf, s, tm, &rounds)
                     : run_rounds<uint16_t, 32>(static_cast<uint16_t*>(W), ldw, n, N, pivot_buf, s, tm, &rounds);
    else
        rc = T == 64 ? run_rounds<uint32_t, 64>(static_cast<uint32_t*>(W), ldw, n, N, pivot_buf, s, tm, &rounds)
                     : run_rounds<uint32_t, 32>(static_cast<uint32_t*>(W), ldw, n, N, pivot_buf, s, tm, &rounds);
    if (rc != SE_OK) return rc;

    dim3 go(n);
    tm.mark(kKindOut);
    if (use16)
        output_kernel<uint16_t><<<go, kThreads, 0, s>>>(static_cast<uint16_t*>(W), ldw, n, out, ldo);
    else
        output_kernel<uint32_t><<<go, kThreads, 0, s>>>(static_cast<uint32_t*>(W), ldw, n, out, ldo);
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
        t.padded_n = N;
        t.indep_cells = (uint64_t)rounds * (uint64_t)(N - T) * (uint64_t)(N - T) * T;
        *timing = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return SE_OK;
}

}  // namespace se
