// count_tiles.h -- workgroup-level pairwise preference counting, shared by the
// single-election counter (count.hip) and the batched counter (batch.hip).
//
// Reference: populate_preferences_from_ranking, scaling_elections.py:37-48.  With ballots
// given as position arrays (positions[v][c] = place of candidate c), every voter adds one
// to d[i][j] exactly when positions[v][i] < positions[v][j]; for a strict order per voter
// d[j][i] = m - d[i][j] (i != j).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace se {

constexpr uint32_t kCT = 128;     // tile side
constexpr uint32_t kVC = 32;      // voters staged per LDS round
constexpr uint32_t kCThreads = 256;

// Lane (ty, tx) owns rows {4 ty + a, 64 + 4 ty + a} and columns {4 tx + b, 64 + 4 tx + b}
// (a, b < 4): each half-row a wave reads from LDS is 256 contiguous bytes.
// d tile (I, J), I <= J, of one election: positions pos (m x n, row pitch ldp) -> d (pitch
// ldd).  Pi/Pj: LDS staging of kVC voters x kCT candidates each.
__device__ __forceinline__ void count_tile(const int32_t* __restrict__ pos, uint32_t m, uint32_t n, size_t ldp,
                                           uint32_t* __restrict__ d, size_t ldd, uint32_t I, uint32_t J,
                                           int32_t (*Pi)[kCT], int32_t (*Pj)[kCT]) {
    const uint32_t I0 = I * kCT, J0 = J * kCT;
    const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;

    uint32_t acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0;

    for (uint32_t v0 = 0; v0 < m; v0 += kVC) {
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < kVC * kCT; q += kCThreads) {
            uint32_t v = q / kCT, c = q % kCT;
            bool ok = v0 + v < m;
            Pi[v][c] = (ok && I0 + c < n) ? pos[(size_t)(v0 + v) * ldp + I0 + c] : 0;
            Pj[v][c] = (ok && J0 + c < n) ? pos[(size_t)(v0 + v) * ldp + J0 + c] : 0;
        }
        __syncthreads();
        const uint32_t vend = min(kVC, m - v0);
        for (uint32_t v = 0; v < vend; ++v) {
            const int4 i0 = *reinterpret_cast<const int4*>(&Pi[v][4 * ty]);
            const int4 i1 = *reinterpret_cast<const int4*>(&Pi[v][64 + 4 * ty]);
            const int4 j0 = *reinterpret_cast<const int4*>(&Pj[v][4 * tx]);
            const int4 j1 = *reinterpret_cast<const int4*>(&Pj[v][64 + 4 * tx]);
            const int32_t pi[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
            const int32_t pj[8] = {j0.x, j0.y, j0.z, j0.w, j1.x, j1.y, j1.z, j1.w};
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] += (uint32_t)(pi[a] < pj[b]);
        }
    }

#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const uint32_t i = I0 + (a < 4 ? 4 * ty + a : 64 + 4 * ty + (a - 4));
        if (i >= n) continue;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint32_t j = J0 + (b < 4 ? 4 * tx + b : 64 + 4 * tx + (b - 4));
            if (j >= n) continue;
            d[(size_t)i * ldd + j] = acc[a][b];
            if (I != J) d[(size_t)j * ldd + i] = m - acc[a][b];
        }
    }
}

// Voter-parallel counting of one sub-block of d: rows [I0, I0 + ri) x columns [J0, J0 + rj)
// (ri, rj <= 128) of an election with n candidates.  The sub-block is cut into 8 x 8 blocks
// and the 256 lanes into G = 256 / blocks voter groups; group g counts voters g, g + G, ...
// and the partial blocks are summed through LDS (`red`, at least 256 * 8 words).  On a
// diagonal sub-block (I0 == J0) only blocks bi <= bj are counted; a strict order per voter
// gives the lower ones by complement, d[j][i] = m - d[i][j] (the diagonal counts itself: 0).
// Off-diagonal sub-blocks write d[i][j] and, by complement, d[j][i].
__device__ __forceinline__ void count_block(const int32_t* __restrict__ pos, uint32_t m, uint32_t n, size_t ldp,
                                            uint32_t I0, uint32_t J0, uint32_t ri, uint32_t rj,
                                            uint32_t* __restrict__ d, size_t ldd, int32_t* red) {
    const uint32_t t = threadIdx.x;
    const bool diag = I0 == J0;
    const uint32_t nbi = (ri + 7) / 8, nbj = (rj + 7) / 8;
    const uint32_t blocks = diag ? nbi * (nbi + 1) / 2 : nbi * nbj;
    const uint32_t G = kCThreads / blocks;
    const uint32_t b = t % blocks, g = t / blocks;
    uint32_t bi, bj;
    if (diag) {  // b-th pair bi <= bj, row-major over the upper triangle of blocks
        uint32_t r = b;
        bi = 0;
        while (r >= nbi - bi) {
            r -= nbi - bi;
            ++bi;
        }
        bj = bi + r;
    } else {
        bi = b / nbj;
        bj = b % nbj;
    }
    uint32_t acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[a][c] = 0;
    if (g < G) {
        for (uint32_t v = g; v < m; v += G) {
            const int32_t* row = pos + (size_t)v * ldp;
            int32_t pi[8], pj[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const uint32_t ci = 8 * bi + a, cj = 8 * bj + a;
                pi[a] = ci < ri ? row[I0 + ci] : 0;
                pj[a] = cj < rj ? row[J0 + cj] : 0;
            }
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[a][c] += (uint32_t)(pi[a] < pj[c]);
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 8; ++c) red[t * 8 + c] = (int32_t)acc[a][c];
        __syncthreads();
        for (uint32_t q = t; q < blocks * 8; q += kCThreads) {
            const uint32_t bb = q / 8, c = q % 8;
            uint32_t sum = 0;
            for (uint32_t gg = 0; gg < G; ++gg) sum += (uint32_t)red[(gg * blocks + bb) * 8 + c];
            uint32_t qi, qj;  // block coordinates of bb
            if (diag) {
                uint32_t r = bb;
                qi = 0;
                while (r >= nbi - qi) {
                    r -= nbi - qi;
                    ++qi;
                }
                qj = qi + r;
            } else {
                qi = bb / nbj;
                qj = bb % nbj;
            }
            const uint32_t li = 8 * qi + a, lj = 8 * qj + c;
            if (li >= ri || lj >= rj) continue;
            const uint32_t i = I0 + li, j = J0 + lj;
            if (diag && qi == qj) {
                d[(size_t)i * ldd + j] = sum;  // diagonal block: both triangles counted
            } else {
                d[(size_t)i * ldd + j] = sum;
                d[(size_t)j * ldd + i] = m - sum;
            }
        }
    }
    (void)n;
}

// Whole election with n <= 128 candidates (one diagonal sub-block).
__device__ __forceinline__ void count_small(const int32_t* __restrict__ pos, uint32_t m, uint32_t n, size_t ldp,
                                            uint32_t* __restrict__ d, size_t ldd, int32_t* red) {
    count_block(pos, m, n, ldp, 0, 0, n, n, d, ldd, red);
}

}  // namespace se
