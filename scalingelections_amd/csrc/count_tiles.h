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

}  // namespace se
