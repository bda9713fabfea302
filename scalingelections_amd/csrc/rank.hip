// rank.hip -- win counts of the Schulze ranking on the GPU.
//
// Reference: get_winner_and_ranking, scaling_elections.py:287-290:
//     wins[i] = #{j != i : p[i][j] > p[j][i]}
// One workgroup per 64-row block walks all 64-column blocks; the transposed operand
// p[j][i] is staged through LDS so both reads are coalesced.  No atomics: each row's
// count is owned by one workgroup.  The stable descending sort (:292) is done on the
// host over n words.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "se_internal.h"

namespace se {

namespace {

constexpr uint32_t kRT = 64;

__global__ __launch_bounds__(256) void wins_kernel(const uint32_t* __restrict__ p, uint32_t n, size_t ld,
                                                   uint32_t* __restrict__ wins) {
    __shared__ uint32_t tr[kRT][kRT + 1];
    __shared__ uint32_t part[4][kRT];
    const uint32_t I0 = blockIdx.x * kRT;
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // lane (ty, tx) counts for rows I0 + a (a = ty, ty + 4, ...) at column J0 + tx, then the
    // row totals are reduced across tx with wave shuffles.
    uint32_t cnt[kRT / 4];
#pragma unroll
    for (uint32_t u = 0; u < kRT / 4; ++u) cnt[u] = 0;
    for (uint32_t J0 = 0; J0 < n; J0 += kRT) {
        __syncthreads();
        for (uint32_t b = ty; b < kRT; b += 4) {  // tr[b][a] = p[J0 + b][I0 + a]
            uint32_t r = J0 + b, c = I0 + tx;
            tr[b][tx] = (r < n && c < n) ? p[(size_t)r * ld + c] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kRT / 4; ++u) {
            uint32_t a = ty + 4 * u;
            uint32_t i = I0 + a, j = J0 + tx;
            if (i < n && j < n && i != j) cnt[u] += p[(size_t)i * ld + j] > tr[tx][a] ? 1u : 0u;
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < kRT / 4; ++u) {
        uint32_t c = cnt[u];
        for (int off = 32; off > 0; off >>= 1) c += (uint32_t)__shfl_xor((int)c, off, 64);
        if (tx == 0) part[ty][u] = c;
    }
    __syncthreads();
    if (threadIdx.x < kRT) {
        uint32_t a = threadIdx.x;
        uint32_t i = I0 + a;
        if (i < n) wins[i] = part[a % 4][a / 4];
    }
}

}  // namespace

int wins_device_impl(DeviceContext* ctx, const uint32_t* p, uint32_t n, size_t ld, uint32_t* wins,
                     hipStream_t s) {
    (void)ctx;
    if (n == 0) return SE_OK;
    wins_kernel<<<(n + kRT - 1) / kRT, 256, 0, s>>>(p, n, ld, wins);
    SE_CHECK_LAUNCH("wins launch failed");
    return SE_OK;
}

}  // namespace se
