// rank.hip -- win counts of the Schulze ranking on the GPU.
//
// Reference: get_winner_and_ranking, scaling_elections.py:287-290:
//     wins[i] = #{j != i : p[i][j] > p[j][i]}
// Workgroup (I, seg) takes the 64-row block I against column blocks [seg * per, (seg + 1)
// * per): the row blocks alone are too few workgroups for the GPU (64 at 4096 candidates,
// 557 us), so the columns are split into segments for about 2048 workgroups.  The
// transposed operand p[j][i] is staged through LDS so both reads are coalesced; each
// workgroup adds its rows' counts atomically into the zeroed wins.  The stable descending
// sort (:292) is done on the host over n words.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "se_internal.h"

namespace se {

namespace {

constexpr uint32_t kRT = 64;

__global__ __launch_bounds__(256) void wins_kernel(const uint32_t* __restrict__ p, uint32_t n, size_t ld,
                                                   uint32_t per, uint32_t* __restrict__ wins) {
    __shared__ uint32_t tr[kRT][kRT + 1];
    __shared__ uint32_t part[4][kRT];
    const uint32_t I0 = blockIdx.x * kRT;
    const uint32_t nbc = (n + kRT - 1) / kRT;
    const uint32_t jb0 = blockIdx.y * per, jb1 = min(nbc, jb0 + per);
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // lane (ty, tx) counts for rows I0 + a (a = ty, ty + 4, ...) at column J0 + tx, then the
    // row totals are reduced across tx with wave shuffles
    uint32_t cnt[kRT / 4];
#pragma unroll
    for (uint32_t u = 0; u < kRT / 4; ++u) cnt[u] = 0;
    for (uint32_t jb = jb0; jb < jb1; ++jb) {
        const uint32_t J0 = jb * kRT;
        // unconditional loads at clamped indices (cells outside n x n are not counted), so
        // a thread's loads issue together
        uint32_t x[kRT / 4], y[kRT / 4];
#pragma unroll
        for (uint32_t u = 0; u < kRT / 4; ++u) {
            const uint32_t a = ty + 4 * u;
            x[u] = p[(size_t)min(I0 + a, n - 1) * ld + min(J0 + tx, n - 1)];
            y[u] = p[(size_t)min(J0 + a, n - 1) * ld + min(I0 + tx, n - 1)];  // p[J0 + a][I0 + tx]
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kRT / 4; ++u) tr[ty + 4 * u][tx] = y[u];
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kRT / 4; ++u) {
            const uint32_t a = ty + 4 * u;
            const uint32_t i = I0 + a, j = J0 + tx;
            if (i < n && j < n && i != j) cnt[u] += x[u] > tr[tx][a] ? 1u : 0u;  // p[j][i] = tr[tx][a]
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
        const uint32_t a = threadIdx.x, i = I0 + a;
        const uint32_t c = part[a % 4][a / 4];
        if (i < n && c) atomicAdd(&wins[i], c);
    }
}

}  // namespace

int wins_device_impl(DeviceContext* ctx, const uint32_t* p, uint32_t n, size_t ld, uint32_t* wins,
                     hipStream_t s) {
    (void)ctx;
    if (n == 0) return SE_OK;
    const uint32_t nb = (n + kRT - 1) / kRT;
    const uint32_t segs = std::max<uint32_t>(1, std::min<uint32_t>(nb, (2048 + nb - 1) / nb));
    const uint32_t per = (nb + segs - 1) / segs;
    if (!hip_ok(hipMemsetAsync(wins, 0, (size_t)n * sizeof(uint32_t), s), "Failed to clear device memory"))
        return SE_ERR_DEVICE;
    wins_kernel<<<dim3(nb, (nb + per - 1) / per), 256, 0, s>>>(p, n, ld, per, wins);
    SE_CHECK_LAUNCH("wins launch failed");
    return SE_OK;
}

}  // namespace se
