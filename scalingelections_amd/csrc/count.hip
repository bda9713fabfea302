// count.hip -- the pairwise preference counter on gfx950.
//
// Reference: populate_preferences_from_ranking / build_pairwise_preferences,
// scaling_elections.py:37-88, which scatters prefs[r[i]][r[j]] += 1 for every i < j of
// every ballot on one CPU thread.  Here each ballot arrives as a position array
// (positions[v][c] = place of candidate c), so
//     d[i][j] = #{v : positions[v][i] < positions[v][j]},
// a dense compare-and-count over 128 x 128 tiles of d with no atomics.  Only tiles with
// I <= J are computed: for a strict order per voter d[j][i] = m - d[i][j] (i != j).
// Per voter a lane holds 8 + 8 positions and does 64 compare+add pairs (v_cmp + v_addc).
//
// General ballots (duplicates allowed, exactly the reference's scatter) go through
// count_rankings_kernel: one workgroup per ballot, atomic increments.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "count_tiles.h"
#include "se_internal.h"

namespace se {

namespace {

// Lane (ty, tx) layout: see count_tile.  Linear block id -> (I, J), I <= J, row-major over
// the upper triangle of 128 x 128 tiles.
__global__ __launch_bounds__(256) void count_kernel(const int32_t* __restrict__ pos, uint32_t m, uint32_t n,
                                                    size_t ldp, uint32_t* __restrict__ d, size_t ldd,
                                                    uint32_t nb) {
    __shared__ __align__(16) int32_t Pi[kVC][kCT];
    __shared__ __align__(16) int32_t Pj[kVC][kCT];
    uint32_t t = blockIdx.x, I = 0;
    while (t >= nb - I) {
        t -= nb - I;
        ++I;
    }
    count_tile(pos, m, n, ldp, d, ldd, I, I + t, Pi, Pj);
}

__global__ void zero_matrix_kernel(uint32_t* d, uint32_t n, size_t ldd) {
    const uint32_t i = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) d[(size_t)i * ldd + j] = 0u;
}

// One workgroup per ballot: prefs[r[i]][r[j]] += 1 for all i < j (scaling_elections.py:46-48).
__global__ __launch_bounds__(256) void count_rankings_kernel(const int32_t* __restrict__ ids,
                                                             const int64_t* __restrict__ offsets,
                                                             uint32_t* __restrict__ d, size_t ldd) {
    const int64_t b = offsets[blockIdx.x], e = offsets[blockIdx.x + 1];
    const int64_t len = e - b;
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x) {
        uint32_t* row = d + (size_t)ids[b + i] * ldd;
        for (int64_t j = i + 1; j < len; ++j) atomicAdd(row + ids[b + j], 1u);
    }
}

}  // namespace

int count_device_impl(DeviceContext* ctx, const int32_t* pos, uint32_t m, uint32_t n, uint32_t* d, size_t ldd,
                      hipStream_t s, double* kernel_ms) {
    (void)ctx;
    if (n == 0) return SE_OK;
    const uint32_t nb = (n + kCT - 1) / kCT;
    const uint32_t blocks = nb * (nb + 1) / 2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms) {
        if (!hip_ok(hipEventCreate(&e0), "Failed to create events") ||
            !hip_ok(hipEventCreate(&e1), "Failed to create events"))
            return SE_ERR_DEVICE;
        (void)hipEventRecord(e0, s);
    }
    count_kernel<<<blocks, kCThreads, 0, s>>>(pos, m, n, n, d, ldd, nb);
    SE_CHECK_LAUNCH("counter launch failed");
    if (kernel_ms) {
        (void)hipEventRecord(e1, s);
        if (!hip_ok(hipEventSynchronize(e1), "CUDA operations did not complete successfully")) return SE_ERR_DEVICE;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *kernel_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return SE_OK;
}

int count_rankings_device_impl(DeviceContext* ctx, const int32_t* ids, const int64_t* offsets, uint32_t m,
                               uint32_t n, uint32_t* d, size_t ldd, hipStream_t s) {
    (void)ctx;
    if (n == 0) return SE_OK;
    zero_matrix_kernel<<<n, 256, 0, s>>>(d, n, ldd);
    SE_CHECK_LAUNCH("counter launch failed");
    if (m > 0) {
        count_rankings_kernel<<<m, 256, 0, s>>>(ids, offsets, d, ldd);
        SE_CHECK_LAUNCH("counter launch failed");
    }
    return SE_OK;
}

}  // namespace se
