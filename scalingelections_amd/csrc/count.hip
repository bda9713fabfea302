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

#include <algorithm>
#include <cstdint>

#include "count_tiles.h"
#include "se_internal.h"

namespace se {

namespace {

// One workgroup per (tile pair, voter slice).  Tile pairs (I, J), I <= J, are numbered
// row-major over the upper triangle of 128 x 128 tiles (n <= 128: one voter-parallel item).
// With several slices each writes partial counts of its voters to its own n x n plane of
// `out` (the complement d[j][i] = m_slice - d[i][j] holds per slice); reduce_slices_kernel
// sums the planes.
__global__ __launch_bounds__(256) void count_kernel(const int32_t* __restrict__ pos, uint32_t m, uint32_t n,
                                                    size_t ldp, uint32_t* __restrict__ out, size_t ldo,
                                                    uint32_t nb, uint32_t slices, uint32_t slice_len,
                                                    size_t plane) {
    __shared__ __align__(16) int32_t Pi[kVC][kCT];
    __shared__ __align__(16) int32_t Pj[kVC][kCT];
    const uint32_t slice = blockIdx.x % slices;
    uint32_t t = blockIdx.x / slices;
    const uint32_t vb = min(m, slice * slice_len), ve = min(m, vb + slice_len);
    const int32_t* p = pos + (size_t)vb * ldp;
    uint32_t* o = out + (size_t)slice * plane;
    if (n <= kCT) {
        count_small(p, ve - vb, n, ldp, o, ldo, &Pi[0][0]);
        return;
    }
    uint32_t I = 0;
    while (t >= nb - I) {
        t -= nb - I;
        ++I;
    }
    count_tile(p, ve - vb, n, ldp, o, ldo, I, I + t, Pi, Pj);
}

__global__ __launch_bounds__(256) void reduce_slices_kernel(const uint32_t* __restrict__ part, uint32_t slices,
                                                            size_t plane, uint32_t n, uint32_t* __restrict__ d,
                                                            size_t ldd) {
    const uint32_t i = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        uint32_t s = 0;
        for (uint32_t k = 0; k < slices; ++k) s += part[(size_t)k * plane + (size_t)i * n + j];
        d[(size_t)i * ldd + j] = s;
    }
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
    if (n == 0) return SE_OK;
    const uint32_t nb = (n + kCT - 1) / kCT;
    const uint32_t tiles = n <= kCT ? 1 : nb * (nb + 1) / 2;
    // voter slices: aim at ~1024 workgroups (4 per CU), at least 32 voters per slice, and
    // partial planes within 1 GiB
    uint64_t want = (1024 + tiles - 1) / tiles;
    want = std::min<uint64_t>(want, std::max<uint32_t>(1, m / 32));
    want = std::min<uint64_t>(want, std::max<uint64_t>(1, (1ull << 30) / ((uint64_t)n * n * 4)));
    const uint32_t slices = (uint32_t)std::max<uint64_t>(1, want);
    const uint32_t slice_len = slices > 1 ? (m + slices - 1) / slices : m;
    uint32_t* out = d;
    size_t ldo = ldd, plane = 0;
    if (slices > 1) {
        plane = (size_t)n * n;
        out = static_cast<uint32_t*>(scratch(ctx, kSlotCountPart, plane * slices * sizeof(uint32_t)));
        if (!out) return SE_ERR_ALLOC;
        ldo = n;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms) {
        if (!hip_ok(hipEventCreate(&e0), "Failed to create events") ||
            !hip_ok(hipEventCreate(&e1), "Failed to create events"))
            return SE_ERR_DEVICE;
        (void)hipEventRecord(e0, s);
    }
    count_kernel<<<tiles * slices, kCThreads, 0, s>>>(pos, m, n, n, out, ldo, nb, slices, slice_len, plane);
    SE_CHECK_LAUNCH("counter launch failed");
    if (slices > 1) {
        reduce_slices_kernel<<<n, 256, 0, s>>>(out, slices, plane, n, d, ldd);
        SE_CHECK_LAUNCH("counter launch failed");
    }
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
