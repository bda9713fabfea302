// count.hip -- the pairwise preference counter on gfx950.
//
// Reference: populate_preferences_from_ranking / build_pairwise_preferences,
// scaling_elections.py:37-88, which scatters prefs[r[i]][r[j]] += 1 for every i < j of
// every ballot on one CPU thread.  Here each ballot arrives as a position array
// (positions[v][c] = place of candidate c), so
//     d[i][j] = #{v : positions[v][i] < positions[v][j]},
// a dense compare-and-count over 128 x 128 tiles of d with no atomics.  Only tiles with
// I <= J are computed: for a strict order per voter d[j][i] = m - d[i][j] (i != j).
// Positions are packed 4 or 2 voters to a word in LDS and compared with one subtract and
// one signed dot product per word (count_tiles.h); above 32768 candidates, one voter per
// word with compare-and-add.
//
// General ballots (duplicates allowed, exactly the reference's scatter) go through
// count_rankings_kernel: one workgroup per ballot, atomic increments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>

#include "count_tiles.h"
#include "se_internal.h"

namespace se {

namespace {

// One workgroup per (tile pair, voter slice).  Tile pairs (I, J), I <= J, are numbered
// row-major over the upper triangle of 128 x 128 tiles (n <= 128: one voter-parallel item).
// Slice s covers voters [v0 + s slice_len, v0 + (s + 1) slice_len) of [v0, v1).  With several
// slices each writes partial counts of its voters to its own n x n plane of `out` (the
// complement d[j][i] = m_slice - d[i][j] holds per slice); reduce_slices_kernel sums the
// planes.  Out-of-range positions for the packing width set *flag.
template <uint32_t K, typename P>
__global__ __launch_bounds__(256) void count_kernel(const P* __restrict__ pos, uint32_t v0, uint32_t v1,
                                                    uint32_t n, size_t ldp, uint32_t* __restrict__ out, size_t ldo,
                                                    uint32_t nb, uint32_t slices, uint32_t slice_len, size_t plane,
                                                    uint32_t accumulate, uint32_t* __restrict__ flag) {
    __shared__ __align__(16) uint32_t PI[kVW][kCT];
    __shared__ __align__(16) uint32_t PJ[kVW][kCT];
    const uint32_t slice = blockIdx.x % slices;
    uint32_t t = blockIdx.x / slices;
    const uint32_t vb = min(v1, v0 + slice * slice_len), ve = min(v1, vb + slice_len);
    uint32_t* o = out + (size_t)slice * plane;
    uint32_t bits;
    if (n <= kCT) {
        bits = count_block<K, P>(pos, vb, ve, ldp, 0, 0, n, n, o, ldo, PI, PJ, &PI[0][0], accumulate ? kPutAdd : kPutStore);
    } else {
        uint32_t I = 0;
        while (t >= nb - I) {
            t -= nb - I;
            ++I;
        }
        bits = count_tile<K, P>(pos, vb, ve, n, ldp, o, ldo, I, I + t, PI, PJ, &PI[0][0], accumulate ? kPutAdd : kPutStore);
    }
    if (flag && (bits & PackTraits<K>::range_mask)) atomicOr(flag, 1u);
}

__global__ __launch_bounds__(256) void reduce_slices_kernel(const uint32_t* __restrict__ part, uint32_t slices,
                                                            size_t plane, uint32_t n, uint32_t* __restrict__ d,
                                                            size_t ldd, uint32_t accumulate) {
    const uint32_t i = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        uint32_t s = accumulate ? d[(size_t)i * ldd + j] : 0u;
        // small elections sum many planes (one per 32+ voters): keep 8 loads in flight
#pragma unroll 8
        for (uint32_t k = 0; k < slices; ++k) s += part[(size_t)k * plane + (size_t)i * n + j];
        d[(size_t)i * ldd + j] = s;
    }
}

// The same sum four columns at a time (n % 4 == 0; d 16-byte aligned with ldd % 4 == 0):
// 16-byte loads keep more of each plane in flight (4096 x 4096, 5 planes: 125 us with one
// column per load).
__global__ __launch_bounds__(256) void reduce_slices4_kernel(const uint4* __restrict__ part, uint32_t slices,
                                                             size_t plane4, uint32_t n4, uint4* __restrict__ d,
                                                             size_t ldd4, uint32_t accumulate) {
    const uint32_t i = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < n4; j += blockDim.x) {
        uint4 s = accumulate ? d[(size_t)i * ldd4 + j] : make_uint4(0, 0, 0, 0);
#pragma unroll 8
        for (uint32_t k = 0; k < slices; ++k) {
            const uint4 v = part[(size_t)k * plane4 + (size_t)i * n4 + j];
            s = make_uint4(s.x + v.x, s.y + v.y, s.z + v.z, s.w + v.w);
        }
        d[(size_t)i * ldd4 + j] = s;
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

template <uint32_t K, typename P>
void launch_count(uint32_t grid, hipStream_t s, const P* pos, uint32_t v0, uint32_t v1, uint32_t n, size_t ldp,
                  uint32_t* out, size_t ldo, uint32_t nb, uint32_t slices, uint32_t slice_len, size_t plane,
                  uint32_t accumulate, uint32_t* flag) {
    count_kernel<K, P><<<grid, kCThreads, 0, s>>>(pos, v0, v1, n, ldp, out, ldo, nb, slices, slice_len, plane,
                                                  accumulate, flag);
}

template <typename P>
void launch_count_k(uint32_t K, uint32_t grid, hipStream_t s, const P* pos, uint32_t v0, uint32_t v1, uint32_t n,
                    uint32_t* out, size_t ldo, uint32_t nb, uint32_t slices, uint32_t slice_len, size_t plane,
                    uint32_t accumulate, uint32_t* flag) {
    if (K == 4)
        launch_count<4, P>(grid, s, pos, v0, v1, n, n, out, ldo, nb, slices, slice_len, plane, accumulate, flag);
    else if (K == 2)
        launch_count<2, P>(grid, s, pos, v0, v1, n, n, out, ldo, nb, slices, slice_len, plane, accumulate, flag);
    else
        launch_count<1, P>(grid, s, pos, v0, v1, n, n, out, ldo, nb, slices, slice_len, plane, accumulate, flag);
}

}  // namespace

// Counter plan: packing width K from n (bytes up to 128 candidates, 16-bit halves up to
// 32768, int32 above), voter slices for parallelism (see below) capped at the largest
// slice a K-packed accumulation keeps exact, and passes of at most `budget` bytes of
// partial planes, each pass adding into d.  (Slices adding atomically into d instead
// measured slower: at 512 candidates x 10000 voters about 100 slices pile some 64 atomic
// adds onto every cell, 0.378 against 0.137 ms.)  Positions: int32, or uint16 with pos16.
int count_device_impl(DeviceContext* ctx, const void* pos, bool pos16, uint32_t m, uint32_t n, uint32_t* d,
                      size_t ldd, hipStream_t s, double* kernel_ms, uint32_t* flag) {
    if (n == 0) return SE_OK;
    const uint32_t nb = (n + kCT - 1) / kCT;
    const uint32_t tiles = n <= kCT ? 1 : nb * (nb + 1) / 2;
    const uint32_t K = n <= 128 ? 4u : n <= 32768 ? 2u : 1u;
    const uint64_t plane_bytes = (uint64_t)n * n * 4;
    // partial-plane budget; SE_COUNT_PARTIAL_BYTES overrides it (tests use it to force passes)
    uint64_t budget = 1ull << 30;
    if (const char* env = std::getenv("SE_COUNT_PARTIAL_BYTES")) budget = std::max<uint64_t>(1, std::strtoull(env, nullptr, 10));
    const uint64_t fit = std::max<uint64_t>(1, budget / plane_bytes);
    // voter slices: about 2560 workgroups (10 per CU, so the last wave is a small share)
    // while every slice keeps at least 512 voters; otherwise about 1024 workgroups with
    // at least 32 voters per slice
    uint64_t want = (2560 + tiles - 1) / tiles;
    if (m / want < 512) {
        want = (1024 + tiles - 1) / tiles;
        want = std::min<uint64_t>(want, std::max<uint32_t>(1, m / 32));
    }
    want = std::min<uint64_t>(want, fit);
    want = std::max<uint64_t>(want, 1);
    uint64_t slice_len = m ? (m + want - 1) / want : 1;
    slice_len = round_up(slice_len, K);
    slice_len = std::min<uint64_t>(slice_len, count_max_slice(K));
    const uint64_t total_slices = m ? (m + slice_len - 1) / slice_len : 1;
    const uint64_t per_pass = std::min<uint64_t>(total_slices, fit);
    const uint64_t passes = (total_slices + per_pass - 1) / per_pass;

    uint32_t* part = nullptr;
    if (per_pass > 1) {
        part = static_cast<uint32_t*>(scratch(ctx, kSlotCountPart, plane_bytes * per_pass));
        if (!part) return SE_ERR_ALLOC;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms) {
        if (!hip_ok(hipEventCreate(&e0), "Failed to create events") ||
            !hip_ok(hipEventCreate(&e1), "Failed to create events"))
            return SE_ERR_DEVICE;
        (void)hipEventRecord(e0, s);
    }
    for (uint64_t p = 0; p < passes; ++p) {
        const uint64_t first = p * per_pass;
        const uint32_t sl = (uint32_t)std::min<uint64_t>(per_pass, total_slices - first);
        const uint32_t v0 = (uint32_t)std::min<uint64_t>(m, first * slice_len);
        const uint32_t v1 = (uint32_t)std::min<uint64_t>(m, (first + sl) * slice_len);
        uint32_t* out = sl > 1 ? part : d;
        const size_t ldo = sl > 1 ? n : ldd, plane = sl > 1 ? (size_t)n * n : 0;
        const uint32_t acc = sl > 1 ? 0u : (p > 0 ? 1u : 0u);
        const uint32_t grid = tiles * sl;
        if (pos16)
            launch_count_k<uint16_t>(K, grid, s, static_cast<const uint16_t*>(pos), v0, v1, n, out, ldo, nb, sl,
                                     (uint32_t)slice_len, plane, acc, flag);
        else
            launch_count_k<int32_t>(K, grid, s, static_cast<const int32_t*>(pos), v0, v1, n, out, ldo, nb, sl,
                                    (uint32_t)slice_len, plane, acc, flag);
        SE_CHECK_LAUNCH("counter launch failed");
        if (sl > 1) {
            if (n % 4 == 0 && ldd % 4 == 0 && ((uintptr_t)d & 15) == 0)
                reduce_slices4_kernel<<<n, 256, 0, s>>>(reinterpret_cast<const uint4*>(part), sl, plane / 4, n / 4,
                                                        reinterpret_cast<uint4*>(d), ldd / 4, p > 0 ? 1u : 0u);
            else
                reduce_slices_kernel<<<n, 256, 0, s>>>(part, sl, plane, n, d, ldd, p > 0 ? 1u : 0u);
            SE_CHECK_LAUNCH("counter launch failed");
        }
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

// d (n x n u32, row stride ldd) -> out (n x n u16, packed) for the host entry point's
// 16-bit transit (every count is at most m <= 65535).
__global__ __launch_bounds__(256) void narrow16_kernel(const uint32_t* __restrict__ d, uint32_t n, size_t ldd,
                                                       uint16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) out[(size_t)i * n + j] = (uint16_t)d[(size_t)i * ldd + j];
}

int narrow_counts16(const uint32_t* d, uint32_t n, size_t ldd, uint16_t* out, hipStream_t s) {
    if (n == 0) return SE_OK;
    narrow16_kernel<<<n, 256, 0, s>>>(d, n, ldd, out);
    SE_CHECK_LAUNCH("counter launch failed");
    return SE_OK;
}

// int32 positions -> uint16 (batches with SE_BATCH_POS16); *flag records a value outside
// 0..65535 (its low 16 bits are stored).
__global__ __launch_bounds__(256) void narrow_positions_kernel(const int32_t* __restrict__ src, size_t count,
                                                               uint16_t* __restrict__ dst, uint32_t* __restrict__ flag) {
    uint32_t bad = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)src[i];
        bad |= v >> 16;
        dst[i] = (uint16_t)v;
    }
    if (bad) atomicOr(flag, 1u);
}

int narrow_positions16(const int32_t* src, size_t count, uint16_t* dst, uint32_t* flag, hipStream_t s) {
    if (count == 0) return SE_OK;
    const size_t blocks = std::min<size_t>(4096, (count + 255) / 256);
    narrow_positions_kernel<<<(uint32_t)blocks, 256, 0, s>>>(src, count, dst, flag);
    SE_CHECK_LAUNCH("position conversion launch failed");
    return SE_OK;
}

// Clears a device word (a kernel, not hipMemsetAsync, so the sequence stays capturable).
int clear_flag(uint32_t* flag, hipStream_t s) {
    zero_matrix_kernel<<<1, 64, 0, s>>>(flag, 1, 1);
    SE_CHECK_LAUNCH("counter launch failed");
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
