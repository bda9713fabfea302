// api.hip -- the extern "C" boundary: error state, per-device contexts and the entry
// points declared in include/scaling_elections_amd.h.  No C++ exception crosses it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <csignal>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "se_internal.h"

namespace se {

// Implemented in closure.hip / count.hip / rank.hip / cpu_closure.cpp.
int closure_device_impl(DeviceContext* ctx, const uint32_t* d, uint32_t n, size_t ld, uint32_t* out,
                        size_t ldo, uint32_t flags, hipStream_t s, se_timing* timing);
int closure_device_ex(DeviceContext* ctx, const void* d_any, bool d16, uint32_t n, size_t ld, void* out_any,
                      size_t ldo, bool* out16, uint32_t flags, hipStream_t s, se_timing* timing);
int count_device_impl(DeviceContext* ctx, const void* pos, bool pos16, uint32_t m, uint32_t n, uint32_t* d,
                      size_t ldd, hipStream_t s, double* kernel_ms, uint32_t* flag);
int narrow_counts16(const uint32_t* d, uint32_t n, size_t ldd, uint16_t* out, hipStream_t s);
int narrow_positions16(const int32_t* src, size_t count, uint16_t* dst, uint32_t* flag, hipStream_t s);
int clear_flag(uint32_t* flag, hipStream_t s);
int count_rankings_device_impl(DeviceContext* ctx, const int32_t* ids, const int64_t* offsets,
                               uint32_t m, uint32_t n, uint32_t* d, size_t ldd, hipStream_t s);
int wins_device_impl(DeviceContext* ctx, const uint32_t* p, uint32_t n, size_t ld, uint32_t* wins,
                     hipStream_t s);
int batch_device_impl(DeviceContext* ctx, uint32_t E, const uint32_t* Cs, const uint32_t* ms, const void* pos,
                      uint32_t* paths, uint32_t* ranking, uint32_t flags, hipStream_t s, se_batch_timing* timing);
int closure_cpu_impl(const uint32_t* d, uint32_t n, size_t ld, uint32_t* p, uint32_t tile, int threads);
int closure_cpu_serial_impl(const uint32_t* d, uint32_t n, size_t ld, uint32_t* p);

namespace {
thread_local std::string t_last_error;
std::mutex g_ctx_mutex;
std::unique_ptr<DeviceContext> g_ctx[64];
// contexts of caller streams: own scratch and lock, so calls on different streams overlap
std::map<std::pair<int, void*>, std::unique_ptr<DeviceContext>> g_stream_ctx;
struct sigaction g_old_sigint;
std::mutex g_sig_mutex;
int g_sigint_depth = 0;
}  // namespace

std::atomic<int> g_interrupted{0};

void set_error(const std::string& msg) { t_last_error = msg; }
void clear_error() { t_last_error.clear(); }

DeviceContext* current_context() {
    int dev = -1;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        set_error("No HIP devices available");
        return nullptr;
    }
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error("Failed to get current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_ctx_mutex);
    if (!g_ctx[dev]) {
        auto ctx = std::make_unique<DeviceContext>();
        ctx->device = dev;
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            set_error("Failed to create a HIP stream");
            return nullptr;
        }
        g_ctx[dev] = std::move(ctx);
    }
    return g_ctx[dev].get();
}

void* scratch(DeviceContext* ctx, Slot s, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (ctx->cap[s] >= bytes) return ctx->ptr[s];
    if (ctx->ptr[s]) {
        // the previous contents are dead; make sure no queued kernel still uses them
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->ptr[s]);
        ctx->ptr[s] = nullptr;
        ctx->cap[s] = 0;
    }
    void* p = nullptr;
    size_t want = round_up(bytes, (size_t)1 << 20);
    if (hipMalloc(&p, want) != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        set_error("Failed to allocate memory on device");
        return nullptr;
    }
    ctx->ptr[s] = p;
    ctx->cap[s] = want;
    return p;
}

// Context for a caller stream (NULL: the library's context of the current device).
DeviceContext* context_for(void* stream) {
    if (!stream) return current_context();
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        set_error("Failed to get current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_ctx_mutex);
    auto& slot = g_stream_ctx[std::make_pair(dev, stream)];
    if (!slot) {
        slot = std::make_unique<DeviceContext>();
        slot->device = dev;
        slot->stream = static_cast<hipStream_t>(stream);
        slot->owns_stream = false;
    }
    return slot.get();
}

static void free_context_scratch(DeviceContext* ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < kSlotCount; ++i) {
        if (ctx->ptr[i]) (void)hipFree(ctx->ptr[i]);
        ctx->ptr[i] = nullptr;
        ctx->cap[i] = 0;
    }
    free_pinned(ctx);
}

}  // namespace se

using namespace se;

static void sigint_handler(int) { g_interrupted.store(1); }

extern "C" {

const char* se_last_error(void) { return t_last_error.c_str(); }

int se_version(void) { return 10000; }

int se_device_count(int* count) {
    if (!count) {
        set_error("count must not be NULL");
        return SE_ERR_INVALID;
    }
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        set_error("No HIP devices available");
        return SE_ERR_NO_DEVICE;
    }
    *count = c;
    return SE_OK;
}

int se_device_info(int device, char* name, size_t name_len, int* compute_units, size_t* global_mem_bytes,
                   int* major, int* minor, char* arch, size_t arch_len) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        (void)hipGetLastError();
        set_error("Failed to get device properties");
        return SE_ERR_NO_DEVICE;
    }
    if (name && name_len) {
        std::strncpy(name, prop.name, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (arch && arch_len) {
        std::strncpy(arch, prop.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (global_mem_bytes) *global_mem_bytes = prop.totalGlobalMem;
    if (major) *major = prop.major;
    if (minor) *minor = prop.minor;
    return SE_OK;
}

static bool gpu_tile_ok(uint32_t tile) { return tile == 4 || tile == 8 || tile == 16 || tile == 32; }

int se_closure(const uint32_t* prefs, uint32_t n, size_t row_stride, uint32_t* paths, uint32_t tile_size,
               uint32_t flags, se_timing* timing) {
    clear_error();
    if (!gpu_tile_ok(tile_size)) {
        set_error("Unsupported tile size");
        return SE_ERR_TILE;
    }
    if (n > 0 && (!prefs || !paths || row_stride < n)) {
        set_error("invalid preferences/paths buffers");
        return SE_ERR_INVALID;
    }
    g_interrupted.store(0);
    DeviceContext* ctx = current_context();
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    if (n == 0) {
        if (timing) *timing = se_timing{};
        return SE_OK;
    }
    const size_t bytes = (size_t)n * n * sizeof(uint32_t);
    uint32_t* d = static_cast<uint32_t*>(scratch(ctx, kSlotPrefs, bytes));
    if (!d) return SE_ERR_ALLOC;
    hipStream_t s = ctx->stream;
    // counts up to 65535 cross PCIe as 16-bit values (narrowed and widened on the host
    // while staged), larger ones as 32-bit; a forced 32-bit closure keeps 32-bit transit
    bool d16 = false;
    int rc = SE_NARROW_FAILED;
    if (!(flags & SE_FLAG_FORCE_U32)) {
        rc = upload_rows_narrow16(ctx, s, reinterpret_cast<uint16_t*>(d), prefs, row_stride * sizeof(uint32_t), n, n);
        d16 = rc == SE_OK;
    }
    if (rc == SE_NARROW_FAILED) rc = upload_rows(ctx, s, d, prefs, row_stride * sizeof(uint32_t), n, n * sizeof(uint32_t));
    if (rc != SE_OK) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    // the staged input is dead after p0 init, so the output reuses its buffer
    bool out16 = false;
    rc = closure_device_ex(ctx, d, d16, n, n, d, n, &out16, flags, s, timing);
    if (rc != SE_OK) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (!hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully")) return SE_ERR_DEVICE;
    if (out16) return download_rows_widen16(ctx, s, paths, n * sizeof(uint32_t), reinterpret_cast<uint16_t*>(d), n, n);
    return download_rows(ctx, s, paths, n * sizeof(uint32_t), d, n, n * sizeof(uint32_t));
}

int se_closure_device(const uint32_t* prefs_dev, uint32_t n, size_t row_stride, uint32_t* paths_dev,
                      size_t paths_stride, uint32_t flags, void* stream, se_timing* timing) {
    clear_error();
    if (n > 0 && (!prefs_dev || !paths_dev || row_stride < n || paths_stride < n)) {
        set_error("invalid preferences/paths device buffers");
        return SE_ERR_INVALID;
    }
    g_interrupted.store(0);
    DeviceContext* ctx = context_for(stream);
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    return closure_device_impl(ctx, prefs_dev, n, row_stride, paths_dev, paths_stride, flags, ctx->stream, timing);
}

int se_closure_cpu(const uint32_t* prefs, uint32_t n, size_t row_stride, uint32_t* paths, uint32_t tile_size,
                   int threads) {
    clear_error();
    if (n > 0 && (!prefs || !paths || row_stride < n)) {
        set_error("invalid preferences/paths buffers");
        return SE_ERR_INVALID;
    }
    // tile resolution, scaling_elections.cu:943-980
    static const uint32_t kTiles[] = {4, 8, 16, 32, 64, 128};
    if (tile_size == 0) {
        for (uint32_t t : kTiles)
            if (n >= t) {
                tile_size = t;
                break;
            }
        if (tile_size == 0) {
            set_error("Number of candidates should be at least 4, ideally divisible by 4");
            return SE_ERR_TILE;
        }
    } else {
        if (tile_size > n) {
            set_error("Tile size should be less than or equal to the number of candidates");
            return SE_ERR_TILE;
        }
        if (std::find(std::begin(kTiles), std::end(kTiles), tile_size) == std::end(kTiles)) {
            set_error("Unsupported tile size");
            return SE_ERR_TILE;
        }
    }
    g_interrupted.store(0);
    return closure_cpu_impl(prefs, n, row_stride, paths, tile_size, threads);
}

int se_closure_cpu_serial(const uint32_t* prefs, uint32_t n, size_t row_stride, uint32_t* paths) {
    clear_error();
    if (n > 0 && (!prefs || !paths || row_stride < n)) {
        set_error("invalid preferences/paths buffers");
        return SE_ERR_INVALID;
    }
    g_interrupted.store(0);
    return closure_cpu_serial_impl(prefs, n, row_stride, paths);
}

int se_count(const int32_t* positions, uint32_t m, uint32_t n, uint32_t* prefs, size_t prefs_stride) {
    clear_error();
    if (n == 0) return SE_OK;
    if (!prefs || prefs_stride < n || (m > 0 && !positions)) {
        set_error("invalid positions/preferences buffers");
        return SE_ERR_INVALID;
    }
    DeviceContext* ctx = current_context();
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    hipStream_t s = ctx->stream;
    // PCIe transit in 16 bits where the values allow: positions (a permutation of 0..n-1
    // fits when n <= 65536; anything else fails the narrowing and goes up as int32, where
    // the counter's range check reports it) and counts (at most m)
    const size_t pbytes = (size_t)m * n * sizeof(int32_t);
    const size_t dbytes = (size_t)n * n * sizeof(uint32_t);
    void* pos = scratch(ctx, kSlotAux, pbytes);
    if (!pos) return SE_ERR_ALLOC;
    uint32_t* d = static_cast<uint32_t*>(scratch(ctx, kSlotPrefs, dbytes));
    if (!d) return SE_ERR_ALLOC;
    uint32_t* flag = static_cast<uint32_t*>(scratch(ctx, kSlotCountFlag, 256));
    if (!flag) return SE_ERR_ALLOC;
    const bool d16 = m <= 0xFFFFu;
    uint16_t* dn = d16 ? static_cast<uint16_t*>(scratch(ctx, kSlotPaths, (size_t)n * n * sizeof(uint16_t))) : nullptr;
    if (d16 && !dn) return SE_ERR_ALLOC;
    bool pos16 = false;
    int rc = SE_OK;
    if (m > 0) {
        rc = upload_rows_narrow16(ctx, s, static_cast<uint16_t*>(pos), reinterpret_cast<const uint32_t*>(positions),
                                  n * sizeof(int32_t), m, n);
        pos16 = rc == SE_OK;
        if (rc == SE_NARROW_FAILED)
            rc = upload_rows(ctx, s, pos, positions, n * sizeof(int32_t), m, n * sizeof(int32_t));
    }
    if (rc == SE_OK) rc = clear_flag(flag, s);
    if (rc == SE_OK) rc = count_device_impl(ctx, pos, pos16, m, n, d, n, s, nullptr, flag);
    if (rc == SE_OK && d16) rc = narrow_counts16(d, n, n, dn, s);
    if (rc != SE_OK) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    uint32_t* bad = small_pinned(ctx);  // pinned: a pageable 4-byte read-back is staged by the runtime
    if (!bad) return SE_ERR_ALLOC;
    if (!hip_ok(hipMemcpyAsync(bad, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "Failed to copy data from device to host"))
        return SE_ERR_COPY;
    rc = d16 ? download_rows_widen16(ctx, s, prefs, prefs_stride * sizeof(uint32_t), dn, n, n)
             : download_rows(ctx, s, prefs, prefs_stride * sizeof(uint32_t), d, n, n * sizeof(uint32_t));
    if (rc != SE_OK) return rc;
    if (!hip_ok(hipStreamSynchronize(s), "CUDA transfers did not complete successfully")) return SE_ERR_COPY;
    if (*bad) {
        set_error("positions out of range: every ballot row must be a permutation of 0..n-1");
        return SE_ERR_INVALID;
    }
    return SE_OK;
}

int se_count_device(const int32_t* positions_dev, uint32_t m, uint32_t n, uint32_t* prefs_dev,
                    size_t prefs_stride, void* stream, double* kernel_ms) {
    clear_error();
    if (n == 0) return SE_OK;
    if (!prefs_dev || prefs_stride < n || (m > 0 && !positions_dev)) {
        set_error("invalid positions/preferences device buffers");
        return SE_ERR_INVALID;
    }
    DeviceContext* ctx = context_for(stream);
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    return count_device_impl(ctx, positions_dev, false, m, n, prefs_dev, prefs_stride, ctx->stream, kernel_ms,
                             nullptr);
}

int se_count_rankings(const int32_t* ids, const int64_t* offsets, uint32_t m, uint32_t n, uint32_t* prefs,
                      size_t prefs_stride) {
    clear_error();
    if (n == 0) return SE_OK;
    if (!prefs || prefs_stride < n || !offsets || (m > 0 && offsets[m] > 0 && !ids)) {
        set_error("invalid ballots/preferences buffers");
        return SE_ERR_INVALID;
    }
    const int64_t total = offsets[m];
    for (uint32_t v = 0; v < m; ++v)
        if (offsets[v] > offsets[v + 1] || offsets[v] < 0) {
            set_error("ballot offsets must be non-decreasing");
            return SE_ERR_INVALID;
        }
    for (int64_t i = 0; i < total; ++i)
        if (ids[i] < 0 || (uint32_t)ids[i] >= n) {
            set_error("candidate ids must lie in [0, n)");
            return SE_ERR_INVALID;
        }
    DeviceContext* ctx = current_context();
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    hipStream_t s = ctx->stream;
    const size_t ibytes = (size_t)std::max<int64_t>(total, 1) * sizeof(int32_t);
    const size_t obytes = (size_t)(m + 1) * sizeof(int64_t);
    char* aux = static_cast<char*>(scratch(ctx, kSlotAux, round_up(ibytes, 256) + obytes));
    if (!aux) return SE_ERR_ALLOC;
    int32_t* ids_dev = reinterpret_cast<int32_t*>(aux);
    int64_t* off_dev = reinterpret_cast<int64_t*>(aux + round_up(ibytes, 256));
    uint32_t* d = static_cast<uint32_t*>(scratch(ctx, kSlotPrefs, (size_t)n * n * sizeof(uint32_t)));
    if (!d) return SE_ERR_ALLOC;
    if ((total > 0 && !hip_ok(hipMemcpyAsync(ids_dev, ids, (size_t)total * sizeof(int32_t), hipMemcpyHostToDevice, s),
                              "Failed to copy data from host to device")) ||
        !hip_ok(hipMemcpyAsync(off_dev, offsets, obytes, hipMemcpyHostToDevice, s),
                "Failed to copy data from host to device"))
        return SE_ERR_COPY;
    int rc = count_rankings_device_impl(ctx, ids_dev, off_dev, m, n, d, n, s);
    if (rc != SE_OK) return rc;
    if (!hip_ok(hipMemcpy2DAsync(prefs, prefs_stride * sizeof(uint32_t), d, n * sizeof(uint32_t),
                                 n * sizeof(uint32_t), n, hipMemcpyDeviceToHost, s),
                "Failed to copy data from device to host"))
        return SE_ERR_COPY;
    if (!hip_ok(hipStreamSynchronize(s), "CUDA transfers did not complete successfully")) return SE_ERR_COPY;
    return SE_OK;
}

static void stable_rank(const uint32_t* wins, uint32_t n, uint32_t* ranking) {
    std::iota(ranking, ranking + n, 0u);
    std::stable_sort(ranking, ranking + n, [wins](uint32_t a, uint32_t b) { return wins[a] > wins[b]; });
}

int se_rank_device(const uint32_t* paths_dev, uint32_t n, size_t row_stride, uint32_t* wins_host,
                   uint32_t* ranking_host, void* stream) {
    clear_error();
    if (n == 0) return SE_OK;
    if (!paths_dev || row_stride < n || !wins_host) {
        set_error("invalid paths/wins buffers");
        return SE_ERR_INVALID;
    }
    DeviceContext* ctx = context_for(stream);
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    hipStream_t s = ctx->stream;
    uint32_t* w = static_cast<uint32_t*>(scratch(ctx, kSlotSmall, (size_t)n * sizeof(uint32_t)));
    if (!w) return SE_ERR_ALLOC;
    int rc = wins_device_impl(ctx, paths_dev, n, row_stride, w, s);
    if (rc != SE_OK) return rc;
    if (!hip_ok(hipMemcpyAsync(wins_host, w, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "Failed to copy data from device to host") ||
        !hip_ok(hipStreamSynchronize(s), "CUDA transfers did not complete successfully"))
        return SE_ERR_COPY;
    if (ranking_host) stable_rank(wins_host, n, ranking_host);
    return SE_OK;
}

int se_rank(const uint32_t* paths, uint32_t n, size_t row_stride, uint32_t* wins, uint32_t* ranking) {
    clear_error();
    if (n == 0) return SE_OK;
    if (!paths || row_stride < n || !wins) {
        set_error("invalid paths/wins buffers");
        return SE_ERR_INVALID;
    }
    DeviceContext* ctx = current_context();
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    hipStream_t s = ctx->stream;
    uint32_t* p = static_cast<uint32_t*>(scratch(ctx, kSlotPaths, (size_t)n * n * sizeof(uint32_t)));
    uint32_t* w = static_cast<uint32_t*>(scratch(ctx, kSlotSmall, (size_t)n * sizeof(uint32_t)));
    if (!p || !w) return SE_ERR_ALLOC;
    if (!hip_ok(hipMemcpy2DAsync(p, n * sizeof(uint32_t), paths, row_stride * sizeof(uint32_t),
                                 n * sizeof(uint32_t), n, hipMemcpyHostToDevice, s),
                "Failed to copy data from host to device"))
        return SE_ERR_COPY;
    int rc = wins_device_impl(ctx, p, n, n, w, s);
    if (rc != SE_OK) return rc;
    if (!hip_ok(hipMemcpyAsync(wins, w, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "Failed to copy data from device to host") ||
        !hip_ok(hipStreamSynchronize(s), "CUDA transfers did not complete successfully"))
        return SE_ERR_COPY;
    if (ranking) stable_rank(wins, n, ranking);
    return SE_OK;
}

int se_batch_device(uint32_t elections, const uint32_t* candidates, const uint32_t* voters,
                    const void* positions_dev, uint32_t* paths_dev, uint32_t* ranking_dev, uint32_t flags,
                    void* stream, se_batch_timing* timing) {
    clear_error();
    if (timing) *timing = se_batch_timing{};
    if (elections == 0) return SE_OK;
    if (!candidates || !voters || !paths_dev || !ranking_dev) {
        set_error("invalid batch buffers");
        return SE_ERR_INVALID;
    }
    uint64_t total_pos = 0;
    for (uint32_t e = 0; e < elections; ++e) total_pos += (uint64_t)candidates[e] * voters[e];
    if (total_pos && !positions_dev) {
        set_error("invalid batch buffers");
        return SE_ERR_INVALID;
    }
    DeviceContext* ctx = context_for(stream);
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    return batch_device_impl(ctx, elections, candidates, voters, positions_dev, paths_dev, ranking_dev, flags,
                             ctx->stream, timing);
}

int se_positions16_device(const int32_t* src_dev, size_t count, uint16_t* dst_dev, void* stream) {
    clear_error();
    if (count == 0) return SE_OK;
    if (!src_dev || !dst_dev) {
        set_error("positions buffers must not be NULL");
        return SE_ERR_INVALID;
    }
    DeviceContext* ctx = context_for(stream);
    if (!ctx) return SE_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> g(ctx->lock);
    hipStream_t s = ctx->stream;
    uint32_t* flag = static_cast<uint32_t*>(scratch(ctx, kSlotCountFlag, 256));
    if (!flag) return SE_ERR_ALLOC;
    int rc = clear_flag(flag, s);
    if (rc == SE_OK) rc = narrow_positions16(src_dev, count, dst_dev, flag, s);
    if (rc != SE_OK) return rc;
    uint32_t* bad = small_pinned(ctx);
    if (!bad) return SE_ERR_ALLOC;
    if (!hip_ok(hipMemcpyAsync(bad, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "Failed to copy data from device to host") ||
        !hip_ok(hipStreamSynchronize(s), "CUDA operations did not complete successfully"))
        return SE_ERR_COPY;
    if (*bad) {
        set_error("positions out of range: a value lies outside 0..65535");
        return SE_ERR_INVALID;
    }
    return SE_OK;
}

float se_reduce_f32(const float* data, size_t count) {
    float acc = 0.0f;
    for (size_t i = 0; i < count; ++i) acc = acc + data[i];
    return acc;
}

int se_stream_release(void* stream) {
    clear_error();
    if (!stream) {
        set_error("stream must not be NULL");
        return SE_ERR_INVALID;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("Failed to get current device");
        return SE_ERR_NO_DEVICE;
    }
    std::unique_ptr<DeviceContext> ctx;
    {
        std::lock_guard<std::mutex> g(g_ctx_mutex);
        auto it = g_stream_ctx.find(std::make_pair(dev, stream));
        if (it == g_stream_ctx.end()) return SE_OK;
        ctx = std::move(it->second);
        g_stream_ctx.erase(it);
    }
    std::lock_guard<std::mutex> g(ctx->lock);
    free_context_scratch(ctx.get());
    return SE_OK;
}

void se_interrupt(void) { g_interrupted.store(1); }
void se_clear_interrupt(void) { g_interrupted.store(0); }

// Reference-counted so that concurrent callers restore the original handler exactly once.
int se_install_sigint(void) {
    std::lock_guard<std::mutex> g(g_sig_mutex);
    if (g_sigint_depth++ > 0) return SE_OK;
    struct sigaction sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.sa_handler = sigint_handler;
    sigemptyset(&sa.sa_mask);
    if (sigaction(SIGINT, &sa, &g_old_sigint) != 0) {
        --g_sigint_depth;
        return SE_ERR_INVALID;
    }
    return SE_OK;
}

void se_restore_sigint(void) {
    std::lock_guard<std::mutex> g(g_sig_mutex);
    if (g_sigint_depth > 0 && --g_sigint_depth == 0) (void)sigaction(SIGINT, &g_old_sigint, nullptr);
}

}  // extern "C"
