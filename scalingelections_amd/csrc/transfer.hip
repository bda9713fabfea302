// transfer.hip -- host <-> device copies of n x n matrices through pinned staging.
//
// The host entry points receive and return pageable caller buffers (the drop-in
// compute_strongest_paths returns a fresh numpy array).  A plain hipMemcpy from pageable
// memory moves about 10 GB/s; here each matrix goes through two pinned chunks (about an
// eighth of the matrix each, 2..64 MiB): host worker threads copy rows between the
// caller's buffer and one chunk while the DMA engine moves the other, so the copy runs at
// the DMA rate.  The chunks are
// allocated once per device context (hipHostMalloc) and reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "se_internal.h"

namespace se {

namespace {

constexpr size_t kChunkBytes = 64u << 20;
constexpr size_t kTaskBytes = 256u << 10;  // host rows per worker task

// Rows per staged chunk: about an eighth of the matrix, so that the host copies of one
// chunk overlap the DMA of the other even for small matrices, between 2 MiB and the
// 64 MiB chunk (row_bytes <= kChunkBytes is checked by the callers).
size_t rows_per_chunk(size_t rows, size_t row_bytes) {
    const size_t target = std::min(kChunkBytes, std::max<size_t>(2u << 20, rows * row_bytes / 8));
    return std::max<size_t>(1, target / row_bytes);
}

// Rows per worker task when `rows` rows are copied: everything in one task (the calling
// thread) up to 1 MiB, where waking the pool costs more than it saves.
size_t rows_per_task(size_t rows, size_t row_bytes) {
    if (rows * row_bytes <= (1u << 20)) return std::max<size_t>(1, rows);
    return std::max<size_t>(1, kTaskBytes / std::max<size_t>(1, row_bytes));
}

// Fixed pool of host threads for row copies.  run(count, fn) calls fn(i) for i < count
// from the workers and the calling thread and returns when all are done.
class RowPool {
  public:
    static RowPool& get() {
        static RowPool pool;
        return pool;
    }

    void run(size_t count, const std::function<void(size_t)>& fn) {
        std::lock_guard<std::mutex> serial(call_mutex_);  // one job at a time
        if (count == 0) return;
        if (workers_.empty() || count < 4) {
            for (size_t i = 0; i < count; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &fn;
            count_ = count;
            next_.store(0);
            active_ = workers_.size();
            ++generation_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> g(m_);
        done_cv_.wait(g, [&] { return active_ == 0; });
        job_ = nullptr;
    }

  private:
    RowPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned n = std::min(15u, hw > 1 ? hw - 1 : 0u);
        for (unsigned i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~RowPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            ++generation_;
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
    }

    void drain() {
        for (;;) {
            const size_t i = next_.fetch_add(1);
            if (i >= count_) break;
            (*job_)(i);
        }
    }

    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
            }
            drain();
            std::lock_guard<std::mutex> g(m_);
            if (--active_ == 0) done_cv_.notify_all();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex call_mutex_, m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)>* job_ = nullptr;
    size_t count_ = 0;
    std::atomic<size_t> next_{0};
    size_t active_ = 0;
    uint64_t generation_ = 0;
    bool stop_ = false;
};

// Copies `rows` rows of `row_bytes` between strided host memory and a packed buffer, in
// parallel blocks of rows.
void copy_rows(char* dst, size_t dst_pitch, const char* src, size_t src_pitch, size_t rows, size_t row_bytes) {
    const size_t per = rows_per_task(rows, row_bytes);
    const size_t tasks = (rows + per - 1) / per;
    RowPool::get().run(tasks, [&](size_t t) {
        const size_t r0 = t * per, r1 = std::min(rows, r0 + per);
        if (dst_pitch == row_bytes && src_pitch == row_bytes) {
            std::memcpy(dst + r0 * dst_pitch, src + r0 * src_pitch, (r1 - r0) * row_bytes);
        } else {
            for (size_t r = r0; r < r1; ++r) std::memcpy(dst + r * dst_pitch, src + r * src_pitch, row_bytes);
        }
    });
}

bool ensure_pinned(DeviceContext* ctx) {
    if (ctx->pinned[0]) return true;
    for (int b = 0; b < 2; ++b) {
        if (!hip_ok(hipHostMalloc(&ctx->pinned[b], kChunkBytes, hipHostMallocDefault),
                    "Failed to allocate pinned host memory") ||
            !hip_ok(hipEventCreateWithFlags(&ctx->pinned_ev[b], hipEventDisableTiming), "Failed to create events"))
            return false;
    }
    return true;
}

}  // namespace

// host (rows x row_bytes, pitch host_pitch) -> device (packed, pitch row_bytes)
int upload_rows(DeviceContext* ctx, hipStream_t s, void* dev, const void* host, size_t host_pitch, size_t rows,
                size_t row_bytes) {
    if (rows == 0 || row_bytes == 0) return SE_OK;
    if (row_bytes > kChunkBytes || !ensure_pinned(ctx)) {  // fall back to one plain copy
        return hip_ok(hipMemcpy2DAsync(dev, row_bytes, host, host_pitch, row_bytes, rows, hipMemcpyHostToDevice, s),
                      "Failed to copy data from host to device")
                   ? SE_OK
                   : SE_ERR_COPY;
    }
    const size_t per = rows_per_chunk(rows, row_bytes);
    bool used[2] = {false, false};
    for (size_t r0 = 0, c = 0; r0 < rows; r0 += per, ++c) {
        const size_t nr = std::min(per, rows - r0);
        const int b = (int)(c & 1);
        if (used[b] && !hip_ok(hipEventSynchronize(ctx->pinned_ev[b]), "Failed to copy data from host to device"))
            return SE_ERR_COPY;
        copy_rows(static_cast<char*>(ctx->pinned[b]), row_bytes, static_cast<const char*>(host) + r0 * host_pitch,
                  host_pitch, nr, row_bytes);
        if (!hip_ok(hipMemcpyAsync(static_cast<char*>(dev) + r0 * row_bytes, ctx->pinned[b], nr * row_bytes,
                                   hipMemcpyHostToDevice, s),
                    "Failed to copy data from host to device") ||
            !hip_ok(hipEventRecord(ctx->pinned_ev[b], s), "Failed to copy data from host to device"))
            return SE_ERR_COPY;
        used[b] = true;
    }
    return SE_OK;
}

// device (packed, pitch row_bytes) -> host (pitch host_pitch); returns after the data landed
int download_rows(DeviceContext* ctx, hipStream_t s, void* host, size_t host_pitch, const void* dev, size_t rows,
                  size_t row_bytes) {
    if (rows == 0 || row_bytes == 0) return SE_OK;
    if (row_bytes > kChunkBytes || !ensure_pinned(ctx)) {
        if (!hip_ok(hipMemcpy2DAsync(host, host_pitch, dev, row_bytes, row_bytes, rows, hipMemcpyDeviceToHost, s),
                    "Failed to copy data from device to host"))
            return SE_ERR_COPY;
        return hip_ok(hipStreamSynchronize(s), "CUDA transfers did not complete successfully") ? SE_OK : SE_ERR_COPY;
    }
    const size_t per = rows_per_chunk(rows, row_bytes);
    const size_t chunks = (rows + per - 1) / per;
    auto issue = [&](size_t c) {
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        const int b = (int)(c & 1);
        return hip_ok(hipMemcpyAsync(ctx->pinned[b], static_cast<const char*>(dev) + r0 * row_bytes, nr * row_bytes,
                                     hipMemcpyDeviceToHost, s),
                      "Failed to copy data from device to host") &&
               hip_ok(hipEventRecord(ctx->pinned_ev[b], s), "Failed to copy data from device to host");
    };
    if (!issue(0)) return SE_ERR_COPY;
    for (size_t c = 0; c < chunks; ++c) {
        // the other buffer's host copy finished in the previous iteration: refill it
        if (c + 1 < chunks && !issue(c + 1)) return SE_ERR_COPY;
        const int b = (int)(c & 1);
        if (!hip_ok(hipEventSynchronize(ctx->pinned_ev[b]), "CUDA transfers did not complete successfully"))
            return SE_ERR_COPY;
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        copy_rows(static_cast<char*>(host) + r0 * host_pitch, host_pitch, static_cast<const char*>(ctx->pinned[b]),
                  row_bytes, nr, row_bytes);
    }
    return SE_OK;
}

// host u32 rows -> device u16 rows (n values each).  The values are narrowed while they
// are staged; SE_NARROW_FAILED when one exceeds 65535 (the device copy is then incomplete
// and the caller uploads 32-bit values instead).
int upload_rows_narrow16(DeviceContext* ctx, hipStream_t s, uint16_t* dev, const uint32_t* host, size_t host_pitch,
                         size_t rows, size_t n) {
    const size_t row_bytes = n * sizeof(uint16_t);
    if (rows == 0 || n == 0) return SE_OK;
    if (row_bytes > kChunkBytes || !ensure_pinned(ctx)) return SE_NARROW_FAILED;
    const size_t per = rows_per_chunk(rows, row_bytes);
    bool used[2] = {false, false};
    std::atomic<uint32_t> high{0};
    for (size_t r0 = 0, c = 0; r0 < rows; r0 += per, ++c) {
        const size_t nr = std::min(per, rows - r0);
        const int b = (int)(c & 1);
        if (used[b] && !hip_ok(hipEventSynchronize(ctx->pinned_ev[b]), "Failed to copy data from host to device"))
            return SE_ERR_COPY;
        uint16_t* stage = static_cast<uint16_t*>(ctx->pinned[b]);
        const size_t rpt = rows_per_task(nr, row_bytes);
        RowPool::get().run((nr + rpt - 1) / rpt, [&](size_t t) {
            const size_t a0 = t * rpt, a1 = std::min(nr, a0 + rpt);
            uint32_t acc = 0;
            for (size_t r = a0; r < a1; ++r) {
                const uint32_t* src =
                    reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(host) + (r0 + r) * host_pitch);
                uint16_t* dst = stage + r * n;
                for (size_t j = 0; j < n; ++j) {
                    acc |= src[j];
                    dst[j] = (uint16_t)src[j];
                }
            }
            if (acc >> 16) high.fetch_or(1);
        });
        if (high.load()) {
            (void)hipStreamSynchronize(s);
            return SE_NARROW_FAILED;
        }
        if (!hip_ok(hipMemcpyAsync(dev + r0 * n, stage, nr * row_bytes, hipMemcpyHostToDevice, s),
                    "Failed to copy data from host to device") ||
            !hip_ok(hipEventRecord(ctx->pinned_ev[b], s), "Failed to copy data from host to device"))
            return SE_ERR_COPY;
        used[b] = true;
    }
    return SE_OK;
}

// device u16 rows (n values each) -> host u32 rows (pitch host_pitch), widened while they
// are unstaged; returns after the data is in host memory.
int download_rows_widen16(DeviceContext* ctx, hipStream_t s, uint32_t* host, size_t host_pitch, const uint16_t* dev,
                          size_t rows, size_t n) {
    const size_t row_bytes = n * sizeof(uint16_t);
    if (rows == 0 || n == 0) return SE_OK;
    if (row_bytes > kChunkBytes || !ensure_pinned(ctx)) {
        set_error("Failed to copy data from device to host");
        return SE_ERR_COPY;
    }
    const size_t per = rows_per_chunk(rows, row_bytes);
    const size_t chunks = (rows + per - 1) / per;
    auto issue = [&](size_t c) {
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        const int b = (int)(c & 1);
        return hip_ok(hipMemcpyAsync(ctx->pinned[b], dev + r0 * n, nr * row_bytes, hipMemcpyDeviceToHost, s),
                      "Failed to copy data from device to host") &&
               hip_ok(hipEventRecord(ctx->pinned_ev[b], s), "Failed to copy data from device to host");
    };
    if (!issue(0)) return SE_ERR_COPY;
    for (size_t c = 0; c < chunks; ++c) {
        if (c + 1 < chunks && !issue(c + 1)) return SE_ERR_COPY;
        const int b = (int)(c & 1);
        if (!hip_ok(hipEventSynchronize(ctx->pinned_ev[b]), "CUDA transfers did not complete successfully"))
            return SE_ERR_COPY;
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        const uint16_t* stage = static_cast<const uint16_t*>(ctx->pinned[b]);
        const size_t rpt = rows_per_task(nr, row_bytes);
        RowPool::get().run((nr + rpt - 1) / rpt, [&](size_t t) {
            const size_t a0 = t * rpt, a1 = std::min(nr, a0 + rpt);
            for (size_t r = a0; r < a1; ++r) {
                uint32_t* dst = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(host) + (r0 + r) * host_pitch);
                const uint16_t* src = stage + r * n;
                for (size_t j = 0; j < n; ++j) dst[j] = src[j];
            }
        });
    }
    return SE_OK;
}

uint32_t* small_pinned(DeviceContext* ctx) {
    if (!ctx->host_small &&
        !hip_ok(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_small), kSmallPinnedBytes, hipHostMallocDefault),
                "Failed to allocate pinned host memory")) {
        ctx->host_small = nullptr;
        return nullptr;
    }
    return ctx->host_small;
}

void free_pinned(DeviceContext* ctx) {
    if (ctx->host_small) (void)hipHostFree(ctx->host_small);
    ctx->host_small = nullptr;
    for (int b = 0; b < 2; ++b) {
        if (ctx->pinned[b]) (void)hipHostFree(ctx->pinned[b]);
        if (ctx->pinned_ev[b]) (void)hipEventDestroy(ctx->pinned_ev[b]);
        ctx->pinned[b] = nullptr;
        ctx->pinned_ev[b] = nullptr;
    }
}

}  // namespace se
