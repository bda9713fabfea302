// se_internal.h -- shared host-side plumbing of the HIP library: error state, per-device
// context (stream, lock, grow-only scratch cache) and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/scaling_elections_amd.h"

namespace se {

// Thread-local message of the last failing call (returned by se_last_error()).
void set_error(const std::string& msg);
void clear_error();

// Interruption flag shared by the CPU and GPU drivers.
extern std::atomic<int> g_interrupted;

// Scratch slots of a device context.  Each slot is a grow-only allocation reused by
// later calls on the same device.
enum Slot : int {
    kSlotWork = 0,     // padded closure matrix (u16 or u32)
    kSlotPrefs = 1,    // staged input matrix (host entry points)
    kSlotPaths = 2,    // staged output matrix (host entry points; se_count's 16-bit counts)
    kSlotSmall = 3,    // scalars: max(p0), counters
    kSlotPivot = 4,    // pivot-tile hand-off between kernels
    kSlotAux = 5,      // positions / wins staging
    kSlotBatchD = 6,   // batched preference matrices
    kSlotBatchMeta = 7,  // batch metadata / per-launch wins
    kSlotCountPart = 8,     // per-slice partial preference matrices
    kSlotCountFlag = 9,     // counter range-check flag
    kSlotBatchPivot = 10,   // per-election pivot hand-off of the round-closed batch elections
    kSlotBatchRounds = 11,  // their per-round work lists
    kSlotCount = 12
};

struct DeviceContext {
    int device = -1;
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    std::mutex lock;
    void* ptr[kSlotCount] = {};
    size_t cap[kSlotCount] = {};
    void* pinned[2] = {nullptr, nullptr};          // pinned staging chunks (transfer.hip)
    hipEvent_t pinned_ev[2] = {nullptr, nullptr};  // last DMA of each chunk
    uint32_t* host_small = nullptr;                // pinned 4 KB for small read-backs (transfer.hip)
};

// Host <-> device matrix copies through the context's pinned staging chunks
// (transfer.hip).  upload_rows returns once the last chunk is queued on s; download_rows
// returns after the data is in host memory.
int upload_rows(DeviceContext* ctx, hipStream_t s, void* dev, const void* host, size_t host_pitch, size_t rows,
                size_t row_bytes);
int download_rows(DeviceContext* ctx, hipStream_t s, void* host, size_t host_pitch, const void* dev, size_t rows,
                  size_t row_bytes);
void free_pinned(DeviceContext* ctx);
// The context's pinned 4 KB read-back buffer (allocated on first use), or nullptr with the
// error set.
uint32_t* small_pinned(DeviceContext* ctx);
constexpr size_t kSmallPinnedBytes = 4096;
// 16-bit transit: u32 host matrix narrowed into a u16 device matrix (SE_NARROW_FAILED when
// a value exceeds 65535), and a u16 device matrix widened into a u32 host matrix.
constexpr int SE_NARROW_FAILED = 1;
int upload_rows_narrow16(DeviceContext* ctx, hipStream_t s, uint16_t* dev, const uint32_t* host, size_t host_pitch,
                         size_t rows, size_t n);
int download_rows_widen16(DeviceContext* ctx, hipStream_t s, uint32_t* host, size_t host_pitch, const uint16_t* dev,
                          size_t rows, size_t n);

// Look-ahead modes of the closure schedule (the tuning table's lookahead values):
// off; the next pivot's strips in a launch of their own; those strips updated by the
// panel workgroups inside the round's one launch (16-bit, pivot 64).
constexpr uint32_t kLaOff = 0, kLaStrips = 1, kLaFused = 2;

// Launch shapes from the tuning table (tune.cpp).  width: SE_TUNE_F16/U16/U32; N: padded
// side.  The batch threshold: elections with at least that many candidates share
// per-round launches.
void tuned_closure_shape(uint32_t N, uint32_t width, uint32_t* pivot, uint32_t* lookahead);
uint32_t tuned_batch_rounds_min_c();

// Context of the current HIP device (created on first use).  Returns nullptr and sets
// the error when no device is usable.
DeviceContext* current_context();

// Ensures slot `s` holds at least `bytes`; returns the device pointer or nullptr (error
// set to "Failed to allocate memory on device").
void* scratch(DeviceContext* ctx, Slot s, size_t bytes);

inline bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::string msg = what;
    if (msg.empty()) msg = hipGetErrorString(e);
    set_error(msg);
    return false;
}

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

}  // namespace se

#define SE_CHECK_LAUNCH(what)                                      \
    do {                                                           \
        hipError_t _e = hipGetLastError();                         \
        if (_e != hipSuccess) {                                    \
            se::set_error(std::string(what) + ": " + hipGetErrorString(_e)); \
            return SE_ERR_DEVICE;                                  \
        }                                                          \
    } while (0)
