// tune.cpp -- the launch-shape table: pivot depth and look-ahead of the closure per value
// width and padded side, and the batch's threshold for shared per-round launches.
//
// No counterpart in the reference, whose GPU path takes one tile size from the caller
// (scaling_elections.cu:895-901).  Here the caller's tile size is validated and ignored;
// the launch shapes come from this table, which never changes a result: every pivot depth
// and schedule computes the same max-min closure.  scalingelections_amd/autotune.py
// measures the choices on the current device and installs them with se_tune_closure.
#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "se_internal.h"

namespace se {

namespace {

// Measured on MI355X by the autotuner (sides 256..16384, synthesized preferences, best of
// every pivot depth and look-ahead mode): on the 16-bit paths pivot 32 wins while rounds
// are short (N <= 2048: 1024 f16 0.50 vs 0.56 ms), pivot 64 above (4096 f16: 2.64 vs
// 3.03 ms); the one-launch-per-round look-ahead wins from 6144 on (8192 f16: 13.56 vs
// 14.24 ms), where a round's tiles outlast the panel workgroups' extra strip update, and
// loses below (4096: 2.83 vs 2.64 ms).  The 32-bit path prefers pivot 32 up to 1024 and
// pivot 64 above (8192: 31.2 vs 37.2 ms); its look-ahead (strip launch, then the next
// panel inside the launch of the round's other tiles) gains 1.4-5% from 2048 on (2048:
// 1.68 vs 1.76 ms; 8192: 30.6 vs 31.4 ms; 32768: 1689 vs 1713 ms).
const std::vector<se_tune_entry>& builtin_table() {
    static const std::vector<se_tune_entry> t = {
        {SE_TUNE_F16, 2048, 32, 1},       {SE_TUNE_F16, 4096, 64, 1},       {SE_TUNE_F16, 0xFFFFFFFFu, 64, 2},
        {SE_TUNE_U16, 2048, 32, 1},       {SE_TUNE_U16, 4096, 64, 1},       {SE_TUNE_U16, 0xFFFFFFFFu, 64, 2},
        {SE_TUNE_U32, 1024, 32, 0},       {SE_TUNE_U32, 0xFFFFFFFFu, 64, 1},
    };
    return t;
}

// Mixed batch (C log-uniform in [8, 512]): the shared per-round launches close the elections
// above 128 candidates in 4.9 ms against 5.3 ms in one workgroup each; for 600..1024
// candidates 11.7 against 19.2 ms.
constexpr uint32_t kBuiltinRoundsMinC = 129;

std::mutex g_tune_mutex;
std::vector<se_tune_entry> g_table;  // empty: the built-in table
uint32_t g_rounds_min_c = 0;         // 0: the built-in value

bool lookup(const std::vector<se_tune_entry>& t, uint32_t N, uint32_t width, se_tune_entry* out) {
    bool found = false;
    for (const se_tune_entry& e : t)
        if (e.width == width && e.max_side >= N && (!found || e.max_side < out->max_side)) {
            *out = e;
            found = true;
        }
    return found;
}

}  // namespace

void tuned_closure_shape(uint32_t N, uint32_t width, uint32_t* pivot, uint32_t* lookahead) {
    se_tune_entry e{width, 0, 64, kLaStrips};
    {
        std::lock_guard<std::mutex> g(g_tune_mutex);
        if (g_table.empty() || !lookup(g_table, N, width, &e)) lookup(builtin_table(), N, width, &e);
    }
    *pivot = e.pivot;
    *lookahead = e.lookahead;
}

uint32_t tuned_batch_rounds_min_c() {
    std::lock_guard<std::mutex> g(g_tune_mutex);
    return g_rounds_min_c ? g_rounds_min_c : kBuiltinRoundsMinC;
}

}  // namespace se

using namespace se;

extern "C" {

int se_tune_closure(const se_tune_entry* entries, uint32_t count) {
    clear_error();
    if (count > SE_TUNE_MAX_ENTRIES || (count && !entries)) {
        set_error("tuning table: at most 64 entries, entries must not be NULL");
        return SE_ERR_INVALID;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const se_tune_entry& e = entries[i];
        if (e.width > SE_TUNE_U32 || (e.pivot != 32 && e.pivot != 64) || e.lookahead > kLaFused ||
            e.max_side == 0 || (e.lookahead == kLaFused && (e.width == SE_TUNE_U32 || e.pivot != 64))) {
            set_error("tuning table: width must be SE_TUNE_F16/U16/U32, pivot 32 or 64, lookahead 0, 1 or 2 "
                      "(2: 16-bit widths with pivot 64), max_side > 0");
            return SE_ERR_INVALID;
        }
    }
    std::lock_guard<std::mutex> g(g_tune_mutex);
    g_table.assign(entries, entries + count);
    return SE_OK;
}

int se_tune_closure_get(se_tune_entry* entries, uint32_t capacity, uint32_t* count) {
    clear_error();
    if (!count || (capacity && !entries)) {
        set_error("count must not be NULL");
        return SE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(g_tune_mutex);
    const std::vector<se_tune_entry>& t = g_table.empty() ? builtin_table() : g_table;
    *count = (uint32_t)t.size();
    std::copy_n(t.begin(), std::min<size_t>(capacity, t.size()), entries);
    return SE_OK;
}

int se_tune_batch(uint32_t rounds_min_candidates) {
    clear_error();
    if (rounds_min_candidates != 0 && (rounds_min_candidates < 129 || rounds_min_candidates > 1025)) {
        set_error("rounds_min_candidates must be 0 or within 129..1025");
        return SE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(g_tune_mutex);
    g_rounds_min_c = rounds_min_candidates;
    return SE_OK;
}

int se_tune_batch_get(uint32_t* rounds_min_candidates) {
    clear_error();
    if (!rounds_min_candidates) {
        set_error("rounds_min_candidates must not be NULL");
        return SE_ERR_INVALID;
    }
    *rounds_min_candidates = tuned_batch_rounds_min_c();
    return SE_OK;
}

}  // extern "C"
