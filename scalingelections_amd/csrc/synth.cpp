// synth.cpp -- seeded Mallows ballots for the mixed-size batch and serving workloads.
//
// Not a reference interface: the reference draws impartial-culture ballots with
// np.random.permutation (scaling_elections.py:478-480).  The batch workloads need Mallows
// ballots with dispersion phi around a reference order, drawn by repeated insertion: item k
// of the reference order is inserted at position j <= k with probability proportional to
// phi^(k - j) (phi = 1 is impartial culture).  Positions are decoded without building the
// list: the last item inserted keeps its insertion index in the final order, and removing
// items from the last back to the first restores each earlier list, so item k lands in the
// ins_k-th free slot once the slots of items k+1.. are taken.  A bitset of free slots and a
// rank-select per item make a ballot O(C * C / 64).
//
// Randomness: splitmix64 streams keyed by (seed, election, voter), so the output does not
// depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "se_internal.h"

namespace se {
namespace {

inline uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double uniform(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// position of the r-th (0-based) set bit
inline uint32_t select_bit(const uint64_t* w, uint32_t words, uint32_t r) {
    for (uint32_t i = 0; i < words; ++i) {
        uint32_t c = (uint32_t)__builtin_popcountll(w[i]);
        if (r < c) {
            uint64_t x = w[i];
            for (uint32_t k = 0; k < r; ++k) x &= x - 1;
            return i * 64 + (uint32_t)__builtin_ctzll(x);
        }
        r -= c;
    }
    return 0xFFFFFFFFu;  // unreachable for r < popcount
}

void election(uint32_t C, uint32_t m, double phi, const int32_t* ref, uint64_t seed, uint32_t e, int32_t* out) {
    const uint32_t words = (C + 63) / 64;
    std::vector<uint64_t> freebits(words);
    std::vector<uint32_t> ins(C);
    std::vector<double> tot(C);
    const bool uniform_ins = phi >= 1.0;
    const double lp = uniform_ins ? 0.0 : std::log(phi);
    double pw = phi;
    for (uint32_t k = 0; k < C; ++k) {  // 1 - phi^(k+1)
        tot[k] = 1.0 - pw;
        pw *= phi;
    }
    for (uint32_t v = 0; v < m; ++v) {
        uint64_t s = seed * 0x100000001B3ull ^ ((uint64_t)e << 32) ^ v;
        (void)splitmix(s);
        for (uint32_t k = 0; k < C; ++k) {
            const double u = uniform(s);
            uint32_t off;
            if (uniform_ins) {
                off = (uint32_t)(u * (double)(k + 1));
            } else {
                const double x = std::floor(std::log1p(-u * tot[k]) / lp);
                off = x < 0 ? 0u : (x > (double)k ? k : (uint32_t)x);
            }
            if (off > k) off = k;
            ins[k] = k - off;
        }
        for (uint32_t i = 0; i < words; ++i) freebits[i] = ~0ull;
        if (C % 64) freebits[words - 1] = (1ull << (C % 64)) - 1;
        int32_t* row = out + (size_t)v * C;
        for (uint32_t k = C; k-- > 0;) {
            const uint32_t slot = select_bit(freebits.data(), words, ins[k]);
            freebits[slot / 64] &= ~(1ull << (slot % 64));
            row[ref[k]] = (int32_t)slot;
        }
    }
}

}  // namespace
}  // namespace se

extern "C" int se_synth_mallows(uint32_t elections, const uint32_t* candidates, const uint32_t* voters,
                                const double* phi, const int32_t* reference, uint64_t seed, int32_t* positions,
                                int threads) {
    se::clear_error();
    if (elections && (!candidates || !voters || !phi || !reference || !positions)) {
        se::set_error("invalid Mallows generator buffers");
        return SE_ERR_INVALID;
    }
    std::vector<uint64_t> pos_off(elections + 1, 0), ref_off(elections + 1, 0);
    for (uint32_t e = 0; e < elections; ++e) {
        pos_off[e + 1] = pos_off[e] + (uint64_t)candidates[e] * voters[e];
        ref_off[e + 1] = ref_off[e] + candidates[e];
        if (!(phi[e] > 0.0) || phi[e] > 1.0) {
            se::set_error("Mallows dispersion must lie in (0, 1]");
            return SE_ERR_INVALID;
        }
    }
    std::vector<uint32_t> order(elections);
    for (uint32_t e = 0; e < elections; ++e) order[e] = e;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return (uint64_t)candidates[a] * candidates[a] * voters[a] > (uint64_t)candidates[b] * candidates[b] * voters[b];
    });
    int P = threads > 0 ? threads : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t i; (i = next.fetch_add(1)) < elections;) {
            const uint32_t e = order[i];
            se::election(candidates[e], voters[e], phi[e], reference + ref_off[e], seed, e, positions + pos_off[e]);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < P; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return SE_OK;
}
