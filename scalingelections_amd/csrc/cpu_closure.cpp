// cpu_closure.cpp -- the CPU branch of compute_strongest_paths (allow_gpu=False) and the
// serial closure the CLI validates against.
//
// Reference: the OpenMP path behind allow_gpu=False, scaling_elections.cu:604-849 and
// :933-988, and the serial kernel scaling_elections.py:91-125.  This is an explicit CPU
// request of the reference API, never a fallback for the GPU path.
//
// Blocked Floyd-Warshall over B x B blocks (B = max(tile, 32); the result does not depend
// on the blocking), zero padded to a multiple of B, run by a pool of std::threads that meet
// at a barrier between phases:
//   pivot block closed by one thread; row/column panels as one max-min product each with the
//   closed pivot (exact once the pivot block is closed); independent blocks in parallel.
// Diagonal masks are dropped (updates through pivot i or j never raise p[i][j]); the
// diagonal is zeroed when the result is written.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "se_internal.h"

namespace se {

namespace {

class Barrier {
   public:
    explicit Barrier(int n) : n_(n) {}
    void wait() {
        std::unique_lock<std::mutex> lk(m_);
        const uint64_t gen = gen_;
        if (++count_ == n_) {
            count_ = 0;
            ++gen_;
            cv_.notify_all();
        } else {
            cv_.wait(lk, [&] { return gen_ != gen; });
        }
    }

   private:
    std::mutex m_;
    std::condition_variable cv_;
    int n_;
    int count_ = 0;
    uint64_t gen_ = 0;
};

// c[i][j] = max(c[i][j], min(a[i][k], b[k][j])) for i, j, k < B, in k-outer order so that
// aliasing views (pivot block: c = a = b) follow Floyd-Warshall order.
inline void block_update(uint32_t* c, const uint32_t* a, const uint32_t* b, size_t ld, uint32_t B) {
    for (uint32_t k = 0; k < B; ++k) {
        const uint32_t* bk = b + (size_t)k * ld;
        for (uint32_t i = 0; i < B; ++i) {
            uint32_t* ci = c + (size_t)i * ld;
            const uint32_t av = a[(size_t)i * ld + k];
            for (uint32_t j = 0; j < B; ++j) {
                uint32_t r = av < bk[j] ? av : bk[j];
                ci[j] = r > ci[j] ? r : ci[j];
            }
        }
    }
}

// One-shot max-min product into c (no aliasing with a or b): c = max(c, a (x) b).
inline void block_product(uint32_t* __restrict c, const uint32_t* __restrict a, const uint32_t* __restrict b,
                          size_t ld, uint32_t B) {
    for (uint32_t i = 0; i < B; ++i) {
        uint32_t* ci = c + (size_t)i * ld;
        for (uint32_t k = 0; k < B; ++k) {
            const uint32_t av = a[(size_t)i * ld + k];
            const uint32_t* bk = b + (size_t)k * ld;
            for (uint32_t j = 0; j < B; ++j) {
                uint32_t r = av < bk[j] ? av : bk[j];
                ci[j] = r > ci[j] ? r : ci[j];
            }
        }
    }
}

void init_p0(const uint32_t* d, uint32_t n, size_t ld, uint32_t* m, size_t N) {
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            uint32_t v = 0;
            if (i < n && j < n && i != j && d[i * ld + j] > d[j * ld + i]) v = d[i * ld + j];
            m[i * N + j] = v;
        }
}

}  // namespace

int closure_cpu_impl(const uint32_t* d, uint32_t n, size_t ld, uint32_t* p, uint32_t tile, int threads) {
    if (n == 0) return SE_OK;
    const uint32_t B = std::max<uint32_t>(tile, 32);
    const size_t tiles = (n + B - 1) / B;
    const size_t N = tiles * B;
    std::vector<uint32_t> buf;
    try {
        buf.resize(N * N);
    } catch (...) {
        set_error("Failed to allocate host memory");
        return SE_ERR_ALLOC;
    }
    uint32_t* m = buf.data();
    init_p0(d, n, ld, m, N);

    int P = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    P = (int)std::min<size_t>((size_t)P, std::max<size_t>(1, tiles * tiles));
    Barrier bar(P);
    std::atomic<int> stop{0};
    auto at = [&](size_t bi, size_t bj) { return m + bi * B * N + bj * B; };

    auto worker = [&](int id) {
        for (size_t k = 0; k < tiles; ++k) {
            if (id == 0) stop.store(g_interrupted.load());
            bar.wait();
            if (stop.load()) return;
            if (id == 0) block_update(at(k, k), at(k, k), at(k, k), N, B);
            bar.wait();
            // panels: (k, j) row panel and (i, k) column panel, 2 * tiles jobs
            for (size_t q = id; q < 2 * tiles; q += P) {
                size_t x = q % tiles;
                if (x == k) continue;
                if (q < tiles) {
                    // c = max(c, P (x) c_old): copy the old panel block first
                    uint32_t tmp[128 * 128];
                    uint32_t* c = at(k, x);
                    for (uint32_t i = 0; i < B; ++i) std::memcpy(tmp + i * B, c + i * N, B * sizeof(uint32_t));
                    for (uint32_t i = 0; i < B; ++i) {
                        uint32_t* ci = c + (size_t)i * N;
                        for (uint32_t kk = 0; kk < B; ++kk) {
                            const uint32_t av = at(k, k)[(size_t)i * N + kk];
                            const uint32_t* bk = tmp + kk * B;
                            for (uint32_t j = 0; j < B; ++j) {
                                uint32_t r = av < bk[j] ? av : bk[j];
                                ci[j] = r > ci[j] ? r : ci[j];
                            }
                        }
                    }
                } else {
                    uint32_t tmp[128 * 128];
                    uint32_t* c = at(x, k);
                    for (uint32_t i = 0; i < B; ++i) std::memcpy(tmp + i * B, c + i * N, B * sizeof(uint32_t));
                    const uint32_t* pv = at(k, k);
                    for (uint32_t i = 0; i < B; ++i) {
                        uint32_t* ci = c + (size_t)i * N;
                        for (uint32_t kk = 0; kk < B; ++kk) {
                            const uint32_t av = tmp[i * B + kk];
                            const uint32_t* bk = pv + (size_t)kk * N;
                            for (uint32_t j = 0; j < B; ++j) {
                                uint32_t r = av < bk[j] ? av : bk[j];
                                ci[j] = r > ci[j] ? r : ci[j];
                            }
                        }
                    }
                }
            }
            bar.wait();
            for (size_t q = id; q < tiles * tiles; q += P) {
                size_t bi = q / tiles, bj = q % tiles;
                if (bi == k || bj == k) continue;
                block_product(at(bi, bj), at(bi, k), at(k, bj), N, B);
            }
            bar.wait();
        }
    };
    if (B > 128) {
        set_error("Unsupported tile size");
        return SE_ERR_TILE;
    }
    std::vector<std::thread> pool;
    for (int t = 1; t < P; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    if (stop.load()) {
        set_error("Stopped by signal");
        return SE_ERR_INTERRUPTED;
    }
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(p + i * n, m + i * N, n * sizeof(uint32_t));
        p[i * n + i] = 0;
    }
    return SE_OK;
}

int closure_cpu_serial_impl(const uint32_t* d, uint32_t n, size_t ld, uint32_t* p) {
    if (n == 0) return SE_OK;
    init_p0(d, n, ld, p, n);
    for (size_t i = 0; i < n; ++i) {
        if ((i & 63) == 0 && g_interrupted.load()) {
            set_error("Stopped by signal");
            return SE_ERR_INTERRUPTED;
        }
        const uint32_t* pi = p + i * n;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            uint32_t* pj = p + j * n;
            const uint32_t a = pj[i];
            for (size_t k = 0; k < n; ++k) {
                uint32_t r = a < pi[k] ? a : pi[k];
                pj[k] = r > pj[k] ? r : pj[k];
            }
        }
    }
    for (size_t i = 0; i < n; ++i) p[i * n + i] = 0;
    return SE_OK;
}

}  // namespace se
