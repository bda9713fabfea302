// count_tiles.h -- workgroup-level pairwise preference counting, shared by the
// single-election counter (count.hip) and the batched counter (batch.hip).
//
// Reference: populate_preferences_from_ranking, scaling_elections.py:37-48.  With ballots
// given as position arrays (positions[v][c] = place of candidate c), every voter adds one
// to d[i][j] exactly when positions[v][i] < positions[v][j]; for a strict order per voter
// d[j][i] = m - d[i][j] (i != j).
//
// Packed counting.  Positions are staged into LDS K voters to a 32-bit word:
//   K = 4 (n <= 128):   one byte per voter (positions < 128),
//   K = 2 (n <= 32768): one 16-bit half per voter (positions < 32768),
//   K = 1 (otherwise):  one int32 per voter, plain compare-and-add.
// For K > 1 the row side carries a guard bit on top of every lane (0x80 / 0x8000), so
//   z = (x | guard) - y,   lane value H + x - y  with H = 128 / 32768,
// lies in [1, 2H) and never borrows from its neighbour; its top bit is [x >= y].  Read as
// a signed lane, z = x - y + H - 2H [x >= y], so the signed dot product of z with ones
// (v_dot4c_i32_i8 / v_dot2c_i32_i16) accumulates, over the M voter slots of a slice,
//   acc = (S_i - S_j) + H M - 2H G,   G = #{slots : x >= y},
// where S_c is the sum of candidate c's positions over the slice.  Hence
//   G = (S_i - S_j + H M - acc) / 2H   (exact in uint32 while 2H G < 2^32),
//   d[i][j] = M - G                    (padding slots hold x = y = 0 and count in G).
// Two VALU instructions (subtract, dot) per K voter-pairs instead of two per voter-pair.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace se {

constexpr uint32_t kCT = 128;     // tile side
constexpr uint32_t kVW = 32;      // packed words per candidate staged per LDS round
constexpr uint32_t kCThreads = 256;

// Largest voter slice (slots) a K-packed accumulation stays exact for.
__host__ __device__ constexpr uint32_t count_max_slice(uint32_t K) {
    return K == 4 ? (1u << 24) - 4 : K == 2 ? 65534u : 0xFFFFFFFFu;
}
// Packing width for an election of n candidates and (per slice) m voters.
__host__ __device__ constexpr uint32_t count_pack(uint32_t n, uint32_t m) {
    return n <= 128 ? 4u : (n <= 32768 && m <= count_max_slice(2)) ? 2u : 1u;
}

template <uint32_t K>
struct PackTraits;
template <>
struct PackTraits<4> {
    static constexpr uint32_t guard = 0x80808080u, H = 128u, shift = 8u, range_mask = ~0x7Fu;
};
template <>
struct PackTraits<2> {
    static constexpr uint32_t guard = 0x80008000u, H = 32768u, shift = 16u, range_mask = ~0x7FFFu;
};
template <>
struct PackTraits<1> {
    static constexpr uint32_t guard = 0u, H = 0u, shift = 0u, range_mask = 0u;
};

typedef short short2_t __attribute__((ext_vector_type(2)));
typedef unsigned short ushort2_t __attribute__((ext_vector_type(2)));

// One K-packed word of row side x against column side y.
template <uint32_t K>
__device__ __forceinline__ void tally(int32_t& acc, uint32_t x, uint32_t y) {
    if constexpr (K == 4) {
        acc = __builtin_amdgcn_sdot4((int)(x - y), 0x01010101, acc, false);
    } else if constexpr (K == 2) {
        const uint32_t z = x - y;
        acc = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, z), short2_t{1, 1}, acc, false);
    } else {
        acc += (int32_t)((int32_t)x < (int32_t)y);
    }
}

// One row of NB (8 or 4) cells: the subtractions first, then the dot products, so no dot
// waits on the instruction just issued (measured 2-3% faster than alternating them).
template <uint32_t K, int NB = 8>
__device__ __forceinline__ void tally_row(int32_t* acc, uint32_t x, const uint32_t* y) {
    if constexpr (K == 1) {
#pragma unroll
        for (int b = 0; b < NB; ++b) tally<1>(acc[b], x, y[b]);
    } else {
        uint32_t z[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) z[b] = x - y[b];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if constexpr (K == 4)
                acc[b] = __builtin_amdgcn_sdot4((int)z[b], 0x01010101, acc[b], false);
            else
                acc[b] = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, z[b]), short2_t{1, 1}, acc[b], false);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// How counts reach d: stored, added (the only writer of those cells in this launch: a
// later voter pass), or added atomically (voter slices of one tile running concurrently
// into a zeroed d).
constexpr uint32_t kPutStore = 0, kPutAdd = 1, kPutAtomic = 2;

__device__ __forceinline__ void put1(uint32_t* p, uint32_t v, uint32_t mode) {
    if (mode == kPutAtomic)
        atomicAdd(p, v);
    else
        *p = mode == kPutAdd ? *p + v : v;
}
__device__ __forceinline__ void put4(uint32_t* p, uint4 v, uint32_t mode) {
    if (mode == kPutAtomic) {
        atomicAdd(p, v.x);
        atomicAdd(p + 1, v.y);
        atomicAdd(p + 2, v.z);
        atomicAdd(p + 3, v.w);
        return;
    }
    uint4* q = reinterpret_cast<uint4*>(p);
    if (mode == kPutAdd) {
        const uint4 o = *q;
        v = make_uint4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
    }
    *q = v;
}

// d[i][j] from the accumulator of one cell (see the header comment).
template <uint32_t K>
__device__ __forceinline__ uint32_t finish(int32_t acc, uint32_t si, uint32_t sj, uint32_t M) {
    if constexpr (K == 1) {
        return (uint32_t)acc;
    } else {
        const uint32_t X = si - sj + PackTraits<K>::H * M - (uint32_t)acc;
        return M - (X >> (PackTraits<K>::shift));
    }
}

// Staging assignment of a thread: side (0 rows, 1 columns), candidate c of that side and
// word phase: the 128 threads of a side split r candidates G = 128 / r ways, thread
// (c, sub) staging words sub, sub + G, ... (all 256 threads load when r is small).
struct StageLane {
    uint32_t side, c, sub, G;
    bool live;
};

__device__ __forceinline__ StageLane stage_lane(uint32_t ri, uint32_t rj, bool split) {
    StageLane L;
    const uint32_t t = threadIdx.x, local = t & (kCT - 1);
    L.side = t >> 7;
    const uint32_t r = L.side ? rj : ri;
    if (!split) {  // thread = candidate
        L.G = 1u;
        L.c = local;
        L.sub = 0u;
        L.live = local < r;
        return L;
    }
    L.G = r ? max(1u, kCT / r) : 1u;
    L.c = r ? local % r : 0u;
    L.sub = r ? local / r : 0u;
    L.live = r != 0 && L.sub < L.G;
    return L;
}

// Stages words [w0, w0 + kVW) of the slice (voters vb..ve, K per word, zero padded) for
// rows I0 + [0, ri) into PI (guarded) and columns J0 + [0, rj) into PJ; columns of the
// staging arrays past ri / rj are left as they are (their cells are never stored).  Each
// thread adds its positions to `s` (its share of the candidate's sum) and ORs them into
// `bits` (range check).
// SPLIT = false: one thread per candidate (G = 1), the loop the full tiles use.
template <uint32_t K, bool SPLIT, typename Pos>
__device__ __forceinline__ void stage_words(const Pos* __restrict__ pos, size_t ldp, uint32_t vb, uint32_t ve,
                                            uint32_t w0, uint32_t I0, uint32_t J0, const StageLane& L,
                                            uint32_t (*PI)[kCT], uint32_t (*PJ)[kCT], uint32_t& s,
                                            uint32_t& bits) {
    if (!L.live) return;
    const uint32_t base = L.side ? J0 : I0;
    uint32_t(*P)[kCT] = L.side ? PJ : PI;
    const uint32_t guard = L.side ? 0u : PackTraits<K>::guard;
    const Pos* col = pos + base + L.c;
    const uint32_t first = SPLIT ? L.sub : 0u, step = SPLIT ? L.G : 1u;
    // Q words per step with every load unconditional, so the loads of a step issue back to
    // back (a guarded load per voter made every word wait for its own round trip).  A round
    // wholly inside the slice (every round but the last) reads straight through; the last
    // one reads voter ve - 1 (in bounds: ve > vb) for voters past the slice and counts them
    // as 0.  The position sum takes one unsigned dot product per packed word; the range
    // check ORs every unpacked position.  Split staging ends with its last (< Q) words one at
    // a time.
    const bool whole = vb + (w0 + kVW) * K <= ve;  // uniform over the workgroup
    auto packed = [&](uint32_t w, auto whole_tag) {
        constexpr bool WHOLE = decltype(whole_tag)::value;
        const uint32_t v0 = vb + (w0 + w) * K;
        const Pos* pv = col + (size_t)v0 * ldp;
        uint32_t word = 0;
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            uint32_t p;
            if constexpr (WHOLE) {
                p = (uint32_t)pv[(size_t)k * ldp];
            } else {
                const uint32_t v = v0 + k;
                const uint32_t raw = (uint32_t)col[(size_t)min(v, ve - 1) * ldp];
                p = v < ve ? raw : 0u;
            }
            if constexpr (K > 1) {
                word |= p << (PackTraits<K>::shift * k);
                bits |= p;
            } else {
                word = p;
            }
        }
        if constexpr (K == 4)
            s = __builtin_amdgcn_udot4(word, 0x01010101u, s, false);
        else if constexpr (K == 2)
            s = __builtin_amdgcn_udot2(__builtin_bit_cast(ushort2_t, word), ushort2_t{1, 1}, s, false);
        return word;
    };
    // words per step (K loads each): on rounds inside the slice all of a step's loads are
    // in flight together; the last round's clamped loads may issue one at a time
    constexpr uint32_t Q = 4u;
    auto run = [&](auto whole_tag) {
        uint32_t w = first;
        for (; w + (Q - 1) * step < kVW; w += Q * step) {
            uint32_t wd[Q];
#pragma unroll
            for (uint32_t u = 0; u < Q; ++u) wd[u] = packed(w + u * step, whole_tag);
#pragma unroll
            for (uint32_t u = 0; u < Q; ++u) P[w + u * step][L.c] = wd[u] | guard;
        }
        if constexpr (SPLIT)
            for (; w < kVW; w += step) P[w][L.c] = packed(w, whole_tag) | guard;
    };
    if (whole)
        run(std::true_type{});
    else
        run(std::false_type{});
}

// Per-candidate position sums from the threads' shares: sums[side * 128 + c] (256 words of
// LDS, not aliasing the staging arrays' live data).  Ends with a barrier.
__device__ __forceinline__ void reduce_sums(uint32_t* sums, const StageLane& L, uint32_t s) {
    __syncthreads();
    sums[threadIdx.x] = 0u;
    __syncthreads();
    if (L.live) atomicAdd(&sums[L.side * kCT + L.c], s);
    __syncthreads();
}

// d tile (I, J), I <= J, full 128 x 128, of one voter slice [vb, ve) of an election with
// n candidates: positions pos (row pitch ldp) -> d (pitch ldd).  Lane (ty, tx) owns rows
// {4 ty + a, 64 + 4 ty + a} and columns {4 tx + b, 64 + 4 tx + b} (a, b < 4): per word a
// wave reads four 16-byte row chunks (broadcast) and 256 contiguous column bytes.
// On a diagonal tile (I == J) every lane's bottom-left quadrant (rows 64 + 4 ty + a,
// columns 4 tx + b) lies below the diagonal: it is not counted, and those cells are
// written as the complement of the top-right quadrant, d[j][i] = m - d[i][j] (a strict
// order per voter), which saves a quarter of the tile's counting.
// PI/PJ: LDS staging of kVW words x kCT candidates; red: 2 x kCT words (may alias PI).
// `put` (kPutStore / kPutAdd / kPutAtomic) says how counts reach d.  Returns the OR of all
// staged positions (range check of the caller).
// Positions P: int32 (any caller) or uint16 (the host entry point's 16-bit transit).
template <uint32_t K, typename P>
__device__ __forceinline__ uint32_t count_tile(const P* __restrict__ pos, uint32_t vb, uint32_t ve,
                                               uint32_t n, size_t ldp, uint32_t* __restrict__ d, size_t ldd,
                                               uint32_t I, uint32_t J, uint32_t (*PI)[kCT], uint32_t (*PJ)[kCT],
                                               uint32_t* red, uint32_t put) {
    const uint32_t I0 = I * kCT, J0 = J * kCT;
    const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const uint32_t ri = min(kCT, n - I0), rj = min(kCT, n - J0);
    const uint32_t words = (ve - vb + K - 1) / K;

    int32_t acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0;
    uint32_t s = 0, bits = 0;
    const StageLane SL = stage_lane(ri, rj, false);

    const bool diag = I == J;
    auto count_words = [&](auto diag_tag) {
        constexpr bool DIAG = decltype(diag_tag)::value;
        for (uint32_t w0 = 0; w0 < words; w0 += kVW) {
            __syncthreads();
            stage_words<K, false, P>(pos, ldp, vb, ve, w0, I0, J0, SL, PI, PJ, s, bits);
            __syncthreads();
            const uint32_t wend = min(kVW, words - w0);
            for (uint32_t w = 0; w < wend; ++w) {
                const uint4 i0 = *reinterpret_cast<const uint4*>(&PI[w][4 * ty]);
                const uint4 i1 = *reinterpret_cast<const uint4*>(&PI[w][64 + 4 * ty]);
                const uint4 j0 = *reinterpret_cast<const uint4*>(&PJ[w][4 * tx]);
                const uint4 j1 = *reinterpret_cast<const uint4*>(&PJ[w][64 + 4 * tx]);
                const uint32_t x[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
                const uint32_t y[8] = {j0.x, j0.y, j0.z, j0.w, j1.x, j1.y, j1.z, j1.w};
#pragma unroll
                for (int a = 0; a < 4; ++a) tally_row<K>(acc[a], x[a], y);
#pragma unroll
                for (int a = 4; a < 8; ++a) {
                    if constexpr (DIAG)
                        tally_row<K, 4>(acc[a] + 4, x[a], y + 4);
                    else
                        tally_row<K>(acc[a], x[a], y);
                }
            }
        }
    };
    if (diag)
        count_words(std::true_type{});
    else
        count_words(std::false_type{});
    reduce_sums(red, SL, s);  // row sums in red[0, 128), column sums in red[128, 256)

    const uint32_t M = words * K, mreal = ve - vb;
    const bool full = ri == kCT && rj == kCT && (ldd & 3) == 0 && ((uintptr_t)d & 15) == 0;
    uint32_t out[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const uint32_t li = a < 4 ? 4 * ty + a : 64 + 4 * ty + (a - 4);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint32_t lj = b < 4 ? 4 * tx + b : 64 + 4 * tx + (b - 4);
            out[a][b] = finish<K>(acc[a][b], red[li], red[kCT + lj], M);
        }
    }
    if (full) {
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const uint32_t i = I0 + (a < 4 ? 4 * ty + a : 64 + 4 * ty + (a - 4));
            if (a < 4 || !diag)
                put4(d + (size_t)i * ldd + J0 + 4 * tx, make_uint4(out[a][0], out[a][1], out[a][2], out[a][3]),
                     put);
            put4(d + (size_t)i * ldd + J0 + 64 + 4 * tx, make_uint4(out[a][4], out[a][5], out[a][6], out[a][7]),
                 put);
        }
        if (diag) {  // bottom-left quadrant: complement of the top-right one
#pragma unroll
            for (int b = 4; b < 8; ++b) {
                const uint32_t j = J0 + 64 + 4 * tx + (b - 4);
                put4(d + (size_t)j * ldd + I0 + 4 * ty,
                     make_uint4(mreal - out[0][b], mreal - out[1][b], mreal - out[2][b], mreal - out[3][b]),
                     put);
            }
        } else {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint32_t j = J0 + (b < 4 ? 4 * tx + b : 64 + 4 * tx + (b - 4));
                put4(d + (size_t)j * ldd + I0 + 4 * ty,
                     make_uint4(mreal - out[0][b], mreal - out[1][b], mreal - out[2][b], mreal - out[3][b]),
                     put);
                put4(d + (size_t)j * ldd + I0 + 64 + 4 * ty,
                     make_uint4(mreal - out[4][b], mreal - out[5][b], mreal - out[6][b], mreal - out[7][b]),
                     put);
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const uint32_t li = a < 4 ? 4 * ty + a : 64 + 4 * ty + (a - 4);
            if (li >= ri) continue;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint32_t lj = b < 4 ? 4 * tx + b : 64 + 4 * tx + (b - 4);
                if (lj >= rj) continue;
                // diagonal tile: the bottom-left quadrant (a >= 4, b < 4) was not counted;
                // it is written from the top-right one (a < 4, b >= 4) by complement
                if (!diag || a < 4 || b >= 4) put1(d + (size_t)(I0 + li) * ldd + J0 + lj, out[a][b], put);
                if (!diag || (a < 4 && b >= 4))
                    put1(d + (size_t)(J0 + lj) * ldd + I0 + li, mreal - out[a][b], put);
            }
        }
    }
    return bits;
}

// Voter-parallel counting of one sub-block of d: rows [I0, I0 + ri) x columns [J0, J0 + rj)
// (ri, rj <= 128) of voter slice [vb, ve).  The sub-block is cut into 8 x 8 blocks and the
// 256 lanes into G = 256 / blocks word groups; group g counts staged words g, g + G, ...
// and the partial blocks are summed through LDS (`red`, at least 256 * 8 + 256 words,
// may alias PI).  On a diagonal sub-block (I0 == J0) only blocks bi <= bj are counted; a
// strict order per voter gives the lower ones by complement, d[j][i] = m - d[i][j] (the
// diagonal counts itself: 0).  Off-diagonal sub-blocks write d[i][j] and, by complement,
// d[j][i].  `put` says how counts reach d (put1).  Returns the OR of all staged
// positions.
template <uint32_t K, typename P>
__device__ __forceinline__ uint32_t count_block(const P* __restrict__ pos, uint32_t vb, uint32_t ve,
                                                size_t ldp, uint32_t I0, uint32_t J0, uint32_t ri, uint32_t rj,
                                                uint32_t* __restrict__ d, size_t ldd, uint32_t (*PI)[kCT],
                                                uint32_t (*PJ)[kCT], uint32_t* red, uint32_t put) {
    const uint32_t t = threadIdx.x;
    const bool diag = I0 == J0;
    const uint32_t nbi = (ri + 7) / 8, nbj = (rj + 7) / 8;
    const uint32_t blocks = diag ? nbi * (nbi + 1) / 2 : nbi * nbj;
    const uint32_t G = kCThreads / blocks;
    const uint32_t b = t % blocks, g = t / blocks;
    uint32_t bi, bj;
    if (diag) {  // b-th pair bi <= bj, row-major over the upper triangle of blocks
        uint32_t r = b;
        bi = 0;
        while (r >= nbi - bi) {
            r -= nbi - bi;
            ++bi;
        }
        bj = bi + r;
    } else {
        bi = b / nbj;
        bj = b % nbj;
    }
    const uint32_t words = (ve - vb + K - 1) / K;
    int32_t acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[a][c] = 0;
    uint32_t s = 0, bits = 0;
    const StageLane SL = stage_lane(ri, rj, true);
    for (uint32_t w0 = 0; w0 < words; w0 += kVW) {
        __syncthreads();
        stage_words<K, true, P>(pos, ldp, vb, ve, w0, I0, J0, SL, PI, PJ, s, bits);
        __syncthreads();
        if (g < G) {
            const uint32_t wend = min(kVW, words - w0);
            for (uint32_t w = g; w < wend; w += G) {
                const uint4 i0 = *reinterpret_cast<const uint4*>(&PI[w][8 * bi]);
                const uint4 i1 = *reinterpret_cast<const uint4*>(&PI[w][8 * bi + 4]);
                const uint4 j0 = *reinterpret_cast<const uint4*>(&PJ[w][8 * bj]);
                const uint4 j1 = *reinterpret_cast<const uint4*>(&PJ[w][8 * bj + 4]);
                const uint32_t x[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
                const uint32_t y[8] = {j0.x, j0.y, j0.z, j0.w, j1.x, j1.y, j1.z, j1.w};
#pragma unroll
                for (int a = 0; a < 8; ++a) tally_row<K>(acc[a], x[a], y);
            }
        }
    }
    uint32_t* sums = red + kCThreads * 8;  // row sums [0, 128), column sums [128, 256)
    reduce_sums(sums, SL, s);
    const uint32_t M = words * K, mreal = ve - vb;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 8; ++c) red[t * 8 + c] = (uint32_t)acc[a][c];
        __syncthreads();
        for (uint32_t q = t; q < blocks * 8; q += kCThreads) {
            const uint32_t bb = q / 8, c = q % 8;
            uint32_t sum = 0;
            for (uint32_t gg = 0; gg < G; ++gg) sum += red[(gg * blocks + bb) * 8 + c];
            uint32_t qi, qj;  // block coordinates of bb
            if (diag) {
                uint32_t r = bb;
                qi = 0;
                while (r >= nbi - qi) {
                    r -= nbi - qi;
                    ++qi;
                }
                qj = qi + r;
            } else {
                qi = bb / nbj;
                qj = bb % nbj;
            }
            const uint32_t li = 8 * qi + a, lj = 8 * qj + c;
            if (li >= ri || lj >= rj) continue;
            const uint32_t v = finish<K>((int32_t)sum, sums[li], sums[kCT + lj], M);
            const uint32_t i = I0 + li, j = J0 + lj;
            put1(d + (size_t)i * ldd + j, v, put);
            if (!(diag && qi == qj)) put1(d + (size_t)j * ldd + i, mreal - v, put);  // diagonal block: both triangles counted
        }
    }
    return bits;
}

}  // namespace se
