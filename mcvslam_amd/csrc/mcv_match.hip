// Hamming matching kernels for gfx950.
//
//   k_knn2_s, k_knn2_lds, k_knn2_m  Matcher::KnnMatch(vector<Mat>, vector<Mat>) top-2
//                 (src/Matcher.cpp:245-302, include/Matcher.hpp:19-33)
//   k_filter      MatchResKnn::FilterRatio + MatchRes::FilterThreshold
//                 (src/Matcher.cpp:100-111, :23-35)
//   k_distinctive MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cpp:101-150)
//   k_filter_orientation  MatchRes::FilterOrientation (src/Matcher.cpp:44-75)
//
// The reference keeps d = {999, 999}, idx = {0, 0} and replaces with strict
// '<' while scanning trains in order, which is the first two entries of a
// stable sort by distance.  Here each candidate becomes the key
// (distance << 22) | train_index; the two smallest keys are that same pair,
// and (999 << 22) | 0 is the reference's sentinel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcv_device.hpp"

namespace mcv {

constexpr uint32_t kIdxBits = 22;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1;
constexpr uint32_t kSentinel = 999u << kIdxBits;

struct MatchPair {
    const uint8_t* q;     // query descriptors (rows of 32 B)
    const uint8_t* t;     // train descriptors
    const int* nq_dev;    // query count in device memory (or null: use nq)
    const int* nt_dev;    // train count in device memory (or null: use nt)
    int nq, nt;           // counts (capacity when the *_dev pointers are set)
    int32_t* out;         // nq x (i0, d0, i1, d1)
};

// keeps the two smallest keys (k0 <= k1): the new second is the median of
// {k0, k1, key}
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

typedef __attribute__((address_space(4))) const uint32_t kconst_u32;

__device__ __forceinline__ void top2_insert(uint32_t& k0, uint32_t& k1, uint32_t key) {
    k1 = med3_u32(k0, k1, key);
    k0 = min(k0, key);
}

// two keys at once: with k0 <= k1, the smallest of {k0, k1, a, b} is
// min(k0, a, b) and the second smallest is min(k1, med3(k0, a, b))
__device__ __forceinline__ void top2_insert2(uint32_t& k0, uint32_t& k1, uint32_t a, uint32_t b) {
    const uint32_t m = med3_u32(k0, a, b);
    k0 = min(min(k0, a), b);
    k1 = min(k1, m);
}

// Variant S<QPL, UNR>.  256 threads; a lane holds QPL queries (64 QPL per
// workgroup) and the 4 waves split the train set.  Train rows are
// wave-uniform and come through scalar loads (SGPR operands of the XORs),
// UNR rows per iteration; the four waves' top-2 keys merge through LDS.
template <int QPL, int UNR>
__global__ __launch_bounds__(256) void k_knn2_s(const MatchPair* pairs) {
    __shared__ uint32_t sk[2][4][64 * QPL];
    const MatchPair P = pairs[blockIdx.y];
    const int nq = P.nq_dev ? min(*P.nq_dev, P.nq) : P.nq;
    const int nt = P.nt_dev ? min(*P.nt_dev, P.nt) : P.nt;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int qb = blockIdx.x * 64 * QPL;
    if (qb >= nq) return;
    uint32_t q[QPL][8], k0[QPL], k1[QPL];
#pragma unroll
    for (int u = 0; u < QPL; u++) {
        const uint4* qrow = reinterpret_cast<const uint4*>(P.q + (long long)min(qb + 64 * u + lane, nq - 1) * 32);
        const uint4 a = qrow[0], b = qrow[1];
        q[u][0] = a.x; q[u][1] = a.y; q[u][2] = a.z; q[u][3] = a.w;
        q[u][4] = b.x; q[u][5] = b.y; q[u][6] = b.z; q[u][7] = b.w;
        k0[u] = kSentinel;
        k1[u] = kSentinel;
    }
    const int per = (nt + 3) / 4;
    const int j0 = __builtin_amdgcn_readfirstlane(wave * per);
    const int j1 = __builtin_amdgcn_readfirstlane(min(nt, j0 + per));
    // train rows through the constant address space: s_load into SGPRs, which
    // the XORs take as operands (wave-uniform data never occupies the vector
    // memory path).  The rows are written by earlier launches only.
    const kconst_u32* tw = (const kconst_u32*)P.t;
    int j = j0;
    for (; j + UNR <= j1; j += UNR) {
        const kconst_u32* ta = tw + (long long)j * 8;
        static_assert(UNR % 2 == 0, "train rows are taken in pairs");
#pragma unroll
        for (int t = 0; t < UNR; t += 2)
#pragma unroll
            for (int u = 0; u < QPL; u++) {
                uint32_t da = 0, db = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    da += __builtin_popcount(q[u][k] ^ ta[8 * t + k]);
                    db += __builtin_popcount(q[u][k] ^ ta[8 * t + 8 + k]);
                }
                top2_insert2(k0[u], k1[u], (da << kIdxBits) | (uint32_t)(j + t), (db << kIdxBits) | (uint32_t)(j + t + 1));
            }
    }
    for (; j < j1; j++) {
        const kconst_u32* ta = tw + (long long)j * 8;
#pragma unroll
        for (int u = 0; u < QPL; u++) {
            uint32_t d = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) d += __builtin_popcount(q[u][k] ^ ta[k]);
            top2_insert(k0[u], k1[u], (d << kIdxBits) | (uint32_t)j);
        }
    }
#pragma unroll
    for (int u = 0; u < QPL; u++) {
        sk[0][wave][64 * u + lane] = k0[u];
        sk[1][wave][64 * u + lane] = k1[u];
    }
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int u = 0; u < QPL; u++) {
        const int qi = qb + 64 * u + lane;
        uint32_t a0 = k0[u], a1 = k1[u];
#pragma unroll
        for (int w = 1; w < 4; w++) {
            const uint32_t b0 = sk[0][w][64 * u + lane], b1 = sk[1][w][64 * u + lane];
            const uint32_t n1 = min(max(a0, b0), min(a1, b1));
            a0 = min(a0, b0);
            a1 = n1;
        }
        if (qi < nq) {
            int4 o;
            o.x = (int)(a0 & kIdxMask);
            o.y = (int)(a0 >> kIdxBits);
            o.z = (int)(a1 & kIdxMask);
            o.w = (int)(a1 >> kIdxBits);
            // the sentinel keeps index 0, as the reference's d_idx init does
            reinterpret_cast<int4*>(P.out)[qi] = o;
        }
    }
}

// Variant L.  256 threads, lane = query: a workgroup holds 256 queries (one 32-byte
// row in 8 VGPRs per lane) and streams the train set through LDS in tiles
// of kKnnTile descriptors, double buffered, loaded with one coalesced 16-byte
// read per thread.  Every wave reads each train row as a broadcast from LDS;
// per (query, train): 8 XOR + 8 popcount-accumulate + key + 2 for the top-2.
constexpr int kKnnTile = 128;
constexpr int kKnnQ = 256;  // queries per workgroup

__global__ __launch_bounds__(256) void k_knn2_lds(const MatchPair* pairs) {
    __shared__ uint4 s_t[2][2 * kKnnTile];
    const MatchPair P = pairs[blockIdx.y];
    const int nq = P.nq_dev ? min(*P.nq_dev, P.nq) : P.nq;
    const int nt = P.nt_dev ? min(*P.nt_dev, P.nt) : P.nt;
    const int q0 = blockIdx.x * kKnnQ;
    if (q0 >= nq) return;
    const int qi = q0 + threadIdx.x;
    uint32_t q[8];
    {
        const uint4* qrow = reinterpret_cast<const uint4*>(P.q + (long long)min(qi, nq - 1) * 32);
        const uint4 a = qrow[0], b = qrow[1];
        q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w;
        q[4] = b.x; q[5] = b.y; q[6] = b.z; q[7] = b.w;
    }
    uint32_t k0 = kSentinel, k1 = kSentinel;
    const int ntiles = (nt + kKnnTile - 1) / kKnnTile;
    const uint4* tsrc = reinterpret_cast<const uint4*>(P.t);
    auto fetch = [&](int tile) -> uint4 {
        const int ti = tile * kKnnTile + (threadIdx.x >> 1);
        return ti < nt ? tsrc[2ll * ti + (threadIdx.x & 1)] : make_uint4(0u, 0u, 0u, 0u);
    };
    uint4 pf = make_uint4(0u, 0u, 0u, 0u);
    if (ntiles > 0) {
        s_t[0][threadIdx.x] = fetch(0);
        if (ntiles > 1) pf = fetch(1);
    }
    __syncthreads();
    for (int t = 0; t < ntiles; t++) {
        const uint4* tb = s_t[t & 1];
        const int cnt = min(kKnnTile, nt - t * kKnnTile);
        const uint32_t base = (uint32_t)(t * kKnnTile);
        auto dist = [&](int j) -> uint32_t {
            const uint4 a = tb[2 * j], b = tb[2 * j + 1];
            uint32_t d = __builtin_popcount(q[0] ^ a.x);
            d += __builtin_popcount(q[1] ^ a.y);
            d += __builtin_popcount(q[2] ^ a.z);
            d += __builtin_popcount(q[3] ^ a.w);
            d += __builtin_popcount(q[4] ^ b.x);
            d += __builtin_popcount(q[5] ^ b.y);
            d += __builtin_popcount(q[6] ^ b.z);
            d += __builtin_popcount(q[7] ^ b.w);
            return d;
        };
        int j = 0;
        for (; j + 4 <= cnt; j += 4) {
            const uint32_t d0 = dist(j), d1 = dist(j + 1), d2 = dist(j + 2), d3 = dist(j + 3);
            top2_insert(k0, k1, (d0 << kIdxBits) | (base + j));
            top2_insert(k0, k1, (d1 << kIdxBits) | (base + j + 1));
            top2_insert(k0, k1, (d2 << kIdxBits) | (base + j + 2));
            top2_insert(k0, k1, (d3 << kIdxBits) | (base + j + 3));
        }
        for (; j < cnt; j++) top2_insert(k0, k1, (dist(j) << kIdxBits) | (base + j));
        if (t + 1 < ntiles) {
            s_t[(t + 1) & 1][threadIdx.x] = pf;
            if (t + 2 < ntiles) pf = fetch(t + 2);
        }
        __syncthreads();
    }
    if (qi < nq) {
        int4 o;
        o.x = (int)(k0 & kIdxMask);
        o.y = (int)(k0 >> kIdxBits);
        o.z = (int)(k1 & kIdxMask);
        o.w = (int)(k1 >> kIdxBits);
        // the sentinel keeps index 0, as the reference's d_idx init does
        reinterpret_cast<int4*>(P.out)[qi] = o;
    }
}

// Variant M<QT>: the distances on the matrix cores.  Over 0/1 vectors of
// the 256 descriptor bits, popcount(q ^ t) = |q| + |t| - 2 q.t, an int8 dot
// product: the A operand holds -2 per set query bit, the B operand 1 per
// set train bit, and a ninth K step adds |q| + |t| as (|q| >> 2) * 4 +
// (|q| & 3) * 1 + 4 * (|t| >> 2) + 1 * (|t| & 3), so nine
// v_mfma_i32_32x32x32_i8 leave the exact distances of a 32 x 32 block in
// the accumulators.  Operand layout (both operands): lane l holds row (A) or
// column (B) l % 32 and 16 bytes of the K slice [32 m + 16 (l / 32), + 16)
// of step m; the result holds column l % 32 and, in accumulator r, row
// 8 (r / 4) + 4 (l / 32) + r % 4.  A wave keeps QT query tiles of 32 rows as
// A operands (expanded once); the workgroup's four waves expand each
// 128-train super-tile into LDS together (the next super-tile's rows in
// flight meanwhile), and each lane keeps the top-2 keys of its 16 rows over
// the trains of its column; the 32 lanes of a half then merge per row.  Per
// distance the vector unit forms the key and updates the top-2.
typedef int mcv_v4i __attribute__((ext_vector_type(4)));
typedef int mcv_v16i __attribute__((ext_vector_type(16)));

// bits 0..15 of h as 16 bytes of 0 or 1 (byte b = bit b): a nibble times
// 0x204081 puts its bit k at bit 8k (the four shifted copies do not overlap)
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t n) { return __umul24(n & 15u, 0x204081u) & 0x01010101u; }

__device__ __forceinline__ mcv_v4i bits16_to_bytes(uint32_t h) {
    mcv_v4i r;
    r.x = (int)nibble_bytes(h);
    r.y = (int)nibble_bytes(h >> 4);
    r.z = (int)nibble_bytes(h >> 8);
    r.w = (int)nibble_bytes(h >> 12);
    return r;
}

// the same bytes times -2 (0xFE)
__device__ __forceinline__ mcv_v4i bits16_to_m2(uint32_t h) {
    mcv_v4i r = bits16_to_bytes(h);
    r.x = (int)(((uint32_t)r.x << 8) - ((uint32_t)r.x << 1));
    r.y = (int)(((uint32_t)r.y << 8) - ((uint32_t)r.y << 1));
    r.z = (int)(((uint32_t)r.z << 8) - ((uint32_t)r.z << 1));
    r.w = (int)(((uint32_t)r.w << 8) - ((uint32_t)r.w << 1));
    return r;
}

template <int QT, bool kTail>
__device__ __forceinline__ void knn_m_tile(const mcv_v4i (&a)[QT][9], const mcv_v4i* sb, uint32_t j, bool colvalid,
                                           uint32_t (&k0)[QT][16], uint32_t (&k1)[QT][16]) {
    mcv_v4i b[9];
#pragma unroll
    for (int m = 0; m < 9; m++) b[m] = sb[64 * m];
#pragma unroll
    for (int t = 0; t < QT; t++) {
        mcv_v16i acc = {};
#pragma unroll
        for (int m = 0; m < 9; m++) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[t][m], b[m], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            uint32_t key = ((uint32_t)acc[r] << kIdxBits) | j;
            if (kTail) key = colvalid ? key : 0xFFFFFFFFu;
            top2_insert(k0[t][r], k1[t][r], key);
        }
    }
}

constexpr int kMSub = 4;  // 32-train sub-tiles per 128-train super-tile

template <int QT>
__global__ __launch_bounds__(256) void k_knn2_m(const MatchPair* pairs) {
    __shared__ mcv_v4i s_b[kMSub][9][64];  // a super-tile as B operands, [sub-tile][step][lane]
    const MatchPair P = pairs[blockIdx.y];
    const int nq = P.nq_dev ? min(*P.nq_dev, P.nq) : P.nq;
    const int nt = P.nt_dev ? min(*P.nt_dev, P.nt) : P.nt;
    const int q0 = blockIdx.x * 128 * QT;
    if (q0 >= nq) return;
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const int wq0 = q0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 32 * QT;
    // A operands: -2 per set bit, then (|q| >> 2, |q| & 3, 4, 1) in the ninth step
    mcv_v4i a[QT][9];
#pragma unroll
    for (int t = 0; t < QT; t++) {
        const uint4* qrow = reinterpret_cast<const uint4*>(P.q + (long long)min(wq0 + 32 * t + (lane & 31), nq - 1) * 32);
        const uint4 x = qrow[0], y = qrow[1];
        const uint32_t w[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
        uint32_t pc = 0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            a[t][m] = bits16_to_m2(w[m] >> (16 * half));
            pc += __builtin_popcount(w[m]);
        }
        a[t][8] = mcv_v4i{half ? 0 : (int)((pc >> 2) | ((pc & 3u) << 8) | (4u << 16) | (1u << 24)), 0, 0, 0};
    }
    uint32_t k0[QT][16], k1[QT][16];
#pragma unroll
    for (int t = 0; t < QT; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            k0[t][r] = kSentinel;
            k1[t][r] = kSentinel;
        }
    // the ninth step's upper half is zero in every sub-tile
    if (threadIdx.x < 32 * kMSub) s_b[threadIdx.x >> 5][8][32 + (threadIdx.x & 31)] = mcv_v4i{0, 0, 0, 0};
    // thread tid loads words 4 (tid & 1) .. + 3 of train (tid >> 1) of a
    // super-tile (one 16-byte read, the next super-tile's in flight while
    // this one is matched) and writes that train's B operands of those four
    // steps, both halves
    const uint4* tv = reinterpret_cast<const uint4*>(P.t);
    auto fetch = [&](int st) -> uint4 {
        const int j = st * 32 * kMSub + (threadIdx.x >> 1);
        return j < nt ? tv[2ll * j + (threadIdx.x & 1)] : make_uint4(0u, 0u, 0u, 0u);
    };
    const int nsuper = (nt + 32 * kMSub - 1) / (32 * kMSub);
    uint4 raw = make_uint4(0u, 0u, 0u, 0u);
    if (nsuper > 0) raw = fetch(0);
    const int jl = threadIdx.x >> 1, sub = jl >> 5, col = jl & 31, m0 = 4 * (threadIdx.x & 1);
    for (int st = 0; st < nsuper; st++) {
        __syncthreads();  // the previous super-tile's operands are read
        {
            const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                s_b[sub][m0 + k][col] = bits16_to_bytes(w[k]);
                s_b[sub][m0 + k][32 + col] = bits16_to_bytes(w[k] >> 16);
            }
            uint32_t pc = __builtin_popcount(w[0]) + __builtin_popcount(w[1]) + __builtin_popcount(w[2]) +
                          __builtin_popcount(w[3]);
            pc += (uint32_t)__shfl_xor((int)pc, 1, 64);
            if ((threadIdx.x & 1) == 0)
                s_b[sub][8][col] = mcv_v4i{(int)(4u | (1u << 8) | ((pc >> 2) << 16) | ((pc & 3u) << 24)), 0, 0, 0};
        }
        __syncthreads();
        if (st + 1 < nsuper) raw = fetch(st + 1);
        const int tj = st * 32 * kMSub;
        if (tj + 32 * kMSub <= nt) {
            // a full super-tile, one sub-tile at a time (not unrolled: one set of
            // accumulators keeps the kernel at 118 VGPRs, four waves per SIMD,
            // whose products and top-2 updates overlap each other's)
#pragma unroll 1
            for (int sb = 0; sb < kMSub; sb++)
                knn_m_tile<QT, false>(a, &s_b[sb][0][lane], (uint32_t)(tj + 32 * sb + (lane & 31)), true, k0, k1);
        } else {
#pragma unroll
            for (int sb = 0; sb < kMSub; sb++) {
                const int tj0 = tj + 32 * sb;
                if (tj0 < nt) {
                    const uint32_t j = (uint32_t)(tj0 + (lane & 31));
                    knn_m_tile<QT, true>(a, &s_b[sb][0][lane], j, (int)j < nt, k0, k1);
                }
            }
        }
    }
    // the 32 lanes of a half hold the same rows: merge their top-2 keys
#pragma unroll
    for (int t = 0; t < QT; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            uint32_t a0 = k0[t][r], a1 = k1[t][r];
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const uint32_t b0 = (uint32_t)__shfl_xor((int)a0, o, 64), b1 = (uint32_t)__shfl_xor((int)a1, o, 64);
                const uint32_t n1 = min(max(a0, b0), min(a1, b1));
                a0 = min(a0, b0);
                a1 = n1;
            }
            const int qi = wq0 + 32 * t + 8 * (r >> 2) + 4 * half + (r & 3);
            if ((lane & 31) == 0 && qi < nq) {
                int4 o;
                o.x = (int)(a0 & kIdxMask);
                o.y = (int)(a0 >> kIdxBits);
                o.z = (int)(a1 & kIdxMask);
                o.w = (int)(a1 >> kIdxBits);
                // the sentinel keeps index 0, as the reference's d_idx init does
                reinterpret_cast<int4*>(P.out)[qi] = o;
            }
        }
}

// Variant F4: the same block products on the fp4 (e2m1) matrix cores,
// v_mfma_scale_f32_32x32x64_f8f6f4 with unit scales, K = 64 per step.
// Operand layout (probed): lane l holds row (A) or column (B) l % 32 and
// the 32 nibbles of K slice [32 (l / 32), + 32) of the step, nibble j at
// bits 4 j of the 16 bytes; the result layout is that of the 32 x 32 i32
// products.  Steps 0..3 take the descriptor bits (A: -2.0 = 0xC per set
// query bit, B: 1.0 = 0x2 per set train bit), bit q + 4 i of a word as
// nibble i of dword q (any order serves that A and B share: this one is a
// shift and a mask per dword); step 4 adds |q| (A: |q| as
// 6.0 nibbles and a remainder, B: all 1.0), step 5 adds |t| (A: all 1.0, B:
// |t| the same way), so the f32 accumulators hold the distances exactly
// (integers <= 256).  Such a float's top 16 bits order it like the integer
// and keep it exactly, so a key is (bits & 0xFFFF0000) | j: one and-or per
// distance, then the med3/min top-2 update; j < 65536 (the launcher checks).
typedef int mcv_v8i __attribute__((ext_vector_type(8)));
typedef float mcv_v16f __attribute__((ext_vector_type(16)));

// bits q, q + 4, .., q + 28 of w as 8 nibbles of 0 or 1 (bit q + 4 i at bit 4 i)
__device__ __forceinline__ uint32_t nibbles_of(uint32_t w, int q) { return (w >> q) & 0x11111111u; }

// the 32 nibbles of a count c <= 256 over K positions [32 h, 32 h + 32):
// c = 6 n + r as n nibbles 6.0 (0x7), then r (1: 0x2, 2: 0x4, 3: 0x5,
// 4: 0x6, 5: 0x6 then 0x2); at most 44 positions
__device__ __forceinline__ mcv_v4i count_nibbles(uint32_t c, int h) {
    const int n6 = (int)(c / 6u), r = (int)(c - 6u * (uint32_t)n6);
    // the remainder's nibbles as one pair (second << 4 | first)
    const uint32_t rem = r == 0 ? 0u : r == 1 ? 0x2u : r == 2 ? 0x4u : r == 3 ? 0x5u : r == 4 ? 0x6u : 0x26u;
    mcv_v4i v;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int p = n6 - (32 * h + 8 * q);  // first position past the 6.0 run, relative to this dword
        const int f = min(max(p, 0), 8);
        uint32_t w = f >= 8 ? 0x77777777u : 0x77777777u & ((1u << (4 * f)) - 1u);
        if (p >= 0 && p < 8) w |= rem << (4 * p);              // a pair straddling the end keeps its first nibble
        if (p == -1) w |= rem >> 4;                            // the second nibble of a pair from the dword before
        v[q] = (int)w;
    }
    return v;
}

__device__ __forceinline__ mcv_v16f mfma_f4(mcv_v4i a, mcv_v4i b, mcv_v16f c) {
    const mcv_v8i A = {a.x, a.y, a.z, a.w, 0, 0, 0, 0}, B = {b.x, b.y, b.z, b.w, 0, 0, 0, 0};
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B, c, 4, 4, 0, 127, 0, 127);
}

constexpr uint32_t kF4Sentinel = 0x44790000u;  // top 16 bits of 999.0f, index 0

template <bool kTail>
__device__ __forceinline__ void knn_f4_tile(const mcv_v4i (&a)[5], mcv_v4i ones, const mcv_v4i* sb, uint32_t j,
                                            bool colvalid, uint32_t (&k0)[16], uint32_t (&k1)[16]) {
    mcv_v4i b[5];
#pragma unroll
    for (int m = 0; m < 5; m++) b[m] = sb[64 * m];
    mcv_v16f acc = {};
#pragma unroll
    for (int m = 0; m < 4; m++) acc = mfma_f4(a[m], b[m], acc);
    acc = mfma_f4(a[4], ones, acc);  // + |q|
    acc = mfma_f4(ones, b[4], acc);  // + |t|
    // the accumulators read as one integer vector: element-wise bit casts of
    // the float vector were folded into acc[0] for every row by the compiler
    const mcv_v16i bits = __builtin_bit_cast(mcv_v16i, acc);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint32_t key = ((uint32_t)bits[r] & 0xFFFF0000u) | j;
        if (kTail) key = colvalid ? key : 0xFFFFFFFFu;
        top2_insert(k0[r], k1[r], key);
    }
}

// kDb: two super-tiles as B operands (4 bit steps, then |t|), the next one
// expanded between the current one's first and second blocks, one barrier
// per super-tile (40 KB of LDS); else one super-tile expanded between two
// barriers (20 KB).  kWaves: waves per SIMD the registers are cut for.
template <bool kDb, int kWaves>
__global__ __launch_bounds__(256, kWaves) void k_knn2_f4(const MatchPair* pairs, const int* order, int np) {
    __shared__ mcv_v4i s_b[kDb ? 2 : 1][kMSub][5][64];
    // with an order: a 1-D grid over the real 128-query tiles of the pairs
    // in that order (largest train sets first), the spare workgroups last;
    // without: (tile, pair) in index order
    int pi = (int)blockIdx.y, tile = (int)blockIdx.x;
    if (order) {
        const int b = (int)blockIdx.x;
        if (b >= order[2 * np]) return;
        int lo = 0, hi = np - 1;  // the last rank whose first tile is <= b
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (order[np + mid] <= b) lo = mid; else hi = mid - 1;
        }
        pi = order[lo];
        tile = b - order[np + lo];
    }
    const MatchPair P = pairs[pi];
    const int nq = P.nq_dev ? min(*P.nq_dev, P.nq) : P.nq;
    const int nt = P.nt_dev ? min(*P.nt_dev, P.nt) : P.nt;
    const int q0 = tile * 128;
    if (q0 >= nq) return;
    const int lane = threadIdx.x & 63, half = lane >> 5;
    const int wq0 = q0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 32;
    const mcv_v4i ones = {0x22222222, 0x22222222, 0x22222222, 0x22222222};
    // A operands: -2.0 per set bit of word 2 m + half in step m; |q| in step 4
    mcv_v4i a[5];
    {
        const uint4* qrow = reinterpret_cast<const uint4*>(P.q + (long long)min(wq0 + (lane & 31), nq - 1) * 32);
        const uint4 x = qrow[0], y = qrow[1];
        const uint32_t w[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
        uint32_t pc = 0;
#pragma unroll
        for (int m = 0; m < 8; m++) pc += __builtin_popcount(w[m]);
        const uint32_t hm = half ? 0xFFFFFFFFu : 0u;  // a blend, not an index (keeps w in registers)
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const uint32_t wm = (w[2 * m] & ~hm) | (w[2 * m + 1] & hm);
#pragma unroll
            for (int q = 0; q < 4; q++) a[m][q] = (int)(nibbles_of(wm, q) * 0xCu);
        }
        a[4] = count_nibbles(pc, half);
    }
    uint32_t k0[16], k1[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        k0[r] = kF4Sentinel;
        k1[r] = kF4Sentinel;
    }
    // thread tid expands words 4 (tid & 1) .. + 3 (steps 2 (tid & 1), + 1,
    // both halves) of train (tid >> 1) of a super-tile, and half (tid & 1)
    // of its |t| step
    const uint4* tv = reinterpret_cast<const uint4*>(P.t);
    auto fetch = [&](int st) -> uint4 {
        const int j = st * 32 * kMSub + (threadIdx.x >> 1);
        return j < nt ? tv[2ll * j + (threadIdx.x & 1)] : make_uint4(0u, 0u, 0u, 0u);
    };
    const int nsuper = (nt + 32 * kMSub - 1) / (32 * kMSub);
    const int jl = threadIdx.x >> 1, sub = jl >> 5, col = jl & 31, th = threadIdx.x & 1;
    auto expand = [&](const uint4 raw, int buf) {
        const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                mcv_v4i v;
#pragma unroll
                for (int q = 0; q < 4; q++) v[q] = (int)(nibbles_of(w[2 * k + h], q) << 1);
                s_b[buf][sub][2 * th + k][32 * h + col] = v;
            }
        uint32_t pc = __builtin_popcount(w[0]) + __builtin_popcount(w[1]) + __builtin_popcount(w[2]) +
                      __builtin_popcount(w[3]);
        pc += (uint32_t)__shfl_xor((int)pc, 1, 64);
        s_b[buf][sub][4][32 * th + col] = count_nibbles(pc, th);
    };
    uint4 raw = make_uint4(0u, 0u, 0u, 0u);
    if (nsuper > 0) {
        expand(fetch(0), 0);
        if (nsuper > 1) raw = fetch(1);
    }
    for (int st = 0; st < nsuper; st++) {
        if constexpr (!kDb) {
            if (st > 0) {
                __syncthreads();  // the previous super-tile's operands are read
                expand(raw, 0);
                if (st + 1 < nsuper) raw = fetch(st + 1);
            }
        }
        // (two buffers) super-tile st is in buffer st & 1; the other buffer
        // was last read in iteration st - 1, which every wave has finished here
        __syncthreads();
        const int cb = kDb ? (st & 1) : 0;
        const int tj = st * 32 * kMSub;
        if (tj + 32 * kMSub <= nt) {
            knn_f4_tile<false>(a, ones, &s_b[cb][0][0][lane], (uint32_t)(tj + (lane & 31)), true, k0, k1);
            if (kDb && st + 1 < nsuper) {
                expand(raw, cb ^ 1);
                if (st + 2 < nsuper) raw = fetch(st + 2);
            }
#pragma unroll 1
            for (int sb = 1; sb < kMSub; sb++)
                knn_f4_tile<false>(a, ones, &s_b[cb][sb][0][lane], (uint32_t)(tj + 32 * sb + (lane & 31)), true, k0,
                                   k1);
        } else {
            // a partial super-tile is the last one: nothing left to expand
#pragma unroll
            for (int sb = 0; sb < kMSub; sb++) {
                const int tj0 = tj + 32 * sb;
                if (tj0 < nt) {
                    const uint32_t j = (uint32_t)(tj0 + (lane & 31));
                    knn_f4_tile<true>(a, ones, &s_b[cb][sb][0][lane], j, (int)j < nt, k0, k1);
                }
            }
        }
    }
    // the 32 lanes of a half hold the same rows: merge their top-2 keys
#pragma unroll
    for (int r = 0; r < 16; r++) {
        uint32_t a0 = k0[r], a1 = k1[r];
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const uint32_t b0 = (uint32_t)__shfl_xor((int)a0, o, 64), b1 = (uint32_t)__shfl_xor((int)a1, o, 64);
            const uint32_t n1 = min(max(a0, b0), min(a1, b1));
            a0 = min(a0, b0);
            a1 = n1;
        }
        const int qi = wq0 + 8 * (r >> 2) + 4 * half + (r & 3);
        if ((lane & 31) == 0 && qi < nq) {
            // the sentinel keeps index 0 and distance 999, as the reference's d_idx init does
            auto dist = [](uint32_t k) {
                return k == kF4Sentinel ? 999 : (int)__builtin_bit_cast(float, k & 0xFFFF0000u);
            };
            int4 o;
            o.x = (int)(a0 & 0xFFFFu);
            o.y = dist(a0);
            o.z = (int)(a1 & 0xFFFFu);
            o.w = dist(a1);
            reinterpret_cast<int4*>(P.out)[qi] = o;
        }
    }
}

// Dispatch order of the pairs for the top-2 kernels: a workgroup's work is
// proportional to its pair's train count, so the pairs with the most trains
// go first (longest-processing-time first) and the launch does not end on a
// few large late workgroups.  One workgroup ranks up to kPairOrderMax pairs
// by (train count descending, index ascending): the keys are distinct, so a
// key's rank (the number of smaller keys, counted over LDS broadcasts) is
// its place.
constexpr int kPairOrderMax = 1024;
__global__ __launch_bounds__(256) void k_pair_order(const MatchPair* pairs, int np, int* order) {
    __shared__ unsigned long long key[kPairOrderMax];
    for (int t = threadIdx.x; t < np; t += 256) {
        const MatchPair P = pairs[t];
        const int nt = P.nt_dev ? min(*P.nt_dev, P.nt) : P.nt;
        key[t] = ((unsigned long long)(0xFFFFFFFFu - (uint32_t)max(nt, 0)) << 32) | (uint32_t)t;
    }
    __syncthreads();
    __shared__ int s_tiles[kPairOrderMax];
    for (int t = threadIdx.x; t < np; t += 256) {
        const unsigned long long k = key[t];
        int rank = 0;
        for (int j = 0; j < np; j++) rank += key[j] < k ? 1 : 0;
        order[rank] = t;
        const MatchPair P = pairs[t];
        const int nq = P.nq_dev ? min(*P.nq_dev, P.nq) : P.nq;
        s_tiles[rank] = (max(nq, 0) + 127) / 128;
    }
    __syncthreads();
    // order[np + r]: the first 128-query tile of the r-th pair in a 1-D grid
    // of the real tiles (order[2 np]: their total)
    if (threadIdx.x == 0) {
        int run = 0;
        for (int r = 0; r < np; r++) {
            order[np + r] = run;
            run += s_tiles[r];
        }
        order[2 * np] = run;
    }
}

// Variants of the same exact computation; the autotuner picks one per
// workload (kKnnVariants of them).
constexpr int kKnnVariants = 12;
inline const char* knn_variant_name(int v) {
    static const char* names[kKnnVariants] = {"S<1,2>", "S<1,4>", "S<2,2>", "S<2,4>", "L",   "S<2,8>",
                                              "S<1,8>", "S<4,2>", "M<1>",  "M<2>",  "F4",    "F4s5"};
    return (v >= 0 && v < kKnnVariants) ? names[v] : "?";
}

// max_nt: the largest train count of the pairs (the fp4 variant keys train
// indices in 16 bits and hands larger sets to M<1>); order: the output of
// k_pair_order (2 npairs + 1 ints), or null (used by the fp4 variant)
inline void launch_knn2(int variant, const MatchPair* pairs, int npairs, int max_nq, hipStream_t s, int max_nt = -1,
                        const int* order = nullptr) {
    if (max_nt < 0) max_nt = max_nq;
    if ((variant == 10 || variant == 11) && max_nt >= 65536) variant = 8;
    switch (variant) {
    case 10:
    case 11: {
        // with an order, a 1-D grid of the real tiles first
        const dim3 grid = order ? dim3((max_nq + 127) / 128 * npairs) : dim3((max_nq + 127) / 128, npairs);
        if (variant == 10)
            hipLaunchKernelGGL((k_knn2_f4<true, 4>), grid, dim3(256), 0, s, pairs, order, npairs);
        else
            hipLaunchKernelGGL((k_knn2_f4<false, 5>), grid, dim3(256), 0, s, pairs, order, npairs);
        break;
    }
    case 1: hipLaunchKernelGGL((k_knn2_s<1, 4>), dim3((max_nq + 63) / 64, npairs), dim3(256), 0, s, pairs); break;
    case 2: hipLaunchKernelGGL((k_knn2_s<2, 2>), dim3((max_nq + 127) / 128, npairs), dim3(256), 0, s, pairs); break;
    case 3: hipLaunchKernelGGL((k_knn2_s<2, 4>), dim3((max_nq + 127) / 128, npairs), dim3(256), 0, s, pairs); break;
    case 4: hipLaunchKernelGGL(k_knn2_lds, dim3((max_nq + kKnnQ - 1) / kKnnQ, npairs), dim3(256), 0, s, pairs); break;
    case 5: hipLaunchKernelGGL((k_knn2_s<2, 8>), dim3((max_nq + 127) / 128, npairs), dim3(256), 0, s, pairs); break;
    case 6: hipLaunchKernelGGL((k_knn2_s<1, 8>), dim3((max_nq + 63) / 64, npairs), dim3(256), 0, s, pairs); break;
    case 7: hipLaunchKernelGGL((k_knn2_s<4, 2>), dim3((max_nq + 255) / 256, npairs), dim3(256), 0, s, pairs); break;
    case 8: hipLaunchKernelGGL((k_knn2_m<1>), dim3((max_nq + 127) / 128, npairs), dim3(256), 0, s, pairs); break;
    case 9: hipLaunchKernelGGL((k_knn2_m<2>), dim3((max_nq + 255) / 256, npairs), dim3(256), 0, s, pairs); break;
    default: hipLaunchKernelGGL((k_knn2_s<1, 2>), dim3((max_nq + 63) / 64, npairs), dim3(256), 0, s, pairs); break;
    }
}

// ---------------------------------------------------------------------------
// MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cpp:101-150), one
// wave per map point with N <= kDistinctMax observed descriptors (staged in
// LDS).  Lane i owns rows i, i + 64, ...; the row median (the value at
// index floor(0.5 (N - 1)) of the sorted row, diagonal 0 included) is the
// smallest v with #{j : d_ij <= v} > floor(0.5 (N - 1)), found by binary
// search over v in [0, 256] without sorting.  The first row with the
// smallest median wins: key (median << 16 | row), wave minimum.
constexpr int kDistinctMax = 256;

__global__ __launch_bounds__(64) void k_distinctive(const uint8_t* desc, const int* ptr, int n, int* best,
                                                    uint8_t* out_desc) {
    __shared__ uint4 s_d[2 * kDistinctMax];
    const int m = blockIdx.x;
    if (m >= n) return;
    const int lane = threadIdx.x;
    const int b = ptr[m], N = ptr[m + 1] - b;
    if (N <= 0) {
        if (lane == 0) best[m] = -1;
        return;
    }
    const uint4* src = reinterpret_cast<const uint4*>(desc + 32ll * b);
    for (int k = lane; k < 2 * N; k += 64) s_d[k] = src[k];
    wave_lds_sync();
    const int kth = (N - 1) / 2;  // (size_t)(0.5 * (N - 1))
    uint32_t key = 0xFFFFFFFFu;
    for (int i = lane; i < N; i += 64) {
        const uint4 a0 = s_d[2 * i], a1 = s_d[2 * i + 1];
        int lo = 0, hi = 256;  // answer in [lo, hi]
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            int cnt = 0;
            for (int j = 0; j < N; j++) {
                const uint4 c0 = s_d[2 * j], c1 = s_d[2 * j + 1];
                const int d = __builtin_popcount(a0.x ^ c0.x) + __builtin_popcount(a0.y ^ c0.y) +
                              __builtin_popcount(a0.z ^ c0.z) + __builtin_popcount(a0.w ^ c0.w) +
                              __builtin_popcount(a1.x ^ c1.x) + __builtin_popcount(a1.y ^ c1.y) +
                              __builtin_popcount(a1.z ^ c1.z) + __builtin_popcount(a1.w ^ c1.w);
                cnt += d <= mid ? 1 : 0;
            }
            if (cnt > kth) hi = mid; else lo = mid + 1;
        }
        key = min(key, ((uint32_t)lo << 16) | (uint32_t)i);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) key = min(key, (uint32_t)__shfl_xor((int)key, o, 64));
    const int bi = (int)(key & 0xFFFFu);
    if (lane == 0) best[m] = bi;
    if (out_desc && lane < 2)
        reinterpret_cast<uint4*>(out_desc + 32ll * m)[lane] = s_d[2 * bi + lane];
}

// ---------------------------------------------------------------------------
// MatchRes::FilterOrientation (src/Matcher.cpp:44-75), one workgroup per
// match list (n <= kOrientMax).  Each match's rotation bin is computed with
// the reference's float expressions; keys (bin << 16 | position) sorted in
// LDS group every bin's matches in list order; the 40 (size, bin) records
// are ordered by one thread with the libstdc++ std::sort replica (size
// descending, the reference's unstable tie order), and the first three bins
// are copied out.
constexpr int kOrientMax = 4096;
constexpr int kHistoLength = 40;

struct OrientJob {
    const int32_t* matches;  // n x (query, train, distance)
    int n;
    const float* angle1;     // by query index
    const float* angle2;     // by train index
    int32_t* out;            // up to n x 3
    int* count;
};

struct OrientLess {
    __device__ bool operator()(uint32_t a, uint32_t b) const { return (a >> 8) > (b >> 8); }
};

__global__ __launch_bounds__(256) void k_filter_orientation(const OrientJob* jobs) {
    __shared__ uint32_t keys[kOrientMax];
    __shared__ int cnt[kHistoLength];
    __shared__ int start[kHistoLength];
    __shared__ uint32_t rec[kHistoLength];
    __shared__ int sort_stack[3 * 16];  // stl_sort's pending ranges: 2 floor(log2 40) + 2 = 12 at most
    const OrientJob J = jobs[blockIdx.x];
    const int n = J.n;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    if (threadIdx.x < kHistoLength) cnt[threadIdx.x] = 0;
    __syncthreads();
    const float factor = 1.0f / (360.0f / kHistoLength);
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        uint32_t k = 0xFFFFFFFFu;
        if (i < n) {
            float rot = J.angle1[J.matches[3 * i]] - J.angle2[J.matches[3 * i + 1]];
            if (rot < 0.0f) rot += 360.0f;
            int bin = (int)roundf(rot * factor);
            if (bin == kHistoLength) bin = 0;
            bin = min(max(bin, 0), kHistoLength - 1);
            k = ((uint32_t)bin << 16) | (uint32_t)i;
            atomicAdd(&cnt[bin], 1);
        }
        keys[i] = k;
    }
    __syncthreads();
    if (n2 > 1) bitonic_sort(keys, n2);
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b < kHistoLength; b++) {
            start[b] = acc;
            acc += cnt[b];
            rec[b] = ((uint32_t)cnt[b] << 8) | (uint32_t)b;
        }
        stl_sort(rec, rec + kHistoLength, OrientLess(), sort_stack);
    }
    __syncthreads();
    int base = 0;
    for (int r = 0; r < 3; r++) {
        const int b = (int)(rec[r] & 0xFF), c = (int)(rec[r] >> 8);
        for (int k = threadIdx.x; k < c; k += blockDim.x) {
            const int i = (int)(keys[start[b] + k] & 0xFFFFu);
            int32_t* o = J.out + 3 * (base + k);
            o[0] = J.matches[3 * i];
            o[1] = J.matches[3 * i + 1];
            o[2] = J.matches[3 * i + 2];
        }
        base += c;
    }
    if (threadIdx.x == 0) *J.count = base;
}

struct FilterJob {
    const int32_t* knn;   // nq x 4
    const int* nq_dev;    // or null
    int nq;
    float ratio;
    int th;
    int32_t* out;         // up to nq x (query, train, distance)
    int* count;           // filtered length
    int* scratch;         // 2 * nq ints
};

__device__ __forceinline__ int block_scan_excl(int v, int* ws, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[wid] = x;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < nw; w++) {
        const int s = ws[w];
        base += (w < wid) ? s : 0;
        total += s;
    }
    __syncthreads();
    return base + x - v;
}

// One workgroup per knn list.  FilterRatio keeps m0 when d0/d1 <= ratio (in
// float; 999 sentinels behave as in the reference, 0/0 is rejected) and keeps
// list order.  FilterThreshold's swap-with-last removal leaves, for k = number
// of survivors: positions p < k that survive keep their element, and the
// m-th non-survivor position below k receives the m-th survivor counted from
// the end of the list among positions >= k.
__device__ __forceinline__ bool passes_th(const FilterJob& J, int q) {
    return !((float)J.knn[4 * q + 1] > (float)J.th);
}

// One list per workgroup of 1024 threads.  J.scratch (2 * nq ints) needs no
// clearing: every entry read is written first.
__device__ void filter_list(const FilterJob& J) {
    __shared__ int ws[16];
    const int nq = J.nq_dev ? min(*J.nq_dev, J.nq) : J.nq;
    int* list = J.scratch;        // ratio survivors (query indices), in order
    int* holes = J.scratch + nq;  // non-survivor positions below k, in order
    int n = 0;
    for (int i0 = 0; i0 < nq; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        int keep = 0;
        if (i < nq) {
            const int4 m = reinterpret_cast<const int4*>(J.knn)[i];
            keep = ((float)m.y / (float)m.w <= J.ratio) ? 1 : 0;
        }
        int tot;
        const int pos = block_scan_excl(keep, ws, tot);
        if (keep) list[n + pos] = i;
        n += tot;
    }
    __syncthreads();
    int k = 0;
    for (int i0 = 0; i0 < n; i0 += blockDim.x) {
        const int p = i0 + threadIdx.x;
        const int good = (p < n && passes_th(J, list[p])) ? 1 : 0;
        int tot;
        block_scan_excl(good, ws, tot);
        k += tot;
    }
    int nb = 0;
    for (int i0 = 0; i0 < k; i0 += blockDim.x) {
        const int p = i0 + threadIdx.x;
        int bad = 0;
        if (p < k) {
            const int q = list[p];
            if (passes_th(J, q)) {
                int* o = J.out + 3 * p;
                o[0] = q;
                o[1] = J.knn[4 * q];
                o[2] = J.knn[4 * q + 1];
            } else {
                bad = 1;
            }
        }
        int tot;
        const int r = block_scan_excl(bad, ws, tot);
        if (bad) holes[nb + r] = p;
        nb += tot;
    }
    __syncthreads();
    int m = 0;
    for (int r0 = 0; r0 < n - k; r0 += blockDim.x) {
        const int r = r0 + threadIdx.x;   // r-th position from the end
        const int p = n - 1 - r;
        int good = 0, q = -1;
        if (r < n - k) {
            q = list[p];
            good = passes_th(J, q) ? 1 : 0;
        }
        int tot;
        const int idx = m + block_scan_excl(good, ws, tot);
        if (good) {
            int* o = J.out + 3 * holes[idx];
            o[0] = q;
            o[1] = J.knn[4 * q];
            o[2] = J.knn[4 * q + 1];
        }
        m += tot;
    }
    if (threadIdx.x == 0) *J.count = k;
}

__global__ __launch_bounds__(1024) void k_filter(const FilterJob* jobs) { filter_list(jobs[blockIdx.x]); }

// the same for one list whose job comes as a kernel argument (nothing to
// upload, so an enqueue allocates nothing)
__global__ __launch_bounds__(1024) void k_filter_one(FilterJob J) { filter_list(J); }

}  // namespace mcv
