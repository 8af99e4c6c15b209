// Hamming matching kernels for gfx950.
//
//   k_knn2        Matcher::KnnMatch(vector<Mat>, vector<Mat>) top-2
//                 (src/Matcher.cpp:245-302, include/Matcher.hpp:19-33)
//   k_filter      MatchResKnn::FilterRatio + MatchRes::FilterThreshold
//                 (src/Matcher.cpp:100-111, :23-35)
//
// The reference keeps d = {999, 999}, idx = {0, 0} and replaces with strict
// '<' while scanning trains in order, which is the first two entries of a
// stable sort by distance.  Here each candidate becomes the key
// (distance << 22) | train_index; the two smallest keys are that same pair,
// and (999 << 22) | 0 is the reference's sentinel.  Keys merge exactly across
// waves, so the train set is split over the 4 waves of a workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

__device__ __forceinline__ void top2_insert(uint32_t& k0, uint32_t& k1, uint32_t key) {
    k1 = min(k1, max(k0, key));
    k0 = min(k0, key);
}

// 256 threads: lane = query, wave = quarter of the train set.  Train rows are
// wave-uniform and read through scalar loads; the query row stays in VGPRs.
__global__ __launch_bounds__(256) void k_knn2(const MatchPair* pairs) {
    __shared__ uint32_t sk[2][4][64];
    const MatchPair P = pairs[blockIdx.y];
    const int nq = P.nq_dev ? min(*P.nq_dev, P.nq) : P.nq;
    const int nt = P.nt_dev ? min(*P.nt_dev, P.nt) : P.nt;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int qi = blockIdx.x * 64 + lane;
    if (blockIdx.x * 64 >= nq) return;
    uint32_t q[8];
    const uint32_t* qrow = reinterpret_cast<const uint32_t*>(P.q + (long long)min(qi, nq - 1) * 32);
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = qrow[k];
    const int per = (nt + 3) / 4;
    const int j0 = __builtin_amdgcn_readfirstlane(wave * per);
    const int j1 = __builtin_amdgcn_readfirstlane(min(nt, j0 + per));
    uint32_t k0 = kSentinel, k1 = kSentinel;
    const uint32_t* tw = reinterpret_cast<const uint32_t*>(P.t);
    int j = j0;
    for (; j + 1 < j1; j += 2) {
        const uint32_t* ta = tw + (long long)j * 8;
        uint32_t da = 0, db = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            da += __builtin_popcount(q[k] ^ ta[k]);
            db += __builtin_popcount(q[k] ^ ta[8 + k]);
        }
        top2_insert(k0, k1, (da << kIdxBits) | (uint32_t)j);
        top2_insert(k0, k1, (db << kIdxBits) | (uint32_t)(j + 1));
    }
    if (j < j1) {
        const uint32_t* ta = tw + (long long)j * 8;
        uint32_t da = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) da += __builtin_popcount(q[k] ^ ta[k]);
        top2_insert(k0, k1, (da << kIdxBits) | (uint32_t)j);
    }
    sk[0][wave][lane] = k0;
    sk[1][wave][lane] = k1;
    __syncthreads();
    if (wave == 0 && qi < nq) {
#pragma unroll
        for (int w = 1; w < 4; w++) {
            const uint32_t b0 = sk[0][w][lane], b1 = sk[1][w][lane];
            const uint32_t n1 = min(max(k0, b0), min(k1, b1));
            k0 = min(k0, b0);
            k1 = n1;
        }
        int4 o;
        o.x = (int)(k0 & kIdxMask);
        o.y = (int)(k0 >> kIdxBits);
        o.z = (int)(k1 & kIdxMask);
        o.w = (int)(k1 >> kIdxBits);
        // the sentinel keeps index 0, as the reference's d_idx init does
        reinterpret_cast<int4*>(P.out)[qi] = o;
    }
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

__global__ __launch_bounds__(1024) void k_filter(const FilterJob* jobs) {
    __shared__ int ws[16];
    const FilterJob J = jobs[blockIdx.x];
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

}  // namespace mcv
