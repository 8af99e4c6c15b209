// Bag-of-words kernels for gfx950 (DBoW3, vendored by the reference under
// modules/DBow3, and Matcher::DBowMatch).
//
//   k_bow_transform  Vocabulary::transform(feature, word_id, weight, nid,
//                    levelsup) (modules/DBow3/src/Vocabulary.cpp:641-672)
//                    for a batch of descriptors: from the root, the child at
//                    the least Hamming distance (DescManip::distance,
//                    DescManip.cpp:92-119), the first of equal distances, to
//                    a leaf; the node reached at level L - levelsup.
//   k_bow_match      Matcher::DBowMatch (src/Matcher.cpp:146-193): per
//                    feature of image 1 in a vocabulary node both images
//                    hold, the two nearest features of image 2 in that node
//                    (strict '<' in list order).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcv {

struct BowVocab {
    const uint32_t* desc;     // n x 8 words (node descriptors, row 0 unused)
    const int32_t* child_ptr; // n + 1
    const int32_t* child_ids;
    const int32_t* word_id;   // -1 for inner nodes
    const double* weight;
    int depth;                // longest root-leaf path
};

// A group of G lanes per descriptor: lane g of the group takes children g,
// g + G, ... of the current node; the key (distance << 16 | child position)
// keeps the first of equal distances, and the group's minimum picks the
// child.  Every lane runs `depth` steps (a group past its leaf idles), so the
// group shuffles always see their whole group.
template <int G>
__global__ __launch_bounds__(256) void k_bow_transform(BowVocab V, const uint32_t* feats, int n, int nid_level,
                                                        int32_t* out_word, double* out_w, int32_t* out_nid) {
    const int gl = threadIdx.x % G;
    const int f = (int)((blockIdx.x * blockDim.x + threadIdx.x) / G);
    const int fc = f < n ? f : n - 1;
    uint32_t q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = feats[8 * (long long)fc + k];
    int node = 0, nid = 0;
    bool done = V.child_ptr[1] == V.child_ptr[0];
    for (int level = 1; level <= V.depth; level++) {
        uint32_t key = 0xFFFFFFFFu;
        int c0 = 0, nch = 0;
        if (!done) {
            c0 = V.child_ptr[node];
            nch = V.child_ptr[node + 1] - c0;
            for (int c = gl; c < nch; c += G) {
                const uint32_t* d = V.desc + 8 * (long long)V.child_ids[c0 + c];
                uint32_t dist = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) dist += __builtin_popcount(q[k] ^ d[k]);
                key = min(key, (dist << 16) | (uint32_t)c);
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) key = min(key, (uint32_t)__shfl_xor((int)key, o, G));
        if (!done) {
            node = V.child_ids[c0 + (int)(key & 0xFFFFu)];
            if (level == nid_level) nid = node;
            done = V.child_ptr[node + 1] == V.child_ptr[node];
        }
    }
    if (f < n && gl == 0) {
        out_word[f] = V.word_id[node];
        out_w[f] = V.weight[node];
        out_nid[f] = nid;
    }
}

// One query of DBowMatch: feature q of image 1 against the features
// feats2[t0 .. t1) of image 2 (one vocabulary node, at least two of them);
// output row (q, train0, d0, train1, d1).
struct BowJob {
    int q, t0, t1, row;
};

__global__ __launch_bounds__(256) void k_bow_match(const BowJob* jobs, int njobs, const uint32_t* d1,
                                                   const uint32_t* d2, const int32_t* feats2, int32_t* out) {
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= njobs) return;
    const BowJob J = jobs[wave];
    uint32_t q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = d1[8 * (long long)J.q + k];
    uint32_t k0 = 999u << 22, k1 = 999u << 22;
    for (int t = J.t0 + lane; t < J.t1; t += 64) {
        const uint32_t* d = d2 + 8 * (long long)feats2[t];
        uint32_t dist = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) dist += __builtin_popcount(q[k] ^ d[k]);
        const uint32_t key = (dist << 22) | (uint32_t)(t - J.t0);
        k1 = min(k1, max(k0, key));
        k0 = min(k0, key);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t b0 = __shfl_xor(k0, off, 64), b1 = __shfl_xor(k1, off, 64);
        const uint32_t n1 = min(max(k0, b0), min(k1, b1));
        k0 = min(k0, b0);
        k1 = n1;
    }
    if (lane == 0) {
        int32_t* o = out + 5 * (long long)J.row;
        o[0] = J.q;
        o[1] = feats2[J.t0 + (int)(k0 & ((1u << 22) - 1))];
        o[2] = (int)(k0 >> 22);
        o[3] = feats2[J.t0 + (int)(k1 & ((1u << 22) - 1))];
        o[4] = (int)(k1 >> 22);
    }
}

}  // namespace mcv
