// Projection-guided window search for gfx950.
//
//   k_grid_build   Object::AssignFeaturesToGrid / PosInGrid (src/Object.cpp:
//                  182-201, :249-257): 30 x 30 cells, cell = round(x * 30/cols),
//                  keypoints outside [0,30) dropped, in-cell order = index.
//   k_window_knn   Object::ProjectBunchMapPoints (:208-236) per map point:
//                  R*Pw + t, z < 0 skip, pinhole project, GetFeaturesInArea
//                  (:263-308) in ix-major / iy / in-cell order, top-2 Hamming
//                  in that candidate order, FilterRatio + FilterThreshold.
//                  With direct positions instead of map points it is the
//                  query of Tracker::Wnd_Track (src/Tracker.cpp:339-353):
//                  GetFeaturesInArea(kp.pt, r, minLevel, maxLevel) then the
//                  same top-2 and filters.
// Included by mcv_api.hip (shares bitonic_sort with the ORB kernels).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcv {

constexpr int kGrid = 30;
constexpr int kGridCells = kGrid * kGrid;
constexpr int kGridSortCap = 16384;

struct GridView {
    const float* kxy;       // n x 2
    const int32_t* koct;    // n octaves (read only by a level-filtered query)
    const uint8_t* kdesc;   // n x 32
    int n;
    int width, height;
    float inv_w, inv_h;     // (float)(30.0 / cols), (float)(30.0 / rows)
    int* cell_start;        // 901 entries
    int* cell_items;        // n entries (keypoint indices grouped by cell)
};

__global__ __launch_bounds__(1024) void k_grid_build(const GridView* views) {
    __shared__ uint32_t keys[kGridSortCap];
    __shared__ int cnt[kGridCells + 1];
    const GridView G = views[blockIdx.x];
    int n2 = 1;
    while (n2 < G.n) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        uint32_t k = 0xFFFFFFFFu;
        if (i < G.n) {
            const float x = G.kxy[2 * i], y = G.kxy[2 * i + 1];
            const int px = (int)roundf((x - 0.f) * G.inv_w);
            const int py = (int)roundf((y - 0.f) * G.inv_h);
            if (px >= 0 && px < kGrid && py >= 0 && py < kGrid) k = ((uint32_t)(px * kGrid + py) << 16) | (uint32_t)i;
        }
        keys[i] = k;
    }
    for (int c = threadIdx.x; c <= kGridCells; c += blockDim.x) cnt[c] = 0;
    __syncthreads();
    bitonic_sort(keys, n2);
    for (int i = threadIdx.x; i < G.n; i += blockDim.x) {
        const uint32_t k = keys[i];
        if (k != 0xFFFFFFFFu) {
            G.cell_items[i] = (int)(k & 0xFFFF);
            atomicAdd(&cnt[k >> 16], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int c = 0; c < kGridCells; c++) {
            G.cell_start[c] = acc;
            acc += cnt[c];
        }
        G.cell_start[kGridCells] = acc;
    }
}

struct WindowCam {
    GridView grid;
    float R[9], t[3];
    float fx, fy, cx, cy;
    int32_t* out;           // n_points x 6
};

struct WindowPoints {
    const float* pts;       // P x 3 map points, or null for direct positions
    const float* uv;        // P x 2 query positions (when pts is null)
    const uint8_t* pdesc;   // P x 32
    const int32_t* plevel;  // P (map points only)
    int P;
    float radius, ratio;
    int th;
    int nl;                 // scale-table entries (map points only)
    int min_level, max_level;  // GetFeaturesInArea's level filter (direct positions only)
    float scales[16];
};

// One wave per (camera, query).  Output row: (kp0, d0, kp1, d1, n_candidates,
// match); a map point whose level is outside [0, nl) gets n_candidates = -1
// and no match (the reference indexes mvScaleFactor with it).
__global__ __launch_bounds__(256) void k_window_knn(const WindowCam* cams, WindowPoints W) {
    const int wave = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p = wave;
    if (p >= W.P) return;
    const WindowCam& C = cams[blockIdx.y];
    const GridView& G = C.grid;
    int32_t* o = C.out + 6 * (long long)p;
    float u = 0.f, v = 0.f, r = 0.f;
    bool valid;
    if (W.pts) {
        // Pc = R * Pw + t: cv::Mat float gemm, products and sums in double,
        // one rounding to float per row
        const float* pw = W.pts + 3 * p;
        float Pc[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
            s += (double)C.R[3 * k] * (double)pw[0];
            s += (double)C.R[3 * k + 1] * (double)pw[1];
            s += (double)C.R[3 * k + 2] * (double)pw[2];
            Pc[k] = (float)(s + (double)C.t[k]);
        }
        const int lvl = W.plevel[p];
        if (lvl < 0 || lvl >= W.nl) {
            if (lane == 0) {
                o[0] = -1; o[1] = 999; o[2] = -1; o[3] = 999; o[4] = -1; o[5] = -1;
            }
            return;
        }
        valid = !(Pc[2] < 0.f);
        if (valid) {
            u = C.fx * Pc[0] / Pc[2] + C.cx;
            v = C.fy * Pc[1] / Pc[2] + C.cy;
            r = W.radius * W.scales[lvl];
        }
    } else {
        u = W.uv[2 * p];
        v = W.uv[2 * p + 1];
        r = W.radius;
        valid = true;
    }
    const bool check_levels = W.pts == nullptr && (W.min_level > 0 || W.max_level >= 0);
    uint32_t k0 = 999u << 22, k1 = 999u << 22;
    int ncand = 0;
    int nMinX = 0, nMaxX = -1, nMinY = 0, nMaxY = -1;
    if (valid && !(u != u || v != v) && !(u < 0 || v < 0 || u >= (float)G.width || v >= (float)G.height)) {
        nMinX = max(0, (int)floorf((u - 0.f - r) * G.inv_w));
        nMaxX = min(kGrid - 1, (int)ceilf((u - 0.f + r) * G.inv_w));
        nMinY = max(0, (int)floorf((v - 0.f - r) * G.inv_h));
        nMaxY = min(kGrid - 1, (int)ceilf((v - 0.f + r) * G.inv_h));
        if (!(nMinX < kGrid && nMaxX >= 0 && nMinY < kGrid && nMaxY >= 0)) nMaxX = -1;
    }
    if (nMaxX >= nMinX && nMaxY >= nMinY) {
        uint32_t q[8];
        const uint32_t* qd = reinterpret_cast<const uint32_t*>(W.pdesc + 32 * (long long)p);
#pragma unroll
        for (int k = 0; k < 8; k++) q[k] = qd[k];
        // enumerate cells ix-major, then iy; candidate order index c
        int base = 0;
        const int ny = nMaxY - nMinY + 1;
        const int ncells = (nMaxX - nMinX + 1) * ny;
        for (int cc = 0; cc < ncells; cc++) {
            const int ix = nMinX + cc / ny, iy = nMinY + cc % ny;
            const int cell = ix * kGrid + iy;
            const int s0 = G.cell_start[cell], s1 = G.cell_start[cell + 1];
            for (int s = s0 + lane; s < s1; s += 64) {
                const int idx = G.cell_items[s];
                const float dx = G.kxy[2 * idx] - u, dy = G.kxy[2 * idx + 1] - v;
                bool in = fabsf(dx) < r && fabsf(dy) < r;
                if (check_levels) {
                    const int oct = G.koct[idx];
                    if (oct < W.min_level || (W.max_level >= 0 && oct > W.max_level)) in = false;
                }
                if (in) {
                    const uint32_t* kd = reinterpret_cast<const uint32_t*>(G.kdesc + 32 * (long long)idx);
                    uint32_t d = 0;
#pragma unroll
                    for (int k = 0; k < 8; k++) d += __builtin_popcount(q[k] ^ kd[k]);
                    const uint32_t key = (d << 22) | (uint32_t)(base + (s - s0));
                    k1 = min(k1, max(k0, key));
                    k0 = min(k0, key);
                    ncand++;
                }
            }
            base += s1 - s0;
        }
    }
    // merge the lanes' top-2 keys and candidate counts
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t b0 = __shfl_xor(k0, off, 64), b1 = __shfl_xor(k1, off, 64);
        const uint32_t n1 = min(max(k0, b0), min(k1, b1));
        k0 = min(k0, b0);
        k1 = n1;
        ncand += __shfl_xor(ncand, off, 64);
    }
    if (lane != 0) return;
    // candidate order index -> keypoint index (the same cell walk)
    auto kp_of = [&](uint32_t key) -> int {
        if ((key >> 22) == 999u) return -1;
        const int c = (int)(key & ((1u << 22) - 1));
        const int ny = nMaxY - nMinY + 1;
        int base = 0;
        for (int cc = 0; cc < (nMaxX - nMinX + 1) * ny; cc++) {
            const int cell = (nMinX + cc / ny) * kGrid + nMinY + cc % ny;
            const int s0 = G.cell_start[cell], s1 = G.cell_start[cell + 1];
            if (c < base + (s1 - s0)) return G.cell_items[s0 + (c - base)];
            base += s1 - s0;
        }
        return -1;
    };
    const int d0 = (int)(k0 >> 22), d1 = (int)(k1 >> 22);
    const int kp0 = kp_of(k0), kp1 = kp_of(k1);
    int match = -1;
    if (ncand > 0 && (float)d0 / (float)d1 <= W.ratio && !((float)d0 > (float)W.th)) match = kp0;
    o[0] = kp0;
    o[1] = d0;
    o[2] = kp1;
    o[3] = d1;
    o[4] = ncand;
    o[5] = match;
}

}  // namespace mcv
