// ORB extraction kernels for gfx950 (ORBextractor::operator(),
// modules/local_feature/ORB/orb3_extractor/ORBextractor.cc:831-899).
//
// One plan covers a batch of images of any sizes; each kernel runs over the
// whole batch through flattened work tables (tiles, cells, levels, keypoint
// slots), so a batch of rig frames is the same launch sequence as one image.
//
//   k_copy_level0   input image -> pyramid level 0
//   k_resize        level l-1 -> level l, cv::resize INTER_LINEAR 8U     (ORBextractor.cc:911)
//   k_blur          GaussianBlur 7x7 sigma 2 REFLECT_101, fixed point   (:875)
//   k_fast_cells    per-cell FAST 9/16 + windowed 3x3 NMS + iniTh/minTh
//                   fallback + row-major compaction                     (:596-633)
//   k_octree_build  quadtree path codes, sort, node tables               (:524-580)
//   k_octree_heap   libstdc++ priority_queue replay, pop order           (:530-578)
//   k_keypoints     node argmax, IC angle, rBRIEF, coordinate scaling    (:75-141, :640-653, :845-896)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcv_device.hpp"
#include "mcv_plan.hpp"
#include "orb_pattern.h"

namespace mcv {

struct Tile {
    int lvl;     // LevelDesc index
    int x0, y0;  // tile origin (64 x 16 pixels)
};
constexpr int TW = 64, TH = 16;

struct BlurTaps {
    int t[7];
};

struct SinCosTables {
    const uint32_t* cos_key;  // sorted float bit patterns where the base result differs
    const float* cos_val;     // host libm cosf at those inputs
    const uint32_t* cos_idx;  // 65537 range starts indexed by (bits >> 16)
    const uint32_t* sin_key;
    const float* sin_val;
    const uint32_t* sin_idx;
    uint32_t max_bits;        // inputs with bits <= max_bits are covered
};

struct OrbGlobals {
    const LevelDesc* lv;
    const CellDesc* cells;
    uint8_t* pyr;
    uint8_t* blur;
    uint32_t* cell_cand;    // per-cell candidate slots
    int* cell_n;            // candidates kept per cell
    uint32_t* lvl_cand;     // per-level contiguous candidates
    OctLevel* oct;
    uint32_t* pool;         // node tables and sorted indices (u32 units)
    unsigned long long* pool_used;
    long long pool_cap;     // u32 units
    uint32_t* nodes_out;    // per keypoint slot: node id in heap-pop order
    const float* seeds;     // rows of (x, y, size, angle, response, octave)
    int umax[16];
    int ini_th, min_th;
    int* err;               // bit 0: pool overflow, bit 1: level too large
};

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy_level0(const Tile* tiles, const LevelDesc* lv,
                                                     const uint8_t* const* src, const int* src_stride, uint8_t* pyr) {
    const Tile t = tiles[blockIdx.x];
    const LevelDesc L = lv[t.lvl];
    const uint8_t* s = src[L.img];
    const int stride = src_stride[L.img];
    for (int i = threadIdx.x; i < TW * TH; i += 256) {
        const int x = t.x0 + (i & (TW - 1)), y = t.y0 + i / TW;
        if (x < L.w && y < L.h) pyr[L.pix + (long long)y * L.w + x] = s[(long long)y * stride + x];
    }
}

// cv::resize INTER_LINEAR, CV_8U: horizontal int taps (11-bit), vertical
// ((b0*(r0>>4))>>16) + ((b1*(r1>>4))>>16) + 2) >> 2.
__global__ __launch_bounds__(256) void k_resize(const Tile* tiles, const LevelDesc* lv, const int2* coef,
                                                uint8_t* pyr) {
    const Tile t = tiles[blockIdx.x];
    const LevelDesc L = lv[t.lvl];
    const LevelDesc S = lv[L.src];
    const uint8_t* sp = pyr + S.pix;
    for (int i = threadIdx.x; i < TW * TH; i += 256) {
        const int x = t.x0 + (i & (TW - 1)), y = t.y0 + i / TW;
        if (x >= L.w || y >= L.h) continue;
        const int2 cx = coef[L.coef_x + x];
        const int2 cy = coef[L.coef_y + y];
        const int a0 = cx.y & 0xFFFF, a1 = cx.y >> 16;
        const int b0 = cy.y & 0xFFFF, b1 = cy.y >> 16;
        const int sx = cx.x, sx1 = min(cx.x + 1, S.w - 1);
        const int y0 = min(max(cy.x, 0), S.h - 1), y1 = min(max(cy.x + 1, 0), S.h - 1);
        const uint8_t* r0p = sp + (long long)y0 * S.w;
        const uint8_t* r1p = sp + (long long)y1 * S.w;
        const int r0 = r0p[sx] * a0 + r0p[sx1] * a1;
        const int r1 = r1p[sx] * a0 + r1p[sx1] * a1;
        pyr[L.pix + (long long)y * L.w + x] = (uint8_t)((((b0 * (r0 >> 4)) >> 16) + ((b1 * (r1 >> 4)) >> 16) + 2) >> 2);
    }
}

__device__ __forceinline__ int reflect101(int i, int n) {
    i = i < 0 ? -i : i;
    i = i >= n ? 2 * n - 2 - i : i;
    return i;
}

// GaussianBlur 7x7 on CV_8U through the fixed-point path: Q8 taps,
// h = sum t*src (<= 65280), out = (sum t*h + 2^15) >> 16.
__global__ __launch_bounds__(256) void k_blur(const Tile* tiles, const LevelDesc* lv, const uint8_t* pyr,
                                              uint8_t* blur, BlurTaps taps) {
    __shared__ uint8_t in[(TH + 6) * (TW + 6)];
    __shared__ uint32_t hb[(TH + 6) * TW];
    const Tile t = tiles[blockIdx.x];
    const LevelDesc L = lv[t.lvl];
    const uint8_t* src = pyr + L.pix;
    for (int i = threadIdx.x; i < (TH + 6) * (TW + 6); i += 256) {
        const int ly = i / (TW + 6), lx = i - ly * (TW + 6);
        const int gx = reflect101(min(t.x0 + lx - 3, L.w + 2), L.w);
        const int gy = reflect101(min(t.y0 + ly - 3, L.h + 2), L.h);
        in[i] = src[(long long)gy * L.w + gx];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (TH + 6) * TW; i += 256) {
        const int ly = i / TW, lx = i - ly * TW;
        const uint8_t* r = in + ly * (TW + 6) + lx;
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 7; k++) acc += (uint32_t)taps.t[k] * r[k];
        hb[i] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TH * TW; i += 256) {
        const int ly = i / TW, lx = i - ly * TW;
        const int x = t.x0 + lx, y = t.y0 + ly;
        if (x >= L.w || y >= L.h) continue;
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 7; k++) acc += (uint32_t)taps.t[k] * hb[(ly + k) * TW + lx];
        const uint32_t v = (acc + (1u << 15)) >> 16;
        blur[L.pix + (long long)y * L.w + x] = (uint8_t)min(v, 255u);
    }
}

// ---------------------------------------------------------------------------
// FAST 9/16.  best(p) = max over the 16 arcs of 9 contiguous circle pixels
// of min(v - p_k) (darker arc) or min(p_k - v) (brighter arc).  A pixel is a
// corner at threshold t iff best > t, and cornerScore<16> returns best - 1.
__device__ __forceinline__ int fast_best(const uint8_t* win, int W, int x, int y) {
    const uint8_t* c = win + y * W + x;
    const int v = c[0];
    int d[16];
    d[0] = v - c[3 * W];
    d[1] = v - c[3 * W + 1];
    d[2] = v - c[2 * W + 2];
    d[3] = v - c[W + 3];
    d[4] = v - c[3];
    d[5] = v - c[-W + 3];
    d[6] = v - c[-2 * W + 2];
    d[7] = v - c[-3 * W + 1];
    d[8] = v - c[-3 * W];
    d[9] = v - c[-3 * W - 1];
    d[10] = v - c[-2 * W - 2];
    d[11] = v - c[-W - 3];
    d[12] = v - c[-3];
    d[13] = v - c[W - 3];
    d[14] = v - c[2 * W - 2];
    d[15] = v - c[3 * W - 1];
    // quick reject: any 9-arc covers two consecutive compass points (0,4,8,12)
    int lo0 = min(d[0], d[4]), lo1 = min(d[4], d[8]), lo2 = min(d[8], d[12]), lo3 = min(d[12], d[0]);
    int hi0 = max(d[0], d[4]), hi1 = max(d[4], d[8]), hi2 = max(d[8], d[12]), hi3 = max(d[12], d[0]);
    const int bound = max(max(max(lo0, lo1), max(lo2, lo3)), -min(min(hi0, hi1), min(hi2, hi3)));
    if (bound <= 0) return 0;
    int m2[16], m4[16];
#pragma unroll
    for (int k = 0; k < 16; k++) m2[k] = min(d[k], d[(k + 1) & 15]);
#pragma unroll
    for (int k = 0; k < 16; k++) m4[k] = min(m2[k], m2[(k + 2) & 15]);
    int dark = -1024;
#pragma unroll
    for (int k = 0; k < 16; k++) dark = max(dark, min(min(m4[k], m4[(k + 4) & 15]), d[(k + 8) & 15]));
#pragma unroll
    for (int k = 0; k < 16; k++) m2[k] = max(d[k], d[(k + 1) & 15]);
#pragma unroll
    for (int k = 0; k < 16; k++) m4[k] = max(m2[k], m2[(k + 2) & 15]);
    int bright = -1024;
#pragma unroll
    for (int k = 0; k < 16; k++) bright = max(bright, -max(max(m4[k], m4[(k + 4) & 15]), d[(k + 8) & 15]));
    const int b = max(dark, bright);
    return b < 0 ? 0 : b;
}

__device__ __forceinline__ int block_excl_scan256(int v, int* ws, int& total) {
    // ws: 4 ints of LDS; returns exclusive prefix of v over the block
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
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
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const int s = ws[w];
        base += (w < wid) ? s : 0;
        total += s;
    }
    __syncthreads();
    return base + x - v;
}

// One workgroup per FAST cell window.  The window is staged in LDS, best
// scores computed for the detection area [3, W-3) x [3, H-3) (elsewhere 0,
// as the 3-row ring buffer of FAST_t leaves them), then the NMS/compaction
// pass runs at iniTh and, if it keeps nothing, again at minTh.
__global__ __launch_bounds__(256) void k_fast_cells(OrbGlobals g) {
    __shared__ uint8_t win[kMaxWin * kMaxWin];
    __shared__ uint8_t best[kMaxWin * kMaxWin];
    __shared__ int ws[4];
    const CellDesc c = g.cells[blockIdx.x];
    const LevelDesc L = g.lv[c.lvl];
    const int W = c.x1 - c.x0, H = c.y1 - c.y0;
    const uint8_t* base = g.pyr + L.pix + (long long)c.y0 * L.w + c.x0;
    for (int i = threadIdx.x; i < W * H; i += 256) {
        const int y = i / W, x = i - y * W;
        win[i] = base[(long long)y * L.w + x];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < W * H; i += 256) {
        const int y = i / W, x = i - y * W;
        int b = 0;
        if (x >= 3 && x < W - 3 && y >= 3 && y < H - 3) b = fast_best(win, W, x, y);
        best[i] = (uint8_t)b;
    }
    __syncthreads();
    int total = 0;
    for (int pass = 0; pass < 2; pass++) {
        const int t = pass == 0 ? g.ini_th : g.min_th;
        int written = 0;
        for (int i0 = 0; i0 < W * H; i0 += 256) {
            const int i = i0 + threadIdx.x;
            int keep = 0;
            uint32_t rec = 0;
            if (i < W * H) {
                const int y = i / W, x = i - y * W;
                const int b = best[i];
                if (b > t && x >= 3 && x < W - 3 && y >= 3 && y < H - 3) {
                    const int s = b - 1;
                    keep = 1;
#pragma unroll
                    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                        for (int dx = -1; dx <= 1; dx++) {
                            if (dx == 0 && dy == 0) continue;
                            const int bq = best[i + dy * W + dx];
                            const int mq = bq > t ? bq - 1 : 0;
                            keep &= (s > mq);
                        }
                    rec = ((uint32_t)s << 24) | ((uint32_t)(y + c.offy) << 12) | (uint32_t)(x + c.offx);
                }
            }
            int chunk_total;
            const int pos = block_excl_scan256(keep, ws, chunk_total);
            if (keep) g.cell_cand[c.cap_base + written + pos] = rec;
            written += chunk_total;
        }
        total = written;
        if (total > 0) break;
    }
    if (threadIdx.x == 0) g.cell_n[blockIdx.x] = total;
}

// ---------------------------------------------------------------------------
// Octree.  Every candidate gets its quadtree path: the root index
// (init_nodes[x / hX], ORBextractor.cc:546) and, per depth, the quadrant
// DivideNode sends it to (n1..n4 = 0..3, :506-514) computed from the node
// rectangle exactly as DivideNode does (:470-503).  Sorting by (path, index)
// makes every quadtree node a contiguous range, in the original candidate
// order inside each range.
__device__ __forceinline__ unsigned long long path_code(int x, int y, const LevelDesc& L) {
    const int root = (int)((float)x / L.hX);
    int ulx = (int)(L.hX * (float)root), urx = (int)(L.hX * (float)(root + 1));
    int uly = 0, by = L.maxBY - L.minBY;
    unsigned long long code = (unsigned long long)root;
    for (int d = 1; d <= kOctreeDepth; d++) {
        const int halfX = (int)ceilf((float)(urx - ulx) / 2.f);
        const int halfY = (int)ceilf((float)(by - uly) / 2.f);
        const int mx = ulx + halfX, my = uly + halfY;
        const int qx = ((float)x < (float)mx) ? 0 : 1;
        const int qy = ((float)y < (float)my) ? 0 : 1;
        if (qx) ulx = mx; else urx = mx;
        if (qy) uly = my; else by = my;
        code = (code << 2) | (unsigned long long)(qx + 2 * qy);
    }
    return code;
}

// depth at which two path codes first differ (0 = different roots)
__device__ __forceinline__ int code_div(unsigned long long a, unsigned long long b) {
    const unsigned long long x = a ^ b;
    if ((x >> (2 * kOctreeDepth)) != 0) return 0;
    const int hb = 63 - __clzll(x);          // highest differing bit
    return kOctreeDepth - hb / 2;
}

template <typename T>
__device__ void bitonic_sort(T* a, int n2) {
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const T x = a[i], y = a[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

constexpr int kOctThreads = 256;

__global__ __launch_bounds__(kOctThreads) void k_octree_build(OrbGlobals g) {
    __shared__ unsigned long long s_keys[kSortCapLds];
    __shared__ uint8_t s_div[kSortCapLds];
    __shared__ int ws[4];
    __shared__ int s_hist[kOctreeDepth + 2];
    __shared__ long long s_off;
    __shared__ int s_n, s_maxdiv;
    const int li = blockIdx.x;
    const LevelDesc L = g.lv[li];
    OctLevel* O = g.oct + li;
    // 1. gather the level's candidates in cell order (cells are stored
    //    row-major, i then j, as the reference's nested loops visit them)
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    int run = 0;
    for (int c0 = 0; c0 < L.cell_count; c0 += kOctThreads) {
        const int ci = c0 + threadIdx.x;
        const int cnt = ci < L.cell_count ? g.cell_n[L.cell_first + ci] : 0;
        int tot;
        const int pos = block_excl_scan256(cnt, ws, tot);
        if (ci < L.cell_count) {
            const CellDesc& c = g.cells[L.cell_first + ci];
            for (int k = 0; k < cnt; k++) {
                const int dst = run + pos + k;
                if (dst < L.cand_cap) g.lvl_cand[L.cand_base + dst] = g.cell_cand[c.cap_base + k];
            }
        }
        run += tot;
    }
    const int n = run;
    if (n > L.cand_cap) {
        if (threadIdx.x == 0) {
            O->n = n; O->failed = 2; O->n_roots = 0; O->depth = 0; O->total = 0; O->nout = 0;
            atomicOr(g.err, 2);
        }
        return;
    }
    if (n == 0) {
        if (threadIdx.x == 0) { O->n = 0; O->failed = 0; O->n_roots = 0; O->depth = 0; O->total = 0; O->nout = 0; }
        return;
    }
    __syncthreads();
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    // keys and divergence bytes live in LDS; a level with more candidates
    // than the LDS sort holds uses pool memory instead (same code, flat
    // pointers; the workgroup's waves share one L1, so barriers order it)
    unsigned long long* keys = s_keys;
    uint8_t* dv = s_div;
    if (n > kSortCapLds) {
        if (threadIdx.x == 0) {
            const long long need = 2ll * n2 + (n + 3) / 4 + 2;
            const unsigned long long off = atomicAdd(g.pool_used, (unsigned long long)need);
            s_off = ((long long)off + need <= g.pool_cap) ? (long long)off : -1;
        }
        __syncthreads();
        if (s_off < 0) {
            if (threadIdx.x == 0) {
                O->n = n; O->failed = 1; O->n_roots = 0; O->depth = 0; O->total = 0; O->nout = 0;
                atomicOr(g.err, 1);
            }
            return;
        }
        const long long base = (s_off + 1) & ~1ll;  // 8-byte aligned
        keys = reinterpret_cast<unsigned long long*>(g.pool + base);
        dv = reinterpret_cast<uint8_t*>(g.pool + base + 2ll * n2);
        __syncthreads();
    }
    // 2. keys = path << 24 | candidate index; sort
    for (int i = threadIdx.x; i < n2; i += kOctThreads) {
        unsigned long long k = ~0ull;
        if (i < n) {
            const uint32_t r = g.lvl_cand[L.cand_base + i];
            const int x = r & 0xFFF, y = (r >> 12) & 0xFFF;
            k = (path_code(x, y, L) << 24) | (unsigned long long)i;
        }
        keys[i] = k;
    }
    __syncthreads();
    bitonic_sort(keys, n2);
    // 3. divergence depth of each sorted position, histogram, deepest depth
    if (threadIdx.x < kOctreeDepth + 2) s_hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_maxdiv = 0;
    __syncthreads();
    int mymax = 0;
    for (int i = threadIdx.x; i < n; i += kOctThreads) {
        const int d = i == 0 ? 0 : code_div(keys[i] >> 24, keys[i - 1] >> 24);
        dv[i] = (uint8_t)d;
        atomicAdd(&s_hist[d], 1);
        mymax = max(mymax, d);
    }
    atomicMax(&s_maxdiv, mymax);
    __syncthreads();
    const int D = s_maxdiv;
    // K_d = #{i : div[i] <= d}; T = sum (K_d + 1)
    int K[kOctreeDepth + 1];
    int T = 0;
    {
        int acc = 0;
        for (int d = 0; d <= kOctreeDepth; d++) {
            acc += s_hist[d];
            K[d] = acc;
        }
        for (int d = 0; d <= D; d++) T += K[d] + 1;
    }
    if (threadIdx.x == 0) {
        const long long need = 2ll * T + n;
        const unsigned long long off = atomicAdd(g.pool_used, (unsigned long long)need);
        s_off = ((long long)off + need <= g.pool_cap) ? (long long)off : -1;
    }
    __syncthreads();
    const long long off = s_off;
    if (off < 0) {
        if (threadIdx.x == 0) {
            O->n = n; O->failed = 1; O->n_roots = 0; O->depth = 0; O->total = 0; O->nout = 0;
            atomicOr(g.err, 1);
        }
        return;
    }
    uint32_t* start = g.pool + off;
    uint32_t* child = start + T;
    uint32_t* sorted = child + T;
    for (int i = threadIdx.x; i < n; i += kOctThreads) sorted[i] = (uint32_t)(keys[i] & 0xFFFFFF);
    __syncthreads();
    // 4. node tables, depth by depth.  rank[] (reusing the key storage) holds
    //    each position's node index at the previous depth.
    int* rank = reinterpret_cast<int*>(keys);
    int doff = 0, prev_off = 0;
    for (int d = 0; d <= D; d++) {
        int run2 = 0;
        for (int i0 = 0; i0 < n; i0 += kOctThreads) {
            const int i = i0 + threadIdx.x;
            const int dvi = i < n ? (int)dv[i] : 255;
            const int f = dvi <= d ? 1 : 0;
            int tot;
            const int r = run2 + block_excl_scan256(f, ws, tot);
            if (f) {
                start[doff + r] = (uint32_t)i;
                if (d > 0 && dvi <= d - 1) child[prev_off + rank[i]] = (uint32_t)(doff + r);
                rank[i] = r;  // only read by the same thread, at the next depth
            }
            run2 += tot;
        }
        if (threadIdx.x == 0) {
            start[doff + K[d]] = (uint32_t)n;
            if (d > 0) child[prev_off + K[d - 1]] = (uint32_t)(doff + K[d]);
        }
        prev_off = doff;
        doff += K[d] + 1;
    }
    // deepest depth: no children
    for (int k = threadIdx.x; k <= K[D]; k += kOctThreads) child[prev_off + k] = 0;
    if (threadIdx.x == 0) {
        O->n = n;
        O->failed = 0;
        O->n_roots = K[0];
        O->depth = D;
        O->tab = off;
        O->total = T;
        O->sorted = off + 2ll * T;
        O->nout = 0;
    }
}

// Replays DistributeOctTree's priority_queue (ORBextractor.cc:530-578) on
// node counts: push the non-empty roots, split the largest node until the
// queue holds N nodes or the top holds a single keypoint, then pop
// everything.  One lane per level; the heap lives in LDS.
constexpr int kHeapCap = 4096;
__global__ __launch_bounds__(64) void k_octree_heap(OrbGlobals g) {
    __shared__ unsigned long long heap[kHeapCap];
    if (threadIdx.x != 0) return;
    const int li = blockIdx.x;
    const LevelDesc L = g.lv[li];
    OctLevel* O = g.oct + li;
    if (O->failed || O->n == 0) {
        O->nout = 0;
        return;
    }
    const uint32_t* start = g.pool + O->tab;
    const uint32_t* child = start + O->total;
    int len = 0;
    for (int k = 0; k < O->n_roots; k++) {
        const uint32_t cnt = start[k + 1] - start[k];
        heap_push(heap, len, ((unsigned long long)cnt << 32) | (unsigned)k);
    }
    const int N = min(L.nfeat, kHeapCap - 4);
    while (len < N && len > 0) {
        const unsigned long long top = heap[0];
        if ((top >> 32) == 1) break;
        heap_pop(heap, len);
        const uint32_t gid = (uint32_t)top;
        const uint32_t c0 = child[gid], c1 = child[gid + 1];
        for (uint32_t c = c0; c < c1; c++) {
            const uint32_t cnt = start[c + 1] - start[c];
            heap_push(heap, len, ((unsigned long long)cnt << 32) | c);
        }
    }
    int j = 0;
    while (len > 0) {
        if (j < L.out_cap) g.nodes_out[L.kp_slot + j] = (uint32_t)heap[0];
        j++;
        heap_pop(heap, len);
    }
    O->nout = min(j, L.out_cap);
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void sincos_patched(float x, const SinCosTables& tb, float* c, float* s) {
    const bool neg = x < 0.f;
    const float ax = neg ? -x : x;
    sincos_base(ax, c, s);
    const uint32_t bits = __float_as_uint(ax);
    if (bits <= tb.max_bits) {
        const uint32_t hi = bits >> 16;
        {
            uint32_t lo = tb.cos_idx[hi], up = tb.cos_idx[hi + 1];
            while (lo < up) {
                const uint32_t mid = (lo + up) >> 1;
                const uint32_t k = tb.cos_key[mid];
                if (k < bits) lo = mid + 1; else up = mid;
            }
            if (lo < tb.cos_idx[hi + 1] && tb.cos_key[lo] == bits) *c = tb.cos_val[lo];
        }
        {
            uint32_t lo = tb.sin_idx[hi], up = tb.sin_idx[hi + 1];
            while (lo < up) {
                const uint32_t mid = (lo + up) >> 1;
                const uint32_t k = tb.sin_key[mid];
                if (k < bits) lo = mid + 1; else up = mid;
            }
            if (lo < tb.sin_idx[hi + 1] && tb.sin_key[lo] == bits) *s = tb.sin_val[lo];
        }
    }
    if (neg) *s = -*s;
}

// One wave per keypoint slot of a level: octree slots take the first
// maximum-response candidate of their node (ORBextractor.cc:568-578), get
// the IC angle on the unblurred level (:75-98) and the rBRIEF descriptor on
// the blurred level (:101-141); seed slots keep their angle (:845-847).
// Level > 0 coordinates are then multiplied by the level scale (:886-890).
__global__ __launch_bounds__(256) void k_keypoints(OrbGlobals g, const int* slot_lvl, int n_slots,
                                                   SinCosTables tb, float* out_kps, uint8_t* out_desc, int cap,
                                                   int* out_count) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= n_slots) return;
    const int li = slot_lvl[2 * wave];
    const int j = slot_lvl[2 * wave + 1];
    const LevelDesc L = g.lv[li];
    const OctLevel* O = g.oct + li;
    const int nout = O->nout;
    // image keypoint count, written once per image by slot 0 of level 0
    if (j == 0 && L.level == 0 && lane == 0) {
        int tot = 0;
        for (int l = 0; l < L.nlevels; l++) {
            const int lj = L.first_level + l;
            tot += g.oct[lj].nout + g.lv[lj].seed_count;
        }
        out_count[L.img] = tot;
    }
    if (j >= nout + L.seed_count) return;
    int off = 0;
    for (int l = 0; l < L.level; l++) {
        const int lj = L.first_level + l;
        off += g.oct[lj].nout + g.lv[lj].seed_count;
    }
    const int pos = off + j;
    if (pos >= cap) return;
    float x, y, resp, angle, size;
    bool seed = j >= nout;
    if (!seed) {
        const uint32_t gid = g.nodes_out[L.kp_slot + j];
        const uint32_t* start = g.pool + O->tab;
        const uint32_t* sorted = g.pool + O->sorted;
        const uint32_t s0 = start[gid], s1 = start[gid + 1];
        unsigned long long bestk = 0;
        for (uint32_t p = s0 + lane; p < s1; p += 64) {
            const uint32_t idx = sorted[p];
            const uint32_t r = g.lvl_cand[L.cand_base + idx];
            const unsigned long long k = ((unsigned long long)(r >> 24) << 32) | (0xFFFFFFu - idx);
            bestk = k > bestk ? k : bestk;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long v = __shfl_xor(bestk, o, 64);
            bestk = v > bestk ? v : bestk;
        }
        const uint32_t idx = 0xFFFFFFu - (uint32_t)(bestk & 0xFFFFFFFFu);
        const uint32_t r = g.lvl_cand[L.cand_base + idx];
        x = (float)((int)(r & 0xFFF) + L.minBX);
        y = (float)((int)((r >> 12) & 0xFFF) + L.minBY);
        resp = (float)(r >> 24);
        size = L.patch;
        // IC_Angle: lanes 0..30 own rows v = lane - 15
        const uint8_t* img = g.pyr + L.pix;
        const int cx = (int)rintf(x), cy = (int)rintf(y);
        int m10 = 0, m01 = 0;
        if (lane < 31) {
            const int v = lane - 15;
            const int du = g.umax[v < 0 ? -v : v];
            const uint8_t* row = img + (long long)(cy + v) * L.w + cx;
            for (int u = -du; u <= du; u++) {
                const int val = row[u];
                m10 += u * val;
                m01 += v * val;
            }
        }
        m10 = wave_sum(m10);
        m01 = wave_sum(m01);
        angle = fast_atan2_deg((float)m01, (float)m10);
    } else {
        const float* sd = g.seeds + 6 * (L.seed_first + (j - nout));
        x = sd[0];
        y = sd[1];
        size = sd[2];
        angle = sd[3];
        resp = sd[4];
    }
    // rBRIEF on the blurred level
    const float factorPI = (float)(3.14159265358979323846 / 180.f);
    float a, b;
    sincos_patched(angle * factorPI, tb, &a, &b);
    const uint8_t* bl = g.blur + L.pix;
    const int cx = (int)rintf(x), cy = (int)rintf(y);
    unsigned long long words[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const signed char* p = kOrbPattern[64 * r + lane];
        int t[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const float px = (float)p[2 * q], py = (float)p[2 * q + 1];
            int yy = cy + (int)rintf(px * b + py * a);
            int xx = cx + (int)rintf(px * a - py * b);
            yy = min(max(yy, 0), L.h - 1);
            xx = min(max(xx, 0), L.w - 1);
            t[q] = bl[(long long)yy * L.w + xx];
        }
        words[r] = __ballot(t[0] < t[1]);
    }
    if (lane == 0) {
        unsigned long long* dst = reinterpret_cast<unsigned long long*>(out_desc + ((long long)L.img * cap + pos) * 32);
        dst[0] = words[0];
        dst[1] = words[1];
        dst[2] = words[2];
        dst[3] = words[3];
    }
    if (lane == 0) {
        if (L.level != 0) {
            x = x * L.scale;
            y = y * L.scale;
        }
        float* o = out_kps + ((long long)L.img * cap + pos) * 6;
        o[0] = x;
        o[1] = y;
        o[2] = size;
        o[3] = angle;
        o[4] = resp;
        reinterpret_cast<int*>(o)[5] = L.level;
    }
}

}  // namespace mcv
