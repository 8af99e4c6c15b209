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

struct Strip {
    int lvl;     // LevelDesc index
    int x0, y0;  // strip origin: kStripW x kStripH pixels, one wave
};

struct BlurTaps {
    int t[7];
    uint32_t lo4, hi3;  // taps 0..3 and 4..6 packed as u8x4 for udot4
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
    uint8_t* score;         // FAST best-arc score per pixel (0 at or below the low threshold)
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
    int* perm_out;          // keypoint-slot processing order, written by the octree heap replay
    const int* perm;        // the order the keypoint kernels follow (null: slot order)
};

// two 16-bit lanes in one VGPR (packed u16 operations)
typedef unsigned short mcv_us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ mcv_us2 as_us2(uint32_t v) { return __builtin_bit_cast(mcv_us2, v); }
__device__ __forceinline__ uint32_t as_u32(mcv_us2 v) { return __builtin_bit_cast(uint32_t, v); }

__device__ __forceinline__ int reflect101(int i, int n) {
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i;
}

// ---------------------------------------------------------------------------
// Strip kernels: a wave owns kStripW x kStripH pixels of one level, each lane
// 4 horizontally adjacent pixels (one u32 of a pitched row).
// kXcd: consecutive workgroups land on one XCD (shared L2 for the row halos
// of neighbouring strips); without it the dispatcher's round robin spreads
// consecutive strips over the 8 XCDs, which balances uneven per-strip work.
#define STRIP_SETUP(tab, n, kXcd)                                   \
    const int xb = (kXcd) ? xcd_block(((n) + 3) / 4) : (int)blockIdx.x; \
    if (xb < 0) return;                                             \
    const int sidx = xb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); \
    if (sidx >= (n)) return;                                        \
    const int lane = threadIdx.x & 63;                              \
    const Strip S = (tab)[sidx];                                    \
    const LevelDesc L = lv[S.lvl];                                  \
    const int x = S.x0 + 4 * lane;                                  \
    const int yend = min(S.y0 + kStripH, L.h);

__global__ __launch_bounds__(256) void k_copy_level0(const Strip* st, int n, const LevelDesc* lv,
                                                     const uint8_t* const* src, const int* src_stride,
                                                     uint8_t* __restrict__ pyr) {
    STRIP_SETUP(st, n, true)
    if (x >= L.w) return;
    const uint8_t* s = src[L.img];
    const long long stride = src_stride[L.img];
    uint8_t* d = pyr + L.pix;  // row y, column x at byte y * pitch + x
    const bool vec = x + 4 <= L.w && ((reinterpret_cast<uintptr_t>(s) | (uintptr_t)stride | (uintptr_t)x) & 3) == 0;
    // 8 rows in flight: all loads of a group before its stores
    for (int y0 = S.y0; y0 < yend; y0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int y = y0 + u;
            v[u] = 0;
            if (y < yend) {
                const uint8_t* sr = s + y * stride + x;
                if (vec) {
                    v[u] = *reinterpret_cast<const uint32_t*>(sr);
                } else {
                    for (int k = 0; k < 4 && x + k < L.w; k++) v[u] |= (uint32_t)sr[k] << (8 * k);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (y0 + u < yend) *reinterpret_cast<uint32_t*>(d + (uint32_t)((y0 + u) * L.pitch + x)) = v[u];
    }
}

// cv::resize INTER_LINEAR, CV_8U: horizontal int taps (coef 2048 scale) on
// the two source rows, then the vertical pass
// (((b0 * (r0 >> 4)) >> 16) + ((b1 * (r1 >> 4)) >> 16) + 2) >> 2.
// One wave per strip (64-thread workgroups), no LDS: a lane's 4 outputs
// read source bytes from at most 12 consecutive bytes of a row from its
// aligned word `wb` on (the plan checks it per level), loaded as three u32.  The two horizontal taps of an
// output column are one byte permute (a u16 pair) and one v_dot2, computed
// once per source row.  The wave walks the strip's source rows in order, 4
// rows per step with the next 4 rows' loads in flight, and stores an output
// row once both of its source rows are done.
__global__ __launch_bounds__(64) void k_resize(const Strip* st, int n, const LevelDesc* lv, const int2* coef,
                                               uint8_t* pyr) {
    const int sidx = xcd_block(n);
    if (sidx < 0) return;
    const int lane = threadIdx.x;
    const Strip S = st[sidx];
    const LevelDesc L = lv[S.lvl];
    const LevelDesc SL = lv[L.src];
    const int x = S.x0 + 4 * lane;
    const int ny = min(S.y0 + kStripH, L.h) - S.y0;
    const uint8_t* sp = pyr + SL.pix;
    // the strip's row coefficients, lane k for output row S.y0 + k
    int2 cyl = make_int2(0, 0);
    if (lane < ny) cyl = coef[L.coef_y + S.y0 + lane];
    // column taps: output j takes window bytes off_j and off_j + 1 (off_0 < 4)
    const int wb = coef[L.coef_x + min(x, L.w - 1)].x & ~3;
    uint32_t sel[4], ab[4];
    bool hiw[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int2 cx = coef[L.coef_x + min(x + j, L.w - 1)];
        const int off = cx.x - wb;
        hiw[j] = j > 0 && off >= 4;
        const uint32_t o = (uint32_t)(hiw[j] ? off - 4 : off);
        sel[j] = o | (0x0cu << 8) | ((o + 1) << 16) | (0x0cu << 24);
        ab[j] = (uint32_t)cx.y;  // a0 | a1 << 16, both in [0, 2048]
    }
    const bool own = x < L.w;  // every lane stays: lane k holds row k's coefficients
    uint8_t* dp = pyr + L.pix;  // output row y, column x at byte y * pitch + x
    auto row_a = [&](int k) { return min(max(__builtin_amdgcn_readlane(cyl.x, k), 0), SL.h - 1); };
    auto row_b = [&](int k) { return min(max(__builtin_amdgcn_readlane(cyl.x, k) + 1, 0), SL.h - 1); };
    const int r_lo = row_a(0), r_hi = row_b(ny - 1);
    auto load4 = [&](int r0, uint32_t (&w)[4][3]) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t* rp = reinterpret_cast<const uint32_t*>(
                sp + (uint32_t)(min(r0 + u, SL.h - 1) * SL.pitch + wb));
            w[u][0] = rp[0];
            w[u][1] = rp[1];
            w[u][2] = rp[2];
        }
    };
    // horizontal taps of one loaded source row
    auto hrow = [&](const uint32_t (&w)[3], uint32_t (&h)[4]) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t pr = hiw[j] ? __builtin_amdgcn_perm(w[2], w[1], sel[j]) : __builtin_amdgcn_perm(w[1], w[0], sel[j]);
            h[j] = __builtin_amdgcn_udot2(as_us2(pr), as_us2(ab[j]), 0u, false);
        }
    };
    uint32_t ha[4] = {0, 0, 0, 0}, hb[4] = {0, 0, 0, 0};  // source rows r - 1 and r
    uint32_t wc[4][3], wn[4][3];
    load4(r_lo, wc);
    int yo = 0;
    for (int r0 = r_lo; r0 <= r_hi; r0 += 4) {
        if (r0 + 4 <= r_hi) load4(r0 + 4, wn);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = r0 + u;
            if (r <= r_hi) {
#pragma unroll
                for (int j = 0; j < 4; j++) ha[j] = hb[j];
                hrow(wc[u], hb);
                // the output rows whose lower source row is r (their upper one is r - 1, or r where clamped)
                while (yo < ny && row_b(yo) == r) {
                    const bool same = row_a(yo) == r;
                    const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane(cyl.y, yo);
                    const uint32_t b0 = cb & 0xFFFFu, b1 = cb >> 16;
                    uint32_t v = 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const uint32_t r0v = same ? hb[j] : ha[j];
                        // b < 2^12 and r >> 4 < 2^15: 24-bit multiplies, same products
                        const uint32_t o = ((__umul24(b0, r0v >> 4) >> 16) + (__umul24(b1, hb[j] >> 4) >> 16) + 2) >> 2;
                        v |= (o & 0xFF) << (8 * j);
                    }
                    if (own) *reinterpret_cast<uint32_t*>(dp + (uint32_t)((S.y0 + yo) * L.pitch + x)) = v;
                    yo++;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int q = 0; q < 3; q++) wc[u][q] = wn[u][q];
    }
}

// Lanes near a level's left or right edge: the 12 source bytes x-4 .. x+7
// of a row (REFLECT_101 on indices clamped to [-3, w+2], the range any
// stored output reaches) lie in the 3 aligned words at `base` (x - 4 clamped
// into [0, pitch - 12]); each of the lane's 3 words is one byte permute of
// two consecutive loaded words (pair 0-1 or 1-2).  Fixed per lane for the
// whole strip, so an edge wave runs the same instructions on every lane.
struct EdgeGather {
    int base;
    uint32_t sel[3];
    bool hi[3];  // pair 1-2 instead of 0-1
};

__device__ __forceinline__ EdgeGather edge_gather(int x, int w, int pitch) {
    EdgeGather e;
    e.base = min(max(x - 4, 0), pitch - 12);
#pragma unroll
    for (int o = 0; o < 3; o++) {
        int p[4], pmin = 12;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            p[b] = reflect101(min(max(x - 4 + 4 * o + b, -3), w + 2), w) - e.base;
            pmin = min(pmin, p[b]);
        }
        const int j = min(pmin >> 2, 1);
        uint32_t sel = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) sel |= (uint32_t)(p[b] - 4 * j) << (8 * b);
        e.sel[o] = sel;
        e.hi[o] = j == 1;
    }
    return e;
}

// GaussianBlur 7x7, sigma 2, BORDER_REFLECT_101, CV_8U fixed point: the
// 7 taps (sum 256) horizontally with udot4, then vertically (tap times a
// horizontal sum < 2^16: 24-bit multiplies), out = (sum + 2^15) >> 16.  Each lane keeps the last 7 horizontal rows of its
// 4 columns in registers while walking down the strip.  Two walks compute
// the same bytes: kFixed = false stops at the level's last row; kFixed =
// true always covers kStripH + 6 rows (row indices reflect, stores past the
// level end are skipped).  kEdge: the wave has a lane near a level edge and
// reads its rows through EdgeGather.
template <bool kFixed, bool kEdge>
__device__ __forceinline__ void blur_load_row(const uint8_t* __restrict__ src, const LevelDesc& L, int x, int yi,
                                              const EdgeGather& eg, uint32_t& w0, uint32_t& w1, uint32_t& w2) {
    const int ry = reflect101(min(max(yi, -3), L.h + 2), L.h);
    const int ro = ry * L.pitch;
    if (!kEdge) {
        // interior lanes: x >= 4
        const uint32_t* rw = reinterpret_cast<const uint32_t*>(src + (uint32_t)(ro + x - 4));
        w0 = rw[0];
        w1 = rw[1];
        w2 = rw[2];
    } else {
        const uint32_t* rw = reinterpret_cast<const uint32_t*>(src + (uint32_t)(ro + eg.base));
        const uint32_t r0 = rw[0], r1 = rw[1], r2 = rw[2];
        w0 = __builtin_amdgcn_perm(eg.hi[0] ? r2 : r1, eg.hi[0] ? r1 : r0, eg.sel[0]);
        w1 = __builtin_amdgcn_perm(eg.hi[1] ? r2 : r1, eg.hi[1] ? r1 : r0, eg.sel[1]);
        w2 = __builtin_amdgcn_perm(eg.hi[2] ? r2 : r1, eg.hi[2] ? r1 : r0, eg.sel[2]);
    }
}

template <bool kFixed, bool kEdge>
__device__ __forceinline__ void blur_strip(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                           const LevelDesc& L, int x, int y0, int yend, const BlurTaps& taps,
                                           const EdgeGather& eg) {
    // vertical taps by u16-pair dot products: pr[s % 6][j] holds the
    // horizontal sums of rows s and s + 1 of column j (each < 2^16) packed,
    // so rows s .. s + 6 take three udot2 and one multiply-add
    uint32_t hp[4] = {0, 0, 0, 0}, pr[6][4];
    const mcv_us2 t01 = {(unsigned short)taps.t[0], (unsigned short)taps.t[1]};
    const mcv_us2 t23 = {(unsigned short)taps.t[2], (unsigned short)taps.t[3]};
    const mcv_us2 t45 = {(unsigned short)taps.t[4], (unsigned short)taps.t[5]};
    const uint32_t t6 = (uint32_t)taps.t[6];
    // one row ahead: the next row's words are requested before this row's
    // taps are computed
    uint32_t n0, n1, n2;
    blur_load_row<kFixed, kEdge>(src, L, x, y0 - 3, eg, n0, n1, n2);
#pragma unroll
    for (int r = 0; r < kStripH + 6; r++) {
        const int yi = y0 - 3 + r;
        if (!kFixed && yi - 3 >= yend) break;
        const uint32_t w0 = n0, w1 = n1, w2 = n2;
        if (r + 1 < kStripH + 6) blur_load_row<kFixed, kEdge>(src, L, x, yi + 1, eg, n0, n1, n2);
        uint32_t h[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t lo = j == 3 ? w1 : __builtin_amdgcn_alignbyte(w1, w0, j + 1);
            const uint32_t hi = j == 3 ? w2 : __builtin_amdgcn_alignbyte(w2, w1, j + 1);
            h[j] = __builtin_amdgcn_udot4(lo, taps.lo4, __builtin_amdgcn_udot4(hi, taps.hi3, 0u, false), false);
        }
        if (r >= 1) {
#pragma unroll
            for (int j = 0; j < 4; j++) pr[(r - 1) % 6][j] = hp[j] | (h[j] << 16);
        }
        if (r >= 6) {
            const int y = yi - 3;
            uint32_t v = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint32_t acc = __umul24(t6, h[j]);
                acc = __builtin_amdgcn_udot2(as_us2(pr[(r - 2) % 6][j]), t45, acc, false);
                acc = __builtin_amdgcn_udot2(as_us2(pr[(r - 4) % 6][j]), t23, acc, false);
                acc = __builtin_amdgcn_udot2(as_us2(pr[(r - 6) % 6][j]), t01, acc, false);
                v |= min((acc + (1u << 15)) >> 16, 255u) << (8 * j);
            }
            if (!kFixed || y < yend) *reinterpret_cast<uint32_t*>(dst + (uint32_t)(y * L.pitch + x)) = v;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) hp[j] = h[j];
    }
}

template <bool kFixed>
__global__ __launch_bounds__(256) void k_blur(const Strip* st, int n, const LevelDesc* lv,
                                              const uint8_t* __restrict__ pyr, uint8_t* __restrict__ blur,
                                              BlurTaps taps) {
    STRIP_SETUP(st, n, true)
    if (x >= L.w) return;
    // a wave with no lane near a level edge reads its 3 words directly
    const bool interior = x >= 4 && x + 8 <= L.pitch && x + 6 < L.w;
    if (__all(interior)) {
        blur_strip<kFixed, false>(pyr + L.pix, blur + L.pix, L, x, S.y0, yend, taps, EdgeGather{});
    } else {
        const EdgeGather eg = edge_gather(x, L.w, L.pitch);
        blur_strip<kFixed, true>(pyr + L.pix, blur + L.pix, L, x, S.y0, yend, taps, eg);
    }
}

// Launch variants of the blur (same bytes): the walk kind.
// blur variants: bit 0 the walk (stop at the level end / fixed length), bit 1
// where a captured graph runs it (before FAST / on a branch beside the
// octree heap replay, which leaves most of the chip idle; the blurred levels
// are first read by the keypoint kernel)
constexpr int kBlurVariants = 4;
inline const char* blur_variant_name(int v) {
    static const char* names[4] = {"stop-at-end", "fixed-walk", "stop-at-end beside the heap",
                                   "fixed-walk beside the heap"};
    return (v >= 0 && v < 4) ? names[v] : "?";
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

__device__ __forceinline__ int byte_of(uint32_t w, int j) { return (int)((w >> (8 * j)) & 0xFF); }

// Score plane: best(p) where best > t_lo = min(iniTh, minTh) and the circle
// fits the level, else 0 (values at or below both thresholds never change a
// corner test or a 3x3 suppression).
//
// A wave walks its strip down in groups of 4 rows.  The rows it needs sit in
// an LDS ring (16 rows x 264 bytes: the strip plus 4 bytes either side),
// fetched with one coalesced u32 per lane per row: the 4 rows the next group
// adds are loaded into registers when a group starts and stored to the ring
// when it ends, so a group's compute covers the load latency (the prologue
// issues its 10 rows at once).  A compass
// bound rejects most pixels: every 9-arc holds two consecutive compass
// points (0,4,8,12), so a corner at t needs such a pair both darker than
// v - t or both brighter than v + t; the test runs on 2 pixels per packed
// u16 operation.  Survivors queue up in LDS and are scored 64 at a time by
// the whole wave from the ring; the group's score rows are zeroed when the
// group starts and each score above t_lo is stored as a byte (no LDS tile,
// so that six workgroups fit a CU's LDS).

typedef short mcv_s2 __attribute__((ext_vector_type(2)));

// compass test for two pixels held in the 16-bit lanes of the arguments
// (values 0..255).  With d = p - v, a pair of neighbouring compass points is
// darker than v - t iff max(d_a, d_b) < -t and brighter than v + t iff
// min(d_a, d_b) > t; a lane of the result is > 0 iff some pair qualifies.
// nt = -t in both lanes.
__device__ __forceinline__ uint32_t compass2(mcv_us2 v, mcv_us2 s, mcv_us2 e, mcv_us2 n, mcv_us2 w, mcv_s2 t,
                                             mcv_s2 nt) {
    const mcv_s2 vs = __builtin_bit_cast(mcv_s2, v);
    const mcv_s2 ds = __builtin_bit_cast(mcv_s2, s) - vs, de = __builtin_bit_cast(mcv_s2, e) - vs;
    const mcv_s2 dn = __builtin_bit_cast(mcv_s2, n) - vs, dw = __builtin_bit_cast(mcv_s2, w) - vs;
    const mcv_s2 mx = __builtin_elementwise_min(
        __builtin_elementwise_min(__builtin_elementwise_max(ds, de), __builtin_elementwise_max(de, dn)),
        __builtin_elementwise_min(__builtin_elementwise_max(dn, dw), __builtin_elementwise_max(dw, ds)));
    const mcv_s2 mn = __builtin_elementwise_max(
        __builtin_elementwise_max(__builtin_elementwise_min(ds, de), __builtin_elementwise_min(de, dn)),
        __builtin_elementwise_max(__builtin_elementwise_min(dn, dw), __builtin_elementwise_min(dw, ds)));
    const mcv_s2 r = __builtin_elementwise_max(__builtin_elementwise_max(mn - t, nt - mx), mcv_s2{0, 0});
    return __builtin_bit_cast(uint32_t, r);
}

// exact best of the pixel at base[3 * kRowB]: base points into row -3 of 7
// consecutive rows kRowB bytes apart (the circle's rows -3 .. +3), so every
// circle byte is one LDS read at an immediate offset from one address.
// Both polarities run in one packed-i16 min/max tree: lane half 0 carries
// v - p_k (darker arcs), half 1 carries p_k - v (brighter arcs).  The 16
// arcs of 9 are taken in pairs: for odd k, arcs k - 1 and k share
// m8 = min(e[k .. k + 7]), so their maximum is min(m8, max(e[k - 1], e[k + 8])).
template <int kRowB>
__device__ __forceinline__ int fast_best_at(const uint8_t* base) {
    const int v = base[3 * kRowB];
    const int p[16] = {base[6 * kRowB],     base[6 * kRowB + 1], base[5 * kRowB + 2], base[4 * kRowB + 3],
                       base[3 * kRowB + 3], base[2 * kRowB + 3], base[kRowB + 2],     base[1],
                       base[0],             base[-1],            base[kRowB - 2],     base[2 * kRowB - 3],
                       base[3 * kRowB - 3], base[4 * kRowB - 3], base[5 * kRowB - 2], base[6 * kRowB - 1]};
    mcv_s2 e[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        // (p, p) * (-1, 1) + (v, -v) in 16-bit lanes: one packed multiply-add
        const mcv_s2 pp = {(short)p[k], (short)p[k]};
        e[k] = pp * mcv_s2{-1, 1} + mcv_s2{(short)v, (short)-v};
    }
    // i = 0..7 stands for k = 2 i + 1: m2[i] = min e[k .. k + 1], m4[i] = min e[k .. k + 3]
    mcv_s2 m2[8], m4[8], a[8];
#pragma unroll
    for (int i = 0; i < 8; i++) m2[i] = __builtin_elementwise_min(e[2 * i + 1], e[(2 * i + 2) & 15]);
#pragma unroll
    for (int i = 0; i < 8; i++) m4[i] = __builtin_elementwise_min(m2[i], m2[(i + 1) & 7]);
#pragma unroll
    for (int i = 0; i < 8; i++)
        a[i] = __builtin_elementwise_min(__builtin_elementwise_min(m4[i], m4[(i + 2) & 7]),
                                         __builtin_elementwise_max(e[2 * i], e[(2 * i + 9) & 15]));
#pragma unroll
    for (int h = 4; h >= 1; h >>= 1)
#pragma unroll
        for (int i = 0; i < h; i++) a[i] = __builtin_elementwise_max(a[i], a[i + h]);
    const int b = max((int)a[0].x, (int)a[0].y);
    return b < 0 ? 0 : b;
}

constexpr int kRing = 16;                 // ring rows per wave (power of two)
constexpr int kRingG = 6;                 // guard rows: slots 0..5 repeated after slot 15
constexpr int kRingW = kStripW / 4 + 2;   // u32 per ring row: strip + one word either side
constexpr int kBlk = 4;                   // rows per group (prefetch) and per score block
// the ring holds a group's rows -3 .. +3 (10) and the next group's 4 new
// rows (stored at the group's end) without overwriting a row the group reads
static_assert(2 * kBlk + 6 <= kRing, "FAST ring too small for the prefetch group");

__global__ __launch_bounds__(256) void k_fast_score(const Strip* st, int n, const LevelDesc* lv, const uint8_t* pyr,
                                                    uint8_t* score, int t_lo) {
    __shared__ uint32_t s_ring[4][kRing + kRingG][kRingW];
    __shared__ uint16_t s_queue[4][64 + kStripW];
    STRIP_SETUP(st, n, false)
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t (*ring)[kRingW] = s_ring[wv];
    uint16_t* queue = s_queue[wv];
    const uint8_t* img = pyr + L.pix;
    const bool own = x < L.w;
    const mcv_s2 tt = {(short)t_lo, (short)t_lo}, ntt = {(short)-t_lo, (short)-t_lo};
    // this lane's columns inside the detection range [3, w - 3)
    int bmask = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (x + j >= 3 && x + j < L.w - 3) bmask |= 1 << j;
    // one level row into registers: this lane's u32, plus the word left of
    // the strip (lane 0) or right of it (lane 63); rows clamp to the level
    auto load_row = [&](int r, uint32_t& a, uint32_t& e) {
        const uint32_t ro = (uint32_t)(min(max(r, 0), L.h - 1) * L.pitch);
        const uint8_t* row = img + ro;
        a = x < L.pitch ? *reinterpret_cast<const uint32_t*>(img + (ro + x)) : 0u;
        e = 0u;
        if (lane == 0 && S.x0 >= 4) e = *reinterpret_cast<const uint32_t*>(row + S.x0 - 4);
        if (lane == 63 && S.x0 + kStripW < L.pitch) e = *reinterpret_cast<const uint32_t*>(row + S.x0 + kStripW);
    };
    // slot r mod 16, and slot r mod 16 + 16 too when that is a guard row, so
    // any 7 rows in a row of slots are consecutive in LDS
    auto store_row = [&](int r, uint32_t a, uint32_t e) {
        const int sl = r & (kRing - 1);
        uint32_t* R = ring[sl];
        R[lane + 1] = a;
        if (lane == 0) R[0] = e;
        if (lane == 63) R[kRingW - 1] = e;
        if (sl < kRingG) {
            uint32_t* G = ring[sl + kRing];
            G[lane + 1] = a;
            if (lane == 0) G[0] = e;
            if (lane == 63) G[kRingW - 1] = e;
        }
    };
    {
        // rows y0 - 3 .. y0 + kBlk + 2: all loads in flight, then the stores
        uint32_t a0[kBlk + 6], e0[kBlk + 6];
#pragma unroll
        for (int k = 0; k < kBlk + 6; k++) load_row(S.y0 - 3 + k, a0[k], e0[k]);
#pragma unroll
        for (int k = 0; k < kBlk + 6; k++) store_row(S.y0 - 3 + k, a0[k], e0[k]);
    }
    int qlen = 0;
    for (int yg = S.y0; yg < yend; yg += kBlk) {
        // the next group's new rows yg + kBlk + 3 .. yg + 2 kBlk + 2
        const bool more = yg + kBlk < yend;
        uint32_t na[kBlk], ne[kBlk];
        if (more) {
#pragma unroll
            for (int k = 0; k < kBlk; k++) load_row(yg + kBlk + 3 + k, na[k], ne[k]);
        }
        const int ye = min(yg + kBlk, yend);
        // the group's score rows start at 0; the drain below writes the
        // scores above t_lo into them (this wave's stores, in program order)
        if (own)
            for (int k = 0; k < ye - yg; k++)
                *reinterpret_cast<uint32_t*>(score + L.pix + (uint32_t)((yg + k) * L.pitch + x)) = 0u;
        for (int y = yg; y < ye; y++) {
            const int tb = y - yg;
            wave_lds_sync();
            int surv = 0;
            if (y >= 3 && y < L.h - 3) {
                const uint32_t* C = ring[y & (kRing - 1)];
                const uint32_t c0 = C[lane], c1 = C[lane + 1], c2 = C[lane + 2];
                const uint32_t up = ring[(y - 3) & (kRing - 1)][lane + 1];
                const uint32_t dn = ring[(y + 3) & (kRing - 1)][lane + 1];
                // pixels 0, 2 (even bytes) and 1, 3 (odd bytes) of the lane's
                // word in u16 lanes; byte permutes pick the centre, north and
                // south bytes and the east (x+3) and west (x-3) neighbours out of
                // the adjacent words (selector 0x0c: zero byte)
                const uint32_t m = 0x00FF00FFu, odd = 0x0c030c01u;
                const uint32_t ev = compass2(as_us2(c1 & m), as_us2(dn & m), as_us2(__builtin_amdgcn_perm(c2, c1, 0x0c050c03u)),
                                             as_us2(up & m), as_us2(__builtin_amdgcn_perm(c1, c0, 0x0c030c01u)), tt, ntt);
                const uint32_t od = compass2(as_us2(__builtin_amdgcn_perm(0u, c1, odd)), as_us2(__builtin_amdgcn_perm(0u, dn, odd)),
                                             as_us2(__builtin_amdgcn_perm(c2, c1, 0x0c060c04u)),
                                             as_us2(__builtin_amdgcn_perm(0u, up, odd)),
                                             as_us2(__builtin_amdgcn_perm(c1, c0, 0x0c040c02u)), tt, ntt);
                // each 16-bit lane capped at 1, then bit 0 / 16 of ev and od
                // moved to bits 0..3 (pixel order 0, 1, 2, 3)
                const uint32_t e1 = as_u32(__builtin_elementwise_min(as_us2(ev), mcv_us2{1, 1}));
                const uint32_t o1 = as_u32(__builtin_elementwise_min(as_us2(od), mcv_us2{1, 1}));
                const uint32_t t2 = e1 | (o1 << 1);
                surv = (int)(((t2 & 3u) | ((t2 >> 14) & 12u)) & (uint32_t)bmask);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const unsigned long long b = __ballot((surv >> j) & 1);
                if ((surv >> j) & 1) queue[qlen + lanes_below(b)] = (uint16_t)((tb << 8) | (4 * lane + j));
                qlen += __popcll(b);
            }
            const bool block_end = y == ye - 1;
            while (qlen >= 64 || (block_end && qlen > 0)) {
                wave_lds_sync();
                const int cnt = min(qlen, 64);
                if (lane < cnt) {
                    const uint32_t e = queue[qlen - cnt + lane];
                    const int et = (int)(e >> 8), pos = (int)(e & 0xFF);
                    const int yy = y - tb + et;
                    // rows yy - 3 .. yy + 3 from slot (yy - 3) mod 16 on (guard rows past 15)
                    const uint8_t* base = reinterpret_cast<const uint8_t*>(ring[(yy - 3) & (kRing - 1)]) + 4 + pos;
                    const int b = fast_best_at<kRingW * 4>(base);
                    if (b > t_lo) score[L.pix + (uint32_t)(yy * L.pitch + S.x0 + pos)] = (uint8_t)b;
                }
                qlen -= cnt;
            }
        }
        if (more) {
            // the ring slots of rows yg + kBlk + 3 + k last held rows yg - 13 + kBlk + k < yg - 3
#pragma unroll
            for (int k = 0; k < kBlk; k++) store_row(yg + kBlk + 3 + k, na[k], ne[k]);
        }
    }
}

// 3x3 suppression and compaction, one wave per FAST cell window.  Pixels of
// the detection area [3, W-3) x [3, H-3) are visited in row-major order;
// neighbours outside it count as 0, as in FAST_t's zeroed ring rows and
// border columns.  A pass at iniTh keeps s = best - 1 where s exceeds every
// neighbour's (best > t ? best - 1 : 0); a cell with no keeper is redone at
// minTh (ORBextractor.cc:604-616).
//
// The area is walked in bands of kCellBand rows.  A band and its one-pixel
// frame are staged in LDS with coalesced u32 loads that keep the global word
// alignment (the frame and any byte outside the area are stored as 0), so
// the centre and neighbour reads of the suppression are LDS reads with no
// bounds tests.  A cell of at most kCellBand rows (the 35-pixel grid's cells
// are 35-52 rows) is staged once and serves both the iniTh and the minTh pass.
constexpr int kCellBand = 46;
constexpr int kCellSW = kMaxWin;  // LDS bytes per staged row: >= 4 * ceil((3 + W - 4) / 4)

// ceil(2^16 / n) for n = 1 .. 24 (u32 per staged row): i / n ==
// (i * kDiv16[n]) >> 16 for every index i < (kCellBand + 2) * 24 used here
__constant__ uint32_t kDiv16[25] = {0,    65536, 32768, 21846, 16384, 13108, 10923, 9363, 8192,
                                    7282, 6554,  5958,  5462,  5042,  4682,  4370,  4096, 3856,
                                    3641, 3450,  3277,  3121,  2979,  2850,  2731};

struct CellWin {
    const uint8_t* wbase;  // word holding detection column -1 of row 0
    int lead;              // byte of column -1 in that word
    int nwr;               // u32 per staged row
    uint32_t wmagic;       // t / nwr for t < (kCellBand + 2) * 24
    int pitch, DW, DH;
};

// staged in batches of 6 words per lane: the batch's loads are all in
// flight before the first store
constexpr int kStageBatch = 6;
__device__ __forceinline__ void cell_stage(const CellWin& C, uint32_t* __restrict__ band, int y0, int nb) {
    const int lane = threadIdx.x & 63;
    const int tot = (nb + 2) * C.nwr;
    for (int k0 = 0; k0 < tot; k0 += 64 * kStageBatch) {
        uint32_t v[kStageBatch], keep[kStageBatch];
        int dst[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; u++) {
            const int k = k0 + 64 * u + lane;
            v[u] = 0;
            keep[u] = 0;
            dst[u] = -1;
            if (k < tot) {
                const int r = (int)(__umul24((uint32_t)k, C.wmagic) >> 16), w = k - __mul24(r, C.nwr);
                const int ry = y0 - 1 + r;
                dst[u] = r * (kCellSW / 4) + w;
                if (ry >= 0 && ry < C.DH) {
                    v[u] = *reinterpret_cast<const uint32_t*>(C.wbase + (__umul24(ry, C.pitch) + 4 * w));
                    // keep the bytes of detection columns [0, DW)
                    const int lo = min(max(C.lead + 1 - 4 * w, 0), 4), hi = min(max(C.lead + 1 + C.DW - 4 * w, 0), 4);
                    keep[u] = hi > lo ? (0xFFFFFFFFu >> (32 - 8 * (hi - lo))) << (8 * lo) : 0u;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; u++)
            if (dst[u] >= 0) band[dst[u]] = v[u] & keep[u];
    }
}

// suppression over the staged rows y0 .. y0 + nb - 1; appends the keepers.
// With t >= 1 the keep rule s > max over the 8 neighbours of (b > t ? b - 1
// : 0), s = best - 1, best > t, is the strict local maximum best > every
// neighbour's byte (a neighbour at or below t is below best anyway).  The
// staged words (4 columns each) are scanned 64 at a time in row-major order;
// those with a byte above t (a minority) join a per-wave queue in that order,
// and the queue is suppressed 64 words at a time: a lane tests its word's 4
// bytes at once, the 8 neighbour words (the left and right neighbours
// rebuilt from the adjacent words with byte shifts) maxed as even and odd
// bytes in packed u16 halves, and a byte is kept when it exceeds both that
// maximum and t.  The keepers of a step are compacted in (word, byte) order:
// the lane's offset is the count of keep bits in the lanes below (one
// ballot per byte position).
constexpr int kCellQ = 128;  // queued word indices per wave (a ring)

__device__ __forceinline__ int cell_scan(const CellWin& C, const uint32_t* __restrict__ band, int y0, int nb,
                                         int t, int offx, int offy, uint32_t* __restrict__ out, int written,
                                         uint32_t* __restrict__ q) {
    const int lane = threadIdx.x & 63;
    const int wf = (C.lead + 1) >> 2;                 // first word holding an area byte
    const int nwa = ((C.lead + C.DW) >> 2) - wf + 1;  // words per row holding area bytes
    const int wlast = (C.lead + C.DW + 1) >> 2;       // last staged word a right neighbour can need
    const uint32_t amagic = kDiv16[nwa];  // i / nwa
    // a byte is kept when it exceeds the neighbour maximum and tk; t lies in
    // [0, 255] (clamped on the host as FAST_t does), and for t = 0 a score
    // s = 0 never beats a non-corner neighbour, so best must exceed 1
    const int tk = max(t, 1);
    const mcv_us2 tt = {(unsigned short)tk, (unsigned short)tk};
    const uint32_t addt = 0x01010101u * (uint32_t)(0x7F - tk);
    const int nit = nb * nwa;
    // suppression of queue entries head .. head + cnt - 1 (word indices in
    // increasing order), keepers appended in (word, byte) order
    int head = 0, tail = 0;
    auto drain = [&](int cnt) {
        int kb = 0, r = 0, w = 0;
        uint32_t W = 0;
        if (lane < cnt) {
            const int i = (int)q[(head + lane) & (kCellQ - 1)];
            r = (int)(__umul24((uint32_t)i, amagic) >> 16);
            w = wf + i - __mul24(r, nwa);
            const int wl = max(w - 1, 0), wr = min(w + 1, wlast);
            // staged rows r (above), r + 1 (the word's row), r + 2 (below)
            const uint32_t* R0 = band + r * (kCellSW / 4);
            const uint32_t* R1 = R0 + kCellSW / 4;
            const uint32_t* R2 = R1 + kCellSW / 4;
            W = R1[w];
            const uint32_t n[8] = {__builtin_amdgcn_alignbyte(R0[w], R0[wl], 3), R0[w],
                                   __builtin_amdgcn_alignbyte(R0[wr], R0[w], 1), __builtin_amdgcn_alignbyte(W, R1[wl], 3),
                                   __builtin_amdgcn_alignbyte(R1[wr], W, 1), __builtin_amdgcn_alignbyte(R2[w], R2[wl], 3),
                                   R2[w], __builtin_amdgcn_alignbyte(R2[wr], R2[w], 1)};
            const uint32_t me = 0x00FF00FFu;
            mcv_us2 mxe = as_us2(n[0] & me), mxo = as_us2((n[0] >> 8) & me);
#pragma unroll
            for (int k = 1; k < 8; k++) {
                mxe = __builtin_elementwise_max(mxe, as_us2(n[k] & me));
                mxo = __builtin_elementwise_max(mxo, as_us2((n[k] >> 8) & me));
            }
            const mcv_us2 ce = as_us2(W & me), co = as_us2((W >> 8) & me);
            const uint32_t ke = as_u32(__builtin_elementwise_min(__builtin_elementwise_sub_sat(ce, mxe),
                                                                 __builtin_elementwise_sub_sat(ce, tt)));
            const uint32_t ko = as_u32(__builtin_elementwise_min(__builtin_elementwise_sub_sat(co, mxo),
                                                                 __builtin_elementwise_sub_sat(co, tt)));
            kb = ((ke & 0xFFFFu) ? 1 : 0) | ((ko & 0xFFFFu) ? 2 : 0) | ((ke >> 16) ? 4 : 0) | ((ko >> 16) ? 8 : 0);
            // bytes outside the area are 0 in the staged rows, so never exceed t
        }
        if (__any(kb != 0)) {
            int below = 0, total = 0;
#pragma unroll
            for (int bt = 0; bt < 4; bt++) {
                const unsigned long long bm = __ballot((kb >> bt) & 1);
                below += lanes_below(bm);
                total += __popcll(bm);
            }
            int k = written + below;
#pragma unroll
            for (int bt = 0; bt < 4; bt++)
                if ((kb >> bt) & 1) {
                    const uint32_t s = ((W >> (8 * bt)) & 0xFFu) - 1u;
                    const int x = 4 * w + bt - (C.lead + 1);  // detection column
                    out[k++] = (s << 24) | ((uint32_t)(y0 + r + 3 + offy) << 12) | (uint32_t)(x + 3 + offx);
                }
            written += total;
        }
        head += cnt;
    };
    for (int i0 = 0; i0 < nit; i0 += 64) {
        const int i = i0 + lane;
        uint32_t W = 0;
        if (i < nit) {
            const int r = (int)(__umul24((uint32_t)i, amagic) >> 16);
            W = band[(r + 1) * (kCellSW / 4) + wf + i - __mul24(r, nwa)];
        }
        // t <= 127: a byte > t sets its high bit in (W + 0x01010101 * (127 - t)) | W
        const bool any = tk <= 127 ? ((((W + addt) | W) & 0x80808080u) != 0u)
                                   : (int)max(max(W & 0xFFu, (W >> 8) & 0xFFu), max((W >> 16) & 0xFFu, W >> 24)) > tk;
        // words with a byte above t join the queue (fewer than 64 are queued
        // before, so the ring never holds more than 127)
        const unsigned long long b = __ballot(any);
        wave_lds_sync();
        if (any) q[(tail + (int)lanes_below(b)) & (kCellQ - 1)] = (uint32_t)i;
        tail += __popcll(b);
        wave_lds_sync();
        if (tail - head >= 64) drain(64);
    }
    if (tail > head) drain(tail - head);
    return written;
}

__global__ __launch_bounds__(256) void k_fast_cells(OrbGlobals g, int ncells) {
    __shared__ uint32_t s_band[4][(kCellBand + 2) * kCellSW / 4];
    __shared__ uint32_t s_q[4][kCellQ];
    const int ci = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (ci >= ncells) return;
    const int lane = threadIdx.x & 63;
    const CellDesc c = g.cells[ci];
    const LevelDesc L = g.lv[c.lvl];
    const int DW = c.x1 - c.x0 - 6, DH = c.y1 - c.y0 - 6;
    int total = 0;
    if (DW > 0 && DH > 0) {
        const uint8_t* sp = g.score + L.pix + (long long)(c.y0 + 3) * L.pitch + (c.x0 + 3);
        CellWin C;
        C.wbase = reinterpret_cast<const uint8_t*>(reinterpret_cast<uintptr_t>(sp - 1) & ~(uintptr_t)3);
        C.lead = (int)((sp - 1) - C.wbase);
        C.nwr = (C.lead + DW + 2 + 3) >> 2;
        C.wmagic = kDiv16[C.nwr];
        C.pitch = L.pitch;
        C.DW = DW;
        C.DH = DH;
        uint32_t* out = g.cell_cand + c.cap_base;
        uint32_t* band = s_band[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
        const bool once = DH <= kCellBand;
        if (once) {
            cell_stage(C, band, 0, DH);
            wave_lds_sync();
        }
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            const int t = pass == 0 ? g.ini_th : g.min_th;
            int written = 0;
            for (int y0 = 0; y0 < DH; y0 += kCellBand) {
                const int nb = min(kCellBand, DH - y0);
                if (!once) {
                    cell_stage(C, band, y0, nb);
                    wave_lds_sync();
                }
                written = cell_scan(C, band, y0, nb, t, c.offx, c.offy, out, written,
                                    s_q[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)]);
                if (!once) wave_lds_sync();
            }
            total = written;
            if (total > 0) break;
        }
    }
    if (lane == 0) g.cell_n[ci] = total;
}

// Exclusive scan of v over a workgroup of kWaves waves; ws: kWaves ints of
// LDS; total receives the block sum.
template <int kWaves>
__device__ __forceinline__ int block_excl_scan(int v, int* ws, int& total) {
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
    for (int w = 0; w < kWaves; w++) {
        const int s = ws[w];
        base += (w < wid) ? s : 0;
        total += s;
    }
    __syncthreads();
    return base + x - v;
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

// Exclusive scan over a workgroup of kWaves waves of every depth's count at
// once; total[d] receives the block sum.  ws: (kWaves + 1) x (kOctreeDepth +
// 1) ints of LDS.
template <int kWaves>
__device__ __forceinline__ void block_excl_scan_depths(int (&v)[kOctreeDepth + 1], int (&total)[kOctreeDepth + 1],
                                                       int (*ws)[kOctreeDepth + 1]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x[kOctreeDepth + 1];
#pragma unroll
    for (int d = 0; d <= kOctreeDepth; d++) x[d] = v[d];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int d = 0; d <= kOctreeDepth; d++) {
            const int y = __shfl_up(x[d], o, 64);
            if (lane >= o) x[d] += y;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int d = 0; d <= kOctreeDepth; d++) ws[wid][d] = x[d];
    }
    __syncthreads();
    // one thread per depth turns the wave sums into exclusive wave offsets
    // (row kWaves receives the totals)
    if (threadIdx.x <= kOctreeDepth) {
        const int d = threadIdx.x;
        int run = 0;
        for (int w = 0; w < kWaves; w++) {
            const int t = ws[w][d];
            ws[w][d] = run;
            run += t;
        }
        ws[kWaves][d] = run;
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d <= kOctreeDepth; d++) {
        v[d] = ws[wid][d] + x[d] - v[d];
        total[d] = ws[kWaves][d];
    }
    __syncthreads();
}

// kOctThreads: 256 for most levels, 1024 for the large ones (the bitonic
// stages and the per-thread chunks then spread over 16 waves); `levels`
// lists the levels of this launch.
template <int kOctThreads>
__global__ __launch_bounds__(kOctThreads) void k_octree_build(OrbGlobals g, const int* levels) {
    constexpr int kWaves = kOctThreads / 64;
    // candidates sorted in LDS: the 256-thread workgroups (the small levels,
    // at most a few hundred candidates) take half the capacity, so that more
    // of them fit on a CU at once
    constexpr int kCap = kOctThreads >= 1024 ? kSortCapLds : kSortCapLds / 2;
    __shared__ unsigned long long s_keys[kCap];
    __shared__ uint8_t s_div[kCap];
    __shared__ int ws[kWaves];
    __shared__ int s_ws[kWaves + 1][kOctreeDepth + 1];
    __shared__ long long s_off;
    const int li = levels[blockIdx.x];
    const LevelDesc L = g.lv[li];
    OctLevel* O = g.oct + li;
    // 1. gather the level's candidates in cell order (cells are stored
    //    row-major, i then j, as the reference's nested loops visit them)
    int run = 0;
    for (int c0 = 0; c0 < L.cell_count; c0 += kOctThreads) {
        const int ci = c0 + threadIdx.x;
        const bool has = ci < L.cell_count;
        const int cnt = has ? g.cell_n[L.cell_first + ci] : 0;
        const int cap_base = has ? g.cells[L.cell_first + ci].cap_base : 0;
        int tot;
        const int pos = block_excl_scan<kWaves>(cnt, ws, tot);
        // 8 candidate words in flight per step
        for (int k0 = 0; k0 < cnt; k0 += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k0 + u < cnt) v[u] = g.cell_cand[cap_base + k0 + u];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int dst = run + pos + k0 + u;
                if (k0 + u < cnt && dst < L.cand_cap) g.lvl_cand[L.cand_base + dst] = v[u];
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
    if (n > kCap) {
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
    bitonic_sort<unsigned long long, (kCap / 2 + kOctThreads - 1) / kOctThreads>(keys, n2);
    // 3. divergence depth of each sorted position
    for (int i = threadIdx.x; i < n; i += kOctThreads)
        dv[i] = (uint8_t)(i == 0 ? 0 : code_div(keys[i] >> 24, keys[i - 1] >> 24));
    __syncthreads();
    // 4. node tables.  Position i starts a node at every depth d >= div[i];
    //    its node index at depth d is rank_d(i) = #{j < i : div[j] <= d}.
    //    Thread t owns the contiguous positions [i0, i1): per-depth counts,
    //    one block scan of all depths, then the writes in position order.
    const int C = (n + kOctThreads - 1) / kOctThreads;
    const int i0 = min(n, (int)threadIdx.x * C), i1 = min(n, i0 + C);
    int rank[kOctreeDepth + 1];
#pragma unroll
    for (int d = 0; d <= kOctreeDepth; d++) rank[d] = 0;
    for (int i = i0; i < i1; i++) {
        const int dvi = dv[i];
#pragma unroll
        for (int d = 0; d <= kOctreeDepth; d++) rank[d] += dvi <= d ? 1 : 0;
    }
    int K[kOctreeDepth + 1];
    block_excl_scan_depths<kWaves>(rank, K, s_ws);
    // K[d] = nodes at depth d; the deepest depth D is the first with K == n
    int D = kOctreeDepth, T = 0;
    int off[kOctreeDepth + 1];
#pragma unroll
    for (int d = kOctreeDepth; d >= 0; d--)
        if (K[d] == n) D = d;
#pragma unroll
    for (int d = 0; d <= kOctreeDepth; d++) {
        off[d] = T;
        if (d <= D) T += K[d] + 1;
    }
    if (threadIdx.x == 0) {
        const long long need = 2ll * T + n;
        const unsigned long long o = atomicAdd(g.pool_used, (unsigned long long)need);
        s_off = ((long long)o + need <= g.pool_cap) ? (long long)o : -1;
    }
    __syncthreads();
    const long long tab = s_off;
    if (tab < 0) {
        if (threadIdx.x == 0) {
            O->n = n; O->failed = 1; O->n_roots = 0; O->depth = 0; O->total = 0; O->nout = 0;
            atomicOr(g.err, 1);
        }
        return;
    }
    uint32_t* start = g.pool + tab;
    uint32_t* child = start + T;
    uint32_t* sorted = child + T;
    for (int i = threadIdx.x; i < n; i += kOctThreads) sorted[i] = (uint32_t)(keys[i] & 0xFFFFFF);
    if (n <= kCap) {
        // depth by depth: the entries of depth d (start) and of depth d - 1
        // (child) are assembled in LDS over the dead sort keys, then stored
        // as contiguous runs.  A thread's entries of one depth are the ranks
        // rank[d] .. of its positions that start a node there.
        __syncthreads();
        uint32_t* st_start = reinterpret_cast<uint32_t*>(s_keys);
        uint32_t* st_child = st_start + kCap;
#pragma unroll
        for (int d = 0; d <= kOctreeDepth; d++) {
            if (d <= D) {
                int r = rank[d], rp = d > 0 ? rank[d - 1] : 0;
                for (int i = i0; i < i1; i++) {
                    const int dvi = dv[i];
                    if (dvi <= d) {
                        st_start[r] = (uint32_t)i;
                        // the depth d-1 node starting at i has this depth-d node as first child
                        if (d > 0 && dvi <= d - 1) st_child[rp++] = (uint32_t)(off[d] + r);
                        r++;
                    }
                }
                __syncthreads();
                for (int k = threadIdx.x; k < K[d]; k += kOctThreads) start[off[d] + k] = st_start[k];
                if (d > 0)
                    for (int k = threadIdx.x; k < K[d - 1]; k += kOctThreads) child[off[d - 1] + k] = st_child[k];
                __syncthreads();
            }
        }
    } else {
        for (int i = i0; i < i1; i++) {
            const int dvi = dv[i];
#pragma unroll
            for (int d = 0; d <= kOctreeDepth; d++) {
                if (d <= D && dvi <= d) {
                    const int r = rank[d]++;
                    start[off[d] + r] = (uint32_t)i;
                    // the depth d-1 node starting at i was given rank[d-1] - 1 just before
                    if (d > 0 && dvi <= d - 1) child[off[d - 1] + rank[d - 1] - 1] = (uint32_t)(off[d] + r);
                }
            }
        }
    }
    // sentinels: start[K_d] = n; child of the depth d-1 sentinel = depth d
    // sentinel; the deepest depth has no children
    if (threadIdx.x <= D) {
        const int d = threadIdx.x;
        int od = 0, Kd = 0, op = 0, Kp = 0;
#pragma unroll
        for (int e = 0; e <= kOctreeDepth; e++)
            if (e == d) { od = off[e]; Kd = K[e]; }
#pragma unroll
        for (int e = 0; e <= kOctreeDepth; e++)
            if (e == d - 1) { op = off[e]; Kp = K[e]; }
        start[od + Kd] = (uint32_t)n;
        if (d > 0) child[op + Kp] = (uint32_t)(od + Kd);
    }
    {
        int oD = 0, KD = 0;
#pragma unroll
        for (int e = 0; e <= kOctreeDepth; e++)
            if (e == D) { oD = off[e]; KD = K[e]; }
        for (int k = threadIdx.x; k <= KD; k += kOctThreads) child[oD + k] = 0;
    }
    if (threadIdx.x == 0) {
        O->n = n;
        O->failed = 0;
        O->n_roots = K[0];
        O->depth = D;
        O->tab = tab;
        O->total = T;
        O->sorted = tab + 2ll * T;
        O->nout = 0;
    }
}

// Replays DistributeOctTree's priority_queue (ORBextractor.cc:530-578) on
// node counts: push the non-empty roots, split the largest node until the
// queue holds N nodes or the top holds a single keypoint, then pop
// everything.  An entry is {count << 3 | children, first child} (a node
// without children, at the deepest depth, carries its own index instead),
// so a popped node's children are known without another lookup.
//
// The whole wave runs the heap (ParHeap: one LDS round trip per gathered
// chain or 5-level subtree), and lanes 0..4 fetch the counts and child
// ranges of a popped node's children together from the node tables in
// global memory.  The workgroup's LDS is the heap alone (8 bytes per entry),
// so workgroups of other kernels still fit on a CU while the replays run.
// Workgroups take levels in `order` (largest feature targets first).  Output
// per keypoint slot: the node's winning candidate word.
// compiler-only ordering of this wave's LDS accesses (the LDS executes one
// wave's DS instructions in order)
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

struct NodeTabs {
    const uint32_t* start;
    const uint32_t* child;
    __device__ __forceinline__ uint32_t st(uint32_t i) const { return start[i]; }
    __device__ __forceinline__ uint32_t ch(uint32_t i) const { return child[i]; }
};

// The first entries of a level's node tables copied to LDS (the shallow
// depths, where the heap replay splits its nodes): a split then reads the
// popped node's children from LDS instead of waiting on global memory.
struct TabPrefix {
    uint32_t* st;
    uint32_t* ch;
    uint32_t n;  // entries held
};

__device__ __forceinline__ TabPrefix tab_prefix_load(const NodeTabs& nt, uint32_t* lds, int cap, uint32_t total) {
    const int lane = threadIdx.x & 63;
    TabPrefix P{lds, lds + cap, min(total, (uint32_t)cap)};
    // 8 entries of each table per lane in flight
    for (uint32_t i0 = 0; i0 < P.n; i0 += 512) {
        uint32_t a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t i = i0 + 64 * u + lane;
            a[u] = i < P.n ? nt.start[i] : 0;
            b[u] = i < P.n ? nt.child[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const uint32_t i = i0 + 64 * u + lane;
            if (i < P.n) {
                P.st[i] = a[u];
                P.ch[i] = b[u];
            }
        }
    }
    wave_lds_sync();
    return P;
}

// lanes 0..nch read start[c0 + lane] and child[c0 + lane] (from LDS when
// the prefix holds all of them)
__device__ __forceinline__ void tab_children(const NodeTabs& nt, const TabPrefix& P, uint32_t c0, uint32_t nch,
                                             uint32_t& s_l, uint32_t& c_l) {
    const int lane = threadIdx.x & 63;
    s_l = 0;
    c_l = 0;
    if (lane <= (int)nch) {
        if (c0 + nch < P.n) {
            s_l = P.st[c0 + lane];
            c_l = P.ch[c0 + lane];
        } else {
            s_l = nt.st(c0 + lane);
            c_l = nt.ch(c0 + lane);
        }
    }
}

// libstdc++ binary heap (push_heap / pop_heap, keys compared alone) run by a
// whole wave on LDS arrays: every sift gathers, in one round trip, either the
// ancestor chain of the hole (sift-up) or the 5-level subtree below it
// (62 nodes, sift-down), and the path is then decided in registers, so a pop
// over a 1000-entry heap costs about three LDS round trips instead of one per
// level.  Entries: key (count << 3 | children) and payload in two arrays.
struct ParHeap {
    uint32_t* key;
    uint32_t* pay;
};

__device__ __forceinline__ uint32_t hkey(uint32_t k) { return k >> 3; }

// push_heap's __push_heap(first, hole, 0, value): move parents down while
// parent < value, then place value.  Lane l reads the l-th ancestor of the
// hole; the run of ancestors smaller than the value (ancestor keys never
// decrease upwards) moves down one step each.  All lanes take part; the
// wave's LDS accesses execute in order, so only the compiler is fenced.
__device__ __forceinline__ void par_sift_up(ParHeap& H, int hole, uint32_t vk, uint32_t vp) {
    const int lane = threadIdx.x & 63;
    // c_l = ((hole + 1) >> l) - 1 is the l-th node of the chain (c_0 = hole)
    const int cl = ((hole + 1) >> lane) - 1;
    const int cp = ((hole + 1) >> (lane + 1)) - 1;
    uint32_t pk = 0xFFFFFFFFu, pp = 0;
    if (lane < 31 && cl >= 1) {
        pk = H.key[cp];
        pp = H.pay[cp];
    }
    const unsigned long long go = __ballot(hkey(pk) < hkey(vk));
    const int L = (int)__builtin_ctzll(~go);  // first step that does not move
    if (lane < L) {
        H.key[cl] = pk;
        H.pay[cl] = pp;
    }
    if (lane == 0) {
        const int final_hole = ((hole + 1) >> L) - 1;
        H.key[final_hole] = vk;
        H.pay[final_hole] = vp;
    }
    lds_order();
}

__device__ __forceinline__ void par_push(ParHeap& H, int& len, uint32_t vk, uint32_t vp) {
    par_sift_up(H, len, vk, vp);
    len++;
}

// 5-level subtrees gathered in one LDS round trip: lane (2^d - 2) + k holds
// node k of depth d (d in 1..5) below the subtree's root.
// lane -> (depth d in 1..5, offset k) of a 5-level subtree
__device__ __forceinline__ int subtree_depth(int lane) {
    return lane < 2 ? 1 : lane < 6 ? 2 : lane < 14 ? 3 : lane < 30 ? 4 : 5;
}

// The reads a pop starts with, in one LDS round trip: the top, the last
// entry and the 5-level subtree below the root.
struct PopReads {
    uint32_t tk, tp, lk, lp, sk, sp;
};

__device__ __forceinline__ PopReads par_pop_reads(const ParHeap& H, int len) {
    const int lane = threadIdx.x & 63;
    const int d = subtree_depth(lane);
    const int idx = (1 << d) - 1 + (lane - ((1 << d) - 2));
    const int last = len - 1;
    PopReads r;
    r.tk = H.key[0];
    r.tp = H.pay[0];
    r.lk = H.key[last];
    r.lp = H.pay[last];
    r.sk = 0;
    r.sp = 0;
    if (lane < 62 && idx < last) {
        r.sk = H.key[idx];
        r.sp = H.pay[idx];
    }
    r.tk = __builtin_amdgcn_readfirstlane(r.tk);
    r.tp = __builtin_amdgcn_readfirstlane(r.tp);
    r.lk = __builtin_amdgcn_readfirstlane(r.lk);
    r.lp = __builtin_amdgcn_readfirstlane(r.lp);
    return r;
}

// Per-lane constants of the 5-level subtree layout: the lane's depth and
// offset, whether it is a right child, and its own and its ancestors' lanes
// as a 64-bit lane mask (lanes 62 and 63 hold no node: their mask names
// lane 63, which no ballot below ever sets).
struct SubLanes {
    int d, k;
    bool valid, right;
    uint32_t am_lo, am_hi;
};

__device__ __forceinline__ SubLanes sub_lanes() {
    const int lane = threadIdx.x & 63;
    SubLanes S;
    S.d = subtree_depth(lane);
    S.k = lane - ((1 << S.d) - 2);
    S.valid = lane < 62;
    S.right = (lane & 1) != 0;
    unsigned long long am = 1ull << 63;
    if (S.valid) {
        am = 0;
        for (int e = 1; e <= S.d; e++) am |= 1ull << ((1 << e) - 2 + (S.k >> (S.d - e)));
    }
    S.am_lo = (uint32_t)am;
    S.am_hi = (uint32_t)(am >> 32);
    return S;
}

// pop_heap with its reads done, as one top-down pass.  pop_heap moves the
// hole from the root to a leaf along the max-child path of the heap without
// its last element x (the right child unless it is smaller than the left;
// the single child at the end of an even-length heap) and then pushes x up
// from that leaf while the parent is smaller.  The path values never
// increase, so the same result comes from walking the path down while the
// next path value is >= x, moving each such value up one level, and placing
// x where the walk stops.  Each round gathers the 5-level subtree below the
// hole (the root's came with the reads; absent nodes read as key 0, below
// every real key).  A lane is on the walk when it and each of its ancestors
// in the subtree is the chosen child of its parent and >= x: one ballot and
// a compare against the lane's ancestor mask, no per-level walk.
__device__ __forceinline__ void par_pop(ParHeap& H, int& len, const PopReads& r, const SubLanes& S) {
    const int L = len - 1;  // entries left; x was entry L
    len = L;
    if (L == 0) return;
    constexpr unsigned long long kValid = (1ull << 62) - 1, kRight = 0xAAAAAAAAAAAAAAAAull & kValid;
    const uint32_t xk = hkey(r.lk);
    uint32_t sk = r.sk, sp = r.sp;
    int hole = 0;
    for (bool first = true;; first = false) {
        const int idx = ((hole + 1) << S.d) - 1 + S.k;
        if (!first) {
            sk = 0;
            sp = 0;
            if (S.valid && idx < L) {
                sk = H.key[idx];
                sp = H.pay[idx];
            }
        }
        const uint32_t hk = hkey(sk);
        // the sibling's key (lane swap within pairs: DPP quad_perm [1, 0, 3, 2])
        const uint32_t hs = (uint32_t)__builtin_amdgcn_mov_dpp((int)hk, 0xB1, 0xF, 0xF, false);
        // chosen child: the right one unless it is smaller than the left
        const unsigned long long chosen = (__ballot(hk >= hs) & kRight) | (__ballot(hk > hs) & ~kRight);
        const unsigned long long M = chosen & __ballot(hk >= xk) & kValid;
        const unsigned long long P =
            __ballot(((uint32_t)M & S.am_lo) == S.am_lo) & __ballot(((uint32_t)(M >> 32) & S.am_hi) == S.am_hi);
        if ((P >> (threadIdx.x & 63)) & 1ull) {
            const int parent = (idx - 1) >> 1;
            H.key[parent] = sk;
            H.pay[parent] = sp;
        }
        if (P == 0) break;
        hole = __builtin_amdgcn_readlane(idx, 63 - __builtin_clzll(P));  // the deepest node of the walk
        if (__builtin_popcountll(P) < 5) break;
    }
    if ((threadIdx.x & 63) == 0) {
        H.key[hole] = r.lk;
        H.pay[hole] = r.lp;
    }
    lds_order();
}

// heap entry of node c from start[c], start[c+1], child[c], child[c+1]
__device__ __forceinline__ uint2 node_entry(uint32_t c, uint32_t s0, uint32_t s1, uint32_t c0, uint32_t c1) {
    const uint32_t nch = c1 - c0;
    return make_uint2(((s1 - s0) << 3) | nch, nch ? c0 : c);
}

__device__ __forceinline__ uint32_t lane_val(uint32_t v, int k) { return __builtin_amdgcn_readlane(v, k); }

// The order the keypoint kernels visit a level's slots in: the kept nodes
// (candidate words in `words`, LDS) by 16-row band of the level (then in any
// order within a band), seed and unused slots last.  Neighbouring keypoints
// are then processed by neighbouring waves, which share their disk and patch
// rows in L2.  Only the processing order: every slot's outputs stay where
// they are.  One wave; the LDS executes the wave's DS instructions in order.
__device__ __forceinline__ void slot_order(const OrbGlobals& g, const LevelDesc& L, int nkeep, const uint32_t* words) {
    const int lane = threadIdx.x & 63;
    __shared__ int s_band[256];
    for (int b = lane; b < 256; b += 64) s_band[b] = 0;
    lds_order();
    for (int jj = lane; jj < nkeep; jj += 64) atomicAdd(&s_band[(words[jj] >> 16) & 0xFF], 1);
    lds_order();
    {
        int c4[4], sum = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            c4[u] = s_band[4 * lane + u];
            sum += c4[u];
        }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        int run = incl - sum;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            s_band[4 * lane + u] = run;
            run += c4[u];
        }
    }
    lds_order();
    const int nslot = L.out_cap + L.seed_count;
    for (int jj = lane; jj < nslot; jj += 64) {
        int pos = jj;
        if (jj < nkeep) pos = atomicAdd(&s_band[(words[jj] >> 16) & 0xFF], 1);
        g.perm_out[L.kp_slot + pos] = L.kp_slot + jj;
    }
}

// a level without output: its slots in slot order
__device__ __forceinline__ void slot_order_identity(const OrbGlobals& g, const LevelDesc& L) {
    for (int j = threadIdx.x & 63; j < L.out_cap + L.seed_count; j += 64) g.perm_out[L.kp_slot + j] = L.kp_slot + j;
}



__global__ __launch_bounds__(64) void k_octree_heap(OrbGlobals g, const int* order, int cap, int tab_cap) {
    extern __shared__ uint32_t heap_lds[];
    ParHeap H{heap_lds, heap_lds + cap};
    const SubLanes SL = sub_lanes();
    const int lane = threadIdx.x;
    const int li = order[blockIdx.x];
    const LevelDesc L = g.lv[li];
    OctLevel* O = g.oct + li;
    if (O->failed || O->n == 0) {
        if (lane == 0) O->nout = 0;
        slot_order_identity(g, L);
        return;
    }
    NodeTabs nt;
    nt.start = g.pool + O->tab;
    nt.child = nt.start + O->total;
    const TabPrefix TP = tab_prefix_load(nt, heap_lds + 2 * cap, tab_cap, (uint32_t)O->total);
    const int nr = O->n_roots;
    int len = 0;
    {
        uint32_t s_l = 0, c_l = 0;
        if (lane <= nr) {
            s_l = nt.st(lane);
            c_l = nt.ch(lane);
        }
        for (int k = 0; k < nr; k++) {
            const uint2 e = node_entry((uint32_t)k, lane_val(s_l, k), lane_val(s_l, k + 1), lane_val(c_l, k),
                                       lane_val(c_l, k + 1));
            par_push(H, len, e.x, e.y);
        }
    }
    const int N = min(L.nfeat, cap - 8);
    while (len < N && len > 0) {
        const PopReads pr = par_pop_reads(H, len);
        const uint32_t tk = pr.tk, tp = pr.tp;
        if (hkey(tk) == 1) break;
        const uint32_t c0 = tp, nch = tk & 7u;
        uint32_t s_l, c_l;
        // the children's table entries are requested before the pop and used
        // after it: their LDS latency passes during the pop's first round,
        // which needs nothing new from LDS
        tab_children(nt, TP, c0, nch, s_l, c_l);
        par_pop(H, len, pr, SL);
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            if (k < nch) {
                const uint2 e = node_entry(c0 + k, lane_val(s_l, k), lane_val(s_l, k + 1), lane_val(c_l, k),
                                           lane_val(c_l, k + 1));
                par_push(H, len, e.x, e.y);
            }
        }
    }
    // drain: the j-th popped node's word goes to key[len] once the pop has
    // freed that slot, i.e. node j ends at key[n0 - 1 - j]
    const int n0 = len;
    while (len > 0) {
        const PopReads pr = par_pop_reads(H, len);
        const uint32_t w = (pr.tp << 3) | (pr.tk & 7u);
        par_pop(H, len, pr, SL);
        if (lane == 0) H.key[len] = w;
    }
    wave_lds_sync();
    const int nkeep = min(n0, L.out_cap);
    if (lane == 0) O->nout = nkeep;
    // each kept node's keypoint, one node per lane: the candidate with the
    // largest response, the first in candidate order among equals
    // (DistributeOctTree's scan, ORBextractor.cc:569-574); its candidate
    // word (response << 24 | y << 12 | x) is the slot's output
    const uint32_t* sorted = g.pool + O->sorted;
    const uint32_t* cand = g.lvl_cand + L.cand_base;
    for (int jj = lane; jj < nkeep; jj += 64) {
        const uint32_t nw = H.key[n0 - 1 - jj];
        const uint32_t ny = nw >> 3, nch = max(nw & 7u, 1u);
        const uint32_t s0 = nt.st(ny), s1 = nt.st(ny + nch);
        uint32_t best = 0, bi = 0;
        for (uint32_t p = s0; p < s1; p++) {
            const uint32_t idx = sorted[p];
            const uint32_t r = cand[idx];
            if (p == s0 || (r >> 24) > (best >> 24) || ((r >> 24) == (best >> 24) && idx < bi)) {
                best = r;
                bi = idx;
            }
        }
        g.nodes_out[L.kp_slot + jj] = best;
        H.pay[jj] = best;  // the drained heap's payload array is free
    }
    slot_order(g, L, nkeep, H.pay);
}

// ---------------------------------------------------------------------------
// The same replay with the top of the heap in registers.  Nodes are numbered
// from 1 (children of node j: 2j and 2j + 1; libstdc++'s index is j - 1).
// Nodes 1..63, the six levels every sift passes through, live in two VGPRs
// of the wave (lane j holds node j; an absent node reads as key 0); deeper
// nodes live in LDS.  The sifts are the formulations of the LDS replay (a
// pop walks the max-child path down while the next value is >= the last
// element x, each such value moving up one level, x where the walk stops;
// a push moves the run of ancestors smaller than the value down one level),
// found lane-parallel in the registers:
//   pop: every lane of the top tests "chosen child of its parent and >= x"
//     (siblings swapped by DPP), one ballot, and a lane is on the walk when
//     that holds for it and every ancestor below the root (its precomputed
//     ancestor mask); the walk's values move up with one lane shuffle.  A
//     walk that reaches depth 5 continues through 5-level LDS subtree gathers
//     as in the LDS replay.
//   push: the chain of ancestors is a lane predicate; the run smaller than
//     the value is one ballot, moved down with one lane shuffle (and, below
//     depth 5, the LDS part gathered and stored as in the LDS replay).
// A heap of fewer than 64 nodes never touches LDS.  LDS holds 8 bytes per
// heap entry, as the LDS variant.
__device__ __forceinline__ uint32_t rdl(uint32_t v, int j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}
// lane j of v takes the (wave-uniform) value x: a compare and a select
__device__ __forceinline__ uint32_t wrl(uint32_t v, uint32_t x, int j) {
    return (int)(threadIdx.x & 63) == j ? x : v;
}
__device__ __forceinline__ int log2i(int j) { return 31 - __builtin_clz((uint32_t)j); }

struct RegHeap {
    uint32_t tk, tp;  // nodes 1..63: lane j holds node j (key, payload); key 0: absent
    uint32_t* k;      // LDS keys, indexed by node (nodes >= 64)
    uint32_t* p;      // LDS payloads
    int len;
    uint32_t tam_lo, tam_hi;  // this lane's node and its ancestors below the root, as lane bits
};

__device__ __forceinline__ void rh_init(RegHeap& H) {
    const int j = threadIdx.x & 63;
    unsigned long long m = 1ull;  // lanes 0, 1 hold no walk node: bit 0 is never set in a walk ballot
    if (j >= 2) {
        m = 0;
        for (int q = j; q >= 2; q >>= 1) m |= 1ull << q;
    }
    H.tam_lo = (uint32_t)m;
    H.tam_hi = (uint32_t)(m >> 32);
    H.tk = 0;
    H.tp = 0;
    H.len = 0;
}

// pop_heap; the top comes back in (ok, op)
__device__ __forceinline__ void rh_pop(RegHeap& H, const SubLanes& S, uint32_t& ok, uint32_t& op) {
    constexpr unsigned long long kTop = ~3ull;  // lanes 2..63
    constexpr unsigned long long kRight = 0xAAAAAAAAAAAAAAAAull;
    const int lane = threadIdx.x & 63;
    ok = rdl(H.tk, 1);
    op = rdl(H.tp, 1);
    const int n = H.len;
    uint32_t xk, xp;
    if (n < 64) {
        xk = rdl(H.tk, n);
        xp = rdl(H.tp, n);
        H.tk = lane == n ? 0u : H.tk;
    } else {
        xk = __builtin_amdgcn_readfirstlane(H.k[n]);
        xp = __builtin_amdgcn_readfirstlane(H.p[n]);
    }
    const int L = n - 1;
    H.len = L;
    if (L == 0) return;
    const uint32_t hx = hkey(xk);
    // the walk through the top
    const uint32_t hk = hkey(H.tk);
    const uint32_t hs = (uint32_t)__builtin_amdgcn_mov_dpp((int)hk, 0xB1, 0xF, 0xF, false);
    const unsigned long long M =
        ((__ballot(hk >= hs) & kRight) | (__ballot(hk > hs) & ~kRight)) & __ballot(hk >= hx) & kTop;
    const unsigned long long P =
        __ballot(((uint32_t)M & H.tam_lo) == H.tam_lo) & __ballot(((uint32_t)(M >> 32) & H.tam_hi) == H.tam_hi);
    // lane q takes the value of its child on the walk
    int src = lane;
    if (lane < 32) {
        if ((P >> (2 * lane + 1)) & 1ull) src = 2 * lane + 1;
        else if ((P >> (2 * lane)) & 1ull) src = 2 * lane;
    }
    H.tk = (uint32_t)__shfl((int)H.tk, src, 64);
    H.tp = (uint32_t)__shfl((int)H.tp, src, 64);
    int hole = P ? 63 - __builtin_clzll(P) : 1;
    if (__builtin_popcountll(P) == 5 && 2 * hole <= L) {
        // the walk goes on below depth 5: 5-level subtrees from LDS
        for (bool first = true;; first = false) {
            const int idx = (hole << S.d) + S.k;
            uint32_t sk = 0, sp = 0;
            if (S.valid && idx <= L) {
                sk = H.k[idx];
                sp = H.p[idx];
            }
            const uint32_t h2 = hkey(sk);
            const uint32_t s2 = (uint32_t)__builtin_amdgcn_mov_dpp((int)h2, 0xB1, 0xF, 0xF, false);
            constexpr unsigned long long kValid = (1ull << 62) - 1, kR = 0xAAAAAAAAAAAAAAAAull & kValid;
            const unsigned long long M2 =
                ((__ballot(h2 >= s2) & kR) | (__ballot(h2 > s2) & ~kR)) & __ballot(h2 >= hx) & kValid;
            const unsigned long long P2 =
                __ballot(((uint32_t)M2 & S.am_lo) == S.am_lo) & __ballot(((uint32_t)(M2 >> 32) & S.am_hi) == S.am_hi);
            if (P2 == 0) break;
            if (first) {
                // the depth-1 node of the walk moves into the register node `hole`
                const int l1 = (P2 & 1ull) ? 0 : 1;
                const uint32_t vk = rdl(sk, l1), vp = rdl(sp, l1);
                H.tk = wrl(H.tk, vk, hole);
                H.tp = wrl(H.tp, vp, hole);
                if (((P2 >> lane) & 1ull) && S.d > 1) {
                    H.k[idx >> 1] = sk;
                    H.p[idx >> 1] = sp;
                }
            } else if ((P2 >> lane) & 1ull) {
                H.k[idx >> 1] = sk;
                H.p[idx >> 1] = sp;
            }
            lds_order();
            hole = __builtin_amdgcn_readlane(idx, 63 - __builtin_clzll(P2));
            if (__builtin_popcountll(P2) < 5) break;
        }
    }
    if (hole < 64) {
        H.tk = wrl(H.tk, xk, hole);
        H.tp = wrl(H.tp, xp, hole);
    } else if (lane == 0) {
        H.k[hole] = xk;
        H.p[hole] = xp;
    }
    lds_order();
}

// push_heap of (vk, vp)
__device__ __forceinline__ void rh_push(RegHeap& H, uint32_t vk, uint32_t vp) {
    const int lane = threadIdx.x & 63;
    const int h = ++H.len;
    const int D = log2i(h);
    const uint32_t hv = hkey(vk);
    const int lj = lane >= 1 ? log2i(lane) : 0;
    // lane j is an ancestor of h in the top
    const bool anc = lane >= 1 && lj < D && lj <= 5 && (h >> (D - lj)) == lane;
    if (h < 64) {
        const unsigned long long less = __ballot(anc && hkey(H.tk) < hv);
        const int m = __builtin_popcountll(less);
        const int src = (lane >= 2 && (anc || lane == h) && ((less >> (lane >> 1)) & 1ull)) ? lane >> 1 : lane;
        H.tk = (uint32_t)__shfl((int)H.tk, src, 64);
        H.tp = (uint32_t)__shfl((int)H.tp, src, 64);
        H.tk = wrl(H.tk, vk, h >> m);
        H.tp = wrl(H.tp, vp, h >> m);
        return;
    }
    // LDS ancestors: levels 1 .. nl (lane k: level k + 1, node h >> (k + 1))
    const int nl = D - 6;
    uint32_t ck = 0xFFFFFFFFu, cp = 0;
    if (lane < nl) {
        ck = H.k[h >> (lane + 1)];
        cp = H.p[h >> (lane + 1)];
    }
    const int ml = __builtin_popcountll(__ballot(lane < nl && hkey(ck) < hv));
    int m = ml;
    unsigned long long lesst = 0;
    if (ml == nl) {
        lesst = __ballot(anc && hkey(H.tk) < hv);
        m += __builtin_popcountll(lesst);
    }
    // LDS moves: the ancestor at level k + 1 into its chain child (level k)
    if (lane < min(m, nl)) {
        H.k[h >> lane] = ck;
        H.p[h >> lane] = cp;
    }
    if (lesst) {
        // the depth-5 ancestor moves into its depth-6 chain child (LDS), the
        // smaller ones above it down one level in the registers
        const int a5 = h >> (nl + 1);
        if (lane == a5) {
            H.k[h >> nl] = H.tk;
            H.p[h >> nl] = H.tp;
        }
        const int src = (lane >= 2 && anc && ((lesst >> (lane >> 1)) & 1ull)) ? lane >> 1 : lane;
        H.tk = (uint32_t)__shfl((int)H.tk, src, 64);
        H.tp = (uint32_t)__shfl((int)H.tp, src, 64);
    }
    const int dst = h >> m;
    if (dst < 64) {
        H.tk = wrl(H.tk, vk, dst);
        H.tp = wrl(H.tp, vp, dst);
    } else if (lane == 0) {
        H.k[dst] = vk;
        H.p[dst] = vp;
    }
    lds_order();
}

__global__ __launch_bounds__(64) void k_octree_heap_r(OrbGlobals g, const int* order, int cap, int tab_cap) {
    extern __shared__ uint32_t heap_lds[];
    RegHeap H;
    rh_init(H);
    H.k = heap_lds;
    H.p = heap_lds + cap;
    const SubLanes SL = sub_lanes();
    const int lane = threadIdx.x;
    const int li = order[blockIdx.x];
    const LevelDesc L = g.lv[li];
    OctLevel* O = g.oct + li;
    if (O->failed || O->n == 0) {
        if (lane == 0) O->nout = 0;
        slot_order_identity(g, L);
        return;
    }
    NodeTabs nt;
    nt.start = g.pool + O->tab;
    nt.child = nt.start + O->total;
    const uint32_t T = (uint32_t)O->total;
    const uint32_t ncand = (uint32_t)O->n;
    const TabPrefix TP = tab_prefix_load(nt, heap_lds + 2 * cap, tab_cap, T);
    const int nr = O->n_roots;
    {
        uint32_t s_l = 0, c_l = 0;
        if (lane <= nr) {
            s_l = nt.st(lane);
            c_l = nt.ch(lane);
        }
        for (int k = 0; k < nr; k++) {
            const uint2 e = node_entry((uint32_t)k, lane_val(s_l, k), lane_val(s_l, k + 1), lane_val(c_l, k),
                                       lane_val(c_l, k + 1));
            rh_push(H, e.x, e.y);
        }
    }
    // node indices stay below N + 4 <= cap - 4: the split loop runs while
    // len < N and each step pops one node and pushes at most four
    const int N = min(L.nfeat, cap - 8);
    while (H.len < N && H.len > 0) {
        const uint32_t tk = rdl(H.tk, 1), tpay = rdl(H.tp, 1);
        if (hkey(tk) == 1) break;
        // table reads stay inside the level's tables whatever the entry holds
        const uint32_t nch = tk & 7u, c0 = min(tpay, T - 1 - min(nch, T - 1));
        uint32_t s_l, c_l;
        tab_children(nt, TP, c0, nch, s_l, c_l);
        uint32_t ek[4], ep[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint2 e = node_entry(c0 + k, lane_val(s_l, k), lane_val(s_l, k + 1), lane_val(c_l, k),
                                       lane_val(c_l, k + 1));
            ek[k] = e.x;
            ep[k] = e.y;
        }
        uint32_t ok, op;
        rh_pop(H, SL, ok, op);
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (k < nch) rh_push(H, ek[k], ep[k]);
    }
    // drain: the j-th popped node's word goes to node n0 - j, the node the
    // pop has just freed (a register lane below 64, LDS above)
    const int n0 = H.len;
    uint32_t tw = 0;
    while (H.len > 0) {
        uint32_t ok, op;
        rh_pop(H, SL, ok, op);
        const uint32_t w = (op << 3) | (ok & 7u);
        const int n = H.len + 1;
        if (n < 64) tw = wrl(tw, w, n);
        else if (lane == 0) H.k[n] = w;
        lds_order();
    }
    const int nkeep = min(n0, L.out_cap);
    if (lane == 0) O->nout = nkeep;
    const uint32_t* sorted = g.pool + O->sorted;
    const uint32_t* cand = g.lvl_cand + L.cand_base;
    for (int b = 0; b < nkeep; b += 64) {
        const int jj = b + lane;
        const int n = n0 - jj;
        const uint32_t wr = (uint32_t)__shfl((int)tw, (n >= 1 && n < 64) ? n : 0, 64);
        if (jj < nkeep) {
            const uint32_t nw = n < 64 ? wr : H.k[n];
            const uint32_t ny = min(nw >> 3, T - 1), nch = max(nw & 7u, 1u);
            const uint32_t s0 = min(nt.st(ny), ncand), s1 = min(nt.st(min(ny + nch, T - 1)), ncand);
            uint32_t best = 0, bi = 0;
            for (uint32_t p = s0; p < s1; p++) {
                const uint32_t idx = min(sorted[p], ncand - 1);
                const uint32_t r = cand[idx];
                if (p == s0 || (r >> 24) > (best >> 24) || ((r >> 24) == (best >> 24) && idx < bi)) {
                    best = r;
                    bi = idx;
                }
            }
            g.nodes_out[L.kp_slot + jj] = best;
            H.p[jj] = best;
        }
    }
    slot_order(g, L, nkeep, H.p);
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
constexpr int kPatchR = 18;                           // test points lie within 18 px of the centre
constexpr int kPatchRows = 2 * kPatchR + 1;            // 37
constexpr int kPatchWords = 10;                        // u32 per staged row (aligned column below cx - 18)
constexpr int kPatchIt = (kPatchRows * kPatchWords + 63) / 64;
static_assert(kPatchWords == 10 && kPatchIt * 64 <= 384, "patch word division by 6554 / 2^16");

__global__ __launch_bounds__(256) void k_keypoints(OrbGlobals g, const int* slot_lvl, int n_slots,
                                                   SinCosTables tb, float* out_kps, uint8_t* out_desc, int cap,
                                                   int* out_count) {
    __shared__ uint32_t s_patch[4][kPatchRows * kPatchWords];
    const int xb = xcd_block((n_slots + 3) / 4);
    if (xb < 0) return;
    const int wv = xb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wv >= n_slots) return;
    const int wave = g.perm ? __builtin_amdgcn_readfirstlane(g.perm[wv]) : wv;
    // slot `wave` is (level li, j) with L.kp_slot + j == wave, so the heap's
    // winner word for it is requested together with the slot entry
    const int2 sl = reinterpret_cast<const int2*>(slot_lvl)[wave];
    const uint32_t win = g.nodes_out[wave];
    const int li = sl.x, j = sl.y;
    const LevelDesc L = g.lv[li];
    // keypoints per level of this image (octree output + seeds), one lane per level
    int cnt_l = 0;
    if (lane < L.nlevels) cnt_l = g.oct[L.first_level + lane].nout + g.lv[L.first_level + lane].seed_count;
    const int nout = __shfl(cnt_l, L.level) - L.seed_count;
    if (j == 0 && L.level == 0) {
        const int tot = wave_sum(cnt_l);
        if (lane == 0) out_count[L.img] = tot;
    }
    const int off = wave_sum(lane < L.level ? cnt_l : 0);
    if (j >= nout + L.seed_count) return;
    const int pos = off + j;
    if (pos >= cap) return;
    float x, y, resp, angle, size;
    const bool seed = j >= nout;
    if (!seed) {
        // the slot's keypoint as the heap kernel left it: response << 24 | y << 12 | x
        const uint32_t r = win;
        x = (float)((int)(r & 0xFFF) + L.minBX);
        y = (float)((int)((r >> 12) & 0xFFF) + L.minBY);
        resp = (float)(r >> 24);
        size = L.patch;
    } else {
        const float* sd = g.seeds + 6 * (L.seed_first + (j - nout));
        x = sd[0];
        y = sd[1];
        size = sd[2];
        angle = sd[3];
        resp = sd[4];
    }
    const int cx = (int)rintf(x), cy = (int)rintf(y);
    // The rotated test points lie within 18 px of the centre (the pattern's
    // largest radius is 18.38).  When that square is inside the level, the
    // blurred 37 x 40-byte patch around it is staged in LDS with coalesced
    // u32 loads issued together with the moment loads below, and the 512
    // taps are LDS reads; otherwise (seeds near the border) the taps are
    // clamped gathers from the level.
    const bool staged = cx >= kPatchR && cx + kPatchR < L.w && cy >= kPatchR && cy + kPatchR < L.h;
    const uint8_t* bl = g.blur + L.pix;
    const int bx = (cx - kPatchR) & ~3;
    uint32_t pw[kPatchIt];
    if (staged) {
#pragma unroll
        for (int it = 0; it < kPatchIt; it++) {
            const int k = lane + 64 * it, r = k / kPatchWords, w = k - kPatchWords * r;
            pw[it] = 0;
            if (k < kPatchRows * kPatchWords && bx + 4 * w + 4 <= L.pitch)
                pw[it] = *reinterpret_cast<const uint32_t*>(bl + (long long)(cy - kPatchR + r) * L.pitch + bx + 4 * w);
        }
    }
    if (!seed) {
        // IC_Angle over the circular patch (|u| <= umax[|v|], |v| <= 15): the
        // 31 rows are read as u32 words from the aligned column below cx - 15
        // (at most 9 words a row, 279 words over the wave); the bytes of a
        // word inside the disk are masked and summed with two udot4 (sum of
        // values, sum of byte index times value); the integer moments are
        // order independent
        const int base = (cx - 15) & ~3;
        const int nwr = (cx + 15 - base) / 4 + 1;
        unsigned long long um = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) um |= (unsigned long long)(g.umax[k] & 15) << (4 * k);
        const uint8_t* lvl = g.pyr + L.pix;
        int m10 = 0, m01 = 0;
#pragma unroll
        for (int it = 0; it < 5; it++) {
            const int t = lane + 64 * it;
            const int row = t / 9, w = t - 9 * row;
            if (t < 31 * 9 && w < nwr) {
                const int v = row - 15;
                const int du = (int)((um >> (4 * (v < 0 ? -v : v))) & 15u);
                const uint32_t dw =
                    *reinterpret_cast<const uint32_t*>(lvl + (long long)(cy + v) * L.pitch + base + 4 * w);
                const int u0 = base + 4 * w - cx;  // u of byte 0
                const int blo = max(-du - u0, 0), bhi = min(du - u0, 3);
                if (blo <= bhi) {
                    const uint32_t m = dw & (0xFFFFFFFFu << (8 * blo)) & (0xFFFFFFFFu >> (8 * (3 - bhi)));
                    const int s1 = (int)__builtin_amdgcn_udot4(m, 0x01010101u, 0u, false);
                    const int sb = (int)__builtin_amdgcn_udot4(m, 0x03020100u, 0u, false);
                    m10 += u0 * s1 + sb;
                    m01 += v * s1;
                }
            }
        }
        m10 = wave_sum(m10);
        m01 = wave_sum(m01);
        angle = fast_atan2_deg((float)m01, (float)m10);
    }
    uint32_t* patch = s_patch[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    if (staged) {
#pragma unroll
        for (int it = 0; it < kPatchIt; it++) {
            const int k = lane + 64 * it;
            if (k < kPatchRows * kPatchWords) patch[k] = pw[it];
        }
    }
    // rBRIEF on the blurred level.  Only the rounded pattern offsets
    // cvRound(px * b + py * a), cvRound(px * a - py * b) depend on the
    // rotation, so the correctly rounded base sin/cos (at most 1 ulp from the
    // host libm's) gives the same offsets unless some product sum lies within
    // 1e-5 of a rounding boundary (a 1-ulp change moves it by < 5e-6); the
    // wave then redoes the offsets with the libm-patched values.
    const float factorPI = (float)(3.14159265358979323846 / 180.f);
    const float xr = angle * factorPI;
    float a, b;
    sincos_base(fabsf(xr), &a, &b);
    if (xr < 0.f) b = -b;
    int oy[8], ox[8];
    bool near = false;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const signed char* p = kOrbPattern[64 * (k >> 1) + lane] + 2 * (k & 1);
        const float px = (float)p[0], py = (float)p[1];
        const float vy = px * b + py * a, vx = px * a - py * b;
        const float ry = rintf(vy), rx = rintf(vx);
        oy[k] = (int)ry;
        ox[k] = (int)rx;
        // within 1e-5 of a half-integer: |v - rint(v)| is exact and <= 0.5
        near = near || fabsf(vy - ry) > 0.5f - 1e-5f || fabsf(vx - rx) > 0.5f - 1e-5f;
    }
    if (__any(near)) {
        sincos_patched(xr, tb, &a, &b);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const signed char* p = kOrbPattern[64 * (k >> 1) + lane] + 2 * (k & 1);
            const float px = (float)p[0], py = (float)p[1];
            oy[k] = (int)rintf(px * b + py * a);
            ox[k] = (int)rintf(px * a - py * b);
        }
    }
    unsigned long long words[4];
    if (staged) {
        wave_lds_sync();
        const uint8_t* pb = reinterpret_cast<const uint8_t*>(patch);
        const int c0 = cx - bx;  // patch column of the centre
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int t[2];
#pragma unroll
            for (int q = 0; q < 2; q++)
                t[q] = pb[(kPatchR + oy[2 * r + q]) * (4 * kPatchWords) + c0 + ox[2 * r + q]];
            words[r] = __ballot(t[0] < t[1]);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int t[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int yy = min(max(cy + oy[2 * r + q], 0), L.h - 1);
                const int xx = min(max(cx + ox[2 * r + q], 0), L.w - 1);
                t[q] = bl[(long long)yy * L.pitch + xx];
            }
            words[r] = __ballot(t[0] < t[1]);
        }
    }
    if (lane == 0) {
        unsigned long long* dst = reinterpret_cast<unsigned long long*>(out_desc + ((long long)L.img * cap + pos) * 32);
        dst[0] = words[0];
        dst[1] = words[1];
        dst[2] = words[2];
        dst[3] = words[3];
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


// Two integer sums over the wave: quad, half-row and row sums by DPP moves
// (the two sums' steps interleaved, so no DPP move waits on the one before
// it), then the four row sums read out to scalars.  Every lane must be active.
__device__ __forceinline__ void wave_sum2_dpp(int& a, int& b) {
    a += __builtin_amdgcn_update_dpp(0, a, 0xB1, 0xF, 0xF, false);
    b += __builtin_amdgcn_update_dpp(0, b, 0xB1, 0xF, 0xF, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x4E, 0xF, 0xF, false);
    b += __builtin_amdgcn_update_dpp(0, b, 0x4E, 0xF, 0xF, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x141, 0xF, 0xF, false);
    b += __builtin_amdgcn_update_dpp(0, b, 0x141, 0xF, 0xF, false);
    a += __builtin_amdgcn_update_dpp(0, a, 0x140, 0xF, 0xF, false);
    b += __builtin_amdgcn_update_dpp(0, b, 0x140, 0xF, 0xF, false);
    a = __builtin_amdgcn_readlane(a, 0) + __builtin_amdgcn_readlane(a, 16) + __builtin_amdgcn_readlane(a, 32) +
        __builtin_amdgcn_readlane(a, 48);
    b = __builtin_amdgcn_readlane(b, 0) + __builtin_amdgcn_readlane(b, 16) + __builtin_amdgcn_readlane(b, 32) +
        __builtin_amdgcn_readlane(b, 48);
}

__device__ __forceinline__ float readlane_f(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// k_keypoints with K keypoint slots per wave.  The per-keypoint set-up
// (slot entry, level prefix, coordinates), the angle (fastAtan2 of the
// moments) and the rotation (host-libm cosf/sinf: the base values patched
// from the exception table) run lane-parallel, lane k for the wave's k-th
// slot; the moments and the descriptors, which spread one keypoint over the
// wave, then walk the K keypoints with their operands read from lane k and
// the next keypoint's loads issued before the current one is reduced.
// Same arithmetic and outputs as k_keypoints (ORBextractor.cc:75-141,
// :568-578, :845-896).
// Octree heap replay: the top of the heap in registers (k_octree_heap_r)
// or the whole heap in LDS (k_octree_heap); the same pop order.
constexpr int kHeapVariants = 2;
inline const char* heap_variant_name(int v) { return v == 1 ? "lds-heap" : "register-top"; }

constexpr int kKpVariants = 4;  // keypoints per wave: 1 (k_keypoints), 8, 16, 32
constexpr int kKpGroup = 256;   // waves that share slot windows in k_keypoints_k
inline const char* kp_variant_name(int v) {
    return v == 1 ? "8/wave" : v == 2 ? "16/wave" : v == 3 ? "32/wave" : "1/wave";
}

template <int K>
__global__ __launch_bounds__(256, 6) void k_keypoints_k(OrbGlobals g, const int* slot_lvl, int n_slots,
                                                     SinCosTables tb, float* out_kps, uint8_t* out_desc, int cap,
                                                     int* out_count) {
    static_assert(K >= 1 && K <= 64, "one lane per keypoint of the wave");
    __shared__ uint32_t s_patch[4][kPatchRows * kPatchWords];
    const int nwg = (n_slots + 4 * K - 1) / (4 * K);
    const int xb = xcd_block(nwg);
    if (xb < 0) return;
    const int lane = threadIdx.x & 63;
    // Slots are dealt out across groups of kKpGroup consecutive waves (one
    // XCD's share of the grid is contiguous): wave w of a group of gs waves
    // takes slots base + w + k * gs, so the waves of a group work on a
    // contiguous window of slots at each step k and their patches share L2.
    const int wv = xb * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sg = wv / kKpGroup, gs = min(kKpGroup, 4 * nwg - sg * kKpGroup);
    const int sbase = sg * kKpGroup * K + (wv - sg * kKpGroup);
    if (sbase >= n_slots) return;
    // ---- set-up: lane k < K takes slot sbase + k * gs
    bool kvalid = false, kseed = false;
    int kimg = 0, klevel = 0, kpos = 0, kcx = 0, kcy = 0, kw = 0, kh = 0, kpitch = 0;
    long long kpix = 0;
    float kx = 0.f, ky = 0.f, kresp = 0.f, ksize = 0.f, kangle = 0.f, kscale = 1.f;
    const int pslot = sbase + lane * gs;
    const int kslot = (g.perm && pslot < n_slots) ? g.perm[pslot] : pslot;
    if (lane < K && pslot < n_slots) {
        const int2 sl = reinterpret_cast<const int2*>(slot_lvl)[kslot];
        const uint32_t win = g.nodes_out[kslot];
        const LevelDesc* Lp = g.lv + sl.x;
        const int j = sl.y;
        kimg = Lp->img;
        klevel = Lp->level;
        const int first = Lp->first_level, nl = Lp->nlevels;
        // keypoints per level of this image (octree output + seeds)
        int off = 0, tot = 0;
#pragma unroll
        for (int l = 0; l < kMaxLevels; l++)
            if (l < nl) {
                const int c = g.oct[first + l].nout + g.lv[first + l].seed_count;
                off += l < klevel ? c : 0;
                tot += c;
            }
        if (j == 0 && klevel == 0) out_count[kimg] = tot;
        const int nout = g.oct[sl.x].nout, nseed = Lp->seed_count;
        kpos = off + j;
        kvalid = j < nout + nseed && kpos < cap;
        if (kvalid) {
            kseed = j >= nout;
            if (!kseed) {
                // the slot's keypoint as the heap kernel left it: response << 24 | y << 12 | x
                kx = (float)((int)(win & 0xFFF) + Lp->minBX);
                ky = (float)((int)((win >> 12) & 0xFFF) + Lp->minBY);
                kresp = (float)(win >> 24);
                ksize = Lp->patch;
            } else {
                const float* sd = g.seeds + 6 * (Lp->seed_first + (j - nout));
                kx = sd[0];
                ky = sd[1];
                ksize = sd[2];
                kangle = sd[3];
                kresp = sd[4];
            }
            kcx = (int)rintf(kx);
            kcy = (int)rintf(ky);
            kw = Lp->w;
            kh = Lp->h;
            kpitch = Lp->pitch;
            kpix = Lp->pix;
            kscale = Lp->scale;
        }
    }
    const int kpix_lo = (int)(uint32_t)kpix, kpix_hi = (int)(kpix >> 32);
    auto pix_of = [&](int k) -> long long {
        return (long long)(((unsigned long long)(uint32_t)__builtin_amdgcn_readlane(kpix_hi, k) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane(kpix_lo, k));
    };

    // ---- IC_Angle moments (ORBextractor.cc:75-98) of every octree keypoint.
    // The disk's 31 rows are read as u32 words from the aligned column below
    // cx - 15 (at most 9 words a row, 279 words over the wave); a lane's
    // (row, word) pairs do not depend on the keypoint.
    // Per (row, word) pair the bytes inside the disk depend only on the
    // alignment c = (cx - 15) & 3 of the word column: u of byte 0 is
    // 4 w - 15 - c.  mpk[it] holds the 4-bit byte masks of the four
    // alignments (nibble c), so a keypoint's masks are one shift away.
    int mv[5], mw[5], mw4[5];
    uint32_t mpk[5];
    {
        unsigned long long um = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) um |= (unsigned long long)(g.umax[q] & 15) << (4 * q);
#pragma unroll
        for (int it = 0; it < 5; it++) {
            const int t = lane + 64 * it;
            const int row = t / 9;
            mv[it] = row - 15;
            mw[it] = t < 31 * 9 ? t - 9 * row : 99;  // 99: no word
            mw4[it] = 4 * mw[it] - 15;
            const int av = mv[it] < 0 ? -mv[it] : mv[it];
            const int du = (int)((um >> (4 * min(av, 15))) & 15u);
            uint32_t pk = 0;
            if (mw[it] != 99) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int u0 = mw4[it] - c;
                    const int blo = max(-du - u0, 0), bhi = min(du - u0, 3);
                    const uint32_t nib = blo <= bhi ? ((0xFu << blo) & (0xFu >> (3 - bhi))) : 0u;
                    pk |= nib << (4 * c);
                }
            }
            mpk[it] = pk;
        }
    }
    int M10 = 0, M01 = 0;
    unsigned long long todo = __ballot(kvalid && !kseed);
    if (todo) {
        uint32_t dw[5];
        auto mload = [&](int k, uint32_t (&d)[5]) {
            const int cx = __builtin_amdgcn_readlane(kcx, k), cy = __builtin_amdgcn_readlane(kcy, k);
            const int pitch = __builtin_amdgcn_readlane(kpitch, k);
            const uint8_t* lvl = g.pyr + pix_of(k);
            const int base = (cx - 15) & ~3;
            const int nwr = (cx + 15 - base) / 4 + 1;
#pragma unroll
            for (int it = 0; it < 5; it++)
                d[it] = mw[it] < nwr ? *reinterpret_cast<const uint32_t*>(
                                           lvl + (uint32_t)(cy * pitch + __mul24(mv[it], pitch) + base + 4 * mw[it]))
                                     : 0u;
        };
        int k = __builtin_ctzll(todo);
        mload(k, dw);
        for (;;) {
            todo &= todo - 1;
            const int cx = __builtin_amdgcn_readlane(kcx, k);
            uint32_t cur[5];
#pragma unroll
            for (int it = 0; it < 5; it++) cur[it] = dw[it];
            const int kn = todo ? __builtin_ctzll(todo) : -1;
            if (kn >= 0) mload(kn, dw);
            const int c = (cx - 15) & 3;  // wave-uniform
            int m10 = 0, m01 = 0;
#pragma unroll
            for (int it = 0; it < 5; it++) {
                const int u0 = mw4[it] - c;  // u of byte 0
                // the disk's bytes of the word as 0/1 weights, and as weights b
                const uint32_t nib = (mpk[it] >> (4 * c)) & 15u;
                const uint32_t w1 = __umul24(nib, 0x204081u) & 0x01010101u;
                const uint32_t wb = (w1 * 255u) & 0x03020100u;
                const int s1 = (int)__builtin_amdgcn_udot4(cur[it], w1, 0u, false);
                const int sb = (int)__builtin_amdgcn_udot4(cur[it], wb, 0u, false);
                m10 += __mul24(u0, s1) + sb;
                m01 += __mul24(mv[it], s1);
            }
            wave_sum2_dpp(m10, m01);
            if (lane == k) {
                M10 = m10;
                M01 = m01;
            }
            if (kn < 0) break;
            k = kn;
        }
    }

    // ---- angle, rotation and the keypoint record, lane k for keypoint k
    const float factorPI = (float)(3.14159265358979323846 / 180.f);
    float ka = 0.f, kb = 0.f;
    if (kvalid) {
        if (!kseed) kangle = fast_atan2_deg((float)M01, (float)M10);
        sincos_patched(kangle * factorPI, tb, &ka, &kb);
        float* o = out_kps + ((long long)kimg * cap + kpos) * 6;
        o[0] = klevel != 0 ? kx * kscale : kx;
        o[1] = klevel != 0 ? ky * kscale : ky;
        o[2] = ksize;
        o[3] = kangle;
        o[4] = kresp;
        reinterpret_cast<int*>(o)[5] = klevel;
    }

    // ---- rBRIEF (ORBextractor.cc:101-141) on the blurred level: the wave's
    // 64 lanes take 4 test pairs each; a keypoint whose 18-px square lies in
    // the level has its 37 x 40-byte patch staged in LDS (the next
    // keypoint's patch loads are issued before the current taps), the others
    // gather clamped taps from the level.
    float ppx[8], ppy[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const signed char* p = kOrbPattern[64 * (q >> 1) + lane] + 2 * (q & 1);
        ppx[q] = (float)p[0];
        ppy[q] = (float)p[1];
    }
    uint32_t* patch = s_patch[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    todo = __ballot(kvalid);
    if (todo) {
        uint32_t pw[kPatchIt];
        auto staged_of = [&](int k) {
            const int cx = __builtin_amdgcn_readlane(kcx, k), cy = __builtin_amdgcn_readlane(kcy, k);
            return cx >= kPatchR && cx + kPatchR < __builtin_amdgcn_readlane(kw, k) && cy >= kPatchR &&
                   cy + kPatchR < __builtin_amdgcn_readlane(kh, k);
        };
        auto pload = [&](int k) {
            const int cx = __builtin_amdgcn_readlane(kcx, k), cy = __builtin_amdgcn_readlane(kcy, k);
            const int pitch = __builtin_amdgcn_readlane(kpitch, k);
            const uint8_t* bl = g.blur + pix_of(k);
            const int bx = (cx - kPatchR) & ~3;
#pragma unroll
            for (int it = 0; it < kPatchIt; it++) {
                // q / 10 as (q * ceil(2^16 / 10)) >> 16, exact for q < 384
                const int q = lane + 64 * it, r = (int)(__umul24((uint32_t)q, 6554u) >> 16),
                          w = q - kPatchWords * r;
                pw[it] = (q < kPatchRows * kPatchWords && bx + 4 * w + 4 <= pitch)
                             ? *reinterpret_cast<const uint32_t*>(
                                   bl + (uint32_t)((cy - kPatchR) * pitch + __mul24(r, pitch) + bx + 4 * w))
                             : 0u;
            }
        };
        int k = __builtin_ctzll(todo);
        bool st = staged_of(k);
        if (st) pload(k);
        for (;;) {
            todo &= todo - 1;
            const int cx = __builtin_amdgcn_readlane(kcx, k), cy = __builtin_amdgcn_readlane(kcy, k);
            const float a = readlane_f(ka, k), b = readlane_f(kb, k);
            if (st) {
                wave_lds_sync();  // the previous keypoint's taps are done
#pragma unroll
                for (int it = 0; it < kPatchIt; it++) {
                    const int q = lane + 64 * it;
                    if (q < kPatchRows * kPatchWords) patch[q] = pw[it];
                }
            }
            const int kn = todo ? __builtin_ctzll(todo) : -1;
            const bool stn = kn >= 0 && staged_of(kn);
            if (stn) pload(kn);
            int oy[8], ox[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                oy[q] = (int)rintf(ppx[q] * b + ppy[q] * a);
                ox[q] = (int)rintf(ppx[q] * a - ppy[q] * b);
            }
            unsigned long long words[4];
            if (st) {
                wave_lds_sync();
                const uint8_t* pb = reinterpret_cast<const uint8_t*>(patch);
                const int c0 = cx - ((cx - kPatchR) & ~3);  // patch column of the centre
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int t[2];
#pragma unroll
                    for (int q = 0; q < 2; q++)
                        t[q] = pb[(kPatchR + oy[2 * r + q]) * (4 * kPatchWords) + c0 + ox[2 * r + q]];
                    words[r] = __ballot(t[0] < t[1]);
                }
            } else {
                const int w = __builtin_amdgcn_readlane(kw, k), h = __builtin_amdgcn_readlane(kh, k);
                const int pitch = __builtin_amdgcn_readlane(kpitch, k);
                const uint8_t* bl = g.blur + pix_of(k);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int t[2];
#pragma unroll
                    for (int q = 0; q < 2; q++) {
                        const int yy = min(max(cy + oy[2 * r + q], 0), h - 1);
                        const int xx = min(max(cx + ox[2 * r + q], 0), w - 1);
                        t[q] = bl[__umul24(yy, pitch) + xx];
                    }
                    words[r] = __ballot(t[0] < t[1]);
                }
            }
            if (lane == 0) {
                const int img = __builtin_amdgcn_readlane(kimg, k), pos = __builtin_amdgcn_readlane(kpos, k);
                unsigned long long* dst =
                    reinterpret_cast<unsigned long long*>(out_desc + ((long long)img * cap + pos) * 32);
                dst[0] = words[0];
                dst[1] = words[1];
                dst[2] = words[2];
                dst[3] = words[3];
            }
            if (kn < 0) break;
            k = kn;
            st = stn;
        }
    }
}

}  // namespace mcv
