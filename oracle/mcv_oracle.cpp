// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain C++17 restatement of MCVSLAM's hot path, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg to check and to
// time the HIP product path. Nothing in the product links, loads or calls it.
//
// It follows the reference line by line (citations are file:line in the
// reference tree) and restates, from their published sources, the OpenCV
// primitives the reference calls, because OpenCV is not available here:
//   cv::resize INTER_LINEAR 8U      (fixed-point, 11-bit coefficients)
//   cv::FAST TYPE_9_16 + NMS         (FAST_t / cornerScore<16>)
//   cv::GaussianBlur 7x7 8U          (bit-exact fixed-point path, error-diffused taps)
//   cv::fastAtan2                    (7th-order float polynomial, degrees)
//   cv::Mat 3x3*3x1 + 3x1 gemm (f32) (accumulated in double, one rounding)
// glibc cosf/sinf and libstdc++ std::priority_queue are used directly, as the
// reference does.  Parity of those OpenCV restatements with a real OpenCV
// build is unpinned (no OpenCV exists in this environment); every other
// function is pinned by the tests listed in tests/test_oracle_*.py.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no -ffast-math).

#include <algorithm>
#include <array>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <queue>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// OpenCV rounding helpers (cvRound = round-half-even via the SSE conversion,
// cvFloor/cvCeil, saturate_cast<short>(float)).
static inline int cv_round(float v) { return (int)lrintf(v); }
static inline int cv_round(double v) { return (int)lrint(v); }
static inline int cv_floor(float v) { return (int)std::floor(v); }
static inline int cv_floor(double v) { return (int)std::floor(v); }
static inline int cv_ceil(float v) { return (int)std::ceil(v); }
static inline short sat_short(float v) {
    int i = cv_round(v);
    return (short)std::min(32767, std::max(-32768, i));
}

struct Image {
    int w = 0, h = 0;
    std::vector<uint8_t> px;
    uint8_t* row(int y) { return px.data() + (size_t)y * w; }
    const uint8_t* row(int y) const { return px.data() + (size_t)y * w; }
};

struct KP {
    float x, y, size, angle, response;
    int octave;
};

// ---------------------------------------------------------------------------
// cv::resize(src, dst, dsize, 0, 0, INTER_LINEAR) for CV_8UC1, the
// fixed-point generic path (ORBextractor.cc:911 calls it level by level).
// Horizontal taps: int32 = S[sx]*a0 + S[sx+1]*a1 with a = rint(f*2048);
// vertical: ((b0*(R0>>4))>>16) + ((b1*(R1>>4))>>16) + 2) >> 2.
void resize_linear_u8(const Image& src, Image& dst, int dw, int dh) {
    dst.w = dw;
    dst.h = dh;
    dst.px.assign((size_t)dw * dh, 0);
    const int sw = src.w, sh = src.h;
    if (dw == sw && dh == sh) {  // same size: cv::resize copies
        dst.px = src.px;
        return;
    }
    const double inv_sx = (double)dw / sw, inv_sy = (double)dh / sh;
    const double scale_x = 1. / inv_sx, scale_y = 1. / inv_sy;
    std::vector<int> xofs(dw);
    std::vector<short> ia(2 * dw);
    for (int dx = 0; dx < dw; dx++) {
        float fx = (float)((dx + 0.5) * scale_x - 0.5);
        int sx = cv_floor(fx);
        fx -= sx;
        if (sx < 0) { fx = 0; sx = 0; }
        if (sx + 1 >= sw) {
            if (sx >= sw - 1) { fx = 0; sx = sw - 1; }
        }
        xofs[dx] = sx;
        ia[2 * dx] = sat_short((1.f - fx) * 2048);
        ia[2 * dx + 1] = sat_short(fx * 2048);
    }
    std::vector<int> r0(dw), r1(dw);
    auto hpass = [&](int sy, std::vector<int>& out) {
        const uint8_t* S = src.row(sy);
        for (int dx = 0; dx < dw; dx++) {
            int sx = xofs[dx];
            int sx1 = std::min(sx + 1, sw - 1);
            out[dx] = S[sx] * ia[2 * dx] + S[sx1] * ia[2 * dx + 1];
        }
    };
    for (int dy = 0; dy < dh; dy++) {
        float fy = (float)((dy + 0.5) * scale_y - 0.5);
        int sy = cv_floor(fy);
        fy -= sy;
        short b0 = sat_short((1.f - fy) * 2048), b1 = sat_short(fy * 2048);
        int y0 = std::min(std::max(sy, 0), sh - 1);
        int y1 = std::min(std::max(sy + 1, 0), sh - 1);
        hpass(y0, r0);
        hpass(y1, r1);
        uint8_t* D = dst.row(dy);
        for (int x = 0; x < dw; x++)
            D[x] = (uint8_t)((((b0 * (r0[x] >> 4)) >> 16) + ((b1 * (r1[x] >> 4)) >> 16) + 2) >> 2);
    }
}

// ---------------------------------------------------------------------------
// cv::FAST(img, kps, threshold, nonmax=true), TYPE_9_16: FAST_t<16> with
// cornerScore<16>; corners on rows/cols [3, n-3), 3x3 NMS through a 3-row
// ring buffer in which non-corners score 0, output row-major.
static const int kFastOffsets[16][2] = {{0, 3},  {1, 3},   {2, 2},   {3, 1},   {3, 0},  {3, -1},
                                        {2, -2}, {1, -3},  {0, -3},  {-1, -3}, {-2, -2}, {-3, -1},
                                        {-3, 0}, {-3, 1},  {-2, 2},  {-1, 3}};

static int corner_score16(const uint8_t* ptr, const int pixel[], int threshold) {
    const int K = 8, N = K * 3 + 1;
    int k, v = ptr[0];
    short d[N];
    for (k = 0; k < N; k++) d[k] = (short)(v - ptr[pixel[k]]);
    int a0 = threshold;
    for (k = 0; k < 16; k += 2) {
        int a = std::min((int)d[k + 1], (int)d[k + 2]);
        a = std::min(a, (int)d[k + 3]);
        if (a <= a0) continue;
        a = std::min(a, (int)d[k + 4]);
        a = std::min(a, (int)d[k + 5]);
        a = std::min(a, (int)d[k + 6]);
        a = std::min(a, (int)d[k + 7]);
        a = std::min(a, (int)d[k + 8]);
        a0 = std::max(a0, std::min(a, (int)d[k]));
        a0 = std::max(a0, std::min(a, (int)d[k + 9]));
    }
    int b0 = -a0;
    for (k = 0; k < 16; k += 2) {
        int b = std::max((int)d[k + 1], (int)d[k + 2]);
        b = std::max(b, (int)d[k + 3]);
        b = std::max(b, (int)d[k + 4]);
        b = std::max(b, (int)d[k + 5]);
        if (b >= b0) continue;
        b = std::max(b, (int)d[k + 6]);
        b = std::max(b, (int)d[k + 7]);
        b = std::max(b, (int)d[k + 8]);
        b0 = std::min(b0, std::max(b, (int)d[k]));
        b0 = std::min(b0, std::max(b, (int)d[k + 9]));
    }
    return -b0 - 1;
}

void fast16(const uint8_t* img, int stride, int cols, int rows, int threshold, std::vector<KP>& kps) {
    kps.clear();
    if (cols < 7 || rows < 7) return;
    const int K = 8, N = 16 + K + 1;
    int pixel[25];
    for (int k = 0; k < 16; k++) pixel[k] = kFastOffsets[k][0] + kFastOffsets[k][1] * stride;
    for (int k = 16; k < 25; k++) pixel[k] = pixel[k - 16];
    threshold = std::min(std::max(threshold, 0), 255);
    uint8_t threshold_tab[512];
    for (int i = -255; i <= 255; i++) threshold_tab[i + 255] = (uint8_t)(i < -threshold ? 1 : i > threshold ? 2 : 0);

    std::vector<uint8_t> bufv((size_t)cols * 3, 0);
    std::vector<int> cpv((size_t)(cols + 1) * 3, 0);
    uint8_t* buf[3] = {bufv.data(), bufv.data() + cols, bufv.data() + 2 * cols};
    int* cpbuf[3] = {cpv.data() + 1, cpv.data() + 1 + (cols + 1), cpv.data() + 1 + 2 * (cols + 1)};

    for (int i = 3; i < rows - 2; i++) {
        const uint8_t* ptr = img + (size_t)i * stride + 3;
        uint8_t* curr = buf[(i - 3) % 3];
        int* cornerpos = cpbuf[(i - 3) % 3];
        std::memset(curr, 0, cols);
        int ncorners = 0;
        if (i < rows - 3) {
            for (int j = 3; j < cols - 3; j++, ptr++) {
                int v = ptr[0];
                const uint8_t* tab = &threshold_tab[0] - v + 255;
                int d = tab[ptr[pixel[0]]] | tab[ptr[pixel[8]]];
                if (d == 0) continue;
                d &= tab[ptr[pixel[2]]] | tab[ptr[pixel[10]]];
                d &= tab[ptr[pixel[4]]] | tab[ptr[pixel[12]]];
                d &= tab[ptr[pixel[6]]] | tab[ptr[pixel[14]]];
                if (d == 0) continue;
                d &= tab[ptr[pixel[1]]] | tab[ptr[pixel[9]]];
                d &= tab[ptr[pixel[3]]] | tab[ptr[pixel[11]]];
                d &= tab[ptr[pixel[5]]] | tab[ptr[pixel[13]]];
                d &= tab[ptr[pixel[7]]] | tab[ptr[pixel[15]]];
                if (d & 1) {
                    int vt = v - threshold, count = 0;
                    for (int k = 0; k < N; k++) {
                        int x = ptr[pixel[k]];
                        if (x < vt) {
                            if (++count > K) {
                                cornerpos[ncorners++] = j;
                                curr[j] = (uint8_t)corner_score16(ptr, pixel, threshold);
                                break;
                            }
                        } else
                            count = 0;
                    }
                }
                if (d & 2) {
                    int vt = v + threshold, count = 0;
                    for (int k = 0; k < N; k++) {
                        int x = ptr[pixel[k]];
                        if (x > vt) {
                            if (++count > K) {
                                cornerpos[ncorners++] = j;
                                curr[j] = (uint8_t)corner_score16(ptr, pixel, threshold);
                                break;
                            }
                        } else
                            count = 0;
                    }
                }
            }
        }
        cornerpos[-1] = ncorners;
        if (i == 3) continue;
        const uint8_t* prev = buf[(i - 4 + 3) % 3];
        const uint8_t* pprev = buf[(i - 5 + 3) % 3];
        cornerpos = cpbuf[(i - 4 + 3) % 3];
        ncorners = cornerpos[-1];
        for (int k = 0; k < ncorners; k++) {
            int j = cornerpos[k];
            int score = prev[j];
            if (score > prev[j + 1] && score > prev[j - 1] && score > pprev[j - 1] && score > pprev[j] &&
                score > pprev[j + 1] && score > curr[j - 1] && score > curr[j] && score > curr[j + 1]) {
                kps.push_back(KP{(float)j, (float)(i - 1), 7.f, -1.f, (float)score, 0});
            }
        }
    }
}

// ---------------------------------------------------------------------------
// cv::fastAtan2: |err| < 0.3 deg polynomial, result in degrees [0, 360).
static const float atan2_p1 = 0.9997878412794807f * (float)(180 / M_PI);
static const float atan2_p3 = -0.3258083974640975f * (float)(180 / M_PI);
static const float atan2_p5 = 0.1555786518463281f * (float)(180 / M_PI);
static const float atan2_p7 = -0.04432655554792128f * (float)(180 / M_PI);

float fast_atan2(float y, float x) {
    float ax = std::abs(x), ay = std::abs(y);
    float a, c, c2;
    if (ax >= ay) {
        c = ay / (ax + (float)DBL_EPSILON);
        c2 = c * c;
        a = (((atan2_p7 * c2 + atan2_p5) * c2 + atan2_p3) * c2 + atan2_p1) * c;
    } else {
        c = ax / (ay + (float)DBL_EPSILON);
        c2 = c * c;
        a = 90.f - (((atan2_p7 * c2 + atan2_p5) * c2 + atan2_p3) * c2 + atan2_p1) * c;
    }
    if (x < 0) a = 180.f - a;
    if (y < 0) a = 360.f - a;
    return a;
}

// ---------------------------------------------------------------------------
// GaussianBlur(7x7, sigma 2, REFLECT_101) bit-exact fixed-point path:
// softdouble Gaussian normalised to sum 1, then error-diffused to 8
// fractional bits with the centre tap taking the remainder.
void gauss_taps_ed(int n, double sigma, int* taps) {
    const double scale2X = -0.125 / (sigma * sigma);
    const int n2 = (n - 1) / 2;
    std::vector<double> values(n2 + 1);
    double sum = 0;
    for (int i = 0, x = 1 - n; i < n2; i++, x += 2) {
        double t = std::exp((double)(x * x) * scale2X);
        values[i] = t;
        sum += t;
    }
    sum *= 2;
    sum += 1.0;
    const double mul1 = 1.0 / sum;
    std::vector<double> k(n);
    for (int i = 0; i < n2; i++) k[i] = k[n - 1 - i] = values[i] * mul1;
    k[n2] = mul1;
    const double fracMul = 256.0;
    double err = 0;
    int s = 0;
    for (int i = 0; i < n2; i++) {
        double adj = k[i] * fracMul + err;
        int v0 = (int)lrint(adj);
        err = adj - (double)v0;
        taps[i] = taps[n - 1 - i] = v0;
        s += v0;
    }
    s *= 2;
    taps[n2] = (int)lrint(fracMul - (double)s);
}

static inline int reflect101(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) {
        if (i < 0) i = -i;
        if (i >= n) i = 2 * n - 2 - i;
    }
    return i;
}

// h[x] = sum_k t[k]*src[x+k-3]   (u8 x Q8 -> Q8, <= 65280)
// out  = (sum_k t[k]*h[y+k-3] + 2^15) >> 16  (Q8 x Q8 -> Q16, rounded)
void gauss_blur7(const Image& src, Image& dst, const int* t) {
    const int w = src.w, h = src.h;
    dst.w = w;
    dst.h = h;
    dst.px.assign((size_t)w * h, 0);
    std::vector<uint32_t> hb((size_t)w * h);
    for (int y = 0; y < h; y++) {
        const uint8_t* S = src.row(y);
        for (int x = 0; x < w; x++) {
            uint32_t acc = 0;
            for (int k = 0; k < 7; k++) acc += (uint32_t)t[k] * S[reflect101(x + k - 3, w)];
            hb[(size_t)y * w + x] = acc;
        }
    }
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            uint64_t acc = 0;
            for (int k = 0; k < 7; k++) acc += (uint64_t)t[k] * hb[(size_t)reflect101(y + k - 3, h) * w + x];
            uint64_t v = (acc + (1u << 15)) >> 16;
            dst.row(y)[x] = (uint8_t)std::min<uint64_t>(v, 255);
        }
    }
}

// ---------------------------------------------------------------------------
// ORBextractor (modules/local_feature/ORB/orb3_extractor/ORBextractor.cc)
static const int PATCH_SIZE = 31;
static const int HALF_PATCH_SIZE = 15;
static const int EDGE_THRESHOLD = 19;

#include "../mcvslam_amd/csrc/orb_pattern.h"

struct Orb {
    int nfeatures = 0, nlevels = 0, iniTh = 0, minTh = 0;
    float scaleFactor = 1.f;
    std::vector<float> scale, invScale, sigma2, invSigma2;
    std::vector<int> nPerLevel;
    int umax[HALF_PATCH_SIZE + 1];
    int gauss[7];
    std::vector<Image> pyr;   // mvImagePyramid (ROI contents only)
    std::vector<Image> blur;  // GaussianBlur of each level (descriptor source)
    std::vector<std::vector<KP>> cand;  // per-level FAST candidates (before octree)

    // ORBextractor::init, ORBextractor.cc:407-457
    void init(int nf, float sf, int nl, int ini, int mn) {
        nfeatures = nf; scaleFactor = sf; nlevels = nl; iniTh = ini; minTh = mn;
        scale.assign(nl, 0.f); sigma2.assign(nl, 0.f);
        scale[0] = 1.0f; sigma2[0] = 1.0f;
        for (int i = 1; i < nl; i++) {
            scale[i] = scale[i - 1] * sf;
            sigma2[i] = scale[i] * scale[i];
        }
        invScale.assign(nl, 0.f); invSigma2.assign(nl, 0.f);
        for (int i = 0; i < nl; i++) {
            invScale[i] = 1.0f / scale[i];
            invSigma2[i] = 1.0f / sigma2[i];
        }
        pyr.assign(nl, Image());
        blur.assign(nl, Image());
        nPerLevel.assign(nl, 0);
        float factor = 1.0f / sf;
        float nd = nf * (1 - factor) / (1 - (float)std::pow((double)factor, (double)nl));
        int sumF = 0;
        for (int level = 0; level < nl - 1; level++) {
            nPerLevel[level] = cv_round(nd);
            sumF += nPerLevel[level];
            nd *= factor;
        }
        nPerLevel[nl - 1] = std::max(nf - sumF, 0);
        int v, v0, vmax = cv_floor(HALF_PATCH_SIZE * std::sqrt(2.f) / 2 + 1);
        int vmin = cv_ceil(HALF_PATCH_SIZE * std::sqrt(2.f) / 2);
        const double hp2 = HALF_PATCH_SIZE * HALF_PATCH_SIZE;
        for (v = 0; v <= vmax; ++v) umax[v] = cv_round(std::sqrt(hp2 - v * v));
        for (v = HALF_PATCH_SIZE, v0 = 0; v >= vmin; --v) {
            while (umax[v0] == umax[v0 + 1]) ++v0;
            umax[v] = v0;
            ++v0;
        }
        gauss_taps_ed(7, 2.0, gauss);
    }

    // ComputePyramid, ORBextractor.cc:901-919
    void computePyramid(const Image& image) {
        for (int level = 0; level < nlevels; ++level) {
            float s = invScale[level];
            int w = cv_round((float)image.w * s), h = cv_round((float)image.h * s);
            if (level == 0)
                pyr[0] = image;
            else
                resize_linear_u8(pyr[level - 1], pyr[level], w, h);
        }
    }

    // IC_Angle, ORBextractor.cc:75-98
    float icAngle(const Image& image, float px, float py) const {
        int m_01 = 0, m_10 = 0;
        const int cx = cv_round(px), cy = cv_round(py);
        const uint8_t* center = image.row(cy) + cx;
        for (int u = -HALF_PATCH_SIZE; u <= HALF_PATCH_SIZE; ++u) m_10 += u * center[u];
        const int step = image.w;
        for (int v = 1; v <= HALF_PATCH_SIZE; ++v) {
            int v_sum = 0;
            int d = umax[v];
            for (int u = -d; u <= d; ++u) {
                int val_plus = center[u + v * step], val_minus = center[u - v * step];
                v_sum += (val_plus - val_minus);
                m_10 += u * (val_plus + val_minus);
            }
            m_01 += v * v_sum;
        }
        return fast_atan2((float)m_01, (float)m_10);
    }

    // computeOrbDescriptor, ORBextractor.cc:101-141.  Taps that would fall
    // outside the image (only possible for caller-supplied seed keypoints,
    // where the reference reads out of bounds) are clamped to the border.
    void descriptor(const KP& kpt, const Image& img, uint8_t* desc) const {
        const float factorPI = (float)(M_PI / 180.f);
        float angle = (float)kpt.angle * factorPI;
        float a = cosf(angle), b = sinf(angle);
        const int cy = cv_round(kpt.y), cx = cv_round(kpt.x);
        for (int i = 0; i < 32; ++i) {
            int val = 0;
            for (int j = 0; j < 8; j++) {
                const signed char* p = kOrbPattern[8 * i + j];
                int t[2];
                for (int q = 0; q < 2; q++) {
                    float x = (float)p[2 * q], y = (float)p[2 * q + 1];
                    int r = cy + cv_round(x * b + y * a);
                    int c = cx + cv_round(x * a - y * b);
                    r = std::min(std::max(r, 0), img.h - 1);
                    c = std::min(std::max(c, 0), img.w - 1);
                    t[q] = img.row(r)[c];
                }
                val |= (t[0] < t[1]) << j;
            }
            desc[i] = (uint8_t)val;
        }
    }
};

// ExtractorNode / DistributeOctTree, ORBextractor.h:32-42, ORBextractor.cc:469-580
struct Node;
using NodeRef = std::shared_ptr<Node>;
struct IPt { int x, y; };
struct Node {
    std::vector<KP> vKeys;
    IPt UL, UR, BL, BR;
    std::vector<NodeRef> sons;
    int size() const { return (int)vKeys.size(); }
    std::vector<NodeRef>& divide() {
        const int halfX = (int)std::ceil(static_cast<float>(UR.x - UL.x) / 2);
        const int halfY = (int)std::ceil(static_cast<float>(BR.y - UL.y) / 2);
        NodeRef n1 = std::make_shared<Node>(), n2 = std::make_shared<Node>(), n3 = std::make_shared<Node>(),
                n4 = std::make_shared<Node>();
        n1->UL = UL;
        n1->UR = IPt{UL.x + halfX, UL.y};
        n1->BL = IPt{UL.x, UL.y + halfY};
        n1->BR = IPt{UL.x + halfX, UL.y + halfY};
        n2->UL = n1->UR;
        n2->UR = UR;
        n2->BL = n1->BR;
        n2->BR = IPt{UR.x, UL.y + halfY};
        n3->UL = n1->BL;
        n3->UR = n1->BR;
        n3->BL = BL;
        n3->BR = IPt{n1->BR.x, BL.y};
        n4->UL = n3->UR;
        n4->UR = n2->BR;
        n4->BL = n3->BR;
        n4->BR = BR;
        for (const KP& kp : vKeys) {
            if (kp.x < n1->UR.x) {
                if (kp.y < n1->BR.y)
                    n1->vKeys.push_back(kp);
                else
                    n3->vKeys.push_back(kp);
            } else if (kp.y < n1->BR.y)
                n2->vKeys.push_back(kp);
            else
                n4->vKeys.push_back(kp);
        }
        if (n1->size()) sons.push_back(n1);
        if (n2->size()) sons.push_back(n2);
        if (n3->size()) sons.push_back(n3);
        if (n4->size()) sons.push_back(n4);
        return sons;
    }
};

// Empty input (undefined behaviour in the reference: top() of an empty
// heap) is defined here as "no keypoints".  N = 0 splits nothing and emits
// the best keypoint of every root, as the reference's size_t comparison does.
std::vector<KP> distribute_octree(const std::vector<KP>& keys, int minX, int maxX, int minY, int maxY, int N) {
    std::vector<KP> ret;
    if (keys.empty()) return ret;
    const int nIni = (int)std::round(static_cast<float>(maxX - minX) / (maxY - minY));
    const float hX = static_cast<float>(maxX - minX) / nIni;
    auto cmp = [](const NodeRef& a, const NodeRef& b) { return a->vKeys.size() < b->vKeys.size(); };
    std::priority_queue<NodeRef, std::vector<NodeRef>, decltype(cmp)> q(cmp);
    std::vector<NodeRef> init_nodes;
    for (int i = 0; i < nIni; i++) {
        NodeRef p = std::make_shared<Node>();
        p->UL = IPt{(int)(hX * static_cast<float>(i)), 0};
        p->UR = IPt{(int)(hX * static_cast<float>(i + 1)), 0};
        p->BL = IPt{p->UL.x, maxY - minY};
        p->BR = IPt{p->UR.x, maxY - minY};
        init_nodes.push_back(p);
    }
    for (const KP& kp : keys) init_nodes[(size_t)(kp.x / hX)]->vKeys.push_back(kp);
    for (auto& p : init_nodes)
        if (p->size()) q.push(p);
    while (q.size() < (size_t)std::max(N, 0)) {
        auto p = q.top();
        if (p->size() == 1) break;
        q.pop();
        for (auto& c : p->divide()) q.push(c);
    }
    ret.reserve(q.size());
    while (!q.empty()) {
        auto p = q.top();
        q.pop();
        KP best = p->vKeys[0];
        for (int i = 1; i < p->size(); i++)
            if (p->vKeys[i].response > best.response) best = p->vKeys[i];
        ret.push_back(best);
    }
    return ret;
}

// ComputeKeyPointsOctTree, ORBextractor.cc:582-654
void compute_keypoints_octree(Orb& o, std::vector<std::vector<KP>>& all) {
    all.assign(o.nlevels, {});
    o.cand.assign(o.nlevels, {});
    const float W = 35;
    for (int level = 0; level < o.nlevels; ++level) {
        const Image& im = o.pyr[level];
        const int minBorderX = EDGE_THRESHOLD - 3;
        const int minBorderY = minBorderX;
        const int maxBorderX = im.w - EDGE_THRESHOLD + 3;
        const int maxBorderY = im.h - EDGE_THRESHOLD + 3;
        std::vector<KP> toDist;
        const float width = (float)(maxBorderX - minBorderX);
        const float height = (float)(maxBorderY - minBorderY);
        const int nCols = (int)(width / W);
        const int nRows = (int)(height / W);
        if (nCols <= 0 || nRows <= 0) continue;
        const int wCell = (int)std::ceil(width / nCols);
        const int hCell = (int)std::ceil(height / nRows);
        std::vector<KP> cell;
        for (int i = 0; i < nRows; i++) {
            const float iniY = (float)(minBorderY + i * hCell);
            float maxY = iniY + hCell + 6;
            if (iniY >= maxBorderY - 3) continue;
            if (maxY > maxBorderY) maxY = (float)maxBorderY;
            for (int j = 0; j < nCols; j++) {
                const float iniX = (float)(minBorderX + j * wCell);
                float maxX = iniX + wCell + 6;
                if (iniX >= maxBorderX - 6) continue;
                if (maxX > maxBorderX) maxX = (float)maxBorderX;
                const int y0 = (int)iniY, y1 = (int)maxY, x0 = (int)iniX, x1 = (int)maxX;
                const uint8_t* base = im.row(y0) + x0;
                fast16(base, im.w, x1 - x0, y1 - y0, o.iniTh, cell);
                if (cell.empty()) fast16(base, im.w, x1 - x0, y1 - y0, o.minTh, cell);
                for (KP k : cell) {
                    k.x += j * wCell;
                    k.y += i * hCell;
                    toDist.push_back(k);
                }
            }
        }
        o.cand[level] = toDist;
        std::vector<KP>& kps = all[level];
        kps = distribute_octree(toDist, minBorderX, maxBorderX, minBorderY, maxBorderY, o.nPerLevel[level]);
        const int scaledPatchSize = (int)(PATCH_SIZE * o.scale[level]);
        for (KP& k : kps) {
            k.x += minBorderX;
            k.y += minBorderY;
            k.octave = level;
            k.size = (float)scaledPatchSize;
        }
    }
    for (int level = 0; level < o.nlevels; ++level)
        for (KP& k : all[level]) k.angle = o.icAngle(o.pyr[level], k.x, k.y);
}

// ORBextractor::operator(), ORBextractor.cc:831-899.  Returns the keypoint
// count, -1 for an empty image, -2 for a seed whose octave is out of range.
int orb_extract(Orb& o, const Image& image, const std::vector<KP>& seeds, std::vector<KP>& out,
                std::vector<uint8_t>& desc) {
    out.clear();
    desc.clear();
    if (image.w <= 0 || image.h <= 0) return -1;
    for (const KP& s : seeds)
        if (s.octave < 0 || s.octave >= o.nlevels) return -2;
    o.computePyramid(image);
    std::vector<std::vector<KP>> all;
    compute_keypoints_octree(o, all);
    for (const KP& s : seeds) all[s.octave].push_back(s);
    int nkp = 0;
    for (auto& l : all) nkp += (int)l.size();
    out.resize(nkp);
    desc.assign((size_t)nkp * 32, 0);
    int cnt = 0;
    for (int level = 0; level < o.nlevels; ++level) {
        std::vector<KP>& kps = all[level];
        if (kps.empty()) continue;
        gauss_blur7(o.pyr[level], o.blur[level], o.gauss);
        float s = o.scale[level];
        for (KP& k : kps) {
            o.descriptor(k, o.blur[level], &desc[(size_t)cnt * 32]);
            if (level != 0) {
                k.x = k.x * s;
                k.y = k.y * s;
            }
            out[cnt++] = k;
        }
    }
    return cnt;
}

// ---------------------------------------------------------------------------
// Matcher (include/Matcher.hpp:19-33, src/Matcher.cpp:23-35,100-111,245-302)
static inline unsigned hamming(const uint8_t* a8, const uint8_t* b8) {
    unsigned dist = 0;
    for (int i = 0; i < 8; i++) {
        int32_t pa, pb;
        std::memcpy(&pa, a8 + 4 * i, 4);
        std::memcpy(&pb, b8 + 4 * i, 4);
        unsigned v = (unsigned)(pa ^ pb);
        v = v - ((v >> 1) & 0x55555555);
        v = (v & 0x33333333) + ((v >> 2) & 0x33333333);
        dist += (((v + (v >> 4)) & 0xF0F0F0F) * 0x1010101) >> 24;
    }
    return dist;
}

static inline void top2(const uint8_t* q, const uint8_t* t, int nt, const int32_t* sel, int32_t* o) {
    unsigned d[2] = {999, 999};
    unsigned id[2] = {0, 0};
    for (int j = 0; j < nt; j++) {
        unsigned dist = hamming(q, t + (size_t)(sel ? sel[j] : j) * 32);
        if (dist < d[0]) {
            d[1] = d[0];
            id[1] = id[0];
            d[0] = dist;
            id[0] = j;
        } else if (dist < d[1]) {
            d[1] = dist;
            id[1] = j;
        }
    }
    o[0] = (int32_t)id[0];
    o[1] = (int32_t)d[0];
    o[2] = (int32_t)id[1];
    o[3] = (int32_t)d[1];
}

// MatchResKnn::FilterRatio then MatchRes::FilterThreshold.
static int filter_ratio_threshold(const int32_t* top, int nq, float ratio, int th, int32_t* out) {
    std::vector<std::array<int32_t, 3>> res;
    for (int i = 0; i < nq; i++) {
        const int32_t* m = top + 4 * i;
        if ((float)m[1] / (float)m[3] <= ratio) res.push_back({i, m[0], m[1]});
    }
    int i = 0, j = (int)res.size() - 1;
    while (i <= j) {
        if ((float)res[i][2] > th)
            res[i] = res[j--];
        else
            i++;
    }
    res.resize(i);
    for (int k = 0; k < i; k++)
        for (int c = 0; c < 3; c++) out[3 * k + c] = res[k][c];
    return i;
}

// ---------------------------------------------------------------------------
// Object grid + window search (src/Object.cpp:182-201,249-308)
struct Grid {
    int w, h;
    float inv_w, inv_h;
    std::vector<std::vector<int>> cells;  // [ix*30+iy]
};
static const int GRID = 30;

static void grid_build(Grid& g, const float* xy, int n, int w, int h) {
    g.w = w;
    g.h = h;
    g.inv_w = (float)(static_cast<double>(GRID) / w);
    g.inv_h = (float)(static_cast<double>(GRID) / h);
    g.cells.assign(GRID * GRID, {});
    for (int i = 0; i < n; i++) {
        int px = (int)std::round((xy[2 * i] - 0) * g.inv_w);
        int py = (int)std::round((xy[2 * i + 1] - 0) * g.inv_h);
        if (px < 0 || px >= GRID || py < 0 || py >= GRID) continue;
        g.cells[px * GRID + py].push_back(i);
    }
}

static void grid_query(const Grid& g, const float* xy, const int* octave, float x, float y, float r, int minLevel,
                       int maxLevel, std::vector<int>& out) {
    out.clear();
    if (std::isnan(x) || std::isnan(y)) return;
    if (x < 0 || y < 0 || x >= g.w || y >= g.h) return;
    const float fX = r, fY = r;
    const int nMinCellX = std::max(0, (int)std::floor((x - 0 - fX) * g.inv_w));
    if (nMinCellX >= GRID) return;
    const int nMaxCellX = std::min(GRID - 1, (int)std::ceil((x - 0 + fX) * g.inv_w));
    if (nMaxCellX < 0) return;
    const int nMinCellY = std::max(0, (int)std::floor((y - 0 - fY) * g.inv_h));
    if (nMinCellY >= GRID) return;
    const int nMaxCellY = std::min(GRID - 1, (int)std::ceil((y - 0 + fY) * g.inv_h));
    if (nMaxCellY < 0) return;
    const bool bCheck = (minLevel > 0) || (maxLevel >= 0);
    for (int ix = nMinCellX; ix <= nMaxCellX; ix++)
        for (int iy = nMinCellY; iy <= nMaxCellY; iy++)
            for (int idx : g.cells[ix * GRID + iy]) {
                if (bCheck) {
                    if (octave[idx] < minLevel) continue;
                    if (maxLevel >= 0 && octave[idx] > maxLevel) continue;
                }
                const float dx = xy[2 * idx] - x, dy = xy[2 * idx + 1] - y;
                if (std::fabs(dx) < fX && std::fabs(dy) < fY) out.push_back(idx);
            }
}

// ---------------------------------------------------------------------------
// BA: EdgeSE3XYZ2Camera in FP64 (include/OptimizeTypes.hpp:70-117) with g2o
// SE3Quat algebra (ThirdParty/g2o/g2o/types/slam3d/se3quat.h:99-105,212-215,
// 280-285), Pinhole (modules/camera/Pinhole.cpp:55-61,105-115), Huber
// (core/robust_kernel_impl.cpp:65-78) and constructQuadraticForm
// (core/base_binary_edge.hpp:64-136).
struct Quat { double x, y, z, w; };
struct SE3 { Quat r; double t[3]; };

static Quat qmul(const Quat& a, const Quat& b) {
    return Quat{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
                a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
static void qrot(const Quat& q, const double v[3], double o[3]) {
    // Eigen _transformVector: uv = 2 q.vec x v; v + w*uv + q.vec x uv
    double uv[3] = {q.y * v[2] - q.z * v[1], q.z * v[0] - q.x * v[2], q.x * v[1] - q.y * v[0]};
    for (double& c : uv) c += c;
    double cx = q.y * uv[2] - q.z * uv[1], cy = q.z * uv[0] - q.x * uv[2], cz = q.x * uv[1] - q.y * uv[0];
    o[0] = v[0] + q.w * uv[0] + cx;
    o[1] = v[1] + q.w * uv[1] + cy;
    o[2] = v[2] + q.w * uv[2] + cz;
}
static void qnormalize_pos(Quat& q) {
    if (q.w < 0) { q.x = -q.x; q.y = -q.y; q.z = -q.z; q.w = -q.w; }
    double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    q.x /= n; q.y /= n; q.z /= n; q.w /= n;
}
static SE3 se3mul(const SE3& a, const SE3& b) {
    SE3 r;
    double rt[3];
    qrot(a.r, b.t, rt);
    for (int i = 0; i < 3; i++) r.t[i] = a.t[i] + rt[i];
    r.r = qmul(a.r, b.r);
    qnormalize_pos(r.r);
    return r;
}
static void se3map(const SE3& T, const double p[3], double o[3]) {
    qrot(T.r, p, o);
    for (int i = 0; i < 3; i++) o[i] += T.t[i];
}
static void qtoR(const Quat& q, double R[9]) {
    const double tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
    R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
    R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

struct Cam { double fx, fy, cx, cy; SE3 Tcb; double delta; };

// One observation: residual e, chi2, Huber weight, J_pose (2x6, row-major),
// J_point (2x3, row-major).
static void edge_linearize(const SE3& Tbw, const Cam& c, const double Pw[3], double u, double v, double invSigma2,
                           double* e, double* chi2, double* w, double* Jpose, double* Jpoint) {
    // computeError, OptimizeTypes.hpp:80-86
    SE3 Tcw = se3mul(c.Tcb, Tbw);
    double Pc0[3];
    se3map(Tcw, Pw, Pc0);
    e[0] = u - (c.fx * Pc0[0] / Pc0[2] + c.cx);
    e[1] = v - (c.fy * Pc0[1] / Pc0[2] + c.cy);
    // chi2 = e' * (I*invSigma2) * e, base_edge.h:94-97
    const double ch = e[0] * (invSigma2 * e[0]) + e[1] * (invSigma2 * e[1]);
    *chi2 = ch;
    const double dsqr = c.delta * c.delta;
    *w = (ch <= dsqr) ? 1.0 : c.delta / std::sqrt(ch);
    // linearizeOplus, OptimizeTypes.hpp:94-113
    double Pb[3], Pc[3];
    se3map(Tbw, Pw, Pb);
    se3map(c.Tcb, Pb, Pc);
    const double z = Pc[2];
    const double Jp[6] = {c.fx / z, 0.0, -c.fx * Pc[0] / (z * z), 0.0, c.fy / z, -c.fy * Pc[1] / (z * z)};
    double Rcw[9], Rcb[9];
    qtoR(Tcw.r, Rcw);
    qtoR(c.Tcb.r, Rcb);
    for (int r = 0; r < 2; r++)
        for (int k = 0; k < 3; k++) {
            double s = 0;
            for (int m = 0; m < 3; m++) s += Jp[3 * r + m] * Rcw[3 * m + k];
            Jpoint[3 * r + k] = -s;
        }
    const double x = Pb[0], y = Pb[1], zb = Pb[2];
    const double D[18] = {0, zb, -y, 1, 0, 0, -zb, 0, x, 0, 1, 0, y, -x, 0, 0, 0, 1};
    double JR[6];
    for (int r = 0; r < 2; r++)
        for (int k = 0; k < 3; k++) {
            double s = 0;
            for (int m = 0; m < 3; m++) s += Jp[3 * r + m] * Rcb[3 * m + k];
            JR[3 * r + k] = s;
        }
    for (int r = 0; r < 2; r++)
        for (int k = 0; k < 6; k++) {
            double s = 0;
            for (int m = 0; m < 3; m++) s += JR[3 * r + m] * D[6 * m + k];
            Jpose[6 * r + k] = -s;
        }
}

// Stereo edge (EdgeStereoSE3ProjectXYZ, include/OptimizeTypes.hpp:119-157,
// src/OptimizeTypes.cpp:12-58): pose vertex Tcw (the keyframe pose; the left
// camera is the body), residual obs(u, v, u_r) - cam_project_uv_u(P_c, bf),
// where cam_project_uv_u takes a float inverse depth and a float bf:
//   invz = (float)(1.0 / z), u = x * invz * fx + cx, v = y * invz * fy + cy,
//   u_r = u - (float)(bf_f * invz)     (bf_f = (float)bf, a float product)
// J_point 3x3 and J_pose 3x6 (rotation first) in FP64 from z, as in
// linearizeOplus.  Information I * invSigma2, Huber on chi2.
static void stereo_edge_linearize(const SE3& Tcw, double fx, double fy, double cx, double cy, double bf,
                                  double delta, const double Pw[3], double u, double v, double ur, double invSigma2,
                                  double* e, double* chi2, double* w, double* Jpose, double* Jpoint) {
    double P[3];
    se3map(Tcw, Pw, P);
    const float invz = (float)(1.0 / P[2]);
    const double ru = P[0] * (double)invz * fx + cx;
    const double rv = P[1] * (double)invz * fy + cy;
    const float bff = (float)bf;
    const double rr = ru - (double)(bff * invz);
    e[0] = u - ru;
    e[1] = v - rv;
    e[2] = ur - rr;
    const double ch = invSigma2 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    *chi2 = ch;
    *w = (ch <= delta * delta) ? 1.0 : delta / std::sqrt(ch);
    double R[9];
    qtoR(Tcw.r, R);
    const double x = P[0], y = P[1], z = P[2], z2 = z * z;
    for (int k = 0; k < 3; k++) {
        Jpoint[k] = -fx * R[k] / z + fx * x * R[6 + k] / z2;
        Jpoint[3 + k] = -fy * R[3 + k] / z + fy * y * R[6 + k] / z2;
        Jpoint[6 + k] = Jpoint[k] - bf * R[6 + k] / z2;
    }
    double* J0 = Jpose;
    double* J1 = Jpose + 6;
    double* J2 = Jpose + 12;
    J0[0] = x * y / z2 * fx;
    J0[1] = -(1 + (x * x / z2)) * fx;
    J0[2] = y / z * fx;
    J0[3] = -1. / z * fx;
    J0[4] = 0;
    J0[5] = x / z2 * fx;
    J1[0] = (1 + y * y / z2) * fy;
    J1[1] = -x * y / z2 * fy;
    J1[2] = -x / z * fy;
    J1[3] = 0;
    J1[4] = -1. / z * fy;
    J1[5] = y / z2 * fy;
    J2[0] = J0[0] - bf * y / z2;
    J2[1] = J0[1] + bf * x / z2;
    J2[2] = J0[2];
    J2[3] = J0[3];
    J2[4] = 0;
    J2[5] = J0[5] - bf / z2;
}

}  // namespace orc

// ===========================================================================
// C entry points for ctypes.
using namespace orc;

static Image make_image(const uint8_t* px, int w, int h) {
    Image im;
    im.w = w;
    im.h = h;
    im.px.assign(px, px + (size_t)w * h);
    return im;
}

extern "C" {

// ORBextractor::init tables: scale[nl], invScale[nl], sigma2[nl],
// invSigma2[nl], nPerLevel[nl], umax[16], gauss taps[7].
void orc_orb_tables(int nf, float sf, int nl, int ini, int mn, float* scale, float* invScale, float* sigma2,
                    float* invSigma2, int* nPerLevel, int* umax, int* gauss) {
    Orb o;
    o.init(nf, sf, nl, ini, mn);
    for (int i = 0; i < nl; i++) {
        scale[i] = o.scale[i];
        invScale[i] = o.invScale[i];
        sigma2[i] = o.sigma2[i];
        invSigma2[i] = o.invSigma2[i];
        nPerLevel[i] = o.nPerLevel[i];
    }
    for (int i = 0; i < 16; i++) umax[i] = o.umax[i];
    for (int i = 0; i < 7; i++) gauss[i] = o.gauss[i];
}

void orc_resize(const uint8_t* src, int sw, int sh, uint8_t* dst, int dw, int dh) {
    Image s = make_image(src, sw, sh), d;
    resize_linear_u8(s, d, dw, dh);
    std::memcpy(dst, d.px.data(), (size_t)dw * dh);
}

void orc_blur(const uint8_t* src, int w, int h, const int* taps, uint8_t* dst) {
    Image s = make_image(src, w, h), d;
    gauss_blur7(s, d, taps);
    std::memcpy(dst, d.px.data(), (size_t)w * h);
}

// cv::FAST on a w x h window with row stride `stride`; out = (x, y, score)
int orc_fast(const uint8_t* img, int stride, int w, int h, int th, float* out, int cap) {
    std::vector<KP> k;
    fast16(img, stride, w, h, th, k);
    int n = (int)k.size();
    for (int i = 0; i < n && i < cap; i++) {
        out[3 * i] = k[i].x;
        out[3 * i + 1] = k[i].y;
        out[3 * i + 2] = k[i].response;
    }
    return n;
}

float orc_fast_atan2(float y, float x) { return fast_atan2(y, x); }

// DistributeOctTree on (x, y, response) triples; out = (x, y, response)
int orc_octree(const float* xyr, int n, int minX, int maxX, int minY, int maxY, int N, float* out, int cap) {
    std::vector<KP> keys(n);
    for (int i = 0; i < n; i++) keys[i] = KP{xyr[3 * i], xyr[3 * i + 1], 7.f, -1.f, xyr[3 * i + 2], 0};
    std::vector<KP> r = distribute_octree(keys, minX, maxX, minY, maxY, N);
    int m = (int)r.size();
    for (int i = 0; i < m && i < cap; i++) {
        out[3 * i] = r[i].x;
        out[3 * i + 1] = r[i].y;
        out[3 * i + 2] = r[i].response;
    }
    return m;
}

// Full ORB extraction of one image.  seeds/out keypoints are rows of
// (x, y, size, angle, response, octave).  level_info (optional) receives, per
// level: width, height, number of FAST candidates, number of keypoints.
// pyr_out/blur_out (optional) receive the concatenated level images.
int orc_orb_extract(const uint8_t* img, int w, int h, int nf, float sf, int nl, int ini, int mn, const float* seeds,
                    int nseeds, float* out_kps, uint8_t* out_desc, int cap, int* level_info, uint8_t* pyr_out,
                    uint8_t* blur_out, float* cand_out, int cand_cap) {
    Orb o;
    o.init(nf, sf, nl, ini, mn);
    Image im;
    if (w > 0 && h > 0) im = make_image(img, w, h);
    std::vector<KP> sd(nseeds);
    for (int i = 0; i < nseeds; i++)
        sd[i] = KP{seeds[6 * i], seeds[6 * i + 1], seeds[6 * i + 2], seeds[6 * i + 3], seeds[6 * i + 4],
                   (int)seeds[6 * i + 5]};
    std::vector<KP> out;
    std::vector<uint8_t> desc;
    int n = orb_extract(o, im, sd, out, desc);
    if (n < 0) return n;
    for (int i = 0; i < n && i < cap; i++) {
        const KP& k = out[i];
        float* r = out_kps + 6 * i;
        r[0] = k.x; r[1] = k.y; r[2] = k.size; r[3] = k.angle; r[4] = k.response; r[5] = (float)k.octave;
        std::memcpy(out_desc + 32 * (size_t)i, &desc[32 * (size_t)i], 32);
    }
    if (level_info) {
        std::vector<int> perlevel(nl, 0);
        for (const KP& k : out) perlevel[k.octave]++;
        for (int l = 0; l < nl; l++) {
            level_info[4 * l] = o.pyr[l].w;
            level_info[4 * l + 1] = o.pyr[l].h;
            level_info[4 * l + 2] = (int)o.cand[l].size();
            level_info[4 * l + 3] = perlevel[l];
        }
    }
    size_t off = 0;
    for (int l = 0; l < nl; l++) {
        size_t sz = (size_t)o.pyr[l].w * o.pyr[l].h;
        if (pyr_out) std::memcpy(pyr_out + off, o.pyr[l].px.data(), sz);
        if (blur_out && !o.blur[l].px.empty()) std::memcpy(blur_out + off, o.blur[l].px.data(), sz);
        off += sz;
    }
    if (cand_out) {
        int c = 0;
        for (int l = 0; l < nl; l++)
            for (const KP& k : o.cand[l]) {
                if (c < cand_cap) {
                    cand_out[4 * c] = k.x;
                    cand_out[4 * c + 1] = k.y;
                    cand_out[4 * c + 2] = k.response;
                    cand_out[4 * c + 3] = (float)l;
                }
                c++;
            }
    }
    return n;
}

// Frame::ComputeStereoMatch (src/Frame.cpp:150-328) on one left/right pair.
// kl/kr: rows (x, y, size, angle, response, octave) in level-0 coordinates,
// dl/dr: 32-byte descriptors, pyr_l/pyr_r: concatenated pyramid levels
// (sizes lw/lh per level), scale/inv_scale: mvScaleFactor/mvInvScaleFactor.
// Outputs per left keypoint: u_right (-1 when unmatched), depth (-1), and
// the SAD distance of the accepted match (-1).  Where the reference would
// index outside a pyramid level (colRange with a negative start) the match
// is skipped.
void orc_stereo(const float* kl, const uint8_t* dl, int nl, const float* kr, const uint8_t* dr, int nr,
                const uint8_t* pyr_l, const uint8_t* pyr_r, const int* lw, const int* lh, int nlev,
                const float* scale, const float* inv_scale, float bf, float b, float* u_right, float* depth,
                int32_t* sad) {
    std::vector<size_t> off(nlev + 1, 0);
    for (int l = 0; l < nlev; l++) off[l + 1] = off[l] + (size_t)lw[l] * lh[l];
    for (int i = 0; i < nl; i++) { u_right[i] = -1.f; depth[i] = -1.f; sad[i] = -1; }
    const int nRows = lh[0];
    std::vector<std::vector<size_t>> rows(nRows);
    for (int iR = 0; iR < nr; iR++) {
        const float kpY = kr[6 * iR + 1];
        const float r = 10.f * scale[(int)kr[6 * iR + 5]];
        const int maxr = (int)std::ceil(kpY + r);
        const int minr = (int)std::floor(kpY - r);
        for (int yi = minr; yi <= maxr; yi++)
            if (yi >= 0 && yi < nRows) rows[yi].push_back(iR);
    }
    const float minZ = b, minD = 1;
    const float maxD = std::min(bf / minZ, float(1000.0));
    std::vector<std::pair<int, int>> vDistIdx;
    std::vector<int> cand;
    for (int iL = 0; iL < nl; iL++) {
        const int levelL = (int)kl[6 * iL + 5];
        const float vL = kl[6 * iL + 1], uL = kl[6 * iL];
        const std::vector<size_t>& vc = rows[(size_t)vL];
        if (vc.empty()) continue;
        const float minU = uL - maxD, maxU = uL - minD;
        if (maxU < 0) continue;
        cand.clear();
        for (size_t iR : vc) {
            const int oR = (int)kr[6 * iR + 5];
            if (oR < levelL - 1 || oR > levelL + 1) continue;
            const float uR = kr[6 * iR];
            if (uR >= minU && uR <= maxU) cand.push_back((int)iR);
        }
        if (cand.empty()) continue;
        int32_t tp[4], m[3];
        top2(dl + 32 * (size_t)iL, dr, (int)cand.size(), cand.data(), tp);
        if (filter_ratio_threshold(tp, 1, 0.70f, (int)(46 * 0.75), m) != 1) continue;
        const int bestIdxR = m[1];
        const float uR0 = kr[6 * cand[bestIdxR]];
        const float scaleFactor = inv_scale[levelL];
        const float scaleduL = std::round(uL * scaleFactor);
        const float scaledvL = std::round(vL * scaleFactor);
        const float scaleduR0 = std::round(uR0 * scaleFactor);
        const int w = 5, L = 5;
        const int cols = lw[levelL], rws = lh[levelL];
        if (scaledvL - w < 0 || scaledvL + w + 1 >= rws || scaleduL - w < 0 || scaleduL + w + 1 >= cols) continue;
        const float iniu = scaleduR0 + L - w, endu = scaleduR0 + L + w + 1;
        if (iniu < 0 || endu >= cols) continue;
        if (scaleduR0 - L - w < 0) continue;   // the reference's colRange would assert
        const uint8_t* IL = pyr_l + off[levelL];
        const uint8_t* IR = pyr_r + off[levelL];
        const int y0 = (int)scaledvL - w, xl = (int)scaleduL - w;
        const int cl = IL[(size_t)(y0 + w) * cols + xl + w];
        int bestDist = INT32_MAX, bestincR = 0;
        float vDists[2 * L + 1];
        for (int incR = -L; incR <= L; incR++) {
            const int xr = (int)scaleduR0 + incR - w;
            const int cr = IR[(size_t)(y0 + w) * cols + xr + w];
            int acc = 0;
            for (int yy = 0; yy < 2 * w + 1; yy++)
                for (int xx = 0; xx < 2 * w + 1; xx++) {
                    const int a = (int)IL[(size_t)(y0 + yy) * cols + xl + xx] - cl;
                    const int bb = (int)IR[(size_t)(y0 + yy) * cols + xr + xx] - cr;
                    acc += std::abs(a - bb);
                }
            const float dist = (float)acc;
            if (dist < bestDist) {
                bestDist = (int)dist;
                bestincR = incR;
            }
            vDists[L + incR] = dist;
        }
        if (bestincR == -L || bestincR == L) continue;
        const float dist1 = vDists[L + bestincR - 1], dist2 = vDists[L + bestincR], dist3 = vDists[L + bestincR + 1];
        const float deltaR = (dist1 - dist3) / (2.0f * (dist1 + dist3 - 2.0f * dist2));
        if (deltaR < -1 || deltaR > 1) continue;
        const float bestuR = scale[levelL] * ((float)scaleduR0 + (float)bestincR + deltaR);
        const float disparity = uL - bestuR;
        if (disparity >= minD && disparity < maxD) {
            depth[iL] = bf / disparity;
            u_right[iL] = bestuR;
            sad[iL] = bestDist;
            vDistIdx.push_back({bestDist, iL});
        }
    }
    if (vDistIdx.empty()) return;
    std::sort(vDistIdx.begin(), vDistIdx.end());
    const float median = (float)vDistIdx[(size_t)(vDistIdx.size() * 1.0 / 2)].first;
    const float th_max = 1.6f * median;
    const float th_min = 0.4 * median;
    int l = 0, r = (int)vDistIdx.size() - 1;
    for (; r >= 0 && vDistIdx[r].first > th_max; r--) {
        u_right[vDistIdx[r].second] = -1;
        depth[vDistIdx[r].second] = -1;
        sad[vDistIdx[r].second] = -1;
    }
    for (; l < r && vDistIdx[l].first < th_min; l++) {
        u_right[vDistIdx[l].second] = -1;
        depth[vDistIdx[l].second] = -1;
        sad[vDistIdx[l].second] = -1;
    }
}

// IC_Angle of one keypoint on an image (umax of ORBextractor::init)
float orc_ic_angle(const uint8_t* img, int w, int h, float x, float y) {
    Orb o;
    o.init(500, 1.2f, 1, 20, 7);
    Image im = make_image(img, w, h);
    return o.icAngle(im, x, y);
}

// computeOrbDescriptor of one keypoint on an (already blurred) image
void orc_descriptor(const uint8_t* img, int w, int h, float x, float y, float angle, uint8_t* out) {
    Orb o;
    o.init(500, 1.2f, 1, 20, 7);
    Image im = make_image(img, w, h);
    KP k{x, y, 31.f, angle, 0.f, 0};
    o.descriptor(k, im, out);
}

// host libm cosf / sinf over an array (the functions the reference calls)
void orc_libm_sincos(const float* x, int n, float* c, float* s) {
    for (int i = 0; i < n; i++) {
        c[i] = cosf(x[i]);
        s[i] = sinf(x[i]);
    }
}

// Matcher::KnnMatch(vector<Mat>, vector<Mat>) -> rows (i0, d0, i1, d1)
void orc_knn2(const uint8_t* q, int nq, const uint8_t* t, int nt, int32_t* out) {
    for (int i = 0; i < nq; i++) top2(q + 32 * (size_t)i, t, nt, nullptr, out + 4 * i);
}

unsigned orc_hamming(const uint8_t* a, const uint8_t* b) { return hamming(a, b); }

// FilterRatio(ratio).FilterThreshold(th) over knn rows -> (query, train, dist)
int orc_filter(const int32_t* top, int nq, float ratio, int th, int32_t* out) {
    return filter_ratio_threshold(top, nq, ratio, th, out);
}

// Object::AssignFeaturesToGrid + GetFeaturesInArea for one query.
int orc_window(const float* xy, const int* octave, int n, int w, int h, float x, float y, float r, int minL, int maxL,
               int32_t* out, int cap) {
    Grid g;
    grid_build(g, xy, n, w, h);
    std::vector<int> c;
    grid_query(g, xy, octave, x, y, r, minL, maxL, c);
    for (int i = 0; i < (int)c.size() && i < cap; i++) out[i] = c[i];
    return (int)c.size();
}

// Object::ProjectBunchMapPoints for one camera, every map point in order.
// pts: np x 3 world points (f32), pdesc: np x 32, plevel: np; R (3x3
// row-major), t: world-to-camera (f32, as Object::GetRotation/GetTranslation);
// cam: fx, fy, cx, cy; scales: per-level scale factors; kxy: nk x 2 keypoint
// positions, kdesc: nk x 32.  out rows: (kp0, d0, kp1, d1, ncand, match)
// where kp* are keypoint indices (-1 for the 999 sentinel) and match is the
// keypoint accepted by FilterRatio(ratio).FilterThreshold(th) or -1.
void orc_project_search(const float* pts, const uint8_t* pdesc, const int* plevel, int np, const float* R,
                        const float* t, const float* cam, const float* scales, float r_th, float ratio, int th,
                        const float* kxy, const int* koct, const uint8_t* kdesc, int nk, int w, int h, int32_t* out) {
    Grid g;
    grid_build(g, kxy, nk, w, h);
    std::vector<int> cand;
    for (int p = 0; p < np; p++) {
        int32_t* o = out + 6 * p;
        o[0] = -1; o[1] = 999; o[2] = -1; o[3] = 999; o[4] = 0; o[5] = -1;
        float Pc[3];
        for (int i = 0; i < 3; i++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += (double)R[3 * i + k] * (double)pts[3 * p + k];
            Pc[i] = (float)(s + (double)t[i]);
        }
        if (Pc[2] < 0) continue;
        const float u = cam[0] * Pc[0] / Pc[2] + cam[2];
        const float v = cam[1] * Pc[1] / Pc[2] + cam[3];
        const float r = r_th * scales[plevel[p]];
        grid_query(g, kxy, koct, u, v, r, -1, -1, cand);
        o[4] = (int32_t)cand.size();
        if (cand.empty()) continue;
        int32_t tp[4];
        top2(pdesc + 32 * (size_t)p, kdesc, (int)cand.size(), cand.data(), tp);
        o[0] = tp[1] == 999 ? -1 : cand[tp[0]];
        o[1] = tp[1];
        o[2] = tp[3] == 999 ? -1 : cand[tp[2]];
        o[3] = tp[3];
        int32_t m[3];
        if (filter_ratio_threshold(tp, 1, ratio, th, m) == 1) o[5] = cand[m[1]];
    }
}

// FP64 linearization of EdgeSE3XYZ2Camera edges.
//   obs_uv: M x 2 (f32), obs_pt / obs_frame / obs_cam: M, obs_isig: M (f32)
//   points: P x 3 (f64), frames: F x 7 (qx qy qz qw tx ty tz, f64, Tbw)
//   cams: C x 12 (fx fy cx cy  qx qy qz qw tx ty tz  delta), f64
// outputs (f64): e M x 2, chi2 M, w M, Jpose M x 12, Jpoint M x 6
void orc_ba_linearize(const float* obs_uv, const int32_t* obs_pt, const int32_t* obs_frame, const int32_t* obs_cam,
                      const float* obs_isig, int M, const double* points, const double* frames, const double* cams,
                      double* e, double* chi2, double* w, double* Jpose, double* Jpoint) {
    for (int i = 0; i < M; i++) {
        const double* f = frames + 7 * obs_frame[i];
        SE3 Tbw{Quat{f[0], f[1], f[2], f[3]}, {f[4], f[5], f[6]}};
        const double* c = cams + 12 * obs_cam[i];
        Cam cam{c[0], c[1], c[2], c[3], SE3{Quat{c[4], c[5], c[6], c[7]}, {c[8], c[9], c[10]}}, c[11]};
        edge_linearize(Tbw, cam, points + 3 * obs_pt[i], (double)obs_uv[2 * i], (double)obs_uv[2 * i + 1],
                       (double)obs_isig[i], e + 2 * i, chi2 + i, w + i, Jpose + 12 * i, Jpoint + 6 * i);
    }
}

// Stereo edges: obs (M x 3 f32: u, v, u_right), point, frame, inverse level
// sigma^2; frames F x 7 (qx qy qz qw tx ty tz, camera = body pose); cam
// (fx, fy, cx, cy, bf, delta).  Outputs e (M x 3), chi2, w, Jpose (M x 18),
// Jpoint (M x 9).
void orc_ba_stereo_linearize(const float* obs, const int32_t* obs_pt, const int32_t* obs_frame, const float* obs_isig,
                             int M, const double* points, const double* frames, const double* cam, double* e,
                             double* chi2, double* w, double* Jpose, double* Jpoint) {
    for (int i = 0; i < M; i++) {
        const double* f = frames + 7 * obs_frame[i];
        SE3 Tcw{Quat{f[0], f[1], f[2], f[3]}, {f[4], f[5], f[6]}};
        qnormalize_pos(Tcw.r);  // SE3Quat keeps a unit quaternion with w >= 0
        stereo_edge_linearize(Tcw, cam[0], cam[1], cam[2], cam[3], cam[4], cam[5], points + 3 * obs_pt[i],
                              (double)obs[3 * i], (double)obs[3 * i + 1], (double)obs[3 * i + 2], (double)obs_isig[i],
                              e + 3 * i, chi2 + i, w + i, Jpose + 18 * i, Jpoint + 9 * i);
    }
}

// MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cpp:101-150) for n
// map points whose observed descriptors are desc[ptr[m] .. ptr[m+1]) (rows
// of 32 B): all-pairs Hamming distances (diagonal 0), each row sorted, the
// row's value at index (size_t)(0.5 * (N - 1)), and the first row with the
// smallest such median.  best[m] is that row's index within the map point's
// list, -1 for a map point without descriptors (the reference returns).
void orc_distinctive(const uint8_t* desc, const int32_t* ptr, int n, int32_t* best) {
    for (int m = 0; m < n; m++) {
        const int b = ptr[m], N = ptr[m + 1] - ptr[m];
        if (N <= 0) {
            best[m] = -1;
            continue;
        }
        std::vector<std::vector<int>> D(N, std::vector<int>(N, 0));
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++) {
                const int d = hamming(desc + 32 * (size_t)(b + i), desc + 32 * (size_t)(b + j));
                D[i][j] = d;
                D[j][i] = d;
            }
        int bestMedian = INT_MAX, bestIdx = 0;
        for (int i = 0; i < N; i++) {
            std::sort(D[i].begin(), D[i].end());
            const int median = D[i][(size_t)(0.5 * (N - 1))];
            if (median < bestMedian) {
                bestMedian = median;
                bestIdx = i;
            }
        }
        best[m] = bestIdx;
    }
}

// MatchRes::FilterOrientation (src/Matcher.cpp:44-75): rotation histogram of
// 40 bins over kp1.angle - kp2.angle, bins ordered by size with std::sort
// (not stable: the tie order is libstdc++'s), the matches of the first three
// bins kept, bin by bin, each in list order.  matches: n x (query, train,
// distance); angles1 / angles2 indexed by query / train.  Returns the count.
int orc_filter_orientation(const int32_t* matches, int n, const float* angles1, const float* angles2, int32_t* out) {
    const int HISTO_LENGTH = 40;
    const float HISTO_FACTOR = 1.0f / (360.0f / HISTO_LENGTH);
    std::vector<uint32_t> rot_bins[HISTO_LENGTH];
    for (int i = 0; i < n; i++) {
        float rot = angles1[matches[3 * i]] - angles2[matches[3 * i + 1]];
        if (rot < 0.0) rot += 360.0f;
        int bin_id = (int)::round(rot * HISTO_FACTOR);
        if (bin_id == HISTO_LENGTH) bin_id = 0;
        if (bin_id < 0 || bin_id >= HISTO_LENGTH) return -1;
        rot_bins[bin_id].push_back((uint32_t)i);
    }
    auto cmp = [](const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) { return a.size() > b.size(); };
    std::sort(&rot_bins[0], &rot_bins[0] + HISTO_LENGTH, cmp);
    int k = 0;
    for (int b = 0; b < 3; b++)
        for (uint32_t idx : rot_bins[b]) {
            out[3 * k] = matches[3 * idx];
            out[3 * k + 1] = matches[3 * idx + 1];
            out[3 * k + 2] = matches[3 * idx + 2];
            k++;
        }
    return k;
}

// chi2 outlier test of FilterCallBack_Chi2 (src/PoseEstimation.cpp:132-151):
// flag = chi2 >= threshold (float compare of the float chi2).
void orc_ba_outliers(const double* chi2, int M, float threshold, uint8_t* flags) {
    for (int i = 0; i < M; i++) flags[i] = ((float)chi2[i] >= threshold) ? 1 : 0;
}

// FP64 quadratic form + Schur complement (block_solver.hpp:315-449 with the
// LM damping of setLambda).  Frames with frame_fixed[f] != 0 are skipped.
// Free frames are numbered in frame order.  Outputs: S (6Fp x 6Fp, full
// symmetric), bs (6Fp), Hll (P x 9), bl (P x 3), Hpp (6Fp x 6Fp), bp (6Fp).
int orc_ba_schur(const double* e, const double* w, const double* Jpose, const double* Jpoint, const int32_t* obs_pt,
                 const int32_t* obs_frame, const float* obs_isig, int M, int P, int F, const int32_t* frame_fixed,
                 double lambda, double* S, double* bs, double* Hll_out, double* bl_out, double* Hpp_out,
                 double* bp_out) {
    std::vector<int> fidx(F, -1);
    int Fp = 0;
    for (int f = 0; f < F; f++)
        if (!frame_fixed[f]) fidx[f] = Fp++;
    const int n = 6 * Fp;
    std::vector<double> Hpp((size_t)n * n, 0.0), bp(n, 0.0), Hll((size_t)P * 9, 0.0), bl((size_t)P * 3, 0.0);
    std::vector<std::vector<std::pair<int, std::array<double, 18>>>> Hpl(P);  // (free frame, 6x3)
    for (int i = 0; i < M; i++) {
        const double om = (double)obs_isig[i] * w[i];
        const double* A = Jpoint + 6 * i;
        const double* B = Jpose + 12 * i;
        const double r0 = -(double)obs_isig[i] * e[2 * i] * w[i], r1 = -(double)obs_isig[i] * e[2 * i + 1] * w[i];
        const int p = obs_pt[i];
        for (int a = 0; a < 3; a++) {
            bl[3 * p + a] += A[a] * r0 + A[3 + a] * r1;
            for (int b = 0; b < 3; b++) Hll[9 * p + 3 * a + b] += om * (A[a] * A[b] + A[3 + a] * A[3 + b]);
        }
        const int fi = fidx[obs_frame[i]];
        if (fi < 0) continue;
        for (int a = 0; a < 6; a++) {
            bp[6 * fi + a] += B[a] * r0 + B[6 + a] * r1;
            for (int b = 0; b < 6; b++)
                Hpp[(size_t)(6 * fi + a) * n + 6 * fi + b] += om * (B[a] * B[b] + B[6 + a] * B[6 + b]);
        }
        std::array<double, 18> blk;
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 3; b++) blk[3 * a + b] = om * (B[a] * A[b] + B[6 + a] * A[3 + b]);
        bool merged = false;
        for (auto& pr : Hpl[p])
            if (pr.first == fi) {
                for (int k = 0; k < 18; k++) pr.second[k] += blk[k];
                merged = true;
            }
        if (!merged) Hpl[p].push_back({fi, blk});
    }
    if (Hpp_out) std::memcpy(Hpp_out, Hpp.data(), sizeof(double) * n * n);
    if (bp_out) std::memcpy(bp_out, bp.data(), sizeof(double) * n);
    if (Hll_out) std::memcpy(Hll_out, Hll.data(), sizeof(double) * P * 9);
    if (bl_out) std::memcpy(bl_out, bl.data(), sizeof(double) * P * 3);
    for (int k = 0; k < n; k++) Hpp[(size_t)k * n + k] += lambda;
    std::vector<double> Ssum(Hpp), b(bp);
    std::vector<int> nobs(P, 0);
    for (int i = 0; i < M; i++) nobs[obs_pt[i]]++;
    for (int p = 0; p < P; p++) {
        if (nobs[p] == 0) continue;  // a point without edges is not in the optimizer
        double D[9];
        for (int k = 0; k < 9; k++) D[k] = Hll[9 * p + k];
        D[0] += lambda; D[4] += lambda; D[8] += lambda;
        const double det = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                           D[2] * (D[3] * D[7] - D[4] * D[6]);
        if (det == 0) return -1;
        double Di[9] = {(D[4] * D[8] - D[5] * D[7]) / det, (D[2] * D[7] - D[1] * D[8]) / det,
                        (D[1] * D[5] - D[2] * D[4]) / det, (D[5] * D[6] - D[3] * D[8]) / det,
                        (D[0] * D[8] - D[2] * D[6]) / det, (D[2] * D[3] - D[0] * D[5]) / det,
                        (D[3] * D[7] - D[4] * D[6]) / det, (D[1] * D[6] - D[0] * D[7]) / det,
                        (D[0] * D[4] - D[1] * D[3]) / det};
        double db[3];
        for (int a = 0; a < 3; a++) db[a] = Di[3 * a] * bl[3 * p] + Di[3 * a + 1] * bl[3 * p + 1] + Di[3 * a + 2] * bl[3 * p + 2];
        for (auto& pi : Hpl[p]) {
            double BD[18];
            for (int a = 0; a < 6; a++)
                for (int c = 0; c < 3; c++)
                    BD[3 * a + c] = pi.second[3 * a] * Di[c] + pi.second[3 * a + 1] * Di[3 + c] + pi.second[3 * a + 2] * Di[6 + c];
            for (int a = 0; a < 6; a++)
                b[6 * pi.first + a] -= pi.second[3 * a] * db[0] + pi.second[3 * a + 1] * db[1] + pi.second[3 * a + 2] * db[2];
            for (auto& pj : Hpl[p])
                for (int a = 0; a < 6; a++)
                    for (int c = 0; c < 6; c++)
                        Ssum[(size_t)(6 * pi.first + a) * n + 6 * pj.first + c] -=
                            BD[3 * a] * pj.second[3 * c] + BD[3 * a + 1] * pj.second[3 * c + 1] + BD[3 * a + 2] * pj.second[3 * c + 2];
        }
    }
    std::memcpy(S, Ssum.data(), sizeof(double) * n * n);
    std::memcpy(bs, b.data(), sizeof(double) * n);
    return Fp;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Levenberg-Marquardt over the whole BA problem in FP64, the oracle of the
// GPU optimizer (mcv_ba_optimize).  Restates:
//   OptimizationAlgorithmLevenberg::solve, computeLambdaInit, computeScale
//     (ThirdParty/g2o/g2o/core/optimization_algorithm_levenberg.cpp:56-159;
//     tau 1e-5, good-step scale clamped to [1/3, 2/3], 10 trials)
//   SparseOptimizer::computeActiveErrors / activeRobustChi2
//     (core/sparse_optimizer.cpp:63-102; Huber rho, robust_kernel_impl.cpp:65-78)
//   BaseBinaryEdge::constructQuadraticForm (core/base_binary_edge.hpp:64-136):
//     H += J^T w Omega J, b += -J^T w Omega e, blocks of fixed vertices skipped
//   BlockSolver::solve (core/block_solver.hpp:315-449): Schur complement of
//     the points with H_ll + lambda I, dense Cholesky of the reduced system,
//     x_l = D^-1 (b_l - H_lp x_p)
//   VertexSE3Expmap::oplusImpl (types/sba/types_six_dof_expmap.h:98-101):
//     T <- exp(x) * T, SE3Quat::exp (types/slam3d/se3quat.h:218-257) with
//     Eigen's rotation-matrix-to-quaternion conversion; VertexSBAPointXYZ:
//     P <- P + x.
// Edges: mono EdgeSE3XYZ2Camera (edge_linearize) and stereo
// EdgeStereoSE3ProjectXYZ (stereo_edge_linearize); inactive edges (level 1)
// take no part; robust = 0 runs without the Huber kernel.
namespace orc {

struct LMEdges {
    int M;
    const float* uv;
    const int32_t* pt;
    const int32_t* fr;
    const int32_t* cam;
    const float* isig;
    const uint8_t* act;
    int Ms;
    const float* sobs;
    const int32_t* spt;
    const int32_t* sfr;
    const float* sisig;
    const uint8_t* sact;
    const double* scam;  // fx fy cx cy bf delta
    int P, F, C;
    const double* cams;  // C x 12
    const int32_t* ffix;
    const int32_t* pfix;
    int robust;
};

static double huber_rho(double chi2, double delta, bool robust) {
    if (!robust) return chi2;
    const double d2 = delta * delta;
    return chi2 <= d2 ? chi2 : 2.0 * std::sqrt(chi2) * delta - d2;
}

// one linearized edge: r rows (2 mono, 3 stereo)
struct LinEdge {
    int rows, pt, fr;
    double e[3], chi2, w, isig, Jp[18], Jx[9];
};

static void lm_linearize(const LMEdges& E, const double* points, const double* frames, std::vector<LinEdge>& out) {
    out.clear();
    const double inf = INFINITY;
    for (int i = 0; i < E.M; i++) {
        if (E.act && !E.act[i]) continue;
        LinEdge L{};
        L.rows = 2;
        L.pt = E.pt[i];
        L.fr = E.fr[i];
        L.isig = (double)E.isig[i];
        const double* f = frames + 7 * L.fr;
        SE3 Tbw{Quat{f[0], f[1], f[2], f[3]}, {f[4], f[5], f[6]}};
        const double* c = E.cams + 12 * E.cam[i];
        Cam cam{c[0], c[1], c[2], c[3], SE3{Quat{c[4], c[5], c[6], c[7]}, {c[8], c[9], c[10]}}, E.robust ? c[11] : inf};
        double Jp[12], Jx[6];
        edge_linearize(Tbw, cam, points + 3 * L.pt, (double)E.uv[2 * i], (double)E.uv[2 * i + 1], L.isig, L.e,
                       &L.chi2, &L.w, Jp, Jx);
        for (int k = 0; k < 12; k++) L.Jp[k] = Jp[k];
        for (int k = 0; k < 6; k++) L.Jx[k] = Jx[k];
        out.push_back(L);
    }
    for (int j = 0; j < E.Ms; j++) {
        if (E.sact && !E.sact[j]) continue;
        LinEdge L{};
        L.rows = 3;
        L.pt = E.spt[j];
        L.fr = E.sfr[j];
        L.isig = (double)E.sisig[j];
        const double* f = frames + 7 * L.fr;
        SE3 Tcw{Quat{f[0], f[1], f[2], f[3]}, {f[4], f[5], f[6]}};
        qnormalize_pos(Tcw.r);
        stereo_edge_linearize(Tcw, E.scam[0], E.scam[1], E.scam[2], E.scam[3], E.scam[4], E.robust ? E.scam[5] : inf,
                              points + 3 * L.pt, (double)E.sobs[3 * j], (double)E.sobs[3 * j + 1],
                              (double)E.sobs[3 * j + 2], L.isig, L.e, &L.chi2, &L.w, L.Jp, L.Jx);
        out.push_back(L);
    }
}

static double lm_robust_chi2(const LMEdges& E, const std::vector<LinEdge>& L) {
    double s = 0;
    size_t k = 0;
    for (int i = 0; i < E.M; i++) {
        if (E.act && !E.act[i]) continue;
        s += huber_rho(L[k++].chi2, E.cams[12 * E.cam[i] + 11], E.robust);
    }
    for (int j = 0; j < E.Ms; j++) {
        if (E.sact && !E.sact[j]) continue;
        s += huber_rho(L[k++].chi2, E.scam[5], E.robust);
    }
    return s;
}

static void inv3(const double* D, double* Di) {
    const double det = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                       D[2] * (D[3] * D[7] - D[4] * D[6]);
    Di[0] = (D[4] * D[8] - D[5] * D[7]) / det; Di[1] = (D[2] * D[7] - D[1] * D[8]) / det;
    Di[2] = (D[1] * D[5] - D[2] * D[4]) / det; Di[3] = (D[5] * D[6] - D[3] * D[8]) / det;
    Di[4] = (D[0] * D[8] - D[2] * D[6]) / det; Di[5] = (D[2] * D[3] - D[0] * D[5]) / det;
    Di[6] = (D[3] * D[7] - D[4] * D[6]) / det; Di[7] = (D[1] * D[6] - D[0] * D[7]) / det;
    Di[8] = (D[0] * D[4] - D[1] * D[3]) / det;
}

// dense Cholesky solve; false when A is not positive definite
static bool llt_solve(int n, std::vector<double> A, const std::vector<double>& b, std::vector<double>& x) {
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0)) return false;
        const double l = std::sqrt(d);
        A[(size_t)j * n + j] = l;
        for (int i = j + 1; i < n; i++) {
            double v = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) v -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = v / l;
        }
    }
    x.assign(b.begin(), b.end());
    for (int i = 0; i < n; i++) {
        double v = x[i];
        for (int k = 0; k < i; k++) v -= A[(size_t)i * n + k] * x[k];
        x[i] = v / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double v = x[i];
        for (int k = i + 1; k < n; k++) v -= A[(size_t)k * n + i] * x[k];
        x[i] = v / A[(size_t)i * n + i];
    }
    return true;
}

// Eigen::Quaternion(const Matrix3&): trace branch, else the largest diagonal
static Quat quat_from_R(const double* m) {
    Quat q{};
    double t = m[0] + m[4] + m[8];
    if (t > 0) {
        t = std::sqrt(t + 1.0);
        q.w = 0.5 * t;
        t = 0.5 / t;
        q.x = (m[7] - m[5]) * t;
        q.y = (m[2] - m[6]) * t;
        q.z = (m[3] - m[1]) * t;
    } else {
        int i = 0;
        if (m[4] > m[0]) i = 1;
        if (m[8] > m[3 * i + i]) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double c[3];
        t = std::sqrt(m[3 * i + i] - m[3 * j + j] - m[3 * k + k] + 1.0);
        c[i] = 0.5 * t;
        t = 0.5 / t;
        q.w = (m[3 * k + j] - m[3 * j + k]) * t;
        c[j] = (m[3 * j + i] + m[3 * i + j]) * t;
        c[k] = (m[3 * k + i] + m[3 * i + k]) * t;
        q.x = c[0]; q.y = c[1]; q.z = c[2];
    }
    return q;
}

// SE3Quat::exp of (omega, upsilon)
static SE3 se3_exp(const double* xi) {
    const double w[3] = {xi[0], xi[1], xi[2]};
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += W[3 * r + k] * W[3 * k + c];
            W2[3 * r + c] = s;
        }
    double R[9], V[9];
    double a, bb, cc;  // R = I + a W + bb W^2, V = I + bb W + cc W^2
    if (th < 0.00001) {
        a = 1.0; bb = 0.5; cc = 1.0 / 6.0;
        for (int k = 0; k < 9; k++) {
            const double I = (k % 4 == 0) ? 1.0 : 0.0;
            R[k] = I + W[k] + 0.5 * W2[k];
            V[k] = I + 0.5 * W[k] + cc * W2[k];
        }
    } else {
        a = std::sin(th) / th;
        bb = (1 - std::cos(th)) / (th * th);
        cc = (th - std::sin(th)) / std::pow(th, 3);
        for (int k = 0; k < 9; k++) {
            const double I = (k % 4 == 0) ? 1.0 : 0.0;
            R[k] = I + a * W[k] + bb * W2[k];
            V[k] = I + bb * W[k] + cc * W2[k];
        }
    }
    SE3 T;
    T.r = quat_from_R(R);
    qnormalize_pos(T.r);
    for (int r = 0; r < 3; r++) T.t[r] = V[3 * r] * xi[3] + V[3 * r + 1] * xi[4] + V[3 * r + 2] * xi[5];
    return T;
}

}  // namespace orc

extern "C" {

// exp(xi) * T for one frame (qx qy qz qw tx ty tz), in place
void orc_se3_oplus(double* frame, const double* xi) {
    SE3 T{Quat{frame[0], frame[1], frame[2], frame[3]}, {frame[4], frame[5], frame[6]}};
    const SE3 r = se3mul(se3_exp(xi), T);
    frame[0] = r.r.x; frame[1] = r.r.y; frame[2] = r.r.z; frame[3] = r.r.w;
    frame[4] = r.t[0]; frame[5] = r.t[1]; frame[6] = r.t[2];
}

// One LM run of `iterations` iterations (SparseOptimizer::optimize(n) after
// initializeOptimization): points (P x 3) and frames (F x 7) updated in
// place.  stats[4 k ..]: chi2 before iteration k, chi2 after it, lambda
// after it, trials.  Returns the iterations run (g2o stops after a
// Terminate).  x_out (optional): the first iteration's first trial step
// [x_p (6 Fp, free frames in order); x_l (3 P, zero for fixed points)].
int orc_ba_lm(int M, const float* uv, const int32_t* pt, const int32_t* fr, const int32_t* cam, const float* isig,
              const uint8_t* act, int Ms, const float* sobs, const int32_t* spt, const int32_t* sfr,
              const float* sisig, const uint8_t* sact, const double* scam, int P, int F, int C, const double* cams,
              const int32_t* ffix, const int32_t* pfix, int robust, double* points, double* frames, int iterations,
              double tau, int max_trials, double* stats, double* x_out) {
    LMEdges E{M, uv, pt, fr, cam, isig, act, Ms, sobs, spt, sfr, sisig, sact, scam, P, F, C, cams, ffix, pfix, robust};
    std::vector<int> fidx(F, -1);
    int Fp = 0;
    for (int f = 0; f < F; f++)
        if (!ffix || !ffix[f]) fidx[f] = Fp++;
    const int n = 6 * Fp;
    auto pfree = [&](int p) { return !pfix || !pfix[p]; };
    std::vector<LinEdge> L;
    double lambda = 0, ni = 2;
    int done = 0;
    for (int it = 0; it < iterations; it++) {
        lm_linearize(E, points, frames, L);
        double currentChi = lm_robust_chi2(E, L);
        if (stats) stats[4 * it] = currentChi;
        // buildSystem
        std::vector<double> Hpp((size_t)Fp * 36, 0.0), bp(n, 0.0), Hll((size_t)P * 9, 0.0), bl((size_t)P * 3, 0.0);
        std::vector<std::vector<std::pair<int, std::array<double, 18>>>> Hpl(P);
        for (const LinEdge& e : L) {
            const int r = e.rows;
            const double W = e.w * e.isig;
            const int fi = fidx[e.fr];
            const bool pf = pfree(e.pt);
            double om[3];
            for (int k = 0; k < r; k++) om[k] = -e.isig * e.e[k] * e.w;
            if (fi >= 0) {
                for (int a = 0; a < 6; a++) {
                    for (int k = 0; k < r; k++) bp[6 * fi + a] += e.Jp[6 * k + a] * om[k];
                    for (int b = 0; b < 6; b++) {
                        double s = 0;
                        for (int k = 0; k < r; k++) s += e.Jp[6 * k + a] * e.Jp[6 * k + b];
                        Hpp[36 * fi + 6 * a + b] += W * s;
                    }
                }
            }
            if (pf) {
                const int p = e.pt;
                for (int a = 0; a < 3; a++) {
                    for (int k = 0; k < r; k++) bl[3 * p + a] += e.Jx[3 * k + a] * om[k];
                    for (int b = 0; b < 3; b++) {
                        double s = 0;
                        for (int k = 0; k < r; k++) s += e.Jx[3 * k + a] * e.Jx[3 * k + b];
                        Hll[9 * p + 3 * a + b] += W * s;
                    }
                }
            }
            if (fi >= 0 && pf) {
                std::array<double, 18> blk;
                for (int a = 0; a < 6; a++)
                    for (int b = 0; b < 3; b++) {
                        double s = 0;
                        for (int k = 0; k < r; k++) s += e.Jp[6 * k + a] * e.Jx[3 * k + b];
                        blk[3 * a + b] = W * s;
                    }
                bool merged = false;
                for (auto& q : Hpl[e.pt])
                    if (q.first == fi) {
                        for (int k = 0; k < 18; k++) q.second[k] += blk[k];
                        merged = true;
                    }
                if (!merged) Hpl[e.pt].push_back({fi, blk});
            }
        }
        if (it == 0) {
            double mx = 0;
            for (int f = 0; f < Fp; f++)
                for (int a = 0; a < 6; a++) mx = std::max(mx, std::fabs(Hpp[36 * f + 7 * a]));
            for (int p = 0; p < P; p++)
                if (pfree(p))
                    for (int a = 0; a < 3; a++) mx = std::max(mx, std::fabs(Hll[9 * p + 4 * a]));
            lambda = tau * mx;
            ni = 2;
        }
        double rho = 0;
        int qmax = 0;
        do {
            const std::vector<double> pts0(points, points + 3 * (size_t)P), frs0(frames, frames + 7 * (size_t)F);
            // Schur complement with lambda on both diagonals
            std::vector<double> S((size_t)n * n, 0.0), bs(bp);
            for (int f = 0; f < Fp; f++)
                for (int a = 0; a < 6; a++)
                    for (int b = 0; b < 6; b++)
                        S[(size_t)(6 * f + a) * n + 6 * f + b] = Hpp[36 * f + 6 * a + b] + (a == b ? lambda : 0.0);
            std::vector<double> Dinv((size_t)P * 9, 0.0);
            for (int p = 0; p < P; p++) {
                if (!pfree(p)) continue;
                double D[9];
                for (int k = 0; k < 9; k++) D[k] = Hll[9 * p + k];
                D[0] += lambda; D[4] += lambda; D[8] += lambda;
                double* Di = &Dinv[9 * (size_t)p];
                inv3(D, Di);
                double db[3];
                for (int a = 0; a < 3; a++) db[a] = Di[3 * a] * bl[3 * p] + Di[3 * a + 1] * bl[3 * p + 1] + Di[3 * a + 2] * bl[3 * p + 2];
                for (auto& qi : Hpl[p]) {
                    double BD[18];
                    for (int a = 0; a < 6; a++)
                        for (int c = 0; c < 3; c++)
                            BD[3 * a + c] = qi.second[3 * a] * Di[c] + qi.second[3 * a + 1] * Di[3 + c] + qi.second[3 * a + 2] * Di[6 + c];
                    for (int a = 0; a < 6; a++)
                        bs[6 * qi.first + a] -= qi.second[3 * a] * db[0] + qi.second[3 * a + 1] * db[1] + qi.second[3 * a + 2] * db[2];
                    for (auto& qj : Hpl[p])
                        for (int a = 0; a < 6; a++)
                            for (int c = 0; c < 6; c++)
                                S[(size_t)(6 * qi.first + a) * n + 6 * qj.first + c] -=
                                    BD[3 * a] * qj.second[3 * c] + BD[3 * a + 1] * qj.second[3 * c + 1] + BD[3 * a + 2] * qj.second[3 * c + 2];
                }
            }
            std::vector<double> xp;
            const bool ok2 = llt_solve(n, S, bs, xp);
            std::vector<double> xl((size_t)P * 3, 0.0);
            if (ok2) {
                for (int p = 0; p < P; p++) {
                    if (!pfree(p)) continue;
                    double c[3] = {bl[3 * p], bl[3 * p + 1], bl[3 * p + 2]};
                    for (auto& q : Hpl[p])
                        for (int a = 0; a < 3; a++)
                            for (int k = 0; k < 6; k++) c[a] -= q.second[3 * k + a] * xp[6 * q.first + k];
                    const double* Di = &Dinv[9 * (size_t)p];
                    for (int a = 0; a < 3; a++) xl[3 * p + a] = Di[3 * a] * c[0] + Di[3 * a + 1] * c[1] + Di[3 * a + 2] * c[2];
                }
                if (x_out && it == 0 && qmax == 0) {
                    for (int k = 0; k < n; k++) x_out[k] = xp[k];
                    for (size_t k = 0; k < xl.size(); k++) x_out[n + k] = xl[k];
                }
                for (int f = 0; f < F; f++)
                    if (fidx[f] >= 0) orc_se3_oplus(frames + 7 * f, &xp[6 * fidx[f]]);
                for (int p = 0; p < P; p++)
                    if (pfree(p))
                        for (int a = 0; a < 3; a++) points[3 * p + a] += xl[3 * p + a];
            }
            std::vector<LinEdge> L2;
            lm_linearize(E, points, frames, L2);
            double tempChi = lm_robust_chi2(E, L2);
            if (!ok2) tempChi = DBL_MAX;
            double scale = 0;
            if (ok2) {
                for (int k = 0; k < n; k++) scale += xp[k] * (lambda * xp[k] + bp[k]);
                for (int p = 0; p < P; p++)
                    if (pfree(p))
                        for (int a = 0; a < 3; a++) scale += xl[3 * p + a] * (lambda * xl[3 * p + a] + bl[3 * p + a]);
            }
            scale += 1e-3;
            rho = (currentChi - tempChi) / scale;
            if (rho > 0 && std::isfinite(tempChi)) {
                double alpha = 1. - std::pow(2 * rho - 1, 3);
                alpha = std::min(alpha, 2. / 3.);
                lambda *= std::max(1. / 3., alpha);
                ni = 2;
                currentChi = tempChi;
            } else {
                lambda *= ni;
                ni *= 2;
                std::copy(pts0.begin(), pts0.end(), points);
                std::copy(frs0.begin(), frs0.end(), frames);
                if (!std::isfinite(lambda)) break;
            }
            qmax++;
        } while (rho < 0 && qmax < max_trials);
        if (stats) {
            stats[4 * it + 1] = currentChi;
            stats[4 * it + 2] = lambda;
            stats[4 * it + 3] = qmax;
        }
        done++;
        if (qmax == max_trials || rho == 0 || !std::isfinite(lambda)) break;
    }
    return done;
}

// Sum over active edges of the Huber rho of chi2 at the given state
// (activeRobustChi2 after computeActiveErrors), FP64.
double orc_ba_robust_chi2(int M, const float* uv, const int32_t* pt, const int32_t* fr, const int32_t* cam,
                          const float* isig, const uint8_t* act, int Ms, const float* sobs, const int32_t* spt,
                          const int32_t* sfr, const float* sisig, const uint8_t* sact, const double* scam, int P, int F,
                          int C, const double* cams, int robust, const double* points, const double* frames) {
    LMEdges E{M, uv, pt, fr, cam, isig, act, Ms, sobs, spt, sfr, sisig, sact, scam, P, F, C, cams, nullptr, nullptr,
              robust};
    std::vector<LinEdge> L;
    lm_linearize(E, points, frames, L);
    return lm_robust_chi2(E, L);
}

}  // extern "C"

extern "C" {

// ---------------------------------------------------------------------------
// DBoW3 (modules/DBow3/src, vendored by the reference).  Vocabulary tree:
// node 0 is the root; children of node i are child_ids[child_ptr[i] ..
// child_ptr[i + 1]) in DBoW3's children order; a node without children is a
// leaf with word_id[i] and weight[i].
//   Vocabulary::transform(feature, word_id, weight, nid, levelsup)
//   (Vocabulary.cpp:641-672): from the root, the child at the least Hamming
//   distance (DescManip::distance, DescManip.cpp:92-119; the first of equal
//   distances, strict '<'), to a leaf; nid = the node reached at level
//   L - levelsup (the root when that level is <= 0; left 0 when the path ends
//   above it).
void orc_bow_transform(const uint8_t* node_desc, const int32_t* child_ptr, const int32_t* child_ids,
                       const int32_t* word_id, const double* weight, int L, int levelsup, const uint8_t* feats, int n,
                       int32_t* out_word, double* out_weight, int32_t* out_nid) {
    const int nid_level = L - levelsup;
    for (int f = 0; f < n; f++) {
        const uint8_t* x = feats + 32 * (size_t)f;
        int nid = 0;
        int final_id = 0, level = 0;
        do {
            ++level;
            double best_d = DBL_MAX;
            const int c0 = child_ptr[final_id], c1 = child_ptr[final_id + 1];  // the node's children, fixed
            for (int c = c0; c < c1; c++) {
                const int id = child_ids[c];
                const double d = (double)hamming(x, node_desc + 32 * (size_t)id);
                if (d < best_d) {
                    best_d = d;
                    final_id = id;
                }
            }
            if (level == nid_level) nid = final_id;
        } while (child_ptr[final_id + 1] > child_ptr[final_id]);
        out_word[f] = word_id[final_id];
        out_weight[f] = weight[final_id];
        out_nid[f] = nid;
    }
}

// Vocabulary::transform(features, BowVector, FeatureVector, levelsup)
// (Vocabulary.cpp:572-629) from the per-feature transforms: weighting 0
// TF_IDF, 1 TF (BowVector::addWeight, then / v.size() when the scoring does
// not normalise), 2 IDF, 3 BINARY (addIfNotExist); words with weight <= 0
// are stopped; FeatureVector::addFeature(nid, i) (FeatureVector.cpp:31-45).
// scoring 0 L1, 1 L2, 2 chi-square, 3 KL, 4 Bhattacharyya (normalised: L2
// for scoring 1, else L1; BowVector::normalize, BowVector.cpp:62-84), 5 dot
// product (not normalised).  Outputs sorted by word / node id; returns the
// word count, *n_nodes the node count.
int orc_bow_vectors(const int32_t* word, const double* w, const int32_t* nid, int n, int weighting, int scoring,
                    int32_t* bv_words, double* bv_vals, int32_t* fv_nodes, int32_t* fv_ptr, int32_t* fv_feats,
                    int32_t* n_nodes) {
    std::map<int, double> v;
    std::map<int, std::vector<int>> fv;
    const bool must = scoring != 5;
    const bool l2 = scoring == 1;
    for (int i = 0; i < n; i++) {
        if (!(w[i] > 0)) continue;
        if (weighting == 0 || weighting == 1) {
            v[word[i]] += w[i];
        } else if (v.find(word[i]) == v.end()) {
            v[word[i]] = w[i];
        }
        fv[nid[i]].push_back(i);
    }
    if ((weighting == 0 || weighting == 1) && !v.empty() && !must) {
        const double nd = (double)v.size();
        for (auto& kv : v) kv.second /= nd;
    }
    if (must) {
        double norm = 0.0;
        if (!l2) {
            for (auto& kv : v) norm += std::fabs(kv.second);
        } else {
            for (auto& kv : v) norm += kv.second * kv.second;
            norm = std::sqrt(norm);
        }
        if (norm > 0.0)
            for (auto& kv : v) kv.second /= norm;
    }
    int k = 0;
    for (auto& kv : v) {
        bv_words[k] = kv.first;
        bv_vals[k] = kv.second;
        k++;
    }
    int m = 0, t = 0;
    fv_ptr[0] = 0;
    for (auto& kv : fv) {
        fv_nodes[m] = kv.first;
        for (int i : kv.second) fv_feats[t++] = i;
        fv_ptr[++m] = t;
    }
    *n_nodes = m;
    return k;
}

// Matcher::DBowMatch (src/Matcher.cpp:146-193): over the nodes both
// FeatureVectors hold (ascending), for each feature of image 1 in the node,
// the two nearest features of image 2 in the node (strict '<' in list order,
// d = {999, 999}); a row (q, i0, d0, i1, d1) when a second one exists.
int orc_bow_match(const uint8_t* d1, int m1, const int32_t* nodes1, const int32_t* ptr1, const int32_t* feats1,
                  const uint8_t* d2, int m2, const int32_t* nodes2, const int32_t* ptr2, const int32_t* feats2,
                  int32_t* out) {
    int a = 0, b = 0, rows = 0;
    while (a < m1 && b < m2) {
        if (nodes1[a] == nodes2[b]) {
            for (int s = ptr1[a]; s < ptr1[a + 1]; s++) {
                const int q = feats1[s];
                unsigned d[2] = {999, 999};
                int di[2] = {0, 0};
                for (int t = ptr2[b]; t < ptr2[b + 1]; t++) {
                    const int j = feats2[t];
                    const unsigned dist = hamming(d1 + 32 * (size_t)q, d2 + 32 * (size_t)j);
                    if (dist < d[0]) {
                        d[1] = d[0];
                        di[1] = di[0];
                        d[0] = dist;
                        di[0] = j;
                    } else if (dist < d[1]) {
                        d[1] = dist;
                        di[1] = j;
                    }
                }
                if (d[1] != 999) {
                    int32_t* o = out + 5 * rows++;
                    o[0] = q; o[1] = di[0]; o[2] = (int)d[0]; o[3] = di[1]; o[4] = (int)d[1];
                }
            }
            a++;
            b++;
        } else if (nodes1[a] < nodes2[b]) {
            a = (int)(std::lower_bound(nodes1 + a, nodes1 + m1, nodes2[b]) - nodes1);
        } else {
            b = (int)(std::lower_bound(nodes2 + b, nodes2 + m2, nodes1[a]) - nodes2);
        }
    }
    return rows;
}

}  // extern "C"
