// Host side of the C ABI declared in include/mcvslam_gpu.h: contexts, the
// ORB batch planner, the libm sin/cos calibration table, graph capture, and
// the launchers of every kernel.  Kernels live in the included .hip files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <gnu/libc-version.h>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../include/mcvslam_gpu.h"
#include "mcv_ba.hip"
#include "mcv_bow.hip"
#include "mcv_match.hip"
#include "mcv_orb.hip"
#include "mcv_window.hip"
#include "mcv_stereo.hip"

using namespace mcv;

namespace mcv {
// clears counters and flags; used instead of hipMemsetAsync so that the
// sequence stays safe to capture into a graph
__global__ void k_zero_words(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
}  // namespace mcv

// Scratch of mcv_match_filter_noalloc_dev: the ratio survivors and the
// threshold holes of one list, 2 * kNoAllocMaxQueries ints (512 KiB) per GPU,
// held until the process ends.  After set-up host code uses only the device
// address that hipGetSymbolAddress returns for it (once per device, in
// mcv_create); the array's name is not a device address in host code.
constexpr int kNoAllocMaxQueries = 65536;
static __device__ __align__(256) int scratch[2 * kNoAllocMaxQueries];

namespace {

struct Fail {
    int code;
    std::string msg;
};

#define HIP_OK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) throw Fail{MCV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define REQUIRE(cond, code, msg)                   \
    do {                                           \
        if (!(cond)) throw Fail{(code), (msg)};    \
    } while (0)

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) HIP_OK(hipFree(p));
        p = nullptr;
        bytes = 0;
        if (n == 0) return;
        HIP_OK(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
};

template <typename T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    b.ensure(std::max<size_t>(v.size() * sizeof(T), 16));
    if (!v.empty()) HIP_OK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// ---------------------------------------------------------------------------
// cosf / sinf calibration.  The descriptor rotation uses the host libm's
// single-precision cosf/sinf (ORBextractor.cc:102-103).  The device computes
// sincos_base(); this table lists every float input in [0, 6.3] where that
// differs from the host libm, with the libm value.  Built once per process
// with all host threads and cached on disk keyed by the libm fingerprint.
struct SinCosHost {
    std::vector<uint32_t> ck, sk, cidx, sidx;
    std::vector<float> cv, sv;
    uint32_t max_bits = 0;
};

static uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Key of the cached correction table: the libm's cosf/sinf and this build's
// sincos_base at the same sample points, so a table made by another libm or
// another build of sincos_base (source, FP flags) is never reused.
static uint64_t libm_fingerprint() {
    uint64_t h = 1469598103934665603ull;
    const char* v = gnu_get_libc_version();
    h = fnv1a(h, v, strlen(v));
    for (uint32_t u = 0; u < 0x40C9999Au; u += 0x3FFF) {
        float x;
        memcpy(&x, &u, 4);
        const float c = cosf(x), s = sinf(x);
        h = fnv1a(h, &c, 4);
        h = fnv1a(h, &s, 4);
        float bc, bs;
        sincos_base(x, &bc, &bs);
        h = fnv1a(h, &bc, 4);
        h = fnv1a(h, &bs, 4);
    }
    return h;
}

static void build_index(const std::vector<uint32_t>& keys, std::vector<uint32_t>& idx) {
    idx.assign(65537, 0);
    size_t p = 0;
    for (uint32_t h = 0; h <= 65536; h++) {
        while (p < keys.size() && (keys[p] >> 16) < h) p++;
        idx[h] = (uint32_t)p;
    }
}

static std::string cache_path(uint64_t fp) {
    const char* dir = getenv("MCV_CACHE_DIR");
    std::string d;
    if (dir && *dir) {
        d = dir;
    } else {
        const char* home = getenv("HOME");
        d = std::string(home && *home ? home : "/tmp") + "/.cache/mcvslam";
    }
    char name[64];
    snprintf(name, sizeof name, "/sincos_v2_%016llx.bin", (unsigned long long)fp);
    return d + name;
}

static bool load_cache(const std::string& path, SinCosHost& t) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    uint64_t nc = 0, ns = 0;
    f.read(reinterpret_cast<char*>(&t.max_bits), 4);
    f.read(reinterpret_cast<char*>(&nc), 8);
    f.read(reinterpret_cast<char*>(&ns), 8);
    if (!f || nc > (1u << 26) || ns > (1u << 26)) return false;
    t.ck.resize(nc);
    t.cv.resize(nc);
    t.sk.resize(ns);
    t.sv.resize(ns);
    f.read(reinterpret_cast<char*>(t.ck.data()), nc * 4);
    f.read(reinterpret_cast<char*>(t.cv.data()), nc * 4);
    f.read(reinterpret_cast<char*>(t.sk.data()), ns * 4);
    f.read(reinterpret_cast<char*>(t.sv.data()), ns * 4);
    return (bool)f;
}

static void save_cache(const std::string& path, const SinCosHost& t) {
    std::string dir = path.substr(0, path.rfind('/'));
    std::string cmd_dir = dir;
    // create the directory chain without a shell
    for (size_t i = 1; i <= cmd_dir.size(); i++)
        if (i == cmd_dir.size() || cmd_dir[i] == '/') {
            std::string sub = cmd_dir.substr(0, i);
            (void)mkdir(sub.c_str(), 0755);
        }
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    uint64_t nc = t.ck.size(), ns = t.sk.size();
    f.write(reinterpret_cast<const char*>(&t.max_bits), 4);
    f.write(reinterpret_cast<const char*>(&nc), 8);
    f.write(reinterpret_cast<const char*>(&ns), 8);
    f.write(reinterpret_cast<const char*>(t.ck.data()), nc * 4);
    f.write(reinterpret_cast<const char*>(t.cv.data()), nc * 4);
    f.write(reinterpret_cast<const char*>(t.sk.data()), ns * 4);
    f.write(reinterpret_cast<const char*>(t.sv.data()), ns * 4);
    f.close();
    if (f) (void)rename(tmp.c_str(), path.c_str());
}

static const SinCosHost& sincos_host() {
    static std::once_flag once;
    static SinCosHost t;
    std::call_once(once, [] {
        const float xmax = 6.3f;
        memcpy(&t.max_bits, &xmax, 4);
        const uint64_t fp = libm_fingerprint();
        const std::string path = cache_path(fp);
        SinCosHost c;
        if (load_cache(path, c) && c.max_bits == t.max_bits) {
            t.ck.swap(c.ck); t.cv.swap(c.cv); t.sk.swap(c.sk); t.sv.swap(c.sv);
        } else {
            unsigned nt = std::thread::hardware_concurrency();
            if (const char* e = getenv("MCV_SINCOS_THREADS")) nt = (unsigned)atoi(e);
            nt = std::max(1u, std::min(nt ? nt : 1u, 16u));
            const uint32_t nchunk = 256;
            std::vector<std::vector<uint32_t>> ck(nchunk), sk(nchunk);
            std::vector<std::vector<float>> cvv(nchunk), svv(nchunk);
            std::atomic<uint32_t> next{0};
            const uint64_t total = (uint64_t)t.max_bits + 1;
            auto work = [&] {
                for (;;) {
                    const uint32_t ci = next.fetch_add(1);
                    if (ci >= nchunk) return;
                    const uint64_t a = total * ci / nchunk, b = total * (ci + 1) / nchunk;
                    for (uint64_t u = a; u < b; u++) {
                        const uint32_t bits = (uint32_t)u;
                        float x;
                        memcpy(&x, &bits, 4);
                        float bc, bs;
                        sincos_base(x, &bc, &bs);
                        const float lc = cosf(x), ls = sinf(x);
                        if (memcmp(&bc, &lc, 4) != 0) { ck[ci].push_back(bits); cvv[ci].push_back(lc); }
                        if (memcmp(&bs, &ls, 4) != 0) { sk[ci].push_back(bits); svv[ci].push_back(ls); }
                    }
                }
            };
            std::vector<std::thread> th;
            for (unsigned i = 0; i < nt; i++) th.emplace_back(work);
            for (auto& x : th) x.join();
            for (uint32_t i = 0; i < nchunk; i++) {
                t.ck.insert(t.ck.end(), ck[i].begin(), ck[i].end());
                t.cv.insert(t.cv.end(), cvv[i].begin(), cvv[i].end());
                t.sk.insert(t.sk.end(), sk[i].begin(), sk[i].end());
                t.sv.insert(t.sv.end(), svv[i].begin(), svv[i].end());
            }
            save_cache(path, t);
        }
        build_index(t.ck, t.cidx);
        build_index(t.sk, t.sidx);
    });
    return t;
}

struct SinCosDev {
    DevBuf ck, cv, cidx, sk, sv, sidx;
    bool ready = false;
    SinCosTables view() const {
        SinCosTables v;
        v.cos_key = ck.as<uint32_t>();
        v.cos_val = cv.as<float>();
        v.cos_idx = cidx.as<uint32_t>();
        v.sin_key = sk.as<uint32_t>();
        v.sin_val = sv.as<float>();
        v.sin_idx = sidx.as<uint32_t>();
        v.max_bits = sincos_host().max_bits;
        return v;
    }
};

// ---------------------------------------------------------------------------
// ORBextractor::init tables (ORBextractor.cc:407-457).
struct OrbConfig {
    int nfeatures = 0, nlevels = 0, ini = 0, mn = 0;
    float scale_factor = 1.f;
    std::vector<float> scale, inv_scale, sigma2, inv_sigma2;
    std::vector<int> npl;
    int umax[16];
    int gauss[7];
};

static void gauss_taps(int* taps) {
    // GaussianBlur(Size(7,7), 2, 2) on CV_8U: normalised Gaussian, error-
    // diffused to Q8, centre tap takes the remainder (ORBextractor.cc:875)
    const int n = 7;
    const double sigma = 2.0;
    const double scale2X = -0.125 / (sigma * sigma);
    const int n2 = (n - 1) / 2;
    double values[4], sum = 0;
    for (int i = 0, x = 1 - n; i < n2; i++, x += 2) {
        values[i] = std::exp((double)(x * x) * scale2X);
        sum += values[i];
    }
    sum = sum * 2 + 1.0;
    const double mul1 = 1.0 / sum;
    double err = 0;
    int s = 0;
    for (int i = 0; i < n2; i++) {
        const double adj = values[i] * mul1 * 256.0 + err;
        const int v0 = (int)lrint(adj);
        err = adj - v0;
        taps[i] = taps[n - 1 - i] = v0;
        s += v0;
    }
    taps[n2] = 256 - 2 * s;
}

// mnFeaturesPerLevel for nf features (ORBextractor.cc:426-436)
static std::vector<int> feature_quotas(int nf, float sf, int nl) {
    std::vector<int> npl(nl, 0);
    const float factor = 1.0f / sf;
    float nd = nf * (1 - factor) / (1 - (float)std::pow((double)factor, (double)nl));
    int sumF = 0;
    for (int l = 0; l < nl - 1; l++) {
        npl[l] = (int)lrintf(nd);
        sumF += npl[l];
        nd *= factor;
    }
    npl[nl - 1] = std::max(nf - sumF, 0);
    return npl;
}

static void orb_init(OrbConfig& c, int nf, float sf, int nl, int ini, int mn) {
    c.nfeatures = nf; c.scale_factor = sf; c.nlevels = nl; c.ini = ini; c.mn = mn;
    c.scale.assign(nl, 0.f); c.sigma2.assign(nl, 0.f); c.inv_scale.assign(nl, 0.f); c.inv_sigma2.assign(nl, 0.f);
    c.scale[0] = 1.0f; c.sigma2[0] = 1.0f;
    for (int i = 1; i < nl; i++) {
        c.scale[i] = c.scale[i - 1] * sf;
        c.sigma2[i] = c.scale[i] * c.scale[i];
    }
    for (int i = 0; i < nl; i++) {
        c.inv_scale[i] = 1.0f / c.scale[i];
        c.inv_sigma2[i] = 1.0f / c.sigma2[i];
    }
    c.npl = feature_quotas(nf, sf, nl);
    int v, v0, vmax = (int)std::floor(15 * std::sqrt(2.f) / 2 + 1);
    int vmin = (int)std::ceil(15 * std::sqrt(2.f) / 2);
    const double hp2 = 15 * 15;
    for (v = 0; v <= vmax; ++v) c.umax[v] = (int)lrint(std::sqrt(hp2 - v * v));
    for (v = 15, v0 = 0; v >= vmin; --v) {
        while (c.umax[v0] == c.umax[v0 + 1]) ++v0;
        c.umax[v] = v0;
        ++v0;
    }
    gauss_taps(c.gauss);
}

// resize coefficients for one axis (cv::resize INTER_LINEAR, fixed point)
static void resize_axis(int src, int dst, std::vector<int2>& out) {
    const double inv = (double)dst / src;
    const double scale = 1. / inv;
    for (int d = 0; d < dst; d++) {
        float f = (float)((d + 0.5) * scale - 0.5);
        int s = (int)std::floor(f);
        f -= s;
        if (s < 0) { f = 0; s = 0; }
        if (s + 1 >= src && s >= src - 1) { f = 0; s = src - 1; }
        const int a0 = std::min(32767, std::max(-32768, (int)lrintf((1.f - f) * 2048)));
        const int a1 = std::min(32767, std::max(-32768, (int)lrintf(f * 2048)));
        out.push_back(make_int2(s, (a0 & 0xFFFF) | (a1 << 16)));
    }
}

// keypoint slots the octree of one level may fill: DistributeOctTree stops
// below N + 3 nodes, or emits all roots when they alone reach N
static int level_out_cap(int nfeat, int n_ini) { return std::max(nfeat + 3, n_ini + 1); }

// octree root count of a level (ORBextractor.cc:527), 1 when it has no cells
static int level_roots(int w, int h) {
    const float width = (float)(w - 32), height = (float)(h - 32);
    const int nCols = (int)(width / 35.f), nRows = (int)(height / 35.f);
    if (!(nCols > 0 && nRows > 0 && width > 0 && height > 0)) return 1;
    return (int)std::round(width / height);
}

// Launch choices that leave every output bit unchanged; picked by the
// autotuner (mcv_rig_autotune) from timings on the real workload.
// Pieces of OrbPlan::enqueue.  kPartPre: everything up to the octree build
// except the blur; kPartPreA + kPartFastLo + kPartPreB + kPartPreC: the same
// split so that FAST on the large levels (below kFastSplitLevel) can run
// beside the resize of the small ones: the copy and the resizes below the
// split level, FAST on the levels below it, the remaining resizes and FAST
// on those levels, then the cell suppression and the octree build.
enum EnqPart {
    kPartAll = 0,
    kPartPre = 1,
    kPartBlur = 2,
    kPartHeap = 3,
    kPartPost = 4,
    kPartPreA = 5,
    kPartFastLo = 6,
    kPartPreB = 7,
    kPartPreC = 8
};
constexpr int kFastSplitLevel = 3;

struct Tuning {
    int knn = 0;   // launch_knn2 variant
    int blur = 0;  // 0: k_blur<false>, 1: k_blur<true>
    int kp = 2;    // keypoints per wave: 0: one (k_keypoints), 1: 8, 2: 16, 3: 32 (k_keypoints_k<K>)
    int heap = 1;  // 0: k_octree_heap_r (register top), 1: k_octree_heap (LDS), the faster one measured
    bool operator==(const Tuning& o) const { return knn == o.knn && blur == o.blur && kp == o.kp && heap == o.heap; }
};

struct ImageSpec {
    int w, h;
    std::vector<std::vector<mcv_keypoint>> seeds;  // per level, in input order
    int nfeatures = 0;                              // 0: the configured count
};

// A batch plan: geometry and work tables for a list of images.
struct OrbPlan {
    int n_images = 0, nlevels = 0, cap = 0, n_slots = 0;
    std::vector<LevelDesc> lv;
    std::vector<CellDesc> cells;
    std::vector<Strip> strips;
    std::vector<int> strip_begin;  // per level index l: [strip_begin[l], strip_begin[l+1])
    std::vector<int2> coef;
    std::vector<int> slot_lvl;
    std::vector<float> seeds;
    long long pyr_bytes = 0, pool_cap = 0;
    int n_cand = 0, max_slots_img = 0;
    int heap_cap = 0;   // octree heap entries (dynamic LDS of k_octree_heap)
    std::vector<int> heap_order;  // levels by feature target, largest first
    std::vector<int> build_levels;  // large levels (1024-thread octree build) first, then the rest
    int n_build_large = 0;
    DevBuf d_heap_order, d_build_levels, d_lv, d_cells, d_strips, d_coef, d_slot, d_seeds, d_pyr, d_blur, d_score, d_cell_cand, d_cell_n, d_lvl_cand, d_oct,
        d_pool, d_used, d_nodes, d_perm, d_src, d_stride;
    hipGraphExec_t graph = nullptr;
    // the captured graph bakes in every output pointer and the capacity
    struct GraphKey {
        const void* kps = nullptr;
        const void* desc = nullptr;
        const void* counts = nullptr;
        int cap = -1;
        bool operator==(const GraphKey& o) const {
            return kps == o.kps && desc == o.desc && counts == o.counts && cap == o.cap;
        }
    } graph_key;
    Tuning tune;

    ~OrbPlan() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
    }

    void build(const OrbConfig& c, const std::vector<ImageSpec>& imgs, long long pool_words) {
        n_images = (int)imgs.size();
        nlevels = c.nlevels;
        lv.clear(); cells.clear(); strips.clear(); coef.clear(); slot_lvl.clear(); seeds.clear();
        pyr_bytes = 0;
        n_cand = 0;
        int slot = 0;
        std::vector<std::vector<int>> lvl_of(n_images, std::vector<int>(nlevels));
        for (int i = 0; i < n_images; i++) {
            const ImageSpec& im = imgs[i];
            REQUIRE(im.w > 0 && im.h > 0, MCV_ERR_EMPTY, "empty image");
            REQUIRE(im.w <= 4096 && im.h <= 4096, MCV_ERR_INVALID, "image larger than 4096 px");
            int first = (int)lv.size();
            int img_slots = 0;
            const std::vector<int> npl = im.nfeatures > 0 ? feature_quotas(im.nfeatures, c.scale_factor, nlevels) : c.npl;
            for (int l = 0; l < nlevels; l++) {
                LevelDesc d{};
                d.img = i;
                d.level = l;
                d.w = l == 0 ? im.w : (int)lrintf((float)im.w * c.inv_scale[l]);
                d.h = l == 0 ? im.h : (int)lrintf((float)im.h * c.inv_scale[l]);
                REQUIRE(d.w >= 8 && d.h >= 8, MCV_ERR_INVALID, "pyramid level smaller than 8 px");
                d.src = l == 0 ? -1 : (int)lv.size() - 1;
                d.nfeat = npl[l];
                d.pitch = (d.w + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
                d.pix = pyr_bytes;
                pyr_bytes += ((long long)d.pitch * d.h + 255) / 256 * 256;
                d.coef_x = d.coef_y = 0;
                if (l > 0) {
                    const LevelDesc& s = lv[d.src];
                    d.coef_x = (int)coef.size();
                    resize_axis(s.w, d.w, coef);
                    d.coef_y = (int)coef.size();
                    resize_axis(s.h, d.h, coef);
                }
                // ComputeKeyPointsOctTree cell grid (ORBextractor.cc:585-633)
                d.minBX = 16; d.minBY = 16;
                d.maxBX = d.w - 16; d.maxBY = d.h - 16;
                const float width = (float)(d.maxBX - d.minBX), height = (float)(d.maxBY - d.minBY);
                const int nCols = (int)(width / 35.f), nRows = (int)(height / 35.f);
                d.cell_first = (int)cells.size();
                d.cand_base = n_cand;
                if (nCols > 0 && nRows > 0 && width > 0 && height > 0) {
                    const int wCell = (int)std::ceil(width / nCols), hCell = (int)std::ceil(height / nRows);
                    for (int r = 0; r < nRows; r++) {
                        const float iniY = (float)(d.minBY + r * hCell);
                        float maxY = iniY + hCell + 6;
                        if (iniY >= d.maxBY - 3) continue;
                        if (maxY > d.maxBY) maxY = (float)d.maxBY;
                        for (int j = 0; j < nCols; j++) {
                            const float iniX = (float)(d.minBX + j * wCell);
                            float maxX = iniX + wCell + 6;
                            if (iniX >= d.maxBX - 6) continue;
                            if (maxX > d.maxBX) maxX = (float)d.maxBX;
                            CellDesc cd{};
                            cd.lvl = (int)lv.size();
                            cd.x0 = (int)iniX; cd.y0 = (int)iniY; cd.x1 = (int)maxX; cd.y1 = (int)maxY;
                            cd.offx = j * wCell; cd.offy = r * hCell;
                            const int W = cd.x1 - cd.x0, H = cd.y1 - cd.y0;
                            REQUIRE(W <= kMaxWin && H <= kMaxWin, MCV_ERR_INVALID, "FAST cell window too large");
                            cd.cap_base = n_cand;
                            const int cap = (W >= 7 && H >= 7) ? ((W - 6 + 1) / 2) * ((H - 6 + 1) / 2) : 0;
                            n_cand += cap;
                            cells.push_back(cd);
                        }
                    }
                    d.n_ini = (int)std::round(width / (float)(d.maxBY - d.minBY));
                    REQUIRE(d.n_ini >= 1 && d.n_ini <= 15, MCV_ERR_INVALID, "image aspect ratio outside the octree root range");
                    d.hX = width / (float)d.n_ini;
                } else {
                    d.n_ini = 1;
                    d.hX = 1.f;
                }
                d.cell_count = (int)cells.size() - d.cell_first;
                d.cand_cap = n_cand - d.cand_base;
                REQUIRE(d.nfeat <= 4000, MCV_ERR_INVALID, "more than 4000 features on one level");
                d.out_cap = level_out_cap(d.nfeat, d.n_ini);
                d.seed_first = (int)(seeds.size() / 6);
                d.seed_count = (int)im.seeds[l].size();
                for (const mcv_keypoint& k : im.seeds[l]) {
                    seeds.insert(seeds.end(), {k.x, k.y, k.size, k.angle, k.response, (float)k.octave});
                }
                d.kp_slot = slot;
                for (int j = 0; j < d.out_cap + d.seed_count; j++) {
                    slot_lvl.push_back((int)lv.size());
                    slot_lvl.push_back(j);
                }
                slot += d.out_cap + d.seed_count;
                img_slots += d.out_cap + d.seed_count;
                d.scale = c.scale[l];
                d.patch = (float)(int)(31 * c.scale[l]);
                d.first_level = first;
                d.nlevels = nlevels;
                lvl_of[i][l] = (int)lv.size();
                lv.push_back(d);
            }
            max_slots_img = std::max(max_slots_img, img_slots);
        }
        n_slots = slot;
        cap = max_slots_img;
        heap_cap = 16;
        for (const LevelDesc& d : lv) heap_cap = std::max(heap_cap, std::max(d.nfeat, d.n_ini) + 16);
        // levels of at least kOctLargePixels pixels (the most FAST
        // candidates) get 1024-thread octree workgroups
        build_levels.clear();
        for (size_t k = 0; k < lv.size(); k++)
            if ((long long)lv[k].w * lv[k].h >= kOctLargePixels) build_levels.push_back((int)k);
        n_build_large = (int)build_levels.size();
        for (size_t k = 0; k < lv.size(); k++)
            if ((long long)lv[k].w * lv[k].h < kOctLargePixels) build_levels.push_back((int)k);
        // largest levels first in each launch: when a launch takes more than
        // one round of workgroups, the later rounds get the small levels
        auto by_area = [&](int a, int b) { return (long long)lv[a].w * lv[a].h > (long long)lv[b].w * lv[b].h; };
        std::stable_sort(build_levels.begin(), build_levels.begin() + n_build_large, by_area);
        std::stable_sort(build_levels.begin() + n_build_large, build_levels.end(), by_area);
        heap_order.resize(lv.size());
        for (size_t k = 0; k < lv.size(); k++) heap_order[k] = (int)k;
        std::stable_sort(heap_order.begin(), heap_order.end(),
                         [&](int a, int b) { return lv[a].nfeat > lv[b].nfeat; });
        // strips grouped by level index, then image
        strip_begin.assign(nlevels + 1, 0);
        for (int l = 0; l < nlevels; l++) {
            strip_begin[l] = (int)strips.size();
            for (int i = 0; i < n_images; i++) {
                const LevelDesc& d = lv[lvl_of[i][l]];
                for (int y = 0; y < d.h; y += kStripH)
                    for (int x = 0; x < d.w; x += kStripW) strips.push_back(Strip{lvl_of[i][l], x, y});
            }
        }
        strip_begin[nlevels] = (int)strips.size();
        // k_resize: the source bytes of 4 adjacent outputs (both taps)
        // lie in the 12 bytes from the first one's aligned word
        for (const LevelDesc& d : lv) {
            if (d.level == 0) continue;
            const LevelDesc& sl = lv[d.src];
            for (int x = 0; x < d.w; x += 4) {
                const int wb = coef[d.coef_x + x].x & ~3;
                for (int j = 0; j < 4; j++) {
                    const int off = coef[d.coef_x + std::min(x + j, d.w - 1)].x - wb;
                    REQUIRE(off >= 0 && off + 1 < 12 && coef[d.coef_x + std::min(x + j, d.w - 1)].x < sl.w,
                            MCV_ERR_INVALID, "scale factor too large for the resize taps");
                }
            }
        }
        pool_cap = pool_words;
    }

    void upload_all(hipStream_t s) {
        upload(d_lv, lv, s);
        upload(d_cells, cells, s);
        upload(d_strips, strips, s);
        upload(d_heap_order, heap_order, s);
        upload(d_build_levels, build_levels, s);
        upload(d_coef, coef, s);
        upload(d_slot, slot_lvl, s);
        upload(d_seeds, seeds, s);
        d_pyr.ensure(pyr_bytes + 256);
        d_blur.ensure(pyr_bytes + 256);
        d_score.ensure(pyr_bytes + 256);
        d_cell_cand.ensure((size_t)std::max(n_cand, 1) * 4);
        d_cell_n.ensure((size_t)std::max<size_t>(cells.size(), 1) * 4);
        d_lvl_cand.ensure((size_t)std::max(n_cand, 1) * 4);
        d_oct.ensure(lv.size() * sizeof(OctLevel));
        d_pool.ensure((size_t)pool_cap * 4);
        d_used.ensure(64);
        HIP_OK(hipMemsetAsync(d_used.p, 0, 64, s));  // set-up only, never captured
        d_nodes.ensure((size_t)std::max(n_slots, 1) * 4);
        d_perm.ensure((size_t)std::max(n_slots, 1) * 4);
        d_src.ensure((size_t)n_images * sizeof(void*));
        d_stride.ensure((size_t)n_images * sizeof(int));
    }

    OrbGlobals globals(const OrbConfig& c) const {
        OrbGlobals g{};
        g.lv = d_lv.as<LevelDesc>();
        g.cells = d_cells.as<CellDesc>();
        g.pyr = d_pyr.as<uint8_t>();
        g.blur = d_blur.as<uint8_t>();
        g.score = d_score.as<uint8_t>();
        g.cell_cand = d_cell_cand.as<uint32_t>();
        g.cell_n = d_cell_n.as<int>();
        g.lvl_cand = d_lvl_cand.as<uint32_t>();
        g.oct = d_oct.as<OctLevel>();
        g.pool = d_pool.as<uint32_t>();
        g.pool_used = d_used.as<unsigned long long>();
        g.pool_cap = pool_cap;
        g.nodes_out = d_nodes.as<uint32_t>();
        g.seeds = d_seeds.as<float>();
        for (int i = 0; i < 16; i++) g.umax[i] = c.umax[i];
        g.ini_th = std::min(std::max(c.ini, 0), 255);
        g.min_th = std::min(std::max(c.mn, 0), 255);
        g.err = reinterpret_cast<int*>(d_used.as<char>() + 8);
        g.perm_out = d_perm.as<int>();
        g.perm = d_perm.as<int>();
        return g;
    }

    // The whole extraction for the batch, enqueued on s.  With `rec`, an
    // event is recorded after every launch (kernel timing, never in a graph).
    // reset_err = false keeps the error word of earlier enqueues (the plain
    // rig path reuses one plan per camera and checks once per run); the pool
    // counter is reset by every enqueue.
    // side: a second stream for graph capture; with the blur variants that
    // run beside the heap, the blur forks onto it after the octree build and
    // joins before the keypoint kernel.
    // part: the whole sequence, or one piece of it for a caller that runs the
    // blur beside the heap on streams of its own: kPartPre (everything up to
    // the octree build, no blur), kPartBlur, kPartHeap, kPartPost (keypoints
    // and descriptors)
    void enqueue(const OrbConfig& c, const SinCosTables& tb, hipStream_t s, mcv_keypoint* d_kps, uint8_t* d_desc,
                 int out_cap, int* d_counts, struct KernelTimer* rec = nullptr, bool reset_err = true,
                 hipStream_t side = nullptr, int part = 0);
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

};

// Events after every launch of a sequence: kernel i ran between events i
// and i + 1 on the same stream.
struct KernelTimer {
    std::vector<hipEvent_t> ev;
    std::vector<const char*> names;
    hipStream_t s = nullptr;
    void start(hipStream_t st) {
        s = st;
        names.clear();
        mark(nullptr);
    }
    void mark(const char* name) {
        if (name) names.push_back(name);
        const size_t k = names.size();
        if (ev.size() <= k) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            ev.push_back(e);
        }
        HIP_OK(hipEventRecord(ev[k], s));
    }
    ~KernelTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

void OrbPlan::enqueue(const OrbConfig& c, const SinCosTables& tb, hipStream_t s, mcv_keypoint* d_kps, uint8_t* d_desc,
                      int out_cap, int* d_counts, KernelTimer* rec, bool reset_err, hipStream_t side, int part) {
    auto mark = [&](const char* n) {
        if (rec) rec->mark(n);
    };
    const bool fork = part == kPartAll && side != nullptr && rec == nullptr && (tune.blur & 2) != 0;
    const OrbGlobals g = globals(c);
    const Strip* st = d_strips.as<Strip>();
    auto strip_grid = [](int n) { return dim3(xcd_grid((n + 3) / 4)); };
    BlurTaps taps{};
    for (int k = 0; k < 7; k++) taps.t[k] = c.gauss[k];
    taps.lo4 = (uint32_t)c.gauss[0] | (uint32_t)c.gauss[1] << 8 | (uint32_t)c.gauss[2] << 16 | (uint32_t)c.gauss[3] << 24;
    taps.hi3 = (uint32_t)c.gauss[4] | (uint32_t)c.gauss[5] << 8 | (uint32_t)c.gauss[6] << 16;
    const int ns = (int)strips.size();
    auto blur = [&](hipStream_t bs) {
        if (tune.blur & 1)
            hipLaunchKernelGGL(k_blur<true>, strip_grid(ns), dim3(256), 0, bs, st, ns, g.lv, g.pyr, g.blur, taps);
        else
            hipLaunchKernelGGL(k_blur<false>, strip_grid(ns), dim3(256), 0, bs, st, ns, g.lv, g.pyr, g.blur, taps);
        mark("k_blur");
    };
    // dynamic LDS: the heap (8 B per entry), then the node-table prefix (8 B per entry)
    auto heap = [&]() {
        const size_t heap_lds = 8 * (size_t)(heap_cap + kHeapTabLds);
        if (tune.heap == 1)
            hipLaunchKernelGGL(k_octree_heap, dim3(lv.size()), dim3(64), heap_lds, s, g, d_heap_order.as<int>(),
                               heap_cap, kHeapTabLds);
        else
            hipLaunchKernelGGL(k_octree_heap_r, dim3(lv.size()), dim3(64), heap_lds, s, g, d_heap_order.as<int>(),
                               heap_cap, kHeapTabLds);
        mark("k_octree_heap");
    };
    auto keypoints = [&]() {
        float* kps = reinterpret_cast<float*>(d_kps);
        if (tune.kp == 1)
            hipLaunchKernelGGL(k_keypoints_k<8>, dim3(xcd_grid((n_slots + 31) / 32)), dim3(256), 0, s, g,
                               d_slot.as<int>(), n_slots, tb, kps, d_desc, out_cap, d_counts);
        else if (tune.kp == 2)
            hipLaunchKernelGGL(k_keypoints_k<16>, dim3(xcd_grid((n_slots + 63) / 64)), dim3(256), 0, s, g,
                               d_slot.as<int>(), n_slots, tb, kps, d_desc, out_cap, d_counts);
        else if (tune.kp == 3)
            hipLaunchKernelGGL(k_keypoints_k<32>, dim3(xcd_grid((n_slots + 127) / 128)), dim3(256), 0, s, g,
                               d_slot.as<int>(), n_slots, tb, kps, d_desc, out_cap, d_counts);
        else
            hipLaunchKernelGGL(k_keypoints, dim3(xcd_grid((n_slots + 3) / 4)), dim3(256), 0, s, g, d_slot.as<int>(),
                               n_slots, tb, kps, d_desc, out_cap, d_counts);
        mark("k_keypoints");
    };
    if (part == kPartBlur) {
        if (ns > 0) blur(s);
        HIP_OK(hipGetLastError());
        return;
    }
    if (part == kPartHeap) {
        heap();
        HIP_OK(hipGetLastError());
        return;
    }
    if (part == kPartPost) {
        keypoints();
        HIP_OK(hipGetLastError());
        return;
    }
    // words 0-1: pool words requested (u64), words 2-3: error bits; level 0
    auto head = [&]() {
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t*>(d_used.p),
                           reset_err ? 4 : 2);
        mark("k_zero_words");
        const int n0 = strip_begin[1] - strip_begin[0];
        if (n0 > 0) {
            hipLaunchKernelGGL(k_copy_level0, strip_grid(n0), dim3(256), 0, s, st, n0, g.lv,
                               d_src.as<const uint8_t*>(), d_stride.as<int>(), g.pyr);
            mark("k_copy_level0");
        }
    };
    // levels l0 .. l1 - 1, each from the one before
    auto resize_levels = [&](int l0, int l1) {
        for (int l = std::max(l0, 1); l < l1; l++) {
            const int n = strip_begin[l + 1] - strip_begin[l];
            if (n > 0) {
                hipLaunchKernelGGL(k_resize, dim3(xcd_grid(n)), dim3(64), 0, s, st + strip_begin[l], n, g.lv,
                                   d_coef.as<int2>(), g.pyr);
                mark("k_resize");
            }
        }
    };
    // FAST scores of levels l0 .. l1 - 1 (their strips are contiguous)
    auto fast = [&](int l0, int l1) {
        const int b0 = strip_begin[l0], n = strip_begin[l1] - b0;
        if (n > 0) {
            hipLaunchKernelGGL(k_fast_score, strip_grid(n), dim3(256), 0, s, st + b0, n, g.lv, g.pyr, g.score,
                               std::min(g.ini_th, g.min_th));
            mark("k_fast_score");
        }
    };
    auto cells_build = [&]() {
        const int nc = (int)cells.size();
        if (nc > 0) {
            hipLaunchKernelGGL(k_fast_cells, dim3(xcd_grid((nc + 3) / 4)), dim3(256), 0, s, g, nc);
            mark("k_fast_cells");
        }
        if (n_build_large > 0)
            hipLaunchKernelGGL(k_octree_build<1024>, dim3(n_build_large), dim3(1024), 0, s, g,
                               d_build_levels.as<int>());
        if ((int)lv.size() > n_build_large)
            hipLaunchKernelGGL(k_octree_build<256>, dim3((int)lv.size() - n_build_large), dim3(256), 0, s, g,
                               d_build_levels.as<int>() + n_build_large);
        mark("k_octree_build");
    };
    const int lsplit = std::min(kFastSplitLevel, nlevels);
    if (part >= kPartPreA) {
        if (part == kPartPreA) {
            head();
            resize_levels(1, lsplit);
        } else if (part == kPartFastLo) {
            fast(0, lsplit);
        } else if (part == kPartPreB) {
            resize_levels(lsplit, nlevels);
            fast(lsplit, nlevels);
        } else {
            cells_build();
        }
        HIP_OK(hipGetLastError());
        return;
    }
    head();
    resize_levels(1, nlevels);
    if (ns > 0 && !fork && part == kPartAll) blur(s);
    fast(0, nlevels);
    cells_build();
    if (part == kPartPre) {
        HIP_OK(hipGetLastError());
        return;
    }
    if (fork && ns > 0) {
        if (!ev_fork) HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        if (!ev_join) HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        HIP_OK(hipEventRecord(ev_fork, s));
        HIP_OK(hipStreamWaitEvent(side, ev_fork, 0));
        blur(side);
        HIP_OK(hipEventRecord(ev_join, side));
    }
    heap();
    if (fork && ns > 0) HIP_OK(hipStreamWaitEvent(s, ev_join, 0));
    keypoints();
    HIP_OK(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------
struct mcv_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t cap = nullptr;  // graph capture only (the null stream cannot capture)
    hipStream_t side = nullptr; // graph capture only: the blur branch
    int knn_variant = 0;        // for mcv_match_knn2*: the last rig tuning's choice
    std::map<std::string, Tuning> tuned;  // rig shape -> autotuned choices
    hipStream_t stream = nullptr;
    std::string err;
    OrbConfig orb;
    bool orb_ready = false;
    SinCosDev sc;
    void* noalloc_scratch = nullptr;  // this device's address of `scratch`
    std::unique_ptr<OrbPlan> single;  // plan of the last single-image extract
    OrbPlan* last_batch = nullptr;     // plan of the last mcv_orb_extract_batch_dev (owned by plans)
    int single_key_w = -1, single_key_h = -1;
    std::vector<int> single_key_seeds;
    DevBuf in_img, out_kps, out_desc, out_cnt;
    DevBuf m_q, m_t, m_out, m_pairs, f_knn, f_out, f_cnt, f_scr, f_jobs;
    DevBuf d_desc_in, d_ptr, d_best, d_desc_out;
    DevBuf o_m, o_a1, o_a2, o_out, o_cnt, o_jobs;
    DevBuf w_pts, w_pdesc, w_plev, w_kxy, w_kdesc, w_cs, w_ci, w_out, w_view, w_cams;
    DevBuf b_d1, b_d2, b_f2, b_jobs, b_out;
    std::map<std::string, std::unique_ptr<OrbPlan>> plans;

    void set_device() { HIP_OK(hipSetDevice(device)); }

    hipStream_t capture_stream() {
        if (!cap) HIP_OK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        return cap;
    }
    // the second stream of a captured graph's branch (default priority)
    hipStream_t side_stream() {
        if (!side) HIP_OK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        return side;
    }

    const SinCosTables tables() {
        if (!sc.ready) {
            const SinCosHost& h = sincos_host();
            upload(sc.ck, h.ck, stream);
            upload(sc.cv, h.cv, stream);
            upload(sc.cidx, h.cidx, stream);
            upload(sc.sk, h.sk, stream);
            upload(sc.sv, h.sv, stream);
            upload(sc.sidx, h.sidx, stream);
            HIP_OK(hipStreamSynchronize(stream));
            sc.ready = true;
        }
        return sc.view();
    }
};

struct mcv_rig {
    mcv_ctx* ctx;
    int n_frames;
    int w[3], h[3];
    float ratio;
    int th;
    std::unique_ptr<OrbPlan> batch;                 // all 3n images
    std::vector<std::unique_ptr<OrbPlan>> single;   // per camera, one image
    int cap = 0;
    DevBuf kps, desc, counts, knn, matches, mcounts, pairs, jobs, scratch, porder;
    std::vector<MatchPair> hp;
    std::vector<FilterJob> hj;
    // Captured graphs of the batched run, one set per launch-variant set used
    // so far: switching variants back and forth reuses instantiated graphs
    // instead of capturing again.  With the blur beside the octree heap the
    // step is four graphs with launches on the context's side stream between
    // them (rig_launch); otherwise g alone is the step.  cur indexes the
    // current variants' set, -1 until the next run captures it.
    struct RigGraph {
        Tuning t;
        hipGraphExec_t g = nullptr, g2 = nullptr;
        hipGraphExec_t pre_b = nullptr, pre_c = nullptr;  // split steps: see rig_launch
    };
    std::vector<RigGraph> graphs;
    int cur = -1;
    static constexpr size_t kMaxGraphs = 16;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_a = nullptr, ev_b = nullptr;
    std::vector<const uint8_t*> graph_inputs;
    KernelTimer timer;
    bool last_batched = false;
    Tuning tune;
    std::string shape_key;
    DevBuf st_u, st_d, st_sad, st_keys, st_jobs;
    // destroys every captured graph (buffers moved, or the variants were retimed)
    void drop_graphs() {
        for (auto& e : graphs)
            for (hipGraphExec_t x : {e.g, e.g2, e.pre_b, e.pre_c})
                if (x) HIP_OK(hipGraphExecDestroy(x));
        graphs.clear();
        cur = -1;
    }
    // the graphs of the current variants, if they were captured
    void pick_graph() {
        cur = -1;
        for (size_t i = 0; i < graphs.size(); i++)
            if (graphs[i].t == tune) cur = (int)i;
    }
    void keep_graph(const RigGraph& g) {
        if (graphs.size() >= kMaxGraphs) drop_graphs();
        graphs.push_back(g);
        cur = (int)graphs.size() - 1;
    }
    ~mcv_rig() {
        for (auto& e : graphs)
            for (hipGraphExec_t x : {e.g, e.g2, e.pre_b, e.pre_c})
                if (x) (void)hipGraphExecDestroy(x);
        for (hipEvent_t x : {ev_fork, ev_join, ev_a, ev_b})
            if (x) (void)hipEventDestroy(x);
    }
};

struct mcv_ba {
    mcv_ctx* ctx;
    int M, P, F, C;
    DevBuf uv, pt, pair, isig, phi, plo, pairs, planes, fptr, fobs, pptr, gobs;
    // stereo edges (EdgeStereoSE3ProjectXYZ)
    int Ms = 0;
    StereoCam scam{};
    bool have_stereo = false, stereo_linearized = false;
    DevBuf s_obs, s_pt, s_fr, s_isig, s_planes, fposes, flags;
    std::vector<int> h_gptr;  // observations of (frame, camera) group g: gobs[h_gptr[g] .. h_gptr[g + 1])
    std::vector<int> h_pt, h_fr, h_pptr;
    std::vector<int> fixed;
    int variant = 0;  // k_ba_linearize launch shape (kBaVariant)
    bool have_obs = false, have_state = false, linearized = false;
    // J_pose planes valid (a full or pose-only linearization of the mono /
    // stereo edges at the current state)
    bool pose_linearized = false, stereo_pose_linearized = false, stereo_full = false;
    // bumped by every change of the planes' inputs or contents (state,
    // observations, any linearization): a reader that recorded it can tell
    // whether the planes still hold its linearization
    long long gen = 0;
    void touched(int mode) {  // mode: kBaFull / kBaResidual / kBaPoseOnly, or -1 (inputs changed)
        linearized = mode == kBaFull;
        pose_linearized = mode == kBaFull || mode == kBaPoseOnly;
        if (mode < 0) stereo_linearized = stereo_pose_linearized = stereo_full = false;
        schur_done = false;
        gen++;
    }
    // optimizer state (g2o keeps FP64 estimates): frames and cameras on the
    // host, points on the device (p64) with the float pairs phi / plo
    std::vector<double> h_frames, h_cams;
    double h_scam[6] = {0, 0, 0, 0, 0, 0};
    std::vector<int> h_spt, h_sfr;
    std::vector<int> point_fixed;                 // empty: no point fixed
    std::vector<uint8_t> active_m, active_s;      // empty: every edge active
    bool robust = true;                           // Huber kernels on
    DevBuf p64, p64_bak, pfix_d, act_m, act_s, xl, xp, partial, dsum, fmax;
    // Schur structure (depends on the edges, the active sets and the fixed
    // frames and points)
    bool schur_ready = false, schur_done = false;
    int npf = 0, Fp = 0, nblk = 0;
    bool free_points_with_edges = true;
    DevBuf pe_ptr, pe, pf_ptr, pf_obs_ptr, pf_obs, pf_point, pf_frame, sfptr, sfobs, fpf_ptr, fpf, blk_ptr, blk_items,
        blk_ij;
    DevBuf Dinv, bl, db, Hll, Hpl, Hpp, bs, bp, Sout;
};

// Launch shapes of k_ba_linearize (same arithmetic per observation; the
// autotuner mcv_ba_autotune picks one): workgroup size, grid cap,
// observations in flight per thread, streaming stores.
struct BAVariantDesc {
    int threads, grid, ilp;
    bool nt;
    const char* name;
};
constexpr int kBaVariants = 8;
static const BAVariantDesc kBaVariant[kBaVariants] = {
    {256, 1024, 1, false, "256x1024 ilp1"},     {256, 1024, 2, false, "256x1024 ilp2"},
    {512, 1024, 1, false, "512x1024 ilp1"},     {512, 1024, 2, false, "512x1024 ilp2"},
    {512, 1024, 2, true, "512x1024 ilp2 nt"},   {1024, 512, 2, true, "1024x512 ilp2 nt"},
    {256, 1024, 1, true, "256x1024 ilp1 nt"},   {512, 1024, 1, true, "512x1024 ilp1 nt"},
};

template <int kIlp, bool kNt>
static void launch_ba_lds(int blocks, int threads, size_t lds, hipStream_t s, const BAObs& ob, const float4* phi,
                          const float4* plo, const BAPair* pairs, int np, const BAOut& o, int residual_only) {
    hipLaunchKernelGGL((k_ba_linearize<true, kIlp, kNt>), dim3(blocks), dim3(threads), lds, s, ob, phi, plo, pairs, np,
                       o, residual_only);
}

static void ba_enqueue_linearize(mcv_ba* b, int variant, int residual_only, hipStream_t s) {
    if (b->M <= 0) return;
    BAObs ob{b->uv.as<float2>(), b->pt.as<uint32_t>(), b->pair.as<uint32_t>(), b->isig.as<float>(), b->M};
    BAOut o{b->planes.as<float>(), (long long)b->M};
    const int np = b->F * b->C;
    const float4* phi = b->phi.as<float4>();
    const float4* plo = b->plo.as<float4>();
    const BAPair* pairs = b->pairs.as<BAPair>();
    if (np <= kLdsPairs) {
        const BAVariantDesc& V = kBaVariant[variant];
        const long long per = (long long)V.threads * V.ilp;
        const int blocks = (int)std::min<long long>((b->M + per - 1) / per, V.grid);
        const size_t lds = (size_t)np * sizeof(BAPair);
        switch (variant) {
        case 1: case 3: launch_ba_lds<2, false>(blocks, V.threads, lds, s, ob, phi, plo, pairs, np, o, residual_only); break;
        case 4: case 5: launch_ba_lds<2, true>(blocks, V.threads, lds, s, ob, phi, plo, pairs, np, o, residual_only); break;
        case 6: case 7: launch_ba_lds<1, true>(blocks, V.threads, lds, s, ob, phi, plo, pairs, np, o, residual_only); break;
        default: launch_ba_lds<1, false>(blocks, V.threads, lds, s, ob, phi, plo, pairs, np, o, residual_only); break;
        }
    } else {
        hipLaunchKernelGGL((k_ba_linearize<false, 1, false>), dim3((b->M + 255) / 256), dim3(256), 0, s, ob, phi, plo,
                           pairs, np, o, residual_only);
    }
    HIP_OK(hipGetLastError());
}

// Device address of `scratch` on `device` (which must be current): looked
// up once per device; each GPU holds its own copy at its own address.
static void* scratch_address(int device) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    REQUIRE(p != nullptr, MCV_ERR_HIP, "no device address for the no-allocation scratch");
    addr[device] = p;
    return p;
}

static int fail(mcv_ctx* ctx, const Fail& f) {
    if (ctx) ctx->err = f.msg;
    return f.code;
}

#define API_BEGIN(ctxp) \
    mcv_ctx* api_ctx_ = (ctxp); \
    try {
#define API_END                                                   \
    }                                                             \
    catch (const Fail& f) { return fail(api_ctx_, f); }           \
    catch (const std::exception& e) { return fail(api_ctx_, Fail{MCV_ERR_STATE, e.what()}); } \
    return MCV_OK;

// ---------------------------------------------------------------------------
extern "C" {

int mcv_create(mcv_ctx** out, int device) {
    if (!out) return MCV_ERR_INVALID;
    *out = nullptr;
    auto* c = new mcv_ctx();
    try {
        int n = 0;
        HIP_OK(hipGetDeviceCount(&n));
        REQUIRE(device >= 0 && device < n, MCV_ERR_INVALID, "no such device");
        c->device = device;
        c->set_device();
        HIP_OK(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
        c->stream = c->own;
        c->noalloc_scratch = scratch_address(device);
    } catch (const Fail& f) {
        int code = f.code;
        delete c;
        return code;
    }
    *out = c;
    return MCV_OK;
}

void mcv_destroy(mcv_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->own) (void)hipStreamSynchronize(c->own);
    c->plans.clear();
    c->single.reset();
    if (c->own) (void)hipStreamDestroy(c->own);
    if (c->cap) (void)hipStreamDestroy(c->cap);
    if (c->side) (void)hipStreamDestroy(c->side);
    delete c;
}

const char* mcv_last_error(const mcv_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mcv_set_stream(mcv_ctx* c, void* s) {
    API_BEGIN(c)
    REQUIRE(c, MCV_ERR_INVALID, "null context");
    c->stream = static_cast<hipStream_t>(s);   // NULL: the null (default) stream
    API_END
}

int mcv_use_own_stream(mcv_ctx* c) {
    API_BEGIN(c)
    REQUIRE(c, MCV_ERR_INVALID, "null context");
    c->stream = c->own;
    API_END
}

int mcv_synchronize(mcv_ctx* c) {
    API_BEGIN(c)
    REQUIRE(c, MCV_ERR_INVALID, "null context");
    c->set_device();
    HIP_OK(hipStreamSynchronize(c->stream));
    API_END
}

int mcv_copy_to_host(mcv_ctx* c, void* dst, const void* src, uint64_t bytes) {
    API_BEGIN(c)
    REQUIRE(c && (bytes == 0 || (dst && src)), MCV_ERR_INVALID, "null copy buffer");
    c->set_device();
    if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    API_END
}

int mcv_orb_configure(mcv_ctx* c, int nf, float sf, int nl, int ini, int mn) {
    API_BEGIN(c)
    REQUIRE(c, MCV_ERR_INVALID, "null context");
    REQUIRE(nf > 0 && nl >= 1 && nl <= kMaxLevels && sf > 1.f, MCV_ERR_INVALID, "invalid ORB parameters");
    orb_init(c->orb, nf, sf, nl, ini, mn);
    c->orb_ready = true;
    c->single.reset();
    c->plans.clear();
    API_END
}

int mcv_orb_tables(const mcv_ctx* c, float* scale, float* inv_scale, float* sigma2, float* inv_sigma2, int32_t* npl) {
    if (!c || !c->orb_ready) return MCV_ERR_STATE;
    for (int i = 0; i < c->orb.nlevels; i++) {
        if (scale) scale[i] = c->orb.scale[i];
        if (inv_scale) inv_scale[i] = c->orb.inv_scale[i];
        if (sigma2) sigma2[i] = c->orb.sigma2[i];
        if (inv_sigma2) inv_sigma2[i] = c->orb.inv_sigma2[i];
        if (npl) npl[i] = c->orb.npl[i];
    }
    return MCV_OK;
}

int mcv_orb_capacity(const mcv_ctx* c, int w, int h, int n_seeds, int32_t* cap) {
    if (!c || !c->orb_ready || !cap) return MCV_ERR_STATE;
    int total = n_seeds;
    for (int l = 0; l < c->orb.nlevels; l++) {
        const int lw = l == 0 ? w : (int)lrintf((float)w * c->orb.inv_scale[l]);
        const int lh = l == 0 ? h : (int)lrintf((float)h * c->orb.inv_scale[l]);
        total += level_out_cap(c->orb.npl[l], level_roots(lw, lh));
    }
    *cap = total;
    return MCV_OK;
}

static long long pool_words_for(int n_images) { return 4ll * 1024 * 1024 + (long long)n_images * 2 * 1024 * 1024; }

int mcv_orb_extract(mcv_ctx* c, const uint8_t* image, int w, int h, int stride, const mcv_keypoint* seeds, int n_seeds,
                    mcv_keypoint* kps, uint8_t* desc, int capacity, int32_t* count) {
    API_BEGIN(c)
    REQUIRE(c && c->orb_ready, MCV_ERR_STATE, "ORB not configured");
    REQUIRE(count, MCV_ERR_INVALID, "null count");
    *count = 0;
    if (!image || w <= 0 || h <= 0) throw Fail{MCV_ERR_EMPTY, "empty image"};
    REQUIRE(stride >= w, MCV_ERR_INVALID, "stride < width");
    REQUIRE(n_seeds >= 0 && (n_seeds == 0 || seeds), MCV_ERR_INVALID, "bad seeds");
    c->set_device();
    ImageSpec spec{w, h, std::vector<std::vector<mcv_keypoint>>(c->orb.nlevels)};
    std::vector<int> seed_key(c->orb.nlevels, 0);
    for (int i = 0; i < n_seeds; i++) {
        REQUIRE(seeds[i].octave >= 0 && seeds[i].octave < c->orb.nlevels, MCV_ERR_INVALID, "seed octave out of range");
        spec.seeds[seeds[i].octave].push_back(seeds[i]);
        seed_key[seeds[i].octave]++;
    }
    if (!c->single || c->single_key_w != w || c->single_key_h != h || c->single_key_seeds != seed_key || n_seeds > 0) {
        c->single.reset(new OrbPlan());
        c->single->build(c->orb, {spec}, pool_words_for(1));
        c->single->upload_all(c->stream);
        c->single_key_w = w;
        c->single_key_h = h;
        c->single_key_seeds = seed_key;
    }
    OrbPlan& P = *c->single;
    const SinCosTables tb = c->tables();
    c->in_img.ensure((size_t)w * h);
    HIP_OK(hipMemcpy2DAsync(c->in_img.p, w, image, stride, w, h, hipMemcpyHostToDevice, c->stream));
    const uint8_t* src = c->in_img.as<uint8_t>();
    const int st = w;
    HIP_OK(hipMemcpyAsync(P.d_src.p, &src, sizeof(src), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(P.d_stride.p, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
    c->out_kps.ensure((size_t)P.cap * sizeof(mcv_keypoint));
    c->out_desc.ensure((size_t)P.cap * 32);
    c->out_cnt.ensure(16);
    int n = 0;
    for (int attempt = 0;; attempt++) {
        P.enqueue(c->orb, tb, c->stream, c->out_kps.as<mcv_keypoint>(), c->out_desc.as<uint8_t>(), P.cap,
                  c->out_cnt.as<int>());
        unsigned long long st[2] = {0, 0};  // pool words requested, error bits
        HIP_OK(hipMemcpyAsync(&n, c->out_cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(st, P.d_used.p, 16, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        const int err = (int)(st[1] & 0xFFFFFFFFu);
        if (err == 0) break;
        REQUIRE(!(err & 2), MCV_ERR_STATE, "a pyramid level has more FAST candidates than its cell capacity");
        // pool exhausted: the counter holds what the run asked for
        REQUIRE(attempt < 2, MCV_ERR_STATE, "octree node pool exhausted");
        P.pool_cap = (long long)(st[0] + st[0] / 4 + 1024);
        P.d_pool.ensure((size_t)P.pool_cap * 4);
    }
    *count = n;
    if (n > capacity) throw Fail{MCV_ERR_CAPACITY, "keypoint capacity too small"};
    if (n > 0) {
        if (kps) HIP_OK(hipMemcpy(kps, c->out_kps.p, (size_t)n * sizeof(mcv_keypoint), hipMemcpyDeviceToHost));
        if (desc) HIP_OK(hipMemcpy(desc, c->out_desc.p, (size_t)n * 32, hipMemcpyDeviceToHost));
    }
    API_END
}

int mcv_orb_pyramid(mcv_ctx* c, int level, uint8_t* out, int capacity, int32_t* w, int32_t* h) {
    API_BEGIN(c)
    REQUIRE(c && c->single, MCV_ERR_STATE, "no extraction yet");
    REQUIRE(level >= 0 && level < c->single->nlevels, MCV_ERR_INVALID, "level out of range");
    const LevelDesc& d = c->single->lv[level];
    if (w) *w = d.w;
    if (h) *h = d.h;
    if (out) {
        REQUIRE(capacity >= d.w * d.h, MCV_ERR_CAPACITY, "pyramid buffer too small");
        c->set_device();
        HIP_OK(hipMemcpy2DAsync(out, (size_t)d.w, c->single->d_pyr.as<uint8_t>() + d.pix, (size_t)d.pitch, (size_t)d.w,
                                (size_t)d.h, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    API_END
}

}  // extern "C"

static std::string batch_key(int n, const int32_t* widths, const int32_t* heights) {
    std::string key = std::to_string(n);
    for (int i = 0; i < n; i++) key += "," + std::to_string(widths[i]) + "x" + std::to_string(heights[i]);
    return key;
}

// Waits for the plan's last run and reads its error word: returns 0, or
// throws for a level with more FAST candidates than its cell capacity, or
// grows the octree pool (dropping the plan's graph, which holds the old
// pool) and returns 1 when the pool overflowed.
static int plan_check(OrbPlan& P, hipStream_t s) {
    unsigned long long st[2] = {0, 0};  // pool words requested, error bits
    HIP_OK(hipMemcpyAsync(st, P.d_used.p, 16, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const int err = (int)(st[1] & 0xFFFFFFFFu);
    REQUIRE(!(err & 2), MCV_ERR_STATE, "a pyramid level has more FAST candidates than its cell capacity");
    if (!(err & 1)) return 0;
    // the counter holds what the last image asked for; earlier images of a
    // plain run may have asked for more, so grow at least twofold
    P.pool_cap = std::max<long long>((long long)(st[0] + st[0] / 4 + 1024), 2 * P.pool_cap);
    P.d_pool.ensure((size_t)P.pool_cap * 4);
    if (P.graph) {
        HIP_OK(hipGraphExecDestroy(P.graph));
        P.graph = nullptr;
    }
    return 1;
}

extern "C" {

int mcv_orb_extract_batch_dev(mcv_ctx* c, int n, const uint8_t* const* d_images, const int32_t* widths,
                              const int32_t* heights, const int32_t* strides, const mcv_keypoint* const* seeds,
                              const int32_t* n_seeds, mcv_keypoint* d_kps, uint8_t* d_desc, int capacity,
                              int32_t* d_counts, int flags) {
    API_BEGIN(c)
    REQUIRE(c && c->orb_ready, MCV_ERR_STATE, "ORB not configured");
    REQUIRE(n > 0 && d_images && widths && heights, MCV_ERR_INVALID, "bad batch");
    REQUIRE(!n_seeds || seeds, MCV_ERR_INVALID, "seed counts without seed arrays");
    c->set_device();
    const int nl = c->orb.nlevels;
    std::vector<ImageSpec> specs;
    std::string key = batch_key(n, widths, heights);
    for (int i = 0; i < n; i++) {
        specs.push_back(ImageSpec{widths[i], heights[i], std::vector<std::vector<mcv_keypoint>>(nl)});
        const int ns = n_seeds ? n_seeds[i] : 0;
        REQUIRE(ns >= 0 && (ns == 0 || seeds[i]), MCV_ERR_INVALID, "bad seeds");
        for (int k = 0; k < ns; k++) {
            const mcv_keypoint& kp = seeds[i][k];
            REQUIRE(kp.octave >= 0 && kp.octave < nl, MCV_ERR_INVALID, "seed octave out of range");
            specs[i].seeds[kp.octave].push_back(kp);
        }
        if (ns) {
            // the plan's keypoint slots depend on the seed count of every level
            key += ";" + std::to_string(i) + ":";
            for (int l = 0; l < nl; l++) key += std::to_string(specs[i].seeds[l].size()) + ".";
        }
    }
    auto& slot = c->plans[key];
    if (!slot) {
        slot.reset(new OrbPlan());
        slot->build(c->orb, specs, pool_words_for(n));
        slot->upload_all(c->stream);
    } else if (n_seeds) {
        // same counts, new values: rows in the order build() lays them out
        slot->seeds.clear();
        for (const ImageSpec& sp : specs)
            for (int l = 0; l < nl; l++)
                for (const mcv_keypoint& k : sp.seeds[l])
                    slot->seeds.insert(slot->seeds.end(), {k.x, k.y, k.size, k.angle, k.response, (float)k.octave});
        upload(slot->d_seeds, slot->seeds, c->stream);
    }
    OrbPlan& P = *slot;
    c->last_batch = &P;
    REQUIRE(capacity >= P.cap, MCV_ERR_CAPACITY, "capacity below mcv_orb_capacity");
    const SinCosTables tb = c->tables();
    std::vector<int> st(n);
    for (int i = 0; i < n; i++) st[i] = strides ? strides[i] : widths[i];
    HIP_OK(hipMemcpyAsync(P.d_src.p, d_images, sizeof(void*) * n, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(P.d_stride.p, st.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    for (int attempt = 0;; attempt++) {
        if (flags & MCV_FLAG_GRAPH) {
            const OrbPlan::GraphKey gk{d_kps, d_desc, d_counts, capacity};
            if (!P.graph || !(P.graph_key == gk)) {
                if (P.graph) HIP_OK(hipGraphExecDestroy(P.graph));
                P.graph = nullptr;
                hipGraph_t gr;
                hipStream_t cs = c->capture_stream();
                HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                P.enqueue(c->orb, tb, cs, d_kps, d_desc, capacity, d_counts, nullptr, true, c->side_stream());
                HIP_OK(hipStreamEndCapture(cs, &gr));
                HIP_OK(hipGraphInstantiate(&P.graph, gr, nullptr, nullptr, 0));
                HIP_OK(hipGraphDestroy(gr));
                P.graph_key = gk;
            }
            HIP_OK(hipGraphLaunch(P.graph, c->stream));
        } else {
            P.enqueue(c->orb, tb, c->stream, d_kps, d_desc, capacity, d_counts);
        }
        if (!(flags & MCV_FLAG_CHECK)) break;
        // a pool overflow grows the pool (the graph holds the old one) and reruns
        if (plan_check(P, c->stream) == 0) break;
        REQUIRE(attempt < 2, MCV_ERR_STATE, "octree node pool exhausted");
    }
    API_END
}

int mcv_orb_batch_check(mcv_ctx* c) {
    API_BEGIN(c)
    REQUIRE(c && c->last_batch, MCV_ERR_STATE, "no batched extraction yet");
    c->set_device();
    if (plan_check(*c->last_batch, c->stream) != 0)
        throw Fail{MCV_ERR_STATE, "octree node pool exhausted in the last run (grown for the next one)"};
    API_END
}

static void level_view(const OrbPlan& P, int image, int level, int blurred, const uint8_t** d_level, int32_t* pitch,
                       int32_t* w, int32_t* h) {
    REQUIRE(image >= 0 && image < P.n_images && level >= 0 && level < P.nlevels, MCV_ERR_INVALID,
            "image or level out of range");
    const LevelDesc& d = P.lv[(size_t)image * P.nlevels + level];
    if (d_level) *d_level = (blurred ? P.d_blur : P.d_pyr).as<uint8_t>() + d.pix;
    if (pitch) *pitch = d.pitch;
    if (w) *w = d.w;
    if (h) *h = d.h;
}

int mcv_orb_batch_pyramid_dev(mcv_ctx* c, int image, int level, int blurred, const uint8_t** d_level, int32_t* pitch,
                              int32_t* w, int32_t* h) {
    API_BEGIN(c)
    REQUIRE(c && c->last_batch, MCV_ERR_STATE, "no batched extraction yet");
    level_view(*c->last_batch, image, level, blurred, d_level, pitch, w, h);
    API_END
}

// ---- matching ---------------------------------------------------------------
static void launch_knn(hipStream_t s, DevBuf& pbuf, const std::vector<MatchPair>& pairs, int max_nq, int variant) {
    if (pairs.empty() || max_nq <= 0) return;
    upload(pbuf, pairs, s);
    int max_nt = 0;
    for (const MatchPair& p : pairs) max_nt = std::max(max_nt, p.nt);
    launch_knn2(variant, pbuf.as<MatchPair>(), (int)pairs.size(), max_nq, s, max_nt);
    HIP_OK(hipGetLastError());
}

int mcv_match_knn2_dev(mcv_ctx* c, const uint8_t* d_q, int nq, const uint8_t* d_t, int nt, int32_t* d_out) {
    API_BEGIN(c)
    REQUIRE(c && nq >= 0 && nt >= 0 && nt < (1 << 22), MCV_ERR_INVALID, "bad sizes");
    c->set_device();
    MatchPair p{d_q, d_t, nullptr, nullptr, nq, nt, d_out};
    launch_knn(c->stream, c->m_pairs, {p}, nq, c->knn_variant);
    API_END
}

int mcv_distinctive_descriptors(mcv_ctx* c, const uint8_t* desc, const int32_t* ptr, int n, int32_t* best,
                                uint8_t* out_desc) {
    API_BEGIN(c)
    REQUIRE(c && ptr && best && n >= 0, MCV_ERR_INVALID, "bad distinctive-descriptor arguments");
    if (n == 0) return MCV_OK;
    REQUIRE(ptr[0] == 0, MCV_ERR_INVALID, "ptr must start at 0");
    for (int m = 0; m < n; m++)
        REQUIRE(ptr[m + 1] >= ptr[m] && ptr[m + 1] - ptr[m] <= kDistinctMax, MCV_ERR_INVALID,
                "descriptor lists must be ordered and hold at most 256 rows");
    const int total = ptr[n];
    REQUIRE(total == 0 || desc, MCV_ERR_INVALID, "null descriptors");
    c->set_device();
    hipStream_t s = c->stream;
    c->d_desc_in.ensure((size_t)std::max(total, 1) * 32);
    c->d_ptr.ensure((size_t)(n + 1) * 4);
    c->d_best.ensure((size_t)n * 4);
    c->d_desc_out.ensure((size_t)n * 32);
    if (total) HIP_OK(hipMemcpyAsync(c->d_desc_in.p, desc, (size_t)total * 32, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(c->d_ptr.p, ptr, (size_t)(n + 1) * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_distinctive, dim3(n), dim3(64), 0, s, c->d_desc_in.as<uint8_t>(), c->d_ptr.as<int>(), n,
                       c->d_best.as<int>(), out_desc ? c->d_desc_out.as<uint8_t>() : nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(best, c->d_best.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    if (out_desc) HIP_OK(hipMemcpyAsync(out_desc, c->d_desc_out.p, (size_t)n * 32, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    API_END
}

int mcv_match_filter_orientation(mcv_ctx* c, const int32_t* matches, int n, const float* angle1, int n1,
                                 const float* angle2, int n2, int32_t* out, int32_t* out_n) {
    API_BEGIN(c)
    REQUIRE(c && out_n && n >= 0 && n <= kOrientMax, MCV_ERR_INVALID, "bad orientation-filter arguments");
    if (n == 0) {
        *out_n = 0;
        return MCV_OK;
    }
    REQUIRE(matches && angle1 && angle2 && out, MCV_ERR_INVALID, "null orientation-filter arrays");
    for (int i = 0; i < n; i++)
        REQUIRE(matches[3 * i] >= 0 && matches[3 * i] < n1 && matches[3 * i + 1] >= 0 && matches[3 * i + 1] < n2,
                MCV_ERR_INVALID, "match index out of range");
    c->set_device();
    hipStream_t s = c->stream;
    c->o_m.ensure((size_t)n * 12);
    c->o_a1.ensure((size_t)std::max(n1, 1) * 4);
    c->o_a2.ensure((size_t)std::max(n2, 1) * 4);
    c->o_out.ensure((size_t)n * 12);
    c->o_cnt.ensure(4);
    HIP_OK(hipMemcpyAsync(c->o_m.p, matches, (size_t)n * 12, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(c->o_a1.p, angle1, (size_t)n1 * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(c->o_a2.p, angle2, (size_t)n2 * 4, hipMemcpyHostToDevice, s));
    std::vector<OrientJob> jobs{OrientJob{c->o_m.as<int32_t>(), n, c->o_a1.as<float>(), c->o_a2.as<float>(),
                                          c->o_out.as<int32_t>(), c->o_cnt.as<int>()}};
    upload(c->o_jobs, jobs, s);
    hipLaunchKernelGGL(k_filter_orientation, dim3(1), dim3(256), 0, s, c->o_jobs.as<OrientJob>());
    HIP_OK(hipGetLastError());
    int k = 0;
    HIP_OK(hipMemcpyAsync(&k, c->o_cnt.p, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (k > 0) {
        HIP_OK(hipMemcpyAsync(out, c->o_out.p, (size_t)k * 12, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    *out_n = k;
    API_END
}

int mcv_match_set_variant(mcv_ctx* c, int variant) {
    API_BEGIN(c)
    REQUIRE(c && variant >= 0 && variant < kKnnVariants, MCV_ERR_INVALID, "unknown matching variant");
    c->knn_variant = variant;
    API_END
}

int mcv_match_knn2(mcv_ctx* c, const uint8_t* q, int nq, const uint8_t* t, int nt, int32_t* out) {
    API_BEGIN(c)
    REQUIRE(c && nq >= 0 && nt >= 0 && nt < (1 << 22), MCV_ERR_INVALID, "bad sizes");
    REQUIRE(nq == 0 || (q && out), MCV_ERR_INVALID, "null buffer");
    if (nq == 0) return MCV_OK;
    c->set_device();
    c->m_q.ensure((size_t)nq * 32);
    c->m_t.ensure((size_t)std::max(nt, 1) * 32);
    c->m_out.ensure((size_t)nq * 16);
    HIP_OK(hipMemcpyAsync(c->m_q.p, q, (size_t)nq * 32, hipMemcpyHostToDevice, c->stream));
    if (nt) HIP_OK(hipMemcpyAsync(c->m_t.p, t, (size_t)nt * 32, hipMemcpyHostToDevice, c->stream));
    MatchPair p{c->m_q.as<uint8_t>(), c->m_t.as<uint8_t>(), nullptr, nullptr, nq, nt, c->m_out.as<int32_t>()};
    launch_knn(c->stream, c->m_pairs, {p}, nq, c->knn_variant);
    HIP_OK(hipMemcpyAsync(out, c->m_out.p, (size_t)nq * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    API_END
}

int mcv_match_filter(mcv_ctx* c, const int32_t* knn, int nq, float ratio, int th, int32_t* out, int32_t* count) {
    API_BEGIN(c)
    REQUIRE(c && count && nq >= 0, MCV_ERR_INVALID, "bad arguments");
    *count = 0;
    if (nq == 0) return MCV_OK;
    REQUIRE(knn && out, MCV_ERR_INVALID, "null buffer");
    c->set_device();
    c->f_knn.ensure((size_t)nq * 16);
    c->f_out.ensure((size_t)nq * 12);
    c->f_cnt.ensure(16);
    c->f_scr.ensure((size_t)nq * 8);
    HIP_OK(hipMemcpyAsync(c->f_knn.p, knn, (size_t)nq * 16, hipMemcpyHostToDevice, c->stream));
    FilterJob j{c->f_knn.as<int32_t>(), nullptr, nq, ratio, th, c->f_out.as<int32_t>(), c->f_cnt.as<int>(),
                c->f_scr.as<int>()};
    upload(c->f_jobs, std::vector<FilterJob>{j}, c->stream);
    hipLaunchKernelGGL(k_filter, dim3(1), dim3(1024), 0, c->stream, c->f_jobs.as<FilterJob>());
    HIP_OK(hipGetLastError());
    int n = 0;
    HIP_OK(hipMemcpyAsync(&n, c->f_cnt.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (n > 0) HIP_OK(hipMemcpy(out, c->f_out.p, (size_t)n * 12, hipMemcpyDeviceToHost));
    *count = n;
    API_END
}

int mcv_match_filter_noalloc_dev(mcv_ctx* c, const int32_t* d_knn, int nq, float ratio, int th, int32_t* d_out,
                                 int32_t* d_count) {
    API_BEGIN(c)
    REQUIRE(c && nq >= 0 && d_count && (nq == 0 || (d_knn && d_out)), MCV_ERR_INVALID, "bad arguments");
    REQUIRE(nq <= kNoAllocMaxQueries, MCV_ERR_CAPACITY, "more queries than the no-allocation scratch holds (65536)");
    REQUIRE(c->noalloc_scratch, MCV_ERR_STATE, "no-allocation scratch not set up");
    c->set_device();
    // the job travels as a kernel argument; the scratch needs no clearing
    // (k_filter_one writes every entry before reading it)
    const FilterJob j{d_knn, nullptr, nq, ratio, th, d_out, d_count, static_cast<int*>(c->noalloc_scratch)};
    hipLaunchKernelGGL(k_filter_one, dim3(1), dim3(1024), 0, c->stream, j);
    HIP_OK(hipGetLastError());
    API_END
}

// ---- rig frames -------------------------------------------------------------
int mcv_rig_create(mcv_ctx* c, int n_frames, const int32_t* w3, const int32_t* h3, const int32_t* nfeatures,
                   float ratio, int th, mcv_rig** out) {
    API_BEGIN(c)
    REQUIRE(c && c->orb_ready, MCV_ERR_STATE, "ORB not configured");
    REQUIRE(out && n_frames > 0 && w3 && h3, MCV_ERR_INVALID, "bad rig arguments");
    c->set_device();
    std::unique_ptr<mcv_rig> r(new mcv_rig());
    r->ctx = c;
    r->n_frames = n_frames;
    for (int k = 0; k < 3; k++) { r->w[k] = w3[k]; r->h[k] = h3[k]; }
    r->ratio = ratio;
    r->th = th;
    std::vector<ImageSpec> specs;
    for (int f = 0; f < n_frames; f++)
        for (int k = 0; k < 3; k++) {
            ImageSpec sp{w3[k], h3[k], std::vector<std::vector<mcv_keypoint>>(c->orb.nlevels)};
            if (nfeatures) {
                REQUIRE(nfeatures[3 * f + k] > 0, MCV_ERR_INVALID, "feature count must be positive");
                sp.nfeatures = nfeatures[3 * f + k];
            }
            specs.push_back(sp);
        }
    r->batch.reset(new OrbPlan());
    r->batch->build(c->orb, specs, pool_words_for(3 * n_frames));
    r->batch->upload_all(c->stream);
    // plain path: one single-image plan per camera, or per image when the
    // feature counts differ between images
    const int nsingle = nfeatures ? 3 * n_frames : 3;
    for (int k = 0; k < nsingle; k++) {
        r->single.emplace_back(new OrbPlan());
        r->single[k]->build(c->orb, {specs[k]}, pool_words_for(1));
        r->single[k]->upload_all(c->stream);
    }
    r->cap = r->batch->cap;
    {
        std::string key = std::to_string(n_frames);
        for (int k = 0; k < 3; k++) key += "," + std::to_string(w3[k]) + "x" + std::to_string(h3[k]);
        if (nfeatures)
            for (int i = 0; i < 3 * n_frames; i++) key += ":" + std::to_string(nfeatures[i]);
        r->shape_key = key;
        auto it = c->tuned.find(key);
        if (it != c->tuned.end()) r->tune = it->second;
        r->batch->tune = r->tune;
    }
    const int n = 3 * n_frames;
    r->kps.ensure((size_t)n * r->cap * sizeof(mcv_keypoint));
    r->desc.ensure((size_t)n * r->cap * 32);
    r->counts.ensure((size_t)n * 4);
    r->knn.ensure((size_t)n * r->cap * 16);
    r->matches.ensure((size_t)n * r->cap * 12);
    r->mcounts.ensure((size_t)n * 4);
    r->scratch.ensure((size_t)n * r->cap * 8);
    r->porder.ensure((size_t)(2 * n + 1) * sizeof(int));
    // pairs per frame: L-R (query L), W-L (query W), W-R (query W)
    const int qa[3] = {0, 2, 2}, ta[3] = {1, 0, 1};
    for (int f = 0; f < n_frames; f++)
        for (int k = 0; k < 3; k++) {
            const int qi = 3 * f + qa[k], ti = 3 * f + ta[k], pi = 3 * f + k;
            MatchPair p;
            p.q = r->desc.as<uint8_t>() + (size_t)qi * r->cap * 32;
            p.t = r->desc.as<uint8_t>() + (size_t)ti * r->cap * 32;
            p.nq_dev = r->counts.as<int>() + qi;
            p.nt_dev = r->counts.as<int>() + ti;
            p.nq = r->cap;
            p.nt = r->cap;
            p.out = r->knn.as<int32_t>() + (size_t)pi * r->cap * 4;
            r->hp.push_back(p);
            FilterJob j;
            j.knn = p.out;
            j.nq_dev = p.nq_dev;
            j.nq = r->cap;
            j.ratio = ratio;
            j.th = th;
            j.out = r->matches.as<int32_t>() + (size_t)pi * r->cap * 3;
            j.count = r->mcounts.as<int>() + pi;
            j.scratch = r->scratch.as<int>() + (size_t)pi * r->cap * 2;
            r->hj.push_back(j);
        }
    upload(r->pairs, r->hp, c->stream);
    upload(r->jobs, r->hj, c->stream);
    HIP_OK(hipStreamSynchronize(c->stream));
    *out = r.release();
    API_END
}

void mcv_rig_destroy(mcv_rig* r) { delete r; }

// the rig's top-2 launch: the pairs ordered by train count first (the
// counts are on the device), then the chosen variant
static void rig_knn(mcv_rig* r, hipStream_t s) {
    const int np = (int)r->hp.size();
    const int* order = nullptr;
    if (np <= kPairOrderMax) {
        hipLaunchKernelGGL(k_pair_order, dim3(1), dim3(256), 0, s, r->pairs.as<MatchPair>(), np,
                           r->porder.as<int>());
        order = r->porder.as<int>();
    }
    launch_knn2(r->tune.knn, r->pairs.as<MatchPair>(), np, r->cap, s, -1, order);
}

static void rig_enqueue_batched(mcv_rig* r, const SinCosTables& tb, hipStream_t s, KernelTimer* rec = nullptr,
                                hipStream_t side = nullptr) {
    OrbPlan& P = *r->batch;
    P.enqueue(r->ctx->orb, tb, s, r->kps.as<mcv_keypoint>(), r->desc.as<uint8_t>(), r->cap, r->counts.as<int>(), rec,
              true, side);
    const int np = (int)r->hp.size();
    rig_knn(r, s);
    if (rec) rec->mark("k_knn2");
    hipLaunchKernelGGL(k_filter, dim3(np), dim3(1024), 0, s, r->jobs.as<FilterJob>());
    if (rec) rec->mark("k_filter");
    HIP_OK(hipGetLastError());
}

// Captures the current variants' graphs of the batched step (the image
// table copies stay outside: mcv_rig_run enqueues them before every launch).
static void rig_capture(mcv_rig* r, const SinCosTables& tb) {
    mcv_ctx* c = r->ctx;
    OrbPlan& P = *r->batch;
    P.tune = r->tune;
    hipStream_t cs = c->capture_stream();
    auto capture = [&](auto&& body) {
        hipGraph_t gr;
        hipGraphExec_t ge = nullptr;
        HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        body();
        HIP_OK(hipStreamEndCapture(cs, &gr));
        HIP_OK(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
        HIP_OK(hipGraphDestroy(gr));
        return ge;
    };
    mcv_rig::RigGraph G;
    G.t = r->tune;
    auto piece = [&](int part) {
        P.enqueue(c->orb, tb, cs, r->kps.as<mcv_keypoint>(), r->desc.as<uint8_t>(), r->cap, r->counts.as<int>(),
                  nullptr, true, nullptr, part);
    };
    if ((r->tune.blur & 2) == 0) {
        G.g = capture([&] { rig_enqueue_batched(r, tb, cs); });
    } else {
        G.g = capture([&] { piece(kPartPreA); });
        G.pre_b = capture([&] { piece(kPartPreB); });
        G.pre_c = capture([&] { piece(kPartPreC); });
        G.g2 = capture([&] {
            P.enqueue(c->orb, tb, cs, r->kps.as<mcv_keypoint>(), r->desc.as<uint8_t>(), r->cap, r->counts.as<int>(),
                      nullptr, true, nullptr, kPartPost);
            const int np = (int)r->hp.size();
            rig_knn(r, cs);
            hipLaunchKernelGGL(k_filter, dim3(np), dim3(1024), 0, cs, r->jobs.as<FilterJob>());
            HIP_OK(hipGetLastError());
        });
    }
    r->keep_graph(G);
}

// Launches the current variants' graphs on s.  A split step runs on s: the
// copy and the resizes below the split level (g), the other resizes and
// their FAST (pre_b), the cell suppression and the octree build (pre_c), the
// heap replay, then the keypoints, top-2 and filter (g2); on the context's
// side stream: FAST of the large levels once g is done (joined before pre_c)
// and the blur once pre_c is done (joined before g2).  A step's side work
// has finished before its g2 starts, so the next step's pyramid writes on s
// never overtake it.
static void rig_launch(mcv_rig* r, const SinCosTables& tb, hipStream_t s) {
    mcv_ctx* c = r->ctx;
    const mcv_rig::RigGraph& G = r->graphs[(size_t)r->cur];
    if (!G.g2) {
        HIP_OK(hipGraphLaunch(G.g, s));
        return;
    }
    OrbPlan& P = *r->batch;
    P.tune = r->tune;
    hipStream_t side = c->side_stream();
    for (hipEvent_t* e : {&r->ev_fork, &r->ev_join, &r->ev_a, &r->ev_b})
        if (!*e) HIP_OK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    HIP_OK(hipGraphLaunch(G.g, s));
    HIP_OK(hipEventRecord(r->ev_a, s));
    HIP_OK(hipStreamWaitEvent(side, r->ev_a, 0));
    P.enqueue(c->orb, tb, side, r->kps.as<mcv_keypoint>(), r->desc.as<uint8_t>(), r->cap, r->counts.as<int>(), nullptr,
              true, nullptr, kPartFastLo);
    HIP_OK(hipEventRecord(r->ev_b, side));
    HIP_OK(hipGraphLaunch(G.pre_b, s));
    HIP_OK(hipStreamWaitEvent(s, r->ev_b, 0));
    HIP_OK(hipGraphLaunch(G.pre_c, s));
    HIP_OK(hipEventRecord(r->ev_fork, s));
    HIP_OK(hipStreamWaitEvent(side, r->ev_fork, 0));
    P.enqueue(c->orb, tb, side, r->kps.as<mcv_keypoint>(), r->desc.as<uint8_t>(), r->cap, r->counts.as<int>(), nullptr,
              true, nullptr, kPartBlur);
    HIP_OK(hipEventRecord(r->ev_join, side));
    P.enqueue(c->orb, tb, s, r->kps.as<mcv_keypoint>(), r->desc.as<uint8_t>(), r->cap, r->counts.as<int>(), nullptr,
              true, nullptr, kPartHeap);
    HIP_OK(hipStreamWaitEvent(s, r->ev_join, 0));
    HIP_OK(hipGraphLaunch(G.g2, s));
}

// The batched sequence launched eagerly with an event after every kernel;
// ms[i] is the duration of the i-th launch, names[i] its kernel.
int mcv_rig_run_timed(mcv_rig* r, const uint8_t* const* d_images, float* ms, const char** names, int cap, int* n) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r && d_images && ms && n, MCV_ERR_INVALID, "bad timed run");
    c->set_device();
    const SinCosTables tb = c->tables();
    hipStream_t s = c->stream;
    OrbPlan& P = *r->batch;
    const int ni = 3 * r->n_frames;
    std::vector<int> st(ni);
    for (int i = 0; i < ni; i++) st[i] = r->w[i % 3];
    HIP_OK(hipMemcpyAsync(P.d_src.p, d_images, sizeof(void*) * ni, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(P.d_stride.p, st.data(), sizeof(int) * ni, hipMemcpyHostToDevice, s));
    r->timer.start(s);
    rig_enqueue_batched(r, tb, s, &r->timer);
    r->last_batched = true;
    HIP_OK(hipStreamSynchronize(s));
    const int k = (int)r->timer.names.size();
    *n = k;
    for (int i = 0; i < k && i < cap; i++) {
        float t = 0;
        HIP_OK(hipEventElapsedTime(&t, r->timer.ev[i], r->timer.ev[i + 1]));
        ms[i] = t;
        if (names) names[i] = r->timer.names[i];
    }
    API_END
}

// Autotuner: each tunable launch of the batched sequence is timed in every
// variant on the caller's images (median of `reps` eager runs, events
// around each launch), the fastest is kept for this rig and for later rigs
// of the same shape in this context, and the graph is re-captured on the
// next run.  Variants compute identical bytes.
int mcv_rig_autotune(mcv_rig* r, const uint8_t* const* d_images, int reps, char* report, int report_cap) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r && d_images && reps > 0, MCV_ERR_INVALID, "bad autotune arguments");
    c->set_device();
    const SinCosTables tb = c->tables();
    hipStream_t s = c->stream;
    OrbPlan& P = *r->batch;
    const int ni = 3 * r->n_frames;
    std::vector<int> st(ni);
    for (int i = 0; i < ni; i++) st[i] = r->w[i % 3];
    HIP_OK(hipMemcpyAsync(P.d_src.p, d_images, sizeof(void*) * ni, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(P.d_stride.p, st.data(), sizeof(int) * ni, hipMemcpyHostToDevice, s));
    auto time_kernel = [&](const char* name) -> float {
        std::vector<float> v;
        for (int k = 0; k < reps + 1; k++) {
            P.tune = r->tune;
            r->timer.start(s);
            rig_enqueue_batched(r, tb, s, &r->timer);
            HIP_OK(hipStreamSynchronize(s));
            float sum = 0;
            for (size_t i = 0; i < r->timer.names.size(); i++)
                if (strcmp(r->timer.names[i], name) == 0) {
                    float t = 0;
                    HIP_OK(hipEventElapsedTime(&t, r->timer.ev[i], r->timer.ev[i + 1]));
                    sum += t;
                }
            if (k > 0) v.push_back(sum);  // the first run warms caches
        }
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    std::string rep;
    char line[160];
    {
        int best = 0;
        float bt = 1e30f;
        for (int v = 0; v < kKnnVariants; v++) {
            r->tune.knn = v;
            const float t = time_kernel("k_knn2");
            snprintf(line, sizeof line, "k_knn2 %s %.4f ms\n", knn_variant_name(v), t);
            rep += line;
            if (t < bt) { bt = t; best = v; }
        }
        r->tune.knn = best;
    }
    {
        int best = 0;
        float bt = 1e30f;
        for (int v = 0; v < 2; v++) {  // the two walks; where the graph runs the blur comes last
            r->tune.blur = v;
            const float t = time_kernel("k_blur");
            snprintf(line, sizeof line, "k_blur %s %.4f ms\n", blur_variant_name(v), t);
            rep += line;
            if (t < bt) { bt = t; best = v; }
        }
        r->tune.blur = best;
    }
    {
        int best = 0;
        float bt = 1e30f;
        for (int v = 0; v < kKpVariants; v++) {
            r->tune.kp = v;
            const float t = time_kernel("k_keypoints");
            snprintf(line, sizeof line, "k_keypoints %s %.4f ms\n", kp_variant_name(v), t);
            rep += line;
            if (t < bt) { bt = t; best = v; }
        }
        r->tune.kp = best;
    }
    {
        int best = 0;
        float bt = 1e30f;
        for (int v = 0; v < kHeapVariants; v++) {
            r->tune.heap = v;
            const float t = time_kernel("k_octree_heap");
            snprintf(line, sizeof line, "k_octree_heap %s %.4f ms\n", heap_variant_name(v), t);
            rep += line;
            if (t < bt) { bt = t; best = v; }
        }
        r->tune.heap = best;
    }
    {
        // the whole captured graph with the blur before FAST and beside the
        // heap replay: median over reps of the mean of 4 back-to-back replays
        struct Ev2 {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Ev2() {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } ge;
        HIP_OK(hipEventCreate(&ge.e[0]));
        HIP_OK(hipEventCreate(&ge.e[1]));
        // each placement's graph is captured once (kept in the rig's graph
        // set) and timed in a second pass over the placements
        auto time_graph = [&]() -> float {
            P.tune = r->tune;
            r->pick_graph();
            if (r->cur < 0) rig_capture(r, tb);
            // back-to-back runs as mcv_rig_run enqueues them (the image
            // table copies, then the replay)
            constexpr int kBack = 4;
            auto launch = [&]() {
                HIP_OK(hipMemcpyAsync(P.d_src.p, d_images, sizeof(void*) * ni, hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync(P.d_stride.p, st.data(), sizeof(int) * ni, hipMemcpyHostToDevice, s));
                rig_launch(r, tb, s);
            };
            for (int j = 0; j < 3; j++) launch();
            HIP_OK(hipStreamSynchronize(s));
            std::vector<float> v;
            for (int k = 0; k < reps; k++) {
                HIP_OK(hipEventRecord(ge.e[0], s));
                for (int j = 0; j < kBack; j++) launch();
                HIP_OK(hipEventRecord(ge.e[1], s));
                HIP_OK(hipStreamSynchronize(s));
                float t = 0;
                HIP_OK(hipEventElapsedTime(&t, ge.e[0], ge.e[1]));
                v.push_back(t / kBack);
            }
            std::sort(v.begin(), v.end());
            return v[v.size() / 2];
        };
        const int kern = r->tune.blur & 1;
        int best = kern;
        float bt = 1e30f;
        r->drop_graphs();
        for (int pass = 0; pass < 2; pass++)
            for (int v = kern; v < kBlurVariants; v += 2) {
                r->tune.blur = v;
                const float t = time_graph();
                if (pass == 0) continue;
                snprintf(line, sizeof line, "graph, blur %s %.4f ms\n", blur_variant_name(v), t);
                rep += line;
                if (t < bt) { bt = t; best = v; }
            }
        r->tune.blur = best;
    }
    snprintf(line, sizeof line, "chosen: k_knn2 %s, k_blur %s, k_keypoints %s, k_octree_heap %s\n",
             knn_variant_name(r->tune.knn), blur_variant_name(r->tune.blur), kp_variant_name(r->tune.kp),
             heap_variant_name(r->tune.heap));
    rep += line;
    P.tune = r->tune;
    for (auto& sp : r->single) sp->tune = r->tune;
    c->tuned[r->shape_key] = r->tune;
    c->knn_variant = r->tune.knn;
    r->drop_graphs();
    if (report && report_cap > 0) {
        strncpy(report, rep.c_str(), (size_t)report_cap - 1);
        report[report_cap - 1] = 0;
    }
    API_END
}

int mcv_rig_set_variants(mcv_rig* r, int knn_variant, int blur_variant, int kp_variant, int heap_variant) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r && knn_variant >= 0 && knn_variant < kKnnVariants && blur_variant >= 0 && blur_variant < kBlurVariants &&
                kp_variant >= 0 && kp_variant < kKpVariants && heap_variant >= 0 && heap_variant < kHeapVariants,
            MCV_ERR_INVALID, "launch variant out of range");
    r->tune.knn = knn_variant;
    r->tune.blur = blur_variant;
    r->tune.kp = kp_variant;
    r->tune.heap = heap_variant;
    r->batch->tune = r->tune;
    for (auto& sp : r->single) sp->tune = r->tune;
    r->pick_graph();
    API_END
}

int mcv_rig_get_variants(const mcv_rig* r, int32_t* knn_variant, int32_t* blur_variant, int32_t* kp_variant,
                         int32_t* heap_variant) {
    if (!r) return MCV_ERR_INVALID;
    if (knn_variant) *knn_variant = r->tune.knn;
    if (blur_variant) *blur_variant = r->tune.blur;
    if (kp_variant) *kp_variant = r->tune.kp;
    if (heap_variant) *heap_variant = r->tune.heap;
    return MCV_OK;
}

int mcv_rig_run(mcv_rig* r, const uint8_t* const* d_images, int flags) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r && d_images, MCV_ERR_INVALID, "bad rig run");
    c->set_device();
    const SinCosTables tb = c->tables();
    const int n = 3 * r->n_frames;
    hipStream_t s = c->stream;
    if (flags & MCV_FLAG_GRAPH) {
        OrbPlan& P = *r->batch;
        std::vector<int> st(n);
        for (int i = 0; i < n; i++) st[i] = r->w[i % 3];
        HIP_OK(hipMemcpyAsync(P.d_src.p, d_images, sizeof(void*) * n, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(P.d_stride.p, st.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        if (r->cur < 0) rig_capture(r, tb);
        rig_launch(r, tb, s);
        r->last_batched = true;
    } else {
        r->last_batched = false;
        // plain path: one rig frame at a time, every image on its own launch
        // sequence (one single-image plan per camera), then the three pairs
        for (int f = 0; f < r->n_frames; f++) {
            for (int k = 0; k < 3; k++) {
                const int i = 3 * f + k;
                OrbPlan& P = *r->single[r->single.size() == 3 ? k : i];
                HIP_OK(hipMemcpyAsync(P.d_src.p, &d_images[i], sizeof(void*), hipMemcpyHostToDevice, s));
                HIP_OK(hipMemcpyAsync(P.d_stride.p, &r->w[k], sizeof(int), hipMemcpyHostToDevice, s));
                // a plan shared by every frame keeps the error bits of the run's earlier frames
                const bool reset_err = r->single.size() != 3 || f == 0;
                P.enqueue(c->orb, tb, s, r->kps.as<mcv_keypoint>() + (size_t)i * r->cap,
                          r->desc.as<uint8_t>() + (size_t)i * r->cap * 32, r->cap, r->counts.as<int>() + i, nullptr,
                          reset_err);
            }
            for (int k = 0; k < 3; k++) {
                const int p = 3 * f + k;
                launch_knn2(r->tune.knn, r->pairs.as<MatchPair>() + p, 1, r->cap, s);
                hipLaunchKernelGGL(k_filter, dim3(1), dim3(1024), 0, s, r->jobs.as<FilterJob>() + p);
            }
            // the single-image plans are reused by the next frame
            HIP_OK(hipStreamSynchronize(s));
        }
        HIP_OK(hipGetLastError());
    }
    API_END
}

int mcv_rig_stereo(mcv_rig* r, float bf, float baseline) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r && r->last_batched, MCV_ERR_STATE, "stereo needs the pyramids of a batched (graph) run");
    REQUIRE(r->w[0] == r->w[1] && r->h[0] == r->h[1], MCV_ERR_INVALID, "left and right images differ in size");
    REQUIRE(baseline > 0.f, MCV_ERR_INVALID, "baseline must be positive");
    c->set_device();
    hipStream_t s = c->stream;
    const int nf = r->n_frames, cap = r->cap, nl = c->orb.nlevels;
    r->st_u.ensure((size_t)nf * cap * 4);
    r->st_d.ensure((size_t)nf * cap * 4);
    r->st_sad.ensure((size_t)nf * cap * 4);
    r->st_keys.ensure((size_t)nf * cap * 8);
    std::vector<StereoJob> jobs(nf);
    for (int f = 0; f < nf; f++) {
        StereoJob& j = jobs[f];
        const int il = 3 * f, ir = 3 * f + 1;
        j.kl = reinterpret_cast<const float*>(r->kps.as<mcv_keypoint>() + (size_t)il * cap);
        j.dl = r->desc.as<uint8_t>() + (size_t)il * cap * 32;
        j.nl = r->counts.as<int>() + il;
        j.kr = reinterpret_cast<const float*>(r->kps.as<mcv_keypoint>() + (size_t)ir * cap);
        j.dr = r->desc.as<uint8_t>() + (size_t)ir * cap * 32;
        j.nr = r->counts.as<int>() + ir;
        j.lvl_l = il * nl;
        j.lvl_r = ir * nl;
        j.u_right = r->st_u.as<float>() + (size_t)f * cap;
        j.depth = r->st_d.as<float>() + (size_t)f * cap;
        j.sad = r->st_sad.as<int>() + (size_t)f * cap;
        j.keys = r->st_keys.as<unsigned long long>() + (size_t)f * cap;
    }
    upload(r->st_jobs, jobs, s);
    StereoParams sp{};
    for (int l = 0; l < nl; l++) {
        sp.scale[l] = c->orb.scale[l];
        sp.inv_scale[l] = c->orb.inv_scale[l];
    }
    sp.bf = bf;
    sp.b = baseline;
    sp.cap = cap;
    REQUIRE(cap <= kStereoSortCap, MCV_ERR_INVALID, "too many keypoints per image for the stereo median");
    hipLaunchKernelGGL(k_stereo, dim3((cap + 3) / 4, nf), dim3(256), 0, s, r->st_jobs.as<StereoJob>(),
                       r->batch->d_lv.as<LevelDesc>(), r->batch->d_pyr.as<uint8_t>(), sp);
    hipLaunchKernelGGL(k_stereo_median, dim3(nf), dim3(256), 0, s, r->st_jobs.as<StereoJob>(), sp);
    HIP_OK(hipGetLastError());
    API_END
}

int mcv_rig_stereo_outputs(mcv_rig* r, float** u_right, float** depth, int32_t** sad) {
    if (!r) return MCV_ERR_INVALID;
    if (u_right) *u_right = r->st_u.as<float>();
    if (depth) *depth = r->st_d.as<float>();
    if (sad) *sad = r->st_sad.as<int32_t>();
    return MCV_OK;
}

int mcv_rig_check(mcv_rig* r) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r, MCV_ERR_INVALID, "null rig");
    c->set_device();
    HIP_OK(hipStreamSynchronize(c->stream));
    const int k0 = r->last_batched ? -1 : 0;
    const int k1 = r->last_batched ? 0 : (int)r->single.size();
    for (int k = k0; k < k1; k++) {
        OrbPlan& P = k < 0 ? *r->batch : *r->single[k];
        if (plan_check(P, c->stream) != 0) {
            // grown for the next run; the rig's captured graphs hold the old pool
            if (k < 0) r->drop_graphs();
            throw Fail{MCV_ERR_STATE, "octree node pool exhausted in the last run (grown for the next one)"};
        }
    }
    API_END
}

int mcv_rig_octree_stats(mcv_rig* r, int64_t* candidates, int64_t* table_entries, int64_t* nodes) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r, MCV_ERR_INVALID, "null rig");
    c->set_device();
    OrbPlan& P = r->last_batched ? *r->batch : *r->single.back();
    std::vector<OctLevel> o(P.lv.size());
    HIP_OK(hipMemcpyAsync(o.data(), P.d_oct.p, o.size() * sizeof(OctLevel), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    int64_t nc = 0, nt = 0, nn = 0;
    for (const OctLevel& l : o) {
        nc += l.n;
        nt += l.failed ? 0 : l.total;
        nn += l.nout;
    }
    if (candidates) *candidates = nc;
    if (table_entries) *table_entries = nt;
    if (nodes) *nodes = nn;
    API_END
}

int mcv_rig_outputs(mcv_rig* r, mcv_keypoint** kps, uint8_t** desc, int32_t** counts, int32_t** matches,
                    int32_t** mcounts, int32_t** knn, int32_t* cap) {
    if (!r) return MCV_ERR_INVALID;
    if (kps) *kps = r->kps.as<mcv_keypoint>();
    if (desc) *desc = r->desc.as<uint8_t>();
    if (counts) *counts = r->counts.as<int32_t>();
    if (matches) *matches = r->matches.as<int32_t>();
    if (mcounts) *mcounts = r->mcounts.as<int32_t>();
    if (knn) *knn = r->knn.as<int32_t>();
    if (cap) *cap = r->cap;
    return MCV_OK;
}

int mcv_rig_pyramid_dev(mcv_rig* r, int image, int level, int blurred, const uint8_t** d_level, int32_t* pitch,
                        int32_t* w, int32_t* h) {
    mcv_ctx* c = r ? r->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(r && r->last_batched, MCV_ERR_STATE, "the rig's pyramids are resident after a batched (graph) run");
    level_view(*r->batch, image, level, blurred, d_level, pitch, w, h);
    API_END
}

// ---- window search ----------------------------------------------------------
// One launch of k_window_knn over n_cams cameras (each with its own keypoint
// set and grid): map points (d_pts) or direct positions (d_uv).
struct WindowQuery {
    const float* d_pts = nullptr;
    const float* d_uv = nullptr;
    const uint8_t* d_qdesc = nullptr;
    const int32_t* d_plev = nullptr;
    int P = 0;
    const float* Rt = nullptr;      // host: n_cams x 12 (map points)
    const float* cams = nullptr;    // host: n_cams x 4 (map points)
    const float* scales = nullptr;  // host: nl (map points)
    int nl = 0;
    int min_level = -1, max_level = -1;  // direct positions
    float radius = 0.f, ratio = 0.f;
    int th = 0;
};

static void window_run(mcv_ctx* c, int n_cams, const WindowQuery& q, const int32_t* sizes,
                       const float* const* d_kxy, const int32_t* const* d_koct, const uint8_t* const* d_kdesc,
                       const int32_t* n_kps, int32_t* d_out) {
    if (q.d_pts) REQUIRE(q.nl >= 1 && q.nl <= 16, MCV_ERR_INVALID, "nlevels out of range");
    hipStream_t s = c->stream;
    int total_k = 0;
    for (int k = 0; k < n_cams; k++) {
        REQUIRE(n_kps[k] >= 0 && n_kps[k] <= kGridSortCap && n_kps[k] < 65536, MCV_ERR_INVALID, "too many keypoints");
        REQUIRE(n_kps[k] == 0 || (d_kxy[k] && d_kdesc[k]), MCV_ERR_INVALID, "null keypoint arrays");
        REQUIRE(sizes[2 * k] > 0 && sizes[2 * k + 1] > 0, MCV_ERR_INVALID, "bad image size");
        total_k += n_kps[k];
    }
    const bool level_filter = !q.d_pts && (q.min_level > 0 || q.max_level >= 0);
    if (level_filter)
        for (int k = 0; k < n_cams; k++) REQUIRE(n_kps[k] == 0 || (d_koct && d_koct[k]), MCV_ERR_INVALID,
                                                 "a level-filtered query needs keypoint octaves");
    c->w_cs.ensure((size_t)n_cams * (kGridCells + 1) * 4);
    c->w_ci.ensure((size_t)std::max(total_k, 1) * 4);
    std::vector<GridView> views(n_cams);
    std::vector<WindowCam> wc(n_cams);
    int off = 0;
    for (int k = 0; k < n_cams; k++) {
        GridView& g = views[k];
        g.kxy = d_kxy[k];
        g.koct = level_filter ? d_koct[k] : nullptr;
        g.kdesc = d_kdesc[k];
        g.n = n_kps[k];
        g.width = sizes[2 * k];
        g.height = sizes[2 * k + 1];
        g.inv_w = (float)(30.0 / g.width);
        g.inv_h = (float)(30.0 / g.height);
        g.cell_start = c->w_cs.as<int>() + (size_t)k * (kGridCells + 1);
        g.cell_items = c->w_ci.as<int>() + off;
        off += n_kps[k];
        WindowCam& w = wc[k];
        w.grid = g;
        if (q.d_pts) {
            for (int i = 0; i < 9; i++) w.R[i] = q.Rt[12 * k + i];
            for (int i = 0; i < 3; i++) w.t[i] = q.Rt[12 * k + 9 + i];
            w.fx = q.cams[4 * k]; w.fy = q.cams[4 * k + 1]; w.cx = q.cams[4 * k + 2]; w.cy = q.cams[4 * k + 3];
        }
        w.out = d_out + (size_t)k * q.P * 6;
    }
    upload(c->w_view, views, s);
    upload(c->w_cams, wc, s);
    hipLaunchKernelGGL(k_grid_build, dim3(n_cams), dim3(1024), 0, s, c->w_view.as<GridView>());
    WindowPoints wp{};
    wp.pts = q.d_pts;
    wp.uv = q.d_uv;
    wp.pdesc = q.d_qdesc;
    wp.plevel = q.d_plev;
    wp.P = q.P;
    wp.radius = q.radius;
    wp.ratio = q.ratio;
    wp.th = q.th;
    wp.nl = q.nl;
    wp.min_level = q.min_level;
    wp.max_level = q.max_level;
    for (int i = 0; i < q.nl && i < 16; i++) wp.scales[i] = q.scales[i];
    if (q.P > 0)
        hipLaunchKernelGGL(k_window_knn, dim3((q.P + 3) / 4, n_cams), dim3(256), 0, s, c->w_cams.as<WindowCam>(), wp);
    HIP_OK(hipGetLastError());
}

int mcv_window_search_dev(mcv_ctx* c, int n_cams, const float* d_points, const uint8_t* d_pdesc, const int32_t* d_plevel,
                          int n_points, const float* Rt, const float* cams, const int32_t* sizes, const float* scales,
                          int nlevels, float radius, float ratio, int th, const float* const* d_kxy,
                          const uint8_t* const* d_kdesc, const int32_t* n_kps, int32_t* d_out) {
    API_BEGIN(c)
    REQUIRE(c && n_cams > 0 && n_points >= 0 && Rt && cams && sizes && scales && d_kxy && d_kdesc && n_kps,
            MCV_ERR_INVALID, "bad window search");
    REQUIRE(n_points == 0 || (d_points && d_pdesc && d_plevel && d_out), MCV_ERR_INVALID, "null map point arrays");
    c->set_device();
    WindowQuery q;
    q.d_pts = d_points;
    q.d_qdesc = d_pdesc;
    q.d_plev = d_plevel;
    q.P = n_points;
    q.Rt = Rt;
    q.cams = cams;
    q.scales = scales;
    q.nl = nlevels;
    q.radius = radius;
    q.ratio = ratio;
    q.th = th;
    window_run(c, n_cams, q, sizes, d_kxy, nullptr, d_kdesc, n_kps, d_out);
    API_END
}

int mcv_window_search_uv_dev(mcv_ctx* c, const float* d_uv, const uint8_t* d_qdesc, int n, float radius, int min_level,
                             int max_level, float ratio, int th, const float* d_kxy, const int32_t* d_koct,
                             const uint8_t* d_kdesc, int n_kps, int width, int height, int32_t* d_out) {
    API_BEGIN(c)
    REQUIRE(c && n >= 0 && n_kps >= 0, MCV_ERR_INVALID, "bad window query");
    REQUIRE(n == 0 || (d_uv && d_qdesc && d_out), MCV_ERR_INVALID, "null query arrays");
    if (n == 0) return MCV_OK;
    c->set_device();
    WindowQuery q;
    q.d_uv = d_uv;
    q.d_qdesc = d_qdesc;
    q.P = n;
    q.min_level = min_level;
    q.max_level = max_level;
    q.radius = radius;
    q.ratio = ratio;
    q.th = th;
    const int32_t sz[2] = {width, height};
    window_run(c, 1, q, sz, &d_kxy, &d_koct, &d_kdesc, &n_kps, d_out);
    API_END
}

int mcv_window_search(mcv_ctx* c, const float* points, const uint8_t* pdesc, const int32_t* plevel, int P, const float* R,
                      const float* t, const float* cam, const float* scales, int nl, float radius, float ratio, int th,
                      const float* kxy, const uint8_t* kdesc, int K, int width, int height, int32_t* out) {
    API_BEGIN(c)
    REQUIRE(c && P >= 0 && K >= 0, MCV_ERR_INVALID, "bad window search");
    if (P == 0) return MCV_OK;
    REQUIRE(points && pdesc && plevel && R && t && cam && scales && out && (K == 0 || (kxy && kdesc)),
            MCV_ERR_INVALID, "null window-search arrays");
    REQUIRE(nl >= 1 && nl <= 16, MCV_ERR_INVALID, "nlevels out of range");
    for (int i = 0; i < P; i++)
        REQUIRE(plevel[i] >= 0 && plevel[i] < nl, MCV_ERR_INVALID, "map point level outside the scale table");
    c->set_device();
    hipStream_t s = c->stream;
    c->w_pts.ensure((size_t)P * 12);
    c->w_pdesc.ensure((size_t)P * 32);
    c->w_plev.ensure((size_t)P * 4);
    c->w_kxy.ensure((size_t)std::max(K, 1) * 8);
    c->w_kdesc.ensure((size_t)std::max(K, 1) * 32);
    c->w_out.ensure((size_t)P * 24);
    HIP_OK(hipMemcpyAsync(c->w_pts.p, points, (size_t)P * 12, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(c->w_pdesc.p, pdesc, (size_t)P * 32, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(c->w_plev.p, plevel, (size_t)P * 4, hipMemcpyHostToDevice, s));
    if (K) {
        HIP_OK(hipMemcpyAsync(c->w_kxy.p, kxy, (size_t)K * 8, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(c->w_kdesc.p, kdesc, (size_t)K * 32, hipMemcpyHostToDevice, s));
    }
    float Rt[12];
    for (int i = 0; i < 9; i++) Rt[i] = R[i];
    for (int i = 0; i < 3; i++) Rt[9 + i] = t[i];
    const int32_t sz[2] = {width, height};
    const float* kx = c->w_kxy.as<float>();
    const uint8_t* kd = c->w_kdesc.as<uint8_t>();
    WindowQuery q;
    q.d_pts = c->w_pts.as<float>();
    q.d_qdesc = c->w_pdesc.as<uint8_t>();
    q.d_plev = c->w_plev.as<int32_t>();
    q.P = P;
    q.Rt = Rt;
    q.cams = cam;
    q.scales = scales;
    q.nl = nl;
    q.radius = radius;
    q.ratio = ratio;
    q.th = th;
    window_run(c, 1, q, sz, &kx, nullptr, &kd, &K, c->w_out.as<int32_t>());
    HIP_OK(hipMemcpyAsync(out, c->w_out.p, (size_t)P * 24, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    API_END
}

// ---- bag of words (DBoW3) ------------------------------------------------------
}  // extern "C"

struct mcv_bow_vocab {
    mcv_ctx* ctx;
    int n = 0, L = 0, depth = 0, max_children = 0, weighting = 0, scoring = 0;
    DevBuf desc, cptr, cids, wid, wt, feats, word, weight, nid;
};

extern "C" {

int mcv_bow_vocab_create(mcv_ctx* c, int n, const int32_t* child_ptr, const int32_t* child_ids, const uint8_t* desc,
                         const int32_t* word_id, const double* weight, int L, int weighting, int scoring,
                         mcv_bow_vocab** out) {
    API_BEGIN(c)
    REQUIRE(c && out && n >= 1 && child_ptr && desc && word_id && weight, MCV_ERR_INVALID, "bad vocabulary");
    REQUIRE(weighting >= 0 && weighting <= 3 && scoring >= 0 && scoring <= 5, MCV_ERR_INVALID,
            "unknown weighting or scoring type");
    REQUIRE(child_ptr[0] == 0, MCV_ERR_INVALID, "child_ptr must start at 0");
    int maxc = 0;
    for (int i = 0; i < n; i++) {
        REQUIRE(child_ptr[i + 1] >= child_ptr[i], MCV_ERR_INVALID, "child_ptr must not decrease");
        maxc = std::max(maxc, child_ptr[i + 1] - child_ptr[i]);
    }
    const int nc = child_ptr[n];
    REQUIRE(nc == 0 || child_ids, MCV_ERR_INVALID, "null child ids");
    REQUIRE(maxc < 65536, MCV_ERR_INVALID, "more than 65535 children under one node");
    // every node reached once from the root (a tree), depth = longest path
    std::vector<int> depth(n, -1);
    std::vector<int> todo{0};
    depth[0] = 0;
    int dmax = 0;
    for (size_t k = 0; k < todo.size(); k++) {
        const int i = todo[k];
        for (int e = child_ptr[i]; e < child_ptr[i + 1]; e++) {
            const int j = child_ids[e];
            REQUIRE(j > 0 && j < n && depth[j] < 0, MCV_ERR_INVALID, "child ids do not form a tree rooted at 0");
            depth[j] = depth[i] + 1;
            dmax = std::max(dmax, depth[j]);
            todo.push_back(j);
        }
    }
    c->set_device();
    hipStream_t s = c->stream;
    std::unique_ptr<mcv_bow_vocab> v(new mcv_bow_vocab());
    v->ctx = c;
    v->n = n;
    v->L = L;
    v->depth = dmax;
    v->max_children = maxc;
    v->weighting = weighting;
    v->scoring = scoring;
    v->desc.ensure((size_t)n * 32);
    HIP_OK(hipMemcpyAsync(v->desc.p, desc, (size_t)n * 32, hipMemcpyHostToDevice, s));
    upload(v->cptr, std::vector<int32_t>(child_ptr, child_ptr + n + 1), s);
    upload(v->cids, std::vector<int32_t>(child_ids, child_ids + nc), s);
    upload(v->wid, std::vector<int32_t>(word_id, word_id + n), s);
    upload(v->wt, std::vector<double>(weight, weight + n), s);
    HIP_OK(hipStreamSynchronize(s));
    *out = v.release();
    API_END
}

void mcv_bow_vocab_destroy(mcv_bow_vocab* v) { delete v; }

int mcv_bow_transform_dev(mcv_bow_vocab* v, const uint8_t* d_feats, int n, int levelsup, int32_t* d_word,
                          double* d_weight, int32_t* d_nid) {
    mcv_ctx* c = v ? v->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(v && n >= 0, MCV_ERR_INVALID, "bad transform");
    if (n == 0) return MCV_OK;
    REQUIRE(d_feats && d_word && d_weight && d_nid, MCV_ERR_INVALID, "null buffers");
    c->set_device();
    BowVocab V{v->desc.as<uint32_t>(), v->cptr.as<int32_t>(), v->cids.as<int32_t>(), v->wid.as<int32_t>(),
               v->wt.as<double>(), v->depth};
    const int nid_level = v->L - levelsup;
    const uint32_t* f = reinterpret_cast<const uint32_t*>(d_feats);
    if (v->max_children <= 16)
        hipLaunchKernelGGL(k_bow_transform<16>, dim3((n + 15) / 16), dim3(256), 0, c->stream, V, f, n, nid_level,
                           d_word, d_weight, d_nid);
    else
        hipLaunchKernelGGL(k_bow_transform<64>, dim3((n + 3) / 4), dim3(256), 0, c->stream, V, f, n, nid_level,
                           d_word, d_weight, d_nid);
    HIP_OK(hipGetLastError());
    API_END
}

int mcv_bow_transform(mcv_bow_vocab* v, const uint8_t* feats, int n, int levelsup, int32_t* word, double* weight,
                      int32_t* nid, int32_t* bv_words, double* bv_vals, int32_t* n_words, int32_t* fv_nodes,
                      int32_t* fv_ptr, int32_t* fv_feats, int32_t* n_nodes) {
    mcv_ctx* c = v ? v->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(v && n >= 0 && (n == 0 || feats), MCV_ERR_INVALID, "bad transform");
    c->set_device();
    hipStream_t s = c->stream;
    std::vector<int32_t> w(n), nd(n);
    std::vector<double> wt(n);
    if (n > 0) {
        v->feats.ensure((size_t)n * 32);
        v->word.ensure((size_t)n * 4);
        v->weight.ensure((size_t)n * 8);
        v->nid.ensure((size_t)n * 4);
        HIP_OK(hipMemcpyAsync(v->feats.p, feats, (size_t)n * 32, hipMemcpyHostToDevice, s));
        const int rc = mcv_bow_transform_dev(v, v->feats.as<uint8_t>(), n, levelsup, v->word.as<int32_t>(),
                                             v->weight.as<double>(), v->nid.as<int32_t>());
        if (rc != MCV_OK) return rc;
        HIP_OK(hipMemcpyAsync(w.data(), v->word.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(wt.data(), v->weight.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(nd.data(), v->nid.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    if (word) std::copy(w.begin(), w.end(), word);
    if (weight) std::copy(wt.begin(), wt.end(), weight);
    if (nid) std::copy(nd.begin(), nd.end(), nid);
    // Vocabulary::transform(features, BowVector, FeatureVector, levelsup)
    // (Vocabulary.cpp:572-629): TF / TF_IDF accumulate (BowVector::addWeight),
    // IDF / BINARY keep the first (addIfNotExist), stopped words (weight <= 0)
    // skipped; FeatureVector::addFeature (FeatureVector.cpp:31-45); the
    // scoring's normalisation (ScoringObject.h: L2 for L2 scoring, L1 for the
    // others but the dot product) or, without one, TF weights / v.size()
    std::map<int32_t, double> bv;
    std::map<int32_t, std::vector<int32_t>> fv;
    const bool must = v->scoring != 5, l2 = v->scoring == 1;
    for (int i = 0; i < n; i++) {
        if (!(wt[i] > 0)) continue;
        if (v->weighting <= 1) bv[w[i]] += wt[i];
        else bv.emplace(w[i], wt[i]);
        fv[nd[i]].push_back(i);
    }
    if (v->weighting <= 1 && !bv.empty() && !must) {
        const double m = (double)bv.size();
        for (auto& kv : bv) kv.second /= m;
    }
    if (must) {
        double norm = 0.0;
        if (!l2) {
            for (auto& kv : bv) norm += std::fabs(kv.second);
        } else {
            for (auto& kv : bv) norm += kv.second * kv.second;
            norm = std::sqrt(norm);
        }
        if (norm > 0.0)
            for (auto& kv : bv) kv.second /= norm;
    }
    if (n_words) *n_words = (int32_t)bv.size();
    if (bv_words || bv_vals) {
        int k = 0;
        for (auto& kv : bv) {
            if (bv_words) bv_words[k] = kv.first;
            if (bv_vals) bv_vals[k] = kv.second;
            k++;
        }
    }
    if (n_nodes) *n_nodes = (int32_t)fv.size();
    if (fv_nodes || fv_ptr || fv_feats) {
        int m = 0, t = 0;
        if (fv_ptr) fv_ptr[0] = 0;
        for (auto& kv : fv) {
            if (fv_nodes) fv_nodes[m] = kv.first;
            for (int32_t i : kv.second) {
                if (fv_feats) fv_feats[t] = i;
                t++;
            }
            m++;
            if (fv_ptr) fv_ptr[m] = t;
        }
    }
    API_END
}

int mcv_bow_match(mcv_ctx* c, const uint8_t* d1, int n1, int m1, const int32_t* nodes1, const int32_t* ptr1,
                  const int32_t* feats1, const uint8_t* d2, int n2, int m2, const int32_t* nodes2,
                  const int32_t* ptr2, const int32_t* feats2, int32_t* out, int32_t* count) {
    API_BEGIN(c)
    REQUIRE(c && count && m1 >= 0 && m2 >= 0 && n1 >= 0 && n2 >= 0, MCV_ERR_INVALID, "bad DBowMatch arguments");
    *count = 0;
    if (m1 == 0 || m2 == 0) return MCV_OK;
    REQUIRE(d1 && d2 && nodes1 && ptr1 && feats1 && nodes2 && ptr2 && feats2 && out, MCV_ERR_INVALID, "null arrays");
    for (int a = 0; a < m1; a++)
        for (int e = ptr1[a]; e < ptr1[a + 1]; e++)
            REQUIRE(feats1[e] >= 0 && feats1[e] < n1, MCV_ERR_INVALID, "feature index out of range");
    for (int b = 0; b < m2; b++)
        for (int e = ptr2[b]; e < ptr2[b + 1]; e++)
            REQUIRE(feats2[e] >= 0 && feats2[e] < n2, MCV_ERR_INVALID, "feature index out of range");
    // the nodes both vectors hold, ascending (the reference's merge with
    // lower_bound jumps visits exactly these); a row per feature of image 1
    // in a node where image 2 has at least two features
    std::vector<BowJob> jobs;
    int a = 0, b = 0;
    while (a < m1 && b < m2) {
        if (nodes1[a] == nodes2[b]) {
            if (ptr2[b + 1] - ptr2[b] >= 2)
                for (int e = ptr1[a]; e < ptr1[a + 1]; e++)
                    jobs.push_back(BowJob{feats1[e], ptr2[b], ptr2[b + 1], (int)jobs.size()});
            a++;
            b++;
        } else if (nodes1[a] < nodes2[b]) {
            a++;
        } else {
            b++;
        }
    }
    const int nj = (int)jobs.size();
    if (nj == 0) return MCV_OK;
    c->set_device();
    hipStream_t s = c->stream;
    c->b_d1.ensure((size_t)n1 * 32);
    c->b_d2.ensure((size_t)n2 * 32);
    c->b_out.ensure((size_t)nj * 20);
    HIP_OK(hipMemcpyAsync(c->b_d1.p, d1, (size_t)n1 * 32, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(c->b_d2.p, d2, (size_t)n2 * 32, hipMemcpyHostToDevice, s));
    upload(c->b_f2, std::vector<int32_t>(feats2, feats2 + ptr2[m2]), s);
    upload(c->b_jobs, jobs, s);
    hipLaunchKernelGGL(k_bow_match, dim3((nj + 3) / 4), dim3(256), 0, s, c->b_jobs.as<BowJob>(), nj,
                       c->b_d1.as<uint32_t>(), c->b_d2.as<uint32_t>(), c->b_f2.as<int32_t>(), c->b_out.as<int32_t>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, c->b_out.p, (size_t)nj * 20, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    *count = nj;
    API_END
}

// ---- bundle adjustment --------------------------------------------------------
int mcv_ba_create(mcv_ctx* c, int M, int P, int F, int C, mcv_ba** out) {
    API_BEGIN(c)
    REQUIRE(c && out && M >= 0 && P > 0 && F > 0 && C > 0, MCV_ERR_INVALID, "bad BA sizes");
    c->set_device();
    std::unique_ptr<mcv_ba> b(new mcv_ba());
    b->ctx = c;
    b->M = M; b->P = P; b->F = F; b->C = C;
    const size_t m = (size_t)std::max(M, 1);
    b->uv.ensure(m * 8);
    b->pt.ensure(m * 4);
    b->pair.ensure(m * 4);
    b->isig.ensure(m * 4);
    b->phi.ensure((size_t)P * 16);
    b->plo.ensure((size_t)P * 16);
    b->pairs.ensure((size_t)F * C * sizeof(BAPair));
    b->planes.ensure(m * 22 * 4);
    b->pptr.ensure((size_t)(P + 1) * 4);
    b->fptr.ensure((size_t)(F + 1) * 4);
    b->fobs.ensure(m * 4);
    *out = b.release();
    API_END
}

void mcv_ba_destroy(mcv_ba* b) { delete b; }

int mcv_ba_set_observations(mcv_ba* b, const float* uv, const int32_t* pt, const int32_t* fr, const int32_t* cam,
                            const float* isig) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && (b->M == 0 || (uv && pt && fr && cam && isig)), MCV_ERR_INVALID, "null observation arrays");
    c->set_device();
    hipStream_t s = c->stream;
    const int M = b->M;
    std::vector<uint32_t> pair(M);
    std::vector<int> pcount(b->P + 1, 0), fcount(b->F + 1, 0);
    bool sorted = true;
    for (int i = 0; i < M; i++) {
        REQUIRE(pt[i] >= 0 && pt[i] < b->P && fr[i] >= 0 && fr[i] < b->F && cam[i] >= 0 && cam[i] < b->C,
                MCV_ERR_INVALID, "observation index out of range");
        pair[i] = (uint32_t)(fr[i] * b->C + cam[i]);
        pcount[pt[i] + 1]++;
        fcount[fr[i] + 1]++;
        if (i && pt[i] < pt[i - 1]) sorted = false;
    }
    for (int p = 0; p < b->P; p++) pcount[p + 1] += pcount[p];
    for (int f = 0; f < b->F; f++) fcount[f + 1] += fcount[f];
    std::vector<int> fobs(std::max(M, 1));
    {
        std::vector<int> pos(fcount.begin(), fcount.end() - 1);
        for (int i = 0; i < M; i++) fobs[pos[fr[i]]++] = i;
    }
    const int G = b->F * b->C;
    b->h_gptr.assign(G + 1, 0);
    for (int i = 0; i < M; i++) b->h_gptr[pair[i] + 1]++;
    for (int g = 0; g < G; g++) b->h_gptr[g + 1] += b->h_gptr[g];
    std::vector<int> gobs(std::max(M, 1));
    {
        std::vector<int> pos(b->h_gptr.begin(), b->h_gptr.end() - 1);
        for (int i = 0; i < M; i++) gobs[pos[pair[i]]++] = i;
    }
    b->gobs.ensure((size_t)std::max(M, 1) * 4);
    if (M) {
        HIP_OK(hipMemcpyAsync(b->uv.p, uv, (size_t)M * 8, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->pt.p, pt, (size_t)M * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->pair.p, pair.data(), (size_t)M * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->isig.p, isig, (size_t)M * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->fobs.p, fobs.data(), (size_t)M * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->gobs.p, gobs.data(), (size_t)M * 4, hipMemcpyHostToDevice, s));
    }
    HIP_OK(hipMemcpyAsync(b->fptr.p, fcount.data(), (size_t)(b->F + 1) * 4, hipMemcpyHostToDevice, s));
    if (sorted) HIP_OK(hipMemcpyAsync(b->pptr.p, pcount.data(), (size_t)(b->P + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    b->have_obs = sorted;
    REQUIRE(sorted, MCV_ERR_INVALID, "observations must be ordered by point");
    b->h_pt.assign(pt, pt + M);
    b->h_fr.assign(fr, fr + M);
    b->h_pptr = pcount;
    b->schur_ready = false;
    b->touched(-1);
    API_END
}

// double -> (hi, lo) float pair
static inline void split(double v, float& hi, float& lo) {
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

static void quat_to_R(const double* q, double R[9]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const double twx = tx * w, twy = ty * w, twz = tz * w;
    const double txx = tx * x, txy = ty * x, txz = tz * x;
    const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R[0] = 1 - (tyy + tzz); R[1] = txy - twz; R[2] = txz + twy;
    R[3] = txy + twz; R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
    R[6] = txz - twy; R[7] = tyz + twx; R[8] = 1 - (txx + tyy);
}

// The (frame, camera) table of the mono edges and the frame poses of the
// stereo edges from the host FP64 frames and cameras; without the robust
// kernel the Huber deltas are infinite (w = 1, rho = chi2).
static void ba_upload_frames(mcv_ba* b, hipStream_t s) {
    const double* frames = b->h_frames.data();
    const double* cams = b->h_cams.data();
    std::vector<BAPair> pr((size_t)b->F * b->C);
    std::vector<FramePose> fp(b->F);
    for (int f = 0; f < b->F; f++) {
        double Rbw[9];
        const double* fq = frames + 7 * f;
        // g2o SE3Quat keeps a normalised quaternion with w >= 0 (se3quat.h:280-285)
        double q[4] = {fq[0], fq[1], fq[2], fq[3]};
        double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (double& v : q) v /= qn;
        quat_to_R(q, Rbw);
        const double* tbw = fq + 4;
        for (int i = 0; i < 9; i++) split(Rbw[i], fp[f].Rh[i], fp[f].Rl[i]);
        for (int i = 0; i < 3; i++) split(tbw[i], fp[f].th[i], fp[f].tl[i]);
        for (int k = 0; k < b->C; k++) {
            const double* cm = cams + 12 * k;
            double Rcb[9];
            double qc[4] = {cm[4], cm[5], cm[6], cm[7]};
            double qcn = std::sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2] + qc[3] * qc[3]);
            for (double& v : qc) v /= qcn;
            quat_to_R(qc, Rcb);
            const double* tcb = cm + 8;
            BAPair& P = pr[(size_t)f * b->C + k];
            // Tcw = Tcb * Tbw
            for (int r = 0; r < 3; r++) {
                for (int cc = 0; cc < 3; cc++) {
                    double v = 0;
                    for (int m = 0; m < 3; m++) v += Rcb[3 * r + m] * Rbw[3 * m + cc];
                    split(v, P.Rh[3 * r + cc], P.Rl[3 * r + cc]);
                }
                double tv = tcb[r];
                for (int m = 0; m < 3; m++) tv += Rcb[3 * r + m] * tbw[m];
                split(tv, P.th[r], P.tl[r]);
            }
            for (int i = 0; i < 9; i++) { P.Rbw[i] = (float)Rbw[i]; P.Rcb[i] = (float)Rcb[i]; }
            for (int i = 0; i < 3; i++) P.tbw[i] = (float)tbw[i];
            P.fx = (float)cm[0]; P.fy = (float)cm[1]; P.cx = (float)cm[2]; P.cy = (float)cm[3];
            P.delta = b->robust ? (float)cm[11] : INFINITY;
            P.pad[0] = P.pad[1] = P.pad[2] = 0.f;
        }
    }
    upload(b->pairs, pr, s);
    upload(b->fposes, fp, s);
    const double* sc = b->h_scam;
    b->scam = StereoCam{(float)sc[0], (float)sc[1], (float)sc[2], (float)sc[3], (float)sc[4],
                        b->robust ? (float)sc[5] : INFINITY};
}

// FP64 points to the device and their float pairs
static void ba_upload_points(mcv_ba* b, const double* points, hipStream_t s) {
    b->p64.ensure((size_t)b->P * 24);
    b->phi.ensure((size_t)b->P * 16);
    b->plo.ensure((size_t)b->P * 16);
    HIP_OK(hipMemcpyAsync(b->p64.p, points, (size_t)b->P * 24, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_ba_points_apply, dim3((b->P + 255) / 256), dim3(256), 0, s, b->p64.as<double>(),
                       (const float*)nullptr, (const uint8_t*)nullptr, b->P, b->phi.as<float4>(), b->plo.as<float4>());
    HIP_OK(hipGetLastError());
}

int mcv_ba_set_state(mcv_ba* b, const double* points, const double* frames, const double* cams) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && points && frames && cams, MCV_ERR_INVALID, "null state arrays");
    c->set_device();
    hipStream_t s = c->stream;
    b->h_frames.assign(frames, frames + 7 * (size_t)b->F);
    b->h_cams.assign(cams, cams + 12 * (size_t)b->C);
    ba_upload_points(b, points, s);
    ba_upload_frames(b, s);
    HIP_OK(hipStreamSynchronize(s));
    b->have_state = true;
    b->touched(-1);
    API_END
}

int mcv_ba_linearize(mcv_ba* b, int residual_only) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->have_obs && b->have_state, MCV_ERR_STATE, "observations or state not set");
    c->set_device();
    REQUIRE(residual_only >= kBaFull && residual_only <= kBaPoseOnly, MCV_ERR_INVALID, "unknown linearization mode");
    ba_enqueue_linearize(b, b->variant, residual_only, c->stream);
    b->touched(residual_only);
    API_END
}

int mcv_ba_set_variant(mcv_ba* b, int variant) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && variant >= 0 && variant < kBaVariants, MCV_ERR_INVALID, "BA launch variant out of range");
    b->variant = variant;
    API_END
}

int mcv_ba_autotune(mcv_ba* b, int reps, char* report, int report_cap) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && reps > 0, MCV_ERR_INVALID, "bad autotune arguments");
    REQUIRE(b->have_obs && b->have_state, MCV_ERR_STATE, "observations or state not set");
    c->set_device();
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    std::string rep;
    char line[160];
    int best = 0;
    float bt = 1e30f;
    for (int v = 0; v < kBaVariants; v++) {
        std::vector<float> ts;
        for (int k = 0; k < reps + 1; k++) {
            HIP_OK(hipEventRecord(e0, s));
            ba_enqueue_linearize(b, v, 0, s);
            HIP_OK(hipEventRecord(e1, s));
            HIP_OK(hipEventSynchronize(e1));
            float t = 0;
            HIP_OK(hipEventElapsedTime(&t, e0, e1));
            if (k > 0) ts.push_back(t);  // the first launch warms caches
        }
        std::sort(ts.begin(), ts.end());
        const float t = ts[ts.size() / 2];
        snprintf(line, sizeof line, "k_ba_linearize %s %.4f ms\n", kBaVariant[v].name, t);
        rep += line;
        if (t < bt) { bt = t; best = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->variant = best;
    snprintf(line, sizeof line, "chosen: k_ba_linearize %s\n", kBaVariant[best].name);
    rep += line;
    b->touched(kBaFull);
    if (report && report_cap > 0) {
        strncpy(report, rep.c_str(), (size_t)report_cap - 1);
        report[report_cap - 1] = 0;
    }
    API_END
}

int mcv_ba_set_stereo(mcv_ba* b, int Ms, const float* obs, const int32_t* pt, const int32_t* fr, const float* isig,
                      const double* cam) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && Ms >= 0 && cam && (Ms == 0 || (obs && pt && fr && isig)), MCV_ERR_INVALID, "bad stereo edges");
    c->set_device();
    hipStream_t s = c->stream;
    for (int i = 0; i < Ms; i++)
        REQUIRE(pt[i] >= 0 && pt[i] < b->P && fr[i] >= 0 && fr[i] < b->F, MCV_ERR_INVALID,
                "stereo edge index out of range");
    b->Ms = Ms;
    for (int k = 0; k < 6; k++) b->h_scam[k] = cam[k];
    b->scam = StereoCam{(float)cam[0], (float)cam[1], (float)cam[2], (float)cam[3], (float)cam[4],
                        b->robust ? (float)cam[5] : INFINITY};
    b->h_spt.assign(pt, pt + Ms);
    b->h_sfr.assign(fr, fr + Ms);
    b->active_s.clear();
    b->schur_ready = false;
    const size_t m = (size_t)std::max(Ms, 1);
    b->s_obs.ensure(m * 12);
    b->s_pt.ensure(m * 4);
    b->s_fr.ensure(m * 4);
    b->s_isig.ensure(m * 4);
    b->s_planes.ensure(m * 32 * 4);
    if (Ms) {
        HIP_OK(hipMemcpyAsync(b->s_obs.p, obs, (size_t)Ms * 12, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->s_pt.p, pt, (size_t)Ms * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->s_fr.p, fr, (size_t)Ms * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(b->s_isig.p, isig, (size_t)Ms * 4, hipMemcpyHostToDevice, s));
    }
    HIP_OK(hipStreamSynchronize(s));
    b->have_stereo = true;
    b->stereo_linearized = b->stereo_pose_linearized = b->stereo_full = false;
    API_END
}

static void ba_enqueue_stereo(mcv_ba* b, int mode, hipStream_t s) {
    if (b->Ms > 0) {
        BAStereoObs ob{b->s_obs.as<float>(), b->s_pt.as<uint32_t>(), b->s_fr.as<uint32_t>(), b->s_isig.as<float>(),
                       b->Ms};
        BAOut o{b->s_planes.as<float>(), (long long)b->Ms};
        hipLaunchKernelGGL(k_ba_stereo_linearize, dim3((b->Ms + 255) / 256), dim3(256), 0, s, ob, b->phi.as<float4>(),
                           b->plo.as<float4>(), b->fposes.as<FramePose>(), b->scam, o, mode);
        HIP_OK(hipGetLastError());
    }
    b->stereo_linearized = true;
    b->stereo_pose_linearized = mode == kBaFull || mode == kBaPoseOnly;
    b->stereo_full = mode == kBaFull;
    b->schur_done = false;
}

int mcv_ba_linearize_stereo(mcv_ba* b, int residual_only) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->have_stereo && b->have_state, MCV_ERR_STATE, "stereo edges or state not set");
    REQUIRE(residual_only >= kBaFull && residual_only <= kBaPoseOnly, MCV_ERR_INVALID, "unknown linearization mode");
    c->set_device();
    ba_enqueue_stereo(b, residual_only, c->stream);
    API_END
}

int mcv_ba_read_stereo_planes(mcv_ba* b, int first, int count, float* host) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->stereo_linearized, MCV_ERR_STATE, "no stereo linearization yet");
    REQUIRE(host && first >= 0 && count >= 0 && first + count <= 32, MCV_ERR_INVALID, "plane range");
    c->set_device();
    if (b->Ms > 0 && count > 0)
        HIP_OK(hipMemcpyAsync(host, b->s_planes.as<float>() + (size_t)first * b->Ms, (size_t)count * b->Ms * 4,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    API_END
}

int mcv_ba_outliers(mcv_ba* b, float thr_mono, float thr_stereo, uint8_t* flags_mono, uint8_t* flags_stereo) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b, MCV_ERR_INVALID, "null BA object");
    REQUIRE(!flags_mono || b->have_obs, MCV_ERR_STATE, "no mono edges");
    REQUIRE(!flags_stereo || b->stereo_linearized, MCV_ERR_STATE, "no stereo linearization yet");
    c->set_device();
    hipStream_t s = c->stream;
    b->flags.ensure((size_t)std::max(std::max(b->M, b->Ms), 1));
    if (flags_mono && b->M > 0) {
        hipLaunchKernelGGL(k_ba_outliers, dim3((b->M + 255) / 256), dim3(256), 0, s,
                           b->planes.as<float>() + 2 * (size_t)b->M, b->M, thr_mono, b->flags.as<uint8_t>());
        HIP_OK(hipMemcpyAsync(flags_mono, b->flags.p, (size_t)b->M, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    if (flags_stereo && b->Ms > 0) {
        hipLaunchKernelGGL(k_ba_outliers, dim3((b->Ms + 255) / 256), dim3(256), 0, s,
                           b->s_planes.as<float>() + 3 * (size_t)b->Ms, b->Ms, thr_stereo, b->flags.as<uint8_t>());
        HIP_OK(hipMemcpyAsync(flags_stereo, b->flags.p, (size_t)b->Ms, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    HIP_OK(hipGetLastError());
    API_END
}

int mcv_ba_linearize_plain(mcv_ba* b, int residual_only) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->have_obs && b->have_state, MCV_ERR_STATE, "observations or state not set");
    c->set_device();
    BAObs ob{b->uv.as<float2>(), b->pt.as<uint32_t>(), b->pair.as<uint32_t>(), b->isig.as<float>(), b->M};
    BAOut o{b->planes.as<float>(), (long long)b->M};
    for (size_t g = 0; g + 1 < b->h_gptr.size(); g++) {
        const int n = b->h_gptr[g + 1] - b->h_gptr[g];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_ba_linearize_group, dim3((n + 255) / 256), dim3(256), 0, c->stream, ob,
                           b->phi.as<float4>(), b->plo.as<float4>(), b->pairs.as<BAPair>(),
                           b->gobs.as<int>() + b->h_gptr[g], n, o, residual_only);
    }
    HIP_OK(hipGetLastError());
    b->touched(residual_only);
    API_END
}

int mcv_ba_generation(mcv_ba* b, int64_t* gen) {
    if (!b || !gen) return MCV_ERR_INVALID;
    *gen = b->gen;
    return MCV_OK;
}

// Host side of the Schur structure: (point, free frame) groups of
// observations, free-frame observation lists, and for every free-frame pair
// block (i <= j) the list of (pf_i, pf_j) products, all in point order.
// Edge lists of the Schur complement over the active mono and stereo edges
// (index i < M mono, M + j stereo): per free point its edges, per (free
// point, free frame) pair its edges, per free frame its edges (points fixed
// or not) and its pairs, and the pair lists of every frame-pair block.
static void ba_build_schur(mcv_ba* b, hipStream_t s) {
    const int P = b->P, F = b->F, M = b->M;
    const int Ms = b->have_stereo ? b->Ms : 0;
    if (b->fixed.empty()) b->fixed.assign(F, 0);
    std::vector<int> fidx(F, -1);
    int Fp = 0;
    for (int f = 0; f < F; f++)
        if (!b->fixed[f]) fidx[f] = Fp++;
    auto act_m = [&](int i) { return b->active_m.empty() || b->active_m[i]; };
    auto act_s = [&](int j) { return b->active_s.empty() || b->active_s[j]; };
    auto pfree = [&](int p) { return b->point_fixed.empty() || !b->point_fixed[p]; };
    // per point: active edges (mono in index order, then stereo)
    std::vector<int> pe_ptr(P + 1, 0);
    for (int i = 0; i < M; i++)
        if (act_m(i) && pfree(b->h_pt[i])) pe_ptr[b->h_pt[i] + 1]++;
    for (int j = 0; j < Ms; j++)
        if (act_s(j) && pfree(b->h_spt[j])) pe_ptr[b->h_spt[j] + 1]++;
    for (int p = 0; p < P; p++) pe_ptr[p + 1] += pe_ptr[p];
    std::vector<int> pe(std::max(pe_ptr[P], 1));
    {
        std::vector<int> a(pe_ptr.begin(), pe_ptr.end() - 1);
        for (int i = 0; i < M; i++)
            if (act_m(i) && pfree(b->h_pt[i])) pe[a[b->h_pt[i]]++] = i;
        for (int j = 0; j < Ms; j++)
            if (act_s(j) && pfree(b->h_spt[j])) pe[a[b->h_spt[j]]++] = M + j;
    }
    auto frame_of = [&](int e) { return e < M ? b->h_fr[e] : b->h_sfr[e - M]; };
    std::vector<int> pf_ptr(P + 1, 0), pf_obs_ptr(1, 0), pf_obs, pf_point, pf_free;
    std::vector<std::pair<int, int>> tmp;
    bool any_free_edges = false;
    for (int p = 0; p < P; p++) {
        tmp.clear();
        if (pe_ptr[p + 1] > pe_ptr[p]) any_free_edges = true;
        for (int t = pe_ptr[p]; t < pe_ptr[p + 1]; t++)
            if (fidx[frame_of(pe[t])] >= 0) tmp.push_back({fidx[frame_of(pe[t])], pe[t]});
        std::stable_sort(tmp.begin(), tmp.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
            return x.first < y.first;
        });
        for (size_t k = 0; k < tmp.size(); k++) {
            if (k == 0 || tmp[k].first != tmp[k - 1].first) {
                if (k) pf_obs_ptr.push_back((int)pf_obs.size());
                pf_point.push_back(p);
                pf_free.push_back(tmp[k].first);
            }
            pf_obs.push_back(tmp[k].second);
        }
        if (!tmp.empty()) pf_obs_ptr.push_back((int)pf_obs.size());
        pf_ptr[p + 1] = (int)pf_point.size();
    }
    const int npf = (int)pf_point.size();
    // free frame -> active edges, free frame -> pf entries
    std::vector<int> sfptr(Fp + 1, 0), fpf_ptr(Fp + 1, 0);
    for (int i = 0; i < M; i++)
        if (act_m(i) && fidx[b->h_fr[i]] >= 0) sfptr[fidx[b->h_fr[i]] + 1]++;
    for (int j = 0; j < Ms; j++)
        if (act_s(j) && fidx[b->h_sfr[j]] >= 0) sfptr[fidx[b->h_sfr[j]] + 1]++;
    for (int q = 0; q < npf; q++) fpf_ptr[pf_free[q] + 1]++;
    for (int f = 0; f < Fp; f++) {
        sfptr[f + 1] += sfptr[f];
        fpf_ptr[f + 1] += fpf_ptr[f];
    }
    std::vector<int> sfobs(std::max(sfptr[Fp], 1)), fpf(std::max(npf, 1));
    {
        std::vector<int> a(sfptr.begin(), sfptr.end() - 1), c(fpf_ptr.begin(), fpf_ptr.end() - 1);
        for (int i = 0; i < M; i++)
            if (act_m(i) && fidx[b->h_fr[i]] >= 0) sfobs[a[fidx[b->h_fr[i]]]++] = i;
        for (int j = 0; j < Ms; j++)
            if (act_s(j) && fidx[b->h_sfr[j]] >= 0) sfobs[a[fidx[b->h_sfr[j]]]++] = M + j;
        for (int q = 0; q < npf; q++) fpf[c[pf_free[q]]++] = q;
    }
    // pair blocks, upper triangle row-major
    const int nblk = Fp * (Fp + 1) / 2;
    auto bidx = [Fp](int i, int j) { return i * Fp - i * (i - 1) / 2 + (j - i); };
    std::vector<int> blk_ptr(nblk + 1, 0);
    for (int p = 0; p < P; p++)
        for (int qa = pf_ptr[p]; qa < pf_ptr[p + 1]; qa++)
            for (int qb = qa; qb < pf_ptr[p + 1]; qb++) blk_ptr[bidx(pf_free[qa], pf_free[qb]) + 1]++;
    for (int k = 0; k < nblk; k++) blk_ptr[k + 1] += blk_ptr[k];
    std::vector<int2> items(std::max(blk_ptr[nblk], 1)), ij(std::max(nblk, 1));
    {
        std::vector<int> a(blk_ptr.begin(), blk_ptr.end() - 1);
        for (int p = 0; p < P; p++)
            for (int qa = pf_ptr[p]; qa < pf_ptr[p + 1]; qa++)
                for (int qb = qa; qb < pf_ptr[p + 1]; qb++)
                    items[a[bidx(pf_free[qa], pf_free[qb])]++] = make_int2(qa, qb);
        for (int i = 0; i < Fp; i++)
            for (int j = i; j < Fp; j++) ij[bidx(i, j)] = make_int2(i, j);
    }
    upload(b->pe_ptr, pe_ptr, s);
    upload(b->pe, pe, s);
    upload(b->pf_ptr, pf_ptr, s);
    upload(b->pf_obs_ptr, pf_obs_ptr, s);
    upload(b->pf_obs, pf_obs, s);
    upload(b->pf_point, pf_point, s);
    upload(b->pf_frame, pf_free, s);
    upload(b->sfptr, sfptr, s);
    upload(b->sfobs, sfobs, s);
    upload(b->fpf_ptr, fpf_ptr, s);
    upload(b->fpf, fpf, s);
    upload(b->blk_ptr, blk_ptr, s);
    upload(b->blk_items, items, s);
    upload(b->blk_ij, ij, s);
    b->Dinv.ensure((size_t)P * 36);
    b->bl.ensure((size_t)P * 12);
    b->db.ensure((size_t)P * 12);
    b->Hll.ensure((size_t)P * 36);
    b->Hpl.ensure((size_t)std::max(npf, 1) * 72);
    b->Hpp.ensure((size_t)std::max(Fp, 1) * 144);
    b->bs.ensure((size_t)std::max(Fp, 1) * 24);
    b->bp.ensure((size_t)std::max(Fp, 1) * 24);
    b->Sout.ensure((size_t)std::max(36 * Fp * Fp, 1) * 4);
    HIP_OK(hipStreamSynchronize(s));
    b->npf = npf;
    b->Fp = Fp;
    b->nblk = nblk;
    b->free_points_with_edges = any_free_edges;
    b->schur_ready = true;
}

static SchurArgs ba_schur_args(mcv_ba* b, float lambda) {
    SchurArgs a{};
    a.planes = b->planes.as<float>();
    a.S = b->M;
    a.isig = b->isig.as<float>();
    a.M = b->M;
    a.splanes = b->s_planes.as<float>();
    a.SS = b->Ms;
    a.sisig = b->s_isig.as<float>();
    a.pe_ptr = b->pe_ptr.as<int>();
    a.pe = b->pe.as<int>();
    a.pf_ptr = b->pf_ptr.as<int>();
    a.pf_obs_ptr = b->pf_obs_ptr.as<int>();
    a.pf_obs = b->pf_obs.as<int>();
    a.pf_point = b->pf_point.as<int>();
    a.pf_frame = b->pf_frame.as<int>();
    a.fptr = b->sfptr.as<int>();
    a.fobs = b->sfobs.as<int>();
    a.fpf_ptr = b->fpf_ptr.as<int>();
    a.fpf = b->fpf.as<int>();
    a.blk_ptr = b->blk_ptr.as<int>();
    a.blk_items = b->blk_items.as<int2>();
    a.blk_ij = b->blk_ij.as<int2>();
    a.P = b->P;
    a.Fp = b->Fp;
    a.nblk = b->nblk;
    a.lambda = lambda;
    a.Dinv = b->Dinv.as<float>();
    a.bl = b->bl.as<float>();
    a.db = b->db.as<float>();
    a.Hll = b->Hll.as<float>();
    a.Hpl = b->Hpl.as<float>();
    a.Hpp = b->Hpp.as<float>();
    a.bs = b->bs.as<float>();
    a.bp = b->bp.as<float>();
    a.Sout = b->Sout.as<float>();
    return a;
}

// the Schur launches (structure ready); with blocks = false only H_ll, H_pl,
// H_pp, b_p (what the largest-diagonal search needs)
static void ba_enqueue_schur(mcv_ba* b, float lambda, hipStream_t s, bool blocks = true) {
    const SchurArgs a = ba_schur_args(b, lambda);
    hipLaunchKernelGGL(k_schur_points, dim3((b->P + 255) / 256), dim3(256), 0, s, a);
    if (b->Fp > 0) {
        hipLaunchKernelGGL(k_schur_frames, dim3(b->Fp), dim3(256), 0, s, a);
        if (blocks) hipLaunchKernelGGL(k_schur_blocks, dim3(b->nblk), dim3(256), 0, s, a);
    }
    HIP_OK(hipGetLastError());
}

static void ba_require_schur_inputs(mcv_ba* b) {
    REQUIRE(b->pose_linearized, MCV_ERR_STATE, "linearize before the Schur complement");
    REQUIRE(b->linearized || !b->free_points_with_edges, MCV_ERR_STATE,
            "a pose-only linearization leaves no J_point for the free points");
    const bool stereo_edges = b->have_stereo && b->Ms > 0;
    REQUIRE(!stereo_edges || b->stereo_pose_linearized, MCV_ERR_STATE,
            "linearize the stereo edges before the Schur complement");
    REQUIRE(!stereo_edges || !b->free_points_with_edges || b->stereo_full, MCV_ERR_STATE,
            "a pose-only stereo linearization leaves no J_point for the free points");
}

int mcv_ba_set_fixed(mcv_ba* b, const int32_t* frame_fixed) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && frame_fixed, MCV_ERR_INVALID, "null fixed flags");
    b->fixed.assign(frame_fixed, frame_fixed + b->F);
    b->schur_ready = false;
    API_END
}

int mcv_ba_set_point_fixed(mcv_ba* b, const int32_t* point_fixed) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b, MCV_ERR_INVALID, "null BA object");
    c->set_device();
    if (point_fixed) b->point_fixed.assign(point_fixed, point_fixed + b->P);
    else b->point_fixed.clear();
    std::vector<uint8_t> f(b->P, 0);
    for (int p = 0; p < b->P && !b->point_fixed.empty(); p++) f[p] = b->point_fixed[p] ? 1 : 0;
    upload(b->pfix_d, f, c->stream);
    HIP_OK(hipStreamSynchronize(c->stream));
    b->schur_ready = false;
    API_END
}

int mcv_ba_set_active(mcv_ba* b, const uint8_t* active_mono, const uint8_t* active_stereo) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b, MCV_ERR_INVALID, "null BA object");
    c->set_device();
    if (active_mono) b->active_m.assign(active_mono, active_mono + b->M);
    else b->active_m.clear();
    if (active_stereo) b->active_s.assign(active_stereo, active_stereo + b->Ms);
    else b->active_s.clear();
    if (!b->active_m.empty()) upload(b->act_m, b->active_m, c->stream);
    if (!b->active_s.empty()) upload(b->act_s, b->active_s, c->stream);
    HIP_OK(hipStreamSynchronize(c->stream));
    b->schur_ready = false;
    API_END
}

int mcv_ba_set_robust(mcv_ba* b, int enabled) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->have_state, MCV_ERR_STATE, "set the state first");
    c->set_device();
    b->robust = enabled != 0;
    ba_upload_frames(b, c->stream);
    HIP_OK(hipStreamSynchronize(c->stream));
    b->touched(-1);
    API_END
}

int mcv_ba_schur(mcv_ba* b, float lambda) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b, MCV_ERR_INVALID, "null BA object");
    c->set_device();
    hipStream_t s = c->stream;
    if (!b->schur_ready) ba_build_schur(b, s);
    ba_require_schur_inputs(b);
    ba_enqueue_schur(b, lambda, s);
    b->schur_done = true;
    API_END
}

int mcv_ba_read_schur(mcv_ba* b, double* S, double* bs, int32_t* n) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->schur_done, MCV_ERR_STATE, "no Schur complement computed");
    c->set_device();
    const int m = 6 * b->Fp;
    if (n) *n = m;
    std::vector<float> hs((size_t)m * m), hb(m);
    if (m) {
        HIP_OK(hipMemcpyAsync(hs.data(), b->Sout.p, hs.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(hb.data(), b->bs.p, hb.size() * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    if (S) for (size_t k = 0; k < hs.size(); k++) S[k] = hs[k];
    if (bs) for (int k = 0; k < m; k++) bs[k] = hb[k];
    API_END
}

// FP64 dense Cholesky solve of the reduced camera system on the host (the
// part the reference leaves to g2o's LinearSolverEigen,
// solvers/eigen/linear_solver_eigen.h:55).  Returns MCV_ERR_STATE when S is
// not positive definite.
int mcv_solve_llt(int n, const double* S, const double* b, double* x) {
    if (n < 0 || (n && (!S || !b || !x))) return MCV_ERR_INVALID;
    std::vector<double> L((size_t)n * n, 0.0);
    for (int j = 0; j < n; j++) {
        double d = S[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(d > 0)) return MCV_ERR_STATE;
        const double ljj = std::sqrt(d);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            double v = S[(size_t)i * n + j];
            for (int k = 0; k < j; k++) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = v / ljj;
        }
    }
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
        double v = b[i];
        for (int k = 0; k < i; k++) v -= L[(size_t)i * n + k] * y[k];
        y[i] = v / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double v = y[i];
        for (int k = i + 1; k < n; k++) v -= L[(size_t)k * n + i] * x[k];
        x[i] = v / L[(size_t)i * n + i];
    }
    return MCV_OK;
}

}  // extern "C"

// ---- Levenberg-Marquardt over the GPU data ------------------------------------
// Pose increment T <- exp(x) * T of VertexSE3Expmap::oplusImpl
// (types/sba/types_six_dof_expmap.h:98-101): SE3Quat::exp
// (types/slam3d/se3quat.h:218-257) with Eigen's matrix -> quaternion
// conversion, then the product with a normalised (w >= 0) quaternion.
struct Qd { double x, y, z, w; };

static Qd qd_mul(const Qd& a, const Qd& b) {
    return Qd{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
              a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

static void qd_rotate(const Qd& q, const double v[3], double o[3]) {
    double uv[3] = {q.y * v[2] - q.z * v[1], q.z * v[0] - q.x * v[2], q.x * v[1] - q.y * v[0]};
    for (double& c : uv) c += c;
    o[0] = v[0] + q.w * uv[0] + (q.y * uv[2] - q.z * uv[1]);
    o[1] = v[1] + q.w * uv[1] + (q.z * uv[0] - q.x * uv[2]);
    o[2] = v[2] + q.w * uv[2] + (q.x * uv[1] - q.y * uv[0]);
}

static void qd_normalize(Qd& q) {
    if (q.w < 0) q = Qd{-q.x, -q.y, -q.z, -q.w};
    const double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    q = Qd{q.x / n, q.y / n, q.z / n, q.w / n};
}

static Qd qd_from_R(const double* m) {
    Qd q{};
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
        if (m[8] > m[4 * i]) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double c[3];
        t = std::sqrt(m[4 * i] - m[4 * j] - m[4 * k] + 1.0);
        c[i] = 0.5 * t;
        t = 0.5 / t;
        q.w = (m[3 * k + j] - m[3 * j + k]) * t;
        c[j] = (m[3 * j + i] + m[3 * i + j]) * t;
        c[k] = (m[3 * k + i] + m[3 * i + k]) * t;
        q.x = c[0]; q.y = c[1]; q.z = c[2];
    }
    return q;
}

static void se3_oplus(double* frame, const double* xi) {
    const double w[3] = {xi[0], xi[1], xi[2]};
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) W2[3 * r + c] = W[3 * r] * W[c] + W[3 * r + 1] * W[3 + c] + W[3 * r + 2] * W[6 + c];
    double R[9], V[9];
    for (int k = 0; k < 9; k++) {
        const double I = (k % 4 == 0) ? 1.0 : 0.0;
        if (th < 0.00001) {
            R[k] = I + W[k] + 0.5 * W2[k];
            V[k] = I + 0.5 * W[k] + (1.0 / 6.0) * W2[k];
        } else {
            R[k] = I + std::sin(th) / th * W[k] + (1 - std::cos(th)) / (th * th) * W2[k];
            V[k] = I + (1 - std::cos(th)) / (th * th) * W[k] + (th - std::sin(th)) / std::pow(th, 3) * W2[k];
        }
    }
    Qd dq = qd_from_R(R);
    qd_normalize(dq);
    double dt[3];
    for (int r = 0; r < 3; r++) dt[r] = V[3 * r] * xi[3] + V[3 * r + 1] * xi[4] + V[3 * r + 2] * xi[5];
    // exp(x) * T
    const Qd q{frame[0], frame[1], frame[2], frame[3]};
    double rt[3];
    qd_rotate(dq, frame + 4, rt);
    Qd nq = qd_mul(dq, q);
    qd_normalize(nq);
    frame[0] = nq.x; frame[1] = nq.y; frame[2] = nq.z; frame[3] = nq.w;
    for (int r = 0; r < 3; r++) frame[4 + r] = dt[r] + rt[r];
}

// both edge sets linearized at the current state (mode as k_ba_linearize)
static void ba_lin_all(mcv_ba* b, int mode, hipStream_t s) {
    ba_enqueue_linearize(b, b->variant, mode, s);
    b->touched(mode);
    if (b->have_stereo) ba_enqueue_stereo(b, mode, s);
}

// activeRobustChi2 of the last linearization (sum over active edges of the
// Huber rho, FP64 partial sums in a fixed order); enqueued, result in dsum[slot]
static void ba_enqueue_rho(mcv_ba* b, hipStream_t s, int slot) {
    const int Ms = b->have_stereo ? b->Ms : 0;
    const int n = b->M + Ms;
    const int blocks = std::max((n + 255) / 256, 1);
    b->partial.ensure((size_t)std::max(blocks, (b->P + 255) / 256 + 1) * 8);
    b->dsum.ensure(64);
    hipLaunchKernelGGL(k_ba_rho_sum, dim3(blocks), dim3(256), 0, s, b->planes.as<float>() + 2 * (size_t)b->M,
                       b->pair.as<uint32_t>(), b->pairs.as<BAPair>(),
                       b->active_m.empty() ? nullptr : b->act_m.as<uint8_t>(), b->M,
                       b->s_planes.as<float>() + 3 * (size_t)b->Ms,
                       b->active_s.empty() ? nullptr : b->act_s.as<uint8_t>(), Ms, b->scam.delta,
                       b->partial.as<double>());
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, b->partial.as<double>(), blocks,
                       b->dsum.as<double>() + slot);
    HIP_OK(hipGetLastError());
}

static double ba_read_double(mcv_ba* b, hipStream_t s, int slot) {
    double v = 0;
    HIP_OK(hipMemcpyAsync(&v, b->dsum.as<double>() + slot, 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return v;
}

extern "C" {

int mcv_ba_robust_chi2(mcv_ba* b, double* chi2) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && chi2, MCV_ERR_INVALID, "null argument");
    REQUIRE(b->gen > 0 && b->have_state && b->have_obs, MCV_ERR_STATE, "no linearization");
    REQUIRE(!(b->have_stereo && b->Ms > 0) || b->stereo_linearized, MCV_ERR_STATE, "stereo edges not linearized");
    c->set_device();
    ba_enqueue_rho(b, c->stream, 0);
    *chi2 = ba_read_double(b, c->stream, 0);
    API_END
}

int mcv_ba_optimize(mcv_ba* b, int iterations, double tau, int max_trials, double* stats, int32_t* done) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->have_obs && b->have_state, MCV_ERR_STATE, "observations or state not set");
    REQUIRE(iterations >= 0 && max_trials >= 1 && tau >= 0, MCV_ERR_INVALID, "bad optimizer settings");
    c->set_device();
    hipStream_t s = c->stream;
    if (!b->schur_ready) ba_build_schur(b, s);
    const int Fp = b->Fp, n = 6 * Fp, P = b->P;
    // free points without edges need no J_point: a pose-only linearization
    const int full_mode = b->free_points_with_edges ? kBaFull : kBaPoseOnly;
    std::vector<int> fidx(b->F, -1);
    for (int f = 0, k = 0; f < b->F; f++)
        if (!b->fixed[f]) fidx[f] = k++;
    b->p64_bak.ensure((size_t)P * 24);
    b->xl.ensure((size_t)std::max(P, 1) * 12);
    b->xp.ensure((size_t)std::max(n, 1) * 4);
    b->fmax.ensure(16);
    const uint8_t* pfix = b->point_fixed.empty() ? nullptr : b->pfix_d.as<uint8_t>();
    const int pblocks = (P + 255) / 256;
    std::vector<float> Sf((size_t)n * n), bsf(n), bpf(n), xpf(n);
    std::vector<double> S((size_t)n * n), bs(n), bp(n), xp(n);
    double lambda = 0, ni = 2;
    int it_done = 0;
    for (int it = 0; it < iterations; it++) {
        // computeActiveErrors + activeRobustChi2 + buildSystem's linearization
        ba_lin_all(b, full_mode, s);
        ba_enqueue_rho(b, s, 0);
        double currentChi = ba_read_double(b, s, 0);
        if (stats) stats[4 * it] = currentChi;
        if (it == 0) {
            // computeLambdaInit: tau * the largest diagonal entry of H
            ba_enqueue_schur(b, 1.f, s, false);
            const int mblocks = std::max(1, std::min(1024, (3 * std::max(P, 2 * Fp) + 255) / 256));
            b->partial.ensure((size_t)mblocks * 8);
            hipLaunchKernelGGL(k_ba_max_diag, dim3(mblocks), dim3(256), 0, s, b->Hpp.as<float>(), Fp,
                               b->Hll.as<float>(), b->pe_ptr.as<int>(), P, reinterpret_cast<float*>(b->partial.p));
            hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, s, reinterpret_cast<const float*>(b->partial.p),
                               mblocks, b->fmax.as<float>());
            HIP_OK(hipGetLastError());
            float mx = 0;
            HIP_OK(hipMemcpyAsync(&mx, b->fmax.p, 4, hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            lambda = tau * (double)mx;
            ni = 2;
        }
        double rho = 0;
        int qmax = 0;
        do {
            // push
            const std::vector<double> frames0 = b->h_frames;
            HIP_OK(hipMemcpyAsync(b->p64_bak.p, b->p64.p, (size_t)P * 24, hipMemcpyDeviceToDevice, s));
            ba_enqueue_schur(b, (float)lambda, s);
            if (n) {
                HIP_OK(hipMemcpyAsync(Sf.data(), b->Sout.p, Sf.size() * 4, hipMemcpyDeviceToHost, s));
                HIP_OK(hipMemcpyAsync(bsf.data(), b->bs.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
                HIP_OK(hipMemcpyAsync(bpf.data(), b->bp.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
            }
            HIP_OK(hipStreamSynchronize(s));
            for (size_t k = 0; k < Sf.size(); k++) S[k] = Sf[k];
            for (int k = 0; k < n; k++) { bs[k] = bsf[k]; bp[k] = bpf[k]; }
            // FP64 Cholesky of the reduced camera system
            const bool ok2 = n == 0 || mcv_solve_llt(n, S.data(), bs.data(), xp.data()) == MCV_OK;
            double scale = 0;
            if (ok2) {
                for (int k = 0; k < n; k++) {
                    xpf[k] = (float)xp[k];
                    scale += xp[k] * (lambda * xp[k] + bp[k]);
                }
                if (n) HIP_OK(hipMemcpyAsync(b->xp.p, xpf.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
                const SchurArgs a = ba_schur_args(b, (float)lambda);
                hipLaunchKernelGGL(k_schur_backsub, dim3(std::max(pblocks, 1)), dim3(256), 0, s, a, b->xp.as<float>(),
                                   b->xl.as<float>(), b->partial.as<double>());
                hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, b->partial.as<double>(), pblocks,
                                   b->dsum.as<double>() + 1);
                // update: points on the device, frames on the host
                hipLaunchKernelGGL(k_ba_points_apply, dim3(std::max(pblocks, 1)), dim3(256), 0, s,
                                   b->p64.as<double>(), b->xl.as<float>(), pfix, P, b->phi.as<float4>(),
                                   b->plo.as<float4>());
                HIP_OK(hipGetLastError());
                for (int f = 0; f < b->F; f++)
                    if (fidx[f] >= 0) se3_oplus(&b->h_frames[7 * (size_t)f], &xp[6 * (size_t)fidx[f]]);
                ba_upload_frames(b, s);
                scale += ba_read_double(b, s, 1);
            }
            // computeActiveErrors at the new state
            ba_lin_all(b, kBaResidual, s);
            ba_enqueue_rho(b, s, 2);
            double tempChi = ba_read_double(b, s, 2);
            if (!ok2) tempChi = DBL_MAX;
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
                // pop
                b->h_frames = frames0;
                ba_upload_frames(b, s);
                HIP_OK(hipMemcpyAsync(b->p64.p, b->p64_bak.p, (size_t)P * 24, hipMemcpyDeviceToDevice, s));
                hipLaunchKernelGGL(k_ba_points_apply, dim3(std::max(pblocks, 1)), dim3(256), 0, s,
                                   b->p64.as<double>(), (const float*)nullptr, (const uint8_t*)nullptr, P,
                                   b->phi.as<float4>(), b->plo.as<float4>());
                HIP_OK(hipGetLastError());
                if (!std::isfinite(lambda)) break;
            }
            qmax++;
        } while (rho < 0 && qmax < max_trials);
        if (stats) {
            stats[4 * it + 1] = currentChi;
            stats[4 * it + 2] = lambda;
            stats[4 * it + 3] = qmax;
        }
        it_done++;
        if (qmax == max_trials || rho == 0 || !std::isfinite(lambda)) break;
    }
    // the residual planes of the final (accepted) state, for the outlier
    // test that follows an optimization (FilterCallBack_Chi2 calls
    // computeError first)
    ba_lin_all(b, kBaResidual, s);
    HIP_OK(hipStreamSynchronize(s));
    if (done) *done = it_done;
    API_END
}

int mcv_ba_get_state(mcv_ba* b, double* points, double* frames) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->have_state, MCV_ERR_STATE, "no state");
    c->set_device();
    if (points) {
        HIP_OK(hipMemcpyAsync(points, b->p64.p, (size_t)b->P * 24, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    if (frames) std::memcpy(frames, b->h_frames.data(), b->h_frames.size() * sizeof(double));
    API_END
}

int mcv_ba_planes(mcv_ba* b, float** d_planes, int64_t* stride) {
    if (!b) return MCV_ERR_INVALID;
    if (d_planes) *d_planes = b->planes.as<float>();
    if (stride) *stride = b->M;
    return MCV_OK;
}

int mcv_ba_read_planes(mcv_ba* b, int first, int count, float* host) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && host && first >= 0 && count >= 0 && first + count <= 22, MCV_ERR_INVALID, "bad plane range");
    c->set_device();
    if (b->M && count)
        HIP_OK(hipMemcpyAsync(host, b->planes.as<float>() + (size_t)first * b->M, (size_t)count * b->M * 4,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    API_END
}

int mcv_ba_copy_planes_dev(mcv_ba* b, int first, int count, float* d_dst) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && d_dst && first >= 0 && count >= 0 && first + count <= 22, MCV_ERR_INVALID, "bad plane range");
    c->set_device();
    if (b->M && count)
        HIP_OK(hipMemcpyAsync(d_dst, b->planes.as<float>() + (size_t)first * b->M, (size_t)count * b->M * 4,
                              hipMemcpyDeviceToDevice, c->stream));
    API_END
}

int mcv_ba_adjoint_dev(mcv_ba* b, const float* d_g, float* d_gpose, float* d_gpoint) {
    mcv_ctx* c = b ? b->ctx : nullptr;
    API_BEGIN(c)
    REQUIRE(b && b->linearized, MCV_ERR_STATE, "no linearization to take the adjoint of");
    REQUIRE(d_g && d_gpose && d_gpoint, MCV_ERR_INVALID, "null gradient buffer");
    c->set_device();
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_ba_point_adj, dim3((b->P + 255) / 256), dim3(256), 0, s, b->planes.as<float>(), (long long)b->M,
                       reinterpret_cast<const float2*>(d_g), b->pptr.as<int>(), b->P, d_gpoint);
    hipLaunchKernelGGL(k_ba_pose_adj, dim3(b->F), dim3(256), 0, s, b->planes.as<float>(), (long long)b->M,
                       reinterpret_cast<const float2*>(d_g), b->fptr.as<int>(), b->fobs.as<int>(), d_gpose);
    HIP_OK(hipGetLastError());
    API_END
}

}  // extern "C"
