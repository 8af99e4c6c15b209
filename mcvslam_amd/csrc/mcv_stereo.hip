// Stereo matching of a rectified left/right pair for gfx950
// (Frame::ComputeStereoMatch, src/Frame.cpp:150-328):
//   k_stereo        one wave per left keypoint: row-band + octave +
//                   disparity gating of the right keypoints in index order,
//                   top-2 Hamming, ratio 0.70 / threshold 34, 11x11 SAD over
//                   11 shifts on the pyramid level, parabola sub-pixel fit
//   k_stereo_median one workgroup per pair: sort the accepted (SAD, index)
//                   pairs, drop those beyond 1.6x / below 0.4x the median
// Included by mcv_api.hip (uses LevelDesc, bitonic_sort, wave_sum).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcv {

struct StereoJob {
    const float* kl;        // left keypoints (mcv_keypoint rows of 6 x 4 B)
    const uint8_t* dl;
    const int* nl;          // device count
    const float* kr;
    const uint8_t* dr;
    const int* nr;
    int lvl_l, lvl_r;       // LevelDesc index of level 0 of the left / right image
    float* u_right;         // per left keypoint, -1 when unmatched
    float* depth;
    int* sad;
    unsigned long long* keys;  // per left keypoint: sad << 32 | index, or ~0
};

struct StereoParams {
    float scale[16], inv_scale[16];
    float bf, b;
    int cap;
};

__global__ __launch_bounds__(256) void k_stereo(const StereoJob* jobs, const LevelDesc* lv, const uint8_t* pyr,
                                                StereoParams sp) {
    const StereoJob J = jobs[blockIdx.y];
    const int iL = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int nl = min(*J.nl, sp.cap), nr = min(*J.nr, sp.cap);
    if (iL >= nl) return;
    const float* kp = J.kl + 6 * iL;
    const float uL = kp[0], vL = kp[1];
    const int levelL = reinterpret_cast<const int*>(kp)[5];
    bool ok = true;
    if (lane == 0) {
        J.u_right[iL] = -1.f;
        J.depth[iL] = -1.f;
        J.sad[iL] = -1;
        J.keys[iL] = ~0ull;
    }
    const int rowL = (int)vL;
    const float minZ = sp.b, minD = 1.f;
    const float maxD = fminf(sp.bf / minZ, 1000.0f);
    const float minU = uL - maxD, maxU = uL - minD;
    if (maxU < 0) return;
    uint32_t q[8];
    const uint32_t* qd = reinterpret_cast<const uint32_t*>(J.dl + 32ll * iL);
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = qd[k];
    uint32_t k0 = 999u << 22, k1 = 999u << 22;
    int ncand = 0;
    for (int iR = lane; iR < nr; iR += 64) {
        const float* kr = J.kr + 6 * iR;
        const float kpY = kr[1];
        const int oR = reinterpret_cast<const int*>(kr)[5];
        const float r = 10.f * sp.scale[oR];
        const int maxr = (int)ceilf(kpY + r), minr = (int)floorf(kpY - r);
        if (rowL < minr || rowL > maxr) continue;
        if (oR < levelL - 1 || oR > levelL + 1) continue;
        const float uR = kr[0];
        if (!(uR >= minU && uR <= maxU)) continue;
        const uint32_t* td = reinterpret_cast<const uint32_t*>(J.dr + 32ll * iR);
        uint32_t d = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) d += __builtin_popcount(q[k] ^ td[k]);
        const uint32_t key = (d << 22) | (uint32_t)iR;
        k1 = min(k1, max(k0, key));
        k0 = min(k0, key);
        ncand++;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t b0 = __shfl_xor(k0, off, 64), b1 = __shfl_xor(k1, off, 64);
        const uint32_t n1 = min(max(k0, b0), min(k1, b1));
        k0 = min(k0, b0);
        k1 = n1;
        ncand += __shfl_xor(ncand, off, 64);
    }
    const float d0 = (float)(k0 >> 22), d1 = (float)(k1 >> 22);
    if (ncand == 0 || !(d0 / d1 <= 0.70f) || d0 > (float)34) ok = false;
    if (!ok) return;
    const int iR = (int)(k0 & ((1u << 22) - 1));
    const float uR0 = J.kr[6 * iR];
    const float scaleFactor = sp.inv_scale[levelL];
    const float scaleduL = roundf(uL * scaleFactor);
    const float scaledvL = roundf(vL * scaleFactor);
    const float scaleduR0 = roundf(uR0 * scaleFactor);
    const int w = 5, L = 5;
    const LevelDesc DL = lv[J.lvl_l + levelL], DR = lv[J.lvl_r + levelL];
    if (scaledvL - w < 0 || scaledvL + w + 1 >= DL.h || scaleduL - w < 0 || scaleduL + w + 1 >= DL.w) return;
    const float iniu = scaleduR0 + L - w, endu = scaleduR0 + L + w + 1;
    if (iniu < 0 || endu >= DR.w) return;
    if (scaleduR0 - L - w < 0) return;  // the reference's colRange would assert
    const uint8_t* IL = pyr + DL.pix;
    const uint8_t* IR = pyr + DR.pix;
    const int y0 = (int)scaledvL - w, xl = (int)scaleduL - w;
    const int cl = IL[(long long)(y0 + w) * DL.pitch + xl + w];
    int bestDist = 0x7FFFFFFF, bestincR = 0;
    float vd[2 * 5 + 1];
#pragma unroll
    for (int incR = -L; incR <= L; incR++) {
        const int xr = (int)scaleduR0 + incR - w;
        const int cr = IR[(long long)(y0 + w) * DR.pitch + xr + w];
        int acc = 0;
        for (int p = lane; p < 121; p += 64) {
            const int yy = p / 11, xx = p - 11 * yy;
            const int a = (int)IL[(long long)(y0 + yy) * DL.pitch + xl + xx] - cl;
            const int b = (int)IR[(long long)(y0 + yy) * DR.pitch + xr + xx] - cr;
            acc += abs(a - b);
        }
        acc = wave_sum(acc);
        const float dist = (float)acc;
        if (dist < (float)bestDist) {
            bestDist = (int)dist;
            bestincR = incR;
        }
        vd[L + incR] = dist;
    }
    if (lane != 0) return;
    if (bestincR == -L || bestincR == L) return;
    float dist1 = 0, dist2 = 0, dist3 = 0;
#pragma unroll
    for (int k = 1; k < 2 * L; k++)
        if (k == L + bestincR) {
            dist1 = vd[k - 1];
            dist2 = vd[k];
            dist3 = vd[k + 1];
        }
    const float deltaR = (dist1 - dist3) / (2.0f * (dist1 + dist3 - 2.0f * dist2));
    if (deltaR < -1 || deltaR > 1) return;
    const float bestuR = sp.scale[levelL] * ((float)scaleduR0 + (float)bestincR + deltaR);
    const float disparity = uL - bestuR;
    if (disparity >= minD && disparity < maxD) {
        J.depth[iL] = sp.bf / disparity;
        J.u_right[iL] = bestuR;
        J.sad[iL] = bestDist;
        J.keys[iL] = ((unsigned long long)(uint32_t)bestDist << 32) | (unsigned)iL;
    }
}

constexpr int kStereoSortCap = 8192;

__global__ __launch_bounds__(256) void k_stereo_median(const StereoJob* jobs, StereoParams sp) {
    __shared__ unsigned long long s[kStereoSortCap];
    __shared__ int s_n;
    const StereoJob J = jobs[blockIdx.x];
    const int nl = min(min(*J.nl, sp.cap), kStereoSortCap);
    int n2 = 1;
    while (n2 < nl) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += 256) s[i] = i < nl ? J.keys[i] : ~0ull;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    bitonic_sort(s, n2);
    // number of accepted entries = first sentinel position
    for (int i = threadIdx.x; i < n2; i += 256)
        if (s[i] != ~0ull && (i + 1 == n2 || s[i + 1] == ~0ull)) s_n = i + 1;
    __syncthreads();
    const int n = s_n;
    if (n == 0) return;
    const float median = (float)(int)(s[n / 2] >> 32);
    const float th_max = 1.6f * median;
    const float th_min = (float)(0.4 * (double)median);
    // r: last kept from the end; l: first kept from the start (l < r)
    __shared__ int s_r, s_l;
    if (threadIdx.x == 0) {
        int r = n - 1;
        while (r >= 0 && (float)(int)(s[r] >> 32) > th_max) r--;
        int l = 0;
        while (l < r && (float)(int)(s[l] >> 32) < th_min) l++;
        s_r = r;
        s_l = l;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) {
        if (i > s_r || i < s_l) {
            const int idx = (int)(s[i] & 0xFFFFFFFFu);
            J.u_right[idx] = -1.f;
            J.depth[idx] = -1.f;
            J.sad[idx] = -1;
        }
    }
}

}  // namespace mcv
