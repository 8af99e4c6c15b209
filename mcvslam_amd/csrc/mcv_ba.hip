// Bundle-adjustment linearization kernels for gfx950.
//
//   k_ba_linearize  EdgeSE3XYZ2Camera::computeError + linearizeOplus
//                   (include/OptimizeTypes.hpp:80-113), chi2 and the Huber
//                   weight (ThirdParty/g2o/g2o/core/base_edge.h:94-97,
//                   core/robust_kernel_impl.cpp:65-78), all edges at once
//   k_ba_linearize_group  the same for one (frame, camera) edge group
//   k_ba_stereo_linearize EdgeStereoSE3ProjectXYZ (3-row stereo edge)
//   k_ba_outliers   the chi2 outlier test of FilterCallBack_Chi2
//   k_ba_point_adj  sum_obs J_point^T g per point (CSR by point)
//   k_ba_pose_adj   sum_obs J_pose^T g per frame (observation list by frame)
//
// Arithmetic is FP32.  The residual e = (u - cx) - fx * X / Z subtracts two
// numbers of a few hundred pixels to get about one pixel, so the camera-frame
// point and the projection are carried as unevaluated float pairs
// (hi + lo, error-free transforms with fmaf): the point and pose state come
// in as such pairs too, which keeps the FP64 state of the optimizer.  The
// Jacobians, chi2 and weight are plain FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcv {

struct f2 {
    float hi, lo;
};

__device__ __forceinline__ f2 two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ f2 fast_two_sum(float a, float b) {
    const float s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ f2 two_prod(float a, float b) {
    const float p = a * b;
    return {p, __builtin_fmaf(a, b, -p)};
}
__device__ __forceinline__ f2 df_add(f2 a, f2 b) {
    f2 s = two_sum(a.hi, b.hi);
    const f2 t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ f2 df_neg(f2 a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ f2 df_mul(f2 a, f2 b) {
    f2 p = two_prod(a.hi, b.hi);
    p.lo = __builtin_fmaf(a.hi, b.lo, p.lo);
    p.lo = __builtin_fmaf(a.lo, b.hi, p.lo);
    return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ f2 df_mul_f(f2 a, float b) {
    f2 p = two_prod(a.hi, b);
    p.lo = __builtin_fmaf(a.lo, b, p.lo);
    return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ f2 df_div(f2 a, f2 b) {
    const float q1 = a.hi / b.hi;
    f2 r = df_add(a, df_neg(df_mul_f(b, q1)));
    const float q2 = r.hi / b.hi;
    r = df_add(r, df_neg(df_mul_f(b, q2)));
    const float q3 = r.hi / b.hi;
    const f2 q = fast_two_sum(q1, q2);
    return df_add(q, f2{q3, 0.f});
}

// Per (frame, camera) pair.  Tcw = Tcb * Tbw as float pairs for the
// residual; Tbw and Rcb in float for the Jacobians.
struct BAPair {
    float Rh[9], Rl[9];   // R_cw hi / lo, row-major
    float th[3], tl[3];   // t_cw hi / lo
    float Rbw[9], tbw[3]; // body from world
    float Rcb[9];         // camera from body
    float fx, fy, cx, cy, delta;
    float pad[3];
};

struct BAObs {
    const float2* uv;       // M
    const uint32_t* pt;     // M point index
    const uint32_t* pair;   // M (frame, camera) pair index
    const float* isig;      // M inverse level sigma^2
    int M;
};

struct BAOut {
    float* planes;   // 22 planes of M floats: e_u e_v chi2 w Jpose[12] Jpoint[6]
    long long stride;
};

// a / b as a float pair, from inv ~ 1 / b.hi (v_rcp_f32, 1 ulp): three
// quotient digits, each remainder taken exactly, so the result carries the
// accuracy of the pair arithmetic and no division is correctly rounded
__device__ __forceinline__ f2 df_div_rcp(f2 a, f2 b, float inv) {
    const float q1 = a.hi * inv;
    f2 r = df_add(a, df_neg(df_mul_f(b, q1)));
    const float q2 = r.hi * inv;
    r = df_add(r, df_neg(df_mul_f(b, q2)));
    const float q3 = r.hi * inv;
    return df_add(fast_two_sum(q1, q2), f2{q3, 0.f});
}

constexpr int kPairWords = (int)(sizeof(BAPair) / 4);
constexpr int kLdsPairs = 192;  // (frame, camera) pairs that fit the LDS copy

// Inputs of one observation: the streamed record and its gathered point.
struct BAIn {
    float2 uv;
    uint32_t pt, pair;
    float isig;
    float4 ph, pl;
};

__device__ __forceinline__ void ba_load_obs(int i, const BAObs& ob, BAIn& in) {
    in.uv = ob.uv[i];
    in.pt = ob.pt[i];
    in.pair = ob.pair[i];
    in.isig = ob.isig[i];
}

__device__ __forceinline__ void ba_load_point(const float4* __restrict__ pts_hi, const float4* __restrict__ pts_lo,
                                              BAIn& in) {
    in.ph = pts_hi[in.pt];
    in.pl = pts_lo[in.pt];
}

// kNt: streaming (non-temporal) stores for the output planes; this kernel
// never reads them back
// Plane k of observation i: the plane base is uniform (scalar registers) and
// the byte offset 4 i a 32-bit lane value, so each store takes one address
// register instead of a 64-bit pointer per plane.
template <bool kNt>
__device__ __forceinline__ void ba_store(const BAOut& out, int k, uint32_t off, float v) {
    float* p = reinterpret_cast<float*>(reinterpret_cast<char*>(out.planes + (long long)k * out.stride) + off);
    if (kNt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Linearization modes: every plane; e, chi2 and w only (the LM trials'
// computeActiveErrors); everything but J_point (edges whose point is fixed:
// EdgeSE3XYZ2CameraOnlyPose, include/OptimizeTypes.hpp:20-68, has the same
// residual and J_pose, and a fixed point's J_point is never used).
constexpr int kBaFull = 0, kBaResidual = 1, kBaPoseOnly = 2;

template <bool kNt>
__device__ __forceinline__ void ba_edge(int i, const BAIn& in, const BAPair& T, BAOut out, int mode) {
    const float2 uv = in.uv;
    const float isig = in.isig;
    const float4 ph = in.ph, pl = in.pl;
    const f2 P[3] = {{ph.x, pl.x}, {ph.y, pl.y}, {ph.z, pl.z}};
    // P_c = R_cw P_w + t_cw (float pairs)
    f2 C[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        f2 acc = {T.th[r], T.tl[r]};
#pragma unroll
        for (int k = 0; k < 3; k++) acc = df_add(acc, df_mul(f2{T.Rh[3 * r + k], T.Rl[3 * r + k]}, P[k]));
        C[r] = acc;
    }
    // e = (obs - c) - f * X / Z
    const float iz = __builtin_amdgcn_rcpf(C[2].hi);
    const f2 qx = df_div_rcp(C[0], C[2], iz);
    const f2 qy = df_div_rcp(C[1], C[2], iz);
    const f2 eu = df_add(two_sum(uv.x, -T.cx), df_neg(df_mul_f(qx, T.fx)));
    const f2 ev = df_add(two_sum(uv.y, -T.cy), df_neg(df_mul_f(qy, T.fy)));
    const float e0 = eu.hi + eu.lo, e1 = ev.hi + ev.lo;
    const float chi2 = e0 * (isig * e0) + e1 * (isig * e1);
    const float d2 = T.delta * T.delta;
    const float w = chi2 <= d2 ? 1.f : T.delta * __builtin_amdgcn_rsqf(chi2);
    const uint32_t off = (uint32_t)i * 4u;
    ba_store<kNt>(out, 0, off, e0);
    ba_store<kNt>(out, 1, off, e1);
    ba_store<kNt>(out, 2, off, chi2);
    ba_store<kNt>(out, 3, off, w);
    if (mode == kBaResidual) return;
    // Jacobians (FP32).  Jpi = [[fx/z, 0, -fx x/z^2], [0, fy/z, -fy y/z^2]]
    const float x = C[0].hi, y = C[1].hi;
    const float j00 = T.fx * iz, j02 = -T.fx * x * (iz * iz);
    const float j11 = T.fy * iz, j12 = -T.fy * y * (iz * iz);
    // J_point = -Jpi R_cw
    if (mode != kBaPoseOnly) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            ba_store<kNt>(out, 16 + k, off, -(j00 * T.Rh[k] + j02 * T.Rh[6 + k]));
            ba_store<kNt>(out, 19 + k, off, -(j11 * T.Rh[3 + k] + j12 * T.Rh[6 + k]));
        }
    }
    // P_b = R_bw P_w + t_bw; M = Jpi R_cb; J_pose = [M [P_b]x | -M]
    float pb[3];
#pragma unroll
    for (int r = 0; r < 3; r++) pb[r] = T.Rbw[3 * r] * ph.x + T.Rbw[3 * r + 1] * ph.y + T.Rbw[3 * r + 2] * ph.z + T.tbw[r];
    float Mr[6];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Mr[k] = j00 * T.Rcb[k] + j02 * T.Rcb[6 + k];
        Mr[3 + k] = j11 * T.Rcb[3 + k] + j12 * T.Rcb[6 + k];
    }
    float Jp[12];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const float m0 = Mr[3 * r], m1 = Mr[3 * r + 1], m2 = Mr[3 * r + 2];
        // M * [p]x, [p]x = [[0,-z,y],[z,0,-x],[-y,x,0]]
        Jp[6 * r + 0] = m1 * pb[2] - m2 * pb[1];
        Jp[6 * r + 1] = m2 * pb[0] - m0 * pb[2];
        Jp[6 * r + 2] = m0 * pb[1] - m1 * pb[0];
        Jp[6 * r + 3] = -m0;
        Jp[6 * r + 4] = -m1;
        Jp[6 * r + 5] = -m2;
    }
#pragma unroll
    for (int k = 0; k < 12; k++) ba_store<kNt>(out, 4 + k, off, Jp[k]);
}

// One observation per thread per step, grid-stride.  kLds: the (frame,
// camera) table (npairs <= kLdsPairs) is copied to dynamic LDS once per
// workgroup, so the per-observation parameter reads are LDS gathers
// (53-word records: odd stride, distinct records fall in distinct banks)
// instead of ~50 scattered global loads.  kIlp observations are in flight
// per thread: all their loads are issued before the first is computed.
template <bool kLds, int kIlp, bool kNt>
__global__ __launch_bounds__(1024) void k_ba_linearize(BAObs ob, const float4* __restrict__ pts_hi,
                                                       const float4* __restrict__ pts_lo,
                                                       const BAPair* __restrict__ pairs, int npairs, BAOut out,
                                                       int mode) {
    extern __shared__ float s_pairs[];  // npairs * kPairWords floats when kLds
    if (kLds) {
        const float* src = reinterpret_cast<const float*>(pairs);
        for (int k = threadIdx.x; k < npairs * kPairWords; k += blockDim.x) s_pairs[k] = src[k];
        __syncthreads();
    }
    const int stride = gridDim.x * blockDim.x;
    if (kIlp == 1) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ob.M; i += stride) {
            const uint32_t pr = ob.pair[i];
            const BAPair& T = kLds ? *reinterpret_cast<const BAPair*>(s_pairs + pr * kPairWords) : pairs[pr];
            BAIn in;
            ba_load_obs(i, ob, in);
            ba_load_point(pts_hi, pts_lo, in);
            ba_edge<kNt>(i, in, T, out, mode);
        }
        return;
    }
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < ob.M; i0 += kIlp * stride) {
        BAIn in[kIlp];
#pragma unroll
        for (int u = 0; u < kIlp; u++)
            if (i0 + u * stride < ob.M) ba_load_obs(i0 + u * stride, ob, in[u]);
#pragma unroll
        for (int u = 0; u < kIlp; u++)
            if (i0 + u * stride < ob.M) ba_load_point(pts_hi, pts_lo, in[u]);
#pragma unroll
        for (int u = 0; u < kIlp; u++) {
            const int i = i0 + u * stride;
            if (i >= ob.M) break;
            const BAPair& T = kLds ? *reinterpret_cast<const BAPair*>(s_pairs + in[u].pair * kPairWords)
                                   : pairs[in[u].pair];
            ba_edge<kNt>(i, in[u], T, out, mode);
        }
    }
}

// The plain launch pattern: the observations of one (frame, camera) edge
// group (list[0..n)), one launch per group.
__global__ __launch_bounds__(256) void k_ba_linearize_group(BAObs ob, const float4* __restrict__ pts_hi,
                                                            const float4* __restrict__ pts_lo,
                                                            const BAPair* __restrict__ pairs, const int* list, int n,
                                                            BAOut out, int mode) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = list[k];
    BAIn in;
    ba_load_obs(i, ob, in);
    ba_load_point(pts_hi, pts_lo, in);
    ba_edge<false>(i, in, pairs[in.pair], out, mode);
}

// ---------------------------------------------------------------------------
// Stereo edge EdgeStereoSE3ProjectXYZ (include/OptimizeTypes.hpp:119-157,
// src/OptimizeTypes.cpp:12-58).  The pose vertex is the keyframe pose (the
// left camera is the body).  cam_project_uv_u rounds the inverse depth to
// float and forms bf * invz as a float product; the rest is double in the
// reference, float pairs here.  32 output planes of M floats:
// e_u e_v e_ur chi2 w J_pose[3x6] J_point[3x3].
struct FramePose {
    float Rh[9], Rl[9];  // R_cw hi / lo
    float th[3], tl[3];  // t_cw hi / lo
};

struct StereoCam {
    float fx, fy, cx, cy, bf, delta;
};

struct BAStereoObs {
    const float* obs;       // M x 3: u, v, u_right
    const uint32_t* pt;
    const uint32_t* fr;
    const float* isig;
    int M;
};

__global__ __launch_bounds__(256) void k_ba_stereo_linearize(BAStereoObs ob, const float4* __restrict__ pts_hi,
                                                             const float4* __restrict__ pts_lo,
                                                             const FramePose* __restrict__ poses, StereoCam cam,
                                                             BAOut out, int mode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ob.M) return;
    const float ou = ob.obs[3 * i], ov = ob.obs[3 * i + 1], our = ob.obs[3 * i + 2];
    const float isig = ob.isig[i];
    const FramePose& T = poses[ob.fr[i]];
    const uint32_t pi = ob.pt[i];
    const float4 ph = pts_hi[pi], pl = pts_lo[pi];
    const f2 P[3] = {{ph.x, pl.x}, {ph.y, pl.y}, {ph.z, pl.z}};
    f2 C[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        f2 acc = {T.th[r], T.tl[r]};
#pragma unroll
        for (int k = 0; k < 3; k++) acc = df_add(acc, df_mul(f2{T.Rh[3 * r + k], T.Rl[3 * r + k]}, P[k]));
        C[r] = acc;
    }
    // invz = (float)(1 / z), rounded from the pair quotient
    const f2 q = df_div_rcp(f2{1.f, 0.f}, C[2], __builtin_amdgcn_rcpf(C[2].hi));
    const float invz = q.hi + q.lo;
    const f2 ru = df_add(df_mul_f(df_mul_f(C[0], invz), cam.fx), f2{cam.cx, 0.f});
    const f2 rv = df_add(df_mul_f(df_mul_f(C[1], invz), cam.fy), f2{cam.cy, 0.f});
    const float bfi = cam.bf * invz;
    const f2 eu = df_add(f2{ou, 0.f}, df_neg(ru));
    const f2 ev = df_add(f2{ov, 0.f}, df_neg(rv));
    const f2 er = df_add(df_add(f2{our, 0.f}, df_neg(ru)), f2{bfi, 0.f});
    const float e0 = eu.hi + eu.lo, e1 = ev.hi + ev.lo, e2 = er.hi + er.lo;
    const float chi2 = isig * (e0 * e0 + e1 * e1 + e2 * e2);
    const float w = chi2 <= cam.delta * cam.delta ? 1.f : cam.delta * __builtin_amdgcn_rsqf(chi2);
    const long long S = out.stride;
    float* o = out.planes + i;
    o[0] = e0;
    o[S] = e1;
    o[2 * S] = e2;
    o[3 * S] = chi2;
    o[4 * S] = w;
    if (mode == kBaResidual) return;
    const float x = C[0].hi, y = C[1].hi, z = C[2].hi;
    const float iz = __builtin_amdgcn_rcpf(z), iz2 = iz * iz;
    const float fx = cam.fx, fy = cam.fy, bf = cam.bf;
    float J0[6], J1[6], J2[6];
    J0[0] = x * y * iz2 * fx;
    J0[1] = -(1.f + x * x * iz2) * fx;
    J0[2] = y * iz * fx;
    J0[3] = -iz * fx;
    J0[4] = 0.f;
    J0[5] = x * iz2 * fx;
    J1[0] = (1.f + y * y * iz2) * fy;
    J1[1] = -x * y * iz2 * fy;
    J1[2] = -x * iz * fy;
    J1[3] = 0.f;
    J1[4] = -iz * fy;
    J1[5] = y * iz2 * fy;
    J2[0] = J0[0] - bf * y * iz2;
    J2[1] = J0[1] + bf * x * iz2;
    J2[2] = J0[2];
    J2[3] = J0[3];
    J2[4] = 0.f;
    J2[5] = J0[5] - bf * iz2;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        o[(5 + k) * S] = J0[k];
        o[(11 + k) * S] = J1[k];
        o[(17 + k) * S] = J2[k];
    }
    if (mode == kBaPoseOnly) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float a = -fx * T.Rh[k] * iz + fx * x * T.Rh[6 + k] * iz2;
        const float b = -fy * T.Rh[3 + k] * iz + fy * y * T.Rh[6 + k] * iz2;
        o[(23 + k) * S] = a;
        o[(26 + k) * S] = b;
        o[(29 + k) * S] = a - bf * T.Rh[6 + k] * iz2;
    }
}

// FilterCallBack_Chi2 (src/PoseEstimation.cpp:132-151): flag = chi2 >= threshold.
__global__ __launch_bounds__(256) void k_ba_outliers(const float* chi2, int M, float threshold, uint8_t* flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) flags[i] = chi2[i] >= threshold ? 1 : 0;
}

// grad_point[p] = sum over the observations of p (CSR, observation order) of
// J_point^T g.  One thread per point, fixed summation order.
__global__ __launch_bounds__(256) void k_ba_point_adj(const float* planes, long long S, const float2* g,
                                                      const int* ptr, int P, float* gout) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int i = ptr[p]; i < ptr[p + 1]; i++) {
        const float2 gi = g[i];
        a0 += planes[16 * S + i] * gi.x + planes[19 * S + i] * gi.y;
        a1 += planes[17 * S + i] * gi.x + planes[20 * S + i] * gi.y;
        a2 += planes[18 * S + i] * gi.x + planes[21 * S + i] * gi.y;
    }
    gout[3 * p] = a0;
    gout[3 * p + 1] = a1;
    gout[3 * p + 2] = a2;
}

// grad_pose[f] = sum over the observations of frame f of J_pose^T g.  One
// workgroup per frame; each thread sums a strided slice, then a fixed-order
// tree reduction: deterministic for a given launch shape.
__global__ __launch_bounds__(256) void k_ba_pose_adj(const float* planes, long long S, const float2* g,
                                                     const int* fptr, const int* fobs, float* gout) {
    __shared__ float red[6][256];
    const int f = blockIdx.x;
    float a[6] = {0, 0, 0, 0, 0, 0};
    for (int k = fptr[f] + threadIdx.x; k < fptr[f + 1]; k += 256) {
        const int i = fobs[k];
        const float2 gi = g[i];
#pragma unroll
        for (int c = 0; c < 6; c++) a[c] += planes[(4 + c) * S + i] * gi.x + planes[(10 + c) * S + i] * gi.y;
    }
#pragma unroll
    for (int c = 0; c < 6; c++) red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int c = 0; c < 6; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 6) gout[6 * f + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace mcv

namespace mcv {

// ---------------------------------------------------------------------------
// Schur complement of the point blocks (BlockSolver::buildSystem +
// BlockSolver::solve, ThirdParty/g2o/g2o/core/block_solver.hpp:315-449,
// 463-520; quadratic form of BaseBinaryEdge::constructQuadraticForm,
// core/base_binary_edge.hpp:64-136), FP32, fixed summation orders, over the
// active mono (2-row, k_ba_linearize planes) and stereo (3-row,
// k_ba_stereo_linearize planes) edges; edge index i < M is mono edge i,
// i >= M stereo edge i - M.  Fixed points have no edge lists (their edges
// still enter H_pp and b_p through the frame lists); fixed frames have no
// pose block.
//   S1  per point: H_ll = sum W J_x^T J_x, b_l = -sum W J_x^T e,
//       D^-1 = (H_ll + lambda I)^-1, db = D^-1 b_l, and per (point, free
//       frame) H_pl = sum W J_p^T J_x (6x3); W = Huber weight * inv sigma^2
//   S3  per free frame: H_pp = sum W J_p^T J_p, b_p = -sum W J_p^T e,
//       b_schur = b_p - sum H_pl db
//   S2  per free-frame pair block (i <= j): S_ij = [i == j](H_pp + lambda I)
//       - sum over shared points H_pl,i D^-1 H_pl,j^T
//   back-substitution: x_l = D^-1 (b_l - sum H_pl^T x_p)
struct SchurArgs {
    const float* planes;    // mono planes (22 of stride S)
    long long S;
    const float* isig;
    int M;
    const float* splanes;   // stereo planes (32 of stride SS)
    long long SS;
    const float* sisig;
    const int* pe_ptr;      // P + 1: active edges of free point p
    const int* pe;
    const int* pf_ptr;      // per point: first pf entry (P + 1)
    const int* pf_obs_ptr;  // per pf entry: range in pf_obs (npf + 1)
    const int* pf_obs;      // edge indices grouped by pf
    const int* pf_point;    // point of each pf entry
    const int* pf_frame;    // free frame of each pf entry
    const int* fptr;        // free frame -> active edges (Fp + 1)
    const int* fobs;
    const int* fpf_ptr;     // free frame -> pf entries (Fp + 1)
    const int* fpf;
    const int* blk_ptr;     // block (i <= j, row-major upper) -> item range
    const int2* blk_items;  // (pf_i, pf_j)
    const int2* blk_ij;     // block -> (i, j)
    int P, Fp, nblk;
    float lambda;
    float* Dinv;            // P x 9
    float* bl;              // P x 3
    float* db;              // P x 3
    float* Hll;             // P x 9 (without lambda)
    float* Hpl;             // npf x 18 (6x3 row-major)
    float* Hpp;             // Fp x 36 (without lambda)
    float* bs;              // Fp x 6 (b_schur)
    float* bp;              // Fp x 6 (b_p)
    float* Sout;            // (6 Fp)^2 row-major
};

// one edge's planes: e row r at base[r * st], J_pose (r, c) at
// base[(jp + 6 r + c) * st], J_point (r, c) at base[(jx + 3 r + c) * st]
struct EdgeRef {
    const float* base;
    long long st;
    int rows, jp, jx;
    float W;
    __device__ __forceinline__ float e(int r) const { return base[r * st]; }
    __device__ __forceinline__ float Jp(int r, int c) const { return base[(jp + 6 * r + c) * st]; }
    __device__ __forceinline__ float Jx(int r, int c) const { return base[(jx + 3 * r + c) * st]; }
};

__device__ __forceinline__ EdgeRef edge_ref(const SchurArgs& a, int i) {
    if (i < a.M) return EdgeRef{a.planes + i, a.S, 2, 4, 16, a.planes[3 * a.S + i] * a.isig[i]};
    const int j = i - a.M;
    return EdgeRef{a.splanes + j, a.SS, 3, 5, 23, a.splanes[4 * a.SS + j] * a.sisig[j]};
}

__global__ __launch_bounds__(256) void k_schur_points(SchurArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, b[3] = {0, 0, 0};
    const int e0 = a.pe_ptr[p], e1 = a.pe_ptr[p + 1];
    for (int t = e0; t < e1; t++) {
        const EdgeRef E = edge_ref(a, a.pe[t]);
        for (int r = 0; r < E.rows; r++) {
            const float er = E.e(r);
            const float A[3] = {E.Jx(r, 0), E.Jx(r, 1), E.Jx(r, 2)};
#pragma unroll
            for (int i = 0; i < 3; i++) {
                b[i] -= E.W * (A[i] * er);
#pragma unroll
                for (int c = 0; c < 3; c++) H[3 * i + c] += E.W * (A[i] * A[c]);
            }
        }
    }
    float Di[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (e1 > e0) {
        float D[9];
#pragma unroll
        for (int k = 0; k < 9; k++) D[k] = H[k];
        D[0] += a.lambda;
        D[4] += a.lambda;
        D[8] += a.lambda;
        const float c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8], c02 = D[3] * D[7] - D[4] * D[6];
        const float det = D[0] * c00 + D[1] * c01 + D[2] * c02;
        const float id = 1.f / det;
        Di[0] = c00 * id; Di[1] = (D[2] * D[7] - D[1] * D[8]) * id; Di[2] = (D[1] * D[5] - D[2] * D[4]) * id;
        Di[3] = c01 * id; Di[4] = (D[0] * D[8] - D[2] * D[6]) * id; Di[5] = (D[2] * D[3] - D[0] * D[5]) * id;
        Di[6] = c02 * id; Di[7] = (D[1] * D[6] - D[0] * D[7]) * id; Di[8] = (D[0] * D[4] - D[1] * D[3]) * id;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        a.Hll[9 * p + k] = H[k];
        a.Dinv[9 * p + k] = Di[k];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        a.bl[3 * p + r] = b[r];
        a.db[3 * p + r] = Di[3 * r] * b[0] + Di[3 * r + 1] * b[1] + Di[3 * r + 2] * b[2];
    }
    for (int q = a.pf_ptr[p]; q < a.pf_ptr[p + 1]; q++) {
        float Hq[18];
#pragma unroll
        for (int k = 0; k < 18; k++) Hq[k] = 0.f;
        for (int t = a.pf_obs_ptr[q]; t < a.pf_obs_ptr[q + 1]; t++) {
            const EdgeRef E = edge_ref(a, a.pf_obs[t]);
            for (int r = 0; r < E.rows; r++) {
                float B[6], A[3];
#pragma unroll
                for (int k = 0; k < 6; k++) B[k] = E.Jp(r, k);
#pragma unroll
                for (int k = 0; k < 3; k++) A[k] = E.Jx(r, k);
#pragma unroll
                for (int i = 0; i < 6; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) Hq[3 * i + c] += E.W * (B[i] * A[c]);
            }
        }
#pragma unroll
        for (int k = 0; k < 18; k++) a.Hpl[18 * (long long)q + k] = Hq[k];
    }
}

// fixed-order tree reduction of 256 x n partial sums (n <= 36); red[0][k]
// holds the sums
__device__ __forceinline__ void block_sum(const float* acc, float (*red)[36], int n) {
    for (int k = 0; k < n; k++) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < n; k++) red[threadIdx.x][k] += red[threadIdx.x + s][k];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_schur_frames(SchurArgs a) {
    __shared__ float red[256][36];
    const int f = blockIdx.x;
    float acc[36];
    float bp[12];  // b_p (0..5) and the Schur coupling sum H_pl db (6..11)
#pragma unroll
    for (int k = 0; k < 36; k++) acc[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) bp[k] = 0.f;
    for (int t = a.fptr[f] + threadIdx.x; t < a.fptr[f + 1]; t += 256) {
        const EdgeRef E = edge_ref(a, a.fobs[t]);
        for (int r = 0; r < E.rows; r++) {
            const float er = E.e(r);
            float B[6];
#pragma unroll
            for (int k = 0; k < 6; k++) B[k] = E.Jp(r, k);
#pragma unroll
            for (int i = 0; i < 6; i++) {
                bp[i] -= E.W * (B[i] * er);
#pragma unroll
                for (int c = 0; c < 6; c++) acc[6 * i + c] += E.W * (B[i] * B[c]);
            }
        }
    }
    for (int t = a.fpf_ptr[f] + threadIdx.x; t < a.fpf_ptr[f + 1]; t += 256) {
        const int q = a.fpf[t];
        const int p = a.pf_point[q];
        const float* Hq = a.Hpl + 18 * (long long)q;
        const float d0 = a.db[3 * p], d1 = a.db[3 * p + 1], d2 = a.db[3 * p + 2];
#pragma unroll
        for (int r = 0; r < 6; r++) bp[6 + r] += Hq[3 * r] * d0 + Hq[3 * r + 1] * d1 + Hq[3 * r + 2] * d2;
    }
    block_sum(acc, red, 36);
    if (threadIdx.x < 36) a.Hpp[36 * f + threadIdx.x] = red[0][threadIdx.x];
    __syncthreads();
    block_sum(bp, red, 12);
    if (threadIdx.x < 6) {
        a.bp[6 * f + threadIdx.x] = red[0][threadIdx.x];
        a.bs[6 * f + threadIdx.x] = red[0][threadIdx.x] - red[0][6 + threadIdx.x];
    }
}

__global__ __launch_bounds__(256) void k_schur_blocks(SchurArgs a) {
    __shared__ float red[256][36];
    const int blk = blockIdx.x;
    const int2 ij = a.blk_ij[blk];
    float acc[36];
#pragma unroll
    for (int k = 0; k < 36; k++) acc[k] = 0.f;
    for (int t = a.blk_ptr[blk] + threadIdx.x; t < a.blk_ptr[blk + 1]; t += 256) {
        const int2 it = a.blk_items[t];
        const int p = a.pf_point[it.x];
        const float* Hi = a.Hpl + 18 * (long long)it.x;
        const float* Hj = a.Hpl + 18 * (long long)it.y;
        const float* Di = a.Dinv + 9 * p;
        // Y = D^-1 H_j^T (3 x 6)
        float Y[18];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 6; c++)
                Y[6 * r + c] = Di[3 * r] * Hj[3 * c] + Di[3 * r + 1] * Hj[3 * c + 1] + Di[3 * r + 2] * Hj[3 * c + 2];
#pragma unroll
        for (int r = 0; r < 6; r++)
#pragma unroll
            for (int c = 0; c < 6; c++)
                acc[6 * r + c] += Hi[3 * r] * Y[c] + Hi[3 * r + 1] * Y[6 + c] + Hi[3 * r + 2] * Y[12 + c];
    }
    block_sum(acc, red, 36);
    if (threadIdx.x < 36) {
        const int r = threadIdx.x / 6, c = threadIdx.x % 6;
        const int n = 6 * a.Fp;
        float v = -red[0][threadIdx.x];
        if (ij.x == ij.y) v += a.Hpp[36 * ij.x + threadIdx.x] + (r == c ? a.lambda : 0.f);
        a.Sout[(long long)(6 * ij.x + r) * n + 6 * ij.y + c] = v;
        if (ij.x != ij.y) a.Sout[(long long)(6 * ij.y + c) * n + 6 * ij.x + r] = v;
    }
}

// fixed-order sum of one double per thread over the workgroup (256)
__device__ __forceinline__ double block_sum_d(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// x_l = D^-1 (b_l - sum over the point's pf entries of H_pl^T x_p) per point
// (0 for points without edges), and per workgroup the partial sum of
// x_l . (lambda x_l + b_l) (the points' share of the LM scale).
__global__ __launch_bounds__(256) void k_schur_backsub(SchurArgs a, const float* xp, float* xl, double* partial) {
    __shared__ double red[256];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double sc = 0.0;
    if (p < a.P) {
        float c[3] = {a.bl[3 * p], a.bl[3 * p + 1], a.bl[3 * p + 2]};
        for (int q = a.pf_ptr[p]; q < a.pf_ptr[p + 1]; q++) {
            const float* Hq = a.Hpl + 18 * (long long)q;
            const float* x = xp + 6 * a.pf_frame[q];
#pragma unroll
            for (int k = 0; k < 3; k++)
                c[k] -= Hq[k] * x[0] + Hq[3 + k] * x[1] + Hq[6 + k] * x[2] + Hq[9 + k] * x[3] + Hq[12 + k] * x[4] +
                        Hq[15 + k] * x[5];
        }
        const float* Di = a.Dinv + 9 * p;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float v = Di[3 * k] * c[0] + Di[3 * k + 1] * c[1] + Di[3 * k + 2] * c[2];
            xl[3 * p + k] = v;
            sc += (double)v * ((double)a.lambda * (double)v + (double)a.bl[3 * p + k]);
        }
    }
    const double t = block_sum_d(sc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Huber rho of chi2 (core/robust_kernel_impl.cpp:65-78) summed over the
// active edges, per workgroup: mono edge i takes delta from its (frame,
// camera) pair, stereo edges the stereo delta; an infinite delta is the
// edge without a robust kernel (rho = chi2).
__device__ __forceinline__ double huber_rho(float chi2, float delta) {
    const double e = chi2, d = delta;
    return e <= d * d ? e : 2.0 * sqrt(e) * d - d * d;
}

__global__ __launch_bounds__(256) void k_ba_rho_sum(const float* chi2, const uint32_t* pair, const BAPair* pairs,
                                                    const uint8_t* act, int M, const float* schi2,
                                                    const uint8_t* sact, int Ms, float sdelta, double* partial) {
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < M) {
        if (!act || act[i]) v = huber_rho(chi2[i], pairs[pair[i]].delta);
    } else if (i - M < Ms) {
        const int j = i - M;
        if (!sact || sact[j]) v = huber_rho(schi2[j], sdelta);
    }
    const double t = block_sum_d(v, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// out[0] = sum of partial[0 .. n) in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void k_sum_partials(const double* partial, int n, double* out) {
    __shared__ double red[256];
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) v += partial[k];
    const double t = block_sum_d(v, red);
    if (threadIdx.x == 0) out[0] = t;
}

// largest |diagonal entry| of H_pp (free frames) and H_ll (points with
// edges), computeLambdaInit's maxDiagonal: per-workgroup maxima over a
// grid-stride loop into partial[blockIdx.x], then k_max_partials
__global__ __launch_bounds__(256) void k_ba_max_diag(const float* Hpp, int Fp, const float* Hll, const int* pe_ptr,
                                                     int P, float* partial) {
    __shared__ float red[256];
    float m = 0.f;
    const int stride = gridDim.x * 256;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < 6 * Fp; k += stride)
        m = fmaxf(m, fabsf(Hpp[36 * (k / 6) + 7 * (k % 6)]));
    for (int k = blockIdx.x * 256 + threadIdx.x; k < 3 * P; k += stride) {
        const int p = k / 3;
        if (pe_ptr[p + 1] > pe_ptr[p]) m = fmaxf(m, fabsf(Hll[9 * p + 4 * (k % 3)]));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_max_partials(const float* partial, int n, float* out) {
    __shared__ float red[256];
    float m = 0.f;
    for (int k = threadIdx.x; k < n; k += 256) m = fmaxf(m, partial[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// FP64 point estimates -> (hi, lo) float pairs; with xl, first the additive
// update P += x_l of the free points (VertexSBAPointXYZ::oplusImpl)
__global__ __launch_bounds__(256) void k_ba_points_apply(double* p64, const float* xl, const uint8_t* fixed, int P,
                                                         float4* hi, float4* lo) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double v[3] = {p64[3 * p], p64[3 * p + 1], p64[3 * p + 2]};
    if (xl && !(fixed && fixed[p])) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[k] += (double)xl[3 * p + k];
            p64[3 * p + k] = v[k];
        }
    }
    float h[3], l[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        h[k] = (float)v[k];
        l[k] = (float)(v[k] - (double)h[k]);
    }
    hi[p] = make_float4(h[0], h[1], h[2], 0.f);
    lo[p] = make_float4(l[0], l[1], l[2], 0.f);
}

}  // namespace mcv
