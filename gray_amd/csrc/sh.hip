// Spherical-harmonics colour (degree <= 3) forward and backward on gfx950.
//
// Reference: cuda/utils/sh.cu:11-304 evaluates the real SH basis with a
// complex recursion.  The basis is the same polynomial family as the
// closed form below (the CPU oracle restates the recursion; the tests pin the
// two against each other); the closed form is cheaper on the device.
// Ordering of the 15 "rest" coefficients per degree l: m = -l..l
// (reference sh.cu:54-67).  One thread per Gaussian evaluates every view of
// the launch, so coefficients are read from HBM once per launch.

#include <hip/hip_fp16.h>
#include <math.h>

#include <algorithm>

#include "ctx.hpp"

namespace {

// real SH normalisation constants: C0 = 1/(2 sqrt(pi)), C1 = sqrt(3/(4 pi)),
// C2 = {sqrt(15/(4pi)), -sqrt(15/(4pi)), sqrt(5/(16pi)), -sqrt(15/(4pi)), sqrt(15/(16pi))},
// C3 = {-sqrt(35/(32pi)), sqrt(105/(4pi)), -sqrt(21/(32pi)), sqrt(7/(16pi)),
//       -sqrt(21/(32pi)), sqrt(105/(16pi)), -sqrt(35/(32pi))}
constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC20 = 1.0925484305920792f;
constexpr float kC22 = 0.31539156525252005f;
constexpr float kC24 = 0.5462742152960396f;
constexpr float kC30 = 0.5900435899266435f;
constexpr float kC31 = 2.890611442640554f;
constexpr float kC32 = 0.4570457994644658f;
constexpr float kC33 = 0.3731763325901154f;
constexpr float kC35 = 1.445305721320277f;

constexpr int kMaxViews = 8;

struct ShViews {
    const float *origin[kMaxViews];
    float *rgb[kMaxViews];
    const float *rgb_in[kMaxViews];
    const float *grad_rgb[kMaxViews];
    int num;
};

__device__ __forceinline__ void sh_basis(float x, float y, float z, int degree, float *b) {
    b[0] = -kC1 * y;
    b[1] = kC1 * z;
    b[2] = -kC1 * x;
    if (degree < 2) return;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    b[3] = kC20 * xy;
    b[4] = -kC20 * yz;
    b[5] = kC22 * (2.0f * zz - xx - yy);
    b[6] = -kC20 * xz;
    b[7] = kC24 * (xx - yy);
    if (degree < 3) return;
    b[8] = -kC30 * y * (3.0f * xx - yy);
    b[9] = kC31 * xy * z;
    b[10] = -kC32 * y * (4.0f * zz - xx - yy);
    b[11] = kC33 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    b[12] = -kC32 * x * (4.0f * zz - xx - yy);
    b[13] = kC35 * z * (xx - yy);
    b[14] = -kC30 * x * (xx - 3.0f * yy);
}

// d b_k / d(x, y, z) treating the components as independent
__device__ __forceinline__ void sh_basis_grad(float x, float y, float z, int degree, float *bx, float *by, float *bz) {
    bx[0] = 0.0f, by[0] = -kC1, bz[0] = 0.0f;
    bx[1] = 0.0f, by[1] = 0.0f, bz[1] = kC1;
    bx[2] = -kC1, by[2] = 0.0f, bz[2] = 0.0f;
    if (degree < 2) return;
    const float xx = x * x, yy = y * y, zz = z * z;
    bx[3] = kC20 * y, by[3] = kC20 * x, bz[3] = 0.0f;
    bx[4] = 0.0f, by[4] = -kC20 * z, bz[4] = -kC20 * y;
    bx[5] = -2.0f * kC22 * x, by[5] = -2.0f * kC22 * y, bz[5] = 4.0f * kC22 * z;
    bx[6] = -kC20 * z, by[6] = 0.0f, bz[6] = -kC20 * x;
    bx[7] = 2.0f * kC24 * x, by[7] = -2.0f * kC24 * y, bz[7] = 0.0f;
    if (degree < 3) return;
    bx[8] = -kC30 * 6.0f * x * y, by[8] = -kC30 * (3.0f * xx - 3.0f * yy), bz[8] = 0.0f;
    bx[9] = kC31 * y * z, by[9] = kC31 * x * z, bz[9] = kC31 * x * y;
    bx[10] = -kC32 * (-2.0f * x * y), by[10] = -kC32 * (4.0f * zz - xx - 3.0f * yy), bz[10] = -kC32 * 8.0f * y * z;
    bx[11] = kC33 * (-6.0f * x * z), by[11] = kC33 * (-6.0f * y * z), bz[11] = kC33 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
    bx[12] = -kC32 * (4.0f * zz - 3.0f * xx - yy), by[12] = -kC32 * (-2.0f * x * y), bz[12] = -kC32 * 8.0f * x * z;
    bx[13] = kC35 * 2.0f * x * z, by[13] = kC35 * (-2.0f * y * z), bz[13] = kC35 * (xx - yy);
    bx[14] = -kC30 * (3.0f * xx - 3.0f * yy), by[14] = -kC30 * (-6.0f * x * y), bz[14] = 0.0f;
}

template <typename T>
__device__ __forceinline__ float ld(const T *p, size_t i);
template <>
__device__ __forceinline__ float ld<float>(const float *p, size_t i) {
    return p[i];
}
template <>
__device__ __forceinline__ float ld<__half>(const __half *p, size_t i) {
    return __half2float(p[i]);
}

__device__ __forceinline__ int rest_count(int degree) { return (degree + 1) * (degree + 1) - 1; }

// Bodies are instantiated per degree so that every coefficient and basis
// array has compile-time indices and lives in registers; the kernels switch
// once on the (launch-uniform) degree read from the device.
// Three consecutive floats (4-byte aligned): one global_load/store_dwordx3.
struct F3 {
    float x, y, z;
};
// Four consecutive floats at 4-byte alignment: one global_load_dwordx4 (the
// vector memory instructions need dword alignment only).
typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));

// The used coefficients of Gaussian g: fp32 in 16-byte vectors (a Gaussian's
// K*3 floats are contiguous), so a wave issues 12 wide loads instead of 45
// scalar ones over the same cache lines; fp16 element by element.
template <int KC>
__device__ __forceinline__ void load_rest(const float *__restrict__ rest, int g, int K, float (*c)[3]) {
    const float *base = rest + (size_t)g * K * 3;
    constexpr int NF = KC * 3, NV = NF / 4;
    float f[NF > 0 ? NF : 1];
    const f4a4 *p = reinterpret_cast<const f4a4 *>(base);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const f4a4 v = p[i];
        f[4 * i] = v.x, f[4 * i + 1] = v.y, f[4 * i + 2] = v.z, f[4 * i + 3] = v.w;
    }
#pragma unroll
    for (int i = 4 * NV; i < NF; ++i) f[i] = base[i];
#pragma unroll
    for (int k = 0; k < KC; ++k)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) c[k][ch] = f[3 * k + ch];
}
template <int KC>
__device__ __forceinline__ void load_rest(const __half *__restrict__ rest, int g, int K, float (*c)[3]) {
#pragma unroll
    for (int k = 0; k < KC; ++k)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) c[k][ch] = ld(rest, ((size_t)g * K + k) * 3 + ch);
}
__device__ __forceinline__ F3 load3(const float *__restrict__ p, int g) { return reinterpret_cast<const F3 *>(p)[g]; }
__device__ __forceinline__ F3 load3(const __half *__restrict__ p, int g) {
    return F3{ld(p, 3 * (size_t)g), ld(p, 3 * (size_t)g + 1), ld(p, 3 * (size_t)g + 2)};
}

// Colour of Gaussian g for every view of the launch from its DC term, mean
// and higher coefficients c.
template <int DEG>
__device__ __forceinline__ void sh_forward_eval(int g, const F3 &dcv, const F3 &mv, const float (*c)[3],
                                                const ShViews &V);

template <int DEG, typename T>
__device__ __forceinline__ void sh_forward_body(int g, int K, const float *__restrict__ mean, const T *__restrict__ dc,
                                                const T *__restrict__ rest, const ShViews &V) {
    constexpr int KC = DEG > 0 ? (DEG + 1) * (DEG + 1) - 1 : 0;
    float c[KC > 0 ? KC : 1][3];
    load_rest<KC>(rest, g, K, c);
    sh_forward_eval<DEG>(g, load3(dc, g), load3(mean, g), c, V);
}

template <int DEG>
__device__ __forceinline__ void sh_forward_eval(int g, const F3 &dcv, const F3 &mv, const float (*c)[3],
                                                const ShViews &V) {
    constexpr int KC = DEG > 0 ? (DEG + 1) * (DEG + 1) - 1 : 0;
    const float d0 = dcv.x, d1 = dcv.y, d2 = dcv.z;
    const float mx = mv.x, my = mv.y, mz = mv.z;
    for (int v = 0; v < V.num; ++v) {
        const float *o = V.origin[v];
        const float dx = mx - o[0], dy = my - o[1], dz = mz - o[2];
        const float inv = 1.0f / sqrtf(fmaxf(dx * dx + dy * dy + dz * dz, 1e-20f));
        const float x = dx * inv, y = dy * inv, z = dz * inv;
        float r0 = kC0 * d0, r1 = kC0 * d1, r2 = kC0 * d2;
        if constexpr (DEG > 0) {
            float b[15];
            sh_basis(x, y, z, DEG, b);
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                r0 = fmaf(b[k], c[k][0], r0);
                r1 = fmaf(b[k], c[k][1], r1);
                r2 = fmaf(b[k], c[k][2], r2);
            }
        }
        reinterpret_cast<F3 *>(V.rgb[v])[g] = F3{fmaxf(r0 + 0.5f, 0.0f), fmaxf(r1 + 0.5f, 0.0f), fmaxf(r2 + 0.5f, 0.0f)};
    }
}

__device__ __forceinline__ int sh_degree(const int32_t *cur_degree, int K) {
    return min(min(*cur_degree, 3), K >= 15 ? 3 : (K >= 8 ? 2 : (K >= 3 ? 1 : 0)));
}

template <typename T>
__global__ void k_sh_forward(int n, int K, const int32_t *__restrict__ cur_degree, const uint8_t *__restrict__ enable,
                             const float *__restrict__ mean, const T *__restrict__ dc, const T *__restrict__ rest,
                             ShViews V) {
    if (!*enable) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    switch (sh_degree(cur_degree, K)) {
        case 3: sh_forward_body<3>(g, K, mean, dc, rest, V); break;
        case 2: sh_forward_body<2>(g, K, mean, dc, rest, V); break;
        case 1: sh_forward_body<1>(g, K, mean, dc, rest, V); break;
        default: sh_forward_body<0>(g, K, mean, dc, rest, V); break;
    }
}

// fp32 coefficients staged through LDS: each wave owns 64 consecutive
// Gaussians, whose K*3*64 coefficient floats are one contiguous block, and
// copies it into its LDS slice with global_load_lds_dwordx4 (every load
// instruction moves 1 KB of consecutive memory, straight into LDS), then every
// lane reads its own K*3 floats (stride K*3 dwords: odd for K = 15, so no
// bank conflicts).  The last, partial wave of the array loads directly.  The
// staging loads are non-temporal (cache policy nt): the coefficients are read
// once per frame, and without allocating in the caches the kernel streams at
// about 5 TB/s instead of 3.1 (the DC terms and means are read the same way).
constexpr int kShStageAux = 2;  // global_load_lds cache policy bits: nt
constexpr int kShLdsWaves = 4;
constexpr int kShMaxK = 15;

__global__ __launch_bounds__(64 * kShLdsWaves) void k_sh_forward_lds(int n, int K, const int32_t *__restrict__ cur_degree,
                                                                    const uint8_t *__restrict__ enable,
                                                                    const float *__restrict__ mean,
                                                                    const float *__restrict__ dc,
                                                                    const float *__restrict__ rest, ShViews V) {
    __shared__ float4 stage[kShLdsWaves][kShMaxK * 3 * 64 / 4];
    if (!*enable) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g0 = (blockIdx.x * kShLdsWaves + wave) * 64;
    if (g0 >= n) return;
    const int g = g0 + lane;
    const int deg = sh_degree(cur_degree, K);
    if (g0 + 64 > n || deg == 0) {
        if (g >= n) return;
        switch (deg) {
            case 3: sh_forward_body<3>(g, K, mean, dc, rest, V); break;
            case 2: sh_forward_body<2>(g, K, mean, dc, rest, V); break;
            case 1: sh_forward_body<1>(g, K, mean, dc, rest, V); break;
            default: sh_forward_body<0>(g, K, mean, dc, rest, V); break;
        }
        return;
    }
    // DC term and mean, non-temporal like the staging (in flight with it)
    const F3 dcv = {__builtin_nontemporal_load(dc + 3 * g), __builtin_nontemporal_load(dc + 3 * g + 1),
                    __builtin_nontemporal_load(dc + 3 * g + 2)};
    const F3 mv = {__builtin_nontemporal_load(mean + 3 * g), __builtin_nontemporal_load(mean + 3 * g + 1),
                   __builtin_nontemporal_load(mean + 3 * g + 2)};
    // the wave's block: 64 * K * 3 floats = 48 K chunks of 16 bytes
    const float4 *src = reinterpret_cast<const float4 *>(rest + (size_t)g0 * K * 3);
    const int chunks = 48 * K;
    for (int i = 0; i * 64 < chunks; ++i) {
        const int c = i * 64 + lane;
        if (c < chunks)
            __builtin_amdgcn_global_load_lds(src + c, (__attribute__((address_space(3))) void *)&stage[wave][i * 64],
                                             16, 0, kShStageAux);
    }
    __builtin_amdgcn_s_waitcnt(0);
    const float *mine = reinterpret_cast<const float *>(stage[wave]) + lane * K * 3;
    float c[kShMaxK][3];
    switch (deg) {
        case 3:
#pragma unroll
            for (int k = 0; k < 15; ++k)
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) c[k][ch] = mine[3 * k + ch];
            sh_forward_eval<3>(g, dcv, mv, c, V);
            break;
        case 2:
#pragma unroll
            for (int k = 0; k < 8; ++k)
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) c[k][ch] = mine[3 * k + ch];
            sh_forward_eval<2>(g, dcv, mv, c, V);
            break;
        default:
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) c[k][ch] = mine[3 * k + ch];
            sh_forward_eval<1>(g, dcv, mv, c, V);
            break;
    }
}

// reference sh_backward_kernel (sh.cu:249-287), summed over the views
template <int DEG>
__device__ __forceinline__ void sh_backward_body(int g, int K, const float *__restrict__ mean,
                                                 const float *__restrict__ rest, float *__restrict__ grad_dc,
                                                 float *__restrict__ grad_rest, float *__restrict__ grad_mean,
                                                 int accumulate, const ShViews &V) {
    constexpr int KC = DEG > 0 ? (DEG + 1) * (DEG + 1) - 1 : 0;
    constexpr int KA = KC > 0 ? KC : 1;
    float c[KA][3];
#pragma unroll
    for (int k = 0; k < KC; ++k)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) c[k][ch] = rest[((size_t)g * K + k) * 3 + ch];
    const float mx = mean[3 * g], my = mean[3 * g + 1], mz = mean[3 * g + 2];
    float gdc0 = 0.0f, gdc1 = 0.0f, gdc2 = 0.0f;
    float grest[KA][3];
#pragma unroll
    for (int k = 0; k < KC; ++k) grest[k][0] = grest[k][1] = grest[k][2] = 0.0f;
    float gm0 = 0.0f, gm1 = 0.0f, gm2 = 0.0f;
    for (int v = 0; v < V.num; ++v) {
        const float *o = V.origin[v];
        const float *rgb = V.rgb_in[v];
        const float *grgb = V.grad_rgb[v];
        float g0 = grgb[3 * g], g1 = grgb[3 * g + 1], g2 = grgb[3 * g + 2];
        if (rgb[3 * g + 0] <= 0.0f) g0 = 0.0f;
        if (rgb[3 * g + 1] <= 0.0f) g1 = 0.0f;
        if (rgb[3 * g + 2] <= 0.0f) g2 = 0.0f;
        const float dx = mx - o[0], dy = my - o[1], dz = mz - o[2];
        const float len_sq = fmaxf(dx * dx + dy * dy + dz * dz, 1e-20f);
        const float inv = 1.0f / sqrtf(len_sq);
        const float x = dx * inv, y = dy * inv, z = dz * inv;
        if (accumulate) {
            gdc0 += kC0 * g0;
            gdc1 += kC0 * g1;
            gdc2 += kC0 * g2;
        } else {
            gdc0 = kC0 * g0;
            gdc1 = kC0 * g1;
            gdc2 = kC0 * g2;
        }
        float gdx = 0.0f, gdy = 0.0f, gdz = 0.0f;
        if constexpr (DEG > 0) {
            float b[15], bx[15], by[15], bz[15];
            sh_basis(x, y, z, DEG, b);
            sh_basis_grad(x, y, z, DEG, bx, by, bz);
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                if (accumulate) {
                    grest[k][0] += b[k] * g0;
                    grest[k][1] += b[k] * g1;
                    grest[k][2] += b[k] * g2;
                } else {
                    grest[k][0] = b[k] * g0;
                    grest[k][1] = b[k] * g1;
                    grest[k][2] = b[k] * g2;
                }
                const float cg = c[k][0] * g0 + c[k][1] * g1 + c[k][2] * g2;
                gdx += bx[k] * cg;
                gdy += by[k] * cg;
                gdz += bz[k] * cg;
            }
        }
        const float inv3 = inv * inv * inv;
        const float dd = dx * gdx + dy * gdy + dz * gdz;
        gm0 += gdx * inv - dx * (dd * inv3);
        gm1 += gdy * inv - dy * (dd * inv3);
        gm2 += gdz * inv - dz * (dd * inv3);
    }
    if (accumulate) {
        grad_dc[3 * g + 0] += gdc0;
        grad_dc[3 * g + 1] += gdc1;
        grad_dc[3 * g + 2] += gdc2;
#pragma unroll
        for (int k = 0; k < KC; ++k)
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) grad_rest[((size_t)g * K + k) * 3 + ch] += grest[k][ch];
    } else {
        grad_dc[3 * g + 0] = gdc0;
        grad_dc[3 * g + 1] = gdc1;
        grad_dc[3 * g + 2] = gdc2;
#pragma unroll
        for (int k = 0; k < KC; ++k)
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) grad_rest[((size_t)g * K + k) * 3 + ch] = grest[k][ch];
    }
    grad_mean[3 * g + 0] += gm0;
    grad_mean[3 * g + 1] += gm1;
    grad_mean[3 * g + 2] += gm2;
}

// 128-thread launches: the bound lifts the register cap that a 1024-thread
// default would impose (no spill of the 45 coefficient gradients)
__global__ __launch_bounds__(128) void k_sh_backward(int n, int K, const int32_t *__restrict__ cur_degree, const uint8_t *__restrict__ enable,
                              const float *__restrict__ mean, const float *__restrict__ rest,
                              const uint8_t *__restrict__ visible, float *__restrict__ grad_dc,
                              float *__restrict__ grad_rest, float *__restrict__ grad_mean, int accumulate,
                              ShViews V) {
    if (!*enable) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    if (visible && !visible[g]) return;
    switch (sh_degree(cur_degree, K)) {
        case 3: sh_backward_body<3>(g, K, mean, rest, grad_dc, grad_rest, grad_mean, accumulate, V); break;
        case 2: sh_backward_body<2>(g, K, mean, rest, grad_dc, grad_rest, grad_mean, accumulate, V); break;
        case 1: sh_backward_body<1>(g, K, mean, rest, grad_dc, grad_rest, grad_mean, accumulate, V); break;
        default: sh_backward_body<0>(g, K, mean, rest, grad_dc, grad_rest, grad_mean, accumulate, V); break;
    }
}

}  // namespace

extern "C" int gray_sh_forward(gray_ctx *ctx, const gray_params *params, const gray_knobs *knobs, int32_t num_views,
                               const float *const *origins, float *const *rgb_out, void *stream) {
    if (!ctx || !params || !knobs || !knobs->enable_sh || !origins || !rgb_out || num_views <= 0) {
        gray_set_error(GRAY_ERR_INVALID, "sh_forward: null argument or no views");
        return GRAY_ERR_INVALID;
    }
    const int n = params->count;
    if (n <= 0) return GRAY_OK;
    if (!params->mean || !params->sh_dc || (params->num_sh_coeffs > 0 && !params->sh_rest) ||
        !params->current_sh_degree) {
        gray_set_error(GRAY_ERR_INVALID, "sh_forward: null parameter pointer");
        return GRAY_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    for (int first = 0; first < num_views; first += kMaxViews) {
        ShViews V{};
        V.num = std::min(kMaxViews, num_views - first);
        for (int i = 0; i < V.num; ++i) {
            V.origin[i] = origins[first + i];
            V.rgb[i] = rgb_out[first + i];
            if (!V.origin[i] || !V.rgb[i]) {
                gray_set_error(GRAY_ERR_INVALID, "sh_forward: null origin/rgb for view %d", first + i);
                return GRAY_ERR_INVALID;
            }
        }
        if (params->sh_is_fp16)
            hipLaunchKernelGGL(k_sh_forward<__half>, dim3(gray_div_up(n, 128)), dim3(128), 0, s, n,
                               params->num_sh_coeffs, params->current_sh_degree, knobs->enable_sh, params->mean,
                               (const __half *)params->sh_dc, (const __half *)params->sh_rest, V);
        else if (params->num_sh_coeffs > 0 && params->num_sh_coeffs <= kShMaxK &&
                 ((uintptr_t)params->sh_rest & 15u) == 0)
            hipLaunchKernelGGL(k_sh_forward_lds, dim3(gray_div_up(n, 64 * kShLdsWaves)), dim3(64 * kShLdsWaves), 0, s,
                               n, params->num_sh_coeffs, params->current_sh_degree, knobs->enable_sh, params->mean,
                               (const float *)params->sh_dc, (const float *)params->sh_rest, V);
        else
            hipLaunchKernelGGL(k_sh_forward<float>, dim3(gray_div_up(n, 128)), dim3(128), 0, s, n,
                               params->num_sh_coeffs, params->current_sh_degree, knobs->enable_sh, params->mean,
                               (const float *)params->sh_dc, (const float *)params->sh_rest, V);
        GRAY_HIP(hipGetLastError());
    }
    return GRAY_OK;
}

extern "C" int gray_sh_backward(gray_ctx *ctx, const gray_params *params, const gray_knobs *knobs, int32_t num_views,
                                const float *const *origins, const float *const *rgb, const float *const *grad_rgb,
                                const uint8_t *was_visible, float *grad_sh_dc, float *grad_sh_rest, float *grad_mean,
                                int accumulate, void *stream) {
    if (!ctx || !params || !knobs || !knobs->enable_sh || !origins || !rgb || !grad_rgb || num_views <= 0 ||
        !grad_sh_dc || (params->num_sh_coeffs > 0 && !grad_sh_rest) || !grad_mean) {
        gray_set_error(GRAY_ERR_INVALID, "sh_backward: null argument or no views");
        return GRAY_ERR_INVALID;
    }
    if (params->sh_is_fp16) {
        gray_set_error(GRAY_ERR_INVALID, "sh_backward: fp16 (inference-only) coefficients have no gradient");
        return GRAY_ERR_INVALID;
    }
    if (num_views > kMaxViews && !accumulate) {
        gray_set_error(GRAY_ERR_INVALID, "sh_backward: assignment semantics need <= %d views", kMaxViews);
        return GRAY_ERR_INVALID;
    }
    const int n = params->count;
    if (n <= 0) return GRAY_OK;
    hipStream_t s = (hipStream_t)stream;
    for (int first = 0; first < num_views; first += kMaxViews) {
        ShViews V{};
        V.num = std::min(kMaxViews, num_views - first);
        for (int i = 0; i < V.num; ++i) {
            V.origin[i] = origins[first + i];
            V.rgb_in[i] = rgb[first + i];
            V.grad_rgb[i] = grad_rgb[first + i];
        }
        hipLaunchKernelGGL(k_sh_backward, dim3(gray_div_up(n, 128)), dim3(128), 0, s, n, params->num_sh_coeffs,
                           params->current_sh_degree, knobs->enable_sh, params->mean,
                           (const float *)params->sh_rest, was_visible, grad_sh_dc, grad_sh_rest, grad_mean,
                           accumulate, V);
        GRAY_HIP(hipGetLastError());
    }
    return GRAY_OK;
}
