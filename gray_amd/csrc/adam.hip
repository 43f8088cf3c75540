// Fused Adam over every Gaussian tensor in one launch (gfx950).
//
// Reference: cuda/opt/adam.cu:6-84 runs one kernel over mean/rotation/scale/
// opacity(/channels) per Gaussian and a second over the 16 SH coefficients,
// then synchronises the device.  Here the seven tensors are streamed as flat
// fp32 arrays in 16-byte vectors (each workgroup owns a contiguous slice of
// one tensor, so every access is coalesced), and the pruning counter update of
// gray/raytracer.py:361-366 rides along as an eighth segment.

#include <math.h>

#include <algorithm>

#include "ctx.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kVecPerThread = 2;                         // float4 per thread
constexpr int kFloatsPerBlock = kThreads * kVecPerThread * 4;  // 2048

struct AdamArgs {
    float *p[7], *g[7], *m[7], *v[7];
    const float *lr[7];
    int64_t numel[7];
    int32_t block_start[9];  // first block of each segment (segment 7 = pruning counter)
    const float *beta1, *beta2, *eps;
    const uint8_t *visible;
    int32_t *counter;
    int32_t count;
    int32_t host_flags;
    int32_t step;
    float step_plus_one;
    const uint8_t *d_update_channels, *d_zero_grads, *d_enable_sh;
    const int32_t *d_laziness;
    const int32_t *d_step;  // device step count (replaces step when set)
};

// Segment mode from the host flags or the device-resident switches.
// Adam touches every parameter, gradient and moment once per step (1.8 GB at
// 1M Gaussians): non-temporal loads, and non-temporal stores of the gradients
// and moments, keep that stream from evicting the BVH and snapshot data that
// the refit reads next.  The parameters are stored normally: the refit reads
// them.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float *p) {
    const f4v x = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
    return make_float4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void store_stream(const float4 &v, float *p) {
    const f4v x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<f4v *>(p));
}

__device__ __forceinline__ int step_of(const AdamArgs &A) { return A.d_step ? *A.d_step : A.step; }

__device__ __forceinline__ int seg_mode(const AdamArgs &A, int t, bool &zero) {
    const bool upd_ch = A.d_update_channels ? *A.d_update_channels != 0 : (A.host_flags & 1) != 0;
    zero = A.d_zero_grads ? *A.d_zero_grads != 0 : (A.host_flags & 2) != 0;
    bool upd_sh;
    if (A.d_enable_sh) {
        const int lazy = A.d_laziness ? max(*A.d_laziness, 1) : 1;
        upd_sh = *A.d_enable_sh != 0 && (step_of(A) % lazy) == 0;
    } else {
        upd_sh = (A.host_flags & 4) != 0;
    }
    if (A.numel[t] <= 0) return 0;
    if (t == 4) return upd_ch ? 1 : (zero ? 2 : 0);  // channels (adam.cu:35-40)
    if (t >= 5) return upd_sh ? 1 : 0;               // SH coefficients (adam.cu:82-85)
    return 1;
}

__device__ __forceinline__ void adam1(float &p, float &m, float &v, float &g, float lr, float b1, float b2, float bc1,
                                      float bc2, float eps, bool zero) {
    m = b1 * m + (1.0f - b1) * g;
    v = b2 * v + (1.0f - b2) * g * g;
    const float mh = m / bc1;
    const float vh = v / bc2;
    p -= lr * mh / (sqrtf(vh) + eps);
    if (zero) g *= 0.0f;
}

__global__ __launch_bounds__(kThreads) void k_adam(AdamArgs A) {
    const int b = blockIdx.x;
    if (b >= A.block_start[7]) {
        // pruning statistics: counter[i] += was_visible[i]
        const int i = (b - A.block_start[7]) * kThreads + threadIdx.x;
        if (A.visible && A.counter && i < A.count && A.visible[i]) A.counter[i] += 1;
        return;
    }
    int t = 0;
#pragma unroll
    for (int k = 1; k < 7; ++k)
        if (b >= A.block_start[k]) t = k;
    bool zero;
    const int mode = seg_mode(A, t, zero);
    if (mode == 0) return;
    const int64_t n = A.numel[t];
    const int64_t base = (int64_t)(b - A.block_start[t]) * kFloatsPerBlock;
    float *__restrict__ P = A.p[t];
    float *__restrict__ G = A.g[t];
    float *__restrict__ M = A.m[t];
    float *__restrict__ V = A.v[t];
    if (mode == 2) {
        for (int j = 0; j < kVecPerThread * 4; ++j) {
            const int64_t i = base + (int64_t)j * kThreads + threadIdx.x;
            if (i < n) G[i] = 0.0f;  // adam.cu:35: the channel gradient is set to +0
        }
        return;
    }
    const float lr = *A.lr[t];
    const float b1 = *A.beta1, b2 = *A.beta2, eps = *A.eps;
    // adam.cu:16-17: powf(beta, step + 1.0f) with the step converted to float first
    const float step1 = A.d_step ? (float)*A.d_step + 1.0f : A.step_plus_one;
    const float bc1 = 1.0f - powf(b1, step1);
    const float bc2 = 1.0f - powf(b2, step1);
#pragma unroll
    for (int j = 0; j < kVecPerThread; ++j) {
        const int64_t i = base + ((int64_t)j * kThreads + threadIdx.x) * 4;
        if (i + 3 < n) {
            float4 p = load_stream(P + i);
            float4 g = load_stream(G + i);
            float4 m = load_stream(M + i);
            float4 v = load_stream(V + i);
            const float4 g0 = g;
            adam1(p.x, m.x, v.x, g.x, lr, b1, b2, bc1, bc2, eps, zero);
            adam1(p.y, m.y, v.y, g.y, lr, b1, b2, bc1, bc2, eps, zero);
            adam1(p.z, m.z, v.z, g.z, lr, b1, b2, bc1, bc2, eps, zero);
            adam1(p.w, m.w, v.w, g.w, lr, b1, b2, bc1, bc2, eps, zero);
            *reinterpret_cast<float4 *>(P + i) = p;
            store_stream(m, M + i);
            store_stream(v, V + i);
            // zeroing leaves a gradient that is already +-0 bit for bit as it
            // is (x * 0 keeps the sign of a zero): the store is skipped for the
            // Gaussians no ray reached, about two thirds of a training frame's
            const bool same = __float_as_uint(g.x) == __float_as_uint(g0.x) &&
                              __float_as_uint(g.y) == __float_as_uint(g0.y) &&
                              __float_as_uint(g.z) == __float_as_uint(g0.z) &&
                              __float_as_uint(g.w) == __float_as_uint(g0.w);
            if (zero && !same) store_stream(g, G + i);
        } else {
            for (int64_t k = i; k < n; ++k) {
                float p = P[k], g = G[k], m = M[k], v = V[k];
                adam1(p, m, v, g, lr, b1, b2, bc1, bc2, eps, zero);
                P[k] = p;
                M[k] = m;
                V[k] = v;
                if (zero) G[k] = g;
            }
        }
    }
}

}  // namespace

extern "C" int gray_adam_step(gray_ctx *ctx, const gray_adam *adam, int32_t step, int32_t flags, void *stream) {
    if (!ctx || !adam || !adam->beta1 || !adam->beta2 || !adam->epsilon) {
        gray_set_error(GRAY_ERR_INVALID, "adam_step: null argument");
        return GRAY_ERR_INVALID;
    }
    AdamArgs A{};
    int32_t blocks = 0;
    for (int t = 0; t < 7; ++t) {
        A.p[t] = adam->param[t];
        A.g[t] = adam->grad[t];
        A.m[t] = adam->m1[t];
        A.v[t] = adam->m2[t];
        A.lr[t] = adam->lr[t];
        A.numel[t] = adam->numel[t] > 0 ? adam->numel[t] : 0;
        if (A.numel[t] > 0 && (!A.p[t] || !A.g[t] || !A.m[t] || !A.v[t] || !A.lr[t])) {
            gray_set_error(GRAY_ERR_INVALID, "adam_step: tensor %d has a null pointer", t);
            return GRAY_ERR_INVALID;
        }
        const float *ptrs[4] = {A.p[t], A.g[t], A.m[t], A.v[t]};
        for (int k = 0; k < 4 && A.numel[t] > 0; ++k) {
            if (((uintptr_t)ptrs[k]) & 15u) {
                gray_set_error(GRAY_ERR_INVALID, "adam_step: tensor %d is not 16-byte aligned", t);
                return GRAY_ERR_INVALID;
            }
        }
        A.block_start[t] = blocks;
        blocks += (int32_t)gray_div_up(A.numel[t], kFloatsPerBlock);
    }
    A.block_start[7] = blocks;
    A.visible = adam->was_visible;
    A.counter = adam->pruning_counter;
    A.count = adam->count;
    if (A.visible && A.counter && A.count > 0) blocks += (int32_t)gray_div_up(A.count, kThreads);
    A.block_start[8] = blocks;
    A.beta1 = adam->beta1;
    A.beta2 = adam->beta2;
    A.eps = adam->epsilon;
    A.host_flags = flags;
    A.step = step;
    A.step_plus_one = (float)step + 1.0f;
    A.d_update_channels = adam->update_channels;
    A.d_zero_grads = adam->zero_grads;
    A.d_enable_sh = adam->enable_sh;
    A.d_laziness = adam->sh_update_laziness;
    A.d_step = adam->step_device;
    if (blocks == 0) return GRAY_OK;
    hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(kThreads), 0, (hipStream_t)stream, A);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

namespace {

// Streaming copy of n16 16-byte vectors with non-temporal loads and stores,
// grid-stride: the device's sustained read + write rate for a kernel that
// moves each byte once (the reference point of the streaming kernels).
__global__ __launch_bounds__(256) void k_stream_copy(const f4v *__restrict__ src, f4v *__restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += 4 * stride) {
        f4v v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * stride < n16) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * stride < n16) __builtin_nontemporal_store(v[k], dst + i + k * stride);
    }
}

}  // namespace

extern "C" int gray_debug_stream_copy(void *dst, const void *src, int64_t bytes, void *stream) {
    if (!dst || !src || bytes < 0 || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15)) {
        gray_set_error(GRAY_ERR_INVALID, "stream_copy: null pointer, or size/alignment not a multiple of 16 bytes");
        return GRAY_ERR_INVALID;
    }
    const int64_t n16 = bytes / 16;
    if (n16 == 0) return GRAY_OK;
    const int64_t blocks = std::min<int64_t>((n16 + 1023) / 1024, 256 * 64);
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const f4v *>(src), reinterpret_cast<f4v *>(dst), n16);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}
