// Fused L1 + SSIM training loss, forward and backward (gfx950).
//
// Reference: train.py:255-260 computes
//     loss = (1 - lambda_ssim) * l1_loss(render, target) + lambda_ssim * (1 - fused_ssim(render, target))
// in PyTorch plus the third-party fused_ssim extension (git a7c48d6), then
// gray/raytracer.py:175-179 copies the autograd gradient of the [C,H,W]
// render into the [H,W,C] framebuffer.  Here both terms and their gradient
// are two kernels over strided planes, so the gradient is written straight
// into the framebuffer layout.
//
// SSIM is the windowed statistic of Wang et al. (2004) as 3D Gaussian
// Splatting evaluates it: an 11-tap separable Gaussian window (sigma 1.5,
// normalised to sum 1), zero padding at the image border ("same"),
// C1 = 0.01^2, C2 = 0.03^2, mean over every pixel of every plane; "valid"
// averages only the pixels at least 5 from the border.  Per pixel p, with
// mu = G*x, E11 = G*x^2, E12 = G*(x y) and s11 = E11 - mu1^2, s12 = E12 - mu1 mu2:
//     m = (2 mu1 mu2 + C1)(2 s12 + C2) / ((mu1^2 + mu2^2 + C1)(s11 + s22 + C2))
// and the gradient of mean(m) over the N averaged pixels w.r.t. x at q is
//     (1/N) sum_p G(p - q) [A(p) + 2 x(q) B(p) + y(q) C(p)]
// with A = dm/dmu1 (E held), B = dm/dE11, C = dm/dE12.
//
// Pass 1 (k_loss_forward): one 16x16 tile of up to three planes (the channels
// of an image) per workgroup; the tile and its 5-pixel halo of x and y are
// staged in LDS with every load in flight at once, the five windowed
// moments are formed by a horizontal then a vertical 11-tap pass, and the
// workgroup writes its partial sums of |x - y| and m plus (when training) the
// maps A, B, C.  Pass 2 (k_loss_backward): the same tiling over the three
// maps, which are convolved with the (symmetric) window and combined with
// x, y and sign(x - y).  k_loss_finish reduces the partial sums in double in
// a fixed order, so the loss is deterministic.

#include <math.h>

#include <mutex>

#include "ctx.hpp"

// The workspace of gray_loss_noalloc (see static_workspace below).
static __device__ __align__(256) unsigned char scratch[GRAY_LOSS_NOALLOC_MAX_BYTES];

namespace {

constexpr int kTile = 16;                 // output pixels per side of a workgroup tile
constexpr int kRadius = 5;                // window radius (11 taps)
constexpr int kTaps = 2 * kRadius + 1;
constexpr int kSpan = kTile + 2 * kRadius;  // 26: tile plus halo
constexpr int kThreads = kTile * kTile;   // 256, one output pixel per thread
constexpr int kGroup = 3;                 // planes per workgroup (the C = 3 of an image)
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;

struct LossArgs {
    const float *x, *y;
    float *grad;
    float *maps;       // [3][P][H][W]: A, B, C
    float *partials;   // [groups * tiles][2]: sum |x - y|, sum m
    float *loss;       // {loss, mean |x - y|, mean SSIM}
    const float *scale_dev;
    int64_t xs[3], ys[3], gs[3];  // plane, row, column strides (elements)
    int64_t plane_px;  // H * W
    int32_t P, H, W, valid, train;
    int32_t tiles_x, tiles, xcd_chunk, groups;
    float win[kTaps];
    float lambda, scale_host;
    float w_l1, w_ssim;  // gradient weights: (1 - lambda) / N_l1, -lambda / N_ssim
    double n_l1, n_ssim;
};

__device__ __forceinline__ bool in_ssim_region(const LossArgs &A, int gx, int gy) {
    if (gx < 0 || gy < 0 || gx >= A.W || gy >= A.H) return false;
    if (!A.valid) return true;
    return gx >= kRadius && gy >= kRadius && gx < A.W - kRadius && gy < A.H - kRadius;
}

// Tile of workgroup b (XCD-aware: the workgroups one XCD receives, b = xcd + 8k,
// take a contiguous run of tiles so neighbouring halos share that XCD's L2).
__device__ __forceinline__ int tile_of(const LossArgs &A) {
    const int b = blockIdx.x;
    return (b & 7) * A.xcd_chunk + (b >> 3);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Stages the span (tile + halo) of up to kGroup planes of `src` into lds.
// Every lane issues all of its loads before the first LDS write: addresses are
// clamped into the image (and plane indices into [0, P)), values outside the
// image are replaced by zero afterwards, so no load is conditional.
template <int kMaps>
__device__ __forceinline__ void stage_span(float (*lds)[kSpan][kSpan + 1], const float *const *src,
                                           const int64_t *plane_stride, const int64_t *row_stride,
                                           const int64_t *col_stride, int g0, int P, int W, int H, int x0, int y0,
                                           int tid) {
    constexpr int kIters = (kSpan * kSpan + kThreads - 1) / kThreads;  // 3
    float v[kIters][kMaps][kGroup];
    bool in[kIters];
#pragma unroll
    for (int j = 0; j < kIters; ++j) {
        const int i = tid + j * kThreads;
        const int ii = i < kSpan * kSpan ? i : 0;
        const int r = ii / kSpan, c = ii - r * kSpan;
        const int gx = x0 + c, gy = y0 + r;
        in[j] = i < kSpan * kSpan && gx >= 0 && gy >= 0 && gx < W && gy < H;
        const int cx = min(max(gx, 0), W - 1), cy = min(max(gy, 0), H - 1);
#pragma unroll
        for (int m = 0; m < kMaps; ++m)
#pragma unroll
            for (int q = 0; q < kGroup; ++q) {
                const int pq = min(g0 + q, P - 1);
                v[j][m][q] = src[m][pq * plane_stride[m] + cy * row_stride[m] + cx * col_stride[m]];
            }
    }
#pragma unroll
    for (int j = 0; j < kIters; ++j) {
        const int i = tid + j * kThreads;
        if (i < kSpan * kSpan) {
            const int r = i / kSpan, c = i - r * kSpan;
#pragma unroll
            for (int m = 0; m < kMaps; ++m)
#pragma unroll
                for (int q = 0; q < kGroup; ++q) lds[m * kGroup + q][r][c] = in[j] ? v[j][m][q] : 0.0f;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_loss_forward(LossArgs A) {
    __shared__ float sxy[2 * kGroup][kSpan][kSpan + 1];  // x of each plane, then y of each plane
    __shared__ float hs[5][kSpan][kTile + 1];
    __shared__ float red[2][kThreads / 64];

    const int tile = tile_of(A);
    if (tile >= A.tiles) return;
    const int g0 = blockIdx.y * kGroup;
    const int np = min(kGroup, A.P - g0);
    const int tid = threadIdx.x;
    const int bx = tile % A.tiles_x, by = tile / A.tiles_x;
    const int x0 = bx * kTile - kRadius, y0 = by * kTile - kRadius;
    {
        const float *src[2] = {A.x, A.y};
        const int64_t ps[2] = {A.xs[0], A.ys[0]}, rs[2] = {A.xs[1], A.ys[1]}, cs[2] = {A.xs[2], A.ys[2]};
        stage_span<2>(sxy, src, ps, rs, cs, g0, A.P, A.W, A.H, x0, y0, tid);
    }
    __syncthreads();

    const int tx = tid % kTile, ty = tid / kTile;
    const int gx = bx * kTile + tx, gy = by * kTile + ty;
    const bool pix = gx < A.W && gy < A.H;
    const bool reg = in_ssim_region(A, gx, gy);
    const int64_t plane3 = (int64_t)A.P * A.plane_px;
    float l1 = 0.0f, ms = 0.0f;
    for (int q = 0; q < np; ++q) {
        float(*sx)[kSpan + 1] = sxy[q];
        float(*sy)[kSpan + 1] = sxy[kGroup + q];
        // horizontal pass: rows of the span, columns of the tile
        for (int i = tid; i < kSpan * kTile; i += kThreads) {
            const int r = i / kTile, c = i - r * kTile;
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, a4 = 0.0f;
#pragma unroll
            for (int k = 0; k < kTaps; ++k) {
                const float xv = sx[r][c + k], yv = sy[r][c + k], g = A.win[k];
                a0 += g * xv;
                a1 += g * yv;
                a2 += g * (xv * xv);
                a3 += g * (yv * yv);
                a4 += g * (xv * yv);
            }
            hs[0][r][c] = a0;
            hs[1][r][c] = a1;
            hs[2][r][c] = a2;
            hs[3][r][c] = a3;
            hs[4][r][c] = a4;
        }
        __syncthreads();

        float mu1 = 0.0f, mu2 = 0.0f, e11 = 0.0f, e22 = 0.0f, e12 = 0.0f;
#pragma unroll
        for (int k = 0; k < kTaps; ++k) {
            const float g = A.win[k];
            mu1 += g * hs[0][ty + k][tx];
            mu2 += g * hs[1][ty + k][tx];
            e11 += g * hs[2][ty + k][tx];
            e22 += g * hs[3][ty + k][tx];
            e12 += g * hs[4][ty + k][tx];
        }
        const float s11 = e11 - mu1 * mu1, s22 = e22 - mu2 * mu2, s12 = e12 - mu1 * mu2;
        const float n1 = 2.0f * mu1 * mu2 + kC1, n2 = 2.0f * s12 + kC2;
        const float d1 = mu1 * mu1 + mu2 * mu2 + kC1, d2 = s11 + s22 + kC2;
        const float inv = 1.0f / (d1 * d2);
        const float m = n1 * n2 * inv;
        if (A.train && pix) {
            float a = 0.0f, b = 0.0f, c = 0.0f;
            if (reg) {
                const float dmu1 = 2.0f * mu2 * n2 * inv - m * 2.0f * mu1 / d1;  // s11, s12 held
                b = -m / d2;                                                     // dm/ds11
                c = 2.0f * n1 * inv;                                             // dm/ds12
                a = dmu1 - 2.0f * mu1 * b - mu2 * c;                             // dm/dmu1, E11 and E12 held
            }
            const int64_t o = (int64_t)(g0 + q) * A.plane_px + (int64_t)gy * A.W + gx;
            A.maps[o] = a;
            A.maps[plane3 + o] = b;
            A.maps[2 * plane3 + o] = c;
        }
        if (pix) l1 += fabsf(sx[ty + kRadius][tx + kRadius] - sy[ty + kRadius][tx + kRadius]);
        if (reg) ms += m;
        if (q + 1 < np) __syncthreads();  // hs is rewritten by the next plane
    }

    l1 = wave_sum(l1);
    ms = wave_sum(ms);
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = l1;
        red[1][tid >> 6] = ms;
    }
    __syncthreads();
    if (tid == 0) {
        const int64_t slot = (int64_t)blockIdx.y * A.tiles + tile;
        A.partials[2 * slot] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        A.partials[2 * slot + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

constexpr int kFinishThreads = 1024;

__global__ __launch_bounds__(kFinishThreads) void k_loss_finish(LossArgs A) {
    __shared__ double red[2][kFinishThreads / 64];
    const int64_t n = (int64_t)A.groups * A.tiles;
    const float2 *__restrict__ part = reinterpret_cast<const float2 *>(A.partials);
    double l1 = 0.0, ms = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kFinishThreads) {
        const float2 v = part[i];
        l1 += (double)v.x;
        ms += (double)v.y;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        l1 += __shfl_xor(l1, o, 64);
        ms += __shfl_xor(ms, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = l1;
        red[1][threadIdx.x >> 6] = ms;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kFinishThreads / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        const double l1m = a / A.n_l1, ssim = b / A.n_ssim;
        const double lam = (double)A.lambda;
        A.loss[0] = (float)((1.0 - lam) * l1m + lam * (1.0 - ssim));
        A.loss[1] = (float)l1m;
        A.loss[2] = (float)ssim;
    }
}

__global__ __launch_bounds__(kThreads) void k_loss_backward(LossArgs A) {
    __shared__ float sm[3 * kGroup][kSpan][kSpan + 1];  // map-major: A of each plane, B of each, C of each
    __shared__ float hs[3][kSpan][kTile + 1];

    const int tile = tile_of(A);
    if (tile >= A.tiles) return;
    const int g0 = blockIdx.y * kGroup;
    const int np = min(kGroup, A.P - g0);
    const int tid = threadIdx.x;
    const int bx = tile % A.tiles_x, by = tile / A.tiles_x;
    const int x0 = bx * kTile - kRadius, y0 = by * kTile - kRadius;
    const int64_t plane3 = (int64_t)A.P * A.plane_px;
    {
        const float *src[3] = {A.maps, A.maps + plane3, A.maps + 2 * plane3};
        const int64_t ps[3] = {A.plane_px, A.plane_px, A.plane_px}, rs[3] = {A.W, A.W, A.W}, cs[3] = {1, 1, 1};
        stage_span<3>(sm, src, ps, rs, cs, g0, A.P, A.W, A.H, x0, y0, tid);
    }
    __syncthreads();

    const int tx = tid % kTile, ty = tid / kTile;
    const int gx = bx * kTile + tx, gy = by * kTile + ty;
    const bool pix = gx < A.W && gy < A.H;
    const float scale = A.scale_host * (A.scale_dev ? *A.scale_dev : 1.0f);
    for (int q = 0; q < np; ++q) {
        for (int i = tid; i < kSpan * kTile; i += kThreads) {
            const int r = i / kTile, c = i - r * kTile;
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
            for (int k = 0; k < kTaps; ++k) {
                const float g = A.win[k];
                a0 += g * sm[q][r][c + k];
                a1 += g * sm[kGroup + q][r][c + k];
                a2 += g * sm[2 * kGroup + q][r][c + k];
            }
            hs[0][r][c] = a0;
            hs[1][r][c] = a1;
            hs[2][r][c] = a2;
        }
        __syncthreads();
        if (pix) {
            float ga = 0.0f, gb = 0.0f, gc = 0.0f;
#pragma unroll
            for (int k = 0; k < kTaps; ++k) {
                const float g = A.win[k];
                ga += g * hs[0][ty + k][tx];
                gb += g * hs[1][ty + k][tx];
                gc += g * hs[2][ty + k][tx];
            }
            const int64_t p = g0 + q;
            const float xv = A.x[p * A.xs[0] + gy * A.xs[1] + gx * A.xs[2]];
            const float yv = A.y[p * A.ys[0] + gy * A.ys[1] + gx * A.ys[2]];
            const float d = xv - yv;
            const float sgn = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
            const float g = A.w_l1 * sgn + A.w_ssim * (ga + 2.0f * xv * gb + yv * gc);
            A.grad[p * A.gs[0] + gy * A.gs[1] + gx * A.gs[2]] = scale * g;
        }
        if (q + 1 < np) __syncthreads();
    }
}

int64_t maps_bytes(int64_t P, int64_t H, int64_t W) { return (3 * P * H * W * 4 + 255) / 256 * 256; }

int64_t tiles_of(int64_t H, int64_t W) { return ((W + kTile - 1) / kTile) * ((H + kTile - 1) / kTile); }

int setup(const gray_loss_args *a, LossArgs &A) {
    if (!a || !a->render || !a->target || !a->workspace) {
        gray_set_error(GRAY_ERR_INVALID, "loss: null args, render, target or workspace");
        return GRAY_ERR_INVALID;
    }
    if (a->planes <= 0 || a->height <= 0 || a->width <= 0 || a->planes > 3 * 65535) {
        gray_set_error(GRAY_ERR_INVALID, "loss: bad shape planes=%d height=%d width=%d", a->planes, a->height,
                       a->width);
        return GRAY_ERR_INVALID;
    }
    if (a->valid_padding && (a->height <= 2 * kRadius || a->width <= 2 * kRadius)) {
        gray_set_error(GRAY_ERR_INVALID, "loss: valid padding needs height and width above %d (got %dx%d)",
                       2 * kRadius, a->width, a->height);
        return GRAY_ERR_INVALID;
    }
    A = LossArgs{};
    A.x = a->render;
    A.y = a->target;
    A.grad = a->grad;
    A.maps = (float *)a->workspace;
    A.partials = (float *)((char *)a->workspace + maps_bytes(a->planes, a->height, a->width));
    A.loss = a->loss;
    A.scale_dev = a->grad_scale_device;
    for (int i = 0; i < 3; ++i) {
        A.xs[i] = a->render_stride[i];
        A.ys[i] = a->target_stride[i];
        A.gs[i] = a->grad_stride[i];
    }
    A.P = a->planes;
    A.H = a->height;
    A.W = a->width;
    A.plane_px = (int64_t)a->height * a->width;
    A.valid = a->valid_padding ? 1 : 0;
    A.train = a->train ? 1 : 0;
    A.tiles_x = (a->width + kTile - 1) / kTile;
    A.tiles = (int32_t)tiles_of(a->height, a->width);
    A.xcd_chunk = (A.tiles + 7) / 8;
    A.groups = (a->planes + kGroup - 1) / kGroup;
    // normalised Gaussian window, sigma 1.5, formed in double
    double w[kTaps], sum = 0.0;
    for (int k = 0; k < kTaps; ++k) {
        const double t = (double)(k - kRadius);
        w[k] = exp(-t * t / (2.0 * 1.5 * 1.5));
        sum += w[k];
    }
    for (int k = 0; k < kTaps; ++k) A.win[k] = (float)(w[k] / sum);
    A.lambda = a->lambda_ssim;
    A.scale_host = a->grad_scale;
    A.n_l1 = (double)a->planes * a->height * a->width;
    const int64_t vh = A.valid ? a->height - 2 * kRadius : a->height;
    const int64_t vw = A.valid ? a->width - 2 * kRadius : a->width;
    A.n_ssim = (double)a->planes * vh * vw;
    A.w_l1 = (float)((1.0 - (double)a->lambda_ssim) / A.n_l1);
    A.w_ssim = (float)(-(double)a->lambda_ssim / A.n_ssim);
    return GRAY_OK;
}

int launch_forward(const LossArgs &A, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_forward, dim3(8 * A.xcd_chunk, A.groups), dim3(kThreads), 0, s, A);
    GRAY_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(kFinishThreads), 0, s, A);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

int launch_backward(const LossArgs &A, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_backward, dim3(8 * A.xcd_chunk, A.groups), dim3(kThreads), 0, s, A);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

int64_t workspace_bytes(int64_t P, int64_t H, int64_t W) {
    return maps_bytes(P, H, W) + 2 * 4 * ((P + kGroup - 1) / kGroup) * tiles_of(H, W);
}

// Workspace of gray_loss_noalloc: one file-scope array per GPU, held from the
// library's load until the process ends.  Host code never uses the name as an
// address (on the host it is only a stand-in for the symbol): the device
// address of each GPU's copy comes from hipGetSymbolAddress with that device
// current, on the first call on that device, and only that address reaches the
// kernels.  Nothing in it has to start at zero: the forward writes every
// partial sum and every derivative map entry that the finish and backward
// kernels read.
constexpr int64_t kStaticWorkspaceBytes = GRAY_LOSS_NOALLOC_MAX_BYTES;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int static_workspace(void **out) {
    int dev = -1;
    GRAY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        gray_set_error(GRAY_ERR_INVALID, "loss_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return GRAY_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return gray_check_hip(e, "loss_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) {
            gray_set_error(GRAY_ERR_HIP, "loss_noalloc: null device address of the static workspace");
            return GRAY_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GRAY_OK;
}

}  // namespace

extern "C" int gray_loss_noalloc(const gray_loss_args *args, void *stream) {
    if (!args) {
        gray_set_error(GRAY_ERR_INVALID, "loss_noalloc: null args");
        return GRAY_ERR_INVALID;
    }
    if (args->workspace) {
        gray_set_error(GRAY_ERR_INVALID, "loss_noalloc: workspace must be NULL (the library holds it)");
        return GRAY_ERR_INVALID;
    }
    if (args->planes <= 0 || args->height <= 0 || args->width <= 0) {
        gray_set_error(GRAY_ERR_INVALID, "loss_noalloc: bad shape planes=%d height=%d width=%d", args->planes,
                       args->height, args->width);
        return GRAY_ERR_INVALID;
    }
    const int64_t need = workspace_bytes(args->planes, args->height, args->width);
    if (need > kStaticWorkspaceBytes) {
        gray_set_error(GRAY_ERR_INVALID,
                       "loss_noalloc: %d planes of %dx%d need %lld workspace bytes, above the static %lld",
                       args->planes, args->width, args->height, (long long)need, (long long)kStaticWorkspaceBytes);
        return GRAY_ERR_INVALID;
    }
    void *ws = nullptr;
    if (int rc = static_workspace(&ws)) return rc;
    gray_loss_args a = *args;
    a.workspace = ws;
    a.train = args->grad ? 1 : 0;
    LossArgs A;
    if (int rc = setup(&a, A)) return rc;
    if (!A.loss) {
        gray_set_error(GRAY_ERR_INVALID, "loss_noalloc: null loss output");
        return GRAY_ERR_INVALID;
    }
    const hipStream_t s = (hipStream_t)stream;
    if (int rc = launch_forward(A, s)) return rc;
    if (A.grad)
        if (int rc = launch_backward(A, s)) return rc;
    return GRAY_OK;
}

extern "C" int gray_loss_workspace_bytes(int32_t planes, int32_t height, int32_t width, int64_t *bytes) {
    if (!bytes || planes <= 0 || height <= 0 || width <= 0) {
        gray_set_error(GRAY_ERR_INVALID, "loss workspace: bad shape or null output");
        return GRAY_ERR_INVALID;
    }
    *bytes = workspace_bytes(planes, height, width);
    return GRAY_OK;
}

extern "C" int gray_loss_forward(const gray_loss_args *args, void *stream) {
    LossArgs A;
    if (int rc = setup(args, A)) return rc;
    if (!A.loss) {
        gray_set_error(GRAY_ERR_INVALID, "loss forward: null loss output");
        return GRAY_ERR_INVALID;
    }
    return launch_forward(A, (hipStream_t)stream);
}

extern "C" int gray_loss_backward(const gray_loss_args *args, void *stream) {
    LossArgs A;
    if (int rc = setup(args, A)) return rc;
    if (!A.grad) {
        gray_set_error(GRAY_ERR_INVALID, "loss backward: null gradient output");
        return GRAY_ERR_INVALID;
    }
    return launch_backward(A, (hipStream_t)stream);
}
