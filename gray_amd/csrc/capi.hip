// C ABI plumbing: context lifetime and error reporting.

#include <stdarg.h>
#include <stdio.h>

#include "ctx.hpp"

namespace {
thread_local char g_error[512] = "";
}

void gray_set_error(int code, const char *fmt, ...) {
    char buf[448];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    snprintf(g_error, sizeof(g_error), "[gray %d] %s", code, buf);
}

int gray_check_hip(hipError_t e, const char *what) {
    gray_set_error(GRAY_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    return GRAY_ERR_HIP;
}

extern "C" const char *gray_last_error(void) { return g_error; }

extern "C" int gray_abi_version(void) { return GRAY_ABI_VERSION; }

extern "C" int gray_ctx_create(gray_ctx **out) {
    if (!out) {
        gray_set_error(GRAY_ERR_INVALID, "ctx_create: null output");
        return GRAY_ERR_INVALID;
    }
    *out = nullptr;
    gray_ctx *ctx = new gray_ctx();
    hipError_t e = hipGetDevice(&ctx->device);
    if (e == hipSuccess) e = hipMalloc((void **)&ctx->err, sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&ctx->top_count, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(ctx->err, 0, sizeof(int32_t));
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    for (hipEvent_t &fe : ctx->fev)
        if (e == hipSuccess) e = hipEventCreate(&fe);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_sh, hipEventDisableTiming);
    if (e != hipSuccess) {
        gray_check_hip(e, "ctx_create");
        gray_ctx_destroy(ctx);
        return GRAY_ERR_HIP;
    }
    *out = ctx;
    return GRAY_OK;
}

extern "C" void gray_ctx_destroy(gray_ctx *ctx) {
    if (!ctx) return;
    void *bufs[] = {ctx->geo,      ctx->l2w,     ctx->aabb,       ctx->nodes,    ctx->parent_int, ctx->parent_leaf,
                    ctx->range_lo, ctx->range_hi, ctx->sorted_ids, ctx->keys,     ctx->keys_alt,   ctx->vals_alt,
                    ctx->hist,     ctx->block_sums, ctx->bounds_partial, ctx->bounds, ctx->prefix, ctx->suffix,
                    ctx->sparse,   ctx->top_nodes, ctx->top_count, ctx->err,    ctx->list_rec,   ctx->list_n,
                    ctx->tasks[0], ctx->tasks[1], ctx->task_counters, ctx->need_ends, ctx->node_info,
                    ctx->wide,     ctx->wide_flag,  ctx->det_rec, ctx->det_keys[0], ctx->det_keys[1],
                    ctx->det_vals[0], ctx->det_vals[1], ctx->det_off, ctx->det_hist, ctx->det_sums,
                    ctx->raw,      ctx->raw_touched, ctx->bwd_overflow, ctx->tile_cost, ctx->tile_order,
                    ctx->tile_next, ctx->bwd_rec, ctx->bwd_cnt, ctx->bwd_lst, ctx->tile_open,
                    ctx->bin_tiles, ctx->bin_items};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->fev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_sh) (void)hipEventDestroy(ctx->ev_sh);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    delete ctx;
}

extern "C" int gray_ctx_share_scene(gray_ctx *ctx, gray_ctx *scene) {
    if (!ctx || ctx == scene) {
        gray_set_error(GRAY_ERR_INVALID, "ctx_share_scene: null context or a context sharing itself");
        return GRAY_ERR_INVALID;
    }
    if (scene && (scene->scene || scene->device != ctx->device)) {
        gray_set_error(GRAY_ERR_INVALID, "ctx_share_scene: the scene must own its BVH and live on the same device");
        return GRAY_ERR_INVALID;
    }
    ctx->scene = scene;
    return GRAY_OK;
}

// One thread writes the camera, flag and counter fields; every thread clears
// 16 bytes of was_visible (bytewise at an unaligned base or the tail).
__global__ void k_frame_setup(gray_frame_args a, int aligned16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        if (a.origin)
            for (int c = 0; c < 3; ++c) a.origin[c] = a.origin_value[c];
        if (a.rotation_c2w)
            for (int c = 0; c < 9; ++c) a.rotation_c2w[c] = a.c2w_value[c];
        if (a.rotation_w2c)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) a.rotation_w2c[3 * r + c] = a.c2w_value[3 * c + r];
        if (a.vertical_fov) *a.vertical_fov = a.fov_value;
        if (a.znear) *a.znear = a.znear_value;
        if (a.zfar) *a.zfar = a.zfar_value;
        if (a.rays_from_python) *a.rays_from_python = 0;
        if (a.grads_enabled) *a.grads_enabled = a.grads_enabled_value != 0 ? 1 : 0;
        if (a.total_num_calls) *a.total_num_calls += 1;
    }
    if (a.was_visible) {
        const int64_t b = i * 16;
        if (aligned16 && b + 16 <= a.count) {
            *reinterpret_cast<uint4 *>(a.was_visible + b) = make_uint4(0u, 0u, 0u, 0u);
        } else {
            for (int64_t j = b; j < b + 16 && j < a.count; ++j) a.was_visible[j] = 0;
        }
    }
}

extern "C" int gray_frame_setup(const gray_frame_args *args, void *stream) {
    if (!args || args->count < 0) {
        gray_set_error(GRAY_ERR_INVALID, "frame_setup: null arguments or negative count");
        return GRAY_ERR_INVALID;
    }
    const int64_t count = args->was_visible ? args->count : 0;
    const int64_t threads = count > 0 ? (count + 15) / 16 : 1;
    const int aligned16 = (((uintptr_t)args->was_visible) & 15u) == 0;
    hipLaunchKernelGGL(k_frame_setup, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       *args, aligned16);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

// Reads (and clears) the device-side error flag; synchronises the stream.
extern "C" int gray_device_status(gray_ctx *ctx, void *stream) {
    if (!ctx) return GRAY_ERR_INVALID;
    int32_t flag = 0;
    hipStream_t s = (hipStream_t)stream;
    GRAY_HIP(hipMemcpyAsync(&flag, ctx->err, sizeof(flag), hipMemcpyDeviceToHost, s));
    GRAY_HIP(hipStreamSynchronize(s));
    if (flag) {
        GRAY_HIP(hipMemsetAsync(ctx->err, 0, sizeof(int32_t), s));
        GRAY_HIP(hipStreamSynchronize(s));
        if (flag & (1 << GRAY_ERR_OVERFLOW)) {
            gray_set_error(GRAY_ERR_OVERFLOW, "device: traversal stack overflow (BVH deeper than %d)", GRAY_STACK);
            return GRAY_ERR_OVERFLOW;
        }
        gray_set_error(GRAY_ERR_STATE, "device: error flag 0x%x", flag);
        return GRAY_ERR_STATE;
    }
    return GRAY_OK;
}
