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
                    ctx->wide,     ctx->wide_flag};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->fev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
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
