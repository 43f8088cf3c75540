// Shared device/host definitions of the gfx950 Gaussian ray tracer.
//
// The per-(ray, Gaussian) hit criterion below is the contract that the CPU
// oracle (oracle/gray_oracle.c) restates operation by operation: it uses only
// IEEE-754 single-precision + - * / in a fixed order with FMA contraction
// disabled, so a given (ray, snapshot) pair yields bit-identical hit sets,
// distances and order on the GPU and on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gray_hip.h"

#define GRAY_WAVE 64
#define GRAY_TILE 8             // 8x8 rays per wave64
#define GRAY_KB 16              // k-buffer entries per ray; divides the reference batch (32)
// Traversal stack entries per lane (LDS).  Every internal node of the Karras
// tree splits its range on a strictly longer common prefix of the 30-bit
// Morton code extended by the 32-bit index, so the depth is at most
// 30 + ceil(log2 N) + 1 and a near-first stack never holds more than depth - 1
// entries: 64 suffices for N < 2^33.
#define GRAY_STACK 64
#define GRAY_MAX_VIEWS 8        // views per launch (kernel-argument table)
#define GRAY_LEAF 0x80000000u   // child reference flag: leaf, payload = Gaussian id
#define GRAY_EMPTY 0xFFFFFFFFu  // child reference: nothing

// Per-Gaussian geometry snapshot (64 B: four 16-byte loads per leaf test).
// m = world-to-local linear part W2L = diag(1/(s*clip*gsf)) * R^T, row-major;
// op = sigmoid(opacity) (or -1 when invisible), clip = adaptive clamp factor
// (reference kernel.cu:3-22).
struct __attribute__((aligned(16))) GeoRecord {
    float m[9];
    float mu[3];
    float op;
    float clip;
    float pad0, pad1;
};

// BVH2 internal node with both child boxes inline (64 B: four 16-byte loads).
struct __attribute__((aligned(16))) BvhNode {
    float lb[6];  // left child box: min xyz, max xyz
    float rb[6];  // right child box
    uint32_t left, right;  // child refs: internal index, or GRAY_LEAF|gaussian id
    uint32_t pad0, pad1;
};

// 4-wide node of the collect pass (64 B: four 16-byte loads).  It collapses a
// binary node at even depth with its internal children: the children are the
// grandchildren reached through internal children and the node's own leaf
// children.  Child boxes are stored by coordinate as half-precision pairs
// rounded outward (min toward -inf, max toward +inf), so the cull stays
// conservative; an unused child has ref GRAY_EMPTY and an empty box.
struct __attribute__((aligned(64))) Bvh4Node {
    uint32_t lo[3][2];  // min x, y, z: children (0,1) and (2,3) as half pairs, child 0/2 in the low half
    uint32_t hi[3][2];  // max x, y, z
    uint32_t child[4];
};
static_assert(sizeof(Bvh4Node) == 64, "Bvh4Node is 64 bytes");

// Device copy of gray_knobs (device pointers to 1-element tensors).
struct DevKnobs {
    const float *exp_power, *alpha_threshold, *t_threshold, *gsf, *eps_fwd, *eps_scale_grad, *bg;
    const uint8_t *render_depth, *enable_sh;
    const uint8_t *jitter;  // may be NULL
    const int32_t *calls;   // may be NULL
    const uint8_t *needs_ray;  // may be NULL
};

// Sub-pixel jitter stream of the reference camera (random.h, camera.h:17-36):
// the seed is the Tiny Encryption Algorithm (Wheeler and Needham, 1994) with
// 4 rounds on (pixel, call index) under the key of its published test, the
// offsets are draws of the Numerical Recipes LCG (a = 1664525,
// c = 1013904223), 24 bits / 2^24.
__host__ __device__ __forceinline__ uint32_t gray_tea4(uint32_t v0, uint32_t v1) {
    uint32_t sum = 0;
    for (int n = 0; n < 4; ++n) {
        sum += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

__host__ __device__ __forceinline__ float gray_lcg_unit(uint32_t &state) {
    state = 1664525u * state + 1013904223u;
    return (float)(state & 0x00FFFFFFu) / (float)0x01000000;
}

struct DevView {
    const float *origin, *c2w, *fov, *znear, *zfar, *rays_o, *rays_d, *rgb;
    const uint8_t *rays_en;
    float *out_c, *out_depth, *out_T, *out_F, *out_est, *out_ro, *out_rd;
    int32_t *stat_hits, *stat_acc;
    uint32_t *hit_ids;
    float *hit_T;
    uint32_t *n_comp;
    float *comp_rgb;
    const float *grad_out;
    float *grad_rgb;
    int32_t width, height, stride, alloc_height;
    int32_t tiles_x, tile_base, ray_count, pad;
};

struct TraceArgs {
    DevKnobs knobs;
    DevView views[GRAY_MAX_VIEWS];
    int32_t num_views, total_tiles, save, capacity;
    int32_t use_bvh, num_gaussians, xcd_chunk, pad;
    const GeoRecord *geo;
    const BvhNode *nodes;
    const float *l2w;          // [N,12] local-to-world 3x3 (row-major, padded) of the snapshot
    // backward-only
    const float *rotation, *scale, *opacity;  // live parameters
    float *g_mean, *g_rot, *g_scale, *g_op, *g_prune;
    uint8_t *visible;
    // debug hit lists
    uint32_t *dbg_ids;
    float *dbg_dist;
    int32_t *dbg_count;
    int32_t dbg_max, pad2;
    int32_t *error_flag;       // set to a GRAY_ERR_* code on device-side failure
    int32_t *tile_stats;       // optional [tiles, 4]: node visits, leaf tests, rounds, hits
    int32_t tile_base_launch, pad3;  // first tile index of this launch (views beyond 8 are launched in chunks)
    // hit lists of the collect pass, slot-major records {dist, id bits, alpha, 0}:
    // entry k of lane g at k*list_stride+g, g = launch tile * 64 + lane;
    // list_n[g] = hits, < 0 = incomplete (the ray is re-traversed)
    float4 *list_rec;
    int32_t *list_n;
    uint8_t *tile_open;        // [tiles] 1: a ray of the tile handed work on as tasks (its lists are not final)
    int64_t list_stride;
    int32_t list_cap;
    int32_t budget;            // iterations of a collect walk before it hands its remaining work on
    // continuation tasks of the collect pass, {lane g, node or leaf reference}
    uint2 *task_in, *task_out;
    int32_t *task_count_in, *task_count_out, *task_head;
    int32_t task_cap;
    int32_t fp16;  // fp16_hits: distances and compositing alphas rounded to fp16 as the reference stores them
    // tile bins of the collect pass (collect_mode 0): bin_tiles holds per
    // launch tile its pair count [T], then the fill cursors [T + 1] (after the
    // fill: the end of the tile's pairs), then the first work item of each tile
    // [T + 1]; bins the Gaussian ids grouped by tile; bin_items a (tile, chunk)
    // pair per work item; *bin_mode 0 while the bins serve the launch, 1 when
    // it falls back to the walk (explicit rays, or more pairs than bin_cap)
    uint32_t *bins;
    int32_t *bin_tiles;
    int2 *bin_items;
    int32_t *bin_mode;
    int64_t bin_cap;
    int32_t bin_item_cap, pad_bins;
    // rays whose list is composited by a whole wave: lists longer than
    // wave_list_min (< 0: none), queued by the per-lane compositing kernel
    int32_t *heavy_count;
    uint32_t *heavy_rays;
    // lists on: set by the per-lane compositing kernel when some ray's list
    // overflowed, so the re-traversal kernel has work (NULL: lists off)
    int32_t *retrace_any;
    int32_t wave_list_min, pad5;
    // deterministic gradient reduction (one view per launch): the backward
    // writes the terms of the composited hit with index k of pixel pix to
    // record det_off[pix] + k of det_rec (16 floats: the gradient terms, then
    // the Gaussian id bits) and the Gaussian id to det_keys
    float *det_rec;
    uint32_t *det_keys;
    const uint32_t *det_off;
    // stored-list backward: per-Gaussian rows of raw geometric sums (16
    // floats: ga*g, pruning weight, v 3, S 6, pad) that the tiles' flushes add
    // to, and a byte per Gaussian marking a touched row (k_raw_apply turns the
    // rows into gradients and clears both)
    float *raw;
    uint8_t *raw_touched;
    // per launch tile: set by the stored-list backward when the tile held more
    // distinct Gaussians than its table; the recompute kernel then redoes the
    // tile's stored lists hit by hit and clears the mark
    uint8_t *bwd_overflow;
    // stored-list backward: each tile's per-Gaussian sums as 64-byte records
    // (colour 3, ga*g, pruning weight, v 3, S 6, view index), indexed tile *
    // 128 + slot; per Gaussian the count of its records and the first
    // kRecList record indices (later ones go to the raw row by atomics)
    float *bwd_rec;
    uint32_t *bwd_cnt;
    uint32_t *bwd_lst;
};

// ---------------------------------------------------------------------------
// Exact ray / ellipsoid-proxy test (reference shaders.cu:32-47).
//
// (o, d) world ray; m, mu snapshot.  o' = W2L (o - mu), d' = W2L d.  Reject if
// o'.d' > 0 (closest approach behind the origin).  The closest approach of the
// line to the proxy centre is at world parameter t = -(o'.d')/(d'.d') (the
// reference's t_local/|d'| with t_local = -o'.(d'/|d'|)), local point
// x = o' + t d'.  Reject if |x|^2 > 1.  Returns t (the reference "distance").
// ---------------------------------------------------------------------------
// The ray origin in the Gaussian's local frame, o' = M (o - mu): the first
// half of the hit test.  Rays that share an origin (a pinhole view) share it.
__device__ __forceinline__ void gray_local_origin(const float *__restrict__ m, float mux, float muy, float muz,
                                                  float ox, float oy, float oz, float &lox, float &loy,
                                                  float &loz) {
#pragma clang fp contract(off)
    const float px = ox - mux, py = oy - muy, pz = oz - muz;
    lox = (m[0] * px + m[1] * py) + m[2] * pz;
    loy = (m[3] * px + m[4] * py) + m[5] * pz;
    loz = (m[6] * px + m[7] * py) + m[8] * pz;
}

// The second half, from o': the same operations in the same order, so the
// split test and gray_hit_test agree bit for bit.
__device__ __forceinline__ bool gray_hit_test_local(const float *__restrict__ m, float lox, float loy, float loz,
                                                    float dx, float dy, float dz, float &dist, float &hx,
                                                    float &hy, float &hz) {
#pragma clang fp contract(off)
    const float ldx = (m[0] * dx + m[1] * dy) + m[2] * dz;
    const float ldy = (m[3] * dx + m[4] * dy) + m[5] * dz;
    const float ldz = (m[6] * dx + m[7] * dy) + m[8] * dz;
    const float od = (lox * ldx + loy * ldy) + loz * ldz;
    if (od > 0.0f) return false;
    const float dd = (ldx * ldx + ldy * ldy) + ldz * ldz;
    const float t = (-od) / dd;
    hx = lox + t * ldx;
    hy = loy + t * ldy;
    hz = loz + t * ldz;
    const float sq = (hx * hx + hy * hy) + hz * hz;
    if (sq > 1.0f) return false;
    dist = t;
    return true;
}

__device__ __forceinline__ bool gray_hit_test(const float *__restrict__ m, float mux, float muy, float muz,
                                              float ox, float oy, float oz, float dx, float dy, float dz,
                                              float &dist, float &hx, float &hy, float &hz) {
    float lox, loy, loz;
    gray_local_origin(m, mux, muy, muz, ox, oy, oz, lox, loy, loz);
    return gray_hit_test_local(m, lox, loy, loz, dx, dy, dz, dist, hx, hy, hz);
}

// Display test of the proxy ellipsoid (the unit sphere in local coordinates),
// reference ellipsoid_viewer.cu:21-66: stable closest-approach form, reject
// when the closest approach lies behind the origin, the entry point if it lies
// in [tlo, thi] else the exit point if it does.  t is world-space; n is the
// (unnormalised) local surface point.
__device__ __forceinline__ bool gray_ellipsoid_hit(const float *__restrict__ m, float mux, float muy, float muz,
                                                   float ox, float oy, float oz, float dx, float dy, float dz,
                                                   float tlo, float thi, float &t, float &nx, float &ny, float &nz) {
#pragma clang fp contract(off)
    const float px = ox - mux, py = oy - muy, pz = oz - muz;
    const float lox = (m[0] * px + m[1] * py) + m[2] * pz;
    const float loy = (m[3] * px + m[4] * py) + m[5] * pz;
    const float loz = (m[6] * px + m[7] * py) + m[8] * pz;
    const float ldx = (m[0] * dx + m[1] * dy) + m[2] * dz;
    const float ldy = (m[3] * dx + m[4] * dy) + m[5] * dz;
    const float ldz = (m[6] * dx + m[7] * dy) + m[8] * dz;
    if ((lox * ldx + loy * ldy) + loz * ldz > 0.0f) return false;
    const float norm = sqrtf((ldx * ldx + ldy * ldy) + ldz * ldz);
    const float ux = ldx / norm, uy = ldy / norm, uz = ldz / norm;
    const float smid = -((lox * ux + loy * uy) + loz * uz);
    const float cx = lox + smid * ux, cy = loy + smid * uy, cz = loz + smid * uz;
    const float sq = (cx * cx + cy * cy) + cz * cz;
    if (sq > 1.0f) return false;
    const float half = sqrtf(1.0f - sq);
    const float t0 = (smid - half) / norm, t1 = (smid + half) / norm;
    if (t0 >= tlo && t0 <= thi)
        t = t0;
    else if (t1 >= tlo && t1 <= thi)
        t = t1;
    else
        return false;
    const float tn = t * norm;
    nx = lox + tn * ux;
    ny = loy + tn * uy;
    nz = loz + tn * uz;
    return true;
}

// Generalised Gaussian kernel (reference kernel.cu:24-38) evaluated at the
// local hit scaled by the clip factor; returns alpha (kernel.cu:40-42).
// q_out = |x*clip|^2 and g_out = kernel value are returned for the backward.
__device__ __forceinline__ float gray_alpha(float hx, float hy, float hz, float clip, float op, float p,
                                            float &g_out, float &q_out) {
#pragma clang fp contract(off)
    const float lx = hx * clip, ly = hy * clip, lz = hz * clip;
    const float q = (lx * lx + ly * ly) + lz * lz;
    float g;
    if (p == 1.0f) {
        g = __expf(-0.5f * q);
    } else if (p == 2.0f) {
        g = __expf((-0.25f * q) * q);
    } else if (p == 3.0f) {
        g = __expf(-((q * q) * q) / 6.0f);
    } else {
        g = __expf(-__powf(q, p) / (2.0f * p));
    }
    g_out = g;
    q_out = q;
    return (GRAY_MAX_ALPHA * g) * op;
}

// Adaptive clamp factor (reference kernel.cu:3-22).
__device__ __forceinline__ float gray_clip_factor(float op, float athr, float p) {
    if (op <= athr) return 0.0f;
    const float lr = logf(op / athr);
    if (p == 1.0f) return sqrtf(2.0f * lr);
    if (p == 2.0f) return sqrtf(sqrtf(4.0f * lr));
    if (p == 3.0f) return cbrtf(sqrtf(6.0f * lr));
    const float k = 2.0f * p;
    return powf(k * lr, 1.0f / k);
}

__device__ __forceinline__ float gray_sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }

// Host-side error reporting shared by all translation units.
void gray_set_error(int code, const char *fmt, ...);
int gray_check_hip(hipError_t e, const char *what);

#define GRAY_HIP(call)                                                   \
    do {                                                                 \
        hipError_t _e = (call);                                          \
        if (_e != hipSuccess) return gray_check_hip(_e, #call);          \
    } while (0)

static inline unsigned gray_div_up(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
