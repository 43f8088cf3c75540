// Gaussian ray tracing on gfx950: forward, backward, debug.
//
// Rays are grouped in 8x8 tiles of one view; a wave64 takes a tile, every lane
// one ray.  The forward runs in passes:
//   collect    every lane walks the LBVH for its own ray and appends every hit
//              (distance, id, alpha) to the ray's list; walks that run long
//              hand their remaining subtrees on as tasks, finished by
//              persistent task rounds (load balance of grazing rays);
//   composite  per ray, the reference batch semantics over the list
//              (forward_pass.cu:150-230): each batch takes the 32 nearest hits
//              strictly beyond tmin (served in sub-rounds of GRAY_KB from a
//              register k-buffer), composites front to back and stops at
//              T < t_threshold;
//   retraverse the same forward with the BVH as the hit source, for rays whose
//              list overflowed (and for every ray when lists are off).
// The backward replays each ray's stored composited list back to front
// (re-traversing rays whose list exceeded the store) and sums the gradients
// of a tile's rays per Gaussian in LDS before the global atomics.
//
// Truncation bookkeeping.  The reference estimates the colour of the hits it
// never composited as (sum_all a*c - sum_comp a*c) / (sum_all a - sum_comp a)
// and the full transmittance as prod_all (1-a) (forward_pass.cu:53-57,
// 206-208, 238-240; shaders.cu:54-57).  The same quantities are formed here
// directly over the non-composited set ("skip set"): the hits the k-buffer
// never handed to compositing, hits beyond zfar and hits after termination.
// F = T * prod_skip (1-a).  This is the same mathematics without the fp32
// cancellation of sum_all - sum_comp.

#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>

#include <hip/hip_fp16.h>

#include <vector>

#include "ctx.hpp"

namespace {

struct LaneRay {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    float one;  // 1.0f, opaque to the compiler (set by make_ray): keeps half_minus a v_fma_mix_f32
};

__device__ __forceinline__ bool slab_hit(const LaneRay &r, float x0, float y0, float z0, float x1, float y1, float z1,
                                         float tlo) {
    const float tx0 = (x0 - r.ox) * r.ix, tx1 = (x1 - r.ox) * r.ix;
    const float ty0 = (y0 - r.oy) * r.iy, ty1 = (y1 - r.oy) * r.iy;
    const float tz0 = (z0 - r.oz) * r.iz, tz1 = (z1 - r.oz) * r.iz;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    // relative slack keeps the cull conservative under rounding; clamped so
    // an infinite slab bound (a zero direction component) cannot make it
    // infinite and accept every box
    const float slack = 1e-5f * fminf(fmaxf(fabsf(tn), fabsf(tf)), 1e30f) + 1e-7f;
    return (x0 <= x1) && (tn <= tf + slack) && (tf + slack >= tlo);
}

__device__ __forceinline__ bool slab_test(const LaneRay &r, const float *b, float tlo) {
    return slab_hit(r, b[0], b[1], b[2], b[3], b[4], b[5], tlo);
}

// XCD-aware block -> tile mapping: workgroup b runs on XCD (b % 8); give each
// XCD a contiguous band of tiles so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ void tile_coords(const TraceArgs &A, int tile, int &view, int &tx, int &ty) {
    int v = 0;
#pragma unroll
    for (int k = 1; k < GRAY_MAX_VIEWS; ++k)
        if (k < A.num_views && tile >= A.views[k].tile_base) v = k;
    view = v;
    const int lt = tile - A.views[v].tile_base;
    tx = lt % A.views[v].tiles_x;
    ty = lt / A.views[v].tiles_x;
}

__device__ __forceinline__ bool block_tile(const TraceArgs &A, int &view, int &tx, int &ty, int *tile_out = nullptr) {
    const int b = blockIdx.x;
    const int tile = (b & 7) * A.xcd_chunk + (b >> 3);
    if (tile >= A.total_tiles) return false;
    if (tile_out) *tile_out = tile;
    int v = 0;
#pragma unroll
    for (int k = 1; k < GRAY_MAX_VIEWS; ++k)
        if (k < A.num_views && tile >= A.views[k].tile_base) v = k;
    view = v;
    const int lt = tile - A.views[v].tile_base;
    tx = lt % A.views[v].tiles_x;
    ty = lt / A.views[v].tiles_x;
    return true;
}

// Primary ray of pixel (x, y): reference Camera::compute_primary_ray_direction
// (camera.h:17-36) without jitter; explicit rays when the view provides them
// (shaders.cu:80-85).  Returns false for an inactive pixel (zero direction).
__device__ __forceinline__ bool make_ray(const DevKnobs &kn, const DevView &V, int x, int y, LaneRay &r) {
#pragma clang fp contract(off)
    const int pix = y * V.stride + x;
    float dx, dy, dz;
    if (V.rays_d && (!V.rays_en || *V.rays_en)) {
        r.ox = V.rays_o[3 * pix + 0];
        r.oy = V.rays_o[3 * pix + 1];
        r.oz = V.rays_o[3 * pix + 2];
        dx = V.rays_d[3 * pix + 0];
        dy = V.rays_d[3 * pix + 1];
        dz = V.rays_d[3 * pix + 2];
    } else {
        r.ox = V.origin[0];
        r.oy = V.origin[1];
        r.oz = V.origin[2];
        const float vs = tanf(*V.fov / 2.0f);
        const float aspect = (float)V.width / (float)V.height;
        float px = (float)x, py = (float)y;
        if (kn.jitter && *kn.jitter) {
            uint32_t seed = gray_tea4((uint32_t)pix, kn.calls ? (uint32_t)*kn.calls : 0u);
            px += gray_lcg_unit(seed) - 0.5f;
            py += gray_lcg_unit(seed) - 0.5f;
        }
        const float yf = vs * (1.0f - (2.0f * (py + 0.5f)) / (float)V.height);
        const float xf = (aspect * vs) * ((2.0f * (px + 0.5f)) / (float)V.width - 1.0f);
        const float *c = V.c2w;  // row-major; direction = col0*x + col1*y - col2
        dx = (c[0] * xf + c[1] * yf) - c[2];
        dy = (c[3] * xf + c[4] * yf) - c[5];
        dz = (c[6] * xf + c[7] * yf) - c[8];
        const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
        dx = dx / len;
        dy = dy / len;
        dz = dz / len;
    }
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    r.ix = 1.0f / dx;
    r.iy = 1.0f / dy;
    r.iz = 1.0f / dz;
    float one = 1.0f;
    asm volatile("" : "+v"(one));
    r.one = one;
    const float l2 = (dx * dx + dy * dy) + dz * dz;
    return l2 != 0.0f;
}

// Node and Gaussian records are fetched with 16-byte vector loads (four per
// 64-byte record).  Vector loads complete in order (vmcnt), so a prefetched
// record does not force a wait on LDS traffic the way an out-of-order scalar
// load (lgkmcnt) would; in_vgpr keeps an address that happens to be
// wave-uniform in a vector register for that reason.
struct NodeV {
    float lb[6], rb[6];
    uint32_t left, right;
};

struct GeoV {
    float m[9], mu[3], op, clip;
};

__device__ __forceinline__ uint32_t in_vgpr(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ NodeV load_node(const BvhNode *__restrict__ nodes, uint32_t idx) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + in_vgpr(idx));
    const float4 a = p[0], b = p[1], c = p[2], d = p[3];
    NodeV n;
    n.lb[0] = a.x, n.lb[1] = a.y, n.lb[2] = a.z, n.lb[3] = a.w, n.lb[4] = b.x, n.lb[5] = b.y;
    n.rb[0] = b.z, n.rb[1] = b.w, n.rb[2] = c.x, n.rb[3] = c.y, n.rb[4] = c.z, n.rb[5] = c.w;
    n.left = __float_as_uint(d.x);
    n.right = __float_as_uint(d.y);
    return n;
}

// A 4-wide node is kept as loaded (16 VGPRs) across the loop and decoded
// where its boxes are tested.
struct Node4V {
    uint4 a, b, c, ch;
};

__device__ __forceinline__ Node4V load_node4(const Bvh4Node *__restrict__ wide, uint32_t idx) {
    const uint4 *q = reinterpret_cast<const uint4 *>(wide + in_vgpr(idx));
    Node4V n;
    n.a = q[0], n.b = q[1], n.c = q[2], n.ch = q[3];
    return n;
}

// Half-precision box bound (low or high half of a word) minus a ray origin
// coordinate: the conversion is exact, the difference rounded once; one
// v_fma_mix_f32 (half operand selected by op_sel) per bound.
template <bool kHigh>
__device__ __forceinline__ float half_minus(uint32_t v, float o, float one) {
    const _Float16 h = __builtin_bit_cast(_Float16, (unsigned short)(kHigh ? (v >> 16) : (v & 0xffffu)));
    return __builtin_fmaf((float)h, one, -o);
}

// Child k's box test; the box words of child k are word k/2 of each pair, its
// half is k%2.  The same slab test as slab_hit (differences to the origin
// times the inverse direction; relative slack, so the cull stays conservative
// under rounding at any distance), written without branches: the slack is
// 1e-5 |tf| + 1e-7 on the exit parameter, and an infinite tf gives an
// infinite (tf = +inf) or NaN (tf = -inf: no hit) slackened exit.
template <int k>
__device__ __forceinline__ bool child_hit(const LaneRay &r, const Node4V &n) {
    constexpr bool h = (k & 1) != 0;
    const uint32_t lx = k < 2 ? n.a.x : n.a.y, ly = k < 2 ? n.a.z : n.a.w, lz = k < 2 ? n.b.x : n.b.y;
    const uint32_t hx = k < 2 ? n.b.z : n.b.w, hy = k < 2 ? n.c.x : n.c.y, hz = k < 2 ? n.c.z : n.c.w;
    const float ax0 = half_minus<h>(lx, r.ox, r.one), ax1 = half_minus<h>(hx, r.ox, r.one);
    const float tx0 = ax0 * r.ix, tx1 = ax1 * r.ix;
    const float ty0 = half_minus<h>(ly, r.oy, r.one) * r.iy, ty1 = half_minus<h>(hy, r.oy, r.one) * r.iy;
    const float tz0 = half_minus<h>(lz, r.oz, r.one) * r.iz, tz1 = half_minus<h>(hz, r.oz, r.one) * r.iz;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    const float tfs = __builtin_fmaf(fabsf(tf), 1e-5f, tf) + 1e-7f;
    // an empty box (min above max, e.g. a subtree of invisible Gaussians) is no hit
    return (ax0 <= ax1) & (tn <= tfs) & (tfs >= 0.0f);
}

__device__ __forceinline__ GeoV load_geo(const GeoRecord *__restrict__ geo, uint32_t gid) {
    const float4 *p = reinterpret_cast<const float4 *>(geo + in_vgpr(gid));
    const float4 a = p[0], b = p[1], c = p[2], d = p[3];
    GeoV g;
    g.m[0] = a.x, g.m[1] = a.y, g.m[2] = a.z, g.m[3] = a.w, g.m[4] = b.x, g.m[5] = b.y, g.m[6] = b.z;
    g.m[7] = b.w, g.m[8] = c.x, g.mu[0] = c.y, g.mu[1] = c.z, g.mu[2] = c.w, g.op = d.x, g.clip = d.y;
    return g;
}

__device__ __forceinline__ void load_rgb(const float *__restrict__ rgb, uint32_t gid, float *c) {
    const float *q = rgb + 3 * (size_t)in_vgpr(gid);
    c[0] = q[0];
    c[1] = q[1];
    c[2] = q[2];
}

// Per-ray traversal.  Every lane walks the LBVH for its own ray with its own
// stack ([level][lane] in LDS, near child first along the ray); the wave
// loops until every lane is done (while-while).  In one iteration a lane
// visits one node, then tests the (up to two) leaves among that node's hit
// children: their records are requested before the next node is, so the
// node load overlaps the leaf tests.  Calls leaf(gid, record, c0, c1, c2) for
// every Gaussian whose box the lane's ray interval [tlo, inf) overlaps (or for
// every Gaussian when use_bvh == 0), c = colour when want_rgb.  Must be
// called by the whole wave.
template <typename Leaf>
__device__ __forceinline__ void traverse(const TraceArgs &A, const BvhNode *__restrict__ nodes,
                                         const GeoRecord *__restrict__ geo, const float *__restrict__ rgb,
                                         bool want_rgb, const LaneRay &r, bool part, float tlo,
                                         uint32_t *__restrict__ stack, int lane, Leaf &&leaf,
                                         int *visits = nullptr) {
    const int n = A.num_gaussians;
    if (n <= 0) return;
    if (!A.use_bvh) {
        for (int g = 0; g < n; ++g) {
            const GeoV rec = load_geo(geo, (uint32_t)g);
            float c[3] = {0.0f, 0.0f, 0.0f};
            if (want_rgb) load_rgb(rgb, (uint32_t)g, c);
            if (part) leaf((uint32_t)g, rec, c[0], c[1], c[2]);
        }
        return;
    }
    uint32_t node = part ? 0u : GRAY_EMPTY;
    NodeV nd;
    if (node != GRAY_EMPTY) nd = load_node(nodes, node);
    int sp = 0;
    const int max_iters = 2 * n + 16;  // a lane visits each node at most once
    for (int it = 0;; ++it) {
        if (!__ballot(node != GRAY_EMPTY)) break;
        if (it > max_iters) {
            if (A.error_flag) atomicOr(A.error_flag, 1 << GRAY_ERR_STATE);
            return;
        }
        uint32_t la = GRAY_EMPTY, lb = GRAY_EMPTY, next = GRAY_EMPTY;
        if (node != GRAY_EMPTY) {
            if (visits) ++visits[0];
            const bool hl = nd.left != GRAY_EMPTY && slab_test(r, nd.lb, tlo);
            const bool hr = nd.right != GRAY_EMPTY && slab_test(r, nd.rb, tlo);
            uint32_t other = GRAY_EMPTY;
            if (hl) {
                if (nd.left & GRAY_LEAF)
                    la = nd.left & ~GRAY_LEAF;
                else
                    next = nd.left;
            }
            if (hr) {
                if (nd.right & GRAY_LEAF)
                    lb = nd.right & ~GRAY_LEAF;
                else if (next == GRAY_EMPTY)
                    next = nd.right;
                else
                    other = nd.right;
            }
            if (other != GRAY_EMPTY) {
                const float cl = ((nd.lb[0] + nd.lb[3]) * r.dx + (nd.lb[1] + nd.lb[4]) * r.dy) +
                                 (nd.lb[2] + nd.lb[5]) * r.dz;
                const float cr = ((nd.rb[0] + nd.rb[3]) * r.dx + (nd.rb[1] + nd.rb[4]) * r.dy) +
                                 (nd.rb[2] + nd.rb[5]) * r.dz;
                if (cr < cl) {
                    const uint32_t t = next;
                    next = other;
                    other = t;
                }
                if (sp < GRAY_STACK) {
                    stack[sp * GRAY_WAVE + lane] = other;
                    ++sp;
                } else if (A.error_flag) {
                    atomicOr(A.error_flag, 1 << GRAY_ERR_OVERFLOW);
                }
            }
            if (next == GRAY_EMPTY && sp > 0) {
                --sp;
                next = stack[sp * GRAY_WAVE + lane];
            }
        }
        // leaf records first, then the next node: the leaf tests wait only on
        // the former (vector loads complete in order)
        GeoV ga, gb;
        float ca[3] = {0.0f, 0.0f, 0.0f}, cb[3] = {0.0f, 0.0f, 0.0f};
        if (la != GRAY_EMPTY) {
            ga = load_geo(geo, la);
            if (want_rgb) load_rgb(rgb, la, ca);
        }
        if (lb != GRAY_EMPTY) {
            gb = load_geo(geo, lb);
            if (want_rgb) load_rgb(rgb, lb, cb);
        }
        node = next;
        if (node != GRAY_EMPTY) nd = load_node(nodes, node);
        if (la != GRAY_EMPTY) {
            if (visits) ++visits[1];
            leaf(la, ga, ca[0], ca[1], ca[2]);
        }
        if (lb != GRAY_EMPTY) {
            if (visits) ++visits[1];
            leaf(lb, gb, cb[0], cb[1], cb[2]);
        }
    }
}

// Sorted per-lane k-buffer in LDS, layout [slot][lane].
struct KBuf {
    float *d;
    uint32_t *id;
    float *a;
    int lane;
    __device__ __forceinline__ float dist(int k) const { return d[k * GRAY_WAVE + lane]; }
    __device__ __forceinline__ uint32_t gid(int k) const { return id[k * GRAY_WAVE + lane]; }
    __device__ __forceinline__ float alpha(int k) const { return a[k * GRAY_WAVE + lane]; }
};

__device__ __forceinline__ bool key_less(float d0, uint32_t i0, float d1, uint32_t i1) {
    return d0 < d1 || (d0 == d1 && i0 < i1);
}

// Insert (dist, id, alpha).  Returns 0 = rejected, 1 = inserted, 2 = inserted
// and the previous last entry (ev_id, ev_a) was evicted.
__device__ __forceinline__ int kb_insert(const KBuf &kb, int &count, float &lastd, uint32_t &lasti, float d,
                                         uint32_t gid, float a, uint32_t &ev_id, float &ev_a) {
    int pos;
    int res = 1;
    if (count == GRAY_KB) {
        if (!key_less(d, gid, lastd, lasti)) return 0;
        ev_id = kb.gid(GRAY_KB - 1);
        ev_a = kb.alpha(GRAY_KB - 1);
        pos = GRAY_KB - 1;
        res = 2;
    } else {
        pos = count;
        ++count;
    }
    while (pos > 0) {
        const float pd = kb.dist(pos - 1);
        const uint32_t pi = kb.gid(pos - 1);
        if (!key_less(d, gid, pd, pi)) break;
        kb.d[pos * GRAY_WAVE + kb.lane] = pd;
        kb.id[pos * GRAY_WAVE + kb.lane] = pi;
        kb.a[pos * GRAY_WAVE + kb.lane] = kb.alpha(pos - 1);
        --pos;
    }
    kb.d[pos * GRAY_WAVE + kb.lane] = d;
    kb.id[pos * GRAY_WAVE + kb.lane] = gid;
    kb.a[pos * GRAY_WAVE + kb.lane] = a;
    lastd = kb.dist(count - 1);
    lasti = kb.gid(count - 1);
    return res;
}

// ---------------------------------------------------------------------------
// Collect pass: every hit of every ray into its list
// ---------------------------------------------------------------------------
// Tuning of the collect pass (compile-time; -D overrides for sweeps).
// Stack entries per lane (LDS).  When it is full, a walk spills the entry as
// a task; the last task round (pooled, below) falls back to re-traversal when
// its stack overflows, so the bounds cost speed, never results.
#ifndef GRAY_COLLECT_STACK
#define GRAY_COLLECT_STACK 16
#endif
// Leaves tested per lane and iteration.
#ifndef GRAY_COLLECT_LEAVES
#define GRAY_COLLECT_LEAVES 1
#endif
// Leaf queue per lane (LDS ring, a power of two >= 4): leaves found by node
// visits wait here for their exact test.
#ifndef GRAY_LEAF_QUEUE
#define GRAY_LEAF_QUEUE 8
#endif
// Iterations the per-tile walk runs before it hands its remaining work on as
// tasks.
#ifndef GRAY_COLLECT_BUDGET
#define GRAY_COLLECT_BUDGET 128
#endif
// Pooled task rounds after the first walk (k_collect_pool).  A round finishes
// every task it takes; entries that overflow its stacks are handed to the
// next round, and the last one sends their rays to re-traversal instead.
#ifndef GRAY_POOL_ROUNDS
#define GRAY_POOL_ROUNDS 2
#endif
// Nodes and leaves each lane takes per step of a pooled round (two of each:
// two independent records in flight per lane and kind).
#ifndef GRAY_POOL_NODES
#define GRAY_POOL_NODES 2
#endif
#ifndef GRAY_POOL_LEAVES
#define GRAY_POOL_LEAVES 2
#endif
// Tasks a wave of a pooled round takes at a time, and its LDS stacks.
#ifndef GRAY_POOL_TAKE
#define GRAY_POOL_TAKE 32
#endif
#ifndef GRAY_POOL_STACK
#define GRAY_POOL_STACK 1024
#endif

constexpr int kCollectStack = GRAY_COLLECT_STACK;
constexpr int kCollectLeaves = GRAY_COLLECT_LEAVES;
constexpr int kLeafQueue = GRAY_LEAF_QUEUE;
constexpr int kTaskRounds = GRAY_POOL_ROUNDS;
static_assert(kTaskRounds >= 1, "at least one pooled task round");
#ifndef GRAY_TASK_WAVES
#define GRAY_TASK_WAVES 4096
#endif
constexpr int kTaskWaves = GRAY_TASK_WAVES;  // persistent waves of a task round
constexpr int kPoolSpillWaves = 256;          // persistent waves of a pooled round after the first
constexpr int32_t kRayDead = -(1 << 30);   // list_n of a ray whose list or tasks overflowed
static_assert((kLeafQueue & (kLeafQueue - 1)) == 0, "leaf queue size: a power of two");

// fp16_hits: the reference keeps each hit's distance and alpha as fp16
// (per_pixel_linked_list.h:28-29, __float2half_rn) and orders, cuts and
// composites with those values (forward_pass.cu:53-110); its full
// transmittance multiplies the fp32 alpha of the hit program.  Here the
// distance is rounded where a hit is produced and the alpha where it is
// composited or summed, so the fp32 alpha stays available for the full
// transmittance.
__device__ __forceinline__ float fp16_round(float x, bool on) { return on ? __half2float(__float2half_rn(x)) : x; }

// Hit records are written once and read back only by the compositing pass:
// a non-temporal store (a wave writes 1 KB of whole lines per slot) keeps them
// from evicting the nodes and leaf records that the walk re-reads.
// Hit-list layout.  Entries below kSlotMajor are slot-major: entry k of lane
// g at k * S + g (S = lanes of the launch), so a wave reading one entry of each
// of its 64 rays reads 1 KB of whole lines.  Entries from kSlotMajor on, which
// only long lists have, follow ray-major: entry k of lane g at
// kSlotMajor * S + g * (cap - kSlotMajor) + (k - kSlotMajor), so the wave that
// composites one long list reads whole lines of it.
#ifndef GRAY_SLOT_MAJOR
#define GRAY_SLOT_MAJOR 64
#endif
constexpr int kSlotMajor = GRAY_SLOT_MAJOR;

__device__ __forceinline__ size_t list_at(const TraceArgs &A, size_t g, int k) {
    const size_t S = (size_t)A.list_stride;
    if (k < kSlotMajor) return (size_t)k * S + g;
    return (size_t)kSlotMajor * S + g * (size_t)(A.list_cap - kSlotMajor) + (size_t)(k - kSlotMajor);
}

typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void put_hit(const TraceArgs &A, size_t g, int slot, float dist, uint32_t gid, float a) {
    const size_t e = list_at(A, g, slot);
    const f4v rec = {dist, __uint_as_float(gid), a, 0.0f};
    __builtin_nontemporal_store(rec, reinterpret_cast<f4v *>(A.list_rec + e));
}

// Exact test of leaf gid for the lane's ray; a hit goes to the ray's list at
// slot cnt (the per-tile pass owns the ray) or at a slot taken from list_n[g]
// with an atomic (tasks of one ray run concurrently).  ok turns false when the
// list overflows.
template <bool kTask>
__device__ __forceinline__ void collect_leaf(const TraceArgs &A, const GeoV &q, uint32_t gid, const LaneRay &r,
                                             size_t g, int cap, float athr, float p, int &cnt, bool &ok) {
    if (q.op <= athr) return;
    float dist, hx, hy, hz;
    if (!gray_hit_test(q.m, q.mu[0], q.mu[1], q.mu[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, dist, hx, hy, hz)) return;
    float gv, qq;
    const float a = gray_alpha(hx, hy, hz, q.clip, q.op, p, gv, qq);
    dist = fp16_round(dist, A.fp16 != 0);
    if (kTask) {
        const int slot = atomicAdd(A.list_n + g, 1);
        if (slot >= 0 && slot < cap) put_hit(A, g, slot, dist, gid, a);
        else ok = false;
    } else {
        if (cnt < cap) put_hit(A, g, cnt, dist, gid, a);
        ++cnt;
        if (cnt > cap) ok = false;
    }
}

// One lane's walk of the 4-wide tree for its ray (the order of the walk does
// not matter here: the list is composited in distance order afterwards), from
// a node or leaf reference (binary node indices of even depth; the root is 0).
// One iteration: the exact test of up to kCollectLeaves queued leaves, whose
// records were requested in the previous iteration, and the visit of one
// node, which tests its four child boxes; hit leaves join the queue, the
// first hit internal child is the next node and the others go onto the stack
// (or, when it is full, become tasks).  Queuing keeps the expensive leaf test
// dense across the wave.  Node visits pause while the queue could not take
// four leaves.
//
// Every hit goes to the ray's list at slot cnt (the per-tile pass owns the
// ray).  After `budget` iterations (< 0: none) a lane that still has work
// appends it to the task buffer as {g, reference} pairs and stops, so no wave
// runs longer than the budget however many boxes its rays cross.  ok turns
// false when the stack, the list or the task buffer overflows: the ray is
// then re-traversed instead.  Must be called by the whole wave.
template <int kStack>
__device__ __forceinline__ void collect_walk(const TraceArgs &A, const GeoRecord *__restrict__ geo,
                                                  const Bvh4Node *__restrict__ wide, const LaneRay &r, size_t g,
                                                  uint32_t start, int budget, uint32_t *__restrict__ s_stack,
                                                  uint32_t *__restrict__ s_leaf, int lane, int &cnt, bool &ok,
                                                  bool &handed, int &visits, int &leaves) {
    constexpr int NL = kCollectLeaves;
    const int cap = A.list_cap;
    const float athr = *A.knobs.alpha_threshold, p = *A.knobs.exp_power;
    uint32_t h = GRAY_EMPTY;
    Node4V nd;
    uint32_t pl[NL];
    GeoV pg[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) pl[l] = GRAY_EMPTY;
    if (start != GRAY_EMPTY) {
        if (start & GRAY_LEAF) {
            pl[0] = start & ~GRAY_LEAF;
            pg[0] = load_geo(geo, pl[0]);
        } else {
            h = start;
            nd = load_node4(wide, start);
        }
    }
    int sp = 0, qh = 0, qn = 0;
    for (int it = 0;; ++it) {
        if (!__ballot(h != GRAY_EMPTY || pl[0] != GRAY_EMPTY)) break;
        if (it == budget) {
            // the wave's remaining work (per lane: stack from the bottom, node,
            // pending and queued leaves) goes to the task buffer slot-major,
            // entry k of every lane before entry k + 1 of any, so a wave of
            // the next round takes the same stack slot of neighbouring rays,
            // which mostly walk the same subtrees.  Entries past the buffer's
            // end are dropped and their ray is re-traversed.
            int m = sp + qn + (h != GRAY_EMPTY);
#pragma unroll
            for (int l = 0; l < NL; ++l) m += pl[l] != GRAY_EMPTY;
            int mmax = m, tot = m;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                mmax = max(mmax, __shfl_xor(mmax, off));
                tot += __shfl_xor(tot, off);
            }
            handed = handed || m > 0;
            if (tot > 0) {
                int base = 0;
                if (lane == 0) base = atomicAdd(A.task_count_out, tot);
                base = __shfl(base, 0);
                const uint64_t below = (1ull << lane) - 1ull;
                int acc = 0;
                for (int k = 0; k < mmax; ++k) {
                    const uint64_t has = __ballot(m > k);
                    if (m > k) {
                        uint32_t ref = GRAY_EMPTY;
                        if (k < sp) {
                            ref = s_stack[k * GRAY_WAVE + lane];
                        } else {
                            int j = k - sp;
                            if (h != GRAY_EMPTY) {
                                if (j == 0) ref = h;
                                --j;
                            }
#pragma unroll
                            for (int l = 0; l < NL; ++l) {
                                if (pl[l] != GRAY_EMPTY) {
                                    if (j == 0) ref = pl[l] | GRAY_LEAF;
                                    --j;
                                }
                            }
                            if (ref == GRAY_EMPTY) ref = s_leaf[((qh + j) & (kLeafQueue - 1)) * GRAY_WAVE + lane] | GRAY_LEAF;
                        }
                        const int o = base + acc + __popcll(has & below);
                        if (o < A.task_cap) A.task_out[o] = make_uint2((uint32_t)g, ref);
                        else ok = false;
                    }
                    acc += __popcll(has);
                }
            }
            break;
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            if (pl[l] == GRAY_EMPTY) continue;
            ++leaves;
            collect_leaf<false>(A, pg[l], pl[l], r, g, cap, athr, p, cnt, ok);
        }
        uint32_t nh = h;
        bool moved = false;
        if (h != GRAY_EMPTY && qn <= kLeafQueue - 4) {
            moved = true;
            ++visits;
            // an unused child (ref GRAY_EMPTY) stays GRAY_EMPTY whatever its box gives
            uint32_t ch[4];
            ch[0] = child_hit<0>(r, nd) ? nd.ch.x : GRAY_EMPTY;
            ch[1] = child_hit<1>(r, nd) ? nd.ch.y : GRAY_EMPTY;
            ch[2] = child_hit<2>(r, nd) ? nd.ch.z : GRAY_EMPTY;
            ch[3] = child_hit<3>(r, nd) ? nd.ch.w : GRAY_EMPTY;
            nh = GRAY_EMPTY;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t c = ch[k];
                if (c == GRAY_EMPTY) continue;
                if (c & GRAY_LEAF) {
                    s_leaf[((qh + qn) & (kLeafQueue - 1)) * GRAY_WAVE + lane] = c & ~GRAY_LEAF;
                    ++qn;
                } else if (nh == GRAY_EMPTY) {
                    nh = c;
                } else if (sp < kStack) {
                    s_stack[sp * GRAY_WAVE + lane] = c;
                    ++sp;
                } else if (budget >= 0) {
                    handed = true;
                    const int o = atomicAdd(A.task_count_out, 1);
                    if (o < A.task_cap) A.task_out[o] = make_uint2((uint32_t)g, c);
                    else ok = false;
                } else {
                    ok = false;
                }
            }
            if (nh == GRAY_EMPTY && sp > 0) {
                --sp;
                nh = s_stack[sp * GRAY_WAVE + lane];
            }
        }
        if (!ok) {
            nh = GRAY_EMPTY;
            sp = qn = 0;
        }
        // requests: the next leaf records first, then the next node
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            pl[l] = GRAY_EMPTY;
            if (qn > 0) {
                pl[l] = s_leaf[qh * GRAY_WAVE + lane];
                qh = (qh + 1) & (kLeafQueue - 1);
                --qn;
                pg[l] = load_geo(geo, pl[l]);
            }
        }
        h = nh;
        if (moved && h != GRAY_EMPTY) nd = load_node4(wide, h);
    }
}

static_assert(kLeafQueue >= 4, "a 4-wide visit may queue four leaves");

// Per-tile collect pass: every ray of the tile walks from the root.  kStack
// is kCollectStack, or kDetStack in the deterministic mode, whose walks run to
// their end without handing work on.
constexpr int kDetStack = 64;
template <int kStack>
__global__ __launch_bounds__(GRAY_WAVE) void k_collect(TraceArgs A, const GeoRecord *__restrict__ geo,
                                                       const Bvh4Node *__restrict__ nodes) {
    __shared__ uint32_t s_stack[kStack * GRAY_WAVE];
    __shared__ uint32_t s_leaf[kLeafQueue * GRAY_WAVE];
    if (A.bin_mode && *A.bin_mode == 0) return;  // the tile bins serve this launch
    int vi, tx, ty, tile;
    if (!block_tile(A, vi, tx, ty, &tile)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    LaneRay r{};
    bool active = false;
    if (x < V.width && y < V.height) active = make_ray(A.knobs, V, x, y, r);
    const size_t g = (size_t)tile * GRAY_WAVE + lane;
    int cnt = 0, visits = 0, leaves = 0;
    bool ok = true, handed = false;
    collect_walk<kStack>(A, geo, nodes, r, g, (active && A.num_gaussians > 0) ? 0u : GRAY_EMPTY,
                                      A.budget, s_stack, s_leaf, lane, cnt, ok, handed, visits, leaves);
    A.list_n[g] = ok ? cnt : -1;
    const bool open = __ballot(handed) != 0;
    if (lane == 0 && A.tile_open) A.tile_open[tile] = open ? 1 : 0;
    if (A.tile_stats) {
        int hsum = ok ? cnt : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            hsum += __shfl_xor(hsum, off);
            visits += __shfl_xor(visits, off);
            leaves += __shfl_xor(leaves, off);
        }
        if (lane == 0) {
            int32_t *ts = A.tile_stats + 4 * (size_t)(A.tile_base_launch + tile);
            ts[0] = visits;
            ts[1] = leaves;
            ts[3] = hsum;
        }
    }
}

// Tile, view and pixel of launch lane g (the inverse of the tile numbering).
__device__ __forceinline__ const DevView &lane_pixel(const TraceArgs &A, uint32_t g, int &x, int &y) {
    const int tile = (int)(g / GRAY_WAVE), lane = (int)(g % GRAY_WAVE);
    int v = 0;
#pragma unroll
    for (int k = 1; k < GRAY_MAX_VIEWS; ++k)
        if (k < A.num_views && tile >= A.views[k].tile_base) v = k;
    const DevView &V = A.views[v];
    const int lt = tile - V.tile_base;
    x = (lt % V.tiles_x) * GRAY_TILE + (lane & 7);
    y = (lt / V.tiles_x) * GRAY_TILE + (lane >> 3);
    return V;
}

// Pooled task round: a wave takes kPoolTake tasks and its 64 lanes share
// their work.  The wave's frontier lives in two LDS stacks, one of nodes and
// one of leaves, each entry tagged with its task's slot.  Every step a lane
// takes kPoolNodes nodes (four child box tests each; hit internal children
// join the node stack, hit leaves the leaf stack) and kPoolLeaves leaves (the
// exact test; a hit is appended to the ray's list with an atomic slot),
// whatever task they belong to, as iterations of the per-lane walk do for its
// own ray.  So
// a task with a large subtree is walked 64 entries at a time instead of by one
// lane, and the round's length follows the tree depth and the wave's total
// work rather than its longest walk.  An entry that does not fit its stack
// becomes a task of the next pooled round; in the last round it marks its ray
// for re-traversal.
constexpr int kPoolTake = GRAY_POOL_TAKE;
constexpr int kPoolStack = GRAY_POOL_STACK;
constexpr int kPoolNodes = GRAY_POOL_NODES;
constexpr int kPoolLeaves = GRAY_POOL_LEAVES;
static_assert(kPoolTake >= 1 && kPoolTake <= GRAY_WAVE, "pool take: 1 to 64 tasks");
static_assert(kPoolStack >= 4 * GRAY_POOL_NODES * GRAY_WAVE, "pool stack: room for one step's pushes");

// The ray of task slot `slot` from the pool's LDS table ([component][slot]).
__device__ __forceinline__ LaneRay pool_ray(const float *s_ray, int slot, float one) {
    LaneRay q;
    q.ox = s_ray[0 * kPoolTake + slot];
    q.oy = s_ray[1 * kPoolTake + slot];
    q.oz = s_ray[2 * kPoolTake + slot];
    q.dx = s_ray[3 * kPoolTake + slot];
    q.dy = s_ray[4 * kPoolTake + slot];
    q.dz = s_ray[5 * kPoolTake + slot];
    q.ix = s_ray[6 * kPoolTake + slot];
    q.iy = s_ray[7 * kPoolTake + slot];
    q.iz = s_ray[8 * kPoolTake + slot];
    q.one = one;
    return q;
}

__global__ __launch_bounds__(GRAY_WAVE) void k_collect_pool(TraceArgs A, const GeoRecord *__restrict__ geo,
                                                            const Bvh4Node *__restrict__ wide) {
    __shared__ uint32_t s_node[kPoolStack], s_leaf[kPoolStack];
    __shared__ uint8_t s_nslot[kPoolStack], s_lslot[kPoolStack];
    __shared__ float s_ray[9 * kPoolTake];  // [component][slot]: origin, direction, inverse direction
    __shared__ uint32_t s_g[kPoolTake];
    const int lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const int total = min(*A.task_count_in, A.task_cap);
    if (total <= 0) return;  // nothing handed on (or the tile bins served the launch)
    const int cap = A.list_cap;
    const float athr = *A.knobs.alpha_threshold, p = *A.knobs.exp_power;
    float one = 1.0f;
    asm volatile("" : "+v"(one));
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(A.task_head, kPoolTake);
        base = __shfl(base, 0);
        if (base >= total) break;
        uint32_t start = GRAY_EMPTY;
        if (lane < kPoolTake) {
            const int t = base + lane;
            uint32_t g = 0;
            LaneRay r{};
            if (t < total) {
                const uint2 task = A.task_in[t];
                if (task.x < (uint64_t)A.list_stride) {
                    g = task.x;
                    int x, y;
                    const DevView &V = lane_pixel(A, g, x, y);
                    if (x < V.width && y < V.height && A.list_n[g] >= 0 && make_ray(A.knobs, V, x, y, r))
                        start = task.y;
                }
            }
            s_ray[0 * kPoolTake + lane] = r.ox;
            s_ray[1 * kPoolTake + lane] = r.oy;
            s_ray[2 * kPoolTake + lane] = r.oz;
            s_ray[3 * kPoolTake + lane] = r.dx;
            s_ray[4 * kPoolTake + lane] = r.dy;
            s_ray[5 * kPoolTake + lane] = r.dz;
            s_ray[6 * kPoolTake + lane] = r.ix;
            s_ray[7 * kPoolTake + lane] = r.iy;
            s_ray[8 * kPoolTake + lane] = r.iz;
            s_g[lane] = g;
        }
        const bool sn0 = start != GRAY_EMPTY && !(start & GRAY_LEAF);
        const bool sl0 = start != GRAY_EMPTY && (start & GRAY_LEAF);
        const uint64_t hn = __ballot(sn0), hl = __ballot(sl0);
        if (sn0) {
            const int o = __popcll(hn & below);
            s_node[o] = start;
            s_nslot[o] = (uint8_t)lane;
        }
        if (sl0) {
            const int o = __popcll(hl & below);
            s_leaf[o] = start & ~GRAY_LEAF;
            s_lslot[o] = (uint8_t)lane;
        }
        int sn = __popcll(hn), sl = __popcll(hl);
        __syncthreads();
        while (sn + sl > 0) {
            // take the top entries of each stack: kPoolNodes nodes and
            // kPoolLeaves leaves per lane
            constexpr int PN = kPoolNodes, PL = kPoolLeaves;
            const int nn = min(sn, PN * GRAY_WAVE), nl = min(sl, PL * GRAY_WAVE);
            sn -= nn;
            sl -= nl;
            uint32_t node[PN], leaf[PL];
            int nslot[PN], lslot[PL];
#pragma unroll
            for (int j = 0; j < PN; ++j) {
                node[j] = GRAY_EMPTY;
                nslot[j] = 0;
                if (lane + j * GRAY_WAVE < nn) {
                    node[j] = s_node[sn + lane + j * GRAY_WAVE];
                    nslot[j] = s_nslot[sn + lane + j * GRAY_WAVE];
                }
            }
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                leaf[j] = GRAY_EMPTY;
                lslot[j] = 0;
                if (lane + j * GRAY_WAVE < nl) {
                    leaf[j] = s_leaf[sl + lane + j * GRAY_WAVE];
                    lslot[j] = s_lslot[sl + lane + j * GRAY_WAVE];
                }
            }
            // every record requested before any is used
            Node4V nd[PN];
            GeoV gv[PL];
#pragma unroll
            for (int j = 0; j < PN; ++j)
                if (node[j] != GRAY_EMPTY) nd[j] = load_node4(wide, node[j]);
#pragma unroll
            for (int j = 0; j < PL; ++j)
                if (leaf[j] != GRAY_EMPTY) gv[j] = load_geo(geo, leaf[j]);
            uint32_t ch[PN][4];
            uint32_t ng[PN];
#pragma unroll
            for (int j = 0; j < PN; ++j) {
                ch[j][0] = ch[j][1] = ch[j][2] = ch[j][3] = GRAY_EMPTY;
                ng[j] = 0;
                if (node[j] != GRAY_EMPTY) {
                    const LaneRay q = pool_ray(s_ray, nslot[j], one);
                    ng[j] = s_g[nslot[j]];
                    ch[j][0] = child_hit<0>(q, nd[j]) ? nd[j].ch.x : GRAY_EMPTY;
                    ch[j][1] = child_hit<1>(q, nd[j]) ? nd[j].ch.y : GRAY_EMPTY;
                    ch[j][2] = child_hit<2>(q, nd[j]) ? nd[j].ch.z : GRAY_EMPTY;
                    ch[j][3] = child_hit<3>(q, nd[j]) ? nd[j].ch.w : GRAY_EMPTY;
                }
            }
#pragma unroll
            for (int j = 0; j < PL; ++j) {
                if (leaf[j] == GRAY_EMPTY) continue;
                const LaneRay q = pool_ray(s_ray, lslot[j], one);
                const uint32_t lg = s_g[lslot[j]];
                int cnt = 0;
                bool ok = true;
                collect_leaf<true>(A, gv[j], leaf[j], q, lg, cap, athr, p, cnt, ok);
                if (!ok) atomicExch(A.list_n + lg, kRayDead);
            }
            // every lane has read its entries before any push overwrites them
            __syncthreads();
            // an entry that does not fit becomes a task of the next round
            // (A.budget >= 0) or, in the last round, sends its ray to
            // re-traversal
#pragma unroll
            for (int j = 0; j < PN; ++j) {
                bool lost = false;
                auto spill = [&](uint32_t ref) {
                    if (A.budget >= 0) {
                        const int o = atomicAdd(A.task_count_out, 1);
                        if (o < A.task_cap) {
                            A.task_out[o] = make_uint2(ng[j], ref);
                            return;
                        }
                    }
                    lost = true;
                };
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t c = ch[j][k];
                    const bool isn = c != GRAY_EMPTY && !(c & GRAY_LEAF);
                    const bool isl = c != GRAY_EMPTY && (c & GRAY_LEAF);
                    const uint64_t mn = __ballot(isn), ml = __ballot(isl);
                    if (isn) {
                        const int o = sn + __popcll(mn & below);
                        if (o < kPoolStack) {
                            s_node[o] = c;
                            s_nslot[o] = (uint8_t)nslot[j];
                        } else {
                            spill(c);
                        }
                    }
                    if (isl) {
                        const int o = sl + __popcll(ml & below);
                        if (o < kPoolStack) {
                            s_leaf[o] = c & ~GRAY_LEAF;
                            s_lslot[o] = (uint8_t)nslot[j];
                        } else {
                            spill(c);
                        }
                    }
                    sn = min(sn + __popcll(mn), kPoolStack);
                    sl = min(sl + __popcll(ml), kPoolStack);
                }
                if (lost) atomicExch(A.list_n + ng[j], kRayDead);
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Tile-binned collect pass (collect_mode 0).  Every ray of a pinhole view
// leaves the view's origin, so the rays that can meet a Gaussian are those of
// the pixels its box projects to.  Each Gaussian's box (k_snapshot's, padded
// past the ellipsoid) is projected into every view of the launch; the (8x8
// tile, Gaussian) pairs are counted per tile, scanned and filled into the
// tiles' bins; then the 64 rays of a tile test every Gaussian of its bin with
// the walk's exact test (gray_local_origin once per Gaussian and view, then
// gray_hit_test_local per ray: the operations of gray_hit_test), so each
// list holds the hits the walk finds.  The walk is a dependent chain of node
// fetches per lane; here a wave tests one staged record at a time with all
// 64 lanes, so the pass is bound by VALU work (about 50 operations per ray
// and bin entry).  A launch with explicit rays, a degenerate camera or more
// pairs than the bins hold sets *bin_mode and the walk runs instead.
// ---------------------------------------------------------------------------
constexpr int kBinChunk = 512;     // bin entries per work item (a tile with more is split over several)
constexpr int kBinThreads = 256;   // threads of the projection kernels (one Gaussian each)
constexpr int kBinScanThreads = 1024;

// View projection: (a, b, w) = inv (p - origin) gives the pixel coordinates
// xf = a / w, yf = b / w of make_ray, whose ray direction is
// B (xf, yf, 1) with B = [c0, c1, -c2] (camera-to-world columns);
// xf = sx (2 (px + 0.5) / W - 1), yf = sy (1 - 2 (py + 0.5) / H).  fp32: the
// rounding (about 1e-6 relative) stays far inside the pixel slack below.
struct BinProj {
    float inv[9];
    float sx, sy;
    float margin;  // pixels of slack around the projected box
    float o[3];
    int w, h, tiles_x, tile_base;
    bool ok;
};

__device__ void bin_proj(const DevView &V, BinProj &P) {
    const float *c = V.c2w;
    const float b0 = c[0], b1 = c[1], b2 = -c[2];
    const float b3 = c[3], b4 = c[4], b5 = -c[5];
    const float b6 = c[6], b7 = c[7], b8 = -c[8];
    const float k0 = b4 * b8 - b5 * b7, k1 = b5 * b6 - b3 * b8, k2 = b3 * b7 - b4 * b6;
    const float det = b0 * k0 + b1 * k1 + b2 * k2;
    const float r = 1.0f / det;
    P.inv[0] = k0 * r, P.inv[1] = (b2 * b7 - b1 * b8) * r, P.inv[2] = (b1 * b5 - b2 * b4) * r;
    P.inv[3] = k1 * r, P.inv[4] = (b0 * b8 - b2 * b6) * r, P.inv[5] = (b2 * b3 - b0 * b5) * r;
    P.inv[6] = k2 * r, P.inv[7] = (b1 * b6 - b0 * b7) * r, P.inv[8] = (b0 * b4 - b1 * b3) * r;
    // the scales as make_ray rounds them
    const float vs = tanf(*V.fov / 2.0f);
    const float aspect = (float)V.width / (float)V.height;
    P.sx = aspect * vs;
    P.sy = vs;
    P.o[0] = V.origin[0], P.o[1] = V.origin[1], P.o[2] = V.origin[2];
    P.w = V.width, P.h = V.height, P.tiles_x = V.tiles_x, P.tile_base = V.tile_base;
    // a camera basis far from orthonormal (|det| off 1 by more than 1e-3) is
    // left to the walk, so the rounding bound above holds
    bool ok = fabsf(fabsf(det) - 1.0f) < 1e-3f && P.sx > 0.0f && P.sy > 0.0f && isfinite(P.sx) && isfinite(P.sy);
    for (int i = 0; i < 9; ++i) ok = ok && isfinite(c[i]);
    for (int i = 0; i < 3; ++i) ok = ok && isfinite(P.o[i]);
    P.ok = ok;
}

// Tiles [tx0, tx1] x [ty0, ty1] of the view whose pixels' rays can meet box b
// (min xyz, max xyz); false when no ray can.  Corners in front of the origin
// (w > 0) are projected.  Where the box reaches the origin's plane (w = 0),
// its projection runs off to infinity: along an edge from a front corner
// (a_f, b_f, w_f) to the plane, (a / w, b / w) moves on a straight line from
// the corner's projection in the direction of the crossing point (a0, b0),
// so the rectangle reaches the view's edge on the side of each of its signs
// (both sides where a coordinate is within rounding of 0; a box around the
// origin covers the whole view).  Corners within rounding of the plane count
// as on it.  Pixels count whose ray (through the pixel's centre, or within
// half a pixel of it with jitter) lies within P.margin of the range.  Not inlined: the count and fill passes call one
// compiled copy, so they see the same rectangles (two inlined copies could be
// contracted differently).

__device__ __noinline__ bool bin_rect(const BinProj &P, const float *b, int &tx0, int &tx1, int &ty0, int &ty1) {
    if (!(b[0] <= b[3] && b[1] <= b[4] && b[2] <= b[5])) return false;  // empty (invisible) or NaN box
    const float lx = b[0] - P.o[0], ly = b[1] - P.o[1], lz = b[2] - P.o[2];
    const float ex = b[3] - b[0], ey = b[4] - b[1], ez = b[5] - b[2];
    float base[3], dx[3], dy[3], dz[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        base[k] = (P.inv[3 * k] * lx + P.inv[3 * k + 1] * ly) + P.inv[3 * k + 2] * lz;
        dx[k] = P.inv[3 * k] * ex;
        dy[k] = P.inv[3 * k + 1] * ey;
        dz[k] = P.inv[3 * k + 2] * ez;
    }
    const float inf = __builtin_huge_valf();
    float xlo = inf, xhi = -inf, ylo = inf, yhi = -inf;
    auto extend = [&](float a, float c, float tol) {
        if (a > tol) xhi = inf;
        else if (a < -tol) xlo = -inf;
        else xlo = -inf, xhi = inf;
        if (c > tol) yhi = inf;
        else if (c < -tol) ylo = -inf;
        else ylo = -inf, yhi = inf;
    };
    float ca[8], cb[8], cw[8], ct[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            v[k] = base[k] + ((i & 1) ? dx[k] : 0.0f) + ((i & 2) ? dy[k] : 0.0f) + ((i & 4) ? dz[k] : 0.0f);
        ca[i] = v[0], cb[i] = v[1], cw[i] = v[2];
        // rounding of (a, b, w): a few 1e-7 of |v| and of the box's size
        ct[i] = 1e-5f * (fabsf(v[0]) + fabsf(v[1]) + fabsf(v[2]) + fabsf(lx) + fabsf(ly) + fabsf(lz) + ex + ey + ez) +
                1e-30f;
        if (cw[i] > ct[i]) {
            const float x = ca[i] / cw[i], y = cb[i] / cw[i];
            xlo = fminf(xlo, x), xhi = fmaxf(xhi, x), ylo = fminf(ylo, y), yhi = fmaxf(yhi, y);
        } else if (cw[i] >= -ct[i]) {
            extend(ca[i], cb[i], ct[i]);  // a corner on the plane
        }
    }
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        // edge e: axis bit k = 1 << (e / 4), i the corner without it
        const int k = 1 << (e >> 2);
        const int lowb = e & 3;
        const int i = k == 1 ? (lowb << 1) : k == 2 ? ((lowb & 1) | ((lowb & 2) << 1)) : lowb;
        const int j = i | k;
        const bool fi = cw[i] > ct[i], fj = cw[j] > ct[j];
        if (fi == fj) continue;
        const int f = fi ? i : j, q = fi ? j : i;
        if (cw[q] >= -ct[q]) continue;  // ends on the plane: its corner extended already
        const float lam = cw[f] / (cw[f] - cw[q]);
        extend(ca[f] + lam * (ca[q] - ca[f]), cb[f] + lam * (cb[q] - cb[f]), fmaxf(ct[f], ct[q]));
    }
    if (!(xlo <= xhi) || !(ylo <= yhi)) return false;  // nothing in front of the origin
    // pixel x's ray passes through continuous coordinate x (+ its jitter):
    // the pixels within the margin of the projected range
    const float hw = 0.5f * (float)P.w, hh = 0.5f * (float)P.h, m = P.margin;
    const float x0 = ceilf((xlo / P.sx + 1.0f) * hw - 0.5f - m), x1 = floorf((xhi / P.sx + 1.0f) * hw - 0.5f + m);
    const float y0 = ceilf((1.0f - yhi / P.sy) * hh - 0.5f - m), y1 = floorf((1.0f - ylo / P.sy) * hh - 0.5f + m);
    if (!(x0 <= x1 && y0 <= y1)) return false;  // between pixel centres
    if (!(x1 >= 0.0f && y1 >= 0.0f && x0 <= (float)(P.w - 1) && y0 <= (float)(P.h - 1))) return false;
    tx0 = (x0 <= 0.0f ? 0 : (int)x0) / GRAY_TILE;
    tx1 = (x1 >= (float)(P.w - 1) ? P.w - 1 : (int)x1) / GRAY_TILE;
    ty0 = (y0 <= 0.0f ? 0 : (int)y0) / GRAY_TILE;
    ty1 = (y1 >= (float)(P.h - 1) ? P.h - 1 : (int)y1) / GRAY_TILE;
    return true;
}

// Clears the tiles' pair counts, the lists' counts, the tile flags and the
// statistics, and decides the launch's mode (thread 0).
__global__ void k_bin_setup(TraceArgs A, int64_t lanes) {
    const int T = A.total_tiles;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < T) {
        A.bin_tiles[i] = 0;
        A.tile_open[i] = 0;
    }
    if (i < lanes) A.list_n[i] = 0;
    if (A.tile_stats && i < 4 * (int64_t)T) A.tile_stats[4 * (int64_t)A.tile_base_launch + i] = 0;
    if (i == 0) {
        int mode = 0;
        for (int v = 0; v < A.num_views; ++v) {
            const DevView &V = A.views[v];
            if (V.rays_d && (!V.rays_en || *V.rays_en)) mode = 1;  // explicit rays: no common origin
            BinProj P;
            bin_proj(V, P);
            if (!P.ok) mode = 1;
        }
        *A.bin_mode = mode;
    }
}

// Pairs of the Gaussians of one workgroup, counted (fill 0) or written
// (fill 1).  The workgroup takes 256 consecutive leaves of the BVH's Morton
// order, so its Gaussians lie close together and project to few tiles: pairs
// are first counted per tile in an LDS hash table, and the table's tiles then
// take their counts (count) or reserve their ranges (fill) with one global
// atomic each instead of one per pair (3.9M pairs on 8.5K tiles serialised on
// the tiles' counters).  A Gaussian covering more than kBinBigRect tiles (a box
// reaching the camera's plane) adds its pairs with global atomics, in both
// kernels; when the table fills, the count kernel does the same for the pairs
// that found no slot and the fill kernel writes the whole workgroup's pairs
// that way.  Count and fill compute the same rectangles, so the fill stays
// inside each tile's scanned range.
constexpr int kBinHash = 2048;      // LDS hash slots per workgroup
constexpr int kBinProbes = 16;
constexpr int kBinBigRect = 256;    // tiles above which a Gaussian's pairs bypass the table

__device__ __forceinline__ uint32_t bin_hash(int tile) { return ((uint32_t)tile * 2654435761u) >> 21; }

__global__ __launch_bounds__(kBinThreads) void k_bin_pairs(TraceArgs A, const float *__restrict__ aabb,
                                                           const uint32_t *__restrict__ order, int n, int fill) {
    __shared__ BinProj s_proj[GRAY_MAX_VIEWS];
    __shared__ int32_t s_key[kBinHash], s_cnt[kBinHash], s_base[kBinHash];
    __shared__ int32_t s_full;
    if (*A.bin_mode != 0) return;
    const int t = threadIdx.x;
    if (t < A.num_views) {
        bin_proj(A.views[t], s_proj[t]);
        // half a pixel of jitter when on, and 0.05 pixel for the rounding of
        // make_ray and of the projection (both near 1e-3 pixel)
        s_proj[t].margin = (A.knobs.jitter && *A.knobs.jitter) ? 0.55f : 0.05f;
    }
    for (int i = t; i < kBinHash; i += kBinThreads) s_key[i] = -1, s_cnt[i] = 0;
    if (t == 0) s_full = 0;
    __syncthreads();
    const int k = blockIdx.x * kBinThreads + t;
    const uint32_t g = k < n ? order[k] : 0u;
    float box[6] = {1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f};  // empty
    if (k < n) {
        const float2 *q = reinterpret_cast<const float2 *>(aabb + 6 * (size_t)g);
        const float2 p0 = q[0], p1 = q[1], p2 = q[2];
        box[0] = p0.x, box[1] = p0.y, box[2] = p1.x, box[3] = p1.y, box[4] = p2.x, box[5] = p2.y;
    }
    int32_t *cnt = A.bin_tiles, *cur = A.bin_tiles + A.total_tiles;
    // pass 1: the workgroup's pairs per tile in the table
    for (int v = 0; v < A.num_views; ++v) {
        const BinProj &P = s_proj[v];
        int tx0, tx1, ty0, ty1;
        if (!bin_rect(P, box, tx0, tx1, ty0, ty1)) continue;
        const bool big = (tx1 - tx0 + 1) * (ty1 - ty0 + 1) > kBinBigRect;
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) {
                const int tile = P.tile_base + ty * P.tiles_x + tx;
                if (big) {
                    if (fill) A.bins[atomicAdd(cur + tile, 1)] = g;
                    else atomicAdd(cnt + tile, 1);
                    continue;
                }
                uint32_t h = bin_hash(tile);
                int slot = -1;
                for (int pr = 0; pr < kBinProbes; ++pr, h = (h + 1) & (kBinHash - 1)) {
                    const int old = atomicCAS(&s_key[h], -1, tile);
                    if (old == -1 || old == tile) {
                        slot = (int)h;
                        break;
                    }
                }
                if (slot >= 0) atomicAdd(&s_cnt[slot], 1);
                else if (fill) s_full = 1;
                else atomicAdd(cnt + tile, 1);
            }
    }
    __syncthreads();
    const bool full = fill && s_full != 0;
    // pass 2: one global atomic per tile of the table
    if (!full) {
        for (int i = t; i < kBinHash; i += kBinThreads) {
            const int tile = s_key[i];
            if (tile < 0) continue;
            if (fill) {
                s_base[i] = atomicAdd(cur + tile, s_cnt[i]);
                s_cnt[i] = 0;
            } else {
                atomicAdd(cnt + tile, s_cnt[i]);
            }
        }
    }
    if (!fill) return;
    __syncthreads();
    // pass 3 (fill): every table pair takes a rank in its tile's range
    for (int v = 0; v < A.num_views; ++v) {
        const BinProj &P = s_proj[v];
        int tx0, tx1, ty0, ty1;
        if (!bin_rect(P, box, tx0, tx1, ty0, ty1)) continue;
        if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) > kBinBigRect) continue;  // written in pass 1
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) {
                const int tile = P.tile_base + ty * P.tiles_x + tx;
                if (full) {
                    A.bins[atomicAdd(cur + tile, 1)] = g;
                    continue;
                }
                uint32_t h = bin_hash(tile);
                int pr = 0;
                while (s_key[h] != tile && ++pr < kBinProbes) h = (h + 1) & (kBinHash - 1);  // inserted in pass 1
                if (s_key[h] == tile) A.bins[s_base[h] + atomicAdd(&s_cnt[h], 1)] = g;
            }
    }
}

// One workgroup: exclusive scans of the pair counts (the fill cursors) and of
// the work items per tile (ceil(count / kBinChunk)); totals past the bins or
// the item buffer send the launch to the walk.
__global__ __launch_bounds__(kBinScanThreads) void k_bin_scan(TraceArgs A) {
    __shared__ int64_t s_p[kBinScanThreads], s_i[kBinScanThreads];
    if (*A.bin_mode != 0) return;
    const int T = A.total_tiles, t = threadIdx.x;
    const int per = (T + kBinScanThreads - 1) / kBinScanThreads;
    const int lo = min(T, t * per), hi = min(T, lo + per);
    const int32_t *cnt = A.bin_tiles;
    int32_t *cur = A.bin_tiles + T, *ist = A.bin_tiles + 2 * T + 1;
    int64_t sp = 0, si = 0;
    for (int i = lo; i < hi; ++i) {
        const int c = cnt[i];
        sp += c;
        si += (c + kBinChunk - 1) / kBinChunk;
    }
    s_p[t] = sp;
    s_i[t] = si;
    __syncthreads();
    for (int off = 1; off < kBinScanThreads; off <<= 1) {
        const int64_t a = t >= off ? s_p[t - off] : 0, b = t >= off ? s_i[t - off] : 0;
        __syncthreads();
        s_p[t] += a;
        s_i[t] += b;
        __syncthreads();
    }
    int64_t p = s_p[t] - sp, it = s_i[t] - si;
    const bool fits = s_p[kBinScanThreads - 1] <= A.bin_cap && s_i[kBinScanThreads - 1] <= A.bin_item_cap;
    for (int i = lo; i < hi && fits; ++i) {
        const int c = cnt[i];
        cur[i] = (int32_t)p;
        ist[i] = (int32_t)it;
        p += c;
        it += (c + kBinChunk - 1) / kBinChunk;
    }
    if (t == kBinScanThreads - 1) {
        ist[T] = fits ? (int32_t)s_i[t] : 0;
        if (!fits) *A.bin_mode = 1;
    }
}

// One thread per tile: its work items (tile, chunk), tiles in order.
__global__ void k_bin_items(TraceArgs A) {
    if (*A.bin_mode != 0) return;
    const int T = A.total_tiles;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    const int c = A.bin_tiles[i], first = A.bin_tiles[2 * T + 1 + i];
    for (int k = 0; k * kBinChunk < c; ++k) A.bin_items[first + k] = make_int2(i, k);
}

// One wave per work item: the 64 rays of the item's tile test up to kBinChunk
// entries of its bin, 64 at a time.  The entries' records are staged in LDS
// by one lane each (with o' = M (o - mu) for the view's origin, and those
// below the alpha threshold dropped); every lane then tests its ray against
// each staged record in turn (a broadcast read), keeping a hit mask, and
// writes the hits after the chunk: at the tile's own running count when the
// bin is one item, else at slots taken with one atomic per lane and chunk
// (items of one tile run concurrently).  Items are mapped XCD-aware: XCD k
// takes the k-th eighth of the items, in tile order, so the neighbouring
// tiles whose bins share Gaussians share one L2.
__device__ __forceinline__ void collect_bin_item(const TraceArgs &A, const GeoRecord *__restrict__ geo,
                                                 float4 *__restrict__ s_rec, int item) {
    const int T = A.total_tiles;
    const int2 wi = A.bin_items[item];
    const int tile = wi.x, chunk = wi.y;
    int v = 0;
#pragma unroll
    for (int k = 1; k < GRAY_MAX_VIEWS; ++k)
        if (k < A.num_views && tile >= A.views[k].tile_base) v = k;
    const DevView &V = A.views[v];
    const int lt = tile - V.tile_base;
    const int lane = threadIdx.x;
    const int x = (lt % V.tiles_x) * GRAY_TILE + (lane & 7), y = (lt / V.tiles_x) * GRAY_TILE + (lane >> 3);
    LaneRay r{};
    bool active = false;
    if (x < V.width && y < V.height) active = make_ray(A.knobs, V, x, y, r);
    const size_t g = (size_t)tile * GRAY_WAVE + lane;
    const int cnt_t = A.bin_tiles[tile], end_t = A.bin_tiles[T + tile];
    const int j0 = end_t - cnt_t + chunk * kBinChunk, j1 = min(end_t, j0 + kBinChunk);
    const bool single = cnt_t <= kBinChunk;
    const int cap = A.list_cap;
    const float athr = *A.knobs.alpha_threshold, p = *A.knobs.exp_power;
    const float ox = V.origin[0], oy = V.origin[1], oz = V.origin[2];
    const uint64_t below = (1ull << lane) - 1ull;
    int n = 0, tests = 0;
    bool ok = true;
    // the next chunk's entry and record are requested before the current
    // chunk's tests, so their loads overlap the tests
    uint32_t gid = j0 + lane < j1 ? A.bins[j0 + lane] : GRAY_EMPTY;
    bool valid = gid < (uint32_t)A.num_gaussians;
    GeoV q = valid ? load_geo(geo, gid) : GeoV{};
    for (int j = j0; j < j1; j += GRAY_WAVE) {
        bool keep = false;
        float4 q0, q1, q2, q3;
        if (valid) {
            keep = q.op > athr;
            float lox, loy, loz;
            gray_local_origin(q.m, q.mu[0], q.mu[1], q.mu[2], ox, oy, oz, lox, loy, loz);
            q0 = make_float4(q.m[0], q.m[1], q.m[2], q.m[3]);
            q1 = make_float4(q.m[4], q.m[5], q.m[6], q.m[7]);
            q2 = make_float4(q.m[8], lox, loy, loz);
            q3 = make_float4(q.op, q.clip, __uint_as_float(gid), 0.0f);
        }
        const uint64_t kb = __ballot(keep);
        if (keep) {
            float4 *d = s_rec + 4 * __popcll(kb & below);
            d[0] = q0, d[1] = q1, d[2] = q2, d[3] = q3;
        }
        const int mk = __popcll(kb);
        __syncthreads();  // the staged records are visible to every lane
        {
            const int jn = j + GRAY_WAVE + lane;
            gid = jn < j1 ? A.bins[jn] : GRAY_EMPTY;
            valid = gid < (uint32_t)A.num_gaussians;
            if (valid) q = load_geo(geo, gid);
        }
        uint64_t mask = 0;
        if (active) {
            for (int k = 0; k < mk; ++k) {
                const float4 a = s_rec[4 * k], c = s_rec[4 * k + 1], e = s_rec[4 * k + 2];
                const float mm[9] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w, e.x};
                float dist, hx, hy, hz;
                if (gray_hit_test_local(mm, e.y, e.z, e.w, r.dx, r.dy, r.dz, dist, hx, hy, hz)) mask |= 1ull << k;
            }
            tests += mk;
        }
        const int nh = __popcll(mask);
        int base = n;
        if (!single && nh > 0) base = atomicAdd(A.list_n + g, nh);
        for (int i = 0; mask; ++i) {
            const int k = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            const float4 a = s_rec[4 * k], c = s_rec[4 * k + 1], e = s_rec[4 * k + 2], f = s_rec[4 * k + 3];
            const float mm[9] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w, e.x};
            float dist = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f, gv, qq;
            gray_hit_test_local(mm, e.y, e.z, e.w, r.dx, r.dy, r.dz, dist, hx, hy, hz);
            const float al = gray_alpha(hx, hy, hz, f.y, f.x, p, gv, qq);
            dist = fp16_round(dist, A.fp16 != 0);
            const int slot = base + i;
            if (slot >= 0 && slot < cap) put_hit(A, g, slot, dist, __float_as_uint(f.z), al);
            else ok = false;
        }
        n += nh;
        __syncthreads();  // every lane has read the records before the next chunk is staged
    }
    if (single) {
        if (active) A.list_n[g] = ok ? n : -1;
    } else if (!ok) {
        atomicExch(A.list_n + g, kRayDead);
    }
    if (A.tile_stats) {
        int hs = active ? n : 0, ts = tests;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            hs += __shfl_xor(hs, off);
            ts += __shfl_xor(ts, off);
        }
        if (lane == 0) {
            int32_t *st = A.tile_stats + 4 * (size_t)(A.tile_base_launch + tile);
            atomicAdd(st + 1, ts);
            atomicAdd(st + 3, hs);
        }
    }
}

__global__ __launch_bounds__(GRAY_WAVE) void k_collect_binned(TraceArgs A, const GeoRecord *__restrict__ geo) {
    __shared__ float4 s_rec[GRAY_WAVE * 4];
    if (*A.bin_mode != 0) return;
    const int T = A.total_tiles;
    const int items = A.bin_tiles[3 * T + 1];
    const int per = (items + 7) / 8, b = blockIdx.x, stride = gridDim.x >> 3;
    for (int s8 = b >> 3; s8 < per; s8 += stride) {
        const int item = (b & 7) * per + s8;
        if (item >= items) break;
        collect_bin_item(A, geo, s_rec, item);
    }
}

__global__ void k_zero_i32(int32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Pixel outputs of the forward (reference forward_pass.cu:232-268): colour
// with the truncation estimate and background, depth, and the backward state.
// skc/ska/pnc are the skip-set sums (sum a*c, sum a, prod (1-a)).
__device__ __forceinline__ void store_pixel(const TraceArgs &A, const DevView &V, int pix, const LaneRay &r,
                                            bool depth_on, float T, float Tfull, float pnc, float skc0, float skc1,
                                            float skc2, float ska, float C0, float C1, float C2, float D,
                                            uint32_t ncomp, uint32_t nhits) {
    // full transmittance: composited hits (Tfull, fp32 alphas) times the skip set
    const float F = Tfull * pnc;
    const float rem = T - F;
    float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
    if (ska > 0.0f) {
        e0 = skc0 / ska;
        e1 = skc1 / ska;
        e2 = skc2 / ska;
    }
    const float cc0 = C0, cc1 = C1, cc2 = C2;
    C0 += rem * e0;
    C1 += rem * e1;
    C2 += rem * e2;
    if (depth_on) D = D / fmaxf(1.0f - T, *A.knobs.eps_fwd);
    const float *bg = A.knobs.bg;
    C0 += bg[0] * F;
    C1 += bg[1] * F;
    C2 += bg[2] * F;
    V.out_c[3 * pix + 0] = C0;
    V.out_c[3 * pix + 1] = C1;
    V.out_c[3 * pix + 2] = C2;
    // reference Framebuffer::write (framebuffer.h:33-52): everything when the
    // forward saves state; otherwise depth if asked for or if the ray output
    // is, and the ray direction with the ray output
    const bool ray_out = !A.save && A.knobs.needs_ray && *A.knobs.needs_ray;
    if (V.out_depth && (A.save || depth_on || ray_out)) V.out_depth[pix] = D;
    if (ray_out && V.out_rd) {
        V.out_rd[3 * pix + 0] = r.dx;
        V.out_rd[3 * pix + 1] = r.dy;
        V.out_rd[3 * pix + 2] = r.dz;
    }
    if (A.save) {
        if (V.out_T) V.out_T[pix] = T;
        if (V.out_F) V.out_F[pix] = F;
        if (V.out_est) {
            V.out_est[3 * pix + 0] = e0;
            V.out_est[3 * pix + 1] = e1;
            V.out_est[3 * pix + 2] = e2;
        }
        if (V.out_ro) {
            V.out_ro[3 * pix + 0] = r.ox;
            V.out_ro[3 * pix + 1] = r.oy;
            V.out_ro[3 * pix + 2] = r.oz;
        }
        if (V.out_rd) {
            V.out_rd[3 * pix + 0] = r.dx;
            V.out_rd[3 * pix + 1] = r.dy;
            V.out_rd[3 * pix + 2] = r.dz;
        }
        if (V.n_comp) V.n_comp[pix] = ncomp;
        if (V.comp_rgb) {
            V.comp_rgb[3 * pix + 0] = cc0;
            V.comp_rgb[3 * pix + 1] = cc1;
            V.comp_rgb[3 * pix + 2] = cc2;
        }
        if (V.stat_hits) V.stat_hits[pix] += (int)nhits;
        if (V.stat_acc) V.stat_acc[pix] += (int)ncomp;
    }
}

// ---------------------------------------------------------------------------
// Forward by re-traversal: every round walks the BVH again.  With lists on it
// takes only the lanes whose list overflowed.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GRAY_WAVE) void k_trace_forward(TraceArgs A, const GeoRecord *__restrict__ geo,
                                                             const BvhNode *__restrict__ nodes,
                                                             const float *__restrict__ rgb0) {
    __shared__ float s_kd[GRAY_KB * GRAY_WAVE];
    __shared__ uint32_t s_ki[GRAY_KB * GRAY_WAVE];
    __shared__ float s_ka[GRAY_KB * GRAY_WAVE];
    __shared__ uint32_t s_stack[GRAY_STACK * GRAY_WAVE];

    if (A.list_cap > 0 && A.retrace_any && *A.retrace_any == 0) return;  // no list overflowed
    int vi, tx, ty, tile;
    if (!block_tile(A, vi, tx, ty, &tile)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    const bool inside = x < V.width && y < V.height;
    LaneRay r{};
    bool active = false;
    if (inside) active = make_ray(A.knobs, V, x, y, r);
    const bool lists = A.list_cap > 0;
    const size_t lg = (size_t)tile * GRAY_WAVE + lane;
    const int lcount = lists ? A.list_n[lg] : -1;
    if (lists) active = active && lcount < 0;
    if (!__ballot(active)) return;
    int visits[2] = {0, 0};
    int nrounds = 0;

    const float athr = *A.knobs.alpha_threshold, p = *A.knobs.exp_power, tthr = *A.knobs.t_threshold;
    const float znear = *V.znear, zfar = *V.zfar;
    const bool depth_on = *A.knobs.render_depth != 0;
    const int pix = y * V.stride + x;
    const int R = V.ray_count;
    const float *__restrict__ rgb = rgb0 + (V.rgb - A.views[0].rgb);

    KBuf kb{s_kd, s_ki, s_ka, lane};
    const bool h16 = A.fp16 != 0;
    float T = 1.0f, Tf = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, D = 0.0f;
    float skc0 = 0.0f, skc1 = 0.0f, skc2 = 0.0f, ska = 0.0f, pnc = 1.0f;
    float tmin = znear;
    uint32_t ncomp = 0, nhits = 0;
    bool alive = active && tmin < zfar;
    // Round cursor: a reference batch (32 hits strictly beyond tmin,
    // forward_pass.cu:150-182) is served in sub-rounds of GRAY_KB; inside a
    // batch the next sub-round continues after the last (dist, id) key, at a
    // batch boundary it restarts strictly beyond tmin.
    float cur_d = znear;
    uint32_t cur_id = 0;
    bool strict = true;
    int bpos = 0, batch = 0;

    for (int round = 0; round < 2 * GRAY_MAX_ROUNDS * (GRAY_BATCH / GRAY_KB); ++round) {
        const bool part = alive;
        if (!__ballot(part)) break;
        int count = 0;
        float lastd = 0.0f;
        uint32_t lasti = 0;
        const bool first = round == 0;
        auto found = [&](uint32_t gid, float dist, float a, float c0, float c1, float c2) {
            if (first) ++nhits;
            const bool beyond = strict ? dist > cur_d : (dist > cur_d || (dist == cur_d && gid > cur_id));
            int res = 0;
            uint32_t ev_id = 0;
            float ev_a = 0.0f;
            if (beyond) res = kb_insert(kb, count, lastd, lasti, dist, gid, a, ev_id, ev_a);
            if (!first) return;
            if (res == 0) {
                const float ah = fp16_round(a, h16);
                skc0 += ah * c0;
                skc1 += ah * c1;
                skc2 += ah * c2;
                ska += ah;
                pnc = pnc * (1.0f - a);
            } else if (res == 2) {
                const float ah = fp16_round(ev_a, h16);
                skc0 += ah * rgb[3 * ev_id + 0];
                skc1 += ah * rgb[3 * ev_id + 1];
                skc2 += ah * rgb[3 * ev_id + 2];
                ska += ah;
                pnc = pnc * (1.0f - ev_a);
            }
        };
        {
            auto leaf = [&](uint32_t gid, const GeoV &g, float c0, float c1, float c2) {
                if (g.op <= athr) return;
                float dist, hx, hy, hz;
                if (!gray_hit_test(g.m, g.mu[0], g.mu[1], g.mu[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, dist, hx, hy,
                                   hz))
                    return;
                float gv, q;
                const float a = gray_alpha(hx, hy, hz, g.clip, g.op, p, gv, q);
                found(gid, fp16_round(dist, h16), a, c0, c1, c2);
            };
            traverse(A, nodes, geo, rgb, first, r, part, first ? 0.0f : cur_d, s_stack, lane, leaf,
                     A.tile_stats ? visits : nullptr);
        }
        ++nrounds;

        if (part) {
            bool terminated = false;
            for (int k = 0; k < count; ++k) {
                const float d = kb.dist(k);
                const uint32_t gid = kb.gid(k);
                const float a32 = kb.alpha(k);
                const float a = fp16_round(a32, h16);
                if (terminated) {
                    if (first) {
                        skc0 += a * rgb[3 * gid + 0];
                        skc1 += a * rgb[3 * gid + 1];
                        skc2 += a * rgb[3 * gid + 2];
                        ska += a;
                        pnc = pnc * (1.0f - a32);
                    }
                    continue;
                }
                tmin = fmaxf(d, tmin);
                const float c0 = rgb[3 * gid + 0], c1 = rgb[3 * gid + 1], c2 = rgb[3 * gid + 2];
                if (d < zfar) {
                    if (!first) {
                        skc0 -= a * c0;
                        skc1 -= a * c1;
                        skc2 -= a * c2;
                        ska -= a;
                        pnc = pnc / (1.0f - a32);
                    }
                    const float w = T * a;
                    C0 += c0 * w;
                    C1 += c1 * w;
                    C2 += c2 * w;
                    if (depth_on) D += d * w;
                    if (A.save && (int)ncomp < A.capacity) {
                        V.hit_ids[(size_t)ncomp * R + pix] = gid;
                        V.hit_T[(size_t)ncomp * R + pix] = T;
                    }
                    T = T * (1.0f - a);
                    Tf = Tf * (1.0f - a32);
                    ++ncomp;
                    if (T < tthr) {
                        tmin = zfar;
                        terminated = true;
                    }
                } else if (first) {
                    skc0 += a * c0;
                    skc1 += a * c1;
                    skc2 += a * c2;
                    ska += a;
                    pnc = pnc * (1.0f - a32);
                }
            }
            alive = false;
            if (!terminated) {
                if (count < GRAY_KB) {
                    tmin = FLT_MAX;  // a partial batch ends the loop (reference: FLT_MAX slots)
                } else {
                    bpos += GRAY_KB;
                    if (bpos >= GRAY_BATCH) {
                        bpos = 0;
                        strict = true;
                        cur_d = tmin;
                        ++batch;
                    } else {
                        strict = false;
                        cur_d = kb.dist(GRAY_KB - 1);
                        cur_id = kb.gid(GRAY_KB - 1);
                    }
                    alive = tmin < zfar && batch < GRAY_MAX_ROUNDS;
                }
            }
        }
    }

    if (A.tile_stats) {
        int hsum = (int)nhits;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            hsum += __shfl_xor(hsum, off);
            visits[0] += __shfl_xor(visits[0], off);
            visits[1] += __shfl_xor(visits[1], off);
        }
        if (lane == 0) {
            int32_t *ts = A.tile_stats + 4 * (size_t)(A.tile_base_launch + tile);
            if (!lists) {  // else the collect and list kernels write them
                ts[0] = visits[0];
                ts[1] = visits[1];
                ts[2] = nrounds;
                ts[3] = hsum;
            }
        }
    }
    if (!active) return;
    store_pixel(A, V, pix, r, depth_on, T, Tf, pnc, skc0, skc1, skc2, ska, C0, C1, C2, D, ncomp, nhits);
}

// ---------------------------------------------------------------------------
// Forward from the collected lists (lanes whose list is complete)
// ---------------------------------------------------------------------------
constexpr uint32_t kListMark = 0x80000000u;  // list id flag: entry was composited (entries >= kMaskEntries)
constexpr int kMaskEntries = 128;             // entries whose flag is a register bit instead
constexpr int kListBatch = 8;                 // list entries loaded per step of a scan
constexpr int kWavePer = 8;                   // list entries per lane of the wave-per-ray kernel
constexpr int kWaveListMax = GRAY_WAVE * kWavePer;
#ifndef GRAY_WAVE_LIST_MIN
#define GRAY_WAVE_LIST_MIN 96
#endif

// Sub-round size of the list compositing (a divisor of the reference batch of
// 32): 8 rather than the re-traversal's 16, since most rays composite fewer
// than 8 hits and every list entry of a scan is merged into all slots
// (composite 0.51 -> 0.49 ms per training frame; 4: 0.52).
constexpr int kListKB = 8;
static_assert(GRAY_BATCH % kListKB == 0, "list sub-rounds divide the batch");

// The kListKB nearest entries of a scan, sorted by (distance, id), held in
// registers (every index is a compile-time constant after unrolling).  Empty
// slots hold (+inf, 0xffffffff), so "x precedes the last slot" is also the
// test for a free slot.  Insertion is a branchless merge: with lt[k] = x <
// slot k (false...false, true...true on a sorted array), slot k becomes
// lt[k-1] ? slot k-1 : (lt[k] ? x : slot k).
struct RegKB {
    float d[kListKB];
    uint32_t id[kListKB];
    uint32_t slot[kListKB];

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int k = 0; k < kListKB; ++k) {
            d[k] = INFINITY;
            id[k] = 0xFFFFFFFFu;
            slot[k] = 0;
        }
    }
    __device__ __forceinline__ bool takes(float xd, uint32_t xi) const {
        return key_less(xd, xi, d[kListKB - 1], id[kListKB - 1]);
    }
    __device__ __forceinline__ void insert(float xd, uint32_t xi, uint32_t xs) {
        bool lt[kListKB];
#pragma unroll
        for (int k = 0; k < kListKB; ++k) lt[k] = key_less(xd, xi, d[k], id[k]);
#pragma unroll
        for (int k = kListKB - 1; k > 0; --k) {
            d[k] = lt[k - 1] ? d[k - 1] : (lt[k] ? xd : d[k]);
            id[k] = lt[k - 1] ? id[k - 1] : (lt[k] ? xi : id[k]);
            slot[k] = lt[k - 1] ? slot[k - 1] : (lt[k] ? xs : slot[k]);
        }
        d[0] = lt[0] ? xd : d[0];
        id[0] = lt[0] ? xi : id[0];
        slot[0] = lt[0] ? xs : slot[0];
    }
};

// Long lists (wave_list_min < n <= kWaveListMax) go to the queue of the
// wave-per-ray kernel: one queue slot per lane, taken with one atomic per
// wave.  Rays outside the image or inactive have empty lists.
__device__ __forceinline__ bool heavy_list(const TraceArgs &A, int n) {
    return A.wave_list_min >= 0 && n > A.wave_list_min && n <= kWaveListMax;
}

__global__ __launch_bounds__(GRAY_WAVE) void k_heavy_queue(TraceArgs A) {
    int vi, tx, ty, tile;
    if (!block_tile(A, vi, tx, ty, &tile)) return;
    const int lane = threadIdx.x;
    const size_t lg = (size_t)tile * GRAY_WAVE + lane;
    const bool heavy = heavy_list(A, A.list_n[lg]);
    const uint64_t hb = __ballot(heavy);
    if (!hb) return;
    const int first = __ffsll((long long)hb) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(A.heavy_count, __popcll(hb));
    base = __shfl(base, first);
    if (heavy) A.heavy_rays[base + __popcll(hb & ((1ull << lane) - 1ull))] = (uint32_t)lg;
}

// Same batch semantics as the re-traversal forward, with the list as the hit
// source: a round scans the lane's list (loads issued kListBatch entries at a
// time) into a register k-buffer keyed by (distance, id) with the list slot
// as payload, then composites front to back and flags every composited
// entry: a bit of a 128-bit register mask for the first 128 entries, the id's
// top bit in the list record for later ones.  After the last round one more
// scan sums the unflagged entries, which are exactly the skip set (hits the
// k-buffer never handed to compositing, hits beyond zfar, hits after
// termination).
//
// phase 0 takes every tile; phase 1 only the tiles whose lists the per-tile
// walk finished (it runs beside the pooled task round), phase 2 the others.
__global__ __launch_bounds__(GRAY_WAVE) void k_composite_lists(TraceArgs A, const float *__restrict__ rgb0,
                                                               int phase) {
    int vi, tx, ty, tile;
    if (!block_tile(A, vi, tx, ty, &tile)) return;
    if (phase != 0 && (A.tile_open[tile] != 0) != (phase == 2)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    LaneRay r{};
    bool active = false;
    if (x < V.width && y < V.height) active = make_ray(A.knobs, V, x, y, r);
    const size_t lg = (size_t)tile * GRAY_WAVE + lane;
    const int n = A.list_n[lg];
    // a ray whose list overflowed is re-traversed by k_trace_forward, which
    // returns at once unless some wave raised this flag
    if (__ballot(active && n < 0) && lane == 0) atomicOr(A.retrace_any, 1);
    active = active && n >= 0;
    // long lists are composited by the wave-per-ray kernel (queued by k_heavy_queue)
    active = active && !heavy_list(A, n);
    if (!__ballot(active)) return;

    float4 *__restrict__ L = A.list_rec;
    const float tthr = *A.knobs.t_threshold;
    const float znear = *V.znear, zfar = *V.zfar;
    const bool depth_on = *A.knobs.render_depth != 0;
    const int pix = y * V.stride + x;
    const int R = V.ray_count;
    const float *__restrict__ rgb = rgb0 + (V.rgb - A.views[0].rgb);

    RegKB kb;
    const bool h16 = A.fp16 != 0;
    float T = 1.0f, Tf = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, D = 0.0f;
    float tmin = znear;
    uint32_t ncomp = 0;
    uint64_t mask0 = 0, mask1 = 0;  // composited entries 0-63, 64-127
    const bool traced = active && tmin < zfar;
    bool alive = traced;
    // round cursor, as in the re-traversal forward
    float cur_d = znear;
    uint32_t cur_id = 0;
    bool strict = true;
    int bpos = 0, batch = 0, nrounds = 0;

    for (int round = 0; round < 2 * GRAY_MAX_ROUNDS * (GRAY_BATCH / kListKB); ++round) {
        if (!__ballot(alive)) break;
        ++nrounds;
        if (!alive) continue;
        kb.clear();
        int count = 0;
        // software-pipelined scan: the next batch is in flight while this one
        // is inserted
        float2 nx[kListBatch];
        auto fetch = [&](int i0) {  // unconditional loads (clamped index): exact wait counts
#pragma unroll
            for (int j = 0; j < kListBatch; ++j) {
                const float4 rec = L[list_at(A, lg, min(i0 + j, n - 1))];
                nx[j] = i0 + j < n ? make_float2(rec.x, rec.y) : make_float2(INFINITY, __uint_as_float(0xFFFFFFFFu));
            }
        };
        if (n > 0) fetch(0);
        for (int i0 = 0; i0 < n; i0 += kListBatch) {
            float d[kListBatch];
            uint32_t id[kListBatch];
#pragma unroll
            for (int j = 0; j < kListBatch; ++j) {
                d[j] = nx[j].x;
                id[j] = __float_as_uint(nx[j].y) & ~kListMark;
            }
            if (i0 + kListBatch < n) fetch(i0 + kListBatch);
#pragma unroll
            for (int j = 0; j < kListBatch; ++j) {
                const bool beyond = strict ? d[j] > cur_d : (d[j] > cur_d || (d[j] == cur_d && id[j] > cur_id));
                if (i0 + j < n && beyond && kb.takes(d[j], id[j])) {
                    kb.insert(d[j], id[j], (uint32_t)(i0 + j));
                    count = count < kListKB ? count + 1 : kListKB;
                }
            }
        }
        bool terminated = false;
#pragma unroll
        for (int k = 0; k < kListKB; ++k) {
            const float dk = kb.d[k];
            if (k >= count || terminated) continue;
            tmin = fmaxf(dk, tmin);
            if (!(dk < zfar)) continue;
            const uint32_t gid = kb.id[k];
            const size_t e = list_at(A, lg, (int)kb.slot[k]);
            const float a32 = L[e].z;
            const float a = fp16_round(a32, h16);
            const float c0 = rgb[3 * gid + 0], c1 = rgb[3 * gid + 1], c2 = rgb[3 * gid + 2];
            const float w = T * a;
            C0 += c0 * w;
            C1 += c1 * w;
            C2 += c2 * w;
            if (depth_on) D += dk * w;
            if (A.save && (int)ncomp < A.capacity) {
                V.hit_ids[(size_t)ncomp * R + pix] = gid;
                V.hit_T[(size_t)ncomp * R + pix] = T;
            }
            const uint32_t sl = kb.slot[k];
            if (sl < (uint32_t)kMaskEntries) {
                const uint64_t bit = 1ull << (sl & 63u);
                mask0 |= sl < 64u ? bit : 0ull;
                mask1 |= sl < 64u ? 0ull : bit;
            } else {
                reinterpret_cast<uint32_t *>(L + e)[1] = gid | kListMark;
            }
            T = T * (1.0f - a);
            Tf = Tf * (1.0f - a32);
            ++ncomp;
            if (T < tthr) {
                tmin = zfar;
                terminated = true;
            }
        }
        alive = false;
        if (!terminated) {
            if (count < kListKB) {
                tmin = FLT_MAX;  // a partial batch ends the loop (reference: FLT_MAX slots)
            } else {
                bpos += kListKB;
                if (bpos >= GRAY_BATCH) {
                    bpos = 0;
                    strict = true;
                    cur_d = tmin;
                    ++batch;
                } else {
                    strict = false;
                    cur_d = kb.d[kListKB - 1];
                    cur_id = kb.id[kListKB - 1];
                }
                alive = tmin < zfar && batch < GRAY_MAX_ROUNDS;
            }
        }
    }

    // skip set: every entry compositing did not flag.  Pipelined like the
    // scan: the colours of this batch are gathered (unconditionally; flagged
    // and padding entries get weight 0, which leaves the sums bit for bit as
    // if skipped) while the next batch's ids and alphas are in flight.
    float skc0 = 0.0f, skc1 = 0.0f, skc2 = 0.0f, ska = 0.0f, pnc = 1.0f;
    // (empty when every entry was composited)
    const int nskip = traced && (int)ncomp < n ? n : 0;
    float2 sx[kListBatch];
    auto fetch_skip = [&](int i0) {
#pragma unroll
        for (int j = 0; j < kListBatch; ++j) {
            const float4 rec = L[list_at(A, lg, min(i0 + j, nskip - 1))];
            sx[j] = make_float2(rec.y, i0 + j < nskip ? rec.z : 0.0f);
        }
    };
    if (nskip > 0) fetch_skip(0);
    for (int i0 = 0; i0 < nskip; i0 += kListBatch) {
        float w[kListBatch], c[kListBatch][3];
#pragma unroll
        for (int j = 0; j < kListBatch; ++j) {
            const uint32_t id = __float_as_uint(sx[j].x);
            const int i = i0 + j;
            const uint64_t mw = i < 64 ? mask0 : mask1;
            const bool marked = i < kMaskEntries ? ((mw >> (i & 63)) & 1ull) != 0 : (id & kListMark) != 0;
            w[j] = marked ? 0.0f : sx[j].y;  // fp32 alpha; its fp16 rounding enters the sums below
            // entries of weight 0 read Gaussian 0's colour (a cached line)
            // instead of gathering one they do not use
            const uint32_t gid = w[j] != 0.0f ? (id & ~kListMark) : 0u;
            c[j][0] = rgb[3 * gid + 0];
            c[j][1] = rgb[3 * gid + 1];
            c[j][2] = rgb[3 * gid + 2];
        }
        if (i0 + kListBatch < nskip) fetch_skip(i0 + kListBatch);
#pragma unroll
        for (int j = 0; j < kListBatch; ++j) {
            if (w[j] == 0.0f) continue;
            const float wh = fp16_round(w[j], h16);
            skc0 += wh * c[j][0];
            skc1 += wh * c[j][1];
            skc2 += wh * c[j][2];
            ska += wh;
            pnc = pnc * (1.0f - w[j]);
        }
    }
    if (A.tile_stats && lane == 0) A.tile_stats[4 * (size_t)(A.tile_base_launch + tile) + 2] = nrounds;
    if (!active) return;
    store_pixel(A, V, pix, r, depth_on, T, Tf, pnc, skc0, skc1, skc2, ska, C0, C1, C2, D, ncomp,
                (uint32_t)(traced ? n : 0));
}

// ---------------------------------------------------------------------------
// Forward from long lists: one wave per ray
// ---------------------------------------------------------------------------
// Wave reductions with DPP: steps within 16-lane rows, row broadcasts across
// them; lane 63 ends with the total, which is returned to every lane (the
// tree, and so the rounding, is fixed).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, kCtrl, kRowMask, 0xF, false);
}

template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_f32(float v) {
    return __uint_as_float(dpp_u32<kCtrl, kRowMask>(__float_as_uint(v)));
}

// (d, id) < (d', id') without branches
__device__ __forceinline__ bool key_less_nb(float d0, uint32_t i0, float d1, uint32_t i1) {
    return (d0 < d1) | ((d0 == d1) & (i0 < i1));
}

template <int kCtrl, int kRowMask>
__device__ __forceinline__ void dpp_min_step(float &d, uint32_t &id) {
    const float od = dpp_f32<kCtrl, kRowMask>(d);
    const uint32_t oi = dpp_u32<kCtrl, kRowMask>(id);
    const bool take = key_less_nb(od, oi, d, id);
    d = take ? od : d;
    id = take ? oi : id;
}

// minimum key (d, id) over the wave
__device__ __forceinline__ void wave_min_key(float &d, uint32_t &id) {
    dpp_min_step<0xB1, 0xF>(d, id);   // quad_perm [1,0,3,2]
    dpp_min_step<0x4E, 0xF>(d, id);   // quad_perm [2,3,0,1]
    dpp_min_step<0x141, 0xF>(d, id);  // row_half_mirror
    dpp_min_step<0x140, 0xF>(d, id);  // row_mirror
    dpp_min_step<0x142, 0xA>(d, id);  // row_bcast:15 into rows 1, 3
    dpp_min_step<0x143, 0xC>(d, id);  // row_bcast:31 into rows 2, 3
    d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d), 63));
    id = (uint32_t)__builtin_amdgcn_readlane((int)id, 63);
}

template <bool kProd>
__device__ __forceinline__ float wave_reduce(float v) {
    auto op = [](float a, float b) { return kProd ? a * b : a + b; };
    v = op(v, dpp_f32<0xB1, 0xF>(v));
    v = op(v, dpp_f32<0x4E, 0xF>(v));
    v = op(v, dpp_f32<0x141, 0xF>(v));
    v = op(v, dpp_f32<0x140, 0xF>(v));
    v = op(v, dpp_f32<0x142, 0xA>(v));
    v = op(v, dpp_f32<0x143, 0xC>(v));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// The per-lane kernel's batch semantics for one ray per wave.  Each lane holds
// up to kWavePer of the ray's list entries, their keys (distance, id) sorted in
// registers and their alpha and colour in LDS; the wave takes the nearest
// remaining entry (a wave minimum over the lanes' first keys) one at a time
// and composites it exactly as the per-lane kernel does, so composited hits,
// transmittances and colour sums are the same operations in the same order.
// The batch rules follow from the sorted order: a round of GRAY_KB selects the
// next entries, and after every GRAY_BATCH selected entries those at the same
// distance are dropped (strict cursor).  The skip set (entries not composited)
// is summed per lane and then over the wave.  Rays come from the queue the
// per-lane kernel filled.
__global__ __launch_bounds__(GRAY_WAVE) void k_composite_wave(TraceArgs A, const float *__restrict__ rgb0) {
    constexpr int P = kWavePer;
    __shared__ float4 stage[P][GRAY_WAVE];  // {alpha (fp32), colour} of entry j of each lane
    const int lane = threadIdx.x;
    const int total = *A.heavy_count;
    const float4 *__restrict__ L = A.list_rec;
    const float tthr = *A.knobs.t_threshold;
    const bool depth_on = *A.knobs.render_depth != 0;
    const bool h16 = A.fp16 != 0;
    // XCD-aware: the queue (rays in roughly tile order) is cut into 8
    // contiguous segments and workgroup b, which runs on XCD b % 8, takes
    // rays of segment b % 8.  Neighbouring rays' entries share the 128-byte
    // lines of the slot-major list, so a line one wave fetches serves the
    // waves of its neighbours from the same L2.
    const int seg = (total + 7) >> 3;
    const int xcd = blockIdx.x & 7, per_xcd = (int)(gridDim.x >> 3);
    for (int k = (int)(blockIdx.x >> 3); k < seg; k += per_xcd) {
        const int q = xcd * seg + k;
        if (q >= total) break;
        const uint32_t g = A.heavy_rays[q];
        const int n = A.list_n[g];
        float d[P], a[P];
        uint32_t id[P], js[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int i = lane + GRAY_WAVE * j;
            d[j] = INFINITY;
            id[j] = 0xFFFFFFFFu;
            a[j] = 0.0f;
            js[j] = (uint32_t)j;
            if (i < n) {
                const float4 rec = L[list_at(A, g, i)];
                d[j] = rec.x;
                id[j] = __float_as_uint(rec.y) & ~kListMark;
                a[j] = rec.z;
            }
        }
        int x, y;
        const DevView &V = lane_pixel(A, g, x, y);
        const float *__restrict__ rgb = rgb0 + (V.rgb - A.views[0].rgb);
        // colours in flight while the keys are sorted
        float c[P][3];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            c[j][0] = c[j][1] = c[j][2] = 0.0f;
            if (lane + GRAY_WAVE * j < n) {
                c[j][0] = rgb[3 * id[j] + 0];
                c[j][1] = rgb[3 * id[j] + 1];
                c[j][2] = rgb[3 * id[j] + 2];
            }
        }
        // 19-comparator sorting network for 8 (padding sorts last); the
        // original index travels with the key
        auto cx = [&](int i, int k) {
            const bool sw = key_less_nb(d[k], id[k], d[i], id[i]);
            const float td = d[i];
            const uint32_t ti = id[i], tj = js[i];
            d[i] = sw ? d[k] : d[i], id[i] = sw ? id[k] : id[i], js[i] = sw ? js[k] : js[i];
            d[k] = sw ? td : d[k], id[k] = sw ? ti : id[k], js[k] = sw ? tj : js[k];
        };
        static_assert(P == 8, "sorting network for 8 entries per lane");
        cx(0, 2), cx(1, 3), cx(4, 6), cx(5, 7);
        cx(0, 4), cx(1, 5), cx(2, 6), cx(3, 7);
        cx(0, 1), cx(2, 3), cx(4, 5), cx(6, 7);
        cx(2, 4), cx(3, 5);
        cx(1, 4), cx(3, 6);
        cx(1, 2), cx(3, 4), cx(5, 6);
#pragma unroll
        for (int j = 0; j < P; ++j) stage[j][lane] = make_float4(a[j], c[j][0], c[j][1], c[j][2]);
        __syncthreads();
        LaneRay r{};
        make_ray(A.knobs, V, x, y, r);
        const float znear = *V.znear, zfar = *V.zfar;
        const int pix = y * V.stride + x;
        const int R = V.ray_count;
        int cnt = n > lane ? min(P, (n - lane + GRAY_WAVE - 1) / GRAY_WAVE) : 0;
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f, sp = 1.0f;
        auto skip_add = [&](uint32_t j) {
            const float4 e = stage[j][lane];
            if (e.x == 0.0f) return;
            const float wh = fp16_round(e.x, h16);  // as the per-lane kernel's skip scan
            s0 += wh * e.y;
            s1 += wh * e.z;
            s2 += wh * e.w;
            sa += wh;
            sp = sp * (1.0f - e.x);
        };
        auto pop = [&]() {
#pragma unroll
            for (int j = 0; j + 1 < P; ++j) d[j] = d[j + 1], id[j] = id[j + 1], js[j] = js[j + 1];
            d[P - 1] = INFINITY, id[P - 1] = 0xFFFFFFFFu, js[P - 1] = 0u;
            --cnt;
        };
        // the same for the lanes where `on` holds, as selects in every lane
        // (no branch: the compiler would copy the whole list on both paths)
        auto pop_if = [&](bool on) {
#pragma unroll
            for (int j = 0; j + 1 < P; ++j) {
                d[j] = on ? d[j + 1] : d[j];
                id[j] = on ? id[j + 1] : id[j];
                js[j] = on ? js[j + 1] : js[j];
            }
            d[P - 1] = on ? INFINITY : d[P - 1];
            id[P - 1] = on ? 0xFFFFFFFFu : id[P - 1];
            js[P - 1] = on ? 0u : js[P - 1];
            cnt -= on ? 1 : 0;
        };
        float T = 1.0f, Tf = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, D = 0.0f, tmin = znear;
        uint32_t ncomp = 0;
        const bool traced = tmin < zfar;
        if (traced) {
            // the first cursor is strict: entries at or before znear are never selected
            for (int j = 0; j < P && cnt > 0 && !(d[0] > znear); ++j) {
                skip_add(js[0]);
                pop();
            }
            int in_round = 0, bpos = 0, batch = 0;
            for (int step = 0; step < n; ++step) {
                float md = cnt > 0 ? d[0] : INFINITY;
                uint32_t mi = cnt > 0 ? id[0] : 0xFFFFFFFFu;
                wave_min_key(md, mi);
                if (mi == 0xFFFFFFFFu) break;  // exhausted: the last round was partial
                const uint64_t own = __ballot(cnt > 0 && id[0] == mi);
                const int owner = __ffsll((long long)own) - 1;
                const uint32_t oj = (uint32_t)__builtin_amdgcn_readlane((int)js[0], owner);
                const float4 e = stage[oj][owner];
                tmin = fmaxf(md, tmin);
                if (!(md < zfar)) break;  // this entry and every later one stay in the skip set
                const float a32 = e.x;
                const float av = fp16_round(a32, h16);
                const float w = T * av;
                C0 += e.y * w;
                C1 += e.z * w;
                C2 += e.w * w;
                if (depth_on) D += md * w;
                if (lane == 0 && A.save && (int)ncomp < A.capacity) {
                    V.hit_ids[(size_t)ncomp * R + pix] = mi;
                    V.hit_T[(size_t)ncomp * R + pix] = T;
                }
                T = T * (1.0f - av);
                Tf = Tf * (1.0f - a32);
                ++ncomp;
                pop_if(lane == owner);
                if (T < tthr) break;
                if (++in_round == GRAY_KB) {
                    in_round = 0;
                    bpos += GRAY_KB;
                    if (bpos >= GRAY_BATCH) {
                        bpos = 0;
                        ++batch;
                        for (int j = 0; j < P && cnt > 0 && !(d[0] > tmin); ++j) {
                            skip_add(js[0]);
                            pop();
                        }
                        if (batch >= GRAY_MAX_ROUNDS) break;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < P; ++j)
                if (j < cnt) skip_add(js[j]);
        }
        s0 = wave_reduce<false>(s0);
        s1 = wave_reduce<false>(s1);
        s2 = wave_reduce<false>(s2);
        sa = wave_reduce<false>(sa);
        sp = wave_reduce<true>(sp);
        if (lane == 0)
            store_pixel(A, V, pix, r, depth_on, T, Tf, sp, s0, s1, s2, sa, C0, C1, C2, D, ncomp,
                        traced ? (uint32_t)n : 0u);
        __syncthreads();  // the next ray overwrites the stage
    }
}

// ---------------------------------------------------------------------------
// Backward: gradient of one composited hit (reference backward_pass.cu:15-141)
// ---------------------------------------------------------------------------
struct PixelState {
    float g0, g1, g2;      // dL/d(out)
    float Tf, F;           // final and full transmittance
    float e0, e1, e2;      // truncation estimate
    float bgdot, estdot;   // bg.g, est.g
    float pixels;          // width*height of the launch (pruning weight normaliser)
};

// Gradient terms of one hit, in the order of kGradTerms: colour 3, rotation 4,
// scale 3, mean 3, opacity 1, pruning weight 1.
constexpr int kGradTerms = 15;

// Adds the gradient terms of one Gaussian to the reference gradient tensors.
__device__ __forceinline__ void grad_to_global(const TraceArgs &A, const DevView &V, uint32_t gid, const float *t) {
    atomicAdd(&V.grad_rgb[3 * gid + 0], t[0]);
    atomicAdd(&V.grad_rgb[3 * gid + 1], t[1]);
    atomicAdd(&V.grad_rgb[3 * gid + 2], t[2]);
    atomicAdd(&A.g_rot[4 * gid + 0], t[3]);
    atomicAdd(&A.g_rot[4 * gid + 1], t[4]);
    atomicAdd(&A.g_rot[4 * gid + 2], t[5]);
    atomicAdd(&A.g_rot[4 * gid + 3], t[6]);
    atomicAdd(&A.g_scale[3 * gid + 0], t[7]);
    atomicAdd(&A.g_scale[3 * gid + 1], t[8]);
    atomicAdd(&A.g_scale[3 * gid + 2], t[9]);
    atomicAdd(&A.g_mean[3 * gid + 0], t[10]);
    atomicAdd(&A.g_mean[3 * gid + 1], t[11]);
    atomicAdd(&A.g_mean[3 * gid + 2], t[12]);
    atomicAdd(&A.g_op[gid], t[13]);
    if (A.g_prune) atomicAdd(&A.g_prune[gid], t[14]);
    if (A.visible) A.visible[gid] = 1;
}

// Deterministic mode: the terms of one composited hit go to its record (see
// TraceArgs::det_rec); k_det_reduce sums them per Gaussian in record order.
__device__ __forceinline__ void det_put(const TraceArgs &A, size_t idx, uint32_t gid, const float *t) {
    f4v *o = reinterpret_cast<f4v *>(A.det_rec + 16 * idx);
    o[0] = f4v{t[0], t[1], t[2], t[3]};
    o[1] = f4v{t[4], t[5], t[6], t[7]};
    o[2] = f4v{t[8], t[9], t[10], t[11]};
    o[3] = f4v{t[12], t[13], t[14], __uint_as_float(gid)};
    A.det_keys[idx] = gid;
}

__device__ __forceinline__ void geometry_terms(const TraceArgs &A, uint32_t gid, float clip, float eps_sg,
                                               const float gl[3][3], float gxw0, float gxw1, float gxw2, float *t);

// Raw per-hit sums of the stored-list backward.  Every term of a hit's
// geometric gradient is linear in two per-hit quantities scaled by per-Gaussian
// constants: with c = ga * dg/d(q) (q = |l|^2, l the local hit point times
// clip) and K = MAX_ALPHA * sigmoid(opacity), the world hit gradient is
// gxw = -K clip M^T (c l) and the L2W-row gradient gl = K clip M^T (c l l^T)
// (M = W2L).  The table therefore sums per Gaussian: colour g w (3), ga g (1),
// w (1), v = c l (3) and the symmetric S = c l l^T (xx xy xz yy yz zz, 6), and
// the per-Gaussian transform to rotation, scale, mean and opacity gradients
// runs once per Gaussian and tile at the flush instead of once per hit.
constexpr int kRawTerms = 14;

// Per-wave accumulator in LDS: the rays of a tile composite largely the same
// Gaussians (about 42 distinct ones for 367 hits in an 8x8 tile of a 1M
// frame), so their raw terms are summed here (ds_add_f32) and each Gaussian
// reaches the global tensors once per tile.  Open addressing with a bounded
// probe; a Gaussian that finds no slot is transformed and written at once.
// Terms are stored [term][slot].
constexpr int kGradSlots = 128;

// Rotation, scale, mean, opacity and pruning terms (t[3..14] of the
// kGradTerms order) of one Gaussian from its raw geometric sums r (the row
// layout of TraceArgs::raw: ga*g, pruning weight, v 3, S xx xy xz yy yz zz).
__device__ __forceinline__ void raw_geometry_terms(const TraceArgs &A, uint32_t gid, float eps_sg, const float *r,
                                                   float *t) {
#pragma clang fp contract(off)
    const f4v *g4 = reinterpret_cast<const f4v *>(A.geo + gid);
    const f4v ga = g4[0], gb = g4[1], gc = g4[2], gd = g4[3];
    const float m[9] = {ga[0], ga[1], ga[2], ga[3], gb[0], gb[1], gb[2], gb[3], gc[0]};
    const float clip = gd[1];
    const float op = gray_sigmoid(A.opacity[gid]);
    const float kc = (GRAY_MAX_ALPHA * op) * clip;
    // gl[i][j] = K clip sum_a M[a][i] S[a][j]
    const float S[3][3] = {{r[5], r[6], r[7]}, {r[6], r[8], r[9]}, {r[7], r[9], r[10]}};
    float gl[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) gl[i][j] = kc * ((m[i] * S[0][j] + m[3 + i] * S[1][j]) + m[6 + i] * S[2][j]);
    const float gxw0 = -kc * ((m[0] * r[2] + m[3] * r[3]) + m[6] * r[4]);
    const float gxw1 = -kc * ((m[1] * r[2] + m[4] * r[3]) + m[7] * r[4]);
    const float gxw2 = -kc * ((m[2] * r[2] + m[5] * r[3]) + m[8] * r[4]);
    geometry_terms(A, gid, clip, eps_sg, gl, gxw0, gxw1, gxw2, t);
    t[13] = ((GRAY_MAX_ALPHA * r[0]) * op) * (1.0f - op);
    t[14] = r[1];
}

// The raw sums s (kRawTerms order) of one Gaussian to the device: the colour
// sums to the view's colour gradient, the geometric sums to the Gaussian's raw
// row (one 64-byte line), the pruning weight normalised by the view's pixels.
__device__ __forceinline__ void raw_to_global(const TraceArgs &A, const DevView &V, uint32_t gid, const float *s,
                                              float pixels) {
    atomicAdd(&V.grad_rgb[3 * gid + 0], s[0]);
    atomicAdd(&V.grad_rgb[3 * gid + 1], s[1]);
    atomicAdd(&V.grad_rgb[3 * gid + 2], s[2]);
    float *row = A.raw + 16 * (size_t)gid;
    atomicAdd(&row[0], s[3]);
    atomicAdd(&row[1], s[4] / pixels);
#pragma unroll
    for (int j = 5; j < kRawTerms; ++j) atomicAdd(&row[j - 3], s[j]);
    A.raw_touched[gid] = 1;
}

// floats per slot row: 14 terms + 6 pad.  A stride of 80 B (20 banks) starts
// the rows of 16 consecutive slots on 16 different 4-bank groups, so the
// 16-byte accesses of lanes holding different slots spread over all banks.
constexpr int kRawRow = 20;
constexpr int kClaimProbes = 8;  // slots tried by a claim
constexpr int kRecList = 4;      // records listed per Gaussian (stored-list backward)

// The table of one wave.  It has no path to global memory inside the hit
// loop: a store or atomic there (even on a branch never taken) makes the
// compiler wait for every load in flight at each step, which would undo the
// loop's prefetch.  A Gaussian that finds no slot (the tile holds about as
// many distinct Gaussians as the table has slots) fails the add; the hits
// concerned are added one by one after the loop.
struct RawTable {
    uint32_t *key;   // [kGradSlots], GRAY_EMPTY = free
    float *val;      // [kGradSlots][kRawRow]
    uint16_t *list;  // [kGradSlots] occupied slots, compacted at the flush
    uint32_t *owner; // [kGradSlots] the lane that updates a slot in a step

    __device__ __forceinline__ void clear(int lane) {
        for (int i = lane; i < kGradSlots; i += GRAY_WAVE) key[i] = GRAY_EMPTY;
        for (int i = lane; i < kGradSlots * kRawRow; i += GRAY_WAVE) val[i] = 0.0f;
    }
    // The slot of gid (claimed if new), or -1 when a bounded probe finds none
    // (the table is nearly full: a full table would cost every later new
    // Gaussian a scan of all slots).
    __device__ __forceinline__ int claim(uint32_t gid) const {
        uint32_t slot = (gid * 2654435761u) >> 25;  // 7 bits
        for (int probe = 0; probe < kClaimProbes; ++probe) {
            // most hits find their Gaussian already in the table: a plain
            // read first, the compare-and-swap only for a free slot
            // (a plain read: a volatile one would lose the LDS address space
            // and become a flat access that waits for every load in flight)
            uint32_t k = key[slot];
            if (k == gid) return (int)slot;
            if (k == GRAY_EMPTY) k = atomicCAS(&key[slot], GRAY_EMPTY, gid);
            if (k == GRAY_EMPTY || k == gid) return (int)slot;
            slot = (slot + 1) & (kGradSlots - 1);
        }
        return -1;
    }
    // The slot holding gid, or -1 (read-only: every key is in place).
    __device__ __forceinline__ int lookup(uint32_t gid) const {
        uint32_t slot = (gid * 2654435761u) >> 25;
        for (int probe = 0; probe < kClaimProbes; ++probe) {
            const uint32_t k = key[slot];
            if (k == gid) return (int)slot;
            if (k == GRAY_EMPTY) return -1;
            slot = (slot + 1) & (kGradSlots - 1);
        }
        return -1;
    }
    // Called by the whole wave (uniform control flow); lanes with slot < 0
    // take no part.  One lane per slot and step updates the slot's row by
    // plain 16-byte reads and writes; lanes that share a slot with another
    // lane (the same Gaussian away from the neighbours combined before) take
    // later rounds.  The workgroup is one wave, so its LDS operations execute
    // in order and the barriers below cost no wait on other waves.
    __device__ __forceinline__ void add_at(int slot, const float *s, int lane) {
        bool pend = slot >= 0;
        while (__ballot(pend)) {
            if (pend) owner[slot] = (uint32_t)lane;
            __syncthreads();
            const bool own = pend && owner[slot] == (uint32_t)lane;
            if (own) {
                f4v *row = reinterpret_cast<f4v *>(val + slot * kRawRow);
                const f4v a = row[0], b = row[1], c = row[2], d = row[3];
                row[0] = f4v{a[0] + s[0], a[1] + s[1], a[2] + s[2], a[3] + s[3]};
                row[1] = f4v{b[0] + s[4], b[1] + s[5], b[2] + s[6], b[3] + s[7]};
                row[2] = f4v{c[0] + s[8], c[1] + s[9], c[2] + s[10], c[3] + s[11]};
                row[3] = f4v{d[0] + s[12], d[1] + s[13], d[2], d[3]};
            }
            __syncthreads();
            pend = pend && !own;
        }
    }
    // After a barrier: the occupied slots compacted into list; returns their count.
    __device__ __forceinline__ int compact(int lane) {
        static_assert(kGradSlots == 2 * GRAY_WAVE, "two slots per lane");
        const bool o0 = key[lane] != GRAY_EMPTY, o1 = key[lane + GRAY_WAVE] != GRAY_EMPTY;
        const uint64_t b0 = __ballot(o0), b1 = __ballot(o1);
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const int n0 = __popcll(b0);
        if (o0) list[__popcll(b0 & below)] = (uint16_t)lane;
        if (o1) list[n0 + __popcll(b1 & below)] = (uint16_t)(lane + GRAY_WAVE);
        __syncthreads();
        return n0 + __popcll(b1);
    }
    // After compact(): every occupied slot's sums straight to the device.
    __device__ __forceinline__ void flush_global(const TraceArgs &A, const DevView &V, int lane, float pixels,
                                                 int total) {
        for (int i = lane; i < total; i += GRAY_WAVE) {
            const int slot = list[i];
            float s[kRawTerms];
#pragma unroll
            for (int j = 0; j < kRawTerms; ++j) s[j] = val[slot * kRawRow + j];
            raw_to_global(A, V, key[slot], s, pixels);
        }
    }
    // After compact(): every occupied slot's Gaussian, one per lane.  Its
    // sums become record tile * kGradSlots + i, listed under the Gaussian
    // (one integer atomic instead of a float atomic per sum); a Gaussian with
    // kRecList records already adds to its raw row.
    __device__ __forceinline__ void flush(const TraceArgs &A, const DevView &V, int lane, float pixels, int total,
                                          int tile, int view) {
        for (int i = lane; i < total; i += GRAY_WAVE) {
            const int slot = list[i];
            const uint32_t gid = key[slot];
            float s[kRawTerms];
#pragma unroll
            for (int j = 0; j < kRawTerms; ++j) s[j] = val[slot * kRawRow + j];
            const uint32_t c = atomicAdd(&A.bwd_cnt[gid], 1u);
            if (c < (uint32_t)kRecList) {
                const uint32_t ri = (uint32_t)tile * kGradSlots + (uint32_t)i;
                f4v *o = reinterpret_cast<f4v *>(A.bwd_rec + 16 * (size_t)ri);
                o[0] = f4v{s[0], s[1], s[2], s[3]};
                o[1] = f4v{s[4] / pixels, s[5], s[6], s[7]};
                o[2] = f4v{s[8], s[9], s[10], s[11]};
                o[3] = f4v{s[12], s[13], __int_as_float(view), 0.0f};
                A.bwd_lst[(size_t)gid * kRecList + c] = ri;
            } else {
                raw_to_global(A, V, gid, s, pixels);
            }
        }
    }
};

// Raw terms (kRawTerms order) of one composited hit of the stored-list
// backward: the per-hit half of the reference backward (backward_pass.cu:
// 32-95): the hit test and alpha from the snapshot record g (three 16-byte
// words: W2L and mean, then the opacity and clip pair), the colour c, then the alpha gradient and
// the running gradient gT of the transmittance behind the hit.  Returns false
// if the ray misses the record.  With rebuild set, Ti is rebuilt from Tafter as
// in backward_hit.
__device__ __forceinline__ bool backward_hit_raw(const f4v &ga, const f4v &gb, const f4v &gc, const float2 &gd,
                                                 float c0, float c1, float c2, const LaneRay &r, float p,
                                                 const PixelState &ps, float Ti, float &gT, bool rebuild,
                                                 float &Tafter, float *s) {
#pragma clang fp contract(off)
    const float m[9] = {ga[0], ga[1], ga[2], ga[3], gb[0], gb[1], gb[2], gb[3], gc[0]};
    float dist, hx, hy, hz;
    if (!gray_hit_test(m, gc[1], gc[2], gc[3], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, dist, hx, hy, hz)) return false;
    const float op = gd.x, clip = gd.y;
    float gv, q;
    const float a = gray_alpha(hx, hy, hz, clip, op, p, gv, q);
    if (rebuild) {
        Ti = Tafter / (1.0f - a);
        Tafter = Ti;
    }
    const float w = Ti * a;
    s[0] = ps.g0 * w;
    s[1] = ps.g1 * w;
    s[2] = ps.g2 * w;
    const float gdot = (ps.g0 * c0 + ps.g1 * c1) + ps.g2 * c2;
    const float ioma = 1.0f / (1.0f - a);
    float gal = (gdot - gT) * Ti;
    gal += ((-ps.F) * ioma) * ps.bgdot;
    gal += -((ps.Tf - ps.F) * ioma) * ps.estdot;
    gT += (gdot - gT) * a;
    s[3] = gal * gv;
    s[4] = w;
    float gsq;
    if (p == 1.0f)
        gsq = gv;
    else if (p == 2.0f)
        gsq = gv * q;
    else if (p == 3.0f)
        gsq = (gv * q) * q;
    else
        gsq = gv * powf(q, p - 1.0f);
    const float c = gsq * gal;
    const float lx = hx * clip, ly = hy * clip, lz = hz * clip;
    const float cx = c * lx, cy = c * ly, cz = c * lz;
    s[5] = cx;
    s[6] = cy;
    s[7] = cz;
    s[8] = cx * lx;
    s[9] = cx * ly;
    s[10] = cx * lz;
    s[11] = cy * ly;
    s[12] = cy * lz;
    s[13] = cz * lz;
    return true;
}

// Rotation, scale and mean terms (t[3..12]) of one Gaussian from the gradient
// gl of its L2W rows and the world-space hit gradient gxw, with the live scale
// and rotation (reference backward_pass.cu:96-126, backward_normalize_act of
// helpers.cu:9-34).  Linear in gl and gxw, so it serves one hit or the sums of
// many.
__device__ __forceinline__ void geometry_terms(const TraceArgs &A, uint32_t gid, float clip, float eps_sg,
                                               const float gl[3][3], float gxw0, float gxw1, float gxw2, float *t) {
#pragma clang fp contract(off)
    // scale
    const float s0 = expf(A.scale[3 * gid + 0]), s1 = expf(A.scale[3 * gid + 1]), s2 = expf(A.scale[3 * gid + 2]);
    const float *L = A.l2w + 12 * (size_t)gid;
    // the rotation columns L[:, j] / d_j: one reciprocal per column
    const float i0 = 1.0f / (s0 * clip + eps_sg), i1 = 1.0f / (s1 * clip + eps_sg), i2 = 1.0f / (s2 * clip + eps_sg);
    const float gs0 = ((gl[0][0] * (L[0] * i0) + gl[1][0] * (L[3] * i0)) + gl[2][0] * (L[6] * i0)) * s0;
    const float gs1 = ((gl[0][1] * (L[1] * i1) + gl[1][1] * (L[4] * i1)) + gl[2][1] * (L[7] * i1)) * s1;
    const float gs2 = ((gl[0][2] * (L[2] * i2) + gl[1][2] * (L[5] * i2)) + gl[2][2] * (L[8] * i2)) * s2;
    // rotation matrix rows
    const float g0x = gl[0][0] * s0, g0y = gl[0][1] * s1, g0z = gl[0][2] * s2;
    const float g1x = gl[1][0] * s0, g1y = gl[1][1] * s1, g1z = gl[1][2] * s2;
    const float g2x = gl[2][0] * s0, g2y = gl[2][1] * s1, g2z = gl[2][2] * s2;
    const float qr = A.rotation[4 * gid + 0], qx = A.rotation[4 * gid + 1], qy = A.rotation[4 * gid + 2],
                qz = A.rotation[4 * gid + 3];
    const float qlen = sqrtf(((qr * qr + qx * qx) + qy * qy) + qz * qz);
    const float iq = 1.0f / qlen;
    const float nr = qr * iq, nx = qx * iq, ny = qy * iq, nz = qz * iq;
    const float gr = (2.f * nx * (g2y - g1z) + 2.f * ny * (g0z - g2x)) + 2.f * nz * (g1x - g0y);
    const float gx = ((-4.f * nx * (g1y + g2z) + 2.f * ny * (g0y + g1x)) + 2.f * nz * (g0z + g2x)) + 2.f * nr * (g2y - g1z);
    const float gy = ((2.f * nx * (g0y + g1x) - 4.f * ny * (g0x + g2z)) + 2.f * nz * (g1z + g2y)) + 2.f * nr * (g0z - g2x);
    const float gz = ((2.f * nx * (g0z + g2x) + 2.f * ny * (g1z + g2y)) - 4.f * nz * (g0x + g1y)) + 2.f * nr * (g1x - g0y);
    // reference backward_normalize_act: (g - q (q.g) / |q|^2) / |q|
    const float qdot = ((gr * qr + gx * qx) + gy * qy) + gz * qz;
    const float iq3 = (iq * iq) * iq;
    const float q0 = (qdot * -qr) * iq3 + gr * iq;
    const float q1 = (qdot * -qx) * iq3 + gx * iq;
    const float q2 = (qdot * -qy) * iq3 + gy * iq;
    const float q3 = (qdot * -qz) * iq3 + gz * iq;

    t[3] = q0;
    t[4] = q1;
    t[5] = q2;
    t[6] = q3;
    t[7] = gs0;
    t[8] = gs1;
    t[9] = gs2;
    t[10] = -gxw0;
    t[11] = -gxw1;
    t[12] = -gxw2;
}

// Gradient of one composited hit; the terms go to sink(gid, terms).
// Ti is the transmittance before the hit; with rebuild set it is instead
// Tafter / (1 - alpha) from the recomputed fp32 alpha and Tafter becomes Ti:
// the reference's reconstruction from the final transmittance
// (backward_pass.cu), used when the forward composited fp16 alphas.
template <typename Sink>
__device__ __forceinline__ void backward_hit(const TraceArgs &A, const DevView &V, const LaneRay &r, float p,
                                             float eps_sg, const PixelState &ps, uint32_t gid, float Ti, float &gT,
                                             Sink &&sink, bool rebuild, float &Tafter) {
#pragma clang fp contract(off)
    const GeoRecord rec = A.geo[gid];
    float dist, hx, hy, hz;
    if (!gray_hit_test(rec.m, rec.mu[0], rec.mu[1], rec.mu[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, dist, hx, hy, hz))
        return;
    float gv, q;
    const float a = gray_alpha(hx, hy, hz, rec.clip, rec.op, p, gv, q);
    if (rebuild) {
        Ti = Tafter / (1.0f - a);
        Tafter = Ti;
    }
    const float c0 = V.rgb[3 * gid + 0], c1 = V.rgb[3 * gid + 1], c2 = V.rgb[3 * gid + 2];
    const float w = Ti * a;
    float t[kGradTerms];
    // colour gradient
    t[0] = ps.g0 * w;
    t[1] = ps.g1 * w;
    t[2] = ps.g2 * w;
    // alpha gradient
    const float gdot = (ps.g0 * c0 + ps.g1 * c1) + ps.g2 * c2;
    const float oma = 1.0f - a;
    float ga = (gdot - gT) * Ti;
    ga += ((-ps.F) / oma) * ps.bgdot;
    ga += -((ps.Tf - ps.F) / oma) * ps.estdot;
    gT += (gdot - gT) * a;
    // opacity (live parameters, reference helpers.cu:9-34)
    const float op = gray_sigmoid(A.opacity[gid]);
    float gop = (GRAY_MAX_ALPHA * ga) * gv;
    gop = (gop * op) * (1.0f - op);
    const float ggv = (GRAY_MAX_ALPHA * ga) * op;
    // kernel -> local hit
    const float clip = rec.clip;
    const float lx = hx * clip, ly = hy * clip, lz = hz * clip;
    float gsq;
    if (p == 1.0f)
        gsq = gv;
    else if (p == 2.0f)
        gsq = gv * q;
    else if (p == 3.0f)
        gsq = (gv * q) * q;
    else
        gsq = gv * powf(q, p - 1.0f);
    const float gxl0 = (-lx * gsq) * ggv, gxl1 = (-ly * gsq) * ggv, gxl2 = (-lz * gsq) * ggv;
    // local -> world hit point (W2L^T), times the clip factor
    const float *m = rec.m;
    const float gxw0 = ((m[0] * gxl0 + m[3] * gxl1) + m[6] * gxl2) * clip;
    const float gxw1 = ((m[1] * gxl0 + m[4] * gxl1) + m[7] * gxl2) * clip;
    const float gxw2 = ((m[2] * gxl0 + m[5] * gxl1) + m[8] * gxl2) * clip;
    // d/d(L2W rows)
    const float gl[3][3] = {{-gxw0 * lx, -gxw0 * ly, -gxw0 * lz},
                            {-gxw1 * lx, -gxw1 * ly, -gxw1 * lz},
                            {-gxw2 * lx, -gxw2 * ly, -gxw2 * lz}};
    geometry_terms(A, gid, clip, eps_sg, gl, gxw0, gxw1, gxw2, t);
    t[13] = gop;
    t[14] = w / ps.pixels;
    sink(gid, t);
}

__device__ __forceinline__ bool load_pixel_state(const TraceArgs &A, const DevView &V, int pix, LaneRay &r,
                                                 PixelState &ps) {
    r.ox = V.out_ro[3 * pix + 0];
    r.oy = V.out_ro[3 * pix + 1];
    r.oz = V.out_ro[3 * pix + 2];
    r.dx = V.out_rd[3 * pix + 0];
    r.dy = V.out_rd[3 * pix + 1];
    r.dz = V.out_rd[3 * pix + 2];
    r.ix = 1.0f / r.dx;
    r.iy = 1.0f / r.dy;
    r.iz = 1.0f / r.dz;
    ps.g0 = V.grad_out[3 * pix + 0];
    ps.g1 = V.grad_out[3 * pix + 1];
    ps.g2 = V.grad_out[3 * pix + 2];
    ps.Tf = V.out_T[pix];
    ps.F = V.out_F[pix];
    ps.e0 = V.out_est[3 * pix + 0];
    ps.e1 = V.out_est[3 * pix + 1];
    ps.e2 = V.out_est[3 * pix + 2];
    const float *bg = A.knobs.bg;
    ps.bgdot = (bg[0] * ps.g0 + bg[1] * ps.g1) + bg[2] * ps.g2;
    ps.estdot = (ps.e0 * ps.g0 + ps.e1 * ps.g1) + ps.e2 * ps.g2;
    ps.pixels = (float)(V.width * V.height);
    return ((r.dx * r.dx + r.dy * r.dy) + r.dz * r.dz) != 0.0f;
}

// Rays whose composited list fits the store: replay it back to front.
// The rays of a tile composite largely the same Gaussians (an 8x8 tile of the
// 1M-Gaussian training frame: about 367 hits on 42 distinct Gaussians).  The
// wave first claims a table slot for each distinct Gaussian of its lists and
// stages its snapshot record and colour in LDS once; the back-to-front loop
// then reads records from LDS, and global memory only for its list entries
// (coalesced, prefetched a step ahead), so the gathers per tile fall from one
// per hit to one per Gaussian.  A Gaussian that found no slot is read from
// global memory where it occurs and its hits are added after the loop.
constexpr int kRecRow = 20;  // floats per staged record: W2L 9, mean 3, op, clip, colour 3, pad 3
constexpr int kSavedFails = 12;  // failed hits per lane redone from saved state instead of a replay
struct StoredLds {
    uint32_t key[kGradSlots];
    __attribute__((aligned(16))) float val[kGradSlots * kRawRow];
    __attribute__((aligned(16))) float rec[kGradSlots * kRecRow];
    uint16_t list[kGradSlots];
    uint32_t owner[kGradSlots];
};

__device__ __forceinline__ void stored_tile(const TraceArgs &A, int tile, StoredLds &L) {
    uint32_t *s_key = L.key;
    float *s_val = L.val, *s_rec = L.rec;
    uint16_t *s_list = L.list;
    uint32_t *s_owner = L.owner;
    int vi, tx, ty;
    tile_coords(A, tile, vi, tx, ty);
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    RawTable tab{s_key, s_val, s_list, s_owner};
    tab.clear(lane);
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    const int pix = y * V.stride + x;
    const int n = (x < V.width && y < V.height) ? (int)V.n_comp[pix] : 0;
    LaneRay r;
    PixelState ps;
    // Every lane runs the wave's longest list (iteration it takes entry
    // m - 1 - it, back to front) so that the DPP exchanges below see all 64
    // lanes; lanes past their own list carry no hit.
    const bool live = n > 0 && n <= A.capacity && load_pixel_state(A, V, pix, r, ps);
    const int m = live ? n : 0;
    int mmax = m;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mmax = max(mmax, __shfl_xor(mmax, off));
    if (mmax == 0) return;
    const float p = *A.knobs.exp_power;
    const float pixels = (float)(V.width * V.height);
    const size_t R = (size_t)V.ray_count;
    const uint32_t *ids = V.hit_ids + pix;
    const float *hT = V.hit_T + pix;
    const float *rgb = V.rgb;
    const bool rebuild = A.fp16 != 0;
    __syncthreads();
    // 1. a slot for every Gaussian of the tile's lists (entries read front
    // to back, a step ahead: the front hits, on surfaces many of the tile's
    // rays share, take the slots first when the table fills; a Gaussian that
    // finds none is looked up in vain below and read from global memory)
    {
        size_t off = 0;
        uint32_t g = m >= 1 ? ids[0] : GRAY_EMPTY;
        for (int k = 0; k < m; ++k) {
            off += R;
            const uint32_t cur = g;
            if (k + 1 < m) g = ids[off];
            tab.claim(cur);
        }
    }
    __syncthreads();
    // 2. the record and colour of every claimed Gaussian, once
    const int total = tab.compact(lane);
    for (int i = lane; i < total; i += GRAY_WAVE) {
        const int slot = s_list[i];
        const uint32_t g = s_key[slot];
        const f4v *g4 = reinterpret_cast<const f4v *>(A.geo + g);
        const float2 oc = *reinterpret_cast<const float2 *>(g4 + 3);
        f4v *o = reinterpret_cast<f4v *>(s_rec + slot * kRecRow);
        o[0] = g4[0];
        o[1] = g4[1];
        o[2] = g4[2];
        o[3] = f4v{oc.x, oc.y, rgb[3 * g + 0], rgb[3 * g + 1]};
        o[4] = f4v{rgb[3 * g + 2], 0.0f, 0.0f, 0.0f};
    }
    __syncthreads();
    // The kernel ends with its longest tiles, whose steps are sequential per
    // ray; waves of long lists take precedence at issue, so they run near
    // their own pace and the short ones fill the gaps.
    if (mmax >= 16)
        __builtin_amdgcn_s_setprio(3);
    else if (mmax >= 8)
        __builtin_amdgcn_s_setprio(2);
    else if (mmax >= 4)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
    // 3. back to front
    float gT = 0.0f;
    float Tafter = ps.Tf;  // fp16_hits: transmittances rebuilt from the final one
    // steps whose hit this lane adds itself after the loop (its Gaussian has
    // no slot); a failure past step 63 drops the whole tile instead.  The
    // first kSavedFails of them keep the step and the transmittance gradient
    // (and final transmittance) they were reached with, so each is redone on
    // its own; more send the lane through a replay of its list.
    uint64_t failed = 0;
    bool far = false;
    int nfail = 0;
    int f_it[kSavedFails];
    float f_gT[kSavedFails], f_Ta[kSavedFails];
    // list entries of hits k (id0, T0) and k-1 (id1, T1) are in registers or
    // in flight; off2 is the offset of entry k-2, stepped down each step
    uint32_t id0 = GRAY_EMPTY, id1 = GRAY_EMPTY;
    float T0 = 0.0f, T1 = 0.0f;
    if (m >= 1) {
        id0 = ids[(size_t)(m - 1) * R];
        T0 = hT[(size_t)(m - 1) * R];
    }
    if (m >= 2) {
        id1 = ids[(size_t)(m - 2) * R];
        T1 = hT[(size_t)(m - 2) * R];
    }
    size_t off2 = (size_t)(m - 3) * R;
    for (int it = 0; it < mmax; ++it) {
        const int k = m - 1 - it;
        uint32_t id2 = GRAY_EMPTY;
        float T2 = 0.0f;
        float t[kRawTerms];  // read only where gid is set
        uint32_t gid = GRAY_EMPTY;
        int slot = -1;
        if (k >= 0) {
            slot = tab.lookup(id0);
            f4v ra, rb, rc, rx;
            float c2;
            if (slot >= 0) {
                // (all five words as 16-byte LDS loads: a scalar load here
                // would be merged with the global one below into a flat load)
                const f4v *q = reinterpret_cast<const f4v *>(s_rec + slot * kRecRow);
                ra = q[0], rb = q[1], rc = q[2], rx = q[3];
                const f4v q4 = q[4];
                c2 = q4[0];
            } else {
                const f4v *g4 = reinterpret_cast<const f4v *>(A.geo + id0);
                const float2 oc = *reinterpret_cast<const float2 *>(g4 + 3);
                ra = g4[0], rb = g4[1], rc = g4[2];
                rx = f4v{oc.x, oc.y, rgb[3 * id0 + 0], rgb[3 * id0 + 1]};
                c2 = rgb[3 * id0 + 2];
            }
            const float gT_in = gT, Ta_in = Tafter;
            if (backward_hit_raw(ra, rb, rc, float2{rx[0], rx[1]}, rx[2], rx[3], c2, r, p, ps, T0, gT, rebuild,
                                 Tafter, t)) {
                if (slot >= 0) {
                    gid = id0;
                } else if (it < 64) {
                    failed |= 1ull << it;
#pragma unroll
                    for (int j = 0; j < kSavedFails; ++j) {
                        if (nfail == j) {
                            f_it[j] = it;
                            f_gT[j] = gT_in;
                            f_Ta[j] = Ta_in;
                        }
                    }
                    ++nfail;
                } else {
                    far = true;
                }
            }
        }
        // The list entry two steps on, in flight during this hit's table
        // update and the whole next step.  Issued after the record is in
        // registers: a load issued earlier would be waited for with the
        // record loads of a Gaussian without a slot (one counter, in order).
        if (k >= 2) {
            id2 = ids[off2];
            T2 = hT[off2];
        }
        off2 -= R;
        // Neighbouring pixels often composite the same Gaussian at the same
        // step: combining their terms first saves table updates.  Pairs are
        // symmetric (both lanes hold the same Gaussian, so the same slot), so
        // the adds run with only the paired lanes active and read their
        // partner.  Horizontal neighbours (lane ^ 1): the even one keeps both.
        {
            const uint32_t g1 = dpp_u32<0xB1, 0xF>(gid);
            if (gid != GRAY_EMPTY && g1 == gid) {
#pragma unroll
                for (int j = 0; j < kRawTerms; ++j) t[j] += dpp_f32<0xB1, 0xF>(t[j]);
                if (lane & 1) gid = GRAY_EMPTY;
            }
        }
        // then pixels two apart in the row (lane ^ 2): the left one keeps both
        {
            const uint32_t g3 = dpp_u32<0x4E, 0xF>(gid);
            if (gid != GRAY_EMPTY && g3 == gid) {
#pragma unroll
                for (int j = 0; j < kRawTerms; ++j) t[j] += dpp_f32<0x4E, 0xF>(t[j]);
                if (lane & 2) gid = GRAY_EMPTY;
            }
        }
        // then vertical neighbours (lane ^ 8, same 16-lane row): the upper one keeps both
        {
            const uint32_t g2 = dpp_u32<0x128, 0xF>(gid);
            if (gid != GRAY_EMPTY && g2 == gid) {
#pragma unroll
                for (int j = 0; j < kRawTerms; ++j) t[j] += dpp_f32<0x128, 0xF>(t[j]);
                if (lane & 8) gid = GRAY_EMPTY;
            }
        }
        tab.add_at(gid != GRAY_EMPTY ? slot : -1, t, lane);
        id0 = id1;
        T0 = T1;
        id1 = id2;
        T1 = T2;
    }
    if (__ballot(far)) {
        // a failure past the step mask: the recompute kernel redoes the tile
        if (lane == 0) A.bwd_overflow[tile] = 1;
        return;
    }
    __syncthreads();
    tab.flush(A, V, lane, pixels, total, tile, vi);
    // Hits whose Gaussian found no slot.  The saved ones are redone from
    // their state and summed per Gaussian in the cleared table (a Gaussian
    // that failed is usually hit by many of the tile's rays: one set of
    // atomics per Gaussian instead of one per hit); a lane with more failures
    // replays its list, each failed hit straight to the device.
    const bool saved = failed && nfail <= kSavedFails;
    int jmax = saved ? nfail : 0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) jmax = max(jmax, __shfl_xor(jmax, off));
    if (jmax > 0) {
        __syncthreads();
        tab.clear(lane);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSavedFails; ++j) {
            if (j >= jmax) break;
            float t[kRawTerms];
            int slot = -1;
            if (saved && j < nfail) {
                const int k = m - 1 - f_it[j];
                const uint32_t g = ids[(size_t)k * R];
                const f4v *g4 = reinterpret_cast<const f4v *>(A.geo + g);
                float gTj = f_gT[j], Taj = f_Ta[j];
                if (backward_hit_raw(g4[0], g4[1], g4[2], *reinterpret_cast<const float2 *>(g4 + 3), rgb[3 * g + 0],
                                     rgb[3 * g + 1], rgb[3 * g + 2], r, p, ps, hT[(size_t)k * R], gTj, rebuild, Taj,
                                     t)) {
                    slot = tab.claim(g);
                    if (slot < 0) raw_to_global(A, V, g, t, pixels);
                }
            }
            tab.add_at(slot, t, lane);
        }
        __syncthreads();
        const int total2 = tab.compact(lane);
        tab.flush_global(A, V, lane, pixels, total2);
    }
    if (failed && !saved) {
        float gT2 = 0.0f, Tafter2 = ps.Tf;
        for (int it = 0; it < m && it < 64 && (failed >> it) != 0; ++it) {
            const int k = m - 1 - it;
            const uint32_t g = ids[(size_t)k * R];
            const f4v *g4 = reinterpret_cast<const f4v *>(A.geo + g);
            float t[kRawTerms];
            if (backward_hit_raw(g4[0], g4[1], g4[2], *reinterpret_cast<const float2 *>(g4 + 3), rgb[3 * g + 0],
                                 rgb[3 * g + 1], rgb[3 * g + 2], r, p, ps, hT[(size_t)k * R], gT2, rebuild, Tafter2,
                                 t) &&
                ((failed >> it) & 1))
                raw_to_global(A, V, g, t, pixels);
        }
    }
}

// The stored-list backward runs persistent: as many workgroups as fit on the
// device take tiles from a list ordered by cost (the tile's longest stored
// list, longest first), one atomic per tile.  A tile's steps are sequential,
// so the longest tiles set the kernel's end: started first, they run beside
// the short ones instead of after them, and no workgroup slot waits on the
// dispatcher.
__global__ __launch_bounds__(GRAY_WAVE) void k_trace_backward_stored(TraceArgs A, const uint32_t *__restrict__ order,
                                                                    int32_t *__restrict__ next) {
    __shared__ StoredLds L;
    __shared__ int s_t;
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(next, 1);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= A.total_tiles) break;
        stored_tile(A, (int)order[t], L);
        __syncthreads();
    }
}

// Cost of each tile for the stored-list backward: the longest list of its rays
// that the store holds (0: nothing to do), clamped to 255.
__global__ __launch_bounds__(GRAY_WAVE) void k_tile_cost(TraceArgs A, uint8_t *__restrict__ cost) {
    int vi, tx, ty, tile;
    if (!block_tile(A, vi, tx, ty, &tile)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    int n = (x < V.width && y < V.height) ? (int)V.n_comp[y * V.stride + x] : 0;
    n = n <= A.capacity ? n : 0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) n = max(n, __shfl_xor(n, off));
    if (lane == 0) cost[tile] = (uint8_t)min(n, 255);
}

// Tiles ordered by descending cost (a counting sort in one workgroup), the
// fetch counter of k_trace_backward_stored reset.
constexpr int kOrderThreads = 1024;
__global__ __launch_bounds__(kOrderThreads) void k_tile_order(const uint8_t *__restrict__ cost, int tiles,
                                                            uint32_t *__restrict__ order, int32_t *__restrict__ next) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t base[256];
    const int t = threadIdx.x;
    if (t < 256) hist[t] = 0;
    __syncthreads();
    for (int i = t; i < tiles; i += kOrderThreads) atomicAdd(&hist[255 - cost[i]], 1u);
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int b = 0; b < 256; ++b) {
            base[b] = acc;
            acc += hist[b];
        }
        *next = 0;
    }
    __syncthreads();
    for (int i = t; i < tiles; i += kOrderThreads) order[atomicAdd(&base[255 - cost[i]], 1u)] = (uint32_t)i;
}

// Gradients of the stored-list backward from the per-Gaussian raw rows, one
// thread per Gaussian: a touched row becomes the rotation, scale, mean,
// opacity and pruning terms (raw_geometry_terms, with the live parameters),
// added to the reference gradient tensors; the row and its mark are cleared
// for the next backward.  No other thread writes these elements, so the adds
// are plain read-modify-writes.
__global__ __launch_bounds__(256) void k_raw_apply(TraceArgs A, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = min(A.bwd_cnt[i], (uint32_t)kRecList);
    const bool touched = A.raw_touched[i] != 0;
    if (c == 0 && !touched) return;
    float r[11] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    f4v *row = reinterpret_cast<f4v *>(A.raw + 16 * (size_t)i);
    if (touched) {
        const f4v r0 = row[0], r1 = row[1], r2 = row[2];
        r[0] = r0[0], r[1] = r0[1], r[2] = r0[2], r[3] = r0[3], r[4] = r1[0], r[5] = r1[1];
        r[6] = r1[2], r[7] = r1[3], r[8] = r2[0], r[9] = r2[1], r[10] = r2[2];
    }
    // the tiles' records: colour to the view's colour gradient, the rest summed
    for (uint32_t j = 0; j < c; ++j) {
        const f4v *q = reinterpret_cast<const f4v *>(A.bwd_rec + 16 * (size_t)A.bwd_lst[(size_t)i * kRecList + j]);
        const f4v q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        float *grgb = A.views[__float_as_int(q3[2])].grad_rgb + 3 * (size_t)i;
        grgb[0] += q0[0];
        grgb[1] += q0[1];
        grgb[2] += q0[2];
        r[0] += q0[3];
        r[1] += q1[0];
        r[2] += q1[1], r[3] += q1[2], r[4] += q1[3];
        r[5] += q2[0], r[6] += q2[1], r[7] += q2[2], r[8] += q2[3];
        r[9] += q3[0], r[10] += q3[1];
    }
    float t[kGradTerms];
    raw_geometry_terms(A, (uint32_t)i, *A.knobs.eps_scale_grad, r, t);
    A.g_rot[4 * i + 0] += t[3];
    A.g_rot[4 * i + 1] += t[4];
    A.g_rot[4 * i + 2] += t[5];
    A.g_rot[4 * i + 3] += t[6];
    A.g_scale[3 * i + 0] += t[7];
    A.g_scale[3 * i + 1] += t[8];
    A.g_scale[3 * i + 2] += t[9];
    A.g_mean[3 * i + 0] += t[10];
    A.g_mean[3 * i + 1] += t[11];
    A.g_mean[3 * i + 2] += t[12];
    A.g_op[i] += t[13];
    if (A.g_prune) A.g_prune[i] += t[14];
    if (A.visible) A.visible[i] = 1;
    A.bwd_cnt[i] = 0;
    if (touched) {
        const f4v z = {0.0f, 0.0f, 0.0f, 0.0f};
        row[0] = z;
        row[1] = z;
        row[2] = z;
        A.raw_touched[i] = 0;
    }
}

__global__ void k_zero_u32(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Deterministic mode, rays whose composited list fits the store: each hit's
// terms go to its own record (no neighbour combining, no table, no atomics).
__global__ __launch_bounds__(GRAY_WAVE) void k_trace_backward_stored_det(TraceArgs A) {
    int vi, tx, ty;
    if (!block_tile(A, vi, tx, ty)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    if (x >= V.width || y >= V.height) return;
    const int pix = y * V.stride + x;
    const int n = (int)V.n_comp[pix];
    LaneRay r;
    PixelState ps;
    if (n <= 0 || n > A.capacity || !load_pixel_state(A, V, pix, r, ps)) return;
    const float p = *A.knobs.exp_power, eps_sg = *A.knobs.eps_scale_grad;
    const int R = V.ray_count;
    const size_t base = A.det_off[pix];
    float gT = 0.0f;
    float Tafter = ps.Tf;
    for (int k = n - 1; k >= 0; --k) {
        const uint32_t g = V.hit_ids[(size_t)k * R + pix];
        const float Ti = V.hit_T[(size_t)k * R + pix];
        backward_hit(A, V, r, p, eps_sg, ps, g, Ti, gT,
                     [&](uint32_t gg, const float *t) { det_put(A, base + k, gg, t); }, A.fp16 != 0, Tafter);
    }
}

// Rays whose list exceeded the store (or capacity 0): regenerate each round by
// re-traversal, sweep it forward (T, partial sums) and back (gradients).  The
// gradient flowing into a round from the rounds behind it is
// (g.C_comp - sum_{composited so far} g.c w) / T.
// With kDet the terms of composited hit k (in distance order) go to record
// det_off[pix] + k instead of the gradient tensors.
template <bool kDet>
__global__ __launch_bounds__(GRAY_WAVE) void k_trace_backward_recompute(TraceArgs A, const GeoRecord *__restrict__ geo,
                                                                        const BvhNode *__restrict__ nodes) {
    __shared__ float s_kd[GRAY_KB * GRAY_WAVE];
    __shared__ uint32_t s_ki[GRAY_KB * GRAY_WAVE];
    __shared__ float s_ka[GRAY_KB * GRAY_WAVE];
    __shared__ uint32_t s_stack[GRAY_STACK * GRAY_WAVE];
    int vi, tx, ty, tile;
    if (!block_tile(A, vi, tx, ty, &tile)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    const bool inside = x < V.width && y < V.height;
    const int pix = y * V.stride + x;
    int n = 0;
    if (inside) n = (int)V.n_comp[pix];
    if (!kDet && A.bwd_overflow && A.bwd_overflow[tile]) {
        // the stored-list backward dropped this tile (its table was full):
        // its stored lists, hit by hit, straight to the gradient tensors
        if (n > 0 && n <= A.capacity) {
            LaneRay r0;
            PixelState ps0;
            if (load_pixel_state(A, V, pix, r0, ps0)) {
                const float p0 = *A.knobs.exp_power, eps0 = *A.knobs.eps_scale_grad;
                const size_t R = (size_t)V.ray_count;
                float gT = 0.0f, Tafter = ps0.Tf;
                for (int k = n - 1; k >= 0; --k)
                    backward_hit(A, V, r0, p0, eps0, ps0, V.hit_ids[(size_t)k * R + pix], V.hit_T[(size_t)k * R + pix],
                                 gT, [&](uint32_t g, const float *t) { grad_to_global(A, V, g, t); }, A.fp16 != 0,
                                 Tafter);
            }
        }
        __syncthreads();
        if (lane == 0) A.bwd_overflow[tile] = 0;
    }
    bool need = inside && n > A.capacity;
    if (!__ballot(need)) return;
    LaneRay r{};
    PixelState ps{};
    if (need) need = load_pixel_state(A, V, pix, r, ps);
    const float athr = *A.knobs.alpha_threshold, p = *A.knobs.exp_power, tthr = *A.knobs.t_threshold;
    const float eps_sg = *A.knobs.eps_scale_grad;
    const float znear = *V.znear, zfar = *V.zfar;
    KBuf kb{s_kd, s_ki, s_ka, lane};
    const bool h16 = A.fp16 != 0;
    // The gradient sweep of a round needs the gradient flowing in from the
    // hits behind it: (Atot - P) / T with Atot the colour the forward
    // composited (against g) and P, T the part composited so far.  With fp32
    // alphas the forward's sums serve.  With fp16_hits the reference rebuilds
    // each hit's transmittance from the final one by fp32 alphas
    // (backward_pass.cu), so a first replay forms the fp32-alpha totals: Q =
    // prod(1 - a32) and Atot = sum g.c a32 T32, and the second replay scales
    // T32 by Tfinal / Q.
    float Atot = 0.0f, scale = 1.0f;
    if (need)
        Atot = (ps.g0 * V.comp_rgb[3 * pix + 0] + ps.g1 * V.comp_rgb[3 * pix + 1]) + ps.g2 * V.comp_rgb[3 * pix + 2];
    for (int pass = h16 ? 0 : 1; pass < 2; ++pass) {
    float P = 0.0f, T = 1.0f, T32 = 1.0f, P32 = 0.0f, tmin = znear;
    int done = 0;
    bool alive = need && tmin < zfar;
    // same round cursor as the forward
    float cur_d = znear;
    uint32_t cur_id = 0;
    bool strict = true;
    int bpos = 0, batch = 0;
    for (int round = 0; round < 2 * GRAY_MAX_ROUNDS * (GRAY_BATCH / GRAY_KB); ++round) {
        const bool part = alive;
        if (!__ballot(part)) break;
        int count = 0;
        float lastd = 0.0f;
        uint32_t lasti = 0;
        auto leaf = [&](uint32_t gid, const GeoV &g, float, float, float) {
            if (g.op <= athr) return;
            float dist, hx, hy, hz;
            if (!gray_hit_test(g.m, g.mu[0], g.mu[1], g.mu[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, dist, hx, hy, hz))
                return;
            dist = fp16_round(dist, h16);
            const bool beyond = strict ? dist > cur_d : (dist > cur_d || (dist == cur_d && gid > cur_id));
            if (!beyond) return;
            float gv, q;
            const float a = gray_alpha(hx, hy, hz, g.clip, g.op, p, gv, q);
            uint32_t ev_id;
            float ev_a;
            kb_insert(kb, count, lastd, lasti, dist, gid, a, ev_id, ev_a);
        };
        traverse(A, nodes, geo, V.rgb, false, r, part, cur_d, s_stack, lane, leaf);
        if (part) {
            bool terminated = false;
            int used = 0;
            float last_d = 0.0f;
            uint32_t last_id = 0;
            if (count == GRAY_KB) {
                last_d = kb.dist(GRAY_KB - 1);
                last_id = kb.gid(GRAY_KB - 1);
            }
            for (int k = 0; k < count; ++k) {
                const float d = kb.dist(k);
                const uint32_t gid = kb.gid(k);
                const float a32 = kb.alpha(k);
                const float a = fp16_round(a32, h16);  // as the forward composited it
                if (terminated) {
                    kb.id[k * GRAY_WAVE + lane] = GRAY_EMPTY;
                    continue;
                }
                tmin = fmaxf(d, tmin);
                if (d < zfar) {
                    const float c0 = V.rgb[3 * gid + 0], c1 = V.rgb[3 * gid + 1], c2 = V.rgb[3 * gid + 2];
                    const float gdot = (ps.g0 * c0 + ps.g1 * c1) + ps.g2 * c2;
                    P += gdot * (T * a);
                    P32 += gdot * (T32 * a32);
                    kb.a[k * GRAY_WAVE + lane] = h16 ? scale * T32 : T;  // transmittance before this hit
                    T = T * (1.0f - a);
                    T32 = T32 * (1.0f - a32);
                    ++done;
                    used = k + 1;
                    if (T < tthr) {
                        tmin = zfar;
                        terminated = true;
                    }
                } else {
                    kb.id[k * GRAY_WAVE + lane] = GRAY_EMPTY;
                }
            }
            const bool last = done >= n;
            if (pass == 1) {
                float gT = last ? 0.0f : (h16 ? (Atot - P32) / T32 : (Atot - P) / T);
                // hits 0 .. used-1 of this round were composited, after done - used earlier ones
                const int first = done - used;
                for (int k = used - 1; k >= 0; --k) {
                    const uint32_t gid = kb.gid(k);
                    if (gid == GRAY_EMPTY) continue;
                    float unused = 0.0f;
                    if (kDet) {
                        // the forward composited n hits: records past them have no slot
                        if (first + k >= n) continue;
                        const size_t rec = A.det_off[pix] + (size_t)(first + k);
                        backward_hit(A, V, r, p, eps_sg, ps, gid, kb.alpha(k), gT,
                                     [&](uint32_t g, const float *t) { det_put(A, rec, g, t); }, false, unused);
                    } else {
                        backward_hit(A, V, r, p, eps_sg, ps, gid, kb.alpha(k), gT,
                                     [&](uint32_t g, const float *t) { grad_to_global(A, V, g, t); }, false,
                                     unused);
                    }
                }
            }
            alive = false;
            if (!terminated && !last && count == GRAY_KB) {
                bpos += GRAY_KB;
                if (bpos >= GRAY_BATCH) {
                    bpos = 0;
                    strict = true;
                    cur_d = tmin;
                    ++batch;
                } else {
                    strict = false;
                    cur_d = last_d;
                    cur_id = last_id;
                }
                alive = tmin < zfar && batch < GRAY_MAX_ROUNDS;
            }
        }
    }
    if (pass == 0 && need) {
        Atot = P32;
        scale = T32 > 0.0f ? ps.Tf / T32 : 1.0f;
    }
    }
}

// ---------------------------------------------------------------------------
// Deterministic gradient reduction (one view per launch)
// ---------------------------------------------------------------------------
// Records per pixel of view 0: its composited hits (0 outside the image);
// entry ray_count is 0 so that the exclusive scan ends with the total.
__global__ void k_det_counts(TraceArgs A, uint32_t *__restrict__ cnt) {
    const DevView &V = A.views[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > V.ray_count) return;
    uint32_t c = 0;
    if (i < V.ray_count) {
        const int x = i % V.stride, y = i / V.stride;
        if (x < V.width && y < V.height) c = V.n_comp[i];
    }
    cnt[i] = c;
}

// Records not written by the backward (rays it skips) keep key `none`, which
// sorts after every Gaussian id; values are the record indices.
__global__ void k_det_init(uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, int total, uint32_t none) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    keys[i] = none;
    vals[i] = (uint32_t)i;
}

// One thread per run of equal keys in the sorted records: sums the run's
// terms in record order (pixel, then hit order) and adds them to the
// gradient tensors.  Runs are disjoint, so no atomics are needed.
__global__ void k_det_reduce(TraceArgs A, const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                             int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t gid = keys[i];
    if (gid >= (uint32_t)A.num_gaussians) return;
    if (i > 0 && keys[i - 1] == gid) return;
    float t[kGradTerms];
#pragma unroll
    for (int j = 0; j < kGradTerms; ++j) t[j] = 0.0f;
    for (int j = i; j < total && keys[j] == gid; ++j) {
        const f4v *rec = reinterpret_cast<const f4v *>(A.det_rec + 16 * (size_t)vals[j]);
        const f4v a = rec[0], b = rec[1], c = rec[2], d = rec[3];
        t[0] += a.x; t[1] += a.y; t[2] += a.z; t[3] += a.w;
        t[4] += b.x; t[5] += b.y; t[6] += b.z; t[7] += b.w;
        t[8] += c.x; t[9] += c.y; t[10] += c.z; t[11] += c.w;
        t[12] += d.x; t[13] += d.y; t[14] += d.z;
    }
    const DevView &V = A.views[0];
    V.grad_rgb[3 * gid + 0] += t[0];
    V.grad_rgb[3 * gid + 1] += t[1];
    V.grad_rgb[3 * gid + 2] += t[2];
    A.g_rot[4 * gid + 0] += t[3];
    A.g_rot[4 * gid + 1] += t[4];
    A.g_rot[4 * gid + 2] += t[5];
    A.g_rot[4 * gid + 3] += t[6];
    A.g_scale[3 * gid + 0] += t[7];
    A.g_scale[3 * gid + 1] += t[8];
    A.g_scale[3 * gid + 2] += t[9];
    A.g_mean[3 * gid + 0] += t[10];
    A.g_mean[3 * gid + 1] += t[11];
    A.g_mean[3 * gid + 2] += t[12];
    A.g_op[gid] += t[13];
    if (A.g_prune) A.g_prune[gid] += t[14];
    if (A.visible) A.visible[gid] = 1;
}

// ---------------------------------------------------------------------------
// Debug: every hit of every ray of view 0 (unsorted), then a per-ray sort.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GRAY_WAVE) void k_debug_hits(TraceArgs A, const GeoRecord *__restrict__ geo,
                                                          const BvhNode *__restrict__ nodes) {
    __shared__ uint32_t s_stack[GRAY_STACK * GRAY_WAVE];
    int vi, tx, ty;
    if (!block_tile(A, vi, tx, ty)) return;
    if (vi != 0) return;
    const DevView &V = A.views[0];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    const bool inside = x < V.width && y < V.height;
    LaneRay r{};
    bool active = false;
    if (inside) active = make_ray(A.knobs, V, x, y, r);
    const float athr = *A.knobs.alpha_threshold;
    const int ray = y * V.width + x;
    int cnt = 0;
    auto leaf = [&](uint32_t gid, const GeoV &g, float, float, float) {
        if (g.op <= athr) return;
        float dist, hx, hy, hz;
        if (!gray_hit_test(g.m, g.mu[0], g.mu[1], g.mu[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, dist, hx, hy, hz))
            return;
        if (cnt < A.dbg_max) {
            A.dbg_ids[(size_t)ray * A.dbg_max + cnt] = gid;
            A.dbg_dist[(size_t)ray * A.dbg_max + cnt] = dist;
        }
        ++cnt;
    };
    traverse(A, nodes, geo, V.rgb, false, r, active, 0.0f, s_stack, lane, leaf);
    if (inside) A.dbg_count[ray] = active ? cnt : 0;
}

// Ellipsoid display mode (reference raygen_ellipsoids + __intersection__is_ellipsoid
// + __closesthit__ellipsoid, ellipsoid_viewer.cu:1-94, selected at
// shaders.cu:90-93): the nearest proxy ellipsoid of opacity >= min_op in
// [znear, zfar] (ties: lower id), flat-shaded max(0.5, -d.n) (colour + 0.15),
// depth = hit distance; background and depth 0 on a miss.  Rays with a zero
// direction leave their pixel untouched.
__global__ __launch_bounds__(GRAY_WAVE) void k_ellipsoids(TraceArgs A, const GeoRecord *__restrict__ geo,
                                                          const BvhNode *__restrict__ nodes,
                                                          const float *__restrict__ min_op) {
    __shared__ uint32_t s_stack[GRAY_STACK * GRAY_WAVE];
    int vi, tx, ty;
    if (!block_tile(A, vi, tx, ty)) return;
    const DevView &V = A.views[vi];
    const int lane = threadIdx.x;
    const int x = tx * GRAY_TILE + (lane & 7), y = ty * GRAY_TILE + (lane >> 3);
    const bool inside = x < V.width && y < V.height;
    LaneRay r{};
    bool active = false;
    if (inside) active = make_ray(A.knobs, V, x, y, r);
    const float znear = *V.znear, zfar = *V.zfar, mo = *min_op;
    float best = INFINITY, bnx = 0.0f, bny = 0.0f, bnz = 0.0f;
    uint32_t bid = GRAY_EMPTY;
    auto leaf = [&](uint32_t gid, const GeoV &g, float, float, float) {
        if (!(g.op > 0.0f) || g.op < mo) return;  // not a proxy, or below the display threshold
        float t, nx, ny, nz;
        if (!gray_ellipsoid_hit(g.m, g.mu[0], g.mu[1], g.mu[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, znear, zfar, t,
                                nx, ny, nz))
            return;
        if (t < best || (t == best && gid < bid)) {
            best = t;
            bid = gid;
            bnx = nx;
            bny = ny;
            bnz = nz;
        }
    };
    traverse(A, nodes, geo, V.rgb, false, r, active, 0.0f, s_stack, lane, leaf);
    if (!active) return;
    const int pix = y * V.stride + x;
    const float *bg = A.knobs.bg;
    float c0 = bg[0], c1 = bg[1], c2 = bg[2], depth = 0.0f;
    if (bid != GRAY_EMPTY) {
#pragma clang fp contract(off)
        const GeoV g = load_geo(geo, bid);
        // local normal -> world: the transpose of world-to-local
        const float ln = sqrtf((bnx * bnx + bny * bny) + bnz * bnz);
        const float lx = bnx / ln, ly = bny / ln, lz = bnz / ln;
        const float wx = (g.m[0] * lx + g.m[3] * ly) + g.m[6] * lz;
        const float wy = (g.m[1] * lx + g.m[4] * ly) + g.m[7] * lz;
        const float wz = (g.m[2] * lx + g.m[5] * ly) + g.m[8] * lz;
        const float wn = sqrtf((wx * wx + wy * wy) + wz * wz);
        const float dn = sqrtf((r.dx * r.dx + r.dy * r.dy) + r.dz * r.dz);
        const float align = fmaxf(0.5f, ((-r.dx / dn) * (wx / wn) + (-r.dy / dn) * (wy / wn)) + (-r.dz / dn) * (wz / wn));
        c0 = align * (V.rgb[3 * bid + 0] + 0.15f);
        c1 = align * (V.rgb[3 * bid + 1] + 0.15f);
        c2 = align * (V.rgb[3 * bid + 2] + 0.15f);
        depth = best;
    }
    V.out_c[3 * pix + 0] = c0;
    V.out_c[3 * pix + 1] = c1;
    V.out_c[3 * pix + 2] = c2;
    if (V.out_depth) V.out_depth[pix] = depth;
}

__global__ void k_debug_sort(int rays, int maxh, uint32_t *__restrict__ ids, float *__restrict__ dist,
                             const int32_t *__restrict__ counts) {
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= rays) return;
    const int n = min(counts[ray], maxh);
    uint32_t *id = ids + (size_t)ray * maxh;
    float *ds = dist + (size_t)ray * maxh;
    for (int i = 1; i < n; ++i) {
        const float d = ds[i];
        const uint32_t g = id[i];
        int j = i - 1;
        while (j >= 0 && key_less(d, g, ds[j], id[j])) {
            ds[j + 1] = ds[j];
            id[j + 1] = id[j];
            --j;
        }
        ds[j + 1] = d;
        id[j + 1] = g;
    }
}

int fill_args(gray_ctx *ctx, const gray_render_args *args, int first_view, int nviews, TraceArgs &A) {
    memset(&A, 0, sizeof(A));
    const gray_ctx *S = scene_of(ctx);
    const gray_knobs &k = args->knobs;
    if (!k.exp_power || !k.alpha_threshold || !k.t_threshold || !k.eps_forward_normalization || !k.eps_scale_grad ||
        !k.background_channels || !k.render_depth) {
        gray_set_error(GRAY_ERR_INVALID, "render: null knob pointer");
        return GRAY_ERR_INVALID;
    }
    A.knobs = DevKnobs{k.exp_power, k.alpha_threshold, k.t_threshold, k.global_scale_factor,
                       k.eps_forward_normalization, k.eps_scale_grad, k.background_channels, k.render_depth,
                       k.enable_sh, k.jitter_primary_rays, k.call_index, k.needs_ray_output};
    int tiles = 0;
    for (int i = 0; i < nviews; ++i) {
        const gray_view &v = args->views[first_view + i];
        if (v.width <= 0 || v.height <= 0 || v.stride < v.width || v.alloc_height < v.height) {
            gray_set_error(GRAY_ERR_INVALID, "render: view %d has bad dimensions", first_view + i);
            return GRAY_ERR_INVALID;
        }
        const bool explicit_rays = v.rays_direction != nullptr;
        const bool pinhole_possible = !explicit_rays || v.rays_enable != nullptr;
        if ((explicit_rays && !v.rays_origin) ||
            (pinhole_possible && (!v.origin || !v.rotation_c2w || !v.vertical_fov)) || !v.znear || !v.zfar ||
            !v.out_channels || (S->n > 0 && !v.rgb)) {
            gray_set_error(GRAY_ERR_INVALID, "render: view %d lacks a camera, rgb or output pointer", first_view + i);
            return GRAY_ERR_INVALID;
        }
        DevView &d = A.views[i];
        d.origin = v.origin;
        d.c2w = v.rotation_c2w;
        d.fov = v.vertical_fov;
        d.znear = v.znear;
        d.zfar = v.zfar;
        d.rays_o = explicit_rays ? v.rays_origin : nullptr;
        d.rays_d = explicit_rays ? v.rays_direction : nullptr;
        d.rays_en = explicit_rays ? v.rays_enable : nullptr;
        d.rgb = v.rgb;
        d.out_c = v.out_channels;
        d.out_depth = v.out_depth;
        d.out_T = v.out_transmittance;
        d.out_F = v.out_full_transmittance;
        d.out_est = v.out_remaining_est;
        d.out_ro = v.out_ray_origin;
        d.out_rd = v.out_ray_direction;
        d.stat_hits = v.stat_hits;
        d.stat_acc = v.stat_accumulated;
        d.hit_ids = v.hit_ids;
        d.hit_T = v.hit_trans;
        d.n_comp = v.n_comp;
        d.comp_rgb = v.comp_rgb;
        d.grad_out = v.grad_out;
        d.grad_rgb = v.grad_rgb;
        d.width = v.width;
        d.height = v.height;
        d.stride = v.stride;
        d.alloc_height = v.alloc_height;
        d.tiles_x = (v.width + GRAY_TILE - 1) / GRAY_TILE;
        d.tile_base = tiles;
        d.ray_count = v.stride * v.alloc_height;
        tiles += d.tiles_x * ((v.height + GRAY_TILE - 1) / GRAY_TILE);
    }
    A.num_views = nviews;
    A.total_tiles = tiles;
    A.xcd_chunk = (tiles + 7) / 8;
    A.save = args->save;
    A.fp16 = args->fp16_hits ? 1 : 0;
    A.capacity = args->hit_capacity < 0 ? 0 : args->hit_capacity;
    A.use_bvh = args->use_bvh;
    A.num_gaussians = (int)S->n;
    A.tile_stats = args->tile_stats;
    A.tile_base_launch = 0;
    A.geo = S->geo;
    A.nodes = S->nodes;
    A.l2w = S->l2w;
    A.error_flag = ctx->err;
    A.heavy_count = nullptr;
    A.heavy_rays = nullptr;
    A.retrace_any = nullptr;
    A.wave_list_min = -1;
    return GRAY_OK;
}

// Grows the collect-pass lists to lanes * cap records and the task buffers
// to kTaskPerLane entries per lane (never shrinks).
constexpr int64_t kTaskPerLane = 16;
// Tile bins reuse the second task buffer: Gaussian ids, two per task entry.
inline int64_t bin_capacity(int64_t task_cap) { return std::min<int64_t>(2 * task_cap, INT32_MAX); }
// device counters of a launch: tasks handed to each round, each round's take
// position, the wave-per-ray queue length and a spare
constexpr int kCounters = 2 * kTaskRounds + 5;  // + the tile bins' mode flag
constexpr int kWaveGrid = 8192;  // persistent waves of the wave-per-ray kernel

int reserve_lists(gray_ctx *ctx, int64_t lanes, int cap) {
    if (cap > (1 << 16)) {
        gray_set_error(GRAY_ERR_INVALID, "render: list_capacity %d exceeds 65536", cap);
        return GRAY_ERR_INVALID;
    }
    if (lanes * kTaskPerLane > INT32_MAX) {
        gray_set_error(GRAY_ERR_INVALID, "render: %lld rays in one launch exceed the task buffer index range",
                       (long long)lanes);
        return GRAY_ERR_INVALID;
    }
    auto fail = [&](const char *what, int64_t bytes) {
        (void)hipGetLastError();
        gray_set_error(GRAY_ERR_NOMEM, "render: cannot allocate %s (%lld bytes)", what, (long long)bytes);
        return GRAY_ERR_NOMEM;
    };
    if (!ctx->task_counters) {
        if (hipMalloc(&ctx->task_counters, sizeof(int32_t) * kCounters) != hipSuccess)
            return fail("task counters", sizeof(int32_t) * kCounters);
    }
    if (lanes > ctx->list_lanes_cap) {
        GRAY_HIP(hipFree(ctx->list_n));
        GRAY_HIP(hipFree(ctx->tile_open));
        ctx->tile_open = nullptr;
        GRAY_HIP(hipFree(ctx->bin_tiles));
        GRAY_HIP(hipFree(ctx->bin_items));
        ctx->bin_tiles = nullptr;
        ctx->bin_items = nullptr;
        ctx->bin_item_cap = 0;
        for (uint2 *&t : ctx->tasks) {
            GRAY_HIP(hipFree(t));
            t = nullptr;
        }
        ctx->list_n = nullptr;
        ctx->list_lanes_cap = 0;
        ctx->task_cap = 0;
        if (hipMalloc(&ctx->list_n, sizeof(int32_t) * lanes) != hipSuccess)
            return fail("list counts", sizeof(int32_t) * lanes);
        if (hipMalloc(&ctx->tile_open, (size_t)(lanes / GRAY_WAVE)) != hipSuccess)
            return fail("tile flags", lanes / GRAY_WAVE);
        for (uint2 *&t : ctx->tasks)
            if (hipMalloc(&t, sizeof(uint2) * lanes * kTaskPerLane) != hipSuccess)
                return fail("task buffers", sizeof(uint2) * lanes * kTaskPerLane);
        const int64_t tiles = lanes / GRAY_WAVE;
        if (hipMalloc(&ctx->bin_tiles, sizeof(int32_t) * (3 * tiles + 2)) != hipSuccess)
            return fail("tile bin counters", sizeof(int32_t) * (3 * tiles + 2));
        // work items: at most one partial chunk per tile plus one per full chunk of the bins (tasks[1])
        const int64_t items = tiles + bin_capacity(lanes * kTaskPerLane) / kBinChunk + 1;
        if (hipMalloc(&ctx->bin_items, sizeof(int2) * items) != hipSuccess)
            return fail("tile bin work items", sizeof(int2) * items);
        ctx->bin_item_cap = items;
        ctx->list_lanes_cap = lanes;
        ctx->task_cap = lanes * kTaskPerLane;
    }
    const int64_t entries = lanes * (int64_t)cap;
    if (entries > ctx->list_entries) {
        GRAY_HIP(hipFree(ctx->list_rec));
        ctx->list_rec = nullptr;
        ctx->list_entries = 0;
        if (hipMalloc(&ctx->list_rec, sizeof(float4) * entries) != hipSuccess)
            return fail("hit lists", (int64_t)sizeof(float4) * entries);
        ctx->list_entries = entries;
    }
    return GRAY_OK;
}

// Grows a device buffer to `count` elements (contents are not kept).
template <typename T>
int det_grow(T *&buf, int64_t &cap, int64_t count, const char *what) {
    if (count <= cap) return GRAY_OK;
    GRAY_HIP(hipFree(buf));
    buf = nullptr;
    cap = 0;
    if (hipMalloc(&buf, sizeof(T) * count) != hipSuccess) {
        (void)hipGetLastError();
        gray_set_error(GRAY_ERR_NOMEM, "render_backward(deterministic): cannot allocate %s (%lld bytes)", what,
                       (long long)(sizeof(T) * count));
        return GRAY_ERR_NOMEM;
    }
    cap = count;
    return GRAY_OK;
}

// Deterministic backward of one view (A holds that view alone): per-pixel
// record offsets by a scan of the composited counts, the records of every
// composited hit, a stable radix sort of (Gaussian id, record index) and one
// ordered sum per Gaussian.  Reads the record total back to the host.
int det_backward(gray_ctx *ctx, TraceArgs &A, hipStream_t s) {
    const DevView &V = A.views[0];
    const int64_t m = (int64_t)V.ray_count + 1;
    int rc;
    if ((rc = det_grow(ctx->det_off, ctx->det_px_cap, m, "record offsets"))) return rc;
    if ((rc = det_grow(ctx->det_sums, ctx->det_sums_cap, gray_scan_scratch(m), "scan sums"))) return rc;
    hipLaunchKernelGGL(k_det_counts, dim3(gray_div_up(m, 256)), dim3(256), 0, s, A, ctx->det_off);
    GRAY_HIP(hipGetLastError());
    if ((rc = gray_exclusive_scan_u32(ctx->det_off, (int)m, ctx->det_sums, s))) return rc;
    uint32_t total = 0;
    GRAY_HIP(hipMemcpyAsync(&total, ctx->det_off + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GRAY_HIP(hipStreamSynchronize(s));
    if (total > (uint32_t)INT32_MAX) {
        gray_set_error(GRAY_ERR_INVALID, "render_backward(deterministic): %u composited hits exceed 2^31 - 1", total);
        return GRAY_ERR_INVALID;
    }
    if (total == 0) return GRAY_OK;
    const int64_t recs = total;
    if (recs > ctx->det_cap) {
        int64_t c[5] = {0, 0, 0, 0, 0};
        ctx->det_cap = 0;
        if ((rc = det_grow(ctx->det_rec, c[0], 16 * recs, "gradient records"))) return rc;
        for (int i = 0; i < 2; ++i) {
            if ((rc = det_grow(ctx->det_keys[i], c[1 + i], recs, "record keys"))) return rc;
            if ((rc = det_grow(ctx->det_vals[i], c[3 + i], recs, "record indices"))) return rc;
        }
        ctx->det_cap = recs;
    }
    int64_t hist = 0, sums = 0;
    gray_radix_scratch(recs, &hist, &sums);
    if ((rc = det_grow(ctx->det_hist, ctx->det_hist_cap, hist, "sort histogram"))) return rc;
    if ((rc = det_grow(ctx->det_sums, ctx->det_sums_cap, sums, "scan sums"))) return rc;
    const uint32_t none = (uint32_t)A.num_gaussians;  // sorts after every id
    hipLaunchKernelGGL(k_det_init, dim3(gray_div_up(recs, 256)), dim3(256), 0, s, ctx->det_keys[0], ctx->det_vals[0],
                       (int)recs, none);
    GRAY_HIP(hipGetLastError());
    A.det_rec = ctx->det_rec;
    A.det_keys = ctx->det_keys[0];
    A.det_off = ctx->det_off;
    const dim3 grid(8 * A.xcd_chunk), block(GRAY_WAVE);
    if (A.capacity > 0) {
        hipLaunchKernelGGL(k_trace_backward_stored_det, grid, block, 0, s, A);
        GRAY_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_trace_backward_recompute<true>, grid, block, 0, s, A, A.geo, A.nodes);
    GRAY_HIP(hipGetLastError());
    int bits = 0;
    while (bits < 32 && (none >> bits) != 0) ++bits;
    uint32_t *keys = nullptr, *vals = nullptr;
    if ((rc = gray_radix_sort_pairs(ctx->det_keys[0], ctx->det_vals[0], ctx->det_keys[1], ctx->det_vals[1],
                                    (int)recs, (bits + 7) / 8, ctx->det_hist, ctx->det_sums, s, &keys, &vals)))
        return rc;
    hipLaunchKernelGGL(k_det_reduce, dim3(gray_div_up(recs, 256)), dim3(256), 0, s, A, keys, vals, (int)recs);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

int check_render(gray_ctx *ctx, const gray_render_args *args) {
    if (!ctx || !args || !args->views || args->num_views <= 0) {
        gray_set_error(GRAY_ERR_INVALID, "render: null context/arguments or no views");
        return GRAY_ERR_INVALID;
    }
    if (!scene_of(ctx)->built) {
        gray_set_error(GRAY_ERR_STATE, "render: no BVH (call gray_bvh_update with rebuild=1 first)");
        return GRAY_ERR_STATE;
    }
    return GRAY_OK;
}

}  // namespace

extern "C" int gray_render_forward(gray_ctx *ctx, const gray_render_args *args, void *stream) {
    int rc = check_render(ctx, args);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    // colours from the SH coefficients on the side stream, beside the collect
    // pass (which does not read them); joined before the first kernel that does
    bool sh_pending = false;
    if (args->sh_params) {
        std::vector<const float *> origins(args->num_views);
        std::vector<float *> outs(args->num_views);
        for (int i = 0; i < args->num_views; ++i) {
            origins[i] = args->views[i].origin;
            outs[i] = const_cast<float *>(args->views[i].rgb);
            if (!origins[i] || !outs[i]) {
                gray_set_error(GRAY_ERR_INVALID, "render_forward(sh_params): view %d lacks an origin or rgb", i);
                return GRAY_ERR_INVALID;
            }
        }
        GRAY_HIP(hipEventRecord(ctx->ev_fork, s));
        GRAY_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        if ((rc = gray_sh_forward(ctx, args->sh_params, &args->knobs, args->num_views, origins.data(), outs.data(),
                                  ctx->side)))
            return rc;
        GRAY_HIP(hipEventRecord(ctx->ev_sh, ctx->side));
        sh_pending = true;
    }
    auto join_sh = [&]() -> int {
        if (sh_pending) GRAY_HIP(hipStreamWaitEvent(s, ctx->ev_sh, 0));
        sh_pending = false;
        return GRAY_OK;
    };
    int tiles_before = 0;
    for (int first = 0; first < args->num_views; first += GRAY_MAX_VIEWS) {
        const int nv = std::min(GRAY_MAX_VIEWS, args->num_views - first);
        TraceArgs A;
        if ((rc = fill_args(ctx, args, first, nv, A))) return rc;
        A.tile_base_launch = tiles_before;
        tiles_before += A.total_tiles;
        if (A.save) {
            for (int i = 0; i < nv; ++i) {
                const DevView &d = A.views[i];
                if (!d.out_T || !d.out_F || !d.out_est || !d.out_ro || !d.out_rd || !d.n_comp || !d.comp_rgb ||
                    (A.capacity > 0 && (!d.hit_ids || !d.hit_T))) {
                    gray_set_error(GRAY_ERR_INVALID, "render_forward(save): view %d lacks backward-state buffers",
                                   first + i);
                    return GRAY_ERR_INVALID;
                }
            }
        }
        if (A.total_tiles == 0) continue;
        const dim3 grid(8 * A.xcd_chunk), block(GRAY_WAVE);
        // stage events (first chunk of views only): fev[0..4] bound collect,
        // task rounds, list compositing and re-traversal
        const bool timed = ctx->timing && first == 0;
        auto mark = [&](int i) -> int {
            if (timed) GRAY_HIP(hipEventRecord(ctx->fev[i], s));
            return GRAY_OK;
        };
        if (args->list_capacity > 0 && A.use_bvh && A.num_gaussians > 0) {
            const int64_t lanes = (int64_t)A.total_tiles * GRAY_WAVE;
            if ((rc = reserve_lists(ctx, lanes, args->list_capacity))) return rc;
            A.list_rec = ctx->list_rec;
            A.list_n = ctx->list_n;
            A.tile_open = ctx->tile_open;
            A.list_stride = lanes;
            A.list_cap = args->list_capacity;
            A.task_cap = (int32_t)ctx->task_cap;
            ctx->list_lanes = lanes;
            // counters: count[r] tasks handed to round r, head[r] its take position
            int32_t *count = ctx->task_counters, *head = ctx->task_counters + kTaskRounds + 1;
            hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, ctx->task_counters, kCounters);
            GRAY_HIP(hipGetLastError());
            A.task_out = ctx->tasks[0];
            A.task_count_out = count;
            // deterministic: every walk runs to its end in its own lane with a
            // 64-entry stack (an overflow sends the ray to re-traversal), so each
            // list has one order
            const bool tasks = !args->deterministic;
            // tile bins (collect_mode 0) for pinhole views; the walk below runs
            // instead when the bins' set-up kernel or scan sets *bin_mode
            const bool binned = tasks && args->collect_mode == 0;
            A.budget = !tasks ? -1 : args->collect_budget > 0 ? args->collect_budget : GRAY_COLLECT_BUDGET;
            // compositing: the task buffers are free once the last round has
            // run, and the first then holds the queue of rays composited by a
            // whole wave (one entry per lane)
            A.heavy_count = ctx->task_counters + 2 * kTaskRounds + 2;
            A.heavy_rays = reinterpret_cast<uint32_t *>(ctx->tasks[0]);
            A.wave_list_min = args->wave_list_min == 0 ? GRAY_WAVE_LIST_MIN : args->wave_list_min;
            A.retrace_any = ctx->task_counters + 2 * kTaskRounds + 3;  // zeroed with the task counters
            if ((rc = mark(0))) return rc;
            const Bvh4Node *cnodes = scene_of(ctx)->wide;
            if (binned) {
                A.bins = reinterpret_cast<uint32_t *>(ctx->tasks[1]);
                A.bin_cap = bin_capacity(ctx->task_cap);
                A.bin_tiles = ctx->bin_tiles;
                A.bin_items = ctx->bin_items;
                A.bin_item_cap = (int32_t)std::min<int64_t>(ctx->bin_item_cap, INT32_MAX);
                A.bin_mode = ctx->task_counters + 2 * kTaskRounds + 4;  // zeroed with the task counters
                const int T = A.total_tiles, n = A.num_gaussians;
                const float *aabb = scene_of(ctx)->aabb;
                hipLaunchKernelGGL(k_bin_setup, dim3(gray_div_up(std::max<int64_t>(lanes, 4 * (int64_t)T), 256)),
                                   dim3(256), 0, s, A, lanes);
                const uint32_t *order = scene_of(ctx)->sorted_ids;  // leaves in Morton order
                hipLaunchKernelGGL(k_bin_pairs, dim3(gray_div_up(n, kBinThreads)), dim3(kBinThreads), 0, s, A, aabb,
                                   order, n, 0);
                hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kBinScanThreads), 0, s, A);
                hipLaunchKernelGGL(k_bin_pairs, dim3(gray_div_up(n, kBinThreads)), dim3(kBinThreads), 0, s, A, aabb,
                                   order, n, 1);
                hipLaunchKernelGGL(k_bin_items, dim3(gray_div_up(T, 256)), dim3(256), 0, s, A);
                hipLaunchKernelGGL(k_collect_binned, dim3(8 * std::min(gray_div_up(A.bin_item_cap, 8), 4096u)), block, 0, s, A,
                                   A.geo);
                GRAY_HIP(hipGetLastError());
            }
            if (tasks)
                hipLaunchKernelGGL(k_collect<kCollectStack>, grid, block, 0, s, A, A.geo, cnodes);
            else
                hipLaunchKernelGGL(k_collect<kDetStack>, grid, block, 0, s, A, A.geo, cnodes);
            GRAY_HIP(hipGetLastError());
            if ((rc = mark(1))) return rc;
            // tiles whose rays handed nothing on have final lists: they are
            // composited on the side stream (after the colours) beside the
            // pooled round, which walks only the other tiles' rays
            const bool early = tasks && !binned;
            bool side_work = false;
            if (early) {
                GRAY_HIP(hipEventRecord(ctx->ev_fork, s));
                GRAY_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
                hipLaunchKernelGGL(k_composite_lists, grid, block, 0, ctx->side, A, A.views[0].rgb, 1);
                GRAY_HIP(hipGetLastError());
                side_work = true;
            }
            // pooled rounds; all but the last may hand work on (budget >= 0)
            for (int round = 0; tasks && round < kTaskRounds; ++round) {
                A.task_in = ctx->tasks[round & 1];
                A.task_count_in = count + round;
                A.task_head = head + round;
                A.task_out = ctx->tasks[(round + 1) & 1];
                A.task_count_out = count + round + 1;
                A.budget = round + 1 < kTaskRounds ? 0 : -1;
                // later rounds take only what an overflowing stack handed on
                // (usually nothing): a small grid, so that an empty round costs
                // few atomics on its head (atomics on one address serialise:
                // 4096 waves taking nothing cost about 40 us)
                hipLaunchKernelGGL(k_collect_pool, dim3(round == 0 ? kTaskWaves : kPoolSpillWaves), block, 0, s, A,
                                   A.geo, cnodes);
                GRAY_HIP(hipGetLastError());
            }
            if ((rc = mark(2))) return rc;
            if ((rc = join_sh())) return rc;
            if (A.wave_list_min >= 0) {
                // the long lists are queued first; their wave-per-ray kernel then
                // runs on the side stream beside the per-lane kernel (both are
                // latency-bound) and is joined back before the re-traversal
                hipLaunchKernelGGL(k_heavy_queue, grid, block, 0, s, A);
                GRAY_HIP(hipGetLastError());
                GRAY_HIP(hipEventRecord(ctx->ev_fork, s));
                GRAY_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
                hipLaunchKernelGGL(k_composite_wave, dim3(kWaveGrid), block, 0, ctx->side, A, A.views[0].rgb);
                GRAY_HIP(hipGetLastError());
                side_work = true;
            }
            hipLaunchKernelGGL(k_composite_lists, grid, block, 0, s, A, A.views[0].rgb, early ? 2 : 0);
            GRAY_HIP(hipGetLastError());
            // joined before the re-traversal, and before the next chunk of views
            // reuses the lists
            if (side_work) {
                GRAY_HIP(hipEventRecord(ctx->ev_join, ctx->side));
                GRAY_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
            }
        } else {
            if ((rc = join_sh())) return rc;
            for (int i = 0; i < 3; ++i)
                if ((rc = mark(i))) return rc;
        }
        if ((rc = mark(3))) return rc;
        hipLaunchKernelGGL(k_trace_forward, grid, block, 0, s, A, A.geo, A.nodes, A.views[0].rgb);
        GRAY_HIP(hipGetLastError());
        if ((rc = mark(4))) return rc;
        if (timed) ctx->fev_recorded = true;
    }
    return join_sh();  // also when no view had a tile
}

extern "C" int gray_render_stage_ms(gray_ctx *ctx, float *ms, int32_t max_stages) {
    if (!ctx || !ms || max_stages < 4) {
        gray_set_error(GRAY_ERR_INVALID, "render_stage_ms: null pointer or fewer than 4 stages");
        return GRAY_ERR_INVALID;
    }
    if (!ctx->fev_recorded) {
        gray_set_error(GRAY_ERR_STATE, "render_stage_ms: no timed forward (enable timing first)");
        return GRAY_ERR_STATE;
    }
    GRAY_HIP(hipEventSynchronize(ctx->fev[4]));
    for (int i = 0; i < 4; ++i) GRAY_HIP(hipEventElapsedTime(&ms[i], ctx->fev[i], ctx->fev[i + 1]));
    return GRAY_OK;
}

extern "C" int gray_render_reserve(gray_ctx *ctx, int64_t max_lanes, int32_t list_capacity) {
    if (!ctx || max_lanes < 0 || list_capacity < 0) {
        gray_set_error(GRAY_ERR_INVALID, "render_reserve: null context or negative size");
        return GRAY_ERR_INVALID;
    }
    if (max_lanes == 0 || list_capacity == 0) return GRAY_OK;
    return reserve_lists(ctx, max_lanes, list_capacity);
}

extern "C" int gray_debug_list_counts(gray_ctx *ctx, int32_t *counts, int64_t max_lanes, int64_t *lanes,
                                      void *stream) {
    if (!ctx || !lanes) {
        gray_set_error(GRAY_ERR_INVALID, "debug_list_counts: null context or lanes pointer");
        return GRAY_ERR_INVALID;
    }
    *lanes = ctx->list_lanes;
    if (!counts || ctx->list_lanes == 0) return GRAY_OK;
    if (max_lanes < ctx->list_lanes) {
        gray_set_error(GRAY_ERR_INVALID, "debug_list_counts: %lld entries needed, room for %lld",
                       (long long)ctx->list_lanes, (long long)max_lanes);
        return GRAY_ERR_INVALID;
    }
    GRAY_HIP(hipMemcpyAsync(counts, ctx->list_n, sizeof(int32_t) * ctx->list_lanes, hipMemcpyDeviceToDevice,
                            (hipStream_t)stream));
    return GRAY_OK;
}

extern "C" int gray_render_backward(gray_ctx *ctx, const gray_render_args *args, const gray_params *params,
                                    const gray_grads *grads, void *stream) {
    int rc = check_render(ctx, args);
    if (rc) return rc;
    if (!params || !grads || !grads->grad_mean || !grads->grad_rotation || !grads->grad_scale || !grads->grad_opacity) {
        gray_set_error(GRAY_ERR_INVALID, "render_backward: null params or gradient pointers");
        return GRAY_ERR_INVALID;
    }
    const int64_t n_scene = scene_of(ctx)->n;
    if (params->count != n_scene) {
        gray_set_error(GRAY_ERR_STATE, "render_backward: %lld Gaussians but the snapshot has %lld",
                       (long long)params->count, (long long)n_scene);
        return GRAY_ERR_STATE;
    }
    if (n_scene == 0) return GRAY_OK;
    hipStream_t s = (hipStream_t)stream;
    const bool stored = !args->deterministic && args->hit_capacity > 0;
    if (stored && ctx->raw_cap < n_scene) {
        // rows of raw sums, zero between backward calls (k_raw_apply clears what it reads)
        GRAY_HIP(hipFree(ctx->raw));
        GRAY_HIP(hipFree(ctx->raw_touched));
        GRAY_HIP(hipFree(ctx->bwd_cnt));
        GRAY_HIP(hipFree(ctx->bwd_lst));
        ctx->raw = nullptr;
        ctx->raw_touched = nullptr;
        ctx->bwd_cnt = nullptr;
        ctx->bwd_lst = nullptr;
        ctx->raw_cap = 0;
        const int64_t words = (n_scene + 3) / 4;
        if (hipMalloc(&ctx->raw, sizeof(float) * 16 * n_scene) != hipSuccess ||
            hipMalloc(&ctx->raw_touched, sizeof(uint32_t) * words) != hipSuccess ||
            hipMalloc(&ctx->bwd_cnt, sizeof(uint32_t) * n_scene) != hipSuccess ||
            hipMalloc(&ctx->bwd_lst, sizeof(uint32_t) * kRecList * n_scene) != hipSuccess) {
            (void)hipGetLastError();
            gray_set_error(GRAY_ERR_NOMEM, "render_backward: cannot allocate the raw gradient rows (%lld bytes)",
                           (long long)(64 * n_scene));
            return GRAY_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_zero_u32, dim3(2048), dim3(256), 0, s, reinterpret_cast<uint32_t *>(ctx->raw),
                           16 * n_scene);
        hipLaunchKernelGGL(k_zero_u32, dim3(256), dim3(256), 0, s, reinterpret_cast<uint32_t *>(ctx->raw_touched),
                           words);
        hipLaunchKernelGGL(k_zero_u32, dim3(256), dim3(256), 0, s, ctx->bwd_cnt, n_scene);
        GRAY_HIP(hipGetLastError());
        ctx->raw_cap = n_scene;
    }
    if (stored) {
        // per-tile overflow marks of the launch with the most tiles
        int64_t tiles = 0;
        for (int first = 0; first < args->num_views; first += GRAY_MAX_VIEWS) {
            TraceArgs T;
            if ((rc = fill_args(ctx, args, first, std::min(GRAY_MAX_VIEWS, args->num_views - first), T))) return rc;
            tiles = std::max<int64_t>(tiles, T.total_tiles);
        }
        if (ctx->bwd_overflow_cap < tiles) {
            GRAY_HIP(hipFree(ctx->bwd_overflow));
            GRAY_HIP(hipFree(ctx->tile_cost));
            GRAY_HIP(hipFree(ctx->tile_order));
            GRAY_HIP(hipFree(ctx->tile_next));
            GRAY_HIP(hipFree(ctx->bwd_rec));
            ctx->bwd_rec = nullptr;
            ctx->bwd_overflow = nullptr;
            ctx->tile_cost = nullptr;
            ctx->tile_order = nullptr;
            ctx->tile_next = nullptr;
            ctx->bwd_overflow_cap = 0;
            const int64_t words = (tiles + 3) / 4;
            if (hipMalloc(&ctx->bwd_overflow, sizeof(uint32_t) * words) != hipSuccess ||
                hipMalloc(&ctx->tile_cost, sizeof(uint32_t) * words) != hipSuccess ||
                hipMalloc(&ctx->tile_order, sizeof(uint32_t) * tiles) != hipSuccess ||
                hipMalloc(&ctx->tile_next, sizeof(int32_t)) != hipSuccess ||
                hipMalloc(&ctx->bwd_rec, sizeof(float) * 16 * kGradSlots * tiles) != hipSuccess) {
                (void)hipGetLastError();
                gray_set_error(GRAY_ERR_NOMEM, "render_backward: cannot allocate the tile overflow marks");
                return GRAY_ERR_NOMEM;
            }
            hipLaunchKernelGGL(k_zero_u32, dim3(64), dim3(256), 0, s, reinterpret_cast<uint32_t *>(ctx->bwd_overflow),
                               words);
            GRAY_HIP(hipGetLastError());
            ctx->bwd_overflow_cap = tiles;
        }
    }
    bool apply = false;
    TraceArgs last;
    // deterministic: one view per launch, so each Gaussian's sum runs in one thread
    const int chunk = args->deterministic ? 1 : GRAY_MAX_VIEWS;
    for (int first = 0; first < args->num_views; first += chunk) {
        const int nv = std::min(chunk, args->num_views - first);
        TraceArgs A;
        if ((rc = fill_args(ctx, args, first, nv, A))) return rc;
        A.raw = ctx->raw;
        A.raw_touched = ctx->raw_touched;
        A.bwd_rec = ctx->bwd_rec;
        A.bwd_cnt = ctx->bwd_cnt;
        A.bwd_lst = ctx->bwd_lst;
        A.bwd_overflow = stored ? ctx->bwd_overflow : nullptr;
        for (int i = 0; i < nv; ++i) {
            const DevView &d = A.views[i];
            if (!d.out_T || !d.out_F || !d.out_est || !d.out_ro || !d.out_rd || !d.n_comp || !d.comp_rgb ||
                !d.grad_out || !d.grad_rgb || (A.capacity > 0 && (!d.hit_ids || !d.hit_T))) {
                gray_set_error(GRAY_ERR_INVALID, "render_backward: view %d lacks forward state or grad buffers",
                               first + i);
                return GRAY_ERR_INVALID;
            }
        }
        A.rotation = params->rotation;
        A.scale = params->scale;
        A.opacity = params->opacity;
        A.g_mean = grads->grad_mean;
        A.g_rot = grads->grad_rotation;
        A.g_scale = grads->grad_scale;
        A.g_op = grads->grad_opacity;
        A.g_prune = grads->pruning_weight;
        A.visible = grads->was_visible;
        if (A.total_tiles == 0) continue;
        if (args->deterministic) {
            if ((rc = det_backward(ctx, A, s))) return rc;
            continue;
        }
        if (A.capacity > 0) {
            if (ctx->stored_grid == 0) {
                int per_cu = 0, cus = 0;
                GRAY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_backward_stored, GRAY_WAVE, 0));
                GRAY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
                ctx->stored_grid = std::max(1, per_cu) * std::max(1, cus);
            }
            hipLaunchKernelGGL(k_tile_cost, dim3(8 * A.xcd_chunk), dim3(GRAY_WAVE), 0, s, A, ctx->tile_cost);
            hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderThreads), 0, s, ctx->tile_cost, A.total_tiles,
                               ctx->tile_order, ctx->tile_next);
            hipLaunchKernelGGL(k_trace_backward_stored, dim3(std::min(ctx->stored_grid, A.total_tiles)),
                               dim3(GRAY_WAVE), 0, s, A, ctx->tile_order, ctx->tile_next);
            GRAY_HIP(hipGetLastError());
            apply = true;
            last = A;
        }
        hipLaunchKernelGGL(k_trace_backward_recompute<false>, dim3(8 * A.xcd_chunk), dim3(GRAY_WAVE), 0, s, A, A.geo,
                           A.nodes);
        GRAY_HIP(hipGetLastError());
        if (apply) {
            // this launch's records (their view indices are the launch's)
            hipLaunchKernelGGL(k_raw_apply, dim3(gray_div_up(n_scene, 256)), dim3(256), 0, s, last, (int)n_scene);
            GRAY_HIP(hipGetLastError());
            apply = false;
        }
    }
    return GRAY_OK;
}

extern "C" int gray_render_ellipsoids(gray_ctx *ctx, const gray_render_args *args, const float *min_opacity,
                                      void *stream) {
    int rc = check_render(ctx, args);
    if (rc) return rc;
    if (!min_opacity) {
        gray_set_error(GRAY_ERR_INVALID, "render_ellipsoids: null min_opacity");
        return GRAY_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    // deterministic: one view per launch, so each Gaussian's sum runs in one thread
    const int chunk = args->deterministic ? 1 : GRAY_MAX_VIEWS;
    for (int first = 0; first < args->num_views; first += chunk) {
        const int nv = std::min(chunk, args->num_views - first);
        TraceArgs A;
        if ((rc = fill_args(ctx, args, first, nv, A))) return rc;
        for (int i = 0; i < nv; ++i) {
            if (!A.views[i].out_c || !A.views[i].rgb) {
                gray_set_error(GRAY_ERR_INVALID, "render_ellipsoids: view %d needs rgb and out_channels", first + i);
                return GRAY_ERR_INVALID;
            }
        }
        if (A.total_tiles == 0) continue;
        hipLaunchKernelGGL(k_ellipsoids, dim3(8 * A.xcd_chunk), dim3(GRAY_WAVE), 0, s, A, A.geo, A.nodes, min_opacity);
        GRAY_HIP(hipGetLastError());
    }
    return GRAY_OK;
}

extern "C" int gray_debug_hitlists(gray_ctx *ctx, const gray_render_args *args, int32_t max_hits, uint32_t *ids,
                                   float *dists, int32_t *counts, void *stream) {
    int rc = check_render(ctx, args);
    if (rc) return rc;
    if (!ids || !dists || !counts || max_hits <= 0) {
        gray_set_error(GRAY_ERR_INVALID, "debug_hitlists: null output or max_hits <= 0");
        return GRAY_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    TraceArgs A;
    if ((rc = fill_args(ctx, args, 0, 1, A))) return rc;
    A.dbg_ids = ids;
    A.dbg_dist = dists;
    A.dbg_count = counts;
    A.dbg_max = max_hits;
    if (A.total_tiles == 0) return GRAY_OK;
    hipLaunchKernelGGL(k_debug_hits, dim3(8 * A.xcd_chunk), dim3(GRAY_WAVE), 0, s, A, A.geo, A.nodes);
    GRAY_HIP(hipGetLastError());
    const int rays = A.views[0].width * A.views[0].height;
    hipLaunchKernelGGL(k_debug_sort, dim3(gray_div_up(rays, 64)), dim3(64), 0, s, rays, max_hits, ids, dists, counts);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}
