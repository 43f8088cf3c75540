// Geometry snapshot + LBVH build/refit for gfx950.
//
// Replaces the OptiX instance acceleration structure of the reference
// (cuda/optix/bvh_wrapper.h:7-124, bvh_wrapper.cu:8-53): one ellipsoid proxy
// per Gaussian (visible iff clip > 0 and any scaled axis > 0, bvh_wrapper.cu:36)
// with an exact AABB instead of the transformed unit box.
//
// Build:  snapshot -> centroid bounds -> 30-bit Morton codes -> LSD radix sort
//         (8-bit digits, LDS histograms, wave-ballot stable ranking) ->
//         Karras 2012 topology -> refit.
// Refit:  phase 1: one workgroup per 1024-leaf chunk refits every node whose
//         leaf range lies inside the chunk, bottom-up with LDS counters over
//         LDS-staged node metadata, and writes the chunk prefix/suffix boxes
//         that spanning nodes read; phase 2: sparse table over the
//         chunk boxes (one workgroup); phase 3: every node spanning chunks
//         takes suffix(a) | table(chunks) | prefix(b) independently.  No kernel
//         depends on another workgroup's plain stores within the same launch.

#include <float.h>
#include <math.h>
#include <stdio.h>

#include <algorithm>

#include <hip/hip_fp16.h>

#include "ctx.hpp"

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per workgroup
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ void box_union(float *a, const float *b) {
    a[0] = fminf(a[0], b[0]);
    a[1] = fminf(a[1], b[1]);
    a[2] = fminf(a[2], b[2]);
    a[3] = fmaxf(a[3], b[3]);
    a[4] = fmaxf(a[4], b[4]);
    a[5] = fmaxf(a[5], b[5]);
}

__device__ __forceinline__ void box_empty(float *a) {
    a[0] = a[1] = a[2] = INFINITY;
    a[3] = a[4] = a[5] = -INFINITY;
}

// ---------------------------------------------------------------------------
// Snapshot: reference populate_bvh_kernel + create_transform_matrix
// (bvh_wrapper.cu:8-53) and compute_scaling_factor (kernel.cu:3-22).
// ---------------------------------------------------------------------------
struct SnapF3 {
    float x, y, z;
};

constexpr int kSnapThreads = 256;

__global__ __launch_bounds__(kSnapThreads) void k_snapshot(int n, const float *__restrict__ mean, const float *__restrict__ rotation,
                           const float *__restrict__ scale, const float *__restrict__ opacity, DevKnobs kn,
                           GeoRecord *__restrict__ geo, float *__restrict__ l2w, float *__restrict__ aabb) {
    // the block's records are assembled in LDS and stored by whole lines (a
    // thread's own 64-, 48- and 24-byte records would each cover parts of
    // lines per store instruction): 56 -> 44 us at 1M Gaussians
    __shared__ float4 s_geo[kSnapThreads * 4], s_l2w[kSnapThreads * 3];
    __shared__ float2 s_box[kSnapThreads * 3];
    const int g0 = blockIdx.x * kSnapThreads, t = threadIdx.x;
    const int g = min(g0 + t, n - 1);  // a tail thread recomputes the last record; its copy is not stored
    const float athr = *kn.alpha_threshold, p = *kn.exp_power, gsf = *kn.gsf;
    // the parameters were last written by the optimizer's non-temporal
    // stores and are read once here: non-temporal loads keep them from
    // allocating in the caches (measured: the refit chain 0.180 -> 0.175 ms)
    const SnapF3 ls = {__builtin_nontemporal_load(scale + 3 * g), __builtin_nontemporal_load(scale + 3 * g + 1),
                       __builtin_nontemporal_load(scale + 3 * g + 2)};
    const SnapF3 mv = {__builtin_nontemporal_load(mean + 3 * g), __builtin_nontemporal_load(mean + 3 * g + 1),
                       __builtin_nontemporal_load(mean + 3 * g + 2)};
    const float op = gray_sigmoid(__builtin_nontemporal_load(opacity + g));
    const float clip = gray_clip_factor(op, athr, p);
    const float sx = (expf(ls.x) * clip) * gsf;
    const float sy = (expf(ls.y) * clip) * gsf;
    const float sz = (expf(ls.z) * clip) * gsf;
    const bool visible = clip > 0.0f && (sx > 0.0f || sy > 0.0f || sz > 0.0f);

    typedef float snap_f4 __attribute__((ext_vector_type(4)));
    const snap_f4 qv = __builtin_nontemporal_load(reinterpret_cast<const snap_f4 *>(rotation) + g);
    const float4 q = make_float4(qv.x, qv.y, qv.z, qv.w);
    const float inv_norm = 1.0f / sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    const float r = q.x * inv_norm, x = q.y * inv_norm, y = q.z * inv_norm, z = q.w * inv_norm;
    float R[9];
    R[0] = 1.f - 2.f * (y * y + z * z);
    R[1] = 2.f * (x * y - r * z);
    R[2] = 2.f * (x * z + r * y);
    R[3] = 2.f * (x * y + r * z);
    R[4] = 1.f - 2.f * (x * x + z * z);
    R[5] = 2.f * (y * z - r * x);
    R[6] = 2.f * (x * z - r * y);
    R[7] = 2.f * (y * z + r * x);
    R[8] = 1.f - 2.f * (x * x + y * y);
    const float s[3] = {sx, sy, sz};
    const float inv[3] = {1.0f / sx, 1.0f / sy, 1.0f / sz};

    GeoRecord rec;
    float A[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            A[3 * i + j] = R[3 * i + j] * s[j];      // local-to-world linear part
            rec.m[3 * i + j] = R[3 * j + i] * inv[i];  // its inverse diag(1/s) R^T
        }
    }
    const float mu[3] = {mv.x, mv.y, mv.z};
    rec.mu[0] = mu[0];
    rec.mu[1] = mu[1];
    rec.mu[2] = mu[2];
    rec.op = visible ? op : -1.0f;
    rec.clip = clip;
    rec.pad0 = rec.pad1 = 0.0f;
    {
        const float *rf = reinterpret_cast<const float *>(&rec);
#pragma unroll
        for (int k = 0; k < 4; ++k) s_geo[4 * t + k] = make_float4(rf[4 * k], rf[4 * k + 1], rf[4 * k + 2], rf[4 * k + 3]);
        s_l2w[3 * t + 0] = make_float4(A[0], A[1], A[2], A[3]);
        s_l2w[3 * t + 1] = make_float4(A[4], A[5], A[6], A[7]);
        s_l2w[3 * t + 2] = make_float4(A[8], 0.0f, 0.0f, 0.0f);
    }

    float box[6];
    if (visible) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float h = sqrtf(A[3 * i] * A[3 * i] + A[3 * i + 1] * A[3 * i + 1] + A[3 * i + 2] * A[3 * i + 2]);
            // conservative padding: covers float rounding of the hit test and slab test
            const float pad = 1e-4f * h + 1e-6f * (fabsf(mu[i]) + 1.0f);
            box[i] = mu[i] - (h + pad);
            box[3 + i] = mu[i] + (h + pad);
        }
    } else {
        box_empty(box);
    }
    s_box[3 * t + 0] = make_float2(box[0], box[1]);
    s_box[3 * t + 1] = make_float2(box[2], box[3]);
    s_box[3 * t + 2] = make_float2(box[4], box[5]);
    __syncthreads();
    const int cnt = min(kSnapThreads, n - g0);
    float4 *dg = reinterpret_cast<float4 *>(geo + g0);
    float4 *dl = reinterpret_cast<float4 *>(l2w + 12 * (size_t)g0);
    float2 *db = reinterpret_cast<float2 *>(aabb + 6 * (size_t)g0);
    for (int i = t; i < 4 * cnt; i += kSnapThreads) dg[i] = s_geo[i];
    for (int i = t; i < 3 * cnt; i += kSnapThreads) dl[i] = s_l2w[i];
    for (int i = t; i < 3 * cnt; i += kSnapThreads) db[i] = s_box[i];
}

// ---------------------------------------------------------------------------
// Centroid bounds (two-level reduction) and Morton codes.
// ---------------------------------------------------------------------------
__global__ void k_bounds_partial(int n, const float *__restrict__ mean, float *__restrict__ partial) {
    __shared__ float sh[6][256];
    float b[6];
    box_empty(b);
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        const float x = mean[3 * g], y = mean[3 * g + 1], z = mean[3 * g + 2];
        b[0] = fminf(b[0], x);
        b[1] = fminf(b[1], y);
        b[2] = fminf(b[2], z);
        b[3] = fmaxf(b[3], x);
        b[4] = fmaxf(b[4], y);
        b[5] = fmaxf(b[5], z);
    }
    for (int i = 0; i < 6; ++i) sh[i][threadIdx.x] = b[i];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int i = 0; i < 3; ++i) sh[i][threadIdx.x] = fminf(sh[i][threadIdx.x], sh[i][threadIdx.x + off]);
            for (int i = 3; i < 6; ++i) sh[i][threadIdx.x] = fmaxf(sh[i][threadIdx.x], sh[i][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_bounds_final(int blocks, const float *__restrict__ partial, float *__restrict__ out) {
    __shared__ float sh[6][256];
    float b[6];
    box_empty(b);
    for (int i = threadIdx.x; i < blocks; i += blockDim.x) {
        for (int k = 0; k < 3; ++k) b[k] = fminf(b[k], partial[6 * i + k]);
        for (int k = 3; k < 6; ++k) b[k] = fmaxf(b[k], partial[6 * i + k]);
    }
    for (int i = 0; i < 6; ++i) sh[i][threadIdx.x] = b[i];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int i = 0; i < 3; ++i) sh[i][threadIdx.x] = fminf(sh[i][threadIdx.x], sh[i][threadIdx.x + off]);
            for (int i = 3; i < 6; ++i) sh[i][threadIdx.x] = fmaxf(sh[i][threadIdx.x], sh[i][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = sh[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton(int n, const float *__restrict__ mean, const float *__restrict__ bounds,
                         uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    uint32_t code = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float lo = bounds[i], hi = bounds[3 + i];
        const float ext = hi - lo;
        float u = ext > 0.0f ? (mean[3 * g + i] - lo) / ext : 0.5f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        const uint32_t q = (uint32_t)fminf(u * 1024.0f, 1023.0f);
        code |= spread10(q) << (2 - i);
    }
    keys[g] = code;
    vals[g] = (uint32_t)g;
}

// ---------------------------------------------------------------------------
// LSD radix sort (8-bit digits).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(int n, const uint32_t *__restrict__ keys, int shift,
                                                             uint32_t *__restrict__ hist, int tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int idx = base + j * kSortThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan, three kernels
__global__ __launch_bounds__(kScanThreads) void k_scan_block(uint32_t *__restrict__ data, int m,
                                                             uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[kScanThreads];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int idx = base + j;
        v[j] = idx < m ? data[idx] : 0u;
        const uint32_t t = v[j];
        v[j] = local;
        local += t;
    }
    sh[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    const uint32_t excl = sh[threadIdx.x] - local;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int idx = base + j;
        if (idx < m) data[idx] = v[j] + excl;
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *__restrict__ sums, int count) {
    __shared__ uint32_t sh[1024];
    uint32_t carry = 0;
    for (int base = 0; base < count; base += 1024) {
        const int idx = base + threadIdx.x;
        const uint32_t v = idx < count ? sums[idx] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t add = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (idx < count) sums[idx] = carry + sh[threadIdx.x] - v;
        const uint32_t total = sh[1023];
        __syncthreads();
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t *__restrict__ data, int m,
                                                           const uint32_t *__restrict__ sums) {
    const uint32_t add = sums[blockIdx.x];
    const int base = blockIdx.x * kScanTile;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int idx = base + j * kScanThreads + threadIdx.x;
        if (idx < m) data[idx] += add;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(int n, const uint32_t *__restrict__ keys_in,
                                                                const uint32_t *__restrict__ vals_in,
                                                                uint32_t *__restrict__ keys_out,
                                                                uint32_t *__restrict__ vals_out, int shift,
                                                                const uint32_t *__restrict__ hist, int tiles) {
    __shared__ uint32_t goff[256];
    __shared__ uint32_t dbase[256];
    __shared__ uint32_t wcount[kSortThreads / GRAY_WAVE][256];
    const int tid = threadIdx.x;
    const int wave = tid / GRAY_WAVE;
    const int lane = tid % GRAY_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    goff[tid] = hist[tid * tiles + blockIdx.x];
    dbase[tid] = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / GRAY_WAVE; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
        const int idx = base + j * kSortThreads + tid;
        const bool valid = idx < n;
        const uint32_t key = valid ? keys_in[idx] : 0u;
        const uint32_t val = valid ? vals_in[idx] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) wcount[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = goff[d] + dbase[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcount[w][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kSortThreads / GRAY_WAVE; ++w) {
            add += wcount[w][tid];
            wcount[w][tid] = 0;
        }
        dbase[tid] += add;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Karras 2012 internal-node construction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t *__restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ void k_karras(int n, const uint32_t *__restrict__ keys, const uint32_t *__restrict__ sorted_ids,
                         BvhNode *__restrict__ nodes, uint32_t *__restrict__ parent_int,
                         uint32_t *__restrict__ parent_leaf, uint32_t *__restrict__ range_lo,
                         uint32_t *__restrict__ range_hi, uint32_t *__restrict__ top_nodes,
                         uint32_t *__restrict__ top_count, uint8_t *__restrict__ need_ends,
                         uint4 *__restrict__ node_info) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = min(i, j), hi = max(i, j);
    uint32_t left, right;
    if (lo == gamma) {
        left = GRAY_LEAF | sorted_ids[gamma];
        parent_leaf[gamma] = (uint32_t)i;
    } else {
        left = (uint32_t)gamma;
        parent_int[gamma] = (uint32_t)i;
        node_info[gamma].w = (uint32_t)i;
    }
    if (hi == gamma + 1) {
        right = GRAY_LEAF | sorted_ids[gamma + 1];
        parent_leaf[gamma + 1] = (uint32_t)i | 0x80000000u;
    } else {
        right = (uint32_t)(gamma + 1);
        parent_int[gamma + 1] = (uint32_t)i | 0x80000000u;
        node_info[gamma + 1].w = (uint32_t)i | 0x80000000u;
    }
    // .w is the parent, written by the parent's thread
    node_info[i].x = (uint32_t)lo;
    node_info[i].y = (uint32_t)hi;
    node_info[i].z = (uint32_t)gamma;
    if (i == 0) node_info[0].w = GRAY_EMPTY;
    nodes[i].left = left;
    nodes[i].right = right;
    nodes[i].pad0 = (uint32_t)lo;
    nodes[i].pad1 = (uint32_t)hi;
    range_lo[i] = (uint32_t)lo;
    range_hi[i] = (uint32_t)hi;
    if (i == 0) parent_int[0] = GRAY_EMPTY;
    if (lo / GRAY_CHUNK != hi / GRAY_CHUNK) {
        const uint32_t slot = atomicAdd(top_count, 1u);
        top_nodes[slot] = (uint32_t)i;
        need_ends[2 * lo] = 1;      // suffix of lo within its chunk
        need_ends[2 * hi + 1] = 1;  // prefix of hi within its chunk
    }
}

// N == 1: no sort runs, so the leaf order (sorted_ids, read by the refit) and
// the root are written here.
__global__ void k_single_leaf(BvhNode *__restrict__ nodes, uint32_t *__restrict__ parent_int,
                              uint32_t *__restrict__ parent_leaf, uint32_t *__restrict__ range_lo,
                              uint32_t *__restrict__ range_hi, uint32_t *__restrict__ sorted_ids,
                              uint32_t *__restrict__ keys) {
    sorted_ids[0] = 0u;
    keys[0] = 0u;
    BvhNode nd;
    box_empty(nd.lb);
    box_empty(nd.rb);
    nd.left = GRAY_LEAF | 0u;
    nd.right = GRAY_EMPTY;
    nd.pad0 = nd.pad1 = 0;
    nodes[0] = nd;
    parent_int[0] = GRAY_EMPTY;
    parent_leaf[0] = 0u;
    range_lo[0] = 0;
    range_hi[0] = 0;
}

// ---------------------------------------------------------------------------
// Refit, phase 1: per chunk.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_box(float *dst, const float *b) {
    // node boxes sit at 8-byte aligned offsets (0 and 24 of a 64-byte node)
    float2 *d = reinterpret_cast<float2 *>(dst);
    d[0] = make_float2(b[0], b[1]);
    d[1] = make_float2(b[2], b[3]);
    d[2] = make_float2(b[4], b[5]);
}

__device__ __forceinline__ void shfl_box(float *out, const float *b, int delta, bool up) {
#pragma unroll
    for (int i = 0; i < 6; ++i) out[i] = up ? __shfl_up(b[i], delta, 64) : __shfl_down(b[i], delta, 64);
}

// One workgroup per chunk of GRAY_CHUNK consecutive leaves, one leaf per
// thread.  The chunk's leaf boxes, the metadata of its internal nodes (children,
// leaf range, parent; Karras node i lies in the chunk of leaf i) and the
// arrival counters are staged in LDS, so a walk up from a leaf touches global
// memory only to write the two child boxes of each node it completes (the
// second arrival) and, at the first node spanning chunks, that node's slot.
// The chunk's prefix and suffix unions are formed with wave shuffles and are
// written only at the leaf positions a spanning node reads (need_ends).
__global__ __launch_bounds__(GRAY_CHUNK) void k_refit_chunks(int n, const float *__restrict__ aabb,
                                                             const uint32_t *__restrict__ sorted_ids,
                                                             const uint4 *__restrict__ node_info,
                                                             const uint32_t *__restrict__ parent_leaf,
                                                             const uint8_t *__restrict__ need_ends,
                                                             BvhNode *__restrict__ nodes, float *__restrict__ prefix,
                                                             float *__restrict__ suffix,
                                                             float *__restrict__ chunk_box) {
    constexpr int B = GRAY_CHUNK;
    constexpr int kWaves = B / 64;
    __shared__ float leafb[B][6];
    __shared__ float intb[B][6];
    __shared__ uint4 meta[B];  // range lo, range hi, split, parent of internal node base + k
    __shared__ uint32_t cnt[B];
    __shared__ float wtot[2][kWaves][6];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int base = blockIdx.x * B;
    const int leaf = base + t;
    const bool live = leaf < n;
    float box[6];
    uint32_t pref = GRAY_EMPTY;
    uchar2 need = make_uchar2(0, 0);
    if (live) {
        const float2 *a2 = reinterpret_cast<const float2 *>(aabb + 6 * (size_t)sorted_ids[leaf]);
        const float2 v0 = a2[0], v1 = a2[1], v2 = a2[2];
        box[0] = v0.x;
        box[1] = v0.y;
        box[2] = v1.x;
        box[3] = v1.y;
        box[4] = v2.x;
        box[5] = v2.y;
        pref = parent_leaf[leaf];
        need = reinterpret_cast<const uchar2 *>(need_ends)[leaf];
    } else {
        box_empty(box);
    }
    if (leaf < n - 1) meta[t] = node_info[leaf];
    cnt[t] = 0u;
#pragma unroll
    for (int i = 0; i < 6; ++i) leafb[t][i] = box[i];

    // prefix / suffix unions within the wave, then across the waves
    float pre[6], suf[6], o[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) pre[i] = suf[i] = box[i];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        shfl_box(o, pre, d, true);
        if (lane >= d) box_union(pre, o);
        shfl_box(o, suf, d, false);
        if (lane + d < 64) box_union(suf, o);
    }
    if (lane == 63)
        for (int i = 0; i < 6; ++i) wtot[0][wave][i] = pre[i];
    if (lane == 0)
        for (int i = 0; i < 6; ++i) wtot[1][wave][i] = suf[i];
    __syncthreads();
    for (int w = 0; w < wave; ++w) box_union(pre, wtot[0][w]);
    for (int w = wave + 1; w < kWaves; ++w) box_union(suf, wtot[1][w]);
    if (need.x) store_box(suffix + 6 * (size_t)leaf, suf);
    if (need.y) store_box(prefix + 6 * (size_t)leaf, pre);
    if (t == 0) store_box(chunk_box + 6 * (size_t)blockIdx.x, suf);

    // bottom-up over the nodes whose range lies inside this chunk
    if (!live) return;
    if (n == 1) {  // the root's only child
        float e[6];
        box_empty(e);
        store_box(nodes[0].lb, box);
        store_box(nodes[0].rb, e);
        return;
    }
    float cur[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) cur[i] = box[i];
    bool child_is_leaf = true;
    int child = leaf;
    while (pref != GRAY_EMPTY) {
        const int p = (int)(pref & 0x7fffffffu);
        const bool right = (pref & 0x80000000u) != 0u;
        const bool in_chunk = p >= base && p < base + B;
        const uint4 m = in_chunk ? meta[p - base] : make_uint4(0u, 0u, 0u, 0u);
        if (!in_chunk || (int)m.x < base || (int)m.y >= base + B) {
            // p spans chunks: its slot for this child now; the rest of p is phase 3's
            store_box(right ? nodes[p].rb : nodes[p].lb, cur);
            break;
        }
        if (!child_is_leaf)
            for (int i = 0; i < 6; ++i) intb[child - base][i] = cur[i];
        __threadfence_block();
        if (atomicAdd(&cnt[p - base], 1u) == 0u) break;
        __threadfence_block();
        // second arrival: the sibling's box is in LDS.  Children of a Karras
        // node split its range at m.z: left = leaf m.z if lo == m.z, else node
        // m.z; right = leaf m.z + 1 if hi == m.z + 1, else node m.z + 1.
        const int sidx = right ? (int)m.z : (int)m.z + 1;
        const bool sleaf = right ? m.x == m.z : m.y == m.z + 1;
        const float *sb = sleaf ? leafb[sidx - base] : intb[sidx - base];
        float sbox[6];
        for (int i = 0; i < 6; ++i) sbox[i] = sb[i];
        store_box(nodes[p].lb, right ? sbox : cur);
        store_box(nodes[p].rb, right ? cur : sbox);
        box_union(cur, sbox);
        child_is_leaf = false;
        child = p;
        pref = m.w;
    }
}

// phase 2: sparse table over chunk boxes (single workgroup)
__global__ __launch_bounds__(1024) void k_refit_sparse(int chunks, int levels, float *__restrict__ table) {
    for (int lv = 1; lv < levels; ++lv) {
        const int half = 1 << (lv - 1);
        const float *prev = table + (size_t)(lv - 1) * chunks * 6;
        float *cur = table + (size_t)lv * chunks * 6;
        for (int c = threadIdx.x; c < chunks; c += blockDim.x) {
            float b[6];
            for (int i = 0; i < 6; ++i) b[i] = prev[6 * c + i];
            if (c + half < chunks) box_union(b, prev + 6 * (c + half));
            for (int i = 0; i < 6; ++i) cur[6 * c + i] = b[i];
        }
        __threadfence_block();
        __syncthreads();
    }
}

// phase 3: nodes whose leaf range spans several chunks
__global__ void k_refit_top(int count, const uint32_t *__restrict__ top_nodes, const uint32_t *__restrict__ range_lo,
                            const uint32_t *__restrict__ range_hi, const uint32_t *__restrict__ parent_int,
                            const float *__restrict__ prefix, const float *__restrict__ suffix,
                            const float *__restrict__ table, int chunks, BvhNode *__restrict__ nodes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t node = top_nodes[t];
    const uint32_t pref = parent_int[node];
    if (pref == GRAY_EMPTY) return;  // root: its own box is never read
    const int a = (int)range_lo[node], b = (int)range_hi[node];
    float box[6];
    for (int i = 0; i < 6; ++i) box[i] = suffix[6 * (size_t)a + i];
    box_union(box, prefix + 6 * (size_t)b);
    const int ca = a / GRAY_CHUNK + 1, cb = b / GRAY_CHUNK - 1;
    if (cb >= ca) {
        const int len = cb - ca + 1;
        const int lv = 31 - __clz(len);
        const float *row = table + (size_t)lv * chunks * 6;
        box_union(box, row + 6 * ca);
        box_union(box, row + 6 * (cb - (1 << lv) + 1));
    }
    const uint32_t p = pref & 0x7fffffffu;
    float *slot = (pref & 0x80000000u) ? nodes[p].rb : nodes[p].lb;
    for (int i = 0; i < 6; ++i) slot[i] = box[i];
}

// ---------------------------------------------------------------------------
// 4-wide nodes of the collect pass.  Topology part, after every rebuild: an
// internal node is a 4-wide node when its depth is even (the root's is 0), so
// a 4-wide node's internal children are the binary nodes two levels down.  The
// depth is counted up the parent chain; it is below 64, and the bound of the
// loop only keeps it finite.
// ---------------------------------------------------------------------------
__global__ void k_wide_flags(int count, const uint4 *__restrict__ node_info, uint8_t *__restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t cur = (uint32_t)i;
    int d = 0;
    for (int k = 0; k < 128 && cur != 0u && cur < (uint32_t)count; ++k) {
        cur = node_info[cur].w & 0x7fffffffu;
        ++d;
    }
    flag[i] = (uint8_t)((d & 1) == 0);
}

// Box part, after every refit: the children of 4-wide node i are, per side of
// binary node i, the two children of an internal child (slots 2s, 2s+1) or the
// leaf child itself (slot 2s, slot 2s+1 unused).  The boxes are the refitted
// binary child boxes rounded outward to half precision.
__global__ void k_widen(int count, const uint8_t *__restrict__ flag, const BvhNode *__restrict__ nodes,
                        Bvh4Node *__restrict__ wide) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || !flag[i]) return;
    const BvhNode nd = nodes[i];
    float lo[3][4], hi[3][4];
    uint32_t ch[4];
    auto put = [&](int k, const float *b, uint32_t ref) {
        for (int a = 0; a < 3; ++a) {
            lo[a][k] = b[a];
            hi[a][k] = b[3 + a];
        }
        ch[k] = ref;
    };
    float none[6];
    box_empty(none);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const uint32_t c = side ? nd.right : nd.left;
        if (c != GRAY_EMPTY && !(c & GRAY_LEAF)) {
            const BvhNode cn = nodes[c];
            put(2 * side, cn.lb, cn.left);
            put(2 * side + 1, cn.rb, cn.right);
        } else {
            put(2 * side, side ? nd.rb : nd.lb, c);
            put(2 * side + 1, none, GRAY_EMPTY);
        }
    }
    // round toward -inf / +inf: a box never shrinks (+-inf stay +-inf, so an
    // empty box stays empty)
    auto down = [](float x, float y) {
        return (uint32_t)__half_as_ushort(__float2half_rd(x)) | ((uint32_t)__half_as_ushort(__float2half_rd(y)) << 16);
    };
    auto up = [](float x, float y) {
        return (uint32_t)__half_as_ushort(__float2half_ru(x)) | ((uint32_t)__half_as_ushort(__float2half_ru(y)) << 16);
    };
    uint4 *d = reinterpret_cast<uint4 *>(wide + i);
    d[0] = make_uint4(down(lo[0][0], lo[0][1]), down(lo[0][2], lo[0][3]), down(lo[1][0], lo[1][1]),
                      down(lo[1][2], lo[1][3]));
    d[1] = make_uint4(down(lo[2][0], lo[2][1]), down(lo[2][2], lo[2][3]), up(hi[0][0], hi[0][1]), up(hi[0][2], hi[0][3]));
    d[2] = make_uint4(up(hi[1][0], hi[1][1]), up(hi[1][2], hi[1][3]), up(hi[2][0], hi[2][1]), up(hi[2][2], hi[2][3]));
    d[3] = make_uint4(ch[0], ch[1], ch[2], ch[3]);
}

}  // namespace

// Block sums gray_exclusive_scan_u32 needs for m values (elements).
int64_t gray_scan_scratch(int64_t m) { return gray_div_up(std::max<int64_t>(m, 1), kScanTile) + 1; }

// Scratch sizes of gray_radix_sort_pairs for n items (elements).
void gray_radix_scratch(int64_t n, int64_t *hist_elems, int64_t *sums_elems) {
    const int64_t tiles = gray_div_up(std::max<int64_t>(n, 1), kSortTile);
    *hist_elems = 256 * tiles;
    *sums_elems = gray_scan_scratch(*hist_elems);
}

// Exclusive prefix sum of m uint32 values in place (block_sums: as many as
// gray_radix_scratch gives for a histogram of m elements).
int gray_exclusive_scan_u32(uint32_t *data, int m, uint32_t *block_sums, hipStream_t s) {
    if (m <= 0) return GRAY_OK;
    const int sblocks = (int)gray_div_up(m, kScanTile);
    hipLaunchKernelGGL(k_scan_block, dim3(sblocks), dim3(kScanThreads), 0, s, data, m, block_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, block_sums, sblocks);
    hipLaunchKernelGGL(k_scan_add, dim3(sblocks), dim3(kScanThreads), 0, s, data, m, block_sums);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

// Stable LSD radix sort of n (key, value) pairs by the low 8*passes bits of
// the key, ping-ponging between (keys, vals) and (keys_alt, vals_alt); the
// sorted pairs end in *keys_res / *vals_res.  Equal keys keep their input
// order, so the result does not depend on scheduling.
int gray_radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, int n, int passes,
                          uint32_t *hist, uint32_t *block_sums, hipStream_t s, uint32_t **keys_res,
                          uint32_t **vals_res) {
    uint32_t *kin = keys, *vin = vals, *kout = keys_alt, *vout = vals_alt;
    if (n > 0) {
        const int tiles = (int)gray_div_up(n, kSortTile);
        const int m = 256 * tiles;
        for (int pass = 0; pass < passes; ++pass) {
            const int shift = 8 * pass;
            hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, s, n, kin, shift, hist, tiles);
            GRAY_HIP(hipGetLastError());
            if (int rc = gray_exclusive_scan_u32(hist, m, block_sums, s)) return rc;
            hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, n, kin, vin, kout, vout, shift,
                               hist, tiles);
            GRAY_HIP(hipGetLastError());
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
    }
    *keys_res = kin;
    *vals_res = vin;
    return GRAY_OK;
}

namespace {

template <typename T>
int grow(T **ptr, int64_t count) {
    if (*ptr) {
        hipError_t e = hipFree(*ptr);
        if (e != hipSuccess) return gray_check_hip(e, "hipFree");
    }
    *ptr = nullptr;
    if (count <= 0) return GRAY_OK;
    hipError_t e = hipMalloc((void **)ptr, sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
        gray_set_error(GRAY_ERR_NOMEM, "hipMalloc of %lld bytes failed: %s", (long long)(sizeof(T) * count),
                       hipGetErrorString(e));
        return GRAY_ERR_NOMEM;
    }
    return GRAY_OK;
}

int chunks_of(int64_t n) { return (int)((n + GRAY_CHUNK - 1) / GRAY_CHUNK); }
int levels_of(int chunks) {
    int lv = 1;
    while ((1 << (lv - 1)) < chunks) ++lv;
    return lv;
}

}  // namespace

int gray_ctx_reserve(gray_ctx *ctx, int64_t n) {
    if (n <= ctx->cap) return GRAY_OK;
    int64_t cap = n;
    int rc;
#define G(p, c) \
    if ((rc = grow(&ctx->p, (c))) != GRAY_OK) return rc;
    G(geo, cap);
    G(l2w, cap * 12);
    G(aabb, cap * 6);
    G(nodes, cap);
    G(parent_int, cap);
    G(parent_leaf, cap);
    G(range_lo, cap);
    G(range_hi, cap);
    G(sorted_ids, cap);
    G(keys, cap);
    G(keys_alt, cap);
    G(vals_alt, cap);
    G(prefix, cap * 6);
    G(suffix, cap * 6);
    G(top_nodes, cap);
    G(need_ends, cap * 2);
    G(node_info, cap);
    G(wide, cap);
    G(wide_flag, cap);
    const int64_t tiles = gray_div_up(cap, kSortTile);
    const int64_t hm = 256 * tiles;
    G(hist, hm);
    G(block_sums, gray_div_up(hm, kScanTile) + 1);
    const int chunks = chunks_of(cap);
    const int levels = levels_of(chunks);
    G(sparse, (int64_t)levels * chunks * 6);
    G(bounds_partial, 6 * 1025);
#undef G
    ctx->cap = cap;
    ctx->built = false;
    return GRAY_OK;
}

int gray_refit(gray_ctx *ctx, hipStream_t s) {
    const int n = (int)ctx->n;
    if (n <= 0) return GRAY_OK;
    const int chunks = chunks_of(n);
    hipLaunchKernelGGL(k_refit_chunks, dim3(chunks), dim3(GRAY_CHUNK), 0, s, n, ctx->aabb, ctx->sorted_ids,
                       ctx->node_info, ctx->parent_leaf, ctx->need_ends, ctx->nodes, ctx->prefix, ctx->suffix,
                       ctx->sparse);
    GRAY_HIP(hipGetLastError());
    if (ctx->num_top > 0) {
        const int levels = levels_of(chunks);
        hipLaunchKernelGGL(k_refit_sparse, dim3(1), dim3(1024), 0, s, chunks, levels, ctx->sparse);
        GRAY_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_refit_top, dim3(gray_div_up(ctx->num_top, 256)), dim3(256), 0, s, (int)ctx->num_top,
                           ctx->top_nodes, ctx->range_lo, ctx->range_hi, ctx->parent_int, ctx->prefix, ctx->suffix,
                           ctx->sparse, chunks, ctx->nodes);
        GRAY_HIP(hipGetLastError());
    }
    const int internal = n > 1 ? n - 1 : 1;
    hipLaunchKernelGGL(k_widen, dim3(gray_div_up(internal, 256)), dim3(256), 0, s, internal, ctx->wide_flag,
                       ctx->nodes, ctx->wide);
    GRAY_HIP(hipGetLastError());
    return GRAY_OK;
}

static int build_topology(gray_ctx *ctx, const float *mean, hipStream_t s) {
    const int n = (int)ctx->n;
    GRAY_HIP(hipMemsetAsync(ctx->need_ends, 0, 2 * (size_t)n, s));
    if (n == 1) {
        hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(1), 0, s, ctx->nodes, ctx->parent_int, ctx->parent_leaf,
                           ctx->range_lo, ctx->range_hi, ctx->sorted_ids, ctx->keys);
        hipLaunchKernelGGL(k_wide_flags, dim3(1), dim3(64), 0, s, 1, ctx->node_info, ctx->wide_flag);
        GRAY_HIP(hipGetLastError());
        ctx->num_top = 0;
        return GRAY_OK;
    }
    const int bblocks = (int)std::min<int64_t>(1024, gray_div_up(n, 256));
    hipLaunchKernelGGL(k_bounds_partial, dim3(bblocks), dim3(256), 0, s, n, mean, ctx->bounds_partial);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, s, bblocks, ctx->bounds_partial,
                       ctx->bounds_partial + 6 * 1024);
    hipLaunchKernelGGL(k_morton, dim3(gray_div_up(n, 256)), dim3(256), 0, s, n, mean,
                       ctx->bounds_partial + 6 * 1024, ctx->keys, ctx->sorted_ids);
    GRAY_HIP(hipGetLastError());
    // radix sort (keys, sorted_ids) in 4 passes, ping-pong with *_alt; after
    // an even number of passes the result is back in keys / sorted_ids
    uint32_t *kres = nullptr, *vres = nullptr;
    if (int rc = gray_radix_sort_pairs(ctx->keys, ctx->sorted_ids, ctx->keys_alt, ctx->vals_alt, (int)n, 4, ctx->hist,
                                       ctx->block_sums, s, &kres, &vres))
        return rc;
    GRAY_HIP(hipMemsetAsync(ctx->top_count, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_karras, dim3(gray_div_up(n - 1, 256)), dim3(256), 0, s, n, ctx->keys, ctx->sorted_ids,
                       ctx->nodes, ctx->parent_int, ctx->parent_leaf, ctx->range_lo, ctx->range_hi, ctx->top_nodes,
                       ctx->top_count, ctx->need_ends, ctx->node_info);
    hipLaunchKernelGGL(k_wide_flags, dim3(gray_div_up(n - 1, 256)), dim3(256), 0, s, n - 1, ctx->node_info,
                       ctx->wide_flag);
    GRAY_HIP(hipGetLastError());
    uint32_t top = 0;
    GRAY_HIP(hipMemcpyAsync(&top, ctx->top_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GRAY_HIP(hipStreamSynchronize(s));
    ctx->num_top = top;
    return GRAY_OK;
}

extern "C" int gray_bvh_update(gray_ctx *ctx, const gray_params *params, const gray_knobs *knobs, int rebuild,
                               void *stream) {
    if (!ctx || !params || !knobs) {
        gray_set_error(GRAY_ERR_INVALID, "gray_bvh_update: null argument");
        return GRAY_ERR_INVALID;
    }
    const int64_t n = params->count;
    if (n < 0 || n > 0x7fffffff / 16) {
        gray_set_error(GRAY_ERR_INVALID, "gray_bvh_update: bad Gaussian count %lld", (long long)n);
        return GRAY_ERR_INVALID;
    }
    if (n > 0 && (!params->mean || !params->rotation || !params->scale || !params->opacity)) {
        gray_set_error(GRAY_ERR_INVALID, "gray_bvh_update: null parameter pointer");
        return GRAY_ERR_INVALID;
    }
    if (((uintptr_t)params->rotation & 15u) != 0u) {
        gray_set_error(GRAY_ERR_INVALID, "gray_bvh_update: rotation must be 16-byte aligned (one float4 per Gaussian)");
        return GRAY_ERR_INVALID;
    }
    if (ctx->scene) {
        gray_set_error(GRAY_ERR_STATE, "gray_bvh_update: the context renders a shared scene; update the scene's context");
        return GRAY_ERR_STATE;
    }
    hipStream_t s = (hipStream_t)stream;
    if (!rebuild && (!ctx->built || ctx->n != n)) {
        gray_set_error(GRAY_ERR_STATE, "gray_bvh_update: refit needs a prior rebuild with the same count");
        return GRAY_ERR_STATE;
    }
    int rc = gray_ctx_reserve(ctx, std::max<int64_t>(n, 1));
    if (rc != GRAY_OK) return rc;
    ctx->n = n;
    if (n == 0) {
        ctx->built = true;
        ctx->num_top = 0;
        return GRAY_OK;
    }
    const bool timed = rebuild && ctx->timing;
    if (timed) GRAY_HIP(hipEventRecord(ctx->ev0, s));
    DevKnobs kn{knobs->exp_power, knobs->alpha_threshold, knobs->t_threshold, knobs->global_scale_factor,
                knobs->eps_forward_normalization, knobs->eps_scale_grad, knobs->background_channels,
                knobs->render_depth, knobs->enable_sh, knobs->jitter_primary_rays, knobs->call_index,
                knobs->needs_ray_output};
    hipLaunchKernelGGL(k_snapshot, dim3(gray_div_up(n, kSnapThreads)), dim3(kSnapThreads), 0, s, (int)n, params->mean,
                       params->rotation, params->scale, params->opacity, kn, ctx->geo, ctx->l2w, ctx->aabb);
    GRAY_HIP(hipGetLastError());
    if (rebuild) {
        ctx->built = false;
        if ((rc = build_topology(ctx, params->mean, s)) != GRAY_OK) return rc;
        ctx->built = true;
    }
    if ((rc = gray_refit(ctx, s)) != GRAY_OK) return rc;
    if (timed) {
        GRAY_HIP(hipEventRecord(ctx->ev1, s));
        GRAY_HIP(hipEventSynchronize(ctx->ev1));
        float ms = -1.0f;
        GRAY_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_build_ms = ms;
    }
    return GRAY_OK;
}

extern "C" int gray_bvh_set_timing(gray_ctx *ctx, int enable) {
    if (!ctx) return GRAY_ERR_INVALID;
    ctx->timing = enable != 0;
    return GRAY_OK;
}

extern "C" float gray_bvh_last_build_ms(gray_ctx *ctx) { return ctx ? ctx->last_build_ms : -1.0f; }

extern "C" int gray_bvh_info(gray_ctx *ctx_in, int64_t *num_gaussians, int64_t *num_nodes, int64_t *depth_bound) {
    if (!ctx_in) return GRAY_ERR_INVALID;
    const gray_ctx *ctx = scene_of(ctx_in);
    if (num_gaussians) *num_gaussians = ctx->n;
    if (num_nodes) *num_nodes = ctx->n > 1 ? ctx->n - 1 : (ctx->n == 1 ? 1 : 0);
    if (depth_bound) *depth_bound = GRAY_STACK;
    return GRAY_OK;
}

extern "C" int gray_debug_bvh(gray_ctx *ctx_in, float *nodes, uint32_t *sorted_ids, uint32_t *keys, void *stream) {
    if (!ctx_in) return GRAY_ERR_INVALID;
    const gray_ctx *ctx = scene_of(ctx_in);
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = ctx->n;
    if (n == 0) return GRAY_OK;
    const int64_t nn = n > 1 ? n - 1 : 1;
    if (nodes) GRAY_HIP(hipMemcpyAsync(nodes, ctx->nodes, sizeof(BvhNode) * nn, hipMemcpyDeviceToDevice, s));
    if (sorted_ids) GRAY_HIP(hipMemcpyAsync(sorted_ids, ctx->sorted_ids, 4 * n, hipMemcpyDeviceToDevice, s));
    if (keys) GRAY_HIP(hipMemcpyAsync(keys, ctx->keys, 4 * n, hipMemcpyDeviceToDevice, s));
    return GRAY_OK;
}

extern "C" int gray_debug_wide(gray_ctx *ctx_in, uint32_t *nodes, uint8_t *flags, void *stream) {
    if (!ctx_in) return GRAY_ERR_INVALID;
    const gray_ctx *ctx = scene_of(ctx_in);
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = ctx->n;
    if (n == 0) return GRAY_OK;
    const int64_t nn = n > 1 ? n - 1 : 1;
    if (nodes) GRAY_HIP(hipMemcpyAsync(nodes, ctx->wide, sizeof(Bvh4Node) * nn, hipMemcpyDeviceToDevice, s));
    if (flags) GRAY_HIP(hipMemcpyAsync(flags, ctx->wide_flag, nn, hipMemcpyDeviceToDevice, s));
    return GRAY_OK;
}

extern "C" int gray_debug_snapshot(gray_ctx *ctx_in, float *geo, float *aabb, void *stream) {
    if (!ctx_in) return GRAY_ERR_INVALID;
    const gray_ctx *ctx = scene_of(ctx_in);
    hipStream_t s = (hipStream_t)stream;
    if (ctx->n == 0) return GRAY_OK;
    if (geo) GRAY_HIP(hipMemcpyAsync(geo, ctx->geo, sizeof(GeoRecord) * ctx->n, hipMemcpyDeviceToDevice, s));
    if (aabb) GRAY_HIP(hipMemcpyAsync(aabb, ctx->aabb, sizeof(float) * 6 * ctx->n, hipMemcpyDeviceToDevice, s));
    return GRAY_OK;
}
