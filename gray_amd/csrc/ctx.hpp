// Library context: owns the geometry snapshot, the LBVH and its build scratch.
#pragma once

#include "common.hpp"

#define GRAY_CHUNK 1024  // leaves per refit chunk (phase-1 workgroup)

struct gray_ctx {
    int device = 0;
    // when set (gray_ctx_share_scene), renders read this context's geometry
    // snapshot and BVH; the fields below then keep only the render scratch
    gray_ctx *scene = nullptr;
    int64_t n = 0;         // Gaussians in the current snapshot
    int64_t cap = 0;       // allocated capacity (Gaussians)
    bool built = false;    // a topology exists for n
    // geometry snapshot (original Gaussian order)
    GeoRecord *geo = nullptr;
    float *l2w = nullptr;   // [cap,12]
    float *aabb = nullptr;  // [cap,6]
    // LBVH (Karras layout: internal nodes 0..n-2, root 0)
    BvhNode *nodes = nullptr;       // [max(cap-1,1)]
    Bvh4Node *wide = nullptr;       // [cap]: 4-wide nodes at the binary index of every even-depth internal node
    uint8_t *wide_flag = nullptr;   // [cap]: 1 where an internal node lies at even depth (root: depth 0)
    uint32_t *parent_int = nullptr; // [cap] parent | side<<31 (root: GRAY_EMPTY)
    uint32_t *parent_leaf = nullptr;// [cap]
    uint32_t *range_lo = nullptr;   // [cap]
    uint32_t *range_hi = nullptr;   // [cap]
    uint4 *node_info = nullptr;     // [cap]: leaf range lo, hi, split, parent (bit 31: right child)
    uint32_t *sorted_ids = nullptr; // [cap] leaf k -> Gaussian id
    uint32_t *keys = nullptr;       // [cap] sorted Morton codes
    // radix sort scratch
    uint32_t *keys_alt = nullptr, *vals_alt = nullptr;
    uint32_t *hist = nullptr;       // [256 * tiles]
    uint32_t *block_sums = nullptr;
    int64_t hist_cap = 0;
    float *bounds_partial = nullptr;// [blocks*6]
    float *bounds = nullptr;        // [6]
    // refit scratch
    float *prefix = nullptr, *suffix = nullptr; // [cap,6], written only where need_ends is set
    uint8_t *need_ends = nullptr;   // [cap][2]: suffix / prefix box read by a node spanning chunks
    float *sparse = nullptr;        // [levels, chunks, 6]
    int64_t sparse_cap = 0;
    uint32_t *top_nodes = nullptr;  // [cap]
    uint32_t *top_count = nullptr;  // [1] device
    int64_t num_top = 0;
    int32_t *err = nullptr;         // [1] device error flag
    // hit lists of the forward collect pass (one launch of <= 8 views)
    float4 *list_rec = nullptr;     // [list_entries] {dist, id bits, alpha, 0}
    int32_t *list_n = nullptr;
    uint8_t *tile_open = nullptr;   // [list_lanes_cap / 64]: tile handed work on (k_collect)
    int64_t list_entries = 0;       // allocated records
    int64_t list_lanes_cap = 0;     // allocated entries of list_n
    int64_t list_lanes = 0;         // lanes of the last collecting launch
    // continuation tasks of the collect pass: two ping-pong buffers + counters
    uint2 *tasks[2] = {nullptr, nullptr};
    int64_t task_cap = 0;
    int32_t *task_counters = nullptr;  // [2 * kMaxTaskRounds + 2]
    // tile bins of the collect pass (the bins themselves live in tasks[1]):
    // per tile its pair count, fill cursor and first work item; the work items
    int32_t *bin_tiles = nullptr;      // [3 * list_lanes_cap / 64 + 2]
    int2 *bin_items = nullptr;         // [bin_item_cap]
    int64_t bin_item_cap = 0;
    // deterministic gradient reduction: per-hit records, their sort keys and scratch
    float *det_rec = nullptr;        // [det_cap][16]: 15 gradient terms, Gaussian id bits
    uint32_t *det_keys[2] = {nullptr, nullptr}, *det_vals[2] = {nullptr, nullptr};
    int64_t det_cap = 0;
    uint32_t *det_off = nullptr;     // [det_px_cap] per-pixel record offsets
    int64_t det_px_cap = 0;
    uint32_t *det_hist = nullptr, *det_sums = nullptr;
    int64_t det_hist_cap = 0, det_sums_cap = 0;
    // stored-list backward: raw geometric gradient sums per Gaussian (kept
    // zero between backward calls) and their touched marks
    float *raw = nullptr;            // [raw_cap][16]
    uint8_t *raw_touched = nullptr;  // [raw_cap]
    int64_t raw_cap = 0;
    uint8_t *bwd_overflow = nullptr; // [bwd_overflow_cap] per launch tile, zero between backward calls
    int64_t bwd_overflow_cap = 0;
    // stored-list backward: tile costs, the tiles in descending cost, the fetch counter
    uint8_t *tile_cost = nullptr;    // [bwd_overflow_cap]
    uint32_t *tile_order = nullptr;  // [bwd_overflow_cap]
    int32_t *tile_next = nullptr;    // [1]
    int stored_grid = 0;             // resident workgroups of the stored-list backward (0 = not queried)
    float *bwd_rec = nullptr;        // [bwd_overflow_cap * 128][16] per-tile per-Gaussian sums
    uint32_t *bwd_cnt = nullptr;     // [raw_cap] records per Gaussian, zero between backward calls
    uint32_t *bwd_lst = nullptr;     // [raw_cap][4] record indices per Gaussian
    // timing
    bool timing = false;
    float last_build_ms = -1.0f;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // second stream of the forward (default priority): the wave-per-ray
    // compositing runs on it beside the per-lane compositing, forked from and
    // joined back into the caller's stream with these events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_sh = nullptr;  // the colours computed on the side stream (render args sh_params)
    // forward stage events (recorded when timing is on): collect, task rounds,
    // list compositing, re-traversal
    hipEvent_t fev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool fev_recorded = false;
};

// The context whose snapshot and BVH a render of ctx reads.
inline const gray_ctx *scene_of(const gray_ctx *ctx) { return ctx->scene ? ctx->scene : ctx; }

int gray_ctx_reserve(gray_ctx *ctx, int64_t n);
// sorting and scanning primitives of the BVH build, shared with the
// deterministic gradient reduction
int64_t gray_scan_scratch(int64_t m);
void gray_radix_scratch(int64_t n, int64_t *hist_elems, int64_t *sums_elems);
int gray_exclusive_scan_u32(uint32_t *data, int m, uint32_t *block_sums, hipStream_t s);
int gray_radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, int n, int passes,
                          uint32_t *hist, uint32_t *block_sums, hipStream_t s, uint32_t **keys_res,
                          uint32_t **vals_res);
int gray_refit(gray_ctx *ctx, hipStream_t s);
