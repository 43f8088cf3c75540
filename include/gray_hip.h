/*
 * gray_hip.h -- C ABI of the MI355X (gfx950) differentiable Gaussian ray tracer.
 *
 * Every entry point takes plain device pointers and sizes, a caller-owned HIP
 * stream (passed as void*, NULL = default stream) and returns an int status
 * (GRAY_OK or a GRAY_ERR_* code; gray_last_error() gives the message of the
 * last failure on the calling thread).  Nothing traps on the device: capacity
 * problems come back as GRAY_ERR_OVERFLOW.  All Gaussian tensors use the
 * reference layouts (cuda/core/gaussians.h:7-65):
 *   mean [N,3], rotation [N,4] (r,x,y,z, unnormalised), scale [N,3] (log),
 *   opacity [N,1] (logit), channels [N,3], sh_coeffs_dc [N,1,3],
 *   sh_coeffs_rest [N,15,3] (coefficient-major, channel-minor), fp32.
 * The library owns only the BVH and its build scratch (inside gray_ctx); every
 * other buffer is allocated by the caller.  One host thread per gray_ctx.
 *
 * Each function below names the reference interface it replaces.
 */
#ifndef GRAY_HIP_H
#define GRAY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRAY_ABI_VERSION 14

#define GRAY_OK 0
#define GRAY_ERR_INVALID 1  /* bad argument (null pointer, size, shape) */
#define GRAY_ERR_HIP 2      /* a HIP runtime call failed */
#define GRAY_ERR_OVERFLOW 3 /* a device-side capacity was exceeded (nothing trapped) */
#define GRAY_ERR_NOMEM 4    /* device allocation failed */
#define GRAY_ERR_STATE 5    /* call order violated (e.g. render before any BVH build) */

#define GRAY_MAX_ALPHA 0.9999f /* reference CMakeLists.txt:68-71 (MAX_ALPHA) */
#define GRAY_CHANNELS 3        /* reference CMakeLists.txt:68-71 (CHANNELS) */
#define GRAY_BATCH 32          /* reference forward_pass.cu:4 (BUFFER_SIZE) */
#define GRAY_MAX_ROUNDS 99     /* reference forward_pass.cu:3 (MAX_ITERATIONS) */

typedef struct gray_ctx gray_ctx;

/* Runtime knobs, device pointers to 1-element tensors exactly like the
 * reference ConfigDataHolder (cuda/core/config.h:5-44).  Kernels read them on
 * the device, so changing a knob with fill_() needs no host round trip. */
typedef struct {
    const float *exp_power;                 /* [1] */
    const float *alpha_threshold;           /* [1] */
    const float *t_threshold;               /* [1] */
    const float *global_scale_factor;       /* [1] */
    const float *eps_forward_normalization; /* [1] */
    const float *eps_scale_grad;            /* [1] */
    const float *background_channels;       /* [3] */
    const uint8_t *render_depth;            /* [1] bool */
    const uint8_t *enable_sh;               /* [1] bool */
    const uint8_t *jitter_primary_rays;     /* [1] bool or NULL: sub-pixel jitter of camera rays (camera.h:17-36) */
    const int32_t *call_index;              /* [1] or NULL (0): jitter seed, the reference's metadata
                                               total_num_calls (shaders.cu:76) */
    const uint8_t *needs_ray_output;        /* [1] bool or NULL: without saved state, still write depth and
                                               the ray direction (framebuffer.h:46-51) */
} gray_knobs;

/* Gaussian parameters (read by every stage) */
typedef struct {
    int32_t count;            /* N */
    int32_t num_sh_coeffs;    /* (max_degree+1)^2-1, 15 for degree 3 */
    int32_t sh_is_fp16;       /* inference_only layout (gaussians.h:132) */
    const int32_t *current_sh_degree; /* [1] device */
    const float *mean;        /* [N,3] */
    const float *rotation;    /* [N,4], 16-byte aligned */
    const float *scale;       /* [N,3] log-scale */
    const float *opacity;     /* [N,1] logit */
    const void *sh_dc;        /* [N,1,3] fp32 (or fp16 when sh_is_fp16) */
    const void *sh_rest;      /* [N,K,3] */
} gray_params;

/* One camera view of a (possibly multi-view) launch.  Camera values are
 * device pointers so a caller can update them with fill_/copy_ in place
 * (reference camera.h:8-38).  Per-pixel buffers are [alloc_height, stride, C]
 * with pixel_id = y*stride + x (reference shaders.cu:73). */
typedef struct {
    const float *origin;        /* [3] camera origin */
    const float *rotation_c2w;  /* [3,3] row-major camera-to-world (Blender axes) */
    const float *vertical_fov;  /* [1] radians */
    const float *znear;         /* [1] */
    const float *zfar;          /* [1] */
    const float *rays_origin;   /* [H,stride,3] explicit rays, or NULL for pinhole */
    const float *rays_direction;/* [H,stride,3] explicit rays, or NULL for pinhole */
    const uint8_t *rays_enable; /* [1] device flag or NULL; when given, the explicit rays are used only
                                   while it is nonzero (reference config rays_from_python) */
    int32_t width, height;      /* active render resolution */
    int32_t stride;             /* allocated framebuffer width (pixels) */
    int32_t alloc_height;       /* allocated framebuffer height */
    /* per-view Gaussian colours [N,3] (SH evaluated for this view) */
    const float *rgb;
    /* forward outputs (reference Framebuffer, framebuffer.h:22-53) */
    float *out_channels;        /* [H,stride,3] always written */
    float *out_depth;           /* [H,stride,1] or NULL */
    float *out_transmittance;   /* [H,stride,1] or NULL */
    float *out_full_transmittance; /* [H,stride,1] or NULL */
    float *out_remaining_est;   /* [H,stride,3] or NULL */
    float *out_ray_origin;      /* [H,stride,3] or NULL */
    float *out_ray_direction;   /* [H,stride,3] or NULL */
    int32_t *stat_hits;         /* [H,stride] num_gaussians_hit (+=) or NULL */
    int32_t *stat_accumulated;  /* [H,stride] num_gaussians_accumulated (+=) or NULL */
    /* backward state written by the forward when save != 0 (caller owned):
     * hit_ids/hit_trans are [hit_capacity, alloc_height*stride] (slot-major),
     * n_comp [alloc_height*stride], comp_rgb [alloc_height*stride, 3] */
    uint32_t *hit_ids;
    float *hit_trans;
    uint32_t *n_comp;
    float *comp_rgb;
    /* backward inputs / outputs */
    const float *grad_out;      /* [H,stride,3] dL/d(out_channels) */
    float *grad_rgb;            /* [N,3] dL/d(rgb) for this view (+=) */
} gray_view;

typedef struct {
    gray_knobs knobs;
    int32_t num_views;
    const gray_view *views;     /* host array of num_views entries */
    int32_t save;               /* forward: write backward state + ray outputs */
    int32_t hit_capacity;       /* L: hits stored per ray (0 = recompute all in backward) */
    int32_t use_bvh;            /* 1 = BVH traversal, 0 = test every Gaussian (brute force) */
    int32_t *tile_stats;        /* optional [tiles, 4] traversal statistics per 8x8 tile of the forward:
                                   node visits, leaf tests, rounds, hits (summed over the tile's rays);
                                   tiles are numbered view by view, row-major.  NULL = off */
    int32_t list_capacity;      /* hits collected per ray by the forward's traversal pass before
                                   compositing (0 = composite during traversal, re-traversing per batch).
                                   Rays with more hits fall back to re-traversal.  The context keeps
                                   list_capacity * 16 bytes per ray of one launch (<= 8 views),
                                   plus 256 bytes per ray of continuation tasks. */
    int32_t collect_budget;     /* iterations a ray's walk runs before its remaining subtrees become tasks
                                   that other lanes take (load balance of rays that cross many boxes);
                                   <= 0 = the built-in default */
    int32_t fp16_hits;          /* 1 = the reference's hit storage numerics (per_pixel_linked_list.h:28-29):
                                   hit distances and the alphas used for compositing, depth and the
                                   truncation estimate rounded to fp16 (round to nearest even); the full
                                   transmittance keeps the fp32 alpha (shaders.cu).  0 = fp32 throughout */
    int32_t wave_list_min;      /* collected lists longer than this (and at most 512 hits) are composited by
                                   a whole wave each instead of one lane: same composited hits, transmittance
                                   and colour; the skip-set sums are added in another order.  0 = the
                                   built-in default (96), < 0 = never */
    int32_t deterministic;      /* 1 = reproducible gradients: the forward walks each ray's tree in one lane
                                   (no continuation tasks), so hit lists and truncation sums keep one order,
                                   and the backward sums each Gaussian's terms in pixel and hit order
                                   (records sorted by Gaussian with a stable radix sort) instead of with
                                   atomics.  Two runs on the same inputs give bit-identical images and
                                   gradients.  The backward then synchronizes the stream once per view and
                                   keeps 80 bytes per composited hit.  No counterpart in the reference
                                   (its gradients are summed with atomics); 0 = off */
    const gray_params *sh_params; /* optional: the launch computes each view's colours from these parameters'
                                   SH coefficients itself (as gray_sh_forward into views[].rgb from
                                   views[].origin, with knobs), on the context's second stream while the
                                   collect pass runs, and waits for them before compositing; the reference's
                                   forward_pass evaluates SH inside the pass as well (raytracer.h:95-97).
                                   NULL = views[].rgb already holds the colours */
    int32_t collect_mode;       /* how the collect pass (list_capacity > 0) finds each ray's hits:
                                   0 = tile bins: every Gaussian's box is projected into each pinhole view,
                                   the (8x8 tile, Gaussian) pairs are counting-sorted by tile, and each
                                   tile's rays test its bin (records staged in LDS, the same exact hit test
                                   as the walk, so the same hits); a launch with explicit rays
                                   (rays_direction enabled) or more pairs than the task buffers hold falls
                                   back to the walk on the device.  1 = the BVH walk (per-lane 4-wide walk
                                   plus pooled task rounds).  The deterministic mode always walks.  No
                                   counterpart in the reference (OptiX traverses its BVH) */
} gray_render_args;

/* Gradient targets of the backward (reference gaussians.h:34-62). */
typedef struct {
    float *grad_mean;           /* [N,3] += */
    float *grad_rotation;       /* [N,4] += */
    float *grad_scale;          /* [N,3] += */
    float *grad_opacity;        /* [N,1] += */
    float *pruning_weight;      /* [N,1] += (or NULL) */
    uint8_t *was_visible;       /* [N,1] set to 1 (or NULL) */
} gray_grads;

/* Adam state (reference cuda/opt/adam.cu:6-84, gaussians.h:21-56). */
typedef struct {
    float *param[7];   /* mean, rotation, scale, opacity, channels, sh_dc, sh_rest */
    float *grad[7];
    float *m1[7];
    float *m2[7];
    const float *lr[7];      /* device [1] (first element read, adam.cu:33) */
    int64_t numel[7];        /* floats per tensor */
    const float *beta1;      /* device [1] */
    const float *beta2;      /* device [1] */
    const float *epsilon;    /* device [1] */
    /* pruning statistics fused into the step (gray/raytracer.py:361-366) */
    const uint8_t *was_visible; /* [N] or NULL */
    int32_t *pruning_counter;   /* [N] or NULL */
    int32_t count;              /* N */
    /* optional device-resident switches (reference ConfigDataHolder and
     * GaussianDataHolder fields); when set they override the host flags, so
     * a step needs no device->host read */
    const uint8_t *update_channels;    /* [1] bool or NULL */
    const uint8_t *zero_grads;         /* [1] bool or NULL */
    const uint8_t *enable_sh;          /* [1] bool or NULL */
    const int32_t *sh_update_laziness; /* [1] or NULL */
    /* device [1] or NULL: the step count after its increment (the reference's
     * meta_data.total_num_steps, raytracer.h:127-128); when set it replaces
     * the host `step` argument of gray_adam_step for the bias correction and
     * the SH laziness test, so a caller that restores or edits the counter
     * steers the optimiser without a device->host read */
    const int32_t *step_device;
} gray_adam;

/* Training loss over P strided fp32 planes of H x W pixels (render and
 * target), e.g. a [C,H,W] image (strides {H*W, W, 1}) or the [H,W,C]
 * framebuffer (strides {1, W_alloc*C, C}). */
typedef struct {
    int32_t planes, height, width;
    int32_t valid_padding;     /* 0: SSIM over every pixel, zero padding ("same");
                                  1: only pixels >= 5 from the border ("valid") */
    int32_t train;             /* 1: the forward also writes the SSIM derivative maps the backward reads */
    int32_t pad;
    const float *render;  int64_t render_stride[3];  /* plane, row, column (elements) */
    const float *target;  int64_t target_stride[3];
    float *grad;          int64_t grad_stride[3];    /* backward output, may alias nothing above */
    float lambda_ssim;         /* loss = (1 - lambda) mean|render - target| + lambda (1 - mean SSIM) */
    float grad_scale;          /* grad = grad_scale * (*grad_scale_device) * dloss/drender */
    const float *grad_scale_device; /* device [1] or NULL (= 1) */
    void *workspace;           /* device, gray_loss_workspace_bytes(planes, height, width) bytes */
    float *loss;               /* device [3]: loss, mean |render - target|, mean SSIM */
} gray_loss_args;

/* ---- context -------------------------------------------------------------- */

/* Replaces the torch custom-class ctor gray.Raytracer (cuda/raytracer.h:49-86).
 * Creates the BVH holder on the current HIP device. */
int gray_ctx_create(gray_ctx **out);
void gray_ctx_destroy(gray_ctx *ctx);

/* No reference counterpart (serving: one resident scene, several launches in
 * flight).  Makes renders of `ctx` (forward, backward, ellipsoids, debug
 * copies) read the geometry snapshot and BVH of `scene` instead of its own;
 * ctx keeps only its render scratch (hit lists, task buffers, error flag,
 * stage events), so any number of contexts render one scene held once in
 * HBM, each on its own stream.  gray_bvh_update on ctx is then refused: the
 * scene's context updates the BVH, and the caller orders that update against
 * the renders in flight.  `scene` must own its BVH, live on the same device
 * and outlive the sharing; NULL detaches. */
int gray_ctx_share_scene(gray_ctx *ctx, gray_ctx *scene);
const char *gray_last_error(void);
int gray_abi_version(void);

/* No reference counterpart (the reference traps on the device,
 * per_pixel_linked_list.h:23-26).  Synchronises the stream, then reports and
 * clears the device-side error flag (e.g. GRAY_ERR_OVERFLOW for a traversal
 * stack overflow) of every launch since the previous call. */
int gray_device_status(gray_ctx *ctx, void *stream);

/* No reference counterpart: the per-frame holder updates that precede a
 * forward pass, in one launch instead of about ten small ones.  They are the
 * camera fills and Camera::set_pose of the Python layer
 * (gray/raytracer.py:126-132) and the metadata update and was_visible reset
 * of Raytracer::forward_pass (cuda/raytracer.h:88-93).  Each device pointer
 * may be NULL, which leaves that field alone. */
typedef struct gray_frame_args {
    float *origin;             /* device [3] <- origin_value */
    float *rotation_c2w;       /* device [3,3] <- c2w_value (row-major) */
    float *rotation_w2c;       /* device [3,3] <- the transpose of c2w_value */
    float *vertical_fov;       /* device [1] <- fov_value (radians) */
    float *znear;              /* device [1] <- znear_value */
    float *zfar;               /* device [1] <- zfar_value */
    uint8_t *rays_from_python; /* device bool [1] <- false */
    uint8_t *grads_enabled;    /* device bool [1] <- grads_enabled_value != 0 */
    int32_t *total_num_calls;  /* device [1] += 1 */
    uint8_t *was_visible;      /* device bool [count] <- false */
    int64_t count;
    float origin_value[3];
    float c2w_value[9];
    float fov_value, znear_value, zfar_value;
    int32_t grads_enabled_value;
} gray_frame_args;
int gray_frame_setup(const gray_frame_args *args, void *stream);

/* ---- geometry + BVH ------------------------------------------------------- */

/* Replaces BVHWrapper::rebuild / BVHWrapper::update and populate_bvh_kernel
 * (cuda/optix/bvh_wrapper.h:15-39, bvh_wrapper.cu:8-53, raytracer.h:135-142).
 * Snapshots the per-Gaussian geometry (opacity, clip factor, world-to-local
 * transform, exact ellipsoid AABB) and either rebuilds the LBVH (rebuild != 0:
 * Morton codes, LDS radix sort, Karras topology, refit) or refits the existing
 * topology.  A refit needs an earlier rebuild with the same N. */
int gray_bvh_update(gray_ctx *ctx, const gray_params *params, const gray_knobs *knobs, int rebuild,
                    void *stream);

/* Milliseconds of the last rebuild (device time, measured with HIP events),
 * or -1 if none was timed.  Timing is only recorded when enabled; it also
 * records the forward's stage events (see gray_render_stage_ms). */
int gray_bvh_set_timing(gray_ctx *ctx, int enable);
float gray_bvh_last_build_ms(gray_ctx *ctx);

/* No reference counterpart (serving: allocation up front).  Allocates the
 * collect-pass lists and task buffers for launches of up to max_lanes lanes
 * (sum over the views of ceil(w/8)*ceil(h/8)*64) at list_capacity hits per
 * ray, so that later gray_render_forward calls within those sizes allocate
 * nothing.  Buffers never shrink. */
int gray_render_reserve(gray_ctx *ctx, int64_t max_lanes, int32_t list_capacity);

/* No reference counterpart (profiling aid).  Device milliseconds of the
 * stages of the last timed gray_render_forward (its first 8 views): ms[0] the
 * per-tile collect kernel, ms[1] the collect task rounds, ms[2] list
 * compositing, ms[3] the re-traversal forward (stages a launch skipped read
 * ~0).  Waits for that forward to finish.  max_stages >= 4. */
int gray_render_stage_ms(gray_ctx *ctx, float *ms, int32_t max_stages);

/* Number of Gaussians and internal nodes of the current BVH. */
int gray_bvh_info(gray_ctx *ctx, int64_t *num_gaussians, int64_t *num_nodes, int64_t *depth_bound);

/* ---- spherical harmonics -------------------------------------------------- */

/* Replaces sh_forward_pass / sh_forward_kernel (cuda/utils/sh.cu:224-247,
 * 291-298): rgb_out[v] = max(C0*dc + sum Y_k(dir)*c_k + 0.5, 0) with
 * dir = normalize(mean - origin[v]).  origins: host array of num_views device
 * pointers to [3]; rgb_out: host array of num_views device pointers to [N,3].
 * Returns immediately (no work) when *knobs->enable_sh is 0 on the device. */
int gray_sh_forward(gray_ctx *ctx, const gray_params *params, const gray_knobs *knobs, int32_t num_views,
                    const float *const *origins, float *const *rgb_out, void *stream);

/* Replaces sh_backward_pass / sh_backward_kernel (sh.cu:249-287, 300-304).
 * For Gaussians with was_visible: masks grad_rgb where rgb <= 0, writes
 * dL/d(sh_dc), dL/d(sh_rest) and adds the view-direction term to grad_mean.
 * accumulate == 0 reproduces the reference assignment semantics for the
 * dc/rest gradients (single view); accumulate != 0 sums over the views. */
int gray_sh_backward(gray_ctx *ctx, const gray_params *params, const gray_knobs *knobs, int32_t num_views,
                     const float *const *origins, const float *const *rgb, const float *const *grad_rgb,
                     const uint8_t *was_visible, float *grad_sh_dc, float *grad_sh_rest, float *grad_mean,
                     int accumulate, void *stream);

/* ---- render --------------------------------------------------------------- */

/* Replaces the forward half of Raytracer::forward_pass (raytracer.h:88-102):
 * __raygen__rg + __intersection__is + forward_pass (shaders.cu:10-106,
 * forward_pass.cu:6-151).  Needs a prior gray_bvh_update with the same N.
 * Views are processed in launches of up to 8; rays in 8x8 tiles, one wave64
 * per tile.  With list_capacity > 0 (the default path):
 *   1. collect: each lane walks the 4-wide BVH for its own ray (fp16 child
 *      boxes, short stack in LDS, leaves queued in LDS), runs the exact
 *      ray-ellipsoid test (shaders.cu:22-47) and appends every hit to the
 *      ray's list in the context (16 B per hit, slot-major); after a budget
 *      of iterations the rest of a walk is handed on as {ray, subtree} tasks;
 *      a persistent pooled round gives each wave 32 tasks whose frontier
 *      (node and leaf stacks in LDS) its 64 lanes drain together, two nodes
 *      and two leaves per lane and step, so long walks are split across lanes;
 *   2. composite: one lane per ray scans its list with an 8-entry register
 *      k-buffer, in the reference's batches of 32 beyond tmin (sub-rounds of
 *      8), front to back, stopping at T < t_threshold; lists longer than
 *      wave_list_min (up to 512 hits) are composited by a whole wave each;
 *      the tiles none of whose rays handed work on are composited while the
 *      pooled round still runs;
 *   3. rays whose list or tasks overflowed are re-traversed per batch by a
 *      k-buffer kernel (16 entries in LDS), as is every ray when
 *      list_capacity == 0.
 * The context's second stream runs the SH colours (sh_params) beside the
 * collect pass, the compositing of the finished tiles beside the pooled
 * round and the wave-per-ray compositing beside the per-lane one; all are
 * joined back into `stream` with events before the call returns.
 * Cost model: the collect walk is bound by the dependent latency of its
 * node visits (about 95 4-wide node visits and 44 leaf tests per ray on a
 * 1M-Gaussian surface scene); compositing reads each list once. */
int gray_render_forward(gray_ctx *ctx, const gray_render_args *args, void *stream);

/* Replaces the ellipsoid display mode of the forward launch (config
 * render_ellipsoids, shaders.cu:90-93; ellipsoid_viewer.cu:1-94): per ray the
 * nearest proxy ellipsoid (the clipped unit sphere of the geometry snapshot)
 * with opacity >= *min_opacity (device [1]) in [znear, zfar], ties to the lower
 * id; out_channels = max(0.5, -d.n) (rgb + 0.15) with n the world normal,
 * out_depth = hit distance; background and depth 0 on a miss.  Needs each
 * view's rgb and out_channels; out_depth may be NULL.  Nothing else is
 * written. */
int gray_render_ellipsoids(gray_ctx *ctx, const gray_render_args *args, const float *min_opacity, void *stream);

/* Replaces Raytracer::backward_pass (raytracer.h:116-124, backward_pass.cu:2-142):
 * replays each ray's composited hits (stored by the forward, or recomputed by
 * re-traversal for rays whose list exceeded hit_capacity) back to front and
 * scatters dL/d(mean, rotation, scale, opacity) into grads and dL/d(rgb) into
 * view.grad_rgb.  Uses the geometry snapshot of the forward. */
int gray_render_backward(gray_ctx *ctx, const gray_render_args *args, const gray_params *params,
                         const gray_grads *grads, void *stream);

/* ---- optimiser ------------------------------------------------------------ */

/* Replaces adam_step + adam_step_base_kernel + adam_step_sh_kernel
 * (cuda/opt/adam.cu:25-84, raytracer.h:126-133) and the pruning counter update
 * of Raytracer.update_pruning_stats (gray/raytracer.py:361-366), in ONE
 * launch.  step is total_num_steps after the increment (or *adam->step_device
 * when given); the bias correction uses step+1 like the reference
 * (adam.cu:16-17).  flags: bit0 update_channels,
 * bit1 zero_grads, bit2 update SH (enable_sh && step % sh_update_laziness == 0);
 * each is replaced by the device-resident switch in gray_adam when given. */
int gray_adam_step(gray_ctx *ctx, const gray_adam *adam, int32_t step, int32_t flags, void *stream);

/* ---- training loss ----------------------------------------------------- */

/* Replaces the loss of train.py:255-260 (F.l1_loss + the third-party
 * fused_ssim, git a7c48d6) and the gradient copy into the framebuffer of
 * gray/raytracer.py:175-179.  SSIM: 11-tap Gaussian window (sigma 1.5),
 * C1 = 0.01^2, C2 = 0.03^2.  Workspace needed by the two calls below. */
int gray_loss_workspace_bytes(int32_t planes, int32_t height, int32_t width, int64_t *bytes);

/* Writes args->loss (deterministic: fixed-order double reduction) and, when
 * args->train, the derivative maps into the workspace. */
int gray_loss_forward(const gray_loss_args *args, void *stream);

/* Writes args->grad from the maps of the preceding gray_loss_forward (same
 * args, train = 1, same workspace, not reused in between).  The L1 term's
 * gradient is sign(render - target) / (P H W), sign(0) = 0, as torch's
 * l1_loss. */
int gray_loss_backward(const gray_loss_args *args, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving or training loops,
 * harnesses that forbid hipMalloc): the loss of gray_loss_forward and, when
 * args->grad is not NULL, the gradient of gray_loss_backward, in one call,
 * with the workspace held by the library in a static device array of
 * GRAY_LOSS_NOALLOC_MAX_BYTES per GPU (allocated when the library's code is
 * loaded on that GPU, held until the process ends).  args->workspace must be
 * NULL and args->train is ignored.  Shapes whose workspace
 * (gray_loss_workspace_bytes) exceeds the maximum are rejected with
 * GRAY_ERR_INVALID; the limit admits a 1920x1080 RGB image.  Calls no
 * allocating routine.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (on different streams or host threads) race. */
#define GRAY_LOSS_NOALLOC_MAX_BYTES (80ll << 20)
int gray_loss_noalloc(const gray_loss_args *args, void *stream);

/* ---- debugging / parity --------------------------------------------------- */

/* No reference counterpart (the reference keeps hits in its per-pixel linked
 * lists, per_pixel_linked_list.h:5-64).  For every ray of view 0 writes the
 * sorted (distance, Gaussian id) list of ALL hits passing the hit criterion
 * of shaders.cu:22-47, capped at max_hits: ids/dists are [R, max_hits],
 * counts [R] (true count, may exceed max_hits).  Used for bit-exact hit-set
 * checks against the CPU oracle and between BVH and brute-force traversal. */
int gray_debug_hitlists(gray_ctx *ctx, const gray_render_args *args, int32_t max_hits, uint32_t *ids,
                        float *dists, int32_t *counts, void *stream);

/* Copies the geometry snapshot of the last gray_bvh_update to caller buffers
 * (device): geo [N,16] = {W2L 3x3 row-major, mean 3, opacity, clip, 2 pad},
 * aabb [N,6] = {min xyz, max xyz}.  Either pointer may be NULL. */
int gray_debug_snapshot(gray_ctx *ctx, float *geo, float *aabb, void *stream);

/* No reference counterpart (OptiX's acceleration structure is opaque).
 * Copies the LBVH (device buffers): nodes [max(N-1,1), 16] floats =
 * {left box 6, right box 6, left ref, right ref, range lo, range hi} with refs
 * as uint32 bits (bit 31 = leaf, payload Gaussian id; 0xffffffff = none),
 * sorted_ids [N] (leaf order), keys [N] (30-bit Morton codes, ascending). */
int gray_debug_bvh(gray_ctx *ctx, float *nodes, uint32_t *sorted_ids, uint32_t *keys, void *stream);

/* No reference counterpart (diagnostic of the collect pass's tree).  Copies
 * the 4-wide nodes of the last gray_bvh_update (device buffers): nodes
 * [max(N-1,1), 16] uint32 at the binary node index = {min x, min y, min z,
 * max x, max y, max z as half pairs (children 0,1 then 2,3; the lower-numbered
 * child in the low half), child refs 4}, meaningful where flags [max(N-1,1)]
 * is 1 (internal nodes at even depth, root depth 0).  Either pointer may be
 * NULL. */
int gray_debug_wide(gray_ctx *ctx, uint32_t *nodes, uint8_t *flags, void *stream);

/* No reference counterpart (diagnostic of the collect pass).  Copies the hit
 * counts per ray of the last forward launch that collected lists (device
 * buffer, int32 [tiles*64], tile-major in the launch's tile order, lane =
 * 8*(y%8) + x%8; negative = the ray fell back to re-traversal).  *lanes receives
 * the number of entries; counts may be NULL to query it. */
int gray_debug_list_counts(gray_ctx *ctx, int32_t *counts, int64_t max_lanes, int64_t *lanes, void *stream);

/* No reference counterpart (profiling aid).  Copies `bytes` (a multiple of
 * 16, both pointers 16-byte aligned) from src to dst with non-temporal 16-byte
 * loads and stores: the sustained read + write rate of a kernel that moves
 * each byte once, the reference point quoted beside the streaming kernels. */
int gray_debug_stream_copy(void *dst, const void *src, int64_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GRAY_HIP_H */
