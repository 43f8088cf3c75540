/*
 * gray_oracle.c -- CPU oracle of the differentiable Gaussian ray tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker (or
 * the timed CPU baseline).  The product (gray_amd/) never calls it.
 *
 * A plain-C restatement of the reference algorithm, function by function:
 *   or_snapshot      cuda/optix/bvh_wrapper.cu:8-53, cuda/utils/kernel.cu:3-22,
 *                    cuda/utils/activations.cu:5-70 (sigmoid, exp, normalize)
 *   or_pinhole_rays  cuda/core/camera.h:17-36 (+ gray/camera.py:109-119)
 *   or_hit           cuda/shaders.cu:22-60 (IS: cull, closest approach, clip, alpha)
 *   or_render        cuda/forward_pass.cu:6-151 (batches of 32 beyond tmin,
 *                    front-to-back, T < t_threshold stop, truncation estimate,
 *                    depth normalisation, background)
 *   or_backward      cuda/backward_pass.cu:2-142 (back to front, T/(1-a))
 *   or_sh_forward    cuda/utils/sh.cu:11-78, 224-247 (complex recursion)
 *   or_sh_backward   cuda/utils/sh.cu:80-222, 249-287
 *   or_adam          cuda/opt/adam.cu:13-22
 *
 * The hit test uses IEEE single precision + - * / in the same order as the
 * GPU kernel (gray_amd/csrc/common.hpp), compiled with -ffp-contract=off, so
 * for identical snapshots and rays the hit sets, distances and order agree bit
 * for bit.  Transcendentals come from libm and agree to a few ulps.
 *
 * Truncation bookkeeping: mode 0 forms the estimate over the non-composited
 * hits directly (the same quantity the GPU forms); mode 1 is the literal
 * reference arithmetic sum_all - sum_composited and full_T = prod_all (1-a)
 * in Gaussian-id order (the reference's order is its BVH order).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_ALPHA 0.9999f
#define BATCH 32
#define MAX_ROUNDS 99

typedef struct {
    float exp_power, alpha_threshold, t_threshold, eps_fwd, eps_scale_grad;
    float bg[3];
    float znear, zfar;
    int render_depth;
    int fp16_hits; /* reference hit storage: distances and compositing alphas rounded to fp16 */
} or_knobs;

/* float -> IEEE binary16 -> float, round to nearest even (the reference's
 * __float2half_rn, per_pixel_linked_list.h:28-29), written out on the bits. */
float or_fp16_round(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint32_t sign = u & 0x80000000u;
    uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;      /* inf, nan */
    if (a >= 0x477ff000u) {              /* >= 65520 rounds past 65504 */
        u = sign | 0x7f800000u;
        memcpy(&x, &u, 4);
        return x;
    }
    if (a < 0x38800000u) {               /* below 2^-14: binary16 subnormals, spacing 2^-24 */
        const float q = 5.9604644775390625e-08f;
        const float r = nearbyintf(fabsf(x) / q) * q; /* exact scaling; nearbyint rounds half to even */
        return sign ? -r : r;
    }
    a += 0x0fffu + ((a >> 13) & 1u);     /* drop 13 mantissa bits, ties to even */
    a &= 0xffffe000u;
    u = sign | a;
    memcpy(&x, &u, 4);
    return x;
}

void or_fp16_round_array(int64_t n, const float *x, float *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = or_fp16_round(x[i]);
}

static float sigmoidf_(float x) { return 1.0f / (1.0f + expf(-x)); }

static float clip_factor(float op, float athr, float p) {
    if (op <= athr) return 0.0f;
    const float lr = logf(op / athr);
    if (p == 1.0f) return sqrtf(2.0f * lr);
    if (p == 2.0f) return sqrtf(sqrtf(4.0f * lr));
    if (p == 3.0f) return cbrtf(sqrtf(6.0f * lr));
    const float k = 2.0f * p;
    return powf(k * lr, 1.0f / k);
}

static float eval_kernel(float q, float p) {
    if (p == 1.0f) return expf(-0.5f * q);
    if (p == 2.0f) return expf((-0.25f * q) * q);
    if (p == 3.0f) return expf(-((q * q) * q) / 6.0f);
    return expf(-powf(q, p) / (2.0f * p));
}

/* geo record: m[9] (W2L row-major), mu[3], op (or -1 invisible), clip, pad, pad */
void or_snapshot(int n, const float *mean, const float *rot, const float *scale, const float *opacity, float athr,
                 float p, float gsf, float *geo, float *l2w) {
    for (int g = 0; g < n; ++g) {
        const float op = sigmoidf_(opacity[g]);
        const float clip = clip_factor(op, athr, p);
        float s[3];
        for (int i = 0; i < 3; ++i) s[i] = (expf(scale[3 * g + i]) * clip) * gsf;
        const int visible = clip > 0.0f && (s[0] > 0.0f || s[1] > 0.0f || s[2] > 0.0f);
        float r = rot[4 * g], x = rot[4 * g + 1], y = rot[4 * g + 2], z = rot[4 * g + 3];
        const float norm = sqrtf(r * r + x * x + y * y + z * z);
        r /= norm;
        x /= norm;
        y /= norm;
        z /= norm;
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
        float *G = geo + 16 * (size_t)g;
        for (int i = 0; i < 3; ++i) {
            const float inv = 1.0f / s[i];
            for (int j = 0; j < 3; ++j) {
                G[3 * i + j] = R[3 * j + i] * inv;
                if (l2w) l2w[12 * (size_t)g + 3 * i + j] = R[3 * i + j] * s[j];
            }
        }
        if (l2w) l2w[12 * (size_t)g + 9] = l2w[12 * (size_t)g + 10] = l2w[12 * (size_t)g + 11] = 0.0f;
        G[9] = mean[3 * g];
        G[10] = mean[3 * g + 1];
        G[11] = mean[3 * g + 2];
        G[12] = visible ? op : -1.0f;
        G[13] = clip;
        G[14] = G[15] = 0.0f;
    }
}

/* Tiny Encryption Algorithm (Wheeler and Needham, 1994): `rounds` cycles on
 * (v0, v1) under key k; returns the first word.  The reference seeds its ray
 * jitter with 4 rounds under the key of the published test (random.h). */
uint32_t or_tea(uint32_t v0, uint32_t v1, int rounds, const uint32_t *k) {
    uint32_t sum = 0;
    for (int n = 0; n < rounds; ++n) {
        sum += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + sum) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + sum) ^ ((v0 >> 5) + k[3]);
    }
    return v0;
}

/* Numerical Recipes LCG step (a = 1664525, c = 1013904223, mod 2^32) */
uint32_t or_lcg(uint32_t state) { return 1664525u * state + 1013904223u; }

/* camera.h:17-36 with the sub-pixel jitter: offsets (draw - 0.5) for x then y,
 * draws (lcg & 0xFFFFFF) / 2^24 from the seed tea4(pixel, call index) */
void or_pinhole_rays_jittered(int W, int H, int stride, float fov, const float *origin, const float *c,
                              uint32_t calls, float *ro, float *rd) {
    static const uint32_t key[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
    const float vs = tanf(fov / 2.0f);
    const float aspect = (float)W / (float)H;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            uint32_t st = or_tea((uint32_t)(y * stride + x), calls, 4, key);
            st = or_lcg(st);
            const float jx = (float)(st & 0x00FFFFFFu) / (float)0x01000000 - 0.5f;
            st = or_lcg(st);
            const float jy = (float)(st & 0x00FFFFFFu) / (float)0x01000000 - 0.5f;
            const float px = (float)x + jx, py = (float)y + jy;
            const float yf = vs * (1.0f - (2.0f * (py + 0.5f)) / (float)H);
            const float xf = (aspect * vs) * ((2.0f * (px + 0.5f)) / (float)W - 1.0f);
            float d0 = (c[0] * xf + c[1] * yf) - c[2];
            float d1 = (c[3] * xf + c[4] * yf) - c[5];
            float d2 = (c[6] * xf + c[7] * yf) - c[8];
            const float len = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
            const size_t i = (size_t)y * W + x;
            ro[3 * i] = origin[0];
            ro[3 * i + 1] = origin[1];
            ro[3 * i + 2] = origin[2];
            rd[3 * i] = d0 / len;
            rd[3 * i + 1] = d1 / len;
            rd[3 * i + 2] = d2 / len;
        }
}

void or_pinhole_rays(int W, int H, int stride, float fov, const float *origin, const float *c, float *ro, float *rd) {
    const float vs = tanf(fov / 2.0f);
    const float aspect = (float)W / (float)H;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int pix = y * stride + x;
            const float yf = vs * (1.0f - (2.0f * ((float)y + 0.5f)) / (float)H);
            const float xf = (aspect * vs) * ((2.0f * ((float)x + 0.5f)) / (float)W - 1.0f);
            float dx = (c[0] * xf + c[1] * yf) - c[2];
            float dy = (c[3] * xf + c[4] * yf) - c[5];
            float dz = (c[6] * xf + c[7] * yf) - c[8];
            const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
            rd[3 * pix] = dx / len;
            rd[3 * pix + 1] = dy / len;
            rd[3 * pix + 2] = dz / len;
            ro[3 * pix] = origin[0];
            ro[3 * pix + 1] = origin[1];
            ro[3 * pix + 2] = origin[2];
        }
}

/* IS criterion; returns 1 on a hit with dist and local point h */
static int or_hit(const float *G, const float *o, const float *d, float *dist, float *h) {
    const float px = o[0] - G[9], py = o[1] - G[10], pz = o[2] - G[11];
    const float lox = (G[0] * px + G[1] * py) + G[2] * pz;
    const float loy = (G[3] * px + G[4] * py) + G[5] * pz;
    const float loz = (G[6] * px + G[7] * py) + G[8] * pz;
    const float ldx = (G[0] * d[0] + G[1] * d[1]) + G[2] * d[2];
    const float ldy = (G[3] * d[0] + G[4] * d[1]) + G[5] * d[2];
    const float ldz = (G[6] * d[0] + G[7] * d[1]) + G[8] * d[2];
    const float od = (lox * ldx + loy * ldy) + loz * ldz;
    if (od > 0.0f) return 0;
    const float dd = (ldx * ldx + ldy * ldy) + ldz * ldz;
    const float t = (-od) / dd;
    h[0] = lox + t * ldx;
    h[1] = loy + t * ldy;
    h[2] = loz + t * ldz;
    const float sq = (h[0] * h[0] + h[1] * h[1]) + h[2] * h[2];
    if (sq > 1.0f) return 0;
    *dist = t;
    return 1;
}

static float or_alpha(const float *G, const float *h, float p, float *gval, float *q_out) {
    const float clip = G[13];
    const float lx = h[0] * clip, ly = h[1] * clip, lz = h[2] * clip;
    const float q = (lx * lx + ly * ly) + lz * lz;
    const float g = eval_kernel(q, p);
    if (gval) *gval = g;
    if (q_out) *q_out = q;
    return (MAX_ALPHA * g) * G[12];
}

typedef struct {
    float d;
    uint32_t id;
    float a;
} hitrec;

static int hit_cmp(const void *x, const void *y) {
    const hitrec *a = (const hitrec *)x, *b = (const hitrec *)y;
    if (a->d < b->d) return -1;
    if (a->d > b->d) return 1;
    return (a->id < b->id) ? -1 : (a->id > b->id);
}

/* every hit of one ray, in Gaussian-id order; returns count (<= cap written) */
static int collect(int n, const float *geo, float athr, float p, const float *o, const float *d, hitrec *out,
                   int cap, int fp16) {
    int cnt = 0;
    for (int g = 0; g < n; ++g) {
        const float *G = geo + 16 * (size_t)g;
        if (G[12] <= athr) continue;
        float dist, h[3];
        if (!or_hit(G, o, d, &dist, h)) continue;
        if (cnt < cap) {
            out[cnt].d = fp16 ? or_fp16_round(dist) : dist;
            out[cnt].id = (uint32_t)g;
            out[cnt].a = or_alpha(G, h, p, NULL, NULL);
        }
        ++cnt;
    }
    return cnt;
}

/* sorted (dist, id) list of all hits of each ray (ids/dists [R, maxh]) */
void or_hitlists(int n, const float *geo, float athr, int R, const float *ro, const float *rd, int maxh,
                 uint32_t *ids, float *dists, int32_t *counts) {
#pragma omp parallel
    {
        hitrec *buf = (hitrec *)malloc(sizeof(hitrec) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 16)
        for (int r = 0; r < R; ++r) {
            const int cnt = collect(n, geo, athr, 2.0f, ro + 3 * (size_t)r, rd + 3 * (size_t)r, buf, n, 0);
            qsort(buf, (size_t)cnt, sizeof(hitrec), hit_cmp);
            counts[r] = cnt;
            for (int k = 0; k < cnt && k < maxh; ++k) {
                ids[(size_t)r * maxh + k] = buf[k].id;
                dists[(size_t)r * maxh + k] = buf[k].d;
            }
        }
        free(buf);
    }
}

/*
 * Ellipsoid display mode (ellipsoid_viewer.cu:1-94): per ray the nearest proxy
 * ellipsoid with opacity >= min_op in [znear, zfar] (ties: lower id), the
 * stable closest-approach sphere test in local coordinates, flat shading
 * max(0.5, -d.n) (rgb + 0.15) with the world normal W2L^T n_local; depth = t.
 * A miss leaves bg and depth 0.  Brute force over all Gaussians.
 */
void or_ellipsoids(int n, const float *geo, const float *rgb, int R, const float *ro, const float *rd, float min_op,
                   float znear, float zfar, const float *bg, float *C, float *D) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int r = 0; r < R; ++r) {
        const float *o = ro + 3 * (size_t)r, *d = rd + 3 * (size_t)r;
        float best = INFINITY, bn[3] = {0.0f, 0.0f, 0.0f};
        int bid = -1;
        for (int g = 0; g < n; ++g) {
            const float *G = geo + 16 * (size_t)g;
            if (!(G[12] > 0.0f) || G[12] < min_op) continue;
            const float p[3] = {o[0] - G[9], o[1] - G[10], o[2] - G[11]};
            float lo[3], ld[3];
            for (int i = 0; i < 3; ++i) {
                lo[i] = (G[3 * i] * p[0] + G[3 * i + 1] * p[1]) + G[3 * i + 2] * p[2];
                ld[i] = (G[3 * i] * d[0] + G[3 * i + 1] * d[1]) + G[3 * i + 2] * d[2];
            }
            if ((lo[0] * ld[0] + lo[1] * ld[1]) + lo[2] * ld[2] > 0.0f) continue;
            const float norm = sqrtf((ld[0] * ld[0] + ld[1] * ld[1]) + ld[2] * ld[2]);
            const float u[3] = {ld[0] / norm, ld[1] / norm, ld[2] / norm};
            const float smid = -((lo[0] * u[0] + lo[1] * u[1]) + lo[2] * u[2]);
            const float c[3] = {lo[0] + smid * u[0], lo[1] + smid * u[1], lo[2] + smid * u[2]};
            const float sq = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
            if (sq > 1.0f) continue;
            const float half = sqrtf(1.0f - sq);
            const float t0 = (smid - half) / norm, t1 = (smid + half) / norm;
            float t;
            if (t0 >= znear && t0 <= zfar)
                t = t0;
            else if (t1 >= znear && t1 <= zfar)
                t = t1;
            else
                continue;
            if (t < best || (t == best && g < bid)) {
                const float tn = t * norm;
                best = t;
                bid = g;
                for (int i = 0; i < 3; ++i) bn[i] = lo[i] + tn * u[i];
            }
        }
        float out[3] = {bg[0], bg[1], bg[2]}, depth = 0.0f;
        if (bid >= 0) {
            const float *G = geo + 16 * (size_t)bid;
            const float ln = sqrtf((bn[0] * bn[0] + bn[1] * bn[1]) + bn[2] * bn[2]);
            const float l[3] = {bn[0] / ln, bn[1] / ln, bn[2] / ln};
            float w[3];
            for (int j = 0; j < 3; ++j) w[j] = (G[j] * l[0] + G[3 + j] * l[1]) + G[6 + j] * l[2];
            const float wn = sqrtf((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
            const float dn = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            float al = ((-d[0] / dn) * (w[0] / wn) + (-d[1] / dn) * (w[1] / wn)) + (-d[2] / dn) * (w[2] / wn);
            al = al > 0.5f ? al : 0.5f;
            for (int j = 0; j < 3; ++j) out[j] = al * (rgb[3 * (size_t)bid + j] + 0.15f);
            depth = best;
        }
        for (int j = 0; j < 3; ++j) C[3 * (size_t)r + j] = out[j];
        D[r] = depth;
    }
}

/*
 * Forward for R explicit rays.  Outputs per ray: C[3], D, T, F, est[3],
 * ncomp and the composited list (ids, T before the hit) up to cap entries.
 * rgb is the per-Gaussian colour [n,3].  mode: 0 skip-set, 1 literal.
 */
void or_render(int n, const float *geo, const float *rgb, int R, const float *ro, const float *rd,
               const or_knobs *k, int mode, float *C, float *D, float *Tout, float *Fout, float *est, int32_t *ncomp,
               int cap, uint32_t *list_ids, float *list_T, int32_t *nhits) {
    const float p = k->exp_power, athr = k->alpha_threshold, tthr = k->t_threshold;
#pragma omp parallel
    {
        hitrec *buf = (hitrec *)malloc(sizeof(hitrec) * (size_t)(n > 0 ? n : 1));
        unsigned char *comp = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 16)
        for (int r = 0; r < R; ++r) {
            const float *o = ro + 3 * (size_t)r, *d = rd + 3 * (size_t)r;
            const float dl = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            const int h16 = k->fp16_hits;
            const int cnt = dl != 0.0f ? collect(n, geo, athr, p, o, d, buf, n, h16) : 0;
            if (nhits) nhits[r] = cnt;
            /* literal reference sums over all hits (traversal order = id order here): the
             * full transmittance of the hit program (fp32 alpha), the skipped sums of the
             * stored (fp16 in fp16_hits mode) alphas */
            float full_T = 1.0f, sk0 = 0.0f, sk1 = 0.0f, sk2 = 0.0f, ska = 0.0f;
            for (int i = 0; i < cnt; ++i) {
                full_T *= 1.0f - buf[i].a;
                const float ah = h16 ? or_fp16_round(buf[i].a) : buf[i].a;
                const float *c = rgb + 3 * (size_t)buf[i].id;
                sk0 += ah * c[0];
                sk1 += ah * c[1];
                sk2 += ah * c[2];
                ska += ah;
            }
            qsort(buf, (size_t)cnt, sizeof(hitrec), hit_cmp);
            memset(comp, 0, (size_t)(cnt > 0 ? cnt : 1));
            float T = 1.0f, Tf = 1.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, depth = 0.0f, tmin = k->znear;
            int nc = 0;
            int start = 0; /* first sorted index with dist > tmin */
            for (int it = 0; it < MAX_ROUNDS && tmin < k->zfar; ++it) {
                while (start < cnt && !(buf[start].d > tmin)) ++start;
                if (start >= cnt) break; /* reference: empty batch */
                const int end = start + BATCH < cnt ? start + BATCH : cnt;
                int stop = 0;
                for (int i = start; i < end; ++i) {
                    const float dist = buf[i].d;
                    tmin = dist > tmin ? dist : tmin;
                    if (dist < k->zfar) {
                        const float a32 = buf[i].a;
                        const float a = h16 ? or_fp16_round(a32) : a32;
                        const float *c = rgb + 3 * (size_t)buf[i].id;
                        if (mode == 1) {
                            sk0 -= a * c[0];
                            sk1 -= a * c[1];
                            sk2 -= a * c[2];
                            ska -= a;
                        }
                        const float w = T * a;
                        c0 += c[0] * w;
                        c1 += c[1] * w;
                        c2 += c[2] * w;
                        if (k->render_depth) depth += dist * w;
                        if (nc < cap) {
                            list_ids[(size_t)r * cap + nc] = buf[i].id;
                            list_T[(size_t)r * cap + nc] = T;
                        }
                        T = T * (1.0f - a);
                        Tf = Tf * (1.0f - a32);
                        comp[i] = 1;
                        ++nc;
                        if (T < tthr) {
                            tmin = k->zfar;
                            stop = 1;
                            break;
                        }
                    }
                }
                if (stop) break;
                if (end - start < BATCH) break; /* reference: FLT_MAX slots end the loop */
                start = end;
            }
            float F, e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
            if (mode == 1) {
                F = full_T;
                if (ska > 0.0f) {
                    e0 = sk0 / ska;
                    e1 = sk1 / ska;
                    e2 = sk2 / ska;
                }
            } else {
                float pnc = 1.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, sa = 0.0f;
                for (int i = 0; i < cnt; ++i) {
                    if (comp[i]) continue;
                    const float a32 = buf[i].a;
                    const float a = h16 ? or_fp16_round(a32) : a32;
                    const float *c = rgb + 3 * (size_t)buf[i].id;
                    s0 += a * c[0];
                    s1 += a * c[1];
                    s2 += a * c[2];
                    sa += a;
                    pnc = pnc * (1.0f - a32);
                }
                F = Tf * pnc;
                if (sa > 0.0f) {
                    e0 = s0 / sa;
                    e1 = s1 / sa;
                    e2 = s2 / sa;
                }
            }
            const float rem = T - F;
            c0 += rem * e0;
            c1 += rem * e1;
            c2 += rem * e2;
            if (k->render_depth) {
                const float nrm = (1.0f - T) > k->eps_fwd ? (1.0f - T) : k->eps_fwd;
                depth = depth / nrm;
            }
            c0 += k->bg[0] * F;
            c1 += k->bg[1] * F;
            c2 += k->bg[2] * F;
            C[3 * (size_t)r] = c0;
            C[3 * (size_t)r + 1] = c1;
            C[3 * (size_t)r + 2] = c2;
            if (D) D[r] = depth;
            Tout[r] = T;
            Fout[r] = F;
            est[3 * (size_t)r] = e0;
            est[3 * (size_t)r + 1] = e1;
            est[3 * (size_t)r + 2] = e2;
            ncomp[r] = nc;
        }
        free(buf);
        free(comp);
    }
}

/*
 * Backward (reference backward_pass.cu:15-141) over the composited lists of
 * or_render (requires cap >= ncomp).  Live parameters (rotation raw, log
 * scale, logit opacity) give the activation derivatives; geo/l2w are the
 * snapshot.  Gradients accumulate in double (ray order), then are added to
 * the float outputs.  pixels = launch width*height (pruning normaliser).
 */
void or_backward(int n, const float *geo, const float *l2w, const float *rgb, const float *rotation,
                 const float *scale, const float *opacity, int R, const float *ro, const float *rd,
                 const float *grad_out, const float *Tfin, const float *F, const float *est, const int32_t *ncomp,
                 int cap, const uint32_t *list_ids, const or_knobs *k, float pixels, float *g_mean, float *g_rot,
                 float *g_scale, float *g_op, float *g_rgb, float *g_prune) {
    double *acc = (double *)calloc((size_t)n * 15, sizeof(double));
    const float p = k->exp_power;
    for (int r = 0; r < R; ++r) {
        const float *o = ro + 3 * (size_t)r, *d = rd + 3 * (size_t)r;
        const float g0 = grad_out[3 * (size_t)r], g1 = grad_out[3 * (size_t)r + 1], g2 = grad_out[3 * (size_t)r + 2];
        const float *e = est + 3 * (size_t)r;
        const float bgdot = (k->bg[0] * g0 + k->bg[1] * g1) + k->bg[2] * g2;
        const float estdot = (e[0] * g0 + e[1] * g1) + e[2] * g2;
        float gT = 0.0f;
        float T = Tfin[r];
        for (int i = ncomp[r] - 1; i >= 0; --i) {
            if (i >= cap) continue;
            const uint32_t id = list_ids[(size_t)r * cap + i];
            const float *G = geo + 16 * (size_t)id;
            float dist, h[3], gv, q;
            if (!or_hit(G, o, d, &dist, h)) continue;
            const float a = or_alpha(G, h, p, &gv, &q);
            const float *c = rgb + 3 * (size_t)id;
            const float w = T / (1.0f - a) * a;
            double *A = acc + 15 * (size_t)id;
            A[11] += (double)(g0 * w);
            A[12] += (double)(g1 * w);
            A[13] += (double)(g2 * w);
            const float gdot = (g0 * c[0] + g1 * c[1]) + g2 * c[2];
            float ga = (gdot - gT) * T / (1.0f - a);
            ga += -F[r] / (1.0f - a) * bgdot;
            ga += -((Tfin[r] - F[r]) / (1.0f - a)) * estdot;
            gT += (gdot - gT) * a;
            const float op = sigmoidf_(opacity[id]);
            float gop = MAX_ALPHA * ga * gv;
            gop = gop * op * (1.0f - op);
            const float ggv = MAX_ALPHA * ga * op;
            const float clip = G[13];
            const float l[3] = {h[0] * clip, h[1] * clip, h[2] * clip};
            float gsq;
            if (p == 1.0f)
                gsq = gv;
            else if (p == 2.0f)
                gsq = gv * q;
            else if (p == 3.0f)
                gsq = gv * q * q;
            else
                gsq = gv * powf(q, p - 1.0f);
            const float gxl[3] = {-l[0] * gsq * ggv, -l[1] * gsq * ggv, -l[2] * gsq * ggv};
            float gxw[3];
            for (int j = 0; j < 3; ++j) gxw[j] = (G[0 + j] * gxl[0] + G[3 + j] * gxl[1] + G[6 + j] * gxl[2]) * clip;
            float gl[3][3];
            for (int rr = 0; rr < 3; ++rr)
                for (int j = 0; j < 3; ++j) gl[rr][j] = -gxw[rr] * l[j];
            const float s[3] = {expf(scale[3 * id]), expf(scale[3 * id + 1]), expf(scale[3 * id + 2])};
            const float *Lw = l2w + 12 * (size_t)id;
            float gs[3];
            for (int j = 0; j < 3; ++j) {
                const float den = s[j] * clip + k->eps_scale_grad;
                gs[j] = (gl[0][j] * (Lw[j] / den) + gl[1][j] * (Lw[3 + j] / den) + gl[2][j] * (Lw[6 + j] / den)) * s[j];
            }
            float gR[3][3];
            for (int rr = 0; rr < 3; ++rr)
                for (int j = 0; j < 3; ++j) gR[rr][j] = gl[rr][j] * s[j];
            const float *qraw = rotation + 4 * (size_t)id;
            const float qlen = sqrtf(qraw[0] * qraw[0] + qraw[1] * qraw[1] + qraw[2] * qraw[2] + qraw[3] * qraw[3]);
            const float nr = qraw[0] / qlen, nx = qraw[1] / qlen, ny = qraw[2] / qlen, nz = qraw[3] / qlen;
            const float gr = 2.f * nx * (gR[2][1] - gR[1][2]) + 2.f * ny * (gR[0][2] - gR[2][0]) +
                             2.f * nz * (gR[1][0] - gR[0][1]);
            const float gx = -4.f * nx * (gR[1][1] + gR[2][2]) + 2.f * ny * (gR[0][1] + gR[1][0]) +
                             2.f * nz * (gR[0][2] + gR[2][0]) + 2.f * nr * (gR[2][1] - gR[1][2]);
            const float gy = 2.f * nx * (gR[0][1] + gR[1][0]) - 4.f * ny * (gR[0][0] + gR[2][2]) +
                             2.f * nz * (gR[1][2] + gR[2][1]) + 2.f * nr * (gR[0][2] - gR[2][0]);
            const float gz = 2.f * nx * (gR[0][2] + gR[2][0]) + 2.f * ny * (gR[1][2] + gR[2][1]) -
                             4.f * nz * (gR[0][0] + gR[1][1]) + 2.f * nr * (gR[1][0] - gR[0][1]);
            const float gq[4] = {gr, gx, gy, gz};
            const float qdot = gq[0] * qraw[0] + gq[1] * qraw[1] + gq[2] * qraw[2] + gq[3] * qraw[3];
            const float len3 = powf(qlen, 3.0f);
            for (int j = 0; j < 4; ++j) A[3 + j] += (double)(qdot * -qraw[j] / len3 + gq[j] / qlen);
            for (int j = 0; j < 3; ++j) A[7 + j] += (double)gs[j];
            for (int j = 0; j < 3; ++j) A[j] += (double)(-gxw[j]);
            A[10] += (double)gop;
            A[14] += (double)(w / pixels);
            T = T / (1.0f - a);
        }
    }
    for (int g = 0; g < n; ++g) {
        const double *A = acc + 15 * (size_t)g;
        for (int j = 0; j < 3; ++j) g_mean[3 * g + j] += (float)A[j];
        for (int j = 0; j < 4; ++j) g_rot[4 * g + j] += (float)A[3 + j];
        for (int j = 0; j < 3; ++j) g_scale[3 * g + j] += (float)A[7 + j];
        g_op[g] += (float)A[10];
        for (int j = 0; j < 3; ++j) g_rgb[3 * g + j] += (float)A[11 + j];
        if (g_prune) g_prune[g] += (float)A[14];
    }
    free(acc);
}

/* ---- spherical harmonics: complex recursion (sh.cu:11-222) ----------------- */
typedef struct {
    float re, im;
} cpx;
static cpx cmk(float r, float i) {
    cpx c = {r, i};
    return c;
}
static cpx cadd(cpx a, cpx b) { return cmk(a.re + b.re, a.im + b.im); }
static cpx cmul(cpx a, cpx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cpx cscale(cpx a, float s) { return cmk(a.re * s, a.im * s); }
static cpx cmuli(cpx a) { return cmk(-a.im, a.re); }
static int cidx(int l, int m) { return l * (l + 1) / 2 + m; }

#define SH_C0 0.28209479177387814f
#define SH_SQRT2 1.4142135623730951f

/* basis and (optionally) its x/y/z partials for degree <= 3; b* have 15 entries */
static void sh_basis_rec(int degree, float x, float y, float z, float *b, float *bx, float *by, float *bz) {
    cpx v[10], vx[10], vy[10], vz[10];
    v[0] = cmk(SH_C0, 0.0f);
    vx[0] = vy[0] = vz[0] = cmk(0.0f, 0.0f);
    const cpx xy = cmk(x, y);
    for (int l = 1; l <= degree; ++l) {
        const int dg = cidx(l, l), pd = cidx(l - 1, l - 1);
        const float ds = -sqrtf((2.0f * l + 1.0f) / (2.0f * l));
        const cpx pv = v[pd], pvx = vx[pd], pvy = vy[pd], pvz = vz[pd];
        v[dg] = cscale(cmul(xy, pv), ds);
        vx[dg] = cscale(cadd(pv, cmul(xy, pvx)), ds);
        vy[dg] = cscale(cadd(cmuli(pv), cmul(xy, pvy)), ds);
        vz[dg] = cscale(cmul(xy, pvz), ds);
        const int sd = cidx(l, l - 1);
        const float ss = sqrtf(2.0f * l + 1.0f);
        v[sd] = cscale(pv, ss * z);
        vx[sd] = cscale(pvx, ss * z);
        vy[sd] = cscale(pvy, ss * z);
        vz[sd] = cscale(cadd(pv, cscale(pvz, z)), ss);
        for (int m = l - 2; m >= 0; --m) {
            const int cur = cidx(l, m), p1 = cidx(l - 1, m), p2 = cidx(l - 2, m);
            const float den = (float)(l * l - m * m);
            const float a = sqrtf((4.0f * l * l - 1.0f) / den);
            const float bb = -sqrtf(((2.0f * l + 1.0f) * (((l - 1.0f) * (l - 1.0f)) - m * m)) / ((2.0f * l - 3.0f) * den));
            v[cur] = cadd(cscale(v[p1], a * z), cscale(v[p2], bb));
            vx[cur] = cadd(cscale(vx[p1], a * z), cscale(vx[p2], bb));
            vy[cur] = cadd(cscale(vy[p1], a * z), cscale(vy[p2], bb));
            vz[cur] = cadd(cscale(cadd(v[p1], cscale(vz[p1], z)), a), cscale(vz[p2], bb));
        }
    }
    int bi = 0;
    for (int l = 1; l <= degree; ++l) {
        for (int m = l; m >= 1; --m, ++bi) {
            const int ix = cidx(l, m);
            b[bi] = SH_SQRT2 * v[ix].im;
            if (bx) {
                bx[bi] = SH_SQRT2 * vx[ix].im;
                by[bi] = SH_SQRT2 * vy[ix].im;
                bz[bi] = SH_SQRT2 * vz[ix].im;
            }
        }
        {
            const int ix = cidx(l, 0);
            b[bi] = v[ix].re;
            if (bx) {
                bx[bi] = vx[ix].re;
                by[bi] = vy[ix].re;
                bz[bi] = vz[ix].re;
            }
            ++bi;
        }
        for (int m = 1; m <= l; ++m, ++bi) {
            const int ix = cidx(l, m);
            b[bi] = SH_SQRT2 * v[ix].re;
            if (bx) {
                bx[bi] = SH_SQRT2 * vx[ix].re;
                by[bi] = SH_SQRT2 * vy[ix].re;
                bz[bi] = SH_SQRT2 * vz[ix].re;
            }
        }
    }
}

/* basis only (for tests pinning it against the closed form) */
void or_sh_basis(int degree, float x, float y, float z, float *b, float *bx, float *by, float *bz) {
    sh_basis_rec(degree, x, y, z, b, bx, by, bz);
}

static void safe_dir(const float *mu, const float *o, float *delta, float *dir) {
    for (int i = 0; i < 3; ++i) delta[i] = mu[i] - o[i];
    float ls = delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2];
    if (ls < 1e-20f) ls = 1e-20f;
    const float inv = 1.0f / sqrtf(ls);
    for (int i = 0; i < 3; ++i) dir[i] = delta[i] * inv;
}

void or_sh_forward(int n, int K, int degree, const float *mean, const float *dc, const float *rest,
                   const float *origin, float *rgb) {
    if (degree > 3) degree = 3;
    for (int g = 0; g < n; ++g) {
        float delta[3], dir[3];
        safe_dir(mean + 3 * (size_t)g, origin, delta, dir);
        float res[3];
        for (int c = 0; c < 3; ++c) res[c] = SH_C0 * dc[3 * (size_t)g + c];
        if (degree > 0) {
            float b[15];
            sh_basis_rec(degree, dir[0], dir[1], dir[2], b, NULL, NULL, NULL);
            const int kc = (degree + 1) * (degree + 1) - 1;
            for (int k = 0; k < kc; ++k)
                for (int c = 0; c < 3; ++c) res[c] = fmaf(b[k], rest[((size_t)g * K + k) * 3 + c], res[c]);
        }
        for (int c = 0; c < 3; ++c) {
            const float v = res[c] + 0.5f;
            rgb[3 * (size_t)g + c] = v > 0.0f ? v : 0.0f;
        }
    }
}

/* accumulate: 0 = reference assignment of the dc/rest gradients, 1 = += */
void or_sh_backward(int n, int K, int degree, const float *mean, const float *rest, const float *origin,
                    const float *rgb, const float *grad_rgb, const uint8_t *visible, float *grad_dc, float *grad_rest,
                    float *grad_mean, int accumulate) {
    if (degree > 3) degree = 3;
    for (int g = 0; g < n; ++g) {
        if (visible && !visible[g]) continue;
        float delta[3], dir[3];
        safe_dir(mean + 3 * (size_t)g, origin, delta, dir);
        float gr[3];
        for (int c = 0; c < 3; ++c) gr[c] = rgb[3 * (size_t)g + c] <= 0.0f ? 0.0f : grad_rgb[3 * (size_t)g + c];
        for (int c = 0; c < 3; ++c) {
            if (accumulate)
                grad_dc[3 * (size_t)g + c] += SH_C0 * gr[c];
            else
                grad_dc[3 * (size_t)g + c] = SH_C0 * gr[c];
        }
        float rx[3] = {0, 0, 0}, ry[3] = {0, 0, 0}, rz[3] = {0, 0, 0};
        if (degree > 0) {
            float b[15], bx[15], by[15], bz[15];
            sh_basis_rec(degree, dir[0], dir[1], dir[2], b, bx, by, bz);
            const int kc = (degree + 1) * (degree + 1) - 1;
            for (int k = 0; k < kc; ++k) {
                for (int c = 0; c < 3; ++c) {
                    const float cv = rest[((size_t)g * K + k) * 3 + c];
                    float *dst = grad_rest + ((size_t)g * K + k) * 3 + c;
                    if (accumulate)
                        *dst += b[k] * gr[c];
                    else
                        *dst = b[k] * gr[c];
                    rx[c] = fmaf(bx[k], cv, rx[c]);
                    ry[c] = fmaf(by[k], cv, ry[c]);
                    rz[c] = fmaf(bz[k], cv, rz[c]);
                }
            }
        }
        const float gd[3] = {rx[0] * gr[0] + rx[1] * gr[1] + rx[2] * gr[2], ry[0] * gr[0] + ry[1] * gr[1] + ry[2] * gr[2],
                             rz[0] * gr[0] + rz[1] * gr[1] + rz[2] * gr[2]};
        float ls = delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2];
        if (ls < 1e-20f) ls = 1e-20f;
        const float inv = 1.0f / sqrtf(ls);
        const float inv3 = inv * inv * inv;
        const float dd = delta[0] * gd[0] + delta[1] * gd[1] + delta[2] * gd[2];
        for (int i = 0; i < 3; ++i) grad_mean[3 * (size_t)g + i] += gd[i] * inv - delta[i] * (dd * inv3);
    }
}

/* ---- Adam (adam.cu:13-22); step = total steps after the increment -------- */
void or_adam(int64_t n, float *p, float *g, float *m, float *v, float lr, float b1, float b2, float eps, int step,
             int zero) {
    const float bc1 = 1.0f - powf(b1, (float)step + 1.0f);
    const float bc2 = 1.0f - powf(b2, (float)step + 1.0f);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0f - b1) * g[i];
        v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
        const float mh = m[i] / bc1, vh = v[i] / bc2;
        p[i] -= lr * mh / (sqrtf(vh) + eps);
        if (zero) g[i] *= 0.0f;
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
