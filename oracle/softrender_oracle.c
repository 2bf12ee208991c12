/*
 * softrender_oracle.c — CPU restatement of softrender's rendering pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see softrender_oracle.h).  This is the checker
 * the HIP path is compared against, and, in its N-thread mode, the timed CPU
 * baseline.  The reference crate (Rust, /root/reference) cannot be built in
 * this image (no rustc/cargo, no crate registry), so this file restates its
 * single-threaded semantics from the source read as text.  Every function
 * cites the reference file:line it follows (paths relative to the crate).
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off (no -ffast-math): Rust never
 * contracts a*b+c into an FMA and uses IEEE division/sqrt, so neither may
 * we.  x86-64 evaluates float in single precision (FLT_EVAL_METHOD == 0).
 *
 * What pins each part is listed in tests/test_oracle_pinning.py; parts that
 * only the real crate could pin (nalgebra 0.12 mat-vec accumulation order,
 * num-traits 0.1.39 NumCast on out-of-range floats) are marked
 * "parity unpinned" where they are used.
 */
#define _GNU_SOURCE
#include "softrender_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float p[4];       /* position x,y,z,w */
    float k[OR_MAXK]; /* varyings ("uniforms" K of ClipVertex/ScreenVertex) */
} ovtx;

/* ------------------------------------------------------------------------ */
/* Interpolate (numeric/interpolate.rs:43-64): per f32 component            */
/*   barycentric = ((x1*u) + (x2*v)) + (x3*w)                               */
/*   linear      = ((1-t)*x1) + (t*x2)                                      */
/* The Matrix/Vector impls (interpolate.rs:153-336) and declare_uniforms!    */
/* (macros.rs:42-76) delegate component-wise to these.                      */
/* ------------------------------------------------------------------------ */
static inline float bary1(float u, float a, float v, float b, float w, float c) {
    return ((a * u) + (b * v)) + (c * w);
}
static inline float lerp1(float t, float a, float b) {
    return ((1.0f - t) * a) + (t * b);
}
static void lerp_vtx(float t, const ovtx* a, const ovtx* b, int K, ovtx* o) {
    for (int i = 0; i < 4; i++) o->p[i] = lerp1(t, a->p[i], b->p[i]);
    for (int i = 0; i < K; i++) o->k[i] = lerp1(t, a->k[i], b->k[i]);
}

/* ------------------------------------------------------------------------ */
/* Vertex programs (build-defined stand-ins for the user closure called by  */
/* VertexShader::run, vertex.rs:97-109, shaped like the examples'            */
/* `projection * view * model * position` and `(mit * normal).normalize()`, */
/* full_example/src/shaders.rs:8-31).  Mat-vec accumulates k = 0..3 from    */
/* +0.0 — the nalgebra 0.12 order (parity unpinned: needs the real crate).   */
/* ------------------------------------------------------------------------ */
static inline float dot3(const float* a, const float* b) {
    return ((a[0] * b[0]) + (a[1] * b[1])) + (a[2] * b[2]);
}
static inline void normalize3(float* v) {
    float len = sqrtf(dot3(v, v));
    v[0] = v[0] / len;
    v[1] = v[1] / len;
    v[2] = v[2] / len;
}

static int vp_varyings(int vp) {
    return vp == SR_VP_WORLD4_NORMAL4 ? 8 : vp == SR_VP_MVP_INT ? 1 : (vp == SR_VP_MVP_NORMAL_WORLD ? 6 : 5);
}

/* 4-vectors of the full example (nalgebra Vector4<f32>): dot sums from +0
 * in index order; norm = sqrt(dot(v, v)); normalize = v / norm. */
static inline float dot4(const float* a, const float* b) {
    float acc = 0.0f;
    for (int c = 0; c < 4; c++) acc = acc + a[c] * b[c];
    return acc;
}
static inline void normalize4(float* v) {
    float n = sqrtf(dot4(v, v));
    for (int c = 0; c < 4; c++) v[c] = v[c] / n;
}

static void vs_run(const sr_draw_desc* d, const or_mesh* m, uint32_t i, ovtx* o) {
    const float* U = d->vs_uniforms;
    float pin[4] = {m->v.px[i], m->v.py[i], m->v.pz[i], 1.0f};
    float nin[3] = {m->v.nx ? m->v.nx[i] : 0.0f, m->v.ny ? m->v.ny[i] : 0.0f,
                    m->v.nz ? m->v.nz[i] : 0.0f};
    for (int r = 0; r < 4; r++) {
        float acc = 0.0f;
        for (int k = 0; k < 4; k++) acc = acc + U[r * 4 + k] * pin[k];
        o->p[r] = acc;
    }
    if (d->vertex_program == SR_VP_MVP_INT) {
        /* an i32 member of the uniforms; interpolate.rs:47-49 converts each
         * vertex's value with R::from (i32 -> f32, rounded to nearest) */
        o->k[0] = m->iattr ? (float)m->iattr[i] : 0.0f;
        return;
    }
    if (d->vertex_program == SR_VP_WORLD4_NORMAL4) {
        /* full_example/src/shaders.rs:8-31: world_position = model * p and
         * normal = (mit * n.to_homogeneous()).normalize(), 4-vectors; model
         * rows 0-2 at U[25..37), row 3 at U[40..44); mit rows 0-2 (columns
         * 0-2) at U[16..25), row 3 at U[37..40) (column 3 meets w = 0). */
        for (int r = 0; r < 4; r++) {
            const float* row = r < 3 ? U + 25 + r * 4 : U + 40;
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc = acc + row[k] * pin[k];
            o->k[r] = acc;
        }
        float n4[4];
        for (int r = 0; r < 4; r++) {
            const float* row = r < 3 ? U + 16 + r * 3 : U + 37;
            float acc = 0.0f;
            for (int k = 0; k < 3; k++) acc = acc + row[k] * nin[k];
            n4[r] = acc;
        }
        normalize4(n4);
        for (int c = 0; c < 4; c++) o->k[4 + c] = n4[c];
        return;
    }
    float n[3];
    for (int r = 0; r < 3; r++) {
        float acc = 0.0f;
        for (int k = 0; k < 3; k++) acc = acc + U[16 + r * 3 + k] * nin[k];
        n[r] = acc;
    }
    normalize3(n);
    o->k[0] = n[0];
    o->k[1] = n[1];
    o->k[2] = n[2];
    if (d->vertex_program == SR_VP_MVP_NORMAL_WORLD) {
        for (int r = 0; r < 3; r++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc = acc + U[25 + r * 4 + k] * pin[k];
            o->k[3 + r] = acc;
        }
    } else {
        o->k[3] = m->v.u ? m->v.u[i] : 0.0f;
        o->k[4] = m->v.v ? m->v.v[i] : 0.0f;
    }
}

int or_vertex(const sr_draw_desc* d, const or_mesh* m, uint32_t i, float clip[4], float* k) {
    ovtx o;
    memset(&o, 0, sizeof(o));
    vs_run(d, m, i, &o);
    memcpy(clip, o.p, sizeof(o.p));
    const int K = vp_varyings(d->vertex_program);
    memcpy(k, o.k, sizeof(float) * K);
    return K;
}

/* ------------------------------------------------------------------------ */
/* ClipVertex::normalize (geometry/clipvertex.rs:89-126).                    */
/* ------------------------------------------------------------------------ */
static void normalize_pos(const float in[4], const sr_viewport* vp, float out[4]) {
    float x = in[0], y = in[1], z = in[2], w = in[3];
    float left = vp->x, bottom = vp->y;
    float right = left + vp->width;   /* :100 */
    float top = bottom + vp->height;  /* :101 */
    float nr = vp->near_z, fr = vp->far_z;
    /* viewport_matrix, :107-112 */
    float M[4][4] = {
        {(right - left) / 2.0f, 0.0f, 0.0f, (right + left) / 2.0f},
        {0.0f, (top - bottom) / -2.0f, 0.0f, (top + bottom) / 2.0f},
        {0.0f, 0.0f, (fr - nr) / -2.0f, (fr + nr) / -2.0f},
        {0.0f, 0.0f, 0.0f, 1.0f},
    };
    float v[4] = {x / w, y / w, z / w, 1.0f}; /* :114-119 */
    for (int r = 0; r < 3; r++) {
        float acc = 0.0f;
        for (int k = 0; k < 4; k++) acc = acc + M[r][k] * v[k];
        out[r] = acc;
    }
    out[3] = 1.0f / w; /* :121 */
}

void or_normalize(const float clip[4], const sr_viewport* vp, float out[4]) {
    normalize_pos(clip, vp, out);
}

/* ------------------------------------------------------------------------ */
/* ClippingPlane (geometry/clip.rs:12-62)                                    */
/* ------------------------------------------------------------------------ */
enum { PL_LEFT, PL_RIGHT, PL_TOP, PL_BOTTOM, PL_NEAR, PL_FAR };

static inline int plane_inside(int pl, const ovtx* v) { /* clip.rs:33-44 */
    float x = v->p[0], y = v->p[1], z = v->p[2], w = v->p[3];
    switch (pl) {
        case PL_LEFT: return x >= -w;
        case PL_RIGHT: return x <= w;
        case PL_TOP: return y >= -w;
        case PL_BOTTOM: return y <= w;
        case PL_NEAR: return z >= 0.0f;
        default: return z <= w;
    }
}

static void plane_intersect(int pl, const ovtx* v1, const ovtx* v2, int K, ovtx* o) {
    /* clip.rs:47-62 */
    float x1 = v1->p[0], y1 = v1->p[1], z1 = v1->p[2], w1 = v1->p[3];
    float x2 = v2->p[0], y2 = v2->p[1], z2 = v2->p[2], w2 = v2->p[3];
    float a, b;
    switch (pl) {
        case PL_LEFT: a = w1 + x1; b = w2 + x2; break;
        case PL_RIGHT: a = w1 - x1; b = w2 - x2; break;
        case PL_TOP: a = w1 + y1; b = w2 + y2; break;
        case PL_BOTTOM: a = w1 - y1; b = w2 - y2; break;
        case PL_NEAR: a = z1; b = z2; break;
        default: a = w1 - z1; b = w2 - z2; break;
    }
    float t = a / (a - b);
    lerp_vtx(t, v1, v2, K, o);
}

/* Geometry outputs: flat vertex lists per kind (storage.rs:8-50). */
typedef struct {
    ovtx* v;
    size_t n, cap;
} vlist;

static void vl_push(vlist* l, const ovtx* v) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 1024;
        l->v = (ovtx*)realloc(l->v, l->cap * sizeof(ovtx));
    }
    l->v[l->n++] = *v;
}

/* clip_primitives, triangle branch (geometry.rs:264-299). */
static void clip_tri_reference(const ovtx* a, const ovtx* b, const ovtx* c, int K, vlist* tris) {
    ovtx poly[40];
    int n = 0;
    const ovtx* edges[3][2] = {{a, b}, {b, c}, {c, a}};
    for (int e = 0; e < 3; e++) {
        const ovtx* s = edges[e][0];
        const ovtx* p = edges[e][1];
        for (int pl = 0; pl < 6; pl++) {
            int s_in = plane_inside(pl, s), p_in = plane_inside(pl, p);
            if (s_in != p_in) plane_intersect(pl, s, p, K, &poly[n++]);
            if (p_in) poly[n++] = *p;
        }
    }
    if (n == 3) {
        for (int i = 0; i < 3; i++) vl_push(tris, &poly[i]);
    } else if (n > 3) {
        const ovtx* last = &poly[n - 1];
        for (int i = 0; i < n - 2; i++) { /* emit_triangle(last, poly[i], poly[i+1]) */
            vl_push(tris, last);
            vl_push(tris, &poly[i]);
            vl_push(tris, &poly[i + 1]);
        }
    }
}

/* Build-defined near-plane clip: the same loop and fan as clip_primitives
 * (geometry.rs:270-298) restricted to ClippingPlane::Near, i.e.
 * Sutherland-Hodgman against z >= 0.  A triangle yields 0, 1 or 2. */
static void clip_tri_near(const ovtx* a, const ovtx* b, const ovtx* c, int K, vlist* tris) {
    ovtx poly[8];
    int n = 0;
    const ovtx* edges[3][2] = {{a, b}, {b, c}, {c, a}};
    for (int e = 0; e < 3; e++) {
        const ovtx* s = edges[e][0];
        const ovtx* p = edges[e][1];
        int s_in = plane_inside(PL_NEAR, s), p_in = plane_inside(PL_NEAR, p);
        if (s_in != p_in) plane_intersect(PL_NEAR, s, p, K, &poly[n++]);
        if (p_in) poly[n++] = *p;
    }
    if (n == 3) {
        for (int i = 0; i < 3; i++) vl_push(tris, &poly[i]);
    } else if (n > 3) {
        const ovtx* last = &poly[n - 1];
        for (int i = 0; i < n - 2; i++) {
            vl_push(tris, last);
            vl_push(tris, &poly[i]);
            vl_push(tris, &poly[i + 1]);
        }
    }
}

/* clip_primitives, line branch (geometry.rs:300-328). */
static int clip_line_reference(const ovtx* s0, const ovtx* e0, int K, ovtx* so, ovtx* eo) {
    ovtx s = *s0, e = *e0;
    int intersections = 0;
    for (int pl = 0; pl < 6; pl++) {
        int s_in = plane_inside(pl, &s), p_in = plane_inside(pl, &e);
        if (s_in != p_in) {
            ovtx I;
            plane_intersect(pl, &s, &e, K, &I);
            if (s_in) e = I;
            else if (p_in) s = I;
            intersections += 1;
        } else if (!s_in) {
            return 0;
        }
        if (intersections > 2) break;
    }
    *so = s;
    *eo = e;
    return 1;
}

/* clip_primitives, point branch (geometry.rs:329-333). */
static int clip_point_reference(const ovtx* p) {
    for (int pl = 0; pl < 6; pl++)
        if (!plane_inside(pl, p)) return 0;
    return 1;
}

static void to_ovtx(const float* f, int K, ovtx* o) {
    memset(o, 0, sizeof(*o));
    for (int i = 0; i < 4; i++) o->p[i] = f[i];
    for (int i = 0; i < K; i++) o->k[i] = f[4 + i];
}
static void from_ovtx(const ovtx* o, int K, float* f) {
    for (int i = 0; i < 4; i++) f[i] = o->p[i];
    for (int i = 0; i < K; i++) f[4 + i] = o->k[i];
}

static int clip_tri_entry(int near_only, const float* a, const float* b, const float* c, int K,
                          float* out, int max_tris) {
    ovtx va, vb, vc;
    to_ovtx(a, K, &va);
    to_ovtx(b, K, &vb);
    to_ovtx(c, K, &vc);
    vlist l = {0};
    if (near_only) clip_tri_near(&va, &vb, &vc, K, &l);
    else clip_tri_reference(&va, &vb, &vc, K, &l);
    int nt = (int)(l.n / 3);
    for (int i = 0; i < nt && i < max_tris; i++)
        for (int j = 0; j < 3; j++) from_ovtx(&l.v[i * 3 + j], K, out + (size_t)(i * 3 + j) * (4 + K));
    free(l.v);
    return nt;
}
int or_clip_triangle_reference(const float* a, const float* b, const float* c, int K,
                               float* out, int max_tris) {
    return clip_tri_entry(0, a, b, c, K, out, max_tris);
}
int or_clip_triangle_near(const float* a, const float* b, const float* c, int K, float* out,
                          int max_tris) {
    return clip_tri_entry(1, a, b, c, K, out, max_tris);
}
int or_clip_line_reference(const float* s, const float* e, int K, float* out) {
    ovtx vs, ve, so, eo;
    to_ovtx(s, K, &vs);
    to_ovtx(e, K, &ve);
    if (!clip_line_reference(&vs, &ve, K, &so, &eo)) return 0;
    from_ovtx(&so, K, out);
    from_ovtx(&eo, K, out + 4 + K);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Stencil (stencil.rs:109-159).  `test(value=framebuffer, mask=ref)`.       */
/* ------------------------------------------------------------------------ */
int or_stencil_test(int test, uint32_t value, uint32_t mask) {
    switch (test) {
        case SR_STENCIL_ALWAYS: return 1;
        case SR_STENCIL_NEVER: return 0;
        case SR_STENCIL_LESS: return mask < value;
        case SR_STENCIL_LEQUAL: return mask <= value;
        case SR_STENCIL_GREATER: return mask > value;
        case SR_STENCIL_GEQUAL: return mask >= value;
        case SR_STENCIL_EQUAL: return mask == value;
        default: return mask != value;
    }
}
uint32_t or_stencil_op(int op, int wrap, uint32_t value, uint32_t mask) {
    value &= 0xFFu;
    switch (op) {
        case SR_STENCIL_KEEP: return value;
        case SR_STENCIL_INVERT: return (~value) & 0xFFu;
        case SR_STENCIL_ZERO: return 0;
        case SR_STENCIL_REPLACE: return mask & 0xFFu;
        case SR_STENCIL_INCR: return wrap ? ((value + 1u) & 0xFFu) : (value == 255u ? 255u : value + 1u);
        default: return wrap ? ((value - 1u) & 0xFFu) : (value == 0u ? 0u : value - 1u);
    }
}

/* ------------------------------------------------------------------------ */
/* Colour helpers, shading programs, sampler and blend.  The reference       */
/* library defines none of these (texture sampling is unimplemented!(),     */
/* texture.rs:48-52; blend is a user fn, blend.rs:232-274); they are         */
/* build-defined and written only with IEEE +,-,*,/,sqrt, min/max and floor, */
/* which are reproducible bit-for-bit on the device.                         */
/* ------------------------------------------------------------------------ */
static inline uint32_t quant8(float x) {
    float c = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint32_t)(c * 255.0f + 0.5f);
}
static inline float unorm8(uint32_t c) { return (float)(c & 0xFFu) / 255.0f; }
static inline uint32_t pack4(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return r | (g << 8) | (b << 16) | (a << 24);
}

/* One texel as RGBA floats: RGBA8 texels as unorm (c / 255), RGBAF32
 * texels (a texture buffer plane, texturebuffer.rs:12-49) as stored; the
 * border colour is RGBA8 either way. */
static void texel(const or_texture* t, int ix, int iy, int edge, uint32_t border, float o[4]) {
    int W = (int)t->w, H = (int)t->h;
    int use_border = 0;
    if (edge == SR_EDGE_WRAP) {
        ix = ((ix % W) + W) % W;
        iy = ((iy % H) + H) % H;
    } else if (edge == SR_EDGE_BORDER) {
        use_border = ix < 0 || iy < 0 || ix >= W || iy >= H;
    } else {
        ix = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix);
        iy = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy);
    }
    if (!use_border && !t->rgba) {
        const float* f = t->rgbaf + 4 * ((size_t)iy * W + ix);
        for (int c = 0; c < 4; c++) o[c] = f[c];
        return;
    }
    uint32_t c = use_border ? border : t->rgba[(size_t)iy * W + ix];
    o[0] = unorm8(c);
    o[1] = unorm8(c >> 8);
    o[2] = unorm8(c >> 16);
    o[3] = unorm8(c >> 24);
}

static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* float -> u32 as Rust's `as` casts it: NaN and negatives to 0, saturating. */
static inline uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

/* RgbaImage::get_pixel(x, y) as a Color (full_example/src/texture.rs:38-45:
 * c / 255; RGBAF32 texels as stored).  get_pixel panics past the image;
 * here such an index takes the last texel. */
static void span_texel(const or_texture* t, uint32_t x, uint32_t y, float o[4]) {
    if (x > t->w - 1) x = t->w - 1;
    if (y > t->h - 1) y = t->h - 1;
    size_t i = (size_t)y * t->w + x;
    if (!t->rgba) {
        for (int c = 0; c < 4; c++) o[c] = t->rgbaf[4 * i + c];
        return;
    }
    uint32_t w = t->rgba[i];
    for (int c = 0; c < 4; c++) o[c] = unorm8(w >> (8 * c));
}

/* The full example's Texture::sample (full_example/src/texture.rs:47-84),
 * EdgeBehavior::edge (texture.rs:27-34: Clamp u.min(1).max(0), Wrap
 * u.fract() = u - trunc(u)), then decode_gamma(texel, SRGB_GAMMA)
 * (color.rs:47-56: r, g, b = c.powf(2.2)). */
static void sample_span(const or_texture* t, float u, float v, int filter, int edge, float out[4]) {
    if (edge == SR_EDGE_WRAP) {
        u = u - truncf(u);
        v = v - truncf(v);
    } else {
        u = fmaxf(fminf(u, 1.0f), 0.0f);
        v = fmaxf(fminf(v, 1.0f), 0.0f);
    }
    const float w1 = (float)(t->w - 1), h1 = (float)(t->h - 1);
    if (filter == SR_FILTER_SPAN_NEAREST) {  /* texture.rs:51-56 */
        span_texel(t, f2u_sat(roundf(u * w1)), f2u_sat(roundf(v * h1)), out);
    } else {                                  /* texture.rs:57-80 */
        const float uu = (u * w1) + 0.5f, vv = (v * h1) + 0.5f;
        const uint32_t x = f2u_sat(floorf(uu)), y = f2u_sat(floorf(vv));
        const float ur = uu - (float)x, vr = vv - (float)y;
        const float uo = 1.0f - ur, vo = 1.0f - vr;
        float t00[4], t10[4], t01[4], t11[4];
        span_texel(t, x, y, t00);
        span_texel(t, x + 1, y, t10);
        span_texel(t, x, y + 1, t01);
        span_texel(t, x + 1, y + 1, t11);
        for (int c = 0; c < 4; c++)
            out[c] = (((t00[c] * uo) + (t10[c] * ur)) * vo) + (((t01[c] * uo) + (t11[c] * ur)) * vr);
    }
    for (int c = 0; c < 3; c++) out[c] = powf(out[c], 2.2f);
}

/* Texel centres at (i+0.5)/W (GL convention); Filter/Edge per texture.rs:20-40. */
void or_sample(const or_texture* t, float u, float v, int filter, int edge, uint32_t border,
               float out[4]) {
    if (filter >= SR_FILTER_SPAN_NEAREST) {
        sample_span(t, u, v, filter, edge, out);
        return;
    }
    if (filter == SR_FILTER_NEAREST) {
        float fx = clampf(u * (float)t->w, -1.0e8f, 1.0e8f);
        float fy = clampf(v * (float)t->h, -1.0e8f, 1.0e8f);
        texel(t, (int)floorf(fx), (int)floorf(fy), edge, border, out);
        return;
    }
    float fx = clampf(u * (float)t->w - 0.5f, -1.0e8f, 1.0e8f);
    float fy = clampf(v * (float)t->h - 0.5f, -1.0e8f, 1.0e8f);
    float x0 = floorf(fx), y0 = floorf(fy);
    float ax = fx - x0, ay = fy - y0;
    int ix = (int)x0, iy = (int)y0;
    float t00[4], t10[4], t01[4], t11[4];
    texel(t, ix, iy, edge, border, t00);
    texel(t, ix + 1, iy, edge, border, t10);
    texel(t, ix, iy + 1, edge, border, t01);
    texel(t, ix + 1, iy + 1, edge, border, t11);
    float bx = 1.0f - ax, by = 1.0f - ay;
    for (int c = 0; c < 4; c++) {
        float top = (t00[c] * bx) + (t10[c] * ax);
        float bot = (t01[c] * bx) + (t11[c] * ax);
        out[c] = (top * by) + (bot * ay);
    }
}

/* powi by repeated squaring (r *= x for each set bit, x squared between:
 * x^5 = x * x^4), the order compiler-rt's __powisf2 multiplies in. */
static float powi_sq(float x, int n) {
    float r = 1.0f;
    for (;;) {
        if (n & 1) r = r * x;
        n >>= 1;
        if (n == 0) break;
        x = x * x;
    }
    return r;
}
static inline float saturate1(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); } /* shaders.rs:101-103 */

/* The full example's fragment_shader (full_example/src/shaders.rs:105-161)
 * with fresnel_schlick (shaders.rs:94-99), aces_filmic_tonemap
 * (color.rs:19-38) and encode_gamma (color.rs:40-47).  k = world position
 * (4) then world normal (4); U: camera xyz, four lights (position xyz,
 * colour rgb, intensity) at U[3 + 7i], U[31] the grey of the material. */
static void fresnel4(const float* U, const float* k, float out[4]) {
    const float* pos = k;
    const float* nrm = k + 4;
    float view[4] = {U[0] - pos[0], U[1] - pos[1], U[2] - pos[2], 1.0f - pos[3]};
    normalize4(view);
    const float mat = powf(U[31], 2.2f);  /* decode_gamma(Color {0.25, ..}, SRGB_GAMMA) */
    const float albedo = 0.7f, ior = 1.45f;
    float col[3] = {0.0f, 0.0f, 0.0f};
    for (int l = 0; l < 4; l++) {
        const float* L = U + 3 + 7 * l;
        float d[4] = {L[0] - pos[0], L[1] - pos[1], L[2] - pos[2], 1.0f - pos[3]};
        const float dist = sqrtf(dot4(d, d));
        normalize4(d);
        float h[4] = {d[0] + view[0], d[1] + view[1], d[2] + view[2], d[3] + view[3]};
        normalize4(h);
        const float intensity = L[6] / powi_sq(dist, 2);
        const float ndl = saturate1(dot4(d, nrm));
        const float ndh = saturate1(dot4(nrm, h));
        const float vdh = saturate1(dot4(view, h));
        const float f0 = powi_sq((1.0f - ior) / (1.0f + ior), 2);
        const float f = f0 + (1.0f - f0) * powi_sq(1.0f - vdh, 5);
        const float diffuse = (1.0f - f) * ndl;
        const float specular = f * powi_sq(ndh, 64);  /* shininess * 2 */
        for (int c = 0; c < 3; c++)
            col[c] = col[c] + (intensity * L[3 + c]) * (specular + ((diffuse * albedo) * mat));
    }
    const float inv_gamma = 1.0f / 2.2f;
    for (int c = 0; c < 3; c++) {
        const float x = col[c];
        const float t = (x * ((2.51f * x) + 0.03f)) / ((x * ((2.43f * x) + 0.59f)) + 0.14f);
        out[c] = powf(t, inv_gamma);
    }
    out[3] = 1.0f;
}

/* Fragment programs; k = interpolated varyings.  Each program's colours
 * before quantisation: out[i] is output i (RGBA floats); returns how many
 * outputs the program has (SR_FP_GBUFFER: 4, the others 1).  An RGBA8
 * plane stores quant8 of each channel, an RGBAF32 plane the floats. */
static int fs_run_f(const sr_draw_desc* d, const float* k, const or_texture* tex, float out[4][4]) {
    const float* U = d->fs_uniforms;
    switch (d->fragment_program) {
        case SR_FP_BLINN_PHONG: {
            /* U: L[0..3) eye[3..6) base[6..9) ambient[9] kd[10] ks[11] alpha[12] */
            float n[3] = {k[0], k[1], k[2]};
            normalize3(n);
            float v[3] = {U[3] - k[3], U[4] - k[4], U[5] - k[5]};
            normalize3(v);
            float ndl = dot3(n, U);
            float diff = ndl > 0.0f ? ndl : 0.0f;
            float h[3] = {U[0] + v[0], U[1] + v[1], U[2] + v[2]};
            normalize3(h);
            float ndh = dot3(n, h);
            float s = ndh > 0.0f ? ndh : 0.0f;
            s = s * s; s = s * s; s = s * s; s = s * s; s = s * s; /* ^32 */
            float lit = U[9] + U[10] * diff;
            float spec = U[11] * s;
            out[0][0] = U[6] * lit + spec; out[0][1] = U[7] * lit + spec;
            out[0][2] = U[8] * lit + spec; out[0][3] = U[12];
            return 1;
        }
        case SR_FP_TEX_LAMBERT:
        case SR_FP_GBUFFER: {
            /* U: L[0..3) ambient[3] kd[4] alpha[5] */
            float n[3] = {k[0], k[1], k[2]};
            normalize3(n);
            float ndl = dot3(n, U);
            float diff = ndl > 0.0f ? ndl : 0.0f;
            float lit = U[3] + U[4] * diff;
            float t[4];
            or_sample(tex, k[3], k[4], d->tex_filter, d->tex_edge, d->tex_border, t);
            out[0][0] = t[0] * lit; out[0][1] = t[1] * lit; out[0][2] = t[2] * lit; out[0][3] = t[3] * U[5];
            if (d->fragment_program == SR_FP_TEX_LAMBERT) return 1;
            out[1][0] = n[0] * 0.5f + 0.5f; out[1][1] = n[1] * 0.5f + 0.5f;
            out[1][2] = n[2] * 0.5f + 0.5f; out[1][3] = 1.0f;
            out[2][0] = k[3]; out[2][1] = k[4]; out[2][2] = diff; out[2][3] = 1.0f;
            out[3][0] = lit; out[3][1] = lit; out[3][2] = lit; out[3][3] = 1.0f;
            return 4;
        }
        case SR_FP_UV_COLOR:
            out[0][0] = k[3]; out[0][1] = k[4]; out[0][2] = U[0]; out[0][3] = U[1];
            return 1;
        case SR_FP_FRESNEL4:
            fresnel4(U, k, out[0]);
            return 1;
        case SR_FP_INT_PALETTE: {
            /* <i32 as NumCast>::from(f32) (interpolate.rs:47-49, num-traits
             * 0.1): truncation toward zero; NaN and out-of-range values, where
             * the reference's unwrap panics, saturate as Rust's `as` does
             * (parity unpinned there: the crate would abort). */
            const float x = k[0];
            const int32_t id = x != x ? 0 : (x >= 2147483648.0f ? INT32_MAX : (x < -2147483648.0f ? INT32_MIN : (int32_t)x));
            const int32_t m = ((id % 8) + 8) % 8;
            for (int c = 0; c < 4; c++) out[0][c] = U[4 * m + c];
            return 1;
        }
        default:
            out[0][0] = U[0]; out[0][1] = U[1]; out[0][2] = U[2]; out[0][3] = U[3];
            return 1;
    }
}

/* Output 0 of the fragment program on varyings k (RGBA floats before any
 * quantisation or blend): the programs on their own, for the tests. */
void or_fragment(const sr_draw_desc* d, const float* k, const or_texture* tex, float out[4]) {
    float o[4][4];
    fs_run_f(d, k, tex, o);
    for (int c = 0; c < 4; c++) out[c] = o[0][c];
}

static inline uint32_t pack_q(const float c[4]) {
    return pack4(quant8(c[0]), quant8(c[1]), quant8(c[2]), quant8(c[3]));
}

static uint32_t fs_run(const sr_draw_desc* d, const float* k, const or_texture* tex) {
    float out[4][4];
    fs_run_f(d, k, tex, out);
    return pack_q(out[0]);
}

/* Blend::blend(src, dst) (color/blend.rs:9-31) on RGBA floats: `()`
 * returns src; OVER is the build-defined src-alpha over; OVER_NORMALIZED
 * is the full example's over (full_example/src/color.rs:5-17), Rust's
 * left-to-right evaluation: ((x * a) + ((y * b) * a1)) / (a + (b * a1)). */
void or_blend_f32(int blend, const float s[4], const float d[4], float o[4]) {
    if (blend == SR_BLEND_OVER) {
        float a = s[3], ia = 1.0f - a;
        for (int c = 0; c < 3; c++) o[c] = (s[c] * a) + (d[c] * ia);
        o[3] = a + (d[3] * ia);
    } else if (blend == SR_BLEND_OVER_NORMALIZED) {
        float a = s[3], b = d[3], a1 = 1.0f - a;
        for (int c = 0; c < 3; c++) o[c] = ((s[c] * a) + ((d[c] * b) * a1)) / (a + (b * a1));
        o[3] = a + (b * (1.0f - a));
    } else {
        for (int c = 0; c < 4; c++) o[c] = s[c];
    }
}

/* The blends on RGBA8: OVER as below, OVER_NORMALIZED on unorm floats,
 * quantised back. */
uint32_t or_blend(int blend, uint32_t s, uint32_t dst) {
    if (blend == SR_BLEND_OVER_NORMALIZED) {
        float fs[4] = {unorm8(s), unorm8(s >> 8), unorm8(s >> 16), unorm8(s >> 24)};
        float fd[4] = {unorm8(dst), unorm8(dst >> 8), unorm8(dst >> 16), unorm8(dst >> 24)};
        float o[4];
        or_blend_f32(blend, fs, fd, o);
        return pack_q(o);
    }
    if (blend != SR_BLEND_OVER) return s;
    float a = unorm8(s >> 24), ia = 1.0f - a;
    uint32_t r = quant8((unorm8(s) * a) + (unorm8(dst) * ia));
    uint32_t g = quant8((unorm8(s >> 8) * a) + (unorm8(dst >> 8) * ia));
    uint32_t b = quant8((unorm8(s >> 16) * a) + (unorm8(dst >> 16) * ia));
    uint32_t oa = quant8(a + (unorm8(dst >> 24) * ia));
    return pack4(r, g, b, oa);
}

/* c.mul_alpha(ColorAlpha::from_scalar(alpha)) (line.rs:100).
 * SR_ALPHA_U8: Vector4<u8> (color/predefined.rs:68-90): from_scalar is
 *   NumCast f64->u8 (truncation; out of [0,255] the reference panics —
 *   parity unpinned there, treated as 0), then AlphaMultiply for u8
 *   (helper.rs:31-42): (w as f32 * (alpha as f32 / 255.0)) as u8.
 * SR_ALPHA_F32: #[repr(C)] RGBA8 with Alpha = f32: from_scalar = (f32)alpha,
 *   a' = (a as f32 * alpha) as u8 (saturating cast). */
static inline uint32_t sat_u8(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint32_t)f;
}
uint32_t or_mul_alpha(int mode, uint32_t c, double alpha) {
    uint32_t a = c >> 24;
    uint32_t na;
    if (mode == SR_ALPHA_F32) {
        float af = (float)alpha;
        na = sat_u8((float)a * af);
    } else {
        uint32_t au8 = (alpha >= 0.0 && alpha <= 255.0) ? (uint32_t)alpha : 0u;
        na = sat_u8((float)a * ((float)au8 / 255.0f));
    }
    return (c & 0x00FFFFFFu) | (na << 24);
}

/* Rust f32::hypot -> the platform libm hypotf (glibc here). */
float or_hypotf(float x, float y) { return hypotf(x, y); }
float or_div_f32(float a, float b) { return a / b; }

/* ------------------------------------------------------------------------ */
/* Tile grid (pipeline/stages/fragment.rs:188-216): inclusive rectangles.   */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t x0, y0, x1, y1;
} otile;

static int make_tiles(uint32_t W, uint32_t H, uint32_t tw, uint32_t th, otile* out, int max) {
    int n = 0;
    if (W == 0 || H == 0) return 0;
    uint32_t xmax = W - 1, ymax = H - 1, y = 0;
    while (y < ymax) {
        uint32_t x = 0;
        uint32_t ny = y + th < ymax ? y + th : ymax; /* min(y + h, ymax) */
        while (x < xmax) {
            uint32_t nx = x + tw < xmax ? x + tw : xmax;
            if (out && n < max) {
                out[n].x0 = x; out[n].y0 = y; out[n].x1 = nx; out[n].y1 = ny;
            }
            n++;
            x = nx;
        }
        y = ny;
    }
    return n;
}

int or_tiles(uint32_t w, uint32_t h, uint32_t tw, uint32_t th, uint32_t* out, int max) {
    int n = make_tiles(w, h, tw, th, NULL, 0);
    otile* t = (otile*)malloc(sizeof(otile) * (n > 0 ? n : 1));
    make_tiles(w, h, tw, th, t, n);
    for (int i = 0; i < n && i < max; i++) {
        out[i * 4 + 0] = t[i].x0; out[i * 4 + 1] = t[i].y0;
        out[i * 4 + 2] = t[i].x1; out[i * 4 + 3] = t[i].y1;
    }
    free(t);
    return n;
}

/* ------------------------------------------------------------------------ */
/* Draw state shared by the rasterisers                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    or_fb* fb;
    const sr_draw_desc* d;
    const or_texture* tex;
    int K;
} rstate;

/* Fragment tail shared by all three rasterisers: depth (triangle.rs:117-147,
 * line.rs:79-106, point.rs:300-322).  Depth attachment is f32, so
 * Depth::from_scalar is the identity (depth.rs:26-41). */
/* shade_write on a framebuffer of another format (or_fb.n_color > 0):
 * Depth::from_scalar is exact for f64, the program's outputs go to the
 * planes (outputs past the program's count repeat its last), each blended
 * in its own format. */
static void shade_write_ex(rstate* rs, size_t index, float z, const float* k, uint64_t* shaded) {
    or_fb* fb = rs->fb;
    if (!(z < 0.0f)) return;
    const double dt = fb->depth_fmt == SR_DEPTH_F64 ? fb->depth64[index] : (double)fb->depth[index];
    if (!((double)z >= dt)) return;
    float out[4][4];
    const int n = fs_run_f(rs->d, k, rs->tex, out);
    (*shaded)++;
    for (uint32_t p = 0; p < fb->n_color; p++) {
        const float* c = out[(int)p < n ? (int)p : n - 1];
        if (fb->color_fmt[p] == SR_COLOR_RGBAF32) {
            float* dst = (float*)fb->planes[p] + 4 * index;
            float o[4];
            or_blend_f32(rs->d->blend, c, dst, o);
            for (int q = 0; q < 4; q++) dst[q] = o[q];
        } else {
            uint32_t* dst = (uint32_t*)fb->planes[p] + index;
            *dst = or_blend(rs->d->blend, pack_q(c), *dst);
        }
    }
    if (fb->depth_fmt == SR_DEPTH_F64) fb->depth64[index] = (double)z;
    else fb->depth[index] = z;
}

static inline void shade_write(rstate* rs, size_t index, float z, const float* k, int is_line,
                               double alpha, uint64_t* shaded) {
    if (rs->fb->n_color) {
        shade_write_ex(rs, index, z, k, shaded);
        return;
    }
    if (z < 0.0f) {
        float dt = rs->fb->depth[index];
        if (z >= dt) {
            uint32_t c = fs_run(rs->d, k, rs->tex);
            (*shaded)++;
            if (is_line) c = or_mul_alpha(rs->d->alpha_mode, c, alpha);
            uint32_t p = rs->fb->color[index];
            rs->fb->color[index] = or_blend(rs->d->blend, c, p);
            rs->fb->depth[index] = z;
        }
    }
}

/* clamp_as_int! (triangle.rs:66-72): value < min ? min : value > max ? max
 * : cast(value).  cast of NaN panics in the reference. */
static inline uint32_t clamp_as_int(float v, uint32_t lo, uint32_t hi, int* bad) {
    if (v < (float)lo) return lo;
    if (v > (float)hi) return hi;
    if (v != v) { *bad = 1; return lo; }
    return (uint32_t)v;
}

/* rasterize_triangle (pipeline/stages/rasterization/triangle.rs:23-156). */
static void raster_tri(rstate* rs, const otile* tile, const ovtx* a, const ovtx* b, const ovtx* c,
                       or_stats* st) {
    const sr_draw_desc* d = rs->d;
    or_fb* fb = rs->fb;
    st->tri_calls++;
    float x1 = a->p[0], y1 = a->p[1];
    float x2 = b->p[0], y2 = b->p[1];
    float x3 = c->p[0], y3 = c->p[1];
    if (d->cull != SR_CULL_NONE) { /* :54-61 */
        float A = x1 * y2 + x2 * y3 + x3 * y1 - x2 * y1 - x3 * y2 - x1 * y3;
        int wind = signbit(A) ? SR_CULL_CW : SR_CULL_CCW;
        if (wind == d->cull) { st->tri_culled++; return; }
    }
    float det = (y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3); /* :64 */
    int bad = 0;
    /* :74-78 — f32 min/max are IEEE minNum/maxNum (fminf/fmaxf) */
    uint32_t mnx = clamp_as_int(fminf(fminf(x1, x2), x3), tile->x0, tile->x1, &bad);
    uint32_t mny = clamp_as_int(fminf(fminf(y1, y2), y3), tile->y0, tile->y1, &bad);
    uint32_t mxx = clamp_as_int(fmaxf(fmaxf(x1, x2), x3), tile->x0, tile->x1, &bad);
    uint32_t mxy = clamp_as_int(fmaxf(fmaxf(y1, y2), y3), tile->y0, tile->y1, &bad);
    if (bad) { st->invalid++; return; }
    /* unclamped integer bbox, only to classify strip pixels for the stats */
    float fmnx = fminf(fminf(x1, x2), x3), fmxx = fmaxf(fmaxf(x1, x2), x3);
    float fmny = fminf(fminf(y1, y2), y3), fmxy = fmaxf(fmaxf(y1, y2), y3);
    const uint32_t W = fb->w;
    const uint32_t sref = d->stencil_ref & 0xFFu;
    const int K = rs->K;
    for (uint32_t py = mny; py <= mxy; py++) { /* :82 */
        for (uint32_t px = mnx; px <= mxx; px++) {
            size_t index = (size_t)px + (size_t)py * W; /* coordinate.rs:47-52 */
            st->bbox_tests++;
            float pcx = (float)px, pcy = (float)py;
            int strip = !(pcx + 1.0f > fmnx && pcx <= fmxx && pcy + 1.0f > fmny && pcy <= fmxy);
            if (strip) st->strip_tests++;
            if (!(fb->n_color && fb->stencil_fmt == SR_STENCIL_NONE)) { /* `()`: no test, no op */
                uint32_t sv = fb->stencil[index];
                if (!or_stencil_test(d->stencil_test, sv, sref)) continue; /* :94 */
                fb->stencil[index] = (uint8_t)or_stencil_op(d->stencil_op, d->stencil_wrap, sv, sref);
            }
            float x = (float)px + 0.5f, y = (float)py + 0.5f; /* :104-105 */
            float u = ((y2 - y3) * (x - x3) + (x3 - x2) * (y - y3)) / det; /* :108 */
            float v = ((y3 - y1) * (x - x3) + (x1 - x3) * (y - y3)) / det; /* :109 */
            float w = 1.0f - u - v;                                          /* :110 */
            if (!(u < 0.0f || v < 0.0f || w < 0.0f)) {                      /* :113 */
                st->covered++;
                float z = bary1(u, a->p[2], v, b->p[2], w, c->p[2]); /* :115-117 */
                if (z < 0.0f && z >= fb->depth[index]) {
                    float k[OR_MAXK];
                    for (int i = 0; i < K; i++) k[i] = bary1(u, a->k[i], v, b->k[i], w, c->k[i]);
                    if (strip) st->strip_shaded++;
                    shade_write(rs, index, z, k, 0, 1.0, &st->shaded);
                }
            }
        }
    }
}

/* liang_barsky_iterative (geometry/line.rs:6-54). */
static int liang_barsky(float x1, float y1, float x2, float y2, float xmin, float ymin,
                        float xmax, float ymax, float out[4]) {
    float t0 = 0.0f, t1 = 1.0f;
    float dx = x2 - x1, dy = y2 - y1;
    for (int edge = 0; edge < 4; edge++) {
        float p, q;
        switch (edge) {
            case 0: p = -dx; q = x1 - xmin; break;
            case 1: p = dx; q = xmax - x1; break;
            case 2: p = -dy; q = y1 - ymin; break;
            default: p = dy; q = ymax - y1; break;
        }
        if (p == 0.0f && q < 0.0f) return 0;
        float r = q / p;
        if (p < 0.0f) {
            if (r > t1) return 0;
            else if (r > t0) t0 = r;
        } else if (p > 0.0f) {
            if (r < t0) return 0;
            else if (r < t1) t1 = r;
        }
    }
    out[0] = x1 + t0 * dx;
    out[1] = y1 + t0 * dy;
    out[2] = x1 + t1 * dx;
    out[3] = y1 + t1 * dy;
    return 1;
}
int or_liang_barsky(float x1, float y1, float x2, float y2, float xmin, float ymin, float xmax,
                    float ymax, float out[4]) {
    return liang_barsky(x1, y1, x2, y2, xmin, ymin, xmax, ymax, out);
}

/* Plot sink: either a real fragment (line rasteriser) or a recorder. */
typedef struct {
    int record;
    int64_t* xy;
    double* al;
    int n, max;
    /* fragment context */
    rstate* rs;
    const ovtx *start, *end;
    float cx1, cy1, dlen;
    or_stats* st;
} plotter;

/* The `rasterize_fragment` closure (line.rs:54-109). */
static void plot(plotter* pl, int64_t x, int64_t y, double alpha) {
    if (pl->record) {
        if (pl->n < pl->max) {
            pl->xy[2 * pl->n] = x;
            pl->xy[2 * pl->n + 1] = y;
            pl->al[pl->n] = alpha;
        }
        pl->n++;
        return;
    }
    rstate* rs = pl->rs;
    or_fb* fb = rs->fb;
    or_stats* st = pl->st;
    if (!(x >= 0 && y >= 0)) return; /* :55 */
    /* Coordinate::new(x as u32, y as u32).into_index: x + y*W with no
     * upper bound check; past W*H the reference writes out of bounds (UB),
     * the oracle drops the plot. */
    uint64_t index = (uint64_t)(uint32_t)x + (uint64_t)(uint32_t)y * fb->w;
    if (index >= (uint64_t)fb->w * fb->h) { st->oob_plots++; return; }
    st->line_plots++;
    const sr_draw_desc* d = rs->d;
    uint32_t sref = d->stencil_ref & 0xFFu;
    uint32_t sv = fb->stencil[index];
    if (!or_stencil_test(d->stencil_test, sv, sref)) return;
    fb->stencil[index] = (uint8_t)or_stencil_op(d->stencil_op, d->stencil_wrap, sv, sref);
    float xf = (float)x + 0.5f, yf = (float)y + 0.5f;      /* :72-73 */
    float t = hypotf(pl->cx1 - xf, pl->cy1 - yf) / pl->dlen; /* :75 */
    float z = lerp1(t, pl->start->p[2], pl->end->p[2]);      /* :77 */
    if (z < 0.0f && z >= fb->depth[index]) {
        float k[OR_MAXK];
        for (int i = 0; i < rs->K; i++) k[i] = lerp1(t, pl->start->k[i], pl->end->k[i]);
        shade_write(rs, index, z, k, 1, alpha, &st->line_shaded);
    }
}

/* draw_line_bresenham (line.rs:125-151). */
static void bresenham(int64_t x0, int64_t y0, int64_t x1, int64_t y1, plotter* pl) {
    int64_t dx = x1 - x0 < 0 ? -(x1 - x0) : x1 - x0;
    int64_t dy = -(y1 - y0 < 0 ? -(y1 - y0) : y1 - y0);
    int64_t sx = x0 < x1 ? 1 : -1;
    int64_t sy = y0 < y1 ? 1 : -1;
    int64_t err = dx + dy;
    for (;;) {
        plot(pl, x0, y0, 1.0);
        if (x0 == x1 && y0 == y1) break;
        int64_t e2 = 2 * err;
        if (e2 >= dy) { err += dy; x0 += sx; }
        if (e2 <= dx) { err += dx; y0 += sy; }
    }
}

/* f64 -> i64 `as` cast (saturating, NaN -> 0). */
static inline int64_t f2i64(double v) {
    if (v != v) return 0;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)v;
}
static inline double fract64(double v) { return v - trunc(v); } /* f64::fract */

/* draw_line_xiaolin_wu (line.rs:159-238). */
static void wu(double x0, double y0, double x1, double y1, plotter* pl) {
    int steep = fabs(y1 - y0) > fabs(x1 - x0);
    double t;
    if (steep) { t = x0; x0 = y0; y0 = t; t = x1; x1 = y1; y1 = t; }
    if (x0 > x1) { t = x0; x0 = x1; x1 = t; t = y0; y0 = y1; y1 = t; }
    double dx = x1 - x0, dy = y1 - y0;
    double gradient = dx < 0.0001 ? 1.0 : dy / dx;
#define PLOTF(X, Y, O) plot(pl, f2i64(X), f2i64(Y), (O))
    double xend = round(x0);
    double yend = y0 + gradient * (xend - x0);
    double xgap = 1.0 - fract64(x0 + 0.5);
    double xpxl1 = xend, ypxl1 = trunc(yend);
    if (steep) {
        PLOTF(ypxl1, xpxl1, (1.0 - fract64(yend)) * xgap);
        PLOTF(ypxl1 + 1.0, xpxl1, fract64(yend) * xgap);
    } else {
        PLOTF(xpxl1, ypxl1, (1.0 - fract64(yend)) * xgap);
        PLOTF(xpxl1, ypxl1 + 1.0, fract64(yend) * xgap);
    }
    double intery = yend + gradient;
    xend = round(x1);
    yend = y1 + gradient * (xend - x1);
    xgap = fract64(x1 + 0.5);
    double xpxl2 = xend, ypxl2 = trunc(yend);
    if (steep) {
        PLOTF(ypxl2, xpxl2, (1.0 - fract64(yend)) * xgap);
        PLOTF(ypxl2 + 1.0, xpxl2, fract64(yend) * xgap);
    } else {
        PLOTF(xpxl2, ypxl2, (1.0 - fract64(yend)) * xgap);
        PLOTF(xpxl2, ypxl2 + 1.0, fract64(yend) * xgap);
    }
    double x = xpxl1 + 1.0;
    if (steep) {
        while (x <= (xpxl2 - 1.0)) {
            double yy = trunc(intery);
            PLOTF(yy, x, 1.0 - fract64(intery));
            PLOTF(yy + 1.0, x, fract64(intery));
            intery += gradient;
            x += 1.0;
        }
    } else {
        while (x <= (xpxl2 - 1.0)) {
            double yy = trunc(intery);
            PLOTF(x, yy, 1.0 - fract64(intery));
            PLOTF(x, yy + 1.0, fract64(intery));
            intery += gradient;
            x += 1.0;
        }
    }
#undef PLOTF
}

int or_bresenham(int64_t x0, int64_t y0, int64_t x1, int64_t y1, int64_t* xy, double* a,
                 int max) {
    plotter pl = {0};
    pl.record = 1; pl.xy = xy; pl.al = a; pl.max = max;
    bresenham(x0, y0, x1, y1, &pl);
    return pl.n;
}
int or_wu(double x0, double y0, double x1, double y1, int64_t* xy, double* a, int max) {
    plotter pl = {0};
    pl.record = 1; pl.xy = xy; pl.al = a; pl.max = max;
    wu(x0, y0, x1, y1, &pl);
    return pl.n;
}

/* f32 -> i64 `cast(x).unwrap()` (line.rs:112-116): truncation; NaN or
 * out of range panics in the reference (parity unpinned, treated as 0). */
static inline int64_t f32_to_i64(float v, int* bad) {
    if (!(v > -9.2233720e18f && v < 9.2233720e18f)) { *bad = 1; return 0; }
    return (int64_t)v;
}

/* rasterize_line (pipeline/stages/rasterization/line.rs:22-119). */
static void raster_line(rstate* rs, const otile* tile, const ovtx* s, const ovtx* e,
                        or_stats* st) {
    st->line_calls++;
    float c[4];
    if (!liang_barsky(s->p[0], s->p[1], e->p[0], e->p[1], (float)tile->x0, (float)tile->y0,
                      (float)tile->x1, (float)tile->y1, c))
        return; /* :51 */
    st->line_accepted++;
    plotter pl = {0};
    pl.rs = rs; pl.start = s; pl.end = e; pl.st = st;
    pl.cx1 = c[0]; pl.cy1 = c[1];
    pl.dlen = hypotf(c[0] - c[2], c[1] - c[3]); /* :52 */
    if (rs->d->aa_lines) {
        wu((double)c[0], (double)c[1], (double)c[2], (double)c[3], &pl);
    } else {
        int bad = 0;
        int64_t a0 = f32_to_i64(c[0], &bad), a1 = f32_to_i64(c[1], &bad);
        int64_t b0 = f32_to_i64(c[2], &bad), b1 = f32_to_i64(c[3], &bad);
        if (bad) { st->invalid++; return; }
        bresenham(a0, a1, b0, b1, &pl);
    }
}

/* rasterize_point (pipeline/stages/rasterization/point.rs:21-86). */
static void raster_point(rstate* rs, const otile* tile, const ovtx* p, or_stats* st) {
    st->point_calls++;
    float x = p->p[0], y = p->p[1], z = p->p[2];
    float bx0 = (float)tile->x0, by0 = (float)tile->y0, bx1 = (float)tile->x1, by1 = (float)tile->y1;
    if (!(bx0 <= x && x < bx1 && by0 <= y && y < by1)) return; /* :46 */
    or_fb* fb = rs->fb;
    size_t index = (size_t)(uint32_t)x + (size_t)(uint32_t)y * fb->w; /* :47-49 */
    st->point_hits++;
    const sr_draw_desc* d = rs->d;
    uint32_t sref = d->stencil_ref & 0xFFu;
    uint32_t sv = fb->stencil[index];
    if (!or_stencil_test(d->stencil_test, sv, sref)) return;
    fb->stencil[index] = (uint8_t)or_stencil_op(d->stencil_op, d->stencil_wrap, sv, sref);
    shade_write(rs, index, z, p->k, 0, 1.0, &st->point_shaded);
}

/* ------------------------------------------------------------------------ */
/* The draw (Pipeline::render_mesh -> VertexShader -> [GeometryShader] ->    */
/* finish -> FragmentShader::run).                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    /* indexed screen vertices (None after a geometry program, geometry.rs:255) */
    ovtx* iv;
    size_t niv;
    const uint32_t* idx;
    size_t nidx;
    int prim;
    vlist tris, lines, points; /* generated, screen space */
    otile* tiles;
    int ntiles;
    rstate rs;
    atomic_size_t next_tile;
} drawctx;

static inline uint32_t index_at(const or_mesh* m, size_t i) {
    return m->indices ? m->indices[i] : (uint32_t)i;
}

static void run_tile(drawctx* dc, const otile* tile, or_stats* st) {
    rstate* rs = &dc->rs;
    st->tiles++;
    /* fragment.rs:268-311 */
    if (dc->prim == SR_PRIM_TRIANGLES && dc->iv) {
        for (size_t i = 0; i + 3 <= dc->nidx; i += 3)
            raster_tri(rs, tile, &dc->iv[dc->idx ? dc->idx[i] : i],
                       &dc->iv[dc->idx ? dc->idx[i + 1] : i + 1],
                       &dc->iv[dc->idx ? dc->idx[i + 2] : i + 2], st);
    }
    for (size_t i = 0; i + 3 <= dc->tris.n; i += 3)
        raster_tri(rs, tile, &dc->tris.v[i], &dc->tris.v[i + 1], &dc->tris.v[i + 2], st);
    if (dc->prim == SR_PRIM_LINES && dc->iv) {
        for (size_t i = 0; i + 2 <= dc->nidx; i += 2)
            raster_line(rs, tile, &dc->iv[dc->idx ? dc->idx[i] : i],
                        &dc->iv[dc->idx ? dc->idx[i + 1] : i + 1], st);
    }
    for (size_t i = 0; i + 2 <= dc->lines.n; i += 2)
        raster_line(rs, tile, &dc->lines.v[i], &dc->lines.v[i + 1], st);
    if (dc->prim == SR_PRIM_POINTS && dc->iv) {
        for (size_t i = 0; i < dc->nidx; i++)
            raster_point(rs, tile, &dc->iv[dc->idx ? dc->idx[i] : i], st);
    }
    for (size_t i = 0; i < dc->points.n; i++) raster_point(rs, tile, &dc->points.v[i], st);
}

typedef struct {
    drawctx* dc;
    or_stats st;
} worker_arg;

static void* tile_worker(void* p) {
    worker_arg* wa = (worker_arg*)p;
    drawctx* dc = wa->dc;
    for (;;) { /* fragment.rs:250-253: i = fetch_add(1) */
        size_t i = atomic_fetch_add_explicit(&dc->next_tile, 1, memory_order_relaxed);
        if (i >= (size_t)dc->ntiles) break;
        run_tile(dc, &dc->tiles[i], &wa->st);
    }
    return NULL;
}

typedef struct {
    ovtx* out;
    const or_mesh* m;
    const sr_draw_desc* d;
    int normalize;
    atomic_size_t next;
    size_t n;
} vsjob;

/* Mapper::map (parallel.rs:57-83): chunked atomic map, order preserving. */
static void* vs_worker(void* p) {
    vsjob* j = (vsjob*)p;
    const size_t chunk = 1024;
    for (;;) {
        size_t s = atomic_fetch_add_explicit(&j->next, chunk, memory_order_relaxed);
        if (s >= j->n) break;
        size_t e = s + chunk < j->n ? s + chunk : j->n;
        for (size_t i = s; i < e; i++) {
            ovtx* o = &j->out[i];
            vs_run(j->d, j->m, (uint32_t)i, o);
            if (j->normalize) {
                float sp[4];
                normalize_pos(o->p, &j->d->viewport, sp);
                memcpy(o->p, sp, sizeof(sp));
            }
        }
    }
    return NULL;
}

static void stats_add(or_stats* a, const or_stats* b) {
    uint64_t* pa = (uint64_t*)a;
    const uint64_t* pb = (const uint64_t*)b;
    for (size_t i = 0; i < sizeof(or_stats) / sizeof(uint64_t); i++) pa[i] += pb[i];
}

void or_fb_clear(or_fb* fb, uint32_t rgba, float depth, uint8_t stencil) {
    size_t n = (size_t)fb->w * fb->h;
    for (size_t i = 0; i < n; i++) {
        fb->color[i] = rgba;
        fb->depth[i] = depth;
        fb->stencil[i] = stencil;
    }
}

/* ------------------------------------------------------------------------ */
/* Normal visualisation programs (full_example/src/shaders.rs:35-89).  The    */
/* emitted lines carry new clip positions and the varyings of their vertex   */
/* (vertex normals) or of the centroid (face normals).  Vector4 operands:    */
/* position (wx,wy,wz,1), normal (nx,ny,nz,0) from SR_VP_MVP_NORMAL_WORLD's   */
/* varyings; G[0..16) = projection * view, G[16] = length.  Matrix times      */
/* vector accumulates k = 0..3 from +0.0 like the vertex stage.              */
/* ------------------------------------------------------------------------ */
static void mat4_vec4(const float* M, const float v[4], float out[4]) {
    for (int r = 0; r < 4; r++) {
        float acc = 0.0f;
        for (int k = 0; k < 4; k++) acc = acc + M[r * 4 + k] * v[k];
        out[r] = acc;
    }
}

/* shaders.rs:35-60: per vertex, start = mv * position,
 * end = mv * (position + normal * NORMAL_LENGTH) */
static void gs_vertex_normals(const ovtx* a, const ovtx* b, const ovtx* c, const float* G, vlist* lines) {
    const ovtx* vs[3] = {a, b, c};
    for (int i = 0; i < 3; i++) {
        const ovtx* v = vs[i];
        const float p[4] = {v->k[3], v->k[4], v->k[5], 1.0f};
        const float n[4] = {v->k[0], v->k[1], v->k[2], 0.0f};
        float e[4];
        for (int j = 0; j < 4; j++) e[j] = p[j] + n[j] * G[16];
        ovtx s = *v, t = *v;
        mat4_vec4(G, p, s.p);
        mat4_vec4(G, e, t.p);
        vl_push(lines, &s);
        vl_push(lines, &t);
    }
}

/* shaders.rs:63-89: center = barycentric_interpolate(1/3, a, 1/3, b, 1/3, c)
 * per component ((a*t) + (b*t)) + (c*t); start = mv * center.position,
 * end = mv * (center.position + center.normal.normalize() * NORMAL_LENGTH),
 * normalize = v / sqrt(((x*x + y*y) + z*z) + w*w) */
static void gs_face_normals(const ovtx* a, const ovtx* b, const ovtx* c, int K, const float* G,
                            vlist* lines) {
    const float t = 1.0f / 3.0f;
    ovtx m = *a;
    for (int j = 0; j < K; j++) m.k[j] = ((a->k[j] * t) + (b->k[j] * t)) + (c->k[j] * t);
    const float pw = ((1.0f * t) + (1.0f * t)) + (1.0f * t);
    const float nw = ((0.0f * t) + (0.0f * t)) + (0.0f * t);
    const float p[4] = {m.k[3], m.k[4], m.k[5], pw};
    float n[4] = {m.k[0], m.k[1], m.k[2], nw};
    const float len = sqrtf(((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]) + n[3] * n[3]);
    float e[4];
    for (int j = 0; j < 4; j++) {
        n[j] = n[j] / len;
        e[j] = p[j] + n[j] * G[16];
    }
    ovtx s = m, u = m;
    mat4_vec4(G, p, s.p);
    mat4_vec4(G, e, u.p);
    vl_push(lines, &s);
    vl_push(lines, &u);
}

void or_gs_normal_lines(int program, const float* a_clip, const float* a_k, const float* b_clip,
                        const float* b_k, const float* c_clip, const float* c_k, int K, const float* G,
                        float* out_clip, float* out_k, int* nlines) {
    ovtx a, b, c;
    memcpy(a.p, a_clip, sizeof(a.p)); memcpy(a.k, a_k, sizeof(float) * K);
    memcpy(b.p, b_clip, sizeof(b.p)); memcpy(b.k, b_k, sizeof(float) * K);
    memcpy(c.p, c_clip, sizeof(c.p)); memcpy(c.k, c_k, sizeof(float) * K);
    vlist l;
    memset(&l, 0, sizeof(l));
    if (program == SR_GS_VERTEX_NORMALS) gs_vertex_normals(&a, &b, &c, G, &l);
    else gs_face_normals(&a, &b, &c, K, G, &l);
    for (size_t i = 0; i < l.n; i++) {
        memcpy(out_clip + 4 * i, l.v[i].p, sizeof(float) * 4);
        memcpy(out_k + (size_t)K * i, l.v[i].k, sizeof(float) * K);
    }
    *nlines = (int)(l.n / 2);
    free(l.v);
}

static void normalize_list(vlist* l, const sr_viewport* vp) {
    for (size_t i = 0; i < l->n; i++) {
        float sp[4];
        normalize_pos(l->v[i].p, vp, sp);
        memcpy(l->v[i].p, sp, sizeof(sp));
    }
}

int or_draw(or_fb* fb, const or_mesh* mesh, const sr_draw_desc* d, const or_texture* tex,
            int nthreads, or_stats* stats) {
    if (!fb || !mesh || !d || fb->w == 0 || fb->h == 0) return SR_ERR_INVALID_ARGUMENT;
    int nv = mesh->primitive == SR_PRIM_TRIANGLES ? 3 : mesh->primitive == SR_PRIM_LINES ? 2 : 1;
    size_t nidx = mesh->indices ? mesh->n_indices : mesh->n_vertices;
    if (nidx % (size_t)nv != 0) return SR_ERR_INVALID_ARGUMENT; /* pipeline/mod.rs:148 */
    if (mesh->indices)
        for (size_t i = 0; i < nidx; i++)
            if (mesh->indices[i] >= mesh->n_vertices) return SR_ERR_INVALID_ARGUMENT;
    if ((d->fragment_program == SR_FP_TEX_LAMBERT || d->fragment_program == SR_FP_GBUFFER) &&
        (!tex || !(tex->rgba || tex->rgbaf) || !tex->w || !tex->h))
        return SR_ERR_INVALID_ARGUMENT;
    /* other framebuffer formats take triangle draws only */
    if (fb->n_color && (mesh->primitive != SR_PRIM_TRIANGLES || d->geometry_program == SR_GS_VERTEX_NORMALS ||
                        d->geometry_program == SR_GS_FACE_NORMALS))
        return SR_ERR_UNSUPPORTED;
    if (nthreads < 1) nthreads = 1;
    or_stats local;
    memset(&local, 0, sizeof(local));

    drawctx dc;
    memset(&dc, 0, sizeof(dc));
    dc.rs.fb = fb;
    dc.rs.d = d;
    dc.rs.tex = tex;
    dc.rs.K = vp_varyings(d->vertex_program);
    dc.prim = mesh->primitive;
    dc.idx = mesh->indices;
    dc.nidx = nidx;

    /* Vertex stage (vertex.rs:87-160).  Without a geometry program the
     * viewport transform is applied in the same map (run_to_fragment). */
    size_t nvert = mesh->n_vertices;
    ovtx* clip = (ovtx*)calloc(nvert ? nvert : 1, sizeof(ovtx));
    vsjob job;
    job.out = clip; job.m = mesh; job.d = d; job.n = nvert;
    job.normalize = d->geometry_program == SR_GS_NONE;
    atomic_init(&job.next, 0);
    if (nthreads == 1) {
        vs_worker(&job);
    } else {
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, vs_worker, &job);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        free(th);
    }

    if (d->geometry_program == SR_GS_NONE) {
        dc.iv = clip;
        dc.niv = nvert;
    } else {
        /* GeometryShader::run over the indexed primitives (geometry.rs:199-215),
         * sequential order: i = 0, n, 2n, ...  Then finish (geometry.rs:60-129). */
        const int K = dc.rs.K;
        for (size_t i = 0; i + (size_t)nv <= nidx; i += (size_t)nv) {
            const int normals = d->geometry_program == SR_GS_VERTEX_NORMALS ||
                                d->geometry_program == SR_GS_FACE_NORMALS;
            if (mesh->primitive == SR_PRIM_TRIANGLES) {
                const ovtx* a = &clip[index_at(mesh, i)];
                const ovtx* b = &clip[index_at(mesh, i + 1)];
                const ovtx* c = &clip[index_at(mesh, i + 2)];
                if (d->geometry_program == SR_GS_VERTEX_NORMALS) gs_vertex_normals(a, b, c, d->gs_uniforms, &dc.lines);
                else if (d->geometry_program == SR_GS_FACE_NORMALS)
                    gs_face_normals(a, b, c, K, d->gs_uniforms, &dc.lines);
                else if (d->geometry_program == SR_GS_CLIP_NEAR) clip_tri_near(a, b, c, K, &dc.tris);
                else clip_tri_reference(a, b, c, K, &dc.tris);
            } else if (normals && mesh->primitive == SR_PRIM_LINES) {  /* re_emit, shaders.rs:59,88 */
                vl_push(&dc.lines, &clip[index_at(mesh, i)]);
                vl_push(&dc.lines, &clip[index_at(mesh, i + 1)]);
            } else if (normals) {
                vl_push(&dc.points, &clip[index_at(mesh, i)]);
            } else if (mesh->primitive == SR_PRIM_LINES) {
                const ovtx* s = &clip[index_at(mesh, i)];
                const ovtx* e = &clip[index_at(mesh, i + 1)];
                ovtx so, eo;
                if (d->geometry_program == SR_GS_CLIP_NEAR) {
                    /* near clip of a line: the reference line loop on the Near plane only */
                    ovtx s2 = *s, e2 = *e;
                    int s_in = plane_inside(PL_NEAR, &s2), p_in = plane_inside(PL_NEAR, &e2);
                    if (s_in != p_in) {
                        ovtx I;
                        plane_intersect(PL_NEAR, &s2, &e2, K, &I);
                        if (s_in) e2 = I; else s2 = I;
                    } else if (!s_in) {
                        continue;
                    }
                    vl_push(&dc.lines, &s2);
                    vl_push(&dc.lines, &e2);
                } else if (clip_line_reference(s, e, K, &so, &eo)) {
                    vl_push(&dc.lines, &so);
                    vl_push(&dc.lines, &eo);
                }
            } else {
                const ovtx* p = &clip[index_at(mesh, i)];
                int keep = d->geometry_program == SR_GS_CLIP_NEAR ? plane_inside(PL_NEAR, p)
                                                                  : clip_point_reference(p);
                if (keep) vl_push(&dc.points, p);
            }
        }
        local.gen_tris = dc.tris.n / 3;
        local.gen_lines = dc.lines.n / 2;
        local.gen_points = dc.points.n;
        normalize_list(&dc.tris, &d->viewport);
        normalize_list(&dc.lines, &d->viewport);
        normalize_list(&dc.points, &d->viewport);
        dc.iv = NULL;
    }

    /* Fragment stage (fragment.rs:168-319). */
    uint32_t tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u;
    uint32_t th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
    dc.ntiles = make_tiles(fb->w, fb->h, tw, th, NULL, 0);
    dc.tiles = (otile*)malloc(sizeof(otile) * (dc.ntiles > 0 ? dc.ntiles : 1));
    make_tiles(fb->w, fb->h, tw, th, dc.tiles, dc.ntiles);
    atomic_init(&dc.next_tile, 0);
    if (nthreads == 1) {
        worker_arg wa;
        wa.dc = &dc;
        memset(&wa.st, 0, sizeof(wa.st));
        tile_worker(&wa);
        stats_add(&local, &wa.st);
    } else {
        pthread_t* thr = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
        worker_arg* wa = (worker_arg*)calloc(nthreads, sizeof(worker_arg));
        for (int t = 0; t < nthreads; t++) {
            wa[t].dc = &dc;
            pthread_create(&thr[t], NULL, tile_worker, &wa[t]);
        }
        for (int t = 0; t < nthreads; t++) {
            pthread_join(thr[t], NULL);
            stats_add(&local, &wa[t].st);
        }
        free(thr);
        free(wa);
    }
    free(dc.tiles);
    free(clip);
    free(dc.tris.v);
    free(dc.lines.v);
    free(dc.points.v);
    if (stats) stats_add(stats, &local);
    return local.invalid ? SR_ERR_INVALID_COORDINATE : SR_OK;
}
