// sr_device.hpp — device-side building blocks: exact f32 helpers, the
// fragment programs, the texture sampler, blending and the stencil test/op.
//
// Every arithmetic expression is written in the reference's operation
// order and compiled with -ffp-contract=off, so the device result is the
// same IEEE value the reference computes on the CPU (no FMA contraction,
// correctly rounded division and sqrt).
#pragma once

#include "sr_internal.hpp"

namespace sr {

// f32 quotient num/den formed as RN_f32(RN_f64(num * RN_f64(1/den))).
// This equals the correctly rounded f32 division for every pair whose
// quotient is a normal f32: num and den have 24-bit significands, so num/den
// is never exactly a 25-bit midpoint and lies at least 2^-49 (relative)
// from every one, while the f64 product is within 2^-52 of num/den.
// Zero, infinite and NaN operands follow IEEE (x*inf, 0*inf = NaN).
__device__ __forceinline__ float quot(float num, double R) { return (float)((double)num * R); }

__device__ __forceinline__ uint32_t quant8(float x) {
    float c = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint32_t)(c * 255.0f + 0.5f);
}

__device__ __forceinline__ uint32_t pack4(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return r | (g << 8) | (b << 16) | (a << 24);
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ((ax * bx) + (ay * by)) + (az * bz);
}

// v / |v| per component.  The three f32 quotients x / len are formed as
// quot(x, RN64(1 / len)): the same values as correctly rounded division
// (the argument at quot), for one f64 reciprocal instead of three f32
// division sequences.  len = 0, inf or NaN gives what the division gives
// (x * inf, x * 0, NaN).
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
    const float len = sqrtf(dot3(x, y, z, x, y, z));
    const double R = 1.0 / (double)len;
    x = quot(x, R);
    y = quot(y, R);
    z = quot(z, R);
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) {
    return fminf(fmaxf(x, lo), hi);
}

// StencilTest::test(value = framebuffer, mask = ref), stencil.rs:109-124
__device__ __forceinline__ bool stencil_test(int test, uint32_t value, uint32_t mask) {
    switch (test) {
        case SR_STENCIL_ALWAYS: return true;
        case SR_STENCIL_NEVER: return false;
        case SR_STENCIL_LESS: return mask < value;
        case SR_STENCIL_LEQUAL: return mask <= value;
        case SR_STENCIL_GREATER: return mask > value;
        case SR_STENCIL_GEQUAL: return mask >= value;
        case SR_STENCIL_EQUAL: return mask == value;
        default: return mask != value;
    }
}

// StencilOp::op, stencil.rs:144-159 (u8)
__device__ __forceinline__ uint32_t stencil_op(int op, int wrap, uint32_t v, uint32_t mask) {
    switch (op) {
        case SR_STENCIL_KEEP: return v;
        case SR_STENCIL_INVERT: return (~v) & 0xFFu;
        case SR_STENCIL_ZERO: return 0u;
        case SR_STENCIL_REPLACE: return mask & 0xFFu;
        case SR_STENCIL_INCR: return wrap ? ((v + 1u) & 0xFFu) : (v == 255u ? 255u : v + 1u);
        default: return wrap ? ((v - 1u) & 0xFFu) : (v == 0u ? 0u : v - 1u);
    }
}

// The op applied n times (test Always): the value after n visits.
__device__ __forceinline__ uint32_t stencil_op_n(int op, int wrap, uint32_t v, uint32_t mask,
                                                 uint32_t n) {
    if (n == 0) return v;
    switch (op) {
        case SR_STENCIL_KEEP: return v;
        case SR_STENCIL_INVERT: return (n & 1u) ? ((~v) & 0xFFu) : v;
        case SR_STENCIL_ZERO: return 0u;
        case SR_STENCIL_REPLACE: return mask & 0xFFu;
        case SR_STENCIL_INCR: return wrap ? ((v + n) & 0xFFu) : (n >= 255u - v ? 255u : v + n);
        default: return wrap ? ((v - n) & 0xFFu) : (n >= v ? 0u : v - n);
    }
}

// unorm8 -> f32 through a 256-entry table filled with (float)i / 255.0f.
struct Lut {
    const float* t;
    __device__ __forceinline__ float operator()(uint32_t c) const { return t[c & 0xFFu]; }
};

// Texel (ix, iy) after the edge rule: its index, or -1 for the border
// (textures hold fewer than 2^31 texels).
__device__ __forceinline__ int texel_index(const RasterParams& P, int ix, int iy) {
    const int W = (int)P.texW, H = (int)P.texH;
    if (P.tex_edge == SR_EDGE_WRAP) {
        ix = ((ix % W) + W) % W;
        iy = ((iy % H) + H) % H;
    } else if (P.tex_edge == SR_EDGE_BORDER) {
        if (ix < 0 || iy < 0 || ix >= W || iy >= H) return -1;
    } else {
        ix = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix);
        iy = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy);
    }
    return iy * W + ix;
}

// Channel c of a texel as a float: RGBA8 texels as unorm (c / 255) from
// the packed word, RGBAF32 texels (a texture buffer plane,
// texturebuffer.rs:12-49) as stored; the border colour is RGBA8 either way.
__device__ __forceinline__ float texel_f32(const RasterParams& P, int i, int c, Lut lut) {
    return i >= 0 ? ((const float*)P.tex)[4 * (uint32_t)i + c] : lut(P.tex_border >> (8 * c));
}
__device__ __forceinline__ uint32_t texel_u8(const RasterParams& P, int i) {
    return i >= 0 ? P.tex[(uint32_t)i] : P.tex_border;
}

// Texture sampling (texture.rs:14-62 declares Filter/Edge; the sampling
// itself is build-defined): texel centres at (i + 0.5) / W.  Bilinear
// weights each channel's four texels in turn (an RGBA8 texel's packed word
// is fetched once), so only a channel's four values are live at a time.
// ANY_FMT = false: RGBA8 textures only (the triangle raster, whose draws
// with an RGBAF32 texture take the per-pixel raster instead: the float
// path costs it a wave per SIMD).
// float -> u32 as Rust's `as` casts it: NaN and negatives to 0, saturating.
__device__ __forceinline__ uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// Texel (x, y) of the full example's sampler: RGBA8 as c / 255, RGBAF32 as
// stored; indices past the last texel (a panic there) use the last.
__device__ __forceinline__ void span_texel(const RasterParams& P, uint32_t x, uint32_t y, Lut lut, float o[4]) {
    x = min(x, P.texW - 1u);
    y = min(y, P.texH - 1u);
    const uint32_t i = y * P.texW + x;
    if (P.tex_fmt == SR_COLOR_RGBAF32) {
#pragma unroll
        for (int c = 0; c < 4; c++) o[c] = ((const float*)P.tex)[4 * i + c];
    } else {
        const uint32_t w = P.tex[i];
#pragma unroll
        for (int c = 0; c < 4; c++) o[c] = lut(w >> (8 * c));
    }
}

// The full example's Texture::sample (full_example/src/texture.rs:29-84)
// with its decode_gamma (color.rs:49-56, gamma 2.2), in Rust's evaluation
// order.
__device__ __forceinline__ void sample_span(const RasterParams& P, float u, float v, Lut lut, float out[4]) {
    if (P.tex_edge == SR_EDGE_WRAP) {  // f32::fract
        u = u - truncf(u);
        v = v - truncf(v);
    } else {  // u.min(1.0).max(0.0)
        u = fmaxf(fminf(u, 1.0f), 0.0f);
        v = fmaxf(fminf(v, 1.0f), 0.0f);
    }
    const float w1 = (float)(P.texW - 1u), h1 = (float)(P.texH - 1u);
    if (P.tex_filter == SR_FILTER_SPAN_NEAREST) {
        span_texel(P, f2u_sat(roundf(u * w1)), f2u_sat(roundf(v * h1)), lut, out);
    } else {
        const float uu = (u * w1) + 0.5f, vv = (v * h1) + 0.5f;
        const uint32_t x = f2u_sat(floorf(uu)), y = f2u_sat(floorf(vv));
        const float ur = uu - (float)x, vr = vv - (float)y;
        const float uo = 1.0f - ur, vo = 1.0f - vr;
        float t00[4], t10[4], t01[4], t11[4];
        span_texel(P, x, y, lut, t00);
        span_texel(P, x + 1u, y, lut, t10);
        span_texel(P, x, y + 1u, lut, t01);
        span_texel(P, x + 1u, y + 1u, lut, t11);
#pragma unroll
        for (int c = 0; c < 4; c++)
            out[c] = (((t00[c] * uo) + (t10[c] * ur)) * vo) + (((t01[c] * uo) + (t11[c] * ur)) * vr);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = powf(out[c], 2.2f);
}

// Bilinear weights of texels (x0, y0) .. (x0 + 1, y0 + 1) in each channel:
// the top row, the bottom row, then the two rows.
struct BilinearAt {
    float ax, ay, bx, by;
    __device__ __forceinline__ float operator()(float t00, float t10, float t01, float t11) const {
        float top = (t00 * bx) + (t10 * ax);
        float bot = (t01 * bx) + (t11 * ax);
        return (top * by) + (bot * ay);
    }
};

// The texel grid position of (u, v) for a bilinear sample: the first
// texel (ix, iy) before the edge rule, and the weights.
__device__ __forceinline__ void bilinear_pos(const RasterParams& P, float u, float v, int& ix, int& iy,
                                             BilinearAt& w) {
    float fx = clampf(u * (float)P.texW - 0.5f, -1.0e8f, 1.0e8f);
    float fy = clampf(v * (float)P.texH - 0.5f, -1.0e8f, 1.0e8f);
    float x0 = floorf(fx), y0 = floorf(fy);
    w.ax = fx - x0;
    w.ay = fy - y0;
    ix = (int)x0;
    iy = (int)y0;
    w.bx = 1.0f - w.ax;
    w.by = 1.0f - w.ay;
}

// ... and the indices of the four texels after the edge rule.
__device__ __forceinline__ void bilinear_at(const RasterParams& P, float u, float v, int i[4], BilinearAt& w) {
    int ix, iy;
    bilinear_pos(P, u, v, ix, iy, w);
    i[0] = texel_index(P, ix, iy);
    i[1] = texel_index(P, ix + 1, iy);
    i[2] = texel_index(P, ix, iy + 1);
    i[3] = texel_index(P, ix + 1, iy + 1);
}

// Two adjacent RGBA8 texels in one 8-byte load (4-byte aligned; a global
// load, which leaves the LDS counter alone).
__device__ __forceinline__ uint2 texel_pair(const uint32_t* p) {
    typedef __attribute__((address_space(1))) const uint32_t GU32;
    uint2 r;
    __builtin_memcpy(&r, (GU32*)p, sizeof(r));
    return r;
}

// An RGBA8 sample in two steps: tex_fetch8 loads the texel words and
// tex_mix8 weights them, so a caller can compute something else while the
// loads are in flight.  Together they are the RGBA8 case of sample().  The
// four texels are the two words of row pair r0 and of r1, picked by two
// flags in tex_mix8 (after the wait): texel (x0, y) is r.y when left_hi,
// else r.x; texel (x0 + 1, y) is r.x when right_lo, else r.y.  A Nearest
// sample is the bilinear mix of four copies of its texel with weights
// ax = ay = 0, bx = by = 1, which returns that texel's values exactly
// (t * 1 + t * 0 = t for the finite unorm values), so the mix takes no
// branch on the filter.
struct TexFetch8 {
    uint2 r0, r1;
    bool left_hi, right_lo;
    BilinearAt at;
};
__device__ __forceinline__ TexFetch8 tex_fetch8(const RasterParams& P, float u, float v) {
    TexFetch8 f;
    f.left_hi = f.right_lo = false;
    if (P.tex_filter == SR_FILTER_NEAREST) {
        float fx = clampf(u * (float)P.texW, -1.0e8f, 1.0e8f);
        float fy = clampf(v * (float)P.texH, -1.0e8f, 1.0e8f);
        const uint32_t w = texel_u8(P, texel_index(P, (int)floorf(fx), (int)floorf(fy)));
        f.r0 = make_uint2(w, w);
        f.r1 = f.r0;
        f.at = BilinearAt{0.0f, 0.0f, 1.0f, 1.0f};
        return f;
    }
    int ix, iy;
    bilinear_pos(P, u, v, ix, iy, f.at);
    if (P.tex_edge == SR_EDGE_CLAMP && P.texW >= 2u) {
        // Clamp: texels (ix, y) and (ix + 1, y) after the edge rule both lie
        // in the pair at columns b, b + 1 with b = clamp(ix, 0, W - 2) (the
        // left one is b unless ix >= W - 1, the right one b + 1 unless
        // ix < 0): one 8-byte load per row instead of two 4-byte loads
        const int W = (int)P.texW, H = (int)P.texH;
        const int b = min(max(ix, 0), W - 2);
        const int y0 = min(max(iy, 0), H - 1), y1 = min(max(iy + 1, 0), H - 1);
        f.r0 = texel_pair(P.tex + (uint32_t)(y0 * W + b));
        f.r1 = texel_pair(P.tex + (uint32_t)(y1 * W + b));
        f.left_hi = ix >= W - 1;
        f.right_lo = ix < 0;
        return f;
    }
    f.r0 = make_uint2(texel_u8(P, texel_index(P, ix, iy)), texel_u8(P, texel_index(P, ix + 1, iy)));
    f.r1 = make_uint2(texel_u8(P, texel_index(P, ix, iy + 1)), texel_u8(P, texel_index(P, ix + 1, iy + 1)));
    return f;
}
__device__ __forceinline__ void tex_mix8(const TexFetch8& f, Lut lut, float out[4]) {
    const uint32_t w00 = f.left_hi ? f.r0.y : f.r0.x, w10 = f.right_lo ? f.r0.x : f.r0.y;
    const uint32_t w01 = f.left_hi ? f.r1.y : f.r1.x, w11 = f.right_lo ? f.r1.x : f.r1.y;
#pragma unroll
    for (int c = 0; c < 4; c++)
        out[c] = f.at(lut(w00 >> (8 * c)), lut(w10 >> (8 * c)), lut(w01 >> (8 * c)), lut(w11 >> (8 * c)));
}

template <bool ANY_FMT = true>
__device__ __forceinline__ void sample(const RasterParams& P, float u, float v, Lut lut,
                                       float out[4]) {
    if constexpr (ANY_FMT) {
        if (P.tex_filter >= SR_FILTER_SPAN_NEAREST) {
            sample_span(P, u, v, lut, out);
            return;
        }
        if (P.tex_fmt == SR_COLOR_RGBAF32) {
            if (P.tex_filter == SR_FILTER_NEAREST) {
                float fx = clampf(u * (float)P.texW, -1.0e8f, 1.0e8f);
                float fy = clampf(v * (float)P.texH, -1.0e8f, 1.0e8f);
                const int i = texel_index(P, (int)floorf(fx), (int)floorf(fy));
#pragma unroll
                for (int c = 0; c < 4; c++) out[c] = texel_f32(P, i, c, lut);
                return;
            }
            int i[4];
            BilinearAt at;
            bilinear_at(P, u, v, i, at);
#pragma unroll
            for (int c = 0; c < 4; c++)
                out[c] = at(texel_f32(P, i[0], c, lut), texel_f32(P, i[1], c, lut), texel_f32(P, i[2], c, lut),
                            texel_f32(P, i[3], c, lut));
            return;
        }
    }
    tex_mix8(tex_fetch8(P, u, v), lut, out);
}

// 4-vectors of the full example (nalgebra Vector4<f32>): dot sums from +0
// in index order, norm = sqrt(dot(v, v)), normalize divides each component.
__device__ __forceinline__ float dot4(const float a[4], const float b[4]) {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; c++) acc = acc + a[c] * b[c];
    return acc;
}
__device__ __forceinline__ void normalize4(float v[4]) {
    const float n = sqrtf(dot4(v, v));
    const double R = 1.0 / (double)n;  // v[c] / n as quot(v[c], R), as in normalize3
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = quot(v[c], R);
}
// powi by repeated squaring: r *= x for each set bit, x squared between.
template <int N>
__device__ __forceinline__ float powi_sq(float x) {
    float r = 1.0f;
    int n = N;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if (n & 1) r = r * x;
        n >>= 1;
        if (n == 0) break;
        x = x * x;
    }
    return r;
}
__device__ __forceinline__ float saturate1(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// The full example's fragment_shader (full_example/src/shaders.rs:92-163)
// with aces_filmic_tonemap and encode_gamma (color.rs:19-47): k = world
// position (4) then world normal (4).
__device__ __forceinline__ void fresnel4(const float* U, const float* k, float out[4]) {
    const float pos[4] = {k[0], k[1], k[2], k[3]}, nrm[4] = {k[4], k[5], k[6], k[7]};
    float view[4] = {U[0] - pos[0], U[1] - pos[1], U[2] - pos[2], 1.0f - pos[3]};
    normalize4(view);
    const float mat = powf(U[31], 2.2f);  // decode_gamma of the grey material
    const float albedo = 0.7f, ior = 1.45f;
    float col[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 4; l++) {
        const float* L = U + 3 + 7 * l;
        float d[4] = {L[0] - pos[0], L[1] - pos[1], L[2] - pos[2], 1.0f - pos[3]};
        const float dist = sqrtf(dot4(d, d));
        normalize4(d);  // light_dir
        float h[4] = {d[0] + view[0], d[1] + view[1], d[2] + view[2], d[3] + view[3]};
        normalize4(h);
        const float intensity = L[6] / powi_sq<2>(dist);
        const float ndl = saturate1(dot4(d, nrm));
        const float ndh = saturate1(dot4(nrm, h));
        const float vdh = saturate1(dot4(view, h));
        const float f0 = powi_sq<2>((1.0f - ior) / (1.0f + ior));
        const float f = f0 + (1.0f - f0) * powi_sq<5>(1.0f - vdh);
        const float diffuse = (1.0f - f) * ndl;
        const float specular = f * powi_sq<64>(ndh);
#pragma unroll
        for (int c = 0; c < 3; c++) col[c] = col[c] + (intensity * L[3 + c]) * (specular + ((diffuse * albedo) * mat));
    }
    const float inv_gamma = 1.0f / 2.2f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float x = col[c];
        const float t = (x * ((2.51f * x) + 0.03f)) / ((x * ((2.43f * x) + 0.59f)) + 0.14f);
        out[c] = powf(t, inv_gamma);
    }
    out[3] = 1.0f;
}

// Fragment programs (the closure passed to FragmentShader::run,
// fragment.rs:168-169).  k = interpolated varyings; o[i] = output i as
// RGBA floats before quantisation.  Returns the program's output count
// (SR_FP_GBUFFER: 4, the others 1).
template <int FP, bool ANY_FMT = true>
__device__ __forceinline__ int fragment_rgba(const RasterParams& P, const float* k, Lut lut, float o[][4]) {
    const float* U = P.fsu;
    if constexpr (FP == SR_FP_BLINN_PHONG) {
        float nx = k[0], ny = k[1], nz = k[2];
        normalize3(nx, ny, nz);
        float vx = U[3] - k[3], vy = U[4] - k[4], vz = U[5] - k[5];
        normalize3(vx, vy, vz);
        float ndl = dot3(nx, ny, nz, U[0], U[1], U[2]);
        float diff = ndl > 0.0f ? ndl : 0.0f;
        float hx = U[0] + vx, hy = U[1] + vy, hz = U[2] + vz;
        normalize3(hx, hy, hz);
        float ndh = dot3(nx, ny, nz, hx, hy, hz);
        float s = ndh > 0.0f ? ndh : 0.0f;
        s = s * s; s = s * s; s = s * s; s = s * s; s = s * s;
        float lit = U[9] + U[10] * diff;
        float spec = U[11] * s;
        o[0][0] = U[6] * lit + spec; o[0][1] = U[7] * lit + spec;
        o[0][2] = U[8] * lit + spec; o[0][3] = U[12];
        return 1;
    } else if constexpr (FP == SR_FP_TEX_LAMBERT || FP == SR_FP_GBUFFER) {
        // RGBA8-only callers issue the texel loads first and light the
        // fragment while they are in flight
        TexFetch8 tf;
        if constexpr (!ANY_FMT) tf = tex_fetch8(P, k[3], k[4]);
        float nx = k[0], ny = k[1], nz = k[2];
        normalize3(nx, ny, nz);
        float ndl = dot3(nx, ny, nz, U[0], U[1], U[2]);
        float diff = ndl > 0.0f ? ndl : 0.0f;
        float lit = U[3] + U[4] * diff;
        float t[4];
        if constexpr (!ANY_FMT) {
            // keep the lighting above the texel reads: the scheduler would
            // otherwise hoist the mix's table reads, and the wait for the
            // texels with them, above it
            __builtin_amdgcn_sched_barrier(0);
            tex_mix8(tf, lut, t);
        } else {
            sample<ANY_FMT>(P, k[3], k[4], lut, t);
        }
        o[0][0] = t[0] * lit; o[0][1] = t[1] * lit; o[0][2] = t[2] * lit; o[0][3] = t[3] * U[5];
        if constexpr (FP == SR_FP_TEX_LAMBERT) {
            return 1;
        } else {
            o[1][0] = nx * 0.5f + 0.5f; o[1][1] = ny * 0.5f + 0.5f; o[1][2] = nz * 0.5f + 0.5f; o[1][3] = 1.0f;
            o[2][0] = k[3]; o[2][1] = k[4]; o[2][2] = diff; o[2][3] = 1.0f;
            o[3][0] = lit; o[3][1] = lit; o[3][2] = lit; o[3][3] = 1.0f;
            return 4;
        }
    } else if constexpr (FP == SR_FP_UV_COLOR) {
        o[0][0] = k[3]; o[0][1] = k[4]; o[0][2] = U[0]; o[0][3] = U[1];
        return 1;
    } else if constexpr (FP == SR_FP_FRESNEL4) {
        fresnel4(U, k, o[0]);
        return 1;
    } else if constexpr (FP == SR_FP_INT_PALETTE) {
        // <i32 as NumCast>::from(f32): truncation toward zero (saturating
        // where the reference's unwrap would panic)
        const float x = k[0];
        const int32_t id = !(x == x) ? 0 : (x >= 2147483648.0f ? 2147483647 : (x < -2147483648.0f ? (-2147483647 - 1) : (int32_t)x));
        const int32_t m = ((id % 8) + 8) % 8;
#pragma unroll
        for (int c = 0; c < 4; c++) o[0][c] = U[4 * m + c];
        return 1;
    } else {
        o[0][0] = U[0]; o[0][1] = U[1]; o[0][2] = U[2]; o[0][3] = U[3];
        return 1;
    }
}

__device__ __forceinline__ uint32_t pack_q(const float c[4]) {
    return pack4(quant8(c[0]), quant8(c[1]), quant8(c[2]), quant8(c[3]));
}

// The program's (first) colour on an RGBA8 plane.
template <int FP, bool ANY_FMT = true>
__device__ __forceinline__ uint32_t fragment(const RasterParams& P, const float* k, Lut lut) {
    float o[FP == SR_FP_GBUFFER ? 4 : 1][4];
    fragment_rgba<FP, ANY_FMT>(P, k, lut, o);
    return pack_q(o[0]);
}

template <int FP>
constexpr int fp_varyings() {
    return FP == SR_FP_FRESNEL4     ? 8
           : FP == SR_FP_INT_PALETTE ? 1
           : FP == SR_FP_BLINN_PHONG ? 6
           : (FP == SR_FP_TEX_LAMBERT || FP == SR_FP_UV_COLOR || FP == SR_FP_GBUFFER) ? 5 : 0;
}

// Blend::blend(src, dst) on RGBA floats (RGBAF32 planes): `()` returns
// src, OVER is src-alpha over, OVER_NORMALIZED the full example's over
// (full_example/src/color.rs:5-17) in Rust's left-to-right order:
// ((x * a) + ((y * b) * a1)) / (a + (b * a1)).
template <int BLEND>
__device__ __forceinline__ void blend_f32(const float s[4], const float d[4], float o[4]) {
    if constexpr (BLEND == SR_BLEND_OVER) {
        const float a = s[3], ia = 1.0f - a;
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = (s[c] * a) + (d[c] * ia);
        o[3] = a + (d[3] * ia);
    } else if constexpr (BLEND == SR_BLEND_OVER_NORMALIZED) {
        const float a = s[3], b = d[3], a1 = 1.0f - a;
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = ((s[c] * a) + ((d[c] * b) * a1)) / (a + (b * a1));
        o[3] = a + (b * (1.0f - a));
    } else {
#pragma unroll
        for (int c = 0; c < 4; c++) o[c] = s[c];
    }
}

// Blend::blend(src, dst), color/blend.rs:9-31
template <int BLEND>
__device__ __forceinline__ uint32_t blend(uint32_t s, uint32_t d, Lut lut) {
    if constexpr (BLEND == SR_BLEND_OVER_NORMALIZED) {  // on unorm floats, quantised back
        const float fs[4] = {lut(s), lut(s >> 8), lut(s >> 16), lut(s >> 24)};
        const float fd[4] = {lut(d), lut(d >> 8), lut(d >> 16), lut(d >> 24)};
        float o[4];
        blend_f32<BLEND>(fs, fd, o);
        return pack_q(o);
    } else if constexpr (BLEND == SR_BLEND_OVER) {
        float a = lut(s >> 24), ia = 1.0f - a;
        uint32_t r = quant8((lut(s) * a) + (lut(d) * ia));
        uint32_t g = quant8((lut(s >> 8) * a) + (lut(d >> 8) * ia));
        uint32_t b = quant8((lut(s >> 16) * a) + (lut(d >> 16) * ia));
        uint32_t oa = quant8(a + (lut(d >> 24) * ia));
        return pack4(r, g, b, oa);
    } else {
        return s;
    }
}

// c.mul_alpha(ColorAlpha::from_scalar(alpha)), line.rs:100
__device__ __forceinline__ uint32_t sat_u8(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 255.0f) return 255u;
    return (uint32_t)f;
}
__device__ __forceinline__ uint32_t mul_alpha(int mode, uint32_t c, double alpha) {
    uint32_t a = c >> 24, na;
    if (mode == SR_ALPHA_F32) {
        na = sat_u8((float)a * (float)alpha);
    } else {
        uint32_t au8 = (alpha >= 0.0 && alpha <= 255.0) ? (uint32_t)alpha : 0u;
        na = sat_u8((float)a * ((float)au8 / 255.0f));
    }
    return (c & 0x00FFFFFFu) | (na << 24);
}

// Work-group index -> tile index so that each XCD (work-groups are dealt to
// the 8 XCDs round-robin) gets one contiguous run of tiles and keeps its
// neighbours' vertex data in its own L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, k = b >> 3;
    const uint32_t q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace sr
