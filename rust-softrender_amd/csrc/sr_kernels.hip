// sr_kernels.hip — gfx950 kernels of the draw path.
//
//   k_vertex        vertex stage + viewport transform  (vertex.rs:87-160,
//                   clipvertex.rs:89-126)
//   k_gs_*          geometry-stage clipping             (geometry.rs:261-336)
//   k_tri_setup     cull, det, bbox, bin counts         (triangle.rs:54-78)
//   k_scan_*        exclusive scans (wave prefix sums)
//   k_*_emit        (execution tile, item) pairs in item order
//   k_radix_*       stable LSD radix sort of pairs by tile (ballot ranking)
//   k_ranges        per-tile [start, end) in the sorted pairs
//   k_strips_*      stencil visits on clamped bboxes outside the bbox
//   k_raster_*      one work-group per execution tile: colour/depth/stencil
//                   of its pixels stay on-chip, every item is applied in the
//                   reference's per-pixel order, the tile is written once.
#include <algorithm>

#include "sr_device.hpp"
#include "sr_kernels.hpp"

namespace sr {

__host__ __device__ static inline uint32_t blocks_for(uint64_t n, uint32_t per) {
    return (uint32_t)((n + per - 1) / per);
}

// ===========================================================================
// Vertex stage
// ===========================================================================
// Screen transform of one position: ClipVertex::normalize,
// geometry/clipvertex.rs:89-126.  The viewport matrix product keeps its zero
// terms (0 * inf is NaN in the reference as well).
__device__ __forceinline__ void normalize_pos(float x, float y, float z, float w,
                                              const sr_viewport& vp, float out[4]) {
    const float left = vp.x, bottom = vp.y;
    const float right = left + vp.width, top = bottom + vp.height;
    const float nr = vp.near_z, fr = vp.far_z;
    const float m00 = (right - left) / 2.0f, m03 = (right + left) / 2.0f;
    const float m11 = (top - bottom) / -2.0f, m13 = (top + bottom) / 2.0f;
    const float m22 = (fr - nr) / -2.0f, m23 = (fr + nr) / -2.0f;
    const float xn = x / w, yn = y / w, zn = z / w;
    float a;
    a = 0.0f; a = a + m00 * xn; a = a + 0.0f * yn; a = a + 0.0f * zn; a = a + m03 * 1.0f; out[0] = a;
    a = 0.0f; a = a + 0.0f * xn; a = a + m11 * yn; a = a + 0.0f * zn; a = a + m13 * 1.0f; out[1] = a;
    a = 0.0f; a = a + 0.0f * xn; a = a + 0.0f * yn; a = a + m22 * zn; a = a + m23 * 1.0f; out[2] = a;
    out[3] = 1.0f / w;
}

// One vertex program invocation (the VertexShader::run closure stand-in,
// vertex.rs:97-109), optionally fused with the viewport transform
// (run_to_fragment, vertex.rs:139).  Writes slot `o` of the SoA output.
__device__ __forceinline__ void vertex_compute(const float* U, int program, int normalize,
                                               const sr_viewport& viewport, const float pin[4],
                                               const float nin[3], float u, float v, int32_t ia,
                                               float outp[4], float var[MAXK]) {
    float clip[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) acc = acc + U[r * 4 + k] * pin[k];
        clip[r] = acc;
    }
    if (program == SR_VP_MVP_INT) {  // R::from(i32): rounded to nearest
        var[0] = (float)ia;
        if (normalize) {
            normalize_pos(clip[0], clip[1], clip[2], clip[3], viewport, outp);
        } else {
            outp[0] = clip[0]; outp[1] = clip[1]; outp[2] = clip[2]; outp[3] = clip[3];
        }
        return;
    }
    if (program == SR_VP_WORLD4_NORMAL4) {
        // the full example's vertex_shader (shaders.rs:8-31): model * (p,1)
        // and normalize(mit * (n,0)) as 4-vectors
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float* row = r < 3 ? U + 25 + r * 4 : U + 40;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; k++) acc = acc + row[k] * pin[k];
            var[r] = acc;
        }
        float n4[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float* row = r < 3 ? U + 16 + r * 3 : U + 37;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) acc = acc + row[k] * nin[k];
            n4[r] = acc;
        }
        float len = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; c++) len = len + n4[c] * n4[c];
        len = sqrtf(len);
#pragma unroll
        for (int c = 0; c < 4; c++) var[4 + c] = n4[c] / len;
        if (normalize) {
            normalize_pos(clip[0], clip[1], clip[2], clip[3], viewport, outp);
        } else {
            outp[0] = clip[0]; outp[1] = clip[1]; outp[2] = clip[2]; outp[3] = clip[3];
        }
        return;
    }
    float n[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) acc = acc + U[16 + r * 3 + k] * nin[k];
        n[r] = acc;
    }
    normalize3(n[0], n[1], n[2]);
    var[0] = n[0]; var[1] = n[1]; var[2] = n[2];
    if (program == SR_VP_MVP_NORMAL_WORLD) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; k++) acc = acc + U[25 + r * 4 + k] * pin[k];
            var[3 + r] = acc;
        }
    } else {
        var[3] = u;
        var[4] = v;
        var[5] = 0.0f;
    }
    if (normalize) {
        normalize_pos(clip[0], clip[1], clip[2], clip[3], viewport, outp);
    } else {
        outp[0] = clip[0]; outp[1] = clip[1]; outp[2] = clip[2]; outp[3] = clip[3];
    }
}

// GeometryShader::finish on clip-space vertices a draw chain shares with
// others (geometry.rs:60-129): the same normalisation the vertex kernel
// fuses for run_to_fragment, varyings copied.
__global__ __launch_bounds__(256) void k_finish(VertexBuf in, VertexBuf out, uint32_t n, int K, sr_viewport vp) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float sp[4];
    normalize_pos(in.comp(0)[i], in.comp(1)[i], in.comp(2)[i], in.comp(3)[i], vp, sp);
#pragma unroll
    for (int c = 0; c < 4; c++) out.comp(c)[i] = sp[c];
    const float* const src = in.var(i);
    float* const dst = out.var(i);
    for (int c = 0; c < K; c++) dst[c] = src[c];
}

// The varying rows of a block's vertices [i0, i0 + n) (n <= 256, row
// stride out.kstride <= MAXK), staged in LDS so that consecutive threads store
// consecutive words.  Every thread of the block calls it.
__device__ __forceinline__ void store_rows(float* lds, const VertexBuf& out, uint32_t i0, uint32_t n,
                                           const float var[MAXK], bool live) {
    const uint32_t K = out.kstride;
    if (live)
        for (uint32_t c = 0; c < K; c++) lds[threadIdx.x * K + c] = var[c];
    __syncthreads();
    float* const dst = out.var(i0);
    for (uint32_t j = threadIdx.x; j < n * K; j += 256u) dst[j] = lds[j];
}

// One vertex program invocation per thread.  Inputs are coalesced SoA
// streams; outputs are the SoA position streams (x, y, z, w) and the
// per-vertex varying rows.
__global__ __launch_bounds__(256) void k_vertex(VSParams P) {
    __shared__ float rows[256 * MAXK];
    const uint32_t b0 = blockIdx.x * 256u, i = b0 + threadIdx.x;
    const bool live = i < P.n;
    float var[MAXK] = {};
    if (live) {
        const float pin[4] = {P.px[i], P.py[i], P.pz[i], 1.0f};
        const float nin[3] = {P.nx ? P.nx[i] : 0.0f, P.ny ? P.ny[i] : 0.0f, P.nz ? P.nz[i] : 0.0f};
        float outp[4];
        vertex_compute(P.U, P.program, P.normalize, P.viewport, pin, nin, P.u ? P.u[i] : 0.0f,
                       P.v ? P.v[i] : 0.0f, P.ia ? P.ia[i] : 0, outp, var);
        // a screen-space (normalised) vertex is read for x, y, z only
#pragma unroll
        for (int c = 0; c < 3; c++) P.out.comp(c)[i] = outp[c];
        if (!P.normalize) P.out.comp(3)[i] = outp[3];
    }
    store_rows(rows, P.out, b0, min(256u, P.n - b0), var, live);
}

// scene owning element i of a batch: first s with i < end[s]
__device__ __forceinline__ uint32_t scene_of(const uint32_t* end, uint32_t n, uint32_t i) {
    uint32_t a = 0, b = n;
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        if (end[m] <= i) a = m + 1;
        else b = m;
    }
    return a;
}

// Loads through the constant address space: with a wave-uniform index they
// become scalar loads into SGPRs (a flat load counts as divergent).
__device__ __forceinline__ uint32_t ld_const(const uint32_t* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) uint32_t*)p)[i];
#else
    return p[i];
#endif
}

// scene_of for a wave-uniform element index, by scalar loads
__device__ __forceinline__ uint32_t scene_of_uniform(const uint32_t* end, uint32_t n, uint32_t i) {
    uint32_t a = 0, b = n;
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        if (ld_const(end, m) <= i) a = m + 1;
        else b = m;
    }
    return a;
}

// A scene's geometry record as scalar loads (s wave-uniform).
__device__ __forceinline__ SceneGeom load_scene(const SceneGeom* base, uint32_t s) {
    constexpr int N = sizeof(SceneGeom) / 4;
    const uint32_t* src = (const uint32_t*)(base + s);
    uint32_t w[N];
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = ld_const(src, (uint32_t)k);
    SceneGeom r;
    __builtin_memcpy(&r, w, sizeof(r));
    return r;
}

// Calls f(scene) for the scene owning element i of [0, n): when the wave's
// first and last elements share a scene (the common case: scenes hold
// thousands of elements) the scene is found and read once per wave by
// scalar loads, otherwise per lane.
template <class F>
__device__ __forceinline__ void with_scene(const BatchTables& T, const uint32_t* end, uint32_t n, uint32_t i,
                                           F&& f) {
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * 256u + (threadIdx.x & ~63u));
    if (w0 >= n) return;
    const uint32_t w1 = min(w0 + 63u, n - 1u);
    const uint32_t s0 = scene_of_uniform(end, T.nscenes, w0);
    if (scene_of_uniform(end, T.nscenes, w1) == s0) {
        if (i >= n) return;
        const SceneGeom S = load_scene(T.scenes, s0);
        f(S);
    } else {
        if (i >= n) return;
        f(T.scenes[scene_of(end, T.nscenes, i)]);
    }
}

__global__ __launch_bounds__(256) void k_vertex_batch(BatchTables T, VertexBuf out) {
    __shared__ float rows[256 * MAXK];
    const uint32_t b0 = blockIdx.x * 256u, i = b0 + threadIdx.x;
    float var[MAXK] = {};
    bool live = false;
    with_scene(T, T.vend, T.nv, i, [&](const SceneGeom& S) {
        const uint32_t l = i - S.vbase;
        const float pin[4] = {S.px[l], S.py[l], S.pz[l], 1.0f};
        const float nin[3] = {S.nx ? S.nx[l] : 0.0f, S.ny ? S.ny[l] : 0.0f, S.nz ? S.nz[l] : 0.0f};
        float outp[4];
        vertex_compute(S.U, S.program, 1, S.viewport, pin, nin, S.u ? S.u[l] : 0.0f, S.v ? S.v[l] : 0.0f, 0,
                       outp, var);
        // screen space: x, y, z are all that setup and shading read
#pragma unroll
        for (int c = 0; c < 3; c++) out.comp(c)[i] = outp[c];
        live = true;
    });
    if (b0 < T.nv) store_rows(rows, out, b0, min(256u, T.nv - b0), var, live);
}

// ===========================================================================
// Geometry stage: clipping programs (clip space), fused with finish()
// ===========================================================================
struct CV {
    float p[4];
    float k[6];
};

__device__ __forceinline__ void load_cv(const VertexBuf& v, uint32_t i, int K, CV& o) {
#pragma unroll
    for (int c = 0; c < 4; c++) o.p[c] = v.comp(c)[i];
#pragma unroll
    for (int c = 0; c < 6; c++) o.k[c] = c < K ? v.var(i)[c] : 0.0f;
}

// ClippingPlane::has_inside, geometry/clip.rs:33-44 (Left..Far = 0..5)
__device__ __forceinline__ bool inside(int pl, const CV& v) {
    const float x = v.p[0], y = v.p[1], z = v.p[2], w = v.p[3];
    switch (pl) {
        case 0: return x >= -w;
        case 1: return x <= w;
        case 2: return y >= -w;
        case 3: return y <= w;
        case 4: return z >= 0.0f;
        default: return z <= w;
    }
}

// ClippingPlane::intersect, geometry/clip.rs:47-62; lerp per
// numeric/interpolate.rs:55-58: ((1 - t) * a) + (t * b).
__device__ __forceinline__ void intersect(int pl, const CV& v1, const CV& v2, CV& o) {
    const float x1 = v1.p[0], y1 = v1.p[1], z1 = v1.p[2], w1 = v1.p[3];
    const float x2 = v2.p[0], y2 = v2.p[1], z2 = v2.p[2], w2 = v2.p[3];
    float a, b;
    switch (pl) {
        case 0: a = w1 + x1; b = w2 + x2; break;
        case 1: a = w1 - x1; b = w2 - x2; break;
        case 2: a = w1 + y1; b = w2 + y2; break;
        case 3: a = w1 - y1; b = w2 - y2; break;
        case 4: a = z1; b = z2; break;
        default: a = w1 - z1; b = w2 - z2; break;
    }
    const float t = a / (a - b);
    const float omt = 1.0f - t;
#pragma unroll
    for (int c = 0; c < 4; c++) o.p[c] = (omt * v1.p[c]) + (t * v2.p[c]);
#pragma unroll
    for (int c = 0; c < 6; c++) o.k[c] = (omt * v1.k[c]) + (t * v2.k[c]);
}

__device__ __forceinline__ uint32_t prim_vertex(const uint32_t* idx, uint32_t i) {
    return idx ? idx[i] : i;
}

// Number of output primitives of one input primitive.
template <int MODE, int PRIM>
__device__ uint32_t gs_count_one(const GSParams& P, uint32_t t) {
    if constexpr (PRIM == SR_PRIM_TRIANGLES) {
        CV v[3];
        for (int j = 0; j < 3; j++) load_cv(P.in, prim_vertex(P.idx, 3 * t + j), 0, v[j]);
        int n = 0;
        for (int e = 0; e < 3; e++) {
            const CV& s = v[e];
            const CV& p = v[(e + 1) % 3];
            if constexpr (MODE == SR_GS_CLIP_NEAR) {
                const bool si = inside(4, s), pi = inside(4, p);
                n += (si != pi) + pi;
            } else {
                for (int pl = 0; pl < 6; pl++) {
                    const bool si = inside(pl, s), pi = inside(pl, p);
                    n += (si != pi) + pi;
                }
            }
        }
        return n == 3 ? 1u : (n > 3 ? (uint32_t)(n - 2) : 0u);
    } else if constexpr (PRIM == SR_PRIM_LINES) {
        CV s, e;
        load_cv(P.in, prim_vertex(P.idx, 2 * t), 0, s);
        load_cv(P.in, prim_vertex(P.idx, 2 * t + 1), 0, e);
        if constexpr (MODE == SR_GS_CLIP_NEAR) {
            return (inside(4, s) || inside(4, e)) ? 1u : 0u;
        } else {
            // geometry.rs:300-328; the intersection only moves endpoints
            // along the segment, so run it exactly to decide acceptance.
            int ints = 0;
            for (int pl = 0; pl < 6; pl++) {
                const bool si = inside(pl, s), pi = inside(pl, e);
                if (si != pi) {
                    CV I;
                    intersect(pl, s, e, I);
                    if (si) e = I; else s = I;
                    ints += 1;
                } else if (!si) {
                    return 0u;
                }
                if (ints > 2) break;
            }
            return 1u;
        }
    } else {
        CV p;
        load_cv(P.in, prim_vertex(P.idx, t), 0, p);
        if constexpr (MODE == SR_GS_CLIP_NEAR) return inside(4, p) ? 1u : 0u;
        for (int pl = 0; pl < 6; pl++)
            if (!inside(pl, p)) return 0u;
        return 1u;
    }
}

template <int MODE, int PRIM>
__global__ __launch_bounds__(256) void k_gs_count(GSParams P) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= P.n_in) return;
    P.counts[t] = gs_count_one<MODE, PRIM>(P, t);
}

__device__ __forceinline__ void store_screen(const GSParams& P, uint32_t slot, const CV& v) {
    float s[4];
    normalize_pos(v.p[0], v.p[1], v.p[2], v.p[3], P.viewport, s);
#pragma unroll
    for (int c = 0; c < 4; c++) P.out.comp(c)[slot] = s[c];
    // static indices into v.k (a runtime index would put v in scratch memory)
#pragma unroll
    for (int c = 0; c < 6; c++)
        if (c < P.K) P.out.var(slot)[c] = v.k[c];
}

template <int MODE, int PRIM>
__global__ __launch_bounds__(256) void k_gs_emit(GSParams P) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= P.n_in) return;
    uint32_t o = P.offsets[t];
    if constexpr (PRIM == SR_PRIM_TRIANGLES) {
        if (P.counts[t] == 0) return;
        CV v[3];
#pragma unroll
        for (int j = 0; j < 3; j++) load_cv(P.in, prim_vertex(P.idx, 3 * t + j), P.K, v[j]);
        if constexpr (MODE == SR_GS_CLIP_NEAR) {
            // One plane: the polygon the loop below would build, written out
            // per inside pattern (a runtime-indexed polygon array would live
            // in scratch memory).  Edges (a,b), (b,c), (c,a) each push their
            // intersection if they cross, then their end vertex if inside.
            const CV &a = v[0], &b = v[1], &c = v[2];
            const int m = (inside(4, a) ? 1 : 0) | (inside(4, b) ? 2 : 0) | (inside(4, c) ? 4 : 0);
            if (m == 0) return;  // none inside (its count is 0 too)
            // all three intersections (unused ones are never stored), so that
            // every lane runs the same code below whatever its pattern
            CV i0, i1, i2;
            intersect(4, a, b, i0);
            intersect(4, b, c, i1);
            intersect(4, c, a, i2);
            // The polygon's fan as vertices q0..q3: emit_triangle(q3, q0, q1),
            // then for four vertices (q3, q1, q2).
            //   7: [b, c, a]                       as (q3, q0, q1) = (b, c, a)
            //   1: a inside: [I(a,b), I(c,a), a]    (q3, q0, q1) = (i0, i2, a)
            //   2: b inside: [I(a,b), b, I(b,c)]    (q3, q0, q1) = (i0, b, i1)
            //   4: c inside: [I(b,c), c, I(c,a)]    (q3, q0, q1) = (i1, c, i2)
            //   3: a, b inside: [b, I(b,c), I(c,a), a]
            //   5: a, c inside: [I(a,b), I(b,c), c, a]
            //   6: b, c inside: [I(a,b), b, c, I(c,a)]
            // the cases only copy vertices (cheap under divergence); the
            // normalisation and stores after them run once for every lane
            CV q0, q1, q2, q3;
            switch (m) {
                case 7: q3 = b; q0 = c; q1 = a; q2 = a; break;
                case 1: q3 = i0; q0 = i2; q1 = a; q2 = a; break;
                case 2: q3 = i0; q0 = b; q1 = i1; q2 = i1; break;
                case 4: q3 = i1; q0 = c; q1 = i2; q2 = i2; break;
                case 3: q0 = b; q1 = i1; q2 = i2; q3 = a; break;
                case 5: q0 = i0; q1 = i1; q2 = c; q3 = a; break;
                default: q0 = i0; q1 = b; q2 = c; q3 = i2; break;  // 6
            }
            const bool quad = m == 3 || m == 5 || m == 6;
            store_screen(P, 3 * o + 0, q3);
            store_screen(P, 3 * o + 1, q0);
            store_screen(P, 3 * o + 2, q1);
            if (quad) {
                store_screen(P, 3 * (o + 1) + 0, q3);
                store_screen(P, 3 * (o + 1) + 1, q1);
                store_screen(P, 3 * (o + 1) + 2, q2);
            }
            return;
        }
        CV poly[36];
        int n = 0;
        for (int e = 0; e < 3; e++) {
            const CV& s = v[e];
            const CV& p = v[(e + 1) % 3];
            {
                for (int pl = 0; pl < 6; pl++) {
                    const bool si = inside(pl, s), pi = inside(pl, p);
                    if (si != pi) intersect(pl, s, p, poly[n++]);
                    if (pi) poly[n++] = p;
                }
            }
        }
        if (n == 3) {
            for (int j = 0; j < 3; j++) store_screen(P, 3 * o + j, poly[j]);
        } else if (n > 3) {
            for (int i = 0; i < n - 2; i++) {  // emit_triangle(last, poly[i], poly[i+1])
                store_screen(P, 3 * (o + i) + 0, poly[n - 1]);
                store_screen(P, 3 * (o + i) + 1, poly[i]);
                store_screen(P, 3 * (o + i) + 2, poly[i + 1]);
            }
        }
    } else if constexpr (PRIM == SR_PRIM_LINES) {
        if (P.counts[t] == 0) return;
        CV s, e;
        load_cv(P.in, prim_vertex(P.idx, 2 * t), P.K, s);
        load_cv(P.in, prim_vertex(P.idx, 2 * t + 1), P.K, e);
        if constexpr (MODE == SR_GS_CLIP_NEAR) {
            const bool si = inside(4, s), pi = inside(4, e);
            if (si != pi) {
                CV I;
                intersect(4, s, e, I);
                if (si) e = I; else s = I;
            }
        } else {
            int ints = 0;
            for (int pl = 0; pl < 6; pl++) {
                const bool si = inside(pl, s), pi = inside(pl, e);
                if (si != pi) {
                    CV I;
                    intersect(pl, s, e, I);
                    if (si) e = I; else s = I;
                    ints += 1;
                }
                if (ints > 2) break;
            }
        }
        store_screen(P, 2 * o, s);
        store_screen(P, 2 * o + 1, e);
    } else {
        if (P.counts[t] == 0) return;
        CV p;
        load_cv(P.in, prim_vertex(P.idx, t), P.K, p);
        store_screen(P, o, p);
    }
}

// ===========================================================================
// Triangle setup: cull, det, bounding box, execution-tile counts
// (rasterize_triangle's per-triangle part, triangle.rs:48-78)
// ===========================================================================
// Returns the number of execution tiles the triangle's box covers and
// writes its record (X0 > X1 marks culled / invalid).
// Execution tiles of a triangle's clamped box: [ex0,ex1] x [ey0,ey1] and
// the pixel box [px0,px1] x [py0,py1]; false if the box misses the
// framebuffer.
struct TileSpan {
    int32_t px0, px1, py0, py1;
    uint32_t ex0, ex1, ey0, ey1;
};

template <class Box>  // TriRec or TriBin
__device__ __forceinline__ bool tile_span(const Box& r, uint32_t Wi, uint32_t Hi, uint32_t ets, TileSpan& s) {
    s.px0 = r.X0 > 0 ? r.X0 : 0;
    s.px1 = r.X1 < (int32_t)Wi - 1 ? r.X1 : (int32_t)Wi - 1;
    s.py0 = r.Y0 > 0 ? r.Y0 : 0;
    s.py1 = r.Y1 < (int32_t)Hi - 1 ? r.Y1 : (int32_t)Hi - 1;
    if (s.px0 > s.px1 || s.py0 > s.py1) return false;
    s.ex0 = s.px0 / ets; s.ex1 = s.px1 / ets; s.ey0 = s.py0 / ets; s.ey1 = s.py1 / ets;
    return true;
}

// Conservative coverage test for binning.  The barycentrics of
// rasterize_triangle (triangle.rs:104-113) are linear in the pixel
// coordinates: u(x, y) = ua*x + ub*y + uc at pixel (x, y) (centre x+0.5,
// y+0.5), likewise v, and w = 1 - u - v.  Over a pixel rectangle each
// maximum sits at the corner picked by the signs of the two coefficients,
// so a rectangle is rejected when one maximum stays below minus its margin:
// sixteen times the error bound of the f32 evaluation the raster performs
// (edge products, sums and the correctly rounded quotient:
// |u_f32 - u| <= 2^-22 * (m_u + |u|), m_u = (|e23y dx| + |e23x dy|) / |det|,
// bounded over the whole clamped box), plus a bound on the rounding of this
// f64 evaluation itself.  A triangle with det == 0 never draws (its
// barycentrics are infinite or NaN, so a sign test or the depth test fails).
// Only valid when the draw has no stencil effects, which apply to the whole
// clamped box.
struct CoverPlanes {
    double ua, ub, uc, va, vb, vc, wa, wb, wc;
    double du, dv, dw;
    bool never;
};

__device__ __forceinline__ CoverPlanes cover_planes(const TriRec& r, const TileSpan& sp) {
    CoverPlanes c;
    const double R = r.R;
    c.never = !(fabs(R) < __builtin_huge_val());
    const double ey23 = r.e23y, ex23 = r.e23x, ey31 = r.e31y, ex31 = r.e31x;
    const double ox = 0.5 - (double)r.x3, oy = 0.5 - (double)r.y3;  // dx = x + ox
    c.ua = ey23 * R; c.ub = ex23 * R; c.uc = (ey23 * ox + ex23 * oy) * R;
    c.va = ey31 * R; c.vb = ex31 * R; c.vc = (ey31 * ox + ex31 * oy) * R;
    c.wa = -(c.ua + c.va); c.wb = -(c.ub + c.vb); c.wc = (1.0 - c.uc) - c.vc;
    // f32-evaluation margins, bounded over the clamped box's corners
    double mu = 0.0, mv = 0.0, au = 0.0, av = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double dx = (double)((k & 1) ? sp.px1 : sp.px0) + ox;
        const double dy = (double)((k & 2) ? sp.py1 : sp.py0) + oy;
        const double u = (ey23 * dx + ex23 * dy) * R, v = (ey31 * dx + ex31 * dy) * R;
        mu = fmax(mu, (fabs(ey23 * dx) + fabs(ex23 * dy)) * fabs(R));
        mv = fmax(mv, (fabs(ey31 * dx) + fabs(ex31 * dy)) * fabs(R));
        au = fmax(au, fabs(u));
        av = fmax(av, fabs(v));
    }
    // rounding of the plane evaluation (products, sums, the folded terms)
    const double xm = (double)sp.px1 + 1.0, ym = (double)sp.py1 + 1.0;
    const double eu = 0x1p-48 * (fabs(c.ua) * xm + fabs(c.ub) * ym + fabs(c.uc) + mu + 1.0);
    const double ev = 0x1p-48 * (fabs(c.va) * xm + fabs(c.vb) * ym + fabs(c.vc) + mv + 1.0);
    const double ew = 0x1p-48 * (fabs(c.wa) * xm + fabs(c.wb) * ym + fabs(c.wc)) + eu + ev;
    const double s = 0x1p-18;
    c.du = s * (mu + au + 1.0) + eu;
    c.dv = s * (mv + av + 1.0) + ev;
    c.dw = s * (mu + mv + au + av + 3.0) + ew;
    return c;
}

__device__ __forceinline__ double plane_max(double a, double b, double cst, int x0, int x1, int y0, int y1) {
    return a * (double)(a > 0.0 ? x1 : x0) + b * (double)(b > 0.0 ? y1 : y0) + cst;
}

__device__ __forceinline__ bool span_tile_covered(const CoverPlanes& c, const TileSpan& s, uint32_t ets,
                                                  uint32_t ex, uint32_t ey) {
    if (c.never) return false;
    const int x0 = max(s.px0, (int)(ex * ets)), x1 = min(s.px1, (int)(ex * ets + ets - 1));
    const int y0 = max(s.py0, (int)(ey * ets)), y1 = min(s.py1, (int)(ey * ets + ets - 1));
    return !(plane_max(c.ua, c.ub, c.uc, x0, x1, y0, y1) < -c.du ||
             plane_max(c.va, c.vb, c.vc, x0, x1, y0, y1) < -c.dv ||
             plane_max(c.wa, c.wb, c.wc, x0, x1, y0, y1) < -c.dw);
}

// Strip coverage.  The binned raster tests a triangle only over its box;
// the reference also tests, in every screen tile the box misses, the box
// clamped into that tile (clamp_as_int!, triangle.rs:66-99): pixels outside
// the box, where f32 rounding can pass the coverage test (a sliver whose
// sharp end points at such a pixel).  A pixel centre p can pass only if the
// exact barycentrics of the record's f32 values satisfy u >= -eu, v >= -ev,
// w >= -ew, where eu bounds |u_f32 - u|: the two products and the sum of
// triangle.rs:108 and the difference x - x3 each round once (3 * 2^-24 of
// |e23y dx| + |e23x dy|), the quotient rounds to f32 (2^-24 |u|), so
// |u_f32 - u| <= 2^-22 (T_u + |u|) with T_u = (|e23y| |dx| + |e23x| |dy|) /
// |det|; where all three tests pass, u_f32, v_f32 <= 1, so |u|, |v| <= 2;
// w = (1 - u) - v adds the two subtractions' rounding.  (Bounds are taken
// 4 times larger.)
//
// With |dx|, |dy| <= d (p at Chebyshev distance d from (x3, y3)) the
// margins are affine in d: eu(d) = 2^-20 ((|e23y| + |e23x|) d / |det| + 2),
// ev(d) likewise, ew = eu + ev + 2^-18, and p can pass only inside T(d),
// the triangle with vertices (u, v) = (-eu, -ev), (1 + ev + ew, -ev),
// (-eu, 1 + eu + ew) mapped back to the screen.  T(d) grows with d and its
// vertices lie within c0 + c1 d of (x3, y3).  With c1 < 1/2, no pixel
// beyond D = (c0 + 2) / (1 - c1) lies in its own T(d) (which stays inside
// distance d), so every pixel that can pass lies in T(D).
//
// Returns BIN_BOX when T(D) holds no pixel centre of the framebuffer
// outside the box; BIN_REACH with [rx0, rx1] x [ry0, ry1] = the pixel
// centres of T(D) joined to the box (inside the framebuffer); BIN_CLAMP
// when no such D is found (an extreme sliver), and the triangle goes to
// every execution tile of its clamped boxes.
__device__ __forceinline__ uint32_t strip_reach(const TriRec& r, float det, uint32_t Wi, uint32_t Hi, int& rx0,
                                                int& rx1, int& ry0, int& ry1) {
    if (!(det != 0.0f) || !(fabsf(det) < __builtin_huge_valf())) return BIN_BOX;  // never shades
    const double e23y = r.e23y, e23x = r.e23x, e31y = r.e31y, e31x = r.e31x;
    const double x3 = r.x3, y3 = r.y3, df = (double)det;
    const double dm = e23y * e31x - e23x * e31y;  // the 2x2 determinant of the f32 edges
    if (!(dm != 0.0)) return BIN_CLAMP;
    const double k = df / dm;
    const double q = 0x1p-20;
    const double au = q * (fabs(e23y) + fabs(e23x)) / fabs(df), bu = 2.0 * q;  // eu(d) = au d + bu
    const double av = q * (fabs(e31y) + fabs(e31x)) / fabs(df), bv = 2.0 * q;
    const double aw = au + av, bw = bu + bv + 0x1p-18;
    // the widened vertices' (u, v) = (U0 + U1 d, V0 + V1 d)
    const double U0[3] = {-bu, 1.0 + bv + bw, -bu}, U1[3] = {-au, av + aw, -au};
    const double V0[3] = {-bv, -bv, 1.0 + bu + bw}, V1[3] = {-av, -av, au + aw};
    double c0 = 0.0, c1 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double ax = k * (e31x * U0[i] - e23x * V0[i]), ay = k * (-e31y * U0[i] + e23y * V0[i]);
        const double bx = k * (e31x * U1[i] - e23x * V1[i]), by = k * (-e31y * U1[i] + e23y * V1[i]);
        c0 = fmax(c0, fmax(fabs(ax), fabs(ay)));
        c1 = fmax(c1, fmax(fabs(bx), fabs(by)));
    }
    if (!(c1 < 0.5)) return BIN_CLAMP;
    const double D = (c0 + 2.0) / (1.0 - c1);
    const double eu = au * D + bu, ev = av * D + bv, ew = eu + ev + 0x1p-18;
    if (!(D < 1e9) || !(eu < 0.25 && ev < 0.25)) return BIN_CLAMP;
    const double us[3] = {-eu, 1.0 + ev + ew, -eu}, vs[3] = {-ev, -ev, 1.0 + eu + ew};
    double mnx = __builtin_huge_val(), mxx = -mnx, mny = mnx, mxy = -mnx;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double x = x3 + k * (e31x * us[i] - e23x * vs[i]);
        const double y = y3 + k * (-e31y * us[i] + e23y * vs[i]);
        mnx = fmin(mnx, x); mxx = fmax(mxx, x);
        mny = fmin(mny, y); mxy = fmax(mxy, y);
    }
    if (!(fabs(mnx) + fabs(mxx) + fabs(mny) + fabs(mxy) < 1e300)) return BIN_CLAMP;
    const double m = 1e-3;  // covers the f64 rounding of the vertices
    const int W = (int)Wi, H = (int)Hi;
    if (!((r.X0 >= 1 && !(mnx > (double)r.X0 - 0.5 + m)) || (r.X1 <= W - 2 && !(mxx < (double)r.X1 + 1.5 - m)) ||
          (r.Y0 >= 1 && !(mny > (double)r.Y0 - 0.5 + m)) || (r.Y1 <= H - 2 && !(mxy < (double)r.Y1 + 1.5 - m))))
        return BIN_BOX;
    // pixels i whose centre i + 0.5 lies in [mn - m, mx + m], joined to the box
    rx0 = (int)fmax(fmin(ceil(mnx - 0.5 - m), (double)r.X0), 0.0);
    rx1 = (int)fmin(fmax(floor(mxx - 0.5 + m), (double)r.X1), (double)(W - 1));
    ry0 = (int)fmax(fmin(ceil(mny - 0.5 - m), (double)r.Y0), 0.0);
    ry1 = (int)fmin(fmax(floor(mxy - 0.5 + m), (double)r.Y1), (double)(H - 1));
    return BIN_REACH;
}

// The same claim, sufficient and cheap (no division): taking d as the
// distance to the farthest pixel centre of the framebuffer, every widened
// vertex lies within |k| (|e31x| + |e23x|) E horizontally (|e31y| + |e23y|
// vertically) of its vertex, E = eu + ev + ew bounding each (u, v)
// displacement; when both stay below 0.45 px the widened triangle cannot
// reach a pixel centre outside the box (those lie at least 0.5 px beyond
// the vertices' extremes).  Multiplied through by |det|: nothing to divide.
__device__ __forceinline__ bool strip_box_fast(const TriRec& r, float det, uint32_t Wi, uint32_t Hi) {
    const double e23y = r.e23y, e23x = r.e23x, e31y = r.e31y, e31x = r.e31x;
    const double x3 = r.x3, y3 = r.y3, ad = fabs((double)det);
    const double D = fmax(fmax(fabs(x3), fabs((double)Wi - x3)), fmax(fabs(y3), fabs((double)Hi - y3))) + 1.0;
    const double adm = fabs(e23y * e31x - e23x * e31y);
    const double Lu = fabs(e23y) + fabs(e23x), Lv = fabs(e31y) + fabs(e31x);
    const double dE = 0x1p-19 * ((Lu + Lv) * D + 4.0 * ad) + 0x1p-18 * ad;  // |det| E
    const double Sx = fabs(e31x) + fabs(e23x), Sy = fabs(e31y) + fabs(e23y);
    return Sx * dE < 0.45 * adm && Sy * dE < 0.45 * adm;
}

// The execution-tile columns (or rows) that the box [a0, a1] clamped into
// each screen tile meets along one axis (screen tile j spans [j t,
// min(j t + t, amax)], j < n, as clamp_as_int! sees it): f(lo, hi) for each
// maximal run of them, in increasing order.  The clamped ranges move right
// with j, so runs are merged as they come.
template <class F>
__device__ __forceinline__ void clamp_runs(int a0, int a1, uint32_t t, uint32_t n, int amax, uint32_t ets, F&& f) {
    int lo = 0, hi = -1;
    for (uint32_t j = 0; j < n; j++) {
        const int T0 = (int)(j * t), T1 = min((int)(j * t + t), amax);
        const int c0 = min(max(a0, T0), T1) / (int)ets, c1 = min(max(a1, T0), T1) / (int)ets;
        if (hi >= lo && c0 <= hi + 1) {
            hi = max(hi, c1);
            continue;
        }
        if (hi >= lo) f(lo, hi);
        lo = c0;
        hi = c1;
    }
    if (hi >= lo) f(lo, hi);
}

// A draw's pixel and tile geometry, as binning needs it.
struct BinGeom {
    uint32_t W, H, ets, tw, th;
    __device__ uint32_t scols() const { return W > 1 ? (W - 1 + tw - 1) / tw : 0u; }  // fragment.rs:196
    __device__ uint32_t srows() const { return H > 1 ? (H - 1 + th - 1) / th : 0u; }
};

// BIN_CLAMP: the execution tiles met by the box clamped into every screen
// tile.  The clamped box of screen tile (i, j) is the product of a column
// range that depends on j alone and a row range that depends on i alone,
// so the tiles are the product of the column runs and the row runs.
__device__ __forceinline__ uint32_t clamp_tile_count(const TriBin& b, const BinGeom& g) {
    uint32_t nc = 0, nr = 0;
    clamp_runs(b.X0, b.X1, g.tw, g.scols(), (int)g.W - 1, g.ets, [&](int lo, int hi) { nc += (uint32_t)(hi - lo + 1); });
    clamp_runs(b.Y0, b.Y1, g.th, g.srows(), (int)g.H - 1, g.ets, [&](int lo, int hi) { nr += (uint32_t)(hi - lo + 1); });
    return nc * nr;
}

__device__ __forceinline__ void emit_clamp_pairs(const TriBin& b, const BinGeom& g, uint32_t ntx, uint32_t key_base,
                                                 uint32_t val, uint32_t* keys, uint32_t* vals) {
    uint32_t o = 0;
    clamp_runs(b.Y0, b.Y1, g.th, g.srows(), (int)g.H - 1, g.ets, [&](int ylo, int yhi) {
        for (int ey = ylo; ey <= yhi; ey++)
            clamp_runs(b.X0, b.X1, g.tw, g.scols(), (int)g.W - 1, g.ets, [&](int xlo, int xhi) {
                for (int ex = xlo; ex <= xhi; ex++) {
                    keys[o] = key_base + (uint32_t)ey * ntx + (uint32_t)ex;
                    vals[o] = val;
                    o++;
                }
            });
    });
}

// An upper bound of every depth a fragment of the triangle can write, as a
// depth_key rounded up to even (the raster's hierarchical depth test
// compares it with the least key of a wave's pixels).  A fragment passes
// only with u, v, w >= 0, and then u, v, w <= 1 and |u + v + w - 1| <=
// 2^-23 (w = (1 - u) - v rounds twice), so the exact z1 u + z2 v + z3 w is
// at most M + K 2^-23 (M = max z_i, K = max |z_i|); the three products and
// two sums of triangle.rs:115-117 add at most 8.001 * 2^-24 K + 5 * 2^-150.
// zhi = M + K 2^-19 + 2^-140, rounded up to f32, covers both.  A NaN or
// infinite depth gives the key of +inf.
__device__ __forceinline__ uint32_t depth_bound_key(float z1, float z2, float z3) {
    const double M = fmax(fmax((double)z1, (double)z2), (double)z3);
    const double K = fmax(fmax(fabs((double)z1), fabs((double)z2)), fabs((double)z3));
    const double b = M + K * 0x1p-19 + 0x1p-140;
    uint32_t key = 0xFF800000u;  // depth_key(+inf)
    if (z1 == z1 && z2 == z2 && z3 == z3 && b < (double)INFINITY) {
        float f = (float)b;
        if ((double)f < b) f = nextafterf(f, INFINITY);
        key = depth_key(__float_as_uint(f));
    }
    return key >= 0xFFFFFFFEu ? 0xFFFFFFFEu : (key + 1u) & ~1u;
}

// Setup of the triangle with screen vertices (x1,y1,z1), (x2,y2,z2),
// (x3,y3,z3); returns its pair count.  (The caller stores the vertex ids in
// the triangle's TriVid.)
__device__ __forceinline__ uint32_t tri_setup_xyz(float x1, float y1, float z1, float x2, float y2, float z2,
                                                  float x3, float y3, float z3, int cull, const BinGeom& g,
                                                  TriRec& r, TriBin& b, uint32_t* err, int empty_cull) {
    r.zs = depth_bound_key(z1, z2, z3);
    b = TriBin{1, 0, 1, 0, 0u, BIN_BOX};
    if (cull != SR_CULL_NONE) {
        // shoelace, triangle.rs:56 (left-to-right evaluation)
        const float A = x1 * y2 + x2 * y3 + x3 * y1 - x2 * y1 - x3 * y2 - x1 * y3;
        const int wind = signbit(A) ? SR_CULL_CW : SR_CULL_CCW;
        if (wind == cull) {
            r.X0 = 1; r.X1 = 0; r.Y0 = 1; r.Y1 = 0;
            return 0;
        }
    }
    const float det = (y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3);  // triangle.rs:64
    const float mnx = fminf(fminf(x1, x2), x3), mxx = fmaxf(fmaxf(x1, x2), x3);
    const float mny = fminf(fminf(y1, y2), y3), mxy = fmaxf(fmaxf(y1, y2), y3);
    r.R = 1.0 / (double)det;
    r.x3 = x3; r.y3 = y3;
    r.e23y = y2 - y3; r.e23x = x3 - x2;
    r.e31y = y3 - y1; r.e31x = x1 - x3;
    r.z1 = z1; r.z2 = z2; r.z3 = z3;
    if (mnx != mnx || mxx != mxx || mny != mny || mxy != mxy) {
        // cast(NaN).unwrap() panics in clamp_as_int!, triangle.rs:70
        atomicOr(&err[0], 1u);
        r.X0 = 1; r.X1 = 0; r.Y0 = 1; r.Y1 = 0;
        return 0;
    }
    const float W = (float)g.W, H = (float)g.H;
    r.X0 = (int32_t)floorf(clampf(mnx, -1.0f, W));
    r.X1 = (int32_t)floorf(clampf(mxx, -1.0f, W));
    r.Y0 = (int32_t)floorf(clampf(mny, -1.0f, H));
    r.Y1 = (int32_t)floorf(clampf(mxy, -1.0f, H));
    b = TriBin{r.X0, r.X1, r.Y0, r.Y1, ~0u, BIN_BOX};
    int rx0, rx1, ry0, ry1;
    const uint32_t mode = strip_box_fast(r, det, g.W, g.H) ? (uint32_t)BIN_BOX
                                                           : strip_reach(r, det, g.W, g.H, rx0, rx1, ry0, ry1);
    if (mode != BIN_BOX) {
        // a pixel outside the box may pass: every tile it may lie in, no culling
        atomicOr(&err[0], 2u);
        r.zs |= 1u;
        b.mode = mode;
        if (mode == BIN_CLAMP) return clamp_tile_count(b, g);
        b.X0 = rx0; b.X1 = rx1; b.Y0 = ry0; b.Y1 = ry1;
        TileSpan sp;
        if (!tile_span(b, g.W, g.H, g.ets, sp)) return 0;
        return (sp.ex1 - sp.ex0 + 1) * (sp.ey1 - sp.ey0 + 1);
    }
    TileSpan sp;
    if (!tile_span(r, g.W, g.H, g.ets, sp)) return 0;
    const uint32_t ntile = (sp.ex1 - sp.ex0 + 1) * (sp.ey1 - sp.ey0 + 1);
    if (!empty_cull) return ntile;
    const CoverPlanes cp = cover_planes(r, sp);
    if (cp.never) {
        b.tmask = 0u;
        return 0;
    }
    uint32_t n = 0, m = 0, i = 0;
    for (uint32_t ey = sp.ey0; ey <= sp.ey1; ey++)
        for (uint32_t ex = sp.ex0; ex <= sp.ex1; ex++, i++) {
            const bool c = span_tile_covered(cp, sp, g.ets, ex, ey);
            n += c ? 1u : 0u;
            if (c && i < 32) m |= 1u << i;
        }
    if (ntile <= 32) b.tmask = m;
    return n;
}

__device__ __forceinline__ uint32_t tri_setup_one(const VertexBuf& v, uint32_t i0, uint32_t i1,
                                                  uint32_t i2, int cull, const BinGeom& g, TriRec& r,
                                                  TriBin& b, uint32_t* err, int empty_cull) {
    const float* sx = v.comp(0);
    const float* sy = v.comp(1);
    const float* sz = v.comp(2);
    return tri_setup_xyz(sx[i0], sy[i0], sz[i0], sx[i1], sy[i1], sz[i1], sx[i2], sy[i2], sz[i2], cull, g, r, b,
                         err, empty_cull);
}

// Writes the (tile, item) pairs of one triangle: every tile of its box, or
// with culling the tiles setup found coverable (its tile mask, or the test
// again, on the full record `rec`, for boxes of more than 32 tiles), in
// row-major order.
__device__ __forceinline__ void emit_pairs(const TriBin& b, const TriRec* rec, const TileSpan& sp, uint32_t ets,
                                           uint32_t ntx, int empty_cull, uint32_t key_base, uint32_t val,
                                           uint32_t* keys, uint32_t* vals) {
    const uint32_t w = sp.ex1 - sp.ex0 + 1, h = sp.ey1 - sp.ey0 + 1;
    uint32_t o = 0;
    if (empty_cull && w * h <= 32) {
        for (uint32_t m = b.tmask; m; m &= m - 1) {
            const uint32_t i = __builtin_ctz(m);
            keys[o] = key_base + (sp.ey0 + i / w) * ntx + sp.ex0 + i % w;
            vals[o] = val;
            o++;
        }
        return;
    }
    CoverPlanes cp;
    if (empty_cull) {
        const TriRec r = *rec;
        cp = cover_planes(r, sp);
    }
    for (uint32_t ey = sp.ey0; ey <= sp.ey1; ey++)
        for (uint32_t ex = sp.ex0; ex <= sp.ex1; ex++) {
            if (empty_cull && !span_tile_covered(cp, sp, ets, ex, ey)) continue;
            keys[o] = key_base + ey * ntx + ex;
            vals[o] = val;
            o++;
        }
}

// Setup records of one work-group, staged in LDS so that they leave with
// consecutive threads on consecutive 16-byte (records) and 4-byte (box
// records, vertex ids) addresses rather than one record per thread.
struct SetupStage {
    TriRec rec[256];
    TriBin bin[256];
    TriVid vid[256];
};

__device__ __forceinline__ void stage_setup(SetupStage& st, const TriRec& r, const TriBin& b, uint32_t i0,
                                            uint32_t i1, uint32_t i2) {
    st.rec[threadIdx.x] = r;
    st.bin[threadIdx.x] = b;
    st.vid[threadIdx.x] = TriVid{i0, i1, i2};
}

// Every thread of the work-group calls it (after staging its record, if any).
__device__ __forceinline__ void store_setup(const SetupStage& st, TriRec* recs, TriBin* bins, TriVid* vids,
                                            uint32_t n) {
    __syncthreads();
    const uint32_t b0 = blockIdx.x * 256u;
    if (b0 >= n) return;
    const uint32_t m = min(256u, n - b0);
    constexpr uint32_t RW = sizeof(TriRec) / 16, BW = sizeof(TriBin) / 4, VW = sizeof(TriVid) / 4;
    const uint4* rs = (const uint4*)st.rec;
    uint4* rd = (uint4*)(recs + b0);
    for (uint32_t i = threadIdx.x; i < m * RW; i += 256u) rd[i] = rs[i];
    const uint32_t* bs = (const uint32_t*)st.bin;
    uint32_t* bd = (uint32_t*)(bins + b0);
    for (uint32_t i = threadIdx.x; i < m * BW; i += 256u) bd[i] = bs[i];
    if (!vids) return;  // no draw reads them (vertices 3t, 3t+1, 3t+2)
    const uint32_t* vs = (const uint32_t*)st.vid;
    uint32_t* vd = (uint32_t*)(vids + b0);
    for (uint32_t i = threadIdx.x; i < m * VW; i += 256u) vd[i] = vs[i];
}

__global__ __launch_bounds__(256) void k_tri_setup(SetupParams P) {
    __shared__ SetupStage st;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < P.n) {
        TriRec r;
        TriBin b;
        const uint32_t i0 = prim_vertex(P.idx, 3 * t), i1 = prim_vertex(P.idx, 3 * t + 1),
                       i2 = prim_vertex(P.idx, 3 * t + 2);
        const BinGeom g{P.W, P.H, P.ets, P.tw, P.th};
        P.counts[t] = tri_setup_one(P.v, i0, i1, i2, P.cull, g, r, b, P.err, P.empty_cull);
        stage_setup(st, r, b, i0, i1, i2);
    }
    store_setup(st, P.recs, P.bins, P.vids, P.n);
}

__global__ __launch_bounds__(256) void k_tri_setup_batch(BatchTables T, VertexBuf v, TriRec* recs,
                                                        TriBin* bins, TriVid* vids, uint32_t* counts,
                                                        uint32_t* err) {
    __shared__ SetupStage st;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    with_scene(T, T.tend, T.nt, t, [&](const SceneGeom& S) {
        const uint32_t l = t - S.tbase;
        TriRec r;
        TriBin b;
        const uint32_t i0 = S.vbase + prim_vertex(S.idx, 3 * l), i1 = S.vbase + prim_vertex(S.idx, 3 * l + 1),
                       i2 = S.vbase + prim_vertex(S.idx, 3 * l + 2);
        const BinGeom g{S.W, S.H, S.ets, S.tw, S.th};
        counts[t] = tri_setup_one(v, i0, i1, i2, S.cull, g, r, b, err, S.empty_cull);
        stage_setup(st, r, b, i0, i1, i2);
    });
    store_setup(st, recs, bins, vids, T.nt);
}

// (execution tile, triangle) pairs of a batch; keys are global tile ids.
// Pairs past the capacity are not written and raise the overflow flag.
// Orders LDS accesses of different lanes of one wave (no work-group barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cooperative emission.  A lane's triangle is "direct" when its pairs
// follow from its box alone: every tile of the box (no culling), or the
// set bits of its tile mask (culling, boxes of at most 32 tiles).  The
// wave writes the direct triangles' pairs together, 64 consecutive pairs
// per step (each lane finds the pair's triangle by binary search over the
// wave's count prefix), so the writes coalesce and no lane waits on a
// neighbour's larger count; the rest (culled boxes of more than 32 tiles)
// run emit_pairs on their own lane.
struct EmitWave {
    uint32_t pre[64], cnt[64], o[64], mask[64], ex0[64], ey0[64], w[64], ntx[64], kbase[64], val[64];
};

// Index of the j-th set bit (j < popcount(m)).
__device__ __forceinline__ uint32_t select_bit(uint32_t m, uint32_t j) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t step = 16; step >= 1; step >>= 1) {
        const uint32_t c = (uint32_t)__popc(m & ((1u << step) - 1u));
        if (j >= c) {
            j -= c;
            m >>= step;
            pos += step;
        }
    }
    return pos;
}

__device__ __forceinline__ uint32_t wave_prefix_excl(uint32_t v, uint32_t& total) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// All lanes of the wave call it.  mask = 0xFFFFFFFF with all = true: every
// tile of the box in row-major order.
__device__ __forceinline__ void emit_wave(EmitWave& e, bool direct, uint32_t c, uint32_t o, uint32_t mask,
                                          bool all, const TileSpan& sp, uint32_t ntx, uint32_t kbase,
                                          uint32_t val, uint32_t* keys, uint32_t* vals) {
    const int lane = threadIdx.x & 63;
    const uint32_t cd = direct ? c : 0u;
    uint32_t total;
    const uint32_t ex = wave_prefix_excl(cd, total);
    e.pre[lane] = ex + cd;
    e.cnt[lane] = cd;
    e.o[lane] = o;
    e.mask[lane] = all ? 0u : mask;  // 0: every tile
    e.ex0[lane] = sp.ex0;
    e.ey0[lane] = sp.ey0;
    e.w[lane] = sp.ex1 - sp.ex0 + 1;
    e.ntx[lane] = ntx;
    e.kbase[lane] = kbase;
    e.val[lane] = val;
    wave_lds_sync();
    for (uint32_t k = lane; k < total; k += 64) {
        uint32_t lo = 0, hi = 63;  // the first lane whose inclusive prefix exceeds k
        while (lo < hi) {
            const uint32_t m = (lo + hi) >> 1;
            if (e.pre[m] > k) hi = m;
            else lo = m + 1;
        }
        const uint32_t j = k - (e.pre[lo] - e.cnt[lo]);
        const uint32_t mk = e.mask[lo];
        const uint32_t i = mk ? select_bit(mk, j) : j;
        const uint32_t w = e.w[lo];
        const uint32_t at = e.o[lo] + j;
        keys[at] = e.kbase[lo] + (e.ey0[lo] + i / w) * e.ntx[lo] + e.ex0[lo] + i % w;
        vals[at] = e.val[lo];
    }
    wave_lds_sync();
}

__global__ __launch_bounds__(256) void k_tri_emit_batch(BatchTables T, const TriRec* recs, const TriBin* bins,
                                                       const uint32_t* counts, const uint32_t* offsets,
                                                       uint32_t cap, uint32_t* keys, uint32_t* vals,
                                                       uint32_t* overflow) {
    __shared__ EmitWave ew[4];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (__builtin_amdgcn_readfirstlane(blockIdx.x * 256u + (threadIdx.x & ~63u)) >= T.nt) return;  // whole wave
    // the three loads issue together (none waits on another)
    const bool in = t < T.nt;
    uint32_t c = in ? counts[t] : 0u;
    const uint32_t o = in ? offsets[t] : 0u;
    const TriBin bt = in ? bins[t] : TriBin{1, 0, 1, 0, 0u, BIN_BOX};
    if (c && (uint64_t)o + c > cap) {
        atomicOr(overflow, 1u);
        c = 0;
    }
    TriBin b{1, 0, 1, 0, 0u, BIN_BOX};
    TileSpan sp{0, 0, 0, 0, 0u, 0u, 0u, 0u};
    bool direct = false, all = false;
    uint32_t ntx = 0, kbase = 0;
    BinGeom g{1, 1, 8, 128, 128};
    int empty_cull = 0;
    // the scene's fields, through the scalar cache when the wave's
    // triangles share one scene (nearly always)
    with_scene(T, T.tend, T.nt, t, [&](const SceneGeom& S) {
        ntx = S.ntx; kbase = S.tile_base; empty_cull = S.empty_cull;
        g = BinGeom{S.W, S.H, S.ets, S.tw, S.th};
    });
    if (c) {
        b = bt;
        if (b.mode == BIN_CLAMP) {
            direct = false;
        } else if (tile_span(b, g.W, g.H, g.ets, sp)) {
            const uint32_t nt = (sp.ex1 - sp.ex0 + 1) * (sp.ey1 - sp.ey0 + 1);
            all = !empty_cull || b.mode == BIN_REACH;
            direct = all || nt <= 32;
        } else {
            c = 0;
        }
    }
    emit_wave(ew[threadIdx.x >> 6], direct, c, o, b.tmask, all, sp, ntx, kbase, t, keys, vals);
    if (c && !direct) {
        if (b.mode == BIN_CLAMP) emit_clamp_pairs(b, g, ntx, kbase, t, keys + o, vals + o);
        else emit_pairs(b, recs + t, sp, g.ets, ntx, empty_cull, kbase, t, keys + o, vals + o);
    }
}

__global__ __launch_bounds__(256) void k_tri_emit(EmitParams P) {
    __shared__ EmitWave ew[4];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (__builtin_amdgcn_readfirstlane(blockIdx.x * 256u + (threadIdx.x & ~63u)) >= P.n) return;  // whole wave
    // the three loads issue together (none waits on another)
    const bool in = t < P.n;
    uint32_t c = in ? P.counts[t] : 0u;
    const uint32_t o = in ? P.offsets[t] : 0u;
    const TriBin bt = in ? P.bins[t] : TriBin{1, 0, 1, 0, 0u, BIN_BOX};
    if (c && (uint64_t)o + c > P.cap) c = 0;  // the host emits again into a larger buffer
    TriBin b{1, 0, 1, 0, 0u, BIN_BOX};
    TileSpan sp{0, 0, 0, 0, 0u, 0u, 0u, 0u};
    bool direct = false, all = false;
    if (c) {
        b = bt;
        if (b.mode == BIN_CLAMP) {
            direct = false;
        } else if (tile_span(b, P.W, P.H, P.ets, sp)) {
            const uint32_t nt = (sp.ex1 - sp.ex0 + 1) * (sp.ey1 - sp.ey0 + 1);
            all = !P.empty_cull || b.mode == BIN_REACH;
            direct = all || nt <= 32;
        } else {
            c = 0;
        }
    }
    emit_wave(ew[threadIdx.x >> 6], direct, c, o, b.tmask, all, sp, P.ntx, P.key_base, P.val_base + t, P.keys,
              P.vals);
    if (c && !direct) {
        if (b.mode == BIN_CLAMP)
            emit_clamp_pairs(b, BinGeom{P.W, P.H, P.ets, P.tw, P.th}, P.ntx, P.key_base, P.val_base + t, P.keys + o,
                             P.vals + o);
        else
            emit_pairs(b, P.recs + t, sp, P.ets, P.ntx, P.empty_cull, P.key_base, P.val_base + t, P.keys + o,
                       P.vals + o);
    }
}

// ===========================================================================
// Scans (exclusive, u32).  256 threads, 16 rounds of 256 elements per block.
// ===========================================================================
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t& total, uint32_t* lds4) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t s = lds4[w];
        if (w < wave) off += s;
        total += s;
    }
    __syncthreads();
    return x - v + off;
}

// The scans take an optional device-side element count (`count`, bounded
// by `n`): blocks past it exit, so a launch sized for a capacity costs
// little when the live count is small.
__device__ __forceinline__ uint32_t live_count(uint32_t n, const uint32_t* count) {
    return count ? min(n, *count) : n;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* sums,
                                                    const uint32_t* count) {
    __shared__ uint32_t lds4[4];
    n = live_count(n, count);
    if (blockIdx.x >= blocks_for(n, SCAN_TILE)) return;
    // 16 consecutive elements per thread (16-byte loads when aligned)
    const uint32_t e0 = blockIdx.x * SCAN_TILE + threadIdx.x * 16u;
    uint32_t acc = 0;
    if ((((uintptr_t)in) & 15u) == 0 && e0 + 15u < n) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 x = ((const uint4*)(in + e0))[q];
            acc += (x.x + x.y) + (x.z + x.w);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) acc += e0 + j < n ? in[e0 + j] : 0u;
    }
    uint32_t tot;
    block_scan_256(acc, tot, lds4);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single work-group scan of the block sums; also writes the grand total.
__global__ __launch_bounds__(256) void k_scan_top(uint32_t* sums, uint32_t n, uint32_t* total_out,
                                                 const uint32_t* count) {
    __shared__ uint32_t lds4[4];
    const uint32_t nb = blocks_for(live_count(n, count), SCAN_TILE);
    uint32_t run = 0;
    for (uint32_t b = 0; b < nb; b += 256) {
        const uint32_t e = b + threadIdx.x;
        const uint32_t v = e < nb ? sums[e] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan_256(v, tot, lds4);
        if (e < nb) sums[e] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = run;
}

// Each thread scans 16 consecutive elements in registers (16-byte loads and
// stores when the arrays are 16-byte aligned), so the block needs one
// block scan of the threads' totals.  In place is fine: a thread writes
// only the elements it read.
// block_off: the scanned block sums (k_scan_top), or with `sums` the raw
// block sums, which each block then adds up below itself (at most
// SCAN_FUSED_BLOCKS of them; the last block also writes the grand total):
// one launch fewer for small scans.
constexpr uint32_t SCAN_FUSED_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_scan_down(const uint32_t* in, uint32_t* out, uint32_t n,
                                                  const uint32_t* block_off, const uint32_t* count,
                                                  const uint32_t* sums, uint32_t* total_out) {
    __shared__ uint32_t lds4[4];
    n = live_count(n, count);
    const uint32_t nb = blocks_for(n, SCAN_TILE);
    if (blockIdx.x >= nb) {
        if (nb == 0 && blockIdx.x == 0 && threadIdx.x == 0 && sums && total_out) *total_out = 0u;
        return;
    }
    uint32_t base_off;
    if (sums) {
        uint32_t acc = 0;
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += 256u) acc += sums[i];
        uint32_t below;
        block_scan_256(acc, below, lds4);
        base_off = below;
        if (blockIdx.x == nb - 1u && threadIdx.x == 0 && total_out) *total_out = below + sums[blockIdx.x];
    } else {
        base_off = block_off[blockIdx.x];
    }
    const uint32_t e0 = blockIdx.x * SCAN_TILE + threadIdx.x * 16u;
    const bool vec = ((((uintptr_t)in) | ((uintptr_t)out)) & 15u) == 0 && e0 + 15u < n;
    uint32_t v[16];
    if (vec) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 x = ((const uint4*)(in + e0))[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = e0 + j < n ? in[e0 + j] : 0u;
    }
    uint32_t local = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const uint32_t x = v[j];
        v[j] = local;
        local += x;
    }
    uint32_t tot;
    const uint32_t off = base_off + block_scan_256(local, tot, lds4);
    if (vec) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            ((uint4*)(out + e0))[q] = make_uint4(off + v[4 * q], off + v[4 * q + 1], off + v[4 * q + 2],
                                                 off + v[4 * q + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (e0 + j < n) out[e0 + j] = off + v[j];
    }
}

// ===========================================================================
// Stable LSD radix sort of (key, value) by key, 8-bit digits.
// ===========================================================================
// Sort segments (segmented mode): segment g holds elements [lo[g], hi[g]),
// its blocks are first[g] .. first[g+1] - 1, its digits are taken from
// key - kbase[g], and its histograms occupy hist[256 * first[g] ..) digit-
// major, so one exclusive scan over all of them yields, for every segment,
// offsets that start at lo[g] (the segments tile [0, live) in order).
struct SortSegs {
    const uint32_t* first;  // [nseg + 1]; first[nseg] = blocks of all segments
    const uint32_t* lo;
    const uint32_t* hi;
    const uint32_t* kbase;
    uint32_t nseg;          // 0: unsegmented
};

// The elements, histogram column and key base of one block.
struct BlockSpan {
    uint32_t lo, hi;       // elements [lo, hi)
    uint32_t hbase, hstride;  // its count of digit d is hist[d * hstride + hbase]
    uint32_t kbase;
    bool live;
};

__device__ __forceinline__ BlockSpan block_span(uint32_t b, uint32_t n, uint32_t nblocks, const uint32_t* n_dev,
                                                const SortSegs& S) {
    BlockSpan r;
    if (S.nseg == 0) {
        if (n_dev) {
            n = min(n, *n_dev);
            nblocks = blocks_for(n, SCAN_TILE);
        }
        r.live = b < nblocks;
        r.lo = b * SCAN_TILE;
        r.hi = min(r.lo + SCAN_TILE, n);
        r.hbase = b;
        r.hstride = nblocks;
        r.kbase = 0;
        return r;
    }
    r.live = b < S.first[S.nseg];
    if (!r.live) return r;
    uint32_t a = 0, z = S.nseg;  // first[a] <= b < first[z]: the last segment starting at or before b
    while (z - a > 1) {
        const uint32_t m = (a + z) >> 1;
        if (S.first[m] <= b) a = m;
        else z = m;
    }
    const uint32_t i = b - S.first[a];
    r.lo = S.lo[a] + i * SCAN_TILE;
    r.hi = min(r.lo + SCAN_TILE, S.hi[a]);
    r.hbase = 256u * S.first[a] + i;
    r.hstride = S.first[a + 1] - S.first[a];
    r.kbase = S.kbase[a];
    return r;
}

// The segment table of a batch's groups (one work-group): group g's pairs
// are those of its triangles [t0, t1), [offsets[t0], offsets[t1]) clamped
// to the live count; hist_count = 256 * all blocks, the scan's length.
__global__ __launch_bounds__(256) void k_seg_table(BatchGroups G, const uint32_t* offsets, const uint32_t* live,
                                                  uint32_t* first, uint32_t* lo, uint32_t* hi, uint32_t* kbase,
                                                  uint32_t* hist_count) {
    __shared__ uint32_t lds4[4];
    const uint32_t L = *live;
    auto pair_at = [&](uint32_t t) { return t < G.ntris ? min(offsets[t], L) : L; };
    uint32_t run = 0;
    for (uint32_t c = 0; c < G.ngroups; c += 256) {
        const uint32_t g = c + threadIdx.x;
        uint32_t nb = 0;
        if (g < G.ngroups) {
            const uint32_t d0 = G.group_first[g], d1 = G.group_first[g + 1];
            const uint32_t t0 = d0 ? G.tend[d0 - 1] : 0u, t1 = d1 ? G.tend[d1 - 1] : 0u;
            lo[g] = pair_at(t0);
            hi[g] = pair_at(t1);
            kbase[g] = g ? G.group_tile_end[g - 1] : 0u;
            nb = blocks_for(hi[g] - lo[g], SCAN_TILE);
        }
        uint32_t tot;
        const uint32_t ex = block_scan_256(nb, tot, lds4);
        if (g < G.ngroups) first[g] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) {
        first[G.ngroups] = run;
        *hist_count = 256u * run;
    }
}

// Digit histograms, digit-major: hist[d * nb + block].  With a device-side
// element count the layout uses the live block count nb = ceil(count /
// SCAN_TILE) (block 0 stores 256 * nb in hist_count for the scan) and
// blocks past it exit.  Segmented: the layout of SortSegs.
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, uint32_t n, int shift,
                                                   uint32_t* hist, uint32_t nblocks,
                                                   const uint32_t* n_dev, uint32_t* hist_count, SortSegs S) {
    __shared__ uint32_t h[256];
    const BlockSpan B = block_span(blockIdx.x, n, nblocks, n_dev, S);
    if (S.nseg == 0 && n_dev && blockIdx.x == 0 && threadIdx.x == 0) *hist_count = 256u * B.hstride;
    if (!B.live) return;
    h[threadIdx.x] = 0;
    __syncthreads();
    // a full block at a 16-byte aligned start: 16 consecutive keys per
    // thread in 16-byte loads; otherwise (partial blocks, segments at any
    // offset) 16 rounds of 256 consecutive keys, which keep the loads
    // coalesced (the counts do not depend on the order)
    const uint32_t e0 = B.lo + threadIdx.x * 16u;
    if ((((uintptr_t)(keys + B.lo)) & 15u) == 0 && B.hi - B.lo == SCAN_TILE) {
        uint32_t k[16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 x = ((const uint4*)(keys + e0))[q];
            k[4 * q] = x.x; k[4 * q + 1] = x.y; k[4 * q + 2] = x.z; k[4 * q + 3] = x.w;
        }
#pragma unroll
        for (int j = 0; j < 16; j++) atomicAdd(&h[((k[j] - B.kbase) >> shift) & 255u], 1u);
    } else {
        for (int r = 0; r < 16; r++) {
            const uint32_t e = B.lo + r * 256u + threadIdx.x;
            if (e < B.hi) atomicAdd(&h[((keys[e] - B.kbase) >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[threadIdx.x * B.hstride + B.hbase] = h[threadIdx.x];
}

// Each wave ranks a contiguous quarter of the block (1024 elements, 16
// rounds of 64 in index order) against its own digit counters in LDS: a
// lane's rank among equal digits in a round comes from 8 ballots (a
// wave-wide match on the digit), the earlier rounds from the wave's
// counters, which only the wave touches (its LDS operations run in order,
// so the rounds need no barrier).  Then the waves' counts give each
// digit's start in the block and each wave's offset within it, and every
// element goes to its place in an LDS copy of the block, which is
// therefore the block stably sorted by digit.  The copy is written out
// with consecutive threads on consecutive addresses of each digit's run
// (the digit's global offset for this block comes from the scanned
// histograms), so the writes coalesce instead of scattering 4-byte words.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* keys_in,
                                                      const uint32_t* vals_in, uint32_t n,
                                                      int shift, const uint32_t* offsets,
                                                      uint32_t nblocks, uint32_t* keys_out,
                                                      uint32_t* vals_out, const uint32_t* n_dev, SortSegs S) {
    __shared__ uint32_t lkeys[SCAN_TILE], lvals[SCAN_TILE];
    __shared__ uint32_t lstart[256], gbase[256];
    __shared__ uint32_t wcnt[4][256];  // per wave: digit counts, then the wave's offset in each digit
    __shared__ uint32_t lds4[4];
    const BlockSpan B = block_span(blockIdx.x, n, nblocks, n_dev, S);
    if (!B.live) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t cnt_blk = B.hi - B.lo;
    const uint32_t wbase = B.lo + (uint32_t)wave * (SCAN_TILE / 4);
    gbase[tid] = offsets[tid * B.hstride + B.hbase];
    uint32_t* const mine = wcnt[wave];
#pragma unroll
    for (int q = 0; q < 4; q++) mine[lane + 64 * q] = 0;
    wave_lds_sync();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t K[16], V[16], R[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t e = wbase + r * 64u + lane;
        const bool valid = e < B.hi;
        K[r] = valid ? keys_in[e] : 0u;
        V[r] = valid ? vals_in[e] : 0u;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t e = wbase + r * 64u + lane;
        const bool valid = e < B.hi;
        const uint32_t d = ((K[r] - B.kbase) >> shift) & 255u;
        uint64_t mask = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            mask &= bit ? bal : ~bal;
        }
        const uint32_t rank = __popcll(mask & lt);
        // the counter is read by every lane of the digit before its first
        // lane (in a later LDS instruction of this wave) advances it
        const uint32_t prior = valid ? mine[d] : 0u;
        R[r] = prior + rank;
        wave_lds_sync();
        if (valid && rank == 0) mine[d] = prior + (uint32_t)__popcll(mask);
        wave_lds_sync();
    }
    __syncthreads();
    // thread tid = digit: block total, the waves' exclusive offsets, the
    // digit's start in the block
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t c = wcnt[w][tid];
        wcnt[w][tid] = tot;
        tot += c;
    }
    uint32_t all;
    lstart[tid] = block_scan_256(tot, all, lds4);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t e = wbase + r * 64u + lane;
        if (e < B.hi) {
            const uint32_t d = ((K[r] - B.kbase) >> shift) & 255u;
            const uint32_t pos = lstart[d] + mine[d] + R[r];
            lkeys[pos] = K[r];
            lvals[pos] = V[r];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < cnt_blk; i += 256u) {
        const uint32_t k = lkeys[i];
        const uint32_t d = ((k - B.kbase) >> shift) & 255u;
        const uint32_t pos = gbase[d] + (i - lstart[d]);
        keys_out[pos] = k;
        vals_out[pos] = lvals[i];
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

// p[i] = pattern[i % period] for i < n words (period 1, 2 or 4): a plane
// cleared to one colour or depth value of any format.
__global__ __launch_bounds__(256) void k_fill_pattern(uint32_t* p, uint4 pat, uint32_t period, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = (uint32_t)(i & (period - 1));
    p[i] = k == 0 ? pat.x : k == 1 ? pat.y : k == 2 ? pat.z : pat.w;
}

__global__ __launch_bounds__(256) void k_iota_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = i;
}

// Persistent-raster partition.  Each tile costs a fixed charge (load or
// clear, resolve) plus one per binned item; the costs are prefix-summed and
// tiles [tile0, tile0 + n) are cut into G runs of near-equal cost: run p
// holds the tiles whose cost prefix lies in [p*total/G, (p+1)*total/G), so
// wg_start[p] is its first tile (local index) and wg_start[G] = n (filled
// beforehand, like the starts of empty runs).
__global__ __launch_bounds__(256) void k_tile_cost(const uint32_t* start, const uint32_t* end, uint32_t n,
                                                  uint32_t fixed, uint32_t* cost) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t > n) return;
    cost[t] = t < n ? fixed + (end[t] - start[t]) : 0u;
}

__global__ __launch_bounds__(256) void k_partition(const uint32_t* prefix, uint32_t tile0, uint32_t n,
                                                  uint32_t G, uint32_t* wg_start) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n) return;
    const uint64_t base = prefix[tile0], tot = (uint64_t)prefix[tile0 + n] - base;
    const uint32_t w = (uint32_t)(((uint64_t)prefix[tile0 + t] - base) * G / tot);
    const uint32_t w0 = t == 0 ? 0u : (uint32_t)(((uint64_t)prefix[tile0 + t - 1] - base) * G / tot) + 1u;
    for (uint32_t q = w0; q <= w; q++) wg_start[q] = t;
}

// Grid-stride over the sorted keys (the grid is bounded; the live count
// may be far below the capacity the launch was sized for).
// Each thread takes four consecutive keys (one 16-byte load when `keys` is
// 16-byte aligned) and their two neighbours.
// Every tile's [start, end) in the sorted keys by binary search: thread t
// finds the first key >= t, which starts tile t and ends tile t - 1, so
// empty tiles need no clearing beforehand.  For a frame's tiles (tens of
// thousands); a batch's millions take k_ranges over the pairs instead.
__global__ __launch_bounds__(256) void k_ranges_search(const uint32_t* keys, uint32_t n, uint32_t nkeys,
                                                       uint32_t* start, uint32_t* end, const uint32_t* n_dev) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t > nkeys) return;
    uint32_t lo = 0, hi = n_dev ? min(*n_dev, n) : n;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (keys[m] < t) lo = m + 1;
        else hi = m;
    }
    if (t < nkeys) start[t] = lo;
    if (t > 0) end[t - 1] = lo;
}

__global__ __launch_bounds__(256) void k_ranges(const uint32_t* keys, uint32_t n, uint32_t key_base,
                                               uint32_t nkeys, uint32_t* start, uint32_t* end,
                                               const uint32_t* n_dev, int aligned) {
    if (n_dev) n = min(n, *n_dev);
    const uint32_t nq = (n + 3u) / 4u;
    for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < nq; q += gridDim.x * 256u) {
        const uint32_t i0 = 4u * q;
        uint32_t k[4];
        if (aligned && i0 + 3u < n) {
            const uint4 v = *(const uint4*)(keys + i0);
            k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) k[j] = i0 + j < n ? keys[i0 + j] : 0u;
        }
        const uint32_t prev = i0 ? keys[i0 - 1] : 0u;
        const uint32_t next = i0 + 4u < n ? keys[i0 + 4u] : 0u;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t i = i0 + j;
            if (i >= n) break;
            const uint32_t kk = k[j];
            if (kk - key_base >= nkeys) continue;  // never for complete pairs; guards the tables
            const uint32_t before = j ? k[j - 1] : prev, after = j < 3 ? k[j + 1] : next;
            if (i == 0 || before != kk) start[kk - key_base] = i;
            if (i == n - 1 || after != kk) end[kk - key_base] = i + 1;
        }
    }
}

// The pair count the sort and the ranges may use: the emitted total, or 0
// when the capacity overflowed (the slots past the first triangle that did
// not fit were never written and hold stale data).
__global__ void k_live_pairs(uint32_t* misc, int overflow_word, int total_word, int live_word) {
    if (blockIdx.x == 0 && threadIdx.x == 0) misc[live_word] = misc[overflow_word] ? 0u : misc[total_word];
}

// A plain draw's bins: misc[0] error flags (bit 0: NaN coordinate),
// misc[1] the pair count; misc[4] = 1 when the count exceeds the pair
// buffers' capacity, misc[5] = the pairs to sort (0 on overflow or NaN).
__global__ void k_bins_live(uint32_t* misc, uint32_t cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool over = misc[1] > cap;
        misc[4] = over ? 1u : 0u;
        misc[5] = (over || (misc[0] & 1u)) ? 0u : misc[1];
    }
}

// ===========================================================================
// Framebuffer fill (Framebuffer::clear, renderbuffer/mod.rs:125-133)
// ===========================================================================
__global__ __launch_bounds__(256) void k_fb_clear(uint32_t* color, float* depth, uint8_t* stencil,
                                                 uint32_t n, uint32_t rgba, float d, uint32_t s) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    color[i] = rgba;
    depth[i] = d;
    stencil[i] = (uint8_t)s;
}

// ===========================================================================
// Stencil visits outside the bounding box (test Always, op != Keep).
//
// rasterize_triangle runs the stencil op on every pixel of the bbox clamped
// into each screen tile, before the coverage test (triangle.rs:66-99).  A
// tile the bbox misses still gets a one-pixel-wide strip (or a corner) on its
// edge.  With test Always the stencil never gates shading and the op
// commutes with itself, so the final value is op^n(s0) with n the number of
// visits.  The raster counts visits inside bboxes; this pass counts the rest
// for every non-culled triangle: per tile column the bbox is Left / Over /
// Right of it, per tile row Above / Over / Below, and the strip cases are
// separable, so they reduce to difference arrays over tile columns x pixel
// rows (L, R), tile rows x pixel columns (A, B) and the tile grid (corners).
// ===========================================================================
__device__ __forceinline__ void strip_status(int lo, int hi, uint32_t t, uint32_t n, uint32_t len,
                                             uint32_t& nbefore, uint32_t& first_after) {
    // tiles entirely before the box (their max < lo) and the first tile
    // entirely after it (its min > hi)
    if (lo >= (int)len) nbefore = n;
    else if (lo <= 0) nbefore = 0;
    else nbefore = min((uint32_t)(lo - 1) / t, n - 1);
    first_after = hi < 0 ? 0u : min((uint32_t)hi / t + 1u, n);
}

// Corner counters live in LDS per work-group when the four arrays fit
// (STRIP_LDS_CORNERS entries each): most triangles add to the same few
// corner cells (e.g. the grid origin), which as global atomics serialise.
// Likewise colR's column 0 and rowB's row 0 (every triangle right of or
// below some tile adds to them): with STRIP_HOT_GROUPS work-groups looping
// over the triangles, each accumulates them in LDS and writes its partial
// row out; k_strips_hot_reduce sums the partials, 32 rows per thread and
// one global add per cell and 32 rows.
constexpr uint32_t STRIP_LDS_CORNERS = 2048;
constexpr uint32_t STRIP_LDS_WORDS = 16384;  // 64 KiB
constexpr uint32_t STRIP_HOT_GROUPS = 512;

__device__ __host__ __forceinline__ uint32_t strip_corner_words(const StripParams& P) {
    const uint32_t NC = (P.srows + 1) * (P.scols + 1);
    return NC <= STRIP_LDS_CORNERS ? 4u * NC : 0u;
}

uint32_t strip_hot_groups(const StripParams& P) {
    const uint64_t words = (uint64_t)strip_corner_words(P) + P.H + 1 + P.W + 1;
    if (P.n == 0 || words > STRIP_LDS_WORDS) return 0;
    return std::min<uint32_t>(blocks_for(P.n, 256), STRIP_HOT_GROUPS);
}

__global__ __launch_bounds__(256) void k_strips_accum(StripParams P) {
    // LDS: 4 x NC corner counters (when they fit), then colR's column 0
    // (HS words) and rowB's row 0 (WS words) when hot_part is set
    extern __shared__ uint32_t sc[];
    const uint32_t HS = P.H + 1, WS = P.W + 1, CS = P.scols + 1;
    const uint32_t NC = (P.srows + 1) * CS;
    const uint32_t ncw = strip_corner_words(P);
    const bool lds = ncw != 0;
    const bool hot = P.hot_part != nullptr;
    uint32_t* const hR = sc + ncw;  // colR[0][y]
    uint32_t* const hB = hR + HS;   // rowB[0][x]
    uint32_t* const corner[4] = {P.cLL, P.cLR, P.cRL, P.cRR};
    const uint32_t nlds = ncw + (hot ? HS + WS : 0u);
    for (uint32_t i = threadIdx.x; i < nlds; i += 256) sc[i] = 0;
    __syncthreads();
    auto cadd = [&](int a, uint32_t idx, uint32_t v) {
        if (lds) atomicAdd(&sc[a * NC + idx], v);
        else atomicAdd(&corner[a][idx], v);
    };
    auto radd = [&](uint32_t y, uint32_t v) {  // colR[0 * HS + y]
        if (hot) atomicAdd(&hR[y], v);
        else atomicAdd(&P.colR[y], v);
    };
    auto badd = [&](uint32_t x, uint32_t v) {  // rowB[0 * WS + x]
        if (hot) atomicAdd(&hB[x], v);
        else atomicAdd(&P.rowB[x], v);
    };
    // hot mode: a grid-stride loop (the grid is at most STRIP_HOT_GROUPS)
    const uint32_t stride = hot ? gridDim.x * 256u : 0xFFFFFFFFu;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x;; t += stride) {
        const bool in = t < P.n;
        if (!hot && !in) break;
        if (hot && __syncthreads_or(in) == 0) break;
        const TriRec* r = in ? &P.recs[t] : nullptr;  // the box itself (a bin may hold a reach)
        const int X0 = r ? r->X0 : 1, X1 = r ? r->X1 : 0, Y0 = r ? r->Y0 : 1, Y1 = r ? r->Y1 : 0;
        if (X0 <= X1) {  // not culled
            uint32_t nR, jL, nB, iA;  // columns right-of-box-status prefix etc.
            strip_status(X0, X1, P.tw, P.scols, P.W, nR, jL);
            strip_status(Y0, Y1, P.th, P.srows, P.H, nB, iA);
            if (jL < nR) jL = nR;
            if (iA < nB) iA = nB;
            const uint32_t xmax = P.W - 1, ymax = P.H - 1;
            // rows over the box: (L,O) and (R,O) strips on tile columns
            for (uint32_t i = nB; i < iA; i++) {
                const int ty0 = (int)(i * P.th), ty1 = (int)min(i * P.th + P.th, ymax);
                const int ylo = max(Y0, ty0), yhi = min(Y1, ty1);
                if (ylo > yhi) continue;
                if (jL < P.scols) {  // columns j >= jL see the box on their left
                    atomicAdd(&P.colL[jL * HS + ylo], 1u);
                    atomicAdd(&P.colL[jL * HS + yhi + 1], 0xFFFFFFFFu);
                }
                if (nR > 0) {  // columns j < nR see it on their right
                    radd(ylo, 1u);
                    radd(yhi + 1, 0xFFFFFFFFu);
                    atomicAdd(&P.colR[nR * HS + ylo], 0xFFFFFFFFu);
                    atomicAdd(&P.colR[nR * HS + yhi + 1], 1u);
                }
            }
            // columns over the box: (O,A) and (O,B) strips on tile rows
            for (uint32_t j = nR; j < jL; j++) {
                const int tx0 = (int)(j * P.tw), tx1 = (int)min(j * P.tw + P.tw, xmax);
                const int xlo = max(X0, tx0), xhi = min(X1, tx1);
                if (xlo > xhi) continue;
                if (iA < P.srows) {
                    atomicAdd(&P.rowT[iA * WS + xlo], 1u);
                    atomicAdd(&P.rowT[iA * WS + xhi + 1], 0xFFFFFFFFu);
                }
                if (nB > 0) {
                    badd(xlo, 1u);
                    badd(xhi + 1, 0xFFFFFFFFu);
                    atomicAdd(&P.rowB[nB * WS + xlo], 0xFFFFFFFFu);
                    atomicAdd(&P.rowB[nB * WS + xhi + 1], 1u);
                }
            }
            // corners: 2D difference arrays over the tile grid [row][col]
            if (jL < P.scols && iA < P.srows) cadd(0, iA * CS + jL, 1u);  // left & above
            if (jL < P.scols && nB > 0) {                                  // left & below
                cadd(1, 0 * CS + jL, 1u);
                cadd(1, nB * CS + jL, 0xFFFFFFFFu);
            }
            if (nR > 0 && iA < P.srows) {  // right & above
                cadd(2, iA * CS + 0, 1u);
                cadd(2, iA * CS + nR, 0xFFFFFFFFu);
            }
            if (nR > 0 && nB > 0) {  // right & below
                cadd(3, 0, 1u);
                cadd(3, nB * CS, 0xFFFFFFFFu);
                cadd(3, nR, 0xFFFFFFFFu);
                cadd(3, nB * CS + nR, 1u);
            }
        }
        if (!hot) break;
    }
    __syncthreads();
    if (lds)  // one global add per non-zero cell and work-group
        for (uint32_t i = threadIdx.x; i < 4 * NC; i += 256)
            if (sc[i]) atomicAdd(&corner[i / NC][i % NC], sc[i]);
    if (hot) {  // this work-group's partial rows
        uint32_t* const part = P.hot_part + (size_t)blockIdx.x * (HS + WS);
        for (uint32_t i = threadIdx.x; i < HS + WS; i += 256) part[i] = hR[i];
    }
}

// colR's column 0 and rowB's row 0 (zero before, and no other term adds to
// them) += the work-groups' partials: blockIdx.y takes partial rows
// [32 y, 32 y + 32).
constexpr uint32_t STRIP_REDUCE_ROWS = 32;
__global__ __launch_bounds__(256) void k_strips_hot_reduce(StripParams P, uint32_t groups) {
    const uint32_t HS = P.H + 1, WS = P.W + 1;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= HS + WS) return;
    const uint32_t g0 = blockIdx.y * STRIP_REDUCE_ROWS, g1 = min(g0 + STRIP_REDUCE_ROWS, groups);
    uint32_t s = 0;
    for (uint32_t g = g0; g < g1; g++) s += P.hot_part[(size_t)g * (HS + WS) + i];
    if (s) atomicAdd(i < HS ? &P.colR[i] : &P.rowB[i - HS], s);
}

// In-place inclusive prefix over `len` elements of each of `rows` rows
// (stride `stride`, element step `step`), one work-group per row.
__global__ __launch_bounds__(256) void k_prefix_rows(uint32_t* a, uint32_t rows, uint32_t len,
                                                    uint32_t stride, uint32_t step) {
    __shared__ uint32_t lds4[4];
    const uint32_t row = blockIdx.x;
    if (row >= rows) return;
    uint32_t* base = a + (size_t)row * stride;
    uint32_t run = 0;
    for (uint32_t c = 0; c < len; c += 256) {
        const uint32_t e = c + threadIdx.x;
        const uint32_t v = e < len ? base[(size_t)e * step] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan_256(v, tot, lds4);
        if (e < len) base[(size_t)e * step] = run + ex + v;
        run += tot;
    }
}

__global__ __launch_bounds__(256) void k_strips_final(StripParams P) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= P.W * P.H) return;
    const uint32_t x = p % P.W, y = p / P.W;
    const uint32_t HS = P.H + 1, WS = P.W + 1, CS = P.scols + 1;
    const uint32_t xmax = P.W - 1, ymax = P.H - 1;
    // tile columns whose left (L) / right (R) edge is x; rows likewise
    int jl = -1, jr0 = -1, jr1 = -1, it = -1, ib0 = -1, ib1 = -1;
    if (x % P.tw == 0 && x / P.tw < P.scols) jl = (int)(x / P.tw);
    if (x % P.tw == 0 && x > 0) {
        const uint32_t j = x / P.tw - 1;
        if (j < P.scols && min((j + 1) * P.tw, xmax) == x) jr0 = (int)j;
    }
    if (min(P.scols * P.tw, xmax) == x && (int)P.scols - 1 != jr0) jr1 = (int)P.scols - 1;
    if (y % P.th == 0 && y / P.th < P.srows) it = (int)(y / P.th);
    if (y % P.th == 0 && y > 0) {
        const uint32_t i = y / P.th - 1;
        if (i < P.srows && min((i + 1) * P.th, ymax) == y) ib0 = (int)i;
    }
    if (min(P.srows * P.th, ymax) == y && (int)P.srows - 1 != ib0) ib1 = (int)P.srows - 1;
    uint32_t n = 0;
    if (jl >= 0) n += P.colL[jl * HS + y];
    if (jr0 >= 0) n += P.colR[jr0 * HS + y];
    if (jr1 >= 0) n += P.colR[jr1 * HS + y];
    if (it >= 0) n += P.rowT[it * WS + x];
    if (ib0 >= 0) n += P.rowB[ib0 * WS + x];
    if (ib1 >= 0) n += P.rowB[ib1 * WS + x];
    const int jrs[2] = {jr0, jr1}, ibs[2] = {ib0, ib1};
    if (jl >= 0 && it >= 0) n += P.cLL[it * CS + jl];
    for (int b = 0; b < 2; b++)
        if (jl >= 0 && ibs[b] >= 0) n += P.cLR[ibs[b] * CS + jl];
    for (int a = 0; a < 2; a++) {
        if (jrs[a] >= 0 && it >= 0) n += P.cRL[it * CS + jrs[a]];
        for (int b = 0; b < 2; b++)
            if (jrs[a] >= 0 && ibs[b] >= 0) n += P.cRR[ibs[b] * CS + jrs[a]];
    }
    P.out[p] = n;
}

// ===========================================================================
// Lines (pipeline/stages/rasterization/line.rs:22-238)
// ===========================================================================
// liang_barsky_iterative, geometry/line.rs:6-54
__device__ __forceinline__ bool liang_barsky(float x1, float y1, float x2, float y2, float xmin,
                                             float ymin, float xmax, float ymax, float out[4]) {
    float t0 = 0.0f, t1 = 1.0f;
    const float dx = x2 - x1, dy = y2 - y1;
#pragma unroll
    for (int edge = 0; edge < 4; edge++) {
        float p, q;
        switch (edge) {
            case 0: p = -dx; q = x1 - xmin; break;
            case 1: p = dx; q = xmax - x1; break;
            case 2: p = -dy; q = y1 - ymin; break;
            default: p = dy; q = ymax - y1; break;
        }
        if (p == 0.0f && q < 0.0f) return false;
        const float r = q / p;
        if (p < 0.0f) {
            if (r > t1) return false;
            else if (r > t0) t0 = r;
        } else if (p > 0.0f) {
            if (r < t0) return false;
            else if (r < t1) t1 = r;
        }
    }
    out[0] = x1 + t0 * dx;
    out[1] = y1 + t0 * dy;
    out[2] = x1 + t1 * dx;
    out[3] = y1 + t1 * dy;
    return true;
}

// Rust f32::hypot is the platform libm hypotf; glibc evaluates it as
// sqrt(x*x + y*y) in double, rounded once to float (infinities first).
__device__ __forceinline__ float hypot_libm(float a, float b) {
    if (__builtin_isinf(a) || __builtin_isinf(b)) return __builtin_inff();
    const double x = (double)a, y = (double)b;
    return (float)sqrt(x * x + y * y);
}

// Plot calls a line item's walk makes (every plot it may keep): Wu makes
// 4 endpoint plots and 2 per interior major coordinate xpxl1+1 ..
// xpxl2-1 (line.rs:159-238, the same rounding as wu_walk); Bresenham makes
// at most |dx| + |dy| + 1 (each step moves x, y or both).
__device__ __forceinline__ uint32_t walk_calls_bound(const float c[4], int aa) {
    if (aa) {
        double x0 = (double)c[0], y0 = (double)c[1], x1 = (double)c[2], y1 = (double)c[3], t;
        if (fabs(y1 - y0) > fabs(x1 - x0)) {
            t = x0; x0 = y0; y0 = t;
            t = x1; x1 = y1; y1 = t;
        }
        if (x0 > x1) { t = x0; x0 = x1; x1 = t; }
        const double inner = round(x1) - round(x0) - 1.0;
        return 4u + 2u * (inner > 0.0 ? (uint32_t)inner : 0u);
    }
    const int64_t dx = (int64_t)c[2] - (int64_t)c[0], dy = (int64_t)c[3] - (int64_t)c[1];
    return (uint32_t)((dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) + 1);
}

// Per line: the screen tiles whose closed rectangle the segment reaches
// (liang_barsky accepts), in the reference's row-major tile order.
__global__ __launch_bounds__(256) void k_line_tiles(LineSetupParams P, int pass) {
    const uint32_t l = blockIdx.x * 256u + threadIdx.x;
    if (l >= P.n_lines) return;
    const uint32_t i0 = prim_vertex(P.idx, 2 * l), i1 = prim_vertex(P.idx, 2 * l + 1);
    const float x1 = P.v.comp(0)[i0], y1 = P.v.comp(1)[i0];
    const float x2 = P.v.comp(0)[i1], y2 = P.v.comp(1)[i1];
    if (!(__builtin_isfinite(x1) && __builtin_isfinite(y1) && __builtin_isfinite(x2) &&
          __builtin_isfinite(y2))) {
        // cast(x).unwrap() on a non-finite coordinate panics, line.rs:112-116
        if (pass == 0) {
            P.counts[l] = 0;
            atomicOr(&P.err[0], 2u);
        }
        return;
    }
    // candidate tiles: the segment's box grown by a pixel (a superset of what
    // liang_barsky can accept; each candidate is then tested exactly)
    const float lo_x = fminf(x1, x2), hi_x = fmaxf(x1, x2);
    const float lo_y = fminf(y1, y2), hi_y = fmaxf(y1, y2);
    const float gx = 1.0f + 1e-6f * fmaxf(fabsf(lo_x), fabsf(hi_x));
    const float gy = 1.0f + 1e-6f * fmaxf(fabsf(lo_y), fabsf(hi_y));
    const float W = (float)P.W, H = (float)P.H;
    const int px0 = (int)floorf(clampf(lo_x - gx, 0.0f, W - 1.0f));
    const int px1 = (int)floorf(clampf(hi_x + gx, -1.0f, W - 1.0f));
    const int py0 = (int)floorf(clampf(lo_y - gy, 0.0f, H - 1.0f));
    const int py1 = (int)floorf(clampf(hi_y + gy, -1.0f, H - 1.0f));
    uint32_t n = 0;
    uint32_t o = pass ? P.offsets[l] : 0u;
    if (px0 <= px1 && py0 <= py1) {
        const uint32_t jlo0 = (uint32_t)px0 / P.tw, jhi = min((uint32_t)px1 / P.tw, P.scols - 1);
        const uint32_t ilo0 = (uint32_t)py0 / P.th, ihi = min((uint32_t)py1 / P.th, P.srows - 1);
        const uint32_t jlo = jlo0 > 0 ? jlo0 - 1 : 0, ilo = ilo0 > 0 ? ilo0 - 1 : 0;
        const uint32_t xmax = P.W - 1, ymax = P.H - 1;
        for (uint32_t i = ilo; i <= ihi; i++) {
            const float ty0 = (float)(i * P.th), ty1 = (float)min(i * P.th + P.th, ymax);
            for (uint32_t j = jlo; j <= jhi; j++) {
                const float tx0 = (float)(j * P.tw), tx1 = (float)min(j * P.tw + P.tw, xmax);
                float c[4];
                if (!liang_barsky(x1, y1, x2, y2, tx0, ty0, tx1, ty1, c)) continue;
                if (pass) {
                    LineItem it;
                    it.line = l;
                    it.tile = i * P.scols + j;
                    it.plot_off = 0;
                    it.plot_cnt = 0;
                    it.cx1 = c[0]; it.cy1 = c[1]; it.cx2 = c[2]; it.cy2 = c[3];
                    it.dlen = hypot_libm(c[0] - c[2], c[1] - c[3]);  // line.rs:52
                    it.bx0 = 1; it.bx1 = 0; it.by0 = 1; it.by1 = 0;
                    it.mbx = it.mby = it.mcols = it.mvalid = 0;
                    it.major0 = 0;
                    it.ranged = it.steep = 0;
                    it.pad = 0;
                    it.mlo = it.mhi = 0;
                    P.items[o + n] = it;
                    const uint32_t ub = walk_calls_bound(c, P.aa);
                    P.plot_ub[o + n] = ub;
                    // longest walks first (ascending key): the long waves
                    // start early and the short ones fill the tail
                    if (P.walk_len) P.walk_len[o + n] = 255u - min(ub / 2u, 255u);
                }
                n++;
            }
        }
    }
    if (pass == 0) P.counts[l] = n;
}

// f64 -> i64 `as` cast (saturating, NaN -> 0)
__device__ __forceinline__ int64_t f2i64(double v) {
    if (v != v) return 0;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)v;
}
__device__ __forceinline__ double fract64(double v) { return v - trunc(v); }  // f64::fract

// draw_line_bresenham, line.rs:125-151
template <class F>
__device__ void bresenham_walk(int64_t x0, int64_t y0, int64_t x1, int64_t y1, F&& plot) {
    const int64_t dx = x1 - x0 < 0 ? -(x1 - x0) : x1 - x0;
    const int64_t dy = -(y1 - y0 < 0 ? -(y1 - y0) : y1 - y0);
    const int64_t sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
    int64_t err = dx + dy;
    for (;;) {
        plot(x0, y0, 1.0);
        if (x0 == x1 && y0 == y1) break;
        const int64_t e2 = 2 * err;
        if (e2 >= dy) { err += dy; x0 += sx; }
        if (e2 <= dx) { err += dx; y0 += sy; }
    }
}

// draw_line_xiaolin_wu, line.rs:159-238 (f64, exactly as written).
// Reports whether the walk is steep and the major coordinate of the first
// interior plot pair.
struct WuShape {
    bool steep;
    double major0;
};
template <class F>
__device__ void wu_walk(double x0, double y0, double x1, double y1, F&& plot, WuShape& shape) {
    const bool steep = fabs(y1 - y0) > fabs(x1 - x0);
    shape.steep = steep;
    double t;
    if (steep) { t = x0; x0 = y0; y0 = t; t = x1; x1 = y1; y1 = t; }
    if (x0 > x1) { t = x0; x0 = x1; x1 = t; t = y0; y0 = y1; y1 = t; }
    const double dx = x1 - x0, dy = y1 - y0;
    const double gradient = dx < 0.0001 ? 1.0 : dy / dx;
    auto pf = [&](double X, double Y, double o) { plot(f2i64(X), f2i64(Y), o); };
    double xend = round(x0);
    double yend = y0 + gradient * (xend - x0);
    double xgap = 1.0 - fract64(x0 + 0.5);
    const double xpxl1 = xend, ypxl1 = trunc(yend);
    if (steep) {
        pf(ypxl1, xpxl1, (1.0 - fract64(yend)) * xgap);
        pf(ypxl1 + 1.0, xpxl1, fract64(yend) * xgap);
    } else {
        pf(xpxl1, ypxl1, (1.0 - fract64(yend)) * xgap);
        pf(xpxl1, ypxl1 + 1.0, fract64(yend) * xgap);
    }
    double intery = yend + gradient;
    xend = round(x1);
    yend = y1 + gradient * (xend - x1);
    xgap = fract64(x1 + 0.5);
    const double xpxl2 = xend, ypxl2 = trunc(yend);
    if (steep) {
        pf(ypxl2, xpxl2, (1.0 - fract64(yend)) * xgap);
        pf(ypxl2 + 1.0, xpxl2, fract64(yend) * xgap);
    } else {
        pf(xpxl2, ypxl2, (1.0 - fract64(yend)) * xgap);
        pf(xpxl2, ypxl2 + 1.0, fract64(yend) * xgap);
    }
    double x = xpxl1 + 1.0;
    shape.major0 = x;
    if (steep) {
        while (x <= (xpxl2 - 1.0)) {
            const double y = trunc(intery);
            pf(y, x, 1.0 - fract64(intery));
            pf(y + 1.0, x, fract64(intery));
            intery += gradient;
            x += 1.0;
        }
    } else {
        while (x <= (xpxl2 - 1.0)) {
            const double y = trunc(intery);
            pf(x, y, 1.0 - fract64(intery));
            pf(x, y + 1.0, fract64(intery));
            intery += gradient;
            x += 1.0;
        }
    }
}

// Walks one line item's plots: pass 0 counts them and the execution tiles
// they touch, pass 1 writes them.  A plot is kept when x >= 0 and y >= 0
// (line.rs:55) and x + y*W lies inside the framebuffer (past its end the
// reference writes out of bounds; such plots are dropped).
__global__ __launch_bounds__(256) void k_line_walk(WalkParams P) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= P.n_items) return;
    // a wave's lanes walk in lock step until the longest walk ends, so the
    // items run grouped by length when an order is given
    const uint32_t it_id = P.order ? P.order[t] : t;
    LineItem& it = P.items[it_id];
    uint32_t n = 0;
    const uint32_t o = P.plot_offsets[it_id], room = P.plot_ub[it_id];
    int bx0 = INT32_MAX, by0 = INT32_MAX, bx1 = -1, by1 = -1;
    // window of execution tiles the plots can reach: the item's screen tile
    // grown by one pixel on each side (Wu plots y+1 next to y)
    const uint32_t Tx0 = (it.tile % P.scols) * P.tw, Ty0 = (it.tile / P.scols) * P.th;
    const int mbx = (int)((Tx0 > 0 ? Tx0 - 1 : 0) / P.ets), mby = (int)((Ty0 > 0 ? Ty0 - 1 : 0) / P.ets);
    const int mcols = (int)(min(Tx0 + P.tw + 1, P.W - 1) / P.ets) - mbx + 1;
    const int mrows = (int)(min(Ty0 + P.th + 1, P.H - 1) / P.ets) - mby + 1;
    bool mvalid = mcols > 0 && mrows > 0 && mcols * mrows <= 128;
    uint64_t mlo = 0, mhi = 0;
    const uint32_t esh = (uint32_t)__builtin_ctz(P.ets);  // execution tiles are a power of two
    uint32_t calls = 0;
    bool wrapped = false;
    auto plot = [&](int64_t x, int64_t y, double alpha) {
        calls++;
        if (!(x >= 0 && y >= 0)) return;
        if (x > 0xFFFFFFFFll || y > 0xFFFFFFFFll) return;
        // index = x + y*W (line.rs:56) as pixel (col, row): x < W except
        // where a plot runs past the right edge and wraps to a later row
        uint32_t col = (uint32_t)x, row = (uint32_t)y;
        if (col >= P.W) {
            wrapped = true;
            row += col / P.W;
            col %= P.W;
            if (row < (uint32_t)y) return;  // wrapped past 2^32 rows: beyond the framebuffer
        }
        if (row >= P.H) return;  // index >= W*H
        if (n >= room) {  // never: the bound counts every call
            atomicOr(&P.err[0], 4u);
            return;
        }
        {
            Plot pl;
            pl.alpha = alpha;
            pl.index = col + row * P.W;
            pl.x = (int32_t)x;
            pl.y = (int32_t)y;
            pl.row = row;
            P.plots[o + n] = pl;
        }
        {
            const int ex = (int)(col >> esh), ey = (int)(row >> esh);
            bx0 = min(bx0, ex); bx1 = max(bx1, ex);
            by0 = min(by0, ey); by1 = max(by1, ey);
            const int rx = ex - mbx, ry = ey - mby;
            if (rx < 0 || rx >= mcols || ry < 0 || ry >= mrows) {
                mvalid = false;  // outside the window: bin the whole box
            } else {
                const int b = ry * mcols + rx;
                if (b < 64) mlo |= 1ull << b;
                else mhi |= 1ull << (b - 64);
            }
        }
        n++;
    };
    WuShape shape{false, 0.0};
    if (P.aa) {
        wu_walk((double)it.cx1, (double)it.cy1, (double)it.cx2, (double)it.cy2, plot, shape);
    } else {
        // cast::<f32, i64>(..).unwrap(): truncation (the clipped points are finite)
        bresenham_walk((int64_t)it.cx1, (int64_t)it.cy1, (int64_t)it.cx2, (int64_t)it.cy2, plot);
    }
    {
        it.plot_off = o;
        it.plot_cnt = n;
        if (n) {
            it.bx0 = (int16_t)bx0; it.bx1 = (int16_t)bx1; it.by0 = (int16_t)by0; it.by1 = (int16_t)by1;
            it.mbx = (uint16_t)mbx; it.mby = (uint16_t)mby; it.mcols = (uint16_t)mcols;
            it.mvalid = mvalid ? 1 : 0;
            it.mlo = mlo; it.mhi = mhi;
            // every plot kept, none wrapped: the interior plots of a major
            // coordinate sit at a known position
            const bool ranged = P.aa && !wrapped && n == calls && n >= 4 && shape.major0 >= 0.0 &&
                                shape.major0 < 2147483647.0;
            it.ranged = ranged ? 1 : 0;
            it.steep = shape.steep ? 1 : 0;
            it.major0 = ranged ? (int32_t)shape.major0 : 0;
            P.bin_counts[it_id] = mvalid ? (uint32_t)(__popcll(mlo) + __popcll(mhi))
                                         : (uint32_t)((bx1 - bx0 + 1) * (by1 - by0 + 1));
        } else {
            it.bx0 = 1; it.bx1 = 0; it.by0 = 1; it.by1 = 0;
            it.mvalid = 0;
            it.ranged = 0;
            P.bin_counts[it_id] = 0;
        }
    }
}

__global__ __launch_bounds__(256) void k_item_emit(ItemEmitParams P) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= P.n || P.counts[t] == 0) return;
    const LineItem& it = P.items[t];
    uint32_t o = P.offsets[t];
    if (it.mvalid) {  // only the tiles holding a plot
        for (int h = 0; h < 2; h++) {
            uint64_t m = h ? it.mhi : it.mlo;
            while (m) {
                const int b = h * 64 + __ffsll((long long)m) - 1;
                m &= m - 1;
                const uint32_t ex = it.mbx + (uint32_t)(b % it.mcols), ey = it.mby + (uint32_t)(b / it.mcols);
                P.keys[o] = ey * P.ntx + ex;
                P.vals[o] = t;
                o++;
            }
        }
        return;
    }
    for (int ey = it.by0; ey <= it.by1; ey++)
        for (int ex = it.bx0; ex <= it.bx1; ex++) {
            P.keys[o] = (uint32_t)ey * P.ntx + (uint32_t)ex;
            P.vals[o] = t;
            o++;
        }
}

// ===========================================================================
// Points (pipeline/stages/rasterization/point.rs:21-86)
// ===========================================================================
// A point is rasterised in the screen tile whose half-open box
// [tx0, tx1) x [ty0, ty1) holds it (point.rs:46).  The boxes of consecutive
// tiles share only their closed ends, so a point lands in at most one tile;
// x >= W-1 or y >= H-1 never lands (the last tile ends at W-1 / H-1).
__global__ __launch_bounds__(256) void k_point_setup(PointSetupParams P) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= P.n) return;
    const uint32_t i = prim_vertex(P.idx, t);
    const float x = P.v.comp(0)[i], y = P.v.comp(1)[i];
    const bool in = x >= 0.0f && x < (float)(P.W - 1) && y >= 0.0f && y < (float)(P.H - 1);
    if (P.pass == 0) {
        P.counts[t] = in ? 1u : 0u;
        return;
    }
    if (!in) return;
    const uint32_t px = (uint32_t)x, py = (uint32_t)y;  // point.rs:47
    const uint32_t o = P.offsets[t];
    P.keys[o] = (py / P.ets) * P.ntx + px / P.ets;
    P.vals[o] = t;
}

// ===========================================================================
// Geometry stage: normal visualisation (full_example/src/shaders.rs:35-89)
// ===========================================================================
// The world position (k[3..6), w = 1) and normal (k[0..3), w = 0) varyings
// of SR_VP_MVP_NORMAL_WORLD become lines in clip space, G[0..16) times the
// Vector4 (k = 0..3 accumulated from +0.0, like the vertex stage), then
// normalised to the screen like finish().  The line vertices carry the
// varyings of their vertex (vertex normals) or of the centroid (face normals).
__device__ __forceinline__ void pv_times(const float* G, const float v[4], float out[4]) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) acc = acc + G[r * 4 + k] * v[k];
        out[r] = acc;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_gs_normals(GSParams P) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= P.n_in) return;
    CV v[3];
    for (int j = 0; j < 3; j++) load_cv(P.in, prim_vertex(P.idx, 3 * t + j), 6, v[j]);
    const float L = P.G[16];
    if constexpr (MODE == SR_GS_VERTEX_NORMALS) {
        // shaders.rs:46-56: start = mv * position, end = mv * (position + normal * L)
        for (int j = 0; j < 3; j++) {
            const float p[4] = {v[j].k[3], v[j].k[4], v[j].k[5], 1.0f};
            const float n[4] = {v[j].k[0], v[j].k[1], v[j].k[2], 0.0f};
            float e[4];
#pragma unroll
            for (int c = 0; c < 4; c++) e[c] = p[c] + n[c] * L;
            CV a = v[j], b = v[j];
            pv_times(P.G, p, a.p);
            pv_times(P.G, e, b.p);
            store_screen(P, 6 * t + 2 * j, a);
            store_screen(P, 6 * t + 2 * j + 1, b);
        }
    } else {
        // shaders.rs:72-85: center = barycentric_interpolate(1/3, a, 1/3, b,
        // 1/3, c), componentwise ((a * t) + (b * t)) + (c * t); the normal is
        // normalised over all four components
        const float third = 1.0f / 3.0f;
        CV m;
#pragma unroll
        for (int c = 0; c < 6; c++) m.k[c] = ((v[0].k[c] * third) + (v[1].k[c] * third)) + (v[2].k[c] * third);
        const float pw = ((1.0f * third) + (1.0f * third)) + (1.0f * third);
        const float nw = ((0.0f * third) + (0.0f * third)) + (0.0f * third);
        const float p[4] = {m.k[3], m.k[4], m.k[5], pw};
        float n[4] = {m.k[0], m.k[1], m.k[2], nw};
        const float len = sqrtf(((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]) + n[3] * n[3]);
        float e[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            n[c] = n[c] / len;
            e[c] = p[c] + n[c] * L;
        }
        CV a = m, b = m;
        pv_times(P.G, p, a.p);
        pv_times(P.G, e, b.p);
        store_screen(P, 2 * t, a);
        store_screen(P, 2 * t + 1, b);
    }
}

// ===========================================================================
// Host-side launchers
// ===========================================================================

void launch_finish(const VertexBuf& in, const VertexBuf& out, uint32_t n, int K, const sr_viewport& vp,
                   hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_finish, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, out, n, K, vp);
}

void launch_vertex(const VSParams& P, hipStream_t s) {
    if (P.n) hipLaunchKernelGGL(k_vertex, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P);
}

void launch_gs_count(int mode, int prim, const GSParams& P, hipStream_t s) {
    if (!P.n_in) return;
    const dim3 g(blocks_for(P.n_in, 256)), b(256);
#define GSC(M, PR) if (mode == M && prim == PR) { hipLaunchKernelGGL((k_gs_count<M, PR>), g, b, 0, s, P); return; }
    GSC(SR_GS_CLIP_NEAR, SR_PRIM_TRIANGLES) GSC(SR_GS_CLIP_NEAR, SR_PRIM_LINES) GSC(SR_GS_CLIP_NEAR, SR_PRIM_POINTS)
    GSC(SR_GS_CLIP_REFERENCE, SR_PRIM_TRIANGLES) GSC(SR_GS_CLIP_REFERENCE, SR_PRIM_LINES)
    GSC(SR_GS_CLIP_REFERENCE, SR_PRIM_POINTS)
#undef GSC
}

void launch_gs_emit(int mode, int prim, const GSParams& P, hipStream_t s) {
    if (!P.n_in) return;
    const dim3 g(blocks_for(P.n_in, 256)), b(256);
#define GSE(M, PR) if (mode == M && prim == PR) { hipLaunchKernelGGL((k_gs_emit<M, PR>), g, b, 0, s, P); return; }
    GSE(SR_GS_CLIP_NEAR, SR_PRIM_TRIANGLES) GSE(SR_GS_CLIP_NEAR, SR_PRIM_LINES) GSE(SR_GS_CLIP_NEAR, SR_PRIM_POINTS)
    GSE(SR_GS_CLIP_REFERENCE, SR_PRIM_TRIANGLES) GSE(SR_GS_CLIP_REFERENCE, SR_PRIM_LINES)
    GSE(SR_GS_CLIP_REFERENCE, SR_PRIM_POINTS)
#undef GSE
}

void launch_gs_normals(int mode, const GSParams& P, hipStream_t s) {
    if (!P.n_in) return;
    const dim3 g(blocks_for(P.n_in, 256)), b(256);
    if (mode == SR_GS_VERTEX_NORMALS) hipLaunchKernelGGL((k_gs_normals<SR_GS_VERTEX_NORMALS>), g, b, 0, s, P);
    else hipLaunchKernelGGL((k_gs_normals<SR_GS_FACE_NORMALS>), g, b, 0, s, P);
}

void launch_tri_setup(const SetupParams& P, hipStream_t s) {
    if (P.n) hipLaunchKernelGGL(k_tri_setup, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P);
}
void launch_vertex_batch(const BatchTables& T, const VertexBuf& out, hipStream_t s) {
    if (T.nv) hipLaunchKernelGGL(k_vertex_batch, dim3(blocks_for(T.nv, 256)), dim3(256), 0, s, T, out);
}
void launch_tri_setup_batch(const BatchTables& T, const VertexBuf& v, TriRec* recs, TriBin* bins, TriVid* vids,
                            uint32_t* counts, uint32_t* err, hipStream_t s) {
    if (T.nt) hipLaunchKernelGGL(k_tri_setup_batch, dim3(blocks_for(T.nt, 256)), dim3(256), 0, s, T, v,
                                 recs, bins, vids, counts, err);
}
void launch_tri_emit_batch(const BatchTables& T, const TriRec* recs, const TriBin* bins, const uint32_t* counts,
                           const uint32_t* offsets, uint32_t cap, uint32_t* keys, uint32_t* vals,
                           uint32_t* overflow, hipStream_t s) {
    if (T.nt) hipLaunchKernelGGL(k_tri_emit_batch, dim3(blocks_for(T.nt, 256)), dim3(256), 0, s, T, recs,
                                 bins, counts, offsets, cap, keys, vals, overflow);
}
void launch_tri_emit(const EmitParams& P, hipStream_t s) {
    if (P.n) hipLaunchKernelGGL(k_tri_emit, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P);
}

uint32_t scan_tmp_words(uint32_t n) { return blocks_for(n, SCAN_TILE) + 1; }

void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, uint32_t* total,
                 hipStream_t s, const uint32_t* count) {
    const uint32_t nb = blocks_for(n, SCAN_TILE);
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, total, 0u, 1u);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, in, n, tmp, count);
    if (nb <= SCAN_FUSED_BLOCKS) {
        hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, s, in, out, n, (const uint32_t*)nullptr, count,
                           (const uint32_t*)tmp, total);
        return;
    }
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, tmp, n, total, count);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, s, in, out, n, (const uint32_t*)tmp, count,
                       (const uint32_t*)nullptr, (uint32_t*)nullptr);
}

uint32_t radix_hist_words(uint32_t n) { return 256u * blocks_for(n, SCAN_TILE); }

int radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
               int key_bits, uint32_t* hist, uint32_t* scan_tmp, uint32_t* total, hipStream_t s,
               const uint32_t* n_dev) {
    if (n <= 1) return 0;
    const uint32_t nb = blocks_for(n, SCAN_TILE);
    const int passes = (key_bits + 7) / 8;
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    const SortSegs none{nullptr, nullptr, nullptr, nullptr, 0u};
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, s, ki, n, shift, hist, nb, n_dev, total + 1, none);
        launch_scan(hist, hist, 256u * nb, scan_tmp, total, s, n_dev ? total + 1 : nullptr);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, s, ki, vi, n, shift,
                           (const uint32_t*)hist, nb, ko, vo, n_dev, none);
        uint32_t* t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    return passes & 1;  // 1: result is in the tmp buffers
}

int radix_sort_groups(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t cap,
                      int local_key_bits, const BatchGroups& G, const uint32_t* tri_offsets,
                      const uint32_t* live, uint32_t* segs, uint32_t* hist, uint32_t* scan_tmp,
                      uint32_t* total, hipStream_t s) {
    if (cap <= 1 || G.ngroups == 0) return 0;
    // at most one partial block per group beyond the full ones
    const uint32_t nb = blocks_for(cap, SCAN_TILE) + G.ngroups;
    const int passes = (local_key_bits + 7) / 8;
    uint32_t* first = segs;
    uint32_t* lo = first + G.ngroups + 1;
    uint32_t* hi = lo + G.ngroups;
    uint32_t* kb = hi + G.ngroups;
    uint32_t* hist_count = kb + G.ngroups;
    hipLaunchKernelGGL(k_seg_table, dim3(1), dim3(256), 0, s, G, tri_offsets, live, first, lo, hi, kb, hist_count);
    const SortSegs S{first, lo, hi, kb, G.ngroups};
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, s, ki, cap, shift, hist, nb, live, hist_count, S);
        launch_scan(hist, hist, 256u * nb, scan_tmp, total, s, hist_count);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, s, ki, vi, cap, shift,
                           (const uint32_t*)hist, nb, ko, vo, live, S);
        uint32_t* t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    return passes & 1;
}

void launch_fill(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_pattern(void* p, const uint32_t pat[4], uint32_t period, size_t n_words, hipStream_t s) {
    if (n_words)
        hipLaunchKernelGGL(k_fill_pattern, dim3((uint32_t)((n_words + 255) / 256)), dim3(256), 0, s, (uint32_t*)p,
                           make_uint4(pat[0], pat[1], pat[2], pat[3]), period, n_words);
}

void launch_iota(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_iota_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
}

void launch_ranges_search(const uint32_t* keys, uint32_t n, uint32_t nkeys, uint32_t* start, uint32_t* end,
                          hipStream_t s, const uint32_t* n_dev) {
    hipLaunchKernelGGL(k_ranges_search, dim3(blocks_for((uint64_t)nkeys + 1, 256)), dim3(256), 0, s, keys, n, nkeys,
                       start, end, n_dev);
}

void launch_ranges(const uint32_t* keys, uint32_t n, uint32_t key_base, uint32_t nkeys, uint32_t* start,
                   uint32_t* end, hipStream_t s, const uint32_t* n_dev) {
    const int aligned = ((uintptr_t)keys & 15u) == 0;
    if (n) hipLaunchKernelGGL(k_ranges, dim3(std::min<uint32_t>(blocks_for(n, 1024), 8192u)), dim3(256), 0, s,
                              keys, n, key_base, nkeys, start, end, n_dev, aligned);
}
void launch_bins_live(uint32_t* misc, uint32_t cap, hipStream_t s) {
    hipLaunchKernelGGL(k_bins_live, dim3(1), dim3(64), 0, s, misc, cap);
}

void launch_live_pairs(uint32_t* misc, int overflow_word, int total_word, int live_word, hipStream_t s) {
    hipLaunchKernelGGL(k_live_pairs, dim3(1), dim3(64), 0, s, misc, overflow_word, total_word, live_word);
}

void launch_tile_cost(const uint32_t* start, const uint32_t* end, uint32_t n, uint32_t fixed, uint32_t* cost,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_tile_cost, dim3(blocks_for((uint64_t)n + 1, 256)), dim3(256), 0, s, start, end, n,
                       fixed, cost);
}

void launch_partition(const uint32_t* prefix, uint32_t tile0, uint32_t n, uint32_t G, uint32_t* wg_start,
                      hipStream_t s) {
    launch_fill(wg_start, n, G + 1, s);
    if (n) hipLaunchKernelGGL(k_partition, dim3(blocks_for(n, 256)), dim3(256), 0, s, prefix, tile0, n, G,
                              wg_start);
}

void launch_fb_clear(uint32_t* color, float* depth, uint8_t* stencil, uint32_t n, uint32_t rgba,
                     float d, uint32_t st, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fb_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, color, depth,
                              stencil, n, rgba, d, st);
}

void launch_raster_tris(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    switch (P.one.ets) {
        case 8: launch_raster_tris_8(P, ntiles, s); break;
        case 32: launch_raster_tris_32(P, ntiles, s); break;
        default: launch_raster_tris_16(P, ntiles, s); break;
    }
}

void launch_strips(const StripParams& P, hipStream_t s) {
    const uint32_t HS = P.H + 1, WS = P.W + 1, CS = P.scols + 1, RS = P.srows + 1;
    const uint32_t G = P.hot_part ? strip_hot_groups(P) : 0u;
    StripParams Q = P;
    if (G == 0) Q.hot_part = nullptr;
    const size_t lds_bytes = sizeof(uint32_t) * (strip_corner_words(Q) + (G ? HS + WS : 0u));
    if (P.n) hipLaunchKernelGGL(k_strips_accum, dim3(G ? G : blocks_for(P.n, 256)), dim3(256), lds_bytes, s, Q);
    if (G)
        hipLaunchKernelGGL(k_strips_hot_reduce, dim3(blocks_for(HS + WS, 256), blocks_for(G, STRIP_REDUCE_ROWS)),
                           dim3(256), 0, s, Q, G);
    // colL/colR: prefix along pixel rows (per tile column), then across tile columns
    hipLaunchKernelGGL(k_prefix_rows, dim3(CS), dim3(256), 0, s, P.colL, CS, HS, HS, 1u);
    hipLaunchKernelGGL(k_prefix_rows, dim3(HS), dim3(256), 0, s, P.colL, HS, CS, 1u, HS);
    hipLaunchKernelGGL(k_prefix_rows, dim3(CS), dim3(256), 0, s, P.colR, CS, HS, HS, 1u);
    hipLaunchKernelGGL(k_prefix_rows, dim3(HS), dim3(256), 0, s, P.colR, HS, CS, 1u, HS);
    hipLaunchKernelGGL(k_prefix_rows, dim3(RS), dim3(256), 0, s, P.rowT, RS, WS, WS, 1u);
    hipLaunchKernelGGL(k_prefix_rows, dim3(WS), dim3(256), 0, s, P.rowT, WS, RS, 1u, WS);
    hipLaunchKernelGGL(k_prefix_rows, dim3(RS), dim3(256), 0, s, P.rowB, RS, WS, WS, 1u);
    hipLaunchKernelGGL(k_prefix_rows, dim3(WS), dim3(256), 0, s, P.rowB, WS, RS, 1u, WS);
    uint32_t* corners[4] = {P.cLL, P.cLR, P.cRL, P.cRR};
    for (uint32_t* c : corners) {
        hipLaunchKernelGGL(k_prefix_rows, dim3(RS), dim3(256), 0, s, c, RS, CS, CS, 1u);
        hipLaunchKernelGGL(k_prefix_rows, dim3(CS), dim3(256), 0, s, c, CS, RS, 1u, CS);
    }
    hipLaunchKernelGGL(k_strips_final, dim3(blocks_for((uint64_t)P.W * P.H, 256)), dim3(256), 0, s, P);
}

void launch_line_tiles(const LineSetupParams& P, int pass, hipStream_t s) {
    if (P.n_lines) hipLaunchKernelGGL(k_line_tiles, dim3(blocks_for(P.n_lines, 256)), dim3(256), 0, s, P, pass);
}
void launch_line_walk(const WalkParams& P, hipStream_t s) {
    if (P.n_items) hipLaunchKernelGGL(k_line_walk, dim3(blocks_for(P.n_items, 256)), dim3(256), 0, s, P);
}
void launch_item_emit(const ItemEmitParams& P, hipStream_t s) {
    if (P.n) hipLaunchKernelGGL(k_item_emit, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P);
}
void launch_point_setup(const PointSetupParams& P, hipStream_t s) {
    if (P.n) hipLaunchKernelGGL(k_point_setup, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P);
}

}  // namespace sr
