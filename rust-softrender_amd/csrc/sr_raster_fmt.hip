// sr_raster_fmt.hip — the raster for every framebuffer format
// (sr_fb_create_ex: up to four colour planes of RGBA8 or RGBAF32 colour, f32
// or f64 depth, u8 or `()` stencil), for the multi-output program and for
// the normalised over-blend.  It takes the binned tile lists of an 8x8
// execution tile like the main raster (or every triangle, in the exhaustive
// mode; a triangle with the strip flag is tested in its clamped boxes in
// either) and runs the reference's per-pixel sequence (triangle.rs:82-155)
// with each lane owning one pixel: stencil test and op in order, coverage,
// depth, then the fragment program and the blend of every plane, at once
// and in order.  No shading queue: this path is for the formats, not for
// throughput.
#include "sr_kernels.hpp"
#include "sr_raster_common.hpp"

namespace sr {

template <int FP>
__device__ __forceinline__ void shade_varyings(const RasterParams& P, uint32_t tri, float u, float v,
                                               float* k) {
    constexpr int KV = fp_varyings<FP>();
    if constexpr (KV > 0) {
        uint32_t i0 = 3u * tri, i1 = i0 + 1u, i2 = i0 + 2u;  // no indices: 3t, 3t+1, 3t+2
        if (P.vids) {
            const TriVid& r = P.vids[tri];
            i0 = r.i0; i1 = r.i1; i2 = r.i2;
        }
        const float w = (1.0f - u) - v;
        const float *r0 = P.verts.var(i0), *r1 = P.verts.var(i1), *r2 = P.verts.var(i2);
#pragma unroll
        for (int c = 0; c < KV; c++) k[c] = ((r0[c] * u) + (r1[c] * v)) + (r2[c] * w);  // interpolate.rs:43-64
    }
}

template <int FP, int BLEND>
__global__ __launch_bounds__(64) void k_raster_fmt(RasterParams P, FmtParams F) {
    __shared__ float lut_s[256];
    if (P.skip_if && ((load_item(P.skip_if, 0) & 1u) | load_item(P.skip_if, 4))) return;  // NaN, overflow
    for (int i = threadIdx.x; i < 256; i += 64) lut_s[i] = (float)i / 255.0f;
    __syncthreads();
    const Lut lut{lut_s};
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t beg = load_item(P.tile_start, tile), end = load_item(P.tile_end, tile);
    if (beg >= end) return;
    const int lane = threadIdx.x;
    const uint32_t etx = tile % P.ntx, ety = tile / P.ntx;
    const int wx0 = (int)(etx * 8), wy0 = (int)(ety * 8);
    const int wx1 = min(wx0 + 7, (int)P.W - 1), wy1 = min(wy0 + 7, (int)P.H - 1);
    const int px = wx0 + (lane & 7), py = wy0 + (lane >> 3);
    const bool live = px < (int)P.W && py < (int)P.H;
    const size_t idx = live ? (size_t)py * P.W + px : 0;

    // the pixel's planes; an RGBA8 plane keeps its packed value in c[p][0]
    float c[SR_MAX_COLOR_PLANES][4] = {};
#pragma unroll
    for (int p = 0; p < SR_MAX_COLOR_PLANES; p++) {
        if (p >= (int)F.n_color || !live) continue;
        if (F.color_fmt[p] == SR_COLOR_RGBAF32) {
            const float4 v = ((const float4*)F.planes[p])[idx];
            c[p][0] = v.x; c[p][1] = v.y; c[p][2] = v.z; c[p][3] = v.w;
        } else {
            c[p][0] = __builtin_bit_cast(float, ((const uint32_t*)F.planes[p])[idx]);
        }
    }
    double dep = live ? (F.depth64 ? F.depth64[idx] : (double)P.depth[idx]) : 0.0;
    uint32_t st = (live && F.has_stencil) ? P.stencil[idx] : 0u;
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;  // triangle.rs:104-105

    uint32_t jlo, jhi, ilo, ihi;
    sem_range((uint32_t)wx0, (uint32_t)wx1, P.tw, P.scols, jlo, jhi);
    sem_range((uint32_t)wy0, (uint32_t)wy1, P.th, P.srows, ilo, ihi);
    const uint32_t xmax = P.W - 1, ymax = P.H - 1;
    for (uint32_t si = ilo; si <= ihi; si++) {
        const int Ty0 = (int)(si * P.th), Ty1 = (int)min(si * P.th + P.th, ymax);
        for (uint32_t sj = jlo; sj <= jhi; sj++) {
            const int Tx0 = (int)(sj * P.tw), Tx1 = (int)min(sj * P.tw + P.tw, xmax);
            if (max(wx0, Tx0) > min(wx1, Tx1) || max(wy0, Ty0) > min(wy1, Ty1)) continue;
            const bool inT = live && px >= Tx0 && px <= Tx1 && py >= Ty0 && py <= Ty1;
            for (uint32_t it = beg; it < end; it++) {
                const uint32_t id = load_item(P.items, it);
                const RecLane rc = rec_s(P.tris, id);
                int cx0 = rc.X0, cx1 = rc.X1, cy0 = rc.Y0, cy1 = rc.Y1;
                if (P.exhaustive || (rc.zs & 1u)) {  // every box, or one a pixel outside may pass
                    if (rc.X0 > rc.X1) continue;  // culled or invalid
                    // clamp_as_int! into the screen tile, triangle.rs:66-78
                    cx0 = min(max(rc.X0, Tx0), Tx1); cx1 = min(max(rc.X1, Tx0), Tx1);
                    cy0 = min(max(rc.Y0, Ty0), Ty1); cy1 = min(max(rc.Y1, Ty0), Ty1);
                }
                bool ok = inT && px >= cx0 && px <= cx1 && py >= cy0 && py <= cy1;
                if (F.track_stencil) {  // stencil before coverage, triangle.rs:90-99
                    const bool sp = ok && stencil_test(P.s_test, st, P.s_ref);
                    st = sp ? stencil_op(P.s_op, P.s_wrap, st, P.s_ref) : st;
                    ok = sp;
                }
                const float dx = fx - rc.x3, dy = fy - rc.y3;
                const float un = (rc.e23y * dx) + (rc.e23x * dy);  // triangle.rs:108
                const float vn = (rc.e31y * dx) + (rc.e31x * dy);  // triangle.rs:109
                const float u = quot(un, rc.R), v = quot(vn, rc.R);
                const float w = (1.0f - u) - v;                    // triangle.rs:110
                const float z = ((rc.z1 * u) + (rc.z2 * v)) + (rc.z3 * w);
                // Depth::from_scalar(z) is exact for f64 (depth.rs:26-41), so
                // z >= dt compares the same values either way
                if (!(ok && !(u < 0.0f) && !(v < 0.0f) && !(w < 0.0f) && z < 0.0f && (double)z >= dep)) continue;
                dep = (double)z;
                float k[8];
                shade_varyings<FP>(P, id, u, v, k);
                float o[FP == SR_FP_GBUFFER ? 4 : 1][4];
                constexpr int NO = FP == SR_FP_GBUFFER ? 4 : 1;
                fragment_rgba<FP>(P, k, lut, o);
#pragma unroll
                for (int p = 0; p < SR_MAX_COLOR_PLANES; p++) {
                    if (p >= (int)F.n_color) continue;
                    const float* src = o[p < NO ? p : NO - 1];
                    if (F.color_fmt[p] == SR_COLOR_RGBAF32) {
                        float out[4];
                        blend_f32<BLEND>(src, c[p], out);
#pragma unroll
                        for (int q = 0; q < 4; q++) c[p][q] = out[q];
                    } else {
                        const uint32_t d = __builtin_bit_cast(uint32_t, c[p][0]);
                        c[p][0] = __builtin_bit_cast(float, blend<BLEND>(pack_q(src), d, lut));
                    }
                }
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int p = 0; p < SR_MAX_COLOR_PLANES; p++) {
        if (p >= (int)F.n_color) continue;
        if (F.color_fmt[p] == SR_COLOR_RGBAF32)
            ((float4*)F.planes[p])[idx] = make_float4(c[p][0], c[p][1], c[p][2], c[p][3]);
        else
            ((uint32_t*)F.planes[p])[idx] = __builtin_bit_cast(uint32_t, c[p][0]);
    }
    if (F.depth64) F.depth64[idx] = dep;
    else P.depth[idx] = (float)dep;
    if (F.has_stencil) P.stencil[idx] = (uint8_t)st;
}

template <int FP>
static void raster_fmt_blend(const RasterParams& P, const FmtParams& F, uint32_t ntiles, hipStream_t s) {
    const dim3 g(ntiles), b(64);
    if (P.blend == SR_BLEND_OVER) hipLaunchKernelGGL((k_raster_fmt<FP, SR_BLEND_OVER>), g, b, 0, s, P, F);
    else if (P.blend == SR_BLEND_OVER_NORMALIZED)
        hipLaunchKernelGGL((k_raster_fmt<FP, SR_BLEND_OVER_NORMALIZED>), g, b, 0, s, P, F);
    else hipLaunchKernelGGL((k_raster_fmt<FP, SR_BLEND_REPLACE>), g, b, 0, s, P, F);
}

void launch_raster_fmt(const RasterParams& P, const FmtParams& F, uint32_t ntiles, hipStream_t s) {
    if (!ntiles) return;
    switch (P.fp) {
        case SR_FP_BLINN_PHONG: raster_fmt_blend<SR_FP_BLINN_PHONG>(P, F, ntiles, s); break;
        case SR_FP_TEX_LAMBERT: raster_fmt_blend<SR_FP_TEX_LAMBERT>(P, F, ntiles, s); break;
        case SR_FP_UV_COLOR: raster_fmt_blend<SR_FP_UV_COLOR>(P, F, ntiles, s); break;
        case SR_FP_GBUFFER: raster_fmt_blend<SR_FP_GBUFFER>(P, F, ntiles, s); break;
        case SR_FP_FRESNEL4: raster_fmt_blend<SR_FP_FRESNEL4>(P, F, ntiles, s); break;
        case SR_FP_INT_PALETTE: raster_fmt_blend<SR_FP_INT_PALETTE>(P, F, ntiles, s); break;
        default: raster_fmt_blend<SR_FP_FLAT>(P, F, ntiles, s); break;
    }
}

}  // namespace sr
