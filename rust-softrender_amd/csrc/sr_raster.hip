// sr_raster.hip — the raster work-groups.  Compiled once per execution
// tile size (-DSR_ETS=8|16|32|64) so the template instances build in
// parallel; each object exports launch_raster_*_<ETS>.
#include "sr_device.hpp"
#include "sr_kernels.hpp"

#ifndef SR_ETS
#define SR_ETS 16
#endif

namespace sr {

// ===========================================================================
// Raster: triangles
// ===========================================================================
// Screen tile grid of the reference (fragment.rs:188-216): column j spans
// [j*tw, min((j+1)*tw, W-1)], rows likewise; both ends inclusive.
__device__ __forceinline__ void sem_range(uint32_t lo_px, uint32_t hi_px, uint32_t t, uint32_t n,
                                          uint32_t& jlo, uint32_t& jhi) {
    // tiles whose inclusive span meets [lo_px, hi_px]
    const uint32_t a = (lo_px + t - 1) / t;  // ceil(lo/t)
    jlo = a > 0 ? a - 1 : 0;
    jhi = hi_px / t;
    if (jhi > n - 1) jhi = n - 1;
}

// Pixel ownership: an ETS x ETS execution tile, 256 threads (64 for ETS 8),
// ETS*ETS/threads pixels per thread.  Wave w owns rows [w*ETS/4, (w+1)*ETS/4)
// so a triangle whose box misses those rows is skipped by the whole wave.
template <int ETS>
constexpr int raster_threads() { return ETS == 8 ? 64 : 256; }

template <int ETS, int FP, int BLEND, int SMODE>
__global__ __launch_bounds__(256) void k_raster_tris(RasterParams P) {
    constexpr int NT = raster_threads<ETS>();
    constexpr int PPT = (ETS * ETS) / NT;
    constexpr int KV = fp_varyings<FP>();
    constexpr int CH = NT;  // triangle records staged per chunk

    __shared__ TriRec srec[CH];
    __shared__ float slut[256];

    const int tid = threadIdx.x;
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t etx = tile % P.ntx, ety = tile / P.ntx;
    uint32_t beg, end;
    if (P.exhaustive_n) {
        beg = 0;
        end = P.exhaustive_n;
    } else {
        beg = P.tile_start[tile];
        end = P.tile_end[tile];
    }
    if (beg >= end && SMODE != 1) return;

    for (int i = tid; i < 256; i += NT) slut[i] = (float)i / 255.0f;
    const Lut lut{slut};

    // tile pixel span (clipped to the framebuffer)
    const uint32_t tx0 = etx * ETS, ty0 = ety * ETS;
    const uint32_t tx1 = min(tx0 + ETS, P.W) - 1, ty1 = min(ty0 + ETS, P.H) - 1;

    // this thread's pixels
    int px[PPT], py[PPT];
    bool live[PPT];
    uint32_t col[PPT], cnt[PPT], st[PPT];
    float dep[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        int lx, ly;
        if constexpr (ETS == 8) {
            lx = tid & 7; ly = tid >> 3;
        } else {
            const int wave = tid >> 6, lane = tid & 63;
            constexpr int RPW = ETS / 4;                 // rows per wave
            constexpr int RPP = 64 / ETS > 0 ? 64 / ETS : 1;  // rows per pass
            if constexpr (ETS <= 64) {
                lx = lane % ETS;
                ly = wave * RPW + (lane / ETS) + j * RPP;
            }
        }
        px[j] = (int)tx0 + lx;
        py[j] = (int)ty0 + ly;
        live[j] = (uint32_t)px[j] < P.W && (uint32_t)py[j] < P.H;
        const size_t idx = (size_t)py[j] * P.W + px[j];
        col[j] = live[j] ? P.color[idx] : 0u;
        dep[j] = live[j] ? P.depth[idx] : 0.0f;
        st[j] = (SMODE == 2 && live[j]) ? P.stencil[idx] : 0u;
        cnt[j] = 0;
    }
    // this wave's pixel rows
    int wy0, wy1;
    if constexpr (ETS == 8) {
        wy0 = (int)ty0; wy1 = (int)ty1;
    } else {
        const int wave = tid >> 6;
        wy0 = (int)ty0 + wave * (ETS / 4);
        wy1 = min((int)ty0 + (wave + 1) * (ETS / 4) - 1, (int)ty1);
    }
    const int wx0 = (int)tx0, wx1 = (int)tx1;

    // screen tiles meeting this execution tile, in the reference's row-major order
    uint32_t jlo, jhi, ilo, ihi;
    sem_range(tx0, tx1, P.tw, P.scols, jlo, jhi);
    sem_range(ty0, ty1, P.th, P.srows, ilo, ihi);
    const uint32_t xmax = P.W - 1, ymax = P.H - 1;

    for (uint32_t si = ilo; si <= ihi; si++) {
        const int Ty0 = (int)(si * P.th), Ty1 = (int)min(si * P.th + P.th, ymax);
        for (uint32_t sj = jlo; sj <= jhi; sj++) {
            const int Tx0 = (int)(sj * P.tw), Tx1 = (int)min(sj * P.tw + P.tw, xmax);
            // this wave's rows/cols inside the screen tile
            const int rx0 = max(wx0, Tx0), rx1 = min(wx1, Tx1);
            const int ry0 = max(wy0, Ty0), ry1 = min(wy1, Ty1);
            const bool wave_in = rx0 <= rx1 && ry0 <= ry1;
            bool inT[PPT];
#pragma unroll
            for (int j = 0; j < PPT; j++)
                inT[j] = live[j] && px[j] >= Tx0 && px[j] <= Tx1 && py[j] >= Ty0 && py[j] <= Ty1;

            for (uint32_t c0 = beg; c0 < end; c0 += CH) {
                const uint32_t nc = min((uint32_t)CH, end - c0);
                __syncthreads();
                if ((uint32_t)tid < nc) {
                    const uint32_t item = P.exhaustive_n ? c0 + tid : P.items[c0 + tid];
                    srec[tid] = P.tris[item];
                }
                __syncthreads();
                if (!wave_in) continue;
                for (uint32_t q = 0; q < nc; q++) {
                    const TriRec& r = srec[q];
                    // tile-clamped bbox, clamp_as_int!, triangle.rs:66-78
                    const int X0 = __builtin_amdgcn_readfirstlane(r.X0);
                    const int X1 = __builtin_amdgcn_readfirstlane(r.X1);
                    const int Y0 = __builtin_amdgcn_readfirstlane(r.Y0);
                    const int Y1 = __builtin_amdgcn_readfirstlane(r.Y1);
                    if (X0 > X1) continue;  // culled or invalid
                    const int cx0 = min(max(X0, Tx0), Tx1), cx1 = min(max(X1, Tx0), Tx1);
                    const int cy0 = min(max(Y0, Ty0), Ty1), cy1 = min(max(Y1, Ty0), Ty1);
                    if (cx1 < rx0 || cx0 > rx1 || cy1 < ry0 || cy0 > ry1) continue;
                    const float x3 = r.x3, y3 = r.y3, e23y = r.e23y, e23x = r.e23x,
                                e31y = r.e31y, e31x = r.e31x;
                    const double R = r.R;
                    const float z1 = r.z1, z2 = r.z2, z3 = r.z3;
#pragma unroll
                    for (int j = 0; j < PPT; j++) {
                        if (!(inT[j] && px[j] >= cx0 && px[j] <= cx1 && py[j] >= cy0 &&
                              py[j] <= cy1))
                            continue;
                        if constexpr (SMODE == 2) {  // stencil before coverage, triangle.rs:90-99
                            if (!stencil_test(P.s_test, st[j], P.s_ref)) continue;
                            st[j] = stencil_op(P.s_op, P.s_wrap, st[j], P.s_ref);
                        } else if constexpr (SMODE == 1) {
                            if (px[j] >= X0 && px[j] <= X1 && py[j] >= Y0 && py[j] <= Y1) cnt[j]++;
                        }
                        const float x = (float)px[j] + 0.5f, y = (float)py[j] + 0.5f;
                        const float dx = x - x3, dy = y - y3;
                        const float un = (e23y * dx) + (e23x * dy);  // triangle.rs:108
                        const float vn = (e31y * dx) + (e31x * dy);  // triangle.rs:109
                        const float u = quot(un, R), v = quot(vn, R);
                        const float w = (1.0f - u) - v;              // triangle.rs:110
                        if (u < 0.0f || v < 0.0f || w < 0.0f) continue;  // :113
                        const float z = ((z1 * u) + (z2 * v)) + (z3 * w);  // :115-117
                        if (!(z < 0.0f && z >= dep[j])) continue;          // :120-126
                        float k[KV > 0 ? KV : 1];
                        if constexpr (KV > 0) {
                            const uint32_t i0 = r.i0, i1 = r.i1, i2 = r.i2;
#pragma unroll
                            for (int c = 0; c < KV; c++) {
                                const float* vc = P.verts.comp(4 + c);
                                k[c] = ((vc[i0] * u) + (vc[i1] * v)) + (vc[i2] * w);
                            }
                        }
                        const uint32_t fc = fragment<FP>(P, k, lut);
                        col[j] = blend<BLEND>(fc, col[j], lut);
                        dep[j] = z;
                    }
                }
            }
        }
    }
    // resolve: the tile goes back to HBM once
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        if (!live[j]) continue;
        const size_t idx = (size_t)py[j] * P.W + px[j];
        if (beg < end) {
            P.color[idx] = col[j];
            P.depth[idx] = dep[j];
        }
        if constexpr (SMODE == 2) {
            P.stencil[idx] = (uint8_t)st[j];
        } else if constexpr (SMODE == 1) {
            const uint32_t n = cnt[j] + (P.strip_counts ? P.strip_counts[idx] : 0u);
            if (n) P.stencil[idx] = (uint8_t)stencil_op_n(P.s_op, P.s_wrap, P.stencil[idx], P.s_ref, n);
        }
    }
}

template <int ETS, int FP, int BLEND>
static void raster_tris_smode(const RasterParams& P, uint32_t ntiles, hipStream_t s) {
    const dim3 g(ntiles), b(raster_threads<ETS>());
    if (P.smode == 2) hipLaunchKernelGGL((k_raster_tris<ETS, FP, BLEND, 2>), g, b, 0, s, P);
    else if (P.smode == 1) hipLaunchKernelGGL((k_raster_tris<ETS, FP, BLEND, 1>), g, b, 0, s, P);
    else hipLaunchKernelGGL((k_raster_tris<ETS, FP, BLEND, 0>), g, b, 0, s, P);
}

template <int ETS, int FP>
static void raster_tris_blend(const RasterParams& P, uint32_t ntiles, hipStream_t s) {
    if (P.blend == SR_BLEND_OVER) raster_tris_smode<ETS, FP, SR_BLEND_OVER>(P, ntiles, s);
    else raster_tris_smode<ETS, FP, SR_BLEND_REPLACE>(P, ntiles, s);
}

template <int ETS>
static void raster_tris_fp(const RasterParams& P, uint32_t ntiles, hipStream_t s) {
    switch (P.fp) {
        case SR_FP_BLINN_PHONG: raster_tris_blend<ETS, SR_FP_BLINN_PHONG>(P, ntiles, s); break;
        case SR_FP_TEX_LAMBERT: raster_tris_blend<ETS, SR_FP_TEX_LAMBERT>(P, ntiles, s); break;
        case SR_FP_UV_COLOR: raster_tris_blend<ETS, SR_FP_UV_COLOR>(P, ntiles, s); break;
        default: raster_tris_blend<ETS, SR_FP_FLAT>(P, ntiles, s); break;
    }
}

#define SR_CAT2(a, b) a##b
#define SR_CAT(a, b) SR_CAT2(a, b)
void SR_CAT(launch_raster_tris_, SR_ETS)(const RasterParams& P, uint32_t ntiles, hipStream_t s) {
    if (ntiles) raster_tris_fp<SR_ETS>(P, ntiles, s);
}

}  // namespace sr
