// sr_raster_lp.hip — raster work-groups for line plots and points.
//
// One work-group per 16x16 execution tile; the tile's colour, depth and
// stencil live in LDS.  Each plot (or point) is one stencil/depth/shade/
// blend step on one pixel.  Per pixel those steps must run in the
// reference's order: screen tile by screen tile (row-major,
// fragment.rs:250-253), then line by line, then plot by plot.  Plots of a
// chunk run in parallel; two plots on the same pixel are ordered by claim
// rounds (atomicMin on the plot's rank: the earliest pending plot of each
// pixel goes first).
#include "sr_device.hpp"
#include "sr_kernels.hpp"

namespace sr {

constexpr int LP_ETS = 16;

__device__ __forceinline__ float lp_hypot(float a, float b) {
    // glibc hypotf: sqrt(x*x + y*y) in double, rounded once to float
    if (__builtin_isinf(a) || __builtin_isinf(b)) return __builtin_inff();
    const double x = (double)a, y = (double)b;
    return (float)sqrt(x * x + y * y);
}

__device__ __forceinline__ uint32_t block_min_256(uint32_t v, uint32_t* lds4) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint32_t r = min(min(lds4[0], lds4[1]), min(lds4[2], lds4[3]));
    __syncthreads();
    return r;
}

struct PixState {
    uint32_t* col;
    float* dep;
    uint32_t* st;
};

__device__ __forceinline__ void load_tile(const RasterParams& P, uint32_t etx, uint32_t ety,
                                          PixState s, int& gx, int& gy, bool& live) {
    const int t = threadIdx.x;
    gx = (int)(etx * LP_ETS) + (t % LP_ETS);
    gy = (int)(ety * LP_ETS) + (t / LP_ETS);
    live = (uint32_t)gx < P.W && (uint32_t)gy < P.H;
    const size_t idx = (size_t)gy * P.W + gx;
    s.col[t] = live ? P.color[idx] : 0u;
    s.dep[t] = live ? P.depth[idx] : 0.0f;
    s.st[t] = live ? P.stencil[idx] : 0u;
}

__device__ __forceinline__ void store_tile(const RasterParams& P, PixState s, int gx, int gy, bool live) {
    if (!live) return;
    const int t = threadIdx.x;
    const size_t idx = (size_t)gy * P.W + gx;
    P.color[idx] = s.col[t];
    P.depth[idx] = s.dep[t];
    P.stencil[idx] = (uint8_t)s.st[t];
}

// Runs `apply(rank)` for every pending rank in this chunk, same-pixel ranks
// in increasing order.
template <class F>
__device__ __forceinline__ void ordered_rounds(bool pending, int lp, uint32_t rank, uint32_t* claim,
                                               F&& apply) {
    while (__syncthreads_or(pending)) {
        if (pending) atomicMin(&claim[lp], rank);
        __syncthreads();
        const bool mine = pending && claim[lp] == rank;
        if (mine) {
            apply();
            pending = false;
        }
        __syncthreads();
        if (mine) claim[lp] = 0xFFFFFFFFu;
        __syncthreads();
    }
}

// A work-group OR with one barrier: three LDS flags (zero at the start of
// the kernel, before a barrier) used in turn; each use clears the next
// flag before its barrier, when every read of that flag (two uses back)
// has completed.  All threads call it the same number of times.
struct BlockOr {
    uint32_t* f;     // 3 LDS words
    uint32_t k = 0;  // uniform
    __device__ bool operator()(bool p) {
        const uint32_t nk = k == 2u ? 0u : k + 1u;
        if (p) f[k] = 1u;
        if (threadIdx.x == 0) f[nk] = 0u;
        __syncthreads();
        const bool r = f[k] != 0u;
        k = nk;
        return r;
    }
};

// The same for ranks that are the thread index (lines), with two barriers
// per round and no clearing: a claim is (round << 8) | (255 - thread), so
// every round's claims exceed all earlier ones and atomicMax leaves the
// lowest pending thread of each pixel.  `round` (uniform) carries across
// calls; claim[] starts at zero and restarts at zero before the 24-bit
// round number would wrap.
template <class F>
__device__ __forceinline__ void ordered_rounds_t(bool pending, int lp, uint32_t& round, uint32_t* claim,
                                                 BlockOr& any, F&& apply) {
    const uint32_t me = 255u - (uint32_t)threadIdx.x;
    auto bid = [&]() -> uint32_t {
        if (round == 0xFFFFFFu) {
            __syncthreads();
            claim[threadIdx.x] = 0u;
            __syncthreads();
            round = 0u;
        }
        round++;
        const uint32_t v = (round << 8) | me;
        if (pending) atomicMax(&claim[lp], v);
        return v;
    };
    uint32_t v = bid();
    if (!any(pending)) return;  // its barrier also makes the claims visible
    for (;;) {
        if (pending && claim[lp] == v) {
            apply();
            pending = false;
        }
        if (!any(pending)) return;  // claims read, pixels written
        v = bid();
        __syncthreads();
    }
}

// Exclusive prefix sum over the work-group's 256 threads (in thread order).
__device__ __forceinline__ uint32_t lp_scan(uint32_t v, uint32_t& total, uint32_t* lds4) {
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
        const uint32_t q = lds4[w];
        if (w < wave) off += q;
        total += q;
    }
    __syncthreads();
    return x - v + off;
}

// Lines: per screen tile (ascending), the tile's line items are staged in
// LDS (256 at a time, list = line order) with what their plots need, and
// the concatenated plot sequence of the staged items is processed 256 plots
// per pass: plot g belongs to the first staged item whose plot-count prefix
// exceeds g.  Plots of one pass run in parallel; same-pixel plots are
// ordered by their position in the sequence (ordered_rounds), which is the
// reference's line-then-plot order.
template <int FP, int BLEND>
__global__ __launch_bounds__(256) void k_raster_lines(RasterParams P) {
    constexpr int KV = fp_varyings<FP>();
    __shared__ uint32_t s_col[256], s_st[256], claim[256], lds4[4];
    __shared__ float s_dep[256], slut[256];
    __shared__ float i_cx1[256], i_cy1[256], i_dlen[256], i_zs[256], i_ze[256];
    __shared__ uint32_t i_v0[256], i_v1[256], i_off[256], i_end[256];  // i_end: inclusive plot prefix
    __shared__ uint32_t i_skip[256];  // interior plots before this tile's (ranged items)
    __shared__ uint32_t orf[3];
    const int t = threadIdx.x;
    uint32_t round = 0;  // ordered_rounds_t's claim rounds
    BlockOr any{orf};
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t etx = tile % P.ntx, ety = tile / P.ntx;
    const uint32_t beg = P.tile_start[tile], end = P.tile_end[tile];
    if (beg >= end) return;
    const PixState S{s_col, s_dep, s_st};
    int gx, gy;
    bool live;
    load_tile(P, etx, ety, S, gx, gy, live);
    slut[t] = (float)t / 255.0f;
    claim[t] = 0u;
    if (t < 3) orf[t] = 0u;
    __syncthreads();
    const Lut lut{slut};
    const int ex0 = (int)(etx * LP_ETS), ey0 = (int)(ety * LP_ETS);

    // screen tiles in ascending (row-major) order
    bool first = true;
    uint32_t prev = 0;
    for (;;) {
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t i = beg + t; i < end; i += 256) {
            const uint32_t T = P.line_items[P.items[i]].tile;
            if ((first || T > prev) && T < best) best = T;
        }
        best = block_min_256(best, lds4);
        if (best == 0xFFFFFFFFu) break;
        for (uint32_t ib = beg; ib < end; ib += 256) {
            // stage this screen tile's items among the next 256 of the list
            const uint32_t i = ib + (uint32_t)t;
            LineItem it{};
            bool mine = false;
            if (i < end) {
                it = P.line_items[P.items[i]];
                mine = it.tile == best && it.plot_cnt > 0;
            }
            uint32_t nstaged;
            const uint32_t slot = lp_scan(mine ? 1u : 0u, nstaged, lds4);
            if (nstaged == 0) continue;  // uniform
            // plots to scan: all, or for a ranged Wu item the 4 endpoint
            // plots and the interior pairs whose major coordinate lies in
            // this tile's span (the others fall outside the tile)
            uint32_t scan = mine ? it.plot_cnt : 0u, skip = 0;
            if (mine && it.ranged) {
                const int64_t M0 = it.steep ? ey0 : ex0, M1 = M0 + LP_ETS - 1;
                const int64_t nint = (int64_t)(it.plot_cnt - 4) / 2;
                const int64_t lo = max<int64_t>(M0 - it.major0, 0);
                const int64_t hi = max<int64_t>(min<int64_t>(M1 - it.major0 + 1, nint), lo);
                scan = 4u + 2u * (uint32_t)(hi - lo);
                skip = 2u * (uint32_t)lo;
            }
            uint32_t total;
            const uint32_t pstart = lp_scan(scan, total, lds4);
            if (mine) {
                const uint32_t v0 = P.line_v ? P.line_v[2 * it.line] : 2 * it.line;
                const uint32_t v1 = P.line_v ? P.line_v[2 * it.line + 1] : 2 * it.line + 1;
                i_cx1[slot] = it.cx1;
                i_cy1[slot] = it.cy1;
                i_dlen[slot] = it.dlen;
                i_zs[slot] = P.verts.comp(2)[v0];
                i_ze[slot] = P.verts.comp(2)[v1];
                i_v0[slot] = v0;
                i_v1[slot] = v1;
                i_off[slot] = it.plot_off;
                i_skip[slot] = skip;
                i_end[slot] = pstart + scan;
            }
            __syncthreads();
            for (uint32_t c = 0; c < total; c += 256) {
                const uint32_t g = c + (uint32_t)t;
                bool pending = false;
                int lp = 0;
                uint32_t si = 0;
                Plot pl{};
                if (g < total) {
                    uint32_t lo = 0, hi = nstaged - 1;  // first staged item with i_end > g
                    while (lo < hi) {
                        const uint32_t m = (lo + hi) >> 1;
                        if (i_end[m] > g) hi = m;
                        else lo = m + 1;
                    }
                    si = lo;
                    const uint32_t j = g - (si ? i_end[si - 1] : 0u);
                    pl = P.plots[i_off[si] + (j < 4u ? j : j + i_skip[si])];
                    const int py = (int)pl.row - ey0, px = (int)(pl.index - pl.row * P.W) - ex0;
                    if (px >= 0 && px < LP_ETS && py >= 0 && py < LP_ETS) {
                        pending = true;
                        lp = py * LP_ETS + px;
                    }
                }
                ordered_rounds_t(pending, lp, round, claim, any, [&]() {
                    // the rasterize_fragment closure, line.rs:54-109
                    const uint32_t sv = s_st[lp];
                    if (!stencil_test(P.s_test, sv, P.s_ref)) return;
                    s_st[lp] = stencil_op(P.s_op, P.s_wrap, sv, P.s_ref);
                    const float xf = (float)pl.x + 0.5f, yf = (float)pl.y + 0.5f;
                    const float tt = lp_hypot(i_cx1[si] - xf, i_cy1[si] - yf) / i_dlen[si];  // :75
                    const float omt = 1.0f - tt;
                    const float z = (omt * i_zs[si]) + (tt * i_ze[si]);                     // :77
                    if (!(z < 0.0f && z >= s_dep[lp])) return;
                    float kv[KV > 0 ? KV : 1];
                    const float *r0 = P.verts.var(i_v0[si]), *r1 = P.verts.var(i_v1[si]);
#pragma unroll
                    for (int q = 0; q < KV; q++) kv[q] = (omt * r0[q]) + (tt * r1[q]);
                    uint32_t col = fragment<FP>(P, kv, lut);
                    col = mul_alpha(P.alpha_mode, col, pl.alpha);                         // :100
                    s_col[lp] = blend<BLEND>(col, s_col[lp], lut);
                    s_dep[lp] = z;
                });
            }
            __syncthreads();  // the staging arrays are rewritten next
        }
        prev = best;
        first = false;
    }
    __syncthreads();
    store_tile(P, S, gx, gy, live);
}

// Points are sparse (a few per tile), so a work-group reads and writes the
// pixels its points land on in place rather than staging the whole tile:
// it alone touches its tile's pixels, and the claim rounds' barriers order
// same-pixel points.
template <int FP, int BLEND>
__global__ __launch_bounds__(256) void k_raster_points(RasterParams P) {
    constexpr int KV = fp_varyings<FP>();
    __shared__ uint32_t claim[256];
    __shared__ float slut[256];
    const int t = threadIdx.x;
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t etx = tile % P.ntx, ety = tile / P.ntx;
    const uint32_t beg = P.tile_start[tile], end = P.tile_end[tile];
    if (beg >= end) return;
    slut[t] = (float)t / 255.0f;
    claim[t] = 0xFFFFFFFFu;
    __syncthreads();
    const Lut lut{slut};
    const int ex0 = (int)(etx * LP_ETS), ey0 = (int)(ety * LP_ETS);
    // Every point on a pixel belongs to the same screen tile, so list
    // (= point) order is the reference's per-pixel order.
    for (uint32_t b = beg; b < end; b += 256) {
        const uint32_t k = b + t;
        bool pending = false;
        int lp = 0;
        uint32_t v = 0;
        float x = 0.f, y = 0.f;
        size_t idx = 0;
        if (k < end) {
            const uint32_t pt = P.items[k];
            v = P.point_v ? P.point_v[pt] : pt;
            x = P.verts.comp(0)[v];
            y = P.verts.comp(1)[v];
            const uint32_t px = (uint32_t)x, py = (uint32_t)y;                  // point.rs:47
            lp = ((int)py - ey0) * LP_ETS + ((int)px - ex0);
            pending = px < P.W && py < P.H;  // binning keeps only such points
            idx = (size_t)py * P.W + px;
        }
        ordered_rounds(pending, lp, k, claim, [&]() {
            const uint32_t sv = P.stencil[idx];
            if (!stencil_test(P.s_test, sv, P.s_ref)) return;  // point.rs:293
            P.stencil[idx] = (uint8_t)stencil_op(P.s_op, P.s_wrap, sv, P.s_ref);
            const float z = P.verts.comp(2)[v];
            if (!(z < 0.0f && z >= P.depth[idx])) return;        // point.rs:300-306
            float kv[KV > 0 ? KV : 1];
#pragma unroll
            for (int c = 0; c < KV; c++) kv[c] = P.verts.var(v)[c];
            const uint32_t c = fragment<FP>(P, kv, lut);
            P.color[idx] = blend<BLEND>(c, P.color[idx], lut);
            P.depth[idx] = z;
        });
    }
}

#define SR_LP_DISPATCH(KERNEL)                                                                   \
    template <int FP>                                                                            \
    static void KERNEL##_blend(const RasterParams& P, uint32_t n, hipStream_t s) {              \
        if (P.blend == SR_BLEND_OVER)                                                            \
            hipLaunchKernelGGL((KERNEL<FP, SR_BLEND_OVER>), dim3(n), dim3(256), 0, s, P);        \
        else                                                                                     \
            hipLaunchKernelGGL((KERNEL<FP, SR_BLEND_REPLACE>), dim3(n), dim3(256), 0, s, P);     \
    }                                                                                            \
    static void KERNEL##_fp(const RasterParams& P, uint32_t n, hipStream_t s) {                 \
        switch (P.fp) {                                                                          \
            case SR_FP_BLINN_PHONG: KERNEL##_blend<SR_FP_BLINN_PHONG>(P, n, s); break;           \
            case SR_FP_TEX_LAMBERT: KERNEL##_blend<SR_FP_TEX_LAMBERT>(P, n, s); break;           \
            case SR_FP_UV_COLOR: KERNEL##_blend<SR_FP_UV_COLOR>(P, n, s); break;                 \
            default: KERNEL##_blend<SR_FP_FLAT>(P, n, s); break;                                 \
        }                                                                                        \
    }

SR_LP_DISPATCH(k_raster_lines)
SR_LP_DISPATCH(k_raster_points)

void launch_raster_lines(const RasterParams& P, uint32_t ntiles, hipStream_t s) {
    if (ntiles) k_raster_lines_fp(P, ntiles, s);
}
void launch_raster_points(const RasterParams& P, uint32_t ntiles, hipStream_t s) {
    if (ntiles) k_raster_points_fp(P, ntiles, s);
}

}  // namespace sr
