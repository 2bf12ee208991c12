// sr_raster.hip — the raster work-groups.  Compiled once per execution
// tile size (-DSR_ETS=8|16|32) so the template instances build in
// parallel; each object exports launch_raster_*_<ETS>.
#include <type_traits>

#include "sr_kernels.hpp"
#include "sr_raster_common.hpp"

#ifndef SR_ETS
#define SR_ETS 16
#endif

// Waves per SIMD the raster kernels are compiled for (the register budget:
// 512 / waves VGPRs).  0 leaves the choice to the compiler.
#ifndef SR_RASTER_WAVES
#define SR_RASTER_WAVES 0
#endif
#if SR_RASTER_WAVES > 0
#define SR_RASTER_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SR_RASTER_WAVES, 8)))
#else
#define SR_RASTER_OCCUPANCY
#endif

namespace sr {

// ===========================================================================
// Raster: triangles
// ===========================================================================
// Pixel ownership: an ETS x ETS execution tile is split among NWV waves
// (1, 2 or 4, chosen independently of the tile size): with 4 each wave owns
// an ETS/2 square, with 2 a half of ETS/2 rows, with 1 the whole tile; a
// lane owns PPT = pixels per wave / 64 pixels, rows 64/RW apart.  Every
// wave walks the tile's triangle list for its own pixels.
template <int ETS, int NWV>
struct WaveRegion {
    static_assert(NWV == 1 || NWV == 2 || NWV == 4, "waves per tile");
    static_assert(ETS >= 16 || NWV == 1, "an 8x8 tile is one wave");
    static constexpr int RW = NWV == 4 ? ETS / 2 : ETS;  // region width
    static constexpr int RH = NWV == 1 ? ETS : ETS / 2;  // region height
    static constexpr int PPT = RW * RH / 64;              // pixels per lane
    static constexpr int RPP = 64 / RW;                   // region rows one pass of 64 lanes covers
};

// Fragments that pass stencil + depth wait in a per-wave LDS queue, in
// order, until some pixel's queue is full; then the wave shades the whole
// queue densely (64 fragments per pass, all lanes busy) and blends each
// pixel's fragments in order.  This reorders nothing the reference can
// observe: the depth test only reads depths written by earlier fragments,
// and every fragment program here returns a colour (never Discard), so the
// depth write (triangle.rs:142) does not depend on shading; blending, the
// only colour-order dependence, runs in queue (= reference) order.
// Queue slots per pixel: LDS holds 64 * PPT * Q queued fragments per wave
// (12 bytes each).  With one pixel per lane the LDS of a wave's queue
// bounds the waves a CU holds (160 KB).
#ifndef SR_QUEUE_DEPTH1
#define SR_QUEUE_DEPTH1 4
#endif
template <int PPT>
constexpr int queue_depth() { return PPT == 1 ? SR_QUEUE_DEPTH1 : (PPT == 2 ? 4 : (PPT == 4 ? 2 : 1)); }

// Triangles between refreshes of the hierarchical depth key.
#ifndef SR_HIZ_PERIOD
#define SR_HIZ_PERIOD 8u
#endif

// The least depth_key of the depths of the wave's pixels in `in` (DPP
// min: row_shr 1, 2, 4, 8 within rows of 16 lanes, shifted-in lanes
// keeping the identity, then row_bcast 15 and 31 as in wave_excl_scan;
// lane 63 ends with the minimum).
template <int PPT>
__device__ __forceinline__ uint32_t wave_min_key(const float (&dep)[PPT], const bool (&in)[PPT]) {
    uint32_t m = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < PPT; j++) m = min(m, in[j] ? depth_key(__float_as_uint(dep[j])) : 0xFFFFFFFFu);
    m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m, 0x111, 0xf, 0xf, false));  // row_shr:1
    m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m, 0x112, 0xf, 0xf, false));  // row_shr:2
    m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m, 0x114, 0xf, 0xf, false));  // row_shr:4
    m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m, 0x118, 0xf, 0xf, false));  // row_shr:8
    m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m, 0x142, 0xa, 0xf, false));  // row_bcast:15
    m = min(m, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)m, 63);
}

struct QRec {
    uint32_t tri;  // triangle id; after shading, the fragment colour
    float u, v;    // barycentrics (w = (1 - u) - v is recomputed exactly)
};

__device__ __forceinline__ void wave_sync_lds() {
    // LDS written by one lane and read by another within the same wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix sum over the 64 lanes, and the total, with DPP moves:
// row_shr 1, 2, 4, 8 scan each row of 16 lanes (lanes shifted in from
// outside the row read 0), then row_bcast 15 adds the end of row 0 to row 1
// and of row 2 to row 3 (row mask 0xa), and row_bcast 31 adds the end of
// row 1 to rows 2 and 3 (row mask 0xc).  No LDS, no lane-address registers.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t& total) {
    uint32_t x = v;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); // row_bcast:31
    total = __builtin_amdgcn_readlane(x, 63);
    return x - v;
}

// Shade one queued fragment: interpolate the varyings (interpolate.rs:43-64
// order) and run the fragment program.
template <int FP>
__device__ __forceinline__ uint32_t shade_rec(const RasterParams& P, const QRec& q, Lut lut) {
    constexpr int KV = fp_varyings<FP>();
    float k[KV > 0 ? KV : 1];
    if constexpr (KV > 0) {
        // a triangle list without indices needs no vertex-id load: its
        // vertices are 3t, 3t+1, 3t+2
        uint32_t i0 = 3u * q.tri, i1 = i0 + 1u, i2 = i0 + 2u;
        if (P.vids) {
            const TriVid& r = P.vids[q.tri];
            i0 = r.i0; i1 = r.i1; i2 = r.i2;
        }
        const float u = q.u, v = q.v, w = (1.0f - u) - v;
        const float *r0 = P.verts.var(i0), *r1 = P.verts.var(i1), *r2 = P.verts.var(i2);
#pragma unroll
        for (int c = 0; c < KV; c++) k[c] = ((r0[c] * u) + (r1[c] * v)) + (r2[c] * w);
    }
    return fragment<FP, false>(P, k, lut);  // RGBA8 textures (RGBAF32 ones go to the per-pixel raster)
}

template <int ETS, int NWV, int FP, int BLEND>
struct RasterShared {
    using WR = WaveRegion<ETS, NWV>;
    static constexpr int NT = 64 * NWV;
    static constexpr int NW = NWV;
    static constexpr int PPT = WR::PPT;               // pixels per lane
    static constexpr int WPIX = 64 * PPT;             // pixels per wave
    static constexpr int Q = BLEND == SR_BLEND_OVER ? queue_depth<PPT>() : 1;
    QRec qbuf[NW][WPIX * Q];  // first: queue stores take their offset in the instruction
    uint4 stage[NW][3 * 64];  // per wave: the block's records, bytes 0-47 (chunk c of item k at c*64 + k)
    float lut[256];
    uint16_t flat[NW][WPIX * Q];
};

// One execution tile: `tile` is the scene-local tile index, `gtile` the
// global one that indexes the bins.
// Scene parameters through the scalar cache (a uniform index).
__device__ __forceinline__ RasterParams load_params(const RasterParams* base, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr int N = sizeof(RasterParams) / 4;
    const __attribute__((address_space(4))) uint32_t* src =
        (const __attribute__((address_space(4))) uint32_t*)(base + s);
    uint32_t w[N];
#pragma unroll
    for (int i = 0; i < N; i++) w[i] = src[i];
    RasterParams r;
    __builtin_memcpy(&r, w, sizeof(r));
    return r;
#else
    return base[s];
#endif
}

// The draws a tile renders, in order: the flat kernel has one; a batch
// group (several draws into one framebuffer) has its draws' parameters in
// the scene table and each draw's items are the tile-list entries whose
// triangle ids lie below the draw's end.
struct OneDraw {
    __device__ uint32_t count() const { return 1u; }
};

struct GroupDraws {
    const RasterParams* table;  // scene (draw) table
    const uint32_t* tri_end;    // exclusive end of each draw's triangle ids
    uint32_t d0, n;             // the group's first draw and draw count
    __device__ uint32_t count() const { return n; }
};

// First position in items[lo, hi) (ascending) whose id is >= key.
__device__ __forceinline__ uint32_t lower_bound_items(const uint32_t* items, uint32_t lo, uint32_t hi,
                                                      uint32_t key) {
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (load_item(items, m) < key) lo = m + 1;
        else hi = m;
    }
    return lo;
}


// CLAMP: the draw runs in the exhaustive mode (every triangle in every
// tile), so a triangle's box is clamped into each screen tile exactly as
// clamp_as_int! does (triangle.rs:66-78), strips of pixels outside the box
// included.  Binned draws test the box itself (stencil counting) or nothing
// (no stencil effects): a triangle reaches only the execution tiles its box
// meets, and on every pixel that can pass, the clamped box, the box and no
// box at all agree -- except for a triangle whose record carries the strip
// flag (a pixel outside its box may pass the coverage test): setup bins it
// to every execution tile such a pixel may lie in, and the raster tests it
// in its clamped boxes.
template <int ETS, int NWV, int FP, int BLEND, int SMODE, bool CLAMP, class Draws>
__device__ __forceinline__ void raster_tile(const RasterParams& P, uint32_t tile, uint32_t gtile,
                                            RasterShared<ETS, NWV, FP, BLEND>& sh, const Draws& draws) {
    using SH = RasterShared<ETS, NWV, FP, BLEND>;
    using WR = typename SH::WR;
    constexpr int RW = WR::RW, RH = WR::RH, PPT = SH::PPT, Q = SH::Q;
    constexpr int RPP = WR::RPP;
    // the record's box is read for every triangle when the mode tests it
    // (stencil effects, the exhaustive mode); else only for strip triangles
    constexpr bool BOX = SMODE != 0 || CLAMP;
    // this wave's record stage: 3 chunks of 64 x 16 bytes
    uint4* const stage = sh.stage[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];

    // The empty asm makes the thread index opaque per tile: in the
    // persistent batch kernel the compiler would otherwise hoist every
    // lane-derived value out of the tile loop and keep it live across the
    // whole loop (133 instead of 115 VGPRs: one wave less per SIMD).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t etx = tile % P.ntx, ety = tile / P.ntx;
    // constant-space loads: a flat load counts as divergent, which would turn
    // the whole triangle loop into per-lane control flow
    const uint32_t beg = load_item(P.tile_start, gtile);
    const uint32_t end = load_item(P.tile_end, gtile);
    if (beg >= end && SMODE != 1 && !P.clear) return;
    const Lut lut{sh.lut};

    // tile and wave pixel spans (clipped to the framebuffer)
    const uint32_t tx0 = etx * ETS, ty0 = ety * ETS;
    const int tx1 = (int)min(tx0 + ETS, P.W) - 1, ty1 = (int)min(ty0 + ETS, P.H) - 1;
    const int qx = NWV == 4 ? (wave & 1) * RW : 0;
    const int qy = NWV == 4 ? (wave >> 1) * RH : (NWV == 2 ? wave * RH : 0);
    const int wx0 = (int)tx0 + qx, wy0 = (int)ty0 + qy;
    const int wx1 = min(wx0 + RW - 1, tx1), wy1 = min(wy0 + RH - 1, ty1);

    int px[PPT], py[PPT];
    bool live[PPT];
    uint32_t col[PPT], cnt[PPT], st[PPT], qn[PPT];
    float dep[PPT], fx[PPT], fy[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        px[j] = wx0 + lane % RW;
        py[j] = wy0 + lane / RW + j * RPP;
        live[j] = (uint32_t)px[j] < P.W && (uint32_t)py[j] < P.H;
        const size_t idx = (size_t)py[j] * P.W + px[j];
        // a pixel outside the framebuffer holds a positive NaN depth: it
        // passes no depth test and lies above every key of wave_min_key
        if (P.clear) {  // Framebuffer::clear: colour, depth far, stencil 0
            col[j] = P.clear_rgba;
            dep[j] = live[j] ? -3.40282347e+38f : __uint_as_float(0x7FFFFFFFu);
            st[j] = 0u;
        } else {
            col[j] = live[j] ? P.color[idx] : 0u;
            dep[j] = live[j] ? P.depth[idx] : __uint_as_float(0x7FFFFFFFu);
            st[j] = (SMODE == 2 && live[j]) ? P.stencil[idx] : 0u;
        }
        cnt[j] = 0;
        qn[j] = 0;
        fx[j] = (float)px[j] + 0.5f;  // pixel centre, triangle.rs:104-105
        fy[j] = (float)py[j] + 0.5f;
    }
    QRec* wq = sh.qbuf[wave];
    // the queue's LDS byte address: slot addresses in the pixel loop are
    // formed with 32-bit arithmetic (a 64-bit multiply-add otherwise)
    typedef __attribute__((address_space(3))) QRec LQRec;
    const uint32_t wq_lds = (uint32_t)(uintptr_t)(LQRec*)wq;
    uint32_t qlane[PPT];  // each pixel's first queue slot
#pragma unroll
    for (int j = 0; j < PPT; j++) qlane[j] = wq_lds + (uint32_t)((lane * PPT + j) * Q) * (uint32_t)sizeof(QRec);
    auto qput = [&](int j, uint32_t k, uint32_t tri, float u, float v) {  // k < Q: a 24-bit multiply-add
        LQRec* const q = (LQRec*)(uintptr_t)(qlane[j] + __umul24(k, (uint32_t)sizeof(QRec)));
        q->tri = tri;
        q->u = u;
        q->v = v;
    };
    uint16_t* wflat = sh.flat[wave];

    // Shade every queued fragment of this wave and blend them in order.
    auto flush = [&](const RasterParams& PD) {
        wave_sync_lds();
        uint32_t mine = 0;
#pragma unroll
        for (int j = 0; j < PPT; j++) mine += qn[j];
        uint32_t total;
        uint32_t o = wave_excl_scan(mine, total);
#pragma unroll
        for (int j = 0; j < PPT; j++)
            for (uint32_t k = 0; k < qn[j]; k++) wflat[o++] = (uint16_t)((lane * PPT + j) * Q + k);
        wave_sync_lds();
        for (uint32_t b = 0; b < total; b += 64) {
            const uint32_t r = b + lane;
            if (r < total) {
                const uint32_t s = wflat[r];
                const QRec q = wq[s];
                wq[s].tri = shade_rec<FP>(PD, q, lut);
            }
        }
        wave_sync_lds();
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            for (uint32_t k = 0; k < qn[j]; k++)
                col[j] = blend<BLEND>(wq[(lane * PPT + j) * Q + k].tri, col[j], lut);
            qn[j] = 0;
        }
        wave_sync_lds();
    };


    // Hierarchical depth (draws without stencil effects): hz_key is a lower
    // bound of the depth keys of this wave's pixels in the screen tile being
    // walked, refreshed every SR_HIZ_PERIOD triangles (depths only grow: a
    // fragment passes with z >= its pixel's depth).  A triangle whose depth
    // bound key lies below it cannot pass the depth test at any of those
    // pixels (the only ones it may touch in this walk), and then changes
    // nothing: it is skipped.
    uint32_t hz_key = 0u, hz_n = 1u;

    // One draw's items [sb, se): the screen tiles meeting this wave's pixels
    // in the reference's row-major order, each walking the items in order.
    auto walk = [&](uint32_t sb, uint32_t se, const RasterParams& PD) {
    if (sb < se && wx0 <= wx1 && wy0 <= wy1) {
        uint32_t jlo, jhi, ilo, ihi;
        sem_range((uint32_t)wx0, (uint32_t)wx1, P.tw, P.scols, jlo, jhi);
        sem_range((uint32_t)wy0, (uint32_t)wy1, P.th, P.srows, ilo, ihi);
        const uint32_t xmax = P.W - 1, ymax = P.H - 1;
        for (uint32_t si = ilo; si <= ihi; si++) {
            const int Ty0 = (int)(si * P.th), Ty1 = (int)min(si * P.th + P.th, ymax);
            for (uint32_t sj = jlo; sj <= jhi; sj++) {
                const int Tx0 = (int)(sj * P.tw), Tx1 = (int)min(sj * P.tw + P.tw, xmax);
                const int rx0 = max(wx0, Tx0), rx1 = min(wx1, Tx1);
                const int ry0 = max(wy0, Ty0), ry1 = min(wy1, Ty1);
                if (rx0 > rx1 || ry0 > ry1) continue;
                bool inT[PPT];
#pragma unroll
                for (int j = 0; j < PPT; j++)
                    inT[j] = live[j] && px[j] >= Tx0 && px[j] <= Tx1 && py[j] >= Ty0 && py[j] <= Ty1;
                // hz_key bounds the depths of this screen tile's pixels only
                // (a pixel on a screen tile seam is walked once per tile it
                // lies in; the others wait): start again per screen tile
                hz_key = 0u;
                hz_n = 1u;

                // One triangle against this wave's pixels (rc: its record,
                // read from the stage; cur: its id).
                auto tri = [&](const RecLane& rc, uint32_t cur) {
                    int X0 = rc.X0, X1 = rc.X1, Y0 = rc.Y0, Y1 = rc.Y1;  // (BOX: else read below)
                    if constexpr (CLAMP) {
                        if (X0 > X1) return;  // culled or invalid (binned draws never see these)
                    }
                    if constexpr (SMODE == 0) {
                        if ((rc.zs | 1u) < hz_key) return;  // behind every pixel here
                    }
                    // The pixels tested: the box clamped into this screen
                    // tile (clamp_as_int!, triangle.rs:66-78) in the
                    // exhaustive mode and for a triangle whose record says a
                    // pixel outside its box may pass (binned to every tile
                    // such a pixel may lie in; rc.zs is wave-uniform, so
                    // this is a scalar branch); the box for stencil
                    // effects; else every pixel: in a binned draw no pixel
                    // centre outside the box of any other triangle can pass
                    // the coverage test, and inside the box the test is the
                    // reference's.  Stencil counting counts the box itself
                    // (the strips outside it are counted apart).
                    const bool clampit = CLAMP || (rc.zs & 1u) != 0u;
                    if constexpr (!BOX) {
                        if (clampit) {  // a strip triangle: its box now
                            RecLane b;
                            rec_box(P.tris, cur, b);
                            X0 = b.X0; X1 = b.X1; Y0 = b.Y0; Y1 = b.Y1;
                        }
                    }
                    int cx0 = X0, cx1 = X1, cy0 = Y0, cy1 = Y1;
                    if (clampit) {
                        cx0 = min(max(X0, Tx0), Tx1); cx1 = min(max(X1, Tx0), Tx1);
                        cy0 = min(max(Y0, Ty0), Ty1); cy1 = min(max(Y1, Ty0), Ty1);
                    }
                    // One pixel: stencil, coverage, depth, queue (ok: the
                    // pixel is in the tested region).
                    auto pixel = [&](int j, bool ok) {
                        if constexpr (SMODE == 2) {  // stencil before coverage, triangle.rs:90-99
                            const bool sp = ok && stencil_test(P.s_test, st[j], P.s_ref);
                            st[j] = sp ? stencil_op(P.s_op, P.s_wrap, st[j], P.s_ref) : st[j];
                            ok = sp;
                        } else if constexpr (SMODE == 1) {
                            const bool inr = !clampit || ((px[j] >= X0) & (px[j] <= X1) & (py[j] >= Y0) & (py[j] <= Y1));
                            cnt[j] += (ok & inr) ? 1u : 0u;
                        }
                        const float dx = fx[j] - rc.x3, dy = fy[j] - rc.y3;
                        const float un = (rc.e23y * dx) + (rc.e23x * dy);  // triangle.rs:108
                        const float vn = (rc.e31y * dx) + (rc.e31x * dy);  // triangle.rs:109
                        const float u = quot(un, rc.R), v = quot(vn, rc.R);
                        const float w = (1.0f - u) - v;                    // triangle.rs:110
                        const float z = ((rc.z1 * u) + (rc.z2 * v)) + (rc.z3 * w);  // :115-117
                        // :113 inside, :120-126 depth
                        const bool pass = ok & !(u < 0.0f) & !(v < 0.0f) & !(w < 0.0f) & (z < 0.0f) & (z >= dep[j]);
                        dep[j] = pass ? z : dep[j];
                        if (pass) {
                            if constexpr (BLEND == SR_BLEND_OVER) {
                                qput(j, qn[j], cur, u, v);
                                qn[j]++;
                            } else {
                                qput(j, 0, cur, u, v);  // `()` blend: the last fragment wins
                                qn[j] = 1;
                            }
                        }
                    };
                    if constexpr (SMODE == 0 && !CLAMP) {
                        // no stencil effects: no test at all for a triangle
                        // that cannot pass outside its box; for one that can,
                        // the lanes outside its clamped box sit out
                        if (clampit) {
#pragma unroll
                            for (int j = 0; j < PPT; j++)
                                if ((px[j] >= cx0) & (px[j] <= cx1) & (py[j] >= cy0) & (py[j] <= cy1)) pixel(j, inT[j]);
                        } else {
#pragma unroll
                            for (int j = 0; j < PPT; j++) pixel(j, inT[j]);
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < PPT; j++)
                            pixel(j, inT[j] & (px[j] >= cx0) & (px[j] <= cx1) & (py[j] >= cy0) & (py[j] <= cy1));
                    }
                    if constexpr (BLEND == SR_BLEND_OVER) {
                        // a full queue on any lane: shade and blend the queues
                        bool full = false;
#pragma unroll
                        for (int j = 0; j < PPT; j++) full |= qn[j] == (uint32_t)Q;
                        if (__ballot(full)) flush(PD);
                    }
                };

                // Walk the tile's triangles in order, 64 items at a time.
                // The item ids come a block ahead.  At a block's start lane
                // k copies the first 48 bytes of item c0 + k's record into
                // this wave's LDS stage (three 16-byte LDS DMA loads, waited
                // for once per block, before the stage's first read); the
                // block's survivors are its items in [sb, se) whose depth
                // bound is not below hz_key, and each survivor's record is
                // read from the stage by three broadcast LDS reads.  Every
                // SR_HIZ_PERIOD triangles tested hz_key is refreshed and the
                // block's remaining survivors filtered by it.  Positions are
                // clamped into [sb, se), so every id loaded is a valid
                // triangle.
                uint32_t id = load_ids(P.items, sb & ~63u, sb, se, lane);
                for (uint32_t c0 = sb & ~63u; c0 < se; c0 += 64) {
                    stage_records(P.tris, id, stage);
                    wait_vm_all();  // the stage has landed (the next ids load after this)
                    const uint32_t idn = load_ids(P.items, c0 + 64, sb, se, lane);
                    const uint32_t zsk = stage[lane].z;  // item c0 + k's depth key word
                    const uint32_t pos = c0 + (uint32_t)lane;
                    bool keep = pos >= sb && pos < se;
                    if constexpr (SMODE == 0) keep = keep && (zsk | 1u) >= hz_key;
                    uint64_t m = __ballot(keep);
                    while (m != 0) {
                        const int k = __builtin_amdgcn_readfirstlane(__builtin_ctzll(m));
                        m &= m - 1;
                        if constexpr (SMODE == 0) {
                            if (--hz_n == 0u) {
                                hz_n = SR_HIZ_PERIOD;
                                hz_key = wave_min_key<PPT>(dep, inT);
                                m &= __ballot((zsk | 1u) >= hz_key);
                            }
                        }
                        const uint32_t cur = lane_u32(id, k);
                        RecLane rc = rec_stage(stage, k);
                        if constexpr (BOX) rec_box(P.tris, cur, rc);
                        tri(rc, cur);
                    }
                    id = idn;
                }
            }
        }
    }
    };

    if constexpr (std::is_same<Draws, OneDraw>::value) {
        walk(beg, end, P);
        flush(P);
    } else {
        // each draw of the group in order: its own pass over the screen
        // tiles, its fragments shaded with its own parameters
        uint32_t sb = beg;
        for (uint32_t d = 0; d < draws.count(); d++) {
            const uint32_t se = d + 1 < draws.count()
                                    ? lower_bound_items(P.items, sb, end, load_item(draws.tri_end, draws.d0 + d))
                                    : end;
            const RasterParams PD = load_params(draws.table, draws.d0 + d);
            walk(sb, se, PD);
            flush(PD);
            sb = se;
        }
    }
    // resolve: the tile goes back to HBM once
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        if (!live[j]) continue;
        const size_t idx = (size_t)py[j] * P.W + px[j];
        if (beg < end || P.clear) {
            P.color[idx] = col[j];
            P.depth[idx] = dep[j];
        }
        if constexpr (SMODE == 2) {
            P.stencil[idx] = (uint8_t)st[j];
        } else if constexpr (SMODE == 1) {
            const uint32_t n = cnt[j] + (P.strip_counts ? P.strip_counts[idx] : 0u);
            const uint32_t s0 = P.clear ? 0u : P.stencil[idx];
            if (n || P.clear) P.stencil[idx] = (uint8_t)stencil_op_n(P.s_op, P.s_wrap, s0, P.s_ref, n);
        } else {
            if (P.clear) P.stencil[idx] = 0;
        }
    }
}


// One scene: one work-group per execution tile.
template <int ETS, int NWV, int FP, int BLEND, int SMODE, bool CLAMP>
__global__ __launch_bounds__(256) SR_RASTER_OCCUPANCY void k_raster_tris(RasterParams P) {
    using SH = RasterShared<ETS, NWV, FP, BLEND>;
    __shared__ SH sh;
    if (P.skip_if && ((load_item(P.skip_if, 0) & 1u) | load_item(P.skip_if, 4))) return;  // NaN, overflow
    for (int i = threadIdx.x; i < 256; i += SH::NT) sh.lut[i] = (float)i / 255.0f;
    __syncthreads();
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    raster_tile<ETS, NWV, FP, BLEND, SMODE, CLAMP>(P, tile, tile, sh, OneDraw{});
}

// A batch: a persistent grid over the scene table (no stencil effects in
// batches, so SMODE is 0).
template <int ETS, int NWV, int FP, int BLEND>
__global__ __launch_bounds__(256) SR_RASTER_OCCUPANCY void k_raster_tris_batch(RasterBatch B) {
    using SH = RasterShared<ETS, NWV, FP, BLEND>;
    __shared__ SH sh;
    if (B.overflow && (B.overflow[0] | (B.overflow[2] & 1u))) return;  // overflow, NaN coordinate
    for (int i = threadIdx.x; i < 256; i += SH::NT) sh.lut[i] = (float)i / 255.0f;
    __syncthreads();
    // Work-group b takes run p of near-equal estimated cost (a contiguous
    // run of tiles; many more runs than resident work-groups, so the
    // hardware's hand-out balances what the estimate misses); runs are
    // numbered so that each XCD holds a contiguous stretch of tiles
    // (neighbouring tiles share records in its L2).
    const uint32_t p = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t t0 = load_item(B.wg_start, p), t1 = load_item(B.wg_start, p + 1);
    uint32_t group = 0, group_end = 0;
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t g = B.tile0 + t;
        if (g >= group_end) {
            // framebuffer group owning global tile g: first s with g < group_tile_end[s]
            uint32_t a = 0, b = B.ngroups;
            while (a < b) {
                const uint32_t m = (a + b) >> 1;
                if (load_item(B.group_tile_end, m) <= g) a = m + 1;
                else b = m;
            }
            group = a;
            group_end = load_item(B.group_tile_end, a);
        }
        // the group's draws, in item order; parameters are reloaded per tile
        // (scalar-cache hits) rather than held in registers
        const uint32_t d0 = load_item(B.group_first, group), d1 = load_item(B.group_first, group + 1);
        const RasterParams P = load_params(B.scenes, d0);
        raster_tile<ETS, NWV, FP, BLEND, 0, false>(P, g - P.tile_base, g, sh,
                                                   GroupDraws{B.scenes, B.tri_end, d0, d1 - d0});
    }
}

template <int ETS, int NWV, int FP, int BLEND>
static void raster_tris_smode(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    const dim3 g(ntiles), b(64 * NWV);
    const bool ex = P.one.exhaustive != 0;
    if (P.scenes) hipLaunchKernelGGL((k_raster_tris_batch<ETS, NWV, FP, BLEND>), g, b, 0, s, P);
    else if (P.one.smode == 2) hipLaunchKernelGGL((k_raster_tris<ETS, NWV, FP, BLEND, 2, true>), g, b, 0, s, P.one);
    else if (P.one.smode == 1 && ex) hipLaunchKernelGGL((k_raster_tris<ETS, NWV, FP, BLEND, 1, true>), g, b, 0, s, P.one);
    else if (P.one.smode == 1) hipLaunchKernelGGL((k_raster_tris<ETS, NWV, FP, BLEND, 1, false>), g, b, 0, s, P.one);
    else if (ex) hipLaunchKernelGGL((k_raster_tris<ETS, NWV, FP, BLEND, 0, true>), g, b, 0, s, P.one);
    else hipLaunchKernelGGL((k_raster_tris<ETS, NWV, FP, BLEND, 0, false>), g, b, 0, s, P.one);
}

template <int ETS, int NWV, int FP>
static void raster_tris_blend(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    if (P.one.blend == SR_BLEND_OVER) raster_tris_smode<ETS, NWV, FP, SR_BLEND_OVER>(P, ntiles, s);
    else raster_tris_smode<ETS, NWV, FP, SR_BLEND_REPLACE>(P, ntiles, s);
}

template <int ETS, int NWV>
static void raster_tris_fp(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    switch (P.one.fp) {
        case SR_FP_BLINN_PHONG: raster_tris_blend<ETS, NWV, SR_FP_BLINN_PHONG>(P, ntiles, s); break;
        case SR_FP_TEX_LAMBERT: raster_tris_blend<ETS, NWV, SR_FP_TEX_LAMBERT>(P, ntiles, s); break;
        case SR_FP_UV_COLOR: raster_tris_blend<ETS, NWV, SR_FP_UV_COLOR>(P, ntiles, s); break;
        default: raster_tris_blend<ETS, NWV, SR_FP_FLAT>(P, ntiles, s); break;
    }
}

// Waves per tile this object instantiates for its tile size: 8x8 one wave;
// 16x16 one, two or four; 32x32 two or four.  Any other count runs the tile
// size's default (the last listed).
#define SR_CAT2(a, b) a##b
#define SR_CAT(a, b) SR_CAT2(a, b)
void SR_CAT(launch_raster_tris_, SR_ETS)(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    if (!ntiles) return;
    const int w = P.one.nwv;
#if SR_ETS == 8
    (void)w;
    raster_tris_fp<8, 1>(P, ntiles, s);
#elif SR_ETS == 16
    if (w == 1) raster_tris_fp<16, 1>(P, ntiles, s);
    else if (w == 2) raster_tris_fp<16, 2>(P, ntiles, s);
    else raster_tris_fp<16, 4>(P, ntiles, s);
#elif SR_ETS == 32
    if (w == 2) raster_tris_fp<32, 2>(P, ntiles, s);
    else raster_tris_fp<32, 4>(P, ntiles, s);
#else
#error "SR_ETS must be 8, 16 or 32"
#endif
}

}  // namespace sr
