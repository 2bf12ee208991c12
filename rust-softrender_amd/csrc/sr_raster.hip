// sr_raster.hip — the raster work-groups.  Compiled once per execution
// tile size (-DSR_ETS=8|16|32|64) so the template instances build in
// parallel; each object exports launch_raster_*_<ETS>.
#include <type_traits>

#include "sr_device.hpp"
#include "sr_kernels.hpp"

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

// Loads through the constant address space: read-only for the kernel's
// lifetime, so a wave-uniform address becomes an s_load into SGPRs.
__device__ __forceinline__ uint32_t load_item(const uint32_t* base, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) uint32_t*)base)[i];
#else
    return base[i];
#endif
}

// 64 consecutive item ids, lane k holding the id at position c0 + k, the
// position clamped into [beg, end) (beg < end) so every lane loads.
__device__ __forceinline__ uint32_t load_ids(const uint32_t* items, uint32_t c0, uint32_t beg,
                                             uint32_t end, int lane) {
    const uint32_t c = min(max(c0 + (uint32_t)lane, beg), end - 1);
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(1))) uint32_t*)items)[c];
#else
    return items[c];
#endif
}

// A block of 64 triangle records, lane k holding the first 60 bytes of
// record k (R, x3, y3, edges, z, bbox).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
struct RecBlock {
    u32x4 a, b, c;
    u32x3 d;
};

__device__ __forceinline__ RecBlock load_block(const TriRec* tris, uint32_t id) {
    RecBlock r;
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(1))) u32x4 G4;
    typedef const __attribute__((address_space(1))) u32x3 G3;
#else
    typedef const u32x4 G4;
    typedef const u32x3 G3;
#endif
    const char* base = (const char*)(tris + id);
    r.a = *(G4*)(base);
    r.b = *(G4*)(base + 16);
    r.c = *(G4*)(base + 32);
    r.d = *(G3*)(base + 48);
    return r;
}

// The fields of record k of a block, in SGPRs (k is wave-uniform).
struct RecLane {
    double R;
    float x3, y3, e23y, e23x, e31y, e31x, z1, z2, z3;
    int X0, X1, Y0, Y1;
};

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}
__device__ __forceinline__ float lane_f32(uint32_t v, int k) {
    return __builtin_bit_cast(float, lane_u32(v, k));
}

__device__ __forceinline__ RecLane rec_at(const RecBlock& b, int k) {
    RecLane r;
    const uint64_t lo = lane_u32(b.a.x, k), hi = lane_u32(b.a.y, k);
    r.R = __builtin_bit_cast(double, lo | (hi << 32));
    r.x3 = lane_f32(b.a.z, k);
    r.y3 = lane_f32(b.a.w, k);
    r.e23y = lane_f32(b.b.x, k);
    r.e23x = lane_f32(b.b.y, k);
    r.e31y = lane_f32(b.b.z, k);
    r.e31x = lane_f32(b.b.w, k);
    r.z1 = lane_f32(b.c.x, k);
    r.z2 = lane_f32(b.c.y, k);
    r.z3 = lane_f32(b.c.z, k);
    r.X0 = (int)lane_u32(b.c.w, k);
    r.X1 = (int)lane_u32(b.d.x, k);
    r.Y0 = (int)lane_u32(b.d.y, k);
    r.Y1 = (int)lane_u32(b.d.z, k);
    return r;
}

__device__ __forceinline__ TriRec load_rec(const TriRec* base, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    const __attribute__((address_space(4))) TriRec* p =
        (const __attribute__((address_space(4))) TriRec*)base + i;
    TriRec r;
    r.R = p->R;
    r.x3 = p->x3; r.y3 = p->y3;
    r.e23y = p->e23y; r.e23x = p->e23x; r.e31y = p->e31y; r.e31x = p->e31x;
    r.z1 = p->z1; r.z2 = p->z2; r.z3 = p->z3;
    r.X0 = p->X0; r.X1 = p->X1; r.Y0 = p->Y0; r.Y1 = p->Y1;
    r.i0 = p->i0; r.i1 = p->i1; r.i2 = p->i2;
    return r;
#else
    return base[i];
#endif
}

// Fragments that pass stencil + depth wait in a per-wave LDS queue, in
// order, until some pixel's queue is full; then the wave shades the whole
// queue densely (64 fragments per pass, all lanes busy) and blends each
// pixel's fragments in order.  This reorders nothing the reference can
// observe: the depth test only reads depths written by earlier fragments,
// and every fragment program here returns a colour (never Discard), so the
// depth write (triangle.rs:142) does not depend on shading; blending, the
// only colour-order dependence, runs in queue (= reference) order.
template <int ETS>
constexpr int queue_depth() { return ETS <= 16 ? 8 : (ETS == 32 ? 2 : 1); }

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

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t& total) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
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
        const TriRec& r = P.tris[q.tri];
        const uint32_t i0 = r.i0, i1 = r.i1, i2 = r.i2;
        const float u = q.u, v = q.v, w = (1.0f - u) - v;
#pragma unroll
        for (int c = 0; c < KV; c++) {
            const float* vc = P.verts.comp(4 + c);
            k[c] = ((vc[i0] * u) + (vc[i1] * v)) + (vc[i2] * w);
        }
    }
    return fragment<FP>(P, k, lut);
}

template <int ETS, int FP, int BLEND>
struct RasterShared {
    static constexpr int NT = raster_threads<ETS>();
    static constexpr int NW = NT / 64;
    static constexpr int S = ETS == 8 ? 8 : ETS / 2;  // each wave owns an S x S quadrant
    static constexpr int PPT = (S * S) / 64;          // pixels per lane
    static constexpr int WPIX = 64 * PPT;             // pixels per wave
    static constexpr int Q = BLEND == SR_BLEND_OVER ? queue_depth<ETS>() : 1;
    QRec qbuf[NW][WPIX * Q];  // first: queue stores take their offset in the instruction
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


template <int ETS, int FP, int BLEND, int SMODE, class Draws>
__device__ __forceinline__ void raster_tile(const RasterParams& P, uint32_t tile, uint32_t gtile,
                                            RasterShared<ETS, FP, BLEND>& sh, const Draws& draws) {
    using SH = RasterShared<ETS, FP, BLEND>;
    constexpr int S = SH::S, PPT = SH::PPT, Q = SH::Q;
    constexpr int RPP = 64 / S;  // quadrant rows covered by one pixel pass

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
    const int qx = ETS == 8 ? 0 : (wave & 1) * S, qy = ETS == 8 ? 0 : (wave >> 1) * S;
    const int wx0 = (int)tx0 + qx, wy0 = (int)ty0 + qy;
    const int wx1 = min(wx0 + S - 1, tx1), wy1 = min(wy0 + S - 1, ty1);

    int px[PPT], py[PPT];
    bool live[PPT];
    uint32_t col[PPT], cnt[PPT], st[PPT], qn[PPT];
    float dep[PPT], fx[PPT], fy[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        px[j] = wx0 + lane % S;
        py[j] = wy0 + lane / S + j * RPP;
        live[j] = (uint32_t)px[j] < P.W && (uint32_t)py[j] < P.H;
        const size_t idx = (size_t)py[j] * P.W + px[j];
        if (P.clear) {  // Framebuffer::clear: colour, depth far, stencil 0
            col[j] = P.clear_rgba;
            dep[j] = -3.40282347e+38f;
            st[j] = 0u;
        } else {
            col[j] = live[j] ? P.color[idx] : 0u;
            dep[j] = live[j] ? P.depth[idx] : 0.0f;
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

                // walk the tile's triangles in order, 64 at a time: lane k
                // gathers the record of item c0 + k one block ahead of its use
                // (item ids two blocks ahead) and the loop hands record fields
                // to the whole wave with v_readlane, so the random record
                // fetches overlap instead of stalling each triangle.  Loads
                // are unconditional, with positions clamped into [beg, end),
                // so no loaded register is copied (which would wait for it).
                const uint32_t b0 = sb & ~63u;
                uint32_t id = load_ids(P.items, b0, sb, se, lane);
                RecBlock blk = load_block(P.tris, id);
                uint32_t idn = load_ids(P.items, b0 + 64, sb, se, lane);
                for (uint32_t c0 = b0; c0 < se; c0 += 64) {
                    RecBlock nxt = load_block(P.tris, idn);
                    const uint32_t idc = idn;
                    idn = load_ids(P.items, c0 + 128, sb, se, lane);
                    const uint32_t lo = max(c0, sb), hi = min(c0 + 64, se);
                    for (uint32_t c = lo; c < hi; c++) {
                    const int k = (int)(c - c0);
                    const RecLane rc = rec_at(blk, k);
                    const uint32_t cur = lane_u32(id, k);
                    // tile-clamped bbox, clamp_as_int!, triangle.rs:66-78
                    const int X0 = rc.X0, X1 = rc.X1, Y0 = rc.Y0, Y1 = rc.Y1;
                    if (X0 > X1) continue;  // culled or invalid
                    const int cx0 = min(max(X0, Tx0), Tx1), cx1 = min(max(X1, Tx0), Tx1);
                    const int cy0 = min(max(Y0, Ty0), Ty1), cy1 = min(max(Y1, Ty0), Ty1);
                    // (no wave-level early-out on the clamped box: binned
                    // triangles nearly always reach the wave's pixels, and
                    // the scalar test cost more than the rare skip saved)
                    bool full = false;
#pragma unroll
                    for (int j = 0; j < PPT; j++) {
                        // straight-line per pixel: every test is a mask, so
                        // the wave runs one exec-mask region per pixel
                        // instead of a nested branch per test
                        bool ok = inT[j] & (px[j] >= cx0) & (px[j] <= cx1) & (py[j] >= cy0) & (py[j] <= cy1);
                        if constexpr (SMODE == 2) {  // stencil before coverage, triangle.rs:90-99
                            const bool sp = ok && stencil_test(P.s_test, st[j], P.s_ref);
                            st[j] = sp ? stencil_op(P.s_op, P.s_wrap, st[j], P.s_ref) : st[j];
                            ok = sp;
                        } else if constexpr (SMODE == 1) {
                            cnt[j] += (ok & (px[j] >= X0) & (px[j] <= X1) & (py[j] >= Y0) & (py[j] <= Y1)) ? 1u : 0u;
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
                        // tested on every lane, outside the branch: the
                        // ballot below then takes the compare's mask as is
                        if constexpr (BLEND == SR_BLEND_OVER) full |= qn[j] == (uint32_t)Q;
                    }
                    if constexpr (BLEND == SR_BLEND_OVER) {
                        if (__ballot(full)) {
                            // the record blocks are gathered again after the
                            // flush rather than kept live across the shading
                            // code: that keeps the kernel ~26 VGPRs smaller
                            // (mid-walk flushes are rare)
                            flush(PD);
                            blk = load_block(P.tris, id);
                            nxt = load_block(P.tris, idc);
                        }
                    }
                    }
                    blk = nxt;
                    id = idc;
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
template <int ETS, int FP, int BLEND, int SMODE>
__global__ __launch_bounds__(256) SR_RASTER_OCCUPANCY void k_raster_tris(RasterParams P) {
    using SH = RasterShared<ETS, FP, BLEND>;
    __shared__ SH sh;
    for (int i = threadIdx.x; i < 256; i += SH::NT) sh.lut[i] = (float)i / 255.0f;
    __syncthreads();
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    raster_tile<ETS, FP, BLEND, SMODE>(P, tile, tile, sh, OneDraw{});
}

// A batch: a persistent grid over the scene table (no stencil effects in
// batches, so SMODE is 0).
template <int ETS, int FP, int BLEND>
__global__ __launch_bounds__(256) SR_RASTER_OCCUPANCY void k_raster_tris_batch(RasterBatch B) {
    using SH = RasterShared<ETS, FP, BLEND>;
    __shared__ SH sh;
    if (B.overflow && *B.overflow) return;
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
        raster_tile<ETS, FP, BLEND, 0>(P, g - P.tile_base, g, sh, GroupDraws{B.scenes, B.tri_end, d0, d1 - d0});
    }
}

template <int ETS, int FP, int BLEND>
static void raster_tris_smode(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    const dim3 g(ntiles), b(raster_threads<ETS>());
    if (P.scenes) hipLaunchKernelGGL((k_raster_tris_batch<ETS, FP, BLEND>), g, b, 0, s, P);
    else if (P.one.smode == 2) hipLaunchKernelGGL((k_raster_tris<ETS, FP, BLEND, 2>), g, b, 0, s, P.one);
    else if (P.one.smode == 1) hipLaunchKernelGGL((k_raster_tris<ETS, FP, BLEND, 1>), g, b, 0, s, P.one);
    else hipLaunchKernelGGL((k_raster_tris<ETS, FP, BLEND, 0>), g, b, 0, s, P.one);
}

template <int ETS, int FP>
static void raster_tris_blend(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    if (P.one.blend == SR_BLEND_OVER) raster_tris_smode<ETS, FP, SR_BLEND_OVER>(P, ntiles, s);
    else raster_tris_smode<ETS, FP, SR_BLEND_REPLACE>(P, ntiles, s);
}

template <int ETS>
static void raster_tris_fp(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    switch (P.one.fp) {
        case SR_FP_BLINN_PHONG: raster_tris_blend<ETS, SR_FP_BLINN_PHONG>(P, ntiles, s); break;
        case SR_FP_TEX_LAMBERT: raster_tris_blend<ETS, SR_FP_TEX_LAMBERT>(P, ntiles, s); break;
        case SR_FP_UV_COLOR: raster_tris_blend<ETS, SR_FP_UV_COLOR>(P, ntiles, s); break;
        default: raster_tris_blend<ETS, SR_FP_FLAT>(P, ntiles, s); break;
    }
}

#define SR_CAT2(a, b) a##b
#define SR_CAT(a, b) SR_CAT2(a, b)
void SR_CAT(launch_raster_tris_, SR_ETS)(const RasterBatch& P, uint32_t ntiles, hipStream_t s) {
    if (ntiles) raster_tris_fp<SR_ETS>(P, ntiles, s);
}

}  // namespace sr
