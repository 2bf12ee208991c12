// sr_runtime.hip — the C ABI (include/softrender_hip.h): context, device
// buffers and the draw orchestration.
//
// A draw is the reference chain
//   Pipeline::render_mesh (pipeline/mod.rs:146) -> VertexShader::run
//   (vertex.rs:87) -> [GeometryShader::run (geometry.rs:132)] -> finish
//   (geometry.rs:60) -> FragmentShader::run (fragment.rs:168)
// executed as: vertex kernel -> [clip kernels] -> setup -> scan -> pair
// emission -> stable radix sort by execution tile -> tile ranges ->
// [stencil strip counts] -> raster (one work-group per execution tile).
#include <mutex>

#include "sr_runtime.hpp"

using namespace sr;
using namespace sr_rt;

// ---------------------------------------------------------------------------
// Static device workspace of sr_draw_static: one fixed region per GPU, held
// from the library's first context on that GPU until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: the
// device address of each GPU's copy is looked up once per device at context
// creation and only that address is ever handed to kernels or copies.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kStaticWorkspaceBytes = (size_t)SR_STATIC_WORKSPACE_MIB << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) uint8_t scratch[kStaticWorkspaceBytes];
std::mutex g_scratch_mu;
uint8_t* g_scratch_dev[kMaxDevices] = {};  // device address of each GPU's `scratch`

// With `device` current.  SR_OK once the device's address is known.
int lookup_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return SR_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return SR_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SR_ERR_DEVICE;
    g_scratch_dev[device] = (uint8_t*)p;
    return SR_OK;
}

uint8_t* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}
}  // namespace

namespace sr_rt {

int fail(sr_context ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int ensure(sr_context ctx, DevBuf& b, size_t bytes, hipStream_t s) {
    if (bytes <= b.bytes) return SR_OK;
    if (ctx->arena) {
        // fixed region: a new 256-byte aligned slice past everything carved
        // so far (nothing is freed; a grown buffer simply takes a new slice)
        const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
        if (bytes > ctx->arena_size || off > ctx->arena_size - bytes)
            return fail(ctx, SR_ERR_WORKSPACE,
                        "draw needs more than the static workspace of " +
                            std::to_string(ctx->arena_size >> 20) + " MiB");
        b.p = ctx->arena + off;
        b.bytes = bytes;
        ctx->arena_used = off + bytes;
        return SR_OK;
    }
    if (b.p) {
        HIP_OK(ctx, hipStreamSynchronize(s));
        HIP_OK(ctx, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "hipMalloc workspace failed");
    }
    b.bytes = want;
    return SR_OK;
}

int bits_for(uint32_t n) {
    int b = 1;
    while (b < 32 && (1ull << b) < n) b++;
    return b;
}

int varyings_of(int vp) {
    return vp == SR_VP_WORLD4_NORMAL4 ? 8 : vp == SR_VP_MVP_INT ? 1 : (vp == SR_VP_MVP_NORMAL_WORLD ? 6 : 5);
}

int policy_tile(sr_context ctx, uint32_t w, uint32_t h) {
    const uint64_t px = (uint64_t)w * h;
    for (const auto& e : ctx->tile_policy)
        if (px <= e.first) return e.second & 0xFF;
    return 8;
}

static int valid_waves(uint32_t ets, int w) {
    if (ets == 8) return 1;
    if (ets == 16) return (w == 1 || w == 2) ? w : 4;
    return w == 2 ? 2 : 4;  // 32
}

int draw_waves(sr_context ctx, const sr_draw_desc* d, uint32_t ets, uint32_t w, uint32_t h) {
    int want = d->exec_waves;
    if (!want && !d->exec_tile) {
        const uint64_t px = (uint64_t)w * h;
        for (const auto& e : ctx->tile_policy)
            if (px <= e.first) {
                want = e.second >> 8;
                break;
            }
    }
    return valid_waves(ets, want);
}

}  // namespace sr_rt

namespace {

void mark(sr_context ctx, bool timing, int stage) {
    if (timing) hipEventRecord(ctx->ev[stage + 1], ctx->stream);
}

int read_words(sr_context ctx, int n) {
    HIP_OK(ctx, hipMemcpyAsync(ctx->h_misc, ctx->misc.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

// exclusive scan of counts[0..n) into offsets, total -> misc[slot]
int scan_counts(sr_context ctx, const uint32_t* counts, uint32_t* offsets, uint32_t n, int slot = 1) {
    int rc = ensure(ctx, ctx->scan_tmp, sizeof(uint32_t) * (scan_tmp_words(n) + 16));
    if (rc) return rc;
    launch_scan(counts, offsets, n, P<uint32_t>(ctx->scan_tmp), P<uint32_t>(ctx->misc) + slot,
                ctx->stream);
    return SR_OK;
}

// Sorts (execution tile, item) pairs stably by tile and fills the per-tile
// ranges.  keys/vals hold npairs entries; returns the sorted item array.
int bin_pairs(sr_context ctx, uint32_t npairs, uint32_t ntiles, const uint32_t** sorted_vals) {
    int rc;
    hipStream_t s = ctx->stream;
    if ((rc = ensure(ctx, ctx->hist, sizeof(uint32_t) * (radix_hist_words(npairs) + 256)))) return rc;
    if ((rc = ensure(ctx, ctx->scan_tmp, sizeof(uint32_t) * (scan_tmp_words(radix_hist_words(npairs) + 256) + 16))))
        return rc;
    if ((rc = ensure(ctx, ctx->tstart, sizeof(uint32_t) * ntiles)) ||
        (rc = ensure(ctx, ctx->tend, sizeof(uint32_t) * ntiles)))
        return rc;
    const int swapped = radix_sort(P<uint32_t>(ctx->keys), P<uint32_t>(ctx->vals), P<uint32_t>(ctx->keys2),
                                   P<uint32_t>(ctx->vals2), npairs, bits_for(ntiles), P<uint32_t>(ctx->hist),
                                   P<uint32_t>(ctx->scan_tmp), P<uint32_t>(ctx->misc) + 8, s);
    const uint32_t* skeys = swapped ? P<uint32_t>(ctx->keys2) : P<uint32_t>(ctx->keys);
    *sorted_vals = swapped ? P<uint32_t>(ctx->vals2) : P<uint32_t>(ctx->vals);
    launch_ranges_search(skeys, npairs, ntiles, P<uint32_t>(ctx->tstart), P<uint32_t>(ctx->tend), s);
    return SR_OK;
}

int ensure_pairs(sr_context ctx, uint32_t npairs) {
    int rc;
    const size_t pb = sizeof(uint32_t) * std::max<uint32_t>(npairs, 1);
    if ((rc = ensure(ctx, ctx->keys, pb)) || (rc = ensure(ctx, ctx->vals, pb)) ||
        (rc = ensure(ctx, ctx->keys2, pb)) || (rc = ensure(ctx, ctx->vals2, pb)))
        return rc;
    return SR_OK;
}

// Pairs the pair buffers, the radix histogram and the scan space hold: the
// largest n with n entries in each pair buffer, radix_hist_words(n) + 256
// histogram words and scan_tmp_words(radix_hist_words(n) + 256) + 16 scan
// words (ensure_pair_capacity(n) makes all three hold).
uint32_t pair_capacity(sr_context ctx) {
    uint64_t n = std::min(std::min(ctx->keys.bytes, ctx->vals.bytes), std::min(ctx->keys2.bytes, ctx->vals2.bytes)) /
                 sizeof(uint32_t);
    n = std::min<uint64_t>(n, 0x7FFFFFFFu);
    // radix_hist_words(n) = 256 * blocks_for(n, SCAN_TILE): the blocks the histogram holds
    const uint64_t hw = ctx->hist.bytes / sizeof(uint32_t);
    const uint64_t hist_blocks = hw > 256 ? (hw - 256) / 256 : 0;
    n = std::min<uint64_t>(n, hist_blocks * SCAN_TILE);
    // the histogram's scan: scan_tmp_words(256 * blocks + 256) + 16 words
    while (n > 0 && sizeof(uint32_t) * ((uint64_t)scan_tmp_words(radix_hist_words((uint32_t)n) + 256) + 16) >
                        ctx->scan_tmp.bytes)
        n = n > SCAN_TILE ? ((n + SCAN_TILE - 1) / SCAN_TILE - 1) * (uint64_t)SCAN_TILE : 0;
    return (uint32_t)n;
}

// Grows the pair buffers, histogram and scan space to hold n pairs.
int ensure_pair_capacity(sr_context ctx, uint64_t n) {
    if (n >= 0x7FFFFFFFull) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "draw: bin pairs exceed 2^31");
    if (pair_capacity(ctx) >= n) return SR_OK;
    int rc;
    const uint32_t m = (uint32_t)n;
    if ((rc = ensure_pairs(ctx, m))) return rc;
    if ((rc = ensure(ctx, ctx->hist, sizeof(uint32_t) * (radix_hist_words(m) + 256)))) return rc;
    return ensure(ctx, ctx->scan_tmp, sizeof(uint32_t) * (scan_tmp_words(radix_hist_words(m) + 256) + 16));
}

struct DrawPlan {
    uint32_t W, H, ets, ntx, nty, ntiles, tw, th, scols, srows;
    int nwv = 1;  // waves per execution tile
    int K, smode;
    bool exhaustive, timing;
    // the raster of sr_raster_fmt.hip (other framebuffer formats, the
    // multi-output program, the normalised over-blend)
    bool general = false;
};

void fill_raster_params(const DrawPlan& pl, sr_framebuffer fb, const sr_draw_desc* d,
                        sr_texture tex, RasterParams& R) {
    memset(&R, 0, sizeof(R));
    R.W = pl.W; R.H = pl.H;
    R.color = fb->color; R.depth = fb->depth; R.stencil = fb->stencil;
    R.tw = pl.tw; R.th = pl.th; R.scols = pl.scols; R.srows = pl.srows;
    R.ets = pl.ets; R.ntx = pl.ntx; R.nty = pl.nty; R.nwv = pl.nwv;
    R.K = pl.K;
    R.fp = d->fragment_program; R.blend = d->blend; R.alpha_mode = d->alpha_mode;
    R.s_test = d->stencil_test; R.s_op = d->stencil_op; R.s_wrap = d->stencil_wrap;
    R.s_ref = d->stencil_ref & 0xFFu;
    R.smode = pl.smode;
    R.exhaustive = pl.exhaustive ? 1 : 0;
    if (tex) { R.tex = tex->d; R.tex_fmt = tex->fmt; R.texW = tex->w; R.texH = tex->h; }
    R.tex_filter = d->tex_filter; R.tex_edge = d->tex_edge; R.tex_border = d->tex_border;
    memcpy(R.fsu, d->fs_uniforms, sizeof(R.fsu));
}

// ---------------------------------------------------------------------------
// The triangle draw
// ---------------------------------------------------------------------------
int launch_tri_raster(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, RasterParams R, uint32_t ntris);

int draw_tris(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, const sr_draw_desc* d,
              sr_texture tex, VertexBuf vb, const uint32_t* idx, uint32_t ntris) {
    int rc;
    hipStream_t s = ctx->stream;
    ctx->last.triangles_setup = ntris;
    // one buffer: the triangle records, their binning records, their vertex ids
    const uint32_t nrec = std::max<uint32_t>(ntris, 1);
    if ((rc = ensure(ctx, ctx->recs, (sizeof(TriRec) + sizeof(TriBin) + sizeof(TriVid)) * nrec))) return rc;
    TriBin* const bins = (TriBin*)(P<TriRec>(ctx->recs) + nrec);
    TriVid* const vids = (TriVid*)(bins + nrec);
    if ((rc = ensure(ctx, ctx->counts, sizeof(uint32_t) * (ntris + 1)))) return rc;
    if ((rc = ensure(ctx, ctx->offsets, sizeof(uint32_t) * (ntris + 1)))) return rc;
    SetupParams sp;
    sp.v = vb; sp.idx = idx; sp.n = ntris; sp.cull = d->cull; sp.W = pl.W; sp.H = pl.H;
    sp.ets = pl.ets; sp.tw = pl.tw; sp.th = pl.th;
    sp.recs = P<TriRec>(ctx->recs); sp.bins = bins; sp.vids = idx ? vids : nullptr;
    sp.counts = P<uint32_t>(ctx->counts);
    sp.err = P<uint32_t>(ctx->misc);
    // with no stencil effects, tiles whose pixels the triangle cannot cover
    // are left out of the bins
    sp.empty_cull = pl.smode == 0 && !pl.exhaustive;
    launch_tri_setup(sp, s);
    mark(ctx, pl.timing, ST_SETUP);

    RasterParams R;
    fill_raster_params(pl, fb, d, tex, R);
    R.tris = P<TriRec>(ctx->recs);
    R.vids = idx ? vids : nullptr;  // no indices: vertices 3t, 3t+1, 3t+2
    R.verts = vb;
    R.skip_if = nullptr;
    if (pl.exhaustive) {
        if ((rc = read_words(ctx, 1))) return rc;
        if (ctx->h_misc[0] & 1u) return fail(ctx, SR_ERR_INVALID_COORDINATE, "NaN screen coordinate");
        // every execution tile walks all triangles in order: identity item
        // list, every tile's range [0, ntris)
        if ((rc = ensure_pairs(ctx, ntris))) return rc;
        if ((rc = ensure(ctx, ctx->tstart, sizeof(uint32_t) * pl.ntiles)) ||
            (rc = ensure(ctx, ctx->tend, sizeof(uint32_t) * pl.ntiles)))
            return rc;
        launch_iota(P<uint32_t>(ctx->vals), ntris, s);
        launch_fill(P<uint32_t>(ctx->tstart), 0u, pl.ntiles, s);
        launch_fill(P<uint32_t>(ctx->tend), ntris, pl.ntiles, s);
        R.tile_start = P<uint32_t>(ctx->tstart);
        R.tile_end = P<uint32_t>(ctx->tend);
        R.items = P<uint32_t>(ctx->vals);
        for (int st = ST_SCAN; st <= ST_RANGES; st++) mark(ctx, pl.timing, st);
    } else {
        if ((rc = scan_counts(ctx, P<uint32_t>(ctx->counts), P<uint32_t>(ctx->offsets), ntris)))
            return rc;
        mark(ctx, pl.timing, ST_SCAN);
        // The error flags and the pair count travel to the host while the
        // bins are built and the raster runs, all sized on the device by
        // the pair count: the pair buffers hold `cap` pairs; a draw with
        // more (misc[4] set) or with a NaN coordinate draws nothing, and
        // the host, once the count arrives, grows the buffers and bins and
        // draws again.  So the host waits for setup only, never between
        // kernels.
        HIP_OK(ctx, hipMemcpyAsync(ctx->h_misc, ctx->misc.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(ctx, hipEventRecord(ctx->ev_read, s));
        if ((rc = ensure(ctx, ctx->tstart, sizeof(uint32_t) * pl.ntiles)) ||
            (rc = ensure(ctx, ctx->tend, sizeof(uint32_t) * pl.ntiles)))
            return rc;
        // a first guess for a context's first draws: 4 pairs per triangle
        // (none in the fixed workspace, which sizes them exactly)
        if (!ctx->arena && (rc = ensure_pair_capacity(ctx, (uint64_t)ntris * 4 + 4096))) return rc;
        auto bin_and_draw = [&]() -> int {
            const uint32_t cap = pair_capacity(ctx);
            EmitParams ep;
            ep.recs = P<TriRec>(ctx->recs); ep.bins = bins; ep.counts = P<uint32_t>(ctx->counts);
            ep.offsets = P<uint32_t>(ctx->offsets); ep.n = ntris; ep.W = pl.W; ep.H = pl.H;
            ep.ets = pl.ets; ep.ntx = pl.ntx; ep.tw = pl.tw; ep.th = pl.th; ep.key_base = 0; ep.val_base = 0;
            ep.empty_cull = sp.empty_cull;
            ep.keys = P<uint32_t>(ctx->keys);
            ep.vals = P<uint32_t>(ctx->vals);
            ep.cap = cap;
            launch_tri_emit(ep, s);
            uint32_t* misc = P<uint32_t>(ctx->misc);
            launch_bins_live(misc, cap, s);  // misc[4]: overflow, misc[5]: pairs to sort (0 on overflow or NaN)
            mark(ctx, pl.timing, ST_EMIT);
            const int swapped = radix_sort(P<uint32_t>(ctx->keys), P<uint32_t>(ctx->vals),
                                           P<uint32_t>(ctx->keys2), P<uint32_t>(ctx->vals2), cap,
                                           bits_for(pl.ntiles), P<uint32_t>(ctx->hist),
                                           P<uint32_t>(ctx->scan_tmp), misc + 2, s, misc + 5);
            const uint32_t* skeys = swapped ? P<uint32_t>(ctx->keys2) : P<uint32_t>(ctx->keys);
            const uint32_t* svals = swapped ? P<uint32_t>(ctx->vals2) : P<uint32_t>(ctx->vals);
            mark(ctx, pl.timing, ST_SORT);
            launch_ranges_search(skeys, cap, pl.ntiles, P<uint32_t>(ctx->tstart), P<uint32_t>(ctx->tend), s,
                                 misc + 5);
            mark(ctx, pl.timing, ST_RANGES);
            R.tile_start = P<uint32_t>(ctx->tstart);
            R.tile_end = P<uint32_t>(ctx->tend);
            R.items = svals;
            R.skip_if = misc;
            return launch_tri_raster(ctx, fb, pl, R, ntris);
        };
        if ((rc = bin_and_draw())) return rc;
        HIP_OK(ctx, hipEventSynchronize(ctx->ev_read));
        if (ctx->h_misc[0] & 1u) return fail(ctx, SR_ERR_INVALID_COORDINATE, "NaN screen coordinate");
        // a triangle that may pass the coverage test outside its box went
        // to every tile such a pixel may lie in (the raster then tests it
        // in its clamped boxes)
        ctx->last.strip_slivers = (ctx->h_misc[0] & 2u) ? 1u : 0u;
        const uint32_t npairs = ctx->h_misc[1];
        ctx->last.bin_pairs = npairs;
        if (npairs > pair_capacity(ctx)) {
            const uint64_t want = ctx->arena ? (uint64_t)npairs : (uint64_t)npairs + npairs / 4 + 4096;
            if ((rc = ensure_pair_capacity(ctx, want))) return rc;
            if (npairs > pair_capacity(ctx)) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "draw: bin capacity did not grow");
            if ((rc = bin_and_draw())) return rc;
        }
        return SR_OK;
    }
    return launch_tri_raster(ctx, fb, pl, R, ntris);
}

// The raster of a triangle draw whose bins R names: the per-pixel raster for
// the general formats, else the stencil strips (counting draws) and the
// tile raster.
int launch_tri_raster(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, RasterParams R, uint32_t ntris) {
    int rc;
    hipStream_t s = ctx->stream;
    if (pl.general) {
        // every format through the per-pixel raster; a stencil with a test
        // or op is tracked there in order (the draw is exhaustive then)
        FmtParams F;
        memset(&F, 0, sizeof(F));
        for (uint32_t p = 0; p < fb->fmt.n_color; p++) {
            F.planes[p] = fb->planes[p];
            F.color_fmt[p] = fb->fmt.color[p];
        }
        if (fb->standard()) {
            F.planes[0] = fb->color;
            F.color_fmt[0] = SR_COLOR_RGBA8;
        }
        F.n_color = fb->fmt.n_color;
        F.depth64 = fb->depth64;
        F.has_stencil = fb->stencil != nullptr;
        F.track_stencil = F.has_stencil && pl.smode != 0;
        mark(ctx, pl.timing, ST_STRIPS);
        launch_raster_fmt(R, F, pl.ntiles, s);
        mark(ctx, pl.timing, ST_RASTER);
        return SR_OK;
    }
    if (pl.smode == 1) {
        if ((rc = ensure(ctx, ctx->strips, sizeof(uint32_t) * (size_t)pl.W * pl.H))) return rc;
        const size_t aux = (size_t)(pl.scols + 1) * (pl.H + 1) * 2 + (size_t)(pl.srows + 1) * (pl.W + 1) * 2 +
                           (size_t)(pl.srows + 1) * (pl.scols + 1) * 4;
        StripParams stp;
        stp.recs = P<TriRec>(ctx->recs); stp.n = ntris; stp.W = pl.W; stp.H = pl.H;
        stp.tw = pl.tw; stp.th = pl.th; stp.scols = pl.scols; stp.srows = pl.srows;
        // after the difference arrays: the work-groups' partial hot rows
        // (written whole by each work-group, so not cleared)
        const uint32_t hot_groups = strip_hot_groups(stp);
        const size_t hot_words = (size_t)hot_groups * (pl.H + 1 + pl.W + 1);
        if ((rc = ensure(ctx, ctx->strip_aux, sizeof(uint32_t) * (aux + hot_words)))) return rc;
        uint32_t* a = P<uint32_t>(ctx->strip_aux);
        stp.hot_part = hot_groups ? a + aux : nullptr;
        stp.colL = a; a += (size_t)(pl.scols + 1) * (pl.H + 1);
        stp.colR = a; a += (size_t)(pl.scols + 1) * (pl.H + 1);
        stp.rowT = a; a += (size_t)(pl.srows + 1) * (pl.W + 1);
        stp.rowB = a; a += (size_t)(pl.srows + 1) * (pl.W + 1);
        const size_t nc = (size_t)(pl.srows + 1) * (pl.scols + 1);
        stp.cLL = a; a += nc;
        stp.cLR = a; a += nc;
        stp.cRL = a; a += nc;
        stp.cRR = a; a += nc;
        stp.out = P<uint32_t>(ctx->strips);
        launch_fill(P<uint32_t>(ctx->strip_aux), 0u, (uint32_t)aux, s);
        launch_strips(stp, s);
        R.strip_counts = P<uint32_t>(ctx->strips);
    }
    mark(ctx, pl.timing, ST_STRIPS);
    RasterBatch RB;
    memset(&RB, 0, sizeof(RB));
    RB.one = R;
    RB.ntiles = pl.ntiles;
    launch_raster_tris(RB, pl.ntiles, s);
    mark(ctx, pl.timing, ST_RASTER);
    return SR_OK;
}

}  // namespace

int draw_lines_impl(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, const sr_draw_desc* d,
                    sr_texture tex, VertexBuf vb, const uint32_t* idx, uint32_t nlines);
int draw_points_impl(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, const sr_draw_desc* d,
                     sr_texture tex, VertexBuf vb, const uint32_t* idx, uint32_t npoints);

int sr_rt::validate_draw(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex) {
    if (!ctx || !fb || !mesh || !d) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: null argument");
    if (fb->ctx != ctx || mesh->ctx != ctx) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: handle from another context");
    if (d->vertex_program < SR_VP_MVP_NORMAL_UV || d->vertex_program > SR_VP_MVP_INT)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown vertex program");
    if (d->vertex_program == SR_VP_WORLD4_NORMAL4 && d->geometry_program != SR_GS_NONE)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: no geometry program takes the 4-vector varyings");
    if (d->vertex_program == SR_VP_MVP_INT && d->geometry_program != SR_GS_NONE)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: no geometry program takes the integer varying");
    if (d->vertex_program == SR_VP_MVP_INT && !mesh->iattr)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: the integer vertex program needs sr_mesh_set_int_attrib");
    if (d->fragment_program == SR_FP_INT_PALETTE && d->vertex_program != SR_VP_MVP_INT)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: the palette program needs the integer varying");
    if (d->vertex_program == SR_VP_MVP_INT && d->fragment_program != SR_FP_INT_PALETTE &&
        d->fragment_program != SR_FP_FLAT)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: the integer varying feeds the palette and flat programs only");
    if (d->fragment_program == SR_FP_FRESNEL4 && d->vertex_program != SR_VP_WORLD4_NORMAL4)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: the Fresnel program needs the 4-vector varyings");
    if (d->vertex_program == SR_VP_WORLD4_NORMAL4 && d->fragment_program != SR_FP_FRESNEL4 &&
        d->fragment_program != SR_FP_FLAT)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: the 4-vector varyings feed the Fresnel and flat programs only");
    if (d->tex_filter < SR_FILTER_NEAREST || d->tex_filter > SR_FILTER_SPAN_BILINEAR)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown texture filter");
    if (d->geometry_program < SR_GS_NONE || d->geometry_program > SR_GS_FACE_NORMALS)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown geometry program");
    if ((d->geometry_program == SR_GS_VERTEX_NORMALS || d->geometry_program == SR_GS_FACE_NORMALS) &&
        d->vertex_program != SR_VP_MVP_NORMAL_WORLD)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: normal visualisation needs world-space varyings");
    if (d->fragment_program < SR_FP_FLAT || d->fragment_program > SR_FP_INT_PALETTE)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown fragment program");
    if (d->blend < SR_BLEND_REPLACE || d->blend > SR_BLEND_OVER_NORMALIZED)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown blend");
    if ((d->fragment_program == SR_FP_TEX_LAMBERT || d->fragment_program == SR_FP_GBUFFER) && !tex)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: fragment program samples a texture");
    if (tex && tex->view && tex->fb == fb)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: samples the framebuffer it draws into");
    if (d->fragment_program == SR_FP_BLINN_PHONG && d->vertex_program != SR_VP_MVP_NORMAL_WORLD)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: Blinn-Phong needs world-space varyings");
    if ((d->fragment_program == SR_FP_TEX_LAMBERT || d->fragment_program == SR_FP_UV_COLOR ||
         d->fragment_program == SR_FP_GBUFFER) &&
        d->vertex_program != SR_VP_MVP_NORMAL_UV)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: program needs uv varyings");
    if (d->stencil_test < 0 || d->stencil_test > 7 || d->stencil_op < 0 || d->stencil_op > 5)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: bad stencil config");
    if (d->exec_tile != 0 && d->exec_tile != 8 && d->exec_tile != 16 && d->exec_tile != 32)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: exec_tile must be 0, 8, 16 or 32");
    if (d->exec_waves != 0 && d->exec_waves != 1 && d->exec_waves != 2 && d->exec_waves != 4)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: exec_waves must be 0, 1, 2 or 4");
    return SR_OK;
}


// ===========================================================================
// Batched rendering: many independent (framebuffer, mesh) triangle draws in
// one pass of the pipeline.  Every stage runs once over the whole batch:
// vertices and triangles are numbered across scenes, execution tiles get
// global ids so one stable sort bins all scenes, and the raster is one
// persistent launch per (tile size, program, blend) class whose
// work-groups each take a run of tiles of near-equal cost (cost = a fixed
// term + the tile's binned items, cut on the device from a prefix sum).
// Sizes that depend on the geometry (bin pairs) use a
// capacity checked on the device: on overflow nothing is drawn, the
// capacity grows and the batch runs again, so the sequence has no host
// round trip; it can be captured once into a hipGraph and replayed while
// the batch is unchanged, or queued directly on a stream of its own.
// ===========================================================================
namespace sr_rt {

int BatchWork::init(sr_context ctx, bool new_stream) {
    own_stream = new_stream;
    if (new_stream) HIP_OK(ctx, hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    else stream = ctx->stream;
    HIP_OK(ctx, hipHostMalloc((void**)&h_misc, 16 * sizeof(uint32_t), hipHostMallocDefault));
    HIP_OK(ctx, hipEventCreate(&ev0));
    HIP_OK(ctx, hipEventCreate(&ev1));
    cap = ctx->batch_cap_hint;
    return SR_OK;
}

void BatchWork::drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
}

void BatchWork::release() {
    if (stream) hipStreamSynchronize(stream);
    drop_graph();
    DevBuf* bufs[] = {&verts, &recs, &counts, &offsets, &keys, &vals, &keys2,
                      &vals2, &tstart, &tend, &hist, &scan, &tables, &misc, &cost, &parts};
    for (DevBuf* b : bufs) {
        if (b->p) hipFree(b->p);
        b->p = nullptr;
        b->bytes = 0;
    }
    if (h_stage) hipHostFree(h_stage);
    if (h_misc) hipHostFree(h_misc);
    h_stage = nullptr;
    h_misc = nullptr;
    h_stage_bytes = 0;
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    ev0 = ev1 = nullptr;
    if (own_stream && stream) hipStreamDestroy(stream);
    stream = nullptr;
}

size_t BatchWork::device_bytes() const {
    const DevBuf* bufs[] = {&verts, &recs, &counts, &offsets, &keys, &vals, &keys2,
                            &vals2, &tstart, &tend, &hist, &scan, &tables, &misc, &cost, &parts};
    size_t t = 0;
    for (const DevBuf* b : bufs) t += b->bytes;
    return t;
}

size_t batch_bytes_estimate(uint64_t tris, uint64_t verts, double pairs_per_tri) {
    // screen vertices (10 words), triangle and binning records, counts +
    // offsets, four pair arrays of the sort, radix histograms; 25% growth
    // headroom
    const double pairs = (double)tris * pairs_per_tri + 65536.0;
    const double b = 40.0 * (double)verts + (double)(sizeof(TriRec) + sizeof(TriBin) + sizeof(TriVid)) * (double)tris +
                     8.0 * (double)tris +
                     16.0 * pairs + 4.0 * (double)radix_hist_words((uint32_t)std::min(pairs, 2e9));
    return (size_t)(b * 1.25);
}

// A texture read only the per-pixel raster's sampler does: an RGBAF32
// texture (a texture buffer plane) or the full example's sampler
bool samples_f32(const sr_draw_desc* d, sr_texture tex) {
    return tex && (tex->fmt == SR_COLOR_RGBAF32 || d->tex_filter >= SR_FILTER_SPAN_NEAREST) &&
           (d->fragment_program == SR_FP_TEX_LAMBERT || d->fragment_program == SR_FP_GBUFFER);
}

bool batchable(sr_mesh m, const sr_draw_desc* d, sr_texture tex) {
    return m->prim == SR_PRIM_TRIANGLES && d->geometry_program == SR_GS_NONE && !samples_f32(d, tex) &&
           d->vertex_program != SR_VP_WORLD4_NORMAL4 && d->vertex_program != SR_VP_MVP_INT &&
           d->stencil_test == SR_STENCIL_ALWAYS && d->stencil_op == SR_STENCIL_KEEP &&
           !(d->exec_flags & SR_EXEC_EXHAUSTIVE) && d->fragment_program != SR_FP_GBUFFER &&
           (d->blend == SR_BLEND_REPLACE || d->blend == SR_BLEND_OVER);
}

int batch_plan(sr_context ctx, const sr_batch_item* items, uint32_t n, BatchPlan& pl) {
    // groups: framebuffers in order of first appearance, items in order
    std::vector<sr_framebuffer> fbs;
    std::vector<std::vector<uint32_t>> members;
    for (uint32_t i = 0; i < n; i++) {
        size_t g = 0;
        while (g < fbs.size() && fbs[g] != items[i].fb) g++;
        if (g == fbs.size()) {
            fbs.push_back(items[i].fb);
            members.emplace_back();
        }
        members[g].push_back(i);
    }
    struct Order {
        uint32_t group;
        uint32_t ets, nwv;
        uint64_t cls;  // (ets, waves, fp, blend)
    };
    std::vector<Order> order(fbs.size());
    for (uint32_t g = 0; g < fbs.size(); g++) {
        const sr_batch_item& it = items[members[g][0]];
        const sr_draw_desc* d = it.desc;
        const uint32_t ets = d->exec_tile ? (uint32_t)d->exec_tile : (uint32_t)policy_tile(ctx, it.fb->w, it.fb->h);
        const uint32_t nwv = (uint32_t)draw_waves(ctx, d, ets, it.fb->w, it.fb->h);
        order[g] = Order{g, ets, nwv,
                         ((uint64_t)ets << 24) | ((uint64_t)nwv << 16) | ((uint64_t)d->fragment_program << 8) |
                             (uint64_t)d->blend};
    }
    std::stable_sort(order.begin(), order.end(), [](const Order& a, const Order& b) { return a.cls < b.cls; });
    pl = BatchPlan();
    pl.items.resize(n);
    pl.descs.resize(n);
    pl.geom.resize(n);
    pl.vend.resize(n);
    pl.tend.resize(n);
    uint32_t k = 0;
    for (const Order& o : order) {
        const sr_framebuffer fb = fbs[o.group];
        const uint32_t W = fb->w, H = fb->h, ets = o.ets;
        const uint32_t ntx = (W + ets - 1) / ets, ntiles = ntx * ((H + ets - 1) / ets);
        const uint32_t tile_base = (uint32_t)pl.NTILES;
        pl.group_first.push_back(k);
        for (uint32_t i : members[o.group]) {
            pl.items[k] = items[i];
            pl.descs[k] = *items[i].desc;
            pl.items[k].desc = nullptr;  // the copy in descs[k] is the one used
            const sr_draw_desc* d = &pl.descs[k];
            const sr_mesh m = pl.items[k].mesh;
            const bool drawable = W > 1 && H > 1;  // W or H == 1: no tiles (fragment.rs:196,201)
            const uint32_t ntri = drawable ? mesh_count(m) / 3 : 0;
            SceneGeom& g = pl.geom[k];
            memset(&g, 0, sizeof(g));
            g.px = m->soa[0]; g.py = m->soa[1]; g.pz = m->soa[2]; g.nx = m->soa[3]; g.ny = m->soa[4];
            g.nz = m->soa[5]; g.u = m->soa[6]; g.v = m->soa[7]; g.idx = m->idx;
            g.nv = ntri ? m->nv : 0; g.vbase = (uint32_t)pl.NV;
            g.nt = ntri; g.tbase = (uint32_t)pl.NT;
            g.program = d->vertex_program; g.K = varyings_of(d->vertex_program); g.cull = d->cull;
            g.W = W; g.H = H; g.ets = ets; g.ntx = ntx;
            g.tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u;
            g.th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
            g.tile_base = tile_base;
            g.empty_cull = 1;  // batched draws have no stencil effects
            g.viewport = d->viewport;
            memcpy(g.U, d->vs_uniforms, sizeof(g.U));
            pl.NV += g.nv; pl.NT += g.nt;
            pl.vend[k] = (uint32_t)pl.NV; pl.tend[k] = (uint32_t)pl.NT;
            k++;
        }
        pl.NTILES += ntiles;
        pl.group_tile_end.push_back((uint32_t)pl.NTILES);
        const sr_draw_desc* d0 = items[members[o.group][0]].desc;
        if (pl.classes.empty() || pl.classes.back().cls != o.cls)
            pl.classes.push_back(BatchClass{o.cls, ets, o.nwv, (uint32_t)d0->fragment_program,
                                            (uint32_t)d0->blend, tile_base, 0});
        pl.classes.back().ntiles += ntiles;
    }
    pl.group_first.push_back(k);
    if (pl.NV >= (1ull << 31) || pl.NT >= (1ull << 31) || pl.NTILES >= (1ull << 31))
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: too large");
    return SR_OK;
}

// Batch raster: runs (work-groups) per class, 32 times the work-groups the
// chip holds at once (ETS 8: one wave each, 16 per CU; larger tiles: four
// waves each, 4 per CU), never more than the tiles.  Runs of equal
// estimated cost still differ in real cost (a tile's fixed part depends on
// its framebuffer and record fetch latency), so many short runs, handed
// out by the hardware as work-groups finish, balance better than one run
// per resident work-group: on the 256-scene batch, 2 runs per slot took
// 22.2 ms, 8 took 18.3, 32 took 17.2 and 128 took 17.5.
constexpr uint32_t TILE_FIXED_COST = 16;  // a tile's load/clear + resolve, in item iterations

static uint32_t raster_runs(const BatchClass& C) {
    // work-groups of 64 * nwv threads: 16 / nwv per CU
    return std::min<uint32_t>(C.ntiles, 256u * (16u / C.nwv) * 32u);
}

static size_t raster_runs_total(const BatchPlan& pl) {
    size_t n = 0;
    for (const BatchClass& C : pl.classes) n += raster_runs(C) + 1;
    return n;
}

int batch_enqueue(sr_context ctx, BatchWork& W, const BatchPlan& pl, bool use_graph) {
    int rc;
    hipStream_t s = W.stream;
    const uint32_t ns = (uint32_t)pl.items.size();
    const uint64_t NV = pl.NV, NT = pl.NT, NTILES = pl.NTILES;
    if (W.cap == 0) W.cap = (uint32_t)std::min<uint64_t>(4 * NT + (1u << 16), 0x7FFFFFFFull);
    const uint32_t cap = W.cap;
    const uint32_t nvb = (uint32_t)std::max<uint64_t>(NV, 1);
    const uint32_t ng = (uint32_t)pl.group_tile_end.size();
    // bins are sorted group by group on local tile ids: the passes follow
    // the largest group's tile count
    uint32_t max_group_tiles = 2;
    for (uint32_t g = 0; g < ng; g++)
        max_group_tiles = std::max(max_group_tiles, pl.group_tile_end[g] - (g ? pl.group_tile_end[g - 1] : 0u));
    const int local_bits = bits_for(max_group_tiles);
    const uint64_t hist_words = radix_hist_words(cap) + 256ull * ng;
    const uint64_t seg_words = 4ull * ng + 2;
    const size_t before = W.device_bytes();
    if ((rc = ensure(ctx, W.verts, sizeof(float) * 10ull * nvb, s)) ||
        (rc = ensure(ctx, W.recs, (sizeof(TriRec) + sizeof(TriBin) + sizeof(TriVid)) * std::max<uint64_t>(NT, 1), s)) ||
        (rc = ensure(ctx, W.counts, sizeof(uint32_t) * (NT + 1), s)) ||
        (rc = ensure(ctx, W.offsets, sizeof(uint32_t) * (NT + 1), s)) ||
        (rc = ensure(ctx, W.keys, sizeof(uint32_t) * (size_t)cap, s)) ||
        (rc = ensure(ctx, W.vals, sizeof(uint32_t) * (size_t)cap, s)) ||
        (rc = ensure(ctx, W.keys2, sizeof(uint32_t) * (size_t)cap, s)) ||
        (rc = ensure(ctx, W.vals2, sizeof(uint32_t) * (size_t)cap, s)) ||
        (rc = ensure(ctx, W.tstart, sizeof(uint32_t) * std::max<uint64_t>(NTILES, 1), s)) ||
        (rc = ensure(ctx, W.tend, sizeof(uint32_t) * std::max<uint64_t>(NTILES, 1), s)) ||
        (rc = ensure(ctx, W.hist, sizeof(uint32_t) * (hist_words + seg_words + 256), s)) ||
        (rc = ensure(ctx, W.scan, sizeof(uint32_t) * (scan_tmp_words((uint32_t)std::max<uint64_t>(
                                      std::max<uint64_t>(hist_words, NT + 1), NTILES + 1)) + 16), s)) ||
        (rc = ensure(ctx, W.misc, sizeof(uint32_t) * 8, s)) ||
        (rc = ensure(ctx, W.cost, sizeof(uint32_t) * (NTILES + 1), s)) ||
        (rc = ensure(ctx, W.parts, sizeof(uint32_t) * (raster_runs_total(pl) + pl.classes.size()), s)))
        return rc;
    const size_t tbytes = sizeof(SceneGeom) * ns + sizeof(RasterParams) * ns +
                          sizeof(uint32_t) * (2 * ns + 2 * ng + 1) + 64;
    if ((rc = ensure(ctx, W.tables, tbytes, s))) return rc;
    if (W.device_bytes() != before) W.drop_graph();  // a buffer moved

    // scene tables (device copies, pointers into the workspace)
    uint8_t* tb = P<uint8_t>(W.tables);
    SceneGeom* d_geom = (SceneGeom*)tb;
    RasterParams* d_rp = (RasterParams*)(tb + sizeof(SceneGeom) * ns);
    uint32_t* d_vend = (uint32_t*)(tb + sizeof(SceneGeom) * ns + sizeof(RasterParams) * ns);
    uint32_t* d_tend = d_vend + ns;
    uint32_t* d_group_tile_end = d_tend + ns;
    uint32_t* d_group_first = d_group_tile_end + ng;
    uint32_t* misc = P<uint32_t>(W.misc);  // [0] overflow [1] pairs [2] err [3..4] sort [5] tile cost total
                                            // [6] pairs to sort (0 on overflow)
    // varying rows as wide as the widest vertex program of the batch needs
    // (at most 6: the workspace holds 10 floats a vertex)
    uint32_t kmax = 1;
    for (const SceneGeom& g : pl.geom) kmax = std::max(kmax, (uint32_t)g.K);
    const VertexBuf vb{P<float>(W.verts), nvb, kmax};
    const int swapped_passes = (local_bits + 7) / 8;
    const uint32_t* sorted_vals = (swapped_passes & 1) ? P<uint32_t>(W.vals2) : P<uint32_t>(W.vals);
    std::vector<RasterParams> rp(ns);
    bool need_vids = false;  // some draw shades through the vertex-id table
    for (uint32_t k = 0; k < ns; k++) {
        const sr_batch_item& it = pl.items[k];
        const sr_draw_desc* d = &pl.descs[k];
        RasterParams& R = rp[k];
        memset(&R, 0, sizeof(R));
        R.W = it.fb->w; R.H = it.fb->h;
        R.color = it.fb->color; R.depth = it.fb->depth; R.stencil = it.fb->stencil;
        R.tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u;
        R.th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
        R.scols = R.W > 1 ? (R.W - 1 + R.tw - 1) / R.tw : 0;
        R.srows = R.H > 1 ? (R.H - 1 + R.th - 1) / R.th : 0;
        R.ets = pl.geom[k].ets; R.ntx = pl.geom[k].ntx; R.nty = (R.H + R.ets - 1) / R.ets;
        R.nwv = (int)draw_waves(ctx, d, R.ets, R.W, R.H);
        R.tile_start = P<uint32_t>(W.tstart); R.tile_end = P<uint32_t>(W.tend);
        R.items = sorted_vals;
        R.tris = P<TriRec>(W.recs);
        // a scene without indices whose vertices start at 3x its first
        // triangle needs no vertex ids: vertex 3t+k
        const bool ident = !it.mesh->indexed && pl.geom[k].vbase == 3ull * pl.geom[k].tbase;
        need_vids = need_vids || !ident;
        R.vids = ident ? nullptr
                       : (const TriVid*)((const TriBin*)(P<TriRec>(W.recs) + std::max<uint64_t>(NT, 1)) +
                                         std::max<uint64_t>(NT, 1));
        R.verts = vb;
        R.K = pl.geom[k].K;
        R.fp = d->fragment_program; R.blend = d->blend; R.alpha_mode = d->alpha_mode;
        R.s_test = SR_STENCIL_ALWAYS; R.s_op = SR_STENCIL_KEEP;
        R.smode = 0;
        if (it.tex) { R.tex = it.tex->d; R.tex_fmt = it.tex->fmt; R.texW = it.tex->w; R.texH = it.tex->h; }
        R.tex_filter = d->tex_filter; R.tex_edge = d->tex_edge; R.tex_border = d->tex_border;
        memcpy(R.fsu, d->fs_uniforms, sizeof(R.fsu));
        R.tile_base = pl.geom[k].tile_base;
        R.clear = it.clear ? 1 : 0;
        R.clear_rgba = it.clear_rgba;
    }
    // the captured sequence depends on the capacity and the tables
    std::vector<uint8_t> key;
    if (use_graph) {
        auto put = [&](const void* p, size_t b) {
            const uint8_t* q = (const uint8_t*)p;
            key.insert(key.end(), q, q + b);
        };
        put(&cap, sizeof(cap));
        put(pl.geom.data(), sizeof(SceneGeom) * ns);
        put(rp.data(), sizeof(RasterParams) * ns);
    }
    const bool replay = use_graph && W.exec && key == W.key;
    if (!replay) {
        W.drop_graph();
        if (W.h_stage_bytes < tbytes) {
            if (W.h_stage) hipHostFree(W.h_stage);
            W.h_stage = nullptr;
            W.h_stage_bytes = 0;
            if (hipHostMalloc((void**)&W.h_stage, tbytes, hipHostMallocDefault) != hipSuccess)
                return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: staging alloc failed");
            W.h_stage_bytes = tbytes;
        }
        uint8_t* hs = W.h_stage;
        memcpy(hs, pl.geom.data(), sizeof(SceneGeom) * ns);
        memcpy(hs + sizeof(SceneGeom) * ns, rp.data(), sizeof(RasterParams) * ns);
        uint32_t* hv = (uint32_t*)(hs + sizeof(SceneGeom) * ns + sizeof(RasterParams) * ns);
        memcpy(hv, pl.vend.data(), 4 * ns);
        memcpy(hv + ns, pl.tend.data(), 4 * ns);
        memcpy(hv + 2 * ns, pl.group_tile_end.data(), 4 * ng);
        memcpy(hv + 2 * ns + ng, pl.group_first.data(), 4 * (ng + 1));
        if (!use_graph) HIP_OK(ctx, hipEventRecord(W.ev0, s));
        HIP_OK(ctx, hipMemcpyAsync(tb, hs, tbytes - 64, hipMemcpyHostToDevice, s));
        if (use_graph) {
            HIP_OK(ctx, hipStreamSynchronize(s));
            HIP_OK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        }
        BatchTables T{d_geom, d_vend, d_tend, ns, (uint32_t)NV, (uint32_t)NT};
        launch_fill(misc, 0u, 8u, s);
        launch_vertex_batch(T, vb, s);
        // the triangle records, then their binning records, in one buffer
        TriBin* const bins = (TriBin*)(P<TriRec>(W.recs) + std::max<uint64_t>(NT, 1));
        TriVid* const vids = (TriVid*)(bins + std::max<uint64_t>(NT, 1));
        launch_tri_setup_batch(T, vb, P<TriRec>(W.recs), bins, need_vids ? vids : nullptr, P<uint32_t>(W.counts),
                               misc + 2, s);
        launch_scan(P<uint32_t>(W.counts), P<uint32_t>(W.offsets), (uint32_t)NT, P<uint32_t>(W.scan), misc + 1, s);
        launch_tri_emit_batch(T, P<TriRec>(W.recs), bins, P<uint32_t>(W.counts), P<uint32_t>(W.offsets), cap,
                              P<uint32_t>(W.keys), P<uint32_t>(W.vals), misc + 0, s);
        // on overflow nothing is sorted or binned (the raster skips the batch
        // and it runs again with a larger capacity)
        launch_live_pairs(misc, 0, 1, 6, s);
        const BatchGroups BG{d_tend, d_group_first, d_group_tile_end, ng, (uint32_t)NT};
        const int sw = radix_sort_groups(P<uint32_t>(W.keys), P<uint32_t>(W.vals), P<uint32_t>(W.keys2),
                                         P<uint32_t>(W.vals2), cap, local_bits, BG, P<uint32_t>(W.offsets),
                                         misc + 6, P<uint32_t>(W.hist) + hist_words, P<uint32_t>(W.hist),
                                         P<uint32_t>(W.scan), misc + 3, s);
        const uint32_t* skeys = sw ? P<uint32_t>(W.keys2) : P<uint32_t>(W.keys);
        launch_fill(P<uint32_t>(W.tstart), 0u, (uint32_t)NTILES, s);
        launch_fill(P<uint32_t>(W.tend), 0u, (uint32_t)NTILES, s);
        launch_ranges(skeys, cap, 0u, (uint32_t)NTILES, P<uint32_t>(W.tstart), P<uint32_t>(W.tend), s, misc + 6);
        // cost-balanced runs of tiles for the persistent raster
        launch_tile_cost(P<uint32_t>(W.tstart), P<uint32_t>(W.tend), (uint32_t)NTILES, TILE_FIXED_COST,
                         P<uint32_t>(W.cost), s);
        launch_scan(P<uint32_t>(W.cost), P<uint32_t>(W.cost), (uint32_t)NTILES + 1, P<uint32_t>(W.scan), misc + 5, s);
        uint32_t* runs = P<uint32_t>(W.parts);
        for (size_t c = 0; c < pl.classes.size(); c++) {
            const BatchClass& C = pl.classes[c];
            if (C.ntiles == 0) continue;
            const uint32_t G = raster_runs(C);
            launch_partition(P<uint32_t>(W.cost), C.tile0, C.ntiles, G, runs, s);
            RasterBatch RB;
            memset(&RB, 0, sizeof(RB));
            RB.one.fp = (int)C.fp; RB.one.blend = (int)C.blend; RB.one.ets = C.ets; RB.one.nwv = (int)C.nwv;
            RB.one.smode = 0;
            RB.scenes = d_rp;
            RB.group_tile_end = d_group_tile_end;
            RB.group_first = d_group_first;
            RB.tri_end = d_tend;
            RB.ngroups = ng;
            RB.tile0 = C.tile0;
            RB.ntiles = C.ntiles;
            RB.wg_start = runs;
            RB.overflow = misc;
            launch_raster_tris(RB, G, s);
            runs += G + 1;
        }
        if (use_graph) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (e != hipSuccess) return fail(ctx, SR_ERR_DEVICE, std::string("batch capture: ") + hipGetErrorString(e));
            W.graph = g;
            HIP_OK(ctx, hipGraphInstantiate(&W.exec, g, nullptr, nullptr, 0));
            W.key.swap(key);
        }
    }
    if (use_graph) {
        HIP_OK(ctx, hipEventRecord(W.ev0, s));
        HIP_OK(ctx, hipGraphLaunch(W.exec, s));
    }
    HIP_OK(ctx, hipMemcpyAsync(W.h_misc, misc, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(ctx, hipEventRecord(W.ev1, s));
    return SR_OK;
}

int batch_result(sr_context ctx, BatchWork& W, const BatchPlan& pl, bool* again) {
    *again = false;
    if (W.h_misc[2] & 1u) return fail(ctx, SR_ERR_INVALID_COORDINATE, "batch: NaN screen coordinate");
    if (!W.h_misc[0]) return SR_OK;
    // binning overflowed: nothing was drawn; grow and run again
    const uint64_t want = (uint64_t)W.h_misc[1] + W.h_misc[1] / 4 + 4096;
    if (want >= 0x7FFFFFFFull) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: bin pairs exceed 2^31");
    W.cap = (uint32_t)want;
    W.drop_graph();
    *again = true;
    return SR_OK;
}

}  // namespace sr_rt

namespace {

// Framebuffer::clear with an RGBA8 colour (a batch item's clear): on other
// formats every RGBA8 plane takes it as is and every RGBAF32 plane its
// unorm values (c / 255).
int clear_rgba8(sr_framebuffer fb, uint32_t rgba) {
    if (fb->standard()) return sr_fb_clear(fb, rgba, -3.40282347e+38f, 0);
    uint8_t colors[16 * SR_MAX_COLOR_PLANES];
    uint8_t* c = colors;
    for (uint32_t p = 0; p < fb->fmt.n_color; p++) {
        if (fb->fmt.color[p] == SR_COLOR_RGBAF32) {
            const float f[4] = {(float)(rgba & 0xFFu) / 255.0f, (float)((rgba >> 8) & 0xFFu) / 255.0f,
                                (float)((rgba >> 16) & 0xFFu) / 255.0f, (float)(rgba >> 24) / 255.0f};
            memcpy(c, f, 16);
            c += 16;
        } else {
            memcpy(c, &rgba, 4);
            c += 4;
        }
    }
    return sr_fb_clear_ex(fb, colors);
}

// Validates a batch and draws, one by one and in item order, the items of
// every framebuffer whose items cannot form a one-pass group: they are all
// batchable, share execution tile, program and blend, and only the first
// clears (framebuffers are independent, so this reorders nothing visible).
// The other items go to `fast`, in order.
int split_batch(sr_context ctx, const sr_batch_item* items, uint32_t n, std::vector<sr_batch_item>& fast) {
    int rc;
    std::vector<sr_framebuffer> seen;  // distinct framebuffers, in order of first appearance
    for (uint32_t i = 0; i < n; i++) {
        const sr_batch_item& it = items[i];
        if ((rc = validate_draw(ctx, it.fb, it.mesh, it.desc, it.tex))) return rc;
        if (std::find(seen.begin(), seen.end(), it.fb) == seen.end()) seen.push_back(it.fb);
    }
    for (uint32_t i = 0; i < n; i++) {
        const sr_texture t = items[i].tex;
        if (t && t->view)
            for (sr_framebuffer f : seen)
                if (f == t->fb)
                    return fail(ctx, SR_ERR_INVALID_ARGUMENT,
                                "batch: an item samples a framebuffer the batch draws into");
    }
    for (sr_framebuffer fb : seen) {
        std::vector<uint32_t> mine;
        for (uint32_t i = 0; i < n; i++)
            if (items[i].fb == fb) mine.push_back(i);
        bool group = true;
        int ets0 = -1, ew0 = -1;
        for (size_t j = 0; j < mine.size() && group; j++) {
            const sr_batch_item& it = items[mine[j]];
            const sr_draw_desc* d = it.desc;
            const int ets = d->exec_tile ? d->exec_tile : policy_tile(ctx, fb->w, fb->h);
            const int ew = draw_waves(ctx, d, (uint32_t)ets, fb->w, fb->h);
            const sr_draw_desc* f = items[mine[0]].desc;
            group = batchable(it.mesh, d, it.tex) && fb->standard() && (j == 0 || !it.clear) &&
                    (j == 0 || (ets == ets0 && ew == ew0)) &&
                    d->fragment_program == f->fragment_program && d->blend == f->blend;
            if (j == 0) {
                ets0 = ets;
                ew0 = ew;
            }
        }
        for (uint32_t i : mine) {
            const sr_batch_item& it = items[i];
            if (group) {
                fast.push_back(it);
                continue;
            }
            if (it.clear && (rc = clear_rgba8(it.fb, it.clear_rgba))) return rc;
            if ((rc = sr_draw(ctx, it.fb, it.mesh, it.desc, it.tex))) return rc;
        }
    }
    return SR_OK;
}

// After a one-pass batch ran on W: its error flags, or an overflow
// (*again: it must run again with the grown capacity).
int finish_batch(sr_context ctx, BatchWork& W, const BatchPlan& plan, bool* again) {
    int rc;
    if ((rc = batch_result(ctx, W, plan, again))) return rc;
    ctx->last.strip_slivers = batch_strip_slivers(W) ? 1u : 0u;
    return SR_OK;
}

int batch_draw(sr_context ctx, const sr_batch_item* items, uint32_t n) {
    int rc;
    hipSetDevice(ctx->device);
    ctx->n_stage = 0;
    std::vector<sr_batch_item> fast;
    if ((rc = split_batch(ctx, items, n, fast))) return rc;
    if (fast.empty()) return SR_OK;
    if (!ctx->batch) {
        BatchWork* W = new (std::nothrow) BatchWork();
        if (!W) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: workspace");
        if ((rc = W->init(ctx, false))) {
            W->release();
            delete W;
            return rc;
        }
        ctx->batch = W;
    }
    BatchWork& W = *ctx->batch;
    BatchPlan plan;
    if ((rc = batch_plan(ctx, fast.data(), (uint32_t)fast.size(), plan))) return rc;
    for (int attempt = 0; attempt < 4; attempt++) {
        if ((rc = batch_enqueue(ctx, W, plan, true))) return rc;
        HIP_OK(ctx, hipEventSynchronize(W.ev1));
        float ms = 0.f;
        hipEventElapsedTime(&ms, W.ev0, W.ev1);
        ctx->stage_ms[0] = ms;
        ctx->n_stage = 1;
        ctx->last.bin_pairs = W.h_misc[1];
        ctx->last.triangles_in = plan.NT;
        bool again = false;
        if ((rc = finish_batch(ctx, W, plan, &again))) return rc;
        if (!again) return SR_OK;
    }
    return fail(ctx, SR_ERR_DEVICE, "batch: capacity did not converge");
}

// ---------------------------------------------------------------------------
// Asynchronous batches: the one-pass group queued on a slot's own stream,
// finished by sr_batch_wait.
// ---------------------------------------------------------------------------
constexpr size_t kAsyncSlots = 4;

// Runs the slot's batch to its end (overflow reruns, the one-by-one
// fallback) and records its status; the slot is free afterwards.
int async_complete(sr_context ctx, AsyncBatch& b) {
    int rc = SR_OK;
    for (int attempt = 0; attempt < 5; attempt++) {
        if (hipEventSynchronize(b.W.ev1) != hipSuccess) {
            rc = fail(ctx, SR_ERR_DEVICE, "batch: device failure");
            break;
        }
        bool again = false;
        if ((rc = finish_batch(ctx, b.W, b.plan, &again)) || !again) break;
        if (attempt == 4) {
            rc = fail(ctx, SR_ERR_DEVICE, "batch: capacity did not converge");
            break;
        }
        if ((rc = batch_enqueue(ctx, b.W, b.plan, false))) break;
    }
    b.busy = false;
    ctx->async_done[b.ticket] = rc;
    return rc;
}

void destroy_async(sr_context ctx) {
    for (AsyncBatch* b : ctx->async) {
        b->W.release();
        delete b;
    }
    ctx->async.clear();
    ctx->async_done.clear();
}

void destroy_batch(sr_context ctx) {
    destroy_async(ctx);
    if (!ctx->batch) return;
    ctx->batch->release();
    delete ctx->batch;
    ctx->batch = nullptr;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int sr_ctx_create(int device, sr_context* out) {
    if (!out) return SR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SR_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SR_ERR_INVALID_ARGUMENT;
    sr_context ctx = new (std::nothrow) sr_context_t();
    if (!ctx) return SR_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_misc, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return SR_ERR_DEVICE;
    }
    for (auto& e : ctx->ev) hipEventCreate(&e);
    if (hipEventCreateWithFlags(&ctx->ev_read, hipEventDisableTiming) != hipSuccess) {
        hipHostFree(ctx->h_misc);
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return SR_ERR_DEVICE;
    }
    if (ensure(ctx, ctx->misc, 64 * sizeof(uint32_t)) != SR_OK) {
        delete ctx;
        return SR_ERR_OUT_OF_MEMORY;
    }
    // the static workspace's address on this device (sr_draw_static reports
    // a failed lookup; every other entry point works without it)
    lookup_scratch(device);
    *out = ctx;
    return SR_OK;
}

int sr_ctx_destroy(sr_context ctx) {
    if (!ctx) return SR_ERR_INVALID_ARGUMENT;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->verts, &ctx->gverts, &ctx->recs, &ctx->counts, &ctx->offsets, &ctx->keys,
                      &ctx->vals, &ctx->keys2, &ctx->vals2, &ctx->tstart, &ctx->tend, &ctx->scan_tmp,
                      &ctx->hist, &ctx->misc, &ctx->strips, &ctx->strip_aux, &ctx->litems,
                      &ctx->plots, &ctx->lcounts, &ctx->loffsets, &ctx->pcounts, &ctx->poffsets};
    for (DevBuf* b : bufs)
        if (b->p) hipFree(b->p);
    for (auto& e : ctx->ev)
        if (e) hipEventDestroy(e);
    if (ctx->ev_read) hipEventDestroy(ctx->ev_read);
    destroy_batch(ctx);
    if (ctx->h_misc) hipHostFree(ctx->h_misc);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return SR_OK;
}

const char* sr_last_error(sr_context ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sr_finish(sr_context ctx) {
    if (!ctx) return SR_ERR_INVALID_ARGUMENT;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

void* sr_ctx_stream(sr_context ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int sr_ctx_stage_times(sr_context ctx, float* ms, int n_max, int* n_out) {
    if (!ctx || !ms || !n_out) return SR_ERR_INVALID_ARGUMENT;
    int n = std::min(n_max, ctx->n_stage);
    for (int i = 0; i < n; i++) ms[i] = ctx->stage_ms[i];
    *n_out = n;
    return SR_OK;
}

int sr_ctx_last_stats(sr_context ctx, sr_draw_stats* out) {
    if (!ctx || !out) return SR_ERR_INVALID_ARGUMENT;
    *out = ctx->last;
    return SR_OK;
}

int sr_mesh_upload(sr_context ctx, const sr_vertex_soa* v, uint32_t nv, const uint32_t* indices,
                   uint32_t ni, int32_t primitive, sr_mesh* out) {
    if (!ctx || !v || !out || !v->px || !v->py || !v->pz)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "mesh: null argument");
    if (primitive != SR_PRIM_POINTS && primitive != SR_PRIM_LINES && primitive != SR_PRIM_TRIANGLES)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "mesh: bad primitive");
    const uint32_t per = (uint32_t)primitive;  // 1, 2, 3 vertices
    const uint32_t n_idx = indices ? ni : nv;
    if (n_idx % per != 0)  // assert_eq!(mesh.indices.len() % T::num_vertices(), 0), pipeline/mod.rs:148
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "mesh: index count is not a multiple of the primitive size");
    if (indices)
        for (uint32_t i = 0; i < ni; i++)
            if (indices[i] >= nv) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "mesh: index out of range");
    hipSetDevice(ctx->device);
    sr_mesh m = new (std::nothrow) sr_mesh_t();
    if (!m) return SR_ERR_OUT_OF_MEMORY;
    m->ctx = ctx; m->nv = nv; m->ni = indices ? ni : 0; m->indexed = indices != nullptr; m->prim = primitive;
    const float* src[8] = {v->px, v->py, v->pz, v->nx, v->ny, v->nz, v->u, v->v};
    for (int c = 0; c < 8; c++) {
        if (!src[c] || nv == 0) continue;
        if (hipMalloc(&m->soa[c], sizeof(float) * nv) != hipSuccess ||
            hipMemcpyAsync(m->soa[c], src[c], sizeof(float) * nv, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
            sr_mesh_free(m);
            return fail(ctx, SR_ERR_OUT_OF_MEMORY, "mesh: upload failed");
        }
    }
    if (indices && ni) {
        if (hipMalloc(&m->idx, sizeof(uint32_t) * ni) != hipSuccess ||
            hipMemcpyAsync(m->idx, indices, sizeof(uint32_t) * ni, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
            sr_mesh_free(m);
            return fail(ctx, SR_ERR_OUT_OF_MEMORY, "mesh: index upload failed");
        }
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        sr_mesh_free(m);
        return fail(ctx, SR_ERR_DEVICE, "mesh: sync failed");
    }
    *out = m;
    return SR_OK;
}

int sr_mesh_free(sr_mesh m) {
    if (!m) return SR_ERR_INVALID_ARGUMENT;
    hipStreamSynchronize(m->ctx->stream);
    for (float* p : m->soa)
        if (p) hipFree(p);
    if (m->idx) hipFree(m->idx);
    if (m->iattr) hipFree(m->iattr);
    delete m;
    return SR_OK;
}

int sr_mesh_set_int_attrib(sr_mesh m, const int32_t* values, uint32_t count) {
    if (!m || (!values && count)) return fail(m ? m->ctx : nullptr, SR_ERR_INVALID_ARGUMENT, "mesh: null argument");
    sr_context ctx = m->ctx;
    if (count != m->nv) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "mesh: one integer attribute per vertex");
    hipSetDevice(ctx->device);
    if (!m->iattr && count && hipMalloc(&m->iattr, sizeof(int32_t) * count) != hipSuccess) {
        m->iattr = nullptr;
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "mesh: attribute upload failed");
    }
    if (count && (hipMemcpyAsync(m->iattr, values, sizeof(int32_t) * count, hipMemcpyHostToDevice, ctx->stream) !=
                      hipSuccess ||
                  hipStreamSynchronize(ctx->stream) != hipSuccess))
        return fail(ctx, SR_ERR_DEVICE, "mesh: attribute upload failed");
    return SR_OK;
}

int sr_texture_upload(sr_context ctx, uint32_t w, uint32_t h, const uint32_t* rgba, sr_texture* out) {
    if (!ctx || !rgba || !out || !w || !h) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "texture: bad argument");
    sr_texture t = new (std::nothrow) sr_texture_t();
    if (!t) return SR_ERR_OUT_OF_MEMORY;
    t->ctx = ctx; t->w = w; t->h = h;
    if (hipMalloc(&t->d, sizeof(uint32_t) * (size_t)w * h) != hipSuccess) {
        delete t;
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "texture: alloc failed");
    }
    if (hipMemcpy(t->d, rgba, sizeof(uint32_t) * (size_t)w * h, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(t->d);
        delete t;
        return fail(ctx, SR_ERR_DEVICE, "texture: upload failed");
    }
    *out = t;
    return SR_OK;
}

int sr_texture_free(sr_texture t) {
    if (!t) return SR_ERR_INVALID_ARGUMENT;
    hipStreamSynchronize(t->ctx->stream);
    if (!t->view) hipFree(t->d);
    delete t;
    return SR_OK;
}

int sr_texture_from_fb(sr_framebuffer fb, sr_texture* out) {
    if (!fb || !out) return SR_ERR_INVALID_ARGUMENT;
    if (fb->fmt.color[0] != SR_COLOR_RGBA8) return fail(fb->ctx, SR_ERR_UNSUPPORTED, "texture_from_fb: use sr_texture_from_fb_plane");
    sr_texture t = new (std::nothrow) sr_texture_t();
    if (!t) return SR_ERR_OUT_OF_MEMORY;
    t->ctx = fb->ctx; t->w = fb->w; t->h = fb->h;
    t->d = fb->color;  // RGBA8 row-major: the texture layout already
    t->view = true;
    t->fb = fb;
    *out = t;
    return SR_OK;
}

int sr_fb_create(sr_context ctx, uint32_t w, uint32_t h, sr_framebuffer* out) {
    if (!ctx || !out) return SR_ERR_INVALID_ARGUMENT;
    // with_framebuffer asserts, pipeline/mod.rs:129-130
    if (w == 0 || h == 0) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "Framebuffer must have a non-zero width and height");
    sr_framebuffer fb = new (std::nothrow) sr_framebuffer_t();
    if (!fb) return SR_ERR_OUT_OF_MEMORY;
    fb->ctx = ctx; fb->w = w; fb->h = h;
    const size_t n = (size_t)w * h;
    if (hipMalloc(&fb->color, 4 * n) != hipSuccess || hipMalloc(&fb->depth, 4 * n) != hipSuccess ||
        hipMalloc(&fb->stencil, n) != hipSuccess) {
        sr_fb_free(fb);
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "framebuffer: alloc failed");
    }
    // RenderBuffer::with_dimensions: {Color::empty(), Depth::far(), 0}, renderbuffer/mod.rs:22-30,58-63
    launch_fb_clear(fb->color, fb->depth, fb->stencil, (uint32_t)n, 0u, -3.40282347e+38f, 0u, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        sr_fb_free(fb);
        return fail(ctx, SR_ERR_DEVICE, "framebuffer: clear failed");
    }
    *out = fb;
    return SR_OK;
}

static size_t plane_bytes(int fmt) { return fmt == SR_COLOR_RGBAF32 ? 16 : 4; }

// The functions that take one RGBA8 colour and f32 depth need plane 0 in
// RGBA8 and f32 depth.
static bool rgba8_depth32(sr_framebuffer fb) {
    return fb->fmt.color[0] == SR_COLOR_RGBA8 && fb->fmt.depth == SR_DEPTH_F32;
}

int sr_fb_create_ex(sr_context ctx, uint32_t w, uint32_t h, const sr_fb_format* f, sr_framebuffer* out) {
    if (!ctx || !out || !f) return SR_ERR_INVALID_ARGUMENT;
    if (w == 0 || h == 0) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "Framebuffer must have a non-zero width and height");
    if (f->n_color < 1 || f->n_color > SR_MAX_COLOR_PLANES) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "framebuffer: 1..4 colour planes");
    for (uint32_t p = 0; p < f->n_color; p++)
        if (f->color[p] != SR_COLOR_RGBA8 && f->color[p] != SR_COLOR_RGBAF32)
            return fail(ctx, SR_ERR_INVALID_ARGUMENT, "framebuffer: unknown colour format");
    if ((f->depth != SR_DEPTH_F32 && f->depth != SR_DEPTH_F64) || (f->stencil != SR_STENCIL_U8 && f->stencil != SR_STENCIL_NONE))
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "framebuffer: unknown depth or stencil format");
    sr_framebuffer fb = new (std::nothrow) sr_framebuffer_t();
    if (!fb) return SR_ERR_OUT_OF_MEMORY;
    fb->ctx = ctx; fb->w = w; fb->h = h;
    fb->fmt = *f;
    for (uint32_t p = f->n_color; p < SR_MAX_COLOR_PLANES; p++) fb->fmt.color[p] = 0;
    const size_t n = (size_t)w * h;
    bool ok = true;
    for (uint32_t p = 0; p < f->n_color && ok; p++) ok = hipMalloc(&fb->planes[p], plane_bytes(f->color[p]) * n) == hipSuccess;
    if (ok && f->depth == SR_DEPTH_F64) ok = hipMalloc(&fb->depth64, 8 * n) == hipSuccess;
    if (ok && f->depth == SR_DEPTH_F32) ok = hipMalloc(&fb->depth, 4 * n) == hipSuccess;
    if (ok && f->stencil == SR_STENCIL_U8) ok = hipMalloc(&fb->stencil, n) == hipSuccess;
    if (!ok) {
        sr_fb_free(fb);
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "framebuffer: alloc failed");
    }
    if (f->color[0] == SR_COLOR_RGBA8) fb->color = (uint32_t*)fb->planes[0];
    // with_dimensions: every plane Color::empty() (zero), depth far, stencil 0
    // (texturebuffer.rs:102-111)
    uint8_t zeros[16 * SR_MAX_COLOR_PLANES] = {};
    int rc = sr_fb_clear_ex(fb, zeros);
    if (rc == SR_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = SR_ERR_DEVICE;
    if (rc != SR_OK) {
        sr_fb_free(fb);
        return fail(ctx, rc, "framebuffer: clear failed");
    }
    *out = fb;
    return SR_OK;
}

int sr_fb_get_format(sr_framebuffer fb, sr_fb_format* out) {
    if (!fb || !out) return SR_ERR_INVALID_ARGUMENT;
    *out = fb->fmt;
    return SR_OK;
}

int sr_fb_clear_ex(sr_framebuffer fb, const void* colors) {
    if (!fb || !colors) return SR_ERR_INVALID_ARGUMENT;
    const size_t n = (size_t)fb->w * fb->h;
    hipStream_t s = fb->ctx->stream;
    const uint8_t* c = (const uint8_t*)colors;
    for (uint32_t p = 0; p < fb->fmt.n_color; p++) {
        uint32_t pat[4] = {};
        const size_t b = plane_bytes(fb->fmt.color[p]);
        memcpy(pat, c, b);
        c += b;
        launch_fill_pattern(fb->planes[p], pat, (uint32_t)(b / 4), n * (b / 4), s);
    }
    if (fb->depth64) {
        const double far = -1.7976931348623157e308;  // f64::MIN, depth.rs:31
        uint32_t pat[4];
        memcpy(pat, &far, 8);
        launch_fill_pattern(fb->depth64, pat, 2, 2 * n, s);
    }
    if (fb->depth) {
        const float far = -3.40282347e+38f;  // f32::MIN
        uint32_t pat[4];
        memcpy(pat, &far, 4);
        launch_fill_pattern(fb->depth, pat, 1, n, s);
    }
    if (fb->stencil) HIP_OK(fb->ctx, hipMemsetAsync(fb->stencil, 0, n, s));
    return SR_OK;
}

int sr_fb_read_ex(sr_framebuffer fb, void* const* planes, void* depth, uint8_t* stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    sr_context ctx = fb->ctx;
    const size_t n = (size_t)fb->w * fb->h;
    for (uint32_t p = 0; planes && p < fb->fmt.n_color; p++)
        if (planes[p])
            HIP_OK(ctx, hipMemcpyAsync(planes[p], fb->planes[p], plane_bytes(fb->fmt.color[p]) * n,
                                       hipMemcpyDeviceToHost, ctx->stream));
    if (depth)
        HIP_OK(ctx, hipMemcpyAsync(depth, fb->depth64 ? (void*)fb->depth64 : (void*)fb->depth,
                                   (fb->depth64 ? 8 : 4) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (stencil && fb->stencil) HIP_OK(ctx, hipMemcpyAsync(stencil, fb->stencil, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_fb_write_ex(sr_framebuffer fb, const void* const* planes, const void* depth, const uint8_t* stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    sr_context ctx = fb->ctx;
    const size_t n = (size_t)fb->w * fb->h;
    for (uint32_t p = 0; planes && p < fb->fmt.n_color; p++)
        if (planes[p])
            HIP_OK(ctx, hipMemcpyAsync(fb->planes[p], planes[p], plane_bytes(fb->fmt.color[p]) * n,
                                       hipMemcpyHostToDevice, ctx->stream));
    if (depth)
        HIP_OK(ctx, hipMemcpyAsync(fb->depth64 ? (void*)fb->depth64 : (void*)fb->depth, depth,
                                   (fb->depth64 ? 8 : 4) * n, hipMemcpyHostToDevice, ctx->stream));
    if (stencil && fb->stencil) HIP_OK(ctx, hipMemcpyAsync(fb->stencil, stencil, n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_texture_from_fb_plane(sr_framebuffer fb, uint32_t plane, sr_texture* out) {
    if (!fb || !out || plane >= fb->fmt.n_color) return SR_ERR_INVALID_ARGUMENT;
    sr_texture t = new (std::nothrow) sr_texture_t();
    if (!t) return SR_ERR_OUT_OF_MEMORY;
    t->ctx = fb->ctx; t->w = fb->w; t->h = fb->h;
    t->d = (uint32_t*)fb->planes[plane];  // row-major texels of the plane's format: the texture layout
    t->fmt = fb->fmt.color[plane];
    t->view = true;
    t->fb = fb;
    *out = t;
    return SR_OK;
}

int sr_fb_free(sr_framebuffer fb) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    hipStreamSynchronize(fb->ctx->stream);
    bool color_in_planes = false;
    for (void* p : fb->planes)
        if (p) {
            color_in_planes |= p == (void*)fb->color;
            hipFree(p);
        }
    if (fb->color && !color_in_planes) hipFree(fb->color);
    if (fb->depth) hipFree(fb->depth);
    if (fb->depth64) hipFree(fb->depth64);
    if (fb->stencil) hipFree(fb->stencil);
    delete fb;
    return SR_OK;
}

int sr_fb_clear(sr_framebuffer fb, uint32_t rgba, float depth, uint8_t stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    if (!fb->standard()) return fail(fb->ctx, SR_ERR_UNSUPPORTED, "fb_clear: use sr_fb_clear_ex for this format");
    launch_fb_clear(fb->color, fb->depth, fb->stencil, fb->w * fb->h, rgba, depth, stencil, fb->ctx->stream);
    return SR_OK;
}

int sr_fb_read(sr_framebuffer fb, uint32_t* rgba, float* depth, uint8_t* stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    if (!rgba8_depth32(fb)) return fail(fb->ctx, SR_ERR_UNSUPPORTED, "fb_read: use sr_fb_read_ex for this format");
    sr_context ctx = fb->ctx;
    const size_t n = (size_t)fb->w * fb->h;
    if (rgba) HIP_OK(ctx, hipMemcpyAsync(rgba, fb->color, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    if (depth) HIP_OK(ctx, hipMemcpyAsync(depth, fb->depth, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    if (stencil && fb->stencil) HIP_OK(ctx, hipMemcpyAsync(stencil, fb->stencil, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_fb_write(sr_framebuffer fb, const uint32_t* rgba, const float* depth, const uint8_t* stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    if (!rgba8_depth32(fb)) return fail(fb->ctx, SR_ERR_UNSUPPORTED, "fb_write: use sr_fb_write_ex for this format");
    sr_context ctx = fb->ctx;
    const size_t n = (size_t)fb->w * fb->h;
    if (rgba) HIP_OK(ctx, hipMemcpyAsync(fb->color, rgba, 4 * n, hipMemcpyHostToDevice, ctx->stream));
    if (depth) HIP_OK(ctx, hipMemcpyAsync(fb->depth, depth, 4 * n, hipMemcpyHostToDevice, ctx->stream));
    if (stencil && fb->stencil) HIP_OK(ctx, hipMemcpyAsync(fb->stencil, stencil, n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_fb_size(sr_framebuffer fb, uint32_t* w, uint32_t* h) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    if (w) *w = fb->w;
    if (h) *h = fb->h;
    return SR_OK;
}

// One draw chain.  `clip` (optional) holds the mesh's vertices already run
// through the chain's vertex program, in clip space: the vertex stage is
// skipped (a chain that shares them with others, GeometryShader::duplicate).
static int draw_one(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex,
                    const VertexBuf* clip) {
    memset(&ctx->last, 0, sizeof(ctx->last));
    ctx->n_stage = 0;

    DrawPlan pl;
    pl.W = fb->w; pl.H = fb->h;
    pl.tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u;
    pl.th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
    pl.ets = d->exec_tile ? (uint32_t)d->exec_tile : (uint32_t)policy_tile(ctx, fb->w, fb->h);
    pl.nwv = draw_waves(ctx, d, pl.ets, fb->w, fb->h);
    pl.ntx = (pl.W + pl.ets - 1) / pl.ets;
    pl.nty = (pl.H + pl.ets - 1) / pl.ets;
    pl.ntiles = pl.ntx * pl.nty;
    pl.K = varyings_of(d->vertex_program);
    pl.timing = (d->exec_flags & SR_EXEC_TIMING) != 0;
    // W == 1 or H == 1: the tile loop never runs (fragment.rs:196,201)
    pl.scols = pl.W > 1 ? (pl.W - 1 + pl.tw - 1) / pl.tw : 0;
    pl.srows = pl.H > 1 ? (pl.H - 1 + pl.th - 1) / pl.th : 0;
    const bool value_test = d->stencil_test != SR_STENCIL_ALWAYS && d->stencil_test != SR_STENCIL_NEVER;
    if (d->stencil_test == SR_STENCIL_ALWAYS) pl.smode = d->stencil_op == SR_STENCIL_KEEP ? 0 : 1;
    else pl.smode = 2;
    pl.exhaustive = (d->exec_flags & SR_EXEC_EXHAUSTIVE) != 0 || (value_test && d->stencil_op != SR_STENCIL_KEEP);
    const bool no_stencil = fb->fmt.stencil == SR_STENCIL_NONE;  // `()`: no test, no op (stencil.rs:65-86)
    const bool normals_gs = d->geometry_program == SR_GS_VERTEX_NORMALS || d->geometry_program == SR_GS_FACE_NORMALS;
    // every format, the normalised blend, the multi-output program and a
    // triangle draw sampling an RGBAF32 texture take the per-pixel raster
    pl.general = !fb->standard() || d->blend == SR_BLEND_OVER_NORMALIZED || d->fragment_program == SR_FP_GBUFFER ||
                 d->fragment_program == SR_FP_FRESNEL4 || d->fragment_program == SR_FP_INT_PALETTE ||
                 (samples_f32(d, tex) && mesh->prim == SR_PRIM_TRIANGLES && !normals_gs);
    if (pl.general) {
        if (mesh->prim != SR_PRIM_TRIANGLES || normals_gs)
            return fail(ctx, SR_ERR_UNSUPPORTED, "draw: lines and points need sr_fb_create's format, RGBA8 programs and blends");
        // 8x8 execution tiles; a stencil with a test or op runs exhaustive
        pl.ets = 8;
        pl.nwv = 1;
        pl.ntx = (pl.W + 7) / 8;
        pl.nty = (pl.H + 7) / 8;
        pl.ntiles = pl.ntx * pl.nty;
        if (no_stencil) pl.smode = 0;
        else if (pl.smode != 0) pl.exhaustive = true;
    }

    const uint32_t per = (uint32_t)mesh->prim;
    const uint32_t nidx = mesh_count(mesh);
    const uint32_t nprim = nidx / per;
    ctx->last.triangles_in = mesh->prim == SR_PRIM_TRIANGLES ? nprim : 0;
    ctx->last.lines_in = mesh->prim == SR_PRIM_LINES ? nprim : 0;
    ctx->last.points_in = mesh->prim == SR_PRIM_POINTS ? nprim : 0;
    // StencilTest::Never: no pixel passes, nothing is written (stencil.rs:115)
    if (pl.scols == 0 || pl.srows == 0 || (d->stencil_test == SR_STENCIL_NEVER && !no_stencil) || nprim == 0)
        return SR_OK;

    int rc;
    hipStream_t s = ctx->stream;
    if (pl.timing) hipEventRecord(ctx->ev[0], s);
    // clear the error word, the pair count and the overflow flag
    launch_fill(P<uint32_t>(ctx->misc), 0u, 6u, s);

    // ---- vertex stage
    const uint32_t nv = mesh->nv;
    if (!clip &&
        (rc = ensure(ctx, ctx->verts, sizeof(float) * (size_t)(4 + pl.K) * std::max<uint32_t>(nv, 1))))
        return rc;
    VertexBuf vb = clip ? *clip : VertexBuf{P<float>(ctx->verts), std::max<uint32_t>(nv, 1), (uint32_t)pl.K};
    VSParams vs;
    memset(&vs, 0, sizeof(vs));
    vs.px = mesh->soa[0]; vs.py = mesh->soa[1]; vs.pz = mesh->soa[2];
    vs.nx = mesh->soa[3]; vs.ny = mesh->soa[4]; vs.nz = mesh->soa[5];
    vs.u = mesh->soa[6]; vs.v = mesh->soa[7]; vs.ia = mesh->iattr;
    vs.n = nv; vs.program = d->vertex_program; vs.K = pl.K;
    // normal visualisation re-emits lines and points unchanged (shaders.rs:59,88):
    // for them it is run_to_fragment
    const bool normals = d->geometry_program == SR_GS_VERTEX_NORMALS || d->geometry_program == SR_GS_FACE_NORMALS;
    const int gs = normals && mesh->prim != SR_PRIM_TRIANGLES ? SR_GS_NONE : d->geometry_program;
    vs.normalize = gs == SR_GS_NONE;
    vs.viewport = d->viewport;
    memcpy(vs.U, d->vs_uniforms, sizeof(vs.U));
    vs.out = vb;
    if (!clip) {
        launch_vertex(vs, s);
    } else if (gs == SR_GS_NONE) {
        // finish the shared clip-space vertices into this chain's own buffer
        if ((rc = ensure(ctx, ctx->gverts, sizeof(float) * (size_t)(4 + pl.K) * std::max<uint32_t>(nv, 1))))
            return rc;
        const VertexBuf sv{P<float>(ctx->gverts), std::max<uint32_t>(nv, 1), (uint32_t)pl.K};
        launch_finish(*clip, sv, nv, pl.K, d->viewport, s);
        vb = sv;
    }
    mark(ctx, pl.timing, ST_VERTEX);

    const uint32_t* idx = mesh->idx;
    uint32_t nwork = nprim;
    int out_prim = mesh->prim;
    if (normals && gs != SR_GS_NONE) {
        // ---- geometry stage: lines at fixed offsets, no count pass
        const uint32_t lines_per = d->geometry_program == SR_GS_VERTEX_NORMALS ? 3u : 1u;
        const uint64_t nl = (uint64_t)nprim * lines_per;
        if (nl > 0x7fffffffull) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: too many normal lines");
        if ((rc = ensure(ctx, ctx->gverts, sizeof(float) * (size_t)(4 + pl.K) * 2 * nl))) return rc;
        GSParams gp;
        memset(&gp, 0, sizeof(gp));
        gp.in = vb; gp.idx = idx; gp.n_in = nprim; gp.K = pl.K;
        gp.viewport = d->viewport;
        memcpy(gp.G, d->gs_uniforms, sizeof(gp.G));
        gp.out = VertexBuf{P<float>(ctx->gverts), (uint32_t)(2 * nl), (uint32_t)pl.K};
        launch_gs_normals(d->geometry_program, gp, s);
        vb = gp.out;
        idx = nullptr;
        nwork = (uint32_t)nl;
        out_prim = SR_PRIM_LINES;
        ctx->last.generated_lines = nwork;
    } else if (gs != SR_GS_NONE) {
        // ---- geometry stage: count, scan, emit (fused with finish)
        if ((rc = ensure(ctx, ctx->counts, sizeof(uint32_t) * (nprim + 1))) ||
            (rc = ensure(ctx, ctx->offsets, sizeof(uint32_t) * (nprim + 1))))
            return rc;
        GSParams gp;
        memset(&gp, 0, sizeof(gp));
        gp.in = vb; gp.idx = idx; gp.n_in = nprim; gp.K = pl.K;
        gp.counts = P<uint32_t>(ctx->counts); gp.offsets = P<uint32_t>(ctx->offsets);
        gp.viewport = d->viewport;
        launch_gs_count(d->geometry_program, mesh->prim, gp, s);
        if ((rc = scan_counts(ctx, gp.counts, P<uint32_t>(ctx->offsets), nprim))) return rc;
        if ((rc = read_words(ctx, 2))) return rc;
        const uint32_t ngen = ctx->h_misc[1];
        if ((rc = ensure(ctx, ctx->gverts, sizeof(float) * (size_t)(4 + pl.K) * std::max<uint32_t>(ngen * per, 1))))
            return rc;
        gp.out = VertexBuf{P<float>(ctx->gverts), std::max<uint32_t>(ngen * per, 1), (uint32_t)pl.K};
        launch_gs_emit(d->geometry_program, mesh->prim, gp, s);
        vb = gp.out;
        idx = nullptr;
        nwork = ngen;
        if (mesh->prim == SR_PRIM_TRIANGLES) ctx->last.generated_tris = ngen;
        else if (mesh->prim == SR_PRIM_LINES) ctx->last.generated_lines = ngen;
        else ctx->last.generated_points = ngen;
    }
    mark(ctx, pl.timing, ST_GEOMETRY);

    if (nwork == 0) {
        rc = SR_OK;
    } else if (out_prim == SR_PRIM_TRIANGLES) {
        rc = draw_tris(ctx, fb, pl, d, tex, vb, idx, nwork);
    } else if (out_prim == SR_PRIM_LINES) {
        rc = draw_lines_impl(ctx, fb, pl, d, tex, vb, idx, nwork);
    } else {
        rc = draw_points_impl(ctx, fb, pl, d, tex, vb, idx, nwork);
    }
    if (rc) return rc;
    HIP_OK(ctx, hipGetLastError());
    if (pl.timing) {
        HIP_OK(ctx, hipStreamSynchronize(s));
        for (int st = 0; st < ST_COUNT; st++) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ctx->ev[st], ctx->ev[st + 1]);
            ctx->stage_ms[st] = ms;
        }
        ctx->n_stage = ST_COUNT;
    }
    return SR_OK;
}

int sr_draw(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex) {
    if (int rc = validate_draw(ctx, fb, mesh, d, tex)) return rc;
    hipSetDevice(ctx->device);
    return draw_one(ctx, fb, mesh, d, tex, nullptr);
}

}  // extern "C"

// sr_draw on a framebuffer of `ctx` with a mesh and texture that another
// context on the same device owns (the request server draws its clients'
// meshes on its own context).  The draw was validated when it was queued.
int sr_rt::draw_foreign(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex) {
    hipSetDevice(ctx->device);
    return draw_one(ctx, fb, mesh, d, tex, nullptr);
}

extern "C" {

int sr_draw_static(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex) {
    if (int rc = validate_draw(ctx, fb, mesh, d, tex)) return rc;
    hipSetDevice(ctx->device);
    uint8_t* const base = scratch_of(ctx->device);
    if (!base) return fail(ctx, SR_ERR_DEVICE, "draw_static: no static workspace on this device");
    // The draw runs exactly as sr_draw does, with every workspace buffer
    // carved from the static region (ensure() allocates nothing while
    // ctx->arena is set); the context's own buffers are set aside meanwhile
    // and restored untouched.  misc (the error and count words) was
    // allocated with the context.
    DevBuf* const bufs[] = {&ctx->verts, &ctx->gverts, &ctx->recs, &ctx->counts, &ctx->offsets, &ctx->keys,
                            &ctx->vals, &ctx->keys2, &ctx->vals2, &ctx->tstart, &ctx->tend, &ctx->scan_tmp,
                            &ctx->hist, &ctx->strips, &ctx->strip_aux, &ctx->litems, &ctx->plots,
                            &ctx->lcounts, &ctx->loffsets, &ctx->pcounts, &ctx->poffsets};
    constexpr size_t NB = sizeof(bufs) / sizeof(bufs[0]);
    DevBuf saved[NB];
    for (size_t i = 0; i < NB; i++) {
        saved[i] = *bufs[i];
        *bufs[i] = DevBuf{};
    }
    ctx->arena = base;
    ctx->arena_size = kStaticWorkspaceBytes;
    ctx->arena_used = 0;
    const int rc = draw_one(ctx, fb, mesh, d, tex, nullptr);
    ctx->arena = nullptr;
    ctx->arena_size = ctx->arena_used = 0;
    for (size_t i = 0; i < NB; i++) *bufs[i] = saved[i];
    return rc;
}

int sr_draw_chains(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* descs, uint32_t n,
                   sr_texture tex) {
    if (!ctx || (!descs && n)) return SR_ERR_INVALID_ARGUMENT;
    if (n == 0) return SR_OK;
    for (uint32_t i = 0; i < n; i++) {
        if (int rc = validate_draw(ctx, fb, mesh, &descs[i], tex)) return rc;
        if (descs[i].vertex_program != descs[0].vertex_program ||
            memcmp(descs[i].vs_uniforms, descs[0].vs_uniforms, sizeof(descs[0].vs_uniforms)) != 0)
            return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw_chains: every chain must share the vertex program");
    }
    hipSetDevice(ctx->device);
    if (n == 1) return draw_one(ctx, fb, mesh, &descs[0], tex, nullptr);
    // the vertex stage once, in clip space (VertexShader::run, vertex.rs:87-120)
    const int K = varyings_of(descs[0].vertex_program);
    const uint32_t nv = mesh->nv;
    int rc;
    if ((rc = ensure(ctx, ctx->verts, sizeof(float) * (size_t)(4 + K) * std::max<uint32_t>(nv, 1)))) return rc;
    const VertexBuf clip{P<float>(ctx->verts), std::max<uint32_t>(nv, 1), (uint32_t)K};
    VSParams vs;
    memset(&vs, 0, sizeof(vs));
    vs.px = mesh->soa[0]; vs.py = mesh->soa[1]; vs.pz = mesh->soa[2];
    vs.nx = mesh->soa[3]; vs.ny = mesh->soa[4]; vs.nz = mesh->soa[5];
    vs.u = mesh->soa[6]; vs.v = mesh->soa[7]; vs.ia = mesh->iattr;
    vs.n = nv; vs.program = descs[0].vertex_program; vs.K = K;
    vs.normalize = 0;
    vs.viewport = descs[0].viewport;
    memcpy(vs.U, descs[0].vs_uniforms, sizeof(vs.U));
    vs.out = clip;
    launch_vertex(vs, ctx->stream);
    for (uint32_t i = 0; i < n; i++)
        if ((rc = draw_one(ctx, fb, mesh, &descs[i], tex, &clip))) return rc;
    return SR_OK;
}

int sr_batch_draw(sr_context ctx, const sr_batch_item* items, uint32_t n) {
    if (!ctx || (!items && n)) return SR_ERR_INVALID_ARGUMENT;
    return batch_draw(ctx, items, n);
}

int sr_batch_submit(sr_context ctx, const sr_batch_item* items, uint32_t n, uint64_t* ticket) {
    if (!ctx || (!items && n) || !ticket) return SR_ERR_INVALID_ARGUMENT;
    int rc;
    hipSetDevice(ctx->device);
    std::vector<sr_batch_item> fast;
    if ((rc = split_batch(ctx, items, n, fast))) return rc;
    const uint64_t t = ctx->next_ticket++;
    if (fast.empty()) {
        ctx->async_done[t] = SR_OK;
        *ticket = t;
        return SR_OK;
    }
    // a free slot, else the oldest one completes first
    AsyncBatch* slot = nullptr;
    for (AsyncBatch* b : ctx->async)
        if (!b->busy) {
            slot = b;
            break;
        }
    if (!slot && ctx->async.size() < kAsyncSlots) {
        slot = new (std::nothrow) AsyncBatch();
        if (!slot) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: slot");
        if ((rc = slot->W.init(ctx, true))) {
            slot->W.release();
            delete slot;
            return rc;
        }
        ctx->async.push_back(slot);
    }
    if (!slot) {
        slot = ctx->async[0];
        for (AsyncBatch* b : ctx->async)
            if (b->ticket < slot->ticket) slot = b;
        async_complete(ctx, *slot);  // its status waits in async_done for its own sr_batch_wait
    }
    slot->plan = BatchPlan();
    if ((rc = batch_plan(ctx, fast.data(), (uint32_t)fast.size(), slot->plan))) return rc;
    // work already queued on the context stream (the caller's clears and
    // uploads, the items drawn one by one above) comes first
    HIP_OK(ctx, hipEventRecord(ctx->ev_read, ctx->stream));
    HIP_OK(ctx, hipStreamWaitEvent(slot->W.stream, ctx->ev_read, 0));
    if ((rc = batch_enqueue(ctx, slot->W, slot->plan, false))) return rc;
    slot->ticket = t;
    slot->busy = true;
    *ticket = t;
    return SR_OK;
}

int sr_batch_query(sr_context ctx, uint64_t ticket) {
    if (!ctx) return SR_ERR_INVALID_ARGUMENT;
    for (AsyncBatch* b : ctx->async)
        if (b->busy && b->ticket == ticket)
            return hipEventQuery(b->W.ev1) == hipErrorNotReady ? SR_ERR_NOT_READY : SR_OK;
    return ctx->async_done.count(ticket) ? SR_OK : fail(ctx, SR_ERR_INVALID_ARGUMENT, "batch: unknown ticket");
}

int sr_batch_wait(sr_context ctx, uint64_t ticket) {
    if (!ctx) return SR_ERR_INVALID_ARGUMENT;
    hipSetDevice(ctx->device);
    for (AsyncBatch* b : ctx->async)
        if (b->busy && b->ticket == ticket) async_complete(ctx, *b);
    auto it = ctx->async_done.find(ticket);
    if (it == ctx->async_done.end()) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "batch: unknown ticket");
    const int rc = it->second;
    ctx->async_done.erase(it);
    return rc;
}

int sr_ctx_set_exec_tile_policy(sr_context ctx, uint32_t n, const uint64_t* max_pixels,
                                const int32_t* exec_tile) {
    return sr_ctx_set_exec_policy(ctx, n, max_pixels, exec_tile, nullptr);
}

int sr_ctx_set_exec_policy(sr_context ctx, uint32_t n, const uint64_t* max_pixels, const int32_t* exec_tile,
                           const int32_t* exec_waves) {
    if (!ctx || !n || !max_pixels || !exec_tile) return SR_ERR_INVALID_ARGUMENT;
    std::vector<std::pair<uint64_t, int>> p;
    for (uint32_t i = 0; i < n; i++) {
        const int e = exec_tile[i], w = exec_waves ? exec_waves[i] : 0;
        if (e != 8 && e != 16 && e != 32)
            return fail(ctx, SR_ERR_INVALID_ARGUMENT, "policy: exec tile must be 8, 16 or 32");
        if (w != 0 && w != 1 && w != 2 && w != 4)
            return fail(ctx, SR_ERR_INVALID_ARGUMENT, "policy: waves must be 0, 1, 2 or 4");
        p.emplace_back(max_pixels[i], e | (w << 8));
    }
    ctx->tile_policy = p;
    if (ctx->batch) ctx->batch->drop_graph();
    return SR_OK;
}

int sr_ctx_set_batch_capacity(sr_context ctx, uint32_t pairs) {
    if (!ctx) return SR_ERR_INVALID_ARGUMENT;
    ctx->batch_cap_hint = std::max(ctx->batch_cap_hint, pairs);
    if (ctx->batch && pairs > ctx->batch->cap) {
        ctx->batch->cap = pairs;
        ctx->batch->drop_graph();
    }
    return SR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Lines: (line, screen tile) items -> plot records -> binned by execution
// tile -> raster.  Points: one record each -> binned -> raster.  Both use
// 16x16 execution tiles.
// ---------------------------------------------------------------------------
static constexpr uint32_t LP_ETS = 16;

int draw_lines_impl(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, const sr_draw_desc* d,
                    sr_texture tex, VertexBuf vb, const uint32_t* idx, uint32_t nlines) {
    int rc;
    hipStream_t s = ctx->stream;
    const uint32_t ntx = (pl.W + LP_ETS - 1) / LP_ETS, nty = (pl.H + LP_ETS - 1) / LP_ETS;
    if ((rc = ensure(ctx, ctx->counts, sizeof(uint32_t) * (nlines + 1))) ||
        (rc = ensure(ctx, ctx->offsets, sizeof(uint32_t) * (nlines + 1))))
        return rc;
    LineSetupParams lp;
    memset(&lp, 0, sizeof(lp));
    lp.v = vb; lp.idx = idx; lp.n_lines = nlines; lp.W = pl.W; lp.H = pl.H;
    lp.tw = pl.tw; lp.th = pl.th; lp.scols = pl.scols; lp.srows = pl.srows;
    lp.counts = P<uint32_t>(ctx->counts); lp.offsets = P<uint32_t>(ctx->offsets);
    lp.err = P<uint32_t>(ctx->misc);
    launch_line_tiles(lp, 0, s);
    mark(ctx, pl.timing, ST_SETUP);
    if ((rc = scan_counts(ctx, lp.counts, P<uint32_t>(ctx->offsets), nlines, 1))) return rc;
    if ((rc = read_words(ctx, 2))) return rc;
    if (ctx->h_misc[0]) return fail(ctx, SR_ERR_INVALID_COORDINATE, "non-finite line coordinate");
    const uint32_t nitems = ctx->h_misc[1];
    if (nitems == 0) return SR_OK;
    if ((rc = ensure(ctx, ctx->litems, sizeof(LineItem) * nitems)) ||
        (rc = ensure(ctx, ctx->lcounts, sizeof(uint32_t) * (nitems + 1))) ||
        (rc = ensure(ctx, ctx->loffsets, sizeof(uint32_t) * (nitems + 1))) ||
        (rc = ensure(ctx, ctx->pcounts, sizeof(uint32_t) * (nitems + 1))) ||
        (rc = ensure(ctx, ctx->poffsets, sizeof(uint32_t) * (nitems + 1))))
        return rc;
    lp.items = P<LineItem>(ctx->litems);
    // items ordered by walk length, longest first (one 8-bit radix pass)
    if ((rc = ensure_pairs(ctx, nitems)) ||
        (rc = ensure(ctx, ctx->hist, sizeof(uint32_t) * (radix_hist_words(nitems) + 256))) ||
        (rc = ensure(ctx, ctx->scan_tmp, sizeof(uint32_t) * (scan_tmp_words(radix_hist_words(nitems) + 256) + 16))))
        return rc;
    lp.walk_len = P<uint32_t>(ctx->keys);
    lp.plot_ub = P<uint32_t>(ctx->lcounts);
    lp.aa = d->aa_lines ? 1 : 0;
    launch_line_tiles(lp, 1, s);
    // room for each item's plots (an upper bound: the walk's plot calls)
    if ((rc = scan_counts(ctx, lp.plot_ub, P<uint32_t>(ctx->loffsets), nitems, 2))) return rc;
    if ((rc = read_words(ctx, 3))) return rc;
    const uint32_t nplots = ctx->h_misc[2];
    if ((rc = ensure(ctx, ctx->plots, sizeof(Plot) * std::max<uint32_t>(nplots, 1)))) return rc;
    launch_iota(P<uint32_t>(ctx->vals), nitems, s);
    const int swapped = radix_sort(P<uint32_t>(ctx->keys), P<uint32_t>(ctx->vals), P<uint32_t>(ctx->keys2),
                                   P<uint32_t>(ctx->vals2), nitems, 8, P<uint32_t>(ctx->hist),
                                   P<uint32_t>(ctx->scan_tmp), P<uint32_t>(ctx->misc) + 8, s);
    WalkParams wp;
    memset(&wp, 0, sizeof(wp));
    wp.order = swapped ? P<uint32_t>(ctx->vals2) : P<uint32_t>(ctx->vals);
    wp.items = P<LineItem>(ctx->litems); wp.n_items = nitems; wp.W = pl.W; wp.H = pl.H;
    wp.aa = d->aa_lines ? 1 : 0; wp.ets = LP_ETS; wp.ntx = ntx; wp.nty = nty;
    wp.tw = pl.tw; wp.th = pl.th; wp.scols = pl.scols;
    wp.plot_ub = P<uint32_t>(ctx->lcounts); wp.bin_counts = P<uint32_t>(ctx->pcounts);
    wp.plot_offsets = P<uint32_t>(ctx->loffsets); wp.err = P<uint32_t>(ctx->misc);
    wp.plots = P<Plot>(ctx->plots);
    launch_line_walk(wp, s);
    if ((rc = scan_counts(ctx, wp.bin_counts, P<uint32_t>(ctx->poffsets), nitems, 3))) return rc;
    mark(ctx, pl.timing, ST_SCAN);
    if ((rc = read_words(ctx, 4))) return rc;
    if (ctx->h_misc[0] & 4u) return fail(ctx, SR_ERR_DEVICE, "line walk: plot bound exceeded");
    const uint32_t npairs = ctx->h_misc[3];
    ctx->last.bin_pairs = npairs;
    if (npairs == 0) return SR_OK;
    if ((rc = ensure_pairs(ctx, npairs))) return rc;
    ItemEmitParams ip;
    ip.items = P<LineItem>(ctx->litems); ip.counts = P<uint32_t>(ctx->pcounts);
    ip.offsets = P<uint32_t>(ctx->poffsets); ip.n = nitems; ip.ntx = ntx;
    ip.keys = P<uint32_t>(ctx->keys); ip.vals = P<uint32_t>(ctx->vals);
    launch_item_emit(ip, s);
    mark(ctx, pl.timing, ST_EMIT);
    const uint32_t* items = nullptr;
    if ((rc = bin_pairs(ctx, npairs, ntx * nty, &items))) return rc;
    mark(ctx, pl.timing, ST_SORT);
    mark(ctx, pl.timing, ST_RANGES);
    mark(ctx, pl.timing, ST_STRIPS);
    RasterParams R;
    fill_raster_params(pl, fb, d, tex, R);
    R.ets = LP_ETS; R.ntx = ntx; R.nty = nty;
    R.tile_start = P<uint32_t>(ctx->tstart); R.tile_end = P<uint32_t>(ctx->tend);
    R.items = items;
    R.line_items = P<LineItem>(ctx->litems);
    R.plots = P<Plot>(ctx->plots);
    R.verts = vb;
    R.line_v = idx;
    launch_raster_lines(R, ntx * nty, s);
    mark(ctx, pl.timing, ST_RASTER);
    return SR_OK;
}

int draw_points_impl(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, const sr_draw_desc* d,
                     sr_texture tex, VertexBuf vb, const uint32_t* idx, uint32_t npoints) {
    int rc;
    hipStream_t s = ctx->stream;
    const uint32_t ntx = (pl.W + LP_ETS - 1) / LP_ETS, nty = (pl.H + LP_ETS - 1) / LP_ETS;
    if ((rc = ensure(ctx, ctx->counts, sizeof(uint32_t) * (npoints + 1))) ||
        (rc = ensure(ctx, ctx->offsets, sizeof(uint32_t) * (npoints + 1))))
        return rc;
    PointSetupParams pp;
    memset(&pp, 0, sizeof(pp));
    pp.v = vb; pp.idx = idx; pp.n = npoints; pp.W = pl.W; pp.H = pl.H; pp.tw = pl.tw; pp.th = pl.th;
    pp.scols = pl.scols; pp.ets = LP_ETS; pp.ntx = ntx;
    pp.counts = P<uint32_t>(ctx->counts); pp.offsets = P<uint32_t>(ctx->offsets);
    pp.pass = 0;
    launch_point_setup(pp, s);
    mark(ctx, pl.timing, ST_SETUP);
    if ((rc = scan_counts(ctx, pp.counts, P<uint32_t>(ctx->offsets), npoints, 1))) return rc;
    mark(ctx, pl.timing, ST_SCAN);
    if ((rc = read_words(ctx, 2))) return rc;
    const uint32_t npairs = ctx->h_misc[1];
    ctx->last.bin_pairs = npairs;
    if (npairs == 0) return SR_OK;
    if ((rc = ensure_pairs(ctx, npairs))) return rc;
    pp.keys = P<uint32_t>(ctx->keys); pp.vals = P<uint32_t>(ctx->vals); pp.pass = 1;
    launch_point_setup(pp, s);
    mark(ctx, pl.timing, ST_EMIT);
    const uint32_t* items = nullptr;
    if ((rc = bin_pairs(ctx, npairs, ntx * nty, &items))) return rc;
    mark(ctx, pl.timing, ST_SORT);
    mark(ctx, pl.timing, ST_RANGES);
    mark(ctx, pl.timing, ST_STRIPS);
    RasterParams R;
    fill_raster_params(pl, fb, d, tex, R);
    R.ets = LP_ETS; R.ntx = ntx; R.nty = nty;
    R.tile_start = P<uint32_t>(ctx->tstart); R.tile_end = P<uint32_t>(ctx->tend);
    R.items = items;
    R.verts = vb;
    R.point_v = idx;
    launch_raster_points(R, ntx * nty, s);
    mark(ctx, pl.timing, ST_RASTER);
    return SR_OK;
}
