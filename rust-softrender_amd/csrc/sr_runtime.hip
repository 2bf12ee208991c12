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
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "sr_kernels.hpp"

using namespace sr;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

enum Stage { ST_VERTEX, ST_GEOMETRY, ST_SETUP, ST_SCAN, ST_EMIT, ST_SORT, ST_RANGES, ST_STRIPS,
             ST_RASTER, ST_COUNT };

}  // namespace

struct sr_context_t {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf verts, gverts, recs, counts, offsets, keys, vals, keys2, vals2, tstart, tend, scan_tmp,
        hist, misc, strips, strip_aux, litems, plots, lcounts, loffsets, pcounts, poffsets;
    uint32_t* h_misc = nullptr;  // pinned readback words
    hipEvent_t ev[ST_COUNT + 1] = {};
    float stage_ms[ST_COUNT] = {};
    int n_stage = 0;
    sr_draw_stats last = {};
    // execution tile chosen when sr_draw_desc.exec_tile == 0: the first
    // entry whose pixel bound holds the framebuffer (set by the autotuner)
    std::vector<std::pair<uint64_t, int>> tile_policy{{~0ull, 8}};
    struct Batch* batch = nullptr;
};

struct sr_mesh_t {
    sr_context ctx;
    float* soa[8] = {};  // px py pz nx ny nz u v (null if absent)
    uint32_t* idx = nullptr;
    uint32_t nv = 0, ni = 0;
    int prim = 0;
};

struct sr_texture_t {
    sr_context ctx;
    uint32_t w = 0, h = 0;
    uint32_t* d = nullptr;
};

struct sr_framebuffer_t {
    sr_context ctx;
    uint32_t w = 0, h = 0;
    uint32_t* color = nullptr;
    float* depth = nullptr;
    uint8_t* stencil = nullptr;
};

namespace {

int fail(sr_context ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIP_OK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail((ctx), SR_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

int ensure(sr_context ctx, DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes) return SR_OK;
    if (b.p) {
        HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
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

template <class T>
T* P(DevBuf& b) { return reinterpret_cast<T*>(b.p); }

int bits_for(uint32_t n) {  // bits to hold values 0..n-1
    int b = 1;
    while (b < 32 && (1ull << b) < n) b++;
    return b;
}

void mark(sr_context ctx, bool timing, int stage) {
    if (timing) hipEventRecord(ctx->ev[stage + 1], ctx->stream);
}

int read_words(sr_context ctx, int n) {
    HIP_OK(ctx, hipMemcpyAsync(ctx->h_misc, ctx->misc.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int varyings_of(int vp) { return vp == SR_VP_MVP_NORMAL_WORLD ? 6 : 5; }

int policy_tile(sr_context ctx, uint32_t w, uint32_t h) {
    const uint64_t px = (uint64_t)w * h;
    for (const auto& e : ctx->tile_policy)
        if (px <= e.first) return e.second;
    return 8;
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
    launch_fill(P<uint32_t>(ctx->tstart), 0u, ntiles, s);
    launch_fill(P<uint32_t>(ctx->tend), 0u, ntiles, s);
    launch_ranges(skeys, npairs, 0u, P<uint32_t>(ctx->tstart), P<uint32_t>(ctx->tend), s);
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

struct DrawPlan {
    uint32_t W, H, ets, ntx, nty, ntiles, tw, th, scols, srows;
    int K, smode;
    bool exhaustive, timing;
};

void fill_raster_params(const DrawPlan& pl, sr_framebuffer fb, const sr_draw_desc* d,
                        sr_texture tex, RasterParams& R) {
    memset(&R, 0, sizeof(R));
    R.W = pl.W; R.H = pl.H;
    R.color = fb->color; R.depth = fb->depth; R.stencil = fb->stencil;
    R.tw = pl.tw; R.th = pl.th; R.scols = pl.scols; R.srows = pl.srows;
    R.ets = pl.ets; R.ntx = pl.ntx; R.nty = pl.nty;
    R.K = pl.K;
    R.fp = d->fragment_program; R.blend = d->blend; R.alpha_mode = d->alpha_mode;
    R.s_test = d->stencil_test; R.s_op = d->stencil_op; R.s_wrap = d->stencil_wrap;
    R.s_ref = d->stencil_ref & 0xFFu;
    R.smode = pl.smode;
    if (tex) { R.tex = tex->d; R.texW = tex->w; R.texH = tex->h; }
    R.tex_filter = d->tex_filter; R.tex_edge = d->tex_edge; R.tex_border = d->tex_border;
    memcpy(R.fsu, d->fs_uniforms, sizeof(R.fsu));
}

// ---------------------------------------------------------------------------
// The triangle draw
// ---------------------------------------------------------------------------
int draw_tris(sr_context ctx, sr_framebuffer fb, const DrawPlan& pl, const sr_draw_desc* d,
              sr_texture tex, VertexBuf vb, const uint32_t* idx, uint32_t ntris) {
    int rc;
    hipStream_t s = ctx->stream;
    ctx->last.triangles_setup = ntris;
    if ((rc = ensure(ctx, ctx->recs, sizeof(TriRec) * std::max<uint32_t>(ntris, 1)))) return rc;
    if ((rc = ensure(ctx, ctx->counts, sizeof(uint32_t) * (ntris + 1)))) return rc;
    if ((rc = ensure(ctx, ctx->offsets, sizeof(uint32_t) * (ntris + 1)))) return rc;
    SetupParams sp;
    sp.v = vb; sp.idx = idx; sp.n = ntris; sp.cull = d->cull; sp.W = pl.W; sp.H = pl.H;
    sp.ets = pl.ets; sp.recs = P<TriRec>(ctx->recs); sp.counts = P<uint32_t>(ctx->counts);
    sp.err = P<uint32_t>(ctx->misc);
    launch_tri_setup(sp, s);
    mark(ctx, pl.timing, ST_SETUP);

    RasterParams R;
    fill_raster_params(pl, fb, d, tex, R);
    R.tris = P<TriRec>(ctx->recs);
    R.verts = vb;
    R.error_flags = P<uint32_t>(ctx->misc);
    if (pl.exhaustive) {
        if ((rc = read_words(ctx, 1))) return rc;
        if (ctx->h_misc[0]) return fail(ctx, SR_ERR_INVALID_COORDINATE, "NaN screen coordinate");
        R.exhaustive_n = ntris;
        for (int st = ST_SCAN; st <= ST_RANGES; st++) mark(ctx, pl.timing, st);
    } else {
        if ((rc = scan_counts(ctx, P<uint32_t>(ctx->counts), P<uint32_t>(ctx->offsets), ntris)))
            return rc;
        mark(ctx, pl.timing, ST_SCAN);
        if ((rc = read_words(ctx, 2))) return rc;
        if (ctx->h_misc[0]) return fail(ctx, SR_ERR_INVALID_COORDINATE, "NaN screen coordinate");
        const uint32_t npairs = ctx->h_misc[1];
        ctx->last.bin_pairs = npairs;
        const size_t pb = sizeof(uint32_t) * std::max<uint32_t>(npairs, 1);
        if ((rc = ensure(ctx, ctx->keys, pb)) || (rc = ensure(ctx, ctx->vals, pb)) ||
            (rc = ensure(ctx, ctx->keys2, pb)) || (rc = ensure(ctx, ctx->vals2, pb)))
            return rc;
        if ((rc = ensure(ctx, ctx->hist, sizeof(uint32_t) * (radix_hist_words(npairs) + 256))))
            return rc;
        if ((rc = ensure(ctx, ctx->scan_tmp,
                         sizeof(uint32_t) * (scan_tmp_words(radix_hist_words(npairs) + 256) + 16))))
            return rc;
        if ((rc = ensure(ctx, ctx->tstart, sizeof(uint32_t) * pl.ntiles)) ||
            (rc = ensure(ctx, ctx->tend, sizeof(uint32_t) * pl.ntiles)))
            return rc;
        EmitParams ep;
        ep.recs = P<TriRec>(ctx->recs); ep.counts = P<uint32_t>(ctx->counts);
        ep.offsets = P<uint32_t>(ctx->offsets); ep.n = ntris; ep.W = pl.W; ep.H = pl.H;
        ep.ets = pl.ets; ep.ntx = pl.ntx; ep.key_base = 0; ep.val_base = 0;
        ep.keys = P<uint32_t>(ctx->keys); ep.vals = P<uint32_t>(ctx->vals);
        launch_tri_emit(ep, s);
        mark(ctx, pl.timing, ST_EMIT);
        const int swapped = radix_sort(P<uint32_t>(ctx->keys), P<uint32_t>(ctx->vals),
                                       P<uint32_t>(ctx->keys2), P<uint32_t>(ctx->vals2), npairs,
                                       bits_for(pl.ntiles), P<uint32_t>(ctx->hist),
                                       P<uint32_t>(ctx->scan_tmp), P<uint32_t>(ctx->misc) + 2, s);
        const uint32_t* skeys = swapped ? P<uint32_t>(ctx->keys2) : P<uint32_t>(ctx->keys);
        const uint32_t* svals = swapped ? P<uint32_t>(ctx->vals2) : P<uint32_t>(ctx->vals);
        mark(ctx, pl.timing, ST_SORT);
        launch_fill(P<uint32_t>(ctx->tstart), 0u, pl.ntiles, s);
        launch_fill(P<uint32_t>(ctx->tend), 0u, pl.ntiles, s);
        launch_ranges(skeys, npairs, 0u, P<uint32_t>(ctx->tstart), P<uint32_t>(ctx->tend), s);
        mark(ctx, pl.timing, ST_RANGES);
        R.tile_start = P<uint32_t>(ctx->tstart);
        R.tile_end = P<uint32_t>(ctx->tend);
        R.items = svals;
    }
    if (pl.smode == 1) {
        if ((rc = ensure(ctx, ctx->strips, sizeof(uint32_t) * (size_t)pl.W * pl.H))) return rc;
        const size_t aux = (size_t)(pl.scols + 1) * (pl.H + 1) * 2 + (size_t)(pl.srows + 1) * (pl.W + 1) * 2 +
                           (size_t)(pl.srows + 1) * (pl.scols + 1) * 4;
        if ((rc = ensure(ctx, ctx->strip_aux, sizeof(uint32_t) * aux))) return rc;
        StripParams stp;
        uint32_t* a = P<uint32_t>(ctx->strip_aux);
        stp.recs = P<TriRec>(ctx->recs); stp.n = ntris; stp.W = pl.W; stp.H = pl.H;
        stp.tw = pl.tw; stp.th = pl.th; stp.scols = pl.scols; stp.srows = pl.srows;
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

int validate_draw(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex) {
    if (!ctx || !fb || !mesh || !d) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: null argument");
    if (fb->ctx != ctx || mesh->ctx != ctx) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: handle from another context");
    if (d->vertex_program != SR_VP_MVP_NORMAL_UV && d->vertex_program != SR_VP_MVP_NORMAL_WORLD)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown vertex program");
    if (d->geometry_program < SR_GS_NONE || d->geometry_program > SR_GS_CLIP_NEAR)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown geometry program");
    if (d->fragment_program < SR_FP_FLAT || d->fragment_program > SR_FP_UV_COLOR)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: unknown fragment program");
    if (d->fragment_program == SR_FP_TEX_LAMBERT && !tex)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: fragment program samples a texture");
    if (d->fragment_program == SR_FP_BLINN_PHONG && d->vertex_program != SR_VP_MVP_NORMAL_WORLD)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: Blinn-Phong needs world-space varyings");
    if ((d->fragment_program == SR_FP_TEX_LAMBERT || d->fragment_program == SR_FP_UV_COLOR) &&
        d->vertex_program != SR_VP_MVP_NORMAL_UV)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: program needs uv varyings");
    if (d->stencil_test < 0 || d->stencil_test > 7 || d->stencil_op < 0 || d->stencil_op > 5)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: bad stencil config");
    if (d->exec_tile != 0 && d->exec_tile != 8 && d->exec_tile != 16 && d->exec_tile != 32 && d->exec_tile != 64)
        return fail(ctx, SR_ERR_INVALID_ARGUMENT, "draw: exec_tile must be 0, 8, 16, 32 or 64");
    return SR_OK;
}


// ===========================================================================
// Batched rendering: many independent (framebuffer, mesh) triangle draws in
// one pass of the pipeline.  Every stage runs once over the whole batch:
// vertices and triangles are numbered across scenes, execution tiles get
// global ids so one stable sort bins all scenes, and the raster is one
// persistent launch per (tile size, program, blend) class pulling tiles
// from per-XCD queues.  Sizes that depend on the geometry (bin pairs) use a
// capacity checked on the device: on overflow nothing is drawn, the
// capacity grows and the batch runs again, so the sequence has no host
// round trip and is captured once into a hipGraph and replayed while the
// batch is unchanged.
// ===========================================================================
struct Batch {
    DevBuf verts, recs, counts, offsets, keys, vals, keys2, vals2, tstart, tend, hist, scan, tables, misc;
    uint32_t cap = 0;          // bin pair capacity
    std::vector<uint8_t> key;  // what the cached graph renders
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* h_stage = nullptr;  // pinned staging for the scene tables
    size_t h_stage_bytes = 0;
};

static void invalidate_batch_graph(sr_context ctx) {
    Batch* B = ctx->batch;
    if (!B) return;
    if (B->exec) hipGraphExecDestroy(B->exec);
    if (B->graph) hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
    B->key.clear();
}

static int set_batch_capacity(sr_context ctx, uint32_t pairs) {
    if (!ctx->batch) ctx->batch = new (std::nothrow) Batch();
    if (!ctx->batch) return SR_ERR_OUT_OF_MEMORY;
    if (pairs > ctx->batch->cap) {
        ctx->batch->cap = pairs;
        invalidate_batch_graph(ctx);
    }
    return SR_OK;
}

static void destroy_batch(sr_context ctx) {
    Batch* B = ctx->batch;
    if (!B) return;
    invalidate_batch_graph(ctx);
    DevBuf* bufs[] = {&B->verts, &B->recs, &B->counts, &B->offsets, &B->keys, &B->vals, &B->keys2,
                      &B->vals2, &B->tstart, &B->tend, &B->hist, &B->scan, &B->tables, &B->misc};
    for (DevBuf* b : bufs)
        if (b->p) hipFree(b->p);
    if (B->h_stage) hipHostFree(B->h_stage);
    delete B;
    ctx->batch = nullptr;
}

// grows a batch buffer; a reallocation invalidates the cached graph
static int bensure(sr_context ctx, DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes) return SR_OK;
    invalidate_batch_graph(ctx);
    return ensure(ctx, b, bytes);
}

static bool batchable(sr_mesh m, const sr_draw_desc* d) {
    return m->prim == SR_PRIM_TRIANGLES && d->geometry_program == SR_GS_NONE &&
           d->stencil_test == SR_STENCIL_ALWAYS && d->stencil_op == SR_STENCIL_KEEP &&
           !(d->exec_flags & SR_EXEC_EXHAUSTIVE);
}

struct BatchScene {
    uint32_t item;
    uint32_t ets;
    uint64_t cls;  // (ets, fp, blend)
};

static int batch_draw(sr_context ctx, const sr_batch_item* items, uint32_t n) {
    int rc;
    hipSetDevice(ctx->device);
    ctx->n_stage = 0;
    std::vector<BatchScene> fast;
    std::vector<sr_framebuffer> seen;
    for (uint32_t i = 0; i < n; i++) {
        const sr_batch_item& it = items[i];
        if ((rc = validate_draw(ctx, it.fb, it.mesh, it.desc, it.tex))) return rc;
        for (sr_framebuffer f : seen)
            if (f == it.fb) return fail(ctx, SR_ERR_INVALID_ARGUMENT, "batch: a framebuffer appears twice");
        seen.push_back(it.fb);
    }
    for (uint32_t i = 0; i < n; i++) {
        const sr_batch_item& it = items[i];
        const sr_draw_desc* d = it.desc;
        const uint32_t per = (uint32_t)it.mesh->prim;
        const uint32_t nprim = (it.mesh->ni ? it.mesh->ni : it.mesh->nv) / per;
        const bool drawable = it.fb->w > 1 && it.fb->h > 1 && nprim > 0 && d->stencil_test != SR_STENCIL_NEVER;
        if (batchable(it.mesh, d)) {
            const uint32_t ets = d->exec_tile ? (uint32_t)d->exec_tile : (uint32_t)policy_tile(ctx, it.fb->w, it.fb->h);
            fast.push_back(BatchScene{i, ets, ((uint64_t)ets << 16) | ((uint64_t)d->fragment_program << 8) |
                                                   (uint64_t)d->blend});
            (void)drawable;
        } else {
            if (it.clear && (rc = sr_fb_clear(it.fb, it.clear_rgba, -3.40282347e+38f, 0))) return rc;
            if ((rc = sr_draw(ctx, it.fb, it.mesh, d, it.tex))) return rc;
        }
    }
    if (fast.empty()) return SR_OK;
    std::stable_sort(fast.begin(), fast.end(), [](const BatchScene& a, const BatchScene& b) { return a.cls < b.cls; });
    if (!ctx->batch && (rc = set_batch_capacity(ctx, 0))) return rc;
    Batch* B = ctx->batch;
    hipStream_t s = ctx->stream;

    // ---- host tables
    const uint32_t ns = (uint32_t)fast.size();
    std::vector<SceneGeom> geom(ns);
    std::vector<RasterParams> rp(ns);
    std::vector<uint32_t> vend(ns), tend(ns), tile_end(ns);
    struct Class { uint64_t cls; uint32_t ets, fp, blend, tile0, ntiles; };
    std::vector<Class> classes;
    uint64_t NV = 0, NT = 0, NTILES = 0;
    for (uint32_t k = 0; k < ns; k++) {
        const sr_batch_item& it = items[fast[k].item];
        const sr_draw_desc* d = it.desc;
        const sr_mesh m = it.mesh;
        const uint32_t W = it.fb->w, H = it.fb->h, ets = fast[k].ets;
        const uint32_t tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u, th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
        const bool drawable = W > 1 && H > 1;  // W or H == 1: no tiles (fragment.rs:196,201)
        const uint32_t ntri = drawable ? (m->ni ? m->ni : m->nv) / 3 : 0;
        SceneGeom& g = geom[k];
        memset(&g, 0, sizeof(g));
        g.px = m->soa[0]; g.py = m->soa[1]; g.pz = m->soa[2]; g.nx = m->soa[3]; g.ny = m->soa[4];
        g.nz = m->soa[5]; g.u = m->soa[6]; g.v = m->soa[7]; g.idx = m->idx;
        g.nv = ntri ? m->nv : 0; g.vbase = (uint32_t)NV;
        g.nt = ntri; g.tbase = (uint32_t)NT;
        g.program = d->vertex_program; g.K = varyings_of(d->vertex_program); g.cull = d->cull;
        g.W = W; g.H = H; g.ets = ets; g.ntx = (W + ets - 1) / ets;
        g.tile_base = (uint32_t)NTILES;
        g.viewport = d->viewport;
        memcpy(g.U, d->vs_uniforms, sizeof(g.U));
        const uint32_t ntiles = g.ntx * ((H + ets - 1) / ets);
        NV += g.nv; NT += g.nt; NTILES += ntiles;
        vend[k] = (uint32_t)NV; tend[k] = (uint32_t)NT; tile_end[k] = (uint32_t)NTILES;
        if (classes.empty() || classes.back().cls != fast[k].cls)
            classes.push_back(Class{fast[k].cls, ets, (uint32_t)d->fragment_program, (uint32_t)d->blend,
                                    g.tile_base, 0});
        classes.back().ntiles += ntiles;
    }
    if (NV >= (1ull << 31) || NT >= (1ull << 31) || NTILES >= (1ull << 31))
        return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: too large");
    if (B->cap == 0) B->cap = (uint32_t)std::min<uint64_t>(4 * NT + (1u << 16), 0x7FFFFFFFull);

    for (int attempt = 0; attempt < 4; attempt++) {
        const uint32_t cap = B->cap;
        const uint32_t nvb = (uint32_t)std::max<uint64_t>(NV, 1);
        if ((rc = bensure(ctx, B->verts, sizeof(float) * 10ull * nvb)) ||
            (rc = bensure(ctx, B->recs, sizeof(TriRec) * std::max<uint64_t>(NT, 1))) ||
            (rc = bensure(ctx, B->counts, sizeof(uint32_t) * (NT + 1))) ||
            (rc = bensure(ctx, B->offsets, sizeof(uint32_t) * (NT + 1))) ||
            (rc = bensure(ctx, B->keys, sizeof(uint32_t) * (size_t)cap)) ||
            (rc = bensure(ctx, B->vals, sizeof(uint32_t) * (size_t)cap)) ||
            (rc = bensure(ctx, B->keys2, sizeof(uint32_t) * (size_t)cap)) ||
            (rc = bensure(ctx, B->vals2, sizeof(uint32_t) * (size_t)cap)) ||
            (rc = bensure(ctx, B->tstart, sizeof(uint32_t) * std::max<uint64_t>(NTILES, 1))) ||
            (rc = bensure(ctx, B->tend, sizeof(uint32_t) * std::max<uint64_t>(NTILES, 1))) ||
            (rc = bensure(ctx, B->hist, sizeof(uint32_t) * (radix_hist_words(cap) + 256))) ||
            (rc = bensure(ctx, B->scan, sizeof(uint32_t) * (scan_tmp_words(std::max<uint32_t>(radix_hist_words(cap) + 256, (uint32_t)NT + 1)) + 16))) ||
            (rc = bensure(ctx, B->misc, sizeof(uint32_t) * (8 + 8 * classes.size()))))
            return rc;
        const size_t tbytes = sizeof(SceneGeom) * ns + sizeof(RasterParams) * ns + sizeof(uint32_t) * 3 * ns + 64;
        if ((rc = bensure(ctx, B->tables, tbytes))) return rc;

        // scene tables (device copies, pointers into the batch buffers)
        uint8_t* tb = P<uint8_t>(B->tables);
        SceneGeom* d_geom = (SceneGeom*)tb;
        RasterParams* d_rp = (RasterParams*)(tb + sizeof(SceneGeom) * ns);
        uint32_t* d_vend = (uint32_t*)(tb + sizeof(SceneGeom) * ns + sizeof(RasterParams) * ns);
        uint32_t* d_tend = d_vend + ns;
        uint32_t* d_tile_end = d_tend + ns;
        uint32_t* misc = P<uint32_t>(B->misc);  // [0] overflow [1] pairs [2] err [3] scratch [8..) queues
        const VertexBuf vb{P<float>(B->verts), nvb};
        const int swapped_passes = (bits_for((uint32_t)std::max<uint64_t>(NTILES, 2)) + 7) / 8;
        const uint32_t* sorted_vals = (swapped_passes & 1) ? P<uint32_t>(B->vals2) : P<uint32_t>(B->vals);
        for (uint32_t k = 0; k < ns; k++) {
            const sr_batch_item& it = items[fast[k].item];
            const sr_draw_desc* d = it.desc;
            RasterParams& R = rp[k];
            memset(&R, 0, sizeof(R));
            R.W = it.fb->w; R.H = it.fb->h;
            R.color = it.fb->color; R.depth = it.fb->depth; R.stencil = it.fb->stencil;
            R.tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u;
            R.th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
            R.scols = R.W > 1 ? (R.W - 1 + R.tw - 1) / R.tw : 0;
            R.srows = R.H > 1 ? (R.H - 1 + R.th - 1) / R.th : 0;
            R.ets = geom[k].ets; R.ntx = geom[k].ntx; R.nty = (R.H + R.ets - 1) / R.ets;
            R.tile_start = P<uint32_t>(B->tstart); R.tile_end = P<uint32_t>(B->tend);
            R.items = sorted_vals;
            R.tris = P<TriRec>(B->recs);
            R.verts = vb;
            R.K = geom[k].K;
            R.fp = d->fragment_program; R.blend = d->blend; R.alpha_mode = d->alpha_mode;
            R.s_test = SR_STENCIL_ALWAYS; R.s_op = SR_STENCIL_KEEP;
            R.smode = 0;
            if (it.tex) { R.tex = it.tex->d; R.texW = it.tex->w; R.texH = it.tex->h; }
            R.tex_filter = d->tex_filter; R.tex_edge = d->tex_edge; R.tex_border = d->tex_border;
            memcpy(R.fsu, d->fs_uniforms, sizeof(R.fsu));
            R.tile_base = geom[k].tile_base;
            R.clear = it.clear ? 1 : 0;
            R.clear_rgba = it.clear_rgba;
        }
        // cache key: everything the captured sequence depends on
        std::vector<uint8_t> key;
        auto put = [&](const void* p, size_t b) {
            const uint8_t* q = (const uint8_t*)p;
            key.insert(key.end(), q, q + b);
        };
        put(&cap, sizeof(cap));
        put(geom.data(), sizeof(SceneGeom) * ns);
        put(rp.data(), sizeof(RasterParams) * ns);
        const bool replay = B->exec && key == B->key;
        if (!replay) {
            invalidate_batch_graph(ctx);
            if (B->h_stage_bytes < tbytes) {
                if (B->h_stage) hipHostFree(B->h_stage);
                B->h_stage = nullptr;
                B->h_stage_bytes = 0;
                if (hipHostMalloc((void**)&B->h_stage, tbytes, hipHostMallocDefault) != hipSuccess)
                    return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: staging alloc failed");
                B->h_stage_bytes = tbytes;
            }
            uint8_t* hs = (uint8_t*)B->h_stage;
            memcpy(hs, geom.data(), sizeof(SceneGeom) * ns);
            memcpy(hs + sizeof(SceneGeom) * ns, rp.data(), sizeof(RasterParams) * ns);
            uint32_t* hv = (uint32_t*)(hs + sizeof(SceneGeom) * ns + sizeof(RasterParams) * ns);
            memcpy(hv, vend.data(), 4 * ns);
            memcpy(hv + ns, tend.data(), 4 * ns);
            memcpy(hv + 2 * ns, tile_end.data(), 4 * ns);
            HIP_OK(ctx, hipMemcpyAsync(tb, hs, tbytes - 64, hipMemcpyHostToDevice, s));
            HIP_OK(ctx, hipStreamSynchronize(s));

            // ---- capture the batch
            HIP_OK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            BatchTables T{d_geom, d_vend, d_tend, ns, (uint32_t)NV, (uint32_t)NT};
            launch_fill(misc, 0u, (uint32_t)(8 + 8 * classes.size()), s);
            launch_vertex_batch(T, vb, s);
            launch_tri_setup_batch(T, vb, P<TriRec>(B->recs), P<uint32_t>(B->counts), misc + 2, s);
            launch_scan(P<uint32_t>(B->counts), P<uint32_t>(B->offsets), (uint32_t)NT, P<uint32_t>(B->scan), misc + 1, s);
            launch_tri_emit_batch(T, P<TriRec>(B->recs), P<uint32_t>(B->counts), P<uint32_t>(B->offsets), cap,
                                  P<uint32_t>(B->keys), P<uint32_t>(B->vals), misc + 0, s);
            const int sw = radix_sort(P<uint32_t>(B->keys), P<uint32_t>(B->vals), P<uint32_t>(B->keys2),
                                      P<uint32_t>(B->vals2), cap, bits_for((uint32_t)std::max<uint64_t>(NTILES, 2)),
                                      P<uint32_t>(B->hist), P<uint32_t>(B->scan), misc + 3, s, misc + 1);
            const uint32_t* skeys = sw ? P<uint32_t>(B->keys2) : P<uint32_t>(B->keys);
            launch_fill(P<uint32_t>(B->tstart), 0u, (uint32_t)NTILES, s);
            launch_fill(P<uint32_t>(B->tend), 0u, (uint32_t)NTILES, s);
            launch_ranges(skeys, cap, 0u, P<uint32_t>(B->tstart), P<uint32_t>(B->tend), s, misc + 1);
            for (size_t c = 0; c < classes.size(); c++) {
                const Class& C = classes[c];
                if (C.ntiles == 0) continue;
                RasterBatch RB;
                memset(&RB, 0, sizeof(RB));
                RB.one.fp = (int)C.fp; RB.one.blend = (int)C.blend; RB.one.ets = C.ets; RB.one.smode = 0;
                RB.scenes = d_rp;
                RB.scene_tile_end = d_tile_end;
                RB.nscenes = ns;
                RB.tile0 = C.tile0;
                RB.ntiles = C.ntiles;
                RB.queue = misc + 8 + 8 * c;
                RB.overflow = misc;
                const uint32_t grid = std::min<uint32_t>(C.ntiles, C.ets == 8 ? 256u * 20u : 256u * 5u);
                launch_raster_tris(RB, grid, s);
            }
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (e != hipSuccess) return fail(ctx, SR_ERR_DEVICE, std::string("batch capture: ") + hipGetErrorString(e));
            B->graph = g;
            HIP_OK(ctx, hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
            B->key = key;
        }
        HIP_OK(ctx, hipEventRecord(ctx->ev[0], s));
        HIP_OK(ctx, hipGraphLaunch(B->exec, s));
        HIP_OK(ctx, hipEventRecord(ctx->ev[1], s));
        HIP_OK(ctx, hipMemcpyAsync(ctx->h_misc, misc, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(ctx, hipStreamSynchronize(s));
        float ms = 0.f;
        hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
        ctx->stage_ms[0] = ms;
        ctx->n_stage = 1;
        ctx->last.bin_pairs = ctx->h_misc[1];
        ctx->last.triangles_in = NT;
        if (ctx->h_misc[2]) return fail(ctx, SR_ERR_INVALID_COORDINATE, "batch: NaN screen coordinate");
        if (!ctx->h_misc[0]) return SR_OK;
        // binning overflowed: nothing was drawn; grow and run again
        const uint64_t want = (uint64_t)ctx->h_misc[1] + ctx->h_misc[1] / 4 + 4096;
        if (want >= 0x7FFFFFFFull) return fail(ctx, SR_ERR_OUT_OF_MEMORY, "batch: bin pairs exceed 2^31");
        B->cap = (uint32_t)want;
        invalidate_batch_graph(ctx);
    }
    return fail(ctx, SR_ERR_DEVICE, "batch: capacity did not converge");
}

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
    if (ensure(ctx, ctx->misc, 64 * sizeof(uint32_t)) != SR_OK) {
        delete ctx;
        return SR_ERR_OUT_OF_MEMORY;
    }
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
    m->ctx = ctx; m->nv = nv; m->ni = indices ? ni : 0; m->prim = primitive;
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
    delete m;
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
    hipFree(t->d);
    delete t;
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

int sr_fb_free(sr_framebuffer fb) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    hipStreamSynchronize(fb->ctx->stream);
    if (fb->color) hipFree(fb->color);
    if (fb->depth) hipFree(fb->depth);
    if (fb->stencil) hipFree(fb->stencil);
    delete fb;
    return SR_OK;
}

int sr_fb_clear(sr_framebuffer fb, uint32_t rgba, float depth, uint8_t stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    launch_fb_clear(fb->color, fb->depth, fb->stencil, fb->w * fb->h, rgba, depth, stencil, fb->ctx->stream);
    return SR_OK;
}

int sr_fb_read(sr_framebuffer fb, uint32_t* rgba, float* depth, uint8_t* stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    sr_context ctx = fb->ctx;
    const size_t n = (size_t)fb->w * fb->h;
    if (rgba) HIP_OK(ctx, hipMemcpyAsync(rgba, fb->color, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    if (depth) HIP_OK(ctx, hipMemcpyAsync(depth, fb->depth, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    if (stencil) HIP_OK(ctx, hipMemcpyAsync(stencil, fb->stencil, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_fb_write(sr_framebuffer fb, const uint32_t* rgba, const float* depth, const uint8_t* stencil) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    sr_context ctx = fb->ctx;
    const size_t n = (size_t)fb->w * fb->h;
    if (rgba) HIP_OK(ctx, hipMemcpyAsync(fb->color, rgba, 4 * n, hipMemcpyHostToDevice, ctx->stream));
    if (depth) HIP_OK(ctx, hipMemcpyAsync(fb->depth, depth, 4 * n, hipMemcpyHostToDevice, ctx->stream));
    if (stencil) HIP_OK(ctx, hipMemcpyAsync(fb->stencil, stencil, n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_fb_size(sr_framebuffer fb, uint32_t* w, uint32_t* h) {
    if (!fb) return SR_ERR_INVALID_ARGUMENT;
    if (w) *w = fb->w;
    if (h) *h = fb->h;
    return SR_OK;
}

int sr_draw(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex) {
    if (int rc = validate_draw(ctx, fb, mesh, d, tex)) return rc;
    hipSetDevice(ctx->device);
    memset(&ctx->last, 0, sizeof(ctx->last));
    ctx->n_stage = 0;

    DrawPlan pl;
    pl.W = fb->w; pl.H = fb->h;
    pl.tw = d->tile_w > 0 ? (uint32_t)d->tile_w : 128u;
    pl.th = d->tile_h > 0 ? (uint32_t)d->tile_h : 128u;
    pl.ets = d->exec_tile ? (uint32_t)d->exec_tile : (uint32_t)policy_tile(ctx, fb->w, fb->h);
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

    const uint32_t per = (uint32_t)mesh->prim;
    const uint32_t nidx = mesh->ni ? mesh->ni : mesh->nv;
    const uint32_t nprim = nidx / per;
    ctx->last.triangles_in = mesh->prim == SR_PRIM_TRIANGLES ? nprim : 0;
    ctx->last.lines_in = mesh->prim == SR_PRIM_LINES ? nprim : 0;
    ctx->last.points_in = mesh->prim == SR_PRIM_POINTS ? nprim : 0;
    // StencilTest::Never: no pixel passes, nothing is written (stencil.rs:115)
    if (pl.scols == 0 || pl.srows == 0 || d->stencil_test == SR_STENCIL_NEVER || nprim == 0) return SR_OK;

    int rc;
    hipStream_t s = ctx->stream;
    if (pl.timing) hipEventRecord(ctx->ev[0], s);
    // clear the error word
    launch_fill(P<uint32_t>(ctx->misc), 0u, 4u, s);

    // ---- vertex stage
    const uint32_t nv = mesh->nv;
    if ((rc = ensure(ctx, ctx->verts, sizeof(float) * (size_t)(4 + pl.K) * std::max<uint32_t>(nv, 1)))) return rc;
    VertexBuf vb{P<float>(ctx->verts), std::max<uint32_t>(nv, 1)};
    VSParams vs;
    memset(&vs, 0, sizeof(vs));
    vs.px = mesh->soa[0]; vs.py = mesh->soa[1]; vs.pz = mesh->soa[2];
    vs.nx = mesh->soa[3]; vs.ny = mesh->soa[4]; vs.nz = mesh->soa[5];
    vs.u = mesh->soa[6]; vs.v = mesh->soa[7];
    vs.n = nv; vs.program = d->vertex_program; vs.K = pl.K;
    vs.normalize = d->geometry_program == SR_GS_NONE;
    vs.viewport = d->viewport;
    memcpy(vs.U, d->vs_uniforms, sizeof(vs.U));
    vs.out = vb;
    launch_vertex(vs, s);
    mark(ctx, pl.timing, ST_VERTEX);

    const uint32_t* idx = mesh->idx;
    uint32_t nwork = nprim;
    if (d->geometry_program != SR_GS_NONE) {
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
        gp.out = VertexBuf{P<float>(ctx->gverts), std::max<uint32_t>(ngen * per, 1)};
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
    } else if (mesh->prim == SR_PRIM_TRIANGLES) {
        rc = draw_tris(ctx, fb, pl, d, tex, vb, idx, nwork);
    } else if (mesh->prim == SR_PRIM_LINES) {
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

int sr_batch_draw(sr_context ctx, const sr_batch_item* items, uint32_t n) {
    if (!ctx || (!items && n)) return SR_ERR_INVALID_ARGUMENT;
    return batch_draw(ctx, items, n);
}

int sr_ctx_set_exec_tile_policy(sr_context ctx, uint32_t n, const uint64_t* max_pixels,
                                const int32_t* exec_tile) {
    if (!ctx || !n || !max_pixels || !exec_tile) return SR_ERR_INVALID_ARGUMENT;
    std::vector<std::pair<uint64_t, int>> p;
    for (uint32_t i = 0; i < n; i++) {
        const int e = exec_tile[i];
        if (e != 8 && e != 16 && e != 32 && e != 64)
            return fail(ctx, SR_ERR_INVALID_ARGUMENT, "policy: exec tile must be 8, 16, 32 or 64");
        p.emplace_back(max_pixels[i], e);
    }
    ctx->tile_policy = p;
    if (ctx->batch) invalidate_batch_graph(ctx);
    return SR_OK;
}

int sr_ctx_set_batch_capacity(sr_context ctx, uint32_t pairs) {
    if (!ctx) return SR_ERR_INVALID_ARGUMENT;
    return set_batch_capacity(ctx, pairs);
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
    launch_line_tiles(lp, 1, s);
    WalkParams wp;
    memset(&wp, 0, sizeof(wp));
    wp.items = P<LineItem>(ctx->litems); wp.n_items = nitems; wp.W = pl.W; wp.H = pl.H;
    wp.aa = d->aa_lines ? 1 : 0; wp.ets = LP_ETS; wp.ntx = ntx; wp.nty = nty;
    wp.plot_counts = P<uint32_t>(ctx->lcounts); wp.bin_counts = P<uint32_t>(ctx->pcounts);
    wp.plot_offsets = P<uint32_t>(ctx->loffsets); wp.err = P<uint32_t>(ctx->misc);
    launch_line_walk(wp, 0, s);
    if ((rc = scan_counts(ctx, wp.plot_counts, P<uint32_t>(ctx->loffsets), nitems, 2))) return rc;
    if ((rc = scan_counts(ctx, wp.bin_counts, P<uint32_t>(ctx->poffsets), nitems, 3))) return rc;
    mark(ctx, pl.timing, ST_SCAN);
    if ((rc = read_words(ctx, 4))) return rc;
    const uint32_t nplots = ctx->h_misc[2], npairs = ctx->h_misc[3];
    ctx->last.bin_pairs = npairs;
    if (npairs == 0) return SR_OK;
    if ((rc = ensure(ctx, ctx->plots, sizeof(Plot) * std::max<uint32_t>(nplots, 1)))) return rc;
    wp.plots = P<Plot>(ctx->plots);
    launch_line_walk(wp, 1, s);
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
