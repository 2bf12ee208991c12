// sr_runtime.hpp — host-side runtime internals shared by the C ABI
// translation units: the context and handle types, device workspaces, and
// the batched pipeline split into plan / enqueue / result so a caller can
// keep several batches in flight on their own streams (the request server)
// or run one batch synchronously through a cached hipGraph (sr_batch_draw).
#pragma once

#include <string.h>

#include <algorithm>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "sr_kernels.hpp"

namespace sr_rt {
using namespace sr;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

enum Stage { ST_VERTEX, ST_GEOMETRY, ST_SETUP, ST_SCAN, ST_EMIT, ST_SORT, ST_RANGES, ST_STRIPS,
             ST_RASTER, ST_COUNT };

struct BatchWork;
struct AsyncBatch;

}  // namespace sr_rt

struct sr_context_t {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    sr_rt::DevBuf verts, gverts, recs, counts, offsets, keys, vals, keys2, vals2, tstart, tend, scan_tmp,
        hist, misc, strips, strip_aux, litems, plots, lcounts, loffsets, pcounts, poffsets;
    uint32_t* h_misc = nullptr;  // pinned readback words
    hipEvent_t ev[sr_rt::ST_COUNT + 1] = {};
    hipEvent_t ev_read = nullptr;  // a readback into h_misc has landed
    float stage_ms[sr_rt::ST_COUNT] = {};
    int n_stage = 0;
    sr_draw_stats last = {};
    // execution tile chosen when sr_draw_desc.exec_tile == 0: the first
    // entry whose pixel bound holds the framebuffer (set by the autotuner)
    // (max pixels, exec tile | waves << 8); waves 0 = the tile size's default
    std::vector<std::pair<uint64_t, int>> tile_policy{{~0ull, 8}};
    uint32_t batch_cap_hint = 0;             // initial bin-pair capacity of new batch workspaces
    sr_rt::BatchWork* batch = nullptr;       // workspace of sr_batch_draw (on ctx->stream)
    std::vector<sr_rt::AsyncBatch*> async;   // slots of sr_batch_submit, each with its own stream
    std::map<uint64_t, int> async_done;      // finished tickets not yet waited for -> status
    uint64_t next_ticket = 1;
    // While sr_draw_static runs: the draw's buffers are carved in order
    // from this fixed device region instead of being allocated.
    uint8_t* arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
};

struct sr_mesh_t {
    sr_context ctx;
    float* soa[8] = {};  // px py pz nx ny nz u v (null if absent)
    int32_t* iattr = nullptr;  // integer attribute (sr_mesh_set_int_attrib), per vertex
    uint32_t* idx = nullptr;
    uint32_t nv = 0, ni = 0;
    bool indexed = false;  // an index list was given (it may be empty: nothing is drawn)
    int prim = 0;
};

// Vertices the mesh's primitives are assembled from: the index list, or
// every vertex in order when there is none.
inline uint32_t mesh_count(const sr_mesh_t* m) { return m->indexed ? m->ni : m->nv; }

struct sr_texture_t {
    sr_context ctx;
    uint32_t w = 0, h = 0;
    uint32_t* d = nullptr;            // texels (RGBAF32: 4 floats each)
    int fmt = SR_COLOR_RGBA8;         // SR_COLOR_*
    bool view = false;              // aliases a framebuffer's colour plane (not owned)
    sr_framebuffer fb = nullptr;    // that framebuffer
};

struct sr_framebuffer_t {
    sr_context ctx;
    uint32_t w = 0, h = 0;
    uint32_t* color = nullptr;    // plane 0 when it is RGBA8
    float* depth = nullptr;       // f32 depth
    uint8_t* stencil = nullptr;   // u8 stencil
    sr_fb_format fmt{1, {SR_COLOR_RGBA8, 0, 0, 0}, SR_DEPTH_F32, SR_STENCIL_U8};
    void* planes[SR_MAX_COLOR_PLANES] = {};
    double* depth64 = nullptr;    // f64 depth
    // sr_fb_create's format: one RGBA8 plane, f32 depth, u8 stencil (the
    // format every draw path takes; the others take triangle draws only)
    bool standard() const {
        return fmt.n_color == 1 && fmt.color[0] == SR_COLOR_RGBA8 && fmt.depth == SR_DEPTH_F32 &&
               fmt.stencil == SR_STENCIL_U8;
    }
};

namespace sr_rt {
using namespace sr;

int fail(sr_context ctx, int code, const std::string& msg);

#define HIP_OK(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return sr_rt::fail((ctx), SR_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// Grows `b` to hold `bytes` (with 25% headroom).  Work queued on `s` that
// may still read the old buffer is drained before it is freed.
int ensure(sr_context ctx, DevBuf& b, size_t bytes, hipStream_t s);
inline int ensure(sr_context ctx, DevBuf& b, size_t bytes) { return ensure(ctx, b, bytes, ctx->stream); }

template <class T>
T* P(DevBuf& b) { return reinterpret_cast<T*>(b.p); }

int bits_for(uint32_t n);  // bits to hold values 0..n-1
int varyings_of(int vp);
int policy_tile(sr_context ctx, uint32_t w, uint32_t h);
// Waves per tile for execution tile `ets`: the desc's, else (auto tile) the
// policy's, made valid for the tile size.
int draw_waves(sr_context ctx, const sr_draw_desc* d, uint32_t ets, uint32_t w, uint32_t h);
int validate_draw(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex);
int draw_foreign(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* d, sr_texture tex);

// True when a draw can join a batch: triangles, no geometry program,
// stencil Always/Keep, not the exhaustive (unbinned) debug mode.
bool samples_f32(const sr_draw_desc* d, sr_texture tex);
bool batchable(sr_mesh m, const sr_draw_desc* d, sr_texture tex);

// ---------------------------------------------------------------------------
// Batched pipeline
// ---------------------------------------------------------------------------
// One batch workspace: device buffers sized by the batch, pinned staging for
// the scene tables and the readback words, a stream, and the cached graph of
// the last synchronously drawn batch.
struct BatchWork {
    DevBuf verts, recs, counts, offsets, keys, vals, keys2, vals2, tstart, tend, hist, scan, tables, misc,
        cost, parts;
    uint32_t cap = 0;          // bin pair capacity
    std::vector<uint8_t> key;  // what the cached graph renders
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint8_t* h_stage = nullptr;  // pinned staging for the scene tables
    size_t h_stage_bytes = 0;
    uint32_t* h_misc = nullptr;  // pinned: [0] overflow [1] bin pairs [2] flags (1 NaN, 2 strip slivers)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    int init(sr_context ctx, bool new_stream);
    void release();
    void drop_graph();
    size_t device_bytes() const;
};

struct BatchClass {
    uint64_t cls;
    uint32_t ets, nwv, fp, blend, tile0, ntiles;
};

// After batch_result: a triangle of the batch may pass the coverage test
// outside its box (setup's strip-risk flag); the batched raster then drew
// nothing and the batch's draws must go through sr_draw one by one.
// A triangle of the batch could pass the coverage test outside its box (it
// was binned to every tile such a pixel may lie in).
inline bool batch_strip_slivers(const BatchWork& W) { return (W.h_misc[2] & 2u) != 0; }

// Host tables of one batch.  `items` are copies (descriptors included), so a
// plan outlives the caller's arrays.  Draws into the same framebuffer form a
// group: consecutive draws (in item order) sharing one range of execution
// tiles.
struct BatchPlan {
    std::vector<sr_batch_item> items;  // per draw, group by group
    std::vector<sr_draw_desc> descs;
    std::vector<SceneGeom> geom;
    std::vector<uint32_t> vend, tend;                  // per draw: prefix of vertices / triangles
    std::vector<uint32_t> group_tile_end, group_first; // per group (+1 entry for group_first)
    std::vector<BatchClass> classes;
    uint64_t NV = 0, NT = 0, NTILES = 0;
};

// One asynchronous batch (sr_batch_submit): its workspace and stream, its
// plan (copies of the items and descriptors), its ticket.
struct AsyncBatch {
    BatchWork W;
    BatchPlan plan;
    uint64_t ticket = 0;
    bool busy = false;
};


// Builds the tables for `items` (all batchable and validated; items on one
// framebuffer share execution tile, program and blend, and only the first
// of them may clear); groups are ordered by (execution tile, program, blend).
int batch_plan(sr_context ctx, const sr_batch_item* items, uint32_t n, BatchPlan& plan);

// Sizes the workspace, uploads the tables and queues the whole batch on
// W.stream, ending with the readback of W.h_misc and W.ev1.  With
// `use_graph` the sequence is captured once and replayed while the batch is
// unchanged (the host waits for the table upload before capturing).
int batch_enqueue(sr_context ctx, BatchWork& W, const BatchPlan& plan, bool use_graph);

// After W.ev1 has completed: SR_OK with *again = false when the batch was
// drawn; *again = true when the bins overflowed (nothing was drawn and the
// capacity has grown: enqueue the same plan again).
int batch_result(sr_context ctx, BatchWork& W, const BatchPlan& plan, bool* again);

// Device bytes a batch of `tris` triangles and `verts` vertices needs in a
// workspace with `pairs_per_tri` bin pairs per triangle.
size_t batch_bytes_estimate(uint64_t tris, uint64_t verts, double pairs_per_tri);

}  // namespace sr_rt
