// sr_kernels.hpp — kernel parameter blocks and host-side launchers.
#pragma once

#include "sr_internal.hpp"

namespace sr {

struct VSParams {
    const float *px, *py, *pz, *nx, *ny, *nz, *u, *v;
    const int32_t* ia;  // integer attribute (SR_VP_MVP_INT), or null
    uint32_t n;
    int program;
    int K;
    int normalize;  // 1: run_to_fragment (viewport fused), 0: keep clip space
    sr_viewport viewport;
    float U[48];
    VertexBuf out;
};

struct GSParams {
    VertexBuf in;        // clip-space vertices
    const uint32_t* idx; // mesh indices (null: implicit)
    uint32_t n_in;       // input primitives
    int K;
    uint32_t* counts;    // outputs per input primitive
    const uint32_t* offsets;
    VertexBuf out;       // generated vertices, screen space (finish fused)
    sr_viewport viewport;
    float G[17];         // normal visualisation: projection * view (row-major), length
};

struct SetupParams {
    VertexBuf v;
    const uint32_t* idx;  // null: vertex 3t+k
    uint32_t n;
    int cull;
    uint32_t W, H, ets;
    uint32_t tw, th;  // screen tiles (clamp_as_int!'s boxes)
    TriRec* recs;
    TriBin* bins;
    TriVid* vids;
    uint32_t* counts;
    uint32_t* err;
    int empty_cull;  // bin only execution tiles the triangle may cover (no stencil effects)
};

struct EmitParams {
    const TriRec* recs;
    const TriBin* bins;
    const uint32_t* counts;
    const uint32_t* offsets;
    uint32_t cap;  // pairs keys/vals hold: a triangle whose pairs would pass it writes none
    uint32_t n, W, H, ets, ntx;
    uint32_t tw, th;  // screen tiles
    uint32_t key_base, val_base;
    uint32_t* keys;
    uint32_t* vals;
    int empty_cull;
};

// One scene of a batch: its vertex program, mesh streams and placement in
// the batch's vertex / triangle / execution-tile numbering.
struct SceneGeom {
    const float *px, *py, *pz, *nx, *ny, *nz, *u, *v;
    const uint32_t* idx;      // null: vertex 3t+k
    uint32_t nv, vbase;       // vertices, first batch vertex
    uint32_t nt, tbase;       // triangles, first batch triangle
    int program, K, cull;
    uint32_t W, H, ets, ntx, tile_base;
    uint32_t tw, th;  // screen tiles
    int empty_cull;
    sr_viewport viewport;
    float U[48];
};

struct BatchTables {
    const SceneGeom* scenes;
    const uint32_t* vend;  // inclusive prefix of vertices per scene
    const uint32_t* tend;  // inclusive prefix of triangles per scene
    uint32_t nscenes, nv, nt;
};

struct LineSetupParams {
    VertexBuf v;
    const uint32_t* idx;   // null: vertices 2l, 2l+1
    uint32_t n_lines;
    uint32_t W, H, tw, th, scols, srows;
    uint32_t* counts;      // per line: accepted screen tiles
    const uint32_t* offsets;
    LineItem* items;
    uint32_t* err;
    int aa;                // Wu (1) or Bresenham (0) walks
    uint32_t* walk_len;    // pass 1 (optional): per item, a sort key, 255 - its walk length (capped)
    uint32_t* plot_ub;     // pass 1: per item, an upper bound of its plots (the walk's plot calls)
};

struct WalkParams {
    LineItem* items;
    uint32_t n_items;
    uint32_t W, H;
    int aa;
    uint32_t ets, ntx, nty;
    uint32_t tw, th, scols;  // screen tiles
    const uint32_t* order;   // optional: thread t walks item order[t] (items of similar length share a wave)
    const uint32_t* plot_ub;       // per item: room for its plots (an upper bound)
    const uint32_t* plot_offsets;  // per item: first plot slot (exclusive scan of plot_ub)
    uint32_t* bin_counts;    // out: execution tiles per item
    Plot* plots;             // out
    uint32_t* err;
};

struct ItemEmitParams {
    const LineItem* items;
    const uint32_t* counts;
    const uint32_t* offsets;
    uint32_t n, ntx;
    uint32_t* keys;
    uint32_t* vals;
};

struct PointSetupParams {
    VertexBuf v;
    const uint32_t* idx;
    uint32_t n;
    uint32_t W, H, tw, th, scols, ets, ntx;
    uint32_t* counts;   // 0/1
    const uint32_t* offsets;
    uint32_t* keys;
    uint32_t* vals;
    int pass;
};

struct StripParams {
    const TriRec* recs;
    uint32_t n;
    uint32_t W, H, tw, th, scols, srows;
    uint32_t* colL;  // [scols+1][H+1]  rows of column tx0_j, difference along y
    uint32_t* colR;  // [scols+1][H+1]
    uint32_t* rowT;  // [srows+1][W+1]
    uint32_t* rowB;  // [srows+1][W+1]
    uint32_t* cLL;   // corners [srows+1][scols+1]
    uint32_t* cLR;
    uint32_t* cRL;
    uint32_t* cRR;
    uint32_t* out;   // per pixel
    // per work-group partials of colR's column 0 and rowB's row 0 (the
    // cells nearly every triangle adds to): strip_hot_groups(P) rows of
    // H + 1 + W + 1 words, or null (then they take global atomics)
    uint32_t* hot_part;
};

void launch_vertex(const VSParams& P, hipStream_t s);
// clip space -> screen space (finish) for n vertices with K varyings
void launch_finish(const VertexBuf& in, const VertexBuf& out, uint32_t n, int K, const sr_viewport& vp,
                   hipStream_t s);
void launch_gs_count(int mode, int prim, const GSParams& P, hipStream_t s);
void launch_gs_emit(int mode, int prim, const GSParams& P, hipStream_t s);
// normal visualisation on a triangle list: lines (2 vertices each) into P.out,
// 3 per triangle (SR_GS_VERTEX_NORMALS) or 1 (SR_GS_FACE_NORMALS)
void launch_gs_normals(int mode, const GSParams& P, hipStream_t s);
void launch_tri_setup(const SetupParams& P, hipStream_t s);
void launch_tri_emit(const EmitParams& P, hipStream_t s);
void launch_vertex_batch(const BatchTables& T, const VertexBuf& out, hipStream_t s);
void launch_tri_setup_batch(const BatchTables& T, const VertexBuf& v, TriRec* recs, TriBin* bins, TriVid* vids,
                            uint32_t* counts, uint32_t* err, hipStream_t s);
void launch_tri_emit_batch(const BatchTables& T, const TriRec* recs, const TriBin* bins, const uint32_t* counts,
                           const uint32_t* offsets, uint32_t cap, uint32_t* keys, uint32_t* vals,
                           uint32_t* overflow, hipStream_t s);
uint32_t scan_tmp_words(uint32_t n);
// exclusive scan of in[0..n) (or of the first *count elements, count <= n)
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, uint32_t* total,
                 hipStream_t s, const uint32_t* count = nullptr);
uint32_t radix_hist_words(uint32_t n);
// returns 1 if the sorted data ended up in the *_tmp buffers
// n_dev (optional): the element count in device memory, at most n; then
// total[1] is used as scratch
int radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
               int key_bits, uint32_t* hist, uint32_t* scan_tmp, uint32_t* total, hipStream_t s,
               const uint32_t* n_dev = nullptr);

// Segmented sort of a batch's (execution tile, triangle) pairs: the pairs of
// each framebuffer group are already contiguous and in group order (they
// are emitted triangle by triangle), so each group's run is sorted apart by
// its local tile (key - the group's first tile).  That takes the passes the
// largest group's tile count needs instead of those of the batch's tile
// count.  `segs` holds 4 * ngroups + 2 words of scratch; the histograms
// need radix_hist_words(cap) + 256 * (ngroups + 1) words.  Sorts the first
// *live of `cap` pairs; returns 1 if the result is in the *_tmp buffers.
struct BatchGroups {
    const uint32_t* tend;            // per draw: inclusive prefix of triangles
    const uint32_t* group_first;     // per group: first draw (+1 entry)
    const uint32_t* group_tile_end;  // per group: end of its execution tiles
    uint32_t ngroups, ntris;
};
int radix_sort_groups(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t cap,
                      int local_key_bits, const BatchGroups& G, const uint32_t* tri_offsets,
                      const uint32_t* live, uint32_t* segs, uint32_t* hist, uint32_t* scan_tmp,
                      uint32_t* total, hipStream_t s);
void launch_fill(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s);
// start/end of tiles [0, nkeys) in keys[0, n) (sorted), by binary search:
// every tile written, no clearing needed.  n_dev: the key count on the
// device (at most n), when only the device knows it.
void launch_ranges_search(const uint32_t* keys, uint32_t n, uint32_t nkeys, uint32_t* start, uint32_t* end,
                          hipStream_t s, const uint32_t* n_dev = nullptr);
// p[i] = pat[i % period] over n_words 32-bit words (period 1, 2 or 4)
void launch_fill_pattern(void* p, const uint32_t pat[4], uint32_t period, size_t n_words, hipStream_t s);
void launch_iota(uint32_t* p, uint32_t n, hipStream_t s);  // p[i] = i
// start/end of each key's run in sorted keys; keys outside [key_base,
// key_base + nkeys) are skipped
void launch_ranges(const uint32_t* keys, uint32_t n, uint32_t key_base, uint32_t nkeys, uint32_t* start,
                   uint32_t* end, hipStream_t s, const uint32_t* n_dev = nullptr);
// misc[live_word] = misc[overflow_word] ? 0 : misc[total_word]
void launch_live_pairs(uint32_t* misc, int overflow_word, int total_word, int live_word, hipStream_t s);
// misc[4] = misc[1] > cap; misc[5] = pairs to sort (0 on overflow or a NaN flag in misc[0])
void launch_bins_live(uint32_t* misc, uint32_t cap, hipStream_t s);
// persistent-raster partition: cost[t] = fixed + items of tile t for
// t < n, cost[n] = 0 (scan it into a prefix of n + 1 entries), then G runs
// of near-equal cost over tiles [tile0, tile0 + n): wg_start[0..G]
void launch_tile_cost(const uint32_t* start, const uint32_t* end, uint32_t n, uint32_t fixed, uint32_t* cost,
                      hipStream_t s);
void launch_partition(const uint32_t* prefix, uint32_t tile0, uint32_t n, uint32_t G, uint32_t* wg_start,
                      hipStream_t s);
void launch_fb_clear(uint32_t* color, float* depth, uint8_t* stencil, uint32_t n, uint32_t rgba,
                     float d, uint32_t st, hipStream_t s);
// grid = ntiles work-groups (inline scene) or a persistent grid (scene table);
// the batch's `one` carries the launch-wide program choice (fp, blend, smode, ets)
void launch_raster_tris(const RasterBatch& P, uint32_t grid, hipStream_t s);
// The raster for other framebuffer formats, the multi-output program and the
// normalised over-blend (sr_raster_fmt.hip): 8x8 execution tiles.
void launch_raster_fmt(const RasterParams& P, const FmtParams& F, uint32_t ntiles, hipStream_t s);
void launch_raster_tris_8(const RasterBatch& P, uint32_t grid, hipStream_t s);
void launch_raster_tris_16(const RasterBatch& P, uint32_t grid, hipStream_t s);
void launch_raster_tris_32(const RasterBatch& P, uint32_t grid, hipStream_t s);

// lines / points / stencil strips
void launch_line_tiles(const LineSetupParams& P, int pass, hipStream_t s);
void launch_line_walk(const WalkParams& P, hipStream_t s);
void launch_item_emit(const ItemEmitParams& P, hipStream_t s);
void launch_point_setup(const PointSetupParams& P, hipStream_t s);
void launch_raster_lines(const RasterParams& P, uint32_t ntiles, hipStream_t s);
void launch_raster_points(const RasterParams& P, uint32_t ntiles, hipStream_t s);
void launch_strips(const StripParams& P, hipStream_t s);
// Work-groups of the strip pass when colR's column 0 and rowB's row 0 are
// accumulated in LDS (P.hot_part then needs that many rows of H + W + 2
// words), or 0 when they do not fit.
uint32_t strip_hot_groups(const StripParams& P);

}  // namespace sr
