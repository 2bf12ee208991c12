// sr_internal.hpp — types shared by the kernels (sr_kernels.hip) and the
// host runtime (sr_runtime.hip).  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/softrender_hip.h"

namespace sr {

constexpr int MAXK = 8;          // varyings per vertex
constexpr int SCAN_TILE = 4096;  // elements per scan / radix block (256 threads x 16)

// Screen- (or clip-) space vertices in one allocation: the position as
// four streams (x, y, z, w: the setup and binning passes read x, y, z
// coalesced), then the varyings interleaved per vertex, `kstride` floats a
// row (a shaded fragment reads each of its three vertices' varyings as one
// short contiguous row instead of one word from each of K streams).
//   comp(c)[i] = position component c of vertex i (c < 4)
//   var(i)[k]  = varying k of vertex i (k < kstride)
// stride = capacity in vertices.
struct VertexBuf {
    float* base;
    uint32_t stride;
    uint32_t kstride;
    __host__ __device__ float* comp(int c) const { return base + (size_t)c * stride; }
    __host__ __device__ float* var(uint32_t i) const { return base + 4 * (size_t)stride + (size_t)i * kstride; }
};

// Per-triangle setup record, written once by k_tri_setup and read by every
// raster workgroup whose tile the triangle touches: 64 bytes, one cache
// line, what the raster's per-triangle loop reads (into SGPRs).  The
// triangle's vertex ids, needed only to shade, are kept apart (TriVid).
struct alignas(64) TriRec {
    // bytes 0-47: what the raster tests every pixel with (staged into LDS
    // 16 bytes at a time); 48-63: the box, read where a mode needs it
    double R;             // RN64(1 / (double)det): f32 quotients are formed as (float)(num * R)
    uint32_t zs;          // bit 0 (strip): a pixel centre outside the box may pass the
                          // coverage test; the raster tests the box clamped into each
                          // screen tile (clamp_as_int!), as the reference does.
                          // Bits 1..31: depth_key(zhi) rounded up to even, zhi an
                          // upper bound of every depth the triangle can write
                          // (depth_bound_key)
    float x3, y3;         // screen position of the third vertex
    float e23y, e23x;     // (y2 - y3), (x3 - x2)   — triangle.rs:108
    float e31y, e31x;     // (y3 - y1), (x1 - x3)   — triangle.rs:109
    float z1, z2, z3;
    int32_t X0, X1, Y0, Y1;  // floor of the f32 bbox, clamped to [-1, W] x [-1, H]
};
static_assert(sizeof(TriRec) == 64, "TriRec layout");

// Order-preserving key of an f32 (a < b  <=>  key(a) < key(b) for all
// non-NaN a, b, with -0 just below +0); a NaN of either sign lies beyond
// the infinity of its sign.
__host__ __device__ __forceinline__ uint32_t depth_key(uint32_t bits) {
    return bits ^ ((uint32_t)((int32_t)bits >> 31) | 0x80000000u);
}

// The vertex ids of a triangle in the screen vertex buffer (the shading
// gathers' addresses).
struct TriVid {
    uint32_t i0, i1, i2;
};
static_assert(sizeof(TriVid) == 12, "TriVid layout");

// Per-triangle binning record: what pair emission reads (the pixel box
// whose execution tiles get the triangle and the coverable-tile mask),
// apart from the full record for culled boxes of more than 32 tiles.
// Kept apart from TriRec so emission streams 24 bytes per triangle
// instead of 88.
enum BinMode : uint32_t {
    BIN_BOX = 0,    // the tiles of the box (or of its mask, with culling)
    BIN_REACH = 1,  // every tile of X0..Y1: the box grown to the pixels that may pass
    BIN_CLAMP = 2,  // every tile meeting the box clamped into some screen tile
};
struct TriBin {
    int32_t X0, X1, Y0, Y1;  // the box (BIN_BOX, BIN_CLAMP) or the reach (BIN_REACH)
    uint32_t tmask;          // bit i = i-th execution tile of the box (row-major) may be
                             // covered; all ones when the box has over 32 tiles
    uint32_t mode;           // BinMode
};
static_assert(sizeof(TriBin) == 24, "TriBin layout");

// Per (line, screen tile) work item: the segment clipped to one screen
// tile by liang_barsky_iterative (geometry/line.rs:6-54), as rasterize_line
// receives it (line.rs:51-52).  80 bytes.
struct alignas(16) LineItem {
    uint32_t line;        // primitive index
    uint32_t tile;        // screen tile index (row-major, the reference's tile order)
    uint32_t plot_off;    // first plot record
    uint32_t plot_cnt;    // number of plot records
    float cx1, cy1;       // clipped start
    float cx2, cy2;       // clipped end
    float dlen;           // hypot of the clipped segment, line.rs:52
    int16_t bx0, by0, bx1, by1;  // execution tiles touched by its plots (bx0 > bx1: none)
    // the execution tiles that hold a plot, as bits (row-major, mcols per
    // row) of a window starting at tile (mbx, mby); mvalid = 0: use the box
    uint16_t mbx, mby, mcols, mvalid;
    // ranged = 1 (Wu walks that dropped or wrapped no plot): plots 0..3 are
    // the endpoints and plot 4 + 2i + {0, 1} lies on major coordinate
    // major0 + i (the pixel column, or row when steep)
    int32_t major0;
    uint16_t ranged, steep;
    uint32_t pad;
    uint64_t mlo, mhi;
};
static_assert(sizeof(LineItem) == 80, "LineItem layout");

// One plot of a line item: framebuffer index (x + y*W, line.rs:56-58), the
// plot coordinates (for the pixel centre in line.rs:72-73) and the
// coverage alpha in f64 as draw_line_xiaolin_wu computes it.
struct alignas(8) Plot {
    double alpha;
    uint32_t index;
    int32_t x, y;
    uint32_t row;  // index / W: the pixel's row (its column is index - row * W)
};
static_assert(sizeof(Plot) == 24, "Plot layout");

// Everything a raster workgroup needs, passed by value.
struct RasterParams {
    // framebuffer
    uint32_t W, H;
    uint32_t* color;
    float* depth;
    uint8_t* stencil;
    const uint32_t* strip_counts;  // per-pixel stencil visits outside bboxes (or null)
    // screen tiles of the reference (inclusive rectangles, fragment.rs:188-216)
    uint32_t tw, th;
    uint32_t scols, srows;
    // execution tiles
    uint32_t ets, ntx, nty;
    int nwv;                // waves sharing an execution tile (1, 2 or 4)
    const uint32_t* tile_start;
    const uint32_t* tile_end;
    const uint32_t* items;  // sorted item ids per execution tile
    // primitives
    const TriRec* tris;
    const TriVid* vids;     // per triangle, beside tris; null: triangle t's vertices are 3t, 3t+1, 3t+2
    const LineItem* line_items;
    const Plot* plots;
    VertexBuf verts;        // screen-space vertices
    const uint32_t* line_v; // 2 vertex ids per line
    const uint32_t* point_v;// vertex id per point
    uint32_t n_prims;
    int K;
    // programs
    int fp, blend, alpha_mode;
    int s_test, s_op, s_wrap;
    uint32_t s_ref;
    int smode;              // 0: stencil untouched, 1: count visits, 2: track in order
    int exhaustive;         // every triangle in every tile: the box is clamped into each
                            // screen tile exactly as clamp_as_int! does
    const uint32_t* tex;    // RGBA8 texels, or RGBAF32 (4 floats each) when tex_fmt says so
    int tex_fmt;            // SR_COLOR_*
    uint32_t texW, texH;
    int tex_filter, tex_edge;
    uint32_t tex_border;
    float fsu[32];
    // non-null: nothing is drawn when skip_if[0] bit 0 (a NaN screen
    // coordinate) or skip_if[4] (the bins overflowed) is set
    const uint32_t* skip_if;
    // batch: this scene's first global execution tile, and a fused clear
    // (Framebuffer::clear, renderbuffer/mod.rs:125-133) applied on load
    uint32_t tile_base;
    int clear;
    uint32_t clear_rgba;
};

// Framebuffers of other formats (sr_fb_create_ex): the colour planes, the
// depth plane of f64 depth, and how stencil behaves.
struct FmtParams {
    void* planes[SR_MAX_COLOR_PLANES];
    int color_fmt[SR_MAX_COLOR_PLANES];
    uint32_t n_color;
    double* depth64;      // f64 depth (else RasterParams::depth)
    int track_stencil;    // u8 stencil with a test or op: tracked per pixel in order
    int has_stencil;      // 0: `()` stencil, no plane
};

// Raster launch: one scene given inline (one work-group per tile), or a
// batch: a table of draws ("scenes") grouped by framebuffer, each group's
// draws consecutive and in item order; the groups' execution tiles are
// numbered globally [group_tile_end[g-1], group_tile_end[g]) and split
// into runs of near-equal cost, one per work-group of a persistent grid.
struct RasterBatch {
    RasterParams one;
    const RasterParams* scenes;      // null: use `one`
    const uint32_t* group_tile_end;  // inclusive prefix of tiles per group
    const uint32_t* group_first;     // first draw of each group; [ngroups] = draw count
    const uint32_t* tri_end;         // per draw: exclusive end of its batch triangle ids
    uint32_t ngroups;
    uint32_t tile0, ntiles;          // global tile range of this launch
    const uint32_t* wg_start;        // run p = tiles [tile0 + wg_start[p], tile0 + wg_start[p+1])
    const uint32_t* overflow;        // [0] binning overflowed, [2] bit 0: a NaN screen coordinate:
                                     // either draws nothing (no fused clear either)
};

}  // namespace sr
