/*
 * softrender_hip.h — C ABI of the MI355X rendering path.
 *
 * This is the drop-in boundary for softrender's draw path.  The reference is a
 * generic Rust builder chain whose "plugins" are monomorphised closures:
 *
 *   Pipeline::render_mesh            src/pipeline/mod.rs:146-155
 *   VertexShader::run / run_to_fragment
 *                                    src/pipeline/stages/vertex.rs:87-160
 *   GeometryShader::run / clip_primitives / finish
 *                                    src/pipeline/stages/geometry.rs:60-336
 *   FragmentShader::{with_faces_culled, with_antialiased_lines,
 *                    with_tile_size, with_blend, run}
 *                                    src/pipeline/stages/fragment.rs:82-319
 *   Framebuffer::clear / pixel + depth + stencil accessors
 *                                    src/framebuffer/renderbuffer/mod.rs:58-133
 *
 * Closures cannot run on the device, so each shader stage is a program id
 * plus a uniform block; the builder state (cull, AA lines, tile size, blend,
 * stencil config, stencil reference value) travels in sr_draw_desc.  One
 * sr_draw() is one `render_mesh(..).run(..)[.clip_primitives()].finish(..)
 * .run(..)` chain: the framebuffer afterwards holds exactly what the
 * reference's single-threaded pipeline would leave in it.
 *
 * Conventions: every function returns an int status (SR_OK or a negative
 * SR_ERR_*); no C++ exception crosses the ABI; handles are owned by the
 * library until their _free or _destroy call; there are no callbacks into the
 * host.  A context is externally synchronised (the reference's stages hold
 * `&mut Pipeline`, src/pipeline/stages/vertex.rs:30-35); separate contexts
 * may be used from separate threads, each with its own HIP stream.
 */
#ifndef SOFTRENDER_HIP_H
#define SOFTRENDER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SR_OK 0
#define SR_ERR_INVALID_ARGUMENT (-1)   /* reference: assert!/assert_eq! panics, pipeline/mod.rs:129-130,148 */
#define SR_ERR_OUT_OF_MEMORY (-2)
#define SR_ERR_DEVICE (-3)             /* a HIP runtime call failed */
#define SR_ERR_INVALID_COORDINATE (-4) /* reference: cast(..).unwrap() panics on NaN, triangle.rs:70, point.rs:47 */
#define SR_ERR_UNSUPPORTED (-5)
#define SR_ERR_BUDGET (-6)             /* serving: request does not fit the HBM budget */
#define SR_ERR_NOT_READY (-7)          /* serving: ticket still in flight */
#define SR_ERR_NO_DEVICE (-8)
#define SR_ERR_CANCELLED (-9)          /* serving: the server was destroyed first */
#define SR_ERR_WORKSPACE (-10)         /* sr_draw_static: the draw needs more than the static workspace */

/* ---- enums (values are part of the ABI) -------------------------------- */
/* primitive.rs:7-167: zero-sized type tags selecting index assembly */
#define SR_PRIM_POINTS 1
#define SR_PRIM_LINES 2
#define SR_PRIM_TRIANGLES 3

/* Vertex programs (stand-ins for the `Fn(&V,&U) -> ClipVertex` closure,
 * vertex.rs:87-89).  Both write clip = MVP * (p,1) and a normalised
 * world-space normal n' = normalize(N * n). */
#define SR_VP_MVP_NORMAL_UV 0    /* varyings: nx ny nz u v        (K=5) */
#define SR_VP_MVP_NORMAL_WORLD 1 /* varyings: nx ny nz wx wy wz   (K=6) */
/* The full example's vertex shader (full_example/src/shaders.rs:8-31):
 * clip = MVP * (p,1); varyings the world position model * (p,1) and the
 * world normal normalize(mit * (n,0)), both as 4-vectors (K=8: px py pz pw
 * nx ny nz nw).  model is vs_uniforms[25..37) (rows 0-2) and [40..44) (row
 * 3); mit (the model matrix's inverse transpose) is [16..25) (rows 0-2,
 * columns 0-2) and [37..40) (row 3, columns 0-2).  The 4-vector norm sums
 * ((x*x + y*y) + z*z) + w*w.  No geometry program takes it. */
#define SR_VP_WORLD4_NORMAL4 2
/* clip = MVP * (p,1) and one varying: the mesh's integer attribute as f32
 * (`R::from(i32)`, rounded to nearest), K=1.  Needs sr_mesh_set_int_attrib;
 * no geometry program takes it (clipping would cast at every clip vertex). */
#define SR_VP_MVP_INT 3

/* Geometry programs (GeometryShader::run, geometry.rs:132-258). */
#define SR_GS_NONE 0           /* run_to_fragment / finish without a program */
#define SR_GS_CLIP_REFERENCE 1 /* clip_primitives, geometry.rs:261-336, bit-exact */
#define SR_GS_CLIP_NEAR 2      /* Sutherland-Hodgman against the Near plane only */
/* Normal visualisation (the full example's programs, full_example/src/
 * shaders.rs:35-89): each triangle emits lines, drawn by the line raster;
 * other primitives pass through unchanged (re_emit).  They read the world
 * position and normal varyings of SR_VP_MVP_NORMAL_WORLD, gs_uniforms[0..16)
 * = projection * view (row-major) and gs_uniforms[16] = line length. */
#define SR_GS_VERTEX_NORMALS 3 /* per vertex: PV*(p,1) -> PV*((p,1) + (n,0)*len) */
#define SR_GS_FACE_NORMALS 4   /* per triangle, from the centroid c = a/3 + b/3 + c/3:
                                  PV*c.p -> PV*(c.p + normalize(c.n)*len) */

/* Fragment programs (the `Fn(&ScreenVertex,&U) -> Fragment` closure,
 * fragment.rs:168-169). */
#define SR_FP_FLAT 0        /* constant colour fs_uniforms[0..3] */
#define SR_FP_BLINN_PHONG 1 /* needs SR_VP_MVP_NORMAL_WORLD */
#define SR_FP_TEX_LAMBERT 2 /* needs SR_VP_MVP_NORMAL_UV and a texture */
#define SR_FP_UV_COLOR 3    /* r=u g=v b=fs[0] a=fs[1]; needs SR_VP_MVP_NORMAL_UV */
/* Several colour outputs at once (a texture buffer's colour tuple,
 * texturebuffer.rs:84-86,133-136), for framebuffers made with
 * sr_fb_create_ex; needs SR_VP_MVP_NORMAL_UV and a texture.  With n the
 * normalised interpolated normal, diff = max(n . L, 0) (L = fs[0..2]),
 * lit = fs[3] + fs[4] * diff and t the texture sample at (u, v):
 *   plane 0: (t.r * lit, t.g * lit, t.b * lit, t.a * fs[5])  (as TEX_LAMBERT)
 *   plane 1: (n.x * 0.5 + 0.5, n.y * 0.5 + 0.5, n.z * 0.5 + 0.5, 1)
 *   plane 2: (u, v, diff, 1)
 *   plane 3: (lit, lit, lit, 1)
 * A framebuffer with fewer planes keeps the first ones; every other
 * program writes its one colour to every plane. */
#define SR_FP_GBUFFER 4
/* The full example's fragment shader (full_example/src/shaders.rs:92-163,
 * color.rs:19-56): four point lights with Fresnel-Schlick (ior 1.45)
 * blended Lambert diffuse (albedo 0.7) and Blinn-Phong specular (NdotH^64),
 * inverse-square falloff, on a grey material decode_gamma(g, 2.2); then
 * ACES filmic tone mapping (A 2.51, B 0.03, C 2.43, D 0.59, E 0.14) and
 * encode_gamma(1/2.2).  Alpha 1.  Needs SR_VP_WORLD4_NORMAL4.
 * fs_uniforms: [0..3) camera position (w = 1); light i (0..3) at
 * [3+7i, 10+7i): position xyz (w = 1), colour rgb, intensity; [31] g.
 * Vectors are 4-vectors as in the example (dot and norm over 4
 * components); powi(x, n) multiplies by repeated squaring (x^5 = x * x^4).
 * Transcendentals (powf, sqrtf) are the platform's: colours agree with the
 * oracle within a few ulps, not bit for bit. */
#define SR_FP_FRESNEL4 5
/* The integer varying back as i32 (`<i32 as NumCast>::from`, truncation
 * toward zero; NaN and values out of i32's range, where the reference's
 * unwrap panics, saturate as Rust's `as` does), its index m = id mod 8 in
 * 0..7, and colour fs_uniforms[4m .. 4m+4) (RGBA floats).  Needs
 * SR_VP_MVP_INT.  Per-pixel raster. */
#define SR_FP_INT_PALETTE 6

/* Blend (color/blend.rs:9-74): `()` replaces, OVER is src-alpha over,
 * OVER_NORMALIZED the full example's blend (full_example/src/color.rs:5-17):
 * per colour channel (x*a + y*b*(1-a)) / (a + b*(1-a)), alpha a + b*(1-a),
 * with x, a the source and y, b the destination colour and alpha.  On RGBA8
 * planes every blend works on unorm floats and quantises; on RGBAF32 planes
 * it works on the floats as they are (no clamp, no quantisation). */
#define SR_BLEND_REPLACE 0
#define SR_BLEND_OVER 1
#define SR_BLEND_OVER_NORMALIZED 2

/* Framebuffer formats (sr_fb_create_ex).  Colour planes: Vector4<u8> and
 * RGBAf32Color = Vector4<f32> (color/predefined.rs:17,68-90); depth
 * (framebuffer/attachments/depth.rs:10-41): f32 or f64, far = the type's
 * minimum, compared as d >= dt after Depth::from_scalar (exact for f64);
 * stencil: u8 or `()` (stencil.rs:65-86: no test, no op). */
#define SR_COLOR_RGBA8 0
#define SR_COLOR_RGBAF32 1
#define SR_DEPTH_F32 0
#define SR_DEPTH_F64 1
#define SR_STENCIL_U8 0
#define SR_STENCIL_NONE 1
#define SR_MAX_COLOR_PLANES 4

/* Face culling (geometry/winding.rs:11-43; triangle.rs:54-61) */
#define SR_CULL_NONE 0
#define SR_CULL_CW 1
#define SR_CULL_CCW 2

/* StencilTest, stencil.rs:90-107 */
#define SR_STENCIL_ALWAYS 0
#define SR_STENCIL_NEVER 1
#define SR_STENCIL_LESS 2      /* ref <  fb */
#define SR_STENCIL_GREATER 3   /* ref >  fb */
#define SR_STENCIL_LEQUAL 4    /* ref <= fb */
#define SR_STENCIL_GEQUAL 5    /* ref >= fb */
#define SR_STENCIL_EQUAL 6
#define SR_STENCIL_NOTEQUAL 7
/* StencilOp, stencil.rs:128-142 */
#define SR_STENCIL_KEEP 0
#define SR_STENCIL_INVERT 1
#define SR_STENCIL_ZERO 2
#define SR_STENCIL_REPLACE 3
#define SR_STENCIL_INCR 4
#define SR_STENCIL_DECR 5

/* Alpha type of the colour attachment (color/mod.rs:22-32, helper.rs:31-73):
 * U8 is the library's Vector4<u8> (Alpha = u8: Wu coverage collapses to 0/1);
 * F32 is a #[repr(C)] RGBA8 colour whose Alpha is f32. */
#define SR_ALPHA_U8 0
#define SR_ALPHA_F32 1

/* texture.rs:20-40 */
#define SR_FILTER_NEAREST 0
#define SR_FILTER_BILINEAR 1
/* The full example's sampler (full_example/src/texture.rs:29-84): texel
 * coordinates u * (W - 1) (nearest: rounded half away from zero; bilinear:
 * + 0.5, floor, the next texel weighted by the fraction), edges Clamp
 * (u, v clamped to [0, 1]) or Wrap (u - trunc(u)); SR_EDGE_BORDER acts as
 * Clamp.  The texel is then decode_gamma(2.2)'d (r, g, b = c^2.2).  Texel
 * indices past the last texel, where the example would panic, use the
 * last.  Per-pixel raster (as RGBAF32 textures are). */
#define SR_FILTER_SPAN_NEAREST 2
#define SR_FILTER_SPAN_BILINEAR 3
#define SR_EDGE_CLAMP 0
#define SR_EDGE_WRAP 1
#define SR_EDGE_BORDER 2

/* sr_draw_desc.exec_flags */
#define SR_EXEC_EXHAUSTIVE 1u /* every triangle is tested in every screen tile (reference order, no pruning) */
#define SR_EXEC_TIMING 2u     /* record per-stage HIP events (read with sr_ctx_stage_times) */

/* ---- plain data -------------------------------------------------------- */
typedef struct sr_viewport {
    /* clipvertex.rs:40-64 Viewport{x,y,width,height,near,far} */
    float x, y, width, height, near_z, far_z;
} sr_viewport;

/* Vertex streams, structure-of-arrays.  mesh.rs:12-64 keeps AoS
 * `Vec<V>` with a Point3 position plus user data; the device wants one
 * coalesced stream per component.  normal/uv streams may be NULL (zeros). */
typedef struct sr_vertex_soa {
    const float *px, *py, *pz;
    const float *nx, *ny, *nz;
    const float *u, *v;
} sr_vertex_soa;

typedef struct sr_draw_desc {
    int32_t vertex_program;   /* SR_VP_* */
    int32_t geometry_program; /* SR_GS_* */
    int32_t fragment_program; /* SR_FP_* */
    int32_t blend;            /* SR_BLEND_* */
    int32_t cull;             /* SR_CULL_*, fragment.rs:82-91 */
    int32_t aa_lines;         /* Xiaolin Wu if non-zero, fragment.rs:96-105 */
    int32_t tile_w, tile_h;   /* screen tile, fragment.rs:29,107-116 (default 128x128) */
    int32_t stencil_test;     /* SR_STENCIL_* test, StencilConfig::get_test */
    int32_t stencil_op;       /* SR_STENCIL_* op,   StencilConfig::get_op */
    int32_t stencil_wrap;     /* Increment/Decrement { wrap } */
    uint32_t stencil_ref;     /* render_mesh(.., Some(v)), pipeline/mod.rs:154 (u8) */
    int32_t alpha_mode;       /* SR_ALPHA_* */
    int32_t tex_filter;       /* SR_FILTER_* */
    int32_t tex_edge;         /* SR_EDGE_* */
    uint32_t tex_border;      /* RGBA8 for SR_EDGE_BORDER */
    sr_viewport viewport;
    /* vs_uniforms: [0..16) MVP row-major, [16..25) normal matrix (3x3,
     * row-major), [25..37) model matrix rows 0..2 (3x4, row-major). */
    float vs_uniforms[48];
    /* fs_uniforms: program specific, documented with each program. */
    float fs_uniforms[32];
    /* gs_uniforms: geometry program uniforms (normal visualisation). */
    float gs_uniforms[20];
    int32_t exec_tile;   /* execution tile edge in pixels: 0 = auto, else 8/16/32 */
    uint32_t exec_flags; /* SR_EXEC_* */
    /* waves (64 lanes each) sharing one execution tile: 0 = auto, else 1, 2
     * or 4 where the tile size has it (8: 1; 16: 1, 2, 4; 32: 2, 4),
     * the tile size's default (4 from 16 up) otherwise */
    int32_t exec_waves;
    int32_t reserved[5];
} sr_draw_desc;

typedef struct sr_draw_stats {
    uint64_t triangles_in, triangles_setup, lines_in, points_in;
    uint64_t bin_pairs;  /* (item, execution tile) pairs */
    uint64_t generated_tris, generated_lines, generated_points;
    /* 1 when some triangle could pass the coverage test outside its own
     * box (a sliver, through f32 rounding): it was binned to every
     * execution tile such a pixel may lie in and tested in its box clamped
     * into each screen tile, as triangle.rs:66-99 does */
    uint64_t strip_slivers;
} sr_draw_stats;

/* ---- opaque handles ----------------------------------------------------- */
typedef struct sr_context_t* sr_context;
typedef struct sr_mesh_t* sr_mesh;
typedef struct sr_texture_t* sr_texture;
typedef struct sr_framebuffer_t* sr_framebuffer;
typedef struct sr_server_t* sr_server;

/* ---- context ------------------------------------------------------------ */
/* Pipeline::new (pipeline/mod.rs:110-117) creates the CPU pool; here the
 * context owns a HIP stream and a grow-only device workspace. */
int sr_ctx_create(int device, sr_context* out);
int sr_ctx_destroy(sr_context ctx);
const char* sr_last_error(sr_context ctx);
int sr_finish(sr_context ctx); /* wait for all queued work */
void* sr_ctx_stream(sr_context ctx); /* the context's hipStream_t */
/* Per-stage device times (ms) of the last draw run with SR_EXEC_TIMING.
 * names: "vertex","geometry","setup","scan","emit","sort","ranges","strips","raster" */
int sr_ctx_stage_times(sr_context ctx, float* ms_out, int n_max, int* n_out);
int sr_ctx_last_stats(sr_context ctx, sr_draw_stats* out);

/* ---- mesh (mesh.rs:12-64) ---------------------------------------------- */
/* indices == NULL means non-indexed (index i = i); a non-NULL list with
 * n_indices == 0 is an empty Mesh::indices and draws nothing.  n_indices
 * must be a multiple of the primitive's vertex count (assert_eq!,
 * pipeline/mod.rs:148). */
int sr_mesh_upload(sr_context ctx, const sr_vertex_soa* vertices, uint32_t n_vertices,
                   const uint32_t* indices, uint32_t n_indices, int32_t primitive,
                   sr_mesh* out);
int sr_mesh_free(sr_mesh mesh);
/* An integer vertex attribute (an i32 member of the vertex's uniforms,
 * interpolated through f32 and cast back: interpolate.rs:43-64), one value
 * per vertex (count == the mesh's n_vertices); replaces an earlier one.
 * SR_VP_MVP_INT reads it. */
int sr_mesh_set_int_attrib(sr_mesh mesh, const int32_t* values, uint32_t count);

/* ---- texture (texture.rs:14-62; sampling is build-defined) -------------- */
int sr_texture_upload(sr_context ctx, uint32_t width, uint32_t height,
                      const uint32_t* rgba8, sr_texture* out);
int sr_texture_free(sr_texture tex);
/* Render to texture: a texture that reads the framebuffer's colour plane
 * (RGBA8, row-major, no copy), for a later draw into another framebuffer.
 * It stays valid while the framebuffer lives and sees its current contents;
 * a draw may not sample the framebuffer it draws into
 * (SR_ERR_INVALID_ARGUMENT).  Replaces a texture buffer declared with
 * declare_texture_buffer! and read through TextureRead::sample
 * (src/framebuffer/texturebuffer.rs:15-231, src/texture.rs:14-62). */
int sr_texture_from_fb(sr_framebuffer fb, sr_texture* out);

/* ---- framebuffer (renderbuffer/mod.rs:14-133) --------------------------- */
/* Width and height must be non-zero (pipeline/mod.rs:129-130).  Planes are
 * row-major, pixel (x,y) at x + y*W (geometry/coordinate.rs:47-52). */
int sr_fb_create(sr_context ctx, uint32_t width, uint32_t height, sr_framebuffer* out);
int sr_fb_free(sr_framebuffer fb);
/* Framebuffer::clear sets colour and resets depth to far (f32::MIN) and
 * stencil to 0 (renderbuffer/mod.rs:125-133); depth/stencil are explicit here. */
int sr_fb_clear(sr_framebuffer fb, uint32_t rgba, float depth, uint8_t stencil);
/* Any of the output pointers may be NULL. */
int sr_fb_read(sr_framebuffer fb, uint32_t* rgba, float* depth, uint8_t* stencil);
int sr_fb_write(sr_framebuffer fb, const uint32_t* rgba, const float* depth,
                const uint8_t* stencil);
int sr_fb_size(sr_framebuffer fb, uint32_t* width, uint32_t* height);

/* Framebuffers of other formats, and texture buffers: `n_color` colour
 * planes stored apart (declare_texture_buffer!, texturebuffer.rs:72-198:
 * one Vec per colour field, so each can be sampled without a copy), each
 * RGBA8 (4 bytes a pixel, r | g<<8 | b<<16 | a<<24) or RGBAF32 (16 bytes:
 * r, g, b, a floats), plus depth and stencil.  sr_fb_create's framebuffer
 * is {1, {RGBA8}, F32, U8}.  The functions above that take one RGBA8 colour
 * (sr_fb_clear, sr_fb_read, sr_fb_write, sr_texture_from_fb) apply to
 * plane 0 when it is RGBA8 and the depth is f32.  Triangle draws (with or
 * without a geometry program) render into any format; line and point draws
 * need sr_fb_create's format (SR_ERR_UNSUPPORTED otherwise). */
typedef struct sr_fb_format {
    uint32_t n_color;                      /* 1..SR_MAX_COLOR_PLANES */
    int32_t color[SR_MAX_COLOR_PLANES];    /* SR_COLOR_* of each plane */
    int32_t depth;                         /* SR_DEPTH_* */
    int32_t stencil;                       /* SR_STENCIL_* */
} sr_fb_format;
int sr_fb_create_ex(sr_context ctx, uint32_t width, uint32_t height, const sr_fb_format* format,
                    sr_framebuffer* out);
int sr_fb_get_format(sr_framebuffer fb, sr_fb_format* out);
/* Framebuffer::clear (renderbuffer/mod.rs:125-133, texturebuffer.rs:
 * 183-198): every plane to its colour, depth to far, stencil to 0.
 * `colors` holds each plane's colour in its format (4 or 16 bytes), packed
 * in plane order. */
int sr_fb_clear_ex(sr_framebuffer fb, const void* colors);
/* Copy planes out / in: planes[i] (NULL to skip) is w*h pixels of plane i's
 * format; depth is w*h floats or doubles; any pointer may be NULL. */
int sr_fb_read_ex(sr_framebuffer fb, void* const* planes, void* depth, uint8_t* stencil);
int sr_fb_write_ex(sr_framebuffer fb, const void* const* planes, const void* depth, const uint8_t* stencil);
/* A texture reading colour plane `plane` of the framebuffer (RGBA8 or
 * RGBAF32 texels, no copy): TextureBufferRef, texturebuffer.rs:12-49. */
int sr_texture_from_fb_plane(sr_framebuffer fb, uint32_t plane, sr_texture* out);

/* ---- ingestion: Wavefront OBJ -------------------------------------------
 * The reference's examples load meshes with tobj 0.1.3's load_obj
 * (examples/suzanne.rs:23-47): each distinct v/vt/vn corner becomes a
 * vertex (in order of first use), polygons become triangle fans from their
 * first corner, and an `o` statement starts a new model. */
typedef struct sr_obj_info {
    uint32_t models;      /* models in the file */
    uint32_t vertices;    /* of the requested model */
    uint32_t indices;     /* 3 per triangle */
    int32_t has_normals, has_uvs;
} sr_obj_info;
/* Host-side parse (no device work).  With null arrays only `info` is
 * filled; otherwise positions (3 floats per vertex), normals (3, when
 * present), uvs (2, when present) and indices are copied out. */
int sr_obj_parse(const char* path, uint32_t model, sr_obj_info* info, float* positions, float* normals,
                 float* uvs, uint32_t* indices);
/* Parses model `model` and uploads it as an indexed triangle mesh. */
int sr_mesh_load_obj(sr_context ctx, const char* path, uint32_t model, sr_mesh* out);

/* ---- draw --------------------------------------------------------------- */
/* One reference draw chain (see top of file).  `tex` may be NULL unless the
 * fragment program samples.  Asynchronous on the context stream except
 * where the workspace has to grow. */
int sr_draw(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* desc,
            sr_texture tex);

/* Size of the static device workspace of sr_draw_static, per GPU. */
#define SR_STATIC_WORKSPACE_MIB 64
/* sr_draw without any device allocation: no counterpart in the reference
 * (whose draws allocate host vectors freely); it is for callers that cannot
 * allocate device memory during a draw (code under stream or graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and result as sr_draw, but every workspace buffer (screen
 * vertices, triangle records, bin pairs, sort buffers, stencil strip
 * counts, line plots) is carved from one fixed region of
 * SR_STATIC_WORKSPACE_MIB MiB that the library holds on each GPU from its
 * first context there until the process ends; the caller's framebuffer,
 * mesh and texture are its only other memory.  A draw that needs more
 * returns SR_ERR_WORKSPACE (sr_last_error names the limit) with the
 * framebuffer unchanged; at 1920x1080 that is roughly 200k triangles.
 * Every call on one GPU shares that GPU's region: calls from different
 * contexts on the same GPU that overlap in time race (calls on one context
 * run in order on its stream).  The context's own workspace is neither
 * used nor changed, so sr_draw keeps working as before. */
int sr_draw_static(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* desc,
                   sr_texture tex);

/* ---- batching: many independent draws in one persistent launch ---------- */
/* Several draw chains over one mesh that share its vertex stage
 * (GeometryShader::duplicate, src/pipeline/stages/geometry.rs:43-58; the
 * full example draws a model lit and then its normals from one vertex
 * pass, full_example/src/main.rs:72-91).  The vertex program runs once, in
 * clip space; every chain then runs its own geometry program (or finish),
 * raster and fragment stages, in order.  Every desc must name the same
 * vertex program and vs_uniforms.  Equal, chain by chain, to sr_draw calls
 * in order.  sr_ctx_last_stats then describes the last chain. */
int sr_draw_chains(sr_context ctx, sr_framebuffer fb, sr_mesh mesh, const sr_draw_desc* descs, uint32_t n,
                   sr_texture tex);

typedef struct sr_batch_item {
    sr_framebuffer fb;
    sr_mesh mesh;
    const sr_draw_desc* desc;
    sr_texture tex;
    int32_t clear;       /* non-zero: Framebuffer::clear(clear_rgba) first (depth far, stencil 0) */
    uint32_t clear_rgba;
} sr_batch_item;
/* Renders n (framebuffer, mesh) draws.  Results equal, item by item in
 * order, an sr_fb_clear (if `clear`) followed by sr_draw.  A framebuffer may
 * appear several times: its items are then drawn into it in item order,
 * like several render_mesh chains on one pipeline (the multi-mesh case of
 * src/lib.rs:41-51).  Triangle draws without a geometry program and with
 * stencil Always/Keep run in one pass over the whole batch (captured in a
 * hipGraph and replayed while the batch is unchanged) when every item on
 * their framebuffer qualifies, shares execution tile, program and blend,
 * and only the first of them clears; other framebuffers' items are drawn
 * one by one (before the pass).  No item may sample a framebuffer the batch
 * draws into.  A NaN screen coordinate in any item of the one-pass group
 * returns SR_ERR_INVALID_COORDINATE (as sr_draw does) with none of that
 * group's framebuffers changed, its clears included.  Synchronous. */
int sr_batch_draw(sr_context ctx, const sr_batch_item* items, uint32_t n);

/* sr_batch_draw without waiting (the multi-mesh submission src/lib.rs:41-51
 * asks for, decoupled from the host): items that cannot join the one-pass
 * group are drawn before it returns, the one-pass group is queued on one
 * of the context's batch streams (after the work already queued on the
 * context's stream) and `ticket` names it.  Until sr_batch_wait returns for
 * that ticket the batch's framebuffers, meshes and textures belong to it:
 * no other call may read or write them.  Up to four batches are in flight;
 * a fifth submit first completes the oldest.  Items and descriptors are
 * copied. */
int sr_batch_submit(sr_context ctx, const sr_batch_item* items, uint32_t n, uint64_t* ticket);
/* SR_OK when the ticket's batch has finished on the device (sr_batch_wait
 * will not block on it), SR_ERR_NOT_READY while it runs. */
int sr_batch_query(sr_context ctx, uint64_t ticket);
/* Blocks until the ticket's batch is complete (reruns after a bin
 * overflow included) and returns its status;
 * afterwards the ticket is unknown. */
int sr_batch_wait(sr_context ctx, uint64_t ticket);

/* Execution tile used when sr_draw_desc.exec_tile == 0: entry i applies to
 * framebuffers of at most max_pixels[i] pixels (first match wins).  Set by
 * the autotuner (rust-softrender_amd/autotune.py). */
int sr_ctx_set_exec_tile_policy(sr_context ctx, uint32_t n, const uint64_t* max_pixels,
                                const int32_t* exec_tile);
/* The same with the waves per tile of each entry (exec_waves NULL: each
 * tile size's default), used when sr_draw_desc.exec_tile == 0. */
int sr_ctx_set_exec_policy(sr_context ctx, uint32_t n, const uint64_t* max_pixels, const int32_t* exec_tile,
                           const int32_t* exec_waves);
/* Initial (execution tile, triangle) pair capacity of the batch path; it
 * grows by itself when a batch overflows it. */
int sr_ctx_set_batch_capacity(sr_context ctx, uint32_t pairs);

/* ---- serving: a request loop with continuous batching -------------------
 * No counterpart in the reference (it renders one frame per call on the
 * calling thread, pipeline/mod.rs:115-155); this is the serving layer built
 * over the same draw.  Meshes and textures stay resident in the context
 * that created them; the server owns a pool of device framebuffers and of
 * pinned host result buffers.  A dispatcher thread admits pending requests
 * in arrival order while their framebuffers plus bin workspaces fit the HBM
 * budget, and whenever one of `max_inflight` batch slots is free it renders
 * everything admitted so far as one batch on that slot's stream, so work
 * that arrives while other batches run joins the next one and overlaps on
 * the device.  Every result equals an sr_fb_clear + sr_draw of the same
 * request on its own.  submit / poll / release / stats are thread-safe. */
typedef struct sr_server_config {
    uint64_t hbm_budget;       /* bytes of framebuffers + bin workspaces in flight; 0: 16 GiB */
    uint32_t max_inflight;     /* batches on the device at once, one stream each; 0: 4 */
    uint32_t max_batch_frames; /* requests per batch; 0: 32 */
    uint64_t max_batch_tris;   /* triangles per batch (a larger request goes alone); 0: 32M */
    float pairs_per_tri;       /* bin pairs per triangle used to size workspaces; 0: 4 */
    int32_t reserved[7];
} sr_server_config;

typedef struct sr_request {
    sr_mesh mesh;              /* triangles; from the server's context */
    sr_texture tex;            /* NULL unless the fragment program samples */
    sr_draw_desc desc;         /* copied; triangle draw without geometry program, stencil Always/Keep */
    uint32_t width, height;    /* framebuffer size */
    uint32_t clear_rgba;       /* the frame starts as Framebuffer::clear(clear_rgba) */
    int32_t want_depth;        /* non-zero: the depth plane is returned as well */
    uint64_t tag;              /* caller's id, returned in the completion */
    double arrival;            /* sr_server_clock() time of arrival; 0: the submit time */
} sr_request;

typedef struct sr_completion {
    uint64_t ticket, tag;
    int32_t status;            /* SR_OK or SR_ERR_* */
    uint32_t width, height;
    uint32_t batch_frames;     /* requests in the batch that drew this one */
    uint64_t batch_id;
    const uint32_t* rgba;      /* pinned host memory, valid until sr_server_release(ticket) */
    const float* depth;        /* NULL unless want_depth */
    double t_arrival, t_admit, t_done; /* sr_server_clock() seconds; t_done = pixels on the host */
} sr_completion;

typedef struct sr_server_stats {
    uint64_t submitted, completed, failed, batches, overflow_reruns;
    uint64_t budget_bytes, in_flight_bytes, peak_in_flight_bytes, device_bytes;
    uint32_t pending, inflight_batches, max_batch_frames_seen, reserved;
} sr_server_stats;

int sr_server_create(sr_context ctx, const sr_server_config* cfg, sr_server* out);
/* Waits for batches on the device; requests not yet admitted complete with
 * SR_ERR_CANCELLED.  Result pointers die with the server. */
int sr_server_destroy(sr_server srv);
/* Queues one request (never blocks on the device).  SR_ERR_BUDGET when the
 * request alone exceeds the budget. */
int sr_server_submit(sr_server srv, const sr_request* req, uint64_t* ticket);
/* Up to `max` completions; waits up to timeout_ms (0: no wait, <0: forever)
 * for the first one. */
int sr_server_poll(sr_server srv, sr_completion* out, uint32_t max, uint32_t* n, int32_t timeout_ms);
/* Returns a completion's host buffers to the pool. */
int sr_server_release(sr_server srv, uint64_t ticket);
int sr_server_get_stats(sr_server srv, sr_server_stats* out);
const char* sr_server_last_error(sr_server srv);
double sr_server_clock(void); /* monotonic seconds */

#ifdef __cplusplus
}
#endif
#endif /* SOFTRENDER_HIP_H */
