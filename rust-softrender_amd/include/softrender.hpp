// softrender.hpp — C++ host layer over the C ABI (include/softrender_hip.h)
// that keeps the reference crate's builder chain:
//
//   Pipeline::from_framebuffer / with_framebuffer   pipeline/mod.rs:110-140
//   Pipeline::render_mesh(prim, mesh, stencil)      pipeline/mod.rs:146-155
//   VertexShader::run / run_to_fragment             pipeline/stages/vertex.rs:87-160
//   GeometryShader::run / clip_primitives / finish  pipeline/stages/geometry.rs:60-336
//   FragmentShader::with_faces_culled / with_antialiased_lines / with_tile_size /
//                   with_blend / run                pipeline/stages/fragment.rs:82-319
//   Framebuffer::clear, pixel/depth/stencil reads   framebuffer/renderbuffer/mod.rs
//
// Shader closures become program ids plus #[repr(C)]-style uniform blocks
// (VertexProgram, GeometryProgram, FragmentProgram below).  Where the
// reference panics (assert!/unwrap), these classes throw softrender::Error
// carrying the C status code.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/softrender_hip.h"

namespace softrender {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(sr_context ctx, int rc) {
    if (rc != SR_OK) throw Error(rc, std::string("softrender: ") + (ctx ? sr_last_error(ctx) : "error"));
}

// primitive.rs: zero-sized tags
struct Point { static constexpr int32_t id = SR_PRIM_POINTS; };
struct Line { static constexpr int32_t id = SR_PRIM_LINES; };
struct Triangle { static constexpr int32_t id = SR_PRIM_TRIANGLES; };

enum class FaceWinding { Clockwise = SR_CULL_CW, CounterClockwise = SR_CULL_CCW };  // winding.rs

struct Dimensions {
    uint32_t width, height;
};

struct Viewport {  // clipvertex.rs:40-64
    float x = 0, y = 0, width = 0, height = 0, near_z = 0, far_z = 1;
    static Viewport from_dimensions(Dimensions d, float near_z = 0.0f, float far_z = 1.0f) {
        return Viewport{0.0f, 0.0f, (float)d.width, (float)d.height, near_z, far_z};
    }
};

enum class StencilTest { Always = SR_STENCIL_ALWAYS, Never = SR_STENCIL_NEVER, LessThan = SR_STENCIL_LESS,
                         GreaterThan = SR_STENCIL_GREATER, LessThanEq = SR_STENCIL_LEQUAL,
                         GreaterThanEq = SR_STENCIL_GEQUAL, Equal = SR_STENCIL_EQUAL,
                         NotEqual = SR_STENCIL_NOTEQUAL };
struct StencilOp {  // stencil.rs:128-142
    int32_t op = SR_STENCIL_KEEP;
    bool wrap = false;
    static StencilOp Keep() { return {SR_STENCIL_KEEP, false}; }
    static StencilOp Invert() { return {SR_STENCIL_INVERT, false}; }
    static StencilOp Zero() { return {SR_STENCIL_ZERO, false}; }
    static StencilOp Replace() { return {SR_STENCIL_REPLACE, false}; }
    static StencilOp Increment(bool wrap) { return {SR_STENCIL_INCR, wrap}; }
    static StencilOp Decrement(bool wrap) { return {SR_STENCIL_DECR, wrap}; }
};
struct StencilConfig {  // GenericStencilConfig, stencil.rs:177-199
    StencilOp op = StencilOp::Keep();
    StencilTest test = StencilTest::Always;
};

// ---- programs ("closures") ------------------------------------------------
struct VertexProgram {
    int32_t id = SR_VP_MVP_NORMAL_UV;
    std::array<float, 16> mvp{};        // row-major
    std::array<float, 9> normal{};      // row-major 3x3
    std::array<float, 12> model{};      // rows 0..2 of the model matrix
    static VertexProgram mvp_normal_uv(const std::array<float, 16>& m) {
        VertexProgram p;
        p.id = SR_VP_MVP_NORMAL_UV;
        p.mvp = m;
        p.normal = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        p.model = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        return p;
    }
    static VertexProgram mvp_normal_world(const std::array<float, 16>& m) {
        VertexProgram p = mvp_normal_uv(m);
        p.id = SR_VP_MVP_NORMAL_WORLD;
        return p;
    }
};

enum class GeometryProgram { None = SR_GS_NONE, ClipReference = SR_GS_CLIP_REFERENCE, ClipNear = SR_GS_CLIP_NEAR,
                             VertexNormals = SR_GS_VERTEX_NORMALS, FaceNormals = SR_GS_FACE_NORMALS };

// Uniforms of the normal visualisation programs (full_example/src/shaders.rs:
// `projection * view` and NORMAL_LENGTH)
struct NormalLines {
    std::array<float, 16> projection_view{};  // row-major
    float length = 0.05f;                     // shaders.rs:33
};

struct FragmentProgram {
    int32_t id = SR_FP_FLAT;
    std::array<float, 32> uniforms{};
    int32_t filter = SR_FILTER_BILINEAR, edge = SR_EDGE_CLAMP;
    uint32_t border = 0;
    static FragmentProgram flat(float r, float g, float b, float a) {
        FragmentProgram p;
        p.uniforms[0] = r; p.uniforms[1] = g; p.uniforms[2] = b; p.uniforms[3] = a;
        return p;
    }
};

enum class Blend { Replace = SR_BLEND_REPLACE, Over = SR_BLEND_OVER };  // `()` / src-alpha over

// ---- resources ------------------------------------------------------------
class Context {
  public:
    explicit Context(int device = 0) { check(nullptr, sr_ctx_create(device, &h_)); }
    ~Context() { if (h_) sr_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    sr_context get() const { return h_; }
    void finish() const { check(h_, sr_finish(h_)); }

  private:
    sr_context h_ = nullptr;
};

// mesh.rs:12-64 — vertices (SoA position/normal/uv) plus indices
class Mesh {
  public:
    Mesh(std::shared_ptr<Context> ctx, int32_t primitive, const std::vector<float>& px,
         const std::vector<float>& py, const std::vector<float>& pz, const std::vector<uint32_t>& indices,
         const std::vector<float>* nx = nullptr, const std::vector<float>* ny = nullptr,
         const std::vector<float>* nz = nullptr, const std::vector<float>* u = nullptr,
         const std::vector<float>* v = nullptr)
        : ctx_(std::move(ctx)), primitive_(primitive) {
        sr_vertex_soa s{px.data(), py.data(), pz.data(), nx ? nx->data() : nullptr, ny ? ny->data() : nullptr,
                        nz ? nz->data() : nullptr, u ? u->data() : nullptr, v ? v->data() : nullptr};
        check(ctx_->get(), sr_mesh_upload(ctx_->get(), &s, (uint32_t)px.size(),
                                          indices.empty() ? nullptr : indices.data(),
                                          (uint32_t)indices.size(), primitive, &h_));
    }
    ~Mesh() { if (h_) sr_mesh_free(h_); }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    sr_mesh get() const { return h_; }
    int32_t primitive() const { return primitive_; }

  private:
    std::shared_ptr<Context> ctx_;
    int32_t primitive_;
    sr_mesh h_ = nullptr;
};

class Texture {
  public:
    Texture(std::shared_ptr<Context> ctx, uint32_t w, uint32_t h, const std::vector<uint32_t>& rgba)
        : ctx_(std::move(ctx)) {
        check(ctx_->get(), sr_texture_upload(ctx_->get(), w, h, rgba.data(), &h_));
    }
    ~Texture() { if (h_) sr_texture_free(h_); }
    Texture(const Texture&) = delete;
    Texture& operator=(const Texture&) = delete;
    sr_texture get() const { return h_; }

  private:
    std::shared_ptr<Context> ctx_;
    sr_texture h_ = nullptr;
};

// RenderBuffer with RGBA8 colour, f32 depth and u8 stencil, resident in HBM
class Framebuffer {
  public:
    Framebuffer(std::shared_ptr<Context> ctx, Dimensions d) : ctx_(std::move(ctx)), dim_(d) {
        check(ctx_->get(), sr_fb_create(ctx_->get(), d.width, d.height, &h_));
    }
    ~Framebuffer() { if (h_) sr_fb_free(h_); }
    Framebuffer(const Framebuffer&) = delete;
    Framebuffer& operator=(const Framebuffer&) = delete;
    Dimensions dimensions() const { return dim_; }
    // Framebuffer::clear(color): depth back to far (f32::MIN), stencil to 0
    void clear(uint32_t rgba) { check(ctx_->get(), sr_fb_clear(h_, rgba, -3.40282347e+38f, 0)); }
    void read(std::vector<uint32_t>* rgba, std::vector<float>* depth, std::vector<uint8_t>* stencil) const {
        const size_t n = (size_t)dim_.width * dim_.height;
        if (rgba) rgba->resize(n);
        if (depth) depth->resize(n);
        if (stencil) stencil->resize(n);
        check(ctx_->get(), sr_fb_read(h_, rgba ? rgba->data() : nullptr, depth ? depth->data() : nullptr,
                                      stencil ? stencil->data() : nullptr));
    }
    sr_framebuffer get() const { return h_; }

  private:
    std::shared_ptr<Context> ctx_;
    Dimensions dim_;
    sr_framebuffer h_ = nullptr;
};

// ---- the builder chain ----------------------------------------------------
class Pipeline;

class FragmentShader {
  public:
    FragmentShader with_faces_culled(std::optional<FaceWinding> w) && {
        desc_.cull = w ? (int32_t)*w : SR_CULL_NONE;
        return std::move(*this);
    }
    FragmentShader with_antialiased_lines(bool enable) && {
        desc_.aa_lines = enable ? 1 : 0;
        return std::move(*this);
    }
    FragmentShader with_tile_size(Dimensions d) && {
        desc_.tile_w = (int32_t)d.width;
        desc_.tile_h = (int32_t)d.height;
        return std::move(*this);
    }
    FragmentShader with_blend(Blend b) && {
        desc_.blend = (int32_t)b;
        return std::move(*this);
    }
    FragmentShader with_color_alpha_f32(bool f32_alpha) && {
        desc_.alpha_mode = f32_alpha ? SR_ALPHA_F32 : SR_ALPHA_U8;
        return std::move(*this);
    }
    // FragmentShader::run — consumes the stage, draws, returns nothing
    void run(const FragmentProgram& fp, const Texture* tex = nullptr) &&;

  private:
    friend class GeometryShader;
    friend class VertexShader;
    FragmentShader(Pipeline* p, std::shared_ptr<Mesh> m, sr_draw_desc d) : pipe_(p), mesh_(std::move(m)), desc_(d) {}
    Pipeline* pipe_;
    std::shared_ptr<Mesh> mesh_;
    sr_draw_desc desc_;
};

class GeometryShader {
  public:
    GeometryShader run(GeometryProgram gs) && {
        if (desc_.geometry_program != SR_GS_NONE && gs != GeometryProgram::None)
            throw Error(SR_ERR_UNSUPPORTED, "one device geometry program per draw");
        if (gs != GeometryProgram::None) desc_.geometry_program = (int32_t)gs;
        return std::move(*this);
    }
    GeometryShader run(GeometryProgram gs, const NormalLines& u) && {
        std::memcpy(desc_.gs_uniforms, u.projection_view.data(), 16 * sizeof(float));
        desc_.gs_uniforms[16] = u.length;
        return std::move(*this).run(gs);
    }
    GeometryShader clip_primitives() && { return std::move(*this).run(GeometryProgram::ClipReference); }
    // duplicate (geometry.rs:43-58): a second chain over the same vertices
    GeometryShader duplicate() const& { return GeometryShader(pipe_, mesh_, desc_); }
    FragmentShader finish(const Viewport& vp) && {
        desc_.viewport = sr_viewport{vp.x, vp.y, vp.width, vp.height, vp.near_z, vp.far_z};
        return FragmentShader(pipe_, std::move(mesh_), desc_);
    }

  private:
    friend class VertexShader;
    GeometryShader(Pipeline* p, std::shared_ptr<Mesh> m, sr_draw_desc d) : pipe_(p), mesh_(std::move(m)), desc_(d) {}
    Pipeline* pipe_;
    std::shared_ptr<Mesh> mesh_;
    sr_draw_desc desc_;
};

class VertexShader {
  public:
    GeometryShader run(const VertexProgram& vp) && {
        set(vp);
        return GeometryShader(pipe_, std::move(mesh_), desc_);
    }
    FragmentShader run_to_fragment(const Viewport& v, const VertexProgram& vp) && {
        set(vp);
        desc_.viewport = sr_viewport{v.x, v.y, v.width, v.height, v.near_z, v.far_z};
        return FragmentShader(pipe_, std::move(mesh_), desc_);
    }

  private:
    friend class Pipeline;
    VertexShader(Pipeline* p, std::shared_ptr<Mesh> m, sr_draw_desc d) : pipe_(p), mesh_(std::move(m)), desc_(d) {}
    void set(const VertexProgram& vp) {
        desc_.vertex_program = vp.id;
        std::memcpy(desc_.vs_uniforms, vp.mvp.data(), 16 * sizeof(float));
        std::memcpy(desc_.vs_uniforms + 16, vp.normal.data(), 9 * sizeof(float));
        std::memcpy(desc_.vs_uniforms + 25, vp.model.data(), 12 * sizeof(float));
    }
    Pipeline* pipe_;
    std::shared_ptr<Mesh> mesh_;
    sr_draw_desc desc_;
};

class Pipeline {
  public:
    // Pipeline::from_framebuffer (pipeline/mod.rs:120-140): asserts non-zero size
    Pipeline(std::shared_ptr<Context> ctx, std::shared_ptr<Framebuffer> fb)
        : ctx_(std::move(ctx)), fb_(std::move(fb)) {
        if (fb_->dimensions().width == 0 || fb_->dimensions().height == 0)
            throw Error(SR_ERR_INVALID_ARGUMENT, "Framebuffer must have a non-zero size");
    }
    Framebuffer& framebuffer() { return *fb_; }
    StencilConfig& stencil_config_mut() { return stencil_; }
    // render_mesh (pipeline/mod.rs:146-155): the index count must be a
    // multiple of the primitive's vertex count (checked at upload)
    template <class Prim>
    VertexShader render_mesh(Prim, std::shared_ptr<Mesh> mesh, std::optional<uint8_t> stencil_value = {}) {
        if (mesh->primitive() != Prim::id) throw Error(SR_ERR_INVALID_ARGUMENT, "mesh primitive mismatch");
        sr_draw_desc d;
        std::memset(&d, 0, sizeof(d));
        d.tile_w = d.tile_h = 128;  // DEFAULT_TILE_SIZE, fragment.rs:29
        d.stencil_test = (int32_t)stencil_.test;
        d.stencil_op = stencil_.op.op;
        d.stencil_wrap = stencil_.op.wrap ? 1 : 0;
        d.stencil_ref = stencil_value.value_or(0);  // unwrap_or_default, pipeline/mod.rs:154
        d.alpha_mode = SR_ALPHA_U8;
        return VertexShader(this, std::move(mesh), d);
    }
    void draw(const sr_draw_desc& d, sr_mesh m, const Texture* tex) {
        check(ctx_->get(), sr_draw(ctx_->get(), fb_->get(), m, &d, tex ? tex->get() : nullptr));
    }

  private:
    std::shared_ptr<Context> ctx_;
    std::shared_ptr<Framebuffer> fb_;
    StencilConfig stencil_;
};

inline void FragmentShader::run(const FragmentProgram& fp, const Texture* tex) && {
    desc_.fragment_program = fp.id;
    std::memcpy(desc_.fs_uniforms, fp.uniforms.data(), sizeof(desc_.fs_uniforms));
    desc_.tex_filter = fp.filter;
    desc_.tex_edge = fp.edge;
    desc_.tex_border = fp.border;
    pipe_->draw(desc_, mesh_->get(), tex);
}

}  // namespace softrender
