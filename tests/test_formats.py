"""Framebuffer formats and texture buffers (sr_fb_create_ex).

RGBA8 and RGBAF32 colour planes (color/predefined.rs:17,68-90), several
planes at once written by one fragment program (declare_texture_buffer!,
texturebuffer.rs:72-198), f64 depth, the `()` stencil (stencil.rs:65-86),
the full example's normalised over-blend (full_example/src/color.rs:5-17)
and render-to-texture from an RGBAF32 plane.

CPU: the oracle against closed forms — the blends against their formulas in
numpy float32 (Rust's left-to-right order), the multi-output program at
pixels whose barycentrics are exactly (1, 0, 0), f64 depth and the `()`
stencil against the f32 / u8 framebuffer, and an RGBAF32 texture read at
texel centres.
GPU: the device against the oracle, bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import scenes
import softrender_hip as sr
from oracle_checks import screen_tri_mesh
from parity import assert_parity, render_ref

f32 = np.float32
RGBA8, F32 = sr.COLOR_RGBA8, sr.COLOR_RGBAF32
NORMAL = np.array([0.3, -0.5, 0.8], np.float64)
NORMAL /= np.linalg.norm(NORMAL)
LIGHT = np.array([0.0, 0.6, 0.8], np.float32)


def gbuffer_uniforms(alpha=0.75):
    u = np.zeros(32, np.float32)
    u[0:3] = LIGHT
    u[3], u[4], u[5] = 0.2, 0.7, alpha
    return u


def _uv_tex(n=16):
    """An RGBA8 gradient texture: texel (x, y) = (16x, 16y, 128, 255)."""
    y, x = np.mgrid[0:n, 0:n]
    return (n, n, ((x * 16) | ((y * 16) << 8) | (128 << 16) | (255 << 24)).astype(np.uint32).ravel())


def _tri_scene(W=32, H=32, fp=sr.FP_GBUFFER, blend=0, alpha=0.75, **kw):
    pts = ((2.5, 3.5), (27.5, 6.5), (9.5, 28.5))
    m = screen_tri_mesh(W, H, pts, (0.5, -0.25, 0.875), u=(0.1, 0.5, 0.9), v=(0.8, 0.2, 0.45), nrm=NORMAL)
    d = dict(vertex_program=0, fragment_program=fp, blend=blend, cull=0, tex_filter=0, tex_edge=0,
             vs_uniforms=scenes.vs_uniforms(np.eye(4)), fs_uniforms=gbuffer_uniforms(alpha),
             viewport=(0.0, 0.0, float(W), float(H), 0.0, 1.0))
    d.update(kw)
    return dict(m, width=W, height=H, desc=d), pts


def _oracle(scene, planes, depth=0, stencil=0, colors=None, tex=None):
    fb = oracle.FramebufferEx(scene["width"], scene["height"], planes, depth, stencil, colors)
    rc, st = oracle.draw(fb, scene, sr.make_desc(**scene["desc"]), tex)
    assert rc == 0, rc
    return fb, st


# ---------------------------------------------------------------------------
# CPU: closed forms
# ---------------------------------------------------------------------------
def test_gbuffer_planes_at_the_vertices():
    """At a vertex's pixel the barycentrics are exactly (1, 0, 0), so every
    interpolated varying is that vertex's value and each plane is the
    program's closed form of it."""
    s, pts = _tri_scene()
    tex = _uv_tex()
    fb, st = _oracle(s, [RGBA8, F32, F32, F32], tex=tex)
    assert st["shaded"] > 100
    def norm3(v):  # normalize3: v / sqrt(((x*x) + (y*y)) + (z*z)), in f32
        ln = np.sqrt(f32(f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2])))
        return np.array([f32(v[0] / ln), f32(v[1] / ln), f32(v[2] / ln)], np.float32)

    # the vertex program normalises the normal, the fragment program again
    nn = norm3(norm3(np.asarray(NORMAL, np.float32)))
    ndl = f32(f32(f32(nn[0] * LIGHT[0]) + f32(nn[1] * LIGHT[1])) + f32(nn[2] * LIGHT[2]))
    diff = max(ndl, f32(0))
    lit = f32(f32(0.2) + f32(f32(0.7) * diff))
    for k, (x, y) in enumerate(pts):
        i = int(x) + int(y) * 32
        uu, vv = f32(s["u"][k]), f32(s["v"][k])
        assert np.array_equal(fb.planes[1][i], np.array([f32(nn[0] * f32(0.5)) + f32(0.5), f32(nn[1] * f32(0.5)) + f32(0.5),
                                                         f32(nn[2] * f32(0.5)) + f32(0.5), 1.0], np.float32))
        assert np.array_equal(fb.planes[2][i], np.array([uu, vv, diff, 1.0], np.float32))
        assert np.array_equal(fb.planes[3][i], np.array([lit, lit, lit, 1.0], np.float32))
        # plane 0: nearest texel of (u, v) times lit, quantised
        tx, ty = int(np.floor(f32(uu * f32(16)))), int(np.floor(f32(vv * f32(16))))
        t = [f32(tx * 16) / f32(255), f32(ty * 16) / f32(255), f32(128) / f32(255), f32(1.0)]
        q = lambda c: int(np.floor(f32(f32(min(max(c, f32(0)), f32(1))) * f32(255)) + f32(0.5)))
        want = q(f32(t[0] * lit)) | (q(f32(t[1] * lit)) << 8) | (q(f32(t[2] * lit)) << 16) | (q(f32(t[3] * f32(0.75))) << 24)
        assert int(fb.planes[0][i]) == want


def _over_normalized(s, d):
    a, b = f32(s[3]), f32(d[3])
    a1 = f32(f32(1) - a)
    rgb = [f32(f32(f32(f32(s[c]) * a) + f32(f32(f32(d[c]) * b) * a1)) / f32(a + f32(b * a1))) for c in range(3)]
    return np.array(rgb + [f32(a + f32(b * f32(f32(1) - a)))], np.float32)


@pytest.mark.parametrize("alpha", [0.75, 0.3])
def test_over_normalized_closed_form(alpha):
    """FLAT colour over a cleared RGBAF32 plane: every covered pixel is the
    formula of the source and the clear colour."""
    dst = [0.2, 0.4, 0.9, 0.5]
    src = [1.0, 0.5, 0.25, alpha]
    fs = np.zeros(32, np.float32)
    fs[0:4] = src
    s, _ = _tri_scene(fp=sr.FP_FLAT, blend=sr.BLEND_OVER_NORMALIZED, fs_uniforms=fs)
    fb, st = _oracle(s, [F32, RGBA8], colors=[dst, 0xFF336699])
    cov = fb.depth > oracle.F32_MIN
    assert cov.sum() == st["shaded"] > 100
    want = _over_normalized(np.array(src, np.float32), np.array(dst, np.float32))
    assert np.array_equal(fb.planes[0][cov], np.tile(want, (cov.sum(), 1)))
    # the RGBA8 plane: the source quantised, then the same blend on unorm
    # floats (c / 255), quantised again
    q = lambda c: int(np.floor(f32(f32(min(max(c, f32(0)), f32(1))) * f32(255)) + f32(0.5)))
    s8 = np.array([q(f32(c)) for c in src], np.float32) / f32(255)
    w8 = _over_normalized(s8.astype(np.float32), np.array([0x99, 0x66, 0x33, 0xFF], np.float32) / f32(255))
    w = [q(c) for c in w8]
    assert set(fb.planes[1][cov].tolist()) == {w[0] | w[1] << 8 | w[2] << 16 | w[3] << 24}


def test_depth_f64_and_none_stencil_match_the_standard_format():
    s = scenes.soup_scene(1500, 96, 64, seed=41)
    ref, _ = render_ref(dict(s, clear=0))
    fb, _ = _oracle(s, [RGBA8], depth=sr.DEPTH_F64, stencil=sr.STENCIL_NONE, tex=s["texture"])
    assert np.array_equal(fb.planes[0], ref.color)
    cov = ref.depth > oracle.F32_MIN
    assert np.array_equal(fb.depth64[cov], ref.depth[cov].astype(np.float64))
    assert np.all(fb.depth64[~cov] == -np.finfo(np.float64).max)
    # StencilTest::Never on a `()` stencil is no test at all: the same frame
    s2 = dict(s, desc=dict(s["desc"], stencil_test=1))
    fb2, _ = _oracle(s2, [RGBA8], stencil=sr.STENCIL_NONE, tex=s["texture"])
    assert np.array_equal(fb2.planes[0], ref.color)


def test_rgbaf32_texture_reads_texel_centres():
    w = h = 8
    data = np.random.default_rng(3).random((w * h, 4)).astype(np.float32)
    t = oracle.OrTexture(w, h, None, data.ctypes.data_as(C.POINTER(C.c_float)))
    out = np.zeros(4, np.float32)
    for ix, iy in ((0, 0), (3, 5), (7, 7)):
        oracle.lib().or_sample(C.byref(t), f32((ix + 0.5) / w), f32((iy + 0.5) / h), 0, 0, 0,
                               out.ctypes.data_as(C.POINTER(C.c_float)))
        assert np.array_equal(out, data[iy * w + ix])


# ---------------------------------------------------------------------------
# GPU: the device against the oracle
# ---------------------------------------------------------------------------
def _gpu(ctx, scene, planes, depth=0, stencil=0, colors=None, tex=None):
    fb = ctx.framebuffer_ex(scene["width"], scene["height"], planes, depth, stencil)
    fb.clear_ex(colors or [0] * len(planes))
    mesh = ctx.mesh(scene)
    try:
        ctx.draw(fb, mesh, sr.make_desc(**scene["desc"]), tex)
        return fb, fb.read_ex()
    finally:
        mesh.free()


def _same(got, ref):
    planes, dep, st = got
    for p, r in zip(planes, ref.planes):
        assert np.array_equal(p.view(np.uint32), np.asarray(r).view(np.uint32))
    if ref.depth64 is not None:
        assert np.array_equal(dep.view(np.uint64), ref.depth64.view(np.uint64))
    else:
        assert np.array_equal(dep.view(np.uint32), ref.depth.view(np.uint32))
    if st is not None:
        assert np.array_equal(st, ref.stencil)


@pytest.mark.gpu
@pytest.mark.parametrize("blend", [0, 1, 2])
def test_gpu_gbuffer_soup(ctx, blend):
    s = scenes.soup_scene(4000, 200, 120, seed=42)
    s["desc"] = dict(s["desc"], fragment_program=sr.FP_GBUFFER, blend=blend, fs_uniforms=gbuffer_uniforms(0.6))
    planes = [RGBA8, F32, F32, RGBA8]
    colors = [0xFF202020, [0.1, 0.2, 0.3, 0.5], [0.0, 0.0, 0.0, 1.0], 0x80FFFFFF]
    ref, st = _oracle(s, planes, depth=sr.DEPTH_F64, colors=colors, tex=s["texture"])
    assert st["shaded"] > 1000
    tex = ctx.texture(*s["texture"])
    fb, got = _gpu(ctx, s, planes, depth=sr.DEPTH_F64, colors=colors, tex=tex)
    fb.free()
    tex.free()
    _same(got, ref)


@pytest.mark.gpu
def test_gpu_over_normalized_on_the_standard_framebuffer(ctx):
    s = scenes.soup_scene(3000, 160, 96, seed=43)
    s["desc"] = dict(s["desc"], blend=sr.BLEND_OVER_NORMALIZED)
    ref, _ = render_ref(s)
    from parity import render_gpu
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("test,op", [(0, 4), (2, 5), (6, 3)])
def test_gpu_stencil_on_other_formats(ctx, test, op):
    s = scenes.soup_scene(900, 120, 80, seed=44)
    s["desc"] = dict(s["desc"], stencil_test=test, stencil_op=op, stencil_wrap=1, stencil_ref=2, tile_w=32,
                     tile_h=32)
    planes = [F32, RGBA8]
    ref, _ = _oracle(s, planes, colors=[[0, 0, 0, 1], 0xFF000000], tex=s["texture"])
    tex = ctx.texture(*s["texture"])
    fb, got = _gpu(ctx, s, planes, colors=[[0, 0, 0, 1], 0xFF000000], tex=tex)
    fb.free()
    tex.free()
    _same(got, ref)
    # the `()` stencil: no test, no op
    ref2, _ = _oracle(s, planes, stencil=sr.STENCIL_NONE, colors=[[0, 0, 0, 1], 0xFF000000], tex=s["texture"])
    tex = ctx.texture(*s["texture"])
    fb, got = _gpu(ctx, s, planes, stencil=sr.STENCIL_NONE, colors=[[0, 0, 0, 1], 0xFF000000], tex=tex)
    fb.free()
    tex.free()
    _same(got, ref2)


@pytest.mark.gpu
def test_gpu_two_pass_render_to_texture_buffer(ctx):
    """Pass 1 writes a texture buffer of an RGBA8 and an RGBAF32 plane;
    pass 2 samples the RGBAF32 plane (no copy) into a standard framebuffer."""
    W = H = 96
    s1 = scenes.soup_scene(2500, W, H, seed=45)
    s1["desc"] = dict(s1["desc"], fragment_program=sr.FP_GBUFFER, blend=0, fs_uniforms=gbuffer_uniforms(1.0))
    ref1, _ = _oracle(s1, [RGBA8, F32], colors=[0xFF000000, [0.25, 0.5, 0.75, 1.0]], tex=s1["texture"])
    tex = ctx.texture(*s1["texture"])
    fb1, got1 = _gpu(ctx, s1, [RGBA8, F32], colors=[0xFF000000, [0.25, 0.5, 0.75, 1.0]], tex=tex)
    _same(got1, ref1)
    s2 = scenes.soup_scene(2000, 128, 80, seed=46)
    s2["desc"] = dict(s2["desc"], tex_filter=1, tex_edge=1)
    ref2, _ = render_ref(dict(s2, texture=oracle.plane_texture(ref1, 1)))
    t1 = ctx.texture_from_fb_plane(fb1, 1)
    fb2 = ctx.framebuffer(128, 80)
    fb2.clear(s2["clear"])
    mesh = ctx.mesh(s2)
    ctx.draw(fb2, mesh, sr.make_desc(**s2["desc"]), t1)
    assert_parity(fb2.read(), ref2)
    for o in (mesh, t1, fb2, fb1, tex):
        o.free()


@pytest.mark.gpu
def test_gpu_formats_reject_lines(ctx):
    fb = ctx.framebuffer_ex(32, 32, [F32])
    line = dict(px=np.array([-0.5, 0.5], np.float32), py=np.zeros(2, np.float32), pz=np.full(2, 0.5, np.float32),
                nx=None, ny=None, nz=None, u=np.zeros(2, np.float32), v=np.zeros(2, np.float32), indices=None,
                primitive=2)
    mesh = ctx.mesh(line)
    with pytest.raises(sr.SoftrenderError) as e:
        ctx.draw(fb, mesh, sr.make_desc(vertex_program=0, fragment_program=0,
                                        vs_uniforms=scenes.vs_uniforms(np.eye(4)),
                                        viewport=(0.0, 0.0, 32.0, 32.0, 0.0, 1.0)))
    assert e.value.code == -5  # SR_ERR_UNSUPPORTED
    assert fb.format() == dict(planes=[F32], depth=0, stencil=0)
    mesh.free()
    fb.free()
