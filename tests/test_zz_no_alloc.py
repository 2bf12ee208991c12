"""sr_draw_static (the draw with its workspace in the library's fixed per-GPU
region) against sr_draw, through the C ABI: identical framebuffers, bit for
bit, on consecutive calls (the second with a smaller input, so stale
contents of the shared region from the first call must not leak into it),
for triangles (binned and with stencil strip counts), lines and points; a
draw too large for the region is refused and leaves the framebuffer as it
was."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr

pytestmark = pytest.mark.gpu


def _draw_both(ctx, scene, overrides=None):
    """The scene drawn by sr_draw into one framebuffer and by sr_draw_static
    into another; returns both (color, depth, stencil) triples."""
    W, H = scene["width"], scene["height"]
    tex = None
    if scene.get("texture") is not None:
        tw, th, rgba = scene["texture"]
        tex = ctx.texture(tw, th, rgba)
    out = []
    for static in (False, True):
        fb = ctx.framebuffer(W, H)
        fb.clear(scene.get("clear", 0))
        for d in scene.get("draws") or [scene]:
            kw = dict(d["desc"])
            if overrides:
                kw.update(overrides)
            mesh = ctx.mesh(d)
            try:
                (ctx.draw_static if static else ctx.draw)(fb, mesh, sr.make_desc(**kw), tex)
            finally:
                mesh.free()
        out.append(fb.read())
        fb.free()
    if tex:
        tex.free()
    return out


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_static_matches_draw_twice_smaller_second(ctx):
    big = scenes.soup_scene(60_000, 640, 360, seed=31)
    small = scenes.soup_scene(5_000, 200, 120, seed=32)
    for s in (big, small):
        plain, static = _draw_both(ctx, s)
        _same(plain, static)
        assert (plain[1] > np.finfo(np.float32).min).sum() > 1000  # something was drawn


def test_static_stencil_strips_then_smaller(ctx):
    ov = dict(stencil_test=0, stencil_op=4, stencil_wrap=1, stencil_ref=0, tile_w=64, tile_h=48)
    for s in (scenes.soup_scene(20_000, 480, 270, seed=33), scenes.soup_scene(1_500, 160, 100, seed=34)):
        plain, static = _draw_both(ctx, s, ov)
        _same(plain, static)
        assert plain[2].any()


def test_static_lines_points_clip(ctx):
    s = scenes.clip_scene(4_000, 2_000, 1_000, width=480, height=270)
    plain, static = _draw_both(ctx, s)
    _same(plain, static)
    s2 = scenes.clip_scene(1_000, 500, 300, width=320, height=180, seed=scenes.SEED_CLIP + 1)
    plain, static = _draw_both(ctx, s2)
    _same(plain, static)


def test_static_refuses_what_does_not_fit(ctx):
    # ~264 B of workspace per soup triangle: 600k triangles need ~150 MiB
    s = scenes.soup_scene(600_000, 640, 360, seed=35)
    fb = ctx.framebuffer(640, 360)
    fb.clear(0xFF204060)
    before = fb.read()
    mesh = ctx.mesh(s)
    tw, th, rgba = s["texture"]
    tex = ctx.texture(tw, th, rgba)
    with pytest.raises(sr.SoftrenderError) as e:
        ctx.draw_static(fb, mesh, sr.make_desc(**s["desc"]), tex)
    assert e.value.code == sr.SR_ERR_WORKSPACE
    assert "static workspace" in str(e.value)
    _same(before, fb.read())
    # the context's own path still works afterwards, and so does a small static draw
    ctx.draw(fb, mesh, sr.make_desc(**s["desc"]), tex)
    mesh.free()
    tex.free()
    fb.free()
    small = scenes.soup_scene(2_000, 128, 96, seed=36)
    plain, static = _draw_both(ctx, small)
    _same(plain, static)
