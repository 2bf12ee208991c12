"""Slivers that the reference shades outside their own box (tests/slivers.py).

CPU: the oracle and the independent numpy-float32 restatement both shade
pixel (512, 512) for each sliver, outside its box (a pixel the reference
reaches only through a screen tile's clamped strip).

GPU: a triangle binned to the execution tiles of its box alone would
never be tested at that pixel; setup finds that a pixel outside the box
may pass, bins the sliver to every tile such a pixel may lie in, and the
raster tests it in its box clamped into each screen tile, so every path
(plain draw, stencil counting, the one-pass batch, the request server, the
static workspace draw) equals the oracle bit for bit.
"""
import numpy as np
import pytest

import np_restatement as npr
import oracle
import softrender_hip as sr
from oracle_checks import desc_dict
from parity import assert_parity, render_gpu, render_ref
from slivers import PIXEL, SLIVERS, W, sliver_scene

IDX = PIXEL[0] + PIXEL[1] * W


@pytest.mark.parametrize("k", range(len(SLIVERS)))
def test_oracle_shades_outside_the_box(k):
    s = sliver_scene(k)
    fb, st = render_ref(s)
    assert st[0]["strip_shaded"] == 1, st
    assert fb.color[IDX] == 0xFF4080FF  # FLAT (1, 0.5, 0.25, 1)
    # the box of the sliver ends before column and row 512
    sx = (s["px"] * np.float32(W / 2)) + np.float32(W / 2)
    sy = (s["py"] * np.float32(-W / 2)) + np.float32(W / 2)
    assert np.floor(sx.max()) < PIXEL[0] and np.floor(sy.max()) < PIXEL[1]


@pytest.mark.parametrize("k", range(len(SLIVERS)))
def test_restatement_agrees(k):
    s = sliver_scene(k)
    d = sr.make_desc(**s["desc"])
    ofb, _ = render_ref(s)
    fb = dict(w=W, h=W, color=np.full(W * W, 0xFF000000, np.uint32),
              depth=np.full(W * W, npr.F32_MIN, np.float32), stencil=np.zeros(W * W, np.uint8))
    npr.draw(fb, s, desc_dict(d))
    assert fb["color"][IDX] == 0xFF4080FF
    assert np.array_equal(fb["color"], ofb.color)
    assert np.array_equal(fb["depth"].view(np.uint32), ofb.depth.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(len(SLIVERS)))
def test_gpu_sliver_alone_and_in_a_soup(ctx, k):
    for extra in (0, 3000):
        s = sliver_scene(k, extra=extra)
        ref, st = render_ref(s)
        assert st[0]["strip_shaded"] >= 1
        got = render_gpu(ctx, s)
        assert ctx.last_stats()["strip_slivers"] == 1
        assert_parity(got, ref)
        assert got[0][IDX] == ref.color[IDX] == 0xFF4080FF or extra


@pytest.mark.gpu
def test_gpu_sliver_with_stencil_counts(ctx):
    ov = dict(stencil_test=0, stencil_op=4, stencil_wrap=1, stencil_ref=0)
    s = sliver_scene(0, extra=1500)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)
    assert ctx.last_stats()["strip_slivers"] == 1


@pytest.mark.gpu
def test_gpu_ordinary_draw_has_no_strip_slivers(ctx):
    import scenes

    s = scenes.soup_scene(20_000, 640, 360, seed=9)
    render_gpu(ctx, s)
    assert ctx.last_stats()["strip_slivers"] == 0


@pytest.mark.gpu
def test_gpu_sliver_in_a_batch(ctx):
    import scenes

    specs = [scenes.soup_scene(2000, 320, 180, seed=61), sliver_scene(1, extra=500),
             scenes.soup_scene(900, 256, 256, seed=62)]
    tex = ctx.texture(*specs[0]["texture"])
    items, keep = [], []
    for s in specs:
        fb = ctx.framebuffer(s["width"], s["height"])
        fb.clear(0xFF00FF00)
        mesh = ctx.mesh(s)
        desc = sr.make_desc(**s["desc"])
        keep.append((fb, mesh, desc))
        items.append((fb, mesh, desc, tex if s.get("texture") else None, 0xFF000000))
    ctx.batch_draw(items)
    assert ctx.last_stats()["strip_slivers"] == 1
    for (fb, *_), s in zip(items, specs):
        ref, _ = render_ref(dict(s, clear=0xFF000000))
        assert_parity(fb.read(), ref)


@pytest.mark.gpu
def test_gpu_sliver_served(ctx):
    s = sliver_scene(2, extra=800)
    mesh = ctx.mesh(s)
    srv = sr.Server(ctx, max_inflight=2, max_batch_frames=4)
    try:
        t = srv.submit(mesh, sr.make_desc(**s["desc"]), W, W, want_depth=True)
        got = []
        while not got:
            got = srv.poll(4, timeout_ms=200)
    finally:
        srv.close()
        mesh.free()
    ref, _ = render_ref(s)
    assert got[0]["status"] == 0 and got[0]["ticket"] == t
    assert np.array_equal(got[0]["rgba"], ref.color)
    assert np.array_equal(got[0]["depth"].view(np.uint32), ref.depth.view(np.uint32))


@pytest.mark.gpu
def test_gpu_sliver_static_workspace(ctx):
    s = sliver_scene(0, extra=400)
    fb = ctx.framebuffer(W, W)
    fb.clear(0xFF000000)
    mesh = ctx.mesh(s)
    ctx.draw_static(fb, mesh, sr.make_desc(**s["desc"]))
    got = fb.read()
    mesh.free()
    fb.free()
    ref, _ = render_ref(s)
    assert_parity(got, ref)
