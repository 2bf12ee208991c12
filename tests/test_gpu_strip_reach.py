"""Views of the serving workload's meshes (balls of randomly oriented
triangles under an orbit camera), where slivers that may pass the coverage
test outside their own box are common: setup bins them to every tile such
a pixel may lie in and the raster tests them in their tile-clamped boxes.
Every path equals the oracle bit for bit: the plain draw (with stencil
counting too), the one-pass batch and the request server."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

VIEWS = [(512, 512, 60_000, 0x53), (1920, 1080, 40_000, 0x52), (640, 360, 120_000, 0x61)]


def ball_scene(W, H, n, seed):
    s = scenes.serve_mesh(n, seed)
    mvp = scenes.orbit_camera(scenes.SplitMix64(seed + 1), W, H)
    return dict(s, width=W, height=H, clear=0xFF000000, texture=(1024, 1024, scenes.checkerboard(1024, 32)),
                desc=scenes.serve_desc(mvp, W, H))


def test_ball_views_plain(ctx):
    flagged = 0
    for W, H, n, seed in VIEWS:
        s = ball_scene(W, H, n, seed)
        print(f"ball view {W}x{H} {n} tris seed {seed:#x}", flush=True)
        ref, _ = render_ref(s)
        got = render_gpu(ctx, s)
        flagged += ctx.last_stats()["strip_slivers"]
        assert_parity(got, ref)
    assert flagged >= 1


def test_ball_view_stencil_counting(ctx):
    s = ball_scene(640, 360, 120_000, 0x61)
    ov = dict(stencil_test=0, stencil_op=4, stencil_wrap=1, stencil_ref=0)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


def test_ball_views_batched(ctx):
    specs = [ball_scene(*v) for v in VIEWS]
    tex = ctx.texture(*specs[0]["texture"])
    items, keep = [], []
    try:
        for s in specs:
            fb = ctx.framebuffer(s["width"], s["height"])
            mesh = ctx.mesh(s)
            keep += [fb, mesh]
            items.append((fb, mesh, sr.make_desc(**s["desc"]), tex, 0xFF000000))
        ctx.batch_draw(items)
        assert ctx.last_stats()["strip_slivers"] == 1
        for (fb, *_), s in zip(items, specs):
            ref, _ = render_ref(s)
            assert_parity(fb.read(), ref)
    finally:
        for o in keep + [tex]:
            o.free()


def test_ball_views_served(ctx):
    specs = [ball_scene(*v) for v in VIEWS]
    tex = ctx.texture(*specs[0]["texture"])
    meshes = [ctx.mesh(s) for s in specs]
    srv = sr.Server(ctx)
    try:
        for k, s in enumerate(specs):
            srv.submit(meshes[k], sr.make_desc(**s["desc"]), s["width"], s["height"], tex=tex, tag=k,
                       want_depth=True)
        got = {}
        while len(got) < len(specs):
            for r in srv.poll(16, timeout_ms=100):
                got[r["tag"]] = r
        for k, s in enumerate(specs):
            assert got[k]["status"] == 0
            ref, _ = render_ref(s)
            assert np.array_equal(got[k]["rgba"].view(np.uint32).ravel(), ref.color)
            assert np.array_equal(got[k]["depth"].view(np.uint32).ravel(), ref.depth.view(np.uint32))
    finally:
        srv.close()
        for m in meshes:
            m.free()
        tex.free()
