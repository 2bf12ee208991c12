"""Execution configurations and the autotuner.

The execution tiling (tile edge, waves sharing a tile) never changes a
result: every configuration equals the oracle bit for bit, for the plain
draw, a stencil-counting draw and a batch.  A tiny autotuner sweep installs
a policy under which the images stay the same.
"""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

CONFIGS = [(8, 1), (16, 1), (16, 2), (16, 4), (32, 2), (32, 4)]


@pytest.mark.parametrize("ets,nwv", CONFIGS)
def test_soup_every_configuration(ctx, ets, nwv):
    s = scenes.soup_scene(6000, 320, 180, seed=11)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s, dict(exec_tile=ets, exec_waves=nwv)), ref)


@pytest.mark.parametrize("ets,nwv", [(16, 1), (16, 2), (32, 2)])
def test_stencil_counts_every_configuration(ctx, ets, nwv):
    s = scenes.soup_scene(2500, 300, 170, seed=13)
    ov = dict(stencil_test=0, stencil_op=4, stencil_wrap=1, stencil_ref=7, tile_w=64, tile_h=48,
              exec_tile=ets, exec_waves=nwv)
    ref, _ = render_ref(s, {k: v for k, v in ov.items()})
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("ets,nwv", [(16, 1), (16, 2), (32, 2)])
def test_batch_every_configuration(ctx, ets, nwv):
    specs = [scenes.soup_scene(1500 + 500 * i, w, h, seed=70 + i) for i, (w, h) in
             enumerate([(256, 256), (320, 180), (97, 61)])]
    tex = ctx.texture(*specs[0]["texture"])
    items, keep = [], []
    for s in specs:
        fb = ctx.framebuffer(s["width"], s["height"])
        fb.clear(0xFF00FF00)
        mesh = ctx.mesh(s)
        desc = sr.make_desc(**dict(s["desc"], exec_tile=ets, exec_waves=nwv))
        keep.append((fb, mesh, desc))
        items.append((fb, mesh, desc, tex, 0xFF000000))
    ctx.batch_draw(items)
    for (fb, *_), s in zip(items, specs):
        ref, _ = render_ref(dict(s, clear=0xFF000000))
        assert_parity(fb.read(), ref)


def test_autotune_tiny_sweep_keeps_images():
    import autotune

    ctx = sr.Context(0)
    try:
        tuned = autotune.tune(ctx, resolutions=[(128, 128), (320, 180)], tris_per_mpx=2e5, scenes_per_class=2,
                              configs=[(8, 1), (16, 1), (16, 4)], capacities=(1.0, 4.0), reps=1)
        assert tuned["images_agree"]
        assert len(tuned["policy"]) == 2
        assert all((e, w) in [(8, 1), (16, 1), (16, 4)] for _, e, w in tuned["policy"])
        assert tuned["pairs_per_tri"] in (1.0, 4.0)
        autotune.apply(ctx, tuned, expected_tris=10_000)
        s = scenes.soup_scene(3000, 320, 180, seed=12)
        ref, _ = render_ref(s)
        assert_parity(render_gpu(ctx, s), ref)
        fb = ctx.framebuffer(320, 180)
        mesh = ctx.mesh(s)
        tex = ctx.texture(*s["texture"])
        ctx.batch_draw([(fb, mesh, sr.make_desc(**s["desc"]), tex, s["clear"])])
        assert_parity(fb.read(), ref)
        for o in (mesh, tex, fb):
            o.free()
    finally:
        ctx.close()
