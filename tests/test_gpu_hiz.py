"""Hierarchical depth: a triangle whose depth bound lies below every depth a
wave's pixels hold is skipped in that wave.  Scenes where most triangles
are skipped (deep overlap at low resolution), framebuffers whose stored
depths hold every special value (NaNs of both signs, zeros of both signs,
infinities, tiles at depth almost 0, pixels that no triangle can pass), the
plain draw at every execution tile, the `()` blend and the batch: every
image equals the oracle's bit for bit."""
import numpy as np
import pytest

import oracle
import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

W, H = 317, 239  # partial execution tiles on the right and bottom edges


def preset(seed):
    """Colour, depth and stencil planes to draw over: per 8x8 block one kind
    of depth content."""
    rng = np.random.default_rng(seed)
    n = W * H
    col = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    st = np.zeros(n, np.uint8)
    special = np.array([np.nan, -np.nan, 0.0, -0.0, np.inf, -np.inf, -3.4028235e38, -1e-30, 0.5],
                       np.float32)
    dep = rng.uniform(-1, 0, n).astype(np.float32)
    y, x = np.divmod(np.arange(n), W)
    block = (y // 8) * ((W + 7) // 8) + x // 8
    kind = block % 6
    pick = rng.integers(0, len(special), n)
    dep = np.where(kind == 1, special[pick], dep)            # specials mixed in
    dep = np.where(kind == 2, np.float32(-1e-30), dep)       # nearly everything hidden
    dep = np.where(kind == 3, np.float32(np.nan), dep)       # nothing passes
    dep = np.where(kind == 4, -np.float32(np.nan), dep)      # nothing passes, least key
    dep = np.where((kind == 5) & (rng.uniform(0, 1, n) < 0.05), -np.float32(np.nan), dep)
    return col, dep.astype(np.float32), st


def ref_over(s, col, dep, st, overrides=None):
    fb = oracle.Framebuffer(s["width"], s["height"])
    fb.color[:], fb.depth[:], fb.stencil[:] = col, dep, st
    desc = dict(s["desc"], **(overrides or {}))
    rc, _ = oracle.draw(fb, s, sr.make_desc(**desc), s.get("texture"), 1)
    assert rc == 0
    return fb


@pytest.mark.parametrize("exec_tile", [8, 16, 32])
def test_preset_depths(ctx, exec_tile):
    s = scenes.soup_scene(40_000, W, H, seed=0x71)
    col, dep, st = preset(0x72)
    fb = ctx.framebuffer(W, H)
    try:
        fb.write(col, dep, st)
        got = render_gpu(ctx, s, dict(exec_tile=exec_tile), fb=fb)
    finally:
        fb.free()
    assert_parity(got, ref_over(s, col, dep, st, dict(exec_tile=exec_tile)))


@pytest.mark.parametrize("exec_tile,blend", [(8, 1), (16, 1), (32, 1), (8, 0)])
def test_deep_overlap(ctx, exec_tile, blend):
    """256x256 with 120k triangles: about 60 layers per pixel, so most
    triangles arrive behind a wave's every pixel."""
    s = scenes.soup_scene(120_000, 256, 256, seed=0x73)
    ov = dict(exec_tile=exec_tile, blend=blend)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


def test_batch_deep_overlap_and_preset(ctx):
    specs = [scenes.soup_scene(n, w, h, seed=0x74 + i)
             for i, (n, w, h) in enumerate([(90_000, 256, 256), (150_000, 512, 512), (30_000, W, H)])]
    tex = ctx.texture(*specs[0]["texture"])
    col, dep, st = preset(0x75)
    items, keep = [], []
    try:
        for s in specs:
            fb = ctx.framebuffer(s["width"], s["height"])
            mesh = ctx.mesh(s)
            keep += [fb, mesh]
            pre = (s["width"], s["height"]) == (W, H)
            if pre:
                fb.write(col, dep, st)
            items.append((fb, mesh, sr.make_desc(**s["desc"]), tex, None if pre else 0xFF000000))
        ctx.batch_draw(items)
        for (fb, *_), s, it in zip(items, specs, items):
            if it[4] is None:
                ref = ref_over(s, col, dep, st)
            else:
                ref, _ = render_ref(dict(s, clear=0xFF000000))
            assert_parity(fb.read(), ref)
    finally:
        for o in keep + [tex]:
            o.free()
