"""Asynchronous batches (sr_batch_submit / sr_batch_query / sr_batch_wait):
several batches in flight at once each equal the oracle; a fifth submit
completes the oldest first; items that cannot join the one-pass group are
drawn at submit; tickets are single-use."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_ref

pytestmark = pytest.mark.gpu


def _items(ctx, specs, tex):
    items, keep = [], []
    for s in specs:
        fb = ctx.framebuffer(s["width"], s["height"])
        fb.clear(0xFF00FF00)
        mesh = ctx.mesh(s)
        desc = sr.make_desc(**s["desc"])
        keep.append((fb, mesh, desc))
        items.append((fb, mesh, desc, tex if s.get("texture") else None, 0xFF000000))
    return items, keep


def test_async_batches_in_flight(ctx):
    groups = [[scenes.soup_scene(800 + 300 * j, 160 + 16 * i, 96, seed=100 + 10 * i + j) for j in range(3)]
              for i in range(5)]
    tex = ctx.texture(*groups[0][0]["texture"])
    batches = [_items(ctx, g, tex) for g in groups]
    tickets = [ctx.batch_submit(items) for items, _ in batches]
    assert len(set(tickets)) == 5
    for t in reversed(tickets):  # any order
        ctx.batch_wait(t)
    for (items, _), g in zip(batches, groups):
        for (fb, *_), s in zip(items, g):
            ref, _ = render_ref(dict(s, clear=0xFF000000))
            assert_parity(fb.read(), ref)
    with pytest.raises(sr.SoftrenderError):
        ctx.batch_wait(tickets[0])  # already waited for


def test_async_batch_query_and_fallback_items(ctx):
    tri = scenes.soup_scene(3000, 256, 144, seed=120)
    stencil = scenes.soup_scene(900, 128, 96, seed=121)
    stencil["desc"].update(stencil_test=0, stencil_op=4, stencil_wrap=0)  # drawn at submit
    specs = [tri, stencil]
    tex = ctx.texture(*tri["texture"])
    items, keep = _items(ctx, specs, tex)
    t = ctx.batch_submit(items)
    while not ctx.batch_ready(t):
        pass
    ctx.batch_wait(t)
    for (fb, *_), s in zip(items, specs):
        ref, _ = render_ref(dict(s, clear=0xFF000000))
        assert_parity(fb.read(), ref)
