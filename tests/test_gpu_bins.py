"""Bin capacity of the plain draw.

The plain draw sizes its (execution tile, triangle) pair buffers before it
knows the pair count (a guess of 4 pairs per triangle on a context's first
draws), builds the bins and launches the raster at once; a draw with more
pairs draws nothing on the device, and once the count reaches the host the
buffers grow and the draw runs again.  Both ways, and a sequence of draws
whose sizes grow and shrink on one context, equal the oracle."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu


def big_triangles(n, W, H, seed, scale=8.0):
    """Soup triangles scaled about their centroids: tens of execution tiles
    per triangle, so the first draw overflows the guessed capacity."""
    s = scenes.soup_scene(n, W, H, seed=seed)
    for c in ("px", "py"):
        v = s[c].reshape(-1, 3).astype(np.float64)
        m = v.mean(1, keepdims=True)
        s[c] = (m + (v - m) * scale).reshape(-1).astype(np.float32)
    return s


def test_overflow_then_redraw_on_a_fresh_context():
    ctx = sr.Context(0)
    try:
        s = big_triangles(3000, 320, 240, seed=41)
        ref, _ = render_ref(s)
        got = render_gpu(ctx, s)
        pairs = ctx.last_stats()["bin_pairs"]
        assert pairs > 4 * 3000 + 4096, pairs  # the first guess was too small
        assert_parity(got, ref)
        # the same draw again: the grown buffers hold it at once
        assert_parity(render_gpu(ctx, s), ref)
    finally:
        ctx.close()


def test_growing_and_shrinking_draws_on_one_context():
    ctx = sr.Context(0)
    try:
        for n, (W, H), seed in [(1_000, (256, 256), 51), (120_000, (512, 512), 52), (400_000, (512, 512), 53),
                                (20_000, (320, 180), 54), (4_000, (256, 256), 55)]:
            s = scenes.soup_scene(n, W, H, seed=seed)
            if n == 4_000:
                s = big_triangles(n, W, H, seed)
            print(f"draw {n} triangles at {W}x{H}", flush=True)
            ref, _ = render_ref(s)
            assert_parity(render_gpu(ctx, s), ref)
    finally:
        ctx.close()
