"""GPU parity for line and point draws (rasterization/line.rs, point.rs) and
the 4K clipping/primitive frame at reduced size, bit for bit."""
import numpy as np
import pytest

import scenes
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu


def line_scene(n, w, h, seed, aa=1, alpha_mode=1, blend=1, tile=(128, 128), margin=2.0, **kw):
    rng = scenes.SplitMix64(seed)
    L = 30.0 * np.exp(0.7 * rng.normal(n))
    ang = rng.uniform(n, 0.0, 2 * np.pi)
    sx = rng.uniform(n, -margin, w - 1 + margin)
    sy = rng.uniform(n, -margin, h - 1 + margin)
    ex, ey = sx + L * np.cos(ang), sy + L * np.sin(ang)
    x = np.stack([sx, ex], 1).reshape(-1) * (2.0 / w) - 1.0
    y = 1.0 - np.stack([sy, ey], 1).reshape(-1) * (2.0 / h)
    z = np.repeat(rng.uniform(n, 0.02, 0.98), 2)
    u, v = rng.uniform(2 * n), rng.uniform(2 * n)
    fs = np.zeros(32, np.float32)
    fs[0], fs[1] = 0.7, 0.9
    desc = dict(vertex_program=0, fragment_program=3, blend=blend, aa_lines=aa, alpha_mode=alpha_mode,
                vs_uniforms=scenes.vs_uniforms(np.eye(4)), fs_uniforms=fs, tile_w=tile[0], tile_h=tile[1],
                viewport=(0.0, 0.0, float(w), float(h), 0.0, 1.0))
    desc.update(kw)
    return dict(width=w, height=h, clear=0xFF000000, texture=None, primitive=2,
                px=x.astype(np.float32), py=y.astype(np.float32), pz=z.astype(np.float32),
                nx=None, ny=None, nz=None, u=u.astype(np.float32), v=v.astype(np.float32),
                indices=None, desc=desc)


def point_scene(n, w, h, seed, blend=0, **kw):
    rng = scenes.SplitMix64(seed)
    fs = np.zeros(32, np.float32)
    fs[0], fs[1] = 0.3, 0.6
    desc = dict(vertex_program=0, fragment_program=3, blend=blend,
                vs_uniforms=scenes.vs_uniforms(np.eye(4)), fs_uniforms=fs,
                viewport=(0.0, 0.0, float(w), float(h), 0.0, 1.0))
    desc.update(kw)
    return dict(width=w, height=h, clear=0xFF000000, texture=None, primitive=1,
                px=rng.uniform(n, -1.05, 1.05).astype(np.float32),
                py=rng.uniform(n, -1.05, 1.05).astype(np.float32),
                pz=rng.uniform(n, 0.01, 0.99).astype(np.float32),
                nx=None, ny=None, nz=None, u=rng.uniform(n).astype(np.float32),
                v=rng.uniform(n).astype(np.float32), indices=None, desc=desc)


@pytest.mark.parametrize("aa", [0, 1])
@pytest.mark.parametrize("blend", [0, 1])
def test_lines(ctx, aa, blend):
    s = line_scene(3000, 300, 200, seed=31, aa=aa, blend=blend)
    ref, st = render_ref(s)
    assert st[0]["line_plots"] > 10000
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.parametrize("tile", [(128, 128), (40, 25), (300, 7)])
def test_lines_tile_sizes(ctx, tile):
    s = line_scene(1500, 257, 129, seed=32, tile=tile)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


def test_lines_u8_alpha_and_stencil(ctx):
    s = line_scene(2000, 256, 160, seed=33, alpha_mode=0, stencil_test=0, stencil_op=4, stencil_wrap=0)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


def test_lines_value_stencil(ctx):
    s = line_scene(2000, 200, 150, seed=34, stencil_test=5, stencil_op=4, stencil_wrap=1, stencil_ref=1)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


def test_lines_at_the_right_edge_wrap(ctx):
    # Wu writes one pixel past the tile; at x == W the index wraps into the
    # next row (line.rs:55-58); vertical lines hugging the right edge
    w, h = 64, 48
    n = 40
    xs = np.full(n, w - 1.2)
    ys0 = np.linspace(1, h - 8, n)
    s = line_scene(n, w, h, seed=35, margin=0.0)
    x = np.stack([xs, xs + 0.6], 1).reshape(-1) * (2.0 / w) - 1.0
    y = 1.0 - np.stack([ys0, ys0 + 6.0], 1).reshape(-1) * (2.0 / h)
    s["px"], s["py"] = x.astype(np.float32), y.astype(np.float32)
    ref, st = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


def test_degenerate_lines(ctx):
    s = line_scene(8, 64, 64, seed=36)
    pts = np.array([[10, 10, 10, 10], [5.5, 5.5, 5.5, 40], [0, 0, 63, 63], [-20, 5, 80, 5],
                    [31.5, 2, 31.5, 2.0001], [127.9, 3, 128.1, 9], [3, 62.9, 60, 62.9], [-5, -5, -1, -1]])
    x = (pts[:, [0, 2]].reshape(-1) * (2.0 / 64) - 1.0).astype(np.float32)
    y = (1.0 - pts[:, [1, 3]].reshape(-1) * (2.0 / 64)).astype(np.float32)
    s["px"], s["py"] = x, y
    for aa in (0, 1):
        s["desc"]["aa_lines"] = aa
        ref, _ = render_ref(s)
        assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.parametrize("blend", [0, 1])
def test_points(ctx, blend):
    s = point_scene(20000, 200, 120, seed=37, blend=blend, stencil_test=0, stencil_op=4, stencil_wrap=1)
    ref, st = render_ref(s)
    assert st[0]["point_hits"] > 1000
    assert_parity(render_gpu(ctx, s), ref)


def test_points_value_stencil_and_tiles(ctx):
    s = point_scene(5000, 130, 129, seed=38, stencil_test=2, stencil_op=3, stencil_ref=3, tile_w=64, tile_h=50)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.parametrize("gs", [1, 2])
def test_clip_programs_lines_points(ctx, gs):
    s = line_scene(1500, 200, 140, seed=39)
    z = s["pz"].copy()
    z[::3] = -0.4
    s["pz"] = z
    s["desc"]["geometry_program"] = gs
    ref, st = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)
    p = point_scene(3000, 200, 140, seed=40, geometry_program=gs)
    p["pz"][::2] = -0.2
    ref, _ = render_ref(p)
    assert_parity(render_gpu(ctx, p), ref)


def test_clip_frame_reduced(ctx):
    # the 4K clipping/primitive frame's three draws, at 640x360
    s = scenes.clip_scene(20_000, 8_000, 4_000, 640, 360, seed=41)
    ref, st = render_ref(s)
    assert st[0]["gen_tris"] > 20_000
    assert_parity(render_gpu(ctx, s), ref)


def test_clip_frame_full_size(ctx):
    # the full 3840x2160 frame: 500k triangles (30% through the near clip),
    # 200k Wu lines with f32 alpha, 100k points, stencil increment; the
    # single-threaded oracle takes ~30 s on one core
    s = scenes.clip_scene()
    ref, st = render_ref(s)
    assert st[0]["gen_tris"] > 500_000 and st[1]["line_plots"] > 10_000_000
    assert_parity(render_gpu(ctx, s), ref)


def test_points_scan_past_1024_blocks():
    """4.5M points: the per-primitive count scan spans more than 1024 blocks
    of 4096 elements, so the scan takes its three-level path (block sums
    scanned by a top-level pass) rather than the fused one."""
    import softrender_hip as sr

    s = point_scene(4_500_000, 256, 192, seed=35, blend=1, tile_w=100, tile_h=64)
    ctx = sr.Context(0)
    try:
        got = render_gpu(ctx, s)
        assert ctx.last_stats()["points_in"] == 4_500_000
    finally:
        ctx.close()
    ref, st = render_ref(s)
    assert_parity(got, ref)
