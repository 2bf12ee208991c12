"""GPU parity: the HIP path through the C ABI against the oracle, bit for bit.

Coverage and depth must match exactly (depth planes compared bitwise, which
also compares coverage); stencil exactly; colour exactly (the tolerance the
contract allows is 1 LSB, the programs here are built to be exact).
"""
import numpy as np
import pytest

import scenes
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu


def test_sphere_512(ctx):
    s = scenes.sphere_scene()
    ref, st = render_ref(s)
    assert st[0]["strip_shaded"] == 0
    r = assert_parity(render_gpu(ctx, s), ref)
    assert st[0]["shaded"] > 50_000, st


@pytest.mark.parametrize("ets", [8, 16, 32])
def test_soup_exec_tile_sizes(ctx, ets):
    s = scenes.soup_scene(6000, 320, 180, seed=11)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s, dict(exec_tile=ets)), ref)


@pytest.mark.parametrize("tile", [(128, 128), (100, 37), (7, 300), (1, 1), (640, 640)])
def test_soup_semantic_tiles(ctx, tile):
    # the reference tile size decides which pixels are rasterised twice
    s = scenes.soup_scene(3000, 257, 131, seed=12)
    ov = dict(tile_w=tile[0], tile_h=tile[1])
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("cull", [0, 1, 2])
def test_cull(ctx, cull):
    s = scenes.sphere_scene(200, 160)
    ov = dict(cull=cull)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("op,wrap", [(4, 0), (4, 1), (5, 1), (1, 0), (3, 0), (2, 0)])
def test_stencil_always(ctx, op, wrap):
    s = scenes.soup_scene(2500, 300, 170, seed=13)
    ov = dict(stencil_test=0, stencil_op=op, stencil_wrap=wrap, stencil_ref=7, tile_w=64, tile_h=48)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


def test_stencil_strips_many_triangles(ctx):
    # more triangles than one pass of the strip kernel's work-groups takes
    # (512 x 256): each work-group accumulates several rounds in LDS
    s = scenes.soup_scene(150_000, 320, 200, seed=21)
    ov = dict(stencil_test=0, stencil_op=4, stencil_wrap=1, stencil_ref=0, tile_w=64, tile_h=48)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


def test_stencil_strips_wide_framebuffer(ctx):
    # W + H too large for the LDS copies of the strip rows: global atomics
    s = scenes.soup_scene(1500, 16500, 6, seed=22)
    ov = dict(stencil_test=0, stencil_op=4, stencil_wrap=1, stencil_ref=0, tile_w=64, tile_h=4)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("test,op", [(2, 5), (6, 3), (7, 4), (4, 0)])
def test_stencil_value_tests(ctx, test, op):
    s = scenes.soup_scene(800, 160, 96, seed=14)
    ov = dict(stencil_test=test, stencil_op=op, stencil_wrap=1, stencil_ref=2, tile_w=32, tile_h=32)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("flt,edge", [(0, 0), (1, 0), (0, 1), (1, 1), (1, 2)])
def test_sampler_modes(ctx, flt, edge):
    s = scenes.soup_scene(3000, 256, 144, seed=15)
    # push uv outside [0,1] so wrap/border/clamp differ
    s["u"] = (s["u"] * 3.0 - 1.0).astype(np.float32)
    s["v"] = (s["v"] * 3.0 - 1.0).astype(np.float32)
    ov = dict(tex_filter=flt, tex_edge=edge, tex_border=0xFF3366CC)
    ref, _ = render_ref(s, ov)
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("tw,th", [(2, 1), (3, 2), (5, 7), (1, 4), (64, 3)])
@pytest.mark.parametrize("flt,edge", [(1, 0), (0, 0), (1, 1), (1, 2)])
def test_sampler_small_random_textures(ctx, tw, th, flt, edge):
    # every texel differs from its neighbours, and uv runs well past [0, 1]:
    # a wrong texel of a bilinear pair at a clamped edge (or the 1-wide case)
    # changes the image
    s = scenes.soup_scene(3000, 256, 144, seed=18)
    s["u"] = (s["u"] * 3.0 - 1.0).astype(np.float32)
    s["v"] = (s["v"] * 3.0 - 1.0).astype(np.float32)
    rng = np.random.default_rng(tw * 31 + th)
    s["texture"] = (tw, th, rng.integers(0, 1 << 32, tw * th, dtype=np.uint64).astype(np.uint32))
    ov = dict(tex_filter=flt, tex_edge=edge, tex_border=0xFF3366CC)
    ref, st = render_ref(s, ov)
    assert st[0]["shaded"] > 1000
    assert_parity(render_gpu(ctx, s, ov), ref)


def test_exhaustive_matches_binned(ctx):
    s = scenes.soup_scene(1500, 200, 120, seed=16)
    a = render_gpu(ctx, s, dict(exec_flags=1))
    b = render_gpu(ctx, s)
    ref, _ = render_ref(s)
    assert_parity(a, ref)
    assert_parity(b, ref)


def test_soup_1080p_100k(ctx):
    s = scenes.soup_scene(100_000, 1920, 1080, seed=17)
    ref, st = render_ref(s)
    assert st[0]["strip_shaded"] == 0
    assert_parity(render_gpu(ctx, s), ref)


def test_soup_1080p_full_size(ctx):
    # the 1920x1080, 1M-triangle frame the benchmark draws (oracle: ~15 s)
    s = scenes.soup_scene(1_000_000, 1920, 1080)
    ref, st = render_ref(s)
    assert st[0]["shaded"] == 6973823, st[0]["shaded"]
    assert_parity(render_gpu(ctx, s), ref)


def test_degenerate_and_tiny(ctx):
    # repeated vertices, collinear triangles, zero-area, and sub-pixel ones
    p = np.array([[-0.5, -0.5, 0.3], [-0.5, -0.5, 0.3], [0.2, 0.4, 0.3],
                  [-0.9, -0.9, 0.1], [0.0, 0.0, 0.1], [0.9, 0.9, 0.1],
                  [0.1, 0.1, 0.5], [0.1001, 0.1, 0.5], [0.1, 0.1003, 0.5],
                  [-1.0, 1.0, 0.2], [1.0, 1.0, 0.2], [-1.0, -1.0, 0.2]], np.float32)
    s = scenes.soup_scene(4, 64, 48, seed=18)
    s.update(px=p[:, 0].copy(), py=p[:, 1].copy(), pz=p[:, 2].copy(),
             nx=np.ones(12, np.float32), ny=np.zeros(12, np.float32), nz=np.zeros(12, np.float32),
             u=np.linspace(0, 1, 12).astype(np.float32), v=np.linspace(1, 0, 12).astype(np.float32))
    for ov in (dict(), dict(stencil_test=0, stencil_op=4, stencil_wrap=0, tile_w=16, tile_h=16)):
        ref, _ = render_ref(s, ov)
        assert_parity(render_gpu(ctx, s, ov), ref)


def _lattice_scene(W=64, H=64, n=600, seed=23):
    """Triangles whose vertices sit on the half-pixel lattice (pixel centres
    and corners), so edges pass exactly through pixel centres (barycentric
    exactly 0), plus slivers from nearly collinear lattice points and
    degenerate ones: the cases where binning's coverage test has no slack."""
    rng = scenes.SplitMix64(seed)
    q = np.floor(rng.uniform(6 * n, 0, 2 * W + 1)).reshape(n, 3, 2) / 2.0  # screen coords, step 0.5
    q[:, :, 1] *= H / W
    k = n // 3
    # slivers: third vertex one half-step off the line through the first two
    q[k:2 * k, 2] = (q[k:2 * k, 0] + q[k:2 * k, 1]) / 2.0
    q[k:2 * k, 2, 0] += 0.5
    q[2 * k:2 * k + 20, 1] = q[2 * k:2 * k + 20, 0]  # repeated vertex (det 0)
    sx = q[:, :, 0].reshape(-1)
    sy = q[:, :, 1].reshape(-1)
    s = scenes.soup_scene(n, W, H, seed=seed)
    # screen -> NDC through the exact inverse of the viewport transform (W, H powers of two)
    s["px"] = (sx * (2.0 / W) - 1.0).astype(np.float32)
    s["py"] = (1.0 - sy * (2.0 / H)).astype(np.float32)
    s["pz"] = np.repeat(rng.uniform(n, 0.05, 0.95), 3).astype(np.float32)
    return s


@pytest.mark.parametrize("ets", [8, 16])
@pytest.mark.parametrize("blend", [0, 1])
def test_lattice_edges_through_pixel_centres(ctx, ets, blend):
    s = _lattice_scene()
    ov = dict(exec_tile=ets, blend=blend)
    ref, st = render_ref(s, ov)
    assert st[0]["covered"] > 1000, st
    assert_parity(render_gpu(ctx, s, ov), ref)


@pytest.mark.parametrize("wh", [(1, 50), (50, 1), (2, 2), (3, 200)])
def test_thin_framebuffers(ctx, wh):
    s = scenes.soup_scene(400, wh[0], wh[1], seed=19)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.parametrize("gs", [1, 2])
def test_clip_programs_triangles(ctx, gs):
    s = scenes.soup_scene(2000, 240, 160, seed=20)
    rng = scenes.SplitMix64(99)
    z = s["pz"].reshape(-1, 3).copy()
    z[:, 0] = np.where(rng.uniform(len(z)) < 0.4, -0.3, z[:, 0])
    x = s["px"].reshape(-1, 3).copy()
    x[:, 1] = np.where(rng.uniform(len(x)) < 0.2, x[:, 1] * 3.0, x[:, 1])
    s["pz"] = z.reshape(-1).astype(np.float32)
    s["px"] = x.reshape(-1).astype(np.float32)
    ov = dict(geometry_program=gs)
    ref, st = render_ref(s, ov)
    assert st[0]["gen_tris"] > 0
    assert_parity(render_gpu(ctx, s, ov), ref)


def test_accumulating_draws(ctx):
    # two draws into one framebuffer: the second reads what the first left
    a = scenes.soup_scene(1500, 200, 130, seed=21)
    b = scenes.soup_scene(1500, 200, 130, seed=22)
    both = dict(width=200, height=130, clear=0xFF101010, texture=a["texture"], draws=[a, b])
    ref, _ = render_ref(both)
    assert_parity(render_gpu(ctx, both), ref)


@pytest.mark.parametrize("flt,edge", [(1, 0), (0, 1), (1, 2)])
def test_render_to_texture_two_passes(ctx, flt, edge):
    # pass 1: the lit sphere into a 512x512 framebuffer; pass 2: a soup that
    # samples pass 1's colour plane through a texture view (no copy).  The
    # oracle runs both passes, feeding its own pass-1 image to pass 2.
    import softrender_hip as sr

    s1 = scenes.sphere_scene()
    ref1, _ = render_ref(s1)
    fb1 = ctx.framebuffer(s1["width"], s1["height"])
    fb1.clear(s1.get("clear", 0))
    m1 = ctx.mesh(s1)
    ctx.draw(fb1, m1, sr.make_desc(**s1["desc"]))
    tex = ctx.texture_from_fb(fb1)
    s2 = scenes.soup_scene(20_000, 320, 240, seed=61)
    s2["desc"] = dict(s2["desc"], tex_filter=flt, tex_edge=edge, tex_border=0xFF3366CC)
    s2["u"] = (s2["u"] * 1.5 - 0.25).astype(np.float32)  # reach past the edges
    s2["texture"] = (s1["width"], s1["height"], ref1.color.copy())
    ref2, st = render_ref(s2)
    assert st[0]["shaded"] > 10_000
    fb2 = ctx.framebuffer(s2["width"], s2["height"])
    fb2.clear(s2["clear"])
    m2 = ctx.mesh(s2)
    ctx.draw(fb2, m2, sr.make_desc(**s2["desc"]), tex)
    assert_parity(fb2.read(), ref2)
    # sampling the framebuffer being drawn is refused
    with pytest.raises(sr.SoftrenderError) as e:
        ctx.draw(fb1, m2, sr.make_desc(**s2["desc"]), tex)
    assert e.value.code == -1
    # so is a batch item sampling a framebuffer the same batch draws
    with pytest.raises(sr.SoftrenderError):
        ctx.batch_draw([(fb1, m1, sr.make_desc(**s1["desc"]), None, 0xFF000000),
                        (fb2, m2, sr.make_desc(**s2["desc"]), tex, 0xFF000000)])
    tex.free()
    for m in (m1, m2):
        m.free()
    for fb in (fb1, fb2):
        fb.free()
