"""Batched rendering (sr_batch_draw): every framebuffer of the batch equals
that scene rendered alone by the oracle, bit for bit — across graph replays,
capacity overflow, mixed classes and items that fall back to sr_draw."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_ref

pytestmark = pytest.mark.gpu


def make_items(ctx, specs, tex):
    items, keep = [], []
    for s in specs:
        fb = ctx.framebuffer(s["width"], s["height"])
        fb.clear(0xFF00FF00)  # a stale frame the fused clear must erase
        mesh = ctx.mesh(s)
        desc = sr.make_desc(**s["desc"])
        keep.append((fb, mesh, desc))
        items.append((fb, mesh, desc, tex if s.get("texture") else None, s.get("clear", 0xFF000000)))
    return items, keep


def check_all(items, specs):
    for (fb, mesh, desc, tex, clear), s in zip(items, specs):
        ref, _ = render_ref(dict(s, clear=clear))
        assert_parity(fb.read(), ref)


def mix(seed=50, n=10):
    out = []
    rng = scenes.SplitMix64(seed)
    res = [(256, 256), (512, 512), (320, 180), (97, 61), (640, 360)]
    for i in range(n):
        w, h = res[i % len(res)]
        nt = int(200 + 4000 * rng.uniform(1)[0])
        s = scenes.soup_scene(nt, w, h, seed=seed + i)
        if i % 3 == 1:
            s["desc"]["blend"] = 0
        if i % 4 == 2:
            s["desc"]["exec_tile"] = 16
        out.append(s)
    return out


def test_batch_matches_oracle(ctx):
    specs = mix()
    tw, th, rgba = specs[0]["texture"]
    tex = ctx.texture(tw, th, rgba)
    items, keep = make_items(ctx, specs, tex)
    ctx.batch_draw(items)
    check_all(items, specs)
    # replay (the cached graph) after dirtying the framebuffers
    for fb, *_ in items:
        fb.clear(0xFF123456)
    ctx.batch_draw(items)
    ctx.batch_draw(items)
    check_all(items, specs)


def test_batch_capacity_overflow_regrows(ctx):
    specs = mix(seed=60, n=6)
    tw, th, rgba = specs[0]["texture"]
    tex = ctx.texture(tw, th, rgba)
    c2 = sr.Context(0)
    try:
        tex2 = c2.texture(tw, th, rgba)
        c2.set_batch_capacity(64)
        items, keep = make_items(c2, specs, tex2)
        c2.batch_draw(items)
        assert c2.last_stats()["bin_pairs"] > 64
        check_all(items, specs)
    finally:
        c2.close()


def test_batch_alone_equals_in_the_mix(ctx):
    specs = mix(seed=70, n=8)
    tw, th, rgba = specs[0]["texture"]
    tex = ctx.texture(tw, th, rgba)
    items, keep = make_items(ctx, specs, tex)
    ctx.batch_draw(items)
    mixed = [fb.read() for fb, *_ in items]
    for k, s in enumerate(specs):
        alone, keep1 = make_items(ctx, [s], tex)
        ctx.batch_draw(alone)
        a = alone[0][0].read()
        for x, y in zip(a, mixed[k]):
            assert np.array_equal(x, y)


def test_batch_with_fallback_items(ctx):
    tri = scenes.soup_scene(1500, 200, 120, seed=80)
    sphere = scenes.sphere_scene(150, 150)  # Blinn-Phong, culled
    stencil = scenes.soup_scene(800, 128, 96, seed=81)
    stencil["desc"].update(stencil_test=0, stencil_op=4, stencil_wrap=0)  # falls back to sr_draw
    clip = scenes.soup_scene(800, 160, 90, seed=82)
    clip["desc"]["geometry_program"] = 2  # falls back
    thin = scenes.soup_scene(50, 1, 40, seed=83)  # no tiles at all
    specs = [tri, sphere, stencil, clip, thin]
    tw, th, rgba = tri["texture"]
    tex = ctx.texture(tw, th, rgba)
    items, keep = make_items(ctx, specs, tex)
    ctx.batch_draw(items)
    check_all(items, specs)


def test_batch_same_mesh_twice_into_one_framebuffer(ctx):
    # the second draw of the same triangles lands on the first: with over
    # blending every covered pixel blends twice, with depth ties (d >= dt)
    s = scenes.soup_scene(400, 64, 64, seed=84)
    tex = ctx.texture(*s["texture"])
    fb = ctx.framebuffer(64, 64)
    mesh = ctx.mesh(s)
    d = sr.make_desc(**s["desc"])
    ctx.batch_draw([(fb, mesh, d, tex, 0xFF000000), (fb, mesh, d, tex)])
    ref, _ = render_ref(dict(s, clear=0xFF000000, draws=[s, s]))
    assert_parity(fb.read(), ref)
    mesh.free()
    fb.free()
    tex.free()


def _multi_draw_scene(W=300, H=200, seed=70):
    # three soups drawn in order into one framebuffer, each with its own
    # alpha (fragment uniforms), over screen-tile seams (tile 64)
    draws = []
    for i, alpha in enumerate((0.8, 0.5, 0.95)):
        s = scenes.soup_scene(3000 + 1500 * i, W, H, seed=seed + i, alpha=alpha)
        s["desc"] = dict(s["desc"], tile_w=64, tile_h=64)
        draws.append(s)
    return dict(width=W, height=H, clear=0xFF102030, texture=draws[0]["texture"], draws=draws)


def test_batch_many_draws_into_one_framebuffer(ctx):
    """Several render_mesh chains into one framebuffer in one batched pass
    equal the same draws one after another (the oracle draws them in order)."""
    ms = _multi_draw_scene()
    other = scenes.soup_scene(2500, 160, 120, seed=80)
    tw, th, rgba = ms["texture"]
    tex = ctx.texture(tw, th, rgba)
    fb = ctx.framebuffer(ms["width"], ms["height"])
    fb.clear(0xFF00FF00)
    fb2 = ctx.framebuffer(other["width"], other["height"])
    meshes = [ctx.mesh(d) for d in ms["draws"]]
    descs = [sr.make_desc(**d["desc"]) for d in ms["draws"]]
    m2, d2 = ctx.mesh(other), sr.make_desc(**other["desc"])
    # interleaved with another framebuffer's item: per framebuffer, item order counts
    items = [(fb, meshes[0], descs[0], tex, ms["clear"]), (fb2, m2, d2, tex, 0xFF000000),
             (fb, meshes[1], descs[1], tex), (fb, meshes[2], descs[2], tex)]
    ctx.batch_draw(items)
    ref, st = render_ref(ms)
    assert sum(x["shaded"] for x in st) > 5000
    assert_parity(fb.read(), ref)
    ref2, _ = render_ref(dict(other, clear=0xFF000000))
    assert_parity(fb2.read(), ref2)
    # replay (cached graph), and a group that cannot be batched (a later
    # item clears, so the group is drawn one by one): same results
    fb.clear(0xFF123456)
    ctx.batch_draw(items)
    assert_parity(fb.read(), ref)
    ctx.batch_draw(items[:3] + [(fb, meshes[2], descs[2], tex, ms["clear"])])
    ref3, _ = render_ref(dict(ms, draws=[ms["draws"][2]]))
    assert_parity(fb.read(), ref3)
    for m in meshes + [m2]:
        m.free()
    for f in (fb, fb2):
        f.free()
    tex.free()


def test_batch_nan_coordinate_draws_nothing(ctx):
    """A NaN screen coordinate in one item: sr_batch_draw reports it like
    sr_draw does, and no framebuffer of the one-pass group changes (the
    fused clears included), on a fresh batch and on a replayed graph."""
    specs = mix(seed=90, n=5)
    bad = dict(specs[2])
    px = bad["px"].copy()
    # every x of the second triangle: f32::min/max skip a NaN operand, so one
    # NaN vertex leaves the box finite (and never shades); three make the
    # box NaN, which the reference's cast panics on (triangle.rs:70)
    px[3:6] = np.float32("nan")
    bad["px"] = px
    specs[2] = bad
    tw, th, rgba = specs[0]["texture"]
    tex = ctx.texture(tw, th, rgba)
    items, keep = make_items(ctx, specs, tex)
    before = [fb.read() for fb, *_ in items]
    for _ in range(2):
        with pytest.raises(sr.SoftrenderError) as e:
            ctx.batch_draw(items)
        assert e.value.code == -4  # SR_ERR_INVALID_COORDINATE
        for (fb, *_), b in zip(items, before):
            for x, y in zip(fb.read(), b):
                assert np.array_equal(x, y)
    # the same batch without the NaN draws normally afterwards
    good = mix(seed=90, n=5)
    items2, keep2 = make_items(ctx, good, tex)
    ctx.batch_draw(items2)
    check_all(items2, good)
