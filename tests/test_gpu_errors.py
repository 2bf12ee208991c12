"""Error paths through the C ABI.  Where the reference panics (assert!,
cast(NaN).unwrap()) the library returns a status and leaves the
framebuffer as the draws before it left it; empty inputs draw nothing.
Every case is rejected on the host or by a device-side flag the host reads,
so no kernel runs out of bounds."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

INVALID_ARGUMENT, INVALID_COORDINATE = -1, -4


def tri_scene(n=200, w=64, h=48, seed=5):
    s = scenes.soup_scene(n, w, h, seed=seed)
    s["desc"] = dict(s["desc"], fragment_program=0, fs_uniforms=np.r_[0.2, 0.6, 0.9, 1.0, np.zeros(28)])
    return s


def draw_status(ctx, fb, s, **over):
    mesh = ctx.mesh(s)
    try:
        ctx.draw(fb, mesh, sr.make_desc(**dict(s["desc"], **over)))
        return 0
    except sr.SoftrenderError as e:
        return e.code
    finally:
        mesh.free()


def test_one_nan_vertex_is_not_an_error(ctx):
    # the box takes IEEE min/max, which skip a NaN (triangle.rs:60-64), so one
    # NaN vertex leaves a finite box: no panic, the triangle covers nothing
    s = tri_scene()
    s["px"] = s["px"].copy()
    s["px"][7] = np.nan
    ref, st = render_ref(s)
    assert st[0]["invalid"] == 0
    assert_parity(render_gpu(ctx, s), ref)


def test_nan_box_is_an_error_and_draws_nothing(ctx):
    # all three x NaN: clamp_as_int! casts NaN, cast(..).unwrap() panics (triangle.rs:70)
    s = tri_scene()
    fb = ctx.framebuffer(s["width"], s["height"])
    fb.clear(0xFF102030)
    before = fb.read()
    s["px"] = s["px"].copy()
    s["px"][6:9] = np.nan
    assert draw_status(ctx, fb, s) == INVALID_COORDINATE
    for a, b in zip(fb.read(), before):
        assert (a.view(np.uint32) == b.view(np.uint32)).all()
    # the context stays usable
    assert draw_status(ctx, fb, tri_scene(seed=6)) == 0
    fb.free()


def test_nan_line_endpoint_is_an_error(ctx):
    # line.rs:112-116: cast of a non-finite coordinate unwraps
    n = 50
    rng = np.random.default_rng(2)
    x = rng.uniform(-0.9, 0.9, 2 * n).astype(np.float32)
    y = rng.uniform(-0.9, 0.9, 2 * n).astype(np.float32)
    x[9] = np.inf
    s = dict(primitive=2, px=x, py=y, pz=np.full(2 * n, 0.5, np.float32), nx=None, ny=None, nz=None,
             u=np.zeros(2 * n, np.float32), v=np.zeros(2 * n, np.float32), indices=None,
             desc=dict(vertex_program=0, fragment_program=0, fs_uniforms=np.r_[1.0, 1.0, 1.0, 1.0, np.zeros(28)],
                       vs_uniforms=scenes.vs_uniforms(np.eye(4)), viewport=(0.0, 0.0, 80.0, 60.0, 0.0, 1.0)))
    fb = ctx.framebuffer(80, 60)
    fb.clear(0)
    assert draw_status(ctx, fb, s) == INVALID_COORDINATE
    fb.free()


def test_empty_index_list_draws_nothing(ctx):
    s = tri_scene()
    s["indices"] = np.zeros(0, np.uint32)
    fb = ctx.framebuffer(s["width"], s["height"])
    fb.clear(0xFF000000)
    before = fb.read()
    try:
        mesh = ctx.mesh(s)
    except sr.SoftrenderError as e:  # an upload may refuse it outright
        assert e.code == INVALID_ARGUMENT
    else:
        ctx.draw(fb, mesh, sr.make_desc(**s["desc"]))
        mesh.free()
        for a, b in zip(fb.read(), before):
            assert (a.view(np.uint32) == b.view(np.uint32)).all()
    fb.free()


def test_index_count_not_a_multiple_of_the_primitive(ctx):
    # pipeline/mod.rs:148: assert_eq!(indices.len() % T::num_vertices(), 0)
    s = tri_scene()
    s["indices"] = np.arange(3 * 10 + 1, dtype=np.uint32)
    with pytest.raises(sr.SoftrenderError) as e:
        ctx.mesh(s)
    assert e.value.code == INVALID_ARGUMENT


def test_zero_sized_framebuffer_is_refused(ctx):
    # pipeline/mod.rs:129-130 asserts a non-zero size
    for w, h in ((0, 16), (16, 0)):
        with pytest.raises(sr.SoftrenderError) as e:
            ctx.framebuffer(w, h)
        assert e.value.code == INVALID_ARGUMENT


@pytest.mark.parametrize("field,value", [("exec_tile", 12), ("exec_tile", 64), ("exec_waves", 3),
                                         ("vertex_program", 9), ("geometry_program", 7),
                                         ("fragment_program", 11), ("stencil_test", 8), ("stencil_op", 6)])
def test_bad_descriptor_fields_are_refused(ctx, field, value):
    s = tri_scene()
    fb = ctx.framebuffer(s["width"], s["height"])
    assert draw_status(ctx, fb, s, **{field: value}) == INVALID_ARGUMENT
    fb.free()


def test_texture_program_needs_a_texture(ctx):
    s = tri_scene()
    fb = ctx.framebuffer(s["width"], s["height"])
    assert draw_status(ctx, fb, s, fragment_program=2) == INVALID_ARGUMENT
    fb.free()


def test_never_stencil_draws_nothing(ctx):
    # stencil.rs:115: Never passes no pixel
    s = tri_scene()
    fb = ctx.framebuffer(s["width"], s["height"])
    fb.clear(0xFF000000)
    before = fb.read()
    assert draw_status(ctx, fb, s, stencil_test=1) == 0
    for a, b in zip(fb.read(), before):
        assert (a.view(np.uint32) == b.view(np.uint32)).all()
    fb.free()
