"""Normal visualisation geometry programs on the device against the oracle:
the full example's frame (full_example/src/main.rs:66-92) — each model lit,
then its vertex or face normals drawn over it as green lines — and the
pass-through of line and point meshes (re_emit, shaders.rs:59,88)."""
import os

import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

SUZANNE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "suzanne.obj")
VN, FN = scenes.GS_VERTEX_NORMALS, scenes.GS_FACE_NORMALS


def suzanne_normals(program, aa, w=320, h=240):
    m = sr.parse_obj(SUZANNE)
    return scenes.normals_scene(m["positions"], m["normals"], m["indices"], w, h, program=program, aa=aa)


@pytest.mark.parametrize("aa", [0, 1])
@pytest.mark.parametrize("program", [VN, FN])
def test_full_example_frame(ctx, program, aa):
    s = suzanne_normals(program, aa)
    ref, st = render_ref(s)
    ntri = len(s["draws"][1]["indices"]) // 3
    assert st[1]["gen_lines"] == (3 if program == VN else 1) * ntri
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.parametrize("program", [VN, FN])
def test_sphere_normals_wide_frame(ctx, program):
    pos, nrm, idx = scenes.uv_sphere(24, 48)
    s = scenes.normals_scene(pos, nrm, idx, 640, 360, program=program, aa=1, length=0.2)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


def test_generated_line_count(ctx):
    s = suzanne_normals(VN, 0, 96, 96)
    d = s["draws"][1]
    fb = ctx.framebuffer(96, 96)
    fb.clear(0)
    mesh = ctx.mesh(d)
    ctx.draw(fb, mesh, sr.make_desc(**d["desc"]))
    assert ctx.last_stats()["generated_lines"] == 3 * (len(d["indices"]) // 3)
    mesh.free()
    fb.free()


@pytest.mark.parametrize("primitive", [1, 2])
@pytest.mark.parametrize("program", [VN, FN])
def test_lines_and_points_pass_through(ctx, primitive, program):
    rng = np.random.default_rng(11)
    n = 120
    pos = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    view = scenes.look_at((0.3, 0.2, 3), (0, 0, 0), (0, 1, 0))
    PV = scenes.perspective_zero_to_one(60.0, 1.5, 0.1, 100.0) @ view
    fs = np.zeros(32, np.float32)
    fs[0:4] = (0.2, 0.9, 0.4, 0.7)
    s = dict(width=150, height=100, clear=0xFF000000, texture=None, primitive=primitive,
             px=pos[:, 0].copy(), py=pos[:, 1].copy(), pz=pos[:, 2].copy(),
             nx=np.zeros(n, np.float32), ny=np.zeros(n, np.float32), nz=np.ones(n, np.float32),
             u=None, v=None, indices=None,
             desc=dict(vertex_program=1, geometry_program=program, fragment_program=0, blend=1,
                       aa_lines=1, alpha_mode=1, fs_uniforms=fs, vs_uniforms=scenes.vs_uniforms(PV),
                       gs_uniforms=np.r_[PV.ravel(), 0.05], viewport=(0.0, 0.0, 150.0, 100.0, 0.0, 1.0)))
    ref, _ = render_ref(s)
    assert (ref.color != 0xFF000000).sum() > 20
    assert_parity(render_gpu(ctx, s), ref)


def test_normals_need_world_varyings(ctx):
    s = suzanne_normals(FN, 0, 32, 32)
    d = s["draws"][1]
    fb = ctx.framebuffer(32, 32)
    mesh = ctx.mesh(d)
    with pytest.raises(sr.SoftrenderError):
        ctx.draw(fb, mesh, sr.make_desc(**dict(d["desc"], vertex_program=0)))
    mesh.free()
    fb.free()


def test_chains_share_one_vertex_pass(ctx):
    """sr_draw_chains (duplicate): the lit chain and the normals chain from one
    vertex pass equal the two draws made one by one, and the oracle."""
    s = suzanne_normals(FN, 1)
    ref, _ = render_ref(s)
    apart = render_gpu(ctx, s)
    fb = ctx.framebuffer(s["width"], s["height"])
    fb.clear(s["clear"])
    mesh = ctx.mesh(s["draws"][0])
    ctx.draw_chains(fb, mesh, [sr.make_desc(**d["desc"]) for d in s["draws"]])
    shared = fb.read()
    for a, b in zip(shared, apart):
        assert (a.view(np.uint32) == b.view(np.uint32)).all()
    assert_parity(shared, ref)
    mesh.free()
    fb.free()


@pytest.mark.parametrize("gs", [0, 1, 2, VN])
def test_chains_with_each_geometry_program(ctx, gs):
    pos, nrm, idx = scenes.uv_sphere(12, 24)
    s = scenes.normals_scene(pos, nrm, idx, 200, 150, program=VN, aa=0)
    lit, lines = s["draws"]
    second = dict(lines, desc=dict(lines["desc"], geometry_program=gs)) if gs == VN else \
        dict(lit, desc=dict(lit["desc"], geometry_program=gs, blend=1, cull=0))
    s["draws"] = [lit, second]
    ref, _ = render_ref(s)
    fb = ctx.framebuffer(200, 150)
    fb.clear(s["clear"])
    mesh = ctx.mesh(lit)
    ctx.draw_chains(fb, mesh, [sr.make_desc(**d["desc"]) for d in s["draws"]])
    assert_parity(fb.read(), ref)
    mesh.free()
    fb.free()


def test_chains_must_share_the_vertex_program(ctx):
    s = suzanne_normals(FN, 0, 32, 32)
    fb = ctx.framebuffer(32, 32)
    mesh = ctx.mesh(s["draws"][0])
    d0 = sr.make_desc(**s["draws"][0]["desc"])
    d1 = sr.make_desc(**dict(s["draws"][1]["desc"], vs_uniforms=np.zeros(48, np.float32)))
    with pytest.raises(sr.SoftrenderError):
        ctx.draw_chains(fb, mesh, [d0, d1])
    mesh.free()
    fb.free()
