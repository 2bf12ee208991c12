"""OBJ ingestion on the device: a model uploaded by sr_mesh_load_obj renders
exactly like the same arrays uploaded by sr_mesh_upload, and like the
oracle drawing them (the suzanne example: lit with world-space normals)."""
import os

import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

SUZANNE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "suzanne.obj")


def suzanne_scene():
    m = sr.parse_obj(SUZANNE)
    s = scenes.sphere_scene(320, 320)
    p, n, uv = m["positions"], m["normals"], m["uvs"]
    s.update(px=p[:, 0].copy(), py=p[:, 1].copy(), pz=p[:, 2].copy(), nx=n[:, 0].copy(),
             ny=n[:, 1].copy(), nz=n[:, 2].copy(), u=uv[:, 0].copy(), v=uv[:, 1].copy(),
             indices=m["indices"].copy(), primitive=3)
    s["desc"] = dict(s["desc"], cull=0)
    return s


def test_obj_mesh_renders_like_the_oracle(ctx):
    s = suzanne_scene()
    ref, st = render_ref(s)
    assert st[0]["shaded"] > 10_000
    fb = ctx.framebuffer(s["width"], s["height"])
    fb.clear(s.get("clear", 0))
    mesh = ctx.mesh_from_obj(SUZANNE)
    ctx.draw(fb, mesh, sr.make_desc(**s["desc"]))
    assert_parity(fb.read(), ref)
    assert_parity(render_gpu(ctx, s), ref)
    mesh.free()
    fb.free()
