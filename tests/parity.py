"""Helpers shared by the parity tests: run a scene on the GPU (through the
C ABI) and on the oracle, and compare framebuffers."""
from __future__ import annotations

import numpy as np

import oracle
import softrender_hip as sr


def render_gpu(ctx, scene, overrides=None, fb=None, keep=False):
    """Clear + every draw of `scene` through sr_draw; returns (color, depth, stencil)."""
    W, H = scene["width"], scene["height"]
    own = fb is None
    if own:
        fb = ctx.framebuffer(W, H)
        fb.clear(scene.get("clear", 0))
    tex = None
    if scene.get("texture") is not None:
        tw, th, rgba = scene["texture"]
        tex = ctx.texture(tw, th, rgba)
    for d in scene.get("draws") or [scene]:
        desc_kw = dict(d["desc"])
        if overrides:
            desc_kw.update(overrides)
        mesh = ctx.mesh(d)
        try:
            ctx.draw(fb, mesh, sr.make_desc(**desc_kw), tex)
        finally:
            mesh.free()
    out = fb.read()
    if tex:
        tex.free()
    if own and not keep:
        fb.free()
    return out


def render_ref(scene, overrides=None, nthreads=1):
    fb = oracle.Framebuffer(scene["width"], scene["height"], scene.get("clear", 0))
    stats = []
    for d in scene.get("draws") or [scene]:
        desc_kw = dict(d["desc"])
        if overrides:
            desc_kw.update(overrides)
        rc, st = oracle.draw(fb, d, sr.make_desc(**desc_kw), scene.get("texture"), nthreads)
        assert rc == 0, rc
        stats.append(st)
    return fb, stats


def diff(gpu, ref):
    """Counts of mismatching pixels: coverage/depth (bitwise), stencil, and
    the max per-channel colour difference."""
    col, dep, st = gpu
    dep_bad = int((dep.view(np.uint32) != ref.depth.view(np.uint32)).sum())
    st_bad = int((st != ref.stencil).sum())
    a = col.view(np.uint8).reshape(-1, 4).astype(np.int32)
    b = ref.color.view(np.uint8).reshape(-1, 4).astype(np.int32)
    d = np.abs(a - b)
    return dict(depth=dep_bad, stencil=st_bad, color_px=int((d.max(1) > 0).sum()),
                color_max=int(d.max()) if d.size else 0)


def assert_parity(gpu, ref, color_tol=0):
    r = diff(gpu, ref)
    assert r["depth"] == 0, r
    assert r["stencil"] == 0, r
    assert r["color_max"] <= color_tol, r
    return r
