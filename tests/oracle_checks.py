"""Independent checks of the oracle's triangle path.

Each check_* function takes an oracle library (oracle.load / oracle.lib)
and raises AssertionError when the library disagrees with something fixed
independently of its code:

  * exact values the reference's arithmetic fixes: a pixel centre on a
    vertex gives barycentrics (1, 0, 0) exactly, so depth and varyings
    there equal that vertex's values bit for bit
    (src/pipeline/stages/rasterization/triangle.rs:108-134,
    src/numeric/interpolate.rs:26-28);
  * float64 evaluations of linear quantities (depth at interior pixels,
    the vertex programs' mat-vecs);
  * hand-derived closed forms: the cull orientation of a known triangle
    (triangle.rs:54-61), Blinn-Phong and texture x Lambert at chosen
    vectors, OVER at alpha 128;
  * an independently written numpy-float32 restatement of the whole
    triangle draw (tests/np_restatement.py), compared bit for bit on
    random tiny scenes.

tests/test_oracle_independent.py runs them on the oracle;
tests/test_oracle_mutants.py runs them on oracle builds with planted
mistakes and requires each mistake to be caught.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import np_restatement as NR
import oracle
import scenes
import softrender_hip as sr

f32 = np.float32
F32MIN = np.float32(np.finfo(np.float32).min)


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def screen_tri_mesh(W, H, pts, zc, u=None, v=None, nrm=(0.0, 0.0, 1.0)):
    """Clip-space vertices (w = 1, identity MVP) that land exactly on the
    screen positions `pts` of a W x H viewport {0,0,W,H,0,1}; zc are clip z
    (screen z = -zc/2 - 1/2)."""
    pts = np.asarray(pts, np.float64)
    n = len(pts)
    return dict(px=(pts[:, 0] * 2.0 / W - 1.0).astype(np.float32),
                py=(1.0 - pts[:, 1] * 2.0 / H).astype(np.float32),
                pz=np.asarray(zc, np.float32),
                nx=np.full(n, nrm[0], np.float32), ny=np.full(n, nrm[1], np.float32),
                nz=np.full(n, nrm[2], np.float32),
                u=np.zeros(n, np.float32) if u is None else np.asarray(u, np.float32),
                v=np.zeros(n, np.float32) if v is None else np.asarray(v, np.float32),
                indices=None, primitive=3)


def desc_dict(d: sr.DrawDesc) -> dict:
    out = {}
    for name, _ in d._fields_:
        val = getattr(d, name)
        if name == "viewport":
            val = (val.x, val.y, val.width, val.height, val.near_z, val.far_z)
        elif name in ("vs_uniforms", "fs_uniforms", "gs_uniforms", "reserved"):
            val = np.array(val[:], np.float32 if name != "reserved" else np.int32)
        out[name] = val
    return out


def run_oracle(L, W, H, mesh, desc, clear=0xFF000000, tex=None, depth=None, stencil=None, color=None):
    fb = oracle.Framebuffer(W, H, clear)
    if depth is not None:
        fb.depth[:] = depth
    if stencil is not None:
        fb.stencil[:] = stencil
    if color is not None:
        fb.color[:] = color
    rc, st = oracle.draw(fb, mesh, desc, tex, 1, L=L)
    assert rc == 0, rc
    return fb, st


def identity_desc(W, H, **kw):
    base = dict(vertex_program=0, fragment_program=sr_fp("UV"), blend=0,
                vs_uniforms=scenes.vs_uniforms(np.eye(4)), viewport=(0.0, 0.0, float(W), float(H), 0.0, 1.0))
    base.update(kw)
    return sr.make_desc(**base)


def sr_fp(name):
    return {"FLAT": 0, "BP": 1, "TEX": 2, "UV": 3}[name]


def q8(x):
    """quant8 of a float64 value (nearest, ties up)."""
    return int(math.floor(min(max(x, 0.0), 1.0) * 255.0 + 0.5))


def channels(c):
    c = int(c)
    return [(c >> s) & 0xFF for s in (0, 8, 16, 24)]


# ---------------------------------------------------------------------------
# (a) barycentric weights: exact at the vertices, linear inside
# ---------------------------------------------------------------------------
BARY_TRIS = [
    # (screen positions on pixel centres, clip z, u, v)
    (((2.5, 3.5), (27.5, 6.5), (9.5, 28.5)), (0.5, -0.25, 0.875), (0.1, 0.5, 0.9), (0.8, 0.2, 0.45)),
    (((29.5, 1.5), (3.5, 12.5), (20.5, 30.5)), (-0.75, 0.25, 0.5), (0.9, 0.05, 0.6), (0.3, 0.7, 0.15)),
    (((5.5, 25.5), (30.5, 28.5), (16.5, 2.5)), (0.125, 0.625, -0.5), (0.33, 0.66, 0.99), (0.95, 0.5, 0.05)),
]


def check_bary_vertices_exact_and_linear_inside(L):
    W = H = 32
    for pts, zc, uu, vv in BARY_TRIS:
        for rot in range(3):  # every vertex in every slot of the triangle
            o = [(i + rot) % 3 for i in range(3)]
            P = [pts[i] for i in o]
            Z = [zc[i] for i in o]
            U = [uu[i] for i in o]
            V = [vv[i] for i in o]
            mesh = screen_tri_mesh(W, H, P, Z, U, V)
            fb, _ = run_oracle(L, W, H, mesh, identity_desc(W, H, fs_uniforms=np.r_[0.5, 1.0, np.zeros(30)]))
            zs = [np.float32(-0.5) * np.float32(z) - np.float32(0.5) for z in Z]  # exact for these z
            col = fb.color.reshape(H, W)
            dep = fb.depth.reshape(H, W)
            for (x, y), z, u, v in zip(P, zs, U, V):
                px, py = int(x), int(y)
                assert dep[py, px] == z, ("vertex depth", (px, py), dep[py, px], z)
                assert channels(col[py, px])[:2] == [q8(np.float32(u)), q8(np.float32(v))], (
                    "vertex varyings", (px, py), channels(col[py, px]), u, v)
            # interior pixels: depth and u, v are affine in the pixel centre
            (x1, y1), (x2, y2), (x3, y3) = P
            det = (y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3)
            covered = 0
            for py in range(H):
                for px in range(W):
                    x, y = px + 0.5, py + 0.5
                    a = ((y2 - y3) * (x - x3) + (x3 - x2) * (y - y3)) / det
                    b = ((y3 - y1) * (x - x3) + (x1 - x3) * (y - y3)) / det
                    c = 1.0 - a - b
                    if min(a, b, c) < 1e-4:
                        continue  # near an edge rounding may go either way
                    covered += 1
                    zw = a * zs[0] + b * zs[1] + c * zs[2]
                    assert abs(float(dep[py, px]) - zw) < 4e-6, ("interior depth", (px, py), dep[py, px], zw)
                    uw = a * U[0] + b * U[1] + c * U[2]
                    vw = a * V[0] + b * V[1] + c * V[2]
                    ch = channels(col[py, px])
                    assert abs(ch[0] - uw * 255.0) <= 0.51 and abs(ch[1] - vw * 255.0) <= 0.51, (
                        "interior varyings", (px, py), ch[:2], uw * 255, vw * 255)
            assert covered > 50


# ---------------------------------------------------------------------------
# (b) culling: (0,0),(8,0),(0,8) on screen has shoelace +64: CounterClockwise
# ---------------------------------------------------------------------------
def check_cull_orientation(L):
    W = H = 16
    ccw = [(0.0, 0.0), (8.0, 0.0), (0.0, 8.0)]
    cw = [ccw[0], ccw[2], ccw[1]]
    for pts, wind in ((ccw, 2), (cw, 1)):
        mesh = screen_tri_mesh(W, H, pts, (0.5, 0.5, 0.5))
        for cull in (0, 1, 2):
            fb, st = run_oracle(L, W, H, mesh, identity_desc(W, H, cull=cull, fragment_program=0,
                                                             fs_uniforms=np.r_[1, 1, 1, 1, np.zeros(28)]))
            drawn = int((fb.color != 0xFF000000).sum())
            if cull == wind:
                assert drawn == 0 and st["tri_culled"] == st["tri_calls"] > 0, (pts, cull, drawn)
            else:
                assert drawn == 36 and st["tri_culled"] == 0, (pts, cull, drawn)


# ---------------------------------------------------------------------------
# (c) vertex programs with non-symmetric matrices against float64
# ---------------------------------------------------------------------------
def check_vertex_programs_float64(L):
    rng = np.random.default_rng(21)
    n = 16
    mesh = dict(px=rng.uniform(-2, 2, n).astype(np.float32), py=rng.uniform(-2, 2, n).astype(np.float32),
                pz=rng.uniform(-2, 2, n).astype(np.float32), nx=rng.uniform(-1, 1, n).astype(np.float32),
                ny=rng.uniform(-1, 1, n).astype(np.float32), nz=rng.uniform(-1, 1, n).astype(np.float32),
                u=rng.uniform(0, 1, n).astype(np.float32), v=rng.uniform(0, 1, n).astype(np.float32),
                indices=None, primitive=3)
    mvp = rng.uniform(-3, 3, (4, 4)).astype(np.float32)
    mvp[3] = (0.1, -0.2, 0.3, 5.0)
    nm = rng.uniform(-2, 2, (3, 3)).astype(np.float32)
    model = rng.uniform(-2, 2, (3, 4)).astype(np.float32)
    assert np.abs(mvp - mvp.T).max() > 0.5 and np.abs(nm - nm.T).max() > 0.5
    for vp in (0, 1):
        d = sr.make_desc(vertex_program=vp, vs_uniforms=scenes.vs_uniforms(mvp, nm, model))
        for i in range(n):
            p = np.array([mesh["px"][i], mesh["py"][i], mesh["pz"][i], 1.0], np.float64)
            nn = np.array([mesh["nx"][i], mesh["ny"][i], mesh["nz"][i]], np.float64)
            clip, k = oracle.vertex(mesh, d, i, L=L)
            want = mvp.astype(np.float64) @ p
            assert np.allclose(clip, want, rtol=1e-5, atol=1e-5), ("clip", vp, i, clip, want)
            nw = nm.astype(np.float64) @ nn
            nw /= np.linalg.norm(nw)
            assert np.allclose(k[:3], nw, rtol=1e-5, atol=1e-6), ("normal", vp, i, k[:3], nw)
            if vp == 1:
                ww = model.astype(np.float64) @ p
                assert np.allclose(k[3:6], ww, rtol=1e-5, atol=1e-5), ("world", i, k[3:6], ww)
            else:
                assert k[3] == mesh["u"][i] and k[4] == mesh["v"][i]


# ---------------------------------------------------------------------------
# (d) the build-defined programs at hand-derived values
# ---------------------------------------------------------------------------
def _program_colour_at_vertex(L, fp, vp, fs, normal, model34=None, uv=(0.0, 0.0), tex=None, **kw):
    """Colour at the pixel whose centre is vertex 0 of a small triangle:
    barycentrics (1, 0, 0) exactly, so the varyings are vertex 0's."""
    W = H = 16
    mesh = screen_tri_mesh(W, H, [(3.5, 3.5), (12.5, 4.5), (5.5, 13.5)], (0.5, 0.5, 0.5),
                           u=[uv[0]] * 3, v=[uv[1]] * 3, nrm=normal)
    vsu = scenes.vs_uniforms(np.eye(4), None, model34)
    fsu = np.zeros(32, np.float32)
    fsu[:len(fs)] = fs
    d = identity_desc(W, H, vertex_program=vp, fragment_program=fp, vs_uniforms=vsu, fs_uniforms=fsu, **kw)
    fb, st = run_oracle(L, W, H, mesh, d, tex=tex)
    return channels(fb.color.reshape(H, W)[3, 3])


def check_blinn_phong_closed_form(L):
    # world position fixed at the origin (model rows = 0): view vector = eye
    model0 = np.zeros((3, 4))
    base, amb, kd, ks = (0.5, 0.25, 0.75), 0.1, 0.6, 0.4
    s32 = 0.9 ** 16  # (n.h)^32 with n.h = 3/sqrt(10)
    cases = [
        # light, eye (= view direction * distance), normal, n.l
        ((0.0, 0.0, 1.0), (0.0, 3.0, 4.0), (0.0, 0.0, 1.0), 1.0),
        ((0.0, 0.6, 0.8), (0.0, 0.0, 2.0), (0.0, 0.0, 1.0), 0.8),
        ((0.0, 0.0, 1.0), (0.0, 0.0, 5.0), (0.0, 0.0, -1.0), 0.0),  # facing away: ambient only
    ]
    for light, eye, nrm, ndl in cases:
        fs = [*light, *eye, *base, amb, kd, ks, 1.0]
        got = _program_colour_at_vertex(L, 1, 1, fs, nrm, model0)
        spec = ks * (s32 if ndl > 0 else 0.0)
        lit = amb + kd * ndl
        want = [q8(b * lit + spec) for b in base] + [255]
        assert all(abs(g - w) <= 1 for g, w in zip(got, want)), ("blinn-phong", light, eye, got, want)


def check_tex_lambert_at_texel_centre(L):
    tw = th = 8
    rgba = np.zeros(tw * th, np.uint32)
    rng = np.random.default_rng(5)
    for i in range(tw * th):
        r, g, b, a = (int(x) for x in rng.integers(0, 256, 4))
        rgba[i] = r | (g << 8) | (b << 16) | (a << 24)
    tex = (tw, th, rgba)
    light = (0.0, 0.6, 0.8)
    amb, kd, alpha = 0.2, 0.5, 0.8
    for (ix, iy) in [(0, 0), (3, 5), (7, 2), (6, 7)]:
        u, v = (ix + 0.5) / tw, (iy + 0.5) / th
        for filt in (0, 1):
            got = _program_colour_at_vertex(L, 2, 0, [*light, amb, kd, alpha], (0.0, 0.0, 1.0), uv=(u, v),
                                            tex=tex, tex_filter=filt)
            t = channels(rgba[iy * tw + ix])
            lit = amb + kd * 0.8
            want = [q8(t[c] / 255.0 * lit) for c in range(3)] + [q8(t[3] / 255.0 * alpha)]
            assert all(abs(g - w) <= 1 for g, w in zip(got, want)), ("tex-lambert", ix, iy, filt, got, want)


def check_over_blend_half_alpha(L):
    B = L.or_blend
    src = 200 | (100 << 8) | (50 << 16) | (128 << 24)
    # r = (200*128 + 10*127)/255 = 105.4, g = 60.2, b = 40.0, a = (128 + 255*127/255) = 255
    assert channels(B(1, src, 10 | (20 << 8) | (30 << 16) | (255 << 24))) == [105, 60, 40, 255]
    # destination alpha 100: a = 128 + 100*127/255 = 177.8
    assert channels(B(1, src, 10 | (20 << 8) | (30 << 16) | (100 << 24))) == [105, 60, 40, 178]
    assert B(0, src, 0xFFFFFFFF) == src


# ---------------------------------------------------------------------------
# (e) random tiny scenes against the numpy-float32 restatement, bit for bit
# ---------------------------------------------------------------------------
def random_scene(seed):
    rng = np.random.default_rng(seed)
    W = int(rng.integers(2, 65))
    H = int(rng.integers(2, 65))
    ntri = int(rng.integers(1, 33))
    nv = 3 * ntri
    indexed = bool(rng.integers(0, 2))
    if indexed:
        nv = int(rng.integers(3, 3 * ntri + 1))
    # clip-space positions through a perspective-like, non-symmetric MVP
    mvp = rng.uniform(-1.2, 1.2, (4, 4))
    mvp[3] = (*rng.uniform(-0.2, 0.2, 3), rng.uniform(1.0, 2.0))
    nm = rng.uniform(-1, 1, (3, 3))
    model = rng.uniform(-1, 1, (3, 4))
    pos = rng.uniform(-1.0, 1.0, (nv, 3))
    if rng.integers(0, 3) == 0:  # snap some vertices to pixel centres under an identity MVP
        mvp = np.eye(4)
        px = rng.integers(0, W, nv) + 0.5
        py = rng.integers(0, H, nv) + 0.5
        pos[:, 0] = px * 2.0 / W - 1.0
        pos[:, 1] = 1.0 - py * 2.0 / H
    if ntri > 2 and not indexed:  # a duplicate triangle: exact depth ties
        pos[3:6] = pos[0:3]
    mesh = dict(px=pos[:, 0].astype(np.float32), py=pos[:, 1].astype(np.float32),
                pz=pos[:, 2].astype(np.float32), nx=rng.uniform(-1, 1, nv).astype(np.float32),
                ny=rng.uniform(-1, 1, nv).astype(np.float32), nz=rng.uniform(-1, 1, nv).astype(np.float32),
                u=rng.uniform(-0.2, 1.2, nv).astype(np.float32), v=rng.uniform(-0.2, 1.2, nv).astype(np.float32),
                indices=rng.integers(0, nv, 3 * ntri).astype(np.uint32) if indexed else None, primitive=3)
    fp = int(rng.integers(0, 4))
    vp = 1 if fp == 1 else (0 if fp in (2, 3) else int(rng.integers(0, 2)))
    fs = rng.uniform(0, 1, 32).astype(np.float32)
    if fp == 1:
        light = rng.standard_normal(3)
        fs[0:3] = light / np.linalg.norm(light)
        fs[3:6] = rng.uniform(-3, 3, 3)
    if fp == 2:
        light = rng.standard_normal(3)
        fs[0:3] = light / np.linalg.norm(light)
    tex = None
    if fp == 2:
        tw, th = int(rng.integers(1, 17)), int(rng.integers(1, 17))
        tex = (tw, th, rng.integers(0, 2 ** 32, tw * th, dtype=np.uint64).astype(np.uint32))
    tile = [(128, 128), (16, 16), (7, 5), (1, 3), (64, 9)][int(rng.integers(0, 5))]
    kw = dict(vertex_program=vp, fragment_program=fp, blend=int(rng.integers(0, 2)),
              cull=int(rng.integers(0, 3)), tile_w=tile[0], tile_h=tile[1],
              stencil_test=int(rng.integers(0, 8)), stencil_op=int(rng.integers(0, 6)),
              stencil_wrap=int(rng.integers(0, 2)), stencil_ref=int(rng.integers(0, 256)),
              tex_filter=int(rng.integers(0, 2)), tex_edge=int(rng.integers(0, 3)),
              tex_border=int(rng.integers(0, 2 ** 32)),
              vs_uniforms=scenes.vs_uniforms(mvp, nm, model), fs_uniforms=fs,
              viewport=(0.0, 0.0, float(W), float(H), 0.0, 1.0))
    if rng.integers(0, 4) == 0:  # a viewport that is not the framebuffer
        kw["viewport"] = (float(rng.uniform(-4, 4)), float(rng.uniform(-4, 4)), float(W * rng.uniform(0.5, 1.5)),
                          float(H * rng.uniform(0.5, 1.5)), float(rng.uniform(0, 0.3)), float(rng.uniform(0.7, 1)))
    n = W * H
    init = dict(color=rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32),
                depth=np.where(rng.random(n) < 0.5, F32MIN, rng.uniform(-1, 0, n)).astype(np.float32),
                stencil=rng.integers(0, 256, n).astype(np.uint8))
    return W, H, mesh, kw, tex, init


def check_numpy_restatement_random_scenes(L, n_scenes=48, seed0=1000):
    differing = []
    shaded_total = 0
    for s in range(seed0, seed0 + n_scenes):
        W, H, mesh, kw, tex, init = random_scene(s)
        desc = sr.make_desc(**kw)
        fb, st = run_oracle(L, W, H, mesh, desc, tex=tex, **init)
        npfb = dict(w=W, h=H, color=init["color"].copy(), depth=init["depth"].copy(),
                    stencil=init["stencil"].copy())
        cnt = NR.draw(npfb, mesh, desc_dict(desc), tex)
        shaded_total += cnt["shaded"]
        same = (np.array_equal(fb.color, npfb["color"]) and
                np.array_equal(fb.depth.view(np.uint32), npfb["depth"].view(np.uint32)) and
                np.array_equal(fb.stencil, npfb["stencil"]) and cnt["shaded"] == st["shaded"])
        if not same:
            differing.append(s)
    assert not differing, f"oracle and numpy restatement differ on scenes {differing}"
    assert shaded_total > 1000  # the scenes do reach the fragment programs


ALL = [check_bary_vertices_exact_and_linear_inside, check_cull_orientation, check_vertex_programs_float64,
       check_blinn_phong_closed_form, check_tex_lambert_at_texel_centre, check_over_blend_half_alpha,
       check_numpy_restatement_random_scenes]
