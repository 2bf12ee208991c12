"""The full example's shading as device programs: its vertex shader
(SR_VP_WORLD4_NORMAL4, full_example/src/shaders.rs:8-31), its four-light
Fresnel fragment shader with ACES tone mapping and gamma encoding
(SR_FP_FRESNEL4, shaders.rs:92-163, color.rs:19-47) and its texture
sampler with gamma decoding (SR_FILTER_SPAN_*, texture.rs:29-84).

CPU: the oracle against float64 restatements of the example written here
from the Rust source (4-vector dot and norm, powi exact in float64), with
f32-sized tolerances, plus closed forms (an unlit point is (0, 0, 0, 1);
nearest sampling at u = 0 and u = 1 gives the first and last texel,
decoded); each check also fails on a plausible slip (the normal's w
dropped, a specular exponent of 32, one light left out, no decode).
GPU: the device against the oracle on the example's frame (Suzanne, RGBAF32
colour with the normalised over-blend; and an RGBA8 framebuffer), and on a
soup sampled with the example's sampler.  Depth and coverage bit for bit;
colours within 1e-5 on RGBAF32 and 1 LSB on RGBA8 (powf and sqrtf are the
platform's on either side).
"""
import os

import numpy as np
import pytest

import oracle
import scenes
import softrender_hip as sr
from parity import assert_parity, render_gpu, render_ref

SUZANNE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "suzanne.obj")


# ---------------------------------------------------------------------------
# float64 restatements of the example (written from the Rust source)
# ---------------------------------------------------------------------------
def fresnel64(cam, lights, grey, pos, nrm, shininess=32, drop_light=None):
    cam4 = np.r_[cam, 1.0]
    view = cam4 - pos
    view = view / np.linalg.norm(view)
    mat = grey ** 2.2
    col = np.zeros(3)
    for i, L in enumerate(lights):
        if i == drop_light:
            continue
        d = np.r_[L[0:3], 1.0] - pos
        dist = np.linalg.norm(d)
        ldir = d / dist
        h = ldir + view
        h = h / np.linalg.norm(h)
        inten = L[6] / dist ** 2
        sat = lambda x: min(max(x, 0.0), 1.0)
        ndl, ndh, vdh = sat(ldir @ nrm), sat(nrm @ h), sat(view @ h)
        f0 = ((1 - 1.45) / (1 + 1.45)) ** 2
        f = f0 + (1 - f0) * (1 - vdh) ** 5
        diffuse = (1 - f) * ndl
        spec = f * ndh ** (shininess * 2)
        col += inten * L[3:6] * (spec + diffuse * 0.7 * mat)
    aces = (col * (2.51 * col + 0.03)) / (col * (2.43 * col + 0.59) + 0.14)
    return np.r_[aces ** (1 / 2.2), 1.0]


def fresnel_desc(cam, lights, grey):
    fs = np.zeros(32, np.float32)
    fs[0:3] = cam
    fs[3:31] = np.asarray(lights, np.float32).reshape(28)
    fs[31] = grey
    return sr.make_desc(vertex_program=2, fragment_program=5, fs_uniforms=fs)


def random_case(rng):
    cam = rng.uniform(-3, 3, 3)
    lights = scenes.example_lights().copy()
    lights[:, 0:3] += rng.uniform(-0.5, 0.5, (4, 3))
    pos = np.r_[rng.uniform(-0.8, 0.8, 3), rng.uniform(0.97, 1.03)]
    nrm = rng.normal(0, 1, 4)
    nrm[3] *= 0.2
    nrm /= np.linalg.norm(nrm) * rng.uniform(0.9, 1.1)  # interpolated normals are not unit
    return cam.astype(np.float32), lights.astype(np.float32), pos.astype(np.float32), nrm.astype(np.float32)


def close(a, b, tol=2e-5):
    return np.all(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) <= tol * (1 + np.abs(b)))


# ---------------------------------------------------------------------------
# CPU: the oracle's programs
# ---------------------------------------------------------------------------
def test_vertex_program_4_vectors():
    rng = np.random.default_rng(3)
    model = scenes.rotation_y(37.0)
    model[:3, 3] = (1.5, -1.0, 2.0)
    mit = np.linalg.inv(model).T
    assert np.abs(mit[3, :3]).max() > 0.5  # the normal's w is not zero
    s = scenes.fresnel_scene(rng.uniform(-1, 1, (12, 3)), rng.normal(0, 1, (12, 3)), np.arange(12), 64, 48,
                             translate=(1.5, -1.0, 2.0), object_deg=37.0)
    d = sr.make_desc(**s["desc"])
    slips = 0
    for i in range(12):
        clip, k = oracle.vertex(s, d, i)
        p = np.r_[s["px"][i], s["py"][i], s["pz"][i], 1.0].astype(np.float64)
        n = np.r_[s["nx"][i], s["ny"][i], s["nz"][i], 0.0].astype(np.float64)
        world = model @ p
        n4 = mit @ n
        n4 = n4 / np.linalg.norm(n4)
        n3 = (mit @ n)[:3] / np.linalg.norm((mit @ n)[:3])  # the slip: w left out of the norm
        assert len(k) == 8
        assert close(k[0:4], world, 1e-6)
        assert close(k[4:8], n4, 1e-6)
        slips += not close(k[4:7], n3, 1e-3)
    assert slips >= 6, slips


def test_fresnel_program_against_float64():
    rng = np.random.default_rng(5)
    slips = {"exponent 32": 0, "one light dropped": 0}
    for _ in range(200):
        cam, lights, pos, nrm = random_case(rng)
        got = oracle.fragment(fresnel_desc(cam, lights, 0.25), np.r_[pos, nrm])
        want = fresnel64(cam.astype(np.float64), lights.astype(np.float64), 0.25, pos.astype(np.float64),
                         nrm.astype(np.float64))
        assert close(got, want), (got, want)
        slips["exponent 32"] += not close(got, fresnel64(cam, lights, 0.25, pos, nrm, shininess=16))
        slips["one light dropped"] += not close(got, fresnel64(cam, lights, 0.25, pos, nrm, drop_light=2))
    assert all(v > 20 for v in slips.values()), slips


def test_unlit_point_is_black():
    cam, lights, pos, nrm = random_case(np.random.default_rng(7))
    lights[:, 6] = 0.0  # no intensity: ACES(0) = 0, 0^(1/2.2) = 0
    out = oracle.fragment(fresnel_desc(cam, lights, 0.25), np.r_[pos, nrm])
    assert np.array_equal(out, np.array([0, 0, 0, 1], np.float32))


def _tex(rng, w=7, h=5):
    return (w, h, rng.integers(0, 2**32, w * h, dtype=np.uint64).astype(np.uint32))


def _texel(tex, x, y):
    w, h, data = tex
    c = int(data[y * w + x])
    return np.array([(c >> (8 * i)) & 0xFF for i in range(4)], np.float64) / 255.0


def _decode(c):
    return np.r_[c[:3] ** 2.2, c[3]]


def test_span_nearest_corners_and_rounding():
    tex = _tex(np.random.default_rng(9))
    w, h, _ = tex
    for edge in (0, 1):  # Clamp, Wrap
        assert close(oracle.sample(tex, 0.0, 0.0, 2, edge), _decode(_texel(tex, 0, 0)), 1e-6)
    assert close(oracle.sample(tex, 1.0, 1.0, 2, 0), _decode(_texel(tex, w - 1, h - 1)), 1e-6)
    assert close(oracle.sample(tex, 7.5, -3.0, 2, 0), _decode(_texel(tex, w - 1, 0)), 1e-6)  # clamped
    # u (W - 1) = 2.5 rounds half away from zero to texel 3
    u = np.float32(2.5) / np.float32(w - 1)
    x = int(np.floor(float(np.float32(u) * np.float32(w - 1)) + 0.5))
    assert close(oracle.sample(tex, float(u), 0.0, 2, 0), _decode(_texel(tex, x, 0)), 1e-6)
    # Wrap: u - trunc(u)
    assert close(oracle.sample(tex, 1.0 + 2.0 / (w - 1), 0.0, 2, 1), oracle.sample(tex, 2.0 / (w - 1), 0.0, 2, 0),
                 1e-6)
    # the slip: no gamma decode
    assert not close(oracle.sample(tex, 0.0, 0.0, 2, 0), _texel(tex, 0, 0), 1e-3)


def test_span_bilinear_weights():
    rng = np.random.default_rng(11)
    tex = _tex(rng)
    w, h, _ = tex
    for _ in range(300):
        u, v = rng.uniform(0, 1, 2)
        uu, vv = u * (w - 1) + 0.5, v * (h - 1) + 0.5
        x, y = int(np.floor(uu)), int(np.floor(vv))
        ur, vr = uu - x, vv - y
        X1, Y1 = min(x + 1, w - 1), min(y + 1, h - 1)
        x, y = min(x, w - 1), min(y, h - 1)
        c = ((_texel(tex, x, y) * (1 - ur) + _texel(tex, X1, y) * ur) * (1 - vr) +
             (_texel(tex, x, Y1) * (1 - ur) + _texel(tex, X1, Y1) * ur) * vr)
        assert close(oracle.sample(tex, u, v, 3, 0), _decode(c), 5e-6)
    # u = 1 takes the last texel twice, where the example would panic
    assert close(oracle.sample(tex, 1.0, 0.0, 3, 0)[:3],
                 _decode((_texel(tex, w - 1, 0) + _texel(tex, w - 1, 1)) / 2)[:3], 5e-6)


# ---------------------------------------------------------------------------
# GPU: the device against the oracle
# ---------------------------------------------------------------------------
def _suzanne_scene(W, H, translate=(0.0, 0.0, 0.0)):
    m = sr.parse_obj(SUZANNE)
    return scenes.fresnel_scene(m["positions"], m["normals"], m["indices"], W, H, translate=translate)


def _oracle_f32(s):
    fb = oracle.FramebufferEx(s["width"], s["height"], [sr.COLOR_RGBAF32], 0, 0, [list(s["clear_f32"])])
    rc, st = oracle.draw(fb, s, sr.make_desc(**s["desc"]), None)
    assert rc == 0
    return fb, st


@pytest.mark.gpu
@pytest.mark.parametrize("translate", [(0.0, 0.0, 0.0), (0.2, -0.1, 0.3)])
def test_gpu_example_frame_rgbaf32(ctx, translate):
    s = _suzanne_scene(320, 180, translate)
    ref, st = _oracle_f32(s)
    assert st["shaded"] > 5000
    fb = ctx.framebuffer_ex(320, 180, [sr.COLOR_RGBAF32], 0, 0)
    mesh = ctx.mesh(s)
    try:
        fb.clear_ex([list(s["clear_f32"])])
        ctx.draw(fb, mesh, sr.make_desc(**s["desc"]))
        planes, dep, _ = fb.read_ex()
    finally:
        mesh.free()
        fb.free()
    assert np.array_equal(dep.view(np.uint32), ref.depth.view(np.uint32))
    got = planes[0].reshape(-1, 4).astype(np.float64)
    want = np.asarray(ref.planes[0], np.float32).reshape(-1, 4).astype(np.float64)
    assert np.abs(got - want).max() <= 1e-5 * (1 + np.abs(want)).max()
    lit = (want[:, 0] > 0.02).sum()
    assert lit > 5000, lit


@pytest.mark.gpu
def test_gpu_example_frame_rgba8(ctx):
    s = _suzanne_scene(256, 192)
    s["desc"] = dict(s["desc"], blend=1)
    s["clear"] = 0xFF030303
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref, color_tol=1)


@pytest.mark.gpu
@pytest.mark.parametrize("filt", [2, 3])
@pytest.mark.parametrize("edge", [0, 1])
def test_gpu_span_sampler(ctx, filt, edge):
    s = scenes.soup_scene(3000, 160, 120, seed=0x51)
    s["texture"] = (64, 48, scenes.checkerboard(64, 8)[: 64 * 48])
    s["u"] = (s["u"] * 3.0 - 1.0).astype(np.float32)  # outside [0, 1] too
    s["desc"] = dict(s["desc"], tex_filter=filt, tex_edge=edge)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref, color_tol=1)


@pytest.mark.gpu
def test_gpu_example_programs_are_checked(ctx):
    s = _suzanne_scene(32, 32)
    fb = ctx.framebuffer(32, 32)
    mesh = ctx.mesh(s)
    try:
        for bad in (dict(vertex_program=0), dict(fragment_program=2), dict(geometry_program=2),
                    dict(tex_filter=4)):
            with pytest.raises(sr.SoftrenderError):
                ctx.draw(fb, mesh, sr.make_desc(**dict(s["desc"], **bad)))
    finally:
        mesh.free()
        fb.free()
