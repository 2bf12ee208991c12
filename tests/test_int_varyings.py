"""Integer varyings: an i32 member of the vertex uniforms, interpolated as
interpolate.rs:43-64 does for integer types (each vertex's value through
R::from into f32, ((x1*u) + (x2*v)) + (x3*w), then <i32 as NumCast>::from,
which truncates toward zero), read by a palette program
(SR_VP_MVP_INT + SR_FP_INT_PALETTE).

CPU: the oracle's cast against closed forms (2.7 -> 2, -2.7 -> -2, not
floor; -9 -> palette 7), the vertex program's i32 -> f32 rounding at 2^24 + 1,
and a draw whose pixel centres sit on the vertices (barycentrics exactly
(1, 0, 0)), where each pixel takes its vertex's own value; a floor in place
of the truncation is caught.
GPU: soups with integer attributes (small, negative, and around 2^24) drawn
by the device, bit for bit against the oracle; the checks on the programs.
"""
import numpy as np
import pytest

import oracle
import scenes
import softrender_hip as sr
from oracle_checks import screen_tri_mesh
from parity import assert_parity, render_gpu, render_ref

PALETTE = np.arange(32, dtype=np.float32) / 31.0  # 8 distinct RGBA colours


def palette_desc(**kw):
    d = dict(vertex_program=sr.VP_MVP_INT, fragment_program=sr.FP_INT_PALETTE, blend=0, cull=0,
             vs_uniforms=scenes.vs_uniforms(np.eye(4)), fs_uniforms=PALETTE)
    d.update(kw)
    return d


def colour_of(i):
    m = ((int(i) % 8) + 8) % 8
    return PALETTE[4 * m:4 * m + 4]


@pytest.mark.parametrize("x,want", [(2.7, 2), (-2.7, -2), (-0.5, 0), (0.99, 0), (9.99, 9), (-9.0, -9),
                                    (16777216.0, 16777216), (-1e12, -2147483648), (1e12, 2147483647)])
def test_cast_truncates_toward_zero(x, want):
    got = oracle.fragment(sr.make_desc(**palette_desc()), [x])
    assert np.array_equal(got, colour_of(want))
    if x != int(x) and x < 0:  # a floor would pick the next index down
        assert not np.array_equal(got, colour_of(np.floor(x)))


def test_vertex_program_rounds_to_f32():
    mesh = dict(primitive=3, px=np.zeros(3, np.float32), py=np.zeros(3, np.float32), pz=np.zeros(3, np.float32),
                nx=None, ny=None, nz=None, u=None, v=None, indices=None,
                iattr=np.array([7, -123456789, 16777217], np.int32))
    d = sr.make_desc(**palette_desc())
    for i, want in enumerate((7.0, -123456792.0, 16777216.0)):
        _, k = oracle.vertex(mesh, d, i)
        assert len(k) == 1 and k[0] == np.float32(want)


def test_pixels_on_vertices_take_their_values():
    W = H = 16
    pts = ((1.5, 2.5), (13.5, 4.5), (5.5, 12.5))  # pixel centres
    m = screen_tri_mesh(W, H, pts, (-0.5, -0.5, -0.5))
    m["iattr"] = np.array([3, -5, 12], np.int32)
    s = dict(m, width=W, height=H, clear=0, desc=dict(palette_desc(), viewport=(0.0, 0.0, W, H, 0.0, 1.0)))
    ref, _ = render_ref(s)
    for (x, y), val in zip(pts, (3, -5, 12)):
        c = ref.color[int(y) * W + int(x)]
        want = [int(round(f * 255)) for f in colour_of(val)]
        assert [(c >> (8 * i)) & 0xFF for i in range(4)] == want


def _int_soup(n, W, H, seed, lo, hi):
    s = scenes.soup_scene(n, W, H, seed=seed)
    rng = np.random.default_rng(seed)
    s["iattr"] = rng.integers(lo, hi, 3 * n).astype(np.int32)
    s["desc"] = dict(s["desc"], vertex_program=sr.VP_MVP_INT, fragment_program=sr.FP_INT_PALETTE,
                     fs_uniforms=PALETTE)
    s["texture"] = None
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("lo,hi,blend", [(-20, 20, 1), (-20, 20, 0), (16777200, 16777240, 1),
                                         (-(2**30), 2**30, 1)])
def test_gpu_integer_soup(ctx, lo, hi, blend):
    s = _int_soup(4000, 200, 150, 0x61 + lo % 7, lo, hi)
    s["desc"]["blend"] = blend
    ref, st = render_ref(s)
    assert st[0]["shaded"] > 1000
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.gpu
def test_gpu_integer_programs_are_checked(ctx):
    s = _int_soup(10, 32, 32, 1, 0, 5)
    fb = ctx.framebuffer(32, 32)
    plain = dict(s)
    plain.pop("iattr")
    m_int, m_plain = ctx.mesh(s), ctx.mesh(plain)
    try:
        d = s["desc"]
        ctx.draw(fb, m_int, sr.make_desc(**d))
        for mesh, bad in ((m_plain, {}), (m_int, dict(fragment_program=2)), (m_int, dict(vertex_program=0)),
                          (m_int, dict(geometry_program=1))):
            with pytest.raises(sr.SoftrenderError):
                ctx.draw(fb, mesh, sr.make_desc(**dict(d, **bad)))
        four = np.zeros(4, np.int32)
        with pytest.raises(sr.SoftrenderError):  # one value per vertex
            ctx._check(sr.lib().sr_mesh_set_int_attrib(m_int.h, four.ctypes.data_as(sr.C.POINTER(sr.C.c_int32)), 4))
    finally:
        m_int.free()
        m_plain.free()
        fb.free()
