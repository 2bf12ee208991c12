"""Pins the oracle's normal visualisation programs (full_example/src/
shaders.rs:35-89) before they check the device.

Each case fixes a fact the programs' definition implies, independently of
the oracle's own arithmetic:
  * start/end points equal `projection * view` applied (in float64) to the
    world position and to position + normal * length, with a non-symmetric
    matrix, so a transposed operand, a dropped length or a flipped normal
    fails;
  * the face program's end point lies `length` away from the centroid in
    world space along the normalised mean normal, whatever the normals'
    magnitudes (normalize() is applied);
  * with the identity matrix and length 0 the lines degenerate to their
    start points bit for bit;
  * lines and points pass through unchanged (re_emit): such draws equal the
    same draw without a geometry program, bit for bit;
  * drawn end to end, the green pixels of a normal lie where the projected
    segment lies.
"""
import numpy as np
import pytest

import oracle
import scenes
import softrender_hip as sr

VN, FN = scenes.GS_VERTEX_NORMALS, scenes.GS_FACE_NORMALS


def camera(w=64, h=64):
    view = scenes.look_at((0.3, 0.2, 3), (0, 0, 0), (0, 1, 0))
    proj = scenes.perspective_zero_to_one(60.0, w / h, 0.1, 100.0)
    return proj @ view


def tri_inputs(rng, scale_normals=False):
    """Three vertices with K = 6 varyings: normal (k0..k2), world pos (k3..k5)."""
    pos = rng.uniform(-1, 1, (3, 3))
    nrm = rng.normal(size=(3, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    if scale_normals:
        nrm *= np.array([[0.5], [2.0], [3.0]])
    k = np.concatenate([nrm, pos], 1).astype(np.float32)
    clip = rng.uniform(-1, 1, (3, 4)).astype(np.float32)  # not read by the programs
    return clip, k


def apply(G, v3, w=1.0):
    return G @ np.r_[np.asarray(v3, np.float64), w]


@pytest.mark.parametrize("seed", range(8))
def test_vertex_normals_against_float64(seed):
    rng = np.random.default_rng(seed)
    clip, k = tri_inputs(rng)
    PV = camera()
    L = 0.05 + 0.1 * seed
    G = np.r_[PV.astype(np.float32).ravel(), np.float32(L)]
    oc, ok = oracle.gs_normal_lines(VN, clip, k, G)
    assert oc.shape == (6, 4)
    P32 = PV.astype(np.float32).astype(np.float64)
    k64 = k.astype(np.float64)
    for j in range(3):
        p, n = k64[j, 3:6], k64[j, 0:3]
        np.testing.assert_allclose(oc[2 * j], apply(P32, p), rtol=2e-6, atol=2e-6)
        np.testing.assert_allclose(oc[2 * j + 1], apply(P32, p + n * np.float32(L)), rtol=2e-6, atol=2e-6)
        # both ends carry the vertex's own varyings
        assert (ok[2 * j] == k[j]).all() and (ok[2 * j + 1] == k[j]).all()
    # the segment is not degenerate: the end moved by the normal
    assert np.abs(oc[1] - oc[0]).max() > 1e-3


@pytest.mark.parametrize("seed", range(8))
def test_face_normal_is_length_from_centroid(seed):
    rng = np.random.default_rng(100 + seed)
    clip, k = tri_inputs(rng, scale_normals=True)
    PV = camera()
    L = 0.25
    G = np.r_[PV.astype(np.float32).ravel(), np.float32(L)]
    oc, ok = oracle.gs_normal_lines(FN, clip, k, G)
    assert oc.shape == (2, 4)
    k64 = k.astype(np.float64)
    centroid = k64[:, 3:6].mean(0)
    mean_n = k64[:, 0:3].mean(0)
    unit = mean_n / np.linalg.norm(mean_n)
    P32 = PV.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(oc[0], apply(P32, centroid), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(oc[1], apply(P32, centroid + unit * L), rtol=1e-5, atol=1e-5)
    # back to world space: the end is `length` from the start along the normal
    inv = np.linalg.inv(P32)
    ws, we = inv @ oc[0].astype(np.float64), inv @ oc[1].astype(np.float64)
    ws, we = ws[:3] / ws[3], we[:3] / we[3]
    assert abs(np.linalg.norm(we - ws) - L) < 1e-4
    assert np.dot((we - ws) / L, unit) > 0.9999
    # both ends carry the centroid's varyings
    np.testing.assert_allclose(ok[0], k64.mean(0), rtol=1e-6, atol=1e-7)
    assert (ok[0] == ok[1]).all()


@pytest.mark.parametrize("program", [VN, FN])
def test_identity_and_zero_length_degenerate(program):
    rng = np.random.default_rng(7)
    clip, k = tri_inputs(rng)
    G = np.r_[np.eye(4, dtype=np.float32).ravel(), np.float32(0.0)]
    oc, _ = oracle.gs_normal_lines(program, clip, k, G)
    assert (oc[0::2] == oc[1::2]).all()
    if program == VN:
        want = np.c_[k[:, 3:6], np.ones(3, np.float32)]
        assert (oc[0::2] == want).all()
    else:
        # (1/3 + 1/3) + 1/3 rounds to exactly 1 in binary32
        assert oc[0, 3] == np.float32(1.0)


def _tiny_mesh(primitive):
    rng = np.random.default_rng(3)
    n = {1: 40, 2: 40}[primitive]
    pos = rng.uniform(-0.8, 0.8, (n, 3)).astype(np.float32)
    nrm = np.tile(np.float32([0, 0, 1]), (n, 1))
    return dict(primitive=primitive, px=pos[:, 0].copy(), py=pos[:, 1].copy(), pz=pos[:, 2].copy(),
                nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(), u=None, v=None,
                indices=None)


@pytest.mark.parametrize("primitive", [1, 2])
@pytest.mark.parametrize("program", [VN, FN])
def test_lines_and_points_re_emit(primitive, program):
    m = _tiny_mesh(primitive)
    PV = camera(48, 40)
    fs = np.zeros(32, np.float32)
    fs[0:4] = (0.2, 0.9, 0.4, 1.0)
    base = dict(vertex_program=1, fragment_program=0, blend=1, aa_lines=1, alpha_mode=1, fs_uniforms=fs,
                vs_uniforms=scenes.vs_uniforms(PV), viewport=(0.0, 0.0, 48.0, 40.0, 0.0, 1.0))
    gsu = np.r_[PV.ravel(), 0.05].astype(np.float32)
    out = []
    for gs in (0, program):
        fb = oracle.Framebuffer(48, 40, 0xFF000000)
        rc, st = oracle.draw(fb, m, sr.make_desc(**dict(base, geometry_program=gs, gs_uniforms=gsu)))
        assert rc == 0
        out.append(fb)
    assert (out[0].color == out[1].color).all() and (out[0].depth.view(np.uint32) == out[1].depth.view(np.uint32)).all()
    assert (out[0].color != 0xFF000000).sum() > 10


def _edge_on_triangle():
    # a triangle in the plane x = 0 facing +x, seen edge-on from +z
    pos = np.float32([[0, -0.3, 0.3], [0, 0.3, 0.3], [0, 0.0, -0.3]])
    nrm = np.tile(np.float32([1, 0, 0]), (3, 1))
    return dict(primitive=3, px=pos[:, 0].copy(), py=pos[:, 1].copy(), pz=pos[:, 2].copy(),
                nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(), u=None, v=None,
                indices=np.arange(3, dtype=np.uint32)), pos


def _screen(PV, p, W, H):
    c = PV @ np.r_[p, 1.0]
    return (c[0] / c[3] + 1) * W / 2, (1 - c[1] / c[3]) * H / 2


@pytest.mark.parametrize("program", [VN, FN])
def test_normal_pixels_follow_the_projected_segment(program):
    W = H = 64
    L = 0.5
    view = scenes.look_at((0, 0, 3), (0, 0, 0), (0, 1, 0))
    PV = scenes.perspective_zero_to_one(60.0, 1.0, 0.1, 100.0) @ view
    m, pos = _edge_on_triangle()
    fs = np.zeros(32, np.float32)
    fs[0:4] = (0, 1, 0, 1)
    desc = dict(vertex_program=1, geometry_program=program, fragment_program=0, blend=0,
                vs_uniforms=scenes.vs_uniforms(PV), fs_uniforms=fs,
                gs_uniforms=np.r_[PV.ravel(), L], viewport=(0.0, 0.0, float(W), float(H), 0.0, 1.0))
    fb = oracle.Framebuffer(W, H, 0xFF000000)
    rc, st = oracle.draw(fb, m, sr.make_desc(**desc))
    assert rc == 0
    assert st["gen_lines"] == (3 if program == VN else 1)
    lit = np.argwhere(fb.color.reshape(H, W) != 0xFF000000)
    starts = pos if program == VN else pos.mean(0, keepdims=True)
    for p in starts:
        sx, sy = _screen(PV, p.astype(np.float64), W, H)
        ex, ey = _screen(PV, p.astype(np.float64) + np.array([L, 0, 0]), W, H)
        assert ex - sx > 5  # the normal points to +x on screen
        row = lit[np.abs(lit[:, 0] - sy) <= 1.0]
        xs = row[:, 1]
        assert xs.size and xs.min() >= np.floor(sx) - 1 and xs.max() <= np.ceil(ex) + 1
        assert xs.max() - xs.min() >= (ex - sx) - 2
    assert len(lit) <= sum(int(np.ceil(_screen(PV, p + np.array([L, 0, 0]), W, H)[0]
                                       - _screen(PV, p, W, H)[0])) + 3 for p in starts.astype(np.float64))


@pytest.mark.parametrize("program", [VN, FN])
def test_full_example_frame_counts(program):
    pos, nrm, idx = scenes.uv_sphere(6, 8)
    s = scenes.normals_scene(pos, nrm, idx, 96, 96, program=program)
    fb, st = oracle.render_scene(s)
    ntri = len(idx) // 3
    assert st[1]["gen_lines"] == (3 if program == VN else 1) * ntri
    green = ((fb.color >> 8) & 0xFF) == 0xFF  # (0, 1, 0, 1) OVER: full green channel
    assert green.sum() > 30
