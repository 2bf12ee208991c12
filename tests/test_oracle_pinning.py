"""Pins the CPU oracle before it is trusted as the checker.

The reference (Rust) cannot be built here, so each oracle function is
checked against what fixes it independently of its own code:
  * the reference's one hot-path known-answer test (pixel index,
    src/geometry/coordinate.rs:66-86);
  * closed forms (viewport transform, tile grid and seam counts, Wu on
    axis-aligned lines, bilinear at texel centres, exact triangle coverage);
  * invariants stated for the reference's own clip_primitives (16 / 13 / 17
    triangles, outside vertices kept) and textbook properties (Sutherland-
    Hodgman area, Bresenham 8-connectivity, Liang-Barsky endpoints);
  * truth tables read off stencil.rs:109-159.
Each case is chosen so a plausible slip (swapped operands, dropped term,
off-by-one tile end, half-open vs closed bounds) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import scenes
import softrender_hip as sr

L = oracle.lib()
F32MIN = np.float32(np.finfo(np.float32).min)


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ---------------------------------------------------------------------------
# Screen-space helper: a viewport where clip (x, y) = (sx/8 - 1, 1 - sy/8)
# lands exactly on screen (sx, sy) for W = H = 16.
# ---------------------------------------------------------------------------
def screen_mesh(tris_screen, z=0.5, uv=None):
    """tris_screen: list of ((sx,sy),(sx,sy),(sx,sy)) in a 16x16 viewport."""
    pts = np.asarray(tris_screen, np.float64).reshape(-1, 2)
    n = len(pts)
    zs = np.broadcast_to(np.asarray(z, np.float64), (n,)) if np.ndim(z) else np.full(n, z)
    return dict(px=(pts[:, 0] / 8.0 - 1.0).astype(np.float32),
                py=(1.0 - pts[:, 1] / 8.0).astype(np.float32),
                pz=np.asarray(zs, np.float32), nx=np.ones(n, np.float32),
                ny=np.zeros(n, np.float32), nz=np.zeros(n, np.float32),
                u=np.zeros(n, np.float32) if uv is None else uv[0],
                v=np.zeros(n, np.float32) if uv is None else uv[1],
                indices=None, primitive=3)


def draw16(mesh, w=16, h=16, clear=0xFF000000, **kw):
    desc = dict(vertex_program=0, fragment_program=0, blend=0,
                vs_uniforms=scenes.vs_uniforms(np.eye(4)),
                viewport=(0.0, 0.0, 16.0, 16.0, 0.0, 1.0))
    fs = np.zeros(32, np.float32)
    fs[0:4] = (1.0, 1.0, 1.0, 1.0)
    desc["fs_uniforms"] = fs
    desc.update(kw)
    fb = oracle.Framebuffer(w, h, clear)
    rc, st = oracle.draw(fb, mesh, sr.make_desc(**desc))
    assert rc == 0
    return fb, st


# ---------------------------------------------------------------------------
# pixel index: the reference's own test, coordinate.rs:66-86 (x + y * W)
# ---------------------------------------------------------------------------
def test_pixel_index_known_answer():
    W, H = 10, 20
    # draw one point per pixel; a point lands at pixel (trunc x, trunc y) of
    # index x + y*W (point.rs:47-49).  Pixels in the last column / row are
    # never drawn (half-open against tile max W-1 / H-1, point.rs:46).
    i = 0
    for y in range(H):
        for x in range(W):
            fb = oracle.Framebuffer(W, H, 0)
            m = dict(px=np.array([(x + 0.5) * 2 / W - 1], np.float32),
                     py=np.array([1 - (y + 0.5) * 2 / H], np.float32),
                     pz=np.array([0.5], np.float32), nx=None, ny=None, nz=None, u=None, v=None,
                     indices=None, primitive=1)
            desc = sr.make_desc(vertex_program=0, fragment_program=0, tile_w=128, tile_h=128,
                                vs_uniforms=scenes.vs_uniforms(np.eye(4)),
                                fs_uniforms=np.r_[np.ones(4), np.zeros(28)].astype(np.float32),
                                viewport=(0, 0, W, H, 0, 1))
            rc, _ = oracle.draw(fb, m, desc)
            assert rc == 0
            hit = np.flatnonzero(fb.color)
            if x < W - 1 and y < H - 1:
                assert list(hit) == [i], (x, y, hit)
                assert i % W == x and i // W == y  # from_index
            else:
                assert len(hit) == 0
            i += 1


# ---------------------------------------------------------------------------
# viewport transform, clipvertex.rs:89-126 (closed form)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("vp", [(0, 0, 512, 512, 0, 1), (0, 0, 1920, 1080, 0, 1), (10, 20, 64, 32, 0.25, 0.75)])
def test_normalize_closed_form(vp):
    v = sr.Viewport(*[float(a) for a in vp])
    out = np.zeros(4, np.float32)
    x, y, w, h, n, f = vp
    for clip, expect in [((-1, 1, 0, 1), (x, y, -(f + n) / 2)),           # top-left corner
                         ((1, -1, 0, 1), (x + w, y + h, -(f + n) / 2)),   # bottom-right
                         ((0, 0, 1, 1), (x + w / 2, y + h / 2, -(f - n) / 2 - (f + n) / 2)),
                         ((2, -2, 0, 2), (x + w, y + h, -(f + n) / 2))]:  # perspective divide
        c = np.asarray(clip, np.float32)
        L.or_normalize(fp(c), C.byref(v), fp(out))
        assert out[0] == np.float32(expect[0]) and out[1] == np.float32(expect[1])
        assert out[2] == np.float32(expect[2])
        assert out[3] == np.float32(1.0) / c[3]  # w := 1/w


# ---------------------------------------------------------------------------
# tile grid, fragment.rs:188-216: inclusive tiles; seams counted twice
# ---------------------------------------------------------------------------
def tiles(w, h, tw=128, th=128):
    buf = (C.c_uint32 * (4 * 70000))()
    n = L.or_tiles(w, h, tw, th, buf, 70000)
    return np.frombuffer(buf, np.uint32)[: 4 * n].reshape(-1, 4)


@pytest.mark.parametrize("w,h,seams", [(512, 512, 3063), (1920, 1080, 30368), (3840, 2160, 123616)])
def test_tile_grid_and_seam_counts(w, h, seams):
    t = tiles(w, h)
    cover = np.zeros((h, w), np.int32)
    for x0, y0, x1, y1 in t:
        cover[y0:y1 + 1, x0:x1 + 1] += 1
    assert cover.min() == 1  # every pixel, including column W-1 and row H-1
    assert int((cover > 1).sum()) == seams
    assert len(t) == math.ceil((w - 1) / 128) * math.ceil((h - 1) / 128)


def test_tile_grid_degenerate():
    assert len(tiles(1, 50)) == 0 and len(tiles(50, 1)) == 0  # W==1 or H==1: no tiles
    t = tiles(129, 300, 128, 128)
    assert t[:, 2].max() == 128 and t[:, 3].max() == 299  # last tile ends at W-1 / H-1


# ---------------------------------------------------------------------------
# clip_primitives invariants (geometry.rs:261-336)
# ---------------------------------------------------------------------------
def V(x, y, z, w=1.0, k=()):
    return np.array([x, y, z, w, *k], np.float32)


def clip_ref(a, b, c, K=0):
    out = np.zeros(40 * 3 * (4 + K), np.float32)
    n = L.or_clip_triangle_reference(fp(a), fp(b), fp(c), K, fp(out), 40)
    return out[: n * 3 * (4 + K)].reshape(n, 3, 4 + K)


def area2(t):
    (x1, y1), (x2, y2), (x3, y3) = t[:, :2].astype(np.float64)
    return (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)


def test_clip_reference_inside_gives_16():
    a, b, c = V(-0.5, -0.5, 0.5), V(0.5, -0.4, 0.5), V(0.0, 0.6, 0.5)
    tr = clip_ref(a, b, c)
    assert len(tr) == 16
    nondeg = [t for t in tr if abs(area2(t)) > 0]
    assert len(nondeg) == 1
    assert {tuple(v[:3]) for v in nondeg[0]} == {tuple(a[:3]), tuple(b[:3]), tuple(c[:3])}


def test_clip_reference_outside_one_plane_keeps_the_triangle():
    # entirely beyond x = w: not removed, 13 triangles including the original
    a, b, c = V(2.0, 0.0, 0.5), V(3.0, 0.1, 0.5), V(2.5, 0.5, 0.5)
    tr = clip_ref(a, b, c)
    assert len(tr) == 13
    assert any({tuple(v[:3]) for v in t} == {tuple(a[:3]), tuple(b[:3]), tuple(c[:3])} for t in tr)


def test_clip_reference_one_vertex_behind_near():
    a, b, c = V(0.0, 0.0, -0.5), V(0.5, 0.0, 0.5), V(0.0, 0.5, 0.5)
    tr = clip_ref(a, b, c)
    assert len(tr) == 17
    assert all(tuple(t[0][:3]) == tuple(a[:3]) for t in tr)  # fanned from the outside vertex
    zs = {float(v[2]) for t in tr for v in t}
    assert 0.0 in zs  # the near-plane intersections (t = a/(a-b) lands exactly on z = 0 here)


def clip_near(a, b, c, K=0):
    out = np.zeros(4 * 3 * (4 + K), np.float32)
    n = L.or_clip_triangle_near(fp(a), fp(b), fp(c), K, fp(out), 4)
    return out[: n * 3 * (4 + K)].reshape(n, 3, 4 + K)


def test_clip_near_textbook():
    rng = np.random.default_rng(3)
    for _ in range(300):
        pts = rng.uniform(-1, 1, (3, 3))
        tri = [V(*p) for p in pts]
        out = clip_near(*tri)
        nin = int((pts[:, 2] >= 0).sum())
        assert len(out) == {0: 0, 1: 1, 2: 2, 3: 1}[nin]
        if nin == 3:
            # the edge loop pushes p of (a,b),(b,c),(c,a): the same triangle,
            # rotated to (b, c, a) — geometry.rs:270-287 with one plane
            assert np.array_equal(out[0], np.stack([tri[1], tri[2], tri[0]]))
        # area of the kept part equals the analytic half-plane area (x, y
        # projected; z is linear over the triangle)
        if nin in (1, 2):
            got = sum(abs(area2(t)) for t in out)
            # exact polygon: clip in float64
            poly = []
            for i in range(3):
                s, p = pts[i], pts[(i + 1) % 3]
                if (s[2] >= 0) != (p[2] >= 0):
                    t = s[2] / (s[2] - p[2])
                    poly.append(s + t * (p - s))
                if p[2] >= 0:
                    poly.append(p)
            poly = np.asarray(poly)
            want = abs(np.dot(poly[:, 0], np.roll(poly[:, 1], -1)) - np.dot(poly[:, 1], np.roll(poly[:, 0], -1)))
            assert abs(got - want) < 1e-5 * max(1.0, want)
            assert all(v[2] >= -1e-6 for t in out for v in t)


def test_clip_reference_line_and_varyings():
    s, e = V(-2.0, 0.0, 0.5, 1.0, (0.0,)), V(0.5, 0.0, 0.5, 1.0, (1.0,))
    out = np.zeros(2 * 5, np.float32)
    assert L.or_clip_line_reference(fp(s), fp(e), 1, fp(out)) == 1
    o = out.reshape(2, 5)
    assert o[0][0] == np.float32(-1.0)          # start moved onto x = -w
    t = np.float32(1.0) / np.float32(2.5)        # a/(a-b) with a = w1+x1 = -1, b = 1.5
    assert o[0][4] == (np.float32(1) - t) * np.float32(0) + t * np.float32(1)
    assert np.array_equal(o[1], e)               # inside endpoint untouched
    s2, e2 = V(2.0, 0.0, 0.5), V(3.0, 0.0, 0.5)
    assert L.or_clip_line_reference(fp(s2), fp(e2), 0, fp(out)) == 0  # both outside one plane


# ---------------------------------------------------------------------------
# Liang-Barsky (geometry/line.rs:6-54)
# ---------------------------------------------------------------------------
def lb(*a):
    out = np.zeros(4, np.float32)
    ok = L.or_liang_barsky(*[C.c_float(v) for v in a], fp(out))
    return ok, out


def test_liang_barsky_known():
    ok, o = lb(1, 2, 5, 6, 0, 0, 8, 8)
    assert ok and list(o) == [1, 2, 5, 6]             # inside: unchanged
    ok, o = lb(-10, 5, 10, 5, 0, 0, 8, 8)
    assert ok and list(o) == [0, 5, 8, 5]             # crossing both x edges
    ok, _ = lb(-10, 9, 10, 9, 0, 0, 8, 8)
    assert not ok                                      # parallel, outside (p == 0, q < 0)
    ok, o = lb(0, 8, 8, 8, 0, 0, 8, 8)
    assert ok and list(o) == [0, 8, 8, 8]             # on the closed boundary: kept
    ok, _ = lb(9, 0, 12, 3, 0, 0, 8, 8)
    assert not ok


# ---------------------------------------------------------------------------
# Bresenham / Wu (rasterization/line.rs:125-238)
# ---------------------------------------------------------------------------
def bres(x0, y0, x1, y1):
    xy = (C.c_int64 * 20000)()
    al = (C.c_double * 10000)()
    n = L.or_bresenham(x0, y0, x1, y1, xy, al, 10000)
    return np.frombuffer(xy, np.int64)[: 2 * n].reshape(-1, 2), np.frombuffer(al, np.float64)[:n]


def wu(x0, y0, x1, y1):
    xy = (C.c_int64 * 20000)()
    al = (C.c_double * 10000)()
    n = L.or_wu(x0, y0, x1, y1, xy, al, 10000)
    return np.frombuffer(xy, np.int64)[: 2 * n].reshape(-1, 2), np.frombuffer(al, np.float64)[:n]


def test_bresenham_properties():
    rng = np.random.default_rng(5)
    for _ in range(500):
        x0, y0, x1, y1 = (int(v) for v in rng.integers(-40, 40, 4))
        p, a = bres(x0, y0, x1, y1)
        assert tuple(p[0]) == (x0, y0) and tuple(p[-1]) == (x1, y1)   # endpoints inclusive
        assert len(p) == max(abs(x1 - x0), abs(y1 - y0)) + 1
        d = np.abs(np.diff(p, axis=0))
        assert d.max(initial=1) <= 1 and (d.sum(1) >= 1).all()          # 8-connected, no repeats
        assert (a == 1.0).all()
        L2 = math.hypot(x1 - x0, y1 - y0)
        if L2:
            dist = np.abs((x1 - x0) * (p[:, 1] - y0) - (y1 - y0) * (p[:, 0] - x0)) / L2
            assert dist.max() <= 0.5 * max(abs(x1 - x0), abs(y1 - y0)) / L2 + 1e-9 + 0.5


def test_wu_horizontal_closed_form():
    p, a = wu(2.0, 5.0, 10.0, 5.0)
    want = [((2, 5), 0.5), ((2, 6), 0.0), ((10, 5), 0.5), ((10, 6), 0.0)]
    for x in range(3, 10):
        want += [((x, 5), 1.0), ((x, 6), 0.0)]
    assert [tuple(q) for q in p] == [w[0] for w in want]
    assert list(a) == [w[1] for w in want]


def test_wu_steep_and_coverage_sums():
    p, a = wu(3.0, 1.0, 3.0, 9.0)   # vertical: steep, plots (y-major swapped back)
    assert all(q[0] in (3, 4) for q in p)
    span = [(tuple(p[i]), a[i]) for i in range(4, len(p))]
    for i in range(0, len(span), 2):
        assert span[i][1] + span[i + 1][1] == 1.0   # the two pixels of a column share coverage
    p, a = wu(0.25, 0.5, 20.75, 7.25)
    for i in range(4, len(p), 2):
        assert abs(a[i] + a[i + 1] - 1.0) < 1e-12
        assert p[i + 1][1] == p[i][1] + 1 and p[i + 1][0] == p[i][0]


# ---------------------------------------------------------------------------
# stencil.rs:109-159
# ---------------------------------------------------------------------------
def test_stencil_truth_tables():
    T = L.or_stencil_test
    # test(value = framebuffer, mask = reference): LessThan passes if ref < fb
    assert T(0, 5, 3) and not T(1, 5, 3)
    assert T(2, 5, 3) and not T(2, 3, 5) and not T(2, 4, 4)      # LessThan
    assert T(3, 3, 5) and not T(3, 5, 3)                          # GreaterThan
    assert T(4, 4, 4) and T(4, 5, 3) and not T(4, 3, 5)           # LessThanEq
    assert T(5, 4, 4) and T(5, 3, 5) and not T(5, 5, 3)           # GreaterThanEq
    assert T(6, 7, 7) and not T(6, 7, 8)
    assert T(7, 7, 8) and not T(7, 7, 7)
    O = L.or_stencil_op
    assert O(0, 0, 9, 3) == 9 and O(1, 0, 0x0F, 3) == 0xF0 and O(2, 0, 9, 3) == 0 and O(3, 0, 9, 3) == 3
    assert O(4, 1, 255, 0) == 0 and O(4, 0, 255, 0) == 255 and O(4, 0, 7, 0) == 8
    assert O(5, 1, 0, 0) == 255 and O(5, 0, 0, 0) == 0 and O(5, 1, 7, 0) == 6


# ---------------------------------------------------------------------------
# triangle coverage, depth, order (triangle.rs:23-156)
# ---------------------------------------------------------------------------
def test_right_triangle_inclusive_coverage():
    fb, st = draw16(screen_mesh([((0, 0), (8, 0), (0, 8))]), tile_w=128, tile_h=128)
    cov = (fb.color.reshape(16, 16) != 0xFF000000)
    ys, xs = np.nonzero(cov)
    # pixel centres (x+.5, y+.5) with x + y + 1 <= 8: all edges inclusive
    assert set(zip(xs, ys)) == {(x, y) for x in range(8) for y in range(8) if x + y <= 7}
    assert st["covered"] == 36


def test_shared_edge_covered_twice_and_ties_go_to_the_later_fragment():
    m = screen_mesh([((0, 0), (8, 0), (0, 8)), ((8, 0), (8, 8), (0, 8))])
    fb, st = draw16(m)
    assert st["covered"] == 36 + 36            # the diagonal belongs to both
    col = fb.color.reshape(16, 16)
    assert (col[:8, :8] != 0xFF000000).all()
    # equal depth everywhere: every diagonal pixel took the second triangle's
    # fragment too (d >= dt), so a two-colour version shows it
    m2 = dict(m)
    m2["u"] = np.array([0, 0, 0, 1, 1, 1], np.float32)
    fb2, _ = draw16(m2, fragment_program=3, fs_uniforms=np.zeros(32, np.float32))
    c2 = fb2.color.reshape(16, 16)
    diag = [c2[y, 7 - y] & 0xFF for y in range(8)]
    assert diag == [255] * 8


def test_depth_reversed_and_z_sign():
    # screen z = -(zn + 1)/2 for n=0, f=1: a clip z of 1 gives -1, z of -1 gives 0
    near = screen_mesh([((0, 0), (12, 0), (0, 12))], z=0.2)   # screen z -0.6
    far = screen_mesh([((0, 0), (12, 0), (0, 12))], z=0.6)    # screen z -0.8
    both = {k: (np.concatenate([near[k], far[k]]) if isinstance(near[k], np.ndarray) else near[k]) for k in near}
    fb, st = draw16(both)
    assert fb.depth.min() == F32MIN and np.float32(-0.6) in fb.depth
    assert st["shaded"] == st["covered"] // 2     # the farther one never passes d >= dt
    behind = screen_mesh([((0, 0), (12, 0), (0, 12))], z=-1.5)  # screen z +0.25 >= 0
    fb, st = draw16(behind)
    assert st["covered"] > 0 and st["shaded"] == 0 and (fb.depth == F32MIN).all()


def test_seam_pixels_blend_twice():
    # an 8x8 square across a 4-pixel tile seam: seam pixels are blended once
    # per containing tile (fragment.rs:196-212 inclusive tiles)
    m = screen_mesh([((0, 0), (8, 0), (0, 8)), ((8, 0), (8, 8), (0, 8))], z=0.5)
    fs = np.zeros(32, np.float32)
    fs[0:4] = (1.0, 1.0, 1.0, 0.5)
    fb, _ = draw16(m, blend=1, tile_w=4, tile_h=16, fs_uniforms=fs, clear=0xFF000000)
    col = fb.color.reshape(16, 16) & 0xFF
    # columns 4 and 8 are seams (tiles [0,4],[4,8],[8,12],[12,15])
    assert col[1, 2] != col[1, 4]
    once = oracle.lib().or_blend(1, 0x80FFFFFF, 0xFF000000) & 0xFF
    assert col[1, 2] == once


# ---------------------------------------------------------------------------
# build-defined programs: sampler, blend, mul_alpha; libm hypotf premise
# ---------------------------------------------------------------------------
def test_bilinear_at_texel_centres_and_midpoints():
    tex_w, tex_h, rgba = 64, 64, scenes.checkerboard(64, 8)
    t = oracle.OrTexture(tex_w, tex_h, rgba.ctypes.data_as(C.POINTER(C.c_uint32)))
    out = np.zeros(4, np.float32)
    for (ix, iy) in [(0, 0), (7, 7), (8, 0), (13, 40)]:
        u, v = (ix + 0.5) / 64, (iy + 0.5) / 64
        L.or_sample(C.byref(t), u, v, 1, 0, 0, fp(out))
        want = 1.0 if ((ix // 8 + iy // 8) % 2) else 0.0
        assert out[0] == want and out[3] == 1.0
    L.or_sample(C.byref(t), 8.0 / 64, 3.5 / 64, 1, 0, 0, fp(out))  # halfway between texel 7 and 8
    assert out[0] == np.float32(0.5)
    L.or_sample(C.byref(t), -3.0, -3.0, 1, 0, 0, fp(out))            # clamp
    assert out[0] == 0.0
    L.or_sample(C.byref(t), -3.0, 0.5, 1, 2, 0xFF0000FF, fp(out))   # border colour (red)
    assert out[0] == 1.0 and out[1] == 0.0


def test_blend_and_alpha():
    B = L.or_blend
    assert B(0, 0x11223344, 0xFFFFFFFF) == 0x11223344             # `()` replaces
    assert B(1, 0xFF102030, 0xFFFFFFFF) == 0xFF102030             # opaque over
    assert B(1, 0x00102030, 0xFF405060) == 0xFF405060             # transparent over
    M = L.or_mul_alpha
    assert M(1, 0xCC000000, 0.5) >> 24 == 102                      # f32 alpha: 204 * 0.5
    assert M(0, 0xCC000000, 0.7) >> 24 == 0                        # u8 alpha collapses (Q14)
    assert M(0, 0xFF000000, 1.0) >> 24 == 1                        # 255 * (1/255) -> 1


def test_hypotf_is_the_double_formula():
    # glibc's hypotf evaluates sqrt(x*x + y*y) in double; the device line
    # rasteriser relies on that being bit-identical.
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(20000) * 300).astype(np.float32)
    y = (rng.standard_normal(20000) * 300).astype(np.float32)
    for a, b in zip(x, y):
        ref = L.or_hypotf(float(a), float(b))
        dbl = np.float32(math.sqrt(float(a) * float(a) + float(b) * float(b)))
        assert np.float32(ref) == dbl


def test_f64_reciprocal_quotient_equals_f32_division():
    # the device forms triangle barycentrics as (float)(num * (1.0 / (double)det))
    rng = np.random.default_rng(11)
    a = ((rng.random(200000) * 4000 - 2000).astype(np.float32) * (rng.random(200000) * 4000 - 2000).astype(np.float32)).astype(np.float32)
    b = ((rng.random(200000) * 60).astype(np.float32) * (rng.random(200000) * 60).astype(np.float32)).astype(np.float32)
    with np.errstate(all="ignore"):
        ref = a / b
        got = (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(np.float32)
    same = (ref.view(np.uint32) == got.view(np.uint32)) | (np.isnan(ref) & np.isnan(got))
    assert same.all()


def _recip_quot(x, d):
    with np.errstate(all="ignore"):
        return (x.astype(np.float64) * (1.0 / d.astype(np.float64))).astype(np.float32)


def _same_f32(a, b):
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


def test_f64_reciprocal_normalize_equals_f32_division():
    # the device normalises vectors as quot(x, 1.0 / (double)len) per
    # component: equal to x / len in f32 over components and lengths of any
    # magnitude, denormals, zeros, infinities and NaN included
    rng = np.random.default_rng(29)
    n = 2_000_000
    v = rng.standard_normal((3, n)) * np.exp2(rng.integers(-140, 120, n))
    v = v.astype(np.float32)
    with np.errstate(all="ignore"):
        ln = np.sqrt(((v[0] * v[0]) + (v[1] * v[1])) + (v[2] * v[2]))
        for c in range(3):
            assert _same_f32(v[c] / ln, _recip_quot(v[c], ln)).all()
        # arbitrary numerators against every significand of one binade of divisors
        d = (np.arange(1 << 23, dtype=np.uint32) | np.uint32(0x3F800000)).view(np.float32)
        for e in (-126, -60, 0, 23, 100):
            x = (rng.random(d.size) * 2.0 - 1.0).astype(np.float32) * np.float32(2.0 ** e)
            assert _same_f32(x / d, _recip_quot(x, d)).all()
        # denormal numerators of either sign
        bits = rng.integers(0, 1 << 23, d.size).astype(np.uint32) | (rng.integers(0, 2, d.size).astype(np.uint32) << 31)
        x = bits.view(np.float32)
        assert _same_f32(x / d, _recip_quot(x, d)).all()
        assert _same_f32(x / (d * np.float32(2.0 ** -100)), _recip_quot(x, d * np.float32(2.0 ** -100))).all()
        sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4028235e38], np.float32)
        xs, ds = np.meshgrid(sp, sp)
        assert _same_f32(xs / ds, _recip_quot(xs, ds)).all()
