"""An independent numpy-float32 restatement of softrender's triangle draw.

TEST INFRASTRUCTURE ONLY.  It is a second, separately written statement of
the reference's single-threaded semantics, used to cross-check the C
oracle bit for bit on tiny scenes.  It is written from the reference's
source text and the build-defined programs documented in
include/softrender_hip.h, not from the oracle's C: every array is float32
and every expression is evaluated one IEEE operation at a time, in the
reference's order (numpy never fuses a multiply and an add).

Covered: the vertex programs (the header's uniform layout), run_to_fragment's
viewport transform, the screen tile grid, rasterize_triangle with culling,
stencil, depth and the fragment programs FLAT / UV_COLOR / TEX_LAMBERT /
BLINN_PHONG, the sampler and both blends.  Lines, points and the geometry
programs have their own pins elsewhere.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
F32_MIN = f32(np.finfo(np.float32).min)  # Bounded::min_value(), depth.rs:31

# include/softrender_hip.h enumerations
VP_NORMAL_UV, VP_NORMAL_WORLD = 0, 1
FP_FLAT, FP_BLINN_PHONG, FP_TEX_LAMBERT, FP_UV_COLOR = 0, 1, 2, 3
BLEND_REPLACE, BLEND_OVER = 0, 1
CULL_NONE, CULL_CW, CULL_CCW = 0, 1, 2


class CoordinateError(Exception):
    """cast(value).unwrap() of a NaN bound panics (triangle.rs:70)."""


# ---------------------------------------------------------------------------
# vertex stage: the header's programs.  Matrix rows times (p, 1), summed
# k = 0..3 starting from +0.0 (the mat-vec order the oracle documents as the
# nalgebra one).
# ---------------------------------------------------------------------------
def _matvec_rows(U, first, rows, cols, vec):
    out = []
    for r in range(rows):
        acc = np.zeros_like(vec[0])
        for k in range(cols):
            acc = acc + f32(U[first + r * cols + k]) * vec[k]
        out.append(acc)
    return out


def _normalize3(x, y, z):
    ln = np.sqrt(((x * x) + (y * y)) + (z * z))
    return x / ln, y / ln, z / ln


def vertex_stage(mesh: dict, vp: int, U: np.ndarray):
    """clip (4, n) and varyings (K, n) of every vertex (vertex.rs:87-120)."""
    n = len(mesh["px"])
    z0 = np.zeros(n, f32)
    get = lambda k: z0 if mesh.get(k) is None else np.asarray(mesh[k], f32)  # noqa: E731
    pin = [get("px"), get("py"), get("pz"), np.ones(n, f32)]
    nin = [get("nx"), get("ny"), get("nz")]
    clip = _matvec_rows(U, 0, 4, 4, pin)
    nrm = _normalize3(*_matvec_rows(U, 16, 3, 3, nin))
    if vp == VP_NORMAL_WORLD:
        k = list(nrm) + _matvec_rows(U, 25, 3, 4, pin)
    else:
        k = list(nrm) + [get("u"), get("v")]
    return np.stack(clip), np.stack(k)


def viewport_transform(clip, vpt):
    """ClipVertex::normalize, clipvertex.rs:89-126."""
    x, y, z, w = clip
    left, bottom, width, height, near, far = (f32(v) for v in vpt)
    right = left + width
    top = bottom + height
    M = [[(right - left) / f32(2.0), f32(0), f32(0), (right + left) / f32(2.0)],
         [f32(0), (top - bottom) / f32(-2.0), f32(0), (top + bottom) / f32(2.0)],
         [f32(0), f32(0), (far - near) / f32(-2.0), (far + near) / f32(-2.0)]]
    v = [x / w, y / w, z / w, np.ones_like(x)]
    out = []
    for r in range(3):
        acc = np.zeros_like(x)
        for k in range(4):
            acc = acc + M[r][k] * v[k]
        out.append(acc)
    out.append(f32(1.0) / w)
    return np.stack(out)


def tile_grid(W, H, tw, th):
    """fragment.rs:188-216: inclusive rectangles in row-major push order."""
    tiles = []
    xmax, ymax = W - 1, H - 1
    y = 0
    while y < ymax:
        ny = min(y + th, ymax)
        x = 0
        while x < xmax:
            nx = min(x + tw, xmax)
            tiles.append((x, y, nx, ny))
            x = nx
        y = ny
    return tiles


# ---------------------------------------------------------------------------
# colour: unorm8 <-> f32, sampler, programs, blend (build-defined; the
# header documents each)
# ---------------------------------------------------------------------------
def unorm(c):
    return (np.asarray(c, np.uint32) & np.uint32(0xFF)).astype(f32) / f32(255.0)


def quant8(x):
    c = np.fmin(np.fmax(x, f32(0.0)), f32(1.0))
    return ((c * f32(255.0)) + f32(0.5)).astype(np.uint32)


def pack(r, g, b, a):
    return r | (g << np.uint32(8)) | (b << np.uint32(16)) | (a << np.uint32(24))


def _texel(tex, ix, iy, edge, border):
    W, H, rgba = tex
    W, H = int(W), int(H)
    if edge == 1:  # wrap
        ix = ((ix % W) + W) % W
        iy = ((iy % H) + H) % H
        c = rgba[iy * W + ix]
    elif edge == 2:  # border colour outside
        inside = (ix >= 0) & (iy >= 0) & (ix < W) & (iy < H)
        c = np.where(inside, rgba[np.clip(iy, 0, H - 1) * W + np.clip(ix, 0, W - 1)], np.uint32(border))
    else:  # clamp
        c = rgba[np.clip(iy, 0, H - 1) * W + np.clip(ix, 0, W - 1)]
    c = np.asarray(c, np.uint32)
    return [unorm(c >> np.uint32(s)) for s in (0, 8, 16, 24)]


def sample(tex, u, v, filt, edge, border):
    """Texel centres at (i + 0.5) / W; coordinates clamped to +-1e8 before
    the integer conversion."""
    W, H = f32(tex[0]), f32(tex[1])
    if filt == 0:
        fx = np.fmin(np.fmax(u * W, f32(-1e8)), f32(1e8))
        fy = np.fmin(np.fmax(v * H, f32(-1e8)), f32(1e8))
        return _texel(tex, np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64), edge, border)
    fx = np.fmin(np.fmax((u * W) - f32(0.5), f32(-1e8)), f32(1e8))
    fy = np.fmin(np.fmax((v * H) - f32(0.5), f32(-1e8)), f32(1e8))
    x0, y0 = np.floor(fx), np.floor(fy)
    ax, ay = fx - x0, fy - y0
    ix, iy = x0.astype(np.int64), y0.astype(np.int64)
    t00 = _texel(tex, ix, iy, edge, border)
    t10 = _texel(tex, ix + 1, iy, edge, border)
    t01 = _texel(tex, ix, iy + 1, edge, border)
    t11 = _texel(tex, ix + 1, iy + 1, edge, border)
    bx, by = f32(1.0) - ax, f32(1.0) - ay
    out = []
    for c in range(4):
        top = (t00[c] * bx) + (t10[c] * ax)
        bot = (t01[c] * bx) + (t11[c] * ax)
        out.append((top * by) + (bot * ay))
    return out


def _dot3(a, b):
    return ((a[0] * b[0]) + (a[1] * b[1])) + (a[2] * b[2])


def fragment(fp, U, k, tex=None, filt=1, edge=0, border=0):
    """The header's fragment programs; k = interpolated varyings (K, n)."""
    U = [f32(x) for x in U]
    n = k.shape[1]
    if fp == FP_BLINN_PHONG:
        # U: L[0..3) eye[3..6) base[6..9) ambient[9] kd[10] ks[11] alpha[12]
        nv = _normalize3(k[0], k[1], k[2])
        vv = _normalize3(U[3] - k[3], U[4] - k[4], U[5] - k[5])
        ndl = _dot3(nv, U[0:3])
        diff = np.where(ndl > f32(0), ndl, f32(0))
        hv = _normalize3(U[0] + vv[0], U[1] + vv[1], U[2] + vv[2])
        ndh = _dot3(nv, hv)
        s = np.where(ndh > f32(0), ndh, f32(0))
        for _ in range(5):  # s^32
            s = s * s
        lit = U[9] + (U[10] * diff)
        spec = U[11] * s
        return pack(quant8((U[6] * lit) + spec), quant8((U[7] * lit) + spec), quant8((U[8] * lit) + spec),
                    np.full(n, quant8(U[12]), np.uint32))
    if fp == FP_TEX_LAMBERT:
        # U: L[0..3) ambient[3] kd[4] alpha[5]
        nv = _normalize3(k[0], k[1], k[2])
        ndl = _dot3(nv, U[0:3])
        diff = np.where(ndl > f32(0), ndl, f32(0))
        lit = U[3] + (U[4] * diff)
        t = sample(tex, k[3], k[4], filt, edge, border)
        return pack(quant8(t[0] * lit), quant8(t[1] * lit), quant8(t[2] * lit), quant8(t[3] * U[5]))
    if fp == FP_UV_COLOR:
        return pack(quant8(k[3]), quant8(k[4]), np.full(n, quant8(U[0]), np.uint32),
                    np.full(n, quant8(U[1]), np.uint32))
    c = pack(quant8(U[0]), quant8(U[1]), quant8(U[2]), quant8(U[3]))
    return np.full(n, c, np.uint32)


def blend(mode, src, dst):
    """`()` replaces (blend.rs:33-40); OVER is src-alpha over in f32."""
    if mode != BLEND_OVER:
        return src
    a = unorm(src >> np.uint32(24))
    ia = f32(1.0) - a
    ch = [quant8((unorm(src >> np.uint32(s)) * a) + (unorm(dst >> np.uint32(s)) * ia)) for s in (0, 8, 16)]
    oa = quant8(a + (unorm(dst >> np.uint32(24)) * ia))
    return pack(ch[0], ch[1], ch[2], oa)


# ---------------------------------------------------------------------------
# stencil (stencil.rs:109-159): test(value = framebuffer, mask = reference)
# ---------------------------------------------------------------------------
def stencil_test(test, value, ref):
    value = value.astype(np.int64)
    return {0: np.ones_like(value, bool), 1: np.zeros_like(value, bool), 2: ref < value,
            3: ref > value, 4: ref <= value, 5: ref >= value, 6: ref == value,
            7: ref != value}[test]


def stencil_op(op, wrap, value, ref):
    v = value.astype(np.int64)
    if op == 0:
        r = v
    elif op == 1:
        r = (~v) & 0xFF
    elif op == 2:
        r = np.zeros_like(v)
    elif op == 3:
        r = np.full_like(v, ref & 0xFF)
    elif op == 4:
        r = (v + 1) & 0xFF if wrap else np.minimum(v + 1, 255)
    else:
        r = (v - 1) & 0xFF if wrap else np.maximum(v - 1, 0)
    return r.astype(np.uint8)


# ---------------------------------------------------------------------------
# rasterize_triangle, triangle.rs:23-156
# ---------------------------------------------------------------------------
def _clamp_as_int(value, lo, hi):
    # triangle.rs:66-72: value < min ? min : value > max ? max : cast(value)
    if value < f32(lo):
        return lo
    if value > f32(hi):
        return hi
    if np.isnan(value):
        raise CoordinateError()
    return int(np.trunc(value))


def raster_triangle(fb, tile, a, b, c, desc, tex, counts):
    x1, y1 = a["p"][0], a["p"][1]
    x2, y2 = b["p"][0], b["p"][1]
    x3, y3 = c["p"][0], c["p"][1]
    cull = desc["cull"]
    if cull != CULL_NONE:  # :54-61, shoelace summed left to right
        area = (((((x1 * y2) + (x2 * y3)) + (x3 * y1)) - (x2 * y1)) - (x3 * y2)) - (x1 * y3)
        winding = CULL_CW if np.signbit(area) else CULL_CCW
        if winding == cull:
            return
    det = ((y2 - y3) * (x1 - x3)) + ((x3 - x2) * (y1 - y3))  # :64
    tx0, ty0, tx1, ty1 = tile
    mnx = _clamp_as_int(np.fmin(np.fmin(x1, x2), x3), tx0, tx1)  # :74-78
    mny = _clamp_as_int(np.fmin(np.fmin(y1, y2), y3), ty0, ty1)
    mxx = _clamp_as_int(np.fmax(np.fmax(x1, x2), x3), tx0, tx1)
    mxy = _clamp_as_int(np.fmax(np.fmax(y1, y2), y3), ty0, ty1)
    if mnx > mxx or mny > mxy:
        return
    W = fb["w"]
    py, px = np.mgrid[mny:mxy + 1, mnx:mxx + 1]
    px, py = px.ravel(), py.ravel()
    idx = px + py * W  # coordinate.rs:47-52
    ref = desc["stencil_ref"] & 0xFF
    sv = fb["stencil"][idx]
    sp = stencil_test(desc["stencil_test"], sv, ref)  # :94
    idx, px, py, sv = idx[sp], px[sp], py[sp], sv[sp]
    fb["stencil"][idx] = stencil_op(desc["stencil_op"], desc["stencil_wrap"], sv, ref)  # :96-99
    x = px.astype(f32) + f32(0.5)  # :104-105
    y = py.astype(f32) + f32(0.5)
    u = (((y2 - y3) * (x - x3)) + ((x3 - x2) * (y - y3))) / det  # :108
    v = (((y3 - y1) * (x - x3)) + ((x1 - x3) * (y - y3))) / det  # :109
    w = (f32(1.0) - u) - v  # :110
    inside = ~((u < 0) | (v < 0) | (w < 0))  # :113
    counts["covered"] += int(inside.sum())
    idx, u, v, w = idx[inside], u[inside], v[inside], w[inside]

    def bary(qa, qb, qc):  # interpolate.rs:26-28, 43-64
        return ((qa * u) + (qb * v)) + (qc * w)

    z = bary(a["p"][2], b["p"][2], c["p"][2])  # :115-117
    front = (z < 0) & (z >= fb["depth"][idx])  # :120-126
    idx, u, v, w, z = idx[front], u[front], v[front], w[front], z[front]
    if len(idx) == 0:
        return
    k = np.stack([bary(a["k"][i], b["k"][i], c["k"][i]) for i in range(len(a["k"]))])
    col = fragment(desc["fragment_program"], desc["fs_uniforms"], k, tex, desc["tex_filter"],
                   desc["tex_edge"], desc["tex_border"])
    counts["shaded"] += len(idx)
    fb["color"][idx] = blend(desc["blend"], col, fb["color"][idx])  # :138-142
    fb["depth"][idx] = z


def draw(fb, mesh, desc, tex=None):
    """One triangle draw without a geometry program (render_mesh ->
    run_to_fragment -> FragmentShader::run).  `fb` is a dict of numpy planes
    (w, h, color, depth, stencil), updated in place."""
    clip, K = vertex_stage(mesh, desc["vertex_program"], np.asarray(desc["vs_uniforms"], f32))
    scr = viewport_transform(clip, desc["viewport"])
    idx = mesh.get("indices")
    n = len(idx) if idx is not None else len(mesh["px"])
    idx = np.arange(n) if idx is None else np.asarray(idx, np.int64)
    verts = [dict(p=scr[:, i], k=K[:, i]) for i in range(scr.shape[1])]
    counts = dict(covered=0, shaded=0)
    if fb["w"] <= 1 or fb["h"] <= 1:
        return counts
    with np.errstate(all="ignore"):  # det = 0 gives inf / NaN barycentrics, as in the reference
        for tile in tile_grid(fb["w"], fb["h"], desc["tile_w"], desc["tile_h"]):
            for t in range(0, n - n % 3, 3):  # fragment.rs:268-274, list order
                raster_triangle(fb, tile, verts[idx[t]], verts[idx[t + 1]], verts[idx[t + 2]], desc, tex,
                                counts)
    return counts
