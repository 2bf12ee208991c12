"""Seeded synthetic scenes for the five workloads (SplitMix64, numpy).

Every scene is a plain dict of numpy arrays plus draw settings, so the same
bytes feed the HIP path and the CPU oracle.  Floats from the generator are
``(u64 >> 40) * 2**-24`` in [0, 1); normals come from Box-Muller.

Workloads (sizes as the reference's benchmark plan names them):
  sphere_512      UV sphere 32 stacks x 64 slices (2,145 vertices, 3,968
                  triangles), 512x512, Lambert + Blinn-Phong, back-face cull.
  soup_1080p      1M-triangle clip-space soup at 1920x1080, log-normal area
                  (median 20 px, sigma 1), 1024^2 checkerboard, bilinear,
                  20% translucent over-blend.
  clip_4k         3840x2160: 500k triangles (30% cross z = 0, near-plane clip
                  in the geometry stage), 200k Wu lines, 100k points, u8
                  stencil increment.
  batch_mix       256 scenes, 1k..500k triangles log-uniform, resolutions
                  {256^2, 512^2, 1920x1080}.
  serve_meshes    64 resident meshes, 1k..1M triangles log-uniform.
"""
from __future__ import annotations

import math
import os

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
F32_MIN = np.float32(np.finfo(np.float32).min)  # f32::MIN, the far depth

SEED_SPHERE = 0x5EED0001
SEED_SOUP = 0x5EED0002
SEED_CLIP = 0x5EED0003
SEED_BATCH = 0x5EED0004
SEED_SERVE = 0x5EED0005


_POOL = None
_CHUNK = 1 << 19


def _pool():
    global _POOL
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
    return _POOL


def _chunked(n: int, fn, dtype):
    """fn(lo, hi) -> array of hi - lo values, evaluated in parallel chunks
    (numpy releases the GIL inside each operation)."""
    if n <= 2 * _CHUNK:
        return fn(0, n)
    out = np.empty(n, dtype)
    bounds = [(lo, min(n, lo + _CHUNK)) for lo in range(0, n, _CHUNK)]

    def run(b):
        out[b[0]:b[1]] = fn(b[0], b[1])

    list(_pool().map(run, bounds))
    return out


class SplitMix64:
    """Vectorised SplitMix64 stream: draw k of the stream is
    mix(state0 + k * GOLDEN), so any range of draws can be computed
    independently (large requests run in parallel chunks)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    @staticmethod
    def _mix(base, lo: int, hi: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(lo + 1, hi + 1, dtype=np.uint64)
            z = base + idx * GOLDEN
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def _advance(self, n: int):
        base = self.state
        with np.errstate(over="ignore"):
            self.state = np.uint64(self.state + np.uint64(n) * GOLDEN)
        return base

    def next_u64(self, n: int) -> np.ndarray:
        base = self._advance(n)
        return _chunked(n, lambda lo, hi: self._mix(base, lo, hi), np.uint64)

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        base = self._advance(n)

        def f(a, b):
            u = (self._mix(base, a, b) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
            return lo + (hi - lo) * u
        return _chunked(n, f, np.float64)

    def normal(self, n: int) -> np.ndarray:
        b1 = self._advance(n)
        b2 = self._advance(n)

        def f(a, b):
            u1 = (self._mix(b1, a, b) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
            u2 = (self._mix(b2, a, b) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
            return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)
        return _chunked(n, f, np.float64)


# ---------------------------------------------------------------------------
# matrices (float64 on the host, handed to the device as float32 uniforms)
# ---------------------------------------------------------------------------
def look_at(eye, target, up) -> np.ndarray:
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def perspective_zero_to_one(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """Right-handed projection with clip z in [0, w]: the near plane maps to
    z = 0, matching ClippingPlane::Near (z >= 0, geometry/clip.rs:41)."""
    f = 1.0 / math.tan(math.radians(fovy_deg) / 2.0)
    p = np.zeros((4, 4))
    p[0, 0] = f / aspect
    p[1, 1] = f
    p[2, 2] = far / (near - far)
    p[2, 3] = near * far / (near - far)
    p[3, 2] = -1.0
    return p


def vs_uniforms(mvp: np.ndarray, normal3: np.ndarray | None = None,
                model34: np.ndarray | None = None) -> np.ndarray:
    u = np.zeros(48, np.float32)
    u[0:16] = np.asarray(mvp, np.float64).reshape(16)
    u[16:25] = (np.eye(3) if normal3 is None else np.asarray(normal3)).reshape(9)
    u[25:37] = (np.eye(4)[:3] if model34 is None else np.asarray(model34)).reshape(12)
    return u


def checkerboard(size: int = 1024, square: int = 32) -> np.ndarray:
    y, x = np.mgrid[0:size, 0:size]
    white = ((x // square + y // square) % 2).astype(bool)
    out = np.full((size, size), 0xFF000000, np.uint32)
    out[white] = 0xFFFFFFFF
    return out.reshape(-1)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# ---------------------------------------------------------------------------
# UV sphere (Lambert + Blinn-Phong, back-face culled)
# ---------------------------------------------------------------------------
def uv_sphere(stacks: int = 32, slices: int = 64, radius: float = 1.0):
    i = np.arange(stacks + 1, dtype=np.float64)[:, None]
    j = np.arange(slices + 1, dtype=np.float64)[None, :]
    theta = math.pi * i / stacks
    phi = 2.0 * math.pi * j / slices
    x = np.sin(theta) * np.cos(phi)
    y = np.cos(theta) * np.ones_like(phi)
    z = np.sin(theta) * np.sin(phi)
    pos = np.stack([x, y, z], -1).reshape(-1, 3) * radius
    nrm = np.stack([x, y, z], -1).reshape(-1, 3)
    idx = []
    row = slices + 1
    for a in range(stacks):
        for b in range(slices):
            v00, v01 = a * row + b, a * row + b + 1
            v10, v11 = (a + 1) * row + b, (a + 1) * row + b + 1
            if a != 0:
                idx += [v00, v10, v01]
            if a != stacks - 1:
                idx += [v01, v10, v11]
    return pos.astype(np.float32), nrm.astype(np.float32), np.asarray(idx, np.uint32)


def screen_winding(mvp, vp_w, vp_h, p3):
    """Sign of the reference's shoelace sum (triangle.rs:56) for one
    triangle given in object space: True if it reads as clockwise."""
    pts = []
    for p in p3:
        c = mvp @ np.append(np.asarray(p, np.float64), 1.0)
        n = c[:3] / c[3]
        pts.append(((n[0] + 1) * vp_w / 2, (1 - n[1]) * vp_h / 2))
    (x1, y1), (x2, y2), (x3, y3) = pts
    a = x1 * y2 + x2 * y3 + x3 * y1 - x2 * y1 - x3 * y2 - x1 * y3
    return a < 0


def sphere_scene(width: int = 512, height: int = 512):
    pos, nrm, idx = uv_sphere()
    view = look_at((0, 0, 3), (0, 0, 0), (0, 1, 0))
    proj = perspective_zero_to_one(60.0, width / height, 0.1, 100.0)
    mvp = proj @ view
    # the front-most triangle of the sphere decides which winding is "front"
    centre = np.argmax(pos[idx].reshape(-1, 3, 3)[:, :, 2].mean(1))
    cw_front = screen_winding(mvp, width, height, pos[idx[3 * centre:3 * centre + 3]])
    light = _unit((1, 1, 1))
    fs = np.zeros(32, np.float32)
    fs[0:3] = light
    fs[3:6] = (0, 0, 3)
    fs[6:9] = (0.8, 0.3, 0.2)
    fs[9], fs[10], fs[11], fs[12] = 0.1, 0.8, 0.5, 1.0
    return dict(
        name="sphere_512", width=width, height=height, primitive=3,
        px=pos[:, 0].copy(), py=pos[:, 1].copy(), pz=pos[:, 2].copy(),
        nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(),
        u=None, v=None, indices=idx,
        desc=dict(vertex_program=1, fragment_program=1, blend=0,
                  cull=2 if cw_front else 1,  # cull the back-facing winding
                  vs_uniforms=vs_uniforms(mvp), fs_uniforms=fs,
                  viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0)),
        clear=0xFF000000, texture=None,
    )


# ---------------------------------------------------------------------------
# The full example's frame (full_example/src/main.rs:66-92): every mesh is
# drawn lit, then its normals are drawn as green lines over it by a geometry
# program fed by the same vertex program.
# ---------------------------------------------------------------------------
GS_VERTEX_NORMALS, GS_FACE_NORMALS = 3, 4
NORMAL_LENGTH = 0.05  # full_example/src/shaders.rs:33


def rotation_y(deg: float) -> np.ndarray:
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float64)


def normals_scene(pos, nrm, idx, width: int = 320, height: int = 320,
                  program: int = GS_FACE_NORMALS, aa: int = 0,
                  length: float = NORMAL_LENGTH, yaw: float = 30.0):
    """Lit mesh plus its vertex or face normals, as the full example draws a
    model; `pos`/`nrm` are (n, 3) arrays and `idx` a triangle list."""
    model = rotation_y(yaw)
    model[:3, 3] = (0.0, 0.1, 0.0)
    view = look_at((0, 0, 3), (0, 0, 0), (0, 1, 0))
    proj = perspective_zero_to_one(60.0, width / height, 0.1, 100.0)
    pv = proj @ view
    mvp = pv @ model
    mit = np.linalg.inv(model[:3, :3]).T
    pos = np.asarray(pos, np.float32)
    nrm = np.asarray(nrm, np.float32)
    idx = np.asarray(idx, np.uint32)
    tri = pos[idx].reshape(-1, 3, 3)
    centre = int(np.argmax(tri[:, :, 2].mean(1)))
    cw_front = screen_winding(mvp, width, height, tri[centre])
    fs = np.zeros(32, np.float32)
    fs[0:3] = _unit((1, 1, 1))
    fs[3:6] = (0, 0, 3)
    fs[6:9] = (0.25, 0.25, 0.25)
    fs[9], fs[10], fs[11], fs[12] = 0.1, 0.8, 0.5, 1.0
    green = np.zeros(32, np.float32)
    green[0:4] = (0.0, 1.0, 0.0, 1.0)  # fragment_shader_green, shaders.rs:97-99
    vsu = vs_uniforms(mvp, mit, model[:3])
    vp = (0.0, 0.0, float(width), float(height), 0.0, 1.0)
    mesh = dict(primitive=3, px=pos[:, 0].copy(), py=pos[:, 1].copy(), pz=pos[:, 2].copy(),
                nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(), u=None, v=None, indices=idx)
    # the example's Color is f32 RGBA (full_example/src/color.rs): f32 alpha
    lit = dict(mesh, desc=dict(vertex_program=1, fragment_program=1, blend=1, alpha_mode=1,
                               cull=2 if cw_front else 1, vs_uniforms=vsu, fs_uniforms=fs, viewport=vp))
    gsu = np.zeros(20, np.float32)
    gsu[:16] = pv.reshape(16)
    gsu[16] = length
    lines = dict(mesh, desc=dict(vertex_program=1, geometry_program=program, fragment_program=0,
                                 blend=1, aa_lines=aa, alpha_mode=1, vs_uniforms=vsu, fs_uniforms=green,
                                 gs_uniforms=gsu, viewport=vp))
    return dict(name=f"normals_{program}_{width}x{height}", width=width, height=height,
                clear=0xFF000000, texture=None, draws=[lit, lines])


def example_lights() -> np.ndarray:
    """The full example's four lights (full_example/src/lib.rs:55-64):
    rows of position xyz, colour rgb, intensity."""
    return np.array([[-1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 9.0],
                     [1.0, 1.0, 1.0, 0.6, 0.6, 1.0, 9.0],
                     [0.0, 3.0, -1.0, 1.0, 0.3, 0.3, 25.0],
                     [-2.0, -1.0, 1.0, 0.7, 1.0, 0.7, 25.0]], np.float64)


def fresnel_scene(pos, nrm, idx, width: int = 320, height: int = 180, translate=(0.0, 0.0, 0.0),
                  object_deg: float = 45.0):
    """The full example's first frame (full_example/src/main.rs:11-42,
    lib.rs:22-66): the model rotated about y (plus `translate`, which makes
    the normal's w non-zero), the camera at (cos 75deg, 1, sin 75deg) * 2
    looking at (0, 0.25, 0), 65deg field of view, four lights, the Fresnel
    program (SR_VP_WORLD4_NORMAL4 + SR_FP_FRESNEL4) and the normalised
    over-blend, on an RGBAF32 framebuffer cleared to (0.01, 0.01, 0.01, 1)."""
    model = rotation_y(object_deg)
    model[:3, 3] = translate
    mit = np.linalg.inv(model).T
    a = np.deg2rad(75.0)
    cam = np.array([np.cos(a) * 2.0, 2.0, np.sin(a) * 2.0])
    view = look_at(cam, (0.0, 0.25, 0.0), (0.0, 1.0, 0.0))
    proj = perspective_zero_to_one(65.0, width / height, 0.1, 1000.0)
    mvp = proj @ view @ model
    vsu = np.zeros(48, np.float32)
    vsu[0:16] = mvp.reshape(16)
    vsu[16:25] = mit[:3, :3].reshape(9)
    vsu[25:37] = model[:3].reshape(12)
    vsu[37:40] = mit[3, :3]
    vsu[40:44] = model[3]
    fs = np.zeros(32, np.float32)
    fs[0:3] = cam
    fs[3:31] = example_lights().reshape(28)
    fs[31] = 0.25
    pos = np.asarray(pos, np.float32)
    nrm = np.asarray(nrm, np.float32)
    mesh = dict(primitive=3, px=pos[:, 0].copy(), py=pos[:, 1].copy(), pz=pos[:, 2].copy(),
                nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(), u=None, v=None,
                indices=np.asarray(idx, np.uint32))
    desc = dict(vertex_program=2, fragment_program=5, blend=2, cull=0, vs_uniforms=vsu, fs_uniforms=fs,
                viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))
    return dict(mesh, name=f"fresnel_{width}x{height}", width=width, height=height, texture=None,
                clear_f32=(0.01, 0.01, 0.01, 1.0), desc=desc)


# ---------------------------------------------------------------------------
# Clip-space triangle soup (the 1080p / 1M-triangle frame)
# ---------------------------------------------------------------------------
def _unit3(x, y, z):
    """(x, y, z) / max(|(x, y, z)|, 1e-12), componentwise."""
    n = np.maximum(np.sqrt((x * x + y * y) + z * z), 1e-12)
    return x / n, y / n, z / n


def _soup_arrays(rng: SplitMix64, n: int, width: int, height: int,
                 median_area: float = 20.0, sigma: float = 1.0,
                 z_range=(-0.99, 0.99), cover=1.0):
    cx = rng.uniform(n, -cover, cover)
    cy = rng.uniform(n, -cover, cover)
    cz = rng.uniform(n, *z_range)
    area = median_area * np.exp(sigma * rng.normal(n))
    theta = rng.uniform(n, 0.0, 2.0 * math.pi)
    side = np.sqrt(4.0 * area / math.sqrt(3.0))
    r = side / math.sqrt(3.0)  # circumradius, pixels
    ang = theta[:, None] + np.array([0.0, 2.0 * math.pi / 3.0, 4.0 * math.pi / 3.0])[None, :]
    x = cx[:, None] + r[:, None] * np.cos(ang) * (2.0 / width)
    y = cy[:, None] + r[:, None] * np.sin(ang) * (2.0 / height)
    z = np.repeat(cz[:, None], 3, 1)
    nv = 3 * n
    nx, ny, nz = _unit3(rng.normal(nv), rng.normal(nv), rng.normal(nv))
    nrm = np.empty((nv, 3), np.float32)
    nrm[:, 0], nrm[:, 1], nrm[:, 2] = nx, ny, nz
    uv = np.empty((nv, 2), np.float32)
    uv[:, 0], uv[:, 1] = rng.uniform(nv), rng.uniform(nv)
    return (x.reshape(-1).astype(np.float32), y.reshape(-1).astype(np.float32),
            z.reshape(-1).astype(np.float32), nrm, uv)


def tex_lambert_uniforms(alpha: float = 0.8) -> np.ndarray:
    fs = np.zeros(32, np.float32)
    fs[0:3] = _unit((0.3, 0.5, 0.8))
    fs[3], fs[4], fs[5] = 0.2, 0.8, alpha
    return fs


def soup_scene(n_tris: int = 1_000_000, width: int = 1920, height: int = 1080,
               seed: int = SEED_SOUP, alpha: float = 0.8):
    rng = SplitMix64(seed)
    x, y, z, nrm, uv = _soup_arrays(rng, n_tris, width, height)
    return dict(
        name=f"soup_{width}x{height}_{n_tris}", width=width, height=height, primitive=3,
        px=x, py=y, pz=z, nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(),
        u=uv[:, 0].copy(), v=uv[:, 1].copy(), indices=None,
        desc=dict(vertex_program=0, fragment_program=2, blend=1, cull=0,
                  vs_uniforms=vs_uniforms(np.eye(4)), fs_uniforms=tex_lambert_uniforms(alpha),
                  tex_filter=1, tex_edge=0,
                  viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0)),
        clear=0xFF000000, texture=(1024, 1024, checkerboard(1024, 32)),
    )


# ---------------------------------------------------------------------------
# 4K clipping / primitive-heavy frame: three draws on one framebuffer
# ---------------------------------------------------------------------------
def clip_scene(n_tris: int = 500_000, n_lines: int = 200_000, n_points: int = 100_000,
               width: int = 3840, height: int = 2160, seed: int = SEED_CLIP,
               straddle: float = 0.3):
    rng = SplitMix64(seed)
    x, y, z, nrm, uv = _soup_arrays(rng, n_tris, width, height, z_range=(0.01, 0.99))
    # 30% of the triangles get 1 or 2 vertices pushed behind the near plane
    ns = int(round(straddle * n_tris))
    pick = rng.uniform(n_tris).argsort()[:ns]
    nbehind = 1 + (rng.uniform(ns) < 0.5).astype(np.int64)
    first = (rng.uniform(ns) * 3).astype(np.int64) % 3
    zz = z.reshape(-1, 3).astype(np.float64)
    depth_behind = rng.uniform(ns, 0.05, 0.6)
    for k in range(2):
        sel = nbehind > k
        col = (first + k) % 3
        zz[pick[sel], col[sel]] = -depth_behind[sel]
    z = zz.reshape(-1).astype(np.float32)
    stencil = dict(stencil_test=0, stencil_op=4, stencil_wrap=0)
    tri_desc = dict(vertex_program=0, geometry_program=2, fragment_program=2, blend=1, cull=0,
                    vs_uniforms=vs_uniforms(np.eye(4)), fs_uniforms=tex_lambert_uniforms(0.8),
                    tex_filter=1, tex_edge=0,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0), **stencil)
    tris = dict(px=x, py=y, pz=z, nx=nrm[:, 0].copy(), ny=nrm[:, 1].copy(), nz=nrm[:, 2].copy(),
                u=uv[:, 0].copy(), v=uv[:, 1].copy(), indices=None, primitive=3, desc=tri_desc)
    # lines: endpoints in the clip cube, log-normal length (median 40 px),
    # kept >= 2 px from the right/bottom edges (the reference writes out of
    # bounds there, line.rs:55-58)
    L = 40.0 * np.exp(0.6 * rng.normal(n_lines))
    ang = rng.uniform(n_lines, 0.0, 2.0 * math.pi)
    mx = width - 3.0
    my = height - 3.0
    sx = rng.uniform(n_lines, 0.0, mx)
    sy = rng.uniform(n_lines, 0.0, my)
    ex = np.clip(sx + L * np.cos(ang), 0.0, mx)
    ey = np.clip(sy + L * np.sin(ang), 0.0, my)
    lx = np.stack([sx, ex], 1).reshape(-1) * (2.0 / width) - 1.0
    ly = 1.0 - np.stack([sy, ey], 1).reshape(-1) * (2.0 / height)
    lz = np.repeat(rng.uniform(n_lines, 0.02, 0.98), 2)
    luv = np.stack([rng.uniform(2 * n_lines), rng.uniform(2 * n_lines)], -1)
    fs_lines = np.zeros(32, np.float32)
    fs_lines[0], fs_lines[1] = 0.9, 1.0
    line_desc = dict(vertex_program=0, geometry_program=0, fragment_program=3, blend=1, cull=0,
                     aa_lines=1, alpha_mode=1, vs_uniforms=vs_uniforms(np.eye(4)),
                     fs_uniforms=fs_lines,
                     viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0), **stencil)
    lines = dict(px=lx.astype(np.float32), py=ly.astype(np.float32), pz=lz.astype(np.float32),
                 nx=None, ny=None, nz=None, u=luv[:, 0].astype(np.float32).copy(),
                 v=luv[:, 1].astype(np.float32).copy(), indices=None, primitive=2, desc=line_desc)
    px = rng.uniform(n_points, -1.0, 1.0)
    py = rng.uniform(n_points, -1.0, 1.0)
    pz = rng.uniform(n_points, 0.01, 0.99)
    fs_pts = np.zeros(32, np.float32)
    fs_pts[0:4] = (1.0, 1.0, 0.2, 1.0)
    point_desc = dict(vertex_program=0, geometry_program=0, fragment_program=0, blend=0, cull=0,
                      vs_uniforms=vs_uniforms(np.eye(4)), fs_uniforms=fs_pts,
                      viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0), **stencil)
    points = dict(px=px.astype(np.float32), py=py.astype(np.float32), pz=pz.astype(np.float32),
                  nx=None, ny=None, nz=None, u=None, v=None, indices=None, primitive=1,
                  desc=point_desc)
    return dict(name=f"clip_{width}x{height}", width=width, height=height, clear=0xFF000000,
                texture=(1024, 1024, checkerboard(1024, 32)), draws=[tris, lines, points])


# ---------------------------------------------------------------------------
# Batched mix and serving meshes
# ---------------------------------------------------------------------------
BATCH_RESOLUTIONS = [(256, 256), (512, 512), (1920, 1080)]


def batch_specs(n_scenes: int = 256, lo: float = 1e3, hi: float = 5e5, seed: int = SEED_BATCH):
    """(n_tris, width, height, seed) per scene: counts log-uniform in [lo, hi]."""
    out = []
    for i in range(n_scenes):
        rng = SplitMix64(seed + i)
        u = float(rng.uniform(1)[0])
        n = int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))
        r = BATCH_RESOLUTIONS[int(rng.uniform(1)[0] * 3) % 3]
        out.append((n, r[0], r[1], seed + i))
    return out


def batch_scene(n_tris: int, width: int, height: int, seed: int):
    s = soup_scene(n_tris, width, height, seed=seed)
    s["name"] = f"batch_{seed:x}_{n_tris}_{width}x{height}"
    return s


def serve_mesh_specs(n_meshes: int = 64, lo: float = 1e3, hi: float = 1e6, seed: int = SEED_SERVE):
    out = []
    for j in range(n_meshes):
        rng = SplitMix64(seed + j)
        u = float(rng.uniform(1)[0])
        out.append((int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo))))), seed + j))
    return out


def serve_mesh(n_tris: int, seed: int):
    """A closed-ish object-space blob: soup triangles scattered in a unit ball
    so an orbit camera at distance 2..4 sees all of it."""
    rng = SplitMix64(seed)
    cx, cy, cz = _unit3(rng.normal(n_tris), rng.normal(n_tris), rng.normal(n_tris))
    rad = np.cbrt(rng.uniform(n_tris)) * 0.8
    cx, cy, cz = cx * rad, cy * rad, cz * rad
    size = 0.02 * np.exp(0.5 * rng.normal(n_tris)) * (1e4 / max(n_tris, 1e4)) ** 0.5
    size3 = np.repeat(size, 3)
    ox, oy, oz = rng.normal(3 * n_tris), rng.normal(3 * n_tris), rng.normal(3 * n_tris)
    px = np.repeat(cx, 3) + ox * size3
    py = np.repeat(cy, 3) + oy * size3
    pz = np.repeat(cz, 3) + oz * size3
    nx, ny, nz = _unit3(rng.normal(3 * n_tris), rng.normal(3 * n_tris), rng.normal(3 * n_tris))
    u, v = rng.uniform(3 * n_tris), rng.uniform(3 * n_tris)
    f = np.float32
    return dict(px=px.astype(f), py=py.astype(f), pz=pz.astype(f), nx=nx.astype(f),
                ny=ny.astype(f), nz=nz.astype(f), u=u.astype(f), v=v.astype(f), indices=None,
                primitive=3)


def orbit_camera(rng: SplitMix64, width: int, height: int):
    az = float(rng.uniform(1, 0.0, 2.0 * math.pi)[0])
    el = float(rng.uniform(1, -math.pi / 3.0, math.pi / 3.0)[0])
    dist = float(rng.uniform(1, 2.0, 4.0)[0])
    eye = (dist * math.cos(el) * math.cos(az), dist * math.sin(el), dist * math.cos(el) * math.sin(az))
    view = look_at(eye, (0, 0, 0), (0, 1, 0))
    proj = perspective_zero_to_one(60.0, width / height, 0.1, 100.0)
    return proj @ view


SERVE_RESOLUTIONS = [(512, 512), (1920, 1080)]


def serve_desc(mvp: np.ndarray, width: int, height: int, alpha: float = 0.8) -> dict:
    """Draw state of a served frame: the soup shading (checkerboard x
    Lambert, over-blend at `alpha`), no culling, full-frame viewport."""
    return dict(vertex_program=0, fragment_program=2, blend=1, cull=0,
                vs_uniforms=vs_uniforms(mvp), fs_uniforms=tex_lambert_uniforms(alpha),
                tex_filter=1, tex_edge=0,
                viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))


def serve_requests(n: int, n_meshes: int = 64, seed: int = SEED_SERVE + 0x10000,
                   resolutions=SERVE_RESOLUTIONS):
    """Per request: mesh id U[0, n_meshes), resolution uniform over
    `resolutions`, orbit camera (azimuth U(0, 2pi), elevation U(-60, 60) deg,
    distance U(2, 4))."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        m = int(float(rng.uniform(1)[0]) * n_meshes) % n_meshes
        w, h = resolutions[int(float(rng.uniform(1)[0]) * len(resolutions)) % len(resolutions)]
        out.append(dict(mesh=m, width=w, height=h, desc=serve_desc(orbit_camera(rng, w, h), w, h)))
    return out


def poisson_arrivals(rate: float, duration: float, seed: int) -> np.ndarray:
    """Arrival times in [0, duration) of a Poisson process of `rate` per second."""
    rng = SplitMix64(seed)
    n = int(rate * duration * 1.5) + 64
    gaps = -np.log1p(-rng.uniform(n)) / rate
    t = np.cumsum(gaps)
    return t[t < duration]
