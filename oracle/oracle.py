"""Python loader for the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker (and the timed CPU baseline).
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "rust-softrender_amd"))
from softrender_hip import DrawDesc, VertexSoA, vertex_soa, make_desc  # noqa: E402  (struct layouts only)

LIB_PATH = os.path.join(HERE, "liboracle.so")
F32_MIN = float(np.finfo(np.float32).min)


class OrFb(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("color", C.POINTER(C.c_uint32)),
                ("depth", C.POINTER(C.c_float)), ("stencil", C.POINTER(C.c_uint8)),
                ("n_color", C.c_uint32), ("color_fmt", C.c_int32 * 4), ("planes", C.c_void_p * 4),
                ("depth_fmt", C.c_int32), ("depth64", C.POINTER(C.c_double)), ("stencil_fmt", C.c_int32)]


class OrMesh(C.Structure):
    _fields_ = [("v", VertexSoA), ("n_vertices", C.c_uint32), ("indices", C.POINTER(C.c_uint32)),
                ("n_indices", C.c_uint32), ("primitive", C.c_int32), ("iattr", C.POINTER(C.c_int32))]


class OrTexture(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("rgba", C.POINTER(C.c_uint32)),
                ("rgbaf", C.POINTER(C.c_float))]


STAT_FIELDS = ["tiles", "tri_calls", "tri_culled", "bbox_tests", "strip_tests", "covered",
               "shaded", "strip_shaded", "line_calls", "line_accepted", "line_plots",
               "line_shaded", "point_calls", "point_hits", "point_shaded", "gen_points",
               "gen_lines", "gen_tris", "oob_plots", "invalid"]


class OrStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in STAT_FIELDS]


_lib = None


def lib():
    """The oracle library built in this folder (make -C oracle)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        _lib = load(LIB_PATH)
    return _lib


def load(path: str):
    """Loads an oracle build from `path` and declares its entry points (the
    pinning tests also load variants with planted mistakes this way)."""
    L = C.CDLL(path)
    L.or_draw.argtypes = [C.POINTER(OrFb), C.POINTER(OrMesh), C.POINTER(DrawDesc),
                          C.POINTER(OrTexture), C.c_int, C.POINTER(OrStats)]
    L.or_draw.restype = C.c_int
    L.or_fb_clear.argtypes = [C.POINTER(OrFb), C.c_uint32, C.c_float, C.c_uint8]
    fp = C.POINTER(C.c_float)
    L.or_normalize.argtypes = [fp, C.POINTER(DrawDesc._fields_[16][1]), fp]
    L.or_clip_triangle_reference.argtypes = [fp, fp, fp, C.c_int, fp, C.c_int]
    L.or_clip_triangle_reference.restype = C.c_int
    L.or_clip_triangle_near.argtypes = [fp, fp, fp, C.c_int, fp, C.c_int]
    L.or_clip_triangle_near.restype = C.c_int
    L.or_clip_line_reference.argtypes = [fp, fp, C.c_int, fp]
    L.or_clip_line_reference.restype = C.c_int
    L.or_tiles.argtypes = [C.c_uint32] * 4 + [C.POINTER(C.c_uint32), C.c_int]
    L.or_tiles.restype = C.c_int
    L.or_liang_barsky.argtypes = [C.c_float] * 8 + [fp]
    L.or_liang_barsky.restype = C.c_int
    i64p, f64p = C.POINTER(C.c_int64), C.POINTER(C.c_double)
    L.or_bresenham.argtypes = [C.c_int64] * 4 + [i64p, f64p, C.c_int]
    L.or_bresenham.restype = C.c_int
    L.or_wu.argtypes = [C.c_double] * 4 + [i64p, f64p, C.c_int]
    L.or_wu.restype = C.c_int
    L.or_stencil_test.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
    L.or_stencil_test.restype = C.c_int
    L.or_stencil_op.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32]
    L.or_stencil_op.restype = C.c_uint32
    L.or_sample.argtypes = [C.POINTER(OrTexture), C.c_float, C.c_float, C.c_int, C.c_int,
                            C.c_uint32, fp]
    L.or_blend.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
    L.or_blend.restype = C.c_uint32
    L.or_mul_alpha.argtypes = [C.c_int, C.c_uint32, C.c_double]
    L.or_mul_alpha.restype = C.c_uint32
    L.or_hypotf.argtypes = [C.c_float, C.c_float]
    L.or_hypotf.restype = C.c_float
    L.or_div_f32.argtypes = [C.c_float, C.c_float]
    L.or_div_f32.restype = C.c_float
    L.or_gs_normal_lines.argtypes = [C.c_int] + [fp] * 6 + [C.c_int, fp, fp, fp, C.POINTER(C.c_int)]
    L.or_vertex.argtypes = [C.POINTER(DrawDesc), C.POINTER(OrMesh), C.c_uint32, fp, fp]
    L.or_vertex.restype = C.c_int
    L.or_fragment.argtypes = [C.POINTER(DrawDesc), fp, C.POINTER(OrTexture), fp]
    L.or_fragment.restype = None
    return L


class Framebuffer:
    """Host framebuffer in the reference's clear state (renderbuffer/mod.rs:22-30)."""

    def __init__(self, w, h, clear=0, depth=F32_MIN, stencil=0):
        self.w, self.h = w, h
        n = w * h
        self.color = np.full(n, clear, np.uint32)
        self.depth = np.full(n, depth, np.float32)
        self.stencil = np.full(n, stencil, np.uint8)

    def struct(self):
        return OrFb(self.w, self.h, self.color.ctypes.data_as(C.POINTER(C.c_uint32)),
                    self.depth.ctypes.data_as(C.POINTER(C.c_float)),
                    self.stencil.ctypes.data_as(C.POINTER(C.c_uint8)))

    def copy(self):
        f = Framebuffer(self.w, self.h)
        f.color, f.depth, f.stencil = self.color.copy(), self.depth.copy(), self.stencil.copy()
        return f


class FramebufferEx:
    """Host framebuffer of another format (sr_fb_create_ex): `planes` lists
    each colour plane's format (0 RGBA8, 1 RGBAF32); cleared to `colors`
    (per plane: RGBA8 uint32 or 4 floats), depth far, stencil 0."""

    def __init__(self, w, h, planes, depth=0, stencil=0, colors=None):
        self.w, self.h, self.fmts, self.depth_fmt, self.stencil_fmt = w, h, list(planes), depth, stencil
        n = w * h
        colors = colors or [0] * len(planes)
        self.planes = []
        for f, c in zip(self.fmts, colors):
            if f == 1:
                self.planes.append(np.tile(np.broadcast_to(np.asarray(c, np.float32), (4,)).reshape(1, 4), (n, 1)))
            else:
                self.planes.append(np.full(n, c, np.uint32))
        self.depth = np.full(n, F32_MIN, np.float32)
        self.depth64 = np.full(n, -np.finfo(np.float64).max, np.float64) if depth == 1 else None
        self.stencil = np.zeros(n, np.uint8)

    def struct(self):
        fb = OrFb()
        fb.w, fb.h = self.w, self.h
        fb.depth = self.depth.ctypes.data_as(C.POINTER(C.c_float))
        fb.stencil = self.stencil.ctypes.data_as(C.POINTER(C.c_uint8))
        fb.n_color = len(self.fmts)
        for i, (f, p) in enumerate(zip(self.fmts, self.planes)):
            fb.color_fmt[i] = f
            fb.planes[i] = p.ctypes.data
        fb.depth_fmt = self.depth_fmt
        if self.depth64 is not None:
            fb.depth64 = self.depth64.ctypes.data_as(C.POINTER(C.c_double))
        fb.stencil_fmt = self.stencil_fmt
        return fb


def plane_texture(fb: FramebufferEx, plane: int):
    """A texture reading one plane of an oracle framebuffer (for draw's `tex`)."""
    return ("plane", fb.w, fb.h, fb.planes[plane], fb.fmts[plane])


def _texture(tex):
    if tex is None:
        return None, None
    if tex[0] == "plane":
        _, w, h, data, fmt = tex
        if fmt == 1:
            a = np.ascontiguousarray(data, np.float32)
            return OrTexture(w, h, None, a.ctypes.data_as(C.POINTER(C.c_float))), a
        data = np.ascontiguousarray(data, np.uint32)
        return OrTexture(w, h, data.ctypes.data_as(C.POINTER(C.c_uint32)), None), data
    w, h, rgba = tex
    rgba = np.ascontiguousarray(rgba, np.uint32)
    return OrTexture(w, h, rgba.ctypes.data_as(C.POINTER(C.c_uint32)), None), rgba


def _or_mesh(mesh: dict):
    soa, keep = vertex_soa(mesh)
    idx = mesh.get("indices")
    n = len(keep["px"])
    ia = mesh.get("iattr")
    ia = None if ia is None else np.ascontiguousarray(ia, np.int32)
    iap = None if ia is None else ia.ctypes.data_as(C.POINTER(C.c_int32))
    if idx is not None:
        idx = np.ascontiguousarray(idx, np.uint32)
        om = OrMesh(soa, n, idx.ctypes.data_as(C.POINTER(C.c_uint32)), len(idx), mesh["primitive"], iap)
    else:
        om = OrMesh(soa, n, None, 0, mesh["primitive"], iap)
    return om, (keep, idx, ia)


def vertex(mesh: dict, desc: DrawDesc, i: int, L=None):
    """The vertex program on vertex i: (clip position (4,), varyings (K,))."""
    om, keep = _or_mesh(mesh)
    clip = np.zeros(4, np.float32)
    k = np.zeros(8, np.float32)
    fp = C.POINTER(C.c_float)
    K = (L or lib()).or_vertex(C.byref(desc), C.byref(om), i, clip.ctypes.data_as(fp), k.ctypes.data_as(fp))
    return clip, k[:K].copy()


def fragment(desc: DrawDesc, k, tex=None, L=None):
    """Output 0 of the fragment program on varyings k: RGBA floats (4,)."""
    k = np.ascontiguousarray(np.r_[np.asarray(k, np.float32), np.zeros(8, np.float32)][:8])
    ot, keep = _texture(tex)
    out = np.zeros(4, np.float32)
    fp = C.POINTER(C.c_float)
    (L or lib()).or_fragment(C.byref(desc), k.ctypes.data_as(fp), C.byref(ot) if ot is not None else None,
                             out.ctypes.data_as(fp))
    return out


def sample(tex, u: float, v: float, filter: int, edge: int, border: int = 0, L=None):
    """or_sample: RGBA floats (4,)."""
    ot, keep = _texture(tex)
    out = np.zeros(4, np.float32)
    (L or lib()).or_sample(C.byref(ot), u, v, filter, edge, border, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def draw(fb: Framebuffer, mesh: dict, desc: DrawDesc, tex=None, nthreads: int = 1, L=None):
    """Runs one draw chain; returns (status, stats dict).  `L`: an oracle
    library from load() (default: lib())."""
    om, keep = _or_mesh(mesh)
    ot, keep_t = _texture(tex)
    st = OrStats()
    fbs = fb.struct()
    rc = (L or lib()).or_draw(C.byref(fbs), C.byref(om), C.byref(desc),
                              C.byref(ot) if ot is not None else None, nthreads, C.byref(st))
    return rc, {k: getattr(st, k) for k in STAT_FIELDS}


def render_scene(scene: dict, nthreads: int = 1):
    """Clear + all draws of a scene dict from scenes.py; returns (fb, stats list)."""
    fb = Framebuffer(scene["width"], scene["height"], scene.get("clear", 0))
    draws = scene.get("draws") or [scene]
    stats = []
    for d in draws:
        desc = make_desc(**d["desc"])
        rc, st = draw(fb, d, desc, scene.get("texture"), nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle draw failed: {rc}")
        stats.append(st)
    return fb, stats


def gs_normal_lines(program: int, clip, k, G):
    """Normal visualisation of one triangle: clip (3, 4) and varyings (3, K)
    in, the emitted line vertices (2n, 4) and their varyings (2n, K) out."""
    L = lib()
    clip = np.ascontiguousarray(clip, np.float32).reshape(3, 4)
    k = np.ascontiguousarray(k, np.float32)
    K = k.shape[1]
    g = np.zeros(20, np.float32)
    flat = np.asarray(G, np.float32).ravel()
    g[:flat.size] = flat
    G = g
    out_c = np.zeros((6, 4), np.float32)
    out_k = np.zeros((6, K), np.float32)
    n = C.c_int(0)
    fp = C.POINTER(C.c_float)
    ptr = lambda a: a.ctypes.data_as(fp)  # noqa: E731
    L.or_gs_normal_lines(program, ptr(clip[0]), ptr(k[0]), ptr(clip[1]), ptr(k[1]), ptr(clip[2]), ptr(k[2]),
                         K, ptr(G), ptr(out_c), ptr(out_k), C.byref(n))
    return out_c[:2 * n.value].copy(), out_k[:2 * n.value].copy()
