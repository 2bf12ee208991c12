"""ctypes bindings for libsoftrender_hip.so (the C ABI in include/softrender_hip.h).

Plumbing for tests, smoke and bench: every call goes through the C ABI into
the HIP kernels.  There is no CPU fallback — if the shared library is
missing this module raises at import.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsoftrender_hip.so")

# status codes (include/softrender_hip.h)
SR_OK = 0
ERRORS = {
    -1: "SR_ERR_INVALID_ARGUMENT", -2: "SR_ERR_OUT_OF_MEMORY", -3: "SR_ERR_DEVICE",
    -4: "SR_ERR_INVALID_COORDINATE", -5: "SR_ERR_UNSUPPORTED", -6: "SR_ERR_BUDGET",
    -7: "SR_ERR_NOT_READY", -8: "SR_ERR_NO_DEVICE", -9: "SR_ERR_CANCELLED",
    -10: "SR_ERR_WORKSPACE",
}
SR_ERR_WORKSPACE = -10

# framebuffer formats (sr_fb_format)
COLOR_RGBA8, COLOR_RGBAF32 = 0, 1
DEPTH_F32, DEPTH_F64 = 0, 1
STENCIL_U8, STENCIL_NONE = 0, 1
BLEND_REPLACE, BLEND_OVER, BLEND_OVER_NORMALIZED = 0, 1, 2
FP_FLAT, FP_BLINN_PHONG, FP_TEX_LAMBERT, FP_UV_COLOR, FP_GBUFFER = 0, 1, 2, 3, 4
FP_FRESNEL4, FP_INT_PALETTE = 5, 6
VP_MVP_NORMAL_UV, VP_MVP_NORMAL_WORLD, VP_WORLD4_NORMAL4, VP_MVP_INT = 0, 1, 2, 3
FILTER_NEAREST, FILTER_BILINEAR, FILTER_SPAN_NEAREST, FILTER_SPAN_BILINEAR = 0, 1, 2, 3
STATIC_WORKSPACE_BYTES = 64 << 20  # SR_STATIC_WORKSPACE_MIB

EXPORTED = [
    "sr_ctx_create", "sr_ctx_destroy", "sr_last_error", "sr_finish", "sr_ctx_stream",
    "sr_ctx_stage_times", "sr_ctx_last_stats", "sr_mesh_upload", "sr_mesh_free", "sr_mesh_set_int_attrib",
    "sr_texture_upload", "sr_texture_free", "sr_texture_from_fb", "sr_obj_parse", "sr_mesh_load_obj",
    "sr_fb_create", "sr_fb_free", "sr_fb_clear", "sr_fb_create_ex", "sr_fb_get_format", "sr_fb_clear_ex",
    "sr_fb_read_ex", "sr_fb_write_ex", "sr_texture_from_fb_plane",
    "sr_fb_read", "sr_fb_write", "sr_fb_size", "sr_draw", "sr_draw_static", "sr_draw_chains", "sr_batch_draw",
    "sr_ctx_set_exec_tile_policy", "sr_ctx_set_exec_policy", "sr_ctx_set_batch_capacity",
    "sr_batch_submit", "sr_batch_query", "sr_batch_wait",
    "sr_server_create", "sr_server_destroy", "sr_server_submit", "sr_server_poll",
    "sr_server_release", "sr_server_get_stats", "sr_server_last_error", "sr_server_clock",
]


class Viewport(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("width", C.c_float),
                ("height", C.c_float), ("near_z", C.c_float), ("far_z", C.c_float)]


class VertexSoA(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in ("px", "py", "pz", "nx", "ny", "nz", "u", "v")]


class DrawDesc(C.Structure):
    _fields_ = [
        ("vertex_program", C.c_int32), ("geometry_program", C.c_int32),
        ("fragment_program", C.c_int32), ("blend", C.c_int32), ("cull", C.c_int32),
        ("aa_lines", C.c_int32), ("tile_w", C.c_int32), ("tile_h", C.c_int32),
        ("stencil_test", C.c_int32), ("stencil_op", C.c_int32), ("stencil_wrap", C.c_int32),
        ("stencil_ref", C.c_uint32), ("alpha_mode", C.c_int32), ("tex_filter", C.c_int32),
        ("tex_edge", C.c_int32), ("tex_border", C.c_uint32), ("viewport", Viewport),
        ("vs_uniforms", C.c_float * 48), ("fs_uniforms", C.c_float * 32),
        ("gs_uniforms", C.c_float * 20),
        ("exec_tile", C.c_int32), ("exec_flags", C.c_uint32), ("exec_waves", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class DrawStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "triangles_in", "triangles_setup", "lines_in", "points_in", "bin_pairs",
        "generated_tris", "generated_lines", "generated_points", "strip_slivers")]


class BatchItem(C.Structure):
    _fields_ = [("fb", C.c_void_p), ("mesh", C.c_void_p), ("desc", C.POINTER(DrawDesc)),
                ("tex", C.c_void_p), ("clear", C.c_int32), ("clear_rgba", C.c_uint32)]


class FbFormat(C.Structure):
    _fields_ = [("n_color", C.c_uint32), ("color", C.c_int32 * 4), ("depth", C.c_int32),
                ("stencil", C.c_int32)]


class ObjInfo(C.Structure):
    _fields_ = [("models", C.c_uint32), ("vertices", C.c_uint32), ("indices", C.c_uint32),
                ("has_normals", C.c_int32), ("has_uvs", C.c_int32)]


def parse_obj(path: str, model: int = 0) -> dict:
    """sr_obj_parse: the host-side OBJ model as numpy arrays (no device work)."""
    info = ObjInfo()
    rc = lib().sr_obj_parse(os.fsencode(path), model, C.byref(info), None, None, None, None)
    if rc != SR_OK:
        raise SoftrenderError(rc, f"sr_obj_parse({path})")
    nv, ni = info.vertices, info.indices
    pos = np.empty(3 * nv, np.float32)
    nrm = np.empty(3 * nv if info.has_normals else 0, np.float32)
    uv = np.empty(2 * nv if info.has_uvs else 0, np.float32)
    idx = np.empty(ni, np.uint32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float)) if a.size else None
    rc = lib().sr_obj_parse(os.fsencode(path), model, C.byref(info), fp(pos), fp(nrm), fp(uv),
                            idx.ctypes.data_as(C.POINTER(C.c_uint32)) if ni else None)
    if rc != SR_OK:
        raise SoftrenderError(rc, f"sr_obj_parse({path})")
    return dict(models=info.models, positions=pos.reshape(-1, 3), normals=nrm.reshape(-1, 3) if nrm.size else None,
                uvs=uv.reshape(-1, 2) if uv.size else None, indices=idx)


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32),
                ("max_batch_frames", C.c_uint32), ("max_batch_tris", C.c_uint64),
                ("pairs_per_tri", C.c_float), ("reserved", C.c_int32 * 7)]


class Request(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("tex", C.c_void_p), ("desc", DrawDesc),
                ("width", C.c_uint32), ("height", C.c_uint32), ("clear_rgba", C.c_uint32),
                ("want_depth", C.c_int32), ("tag", C.c_uint64), ("arrival", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("tag", C.c_uint64), ("status", C.c_int32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("batch_frames", C.c_uint32),
                ("batch_id", C.c_uint64), ("rgba", C.POINTER(C.c_uint32)),
                ("depth", C.POINTER(C.c_float)), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "submitted", "completed", "failed", "batches", "overflow_reruns", "budget_bytes",
        "in_flight_bytes", "peak_in_flight_bytes", "device_bytes")] + [
        (n, C.c_uint32) for n in ("pending", "inflight_batches", "max_batch_frames_seen",
                                  "reserved")]


def make_desc(**kw) -> DrawDesc:
    """DrawDesc with the reference defaults: tile 128x128 (fragment.rs:29),
    no cull, `()` blend, stencil Always/Keep (stencil.rs:169-175)."""
    d = DrawDesc()
    d.tile_w = kw.pop("tile_w", 128)
    d.tile_h = kw.pop("tile_h", 128)
    vp = kw.pop("viewport", (0.0, 0.0, 1.0, 1.0, 0.0, 1.0))
    d.viewport = Viewport(*[float(v) for v in vp])
    vsu = kw.pop("vs_uniforms", None)
    if vsu is not None:
        d.vs_uniforms[:] = [float(v) for v in np.asarray(vsu, np.float32).reshape(48)]
    fsu = kw.pop("fs_uniforms", None)
    if fsu is not None:
        d.fs_uniforms[:] = [float(v) for v in np.asarray(fsu, np.float32).reshape(32)]
    gsu = kw.pop("gs_uniforms", None)
    if gsu is not None:
        g = np.zeros(20, np.float32)
        flat = np.asarray(gsu, np.float32).ravel()
        g[:flat.size] = flat
        d.gs_uniforms[:] = [float(v) for v in g]
    for k, v in kw.items():
        if not hasattr(d, k):
            raise KeyError(k)
        setattr(d, k, v)
    return d


def f32p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))


def vertex_soa(s: dict):
    keep = {k: (None if s.get(k) is None else np.ascontiguousarray(s[k], np.float32))
            for k in ("px", "py", "pz", "nx", "ny", "nz", "u", "v")}
    soa = VertexSoA(*[f32p(keep[k]) for k in ("px", "py", "pz", "nx", "ny", "nz", "u", "v")])
    return soa, keep


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.sr_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.sr_ctx_destroy.argtypes = [vp]
        L.sr_last_error.argtypes = [vp]
        L.sr_last_error.restype = C.c_char_p
        L.sr_finish.argtypes = [vp]
        L.sr_ctx_stream.argtypes = [vp]
        L.sr_ctx_stream.restype = vp
        L.sr_ctx_stage_times.argtypes = [vp, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]
        L.sr_ctx_last_stats.argtypes = [vp, C.POINTER(DrawStats)]
        L.sr_mesh_upload.argtypes = [vp, C.POINTER(VertexSoA), C.c_uint32, C.POINTER(C.c_uint32),
                                     C.c_uint32, C.c_int32, C.POINTER(vp)]
        L.sr_mesh_free.argtypes = [vp]
        L.sr_mesh_set_int_attrib.argtypes = [vp, C.POINTER(C.c_int32), C.c_uint32]
        L.sr_texture_upload.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                        C.POINTER(vp)]
        L.sr_texture_free.argtypes = [vp]
        L.sr_texture_from_fb.argtypes = [vp, C.POINTER(vp)]
        L.sr_obj_parse.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(ObjInfo), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint32)]
        L.sr_mesh_load_obj.argtypes = [vp, C.c_char_p, C.c_uint32, C.POINTER(vp)]
        L.sr_fb_create.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]
        L.sr_fb_free.argtypes = [vp]
        L.sr_fb_clear.argtypes = [vp, C.c_uint32, C.c_float, C.c_uint8]
        L.sr_fb_read.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                                 C.POINTER(C.c_uint8)]
        L.sr_fb_write.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                                  C.POINTER(C.c_uint8)]
        L.sr_fb_size.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.sr_fb_create_ex.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(FbFormat), C.POINTER(vp)]
        L.sr_fb_get_format.argtypes = [vp, C.POINTER(FbFormat)]
        L.sr_fb_clear_ex.argtypes = [vp, vp]
        L.sr_fb_read_ex.argtypes = [vp, C.POINTER(vp), vp, C.POINTER(C.c_uint8)]
        L.sr_fb_write_ex.argtypes = [vp, C.POINTER(vp), vp, C.POINTER(C.c_uint8)]
        L.sr_texture_from_fb_plane.argtypes = [vp, C.c_uint32, C.POINTER(vp)]
        L.sr_draw.argtypes = [vp, vp, vp, C.POINTER(DrawDesc), vp]
        L.sr_draw_static.argtypes = [vp, vp, vp, C.POINTER(DrawDesc), vp]
        L.sr_draw_chains.argtypes = [vp, vp, vp, C.POINTER(DrawDesc), C.c_uint32, vp]
        L.sr_batch_draw.argtypes = [vp, C.POINTER(BatchItem), C.c_uint32]
        L.sr_ctx_set_exec_tile_policy.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint64),
                                                  C.POINTER(C.c_int32)]
        L.sr_ctx_set_batch_capacity.argtypes = [vp, C.c_uint32]
        L.sr_ctx_set_exec_policy.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32)]
        L.sr_batch_submit.argtypes = [vp, C.POINTER(BatchItem), C.c_uint32, C.POINTER(C.c_uint64)]
        L.sr_batch_query.argtypes = [vp, C.c_uint64]
        L.sr_batch_wait.argtypes = [vp, C.c_uint64]
        L.sr_server_create.argtypes = [vp, C.POINTER(ServerConfig), C.POINTER(vp)]
        L.sr_server_destroy.argtypes = [vp]
        L.sr_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.sr_server_poll.argtypes = [vp, C.POINTER(Completion), C.c_uint32, C.POINTER(C.c_uint32),
                                     C.c_int32]
        L.sr_server_release.argtypes = [vp, C.c_uint64]
        L.sr_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.sr_server_last_error.argtypes = [vp]
        L.sr_server_last_error.restype = C.c_char_p
        L.sr_server_clock.argtypes = []
        L.sr_server_clock.restype = C.c_double
        for name in EXPORTED:
            f = getattr(L, name)
            if f.restype is None and name not in ("sr_last_error", "sr_ctx_stream"):
                f.restype = C.c_int
        _lib = L
    return _lib


class SoftrenderError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Context:
    """Owns a HIP stream and device workspace (the Pipeline's role)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        self._check(lib().sr_ctx_create(device, C.byref(self.h)))

    def _check(self, rc):
        if rc != SR_OK:
            msg = lib().sr_last_error(self.h) if self.h else b""
            raise SoftrenderError(rc, (msg or b"").decode())
        return rc

    def close(self):
        if self.h:
            lib().sr_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def finish(self):
        self._check(lib().sr_finish(self.h))

    def stream(self) -> int:
        return lib().sr_ctx_stream(self.h) or 0

    def stage_times(self):
        buf = (C.c_float * 16)()
        n = C.c_int()
        self._check(lib().sr_ctx_stage_times(self.h, buf, 16, C.byref(n)))
        names = ["vertex", "geometry", "setup", "scan", "emit", "sort", "ranges", "strips",
                 "raster"]
        return {names[i]: buf[i] for i in range(min(n.value, len(names)))}

    def last_stats(self) -> dict:
        st = DrawStats()
        self._check(lib().sr_ctx_last_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in DrawStats._fields_}

    def mesh(self, s: dict) -> "Mesh":
        return Mesh(self, s)

    def mesh_from_obj(self, path: str, model: int = 0) -> "Mesh":
        return Mesh(self, obj=path, model=model)

    def texture(self, w, h, rgba) -> "Texture":
        return Texture(self, w, h, rgba)

    def framebuffer(self, w, h) -> "Framebuffer":
        return Framebuffer(self, w, h)

    def framebuffer_ex(self, w, h, planes, depth=DEPTH_F32, stencil=STENCIL_U8) -> "Framebuffer":
        """A framebuffer / texture buffer of other formats (sr_fb_create_ex):
        `planes` lists each colour plane's COLOR_* format."""
        return Framebuffer(self, w, h, planes=planes, depth=depth, stencil=stencil)

    def texture_from_fb(self, fb: "Framebuffer") -> "Texture":
        return Texture(self, fb.w, fb.h_, fb=fb)

    def texture_from_fb_plane(self, fb: "Framebuffer", plane: int) -> "Texture":
        return Texture(self, fb.w, fb.h_, fb=fb, plane=plane)

    def draw(self, fb: "Framebuffer", mesh: "Mesh", desc: DrawDesc, tex: "Texture" = None):
        self._check(lib().sr_draw(self.h, fb.h, mesh.h, C.byref(desc), tex.h if tex else None))

    def draw_static(self, fb: "Framebuffer", mesh: "Mesh", desc: DrawDesc, tex: "Texture" = None):
        """sr_draw_static: the same draw with its workspace in the library's
        fixed per-GPU region (no device allocation)."""
        self._check(lib().sr_draw_static(self.h, fb.h, mesh.h, C.byref(desc), tex.h if tex else None))

    def draw_chains(self, fb: "Framebuffer", mesh: "Mesh", descs, tex: "Texture" = None):
        """Several chains over one mesh sharing one vertex pass (duplicate())."""
        arr = (DrawDesc * len(descs))(*descs)
        self._check(lib().sr_draw_chains(self.h, fb.h, mesh.h, arr, len(descs), tex.h if tex else None))

    @staticmethod
    def _batch_items(items):
        arr = (BatchItem * len(items))()
        keep = []
        for i, it in enumerate(items):
            fb, mesh, desc, tex = it[:4]
            clear = it[4] if len(it) > 4 else None
            keep.append(desc)
            arr[i] = BatchItem(fb.h, mesh.h, C.pointer(desc), tex.h if tex else None,
                               0 if clear is None else 1, 0 if clear is None else clear)
        return arr, keep

    def batch_draw(self, items):
        """items: list of (fb, mesh, desc, tex[, clear_rgba or None])."""
        arr, keep = self._batch_items(items)
        self._check(lib().sr_batch_draw(self.h, arr, len(items)))

    def batch_submit(self, items) -> int:
        """sr_batch_submit: queues the batch and returns its ticket."""
        arr, keep = self._batch_items(items)
        t = C.c_uint64()
        self._check(lib().sr_batch_submit(self.h, arr, len(items), C.byref(t)))
        return t.value

    def batch_ready(self, ticket: int) -> bool:
        rc = lib().sr_batch_query(self.h, ticket)
        if rc == -7:  # SR_ERR_NOT_READY
            return False
        self._check(rc)
        return True

    def batch_wait(self, ticket: int):
        self._check(lib().sr_batch_wait(self.h, ticket))

    def set_exec_tile_policy(self, entries):
        """entries: [(max_pixels, exec_tile), ...] (first match wins)."""
        mp = (C.c_uint64 * len(entries))(*[int(e[0]) for e in entries])
        et = (C.c_int32 * len(entries))(*[int(e[1]) for e in entries])
        self._check(lib().sr_ctx_set_exec_tile_policy(self.h, len(entries), mp, et))

    def set_exec_policy(self, entries):
        """entries: [(max_pixels, exec_tile, exec_waves), ...] (first match wins)."""
        mp = (C.c_uint64 * len(entries))(*[int(e[0]) for e in entries])
        et = (C.c_int32 * len(entries))(*[int(e[1]) for e in entries])
        ew = (C.c_int32 * len(entries))(*[int(e[2]) for e in entries])
        self._check(lib().sr_ctx_set_exec_policy(self.h, len(entries), mp, et, ew))

    def set_batch_capacity(self, pairs: int):
        self._check(lib().sr_ctx_set_batch_capacity(self.h, pairs))


class Server:
    """Continuous-batching request server over a context's resident meshes.

    submit() queues a request and returns its ticket; poll() returns
    completions as dicts whose "rgba" (and "depth") are flat numpy copies of
    the pinned result buffers (the buffers go back to the pool right away).
    """

    def __init__(self, ctx: Context, hbm_budget=0, max_inflight=0, max_batch_frames=0,
                 max_batch_tris=0, pairs_per_tri=0.0):
        self.ctx = ctx
        self.h = C.c_void_p()
        cfg = ServerConfig(hbm_budget, max_inflight, max_batch_frames, max_batch_tris,
                           pairs_per_tri)
        rc = lib().sr_server_create(ctx.h, C.byref(cfg), C.byref(self.h))
        if rc != SR_OK:
            ctx._check(rc)

    def _check(self, rc):
        if rc != SR_OK:
            raise SoftrenderError(rc, (lib().sr_server_last_error(self.h) or b"").decode())
        return rc

    @staticmethod
    def clock() -> float:
        return lib().sr_server_clock()

    def submit(self, mesh: "Mesh", desc: DrawDesc, width: int, height: int, tex: "Texture" = None,
               clear_rgba: int = 0xFF000000, want_depth: bool = False, tag: int = 0,
               arrival: float = 0.0) -> int:
        r = Request(mesh.h, tex.h if tex else None, desc, width, height, clear_rgba,
                    1 if want_depth else 0, tag, arrival)
        t = C.c_uint64()
        self._check(lib().sr_server_submit(self.h, C.byref(r), C.byref(t)))
        return t.value

    def poll(self, max_n: int = 64, timeout_ms: int = 0, copy: bool = True):
        buf = (Completion * max_n)()
        n = C.c_uint32()
        self._check(lib().sr_server_poll(self.h, buf, max_n, C.byref(n), timeout_ms))
        out = []
        for i in range(n.value):
            c = buf[i]
            d = dict(ticket=c.ticket, tag=c.tag, status=c.status, width=c.width, height=c.height,
                     batch_frames=c.batch_frames, batch_id=c.batch_id, t_arrival=c.t_arrival,
                     t_admit=c.t_admit, t_done=c.t_done)
            if copy and c.status == SR_OK:
                px = c.width * c.height
                d["rgba"] = np.ctypeslib.as_array(c.rgba, shape=(px,)).copy()
                if c.depth:
                    d["depth"] = np.ctypeslib.as_array(c.depth, shape=(px,)).copy()
            self._check(lib().sr_server_release(self.h, c.ticket))
            out.append(d)
        return out

    def stats(self) -> dict:
        st = ServerStats()
        self._check(lib().sr_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_ if k != "reserved"}

    def close(self):
        if self.h:
            lib().sr_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    def __init__(self, ctx: Context, s: dict = None, obj: str = None, model: int = 0):
        """Uploads a mesh dict (px..v streams, indices, primitive, and an
        optional per-vertex int32 "iattr") or, with obj=path, model `model`
        of a Wavefront OBJ file."""
        self.ctx = ctx
        self.h = C.c_void_p()
        if obj is not None:
            ctx._check(lib().sr_mesh_load_obj(ctx.h, os.fsencode(obj), model, C.byref(self.h)))
            return
        soa, keep = vertex_soa(s)
        n = len(keep["px"])
        idx = s.get("indices")
        if idx is not None:
            idx = np.ascontiguousarray(idx, np.uint32)
            ip, ni = idx.ctypes.data_as(C.POINTER(C.c_uint32)), len(idx)
        else:
            ip, ni = None, 0
        ctx._check(lib().sr_mesh_upload(ctx.h, C.byref(soa), n, ip, ni, s["primitive"],
                                        C.byref(self.h)))
        if s.get("iattr") is not None:
            ia = np.ascontiguousarray(s["iattr"], np.int32)
            try:
                ctx._check(lib().sr_mesh_set_int_attrib(self.h, ia.ctypes.data_as(C.POINTER(C.c_int32)), len(ia)))
            except Exception:
                self.free()
                raise

    def free(self):
        if self.h:
            lib().sr_mesh_free(self.h)
            self.h = C.c_void_p()


class Texture:
    def __init__(self, ctx: Context, w: int, h: int, rgba=None, fb: "Framebuffer" = None, plane=None):
        """An uploaded RGBA8 image, or (fb=...) a view of a framebuffer's
        colour plane (plane 0, or `plane`) for render-to-texture."""
        self.h = C.c_void_p()
        if fb is not None:
            if plane is None:
                ctx._check(lib().sr_texture_from_fb(fb.h, C.byref(self.h)))
            else:
                ctx._check(lib().sr_texture_from_fb_plane(fb.h, plane, C.byref(self.h)))
            return
        a = np.ascontiguousarray(rgba, np.uint32)
        ctx._check(lib().sr_texture_upload(ctx.h, w, h, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           C.byref(self.h)))

    def free(self):
        if self.h:
            lib().sr_texture_free(self.h)
            self.h = C.c_void_p()


class Framebuffer:
    def __init__(self, ctx: Context, w: int, h: int, planes=None, depth=DEPTH_F32, stencil=STENCIL_U8):
        self.ctx, self.w, self.h_ = ctx, w, h
        self.h = C.c_void_p()
        if planes is None:
            self.planes, self.depth_fmt, self.stencil_fmt = [COLOR_RGBA8], DEPTH_F32, STENCIL_U8
            ctx._check(lib().sr_fb_create(ctx.h, w, h, C.byref(self.h)))
            return
        self.planes, self.depth_fmt, self.stencil_fmt = list(planes), depth, stencil
        col = (C.c_int32 * 4)(*(list(planes) + [0] * (4 - len(planes))))
        fmt = FbFormat(len(planes), col, depth, stencil)
        ctx._check(lib().sr_fb_create_ex(ctx.h, w, h, C.byref(fmt), C.byref(self.h)))

    def format(self) -> dict:
        f = FbFormat()
        self.ctx._check(lib().sr_fb_get_format(self.h, C.byref(f)))
        return dict(planes=list(f.color)[:f.n_color], depth=f.depth, stencil=f.stencil)

    def clear_ex(self, colors):
        """Framebuffer::clear: colors[i] is plane i's colour (an RGBA8 uint32,
        or 4 floats for an RGBAF32 plane); depth far, stencil 0."""
        buf = b""
        for fmt, c in zip(self.planes, colors):
            buf += (np.broadcast_to(np.asarray(c, np.float32), (4,)).tobytes() if fmt == COLOR_RGBAF32
                    else np.uint32(c).tobytes())
        b = C.create_string_buffer(buf, len(buf))
        self.ctx._check(lib().sr_fb_clear_ex(self.h, C.cast(b, C.c_void_p)))

    def read_ex(self):
        """(planes, depth, stencil): RGBA8 planes as (n,) uint32, RGBAF32
        planes as (n, 4) float32, depth float32 or float64, stencil uint8 or
        None."""
        n = self.w * self.h_
        planes = [np.empty((n, 4), np.float32) if f == COLOR_RGBAF32 else np.empty(n, np.uint32)
                  for f in self.planes]
        ptrs = (C.c_void_p * len(planes))(*[p.ctypes.data for p in planes])
        dep = np.empty(n, np.float64 if self.depth_fmt == DEPTH_F64 else np.float32)
        st = np.empty(n, np.uint8) if self.stencil_fmt == STENCIL_U8 else None
        self.ctx._check(lib().sr_fb_read_ex(self.h, ptrs, C.c_void_p(dep.ctypes.data),
                                            st.ctypes.data_as(C.POINTER(C.c_uint8)) if st is not None else None))
        return planes, dep, st

    def write_ex(self, planes, depth=None, stencil=None):
        arrs = [None if p is None else np.ascontiguousarray(
            p, np.float32 if f == COLOR_RGBAF32 else np.uint32) for p, f in zip(planes, self.planes)]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data if a is not None else None for a in arrs])
        dt = np.float64 if self.depth_fmt == DEPTH_F64 else np.float32
        d = None if depth is None else np.ascontiguousarray(depth, dt)
        s = None if stencil is None else np.ascontiguousarray(stencil, np.uint8)
        self.ctx._check(lib().sr_fb_write_ex(self.h, ptrs, C.c_void_p(d.ctypes.data) if d is not None else None,
                                             s.ctypes.data_as(C.POINTER(C.c_uint8)) if s is not None else None))

    def clear(self, rgba: int, depth: float = float(np.finfo(np.float32).min), stencil: int = 0):
        self.ctx._check(lib().sr_fb_clear(self.h, rgba, depth, stencil))

    def read(self):
        n = self.w * self.h_
        col = np.empty(n, np.uint32)
        dep = np.empty(n, np.float32)
        st = np.empty(n, np.uint8)
        self.ctx._check(lib().sr_fb_read(self.h, col.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         dep.ctypes.data_as(C.POINTER(C.c_float)),
                                         st.ctypes.data_as(C.POINTER(C.c_uint8))))
        return col, dep, st

    def read_color_into(self, ptr):
        """Copies the colour plane to host memory at `ptr` (w*h uint32)."""
        self.ctx._check(lib().sr_fb_read(self.h, C.cast(ptr, C.POINTER(C.c_uint32)), None, None))

    def write(self, col, dep, st):
        col = np.ascontiguousarray(col, np.uint32)
        dep = np.ascontiguousarray(dep, np.float32)
        st = np.ascontiguousarray(st, np.uint8)
        self.ctx._check(lib().sr_fb_write(self.h, col.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          dep.ctypes.data_as(C.POINTER(C.c_float)),
                                          st.ctypes.data_as(C.POINTER(C.c_uint8))))

    def free(self):
        if self.h:
            lib().sr_fb_free(self.h)
            self.h = C.c_void_p()
