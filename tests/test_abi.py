"""The C-ABI library loads on the CPU host and exports every function that
include/softrender_hip.h declares (no compute call is made here)."""
import ctypes
import os
import re

import softrender_hip as sr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(sr_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("sr_ctx_create", "sr_mesh_upload", "sr_texture_upload", "sr_fb_create",
                 "sr_fb_clear", "sr_fb_read", "sr_draw", "sr_batch_draw"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sr.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(sr.EXPORTED) <= declared_functions()
    sr.lib()  # binds every prototype


def test_desc_layout_matches_header():
    # sr_draw_desc: 16 int32 + viewport(6 f32) + 48 + 32 + 20 f32 + exec_tile, exec_flags, exec_waves, 5 reserved
    assert ctypes.sizeof(sr.DrawDesc) == 4 * (16 + 6 + 48 + 32 + 20 + 3 + 5)
    assert ctypes.sizeof(sr.VertexSoA) == 8 * 8
    assert ctypes.sizeof(sr.DrawStats) == 9 * 8
