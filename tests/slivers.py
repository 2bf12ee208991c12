"""Adversarial slivers: triangles for which the reference shades a pixel
OUTSIDE the triangle's own bounding box.

rasterize_triangle clamps the box into every screen tile
(triangle.rs:66-78) and runs the coverage test there (:104-113); for a tile
the box misses, that clamped box is a strip of pixels outside the box, and
f32 rounding of u, v, w can let such a pixel centre pass when a sliver's
sharp end points at it.  Each triangle below is a ~1000-pixel-long,
~0.002-pixel-wide sliver (clip space, w = 1, drawn into 1024x1024 with the
identity transform) ending just up-left of pixel (512, 512), the top-left
pixel of screen tile (4, 4) at tile size 128; the reference tests that pixel
(the box clamped into tile (4, 4)) and it passes the coverage test, although
the box ends at column and row 511.  Found by a random search over such
slivers with the f32 expressions of triangle.rs:64-113; the oracle confirms
each (strip_shaded == 1).
"""
import numpy as np

import scenes

W = H = 1024
PIXEL = (512, 512)

# clip-space x and y of the three vertices, exact float32 values
SLIVERS = [
    (["-0x1.fd311cp-13", "-0x1.45be7ep-1", "-0x1.cb4b3cp-2"],
     ["0x1.d3c726p-13", "0x1.4073eep-1", "0x1.c3d31cp-2"]),
    (["-0x1.ba43eep-13", "-0x1.05a9d0p-1", "-0x1.c10352p-2"],
     ["0x1.80c3c4p-11", "0x1.791d66p-1", "0x1.43979ep-1"]),
    (["-0x1.8731a6p-12", "-0x1.62ca1ap-1", "-0x1.cc8f0cp-2"],
     ["0x1.aa6b1ep-14", "0x1.1b58c2p-1", "0x1.6fbe54p-2"]),
]


def sliver_scene(k, extra=0, seed=5, **desc):
    """Sliver k alone (extra=0) or after `extra` ordinary soup triangles; a
    flat colour, no blending unless `desc` says otherwise."""
    xs, ys = SLIVERS[k]
    px = np.array([float.fromhex(v) for v in xs], np.float32)
    py = np.array([float.fromhex(v) for v in ys], np.float32)
    pz = np.full(3, 0.5, np.float32)
    uv = np.zeros(3, np.float32)
    nrm = np.zeros(3, np.float32)
    if extra:
        s = scenes.soup_scene(extra, W, H, seed=seed)
        px = np.concatenate([s["px"], px]); py = np.concatenate([s["py"], py]); pz = np.concatenate([s["pz"], pz])
        uv = np.concatenate([s["u"], uv]); nrm = np.concatenate([s["nx"], nrm])
    fs = np.zeros(32, np.float32)
    fs[0:4] = (1.0, 0.5, 0.25, 1.0)
    d = dict(vertex_program=0, fragment_program=0, blend=0, cull=0,
             vs_uniforms=scenes.vs_uniforms(np.eye(4)), fs_uniforms=fs,
             viewport=(0.0, 0.0, float(W), float(H), 0.0, 1.0))
    d.update(desc)
    n = len(px)
    return dict(width=W, height=H, clear=0xFF000000, texture=None, primitive=3, px=px, py=py, pz=pz,
                nx=nrm, ny=np.zeros(n, np.float32), nz=np.ones(n, np.float32), u=uv, v=uv.copy(),
                indices=None, desc=d)
