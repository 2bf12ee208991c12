"""The C++ builder layer (rust-softrender_amd/include/softrender.hpp) keeps the
reference's Pipeline -> VertexShader -> FragmentShader chain over the C ABI."""
import os
import struct
import subprocess

import numpy as np
import pytest

import scenes
from parity import render_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "builder_sphere.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "builder_sphere")


def test_header_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_builder_chain_matches_oracle(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/builder_sphere is built by __graft_entry__.build()"
    s = scenes.sphere_scene(160, 120)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<4I", s["width"], s["height"], len(s["px"]), len(s["indices"])))
        for k in ("px", "py", "pz", "nx", "ny", "nz"):
            f.write(np.asarray(s[k], np.float32).tobytes())
        f.write(np.asarray(s["indices"], np.uint32).tobytes())
        f.write(np.asarray(s["desc"]["vs_uniforms"], np.float32).tobytes())
        f.write(np.asarray(s["desc"]["fs_uniforms"], np.float32).tobytes())
        f.write(struct.pack("<i", s["desc"]["cull"]))
    out = tmp_path / "out.bin"
    r = subprocess.run([BIN, str(out), str(inp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint32)
    n = s["width"] * s["height"]
    ref, _ = render_ref(s)
    assert np.array_equal(raw[:n], ref.color)
    assert np.array_equal(raw[n:2 * n], ref.depth.view(np.uint32))
