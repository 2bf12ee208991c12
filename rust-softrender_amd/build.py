"""Builds libsoftrender_hip.so (gfx950) in-tree with hipcc.

Flags:
  --offload-arch=gfx950        MI355X only
  -ffp-contract=off            no a*b+c -> FMA fusion: Rust evaluates every
                               product and sum separately, and so must we
  -packed-fp32-ops disabled    no v_pk_{add,mul,fma}_f32 generation
  (default) correctly rounded f32 division and sqrt

Objects compile in parallel; the raster translation unit is built once per
execution tile size.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libsoftrender_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
    "-Wno-unused-result", "-Wno-unknown-warning-option",
]

UNITS = [("sr_kernels.hip", []), ("sr_runtime.hip", []), ("sr_server.hip", []), ("sr_obj.cpp", []),
         ("sr_raster_lp.hip", []), ("sr_raster_fmt.hip", [])] + [
    ("sr_raster.hip", [f"-DSR_ETS={e}"]) for e in (8, 16, 32)]


def _obj(src, defs):
    tag = "".join(d.split("=")[-1] for d in defs)
    return os.path.join(OBJ, os.path.splitext(src)[0] + (f"_{tag}" if tag else "") + ".o")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    for src, defs in UNITS:
        obj = _obj(src, defs)
        cmd = [HIPCC, *COMMON, *defs, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
    failed = False
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src}\n{out}{err}")
        elif verbose:
            msg = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l
                            and "not a recognized feature" not in l)
            if msg.strip():
                sys.stderr.write(msg + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = OUT + ".tmp"
    objs = [_obj(s, d) for s, d in UNITS]
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
