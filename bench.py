"""Benchmark: one frame of the 1920x1080, 1M-triangle soup (clear + draw)
per step, meshes and texture resident in HBM, on one MI355X.

Prints ONE JSON line (rank 0):
  metric/unit      frames/s of the plain draw path (BASELINE.json's metric)
  roofline         the dominant kernel (raster) against MI355X's FP32 vector
                   peak, plus HBM rooflines of the binning and resolve kernels
  cpu_baseline     the CPU oracle in the reference's scheduling (atomic
                   tile queue, every tile visits every triangle) on this
                   host's cores, one full frame
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rust-softrender_amd"))

METRIC = "frames/s and Mtris/s at 1080p/1M tris; % HBM roofline; p50/p99 serve latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

# Algorithmic work per unit (IEEE f32 operations in the reference's
# expressions; a divide or sqrt counts as one operation).
FLOPS_PER_TEST = 14      # triangle.rs:104-113: x,y (2), edge numerators (8), 2 divides, w (2)
FLOPS_PER_COVERED = 5    # z = ((z1*u)+(z2*v))+(z3*w)
FLOPS_PER_SHADED = 145   # 5 varyings x 5, tex+Lambert program, over-blend


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--exec-tile", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, cap) if cap > 0 else n


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import scenes
    import softrender_hip as sr

    scene = scenes.soup_scene(a.tris, 1920, 1080)
    ctx = sr.Context(0)
    fb = ctx.framebuffer(1920, 1080)
    mesh = ctx.mesh(scene)
    tw, th, rgba = scene["texture"]
    tex = ctx.texture(tw, th, rgba)
    kw = dict(scene["desc"])
    if a.exec_tile:
        kw["exec_tile"] = a.exec_tile
    desc = sr.make_desc(**kw)
    tdesc = sr.make_desc(**dict(kw, exec_flags=2))  # SR_EXEC_TIMING

    def step(d):
        fb.clear(scene["clear"])
        ctx.draw(fb, mesh, d, tex)

    for _ in range(a.warmup):
        step(desc)
    ctx.finish()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(desc)
    ctx.finish()
    dt = time.perf_counter() - t0
    ms = 1000.0 * dt / a.steps
    fps = a.steps / dt

    # per-stage device times (HIP events on the context stream)
    stages = {}
    n_t = max(3, min(10, a.steps))
    for _ in range(n_t):
        step(tdesc)
        for k, v in ctx.stage_times().items():
            stages[k] = stages.get(k, 0.0) + v / n_t
    stats = ctx.last_stats()

    out = {
        "metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 soup, 1024^2 checkerboard texture)",
        "config": {"workload": f"soup_1080p_{a.tris // 1000}k: 1920x1080, {a.tris} tris, "
                               "log-normal area (median 20 px), bilinear checkerboard, "
                               "20% over-blend, screen tile 128",
                   "resolution": [1920, 1080], "triangles": a.tris,
                   "exec_tile": desc.exec_tile or 16},
        "mtris_per_s": round(fps * a.tris / 1e6, 2),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "bin_pairs": stats.get("bin_pairs"),
    }

    # HBM rooflines of the binning and resolve-side kernels (algorithmic bytes)
    ntris = a.tris
    pairs = stats.get("bin_pairs") or 0
    W, H = 1920, 1080
    setup_bytes = ntris * (24 + 12 + 80 + 4)        # read x,y,z per vertex; write record + count
    sort_bytes = pairs * 8 * 2 * 2 + pairs * 4 * 2  # 2 radix passes (r+w keys/vals) + hist reads
    if stages.get("setup"):
        out["roofline_setup"] = {"bound": "hbm", "achieved": round(setup_bytes / stages["setup"] / 1e6, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if stages.get("sort"):
        out["roofline_sort"] = {"bound": "hbm", "achieved": round(sort_bytes / stages["sort"] / 1e6, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s"}

    cpu = None
    if not a.no_cpu:
        # oracle as the CPU baseline: one full frame, reference scheduling
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        n = a.cpu_threads or cpu_cores()
        ofb = oracle.Framebuffer(W, H, scene["clear"])
        t0 = time.perf_counter()
        rc, st = oracle.draw(ofb, scene, desc, scene["texture"], n)
        cdt = time.perf_counter() - t0
        cpu = {"value": round(1.0 / cdt, 4), "unit": "frames/s", "cores": n, "kind": "port",
               "sample": f"1 full frame of the same workload, {n} threads, tile queue "
                         f"(every 128x128 tile visits all {ntris} triangles); {cdt:.2f} s"}
        out["cpu_baseline"] = cpu
        inbox = st["bbox_tests"] - st["strip_tests"]
        flops = FLOPS_PER_TEST * inbox + FLOPS_PER_COVERED * st["covered"] + FLOPS_PER_SHADED * st["shaded"]
        out["raster_units"] = {"bbox_tests": inbox, "covered": st["covered"], "shaded": st["shaded"]}
        if stages.get("raster"):
            ach = flops / (stages["raster"] * 1e-3) / 1e12
            out["roofline"] = {"bound": "valu", "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4),
                               "traffic": None, "kernel": "k_raster_tris"}
    print(json.dumps(out), flush=True)
    mesh.free()
    tex.free()
    fb.free()
    ctx.close()


if __name__ == "__main__":
    main()
