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

# Algorithmic work of the raster per unit, as the survey of the reference
# fixes it: a pixel test inside a triangle's tile-clamped box (triangle.rs:
# 104-126: edge numerators, two exact divides, w, inside and depth tests) is
# 35 VALU operations; a shaded fragment (interpolation, shading, bilinear
# sample, blend) is 100.
OPS_PER_BBOX_TEST = 35
OPS_PER_SHADED = 100
# The same work counted operation by operation in the reference's
# expressions (a divide or sqrt as one operation), reported beside it.
FLOPS_PER_TEST = 14      # triangle.rs:104-113: x,y (2), edge numerators (8), 2 divides, w (2)
FLOPS_PER_COVERED = 5    # z = ((z1*u)+(z2*v))+(z3*w)
FLOPS_PER_SHADED = 145   # 5 varyings x 5 (interpolation), bilinear sample + tex x Lambert, over-blend

# Raster work units of the default workload (1920x1080, 1M-triangle soup,
# seed 0x5EED0002) as the single-threaded oracle counts them: pixel tests
# inside the tile-clamped boxes, covered pixels, shaded fragments.  The CPU
# leg (multi-threaded, racing at screen-tile seams like the reference's
# thread pool) recounts the two order-independent ones and checks them.
DEFAULT_UNITS = {"bbox_tests": 86093801, "covered": 33448710, "shaded": 6973823}

MI355X_SIMDS = 256 * 4   # 256 CUs x 4 SIMDs; a wave64 VALU instruction issues over 4 cycles
RASTER_KERNEL = "k_raster_tris<"
COPY_KERNEL = "copyBuffer"
CALIB_MIB = 256


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--exec-tile", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched (256 scenes) and serving (Poisson) measurements")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="seconds per offered rate")
    ap.add_argument("--only", choices=["clip", "batch"], default="",
                    help="run one side workload alone and print its JSON (for profiling): the 4K clip "
                         "frame or the 256-scene batch")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (calibrates PMC byte counters)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (raster HBM traffic and VALU issue)")
    ap.add_argument("--counter-probe", action="store_true",
                    help="(run under rocprofv3 by the counter passes) copy calibration, then 3 frames")
    return ap.parse_args()


def copy_calibration(mib):
    """Five device-to-device hipMemcpy of `mib` MiB (a blit copy kernel that
    reads and writes exactly that many bytes), for scaling FETCH_SIZE and
    WRITE_SIZE into bytes."""
    import ctypes as C
    hip = None
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = C.CDLL(name)
            break
        except OSError:
            continue
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 5}), flush=True)


def counter_probe(a):
    """What the counter passes profile: a device-to-device copy of CALIB_MIB
    MiB (the byte counters' calibration), then three frames of the default
    workload."""
    copy_calibration(CALIB_MIB)
    import scenes
    import softrender_hip as sr

    scene = scenes.soup_scene(a.tris, 1920, 1080)
    ctx = sr.Context(0)
    fb = ctx.framebuffer(1920, 1080)
    mesh = ctx.mesh(scene)
    tex = ctx.texture(*scene["texture"])
    kw = dict(scene["desc"])
    if a.exec_tile:
        kw["exec_tile"] = a.exec_tile
    desc = sr.make_desc(**kw)
    for _ in range(3):
        fb.clear(scene["clear"])
        ctx.draw(fb, mesh, desc, tex)
    ctx.finish()
    mesh.free()
    tex.free()
    fb.free()
    ctx.close()
    print(json.dumps({"counter_probe_frames": 3}), flush=True)


def pmc_pass(a, counters, timeout=180):
    """One `rocprofv3 --pmc <counters>` pass over `bench.py --counter-probe`
    in a child process.  Returns {kernel: {counter: mean value per
    dispatch}} for the raster and the calibration copy, or raises."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="sr_pmc_", dir="/tmp")
    cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, "-u", os.path.abspath(__file__), "--counter-probe", "--tris", str(a.tris)]
    if a.exec_tile:
        cmd += ["--exec-tile", str(a.exec_tile)]
    p = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, start_new_session=True)
    try:
        log, _ = p.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        raise RuntimeError(f"rocprofv3 --pmc {' '.join(counters)} timed out")
    sums = {}
    for fn in glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for row in csv.DictReader(f):
                name = row.get("Kernel_Name", "")
                key = RASTER_KERNEL if RASTER_KERNEL in name else (COPY_KERNEL if COPY_KERNEL in name else None)
                if key is None:
                    continue
                c = sums.setdefault(key, {}).setdefault(row["Counter_Name"], {})
                d = row.get("Dispatch_Id") or row.get("Correlation_Id") or str(len(c))
                c[d] = c.get(d, 0.0) + float(row["Counter_Value"])
    shutil.rmtree(out, ignore_errors=True)
    res = {k: {cn: sum(v.values()) / len(v) for cn, v in cs.items()} for k, cs in sums.items()}
    if RASTER_KERNEL not in res:
        raise RuntimeError(f"rocprofv3 --pmc {' '.join(counters)}: no raster dispatch in its output "
                           f"(exit {p.returncode}): {log[-400:]}")
    return res


def counter_passes(a):
    """HBM traffic and VALU issue of the raster kernel, measured in this run:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by what the
    same counter reads for a device-to-device copy of known size, and the
    SQ instruction counts over the kernel's active cycles."""
    calib = CALIB_MIB << 20
    t0 = time.time()
    f = pmc_pass(a, ["FETCH_SIZE"])
    print(f"bench.py: counter pass FETCH_SIZE done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    w = pmc_pass(a, ["WRITE_SIZE"])
    print(f"bench.py: counter pass WRITE_SIZE done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    q = pmc_pass(a, ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_WAVES", "GRBM_GUI_ACTIVE"])
    print(f"bench.py: counter pass SQ done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    f_scale = calib / (f[COPY_KERNEL]["FETCH_SIZE"] * 1024.0)
    w_scale = calib / (w[COPY_KERNEL]["WRITE_SIZE"] * 1024.0)
    rd = f[RASTER_KERNEL]["FETCH_SIZE"] * 1024.0 * f_scale
    wr = w[RASTER_KERNEL]["WRITE_SIZE"] * 1024.0 * w_scale
    sq = q[RASTER_KERNEL]
    cycles = sq["GRBM_GUI_ACTIVE"] / 8.0  # summed over the 8 XCDs
    return {"traffic": int(rd + wr), "traffic_read": int(rd), "traffic_write": int(wr),
            "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
            "valu_insts": int(sq["SQ_INSTS_VALU"]), "salu_insts": int(sq["SQ_INSTS_SALU"]),
            "valu_issue_frac": round(sq["SQ_INSTS_VALU"] * 4.0 / (MI355X_SIMDS * cycles), 4),
            "source": f"rocprofv3 --pmc passes over bench.py --counter-probe in this run; bytes scaled "
                      f"by a {CALIB_MIB} MiB device-to-device copy"}


T_START = time.time()


def note(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"bench.py: {msg} ({time.time() - T_START:.0f} s)", file=sys.stderr, flush=True)


def cpu_request_latency(nthreads, n_requests=6):
    """Per-request latency of the serving mix on the CPU oracle (one frame at
    a time, reference scheduling on `nthreads` threads): the first
    `n_requests` requests of the seeded mix (64-mesh cache, {512^2, 1080p}),
    mean and max milliseconds."""
    import oracle
    import scenes

    specs = scenes.serve_mesh_specs(64)
    reqs = scenes.serve_requests(n_requests)
    ms = []
    tex = (1024, 1024, scenes.checkerboard(1024, 32))
    for r in reqs:
        n, sd = specs[r["mesh"]]
        mesh = scenes.serve_mesh(n, sd)
        fb = oracle.Framebuffer(r["width"], r["height"], 0xFF000000)
        import softrender_hip as sr
        d = sr.make_desc(**r["desc"])
        t0 = time.perf_counter()
        oracle.draw(fb, mesh, d, tex, nthreads)
        ms.append(1e3 * (time.perf_counter() - t0))
    return {"threads": nthreads, "requests": n_requests, "mean_ms": round(float(sum(ms) / len(ms)), 2),
            "max_ms": round(max(ms), 2),
            "mesh_tris": [specs[r["mesh"]][0] for r in reqs]}


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
              "(sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)
    if a.copy_calib:
        copy_calibration(a.copy_calib)
        return
    if a.only == "clip":
        print(json.dumps(clip_frame(a, steps=a.steps)), flush=True)
        return
    if a.only == "batch":
        import batchbench
        print(json.dumps(batchbench.run(256, reps=a.steps)), flush=True)
        return
    if a.counter_probe:
        counter_probe(a)
        return
    # The counter passes run first, in child processes, before this process
    # touches the GPU.
    counters, counter_err = None, None
    if not a.no_counters:
        try:
            counters = counter_passes(a)
        except Exception as e:  # reported in the JSON line, never silently
            counter_err = str(e)[:300]
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

    # Timed region: every step records HIP events between the pipeline
    # stages on the context stream, so the raster kernel's duration is
    # measured on the very launches that are timed.
    for _ in range(a.warmup):
        step(tdesc)
    ctx.finish()
    stages = {}
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(tdesc)
        for k, v in ctx.stage_times().items():
            stages[k] = stages.get(k, 0.0) + v / a.steps
    ctx.finish()
    dt = time.perf_counter() - t0
    note(f"{a.steps} timed frames done")
    ms = 1000.0 * dt / a.steps
    fps = a.steps / dt
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
                   "exec_tile": desc.exec_tile or 8},
        "mtris_per_s": round(fps * a.tris / 1e6, 2),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "bin_pairs": stats.get("bin_pairs"),
    }

    # HBM rooflines of the binning and resolve-side kernels (algorithmic bytes)
    ntris = a.tris
    pairs = stats.get("bin_pairs") or 0
    W, H = 1920, 1080
    setup_bytes = ntris * (36 + 64 + 24 + 4)        # read x,y,z of 3 vertices; write record, binning record, count
    sort_bytes = pairs * 8 * 2 * 2 + pairs * 4 * 2  # 2 radix passes (r+w keys/vals) + hist reads
    if stages.get("setup"):
        out["roofline_setup"] = {"bound": "hbm", "achieved": round(setup_bytes / stages["setup"] / 1e6, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if stages.get("sort"):
        out["roofline_sort"] = {"bound": "hbm", "achieved": round(sort_bytes / stages["sort"] / 1e6, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s"}

    default_workload = a.tris == 1_000_000
    units = dict(DEFAULT_UNITS) if default_workload else None
    if not a.no_cpu:
        # oracle as the CPU baseline: one full frame, reference scheduling
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        n = a.cpu_threads or cpu_cores()
        ofb = oracle.Framebuffer(W, H, scene["clear"])
        t0 = time.perf_counter()
        rc, st = oracle.draw(ofb, scene, desc, scene["texture"], n)
        cdt = time.perf_counter() - t0
        note(f"CPU baseline frame done ({n} threads)")
        out["cpu_baseline"] = {
            "value": round(1.0 / cdt, 4), "unit": "frames/s", "cores": n, "kind": "port",
            "sample": f"1 full frame of the same workload, {n} threads, tile queue "
                      f"(every 128x128 tile visits all {ntris} triangles); {cdt:.2f} s"}
        # the same frame on one thread (the sequential semantics), for scale
        ofb1 = oracle.Framebuffer(W, H, scene["clear"])
        t0 = time.perf_counter()
        oracle.draw(ofb1, scene, desc, scene["texture"], 1)
        c1 = time.perf_counter() - t0
        note("CPU baseline frame done (1 thread)")
        out["cpu_baseline"]["one_thread"] = {"value": round(1.0 / c1, 4), "unit": "frames/s", "cores": 1,
                                             "seconds": round(c1, 2)}
        out["cpu_baseline"]["serve_request_ms"] = cpu_request_latency(n)
        note("CPU serving requests done")
        counted = {"bbox_tests": st["bbox_tests"] - st["strip_tests"], "covered": st["covered"]}
        if units is None:
            units = dict(counted, shaded=st["shaded"])  # seam races make this one approximate
        elif any(units[k] != v for k, v in counted.items()):
            raise SystemExit(f"bench.py: raster units changed: {counted} vs {units}")
    if units and stages.get("raster"):
        t_launch = stages["raster"] * 1e-3  # s, HIP events on the context stream, averaged over the timed steps
        ops = OPS_PER_BBOX_TEST * units["bbox_tests"] + OPS_PER_SHADED * units["shaded"]
        ach = ops / t_launch / 1e12
        fine = (FLOPS_PER_TEST * units["bbox_tests"] + FLOPS_PER_COVERED * units["covered"]
                + FLOPS_PER_SHADED * units["shaded"]) / t_launch / 1e12
        out["raster_units"] = units
        out["roofline"] = {"bound": "valu", "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4),
                           "traffic": counters["traffic"] if counters else None,
                           "kernel": "k_raster_tris", "avg_launch_ms": round(stages["raster"], 4),
                           "ops_per_launch": ops,
                           "ops_count": f"{OPS_PER_BBOX_TEST} per box pixel test + {OPS_PER_SHADED} per shaded fragment",
                           "achieved_expression_count": round(fine, 3),
                           "expression_count": f"{FLOPS_PER_TEST} per test + {FLOPS_PER_COVERED} per covered pixel + "
                                               f"{FLOPS_PER_SHADED} per shaded fragment"}
        if counters:
            r = out["roofline"]
            r["traffic_read"], r["traffic_write"] = counters["traffic_read"], counters["traffic_write"]
            r["traffic_source"] = counters["source"]
            r["hbm_gbs"] = round(counters["traffic"] / t_launch / 1e9, 1)
            r["valu_issue_frac"] = counters["valu_issue_frac"]
            if pairs:
                r["valu_insts_per_pair"] = round(counters["valu_insts"] / pairs, 1)
                r["salu_insts_per_pair"] = round(counters["salu_insts"] / pairs, 1)
        elif counter_err:
            out["roofline"]["traffic_error"] = counter_err
    mesh.free()
    tex.free()
    fb.free()
    ctx.close()
    if not a.no_extras:
        out.update(extras(a))
    print(json.dumps(out), flush=True)


def extras(a):
    """The batched path (256 scenes of uneven size in one launch vs one
    launch per scene) and the request server under Poisson load (p50/p99
    latency, frames/s) against the plain per-request path."""
    import batchbench
    import serve

    out = {"clip_frame": clip_frame(a)}
    note("clip frame done")
    t = time.time()
    b = batchbench.run(256, reps=3)
    note("batch done")
    batch = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in b.items()}
    batch["wall_s"] = round(time.time() - t, 1)
    t = time.time()
    sv = serve.run(rates=(200, 1000, 2000), duration=a.serve_seconds, capacity_n=384)
    keep = ("fps", "p50_ms", "p99_ms", "max_ms", "mean_batch", "completed", "truncated")
    srv = {"meshes": sv["meshes"], "mean_mesh_tris": round(sv["mean_mesh_tris"]),
           "capacity_served_fps": round(sv["capacity_served"]["fps"], 1),
           "capacity_served_runs": [round(v, 1) for v in sv["capacity_served"]["fps_runs"]],
           "capacity_plain_fps": round(sv["capacity_plain"]["fps"], 1),
           "capacity_ratio": round(sv["capacity_ratio"], 2),
           "copy_ceiling_fps": {k: round(v, 1) for k, v in sv["copy_ceiling_fps"].items()},
           "capacity_of_copy_ceiling": round(sv["capacity_of_copy_ceiling"], 3), "rates": {}}
    for r, row in sv["rates"].items():
        srv["rates"][r] = {
            "served": {k: round(v, 2) if isinstance(v, float) else v
                       for k, v in row["served"].items() if k in keep},
            "plain": {k: round(v, 2) if isinstance(v, float) else v
                      for k, v in row["plain"].items() if k in keep}}
    srv["wall_s"] = round(time.time() - t, 1)
    out.update(batch=batch, serve=srv)
    return out


def clip_frame(a, steps=10):
    """The 3840x2160 clipping- and primitive-heavy frame: 500k triangles (30%
    crossing the near plane, clipped by the geometry program), 200k Wu lines,
    100k points, stencil increment; frames/s of clear + the three draws."""
    import scenes
    import softrender_hip as sr

    s = scenes.clip_scene()
    ctx = sr.Context(0)
    fb = ctx.framebuffer(s["width"], s["height"])
    tex = ctx.texture(*s["texture"])
    draws = [(ctx.mesh(d), sr.make_desc(**d["desc"])) for d in s["draws"]]

    def frame():
        fb.clear(s["clear"])
        for m, d in draws:
            ctx.draw(fb, m, d, tex)

    frame()
    ctx.finish()
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    ctx.finish()
    dt = (time.perf_counter() - t0) / steps
    for m, _ in draws:
        m.free()
    tex.free()
    fb.free()
    ctx.close()
    return {"workload": "clip_4k: 3840x2160, 500k tris (30% near-clipped), 200k Wu lines, 100k points, "
                        "stencil increment", "fps": round(1.0 / dt, 2), "ms_per_frame": round(dt * 1e3, 3)}


if __name__ == "__main__":
    main()
