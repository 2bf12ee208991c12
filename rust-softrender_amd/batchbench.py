"""Batched-rendering workload: 256 seeded scenes (triangles log-uniform
1k..500k, resolutions {256^2, 512^2, 1920x1080}, soup shading), rendered
(a) one frame per launch (sr_fb_clear + sr_draw per scene) and
(b) as one batch (sr_batch_draw with fused clears; the graph is captured on
the first call and replayed).  Reports frames/s of both.

    python batchbench.py [--scenes 256] [--reps 5]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import scenes  # noqa: E402
import softrender_hip as sr  # noqa: E402


class BatchWorkload:
    def __init__(self, ctx, n_scenes=256, seed=scenes.SEED_BATCH, lo=1e3, hi=5e5):
        self.ctx = ctx
        self.specs = scenes.batch_specs(n_scenes, lo, hi, seed)
        first = scenes.batch_scene(*self.specs[0])
        self.tex = ctx.texture(*first["texture"])
        self.fbs, self.meshes, self.descs = [], [], []
        self.tris = 0
        self.pixels = 0
        for (n, w, h, sd) in self.specs:
            s = first if sd == self.specs[0][3] else scenes.batch_scene(n, w, h, sd)
            self.fbs.append(ctx.framebuffer(w, h))
            self.meshes.append(ctx.mesh(s))
            self.descs.append(sr.make_desc(**s["desc"]))
            self.tris += n
            self.pixels += w * h
        self.items = [(fb, m, d, self.tex, 0xFF000000) for fb, m, d in zip(self.fbs, self.meshes, self.descs)]

    def plain(self):
        for fb, m, d in zip(self.fbs, self.meshes, self.descs):
            fb.clear(0xFF000000)
            self.ctx.draw(fb, m, d, self.tex)
        self.ctx.finish()

    def batched(self):
        self.ctx.batch_draw(self.items)

    def close(self):
        for m in self.meshes:
            m.free()
        for fb in self.fbs:
            fb.free()
        self.tex.free()


def timed(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return min(ts), float(np.median(ts))


def frame_digest(fb):
    """BLAKE2b of a framebuffer's colour, depth and stencil planes."""
    h = hashlib.blake2b(digest_size=16)
    for plane in fb.read():
        h.update(plane.tobytes())
    return h.hexdigest()


def run(n_scenes=256, reps=5, tune=None):
    """Times both paths and compares every scene's batched framebuffer with
    its plain one (colour, depth and stencil, by digest)."""
    ctx = sr.Context(0)
    if tune:
        import autotune
        autotune.apply(ctx, tune)
    wl = BatchWorkload(ctx, n_scenes)
    if tune:
        ctx.set_batch_capacity(int(wl.tris * tune["pairs_per_tri"] * 1.25) + 65536)
    p_best, p_med = timed(wl.plain, reps)
    plain_frames = [frame_digest(fb) for fb in wl.fbs]
    b_best, b_med = timed(wl.batched, reps)
    differ = [k for k, fb in enumerate(wl.fbs) if frame_digest(fb) != plain_frames[k]]
    out = {
        "scenes": n_scenes, "triangles": wl.tris, "pixels": wl.pixels,
        "plain_s": p_med, "plain_fps": n_scenes / p_med,
        "batch_s": b_med, "batch_fps": n_scenes / b_med,
        "speedup": p_med / b_med, "batch_device_ms": ctx.stage_times().get("vertex"),
        "batch_mtris_s": wl.tris / b_med / 1e6,
        "scenes_checked": n_scenes, "scenes_differing": len(differ),
        "checked_equal_to_plain": not differ,
    }
    if differ:  # a speedup over different images means nothing
        out["speedup"] = None
        out["differing_scenes"] = differ[:16]
    wl.close()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--tune", default="", help="autotuner JSON to apply")
    a = ap.parse_args()
    tune = json.load(open(a.tune)) if a.tune else None
    print(json.dumps(run(a.scenes, a.reps, tune=tune)), flush=True)


if __name__ == "__main__":
    main()
