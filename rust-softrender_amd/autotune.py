"""Autotuner for the batched renderer.

For each resolution class of a workload it renders a small batch of
representative scenes with every execution tile size (8x8 = one wave per
tile, 16/32/64 = four waves per tile), keeps the fastest, and sizes the bin
capacity from the (execution tile, triangle) pairs per triangle it observed.
The result is installed with sr_ctx_set_exec_tile_policy /
sr_ctx_set_batch_capacity and can be saved as JSON.

    python autotune.py --out policy.json
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import scenes  # noqa: E402
import softrender_hip as sr  # noqa: E402

TILES = (8, 16, 32, 64)


def _batch_time(ctx, items, reps):
    ctx.batch_draw(items)  # capture + warm
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.batch_draw(items)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(ctx, resolutions=scenes.BATCH_RESOLUTIONS, tris_per_mpx=(4e5,), scenes_per_class=8,
         tiles=TILES, reps=3, seed=0x5EED7777):
    """Returns {"policy": [(max_pixels, ets)], "pairs_per_tri": f, "timings": {...}}."""
    results, pairs_per_tri = {}, 0.0
    tex = None
    for (w, h) in sorted(resolutions, key=lambda r: r[0] * r[1]):
        specs = []
        for i in range(scenes_per_class):
            n = int(max(1000, tris_per_mpx[0] * w * h / 1e6 * math.exp(0.5 * (i % 3 - 1))))
            specs.append(scenes.soup_scene(n, w, h, seed=seed + 97 * i + w))
        if tex is None:
            tex = ctx.texture(*specs[0]["texture"])
        fbs = [ctx.framebuffer(w, h) for _ in specs]
        meshes = [ctx.mesh(s) for s in specs]
        times = {}
        for ets in tiles:
            descs = [sr.make_desc(**dict(s["desc"], exec_tile=ets)) for s in specs]
            items = [(fb, m, d, tex, 0xFF000000) for fb, m, d in zip(fbs, meshes, descs)]
            times[ets] = _batch_time(ctx, items, reps)
            st = ctx.last_stats()
            if st["triangles_in"]:
                pairs_per_tri = max(pairs_per_tri, st["bin_pairs"] / st["triangles_in"])
        best = min(times, key=times.get)
        results[(w, h)] = dict(best=best, seconds=times)
        for m in meshes:
            m.free()
        for fb in fbs:
            fb.free()
    ordered = sorted(results, key=lambda r: r[0] * r[1])
    policy = []
    for k, r in enumerate(ordered):
        bound = (r[0] * r[1] + ordered[k + 1][0] * ordered[k + 1][1]) // 2 if k + 1 < len(ordered) else 2 ** 63
        policy.append((int(bound), int(results[r]["best"])))
    if tex is not None:
        tex.free()
    return {"policy": policy, "pairs_per_tri": pairs_per_tri,
            "timings": {f"{w}x{h}": {str(e): t for e, t in v["seconds"].items()}
                        for (w, h), v in results.items()}}


def apply(ctx, tuned, expected_tris=None):
    ctx.set_exec_tile_policy(tuned["policy"])
    if expected_tris:
        ctx.set_batch_capacity(int(expected_tris * tuned["pairs_per_tri"] * 1.25) + 65536)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    ctx = sr.Context(0)
    tuned = tune(ctx, reps=a.reps)
    print(json.dumps(tuned))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(tuned, f, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
