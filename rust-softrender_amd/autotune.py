"""Autotuner for the batched renderer: execution tile size, waves per tile
and bin capacity.

Execution configurations (tile edge, waves sharing a tile): (8, 1), (16, 1),
(16, 2), (16, 4), (32, 2), (32, 4).  For each resolution class of
a workload it renders a small batch of representative scenes (the soup
shading, triangle counts scaled to the class) with every configuration and
keeps the fastest; every configuration must leave the same images (the
execution tiling never changes a result), which it checks.

Bin capacity: the batched pass sizes its (execution tile, triangle) pair
buffers before it knows the pair count and reruns after an overflow.  For
each candidate capacity (pairs per triangle) it times the first batch on a
fresh context, rerun included, and keeps the smallest capacity within 3% of
the fastest.

The result is installed with sr_ctx_set_exec_policy /
sr_ctx_set_batch_capacity and can be saved as JSON.

    python autotune.py --out policy.json
"""
from __future__ import annotations

import argparse
import hashlib
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

CONFIGS = ((8, 1), (16, 1), (16, 2), (16, 4), (32, 2), (32, 4))
CAPACITIES = (1.0, 2.0, 3.0, 4.0, 6.0)  # pairs per triangle


def _digest(fbs):
    h = hashlib.blake2b(digest_size=16)
    for fb in fbs:
        for plane in fb.read():
            h.update(plane.tobytes())
    return h.hexdigest()


def _batch_time(ctx, items, reps):
    ctx.batch_draw(items)  # capture + warm
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.batch_draw(items)
        best = min(best, time.perf_counter() - t0)
    return best


def _class_specs(w, h, n_scenes, tris_per_mpx, seed):
    out = []
    for i in range(n_scenes):
        n = int(max(1000, tris_per_mpx * w * h / 1e6 * math.exp(0.5 * (i % 3 - 1))))
        out.append(scenes.soup_scene(n, w, h, seed=seed + 97 * i + w))
    return out


def tune(ctx, resolutions=scenes.BATCH_RESOLUTIONS, tris_per_mpx=4e5, scenes_per_class=8,
         configs=CONFIGS, capacities=CAPACITIES, reps=3, seed=0x5EED7777):
    """Returns {"policy": [(max_pixels, ets, waves)], "pairs_per_tri": f,
    "timings": {...}, "capacity": {...}, "images_agree": bool}."""
    results, observed = {}, 0.0
    agree = True
    tex = None
    for (w, h) in sorted(resolutions, key=lambda r: r[0] * r[1]):
        specs = _class_specs(w, h, scenes_per_class, tris_per_mpx, seed)
        if tex is None:
            tex = ctx.texture(*specs[0]["texture"])
        fbs = [ctx.framebuffer(w, h) for _ in specs]
        meshes = [ctx.mesh(s) for s in specs]
        times, digests = {}, set()
        for ets, nwv in configs:
            descs = [sr.make_desc(**dict(s["desc"], exec_tile=ets, exec_waves=nwv)) for s in specs]
            items = [(fb, m, d, tex, 0xFF000000) for fb, m, d in zip(fbs, meshes, descs)]
            times[(ets, nwv)] = _batch_time(ctx, items, reps)
            digests.add(_digest(fbs))
            st = ctx.last_stats()
            if st["triangles_in"] and ets == 8:
                observed = max(observed, st["bin_pairs"] / st["triangles_in"])
        agree &= len(digests) == 1
        best = min(times, key=times.get)
        results[(w, h)] = dict(best=best, seconds=times)
        print(f"autotune: {w}x{h} best {best[0]}x{best[0]}/{best[1]}w ({times[best] * 1e3:.2f} ms)",
              file=sys.stderr, flush=True)
        for m in meshes:
            m.free()
        for fb in fbs:
            fb.free()
    ordered = sorted(results, key=lambda r: r[0] * r[1])
    policy = []
    for k, r in enumerate(ordered):
        bound = (r[0] * r[1] + ordered[k + 1][0] * ordered[k + 1][1]) // 2 if k + 1 < len(ordered) else 2 ** 63
        policy.append((int(bound), int(results[r]["best"][0]), int(results[r]["best"][1])))
    if tex is not None:
        tex.free()
    cap = tune_capacity(policy, capacities, resolutions, tris_per_mpx, scenes_per_class, seed) if capacities else {}
    return {"policy": policy, "pairs_per_tri": cap.get("pairs_per_tri", observed * 1.25),
            "observed_pairs_per_tri": observed, "images_agree": agree, "capacity": cap,
            "timings": {f"{w}x{h}": {f"{e}x{e}/{n}w": t for (e, n), t in v["seconds"].items()}
                        for (w, h), v in results.items()}}


def tune_capacity(policy, capacities, resolutions, tris_per_mpx, scenes_per_class, seed):
    """First-batch time (overflow reruns included) per capacity, on a fresh
    context each, with the tuned policy installed."""
    specs = []
    for (w, h) in resolutions:
        specs += _class_specs(w, h, max(1, scenes_per_class // 2), tris_per_mpx, seed + 1)
    ntri = sum(len(s["px"]) // 3 for s in specs)
    first = {}
    for c in capacities:
        ctx = sr.Context(0)
        try:
            apply(ctx, {"policy": policy, "pairs_per_tri": c}, expected_tris=ntri, slack=1.0)
            tex = ctx.texture(*specs[0]["texture"])
            fbs = [ctx.framebuffer(s["width"], s["height"]) for s in specs]
            meshes = [ctx.mesh(s) for s in specs]
            items = [(fb, m, sr.make_desc(**s["desc"]), tex, 0xFF000000) for fb, m, s in zip(fbs, meshes, specs)]
            ctx.finish()
            t0 = time.perf_counter()
            ctx.batch_draw(items)
            first[c] = time.perf_counter() - t0
            for o in meshes + fbs + [tex]:
                o.free()
        finally:
            ctx.close()
    fastest = min(first.values())
    chosen = min(c for c, t in first.items() if t <= 1.03 * fastest)
    return {"pairs_per_tri": chosen, "first_batch_seconds": {str(c): t for c, t in first.items()},
            "triangles": ntri}


def apply(ctx, tuned, expected_tris=None, slack=1.25):
    policy = tuned["policy"]
    ctx.set_exec_policy([(p[0], p[1], p[2] if len(p) > 2 else 0) for p in policy])
    if expected_tris:
        ctx.set_batch_capacity(int(expected_tris * tuned["pairs_per_tri"] * slack) + 65536)


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
