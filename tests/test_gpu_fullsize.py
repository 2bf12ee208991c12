"""The batching and serving workloads at their full sizes.

Batch: the 256 seeded scenes (1k..500k triangles, {256^2, 512^2, 1920x1080})
in one sr_batch_draw; every one of the 256 framebuffers equals sr_fb_clear +
sr_draw of the same scene, bit for bit, and three scenes (one 1080p scene of
at least 100k triangles) equal the oracle.

Serving: the 64-mesh cache (1k..1M triangles) at {512^2, 1920x1080}, all
requests queued at once so they share batches; every completion equals the
same request drawn alone, and one 1080p request of a mesh with at least
500k triangles equals the oracle."""
import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_ref

pytestmark = pytest.mark.gpu


def _equal(a, b):
    return all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(a, b))


def test_batch_256_scenes_full_size(ctx):
    import batchbench

    wl = batchbench.BatchWorkload(ctx, 256)
    try:
        assert wl.tris > 10_000_000
        wl.batched()
        batch = [fb.read() for fb in wl.fbs]
        wl.plain()
        bad = [k for k, fb in enumerate(wl.fbs) if not _equal(batch[k], fb.read())]
        assert not bad, f"scenes differing from sr_draw: {bad[:10]} ({len(bad)} of 256)"
        specs = wl.specs
        big = [k for k, (n, w, h, _) in enumerate(specs) if w == 1920 and n >= 100_000]
        assert big
        picks = [min(big, key=lambda k: specs[k][0])]
        picks += [next(k for k, s in enumerate(specs) if s[1] == 256),
                  next(k for k, s in enumerate(specs) if s[1] == 512)]
        for k in picks:
            s = scenes.batch_scene(*specs[k])
            ref, _ = render_ref(dict(s, clear=0xFF000000))
            assert_parity(batch[k], ref)
    finally:
        wl.close()


def test_serving_full_mesh_cache(ctx):
    import serve

    world = serve.World(ctx, 64)
    try:
        big = [j for j, n in enumerate(world.tris) if n >= 500_000]
        assert big, world.tris
        reqs = scenes.serve_requests(40, n_meshes=64, seed=scenes.SEED_SERVE + 0x30000)
        # the largest meshes at 1080p, whatever the seeded mix drew
        rng = scenes.SplitMix64(scenes.SEED_SERVE + 0x30001)
        for j in big:
            reqs.append(dict(mesh=j, width=1920, height=1080,
                             desc=scenes.serve_desc(scenes.orbit_camera(rng, 1920, 1080), 1920, 1080)))
        srv = sr.Server(ctx, hbm_budget=64 << 30, max_inflight=4, max_batch_frames=32,
                        max_batch_tris=48 << 20)
        try:
            tickets = {}
            for i, r in enumerate(reqs):
                tickets[srv.submit(world.meshes[r["mesh"]], sr.make_desc(**r["desc"]), r["width"],
                                   r["height"], tex=world.tex, want_depth=True, tag=i)] = i
            got = []
            while len(got) < len(reqs):
                got += srv.poll(64, timeout_ms=200)
        finally:
            srv.close()
        assert all(g["status"] == 0 for g in got)
        assert max(g["batch_frames"] for g in got) > 1
        by_tag = {g["tag"]: g for g in got}
        for i, r in enumerate(reqs):
            fb = ctx.framebuffer(r["width"], r["height"])
            fb.clear(0xFF000000)
            ctx.draw(fb, world.meshes[r["mesh"]], sr.make_desc(**r["desc"]), world.tex)
            col, dep, _ = fb.read()
            fb.free()
            g = by_tag[i]
            assert np.array_equal(g["rgba"], col), f"request {i} colour"
            assert np.array_equal(g["depth"].view(np.uint32), dep.view(np.uint32)), f"request {i} depth"
        # one 1080p request of a >= 500k-triangle mesh against the oracle
        i = len(reqs) - len(big)
        r = reqs[i]
        n, sd = world.specs[r["mesh"]]
        mesh = scenes.serve_mesh(n, sd)
        scene = dict(mesh, width=1920, height=1080, clear=0xFF000000, desc=r["desc"],
                     texture=(1024, 1024, scenes.checkerboard(1024, 32)))
        ref, _ = render_ref(scene)
        g = by_tag[i]
        assert np.array_equal(g["rgba"], ref.color)
        assert np.array_equal(g["depth"].view(np.uint32), ref.depth.view(np.uint32))
    finally:
        world.close()
