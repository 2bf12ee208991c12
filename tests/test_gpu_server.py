"""Request server (sr_server_*): every served frame equals the same request
rendered alone (sr_fb_clear + sr_draw), bit for bit, whatever else shared
its batch; small frames also equal the oracle.  Covers continuous batching
(requests joining batches while others run), budget admission, overflow
reruns and argument checks."""
import time

import numpy as np
import pytest

import scenes
import softrender_hip as sr
from parity import assert_parity, render_ref

pytestmark = pytest.mark.gpu

RES = [(256, 256), (320, 180), (97, 61), (512, 512)]


@pytest.fixture(scope="module")
def world(ctx):
    specs = [(n, scenes.SEED_SERVE + 500 + j) for j, n in enumerate((300, 1500, 4000, 9000, 700))]
    meshes, mesh_specs = [], []
    for n, seed in specs:
        m = scenes.serve_mesh(n, seed)
        mesh_specs.append(m)
        meshes.append(ctx.mesh(m))
    tw, th, rgba = 1024, 1024, scenes.checkerboard(1024, 32)
    tex = ctx.texture(tw, th, rgba)
    yield dict(meshes=meshes, mesh_specs=mesh_specs, tex=tex, texture=(tw, th, rgba))
    for m in meshes:
        m.free()
    tex.free()


def requests(n, seed, n_meshes):
    return scenes.serve_requests(n, n_meshes=n_meshes, seed=seed, resolutions=RES)


def alone(ctx, world, rq):
    """The request rendered on its own through the plain path."""
    fb = ctx.framebuffer(rq["width"], rq["height"])
    fb.clear(0xFF000000)
    ctx.draw(fb, world["meshes"][rq["mesh"]], sr.make_desc(**rq["desc"]), world["tex"])
    col, dep, _ = fb.read()
    fb.free()
    return col, dep


def submit_all(srv, world, reqs, want_depth=True):
    tickets = {}
    for i, rq in enumerate(reqs):
        t = srv.submit(world["meshes"][rq["mesh"]], sr.make_desc(**rq["desc"]), rq["width"],
                       rq["height"], tex=world["tex"], want_depth=want_depth, tag=i)
        tickets[t] = i
    return tickets


def collect(srv, n, timeout=60.0):
    got, t0 = [], time.time()
    while len(got) < n:
        got += srv.poll(64, timeout_ms=100)
        assert time.time() - t0 < timeout, f"only {len(got)} of {n} completions"
    return got


def test_served_equals_alone_and_oracle(ctx, world):
    reqs = requests(24, seed=901, n_meshes=len(world["meshes"]))
    srv = sr.Server(ctx, max_inflight=2, max_batch_frames=8)
    try:
        submit_all(srv, world, reqs)
        got = collect(srv, len(reqs))
        st = srv.stats()
    finally:
        srv.close()
    assert all(g["status"] == 0 for g in got)
    # requests were submitted faster than batches drain: some shared a batch
    assert max(g["batch_frames"] for g in got) > 1
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for g in got:
        rq = reqs[g["tag"]]
        col, dep = alone(ctx, world, rq)
        assert np.array_equal(g["rgba"], col), g["tag"]
        assert np.array_equal(g["depth"].view(np.uint32), dep.view(np.uint32)), g["tag"]
        assert g["t_arrival"] <= g["t_admit"] <= g["t_done"]
    # the smallest frames against the oracle
    small = [g for g in got if reqs[g["tag"]]["width"] * reqs[g["tag"]]["height"] <= 97 * 61]
    assert small
    for g in small[:3]:
        rq = reqs[g["tag"]]
        scene = dict(world["mesh_specs"][rq["mesh"]], width=rq["width"], height=rq["height"],
                     desc=rq["desc"], clear=0xFF000000, texture=world["texture"])
        ref, _ = render_ref(scene)
        assert_parity((g["rgba"], g["depth"], ref.stencil), ref)


def test_budget_admission_one_at_a_time(ctx, world):
    # same mesh and resolution everywhere: every request costs the same
    reqs = scenes.serve_requests(6, n_meshes=len(world["meshes"]), seed=902, resolutions=[(256, 256)])
    for rq in reqs:
        rq["mesh"] = 3
    tiny = sr.Server(ctx, hbm_budget=1)
    try:
        with pytest.raises(sr.SoftrenderError) as e:
            tiny.submit(world["meshes"][0], sr.make_desc(**reqs[0]["desc"]), 64, 64, tex=world["tex"])
        assert e.value.code == -6
    finally:
        tiny.close()
    # one request at a time: the peak charge is one request's
    one = sr.Server(ctx, max_inflight=1, max_batch_frames=1)
    try:
        submit_all(one, world, reqs, want_depth=False)
        collect(one, len(reqs))
        largest = one.stats()["peak_in_flight_bytes"]
    finally:
        one.close()
    # a budget of exactly one request: admission has to serialise them
    srv = sr.Server(ctx, hbm_budget=largest, max_inflight=2)
    try:
        submit_all(srv, world, reqs, want_depth=False)
        got = collect(srv, len(reqs))
        st = srv.stats()
    finally:
        srv.close()
    assert all(g["status"] == 0 for g in got)
    assert st["peak_in_flight_bytes"] <= largest
    assert st["max_batch_frames_seen"] == 1
    for g in got:
        col, _ = alone(ctx, world, reqs[g["tag"]])
        assert np.array_equal(g["rgba"], col)


def test_overflow_reruns_keep_results(ctx, world):
    reqs = requests(8, seed=903, n_meshes=len(world["meshes"]))
    srv = sr.Server(ctx, pairs_per_tri=0.001, max_batch_frames=8)
    try:
        submit_all(srv, world, reqs, want_depth=False)
        got = collect(srv, len(reqs))
        st = srv.stats()
    finally:
        srv.close()
    assert all(g["status"] == 0 for g in got)
    for g in got:
        col, _ = alone(ctx, world, reqs[g["tag"]])
        assert np.array_equal(g["rgba"], col)
    assert st["completed"] == len(reqs)
    assert st["overflow_reruns"] > 0


def test_rejects_non_batchable(ctx, world):
    srv = sr.Server(ctx)
    try:
        d = dict(scenes.serve_desc(np.eye(4), 64, 64))
        d["stencil_test"] = 2
        with pytest.raises(sr.SoftrenderError) as e:
            srv.submit(world["meshes"][0], sr.make_desc(**d), 64, 64, tex=world["tex"])
        assert e.value.code == -1
        d = dict(scenes.serve_desc(np.eye(4), 64, 64))
        with pytest.raises(sr.SoftrenderError):
            srv.submit(world["meshes"][0], sr.make_desc(**d), 0, 64, tex=world["tex"])
        with pytest.raises(sr.SoftrenderError):  # samples a texture but has none
            srv.submit(world["meshes"][0], sr.make_desc(**d), 64, 64, tex=None)
    finally:
        srv.close()
