"""Seeded random configurations against the oracle: framebuffer size and
screen tile, execution tile, cull, blend, alpha mode, fragment program,
stencil test and operation, near clipping, triangle counts from 1 to a few
thousand, several draws into one framebuffer.  Each case checks coverage,
depth and stencil bit for bit and colour exactly, like the fixed parity
tests, over combinations those do not spell out one by one."""
import numpy as np
import pytest

import scenes
from parity import assert_parity, render_gpu, render_ref

pytestmark = pytest.mark.gpu

# stencil (test, op) pairs the binned raster takes: Always with any op,
# value tests with Keep, and value tests with an op (exhaustive order)
STENCIL = [(0, 0), (0, 4), (0, 5), (0, 1), (2, 0), (6, 0), (4, 4), (7, 3), (3, 5)]


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    W = int(rng.integers(2, 320))
    H = int(rng.integers(2, 200))
    n = int(rng.choice([1, 7, 60, 500, 2500]))
    s = scenes.soup_scene(n, W, H, seed=3000 + seed)
    if rng.random() < 0.3:  # some triangles cross the near plane
        z = s["pz"].reshape(-1, 3).copy()
        z[:, 0] = np.where(rng.random(len(z)) < 0.4, -0.3, z[:, 0])
        s["pz"] = z.reshape(-1).astype(np.float32)
        gs = int(rng.choice([1, 2]))
    else:
        gs = 0
    test, op = STENCIL[int(rng.integers(len(STENCIL)))]
    fp = int(rng.choice([0, 2, 3]))
    fs = np.zeros(32, np.float32)
    fs[0:4] = rng.random(4)
    desc = dict(s["desc"], geometry_program=gs, cull=int(rng.integers(3)), blend=int(rng.integers(2)),
                alpha_mode=int(rng.integers(2)), exec_tile=int(rng.choice([0, 8, 16, 32, 32])),
                tile_w=int(rng.choice([16, 40, 64, 128])), tile_h=int(rng.choice([16, 33, 64, 128])),
                stencil_test=test, stencil_op=op, stencil_wrap=int(rng.integers(2)),
                stencil_ref=int(rng.integers(0, 4)))
    if fp != 2:
        desc.update(fragment_program=fp, fs_uniforms=fs if fp == 0 else desc["fs_uniforms"])
    s["desc"] = desc
    if rng.random() < 0.3:  # a second draw into the same framebuffer
        t = scenes.soup_scene(max(1, n // 2), W, H, seed=5000 + seed)
        t["desc"] = dict(desc, blend=1 - desc["blend"])
        return dict(width=W, height=H, clear=0xFF203040, texture=s["texture"], draws=[s, t])
    return s


@pytest.mark.parametrize("seed", range(48))
def test_random_configuration(ctx, seed):
    s = random_case(seed)
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)


@pytest.mark.parametrize("seed", range(16))
def test_random_lines_and_points(ctx, seed):
    from test_gpu_lines_points import line_scene, point_scene

    rng = np.random.default_rng(7000 + seed)
    W, H = int(rng.integers(8, 300)), int(rng.integers(8, 200))
    test, op = STENCIL[int(rng.integers(len(STENCIL)))]
    kw = dict(stencil_test=test, stencil_op=op, stencil_wrap=int(rng.integers(2)),
              stencil_ref=int(rng.integers(0, 3)))
    tile = (int(rng.choice([16, 50, 128])), int(rng.choice([16, 37, 128])))
    lines = line_scene(int(rng.choice([1, 40, 600])), W, H, seed=8000 + seed, aa=int(rng.integers(2)),
                       alpha_mode=int(rng.integers(2)), blend=int(rng.integers(2)), tile=tile, **kw)
    points = point_scene(int(rng.choice([1, 300, 3000])), W, H, seed=9000 + seed, blend=int(rng.integers(2)),
                         tile_w=tile[0], tile_h=tile[1], **kw)
    s = dict(width=W, height=H, clear=0xFF000000, texture=None, draws=[lines, points])
    ref, _ = render_ref(s)
    assert_parity(render_gpu(ctx, s), ref)
