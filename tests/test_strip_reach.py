"""The reach bound that bins slivers (setup's strip_reach, sr_kernels.hip).

A triangle's binned raster tests it only in the execution tiles its pairs
name.  Setup claims, per triangle, where a pixel centre can pass the f32
coverage test of triangle.rs:104-113:

  BIN_BOX    nowhere outside the triangle's box,
  BIN_REACH  nowhere outside the reach box it returns,
  BIN_CLAMP  (no claim: the triangle goes to every tile of its clamped boxes).

Here a numpy mirror of strip_reach (test infrastructure, written from the
device function) makes those claims and a brute force checks them: every
pixel centre of the framebuffer goes through the reference's f32
expressions (triangle.rs:104-113, one IEEE operation at a time) and every
pixel that passes must lie inside the claimed region.  The triangles are
the thinnest of a perspective view of a ball of randomly oriented
triangles (the serving workload's kind of mesh, where slivers are common),
random slivers near screen-tile corners, and the three adversarial slivers
of tests/slivers.py.  A planted mistake (the bound without its widening)
must be caught by the same brute force.
"""
import numpy as np
import pytest

import scenes
from slivers import SLIVERS, W as SLIVER_W

f32 = np.float32
BIN_BOX, BIN_REACH, BIN_CLAMP = 0, 1, 2


def strip_reach_np(x, y, W, H, widen=1.0):
    """Mirror of strip_reach for screen vertices x, y (float32, shape (n, 3)):
    (mode, box, reach) per triangle; `widen` scales the margins (1 is the
    device's; 0 plants a mistake)."""
    x1, x2, x3 = x[:, 0], x[:, 1], x[:, 2]
    y1, y2, y3 = y[:, 0], y[:, 1], y[:, 2]
    with np.errstate(all="ignore"):
        det = ((y2 - y3) * (x1 - x3)) + ((x3 - x2) * (y1 - y3))
        e23y, e23x, e31y, e31x = [t.astype(np.float64) for t in (y2 - y3, x3 - x2, y3 - y1, x1 - x3)]
        X3, Y3, df = x3.astype(np.float64), y3.astype(np.float64), det.astype(np.float64)
        dm = e23y * e31x - e23x * e31y
        k = df / dm
        q = 2.0 ** -20 * widen
        au = q * (np.abs(e23y) + np.abs(e23x)) / np.abs(df); bu = 2.0 * q
        av = q * (np.abs(e31y) + np.abs(e31x)) / np.abs(df); bv = 2.0 * q
        aw = au + av; bw = bu + bv + 2.0 ** -18 * widen
        U0 = [-bu, 1.0 + bv + bw, -bu]; U1 = [-au, av + aw, -au]
        V0 = [-bv, -bv, 1.0 + bu + bw]; V1 = [-av, -av, au + aw]
        c0 = np.zeros_like(df); c1 = np.zeros_like(df)
        for i in range(3):
            ax = k * (e31x * U0[i] - e23x * V0[i]); ay = k * (-e31y * U0[i] + e23y * V0[i])
            bx = k * (e31x * U1[i] - e23x * V1[i]); by = k * (-e31y * U1[i] + e23y * V1[i])
            c0 = np.fmax(c0, np.fmax(np.abs(ax), np.abs(ay)))
            c1 = np.fmax(c1, np.fmax(np.abs(bx), np.abs(by)))
        D = (c0 + 2.0) / (1.0 - c1)
        eu = au * D + bu; ev = av * D + bv; ew = eu + ev + 2.0 ** -18 * widen
        us = [-eu, 1.0 + ev + ew, -eu]; vs = [-ev, -ev, 1.0 + eu + ew]
        sx = np.stack([X3 + k * (e31x * u - e23x * v) for u, v in zip(us, vs)], 1)
        sy = np.stack([Y3 + k * (-e31y * u + e23y * v) for u, v in zip(us, vs)], 1)
        mnx, mxx, mny, mxy = sx.min(1), sx.max(1), sy.min(1), sy.max(1)
    never = ~(det != 0) | ~(np.abs(det) < np.inf)
    clamp = ~never & (~(dm != 0) | ~(c1 < 0.5) | ~(D < 1e9) | ~((eu < 0.25) & (ev < 0.25))
                      | ~(np.abs(mnx) + np.abs(mxx) + np.abs(mny) + np.abs(mxy) < 1e300))
    bx0 = np.floor(np.clip(np.fmin(np.fmin(x1, x2), x3), -1, W)).astype(np.int64)
    bx1 = np.floor(np.clip(np.fmax(np.fmax(x1, x2), x3), -1, W)).astype(np.int64)
    by0 = np.floor(np.clip(np.fmin(np.fmin(y1, y2), y3), -1, H)).astype(np.int64)
    by1 = np.floor(np.clip(np.fmax(np.fmax(y1, y2), y3), -1, H)).astype(np.int64)
    m = 1e-3
    with np.errstate(all="ignore"):
        out = (((bx0 >= 1) & ~(mnx > bx0 - 0.5 + m)) | ((bx1 <= W - 2) & ~(mxx < bx1 + 1.5 - m)) |
               ((by0 >= 1) & ~(mny > by0 - 0.5 + m)) | ((by1 <= H - 2) & ~(mxy < by1 + 1.5 - m)))
        rx0 = np.fmax(np.fmin(np.ceil(mnx - 0.5 - m), bx0), 0)
        rx1 = np.fmin(np.fmax(np.floor(mxx - 0.5 + m), bx1), W - 1)
        ry0 = np.fmax(np.fmin(np.ceil(mny - 0.5 - m), by0), 0)
        ry1 = np.fmin(np.fmax(np.floor(mxy - 0.5 + m), by1), H - 1)
    mode = np.where(never, BIN_BOX, np.where(clamp, BIN_CLAMP, np.where(out, BIN_REACH, BIN_BOX)))
    # the cheap sufficient test setup runs first (strip_box_fast)
    with np.errstate(all="ignore"):
        D = np.maximum(np.maximum(np.abs(X3), np.abs(W - X3)), np.maximum(np.abs(Y3), np.abs(H - Y3))) + 1.0
        ad, adm = np.abs(df), np.abs(dm)
        Lu = np.abs(e23y) + np.abs(e23x); Lv = np.abs(e31y) + np.abs(e31x)
        dE = widen * (2.0 ** -19 * ((Lu + Lv) * D + 4.0 * ad) + 2.0 ** -18 * ad)
        fast = ((np.abs(e31x) + np.abs(e23x)) * dE < 0.45 * adm) & ((np.abs(e31y) + np.abs(e23y)) * dE < 0.45 * adm)
    mode = np.where(fast, BIN_BOX, mode)
    return mode, (bx0, bx1, by0, by1), (rx0, rx1, ry0, ry1)


def passing_pixels(x, y, W, H):
    """Every pixel (px, py) of the W x H framebuffer whose centre passes the
    coverage test of triangle.rs:104-113 for the triangle (x[0..2], y[0..2])
    (float32, one IEEE operation at a time)."""
    x1, x2, x3 = (f32(v) for v in x)
    y1, y2, y3 = (f32(v) for v in y)
    det = ((y2 - y3) * (x1 - x3)) + ((x3 - x2) * (y1 - y3))
    py, px = np.mgrid[0:H, 0:W]
    fx = px.astype(f32) + f32(0.5)
    fy = py.astype(f32) + f32(0.5)
    with np.errstate(all="ignore"):
        u = (((y2 - y3) * (fx - x3)) + ((x3 - x2) * (fy - y3))) / det
        v = (((y3 - y1) * (fx - x3)) + ((x1 - x3) * (fy - y3))) / det
        w = (f32(1.0) - u) - v
    ok = ~((u < 0) | (v < 0) | (w < 0)) & ~np.isnan(u) & ~np.isnan(v) & ~np.isnan(w)
    return px[ok], py[ok]


def _check(x, y, W, H, widen=1.0):
    """Brute-force every claim; returns (violations, counts per mode)."""
    mode, box, reach = strip_reach_np(x, y, W, H, widen)
    bad = []
    for t in range(len(x)):
        if mode[t] == BIN_CLAMP:
            continue
        px, py = passing_pixels(x[t], y[t], W, H)
        r = box if mode[t] == BIN_BOX else reach
        inside = (px >= r[0][t]) & (px <= r[1][t]) & (py >= r[2][t]) & (py <= r[3][t])
        if not inside.all():
            bad.append(t)
    return bad, np.bincount(mode, minlength=3)


def _ball_view(n, W, H, seed):
    """Screen vertices of a serving mesh (a ball of randomly oriented
    triangles) under an orbit camera, float32 (the vertex program's shape)."""
    s = scenes.serve_mesh(n, seed)
    mvp = scenes.orbit_camera(scenes.SplitMix64(seed + 1), W, H)
    P = np.stack([s["px"], s["py"], s["pz"], np.ones_like(s["px"])]).astype(f32)
    clip = (mvp.astype(f32) @ P).astype(f32)
    xn, yn = clip[0] / clip[3], clip[1] / clip[3]
    x = (f32(W / 2) * xn + f32(W / 2)).astype(f32).reshape(-1, 3)
    y = (f32(-H / 2) * yn + f32(H / 2)).astype(f32).reshape(-1, 3)
    keep = (clip[3].reshape(-1, 3) > 0).all(1)
    return x[keep], y[keep]


def _thinnest(x, y, n):
    with np.errstate(all="ignore"):
        det = np.abs(((y[:, 1] - y[:, 2]) * (x[:, 0] - x[:, 2])) + ((x[:, 2] - x[:, 1]) * (y[:, 0] - y[:, 2])))
        L2 = np.max([(x[:, a] - x[:, b]) ** 2 + (y[:, a] - y[:, b]) ** 2 for a, b in ((0, 1), (1, 2), (2, 0))], 0)
        thin = det.astype(np.float64) / np.maximum(L2.astype(np.float64), 1e-30)
    order = np.argsort(thin)
    return x[order[:n]], y[order[:n]]


def test_ball_view_slivers_stay_in_their_claims():
    W, H = 384, 256
    x, y = _ball_view(60_000, W, H, seed=0x51)
    x, y = _thinnest(x, y, 1500)
    bad, counts = _check(x, y, W, H)
    assert not bad, f"{len(bad)} triangles pass outside their claimed region: {bad[:5]}"
    assert counts[BIN_REACH] + counts[BIN_CLAMP] > 0 and counts[BIN_BOX] > 1000, counts


def test_slivers_at_screen_tile_corners():
    # thin triangles whose sharp ends sit near multiples of 32 (tile corners)
    rng = np.random.default_rng(7)
    W = H = 128
    n = 1500
    tip = (rng.integers(1, 4, (n, 2)) * 32 + rng.uniform(-1.5, 1.5, (n, 2))).astype(f32)
    ang = rng.uniform(0, 2 * np.pi, n)
    L = rng.uniform(5, 80, n)
    wdt = 10.0 ** rng.uniform(-4, -0.5, n)
    dx, dy = np.cos(ang), np.sin(ang)
    far = np.stack([tip[:, 0] + L * dx, tip[:, 1] + L * dy], 1)
    x = np.stack([tip[:, 0], far[:, 0] - wdt * dy, far[:, 0] + wdt * dy], 1).astype(f32)
    y = np.stack([tip[:, 1], far[:, 1] + wdt * dx, far[:, 1] - wdt * dx], 1).astype(f32)
    bad, counts = _check(x, y, W, H)
    assert not bad, f"{len(bad)} triangles pass outside their claimed region: {bad[:5]}"
    assert counts[BIN_REACH] + counts[BIN_CLAMP] > 0, counts


@pytest.mark.parametrize("k", range(len(SLIVERS)))
def test_adversarial_slivers_are_not_boxed(k):
    # each passes pixel (512, 512) outside its box: the claim must cover it
    xs, ys = SLIVERS[k]
    px = np.array([float.fromhex(v) for v in xs], f32)
    py = np.array([float.fromhex(v) for v in ys], f32)
    W = SLIVER_W
    x = ((px * f32(W / 2)) + f32(W / 2)).reshape(1, 3)
    y = ((py * f32(-W / 2)) + f32(W / 2)).reshape(1, 3)
    mode, box, reach = strip_reach_np(x, y, W, W)
    assert mode[0] != BIN_BOX
    if mode[0] == BIN_REACH:
        assert reach[0][0] <= 512 <= reach[1][0] and reach[2][0] <= 512 <= reach[3][0]
    bad, _ = _check(x, y, W, W)
    assert not bad


def test_planted_mistake_is_caught():
    # without the widening the bound calls the adversarial slivers boxed;
    # the brute force must see them pass outside
    xs = np.array([[float.fromhex(v) for v in s[0]] for s in SLIVERS], f32)
    ys = np.array([[float.fromhex(v) for v in s[1]] for s in SLIVERS], f32)
    W = SLIVER_W
    x = (xs * f32(W / 2)) + f32(W / 2)
    y = (ys * f32(-W / 2)) + f32(W / 2)
    bad, counts = _check(x, y, W, W, widen=0.0)
    assert bad, counts
