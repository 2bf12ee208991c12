"""The depth bound behind the raster's hierarchical depth test, on the CPU.

Setup stores, per triangle, depth_key(zhi) rounded up to even, zhi an
upper bound of every depth the triangle's fragments can carry
(depth_bound_key in csrc/sr_kernels.hip); the raster skips a triangle in a
wave whose least pixel depth key lies above it.  That is exact only if no
fragment's depth, formed as triangle.rs:108-117 forms it (f32, reference
operation order, u, v, w >= 0), exceeds zhi.  These tests restate the bound
and the key in numpy and check them by brute force over barycentrics that
pass the coverage test, including ones at the rounding extremes, and check
that the bound without its slack term fails (so the test can see one)."""
import numpy as np

F32 = np.float32


def depth_key(bits):
    bits = np.asarray(bits, np.uint32)
    sign = (bits.view(np.int32) >> 31).view(np.uint32)
    return bits ^ (sign | np.uint32(0x80000000))


def key_of(f):
    return depth_key(np.asarray(f, F32).view(np.uint32))


def depth_bound_key(z1, z2, z3, slack=True):
    z = np.stack([z1, z2, z3]).astype(np.float64)
    M = z.max(0)
    K = np.abs(z).max(0)
    with np.errstate(invalid="ignore", over="ignore"):
        b = M + (K * 2.0 ** -19 + 2.0 ** -140 if slack else 0.0)
        f = b.astype(F32)
        up = f.astype(np.float64) < b
        f = np.where(up, np.nextafter(f, F32(np.inf)), f)
        ok = np.isfinite(z).all(0) & (b < np.inf)
    key = np.where(ok, key_of(f), np.uint32(0xFF800000))
    return np.where(key >= 0xFFFFFFFE, np.uint32(0xFFFFFFFE), (key + np.uint32(1)) & np.uint32(0xFFFFFFFE))


def fragment_depth(z1, z2, z3, u, v):
    """w and z as the raster forms them (triangle.rs:110,115-117)."""
    w = (F32(1.0) - u) - v
    z = ((z1 * u) + (z2 * v)) + (z3 * w)
    return w, z


def barycentrics(rng, n):
    """Random u, v in [0, 1] that keep w >= 0, a third of them at the
    extremes where w's two roundings add up."""
    u = rng.uniform(0, 1, n).astype(F32)
    v = (rng.uniform(0, 1, n) * (1.0 - u.astype(np.float64))).astype(F32)
    k = n // 3
    # u + v just below 1 with many significant bits: w's rounding errors
    v[:k] = (F32(1.0) - u[:k]) * F32(1 - 2.0 ** -20)
    u[k:2 * k] = F32(1.0) - np.abs(rng.normal(0, 1e-6, k)).astype(F32)
    v[k:2 * k] = np.abs(rng.normal(0, 1e-7, k)).astype(F32)
    return u, v


def depths(rng, n, kind):
    if kind == "equal":  # one depth per triangle, as in the soup scenes
        z = rng.uniform(-1, 0, n).astype(F32)
        return z, z.copy(), z.copy()
    if kind == "near":   # depths a few ulps apart
        z = rng.uniform(-1, 0, n).astype(F32)
        d = rng.integers(-4, 5, (2, n)).astype(np.int32)
        return z, (z.view(np.int32) + d[0]).view(F32), (z.view(np.int32) + d[1]).view(F32)
    if kind == "mixed":  # signs and magnitudes of every sort
        e = rng.uniform(-140, 20, (3, n))
        s = np.where(rng.uniform(0, 1, (3, n)) < 0.5, -1.0, 1.0)
        z = (s * 2.0 ** e).astype(F32)
        return z[0], z[1], z[2]
    if kind == "tiny":   # subnormal products
        z = (-rng.uniform(0, 1, (3, n)) * 2.0 ** -126).astype(F32)
        return z[0], z[1], z[2]
    raise ValueError(kind)


def test_key_orders_floats():
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.normal(0, 1, 20000) * 10.0 ** rng.integers(-30, 30, 20000),
                        [0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-45]]).astype(F32)
    a = np.unique(a)  # sorted, -0 and +0 merged
    ka = key_of(a)
    assert (np.diff(ka.astype(np.int64)) > 0).all()
    assert key_of(F32(-0.0)) == key_of(F32(0.0)) - 1
    assert key_of(F32(np.nan)) > key_of(F32(np.inf))
    assert key_of(-F32(np.nan)) < key_of(F32(-np.inf))


def test_bound_holds_for_passing_fragments():
    rng = np.random.default_rng(7)
    n = 400_000
    for kind in ("equal", "near", "mixed", "tiny"):
        z1, z2, z3 = depths(rng, n, kind)
        u, v = barycentrics(rng, n)
        w, z = fragment_depth(z1, z2, z3, u, v)
        ok = (w >= 0) & ~np.isnan(z)
        assert ok.mean() > 0.9, kind
        kb = depth_bound_key(z1, z2, z3)
        kz = key_of(z)
        bad = ok & (kz > kb)
        assert not bad.any(), (kind, z1[bad][:3], z2[bad][:3], z3[bad][:3], u[bad][:3], v[bad][:3])


def test_bound_without_slack_fails():
    """With zhi = max z_i the rounded sum does exceed it: the slack is needed
    and the brute force sees its absence."""
    rng = np.random.default_rng(11)
    n = 400_000
    z1, z2, z3 = depths(rng, n, "equal")
    u, v = barycentrics(rng, n)
    w, z = fragment_depth(z1, z2, z3, u, v)
    ok = (w >= 0) & ~np.isnan(z)
    assert (ok & (key_of(z) > depth_bound_key(z1, z2, z3, slack=False))).any()


def test_special_depths():
    inf, nan = F32(np.inf), F32(np.nan)
    z = np.array([[nan, -1, -1], [inf, -1, -1], [-inf, -1, -1], [-inf, -inf, -inf]], F32).T
    k = depth_bound_key(*z)
    assert (k == 0xFF800000).all()  # never skipped below a finite or -NaN depth
    big = np.full(3, np.finfo(F32).max)
    assert depth_bound_key(*big[:, None]) == 0xFF800000  # rounds up past FLT_MAX
