"""Pin the oracle's contact lists: its grid broadphase + FP32 narrowphase must
equal an all-pairs brute force written independently in numpy float32 with the
same operation order (collision.cu:39-92 semantics), on small scenes."""
import numpy as np
import pytest

from oracle import orc32
from yarnball_amd import scenes

f32 = np.float32


def l2(a):
    return (a[..., 0] * a[..., 0] + a[..., 1] * a[..., 1]) + a[..., 2] * a[..., 2]


def dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def clamp01(x):
    return np.fmin(np.fmax(x, f32(0)), f32(1))


def mix(a, b, t):
    s = f32(1) - t
    return a * s[..., None] + b * t[..., None]


def brute_contacts(p, pos, flags, cid, dx):
    n = len(flags)
    segs = np.nonzero(flags & 2)[0]
    p0, p1 = pos[segs], pos[segs + 1]
    vp = np.sqrt(np.fmax(l2(dx[segs]), l2(dx[segs + 1])))
    sdr = f32(p.scaledDetectionRadius)
    pad = sdr + vp if p.useVelocityRadius else np.full(len(segs), sdr, f32)
    lo = np.fmin(p0, p1) - pad[:, None]
    hi = np.fmax(p0, p1) + pad[:, None]
    I, J = np.triu_indices(len(segs), 1)
    ov = np.all((lo[I] <= hi[J]) & (lo[J] <= hi[I]), axis=1)
    I, J = I[ov], J[ov]
    x, y = segs[I], segs[J]  # x < y
    keep = (y - x) > 2
    c0, c1 = cid[x], cid[x + 1]
    keep &= ~((c0 == y) | (c1 == y) | (c0 == y + 1) | (c1 == y + 1))
    a0 = pos[x]
    a1 = pos[x + 1] - a0
    b0 = pos[y] - a0
    b1 = pos[y + 1] - a0
    A1 = b1 - b0
    a = l2(a1)
    b = dot(a1, A1)
    c = l2(A1)
    with np.errstate(all="ignore"):
        vy = b * dot(a1, b0) - a * dot(A1, b0)
        vy = vy / (a * c - b * b)
        vy = np.where(np.isfinite(vy), vy, f32(0.5)).astype(f32)
        vy = clamp01(vy)
        ux = clamp01(dot(mix(b0, b1, vy), a1) / a)
        vy = clamp01(dot(a1 * ux[:, None] - b0, A1) / c)
    nrm = a1 * ux[:, None] - mix(b0, b1, vy)
    d2 = l2(nrm)
    r = f32(2) * sdr
    if p.useVelocityRadius:
        r = (r + vp[I]) + vp[J]
    keep &= d2 < r * r
    lists = [[] for _ in range(n)]
    for u, w in zip(x[keep], y[keep]):
        lists[u].append(int(w))
        lists[w].append(int(u))
    return [tuple(sorted(l)) for l in lists]


@pytest.mark.parametrize("make", [
    lambda: scenes.knit_patch(rows=6, per=200),
    lambda: scenes.twisted_cable(strands=7, per=150, bundle_r=4e-4),
    lambda: scenes.random_walk(1200, seed=3),
])
def test_oracle_contacts_equal_brute_force(make):
    sc = make()
    rng = np.random.default_rng(0)
    p = sc.params
    pos = sc.verts["pos"].copy()
    dx = (rng.normal(0, 2e-5, pos.shape)).astype(f32)
    flags = sc.verts["flags"]
    cid = sc.verts["connectionIndex"].copy()
    # glue a few endpoints to exercise the exemption
    cid[5], cid[60] = 60, 5
    ab = orc32.aabb(p, pos, flags, dx)
    nc, cols, err = orc32.contacts(p, pos, flags, cid, dx, ab)
    got = [tuple(sorted(cols[: nc[i], i].tolist())) for i in range(sc.n)]
    want = brute_contacts(p, pos, flags, cid, dx)
    assert got == want
    assert sum(map(len, want)) > 0 or sc.name.startswith("random")


def test_contact_overflow_raises_error_flag():
    """More than 128 contacts on one segment (the reference's error[0])."""
    import ctypes

    from yarnball_amd import _lib
    from yarnball_amd._lib import VERTEX_DTYPE, Params, ptr

    lib = _lib.load()
    # one long segment crossed by 200 short parallel segments
    m = 200
    n = 2 + 2 * m
    pos = np.zeros((n, 3), f32)
    pos[0] = (-1e-2, 0, 0)
    pos[1] = (1e-2, 0, 0)
    for k in range(m):
        x = -9e-3 + k * 9e-5
        pos[2 + 2 * k] = (x, -5e-4, 1e-4)
        pos[3 + 2 * k] = (x, 5e-4, 1e-4)
    flags = np.zeros(n, np.uint32)
    flags[0::2] = 2
    cid = np.full(n, -1, np.int32)
    p = Params()
    lib.yb_default_params(ctypes.byref(p), n, 1e-3)
    p.minSegLen = 1e-3
    lib.yb_validate_params(ctypes.byref(p))
    dx = np.zeros((n, 3), f32)
    ab = orc32.aabb(p, pos, flags, dx)
    nc, cols, err = orc32.contacts(p, pos, flags, cid, dx, ab)
    assert err == 1 and nc[0] == 128
