"""Numerical arguments behind the device corrections, checked on the CPU with IEEE numpy arithmetic.

The default (non-exact) device epilogue forms the two K2P proportions from one FP64 reciprocal of
`valid` and takes one logarithm of x^2 y instead of two; JC forms 1 - p/0.75 with an fma. These
tests pin the claims the kernels rely on: the proportions are the correctly rounded float
quotients even when the reciprocal is off by one ulp either way, and the regrouped logarithms stay
far inside the 1e-6 relative tolerance of the reference's expression (KimuraDistance.cpp:339-348,
JCdistance.cpp:170-183).
"""
import numpy as np

RNG = np.random.default_rng(20261203)


def _cases(k: int):
    v = RNG.integers(1, 1 << 24, size=k, dtype=np.int64)
    a = (RNG.random(k) * (v + 1)).astype(np.int64).clip(0, v)
    edge_v = np.array([1, 2, 3, 7, 3 * 5 * 7 * 11, (1 << 24) - 1, (1 << 24) - 3, 1 << 23, 5000,
                       4999, 10007], dtype=np.int64)
    ev = np.repeat(edge_v, 6)
    ea = np.stack([np.zeros_like(edge_v), np.ones_like(edge_v), edge_v - 1, edge_v, edge_v // 2,
                   edge_v // 3]).T.reshape(-1).clip(0, ev)
    return np.concatenate([a, ea]), np.concatenate([v, ev])


def test_shared_reciprocal_gives_correctly_rounded_proportions():
    a, v = _cases(400_000)
    want = a.astype(np.float32) / v.astype(np.float32)   # IEEE float division (correctly rounded)
    r = 1.0 / v.astype(np.float64)
    for rr in (r, np.nextafter(r, 0.0), np.nextafter(r, np.inf)):
        got = (a.astype(np.float64) * rr).astype(np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def _k2p_reference(ts, tv, valid):
    P = ts.astype(np.float32) / valid.astype(np.float32)
    Q = tv.astype(np.float32) / valid.astype(np.float32)
    x = 1.0 - 2.0 * P.astype(np.float64) - Q.astype(np.float64)
    y = 1.0 - 2.0 * Q.astype(np.float64)
    ok = (x.astype(np.float32) > 0) & (y.astype(np.float32) > 0)
    return x, y, ok


def test_single_log_k2p_within_tolerance():
    k = 400_000
    valid = RNG.integers(1, 30001, size=k)
    frac = RNG.random(k) ** 3 * 0.6                       # many small distances, some large
    tv = (RNG.random(k) * frac * valid).astype(np.int64)
    ts = (RNG.random(k) * frac * valid).astype(np.int64)
    ts = np.minimum(ts, valid - tv)
    ts[:1000], tv[:1000] = 1, 0                           # the smallest nonzero distances
    ts[1000:2000], tv[1000:2000] = 0, 1
    x, y, ok = _k2p_reference(ts, tv, valid)
    ref = (-0.5 * np.log(x[ok]) - 0.25 * np.log(y[ok])).astype(np.float32)
    got = (-0.25 * np.log(x[ok] * x[ok] * y[ok])).astype(np.float32)
    nz = ref != 0
    rel = np.abs(got[nz].astype(np.float64) - ref[nz]) / np.abs(ref[nz])
    assert rel.max() < 1e-8
    assert np.mean(got != ref) < 0.01                     # rare 1-ulp differences only


def test_jc_fma_argument_within_tolerance():
    k = 400_000
    valid = RNG.integers(1, 30001, size=k)
    mism = (RNG.random(k) ** 2 * 0.74 * valid).astype(np.int64)
    mism[:1000] = 1
    edge_valid = np.arange(1, 20001)                      # the saturation boundary p = 0.75
    edge_mism = np.concatenate([np.floor(0.75 * edge_valid), np.ceil(0.75 * edge_valid)])
    valid = np.concatenate([valid, edge_valid, edge_valid])
    mism = np.concatenate([mism, edge_mism.astype(np.int64)])
    p = mism.astype(np.float32) / valid.astype(np.float32)
    q = p.astype(np.float64) / 0.75
    sat = q >= 1.0
    assert np.array_equal(sat, p >= np.float32(0.75))   # the device's saturation test
    keep = ~sat & (mism > 0)
    ref = (-0.75 * np.log(1.0 - q[keep])).astype(np.float32)
    arg = 1.0 - p[keep].astype(np.float64) * (4.0 / 3.0)  # fma(-p, 4/3, 1) up to one rounding
    got = (-0.75 * np.log(arg)).astype(np.float32)
    rel = np.abs(got.astype(np.float64) - ref) / np.abs(ref)
    assert rel.max() < 1e-7
