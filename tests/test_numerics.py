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


def test_k2p_x_and_y_exact_and_saturation_on_the_doubles():
    """Below 2^24 valid sites the reference's x = 1 - 2P - Q and y = 1 - 2Q are exact in double
    (so the device's fma(-2, P, 1 - Q) and fma(-2, Q, 1) give the same doubles), and their float
    roundings are <= 0 exactly when the doubles are: the device tests x <= 0 and y <= 0 directly.
    Cases crowd the saturation boundary 2 ts + tv = valid and 2 tv = valid."""
    from fractions import Fraction
    k = 200_000
    valid = RNG.integers(1, 1 << 24, size=k, dtype=np.int64)
    tv = (RNG.random(k) * valid).astype(np.int64)
    ts = np.maximum((valid - tv) // 2 + RNG.integers(-2, 3, size=k), 0)  # near 2 ts + tv = valid
    ts = np.minimum(ts, valid - tv)
    half = valid[:20_000] // 2 + RNG.integers(-1, 2, size=20_000)        # near 2 tv = valid
    tv[:20_000] = np.clip(half, 0, valid[:20_000])
    ts[:20_000] = 0
    x, y, ok = _k2p_reference(ts, tv, valid)
    assert np.array_equal(ok, (x > 0) & (y > 0))
    assert ((x > 0) != (x.astype(np.float32) > 0)).sum() == 0
    P = (ts.astype(np.float32) / valid.astype(np.float32)).astype(np.float64)
    Q = (tv.astype(np.float32) / valid.astype(np.float32)).astype(np.float64)
    for idx in RNG.choice(k, 3000, replace=False):
        fp, fq = Fraction(P[idx]), Fraction(Q[idx])
        assert Fraction(x[idx]) == 1 - 2 * fp - fq
        assert Fraction(y[idx]) == 1 - 2 * fq
        assert Fraction(1.0 - Q[idx]) == 1 - fq
    assert (np.abs(x[x != 0]) >= 2.0 ** -47).all()


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


def test_jc_protein_saturation_and_fma_argument():
    """JC with alpha = 20 (protein): for every float p in [0.9, 1.05] the reference's test
    (double)p / 0.95 >= 1 equals the device's (double)p >= 0.95, and over [0, 0.95) the device's
    fma(-p, 20/19, 1) stays within 2^-52 of the reference's 1 - p / 0.95."""
    b = 19.0 / 20.0
    lo, hi = np.float32(0.9).view(np.uint32), np.float32(1.05).view(np.uint32)
    p = np.arange(lo, hi + 1, dtype=np.uint32).view(np.float32).astype(np.float64)
    assert np.array_equal(p / b >= 1.0, p >= b)
    q = np.concatenate([RNG.random(300_000).astype(np.float32), p[p < b].astype(np.float32)])
    q = q[q.astype(np.float64) < b].astype(np.float64)
    ref = 1.0 - q / b
    got = 1.0 - q * (20.0 / 19.0)                 # the fma's argument up to one more rounding
    assert np.abs(got - ref).max() <= 2.0 ** -52


# ---- the table-driven FP64 logarithm of the device epilogues (device_common.hpp log_table) ------
def _log_table_entries():
    import os
    path = os.path.join(os.path.dirname(__file__), "..", "rapidnj_amd", "csrc", "log_table.inc")
    inv, ln = [], []
    for line in open(path):
        line = line.strip()
        if line.startswith("{"):
            a, b = line.strip("{},").split(",")
            inv.append(float.fromhex(a.strip()))
            ln.append(float.fromhex(b.strip()))
    return np.array(inv), np.array(ln)


def test_log_table_matches_its_generator():
    """log_table.inc holds 1/c_i and ln(c_i) for c_i = 0.5 + (i+1)/256, each the double nearest
    the exact value (checked here with an independent 60-digit evaluation)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    inv, ln = _log_table_entries()
    assert len(inv) == 128
    for i in range(128):
        c = Decimal(1) / 2 + Decimal(i + 1) / 256
        assert inv[i] == float(Decimal(1) / c)
        assert ln[i] == float(c.ln())


def _log_table_np(z):
    """The device's log_table restated in numpy FP64 (fma taken as multiply-add: the bounds below
    hold with room for that one extra rounding per step)."""
    inv, ln = _log_table_entries()
    m, k = np.frexp(z)                                   # m in [0.5, 1)
    i = ((m.view(np.uint64) >> np.uint64(45)) & np.uint64(127)).astype(np.int64)
    r = m * inv[i] - 1.0                                 # fma(m, 1/c, -1) on the device
    p = r * 0.2 - 0.25
    p = p * r + 1.0 / 3.0
    p = p * r - 0.5
    l1 = (r * r) * p + r
    kd = k.astype(np.float64)
    ln2_hi, ln2_lo = float.fromhex("0x1.62e42ff000000p-1"), float.fromhex("-0x1.718432a1b0e26p-35")
    return kd * ln2_hi + (kd * ln2_lo + (ln[i] + l1))


def test_log_table_accuracy():
    """Against libm's log (within an ulp): relative error below 2e-11 over z in [2^-80, 1),
    densely near 1 (the small distances, where the last interval reduces to z - 1 exactly) and at
    every table interval's ends -- so the float distances it yields are the reference's except in
    ties within ~1e-9 of a float rounding boundary. z = 1 is not special-cased: the reduction
    (1 = 2 * 0.5) leaves a value below 1e-13; the DNA corrections select their exact result for
    every pair whose argument can be 1 (no differences, saturation), and the protein corrections
    call log_table_one, which returns 0 for 1."""
    z = np.concatenate([
        np.exp2(-RNG.random(200_000) * 80.0),
        1.0 - np.exp2(-RNG.random(200_000) * 52.0),
        0.5 + np.arange(1, 256) / 512.0,                 # interval ends and centres
        np.nextafter(0.5 + np.arange(1, 257) / 256.0, 0.0),
        np.array([0.5, 0.75, np.nextafter(1.0, 0.0)]),
    ])
    got = _log_table_np(z)
    want = np.log(z)
    rel = np.abs(got - want) / np.abs(want)
    assert rel.max() < 2e-11, (rel.max(), z[np.argmax(rel)])
    assert abs(_log_table_np(np.array([1.0]))[0]) < 1e-13
