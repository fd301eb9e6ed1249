"""The protein oracle (oracle/rnj_oracle.c) against the reference's protein engine.

tests/golden/protein.npz holds, per case, what the reference's bitDistanceProtein returned for
every pair (data[0..2]) and its post-processed Kimura-protein and JC (alpha = 20) matrices
(tests/golden/make_protein_golden.py). The restatement must give the same u64 counts, including
the estimator's valid count that is 128 short (and wraps) when L is a multiple of 128, and the
same float matrices.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "protein.npz"), allow_pickle=False))


def cases(g):
    return sorted({k.rsplit("_", 1)[0] for k in g if k.endswith("_aln")})


def test_cases_present(golden):
    assert set(CASES) <= set(cases(golden))


CASES = ["pf15224", "p12x1", "p20x128", "p30x256", "p25x300", "p16x640", "p10x200gap"]


@pytest.mark.parametrize("case", CASES)
def test_counts_and_matrices(golden, case):
    aln = golden[f"{case}_aln"]
    d0, d2 = O.protein_counts_triangle(aln)
    np.testing.assert_array_equal(d0, golden[f"{case}_d0"])
    assert not golden[f"{case}_d1"].any()
    np.testing.assert_array_equal(d2, golden[f"{case}_d2"])
    for model, key in (("k2p", "kim"), ("jc", "jc")):
        full, _ = O.protein_matrix(aln, model)
        np.testing.assert_array_equal(full.view(np.uint32), golden[f"{case}_{key}"].view(np.uint32))


@pytest.mark.parametrize("case", CASES)
def test_hamming_counts_and_matrices(golden, case):
    """The -f path (fastdist off): simpleDistanceCalculator::calculateHammingDistance counts
    the residue pairs exactly, so where L is a multiple of 128 its valid count is 128 above
    bitDistanceProtein's, and elsewhere the two agree."""
    aln = golden[f"{case}_aln"]
    d0, d2 = O.protein_counts_triangle(aln, naive=True)
    np.testing.assert_array_equal(d0, golden[f"{case}_nd0"])
    assert not golden[f"{case}_nd1"].any()
    np.testing.assert_array_equal(d2, golden[f"{case}_nd2"])
    np.testing.assert_array_equal(golden[f"{case}_nd0"], golden[f"{case}_d0"])
    shift = np.uint64(128 if aln.shape[1] % 128 == 0 else 0)
    np.testing.assert_array_equal(golden[f"{case}_nd2"] - shift, golden[f"{case}_d2"])
    for model, key in (("k2p", "nkim"), ("jc", "njc")):
        full, _ = O.protein_matrix(aln, model, naive=True)
        np.testing.assert_array_equal(full.view(np.uint32), golden[f"{case}_{key}"].view(np.uint32))


def test_hamming_brute_force():
    """The Hamming counter on tiny rows against a direct count over the gap set."""
    rng = np.random.default_rng(7)
    alpha = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWYacdxX-.?zZbBJj*", dtype=np.uint8)
    gaps = set(b"-.XxzZbBJj?")
    for _ in range(200):
        L = int(rng.integers(1, 40))
        a = alpha[rng.integers(0, len(alpha), L)].tobytes()
        b = alpha[rng.integers(0, len(alpha), L)].tobytes()
        both = [(x, y) for x, y in zip(a, b) if x not in gaps and y not in gaps]
        assert O.count_protein_naive(a, b) == (sum(x != y for x, y in both), 0, len(both))


@pytest.mark.parametrize("case", CASES)
def test_dormant_kernel_counts_against_reference(golden, case):
    """The restatement of the reference's dormant protein kernel (a byte below 64 is a gap) on
    the rows dataloader::initGPUData stages: for pairs whose resolved rows hold no byte below 64
    but '-', its mismatches are bitDistanceProtein's data[0] and its gap positions paddedLength
    minus the Hamming counter's residue pairs, both written by the reference."""
    aln = golden[f"{case}_aln"]
    n, L = aln.shape
    staged, stride, P = O.protein_staging(aln)
    assert P == -(-L // 128) * 128 and stride * 4 >= P
    rows = staged.view(np.uint8).reshape(n, -1)[:, :P]
    res = aln.copy()
    res[np.isin(res, np.frombuffer(b"-.XxzZbBJj?", dtype=np.uint8))] = ord("-")
    clean = [not ((r < 64) & (r != ord("-"))).any() for r in res]
    k = checked = 0
    for i in range(n):
        for j in range(i + 1, n):
            r, gp = O.protein_gpu_counts(rows[i].tobytes(), rows[j].tobytes())
            assert 0 <= r + gp <= P
            if clean[i] and clean[j]:
                assert r == golden[f"{case}_d0"][k] and gp == P - golden[f"{case}_nd2"][k]
                checked += 1
            k += 1
    assert checked >= 3


def test_regimes_covered(golden):
    """The fixtures reach every branch: the Dayhoff table, the 10.0 cap, saturated and empty
    pairs, and the wrapped valid count."""
    kim = np.concatenate([golden[f"{c}_kim"].ravel() for c in cases(golden)])
    assert np.any(kim == np.float32(10.0))
    table = (kim >= 1.95) & (kim <= 9.88)
    assert np.any(table)
    d2 = np.concatenate([golden[f"{c}_d2"] for c in cases(golden)])
    assert np.any(d2 > np.uint64(1) << np.uint64(63))        # 128 short, wrapped
    assert np.any(d2 == 0)


def test_dayhoff_boundaries():
    """p just below 0.75 uses the formula, 0.75 the first table entry, 0.93 the last, above it 10."""
    assert O.kimura_protein(0, 10) == 0.0
    assert O.kimura_protein(749, 1000) == pytest.approx(-np.log(1 - 0.749 - 0.749 ** 2 * 0.2), rel=1e-6)
    assert O.kimura_protein(750, 1000) == np.float32(1.95)
    # 930/1000 in float is just above 0.930: capped. Below it the table is nondecreasing, from
    # 1.95 PAM/100 at p = 0.75; the formula just below 0.75 lands near the same value.
    assert O.kimura_protein(930, 1000) == np.float32(10.0)
    table = [O.kimura_protein(k, 1000) for k in range(750, 930)]
    assert all(a <= b for a, b in zip(table, table[1:]))
    assert 9.0 < table[-1] < 10.0
    assert 1.9 < O.kimura_protein(749, 1000) < 2.0
    assert O.kimura_protein(5, 0) == -1.0
    # 95/100 in float is just below 0.95: finite; above it saturated
    assert O.jc_protein(95, 100) > 17.0 and O.jc_protein(96, 100) == -1.0
    assert O.jc_protein(0, 0) == 0.0
