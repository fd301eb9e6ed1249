"""Pins the CPU oracle (oracle/rnj_oracle.c) before anything is checked against it.

Pins, in order of strength:
* golden vectors written by the reference engine itself (tests/golden/make_golden.py):
  packed words, ts/tv/valid, the naive -f counts and post-processed JC/K2P float bits;
* an independent character-level brute force (purine/pyrimidine classes, no bit tricks);
* the substitution model: distances recovered from expected JC/K2P pair frequencies;
* closed-form special cases (negative zero, empty pairs, saturation) and invariants.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import expand_columns, simulate_alignment

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FULL_CASES = sorted(glob.glob(os.path.join(GOLDEN, "edge_L*.npz"))) + [
    os.path.join(GOLDEN, f) for f in ("c1.npz", "lgt70.npz", "saturated.npz")]


def _load(path):
    return dict(np.load(path, allow_pickle=False))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "edge_L*.npz"))),
                         ids=os.path.basename)
def test_packing_matches_reference_words(path):
    d = _load(path)
    bits, gaps, bsc = O.pack(d["aln"])
    assert bsc == d["aln"].shape[1] // 64 + 6
    np.testing.assert_array_equal(bits, d["bits"])
    np.testing.assert_array_equal(gaps, d["gaps"])


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
def test_counts_match_reference(path):
    d = _load(path)
    ts, tv, va = O.counts_triangle(d["aln"], threads=4)
    np.testing.assert_array_equal(ts, d["ts"])
    np.testing.assert_array_equal(tv, d["tv"])
    np.testing.assert_array_equal(va, d["valid"])


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
def test_counters_agree_with_char_brute_force(path):
    d = _load(path)
    aln = d["aln"]
    bits, gaps, _ = O.pack(aln)
    n = aln.shape[0]
    k = 0
    for i in range(n):
        for j in range(i + 1, n):
            ref = (int(d["ts"][k]), int(d["tv"][k]), int(d["valid"][k]))
            k += 1
            if n > 40 and (i * 7 + j) % 11:
                continue  # sample large cases
            assert O.count_chars(aln[i].tobytes(), aln[j].tobytes()) == ref
            assert O.count_words(bits, gaps, i, j) == ref
            assert O.count_swar(bits, gaps, i, j) == ref


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
def test_naive_counter_matches_reference(path):
    d = _load(path)
    aln = d["aln"]
    n = aln.shape[0]
    iu = np.triu_indices(n, 1)
    got = np.array([O.count_naive(aln[i].tobytes(), aln[j].tobytes()) for i, j in zip(*iu)],
                   dtype=np.int64).reshape(-1, 3)
    np.testing.assert_array_equal(got[:, 0], d["naive_ts"])
    np.testing.assert_array_equal(got[:, 1], d["naive_tv"])
    np.testing.assert_array_equal(got[:, 2], d["naive_valid"])


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
@pytest.mark.parametrize("model", ["jc", "k2p"])
def test_matrix_float_bits_match_reference(path, model):
    d = _load(path)
    _, full, _ = O.distance_matrix(d["aln"], model, threads=3)
    np.testing.assert_array_equal(full.view(np.uint32), d[model].view(np.uint32))


def test_uppercase_acgt_naive_equals_packed():
    aln = simulate_alignment(30, 700, 0.05, 21)
    for i in range(0, 30, 3):
        for j in range(i + 1, 30, 4):
            a, b = aln[i].tobytes(), aln[j].tobytes()
            assert O.count_naive(a, b) == O.count_chars(a, b)


def test_bootstrap_counts_equal_expanded_alignment():
    d = _load(os.path.join(GOLDEN, "bootstrap.npz"))
    for r in range(d["weights"].shape[0]):
        ts, tv, va = O.counts_triangle(expand_columns(d["aln"], d["weights"][r]))
        np.testing.assert_array_equal(ts, d[f"ts{r}"])
        np.testing.assert_array_equal(tv, d[f"tv{r}"])
        np.testing.assert_array_equal(va, d[f"valid{r}"])


def _k2p_expected(dist, kappa):
    """Expected transition / transversion fractions at K2P distance `dist` (alpha + 2 beta = 1)."""
    beta = 1.0 / (kappa + 2.0)
    alpha = kappa * beta
    P = 0.25 + 0.25 * math.exp(-4 * beta * dist) - 0.5 * math.exp(-2 * (alpha + beta) * dist)
    Q = 0.5 - 0.5 * math.exp(-4 * beta * dist)
    return P, Q


@pytest.mark.parametrize("dist", [0.001, 0.02, 0.1, 0.5, 1.0, 2.0])
@pytest.mark.parametrize("kappa", [0.5, 1.0, 2.0, 5.0])
def test_k2p_inverts_the_substitution_model(dist, kappa):
    V = 1 << 23
    P, Q = _k2p_expected(dist, kappa)
    d = O.k2p(round(P * V), round(Q * V), V)
    assert d == pytest.approx(dist, rel=2e-4, abs=1e-6)


@pytest.mark.parametrize("dist", [0.001, 0.05, 0.3, 1.0, 2.5])
def test_jc_inverts_the_substitution_model(dist):
    V = 1 << 23
    p = 0.75 * (1.0 - math.exp(-4.0 * dist / 3.0))
    m = round(p * V)
    assert O.jc(m // 3, m - m // 3, V) == pytest.approx(dist, rel=2e-4, abs=1e-6)
    # JC uses only ts + tv: how they split must not matter
    assert O.jc(m, 0, V) == O.jc(0, m, V) == O.jc(m // 3, m - m // 3, V)


def test_correction_special_cases():
    def bits(x):
        return np.float32(x).view(np.uint32)
    negzero = np.float32(-0.0).view(np.uint32)
    assert bits(O.k2p(0, 0, 1000)) == negzero        # identical pair -> -0.0f in both models
    assert bits(O.jc(0, 0, 1000)) == negzero
    assert bits(O.jc(5, 7, 0)) == negzero            # empty pair: JC gives p = 0
    assert O.k2p(0, 0, 0) == -1.0                    # empty pair: K2P sentinel
    assert O.k2p(250, 500, 1000) == -1.0             # 1 - 2P - Q = 0 -> saturated
    assert O.k2p(0, 500, 1000) == -1.0               # 1 - 2Q = 0 -> saturated
    assert O.jc(750, 0, 1000) == -1.0                # p / 0.75 = 1 -> saturated
    assert O.jc(749, 0, 1000) > 0
    # ts only: the tv term vanishes
    assert O.k2p(100, 0, 1000) == np.float32(-0.5 * math.log(1 - 0.2))


def test_postprocess_invariants():
    d = _load(os.path.join(GOLDEN, "saturated.npz"))
    for model in ("jc", "k2p"):
        tri, full, mx = O.distance_matrix(d["aln"], model)
        n = full.shape[0]
        assert np.array_equal(full, full.T)
        assert np.all(full.diagonal().view(np.uint32) == 0)
        iu = np.triu_indices(n, 1)
        sent = tri == -1
        assert sent.any()
        assert mx == tri[tri > 0].max()
        assert np.all(full[iu][sent] == np.float32(2) * np.float32(mx))
        np.testing.assert_array_equal(O.triangle_to_full(tri, n).view(np.uint32),
                                      full.view(np.uint32))


def test_invariants_on_random_alignment():
    aln = simulate_alignment(60, 333, 0.1, 5)
    ts, tv, va = O.counts_triangle(aln)
    assert np.all(ts.astype(np.int64) + tv <= va)
    assert np.all(va <= 333)
    twin = np.vstack([aln[:1], aln[:1]])
    t2, v2, a2 = O.counts_triangle(twin)
    assert t2[0] == v2[0] == 0 and a2[0] == np.sum(aln[0] != ord("-"))


def test_type_detection():
    assert O.is_dna(b"ACGTNU-acgtnu.?")
    assert not O.is_dna(b"ACGTR")
    assert O.is_dna(bytes([200, 65]))   # high bytes are negative chars: count as DNA
