"""The neighbour-joining oracle (oracle/rnj_oracle_nj.c) against the reference's own trees.

tests/golden/nj.npz holds distance matrices computed by the reference and the Newick text its
rapidNJ builds on them (tests/golden/make_nj_golden.py), with negative_branches off and on;
tests/golden/lgt70.npz holds the reference's trees of its own LGT_70 test alignment. The
oracle's text must be byte-identical: same joins in the same order, same branch lengths to the
printed precision, same layout.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def nj_golden():
    return dict(np.load(os.path.join(GOLDEN, "nj.npz"), allow_pickle=False))


def matrix_of(g, case):
    if f"{case}_matrix" in g:
        return g[f"{case}_matrix"]
    return O.distance_matrix(g[f"{case}_aln"], "k2p")[1]  # the reference's K2P bits


CASES = ["dup", "short", "sat", "sim300", "sim300jc", "sim800"]
# tie-heavy matrices: equal criterion and distance within rows (std::sort's order decides) and
# rows whose scan bound ends them early (make_nj_golden.py)
TIE_CASES = ["eq30", "eq64", "eq200", "eqc120", "int30", "int300", "block60", "block140"]


@pytest.mark.parametrize("case", CASES + TIE_CASES)
@pytest.mark.parametrize("neg", [False, True])
def test_reference_trees(nj_golden, case, neg):
    m = matrix_of(nj_golden, case)
    want = bytes(nj_golden[f"{case}_newick_neg" if neg else f"{case}_newick"])
    assert O.newick(m.shape[0], O.nj(m, neg)) == want


@pytest.mark.parametrize("model", ["k2p", "jc"])
def test_lgt70(model):
    d = np.load(os.path.join(GOLDEN, "lgt70.npz"), allow_pickle=False)
    m = d[model]
    assert O.newick(m.shape[0], O.nj(m)).rstrip(b"\n") == bytes(d[f"newick_{model}"])


def test_redundant_copies_join_at_zero_length(nj_golden):
    """The fixture has identical rows, and copies are joined to each other with both branches of
    length 0 (equal separations, distance 0)."""
    m = nj_golden["dup_matrix"]
    left, right, llen, rlen, last, _ = O.nj(m)
    assert len({tuple(r) for r in m}) < m.shape[0]
    assert np.any((llen == 0) & (rlen == 0))


def test_tie_order_is_not_index_order(nj_golden):
    """The all-equal fixtures are ones where taking tied partners in increasing index order gives
    another tree: the reference's text differs from a star joined in index order. A restatement
    that ignored std::sort's order would fail here (checked on the printed first join)."""
    text = bytes(nj_golden["eq64_newick"])
    m = nj_golden["eq64_matrix"]
    assert O.newick(64, O.nj(m)) == text
    assert len(set(np.unique(m[np.triu_indices(64, 1)]))) == 1


def test_join_invariants():
    """Every id is joined exactly once; new clusters are numbered n, n+1, ..."""
    rng = np.random.default_rng(3)
    n = 40
    x = rng.random((n, 5))
    m = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)).astype(np.float32)
    left, right, llen, rlen, last, _ = O.nj(m)
    ids = list(left) + list(right) + list(last)
    assert sorted(ids) == list(range(2 * n - 2))
    for k in range(n - 2):
        assert left[k] < n + k and right[k] < n + k


# ---- the device NJ's row skipping (rapidnj_amd/csrc/nj.hip), audited along rapidNJ's runs ----
SKIP_CASES = ["dup", "short", "sat", "sim300", "sim300jc", "eq200", "eqc120", "int30", "int300",
              "block60", "block140"]


@pytest.mark.parametrize("case", SKIP_CASES)
@pytest.mark.parametrize("neg", [False, True])
def test_row_skipping_rule_holds(nj_golden, case, neg):
    """At every search of rapidNJ's own run, the device's bound of every row (its last
    evaluation's smallest float(d - sep), less the separations' drift since, less its own
    separation and the margin) is at most each of its current criteria, and no row the rule
    skips is ever a record."""
    r = O.nj_skip_audit(nj_golden[f"{case}_matrix"], neg)
    assert r["bound_violations"] == 0 and r["record_violations"] == 0, r


def test_row_skipping_skips_and_the_audit_catches_a_missing_drift(nj_golden):
    """The rule skips most rows of a simulated matrix, and a bound without the drift term (a
    plausible mistake, planted in the oracle's copy of the rule) breaks both claims."""
    m = nj_golden["sim300_matrix"]
    r = O.nj_skip_audit(m)
    assert r["skipped"] > 3 * r["scanned"]
    bad = O.nj_skip_audit(m, drop_drift=True)
    assert bad["bound_violations"] > 0 and bad["record_violations"] > 0


def test_row_skipping_rule_on_simulated_alignments():
    """Larger and lower-divergence matrices from the simulator, K2P and JC."""
    from rapidnj_amd.workloads import simulate_alignment
    for n, L, seed, model in [(400, 800, 31, "k2p"), (350, 3000, 32, "jc"), (300, 120, 33, "k2p")]:
        aln = simulate_alignment(n, L, 0.04, seed)
        _, m, _ = O.distance_matrix(aln, model)
        r = O.nj_skip_audit(m)
        assert r["bound_violations"] == 0 and r["record_violations"] == 0, (n, L, r)
        assert r["skipped"] > 0
