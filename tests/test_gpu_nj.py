"""Neighbour joining on the device (rnj_nj_device / rnj_nj_alignment) against the reference's
own trees and the oracle.

tests/golden/nj.npz and lgt70.npz hold matrices the reference computed and the Newick text its
rapidNJ built on them (tests/golden/make_nj_golden.py): the device text must be byte-identical,
with negative_branches off and on, across redundant copies, equal distances, saturated pairs and
slot-matrix compactions. Larger trees are compared join by join with the oracle
(oracle/rnj_oracle_nj.c), itself pinned to those trees by tests/test_nj_oracle.py.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import simulate_alignment

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def N():
    from rapidnj_amd import native
    return native


@pytest.fixture(scope="module")
def ctx(N):
    c = N.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def nj_golden():
    return dict(np.load(os.path.join(GOLDEN, "nj.npz"), allow_pickle=False))


def device(m):
    import torch
    return torch.from_numpy(np.ascontiguousarray(m, dtype=np.float32)).to("cuda:0")


def same_joins(tree, ref):
    left, right, ll, rl, last, ld = tree.joins()
    rleft, rright, rll, rrl, rlast, rld = ref
    np.testing.assert_array_equal(left, rleft)
    np.testing.assert_array_equal(right, rright)
    np.testing.assert_array_equal(ll.view(np.uint64), rll.view(np.uint64))
    np.testing.assert_array_equal(rl.view(np.uint64), rrl.view(np.uint64))
    np.testing.assert_array_equal(last, rlast)
    assert np.float32(ld).tobytes() == np.float32(rld).tobytes()


CASES = ["dup", "short", "sat", "sim300", "sim300jc"]
# tie-heavy matrices (make_nj_golden.py): equal criterion and distance within a row, where the
# reference takes std::sort's order, and rows its rounded scan bound ends early
TIE_CASES = ["eq30", "eq64", "eq200", "eqc120", "int30", "int300", "block60", "block140"]


@pytest.mark.parametrize("case", CASES + TIE_CASES)
@pytest.mark.parametrize("neg", [False, True])
def test_reference_trees(ctx, nj_golden, case, neg):
    m = nj_golden[f"{case}_matrix"]
    want = bytes(nj_golden[f"{case}_newick_neg" if neg else f"{case}_newick"])
    print(f"nj {case} n={m.shape[0]} neg={neg}", flush=True)
    tree = ctx.nj(device(m), negative_branches=neg)
    assert tree.newick() == want
    same_joins(tree, O.nj(m, neg))
    if case in TIE_CASES:
        assert tree.host_searches > 0  # the host settled the tied or cut-short searches
    else:
        assert tree.host_searches == 0


def _blocks(n, groups, seed, a=0.2, b=0.7):
    lab = np.random.default_rng(seed).integers(0, groups, n)
    m = np.where(lab[:, None] == lab[None, :], np.float32(a), np.float32(b)).astype(np.float32)
    np.fill_diagonal(m, 0)
    return m


def _integers(n, hi, seed):
    a = np.random.default_rng(seed).integers(1, hi + 1, size=(n, n)).astype(np.float32)
    m = np.tril(a, -1)
    return m + m.T


@pytest.mark.parametrize("case", ["eq1500", "blocks1200", "int1000", "int900dup"])
def test_tie_heavy_larger(ctx, case):
    """Join-by-join equality with the oracle (rapidNJ's sorted lists and scan bound, restated)
    on tie-heavy matrices large enough for many slot-matrix compactions and host searches in
    every enqueue window."""
    if case == "eq1500":
        m = np.ones((1500, 1500), dtype=np.float32)
        np.fill_diagonal(m, 0)
    elif case == "blocks1200":
        m = _blocks(1200, 6, 11)
    elif case == "int1000":
        m = _integers(1000, 4, 12)
    else:  # integers with identical rows (redundant copies)
        m = _integers(900, 3, 13)
        for i, j in ((5, 17), (40, 17), (899, 3), (300, 301)):
            m[i] = m[j]
            m[:, i] = m[:, j]
        np.fill_diagonal(m, 0)
    print(f"nj {case}", flush=True)
    tree = ctx.nj(device(m))
    same_joins(tree, O.nj(m))
    if case != "int900dup":  # (its ties never reach a winning row)
        assert tree.host_searches > 0


@pytest.mark.parametrize("model", ["k2p", "jc"])
def test_lgt70(ctx, model):
    d = np.load(os.path.join(GOLDEN, "lgt70.npz"), allow_pickle=False)
    tree = ctx.nj(device(d[model]))
    assert tree.newick().rstrip(b"\n") == bytes(d[f"newick_{model}"])
    # the whole pipeline from the alignment, matrix kept in HBM
    A = ctx.alignment(d["aln"])
    assert ctx.nj_alignment(A, model).newick().rstrip(b"\n") == bytes(d[f"newick_{model}"])


def test_alignment_pipeline_800(ctx, nj_golden):
    A = ctx.alignment(nj_golden["sim800_aln"])
    for neg in (False, True):
        want = bytes(nj_golden["sim800_newick_neg" if neg else "sim800_newick"])
        assert ctx.nj_alignment(A, "k2p", negative_branches=neg).newick() == want


@pytest.mark.parametrize("n,L,seed", [(1500, 600, 7), (700, 60, 8)])
def test_oracle_larger(ctx, n, L, seed):
    """Join-by-join equality with the oracle on larger trees (many compactions; the short
    alignment has many equal distances)."""
    aln = simulate_alignment(n, L, 0.03, seed)
    aln[5] = aln[17]          # redundant copies
    aln[n - 1] = aln[17]
    _, m, _ = O.distance_matrix(aln, "k2p")
    same_joins(ctx.nj(device(m)), O.nj(m))


@pytest.mark.parametrize("n,L,gap,seed,model,dups", [
    (600, 80, 0.00, 41, "k2p", 0), (900, 400, 0.05, 42, "jc", 3), (1200, 2000, 0.08, 43, "k2p", 5),
    (1000, 60, 0.02, 44, "jc", 0), (1300, 700, 0.03, 45, "k2p", 8)])
def test_row_skipping_join_by_join(ctx, n, L, gap, seed, model, dups):
    """Trees whose scans skip most rows (simulated alignments of several lengths, with
    redundant copies), with many compactions: joins equal to the oracle's, one by one."""
    aln = simulate_alignment(n, L, gap, seed)
    for k in range(dups):
        aln[(7 * k + 3) % n] = aln[(11 * k + 5) % n]
    _, m, _ = O.distance_matrix(aln, model)
    same_joins(ctx.nj(device(m)), O.nj(m))


@pytest.mark.skipif(not O.ref_harness_available(), reason="reference harness not built")
def test_5000_taxa_same_newick_as_reference_rapidnj(ctx, tmp_path):
    """A 5000-taxa K2P matrix (simulated 5000 x 1500 alignment, the oracle's post-processed
    distances): the device tree's Newick text equals, byte for byte, the one the reference's own
    rapidNJ (compiled from its sources) writes for the same matrix."""
    import subprocess
    n = 5000
    aln = simulate_alignment(n, 1500, 0.05, 21)
    _, m, _ = O.distance_matrix(aln, "k2p")
    mf, nwk = tmp_path / "m.bin", tmp_path / "t.nwk"
    with open(mf, "wb") as fh:
        np.array([n], dtype=np.uint32).tofile(fh)
        np.ascontiguousarray(m, dtype=np.float32).tofile(fh)
    subprocess.run([O.REF_HARNESS, "nj", str(mf), str(nwk)], check=True, capture_output=True,
                   timeout=120)
    tree = ctx.nj(device(m))
    assert tree.newick() == nwk.read_bytes()


def test_strided_matrix_and_names(ctx):
    import torch
    d = np.load(os.path.join(GOLDEN, "lgt70.npz"), allow_pickle=False)
    m = d["k2p"]
    n = m.shape[0]
    big = torch.full((n, n + 37), 7.0, dtype=torch.float32, device="cuda:0")
    big[:, :n] = device(m)
    tree = ctx.nj(big)
    assert tree.newick().rstrip(b"\n") == bytes(d["newick_k2p"])
    names = [b"seq_%d" % i for i in range(n)]
    text = tree.newick(names)
    assert b"'seq_5'" in text and b"'t5'" not in text


def test_fasta_names(ctx):
    aln = simulate_alignment(12, 200, 0.0, 9)
    fa = b"".join(b">taxon_%d desc\n" % i + r.tobytes() + b"\n" for i, r in enumerate(aln))
    A = ctx.alignment_from_fasta(fa)
    text = ctx.nj_alignment(A, "k2p").newick()
    for i in range(12):  # a FASTA name runs to the first byte below 32, spaces included
        assert b"'taxon_%d desc'" % i in text


def test_small_and_invalid(ctx, N):
    m = np.array([[0, 1, 2], [1, 0, 3], [2, 3, 0]], dtype=np.float32)
    tree = ctx.nj(device(m))
    assert tree.newick() == O.newick(3, O.nj(m))
    with pytest.raises(N.RnjError):
        ctx.nj(device(np.zeros((2, 2), dtype=np.float32)))


DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "_build", "nj_test")
ADAPTER = os.path.join(os.path.dirname(__file__), "cpp", "_build", "adapter_test")


@pytest.mark.parametrize("case", ["dup", "sat", "sim300"])
@pytest.mark.parametrize("neg", [False, True])
def test_cpp_rapidNJ_class(tmp_path, nj_golden, case, neg):
    """rapidnj_amd::rapidNJ<Reader>(reader, n, neg, pb)->run()->serialize_tree(out), the
    reference's call sequence, over host rows (rnj_nj_rows)."""
    import subprocess
    m = nj_golden[f"{case}_matrix"]
    mf, out = tmp_path / "m.bin", tmp_path / "t.nwk"
    with open(mf, "wb") as fh:
        np.array([m.shape[0]], dtype=np.uint32).tofile(fh)
        m.astype(np.float32).tofile(fh)
    r = subprocess.run([DRIVER, str(mf), str(out)] + (["neg"] if neg else []), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = bytes(nj_golden[f"{case}_newick_neg" if neg else f"{case}_newick"])
    assert out.read_bytes() == want


def test_host_rows(ctx, nj_golden):
    m = nj_golden["sim300_matrix"]
    assert ctx.nj_rows(m).newick() == bytes(nj_golden["sim300_newick"])


@pytest.mark.parametrize("model", ["kim", "jc"])
def test_cpp_tree_from_loader(tmp_path, model):
    """rapidnj_amd::NJFromAlignment on a loader (LGT_70): distances and joins on the device."""
    import subprocess
    d = dict(np.load(os.path.join(GOLDEN, "lgt70.npz"), allow_pickle=False))
    aln = np.ascontiguousarray(d["aln"], dtype=np.uint8)
    n, L = aln.shape
    bits, gaps, _ = O.pack(aln)
    inp, out = tmp_path / "in.bin", tmp_path / "t.nwk"
    with open(inp, "wb") as f:
        np.array([n, L, bits.shape[1] // 4], dtype=np.uint32).tofile(f)
        aln.tofile(f)
        np.ascontiguousarray(bits, dtype=np.uint32).tofile(f)
        np.ascontiguousarray(gaps, dtype=np.uint32).tofile(f)
    r = subprocess.run([ADAPTER, str(inp), str(out), model, "tree", "fast"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    key = "newick_k2p" if model == "kim" else "newick_jc"
    assert out.read_bytes().rstrip(b"\n") == bytes(d[key])
