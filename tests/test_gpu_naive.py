"""The naive character rules (the reference's `-f` path, simpleDistanceCalculator) through every
entry point that takes them: the row hand-offs, the disk-matrix row stream, the FASTA reader and
batch plans; and the entry points that refuse them.

With fastdist off the reference's DNA counter classifies a mismatch only when the row taxon's
character is uppercase A, C, G or T (simpleDistanceCalculator.cpp:58-98), so its counts are not
symmetric. The in-memory path counts j > i and mirrors; the disk-matrix path counts every (i, j)
with i the row and takes the maximum over the whole square (KimuraDistance.cpp:300-320, offset 0
with a disk matrix). tests/golden/naive_rows.npz holds both, written by the reference itself
(tests/golden/make_naive_rows_golden.py); larger cases are checked against the oracle's naive
counter (oracle/rnj_oracle.c), pinned to the reference's naive counts by tests/test_oracle.py.
Distances: 1e-6 relative, identical zero signs and sentinels.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import simulate_alignment

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = ("mixed", "sim")


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
def golden():
    return dict(np.load(os.path.join(GOLDEN, "naive_rows.npz"), allow_pickle=False))


def assert_distances_match(got, ref):
    got = np.asarray(got, dtype=np.float32)
    ref = np.asarray(ref, dtype=np.float32)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=0)
    np.testing.assert_array_equal(np.signbit(got), np.signbit(ref))


def naive_alignment(ctx, aln):
    ctx.set_char_rules(True)
    try:
        return ctx.alignment(aln)
    finally:
        ctx.set_char_rules(False)


def row_side_matrix(aln, model):
    """The disk path's post-processed square from the oracle's naive counter: entry (i, j)
    counted with taxon i as the row, maximum over every entry, -1 -> 2 * max, diagonal 0."""
    n = aln.shape[0]
    rows = [aln[i].tobytes() for i in range(n)]
    corr = O.k2p if model == "k2p" else O.jc
    m = np.zeros((n, n), dtype=np.float32)
    for i in range(n):
        for j in range(n):
            if i != j:
                m[i, j] = corr(*O.count_naive(rows[i], rows[j]))
    mx = np.float32(max(0.0, float(m[m > 0].max()) if (m > 0).any() else 0.0))
    m[m == -1] = mx * np.float32(2)
    np.fill_diagonal(m, 0)
    return m, mx


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("model", ["jc", "k2p"])
def test_stream_rows_match_reference_disk_path(ctx, golden, case, model):
    """rnj_distances_stream_rows under the naive rules gives the reference's disk-matrix rows:
    asymmetric where lowercase or U meets uppercase, with the maximum over the whole square."""
    aln = golden[f"{case}_aln"]
    A = naive_alignment(ctx, aln)
    want = golden[f"{case}_disk_{model}"]
    assert not np.array_equal(want, want.T)
    got, mx = ctx.distance_stream_rows(A, model)
    assert_distances_match(got, want)
    assert mx > 0


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("model", ["jc", "k2p"])
def test_rows_match_reference_memory_path(ctx, golden, case, model):
    """The in-memory hand-offs (full rows and rapidNJMem's lower triangle) stay symmetric, as the
    reference's in-memory path mirrors its upper triangle."""
    aln = golden[f"{case}_aln"]
    A = naive_alignment(ctx, aln)
    want = golden[f"{case}_mem_{model}"]
    mat, mx = ctx.distance_rows(A, model)
    assert_distances_match(mat, want)
    low, mx2 = ctx.distance_rows_lower(A, model)
    assert mx2 == mx
    for i, r in enumerate(low):
        np.testing.assert_array_equal(r.view(np.uint32), mat[i, : i + 1].view(np.uint32))


def test_stream_rows_every_config(ctx, N):
    """A larger mixed-case alignment through the row stream in every tile configuration (with and
    without split word ranges) against the oracle's row-side matrix."""
    aln = simulate_alignment(150, 900, 0.05, 23).copy()
    rng = np.random.default_rng(23)
    aln[rng.random(aln.shape) < 0.3] |= 0x20
    aln[((aln | 0x20) == ord("t")) & (rng.random(aln.shape) < 0.3)] += np.uint8(1)  # T/t -> U/u
    aln[0, :8] = np.frombuffer(b"ACGTacgt", np.uint8)
    A = naive_alignment(ctx, aln)
    for model in ("jc", "k2p"):
        want, mx_want = row_side_matrix(aln, model)
        assert not np.array_equal(want, want.T)
        try:
            for c in range(len(N.tile_configs())):
                ctx.set_tile_config(c)
                got, mx = ctx.distance_stream_rows(A, model)
                assert_distances_match(got, want)
                assert mx == pytest.approx(float(mx_want), rel=1e-6)
        finally:
            ctx.set_tile_config(-1)


def test_bit_rules_stream_rows_unchanged(ctx):
    """Bit-rule alignments stream the symmetric matrix the in-memory path hands over."""
    aln = simulate_alignment(200, 500, 0.05, 4)
    A = ctx.alignment(aln)
    got, mx = ctx.distance_stream_rows(A, "k2p")
    mat, mx2 = ctx.distance_rows(A, "k2p")
    np.testing.assert_array_equal(got.view(np.uint32), mat.view(np.uint32))
    assert mx == mx2


@pytest.mark.parametrize("case", CASES)
def test_fasta_reader_applies_naive_rules(ctx, golden, case, tmp_path):
    """rnj_alignment_from_fasta and _file pack under the naive rules when they are set, the
    reference's usual `-f` input route: counts equal the oracle's naive counter."""
    aln = golden[f"{case}_aln"]
    fa = b"".join(b">t%d\n" % i + aln[i].tobytes() + b"\n" for i in range(len(aln)))
    path = tmp_path / "in.fa"
    path.write_bytes(fa)
    n = len(aln)
    iu = np.triu_indices(n, 1)
    ref = np.array([O.count_naive(aln[i].tobytes(), aln[j].tobytes()) for i, j in zip(*iu)],
                   dtype=np.uint32).reshape(-1, 3)
    bits = np.array([O.count_chars(aln[i].tobytes(), aln[j].tobytes()) for i, j in zip(*iu)],
                    dtype=np.uint32).reshape(-1, 3)
    assert (bits != ref).any()
    ctx.set_char_rules(True)
    try:
        for A in (ctx.alignment_from_fasta(fa), ctx.alignment_from_fasta_file(str(path))):
            got = np.stack([t.cpu().numpy().view(np.uint32) for t in ctx.counts(A)], axis=1)
            np.testing.assert_array_equal(got, ref)
    finally:
        ctx.set_char_rules(False)
    A = ctx.alignment_from_fasta(fa)
    got = np.stack([t.cpu().numpy().view(np.uint32) for t in ctx.counts(A)], axis=1)
    np.testing.assert_array_equal(got, bits)


def test_batch_plan_of_naive_alignments(ctx, N, golden):
    """Naive-rule alignments in one captured batch: each matrix equals the reference's in-memory
    naive matrix; a batch mixing character rules is refused with a message that says so."""
    alns = [naive_alignment(ctx, golden[f"{c}_aln"]) for c in CASES]
    with ctx.batch_plan(alns, ["jc", "k2p"], layout="square") as plan:
        for _ in range(3):
            outs = plan.run()
            for c, model, out in zip(CASES, ("jc", "k2p"), outs):
                assert_distances_match(out.cpu().numpy(), golden[f"{c}_mem_{model}"])
    bit = ctx.alignment(golden["sim_aln"])
    with pytest.raises(N.RnjError, match="naive character rules"):
        ctx.batch_plan([alns[0], bit], "k2p")


def test_server_refuses_naive_rules(ctx, N):
    aln = simulate_alignment(20, 100, 0.0, 1)
    with ctx.server() as srv:
        ctx.set_char_rules(True)
        try:
            with pytest.raises(N.RnjError, match="bit-string rules"):
                srv.submit(aln)
        finally:
            ctx.set_char_rules(False)
        t = srv.submit(aln)
        tri, _ = srv.wait(t, 20)
        assert tri.shape == (190,)
