"""Protein distances on the device against the reference's protein engine and the oracle.

Golden vectors (tests/golden/protein.npz, made by tests/golden/make_protein_golden.py from the
reference's bitDistanceProtein and its Kimura-protein / JC post-processing) are matched exactly:
mismatch and valid counts bit for bit (valid as the low 32 bits of the reference's 64-bit count,
which is 128 short when L is a multiple of 128), distances bit for bit with the library log
(exact=True) and within 1e-6 relative with the default log. Larger alignments, every code width
(5, 6 and 8 code planes), the split-and-finish path, every tile configuration and the FASTA
reader are checked against the oracle (oracle/rnj_oracle.c), itself pinned to the golden vectors
by tests/test_protein_oracle.py.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["pf15224", "p12x1", "p20x128", "p30x256", "p25x300", "p16x640", "p10x200gap"]
AMINO = b"ACDEFGHIKLMNPQRSTVWY"
GAPS = b"-.XxzZbBJj?"


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
    return dict(np.load(os.path.join(GOLDEN, "protein.npz"), allow_pickle=False))


def protein_alignment(n, L, seed, alphabet=AMINO, gap=0.1, rates=(0.0, 0.9)):
    """Mutated copies of a random root over `alphabet`, with gap characters at rate `gap`; the
    first row starts with residues that make the reference's discovery say protein."""
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(alphabet, dtype=np.uint8)
    gaps = np.frombuffer(GAPS, dtype=np.uint8)
    root = alpha[rng.integers(0, len(alpha), L)]
    out = np.empty((n, L), dtype=np.uint8)
    for i, r in enumerate(np.linspace(rates[0], rates[1], n)):
        row = root.copy()
        m = rng.random(L) < r
        row[m] = alpha[rng.integers(0, len(alpha), int(m.sum()))]
        g = rng.random(L) < gap
        row[g] = gaps[rng.integers(0, len(gaps), int(g.sum()))]
        out[i] = row
    out[0, : min(L, 2)] = np.frombuffer(b"EF"[: min(L, 2)], dtype=np.uint8)
    return out


def square(ctx, A, model, exact):
    return ctx.distances(A, model, layout="square", exact=exact).cpu().numpy()


def check_counts(ctx, A, aln):
    d0, d2 = O.protein_counts_triangle(aln)
    ts, tv, valid = (t.cpu().numpy().view(np.uint32) for t in ctx.counts(A))
    np.testing.assert_array_equal(ts, d0.astype(np.uint32))
    assert not tv.any()
    np.testing.assert_array_equal(valid, (d2 & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def check_matrices(ctx, A, aln, exact_only=False):
    for model in ("k2p", "jc"):
        want, _ = O.protein_matrix(aln, model)
        got = square(ctx, A, model, True)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        if not exact_only:
            got = square(ctx, A, model, False)
            np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)


@pytest.mark.parametrize("case", CASES)
def test_golden_counts(ctx, golden, case):
    aln = golden[f"{case}_aln"]
    A = ctx.alignment(aln)
    assert A.is_protein
    ts, tv, valid = (t.cpu().numpy().view(np.uint32) for t in ctx.counts(A))
    np.testing.assert_array_equal(ts, golden[f"{case}_d0"].astype(np.uint32))
    assert not tv.any()
    np.testing.assert_array_equal(valid, (golden[f"{case}_d2"] & np.uint64(0xFFFFFFFF)).astype(np.uint32))


@pytest.mark.parametrize("case", CASES)
def test_golden_distances(ctx, golden, case):
    aln = golden[f"{case}_aln"]
    A = ctx.alignment(aln)
    for model, key in (("k2p", "kim"), ("jc", "jc")):
        want = golden[f"{case}_{key}"]
        got = square(ctx, A, model, True)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        got = square(ctx, A, model, False)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
        assert (np.signbit(got) == np.signbit(want)).all()


@pytest.mark.parametrize("case", CASES)
def test_golden_hamming(ctx, golden, case):
    """Under the naive character rules (the reference's -f path) protein alignments count as
    simpleDistanceCalculator::calculateHammingDistance: the exact residue pairs, with no
    128-position shortfall, for the counts and both models."""
    aln = golden[f"{case}_aln"]
    ctx.set_char_rules(True)
    try:
        A = ctx.alignment(aln)
    finally:
        ctx.set_char_rules(False)
    assert A.is_protein
    ts, tv, valid = (t.cpu().numpy().view(np.uint32) for t in ctx.counts(A))
    np.testing.assert_array_equal(ts, golden[f"{case}_nd0"].astype(np.uint32))
    assert not tv.any()
    np.testing.assert_array_equal(valid, golden[f"{case}_nd2"].astype(np.uint32))
    for model, key in (("k2p", "nkim"), ("jc", "njc")):
        want = golden[f"{case}_{key}"]
        got = square(ctx, A, model, True)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        got = square(ctx, A, model, False)
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
        assert (np.signbit(got) == np.signbit(want)).all()


def _dormant_protein_abi(N, staged, n, stride, P):
    """initialiseProtein_gpu / computeDistancesProtein_gpu / getResultsProtein_gpu through the C
    ABI, as KimuraDistance::computeDistancesProteinGPU calls them (KimuraDistance.cpp:113-128)."""
    L = N.lib()
    dist = np.full(n * n, 0xDEADBEEF, dtype=np.uint32)
    gaps = np.full(n * n, 0xDEADBEEF, dtype=np.uint32)
    staged = np.ascontiguousarray(staged, dtype=np.uint32)
    L.initialiseProtein_gpu(n, stride, P)
    L.computeDistancesProtein_gpu(staged.ctypes.data, dist.ctypes.data, gaps.ctypes.data)
    L.getResultsProtein_gpu(dist.ctypes.data, gaps.ctypes.data)
    return dist.reshape(n, n), gaps.reshape(n, n)


@pytest.mark.parametrize("case", CASES)
def test_reference_protein_gpu_abi(N, golden, case):
    """The reference's dormant protein ABI: per pair the mismatches and the gap positions among
    the paddedLength staged characters (bytes below 64 are gaps), equal to the oracle's
    restatement of computeDistanceProtein on every pair, and for rows free of other bytes below 64
    to the reference-written bitDistanceProtein and Hamming counts."""
    aln = golden[f"{case}_aln"]
    n = aln.shape[0]
    staged, stride, P = O.protein_staging(aln)
    dist, gaps = _dormant_protein_abi(N, staged, n, stride, P)
    rows = staged.view(np.uint8).reshape(n, -1)[:, :P]
    iu = np.triu_indices(n, 1)
    want = np.array([O.protein_gpu_counts(rows[i].tobytes(), rows[j].tobytes())
                     for i, j in zip(*iu)], dtype=np.uint32)
    np.testing.assert_array_equal(dist[iu], want[:, 0])
    np.testing.assert_array_equal(gaps[iu], want[:, 1])
    np.testing.assert_array_equal(dist, dist.T)
    np.testing.assert_array_equal(gaps, gaps.T)
    assert not np.diag(dist).any() and not np.diag(gaps).any()
    if case == "pf15224":  # letters and '-' only: the reference's own CPU counts
        np.testing.assert_array_equal(dist[iu], golden[f"{case}_d0"].astype(np.uint32))
        np.testing.assert_array_equal(gaps[iu], (P - golden[f"{case}_nd2"]).astype(np.uint32))


def test_reference_protein_gpu_abi_large(N):
    """A 500-taxon, 3000-site alignment with every gap character, '*' and lowercase through the
    same ABI, against the oracle's restatement (sampled rows)."""
    aln = protein_alignment(500, 3000, seed=77, alphabet=AMINO + b"*acd", gap=0.1)
    staged, stride, P = O.protein_staging(aln)
    dist, gaps = _dormant_protein_abi(N, staged, 500, stride, P)
    rows = staged.view(np.uint8).reshape(500, -1)[:, :P]
    for i in (0, 1, 250, 498):
        for j in range(i + 1, 500, 7):
            assert (dist[i, j], gaps[i, j]) == O.protein_gpu_counts(rows[i].tobytes(),
                                                                     rows[j].tobytes())


@pytest.mark.parametrize("n,L", [(300, 1000), (257, 1024), (130, 4096 + 77), (64, 33)])
def test_oracle_sizes(ctx, n, L):
    aln = protein_alignment(n, L, seed=n + L)
    A = ctx.alignment(aln)
    assert A.info.nplanes == 6
    check_counts(ctx, A, aln)
    check_matrices(ctx, A, aln)


@pytest.mark.parametrize("width,alphabet", [
    (5, AMINO + b"UOB*"),                                     # 23 residues
    (6, AMINO + AMINO.lower().replace(b"x", b"") + b"*UO"),  # 43 residues
    (8, bytes(v for v in range(256) if v not in GAPS and v != ord('E'))),  # 244 residues
], ids=["w5", "w6", "w8"])
def test_code_widths(ctx, width, alphabet):
    aln = protein_alignment(150, 700, seed=width, alphabet=alphabet, gap=0.05)
    A = ctx.alignment(aln)
    assert A.is_protein and A.info.nplanes == 1 + width
    check_counts(ctx, A, aln)
    check_matrices(ctx, A, aln, exact_only=True)


def test_planes_layout(ctx):
    aln = protein_alignment(70, 100, seed=5, gap=0.2)
    A = ctx.alignment(aln)
    P = A.planes()
    assert P.shape[0] == 6 and P.shape[2] == 70
    gap = np.isin(aln, np.frombuffer(GAPS, dtype=np.uint8))
    def bits(p, t):
        return np.unpackbits(np.ascontiguousarray(P[p, :, t]).view(np.uint8), bitorder="little")

    for t in range(70):
        np.testing.assert_array_equal(bits(0, t)[:100].astype(bool), gap[t])
        assert not bits(0, t)[100:].any()
    # codes: equal bytes <-> equal codes, gaps code 0
    codes = np.zeros((70, 100), dtype=np.uint32)
    for p in range(5):
        b = np.stack([bits(1 + p, t)[:100] for t in range(70)])
        codes |= b.astype(np.uint32) << p
    assert not codes[gap].any()
    assert (codes[~gap] > 0).all()
    for v in np.unique(aln[~gap]):
        assert len(np.unique(codes[aln == v])) == 1


def test_split_finish_path(ctx):
    """Few taxa, long sequences: the launch cuts the words into ranges finished separately."""
    aln = protein_alignment(40, 30000, seed=11)
    A = ctx.alignment(aln)
    check_counts(ctx, A, aln)
    check_matrices(ctx, A, aln)


def test_all_tile_configs(ctx, N):
    aln = protein_alignment(200, 900, seed=12, gap=0.15)
    want, _ = O.protein_matrix(aln, "k2p")
    A = ctx.alignment(aln)
    try:
        for c in range(N.lib().rnj_tile_config_count()):
            ctx.set_tile_config(c)
            got = square(ctx, A, "k2p", True)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"config {c}")
    finally:
        ctx.set_tile_config(-1)


def test_saturated_and_empty_pairs(ctx):
    """Kimura's -1 for pairs without common residues (patched to 2 * max), the Dayhoff table and
    the 10.0 cap, JC's saturation; then a multiple-of-128 length, where the reference's valid
    count of such pairs wraps instead."""
    for L in (300, 256):
        aln = protein_alignment(30, L, seed=13, rates=(0.0, 1.0), gap=0.0)
        aln[4] = ord('-')
        aln[7, : L - 6] = ord('X')
        A = ctx.alignment(aln)
        check_counts(ctx, A, aln)
        check_matrices(ctx, A, aln)
        raw = ctx.distances(A, "k2p", raw=True, exact=True).cpu().numpy()
        assert (raw == 10.0).any() and ((raw >= 1.95) & (raw < 10.0)).any()
        assert (raw == -1.0).any() == (L % 128 != 0)
        raw = ctx.distances(A, "jc", raw=True, exact=True).cpu().numpy()
        assert (raw == -1.0).any()


def test_fasta_and_rows(ctx, golden):
    aln = golden["pf15224_aln"]
    fa = b"".join(b">seq%d\n" % i + row.tobytes() + b"\n" for i, row in enumerate(aln))
    A = ctx.alignment_from_fasta(fa)
    assert A.is_protein and A.names[3] == b"seq3"
    got = square(ctx, A, "k2p", True)
    np.testing.assert_array_equal(got.view(np.uint32), golden["pf15224_kim"].view(np.uint32))
    mat, mx = ctx.distance_rows(A, "k2p")
    np.testing.assert_array_equal(mat.view(np.uint32), golden["pf15224_kim"].view(np.uint32))
    rows, _ = ctx.distance_rows_lower(A, "jc")
    want = golden["pf15224_jc"]
    for i, r in enumerate(rows):
        np.testing.assert_array_equal(r.view(np.uint32), want[i, : i + 1].view(np.uint32))


def test_input_type(ctx, N):
    """The reference's -i d|p: forcing the type overrides the discovery on the first row."""
    dna = np.frombuffer(b"ACGTTGCA-N" * 30, dtype=np.uint8).reshape(3, 100).copy()
    dna[1, ::7] = ord('G')
    try:
        ctx.set_input_type("protein")
        A = ctx.alignment(dna)
        assert A.is_protein
        check_counts(ctx, A, dna)
        check_matrices(ctx, A, dna, exact_only=True)
        ctx.set_input_type("dna")
        prot = protein_alignment(5, 64, seed=3)
        assert not ctx.alignment(prot).is_protein
    finally:
        ctx.set_input_type("auto")
    assert not ctx.alignment(dna).is_protein
    assert ctx.alignment(protein_alignment(5, 64, seed=3)).is_protein


@pytest.mark.parametrize("n,L", [(90, 640), (130, 1000)])
def test_protein_bootstrap_replicates(ctx, n, L):
    """Protein resampling (dataloader::sample_fast_dist at 8 bits per site, dataloader.cpp:16-58):
    every replicate, eager and captured, equals the protein run on its column-expanded alignment
    bit for bit, in both models, including a short replicate and lengths that are multiples of
    128 (bitDistanceProtein's valid count 128 short) and ones that are not; the Hamming counter's
    replicates (naive rules) likewise."""
    from rapidnj_amd.workloads import bootstrap_weights, expand_columns
    aln = protein_alignment(n, L, seed=n + 5, alphabet=AMINO + b"*acd", gap=0.1)
    w = bootstrap_weights(L, 5, seed0=11)
    w[4] = 0
    w[4, 10:74] = 2                                  # 128 sites: a multiple of 128
    for naive in (False, True):
        ctx.set_char_rules(naive)
        try:
            A = ctx.alignment(aln)
            expanded = [ctx.alignment(expand_columns(aln, w[r])) for r in range(5)]
        finally:
            ctx.set_char_rules(False)
        assert A.is_protein and all(e.is_protein for e in expanded)
        for model in ("k2p", "jc"):
            got = ctx.bootstrap(A, w, model, raw=True).cpu().numpy()
            with ctx.bootstrap_plan(A, 5, model, max_length=L, raw=True) as plan:
                for _ in range(2):
                    cap = plan.run(w).cpu().numpy()
            for r in range(5):
                ref = ctx.distances(expanded[r], model, raw=True).cpu().numpy()
                np.testing.assert_array_equal(got[r].view(np.uint32), ref.view(np.uint32))
                np.testing.assert_array_equal(cap[r].view(np.uint32), ref.view(np.uint32))
    # one replicate against the oracle's counter on its expanded alignment
    rep = expand_columns(aln, w[4])
    want, _ = O.protein_matrix(rep, "k2p")
    A = ctx.alignment(aln)
    got = ctx.bootstrap(A, w, "k2p", layout="square", exact=True).cpu().numpy()[4]
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_dna_only_operations_refuse_protein(ctx, N):
    aln = protein_alignment(8, 50, seed=4)
    srv = ctx.server()
    try:
        with pytest.raises(N.RnjError) as e:
            srv.submit(aln)
        assert e.value.status == N.RNJ_E_NOT_DNA
    finally:
        srv.close()


@pytest.mark.parametrize("L", [256, 300])
def test_triangle_every_config_with_wrapped_and_saturated_pairs(ctx, N, L):
    """The upper-triangle output (where tiles strictly above the diagonal take the 32-bit
    straight-line epilogue) in every tile configuration, both models, against the oracle: rows of
    gaps and of X scattered over the blocks give pairs without common residues (Kimura -1, patched
    to 2 * max) and, at a multiple-of-128 length, the reference's wrapped 64-bit valid count
    (distance -0.0 or +0.0)."""
    n = 200
    aln = protein_alignment(n, L, seed=21, rates=(0.0, 1.0), gap=0.05)
    for r in (4, 70, 150, 199):
        aln[r] = ord('-')
    for r in (7, 100, 181):
        aln[r, : L - 6] = ord('X')
    A = ctx.alignment(aln)
    iu = np.triu_indices(n, 1)
    try:
        for model in ("k2p", "jc"):
            want = O.protein_matrix(aln, model)[0][iu]
            for c in range(N.lib().rnj_tile_config_count()):
                ctx.set_tile_config(c)
                got = ctx.distances(A, model).cpu().numpy()
                msg = f"{model} config {c}"
                np.testing.assert_allclose(got, want, rtol=1e-6, atol=0, err_msg=msg)
                assert (np.signbit(got) == np.signbit(want)).all(), msg
                exact = ctx.distances(A, model, exact=True).cpu().numpy()
                np.testing.assert_array_equal(exact.view(np.uint32), want.view(np.uint32), err_msg=msg)
    finally:
        ctx.set_tile_config(-1)
    if L % 128 == 0:  # the wrapped pairs are there: zero distances between gap rows and others
        assert (want == 0).sum() > 0
