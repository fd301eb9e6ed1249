"""Parity of the HIP path (through the C ABI) with the reference engine and the oracle.

Bars: counts (ts, tv, valid) and packed planes bit-exact; distances within 1e-6 relative of the
reference's float values, with identical sign of zero, identical saturated pairs and identical
sentinel value 2*max. Expected values come from the golden vectors the reference engine wrote
(tests/golden/) and, at sizes without fixtures, from the CPU oracle on the same seeded input.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import bootstrap_weights, expand_columns, simulate_alignment

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EDGE = sorted(glob.glob(os.path.join(GOLDEN, "edge_L*.npz")))
FULL_CASES = EDGE + [os.path.join(GOLDEN, f) for f in ("c1.npz", "lgt70.npz", "saturated.npz")]
DIST_RTOL = 1e-6  # relative tolerance on corrected distances (FP64 log on device vs glibc)


def _load(path):
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="module")
def N():
    from rapidnj_amd import native
    return native


@pytest.fixture(scope="module")
def ctx(N):
    c = N.Context(0)
    yield c
    c.close()


def _np(t):
    return t.detach().cpu().numpy()


def _u32(t):
    return _np(t).view(np.uint32)


def assert_distances_match(got, ref):
    """Reference post-processed matrices: 1e-6 relative, same zero signs, same sentinels."""
    got = np.asarray(got, dtype=np.float32)
    ref = np.asarray(ref, dtype=np.float32)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=DIST_RTOL, atol=0)
    np.testing.assert_array_equal(np.signbit(got), np.signbit(ref))


def planes_from_reference_words(bits, gaps, L):
    """The reference's 16-sites-per-word codes/filters -> H/L/V planes [3][ceil(L/32)][n]."""
    n = bits.shape[0]
    s = np.arange(L)
    code = (bits[:, s // 16] >> (2 * (s % 16))) & 3
    val = (gaps[:, s // 16] >> (2 * (s % 16))) & 1
    W = (L + 31) // 32
    out = np.zeros((3, W, n), dtype=np.uint32)
    for p, field in enumerate((code >> 1, code & 1, val)):
        for w in range(W):
            blk = field[:, 32 * w: 32 * w + 32].astype(np.uint64)
            out[p, w] = (blk << np.arange(blk.shape[1], dtype=np.uint64)).sum(axis=1).astype(np.uint32)
    return out


@pytest.mark.parametrize("path", EDGE, ids=os.path.basename)
def test_planes_match_reference_packing(ctx, path):
    d = _load(path)
    aln = d["aln"]
    A = ctx.alignment(aln)
    got = A.planes()
    ref = planes_from_reference_words(d["bits"], d["gaps"], aln.shape[1])
    W = ref.shape[1]
    np.testing.assert_array_equal(got[:, :W], ref)
    assert not got[:, W:].any()  # padding words count nothing


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
def test_counts_bit_exact(ctx, path):
    d = _load(path)
    ts, tv, va = ctx.counts(ctx.alignment(d["aln"]))
    np.testing.assert_array_equal(_u32(ts), d["ts"])
    np.testing.assert_array_equal(_u32(tv), d["tv"])
    np.testing.assert_array_equal(_u32(va), d["valid"])


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
@pytest.mark.parametrize("model", ["jc", "k2p"])
def test_square_matrix_matches_reference(ctx, path, model):
    d = _load(path)
    got = _np(ctx.distances(ctx.alignment(d["aln"]), model, layout="square"))
    assert_distances_match(got, d[model])
    assert np.all(got.diagonal().view(np.uint32) == 0)


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
@pytest.mark.parametrize("model", ["jc", "k2p"])
def test_raw_triangle_matches_oracle(ctx, N, path, model):
    """Before post-processing: identical saturated pairs (-1) and per-pair values."""
    import torch
    d = _load(path)
    tri_ref, full_ref, mx_ref = O.distance_matrix(d["aln"], model, threads=4)
    A = ctx.alignment(d["aln"])
    mx = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    got = _np(ctx.distances(A, model, layout="triangle", raw=True, max_out=mx))
    np.testing.assert_array_equal(got == -1, tri_ref == -1)
    assert_distances_match(got, tri_ref)
    assert float(mx.item()) == pytest.approx(mx_ref, rel=DIST_RTOL)
    post = _np(ctx.distances(A, model, layout="triangle"))
    iu = np.triu_indices(A.n, 1)
    assert_distances_match(post, full_ref[iu])


def test_host_rows_handoff(ctx):
    d = _load(os.path.join(GOLDEN, "lgt70.npz"))
    for model in ("jc", "k2p"):
        mat, mx = ctx.distance_rows(ctx.alignment(d["aln"]), model)
        assert_distances_match(mat, d[model])
        assert np.array_equal(mat, mat.T)
        low, mx2 = ctx.distance_rows_lower(ctx.alignment(d["aln"]), model)
        assert mx2 == mx and [len(r) for r in low] == list(range(1, mat.shape[0] + 1))
        for i, r in enumerate(low):
            np.testing.assert_array_equal(r.view(np.uint32), mat[i, : i + 1].view(np.uint32))


def test_host_rows_handoff_many_slabs(ctx):
    """5000 taxa: the square rows span several pinned slabs, the lower rows two; both equal the
    device square matrix bit for bit."""
    A = ctx.alignment(simulate_alignment(5000, 700, 0.05, 19))
    sq = _np(ctx.distances(A, "k2p", layout="square", exact=True))  # the hand-offs use exact logs
    mat, mx = ctx.distance_rows(A, "k2p")
    np.testing.assert_array_equal(mat.view(np.uint32), sq.view(np.uint32))
    low, _ = ctx.distance_rows_lower(A, "k2p")
    for i in (0, 1, 2047, 2048, 3333, 4999):
        np.testing.assert_array_equal(low[i].view(np.uint32), sq[i, : i + 1].view(np.uint32))
    assert all(len(low[i]) == i + 1 for i in range(5000))


@pytest.mark.parametrize("path", EDGE[:6] + [os.path.join(GOLDEN, "c1.npz")], ids=os.path.basename)
def test_from_reference_bitstrings(ctx, path):
    d = _load(path)
    if "bits" not in d:
        d["bits"], d["gaps"], _ = O.pack(d["aln"])
    A = ctx.alignment_from_bitstrings(d["bits"], d["gaps"])
    ts, tv, va = ctx.counts(A)
    np.testing.assert_array_equal(_u32(ts), d["ts"])
    np.testing.assert_array_equal(_u32(tv), d["tv"])
    np.testing.assert_array_equal(_u32(va), d["valid"])


def _staging(bits, gaps, L):
    """dataloader::initGPUData layout: [n][bsStride], bsStride = roundup(ceil(L/16), 128)."""
    n, words = bits.shape
    bsc = words // 4
    stride = ((L + 15) // 16 + 127) & ~127
    sb = np.zeros((n, stride), dtype=np.uint32)
    sg = np.zeros((n, stride), dtype=np.uint32)
    k = min(stride, words)
    sb[:, :k] = bits[:, :k]
    sg[:, :k] = gaps[:, :k]
    return sb, sg, bsc, stride


def test_reference_gpu_abi_k2p_and_counts(N):
    d = _load(os.path.join(GOLDEN, "lgt70.npz"))
    bits, gaps, _ = O.pack(d["aln"])
    sb, sg, bsc, stride = _staging(bits, gaps, d["aln"].shape[1])
    n = bits.shape[0]
    L = N.lib()
    L.initialiseDNA2_gpu(n, bsc, stride)
    L.storeDataDNA2_gpu(sb.ctypes.data, sg.ctypes.data)
    res = np.zeros((n, n), dtype=np.float32)
    L.computeDistancesDNA2_gpu(res.ctypes.data)
    tri_ref, _, _ = O.distance_matrix(d["aln"], "k2p")
    iu = np.triu_indices(n, 1)
    np.testing.assert_array_equal(res[iu] == -1, tri_ref == -1)
    assert_distances_match(res[iu], tri_ref)
    L.initialiseDNA_gpu(n, bsc, stride)
    L.storeDataDNA_gpu(sb.ctypes.data, sg.ctypes.data)
    ts = np.zeros((n, n), dtype=np.uint32)
    tv = np.zeros_like(ts)
    gp = np.zeros_like(ts)
    L.computeDistancesDNA_gpu(ts.ctypes.data, tv.ctypes.data, gp.ctypes.data)
    L.getResultsDNA_gpu(ts.ctypes.data, tv.ctypes.data, gp.ctypes.data)
    np.testing.assert_array_equal(ts[iu], d["ts"])
    np.testing.assert_array_equal(tv[iu], d["tv"])
    np.testing.assert_array_equal(gp[iu], d["valid"])


def test_every_tile_config_agrees(ctx, N):
    d = _load(os.path.join(GOLDEN, "c1.npz"))
    A = ctx.alignment(d["aln"])
    try:
        for c in range(len(N.tile_configs())):
            ctx.set_tile_config(c)
            ts, tv, va = ctx.counts(A)
            np.testing.assert_array_equal(_u32(ts), d["ts"], err_msg=f"config {c}")
            np.testing.assert_array_equal(_u32(tv), d["tv"], err_msg=f"config {c}")
            np.testing.assert_array_equal(_u32(va), d["valid"], err_msg=f"config {c}")
            for model in ("jc", "k2p"):
                got = _np(ctx.distances(A, model, layout="square"))
                assert_distances_match(got, d[model])
    finally:
        ctx.set_tile_config(-1)


@pytest.mark.parametrize("model", ["jc", "k2p"])
def test_saturation_records_every_config(ctx, N, model):
    """Alignments where most pairs and where some pairs saturate, in every tile configuration,
    both layouts and a raw run: the saturated pairs (marked per wave of the epilogue and mapped
    back through each configuration's tile geometry) become 2*max, every one and only those, and
    the maximum is the oracle's."""
    import torch
    rng = np.random.default_rng(12)
    cases = [np.frombuffer(b"ACGT-", np.uint8)[rng.integers(0, 5, size=(300, 24))],
             simulate_alignment(700, 60, 0.3, 13), simulate_alignment(260, 40, 0.5, 14)]
    try:
        for aln in cases:
            n = aln.shape[0]
            tri_ref, full_ref, mx_ref = O.distance_matrix(aln, model)
            assert (tri_ref == -1).any()
            iu = np.triu_indices(n, 1)
            A = ctx.alignment(aln)
            for c in range(len(N.tile_configs())):
                ctx.set_tile_config(c)
                mx = torch.zeros(1, dtype=torch.float32, device="cuda:0")
                sq = _np(ctx.distances(A, model, layout="square", max_out=mx))
                assert_distances_match(sq, full_ref)
                assert float(mx.item()) == pytest.approx(mx_ref, rel=DIST_RTOL)
                tri = _np(ctx.distances(A, model))
                assert_distances_match(tri, full_ref[iu])
                raw = _np(ctx.distances(A, model, raw=True))
                np.testing.assert_array_equal(raw == -1, tri_ref == -1, err_msg=f"config {c}")
    finally:
        ctx.set_tile_config(-1)


@pytest.mark.parametrize("path", FULL_CASES, ids=os.path.basename)
def test_naive_rules_match_reference_counts(ctx, path):
    """RNJ_CHARS_NAIVE (the reference's `-f` counter): ts, tv and valid equal the naive counts
    the reference's simpleDistanceCalculator wrote for the golden alignments."""
    d = _load(path)
    ctx.set_char_rules(True)
    try:
        A = ctx.alignment(d["aln"])
    finally:
        ctx.set_char_rules(False)
    ts, tv, va = ctx.counts(A)
    np.testing.assert_array_equal(_u32(ts), d["naive_ts"])
    np.testing.assert_array_equal(_u32(tv), d["naive_tv"])
    np.testing.assert_array_equal(_u32(va), d["naive_valid"])


def _mixed_case_alignment(n, L, seed):
    """Lowercase letters, U for T and N among uppercase ACGT and gaps: where the naive counter
    and the bit engine differ."""
    aln = simulate_alignment(n, L, 0.05, seed).copy()
    rng = np.random.default_rng(seed)
    aln[rng.random(aln.shape) < 0.2] |= 0x20          # lowercase ('-' is unchanged)
    t = (aln | 0x20) == ord("t")
    aln[t & (rng.random(aln.shape) < 0.3)] += np.uint8(ord("u") - ord("t"))   # T -> U, t -> u
    aln[rng.random(aln.shape) < 0.02] = ord("N")
    aln[0, :40] = np.frombuffer(b"ACGTacgtuU" * 4, np.uint8)       # DNA by the first row
    return aln


def test_naive_rules_every_config(ctx, N):
    """Mixed-case alignment under the naive rules, in every tile configuration: counts equal the
    oracle's naive counter on every pair, distances equal its corrections in both layouts, and
    the naive alignment is refused where only bit-rule alignments go."""
    n, L = 300, 777
    aln = _mixed_case_alignment(n, L, 41)
    iu = np.triu_indices(n, 1)
    ref = np.array([O.count_naive(aln[i].tobytes(), aln[j].tobytes()) for i, j in zip(*iu)],
                   dtype=np.uint32).reshape(-1, 3)
    # the two rules differ on this alignment
    bits = np.array([O.count_chars(aln[i].tobytes(), aln[j].tobytes())
                     for i, j in zip(iu[0][:2000], iu[1][:2000])], dtype=np.uint32)
    assert (bits != ref[:2000]).any()
    ctx.set_char_rules(True)
    try:
        A = ctx.alignment(aln)
    finally:
        ctx.set_char_rules(False)
    try:
        for c in range(len(N.tile_configs())):
            ctx.set_tile_config(c)
            ts, tv, va = ctx.counts(A)
            np.testing.assert_array_equal(_u32(ts), ref[:, 0], err_msg=f"config {c}")
            np.testing.assert_array_equal(_u32(tv), ref[:, 1], err_msg=f"config {c}")
            np.testing.assert_array_equal(_u32(va), ref[:, 2], err_msg=f"config {c}")
            for model in ("jc", "k2p"):
                _, full, _ = O.correct_triangle(ref[:, 0], ref[:, 1], ref[:, 2], n, model)
                assert_distances_match(_np(ctx.distances(A, model, layout="square")), full)
                assert_distances_match(_np(ctx.distances(A, model)), full[iu])
    finally:
        ctx.set_tile_config(-1)
    with pytest.raises(N.RnjError):
        A.update(aln)


def test_naive_bootstrap_replicates(ctx):
    """The -f path's resampling (dataloader::sample_char, dataloader.cpp:60-69) on a mixed-case
    alignment: every replicate, eager and captured, equals the naive run on its column-expanded
    alignment bit for bit, and its counts are the oracle's naive counter on that alignment."""
    n, L = 120, 700
    aln = _mixed_case_alignment(n, L, 43)
    w = bootstrap_weights(L, 4, seed0=7)
    w[3] = 0
    w[3, 5:40] = 2                                  # a short replicate (70 sites)
    ctx.set_char_rules(True)
    try:
        A = ctx.alignment(aln)
        expanded = [ctx.alignment(expand_columns(aln, w[r])) for r in range(4)]
    finally:
        ctx.set_char_rules(False)
    iu = np.triu_indices(n, 1)
    for model in ("jc", "k2p"):
        got = _np(ctx.bootstrap(A, w, model, raw=True))
        with ctx.bootstrap_plan(A, 4, model, max_length=L, raw=True) as plan:
            cap = _np(plan.run(w))
        for r in range(4):
            ref = _np(ctx.distances(expanded[r], model, raw=True))
            np.testing.assert_array_equal(got[r].view(np.uint32), ref.view(np.uint32))
            np.testing.assert_array_equal(cap[r].view(np.uint32), ref.view(np.uint32))
    rep = expand_columns(aln, w[1])
    cnt = np.array([O.count_naive(rep[i].tobytes(), rep[j].tobytes()) for i, j in zip(*iu)],
                   dtype=np.uint32).reshape(-1, 3)
    tri, _, _ = O.correct_triangle(cnt[:, 0], cnt[:, 1], cnt[:, 2], n, "k2p", with_full=False)
    got = _np(ctx.bootstrap(A, w, "k2p", raw=True))[1]
    np.testing.assert_array_equal(got == -1, tri == -1)
    np.testing.assert_allclose(got, tri, rtol=DIST_RTOL, atol=0)


def test_bootstrap_matches_reference_on_expanded_columns(ctx):
    d = _load(os.path.join(GOLDEN, "bootstrap.npz"))
    A = ctx.alignment(d["aln"])
    got = _np(ctx.bootstrap(A, d["weights"], "k2p", layout="square"))
    for r in range(d["weights"].shape[0]):
        assert_distances_match(got[r], d[f"k2p{r}"])


def test_bootstrap_counts_equal_plain_run_on_expanded(ctx):
    aln = simulate_alignment(70, 900, 0.05, 31)
    w = bootstrap_weights(900, 5, seed0=40)
    w[2] = 0
    w[2, :17] = 3                               # a short replicate (51 sites)
    A = ctx.alignment(aln)
    for model in ("jc", "k2p"):
        got = _np(ctx.bootstrap(A, w, model))
        for r in range(w.shape[0]):
            ref = _np(ctx.distances(ctx.alignment(expand_columns(aln, w[r])), model))
            np.testing.assert_array_equal(got[r].view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("n,L,gap", [(2, 1, 0.0), (3, 5, 0.5), (129, 77, 0.1), (200, 2049, 0.05),
                                     (257, 4000, 0.3)])
def test_shapes_against_oracle(ctx, n, L, gap):
    aln = simulate_alignment(n, L, gap, 1000 + n + L)
    ts_r, tv_r, va_r = O.counts_triangle(aln)
    A = ctx.alignment(aln)
    ts, tv, va = ctx.counts(A)
    np.testing.assert_array_equal(_u32(ts), ts_r)
    np.testing.assert_array_equal(_u32(tv), tv_r)
    np.testing.assert_array_equal(_u32(va), va_r)
    for model in ("jc", "k2p"):
        _, full, _ = O.distance_matrix(aln, model)
        assert_distances_match(_np(ctx.distances(A, model, layout="square")), full)


def test_plain_gpu_config_against_oracle(ctx):
    """2000 taxa x 10000 sites, 5% gaps, K2P (the plain single-job GPU configuration)."""
    aln = simulate_alignment(2000, 10000, 0.05, 2)
    A = ctx.alignment(aln)
    ts_r, tv_r, va_r = O.counts_triangle(aln, threads=os.cpu_count() or 8)
    ts, tv, va = ctx.counts(A)
    np.testing.assert_array_equal(_u32(ts), ts_r)
    np.testing.assert_array_equal(_u32(tv), tv_r)
    np.testing.assert_array_equal(_u32(va), va_r)
    tri_ref, _, _ = O.correct_triangle(ts_r, tv_r, va_r, 2000, "k2p")
    got = _np(ctx.distances(A, "k2p", raw=True))
    np.testing.assert_array_equal(got == -1, tri_ref == -1)
    assert_distances_match(got, tri_ref)


def test_large_properties(ctx, N):
    """6000 x 3000: symmetric, zero diagonal, sentinel = 2*max, sampled pairs exact."""
    import torch
    n, L = 6000, 3000
    aln = simulate_alignment(n, L, 0.05, 77)
    A = ctx.alignment(aln)
    mx = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    sq = ctx.distances(A, "k2p", layout="square", max_out=mx)
    assert torch.equal(sq, sq.T)
    assert torch.all(sq.diagonal() == 0)
    m = float(mx.item())
    assert m > 0 and float(sq.max().item()) <= 2 * m
    ts, tv, va = ctx.counts(A)
    # every tile configuration gives the same counts on every pair
    for c in range(len(N.tile_configs())):
        ctx.set_tile_config(c)
        try:
            ts2, tv2, va2 = ctx.counts(A)
        finally:
            ctx.set_tile_config(-1)
        assert torch.equal(ts2, ts) and torch.equal(tv2, tv) and torch.equal(va2, va)
    bits, gaps, _ = O.pack(aln)
    rng = np.random.default_rng(5)
    ii = rng.integers(0, n, 400)
    jj = rng.integers(0, n, 400)
    tsn, tvn, van = _u32(ts), _u32(tv), _u32(va)
    sqn = _np(sq)
    for i, j in zip(ii, jj):
        if i == j:
            continue
        a, b = min(i, j), max(i, j)
        k = a * n - a * (a + 1) // 2 + (b - a - 1)
        ref = O.count_words(bits, gaps, a, b)
        assert (int(tsn[k]), int(tvn[k]), int(van[k])) == ref
        dk = O.k2p(*ref)
        if dk != -1:
            assert sqn[a, b] == pytest.approx(dk, rel=DIST_RTOL)


def test_context_close_releases_children(N):
    """Closing a context closes what was made from it; later closes and collection are no-ops,
    and a new context works (no stale device state or HIP error is left behind)."""
    import gc
    c = N.Context(0)
    aln = simulate_alignment(40, 300, 0.1, 4)
    A = c.alignment(aln)
    B = c.alignment(aln)
    plan = c.bootstrap_plan(A, 2, "k2p")
    ref = _np(c.distances(A, "k2p"))
    c.close()
    assert A.handle is None and B.handle is None and plan.handle is None
    A.close()
    del A, B, plan
    gc.collect()
    c2 = N.Context(0)
    try:
        got = _np(c2.distances(c2.alignment(aln), "k2p"))
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    finally:
        c2.close()


def test_error_paths(ctx, N):
    with pytest.raises(N.RnjError) as e:
        ctx.alignment(np.frombuffer(b"ACGT", dtype=np.uint8).reshape(1, 4))
    assert e.value.status == N.RNJ_E_INVALID
    prot = np.frombuffer(b"MKVLAAGG" * 2, dtype=np.uint8).reshape(2, 8)
    assert ctx.alignment(prot).is_protein
    assert not ctx.alignment(prot, skip_type_check=True).is_protein
    A = ctx.alignment(simulate_alignment(10, 50, 0.0, 3))
    import torch
    small = torch.empty((10, 9), dtype=torch.float32, device="cuda:0")
    with pytest.raises(N.RnjError):
        ctx.distances(A, "k2p", layout="square", out=small)
    with pytest.raises(KeyError):
        ctx.distances(A, "logdet")
