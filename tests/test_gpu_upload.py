"""Uploads of host characters through the pinned staging chunks (8 MiB each): pitched rows,
inputs that span several chunks, end on a chunk boundary or cross one inside a row, the FASTA
host-bytes path and served jobs. Every result is compared bitwise with the contiguous upload and,
where it is cheap, with the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import simulate_alignment

pytestmark = pytest.mark.gpu

CHUNK = 8 << 20


@pytest.fixture(scope="module")
def N():
    from rapidnj_amd import native
    return native


@pytest.fixture(scope="module")
def ctx(N):
    c = N.Context(0)
    yield c
    c.close()


def _pitched(aln, lead=5, tail=37, fill=ord("N")):
    """The alignment as a column slice of a wider buffer (row pitch L + lead + tail) whose other
    bytes hold characters that would count if they were read."""
    n, L = aln.shape
    base = np.full((n, L + lead + tail), fill, dtype=np.uint8)
    base[:, lead:lead + L] = aln
    view = base[:, lead:lead + L]
    assert view.strides == (L + lead + tail, 1)
    return view


# (n, L): one row crossing the first chunk boundary; exactly one and two chunks; three chunks
# with a boundary inside a row; small (below the threaded-copy size)
SHAPES = [(700, 12345), (1024, 8192), (2048, 8192), (1999, 12601), (37, 1001)]


@pytest.mark.parametrize("n,L", SHAPES)
def test_pitched_rows_equal_contiguous(ctx, n, L):
    aln = simulate_alignment(n, L, 0.05, n + L)
    A = ctx.alignment(np.ascontiguousarray(aln))
    B = ctx.alignment(_pitched(aln))
    np.testing.assert_array_equal(A.planes(), B.planes())
    da = ctx.distances(A, "k2p").cpu().numpy()
    db = ctx.distances(B, "k2p").cpu().numpy()
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))
    A.close()
    B.close()


def test_multi_chunk_counts_against_oracle(ctx):
    aln = simulate_alignment(700, 12345, 0.08, 77)  # 8.6 MB: two chunks, a row split between them
    A = ctx.alignment(_pitched(aln, lead=3, tail=61))
    ts, tv, va = ctx.counts(A)
    rts, rtv, rva = O.counts_triangle(aln, threads=8)
    np.testing.assert_array_equal(ts.cpu().numpy().view(np.uint32), rts)
    np.testing.assert_array_equal(tv.cpu().numpy().view(np.uint32), rtv)
    np.testing.assert_array_equal(va.cpu().numpy().view(np.uint32), rva)
    A.close()


def test_fasta_host_bytes_over_several_chunks(ctx):
    aln = simulate_alignment(1200, 20011, 0.05, 5)  # ~24 MB of FASTA: three chunks and a tail
    lines = []
    for i, row in enumerate(aln):
        lines.append(b">t%d" % i)
        s = row.tobytes()
        lines += [s[k:k + 70] for k in range(0, len(s), 70)]
    data = b"\n".join(lines) + b"\n"
    assert len(data) > 2 * CHUNK
    A = ctx.alignment_from_fasta(data)
    B = ctx.alignment(aln)
    assert A.names == [b"t%d" % i for i in range(len(aln))]
    np.testing.assert_array_equal(A.planes(), B.planes())
    A.close()
    B.close()


def test_served_pitched_jobs_equal_alone(ctx):
    jobs = [(simulate_alignment(n, L, 0.04, 900 + n), m)
            for n, L, m in [(900, 11000, "k2p"), (300, 2999, "jc"), (1300, 7001, "k2p"), (64, 500, "jc")]]
    alone = [ctx.distances(ctx.alignment(a), m).cpu().numpy() for a, m in jobs]
    with ctx.server(slots=2, max_batch_jobs=4) as srv:
        tickets = [srv.submit(_pitched(a, lead=1, tail=14), m) for a, m in jobs]
        got = [srv.wait(t, a.shape[0])[0] for t, (a, _) in zip(tickets, jobs)]
    for g, r in zip(got, alone):
        np.testing.assert_array_equal(g.view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("count", [1, 1000, (1 << 21) + 3, 3 * (1 << 21) + 12345])
def test_copy_to_host_equals_torch_copy(ctx, count):
    """rnj_copy_to_host through the staging chunks: a single float, below the threaded-copy size,
    just over one chunk, and several chunks with a tail."""
    import torch
    t = torch.randn(count, device="cuda:0")
    got = ctx.to_host(t)
    np.testing.assert_array_equal(got.view(np.uint32), t.cpu().numpy().view(np.uint32))


def test_copy_to_host_after_work_on_stream(ctx):
    """The copy waits for the distances enqueued before it on the same stream."""
    import torch
    aln = simulate_alignment(3000, 4000, 0.05, 8)  # 18 MB triangle: three chunks
    A = ctx.alignment(aln)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        d = ctx.distances(A, "k2p", stream=s)
        got = ctx.to_host(d, stream=s)
    ref = ctx.distances(A, "k2p").cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    A.close()


def test_valu_issue_rates(ctx):
    """The issue-rate probe behind the roofline peak: both classes measured, bitwise faster."""
    bitwise, bcnt = ctx.valu_issue_rates()
    assert 1e12 < bcnt < bitwise < 1e15


def test_valu_mix_rates(ctx, N):
    """The count loops' instruction-mix probe behind the roofline peak: every mix is measured,
    sits between the two classes' separate rates (mixed, neither class is faster than alone),
    and invalid mixes or plane counts are refused."""
    bitwise, bcnt = ctx.valu_issue_rates()
    for mix, planes in (("k2p", 0), ("jc", 0), ("naive", 0), ("protein", 5), ("protein", 6),
                        ("protein", 8)):
        rate = ctx.valu_mix_rate(mix, planes)
        assert 0.9 * bcnt < rate < 1.05 * bitwise, (mix, planes, rate, bcnt, bitwise)
    with pytest.raises(N.RnjError):
        ctx.valu_mix_rate("protein", 7)
    import ctypes
    r = ctypes.c_double()
    assert N.lib().rnj_valu_mix_rate(ctx.handle, 9, 0, ctypes.byref(r)) == N.RNJ_E_INVALID
