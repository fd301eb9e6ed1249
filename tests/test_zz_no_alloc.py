"""rnj_distances_device_noalloc, the entry point that allocates nothing: through the C ABI its
output equals rnj_distances_device's on the same alignment (bit for bit: both run the same
kernels on the same counts), on two calls in a row with the second on a smaller input (the
scratch array keeps the first call's contents), in both layouts, with and without the 2 * max
post-process, for a launch cut into word ranges (partial counts in the scratch) and for one that
is not; matrices whose scratch does not fit are refused before anything runs."""
import ctypes

import numpy as np
import pytest

from rapidnj_amd.workloads import simulate_alignment

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def N():
    from rapidnj_amd import native
    return native


@pytest.fixture(scope="module")
def ctx(N):
    c = N.Context(0)
    yield c
    c.close()


def noalloc(N, ctx, A, model, layout, raw=False):
    """Output and maximum of the no-allocation entry point, called through the C ABI."""
    import torch
    n = A.n
    lay = {"triangle": N.TRIANGLE, "square": N.SQUARE}[layout]
    shape = (n * (n - 1) // 2,) if lay == N.TRIANGLE else (n, n)
    out = torch.full(shape, 7.0, dtype=torch.float32, device="cuda:0")
    mx = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    ld = out.stride(0) if lay == N.SQUARE else 0
    N.check(N.lib().rnj_distances_device_noalloc(
        ctx.handle, A.handle, N._model(model), lay, out.data_ptr(), ld, N.FLAG_RAW if raw else 0,
        mx.data_ptr(), torch.cuda.current_stream().cuda_stream))
    return out.cpu().numpy(), float(mx.cpu()[0])


def existing(ctx, A, model, layout, raw=False):
    import torch
    mx = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    out = ctx.distances(A, model, layout=layout, raw=raw, max_out=mx)
    return out.cpu().numpy(), float(mx.cpu()[0])


@pytest.mark.parametrize("layout", ["triangle", "square"])
@pytest.mark.parametrize("raw", [False, True])
def test_two_calls_second_smaller(N, ctx, layout, raw):
    big = ctx.alignment(simulate_alignment(3000, 2000, 0.05, 71))   # no split: 4700 tiles
    small = ctx.alignment(simulate_alignment(150, 700, 0.08, 72))   # split into word ranges
    for A, model in ((big, "k2p"), (small, "jc")):
        got, gmx = noalloc(N, ctx, A, model, layout, raw)
        want, wmx = existing(ctx, A, model, layout, raw)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        assert gmx == wmx > 0


def test_naive_and_protein_alignments(N, ctx):
    aln = simulate_alignment(200, 900, 0.05, 73).copy()
    aln[:, ::3] |= 0x20
    ctx.set_char_rules(True)
    try:
        A = ctx.alignment(aln)
    finally:
        ctx.set_char_rules(False)
    rng = np.random.default_rng(74)
    prot = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY-", np.uint8)[rng.integers(0, 21, (120, 640))]
    prot[0, :2] = np.frombuffer(b"EF", np.uint8)
    P = ctx.alignment(prot)
    assert P.is_protein
    for X in (A, P):
        for model in ("jc", "k2p"):
            got, _ = noalloc(N, ctx, X, model, "square")
            want, _ = existing(ctx, X, model, "square")
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_too_large_is_refused_before_any_launch(N, ctx):
    A = ctx.alignment(simulate_alignment(30000, 32, 0.0, 75))
    import torch
    out = torch.zeros(1, dtype=torch.float32, device="cuda:0")  # never written
    rc = N.lib().rnj_distances_device_noalloc(ctx.handle, A.handle, N.K2P, N.TRIANGLE,
                                              out.data_ptr(), 0, 0, None, None)
    assert rc == N.RNJ_E_INVALID
    assert b"no-allocation" in N.lib().rnj_last_error()
    assert float(out.cpu()[0]) == 0.0
