"""Served jobs: each result is the same alone and in the mix (bitwise: counts are exact and every
job keeps its own maximum and saturation records), admission by HBM budget, error paths."""
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


def _jobs(count, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(2, 400))
        L = int(rng.integers(1, 3000))
        gap = float(rng.uniform(0, 0.3))
        model = "jc" if k % 2 else "k2p"
        out.append((simulate_alignment(n, L, gap, 500 + k), model))
    return out


def test_alone_equals_in_mix(ctx):
    jobs = _jobs(24)
    alone = [ctx.distances(ctx.alignment(a), m).cpu().numpy() for a, m in jobs]
    with ctx.server(slots=2, max_batch_jobs=8, batch_window_us=500_000) as srv:
        tickets = [srv.submit(a, m) for a, m in jobs]
        results = [srv.wait(t, a.shape[0]) for t, (a, _) in zip(tickets, jobs)]
    batched = 0
    for (tri, st), ref in zip(results, alone):
        np.testing.assert_array_equal(tri.view(np.uint32), ref.view(np.uint32))
        assert st.total_s >= st.device_s >= 0 and st.batch_jobs >= 1
        batched += st.batch_jobs > 1
    assert batched == len(jobs)  # inside the batch window every launch held several jobs


def test_jobs_arriving_while_others_run(ctx):
    big = simulate_alignment(1500, 8000, 0.05, 9)
    small = _jobs(6, seed=3)
    ref_big = ctx.distances(ctx.alignment(big), "k2p").cpu().numpy()
    refs = [ctx.distances(ctx.alignment(a), m).cpu().numpy() for a, m in small]
    with ctx.server(slots=3, max_batch_jobs=4) as srv:
        tb = srv.submit(big, "k2p")
        ts = [srv.submit(a, m) for a, m in small]
        got_small = [srv.wait(t, a.shape[0])[0] for t, (a, _) in zip(ts, small)]
        got_big, _ = srv.wait(tb, 1500)
    np.testing.assert_array_equal(got_big.view(np.uint32), ref_big.view(np.uint32))
    for g, r in zip(got_small, refs):
        np.testing.assert_array_equal(g.view(np.uint32), r.view(np.uint32))


def test_admission_rejects_jobs_that_never_fit(ctx, N):
    with ctx.server(slots=1, hbm_budget=1 << 20) as srv:
        ok = srv.submit(simulate_alignment(20, 100, 0.0, 1), "jc")
        too_big = srv.submit(simulate_alignment(600, 4000, 0.0, 2), "k2p")
        srv.wait(ok, 20)
        with pytest.raises(N.RnjError) as e:
            srv.wait(too_big, 600)
        assert e.value.status == N.RNJ_E_ADMISSION


def _size_class(nbytes):
    """The server's device blocks: four size classes per power of two, at least 64 KiB."""
    if nbytes <= 64 << 10:
        return 64 << 10
    step = 1 << ((nbytes - 1).bit_length() - 1 - 2)
    return -(-nbytes // step) * step


# (tile edge, waves per work-group) of the tile configurations
_TILE_WAVES = ((64, 4), (64, 4), (128, 8), (32, 1), (32, 1))


def _job_bytes(n, L):
    """HBM a served job holds: characters, planes and triangle (in the server's size classes)
    and its saturation records (a word, block coordinates and 64 masks per epilogue wave, in
    the configuration that needs the most)."""
    n_pad = -(-n // 128) * 128
    W = -(-max(1, -(-L // 32)) // 16) * 16
    pairs = n * (n - 1) // 2
    sat = max((-(-n // t)) * (-(-n // t) + 1) // 2 * w * 66 * 4 for t, w in _TILE_WAVES)
    return (_size_class(n * L) + _size_class(3 * W * n_pad * 4) + _size_class(pairs * 4) + sat)


def test_size_classes():
    assert [_size_class(b) for b in (1, 65536, 65537, 100000, 1 << 20, (1 << 20) + 1)] == \
        [65536, 65536, 81920, 114688, 1 << 20, 1310720]


def test_budget_serialises_but_completes(ctx):
    """A budget that holds one job at a time still runs every job, in order."""
    jobs = _jobs(5, seed=11)
    need = max(_job_bytes(*a.shape) for a, _ in jobs)
    refs = [ctx.distances(ctx.alignment(a), m).cpu().numpy() for a, m in jobs]
    with ctx.server(slots=2, hbm_budget=need + 1) as srv:
        ts = [srv.submit(a, m) for a, m in jobs]
        for t, (a, _), r in zip(ts, jobs, refs):
            got, _ = srv.wait(t, a.shape[0])
            np.testing.assert_array_equal(got.view(np.uint32), r.view(np.uint32))


def test_server_argument_errors(ctx, N):
    with ctx.server() as srv:
        with pytest.raises(N.RnjError):
            srv.submit(np.zeros((1, 10), dtype=np.uint8) + ord("A"), "jc")
        with pytest.raises(N.RnjError):
            srv.wait(123456, 5)
        with pytest.raises(N.RnjError):
            srv.cancel(987654)
