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


def _protein(n, L, seed):
    rng = np.random.default_rng(seed)
    return np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY-", dtype=np.uint8)[rng.integers(0, 21, (n, L))]


def test_resident_jobs_equal_alone(ctx):
    """Jobs on alignments already in HBM (DNA under both character rules, protein), triangles in
    the caller's device buffers or in server buffers copied by wait: each equals the same
    alignment run alone, bit for bit; jobs of one kind share launches."""
    import torch
    specs = [(simulate_alignment(int(n), int(L), 0.05, 900 + k), m, "bits")
             for k, (n, L, m) in enumerate([(300, 900, "k2p"), (2, 5, "jc"), (120, 4000, "jc"),
                                            (700, 333, "k2p"), (64, 64, "k2p")])]
    specs += [(_protein(150, 640, 3), "k2p", "bits"), (_protein(90, 300, 4), "jc", "bits"),
              (simulate_alignment(200, 500, 0.1, 77), "k2p", "naive")]
    alns = []
    for a, _, rules in specs:
        ctx.set_char_rules(rules == "naive")
        try:
            alns.append(ctx.alignment(a))
        finally:
            ctx.set_char_rules(False)
    refs = [ctx.distances(A, m).cpu().numpy() for A, (_, m, _) in zip(alns, specs)]
    outs = [torch.full((A.n * (A.n - 1) // 2,), -7.0, device="cuda:0") if k % 2 == 0 else None
            for k, A in enumerate(alns)]
    with ctx.server(slots=2, max_batch_jobs=8, batch_window_us=200_000) as srv:
        ts = [srv.submit_resident(A, m, out=o) for A, (_, m, _), o in zip(alns, specs, outs)]
        res = [srv.wait(t, A.n) for t, A in zip(ts, alns)]
    for k, ((host, st), ref) in enumerate(zip(res, refs)):
        np.testing.assert_array_equal(host.view(np.uint32), ref.view(np.uint32))
        if outs[k] is not None:
            np.testing.assert_array_equal(outs[k].cpu().numpy().view(np.uint32), ref.view(np.uint32))
        assert st.upload_s < 0.05 and st.launched_at >= 0 and st.finished_at >= st.launched_at
    assert max(st.batch_jobs for _, st in res) > 1
    for A in alns:
        A.close()


def test_job_admitted_during_running_batch_runs_beside_it(ctx):
    """A job admitted while another batch runs on the device is launched on the other slot
    before that batch finishes, and its work completes first: the two overlap on the device."""
    import torch
    rng = np.random.default_rng(5)
    big = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, (8000, 60000), dtype=np.uint8)]
    A_big = ctx.alignment(big)
    del big
    A_small = ctx.alignment(simulate_alignment(100, 1000, 0.02, 6))
    out_big = torch.empty((8000 * 7999 // 2,), device="cuda:0")
    ref_small = ctx.distances(A_small, "jc").cpu().numpy()
    import time
    with ctx.server(slots=2, max_batch_jobs=8) as srv:
        tb = srv.submit_resident(A_big, "k2p", out=out_big)
        time.sleep(0.003)  # the big job's batch is launched by now
        ts = srv.submit_resident(A_small, "jc")
        small, st_small = srv.wait(ts, 100)
        _, st_big = srv.wait(tb, 8000, copy=False)
    np.testing.assert_array_equal(small.view(np.uint32), ref_small.view(np.uint32))
    print("big", st_big.launched_at, st_big.finished_at, "small", st_small.launched_at,
          st_small.finished_at, flush=True)
    assert st_small.batch_id != st_big.batch_id
    assert st_big.launched_at < st_small.launched_at < st_big.finished_at
    assert st_small.finished_at < st_big.finished_at
    A_big.close()
    A_small.close()


def test_resident_errors(ctx, N):
    other = N.Context(0)
    try:
        A_other = other.alignment(simulate_alignment(10, 50, 0.0, 1))
        with ctx.server() as srv:
            with pytest.raises(N.RnjError):
                srv.submit_resident(A_other, "k2p")
            A = ctx.alignment(simulate_alignment(10, 50, 0.0, 2))
            with pytest.raises(N.RnjError):
                srv.submit_resident(A, 7)
            t = srv.submit_resident(A, "jc")
            srv.wait(t, 10)
            A.close()
    finally:
        other.close()
