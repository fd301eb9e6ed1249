"""Parity at the benchmarked sizes, so that the figures bench.py reports are checked on the
inputs they are measured on.

* 20000 x 5000 K2P, 5 % gaps (seed 3, the default bench workload): every one of the 2e8 pairs'
  ts / tv / valid bit-exact against the oracle's SWAR counter; raw distances (K2P and JC) within
  1e-6 relative of the oracle's corrections with the identical set of saturated pairs (-1) and
  identical zero signs; the post-processed triangle with the identical sentinel 2 * max.
* 1000 x 5000 K2P with the 100 multinomial column-weight vectors of the bootstrap workload
  (seed 4, weights seeds 0..99) through the captured plan (rnj_bootstrap_plan_run): every
  replicate against the oracle on its column-expanded alignment (the reference's resampled
  alignment, dataloader.cpp:16-58), raw and post-processed.
* The served stream's distribution (Poisson 20/s, taxa log-uniform 50-5000, sites 500-30000,
  0-8 % gaps, JC/K2P): its first 64 jobs plus its largest-taxa and longest jobs, submitted at
  their arrival times: every result bitwise equal to the job's own launch alone, and the largest,
  longest and four more against the oracle.

The oracle (oracle/rnj_oracle.c) is pinned to the reference's own outputs by tests/test_oracle.py.
"""
import os
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import (WORKLOADS, bootstrap_weights, expand_columns, served_stream,
                                   simulate_alignment, simulate_many)

pytestmark = pytest.mark.gpu

DIST_RTOL = 1e-6


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.fixture(scope="module")
def N():
    from rapidnj_amd import native
    return native


@pytest.fixture(scope="module")
def ctx(N):
    c = N.Context(0)
    yield c
    c.close()


def assert_raw_match(got, ref):
    """Raw (unpatched) distances: 1e-6 relative, identical -1 set, identical zero signs."""
    np.testing.assert_array_equal(got == -1, ref == -1)
    np.testing.assert_allclose(got, ref, rtol=DIST_RTOL, atol=0)
    np.testing.assert_array_equal(np.signbit(got), np.signbit(ref))


def patched(tri, mx):
    out = tri.copy()
    out[out == -1] = np.float32(mx) * np.float32(2)
    return out


# ---- 20000 x 5000 K2P ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def large(ctx):
    wl = WORKLOADS["large"]
    aln = simulate_alignment(wl.n, wl.L, wl.gap, wl.seed)
    print(f"\n[fullsize] {wl.name}: oracle counts on {host_threads()} threads", flush=True)
    ts, tv, va = O.counts_triangle(aln, threads=host_threads())
    return aln, ts, tv, va


def test_large_counts_bit_exact(ctx, large):
    aln, ts, tv, va = large
    A = ctx.alignment(aln)
    for got, want in zip(ctx.counts(A), (ts, tv, va)):
        np.testing.assert_array_equal(ctx.to_host(got).view(np.uint32), want)
    A.close()


@pytest.mark.parametrize("model", ["k2p", "jc"])
def test_large_distances(ctx, large, model):
    aln, ts, tv, va = large
    tri, _, mx = O.correct_triangle(ts, tv, va, aln.shape[0], model, with_full=False)
    assert (tri == -1).any()  # saturated pairs: the sentinel path is exercised at this size
    A = ctx.alignment(aln)
    raw = ctx.to_host(ctx.distances(A, model, raw=True))
    assert_raw_match(raw, tri)
    post = ctx.to_host(ctx.distances(A, model))
    want = patched(tri, mx)
    np.testing.assert_allclose(post, want, rtol=DIST_RTOL, atol=0)
    np.testing.assert_array_equal(post[tri == -1], np.float32(mx) * np.float32(2))
    A.close()


# ---- 1000 x 5000 x 100 bootstrap replicates, captured ------------------------------------------
def test_bootstrap_plan_all_replicates(ctx):
    wl = WORKLOADS["bootstrap"]
    aln = simulate_alignment(wl.n, wl.L, wl.gap, wl.seed)
    w = bootstrap_weights(wl.L, wl.replicates, seed0=0)
    A = ctx.alignment(aln)
    m = wl.n * (wl.n - 1) // 2
    with ctx.bootstrap_plan(A, wl.replicates, "k2p") as plan, \
            ctx.bootstrap_plan(A, wl.replicates, "k2p", raw=True) as plan_raw:
        post = ctx.to_host(plan.run(w).contiguous())
        raw = ctx.to_host(plan_raw.run(w).contiguous())
    assert post.shape == raw.shape == (wl.replicates, m)
    threads = host_threads()
    for r in range(wl.replicates):
        rep = expand_columns(aln, w[r])
        assert rep.shape == aln.shape  # multinomial draws of L columns
        tri, _, mx = O.correct_triangle(*O.counts_triangle(rep, threads=threads), wl.n, "k2p",
                                        with_full=False)
        assert_raw_match(raw[r], tri)
        np.testing.assert_allclose(post[r], patched(tri, mx), rtol=DIST_RTOL, atol=0)
    A.close()


# ---- the served stream --------------------------------------------------------------------------
def test_served_stream_jobs(ctx):
    stream = served_stream()
    pick = list(stream[:64])
    for extra in (max(stream, key=lambda j: j.n), max(stream, key=lambda j: j.L)):
        if extra not in pick:
            pick.append(extra)
    assert max(j.n for j in pick) > 4500 and max(j.L for j in pick) > 29000
    alns = simulate_many([(j.n, j.L, j.gap, j.seed) for j in pick], procs=host_threads())
    print(f"\n[fullsize] served stream: {len(pick)} jobs generated", flush=True)
    alone = []
    for j, a in zip(pick, alns):
        A = ctx.alignment(a)
        alone.append(ctx.to_host(ctx.distances(A, j.model)))
        A.close()
    tickets = [None] * len(pick)
    with ctx.server(slots=2, max_batch_jobs=64) as srv:
        t0 = time.perf_counter()
        first = pick[0].arrival_s

        def submitter():
            for k, j in enumerate(pick):
                delay = t0 + (j.arrival_s - first) - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[k] = srv.submit(alns[k], j.model)

        th = threading.Thread(target=submitter)
        th.start()
        results = []
        for k, j in enumerate(pick):
            while tickets[k] is None:
                time.sleep(0.0005)
            results.append(srv.wait(tickets[k], j.n)[0])
        th.join()
    for j, got, ref in zip(pick, results, alone):
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32),
                                      err_msg=f"job {j.job_id} ({j.n} x {j.L})")
    by_size = sorted(range(len(pick)), key=lambda k: pick[k].n * pick[k].L)
    sample = sorted({by_size[-1], len(pick) - 2, len(pick) - 1, 0, 7, 21, 40})
    for k in sample:
        j = pick[k]
        tri, _, mx = O.correct_triangle(*O.counts_triangle(alns[k], threads=host_threads()), j.n,
                                        j.model, with_full=False)
        np.testing.assert_allclose(results[k], patched(tri, mx), rtol=DIST_RTOL, atol=0,
                                   err_msg=f"job {j.job_id} ({j.n} x {j.L}, {j.model})")
        np.testing.assert_array_equal(np.signbit(results[k]), np.signbit(patched(tri, mx)))
