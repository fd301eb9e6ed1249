"""Captured batched bootstrap (HIP graph): every replay equals the eager launch bitwise, with
eager work and synchronisation between replays, new weights per replay, short replicates."""
import numpy as np
import pytest

from rapidnj_amd.workloads import bootstrap_weights, expand_columns, simulate_alignment

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


@pytest.mark.parametrize("model,layout", [("k2p", "triangle"), ("jc", "square")])
def test_replays_equal_eager(ctx, model, layout):
    import torch
    aln = simulate_alignment(150, 1300, 0.05, 12)
    A = ctx.alignment(aln)
    R = 6
    sets = [bootstrap_weights(1300, R, seed0=s) for s in (0, 50, 90)]
    with ctx.bootstrap_plan(A, R, model, layout=layout) as plan:
        for rep in range(3):
            for w in sets:
                got = plan.run(w).clone()
                ref = ctx.bootstrap(A, w, model, layout=layout)
                torch.cuda.synchronize()
                assert torch.equal(got, ref), f"replay {rep}"
                # unrelated eager work between replays
                ctx.distances(A, "jc").sum().item()


def test_replicates_shorter_than_the_plan(ctx):
    aln = simulate_alignment(80, 700, 0.1, 3)
    A = ctx.alignment(aln)
    w = bootstrap_weights(700, 3, seed0=7)
    w[1] = 0
    w[1, :10] = 5                      # 50 columns
    with ctx.bootstrap_plan(A, 3, "k2p") as plan:
        got = plan.run(w).cpu().numpy()
    for r in range(3):
        ref = ctx.distances(ctx.alignment(expand_columns(aln, w[r])), "k2p").cpu().numpy()
        np.testing.assert_array_equal(got[r].view(np.uint32), ref.view(np.uint32))


def test_weights_beyond_the_plan_are_rejected(ctx, N):
    aln = simulate_alignment(40, 300, 0.0, 4)
    A = ctx.alignment(aln)
    with ctx.bootstrap_plan(A, 2, "jc", max_length=300) as plan:
        w = np.ones((2, 300), dtype=np.uint16)
        w[0, 0] = 2                     # 301 columns
        with pytest.raises(N.RnjError) as e:
            plan.run(w)
        assert e.value.status == N.RNJ_E_INVALID
        plan.run(np.ones((2, 300), dtype=np.uint16))


# ---- captured multi-alignment launches ---------------------------------------------------------
BATCH = [(300, 4000, 0.05, "k2p"), (57, 1001, 0.0, "jc"), (1200, 250, 0.2, "k2p"),
         (64, 30000, 0.08, "jc"), (2, 33, 0.0, "k2p")]


def _batch_alignments(ctx, seed0):
    alns = [simulate_alignment(n, L, g, seed0 + k) for k, (n, L, g, _) in enumerate(BATCH)]
    return alns, [ctx.alignment(a) for a in alns], [m for *_, m in BATCH]


@pytest.mark.parametrize("layout", ["triangle", "square"])
def test_batch_plan_uneven_alignments_every_replay(ctx, layout):
    """Five alignments of different taxa and site counts, JC and K2P, in one captured launch:
    every replay (with eager work between replays) equals each alignment computed alone."""
    import torch
    alns, A, models = _batch_alignments(ctx, 40)
    refs = [ctx.distances(a, m, layout=layout).cpu().numpy() for a, m in zip(A, models)]
    with ctx.batch_plan(A, models, layout=layout) as plan:
        for rep in range(3):
            for o in plan.outs:
                o.fill_(7.0)
            plan.run()
            torch.cuda.synchronize()
            for o, r in zip(plan.outs, refs):
                np.testing.assert_array_equal(o.cpu().numpy().view(np.uint32), r.view(np.uint32))
            ctx.distances(A[0], "jc")  # eager work between replays
            torch.cuda.synchronize()


def test_batch_plan_sees_updated_characters(ctx):
    """rnj_alignment_update_chars repacks new characters in place; the next replay computes them
    (host characters, and device characters on a side stream ordered before the replay)."""
    import torch
    alns, A, models = _batch_alignments(ctx, 60)
    with ctx.batch_plan(A, models) as plan:
        plan.run()
        torch.cuda.synchronize()
        n, L, g, m = BATCH[0]
        new0 = simulate_alignment(n, L, 0.1, 999)
        n3, L3, g3, m3 = BATCH[3]
        new3 = simulate_alignment(n3, L3, 0.02, 998)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):  # the updates and the replay in order on one stream
            A[0].update(new0, stream=s)
            d3 = torch.from_numpy(new3).to("cuda:0")
            A[3].update(d3, stream=s)
            plan.run(stream=s)
        s.synchronize()
        ref0 = ctx.distances(ctx.alignment(new0), m).cpu().numpy()
        ref3 = ctx.distances(ctx.alignment(new3), m3).cpu().numpy()
        np.testing.assert_array_equal(plan.outs[0].cpu().numpy().view(np.uint32), ref0.view(np.uint32))
        np.testing.assert_array_equal(plan.outs[3].cpu().numpy().view(np.uint32), ref3.view(np.uint32))
        # the untouched alignments still give their own matrices
        ref1 = ctx.distances(A[1], models[1]).cpu().numpy()
        np.testing.assert_array_equal(plan.outs[1].cpu().numpy().view(np.uint32), ref1.view(np.uint32))
        np.testing.assert_array_equal(A[0].planes(), ctx.alignment(new0).planes())


def test_launch_job_limit(ctx, N):
    """More than 65535 jobs in one launch (the maximum and patch kernels index jobs by grid y)
    are refused before any work: replicates of a bootstrap call or plan, alignments of a batch
    plan."""
    import torch
    A = ctx.alignment(simulate_alignment(4, 40, 0.0, 5))
    R = 65536
    w = np.ones((R, 40), dtype=np.uint16)
    out = torch.empty((R, 6), dtype=torch.float32, device="cuda:0")
    with pytest.raises(N.RnjError):
        ctx.bootstrap(A, w, "k2p", out=out)
    with pytest.raises(N.RnjError):
        ctx.bootstrap_plan(A, R, "k2p", out=out)
    with pytest.raises(N.RnjError):
        ctx.batch_plan([A] * R, "jc", outs=[out[k] for k in range(R)])
    # 65535 is accepted
    got = ctx.bootstrap(A, w[:65535], "k2p", out=out[:65535]).cpu().numpy()
    ref = ctx.distances(A, "k2p").cpu().numpy()
    assert np.array_equal(got[65534].view(np.uint32), ref.view(np.uint32))


def test_batch_plan_rejections(ctx, N):
    from rapidnj_amd.workloads import simulate_protein_alignment
    dna = ctx.alignment(simulate_alignment(20, 100, 0.0, 1))
    prot = ctx.alignment(simulate_protein_alignment(20, 100, 0.0, 2))
    with pytest.raises(N.RnjError):
        ctx.batch_plan([dna, prot], "k2p")
    with pytest.raises(N.RnjError):
        prot.update(simulate_protein_alignment(20, 100, 0.0, 3))
    with pytest.raises(N.RnjError):
        ctx.batch_plan([dna], 7)  # not a model
