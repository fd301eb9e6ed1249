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
