"""Batched solves (dfm_batch_*): every system of a batch keeps its own Krylov scalars, so

- its result is bit-identical whether it is solved alone or together with others, admitted at any point
  (continuous batching: systems join while others are mid-solve),
- its FP64 iterates are PCG's: iteration count within 2 of the oracle's LsSol::Solve_PCG restatement and a true
  residual at the tolerance, checked with the oracle's matrix,
- in mixed precision the FP64 true residual meets the tolerance.
"""
import numpy as np
import pytest

from fem_cases import KINDS, setup

pytestmark = pytest.mark.gpu

N = None


@pytest.fixture(scope="module")
def ctx():
    global N
    from delfem_amd import _native

    N = _native
    c = _native.Context(0)
    yield c
    c.close()


# (kind index, large mesh): scalar 2D/3D, 2x2 and 3x3 blocks, two sizes of the scalar problem
CASES = [(0, False), (0, True), (1, False), (3, False), (4, False), (6, True)]


def _systems(ctx):
    out = []
    for k, small in CASES:
        cpu, gs, order, _ = setup(N, ctx, KINDS[k], N.ORDER_BLOCKCOLOR, small=small)
        gs.factor()
        out.append((cpu, gs, order))
    return out


def _run(batch, systems, schedule, chunk=7, max_steps=2000):
    """Admit systems[i] at step schedule[i]; step until all finished. Returns (solutions, iterations, states)."""
    pending = sorted(range(len(systems)), key=lambda i: schedule[i])
    step = 0
    while True:
        while pending and schedule[pending[0]] <= step:
            i = pending.pop(0)
            batch.add(systems[i][1], 1e-8, 3000)
        nrun = batch.step(chunk)
        step += 1
        if not pending and nrun == 0:
            break
        assert step < max_steps
    res = []
    for cpu, gs, order in systems:
        st, conv, it, _ = batch.status(gs)
        res.append((gs.get_vector(-2), it, st, conv))
        batch.remove(gs)
    return res


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp32_ilu"])
def test_alone_equals_mixed_in(ctx, precision):
    prec = {"mixed": N.MIXED, "fp64": N.FP64, "fp32_ilu": N.FP32_ILU}[precision]
    systems = _systems(ctx)
    alone = []
    for s in systems:
        b = N.Batch(ctx, prec, 4)
        alone.append(_run(b, [s], [0])[0])
        b.close()
        s[1].set_vector(-1, s[0].res)  # the solve leaves its final residual in the residual handle
    b = N.Batch(ctx, prec, 16)
    mix = _run(b, systems, [0, 3, 0, 5, 1, 9])
    b.close()
    for (xa, ia, sa, ca), (xm, im, sm, cm) in zip(alone, mix):
        assert sa == sm == 1
        assert ia == im and ca == cm
        assert np.array_equal(xa, xm)


def test_fp64_batch_matches_oracle_pcg(ctx):
    systems = _systems(ctx)
    b = N.Batch(ctx, N.FP64, 16)
    res = _run(b, systems, [0, 0, 2, 2, 4, 4])
    b.close()
    for (cpu, gs, order), (x, it, st, conv) in zip(systems, res):
        f = cpu.ilu0(order)
        ok, cconv, cit, cx, _ = cpu.solve(f, "pcg", 1e-8, 3000)
        assert ok and st == 1
        assert abs(it - cit) <= 2, (it, cit)
        bb = cpu.res
        r = bb - cpu.matvec(1.0, x, 0.0, np.zeros_like(bb))
        assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(bb)
        assert np.linalg.norm(x - cx) <= 1e-4 * np.linalg.norm(cx)


def test_mixed_batch_true_residual(ctx):
    systems = _systems(ctx)
    b = N.Batch(ctx, N.MIXED, 16)
    res = _run(b, systems, [0, 1, 2, 3, 4, 5])
    b.close()
    for (cpu, gs, order), (x, it, st, conv) in zip(systems, res):
        assert st == 1 and conv < 1e-8
        bb = cpu.res
        r = bb - cpu.matvec(1.0, x, 0.0, np.zeros_like(bb))
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(bb)


def test_batch_max_iter_and_removal(ctx):
    systems = _systems(ctx)[:3]
    b = N.Batch(ctx, N.MIXED, 4)
    for _, gs, _ in systems:
        b.add(gs, 1e-14, 5)
    while b.step(3):
        pass
    for _, gs, _ in systems:
        st, conv, it, _ = b.status(gs)
        assert st == 2 and it == 5
    b.remove(systems[0][1])
    with pytest.raises(N.DfmError):
        b.status(systems[0][1])
    systems[1][1].close()  # destroying a system takes it out of its batch
    b.close()
