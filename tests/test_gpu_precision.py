"""Precision modes of the device PCG.

- DFM_FP32_ILU (FP64 loop, FP32 copy of the ILU factor as preconditioner): converges to the FP64 tolerance with
  an iteration count close to the FP64 one; the true residual (oracle matrix) meets the tolerance.
- DFM_MIXED on a problem where FP32 vectors slow the Krylov recurrence (thin clamped elasticity strip, tolerance
  1e-10): slow reliable-update cycles are detected, and the solve continues from its FP64 solution in an FP64 loop
  with the FP32 factor, which converges to the true residual an FP64 loop reaches; the switch costs at most a
  few times an FP32_ILU solve from scratch (the mixed loop's progress is only judged at reliable updates); the
  same in a batch.
"""
import numpy as np
import pytest

from fem_cases import KINDS, setup

pytestmark = pytest.mark.gpu

N = W = None


@pytest.fixture(scope="module")
def ctx():
    global N, W
    from delfem_amd import _native, workloads

    N, W = _native, workloads
    c = _native.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
def test_fp32_ilu_mode(ctx, case):
    cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
    gs.factor()
    ok64, conv64, it64 = gs.solve_pcg(1e-10, 3000, N.FP64)
    gs.set_vector(-1, cpu.res)
    ok, conv, it = gs.solve_pcg(1e-10, 3000, N.FP32_ILU)
    assert ok64 and ok and conv < 1e-10
    assert abs(it - it64) <= max(3, 0.05 * it64), (it, it64)
    x = gs.get_vector(-2)
    b = cpu.res
    r = b - cpu.matvec(1.0, x, 0.0, np.zeros_like(b))
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b)


def test_mixed_auto_resolves_by_block_size(ctx):
    """DFM_MIXED_AUTO: a scalar system solves as DFM_MIXED, a vector system as DFM_FP32_ILU, in a single solve
    and in a batch (bit-identical results)."""
    for case, want in [(KINDS[0], N.MIXED), (KINDS[3], N.FP32_ILU)]:
        cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
        gs.factor()
        ref = gs.solve_pcg(1e-8, 3000, want)
        xr = gs.get_vector(-2)
        gs.set_vector(-1, cpu.res)
        got = gs.solve_pcg(1e-8, 3000, N.MIXED_AUTO)
        assert got == ref and np.array_equal(gs.get_vector(-2), xr)
        res = []
        for prec in (want, N.MIXED_AUTO):
            gs.set_vector(-1, cpu.res)
            b = N.Batch(ctx, prec, 2)
            b.add(gs, 1e-8, 3000)
            while b.step(16):
                pass
            res.append((b.status(gs)[:3], gs.get_vector(-2)))
            b.close()
        assert res[0][0] == res[1][0] and np.array_equal(res[0][1], res[1][1])


def _strip(ctx):
    w = W.make("c1")
    s = W.build_system(ctx, w, N.ORDER_BLOCKCOLOR)
    W.assemble(s, w)
    s.factor()
    return w, s


def test_mixed_stall_falls_back(ctx):
    w, s = _strip(ctx)
    ok, conv, it = s.solve_pcg(1e-10, 20000, N.MIXED)
    nrel, fell_back = s.last_solve_info()
    true_rel = s.true_residual()  # FP64 matrix, the solution left in the update handle
    W.assemble(s, w)
    s.factor()
    ok2, conv2, it2 = s.solve_pcg(1e-10, 20000, N.FP32_ILU)
    true_rel2 = s.true_residual()
    print(f"mixed then FP64 vectors: {it} iterations ({nrel} reliable updates), true residual {true_rel:.3e}; "
          f"FP32_ILU from scratch: {it2} iterations, true residual {true_rel2:.3e}")
    assert ok and conv < 1e-10 and fell_back
    # the recursive FP64 residual of this strip drifts from the true one by about 1e-9 relative, in any FP64 loop
    assert true_rel <= max(1.01e-10, 2.0 * true_rel2)
    assert ok2 and it <= 3 * it2


def test_mixed_stall_falls_back_in_batch(ctx):
    w, s = _strip(ctx)
    ok, conv1, it1 = s.solve_pcg(1e-10, 20000, N.MIXED)
    x1 = s.get_vector(-2)
    W.assemble(s, w)
    s.factor()
    b = N.Batch(ctx, N.MIXED, 4)
    b.add(s, 1e-10, 20000)
    while b.step(64):
        pass
    st, conv, it, _ = b.status(s)
    nrel, fell_back = b.info(s)
    assert st == 1 and conv < 1e-10 and fell_back
    # the batch engine reduces in its own order (one workgroup per system), so the counts are close, not equal
    assert abs(it - it1) <= 0.02 * it1 + 5, (it, it1)
    x = s.get_vector(-2)
    assert np.linalg.norm(x - x1) <= 1e-6 * np.linalg.norm(x1)
    b.close()
