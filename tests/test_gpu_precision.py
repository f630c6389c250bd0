"""Precision modes of the device PCG.

- DFM_FP32_ILU (FP64 loop, FP32 copy of the ILU factor as preconditioner): converges to the FP64 tolerance with
  an iteration count close to the FP64 one; the true residual (oracle matrix) meets the tolerance.
- DFM_MIXED on a problem whose condition number defeats an FP32 matrix (thin clamped elasticity strip, tolerance
  1e-10): the stagnation of the FP64 true residual is detected and the solve is repeated with the FP32 factor
  in an FP64 loop, which converges; the same in a batch.
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
    assert ok and conv < 1e-10
    assert fell_back
    # the fallback is the FP64 loop with the FP32 factor: the iterations of an FP32_ILU solve from scratch
    W.assemble(s, w)
    s.factor()
    ok2, conv2, it2 = s.solve_pcg(1e-10, 20000, N.FP32_ILU)
    assert ok2 and it2 == it


def test_mixed_stall_falls_back_in_batch(ctx):
    w, s = _strip(ctx)
    b = N.Batch(ctx, N.MIXED, 4)
    b.add(s, 1e-10, 20000)
    while b.step(64):
        pass
    st, conv, it, _ = b.status(s)
    nrel, fell_back = b.info(s)
    assert st == 1 and conv < 1e-10 and fell_back
    b.close()
