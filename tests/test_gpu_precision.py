"""Precision modes of the device PCG.

- DFM_FP32_ILU (FP64 loop, FP32 copy of the ILU factor as preconditioner): converges to the FP64 tolerance with
  an iteration count close to the FP64 one; its conv_ratio is the true relative residual (checked at exit, the
  loop restarted from its solution while above the tolerance), and the true residual with the oracle's matrix
  meets the tolerance.
- DFM_MIXED on a problem where FP32 vectors slow the Krylov recurrence (thin clamped elasticity strip, tolerance
  1e-10): slow reliable-update cycles are detected, and the solve continues from its FP64 solution in an FP64 loop
  with the FP32 factor, which converges to the true residual an FP64 loop reaches; the switch costs at most a
  few times an FP32_ILU solve from scratch (the mixed loop's progress is only judged at reliable updates); the
  same in a batch. The reported conv_ratio is the true relative residual (on the strip the recursive residual of
  any FP64 loop drifts about 1e-9 from the true one; checked at exit, the loop restarts from its solution).
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
    assert ok64 and ok and conv <= 1e-10
    assert conv == gs.true_residual()
    assert abs(it - it64) <= max(3, 0.05 * it64), (it, it64)
    x = gs.get_vector(-2)
    b = cpu.res
    r = b - cpu.matvec(1.0, x, 0.0, np.zeros_like(b))
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)


def test_mixed_auto_resolves_by_block_size(ctx):
    """DFM_MIXED_AUTO: a scalar system solves as DFM_MIXED, a vector system as DFM_FP64_TRUE, in a single solve
    and in a batch (bit-identical results)."""
    for case, want in [(KINDS[0], N.MIXED), (KINDS[3], N.FP64_TRUE)]:
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


@pytest.mark.parametrize("tol", [3e-10, 1e-10])
def test_mixed_stall_falls_back(ctx, tol):
    """The strip's FP64 loop (recursive residual) drifts about 1e-9 from its true residual; restarts from the
    solution lower the true residual by a few tens of per cent each, down to about 2e-10 (the FP64 attainable
    accuracy of this system), and stop when one gains less than a fifth: at 3e-10 both modes converge on the true
    residual, at 1e-10 they may report non-convergence, and either way conv_ratio is the true residual."""
    w, s = _strip(ctx)
    ok, conv, it = s.solve_pcg(tol, 20000, N.MIXED)
    nrel, fell_back = s.last_solve_info()
    true_rel = s.true_residual()  # FP64 matrix, the solution left in the update handle
    W.assemble(s, w)
    s.factor()
    ok2, conv2, it2 = s.solve_pcg(tol, 20000, N.FP32_ILU)
    true_rel2 = s.true_residual()
    print(f"tol {tol:.0e}: mixed then FP64 vectors: ok {ok}, {it} iterations ({nrel} reliable updates), true residual "
          f"{true_rel:.3e}; FP32_ILU from scratch: ok {ok2}, {it2} iterations, true residual {true_rel2:.3e}")
    assert fell_back
    assert conv == true_rel and conv2 == true_rel2
    assert true_rel < 2e-9 and true_rel2 < 2e-9
    if tol >= 3e-10:
        assert ok and ok2 and true_rel <= tol and true_rel2 <= tol
        assert it <= 3 * it2
    assert ok == (true_rel <= tol) and ok2 == (true_rel2 <= tol)


def test_mixed_stall_falls_back_in_batch(ctx):
    tol = 3e-10
    w, s = _strip(ctx)
    ok, conv1, it1 = s.solve_pcg(tol, 20000, N.MIXED)
    x1 = s.get_vector(-2)
    W.assemble(s, w)
    s.factor()
    b = N.Batch(ctx, N.MIXED, 4)
    b.add(s, tol, 20000)
    while b.step(64):
        pass
    st, conv, it, _ = b.status(s)
    nrel, fell_back = b.info(s)
    assert st == 1 and conv <= tol and fell_back
    assert conv == s.true_residual()
    # the batch engine reduces in its own order (one workgroup per system), so the counts are close, not equal
    assert abs(it - it1) <= 0.02 * it1 + 5, (it, it1)
    x = s.get_vector(-2)
    assert np.linalg.norm(x - x1) <= 1e-6 * np.linalg.norm(x1)
    b.close()


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
def test_fp64_true_mode(ctx, case):
    """DFM_FP64_TRUE: the reference loop in FP64, its true residual checked at exit. Where the recursive and the
    true residual both meet the tolerance it is the DFM_FP64 solve itself (same iterates, same count); its
    conv_ratio is the true residual."""
    cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
    gs.factor()
    ok64, conv64, it64 = gs.solve_pcg(1e-9, 3000, N.FP64)
    x64 = gs.get_vector(-2)
    true64 = gs.true_residual()
    gs.set_vector(-1, cpu.res)
    ok, conv, it = gs.solve_pcg(1e-9, 3000, N.FP64_TRUE)
    assert ok and conv == gs.true_residual() and conv <= 1e-9
    if true64 <= 1e-9:
        assert it == it64 and np.array_equal(gs.get_vector(-2), x64)
    else:
        assert it > it64


def test_fp64_true_meets_where_fp64_does_not(ctx):
    """The thin clamped strip at 3e-10: the reference's recursive test stops with the true residual above the
    tolerance; DFM_FP64_TRUE restarts and meets it, alone and in a batch (the same result)."""
    w, s = _strip(ctx)
    ok64, conv64, it64 = s.solve_pcg(3e-10, 20000, N.FP64)
    true64 = s.true_residual()
    W.assemble(s, w)
    s.factor()
    ok, conv, it = s.solve_pcg(3e-10, 20000, N.FP64_TRUE)
    x = s.get_vector(-2)
    print(f"FP64: {it64} iterations, recursive {conv64:.2e}, true {true64:.2e}; FP64_TRUE: {it} iterations, "
          f"true {conv:.2e}")
    assert ok64 and ok and true64 > 3e-10 and conv <= 3e-10 and conv == s.true_residual() and it > it64
    W.assemble(s, w)
    s.factor()
    b = N.Batch(ctx, N.FP64_TRUE, 2)
    b.add(s, 3e-10, 20000)
    while b.step(32):
        pass
    st, cb, itb, _ = b.status(s)
    assert st == 1 and cb <= 3e-10 and cb == s.true_residual()
    b.close()


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
@pytest.mark.parametrize("mode", ["fp32_ilu", "fp64_true", "mixed"])
def test_pbicgstab_precision(ctx, case, mode):
    """PBiCGSTAB's precision argument: the FP32 factor (fp32_ilu, and mixed, which runs as fp32_ilu) or the FP64
    factor, with the true residual checked and restarted on b - A x: its conv_ratio is the true residual, at most
    the tolerance, and the solution agrees with the FP64 solve's to within the tolerance times cond(A)."""
    prec = {"fp32_ilu": N.FP32_ILU, "fp64_true": N.FP64_TRUE, "mixed": N.MIXED}[mode]
    cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
    gs.factor()
    ok64, conv64, it64 = gs.solve_pbicgstab(1e-9, 3000, N.FP64)
    x64 = gs.get_vector(-2)
    assert ok64 and conv64 < 1e-9
    gs.set_vector(-1, cpu.res)
    ok, conv, it = gs.solve_pbicgstab(1e-9, 3000, prec)
    tr = gs.true_residual()
    print(f"{mode}: FP64 {it64} iterations; {it} iterations, conv {conv:.2e}, true {tr:.2e}")
    assert ok and conv == tr and conv <= 1e-9
    x = gs.get_vector(-2)
    assert np.linalg.norm(x - x64) <= 1e-4 * np.linalg.norm(x64)
    # the reference's contract on the FP64 loop: not converged within max_iter -> true, iteration = max_iter
    gs.set_vector(-1, cpu.res)
    ok, conv, it = gs.solve_pbicgstab(1e-14, 3, N.FP64)
    assert ok and it == 3 and conv == 1e-14
