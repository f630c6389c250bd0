"""The C2 (2D Poisson, lognormal(0, 1) conductivity) and C3 (3D elasticity, lognormal(0, 0.5) Young's modulus, seeded
top-face traction) workloads on the GPU: at reduced size against the oracle (assembly with the traction loads and
the ILU(0) factor bit for bit under the block-colour ordering, PCG iteration counts, the mixed solution's true
residual with the oracle's matrix and its distance from the FP64 solution), and at their full sizes through
size-independent properties (true residuals on the FP64 operator, mixed against FP64)."""
import numpy as np
import pytest

from oracle import pyref as P

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


def _system(ctx, w):
    gs = N.System(ctx, w["ndim"], w["coords"], w["lnods"], w["len"])
    order, _ = gs.make_ordering(N.ORDER_BLOCKCOLOR)
    gs.set_ordering(order)
    gs.set_bc(w["bc"])
    W.assemble(gs, w)
    return gs, order


def _solve_pair(gs, tol, max_iter, mixed_mode):
    """FP64 (reference semantics, its true residual reported), FP64 with the true-residual check, and the mixed mode
    from the same right-hand side: (conv, it, x, true residual) per mode."""
    b = gs.get_vector(-1)
    out = {}
    for name, mode in [("fp64", N.FP64), ("fp64_true", N.FP64_TRUE), ("mixed", mixed_mode)]:
        gs.set_vector(-1, b)
        ok, conv, it = gs.solve_pcg(tol, max_iter, mode)
        assert ok, (name, conv, it)
        out[name] = (conv, it, gs.get_vector(-2), gs.true_residual())
    return b, out


@pytest.mark.parametrize("name,n", [("c2", 128), ("c3", 12)])
def test_workload_reduced_vs_oracle(ctx, name, n):
    w = W.make(name, n=n)
    gs, order = _system(ctx, w)
    cpu = P.CpuSystem(w["ndim"], w["coords"], w["lnods"], w["len"])
    cpu.assemble(w["kind"], w["params"])
    if w["faces"] is not None:
        cpu.traction(w["faces"], w["trac"])
    cpu.set_bc(w["bc"])
    # assembly, traction loads and Dirichlet treatment: bit for bit
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs), np.abs(crs - cpu.crs).max()
    assert np.array_equal(dia, cpu.dia), np.abs(dia - cpu.dia).max()
    assert np.array_equal(gs.get_vector(-1), cpu.res), np.abs(gs.get_vector(-1) - cpu.res).max()
    if w["faces"] is not None:
        assert np.abs(cpu.res).max() > 0.0
    # ILU(0) under the block-colour ordering: bit for bit
    gs.factor()
    f = cpu.ilu0(order)
    g = gs.get_factor()
    assert np.array_equal(g["fcrs"], f["fcrs"]), np.abs(g["fcrs"] - f["fcrs"]).max()
    assert np.array_equal(g["fdia"], f["fdia"])
    tol = w["tol"]
    ok, conv, it, x_cpu, _ = cpu.solve(f, "pcg", tol, 5000)
    assert ok
    b, out = _solve_pair(gs, tol, 5000, N.MIXED)
    assert np.array_equal(b, cpu.res)
    _, it64, x64, _ = out["fp64"]
    assert abs(it64 - it) <= 2, (it64, it)
    nb = np.linalg.norm(b)
    for mode in ["fp64_true", "mixed"]:
        conv_m, _, xm, tr = out[mode]
        # the true residual with the oracle's FP64 matrix
        r = b - cpu.matvec(1.0, xm, 0.0, np.zeros_like(b))
        rel = np.linalg.norm(r) / nb
        assert rel <= 1.01 * tol, (mode, rel)
        assert abs(rel - tr) <= 1e-3 * tol, (mode, rel, tr)
        assert conv_m == pytest.approx(tr, rel=1e-6)
    err = np.linalg.norm(out["mixed"][2] - x64) / np.linalg.norm(x64)
    err_cpu = np.linalg.norm(x64 - x_cpu) / np.linalg.norm(x_cpu)
    print(f"{name} n={n}: PCG {it64} (oracle {it}) iterations, mixed error vs FP64 {err:.2e}, FP64 vs oracle "
          f"{err_cpu:.2e}")
    assert err <= 1e-5 and err_cpu <= 1e-5


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_workload_full_size_properties(ctx, name):
    w = W.make(name)
    gs, _ = _system(ctx, w)
    gs.factor()
    tol = w["tol"]
    _, out = _solve_pair(gs, tol, 20000, N.MIXED)
    for mode in ["fp64_true", "mixed"]:
        assert out[mode][3] <= tol, (mode, out[mode][3])
    # the reference's recursive test stops near, not below, the tolerance on the true residual
    assert out["fp64"][3] <= 20 * tol, out["fp64"][3]
    x64 = out["fp64_true"][2]
    err = np.linalg.norm(out["mixed"][2] - x64) / np.linalg.norm(x64)
    print(f"{name}: iterations fp64 {out['fp64'][1]} fp64_true {out['fp64_true'][1]} mixed {out['mixed'][1]}; "
          f"true residuals {out['fp64'][3]:.2e} {out['fp64_true'][3]:.2e} {out['mixed'][3]:.2e}; mixed error vs "
          f"FP64 {err:.2e}")
    assert err <= 1e-5
    gs.close()
