"""The lane-parallel numeric ILU factorisation (one wavefront per forward group, one block per lane, the updates
of each block applied in ascending k) against the per-row kernel it replaces and against the oracle:

- bit-identical factors on the benchmark workloads at reduced size (2D Poisson with lognormal conductivity, 3D
  elasticity with lognormal Young's modulus), ILU(0) and ILU(1), natural, multicolour and block-multicolour
  numbering;
- a re-ordered system solves correctly after its first solve: the captured Krylov chunks are re-captured when
  the schedule changes (natural -> RCM -> block multicolour on the same system), checked on the true residual.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from delfem_amd import _native as N
    from delfem_amd import workloads as W

    ctx = N.Context(0)
    yield N, W, ctx
    ctx.close()


def _factor(s, row_kernel):
    if row_kernel:
        os.environ["DFM_ILU_ROWKERNEL"] = "1"
    try:
        s.factor()
        return s.get_factor()
    finally:
        os.environ.pop("DFM_ILU_ROWKERNEL", None)


@pytest.mark.parametrize("wname,size", [("c2", 48), ("c3", 7)])
@pytest.mark.parametrize("lev", [0, 1])
@pytest.mark.parametrize("oname", ["natural", "multicolor", "blockcolor"])
def test_lane_factor_equals_row_factor(mods, wname, size, lev, oname):
    N, W, ctx = mods
    w = W.make(wname, n=size)
    kind = {"natural": N.ORDER_NATURAL, "multicolor": N.ORDER_MULTICOLOR, "blockcolor": N.ORDER_BLOCKCOLOR}[oname]
    s = W.build_system(ctx, w, kind)
    if lev:
        s.set_fill_level(lev)
    W.assemble(s, w)
    a = _factor(s, False)
    b = _factor(s, True)
    for k in a:
        assert np.array_equal(a[k], b[k]), (k, wname, lev, oname)
    s.close()


def test_lane_factor_matches_oracle_c3_shape(mods):
    N, W, ctx = mods
    from oracle import pyref as P

    w = W.make("c3", n=5)
    s = W.build_system(ctx, w, N.ORDER_BLOCKCOLOR)
    W.assemble(s, w)
    g = _factor(s, False)
    order, _ = s.make_ordering(N.ORDER_BLOCKCOLOR)
    cpu = P.CpuSystem(w["ndim"], w["coords"], w["lnods"], w["len"])
    cpu.assemble(w["kind"], w["params"])
    cpu.traction(w["faces"], w["trac"])
    cpu.set_bc(w["bc"])
    f = cpu.ilu0(order)
    assert np.array_equal(g["fcrs"], f["fcrs"]) and np.array_equal(g["fdia"], f["fdia"])
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reorder_after_solve(mods, prec):
    N, W, ctx = mods
    p = {"fp64": N.FP64, "mixed": N.MIXED}[prec]
    w = W.make("c2", n=40)
    s = N.System(ctx, w["ndim"], w["coords"], w["lnods"], w["len"])
    s.set_bc(w["bc"])
    for kind in (N.ORDER_NATURAL, N.ORDER_RCM, N.ORDER_BLOCKCOLOR, N.ORDER_NATURAL):
        order, _ = s.make_ordering(kind)
        s.set_ordering(order)
        W.assemble(s, w)
        s.factor()
        ok, conv, it = s.solve_pcg(1e-9, 5000, p)
        assert ok, (kind, conv, it)
        assert s.true_residual() < 2e-9, (kind, s.true_residual())
    s.close()
