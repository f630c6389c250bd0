"""ILU(k) on the device (dfm_prec_set_fill_level, the reference's CPreconditioner_ILU::SetFillInLevel) against the
oracle's level-of-fill restatement (itself pinned bit-exact to the reference's AddFracPtn + DoILUDecomp):

- the factor pattern and values are bit-exact, natural and block-multicolour numbering, levels 1 and 2, block
  sizes 1, 2, 3; fill levels change the schedule (fill can link tiles of one colour), not the arithmetic;
- the preconditioner apply is bit-exact;
- PCG with ILU(1) matches the oracle's Solve_PCG with the same factor (iteration count within 2, solution within
  the tolerance), and needs no more iterations than ILU(0); the mixed-precision and batched solves converge;
- the assembled matrix survives a change of fill level (no re-assembly), and bad levels are refused.
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


CASES = [KINDS[0], KINDS[3], KINDS[4]]
ORDERS = [("natural", 0), ("blockcolor", 2)]


@pytest.mark.parametrize("case", CASES, ids=["b1", "b2", "b3"])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
@pytest.mark.parametrize("lev", [1, 2])
def test_iluk_factor_and_apply_bit_exact(ctx, case, oname, ordering, lev):
    cpu, gs, order, _ = setup(N, ctx, case, ordering)
    gs.set_fill_level(lev)
    gs.factor()
    f = cpu.ilu(None if ordering == N.ORDER_NATURAL else order, lev)
    g = gs.get_factor()
    assert g["fcolind"].size > cpu.colind.size  # fill
    for k in ["frowptr", "fcolind", "diaind"]:
        assert np.array_equal(g[k], f[k]), k
    assert np.array_equal(g["fcrs"], f["fcrs"]), np.abs(g["fcrs"] - f["fcrs"]).max()
    assert np.array_equal(g["fdia"], f["fdia"])
    r = np.random.default_rng(2).normal(size=cpu.nnode * cpu.len)
    gs.resize_tmp(1)
    gs.set_vector(0, r)
    gs.prec_apply(0)
    assert np.array_equal(gs.get_vector(0), cpu.ilu_apply(f, r))


@pytest.mark.parametrize("case", CASES, ids=["b1", "b2", "b3"])
def test_iluk_pcg_matches_oracle(ctx, case):
    cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
    gs.factor()
    ok0, _, it0 = gs.solve_pcg(1e-8, 3000, N.FP64)
    gs.set_vector(-1, cpu.res)
    gs.set_fill_level(1)
    gs.factor()
    ok, conv, it = gs.solve_pcg(1e-8, 3000, N.FP64)
    x = gs.get_vector(-2)
    f = cpu.ilu(order, 1)
    cok, cconv, cit, cx, _ = cpu.solve(f, "pcg", 1e-8, 3000)
    assert ok and cok and ok0
    assert abs(it - cit) <= 2, (it, cit)
    assert it <= it0, (it, it0)
    b = cpu.res
    r = b - cpu.matvec(1.0, x, 0.0, np.zeros_like(b))
    assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(b)
    assert np.linalg.norm(x - cx) <= 1e-5 * np.linalg.norm(cx)
    for prec in (N.MIXED, N.FP32_ILU):
        gs.set_vector(-1, b)
        okp, convp, itp = gs.solve_pcg(1e-8, 6000, prec)
        assert okp and convp < 1e-8
        xp = gs.get_vector(-2)
        rp = b - cpu.matvec(1.0, xp, 0.0, np.zeros_like(b))
        assert np.linalg.norm(rp) <= 2e-8 * np.linalg.norm(b)


def test_iluk_in_batch(ctx):
    systems = []
    for case in CASES:
        cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
        gs.set_fill_level(1)
        gs.factor()
        systems.append((cpu, gs))
    b = N.Batch(ctx, N.MIXED, 8)
    for _, gs in systems:
        b.add(gs, 1e-8, 6000)
    while b.step(16):
        pass
    for cpu, gs in systems:
        st, conv, it, _ = b.status(gs)
        assert st == 1 and conv < 1e-8
        x = gs.get_vector(-2)
        r = cpu.res - cpu.matvec(1.0, x, 0.0, np.zeros_like(cpu.res))
        assert np.linalg.norm(r) <= 2e-8 * np.linalg.norm(cpu.res)
        b.remove(gs)
    b.close()


def test_fill_level_keeps_matrix_and_refuses_bad_levels(ctx):
    cpu, gs, order, _ = setup(N, ctx, KINDS[3], N.ORDER_BLOCKCOLOR)
    crs0, dia0 = gs.get_matrix()
    gs.set_fill_level(2)
    crs1, dia1 = gs.get_matrix()
    assert np.array_equal(crs0, crs1) and np.array_equal(dia0, dia1)
    m = cpu.nnode * cpu.len
    x = np.random.default_rng(5).normal(size=m)
    gs.resize_tmp(2)
    gs.set_vector(0, x)
    gs.matvec(1.0, 0, 0.0, 1)
    assert np.allclose(gs.get_vector(1), cpu.matvec(1.0, x, 0.0, np.zeros(m)), rtol=1e-12, atol=1e-14)
    with pytest.raises(N.DfmError):
        gs.prec_apply(0)  # the factor must be computed again
    with pytest.raises(N.DfmError):
        gs.set_fill_level(-2)
    gs.set_fill_level(0)
    gs.factor()
    g = gs.get_factor()
    assert np.array_equal(g["fcolind"], cpu.ilu0(order)["fcolind"])
