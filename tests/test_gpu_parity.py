"""HIP product (through the C-ABI) vs the oracle restatement, on the same seeded inputs.

Bit-exact: pattern, assembled matrix and residual (every element kind, per-element materials, natural and
permuted device numbering), ILU(0) factor pattern and values under the same ordering, forward/backward
substitution under every ordering. Tolerance-based: SpMV and dot products
(parallel summation order; rel 1e-13), Krylov solutions (iteration count within 2 of the oracle's, solution
error bounded by the residual tolerance)."""
import numpy as np
import pytest

import meshgen as M  # noqa: F401
from fem_cases import KINDS, mesh as _mesh, setup
from oracle import pyref as P

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


# the last three: COrdering_Blk's orderings as the reference computes them (level-set sweeps)
ORDERS = [("natural", 0), ("multicolor", 1), ("blockcolor", 2), ("rcm", 3), ("rcm2", 4), ("amd", 5)]


@pytest.mark.parametrize("case", KINDS, ids=[k[0] for k in KINDS])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
def test_assembly_bit_exact(ctx, case, oname, ordering):
    cpu, gs, order, rn = setup(N, ctx, case, ordering)
    rp, ci = gs.pattern()
    assert np.array_equal(rp, cpu.rowptr) and np.array_equal(ci, cpu.colind)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs), np.abs(crs - cpu.crs).max()
    assert np.array_equal(dia, cpu.dia), np.abs(dia - cpu.dia).max()
    assert np.array_equal(gs.get_vector(-1), cpu.res)
    assert abs(rn - np.sqrt(np.sum(cpu.res ** 2))) <= 1e-14 * max(rn, 1e-300)


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
def test_ilu0_factor_bit_exact(ctx, case, oname, ordering):
    cpu, gs, order, _ = setup(N, ctx, case, ordering)
    gs.factor()
    f = cpu.ilu0(None if ordering == N.ORDER_NATURAL else order)
    g = gs.get_factor()
    for k in ["frowptr", "fcolind", "diaind"]:
        assert np.array_equal(g[k], f[k]), k
    assert np.array_equal(g["fcrs"], f["fcrs"]), np.abs(g["fcrs"] - f["fcrs"]).max()
    assert np.array_equal(g["fdia"], f["fdia"])
    # preconditioner apply, handle semantics (in place on a temp handle)
    r = np.random.default_rng(2).normal(size=cpu.nnode * cpu.len)
    gs.resize_tmp(1)
    gs.set_vector(0, r)
    gs.prec_apply(0)
    z = gs.get_vector(0)
    zc = cpu.ilu_apply(f, r)
    # every ordering, block multicolour included: the sweeps apply each row's entries in ascending column order
    assert np.array_equal(z, zc), np.abs(z - zc).max()


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
def test_handle_ops(ctx, case):
    cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR)
    m = cpu.nnode * cpu.len
    rng = np.random.default_rng(4)
    x, y = rng.normal(size=m), rng.normal(size=m)
    gs.resize_tmp(2)
    gs.set_vector(0, x)
    gs.set_vector(1, y)
    assert np.isclose(gs.dot(0, 1), x @ y, rtol=1e-13)
    gs.matvec(-0.7, 0, 1.3, 1)
    yc = cpu.matvec(-0.7, x, 1.3, y)
    assert np.allclose(gs.get_vector(1), yc, rtol=1e-13, atol=1e-13 * np.abs(yc).max())
    gs.axpy(0.25, 0, 1)
    assert np.allclose(gs.get_vector(1), yc + 0.25 * x, rtol=1e-13, atol=1e-13 * np.abs(yc).max())
    gs.scal(2.0, 0)
    assert np.array_equal(gs.get_vector(0), 2.0 * x)
    gs.copy(0, 1)
    assert np.array_equal(gs.get_vector(1), 2.0 * x)


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
@pytest.mark.parametrize("method", ["pcg", "pbicgstab"])
def test_krylov_vs_oracle(ctx, case, oname, ordering, method):
    cpu, gs, order, _ = setup(N, ctx, case, ordering, small=False)
    f = cpu.ilu0(None if ordering == N.ORDER_NATURAL else order)
    ok, conv, it, x, _ = cpu.solve(f, method, 1e-10, 2000)
    gs.factor()
    if method == "pcg":
        gok, gconv, git = gs.solve_pcg(1e-10, 2000)
    else:
        gok, gconv, git = gs.solve_pbicgstab(1e-10, 2000)
    assert ok and gok
    assert abs(git - it) <= 2, (git, it)
    assert gconv < 1e-10
    xg = gs.get_vector(-2)
    # the GPU solution's true residual, evaluated with the oracle's matrix (size- and rounding-independent)
    b = cpu.res
    r = b - cpu.matvec(1.0, xg, 0.0, np.zeros_like(b))
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(b)
    # both solutions are within cond(A) * tol of the exact one (cond ~ 1e4-1e5 for these elasticity meshes)
    assert np.linalg.norm(xg - x) <= 1e-5 * np.linalg.norm(x)


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
def test_mixed_precision_true_residual(ctx, case):
    cpu, gs, order, _ = setup(N, ctx, case, N.ORDER_BLOCKCOLOR, small=False)
    gs.factor()
    ok, conv, it = gs.solve_pcg(1e-8, 5000, N.FP64)
    x64 = gs.get_vector(-2)
    assert ok
    gs2 = gs
    gs2.set_vector(-1, cpu.res)
    ok, conv, it = gs2.solve_pcg(1e-8, 5000, N.MIXED)
    assert ok and conv < 1e-8
    assert gs2.true_residual() < 1e-8
    xm = gs2.get_vector(-2)
    # independent check with the oracle's FP64 matrix: b - A xm
    b = cpu.res
    r = b - cpu.matvec(1.0, xm, 0.0, np.zeros_like(b))
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b)
    # against the exact solution of the oracle's matrix (dense LU): each error within what its relative residual
    # allows, kappa_2(A) * ||b - A x|| / ||b||, and the mixed solution no further from it than the FP64 one by more
    # than that bound
    A = cpu.dense()
    xs = np.linalg.solve(A, b)
    sv = np.linalg.svd(A, compute_uv=False)
    kappa = sv[0] / sv[-1]
    for xx in (xm, x64):
        rel = np.linalg.norm(b - A @ xx) / np.linalg.norm(b)
        assert np.linalg.norm(xx - xs) <= 1.01 * kappa * rel * np.linalg.norm(xs)
    e_m = np.linalg.norm(xm - xs) / np.linalg.norm(xs)
    e_64 = np.linalg.norm(x64 - xs) / np.linalg.norm(xs)
    print(f"mixed error {e_m:.2e}, fp64 error {e_64:.2e}, kappa {kappa:.2e}", flush=True)
    assert e_m <= e_64 + kappa * 1.01e-8
    # mixed precision must not cost many more iterations than FP64 (reliable updates keep the Krylov space)
    _, _, it64 = gs.solve_pcg(1e-8, 5000, N.FP64)
    assert it <= 1.5 * it64 + 10, (it, it64)


@pytest.mark.parametrize("oname,ordering", ORDERS[:3], ids=[o[0] for o in ORDERS[:3]])
@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
def test_update_field_bit_exact(ctx, case, oname, ordering):
    """UpdateValueOfField (src/femls/linearsystem_field.cpp:716-794): value += 1.0 * update per dof, which is
    value + update exactly (the reference's CNodeAry::AddValueToNodeSegment, src/femfield/node_ary.cpp:223-260)."""
    cpu, gs, order, _ = setup(N, ctx, case, ordering)
    n = cpu.nnode * cpu.len
    rng = np.random.default_rng(41)
    u = rng.standard_normal(n)
    du = rng.standard_normal(n) * 1e-3
    gs.set_field(u)
    gs.set_vector(-2, du)
    gs.update_field()
    assert np.array_equal(gs.get_field(), u + du)


def test_pcg_not_converged_contract(ctx):
    cpu, gs, order, _ = setup(N, ctx, KINDS[0], N.ORDER_NATURAL, small=False)
    gs.factor()
    ok, conv, it = gs.solve_pcg(1e-14, 3)
    f = cpu.ilu0()
    cok, cconv, cit, _, _ = cpu.solve(f, "pcg", 1e-14, 3)
    assert not ok and not cok and it == cit == 3
    assert np.isclose(conv, cconv, rtol=1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_max_iter_at_chunk_boundaries(ctx, precision):
    """The device loop enqueues iterations in chunks of one or two 16-iteration graphs and reads the status of a
    chunk while the next runs; max_iter must still end the loop exactly as the reference's counter does
    (src/ls/solver_ls_iter.cpp:134-171): converged iff the converging iteration index is below max_iter (the
    index returned), else max_iter returned with the last residual ratio, whatever max_iter is against the
    chunk length. Unconverged FP64 ratios agree with the oracle's loop at the same max_iter."""
    prec = N.FP64 if precision == "fp64" else N.MIXED
    cpu, gs, order, _ = setup(N, ctx, KINDS[0], N.ORDER_NATURAL, small=False)
    gs.factor()
    b = gs.get_vector(-1)  # the solve leaves its final residual in the residual handle, as the reference's does
    tol = 1e-10
    ok, conv, git = gs.solve_pcg(tol, 2000, prec)
    assert ok and git > 17, git
    f = cpu.ilu0()
    near = [1, 2, 15, 16, 17, 31, 32, 33, 47, 48, 49]
    for maxit in sorted({m for m in near if m < git} | {git - 1, git, git + 1, git + 2, git + 17}):
        gs.set_vector(-1, b)
        ok2, conv2, it2 = gs.solve_pcg(tol, maxit, prec)
        if maxit > git:
            assert ok2 and it2 == git and conv2 == conv, (maxit, ok2, it2, conv2, conv)
        else:
            assert not ok2 and it2 == maxit, (maxit, ok2, it2)
            if precision == "fp64" and maxit <= 49:
                cok, cconv, cit, _, _ = cpu.solve(f, "pcg", tol, maxit)
                assert not cok and cit == maxit
                assert np.isclose(conv2, cconv, rtol=1e-6), (maxit, conv2, cconv)


def test_singular_factor_reports(ctx):
    """An unconstrained Poisson matrix with zero conductivity: every pivot is singular (DoILUDecomp false)."""
    co, ln = _mesh(2)
    gs = N.System(ctx, 2, co, ln, 1)
    gs.set_bc(np.zeros(co.size // 2, np.int32))
    gs.assemble(N.POISSON, [0.0, 1.0])
    with pytest.raises(N.DfmError) as e:
        gs.factor()
    assert e.value.code == N.ERR_SINGULAR


@pytest.mark.parametrize("kind", [3, 4, 5], ids=["rcm", "rcm2", "amd"])
def test_reference_orderings_of_a_system(ctx, kind):
    # dfm_sys_make_ordering of the reference's kinds equals the restatement on the system's pattern
    cpu, gs, order, _ = setup(N, ctx, KINDS[0], N.ORDER_NATURAL)
    o, _ = gs.make_ordering(kind)
    assert np.array_equal(o, P.ordering(kind, cpu.rowptr, cpu.colind))


def _fan(ndim, k, seed):
    """k triangles around one centre node (2D), or k tetrahedra around an axis of two nodes (3D): the centre (axis)
    rows hold k (k + 1) off-diagonal blocks, past the row-group kernel's lanes for k >= 64."""
    rng = np.random.default_rng(seed)
    ang = (np.arange(k) + rng.uniform(-0.2, 0.2, k)) * (2 * np.pi / k)
    rad = rng.uniform(0.8, 1.2, k)
    ring = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    if ndim == 2:
        co = np.concatenate([[[0.0, 0.0]], ring]).ravel()
        ln = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32).ravel()
    else:
        co = np.concatenate([[[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]],
                             np.concatenate([ring, rng.uniform(-0.1, 0.1, (k, 1))], 1)]).ravel()
        ln = np.array([[0, 1, 2 + i, 2 + (i + 1) % k] for i in range(k)], np.int32).ravel()
    return co, ln


@pytest.mark.parametrize("ndim,k", [(2, 20), (2, 40), (2, 70), (3, 20), (3, 40), (3, 70)])
@pytest.mark.parametrize("oname,ordering", [("natural", 0), ("blockcolor", 2)])
def test_assembly_high_valence_bit_exact(ctx, ndim, k, oname, ordering):
    """Linear elasticity on rows of every lane count of the row-group kernel (16, 32, 64 lanes) and on rows too long
    for it (the geometry store and per-entry gathers)."""
    co, ln = _fan(ndim, k, 40 + k)
    case = KINDS[3] if ndim == 2 else KINDS[4]
    _, _, length, kind, pfun, _, _ = case
    n = co.size // ndim
    rng = np.random.default_rng(k)
    params = pfun(rng, ln.size // (ndim + 1))
    u = rng.uniform(-0.05, 0.05, n * length)
    fl = np.zeros(n * length, np.int32)
    fl[-length:] = 1
    cpu = P.CpuSystem(ndim, co, ln, length)
    cpu.assemble(kind, params, u, None, (0.05, 0.6))
    cpu.set_bc(fl)
    gs = N.System(ctx, ndim, co, ln, length)
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, None)
    gs.assemble(kind, params, glob=(0.05, 0.6))
    rp, _ = gs.pattern()
    assert np.diff(rp).max() == (k if ndim == 2 else k + 1)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs), np.abs(crs - cpu.crs).max()
    assert np.array_equal(dia, cpu.dia), np.abs(dia - cpu.dia).max()
    assert np.array_equal(gs.get_vector(-1), cpu.res)
    gs.close()


@pytest.mark.parametrize("case", [KINDS[0], KINDS[3], KINDS[4]], ids=["b1", "b2", "b3"])
@pytest.mark.parametrize("oname,ordering", [("natural", 0), ("multicolor", 1), ("rcm", 3)])
def test_system_from_block_pattern(ctx, case, oname, ordering):
    """dfm_sys_create_pattern + dfm_sys_set_matrix (CMatDia_BlkCrs::AddPattern(crs) and the values of host-side
    element routines): the oracle's assembled matrix installed on a mesh-free system has the oracle's pattern and
    ILU factor bit for bit and solves as the mesh system that assembled it on the device (the fused dot partials
    of the sweeps may group differently without the mesh: iterates equal up to rounding)."""
    cpu, gs, order, _ = setup(N, ctx, case, ordering, small=False)
    ps = N.System.from_pattern(ctx, cpu.nnode, cpu.len, cpu.rowptr, cpu.colind)
    rp, ci = ps.pattern()
    assert np.array_equal(rp, cpu.rowptr) and np.array_equal(ci, cpu.colind)
    if ordering != N.ORDER_NATURAL:
        ps.set_ordering(order)
    ps.set_matrix(cpu.crs, cpu.dia)
    crs, dia = ps.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    ps.set_vector(-1, cpu.res)
    ps.factor()
    gs.factor()
    f = cpu.ilu0(None if ordering == N.ORDER_NATURAL else order)
    assert np.array_equal(ps.get_factor()["fcrs"], f["fcrs"])
    ok1, c1, i1 = ps.solve_pcg(1e-9, 3000)
    ok2, c2, i2 = gs.solve_pcg(1e-9, 3000)
    assert ok1 and ok2 and abs(i1 - i2) <= 1 and c1 < 1e-9
    x1, x2 = ps.get_vector(-2), gs.get_vector(-2)
    assert np.linalg.norm(x1 - x2) <= 1e-8 * np.linalg.norm(x2)
    r = cpu.res - cpu.matvec(1.0, x1, 0.0, np.zeros_like(x1))
    assert np.linalg.norm(r) <= 1.01e-9 * np.linalg.norm(cpu.res)
    with pytest.raises(N.DfmError):
        ps.assemble(case[3], np.ones(8))  # no elements behind a pattern system
    ps.close()


def test_block_pattern_rejects_bad_input(ctx):
    rp = np.array([0, 1, 2, 2], np.uint32)
    # row 1 holding its diagonal, row 0 holding its diagonal, a column out of range
    for ci in ([1, 1], [0, 2], [1, 3]):
        with pytest.raises(N.DfmError):
            N.System.from_pattern(ctx, 3, 1, rp, np.array(ci, np.uint32))
    with pytest.raises(N.DfmError):  # a repeated column
        N.System.from_pattern(ctx, 3, 1, np.array([0, 2, 2, 2], np.uint32), np.array([1, 1], np.uint32))
    s = N.System.from_pattern(ctx, 3, 2, np.array([0, 1], np.uint32), np.array([2], np.uint32))  # rows 1, 2 empty
    assert s.pattern()[0].tolist() == [0, 1, 1, 1]
    with pytest.raises(N.DfmError):
        s.make_ordering(N.ORDER_BLOCKCOLOR)  # needs coordinates
    s.close()
