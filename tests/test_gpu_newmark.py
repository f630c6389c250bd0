"""Newmark-beta dynamic linear elasticity on the HIP product vs the oracle restatement, through the C-ABI.

Bit-exact: matrix M + beta dt^2 K and residual (2D with and without the predictor terms, 3D; per-element
materials; natural, block-colour and RCM numbering) and UpdateValueOfField_NewmarkBeta. Tolerance-based: a time
loop with the PCG solves (ILU(1), as the solid equation systems use it) against the oracle's, and energy
conservation of the average-acceleration rule at a larger size."""
import numpy as np
import pytest

from fem_cases import LAM2, LAM3, MU2, MU3, flags, mesh
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


NM = [
    ("solid2d", 2, 6, lambda r, n: np.stack([LAM2 * r.uniform(0.5, 2, n), np.full(n, MU2), r.uniform(0.5, 2, n),
                                             np.zeros(n), np.full(n, -0.01)], 1)),
    ("solid3d", 3, 7, lambda r, n: np.stack([np.full(n, LAM3), MU3 * r.uniform(0.5, 2, n), r.uniform(0.5, 2, n),
                                             np.full(n, 0.01), np.zeros(n), np.full(n, -0.02)], 1)),
]
ORDERS = [("natural", 0), ("blockcolor", 2), ("rcm", 3)]


def build(ctx, case, ordering, initial, small=True):
    name, ndim, kind, pfun = case
    co, ln = mesh(ndim, small)
    n = co.size // ndim
    params = pfun(np.random.default_rng(17), ln.size // (ndim + 1))
    rng = np.random.default_rng(29)
    u, v, a = (rng.uniform(-0.05, 0.05, n * ndim) for _ in range(3))
    fl = flags(co, ndim, ndim)
    glob = (0.03, 0.6, 0.3025, 1.0 if initial else 0.0)
    cpu = P.CpuSystem(ndim, co, ln, ndim)
    cpu.assemble_newmark(params, u, v, a, glob)
    cpu.set_bc(fl)
    gs = N.System(ctx, ndim, co, ln, ndim)
    order = None
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, v)
    gs.set_acceleration(a)
    rn = gs.assemble(kind, params, glob=glob)
    return cpu, gs, (u, v, a), params, glob, fl, order, rn


@pytest.mark.parametrize("case", NM, ids=[c[0] for c in NM])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_newmark_assembly_bit_exact(ctx, case, oname, ordering, initial):
    print(f"newmark {case[0]} {oname} initial={initial}", flush=True)
    cpu, gs, _, _, _, _, _, rn = build(ctx, case, ordering, initial)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs), np.abs(crs - cpu.crs).max()
    assert np.array_equal(dia, cpu.dia)
    assert np.array_equal(gs.get_vector(-1), cpu.res), np.abs(gs.get_vector(-1) - cpu.res).max()
    assert abs(rn - np.sqrt(np.sum(cpu.res ** 2))) <= 1e-14 * rn


@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_newmark_update_bit_exact(ctx, initial):
    cpu, gs, (u, v, a), _, _, _, _, _ = build(ctx, NM[1], 2, True)
    x = np.random.default_rng(5).normal(size=u.size)
    gs.set_vector(-2, x)
    gs.update_field_newmark(0.55, 0.27, 0.02, initial)
    want = P.CpuSystem.update_field_newmark(0.55, 0.27, 0.02, initial, x, u, v, a)
    for g, w in zip(gs.get_fields(), want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("case", NM, ids=[c[0] for c in NM])
def test_newmark_time_loop_vs_oracle(ctx, case):
    """Five steps of the reference's dynamic flow (assemble with is_initial, ILU(1) PCG, update with is_initial)."""
    print(f"newmark time loop {case[0]}", flush=True)
    name, ndim, kind, pfun = case
    co, ln = mesh(ndim, True)
    n = co.size // ndim
    params = pfun(np.random.default_rng(3), ln.size // (ndim + 1))
    fl = flags(co, ndim, ndim)
    dt, gamma, beta = 0.05, 0.6, 0.3025
    glob = (dt, gamma, beta, 1.0)
    gs = N.System(ctx, ndim, co, ln, ndim)
    order, _ = gs.make_ordering(2)
    gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_fill_level(1)
    gs.set_field(None, None)
    gs.set_acceleration(None)
    u, v, a = np.zeros(n * ndim), np.zeros(n * ndim), np.zeros(n * ndim)
    for step in range(5):
        cpu = P.CpuSystem(ndim, co, ln, ndim)
        cpu.assemble_newmark(params, u, v, a, glob)
        cpu.set_bc(fl)
        ok, _, it, x, _ = cpu.solve(cpu.ilu(order, 1), "pcg", 1e-12, 2000)
        gs.assemble(kind, params, glob=glob)
        gs.factor()
        gok, _, git = gs.solve_pcg(1e-12, 2000)
        assert ok and gok and abs(git - it) <= 2
        gs.update_field_newmark(gamma, beta, dt, True)
        u, v, a = P.CpuSystem.update_field_newmark(gamma, beta, dt, True, x, u, v, a)
        gu, gv, ga = gs.get_fields()
        for g, w in ((gu, u), (gv, v), (ga, a)):
            assert np.linalg.norm(g - w) <= 1e-7 * np.linalg.norm(w)
    assert np.abs(u).max() > 0


def test_newmark_energy_conservation_large(ctx):
    """Average acceleration (beta 1/4, gamma 1/2), no load, consistent initial state: the energy
    v'Mv/2 + u'Ku/2 stays constant over 10 steps (to the solver tolerance) on a 2D mesh of 20k dofs."""
    from delfem_amd import _native as N2

    co, ln = N2.rect_tri(160, 60, 2.0, 0.75, 8)
    n = co.size // 2
    x = co.reshape(-1, 2)
    fl = np.zeros((n, 2), np.int32)
    fl[x[:, 0] == x[:, 0].min()] = 1
    fl = fl.ravel()
    par = np.array([LAM2, MU2, 1.0, 0.0, 0.0])
    dt, gamma, beta = 0.01, 0.5, 0.25
    z = np.zeros(2 * n)
    # M and K from the oracle (beta = 0 gives M; rho = 0 with beta dt^2 = 1 gives K)
    cm = P.CpuSystem(2, co, ln, 2)
    cm.assemble_newmark(par, z, z, z, (dt, gamma, 0.0, 1.0))
    ck = P.CpuSystem(2, co, ln, 2)
    ck.assemble_newmark(np.array([LAM2, MU2, 0.0, 0.0, 0.0]), z, z, z, (1.0, gamma, 1.0, 1.0))
    Mv = lambda y: cm.matvec(1.0, y, 0.0, z)
    Kv = lambda y: ck.matvec(1.0, y, 0.0, z)
    u = np.where(fl == 0, 0.002 * np.sin(1.5 * x.repeat(2, axis=0)[:, 0]), 0.0)
    # consistent initial acceleration: M a = -K u on the free dofs (the mass solve with PCG on the GPU)
    gm = N2.System(ctx, 2, co, ln, 2)
    gm.set_bc(fl)
    gm.set_field(None, None)
    gm.set_acceleration(None)
    gm.assemble(6, par, glob=(dt, gamma, 0.0, 1.0))  # matrix M, residual the zero load
    gm.set_vector(-1, np.where(fl == 0, -Kv(u), 0.0))
    gm.factor()
    ok, _, _ = gm.solve_pcg(1e-13, 5000)
    assert ok
    a = gm.get_vector(-2)
    energy = lambda u, v: 0.5 * v @ Mv(v) + 0.5 * u @ Kv(u)
    e0 = energy(u, z)
    print("newmark energy 161x61 nodes", flush=True)
    gs = N2.System(ctx, 2, co, ln, 2)
    order, _ = gs.make_ordering(2)
    gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, None)
    gs.set_acceleration(a)
    for _ in range(10):
        gs.assemble(6, par, glob=(dt, gamma, beta, 1.0))
        gs.factor()
        ok, _, _ = gs.solve_pcg(1e-13, 5000)
        assert ok
        gs.update_field_newmark(gamma, beta, dt, True)
        gu, gv, _ = gs.get_fields()
        assert abs(energy(gu, gv) - e0) <= 1e-8 * e0
    assert np.abs(gv).max() > 0


def test_newmark_errors(ctx):
    co, ln = mesh(2, True)
    gs = N.System(ctx, 2, co, ln, 2)
    with pytest.raises(N.DfmError):
        gs.assemble(6, [LAM2, MU2, 1.0, 0.0, 0.0])  # no (dt, gamma, beta, is_initial)
    with pytest.raises(N.DfmError):
        gs.assemble(7, [LAM3, MU3, 1.0, 0.0, 0.0, 0.0], glob=(0.1, 0.5, 0.25, 1.0))  # 3D kind on a 2D mesh


# ---- backward-Euler dynamic 2D elasticity ------------------------------------------------------------------
@pytest.mark.parametrize("stvk", [False, True], ids=["linear", "stvk"])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_beuler_assembly_and_update_bit_exact(ctx, stvk, oname, ordering, initial):
    print(f"backward euler stvk={stvk} {oname} initial={initial}", flush=True)
    co, ln = mesh(2, True)
    n = co.size // 2
    params = NM[0][3](np.random.default_rng(17), ln.size // 3)
    rng = np.random.default_rng(31)
    u, v, v0 = (rng.uniform(-0.05, 0.05, 2 * n) for _ in range(3))
    fl = flags(co, 2, 2)
    cpu = P.CpuSystem(2, co, ln, 2)
    cpu.assemble_beuler(params, u, v, v0, 0.04, initial, stvk)
    cpu.set_bc(fl)
    gs = N.System(ctx, 2, co, ln, 2)
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, v)
    gs.set_velocity_prev(v0)
    gs.assemble(11 if stvk else 8, params, glob=(0.04, 1.0 if initial else 0.0))
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(gs.get_vector(-1), cpu.res), np.abs(gs.get_vector(-1) - cpu.res).max()
    x = rng.normal(size=2 * n)
    gs.set_vector(-2, x)
    gs.update_field_beuler(0.04, initial)
    wu, wv = P.CpuSystem.update_field_beuler(0.04, initial, x, u, v)
    gu, gv, _ = gs.get_fields()
    assert np.array_equal(gu, wu) and np.array_equal(gv, wv)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_stvk3d_newmark_assembly_bit_exact(ctx, ndim, oname, ordering, initial):
    """Dynamic StVK (the Newton step of the geometrically nonlinear, non-stationary solids), 2D and 3D."""
    print(f"stvk newmark {ndim}d {oname} initial={initial}", flush=True)
    co, ln = mesh(ndim, True)
    n = co.size // ndim
    nelem = ln.size // (ndim + 1)
    r = np.random.default_rng(19)
    params = np.stack([LAM3 * r.uniform(0.5, 2, nelem), np.full(nelem, MU3), r.uniform(0.5, 2, nelem),
                       np.zeros(nelem), np.full(nelem, 0.01), np.full(nelem, -0.05)][: 3 + ndim], 1)
    u, v, a = (r.uniform(-0.05, 0.05, ndim * n) for _ in range(3))
    fl = flags(co, ndim, ndim)
    glob = (0.03, 0.6, 0.3025, 1.0 if initial else 0.0)
    cpu = P.CpuSystem(ndim, co, ln, ndim)
    cpu.assemble_stvk3d_newmark(params, u, v, a, glob)
    cpu.set_bc(fl)
    gs = N.System(ctx, ndim, co, ln, ndim)
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, v)
    gs.set_acceleration(a)
    gs.assemble(9 if ndim == 3 else 10, params, glob=glob)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(gs.get_vector(-1), cpu.res), np.abs(gs.get_vector(-1) - cpu.res).max()
