"""Saved stiffness (CLinearSystem_Save) on the HIP product vs the oracle restatement, through the C-ABI.

Bit-exact: matrix, force, MakeResidual's residual and update under every device numbering (the boundary terms
keep the reference's order), UpdateValueOfField. Tolerance-based: the PCG solve of the saved system (iteration
count within 2 of the oracle's, solution within the solver tolerance), and the saved solve against the
incremental one at the largest size."""
import numpy as np
import pytest

from fem_cases import LAM2, LAM3, MU2, MU3, mesh
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


SAVE = [
    ("poisson2d", 2, 1, 0, lambda r, n: np.stack([r.uniform(0.5, 2, n), np.full(n, 1.0)], 1)),
    ("solid2d", 2, 2, 1, lambda r, n: np.stack([LAM2 * r.uniform(0.5, 2, n), np.full(n, MU2), np.ones(n),
                                                np.full(n, 0.02), np.full(n, -0.01)], 1)),
    ("solid3d", 3, 3, 2, lambda r, n: np.stack([np.full(n, LAM3), MU3 * r.uniform(0.5, 2, n), np.ones(n),
                                                np.full(n, 0.01), np.zeros(n), np.full(n, -0.02)], 1)),
]
ORDERS = [("natural", 0), ("blockcolor", 2), ("rcm", 3)]


def save_flags(co, ndim, length):
    x = co.reshape(-1, ndim)[:, 0]
    f = np.zeros((x.size, length), np.int32)
    f[x == x.min()] = 1
    f[x == x.max(), 0] = 1
    return f.ravel()


def save_field(n, length, fl, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.05, 0.05, n * length)
    u[fl == 1] = rng.uniform(0.01, 0.1, int(fl.sum()))
    return u


def build(ctx, case, ordering, small=True, bc_first=False):
    name, ndim, length, kind, pfun = case
    co, ln = mesh(ndim, small)
    n = co.size // ndim
    params = pfun(np.random.default_rng(17), ln.size // (ndim + 1))
    fl = save_flags(co, ndim, length)
    u = save_field(n, length, fl, 23)
    cpu = P.CpuSystem(ndim, co, ln, length)
    cpu.assemble_save(kind, params, fl)
    gs = N.System(ctx, ndim, co, ln, length)
    if bc_first:
        gs.set_bc(fl)
    order = None
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    if not bc_first:
        gs.set_bc(fl)
    gs.set_field(u)
    gs.assemble_save(kind, params)
    return cpu, gs, u, fl, order, kind


@pytest.mark.parametrize("case", SAVE, ids=[c[0] for c in SAVE])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
def test_saved_assembly_and_residual_bit_exact(ctx, case, oname, ordering):
    print(f"saved stiffness {case[0]} {oname}", flush=True)
    cpu, gs, u, fl, _, _ = build(ctx, case, ordering)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(gs.get_force(), cpu.force)
    assert not gs.get_vector(-1).any()
    norm, res, upd = cpu.make_residual(u)
    gn = gs.make_residual()
    assert np.array_equal(gs.get_vector(-1), res), np.abs(gs.get_vector(-1) - res).max()
    assert np.array_equal(gs.get_vector(-2), upd)
    assert abs(gn - norm) <= 1e-14 * norm
    # the number of kept terms: fixed columns of free rows, diagonal blocks included
    nb = 0
    b = cpu.len
    for i in range(cpu.nnode):
        cols = ([i] if fl[i * b:(i + 1) * b].any() else []) + list(cpu.colind[cpu.rowptr[i]:cpu.rowptr[i + 1]])
        nfree = int((fl[i * b:(i + 1) * b] == 0).sum())
        nb += nfree * sum(int(fl[j * b:(j + 1) * b].sum()) for j in cols)
    assert gs.boundary_terms() == nb


@pytest.mark.parametrize("case", SAVE, ids=[c[0] for c in SAVE])
def test_flags_before_ordering(ctx, case):
    """Dirichlet flags set before an ordering follow it into the new device numbering."""
    cpu, gs, u, fl, _, _ = build(ctx, case, 3, bc_first=True)
    _, res, _ = cpu.make_residual(u)
    gs.make_residual()
    assert np.array_equal(gs.get_vector(-1), res)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)


@pytest.mark.parametrize("case", SAVE, ids=[c[0] for c in SAVE])
@pytest.mark.parametrize("oname,ordering", ORDERS[:2], ids=[o[0] for o in ORDERS[:2]])
def test_saved_solve_update_and_resolve(ctx, case, oname, ordering):
    """MakeResidual -> PCG (ILU(1) for the solids, ILU(0) for the scalar system, as the reference's equation
    systems use them) -> UpdateValueOfField; then new Dirichlet values and a second solve without assembly."""
    print(f"saved solve {case[0]} {oname}", flush=True)
    cpu, gs, u, fl, order, kind = build(ctx, case, ordering, small=False)
    lev = 0 if kind == 0 else 1
    gs.set_fill_level(lev)
    gs.factor()
    f = cpu.ilu(None if ordering == 0 else order, lev)
    for step in range(2):
        if step == 1:  # new prescribed values: only the residual is recomputed
            u = save_field(cpu.nnode, cpu.len, fl, 41)
            gs.set_field(u)
        gs.make_residual()
        cpu.make_residual(u)
        ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 3000)
        gok, gconv, git = gs.solve_pcg(1e-10, 3000)
        assert ok and gok
        assert abs(git - it) <= 2, (git, it)
        xg = gs.get_vector(-2)
        assert np.linalg.norm(xg - x) <= 1e-5 * np.linalg.norm(x)
        gs.update_field_save()
        _, u_cpu = cpu.update_field_save(x, u)
        ug = gs.get_field()
        assert np.array_equal(ug[fl == 1], u[fl == 1])
        assert np.linalg.norm(ug - u_cpu) <= 1e-5 * np.linalg.norm(u_cpu)


def test_saved_equals_incremental_large(ctx):
    """At a larger size, through size-independent properties: the saved solve's field equals the incremental
    formulation's (r = f - K u, K du = r, u += du), and the residual is bit-exact to the oracle's."""
    from delfem_amd import _native as N2

    co, ln = N2.rect_tri(256, 256, 1.0, 1.0, 5)
    n = co.size // 2
    nelem = ln.size // 3
    params = np.stack([np.random.default_rng(3).uniform(0.5, 2, nelem), np.full(nelem, 1.0)], 1)
    x = co.reshape(-1, 2)
    fl = ((x[:, 0] == x[:, 0].min()) | (x[:, 1] == x[:, 1].max())).astype(np.int32)
    u = np.zeros(n)
    u[fl == 1] = np.sin(3.0 * x[fl == 1, 1]) + x[fl == 1, 0]
    print("saved vs incremental 256x256", flush=True)
    gs = N2.System(ctx, 2, co, ln, 1)
    order, _ = gs.make_ordering(2)
    gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u)
    gs.assemble_save(0, params)
    gs.make_residual()
    cpu = P.CpuSystem(2, co, ln, 1)
    cpu.assemble_save(0, params, fl)
    _, res, _ = cpu.make_residual(u)
    assert np.array_equal(gs.get_vector(-1), res)
    gs.factor()
    ok, _, _ = gs.solve_pcg(1e-12, 20000)
    assert ok
    gs.update_field_save()
    u_save = gs.get_field()
    inc = N2.System(ctx, 2, co, ln, 1)
    inc.set_ordering(order)
    inc.set_bc(fl)
    inc.set_field(u)
    inc.assemble(0, params)
    inc.factor()
    ok, _, _ = inc.solve_pcg(1e-12, 20000)
    assert ok
    inc.update_field()
    u_inc = inc.get_field()
    assert np.array_equal(u_save[fl == 1], u[fl == 1])
    assert np.abs(u_save - u_inc).max() <= 1e-8 * np.abs(u_inc).max()
    # and against the oracle's own saved solve (its ILU(0) under the same ordering, PCG to the same tolerance)
    f = cpu.ilu0(order)
    cok, _, _, xc, _ = cpu.solve(f, "pcg", 1e-12, 20000, res=res)
    assert cok
    _, u_cpu = cpu.update_field_save(xc, u)
    assert np.abs(u_save - u_cpu).max() <= 1e-8 * np.abs(u_cpu).max()


def test_saved_edge_cases(ctx):
    """No Dirichlet dof: no boundary terms, r = f. Every dof fixed: no terms either (fixed rows are zeroed)."""
    name, ndim, length, kind, pfun = SAVE[1]
    co, ln = mesh(ndim, True)
    n = co.size // ndim
    params = pfun(np.random.default_rng(17), ln.size // (ndim + 1))
    for fl in [np.zeros(n * length, np.int32), np.ones(n * length, np.int32)]:
        gs = N.System(ctx, ndim, co, ln, length)
        gs.set_bc(fl)
        gs.set_field(save_field(n, length, fl, 5))
        gs.assemble_save(kind, params)
        assert gs.boundary_terms() == 0
        gs.make_residual()
        assert np.array_equal(gs.get_vector(-1), gs.get_force())


def test_saved_errors(ctx):
    """As in the reference: no saved Poisson loop on tetrahedra, no saved StVK / diffusion loops; MakeResidual
    needs a saved assembly."""
    co, ln = mesh(3, True)
    gs = N.System(ctx, 3, co, ln, 1)
    with pytest.raises(N.DfmError):
        gs.make_residual()
    with pytest.raises(N.DfmError):
        gs.assemble_save(0, [1.0, 1.0])
    co2, ln2 = mesh(2, True)
    g2 = N.System(ctx, 2, co2, ln2, 2)
    with pytest.raises(N.DfmError):
        g2.assemble_save(3, [LAM2, MU2, 1.0, 0.0, 0.0])
    g2.set_bc(np.zeros(co2.size, np.int32))
    g2.assemble_save(1, [LAM2, MU2, 1.0, 0.0, -0.01])
    g2.assemble(1, [LAM2, MU2, 1.0, 0.0, -0.01])  # a plain assembly ends the saved state
    with pytest.raises(N.DfmError):
        g2.make_residual()


# ---- lumped-mass saved systems (CLinearSystem_SaveDiaM_Newmark / _NewmarkBeta) --------------------------------
SAVEDIA = [
    # (name, len, kind, params per element, glob)
    ("diffusion2d", 1, 5, lambda r, n: np.stack([r.uniform(0.5, 2, n), r.uniform(0.5, 2, n), np.full(n, 0.3)], 1),
     (0.05, 0.6, 0.0)),
    ("solid2d", 2, 1, lambda r, n: np.stack([LAM2 * r.uniform(0.5, 2, n), np.full(n, MU2), r.uniform(0.5, 2, n),
                                             np.full(n, 0.02), np.full(n, -0.01)], 1), (0.05, 0.6, 0.3025)),
]


def build_savedia(ctx, case, ordering, small=True):
    name, length, kind, pfun, glob = case
    co, ln = mesh(2, small)
    n = co.size // 2
    params = pfun(np.random.default_rng(19), ln.size // 3)
    fl = save_flags(co, 2, length)
    u = save_field(n, length, fl, 29)
    rng = np.random.default_rng(31)
    v = rng.uniform(-0.1, 0.1, n * length)
    a = rng.uniform(-0.2, 0.2, n * length) if length == 2 else None
    cpu = P.CpuSystem(2, co, ln, length)
    cpu.assemble_savedia(kind, params, fl, glob)
    gs = N.System(ctx, 2, co, ln, length)
    order = None
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, v)
    if a is not None:
        gs.set_acceleration(a)
    gs.assemble_savedia(kind, params, glob)
    return cpu, gs, (u, v, a), fl, order, kind


@pytest.mark.parametrize("case", SAVEDIA, ids=[c[0] for c in SAVEDIA])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
def test_savedia_assembly_and_residual_bit_exact(ctx, case, oname, ordering):
    """Matrix (mass + time-scaled stiffness), force, MakeResidual's residual (zero at fixed dofs) and update (the
    Newmark predictor) bit for bit under every device numbering; one boundary term per column of every block of
    the row and the diagonal block, per free dof."""
    print(f"lumped-mass saved system {case[0]} {oname}", flush=True)
    cpu, gs, (u, v, a), fl, _, _ = build_savedia(ctx, case, ordering)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(gs.get_force(), cpu.force)
    assert not gs.get_vector(-1).any()
    norm, res, upd = cpu.make_residual_savedia(u, v, a)
    gn = gs.make_residual()
    assert np.array_equal(gs.get_vector(-1), res), np.abs(gs.get_vector(-1) - res).max()
    assert np.array_equal(gs.get_vector(-2), upd)
    assert abs(gn - norm) <= 1e-14 * norm
    b = cpu.len
    nfree_rows = sum(int((fl[i * b:(i + 1) * b] == 0).sum()) * (int(cpu.rowptr[i + 1] - cpu.rowptr[i]) + 1)
                     for i in range(cpu.nnode))
    assert gs.boundary_terms() == nfree_rows * b


@pytest.mark.parametrize("case", SAVEDIA, ids=[c[0] for c in SAVEDIA])
@pytest.mark.parametrize("oname,ordering", ORDERS[:2], ids=[o[0] for o in ORDERS[:2]])
def test_savedia_time_steps(ctx, case, oname, ordering):
    """Three time steps of MakeResidual -> PCG (ILU(0) for the diffusion system, ILU(1) for the solid, as
    CEqnSystem_Scalar2D / CEqnSystem_Solid2D use them) -> UpdateValueOfField on the saved matrix, against the
    oracle's same steps: solution within the solver tolerance, fixed dofs keep their values (and their velocity or
    acceleration), residual at each step bit-exact to the oracle's from the GPU's own fields."""
    print(f"lumped-mass saved time steps {case[0]} {oname}", flush=True)
    cpu, gs, (u, v, a), fl, order, kind = build_savedia(ctx, case, ordering, small=False)
    lev = 0 if kind == 5 else 1
    gs.set_fill_level(lev)
    gs.factor()
    f = cpu.ilu(None if ordering == 0 else order, lev)
    uc, vc, ac = u.copy(), v.copy(), (None if a is None else a.copy())
    for step in range(3):
        ug, vg, ag = gs.get_fields()
        _, res_g, _ = cpu.make_residual_savedia(ug, vg, ag if kind == 1 else None)
        gs.make_residual()
        assert np.array_equal(gs.get_vector(-1), res_g)
        cpu.make_residual_savedia(uc, vc, ac)
        ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 3000)
        gok, gconv, git = gs.solve_pcg(1e-10, 3000)
        assert ok and gok
        assert abs(git - it) <= 2, (git, it)
        xg = gs.get_vector(-2)
        assert np.linalg.norm(xg - x) <= 1e-6 * np.linalg.norm(x)
        gs.update_field_save()
        _, uc, vc, ac2 = cpu.update_field_savedia(x, uc, vc, ac)
        ac = ac2 if kind == 1 else None
        ug, vg, ag = gs.get_fields()
        assert np.array_equal(ug[fl == 1], u[fl == 1])
        if kind == 5:
            assert np.array_equal(vg[fl == 1], v[fl == 1])
        else:
            assert np.array_equal(ag[fl == 1], a[fl == 1])
        assert np.linalg.norm(ug - uc) <= 1e-6 * np.linalg.norm(uc)
        assert np.linalg.norm(vg - vc) <= 1e-6 * np.linalg.norm(vc)


def test_savedia_rejects_other_kinds(ctx):
    co, ln = mesh(3, True)
    gs = N.System(ctx, 3, co, ln, 3)
    with pytest.raises(N.DfmError):
        gs.assemble_savedia(2, np.ones(6), (0.1, 0.5, 0.25))
    co, ln = mesh(2, True)
    gs = N.System(ctx, 2, co, ln, 1)
    with pytest.raises(N.DfmError):
        gs.assemble_savedia(0, np.ones(3), (0.1, 0.5))
