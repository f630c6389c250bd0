"""Pins the oracle's restatement (oracle/cpu_ref.c) against the reference DelFEM itself, compiled from its
own sources into oracle/_ref/ (make -C oracle ref). Bit-exact comparisons throughout: same inputs, same
arithmetic order. Skipped where the reference sources are not present (the GPU host)."""
import numpy as np
import pytest

import meshgen as M
from oracle import pyref as P

pytestmark = pytest.mark.skipif(not P.ref_available(), reason="oracle/_ref not built (reference sources absent)")

LAM2, MU2 = 1.0 * 0.3 / (1 - 0.3 ** 2), 1.0 / (2 * 1.3)  # plane stress E=1, nu=0.3
LAM3, MU3 = 0.3 / (1.3 * 0.4), 1.0 / 2.6


def mesh2d(nx=8, ny=4, seed=11):
    return M.rect_tri(nx, ny, 1.0, 0.2 * ny / 4, seed)


def mesh3d(nx=3, ny=2, nz=2, seed=12):
    return M.box_tet(nx, ny, nz, 1.0, 0.7, 0.6, seed)


def clamp_flags(coords, ndim, length, axis=0):
    x = coords.reshape(-1, ndim)[:, axis]
    f = np.zeros((x.size, length), np.int32)
    f[x == x.min()] = 1
    return f.ravel()


def both(ndim, coords, lnods, length, groups=None):
    return P.RefSystem(ndim, coords, lnods, length, groups), P.CpuSystem(ndim, coords, lnods, length)


@pytest.mark.parametrize("ndim", [2, 3])
def test_pattern_bit_exact(ndim):
    co, ln = mesh2d(11, 7) if ndim == 2 else mesh3d(4, 3, 3)
    ref, cpu = both(ndim, co, ln, 1)
    rp, ci = ref.pattern()
    assert np.array_equal(rp, cpu.rowptr) and np.array_equal(ci, cpu.colind)
    # diagonal excluded, rows ascending (src/femfield/elem_ary.cpp:838, matdiafrac_blkcrs.cpp:195-207)
    for i in range(cpu.nnode):
        row = cpu.colind[cpu.rowptr[i]:cpu.rowptr[i + 1]]
        assert i not in row and np.all(np.diff(row.astype(np.int64)) > 0)


def _rand_u(n, length, seed, scale):
    return np.random.default_rng(seed).uniform(-scale, scale, n * length)


CASES = [
    # (name, ndim, len, kind, params(per element row), uses u, uses v)
    ("poisson2d", 2, 1, P.POISSON, lambda r: [r.uniform(0.5, 2.0), 1.0], True, False),
    ("poisson3d", 3, 1, P.POISSON, lambda r: [r.uniform(0.5, 2.0), 0.7], True, False),
    ("diffusion2d", 2, 1, P.DIFFUSION2D, lambda r: [1.1, r.uniform(0.5, 2.0), 0.3], True, True),
    ("solid2d", 2, 2, P.SOLID2D, lambda r: [LAM2 * r.uniform(0.5, 2), MU2, 1.0, 0.0, -0.01], True, False),
    ("solid3d", 3, 3, P.SOLID3D, lambda r: [LAM3, MU3 * r.uniform(0.5, 2), 1.0, 0.01, 0.0, -0.02], True, False),
    ("stvk2d", 2, 2, P.STVK2D, lambda r: [LAM2, MU2 * r.uniform(0.5, 2), 1.0, 0.0, -0.05], True, False),
    ("stvk3d", 3, 3, P.STVK3D, lambda r: [LAM3 * r.uniform(0.5, 2), MU3, 1.0, 0.0, 0.0, -0.05], True, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("per_element", [False, True], ids=["uniform", "per_element"])
def test_assembly_bit_exact(case, per_element):
    name, ndim, length, kind, pfun, use_u, use_v = case
    co, ln = mesh2d(4, 2) if ndim == 2 else mesh3d(2, 2, 1)
    if not per_element:
        co, ln = mesh2d(9, 5) if ndim == 2 else mesh3d(3, 3, 2)
    npoel = 3 if ndim == 2 else 4
    nelem = ln.size // npoel
    n = co.size // ndim
    rng = np.random.default_rng(5)
    if per_element:
        params = np.array([pfun(rng) for _ in range(nelem)])
        groups = np.arange(nelem, dtype=np.int32)  # one element array per element, in element order
    else:
        params = np.array(pfun(rng))
        groups = None
    ref, cpu = both(ndim, co, ln, length, groups)
    u = _rand_u(n, length, 7, 0.05) if use_u else None
    v = _rand_u(n, length, 8, 0.05) if use_v else None
    ref.set_field(u, v)
    flags = clamp_flags(co, ndim, length)
    ref.set_bc(flags)
    glob = (0.05, 0.6)
    rn = ref.assemble(kind, params if per_element else params[None, :], glob)
    cpu.assemble(kind, params, u, v, glob)
    rn2 = cpu.set_bc(flags)
    crs, dia = ref.matrix()
    assert np.array_equal(crs, cpu.crs), np.abs(crs - cpu.crs).max()
    assert np.array_equal(dia, cpu.dia), np.abs(dia - cpu.dia).max()
    assert np.array_equal(ref.vector(-1), cpu.res)
    assert rn == rn2


@pytest.mark.parametrize("length", [1, 2, 3])
def test_matvec_bit_exact(length):
    ndim = 2 if length != 3 else 3
    co, ln = mesh2d(7, 5) if ndim == 2 else mesh3d(3, 2, 2)
    ref, cpu = both(ndim, co, ln, length)
    kind = {1: P.POISSON, 2: P.SOLID2D, 3: P.SOLID3D}[length]
    par = {1: [1.0, 1.0], 2: [LAM2, MU2, 1.0, 0.0, -0.01], 3: [LAM3, MU3, 1.0, 0.0, 0.0, -0.01]}[length]
    flags = clamp_flags(co, ndim, length)
    ref.set_bc(flags)
    ref.assemble(kind, np.array(par)[None, :])
    cpu.assemble(kind, par)
    cpu.set_bc(flags)
    rng = np.random.default_rng(3)
    x = rng.normal(size=cpu.nnode * length)
    y = rng.normal(size=cpu.nnode * length)
    for alpha, beta in [(1.0, 0.0), (-0.7, 1.3)]:
        assert np.array_equal(ref.matvec(alpha, x, beta, y), cpu.matvec(alpha, x, beta, y))


def _assembled(ndim, length, nxyz, seed=21):
    co, ln = mesh2d(*nxyz) if ndim == 2 else mesh3d(*nxyz)
    ref, cpu = both(ndim, co, ln, length)
    kind = {1: P.POISSON, 2: P.SOLID2D, 3: P.SOLID3D}[length]
    par = {1: [1.3, 1.0], 2: [LAM2, MU2, 1.0, 0.0, -0.01], 3: [LAM3, MU3, 1.0, 0.0, 0.0, -0.01]}[length]
    flags = clamp_flags(co, ndim, length)
    ref.set_bc(flags)
    ref.assemble(kind, np.array(par)[None, :])
    cpu.assemble(kind, par)
    cpu.set_bc(flags)
    return ref, cpu


ILU_CASES = [(2, 1, (9, 6)), (2, 2, (9, 4)), (3, 3, (3, 3, 2)), (3, 1, (3, 3, 3))]


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
@pytest.mark.parametrize("ordered", [False, True], ids=["natural", "ordered"])
def test_ilu0_bit_exact(ndim, length, nxyz, ordered):
    ref, cpu = _assembled(ndim, length, nxyz)
    order = np.random.default_rng(4).permutation(cpu.nnode).astype(np.int32) if ordered else None
    ok, rp, ci, crs, dia = ref.ilu0(order)
    f = cpu.ilu0(order)
    assert ok and f["ok"]
    assert np.array_equal(rp, f["frowptr"]) and np.array_equal(ci, f["fcolind"])
    assert np.array_equal(crs, f["fcrs"]) and np.array_equal(dia, f["fdia"])
    r = np.random.default_rng(9).normal(size=cpu.nnode * length)
    assert np.array_equal(ref.ilu_apply(r, 0, order), cpu.ilu_apply(f, r))


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
@pytest.mark.parametrize("lev", [1, 2, 3, -1])
def test_iluk_bit_exact(ndim, length, nxyz, lev):
    """ILU(lev) (level-of-fill pattern, A scattered into it, DoILUDecomp) against CMatDiaFrac_BlkCrs(lev, A),
    and the apply against CPreconditioner_ILU with SetFillInLevel(lev)."""
    ref, cpu = _assembled(ndim, length, nxyz)
    ok, rp, ci, crs, dia = ref.ilu(lev)
    f = cpu.ilu(None, lev)
    assert ok == f["ok"]
    assert f["frowptr"][-1] > cpu.rowptr[-1]  # there is fill
    assert np.array_equal(rp, f["frowptr"]) and np.array_equal(ci, f["fcolind"])
    assert np.array_equal(crs, f["fcrs"]) and np.array_equal(dia, f["fdia"])
    r = np.random.default_rng(9).normal(size=cpu.nnode * length)
    assert np.array_equal(ref.ilu_apply(r, lev), cpu.ilu_apply(f, r))


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
def test_iluk_levels(ndim, length, nxyz):
    """Unlimited fill reaches a fixed point (ILU(n) = ILU(-1)) that contains every ILU(k) pattern; ILU(1)'s
    level-0 entries are exactly the matrix pattern and its fill has level 1; ILU(-1) is a complete LU: its apply
    solves A z = r. (ILU(k) need not lie within ILU(k+1): the first insertion of an entry fixes its level, so a
    higher limit can give an entry a level that no longer lets it pivot.)"""
    _, cpu = _assembled(ndim, length, nxyz)
    fs = [cpu.ilu(None, k) for k in range(4)] + [cpu.ilu(None, cpu.nnode), cpu.ilu(None, -1)]
    full = fs[-1]
    for a in fs[:-1]:
        for i in range(cpu.nnode):
            ra = set(a["fcolind"][a["frowptr"][i]:a["frowptr"][i + 1]].tolist())
            rb = set(full["fcolind"][full["frowptr"][i]:full["frowptr"][i + 1]].tolist())
            assert ra <= rb
    assert np.array_equal(fs[-2]["frowptr"], full["frowptr"]) and np.array_equal(fs[-2]["fcolind"], full["fcolind"])
    f1 = fs[1]
    for i in range(cpu.nnode):
        cols = f1["fcolind"][f1["frowptr"][i]:f1["frowptr"][i + 1]]
        lev = f1["flev"][f1["frowptr"][i]:f1["frowptr"][i + 1]]
        acols = cpu.colind[cpu.rowptr[i]:cpu.rowptr[i + 1]]
        assert set(cols[lev == 0].tolist()) == set(acols.tolist())
        assert np.all(lev <= 1)
    lu = fs[-1]
    r = np.random.default_rng(3).normal(size=cpu.nnode * length)
    z = cpu.ilu_apply(lu, r)
    assert np.linalg.norm(cpu.matvec(1.0, z, 0.0, np.zeros_like(r)) - r) <= 1e-9 * np.linalg.norm(r)


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES[:3])
@pytest.mark.parametrize("lev", [1, 2])
def test_iluk_ordered_is_relabelled_reference(ndim, length, nxyz, lev):
    """ILU(lev) under an ordering equals the reference's natural-order ILU(lev) of the same mesh with its nodes
    renumbered by that ordering (the reference applies orderings to ILU(0) only)."""
    co, ln = mesh2d(*nxyz) if ndim == 2 else mesh3d(*nxyz)
    n = co.size // ndim
    order = np.random.default_rng(5).permutation(n).astype(np.int32)
    inv = np.empty(n, np.int32)
    inv[order] = np.arange(n, dtype=np.int32)
    co2 = co.reshape(n, ndim)[order].ravel()
    ln2 = inv[ln]
    _, cpu = both(ndim, co, ln, length)
    ref2, _ = both(ndim, co2, ln2, length)
    kind = {1: P.POISSON, 2: P.SOLID2D, 3: P.SOLID3D}[length]
    par = {1: [1.3, 1.0], 2: [LAM2, MU2, 1.0, 0.0, -0.01], 3: [LAM3, MU3, 1.0, 0.0, 0.0, -0.01]}[length]
    cpu.assemble(kind, par)
    cpu.set_bc(clamp_flags(co, ndim, length))
    ref2.set_bc(clamp_flags(co2, ndim, length))
    ref2.assemble(kind, np.array(par)[None, :])
    f = cpu.ilu(order, lev)
    ok, rp, ci, crs, dia = ref2.ilu(lev)
    assert ok and f["ok"]
    assert np.array_equal(rp, f["frowptr"]) and np.array_equal(ci, f["fcolind"])
    assert np.array_equal(crs, f["fcrs"]) and np.array_equal(dia, f["fdia"])


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
@pytest.mark.parametrize("lev", [1, 2])
def test_iluk_pcg_bit_exact(ndim, length, nxyz, lev):
    ref, cpu = _assembled(ndim, length, nxyz)
    f = cpu.ilu(None, lev)
    ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 1000)
    rok, rconv, rit = ref.solve("pcg", 1e-10, 1000, lev, None)
    assert (ok, conv, it) == (bool(rok), rconv, rit)
    assert np.array_equal(ref.vector(-2), x)


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
@pytest.mark.parametrize("method", ["pcg", "pbicgstab"])
@pytest.mark.parametrize("ordered", [False, True], ids=["natural", "ordered"])
def test_krylov_bit_exact(ndim, length, nxyz, method, ordered):
    ref, cpu = _assembled(ndim, length, nxyz)
    order = np.random.default_rng(4).permutation(cpu.nnode).astype(np.int32) if ordered else None
    f = cpu.ilu0(order)
    ok, conv, it, x, _ = cpu.solve(f, method, 1e-10, 1000)
    rok, rconv, rit = ref.solve(method, 1e-10, 1000, 0, order)
    assert (ok, conv, it) == (bool(rok), rconv, rit)
    assert np.array_equal(ref.vector(-2), x)


def test_pcg_nonconvergence_matches():
    """max_iter too small: Solve_PCG returns false, leaves iteration at max, reports the ratio."""
    ref, cpu = _assembled(2, 1, (9, 6))
    f = cpu.ilu0()
    ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-12, 2)
    rok, rconv, rit = ref.solve("pcg", 1e-12, 2, 0, None)
    assert not ok and rok == 0 and it == rit == 2 and conv == rconv


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_unpreconditioned_krylov_bit_exact(ndim, length, nxyz, method):
    """Solve_CG / Solve_BiCGSTAB (no preconditioner): iterates, iteration count and conv_ratio."""
    ref, cpu = _assembled(ndim, length, nxyz)
    ok, conv, it, x, _ = cpu.solve(None, method, 1e-10, 2000)
    rok, rconv, rit = ref.solve(method, 1e-10, 2000)
    assert (ok, conv, it) == (bool(rok), rconv, rit)
    assert np.array_equal(ref.vector(-2), x)


@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_unpreconditioned_max_iter_quirks(method):
    """Not converged: both return true; CG leaves conv_ratio at its input and iteration at max_iter;
    BiCGSTAB (loop from 1) leaves iteration at max_iter."""
    ref, cpu = _assembled(2, 2, (9, 4))
    ok, conv, it, x, _ = cpu.solve(None, method, 1e-14, 3)
    rok, rconv, rit = ref.solve(method, 1e-14, 3)
    assert ok and rok == 1 and it == rit == 3 and conv == rconv == 1e-14
    assert np.array_equal(ref.vector(-2), x)


@pytest.mark.parametrize("kind", [P.ORDER_RCM, P.ORDER_RCM2, P.ORDER_AMD], ids=["rcm", "rcm2", "amd"])
@pytest.mark.parametrize("ndim,length", [(2, 1), (2, 2), (3, 3)])
def test_orderings_match_reference(kind, ndim, length):
    # COrdering_Blk::MakeOrdering_RCM / _RCM2 / _AMD (src/matvec/ordering_blk.cpp) on the system's matrix, read
    # through NewToOld, against the restatement on the same pattern
    co, ln = mesh2d(13, 9) if ndim == 2 else mesh3d(5, 4, 3)
    ref = P.RefSystem(ndim, co, ln, length)
    rp, ci = ref.pattern()
    r = ref.ordering(kind)
    o = P.ordering(kind, rp, ci)
    assert o is not None and np.array_equal(o, r)
    assert np.array_equal(np.sort(r), np.arange(r.size))


def _ref_mat(ndim, length, nxyz, method, tol, maxit):
    # the reference's matrix-level loops read uninitialised temporaries (see tests/ref_mat_runner.py): run them in a
    # fresh process
    import json
    import os
    import subprocess
    import sys

    runner = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_mat_runner.py")
    # glibc's MALLOC_PERTURB_=255 fills every allocation with zero bytes (perturb byte ^ 0xff), so the
    # uninitialised temporaries read as 0.0 whatever the heap held before
    env = dict(os.environ, MALLOC_PERTURB_="255")
    out = subprocess.run([sys.executable, runner, str(ndim), str(length), *map(str, nxyz), method, repr(tol),
                          str(maxit)], capture_output=True, text=True, check=True, env=env).stdout
    d = json.loads(out.strip().splitlines()[-1])
    return d["ok"], d["conv"], d["it"], np.array(d["x"]), np.array(d["r"])


@pytest.mark.parametrize("ndim,length,nxyz", ILU_CASES)
@pytest.mark.parametrize("method", ["cg", "pcg", "bicgstab", "pbicgstab"])
def test_matrix_level_krylov_bit_exact(ndim, length, nxyz, method):
    """MatVec::Sol::Solve_* (src/matvec/solver_mat_iter.cpp) on the system's matrix: iterates, iteration count and
    conv_ratio, ILU(0) in the natural order for the preconditioned ones."""
    ref, cpu = _assembled(ndim, length, nxyz)
    f = cpu.ilu0() if method in ("pcg", "pbicgstab") else None
    ok, conv, it, x, r = cpu.solve(f, "mat_" + method, 1e-10, 2000)
    rok, rconv, rit, rx, rr = _ref_mat(ndim, length, nxyz, method, 1e-10, 2000)
    assert (ok, conv, it) == (bool(rok), rconv, rit)
    assert np.array_equal(rx, x)
    assert np.array_equal(rr, r)


@pytest.mark.parametrize("method", ["cg", "pcg", "bicgstab", "pbicgstab"])
def test_matrix_level_quirks(method):
    """Not converged within 3 iterations: CG / BiCGSTAB / PBiCGSTAB return true with iteration = max; PCG returns
    false with conv_ratio and iteration untouched. A zero right-hand side gives conv_ratio 0, iteration 0."""
    ref, cpu = _assembled(2, 2, (9, 4))
    f = cpu.ilu0() if method in ("pcg", "pbicgstab") else None
    ok, conv, it, x, _ = cpu.solve(f, "mat_" + method, 1e-14, 3)
    rok, rconv, rit, rx, _ = _ref_mat(2, 2, (9, 4), method, 1e-14, 3)
    assert (ok, conv, it) == (bool(rok), rconv, rit)
    assert ok == (method != "pcg") and it == 3 and conv == 1e-14
    assert np.array_equal(rx, x)
    ok, conv, it, x, _ = cpu.solve(f, "mat_" + method, 1e-10, 100, res=np.zeros(cpu.nnode * cpu.len))
    assert ok and conv == 0.0 and it == 0 and not x.any()


# ---- saved stiffness (CLinearSystem_Save) ---------------------------------------------------------------
SAVE_CASES = [
    ("poisson2d", 2, 1, P.POISSON, lambda r: [r.uniform(0.5, 2.0), 1.0]),
    ("solid2d", 2, 2, P.SOLID2D, lambda r: [LAM2 * r.uniform(0.5, 2), MU2, 1.0, 0.0, -0.01]),
    ("solid3d", 3, 3, P.SOLID3D, lambda r: [LAM3, MU3 * r.uniform(0.5, 2), 1.0, 0.01, 0.0, -0.02]),
]


def save_flags(coords, ndim, length):
    """x = min clamped; at x = max only the first dof of each node fixed (blocks with free and fixed dofs)."""
    x = coords.reshape(-1, ndim)[:, 0]
    f = np.zeros((x.size, length), np.int32)
    f[x == x.min()] = 1
    f[x == x.max(), 0] = 1
    return f.ravel()


def save_field(coords, ndim, length, flags, seed):
    """Field values: prescribed values at fixed dofs, arbitrary values elsewhere (MakeResidual reads both)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.05, 0.05, coords.size // ndim * length)
    u[flags == 1] = rng.uniform(0.01, 0.1, int(flags.sum()))
    return u


def _saved_pair(case, per_element, seed=5):
    name, ndim, length, kind, pfun = case
    if per_element:
        co, ln = mesh2d(4, 2) if ndim == 2 else mesh3d(2, 2, 1)
    else:
        co, ln = mesh2d(9, 5) if ndim == 2 else mesh3d(3, 3, 2)
    nelem = ln.size // (ndim + 1)
    rng = np.random.default_rng(seed)
    if per_element:
        params = np.array([pfun(rng) for _ in range(nelem)])
        groups = np.arange(nelem, dtype=np.int32)
    else:
        params = np.array(pfun(rng))
        groups = None
    ref = P.RefSystem(ndim, co, ln, length, groups, save=True)
    cpu = P.CpuSystem(ndim, co, ln, length)
    flags = save_flags(co, ndim, length)
    u = save_field(co, ndim, length, flags, seed + 1)
    ref.set_field(u)
    ref.set_bc(flags)
    ref.assemble_save(kind, params if per_element else params[None, :])
    cpu.assemble_save(kind, params, flags)
    return ref, cpu, u, kind


@pytest.mark.parametrize("case", SAVE_CASES, ids=[c[0] for c in SAVE_CASES])
@pytest.mark.parametrize("per_element", [False, True], ids=["uniform", "per_element"])
def test_saved_stiffness_bit_exact(case, per_element):
    """Matrix, force, residual of MakeResidual and the update it writes, element for element."""
    ref, cpu, u, _ = _saved_pair(case, per_element)
    crs, dia = ref.matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(ref.force(), cpu.force)
    assert not ref.vector(-1).any() and not cpu.res.any()
    rn = ref.make_residual()
    norm, res, upd = cpu.make_residual(u)
    assert np.array_equal(ref.vector(-1), res), np.abs(ref.vector(-1) - res).max()
    assert np.array_equal(ref.vector(-2), upd)
    assert rn == norm
    assert np.any(res != cpu.force)  # the boundary terms contribute


@pytest.mark.parametrize("case", SAVE_CASES, ids=[c[0] for c in SAVE_CASES])
def test_saved_stiffness_solve_and_update_bit_exact(case):
    """MakeResidual -> Solve_PCG (ILU(1) for the solids, as CEqn_Solid3D_Linear / CEqnSystem_Solid2D use it;
    ILU(0) for the scalar system) -> UpdateValueOfField of CLinearSystem_Save."""
    ref, cpu, u, kind = _saved_pair(case, False)
    lev = 0 if kind == P.POISSON else 1
    ref.make_residual()
    cpu.make_residual(u)
    f = cpu.ilu(None, lev)
    ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 1000)
    rok, rconv, rit = ref.solve("pcg", 1e-10, 1000, lev, None)
    assert (ok, conv, it) == (bool(rok), rconv, rit)
    assert np.array_equal(ref.vector(-2), x)
    ref.update_field_save()
    upd2, u2 = cpu.update_field_save(x, u)
    assert np.array_equal(ref.vector(-2), upd2)
    assert np.array_equal(ref.get_field(), u2)
    flags = cpu.flags
    assert np.array_equal(u2[flags == 1], u[flags == 1])  # Dirichlet values kept


# ---- saved stiffness with a lumped mass (CLinearSystem_SaveDiaM_Newmark / _NewmarkBeta) --------------------
SAVEDIA_CASES = [
    # (name, len, kind, params, gamma, beta)
    ("diffusion2d", 1, P.DIFFUSION2D, lambda r: [r.uniform(0.5, 2.0), r.uniform(0.5, 2.0), r.uniform(-1.0, 1.0)],
     0.6, 0.0),
    ("solid2d", 2, P.SOLID2D, lambda r: [LAM2 * r.uniform(0.5, 2), MU2, r.uniform(0.5, 2), 0.01, -0.02], 0.6, 0.3025),
]


def _savedia_pair(case, per_element, seed=13):
    name, length, kind, pfun, gamma, beta = case
    co, ln = mesh2d(4, 2) if per_element else mesh2d(9, 5)
    n, nelem = co.size // 2, ln.size // 3
    rng = np.random.default_rng(seed)
    if per_element:
        params = np.array([pfun(rng) for _ in range(nelem)])
        groups = np.arange(nelem, dtype=np.int32)
    else:
        params = np.array(pfun(rng))
        groups = None
    ref = P.RefSystem(2, co, ln, length, groups, dynamic=length == 2, savedia=True)
    cpu = P.CpuSystem(2, co, ln, length)
    flags = save_flags(co, 2, length)
    u = save_field(co, 2, length, flags, seed + 1)
    v = _rand_u(n, length, seed + 2, 0.1)
    a = _rand_u(n, length, seed + 3, 0.2) if length == 2 else None
    ref.set_field(u, v)
    if a is not None:
        ref.set_acc(a)
    ref.set_bc(flags)
    glob = (0.05, gamma, beta)
    ref.assemble_savedia(kind, params if per_element else params[None, :], glob)
    cpu.assemble_savedia(kind, params, flags, glob)
    return ref, cpu, (u, v, a), kind


@pytest.mark.parametrize("case", SAVEDIA_CASES, ids=[c[0] for c in SAVEDIA_CASES])
@pytest.mark.parametrize("per_element", [False, True], ids=["uniform", "per_element"])
def test_savedia_bit_exact(case, per_element):
    """Matrix (mass + time-scaled stiffness), force, boundary matrix (SetPatternDia pattern, the stiffness alone,
    fixed rows zeroed), residual and update of MakeResidual and its norm, element for element."""
    ref, cpu, (u, v, a), kind = _savedia_pair(case, per_element)
    crs, dia = ref.matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(ref.force(), cpu.force)
    if kind == P.SOLID2D:  # both gravity components land on the first dof, as the reference adds them
        assert not cpu.force[1::2].any() and cpu.force[0::2].any()
    rp, ci, val = ref.boundary()
    assert np.array_equal(rp, cpu.brp) and np.array_equal(ci, cpu.bci)
    assert np.array_equal(val, cpu.bval)
    rn = ref.make_residual()
    norm, res, upd = cpu.make_residual_savedia(u, v, a)
    assert np.array_equal(ref.vector(-1), res), np.abs(ref.vector(-1) - res).max()
    assert np.array_equal(ref.vector(-2), upd)
    assert rn == norm
    assert not res[cpu.flags == 1].any() and res.any()


def test_savedia_pattern_places_diagonal():
    """SetPatternDia: the diagonal before the first column below the row, last in rows with none below."""
    rp = np.array([0, 2, 4, 5], np.uint32)
    ci = np.array([1, 2, 0, 2, 1], np.uint32)
    lib = P.cpu_ref()
    import ctypes as C
    brp = np.zeros(4, np.uint32)
    bci = np.zeros(8, np.uint32)
    nb = lib.cr_pattern_dia(3, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                            brp.ctypes.data_as(C.c_void_p), bci.ctypes.data_as(C.c_void_p))
    assert nb == 8
    assert brp.tolist() == [0, 3, 6, 8]
    assert bci.tolist() == [1, 2, 0, 1, 0, 2, 2, 1]


@pytest.mark.parametrize("case", SAVEDIA_CASES, ids=[c[0] for c in SAVEDIA_CASES])
def test_savedia_solve_and_update_bit_exact(case):
    """MakeResidual -> Solve_PCG (ILU(0) for the scalar system, ILU(1) for the solid, as CEqnSystem_Scalar2D and
    CEqnSystem_Solid2D use them) -> UpdateValueOfField (VELOCITY / ACCELERATION) of the lumped-mass systems,
    twice in a row (the second step from the first one's fields)."""
    ref, cpu, (u, v, a), kind = _savedia_pair(case, False)
    lev = 0 if kind == P.DIFFUSION2D else 1
    for _ in range(2):
        ref.make_residual()
        cpu.make_residual_savedia(u, v, a)
        f = cpu.ilu(None, lev)
        ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 1000)
        rok, rconv, rit = ref.solve("pcg", 1e-10, 1000, lev, None)
        assert (ok, conv, it) == (bool(rok), rconv, rit)
        assert np.array_equal(ref.vector(-2), x)
        ref.update_field_savedia()
        upd2, u2, v2, a2 = cpu.update_field_savedia(x, u, v, a)
        assert np.array_equal(ref.vector(-2), upd2)
        ru, rv, ra = ref.get_fields(acc=kind == P.SOLID2D)
        assert np.array_equal(ru, u2) and np.array_equal(rv, v2)
        if kind == P.SOLID2D:
            assert np.array_equal(ra, a2)
        fl = cpu.flags == 1
        assert np.array_equal(u2[fl], u[fl])  # fixed dofs keep their values
        u, v, a = u2, v2, (a2 if kind == P.SOLID2D else None)


# ---- Newmark-beta dynamic linear elasticity -----------------------------------------------------------------
NM_CASES = [
    ("solid2d", 2, P.SOLID2D_NEWMARK, lambda r: [LAM2 * r.uniform(0.5, 2), MU2, r.uniform(0.5, 2), 0.0, -0.01]),
    ("solid3d", 3, P.SOLID3D_NEWMARK, lambda r: [LAM3, MU3 * r.uniform(0.5, 2), r.uniform(0.5, 2), 0.01, 0.0, -0.02]),
]


@pytest.mark.parametrize("case", NM_CASES, ids=[c[0] for c in NM_CASES])
@pytest.mark.parametrize("per_element", [False, True], ids=["uniform", "per_element"])
@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_newmark_assembly_bit_exact(case, per_element, initial):
    name, ndim, kind, pfun = case
    if per_element:
        co, ln = mesh2d(4, 2) if ndim == 2 else mesh3d(2, 2, 1)
    else:
        co, ln = mesh2d(9, 5) if ndim == 2 else mesh3d(3, 3, 2)
    n, nelem = co.size // ndim, ln.size // (ndim + 1)
    rng = np.random.default_rng(9)
    if per_element:
        params = np.array([pfun(rng) for _ in range(nelem)])
        groups = np.arange(nelem, dtype=np.int32)
    else:
        params = np.array(pfun(rng))
        groups = None
    ref = P.RefSystem(ndim, co, ln, ndim, groups, dynamic=True)
    cpu = P.CpuSystem(ndim, co, ln, ndim)
    u, v, a = (_rand_u(n, ndim, s, 0.05) for s in (1, 2, 3))
    glob = (0.03, 0.6, 0.3025, 1.0 if initial else 0.0)
    ref.set_field(u, v)
    ref.set_acc(a)
    flags = clamp_flags(co, ndim, ndim)
    ref.set_bc(flags)
    rn = ref.assemble(kind, params if per_element else params[None, :], glob)
    cpu.assemble_newmark(params, u, v, a, glob)
    rn2 = cpu.set_bc(flags)
    crs, dia = ref.matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(ref.vector(-1), cpu.res), np.abs(ref.vector(-1) - cpu.res).max()
    assert rn == rn2


@pytest.mark.parametrize("case", NM_CASES, ids=[c[0] for c in NM_CASES])
def test_newmark_time_steps_bit_exact(case):
    """Three steps of CEqn_Solid3D_Linear / CEqnSystem_Solid2D's dynamic flow: assemble (is_initial), ILU(1) PCG,
    UpdateValueOfField_NewmarkBeta(is_initial) — fields after every step, element for element."""
    name, ndim, kind, pfun = case
    co, ln = mesh2d(9, 4) if ndim == 2 else mesh3d(3, 2, 2)
    n = co.size // ndim
    params = np.array(pfun(np.random.default_rng(2)))
    ref = P.RefSystem(ndim, co, ln, ndim, dynamic=True)
    cpu = P.CpuSystem(ndim, co, ln, ndim)
    flags = clamp_flags(co, ndim, ndim)
    ref.set_bc(flags)
    dt, gamma, beta = 0.05, 0.6, 0.3025
    u, v, a = np.zeros(n * ndim), np.zeros(n * ndim), np.zeros(n * ndim)
    ref.set_field(u, v)
    ref.set_acc(a)
    for step in range(3):
        ref.assemble(kind, params[None, :], (dt, gamma, beta, 1.0))
        cpu.zero()
        cpu.assemble_newmark(params, u, v, a, (dt, gamma, beta, 1.0))
        cpu.set_bc(flags)
        f = cpu.ilu(None, 1)
        ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 1000)
        rok, rconv, rit = ref.solve("pcg", 1e-10, 1000, 1, None)
        assert (ok, conv, it) == (bool(rok), rconv, rit)
        assert np.array_equal(ref.vector(-2), x)
        ref.update_field_newmark(gamma, beta, dt, True)
        u, v, a = cpu.update_field_newmark(gamma, beta, dt, True, x, u, v, a)
        ru, rv, ra = ref.get_fields()
        assert np.array_equal(ru, u) and np.array_equal(rv, v) and np.array_equal(ra, a)
    assert np.abs(u).max() > 0


def test_newmark_update_not_initial_bit_exact():
    co, ln = mesh2d(5, 3)
    n = co.size // 2
    ref = P.RefSystem(2, co, ln, 2, dynamic=True)
    u, v, a, x = (_rand_u(n, 2, s, 0.1) for s in (4, 5, 6, 7))
    ref.set_field(u, v)
    ref.set_acc(a)
    ref.set_vector(-2, x)
    ref.update_field_newmark(0.55, 0.27, 0.02, False)
    got = P.CpuSystem.update_field_newmark(0.55, 0.27, 0.02, False, x, u, v, a)
    for r, c in zip(ref.get_fields(), got):
        assert np.array_equal(r, c)


# ---- backward-Euler dynamic 2D elasticity -------------------------------------------------------------------
@pytest.mark.parametrize("stvk", [False, True], ids=["linear", "stvk"])
@pytest.mark.parametrize("per_element", [False, True], ids=["uniform", "per_element"])
@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_beuler_assembly_bit_exact(stvk, per_element, initial):
    co, ln = mesh2d(4, 2) if per_element else mesh2d(9, 5)
    n, nelem = co.size // 2, ln.size // 3
    rng = np.random.default_rng(12)
    pf = lambda r: [LAM2 * r.uniform(0.5, 2), MU2, r.uniform(0.5, 2), 0.01, -0.02]
    params = np.array([pf(rng) for _ in range(nelem)]) if per_element else np.array(pf(rng))
    groups = np.arange(nelem, dtype=np.int32) if per_element else None
    ref = P.RefSystem(2, co, ln, 2, groups)
    cpu = P.CpuSystem(2, co, ln, 2)
    u, v, v0 = (_rand_u(n, 2, s, 0.05) for s in (1, 2, 3))
    ref.set_field(u, v)
    ref.set_velo_pre(v0)
    flags = clamp_flags(co, 2, 2)
    ref.set_bc(flags)
    kind = P.STVK2D_BEULER if stvk else P.SOLID2D_BEULER
    rn = ref.assemble(kind, params if per_element else params[None, :], (0.04, 1.0 if initial else 0.0))
    cpu.assemble_beuler(params, u, v, v0, 0.04, initial, stvk)
    rn2 = cpu.set_bc(flags)
    crs, dia = ref.matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(ref.vector(-1), cpu.res)
    assert rn == rn2


@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_beuler_update_bit_exact(initial):
    co, ln = mesh2d(5, 3)
    n = co.size // 2
    ref = P.RefSystem(2, co, ln, 2)
    u, v, x = (_rand_u(n, 2, s, 0.1) for s in (4, 5, 7))
    ref.set_field(u, v)
    ref.set_vector(-2, x)
    ref.update_field_beuler(0.03, initial)
    gu, gv = P.CpuSystem.update_field_beuler(0.03, initial, x, u, v)
    ru, rv, _ = ref.get_fields(acc=False)
    assert np.array_equal(ru, gu) and np.array_equal(rv, gv)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("per_element", [False, True], ids=["uniform", "per_element"])
@pytest.mark.parametrize("initial", [True, False], ids=["initial", "not_initial"])
def test_stvk3d_newmark_assembly_bit_exact(ndim, per_element, initial):
    """AddLinSys_StVenant2D/3D_NonStatic_NewmarkBeta: matrix and residual element for element."""
    if ndim == 3:
        co, ln = mesh3d(2, 2, 1) if per_element else mesh3d(3, 3, 2)
    else:
        co, ln = mesh2d(4, 2) if per_element else mesh2d(9, 5)
    n, nelem = co.size // ndim, ln.size // (ndim + 1)
    rng = np.random.default_rng(13)
    pf = lambda r: [LAM3 * r.uniform(0.5, 2), MU3, r.uniform(0.5, 2), 0.0, 0.01, -0.05][: 3 + ndim]
    params = np.array([pf(rng) for _ in range(nelem)]) if per_element else np.array(pf(rng))
    groups = np.arange(nelem, dtype=np.int32) if per_element else None
    ref = P.RefSystem(ndim, co, ln, ndim, groups, dynamic=True)
    cpu = P.CpuSystem(ndim, co, ln, ndim)
    u, v, a = (_rand_u(n, ndim, s, 0.05) for s in (1, 2, 3))
    glob = (0.03, 0.6, 0.3025, 1.0 if initial else 0.0)
    ref.set_field(u, v)
    ref.set_acc(a)
    flags = clamp_flags(co, ndim, ndim)
    ref.set_bc(flags)
    kind = P.STVK3D_NEWMARK if ndim == 3 else P.STVK2D_NEWMARK
    rn = ref.assemble(kind, params if per_element else params[None, :], glob)
    cpu.assemble_stvk3d_newmark(params, u, v, a, glob)
    rn2 = cpu.set_bc(flags)
    crs, dia = ref.matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(ref.vector(-1), cpu.res), np.abs(ref.vector(-1) - cpu.res).max()
    assert rn == rn2


@pytest.mark.parametrize("ndim,length", [(2, 1), (2, 2), (3, 3)])
def test_update_value_of_field_is_value_plus_update(ndim, length):
    """CLinearSystem_Field::UpdateValueOfField (src/femls/linearsystem_field.cpp:716-794) adds 1.0 * update to the
    field values (CNodeAry::AddValueToNodeSegment, src/femfield/node_ary.cpp:223-260): the reference's result is
    value + update bit for bit, which is what dfm_sys_update_field is checked against."""
    co, ln = mesh2d(6, 4) if ndim == 2 else mesh3d(2, 2, 2)
    ref = P.RefSystem(ndim, co, ln, length)
    n = co.size // ndim
    rng = np.random.default_rng(5)
    u = rng.standard_normal(n * length)
    du = rng.standard_normal(n * length) * 1e-3
    ref.set_field(u)
    ref.set_vector(-2, du)
    ref.update_field()
    assert np.array_equal(ref.get_field(), u + du)
