"""HIP product vs the oracle on the reference's own unstructured GiD meshes (tests/golden/gid, read by
dfm_mesh_read_gid): irregular sparsity, rows of uneven length, colouring and level sets of real meshes.

Bit-exact: pattern, assembled matrix and residual, ILU(0) factor and its application, under natural, block
multicolour and the reference's RCM ordering. Krylov: iteration count within 2 of the oracle's, true residual."""
import os

import numpy as np
import pytest

from fem_cases import KINDS
from oracle import pyref as P

pytestmark = pytest.mark.gpu

GID = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gid")
N = None


@pytest.fixture(scope="module")
def ctx():
    global N
    from delfem_amd import _native

    N = _native
    c = _native.Context(0)
    yield c
    c.close()


# (mesh, case) pairs: scalar and vector problems on each dimension's meshes
CASES = [("rect_tri", KINDS[0]), ("hexa_tri", KINDS[0]), ("rect_tri", KINDS[3]), ("hexa_tri", KINDS[5]),
         ("cylinder_tet", KINDS[1]), ("cylinder_tet", KINDS[4])]
ORDERS = [("natural", 0), ("blockcolor", 2), ("rcm", 3)]


def flags(coords, ndim, length):
    # these meshes touch their lowest x in a single node: clamp the slab within 10 % of the extent (a nonsingular
    # elasticity problem)
    x = coords.reshape(-1, ndim)[:, 0]
    f = np.zeros((x.size, length), np.int32)
    f[x <= x.min() + 0.1 * (x.max() - x.min())] = 1
    return f.ravel()


def setup(ctx, mesh, case, ordering):
    name, ndim, length, kind, pfun, use_u, use_v = case
    nd, co, ln = N.read_gid(os.path.join(GID, mesh + ".msh"))
    assert nd == ndim
    n = co.size // ndim
    nelem = ln.size // (ndim + 1)
    rng = np.random.default_rng(23)
    params = pfun(rng, nelem)
    u = rng.uniform(-0.05, 0.05, n * length) if use_u else None
    v = rng.uniform(-0.05, 0.05, n * length) if use_v else None
    fl = flags(co, ndim, length)
    cpu = P.CpuSystem(ndim, co, ln, length)
    cpu.assemble(kind, params, u, v, (0.05, 0.6))
    cpu.set_bc(fl)
    gs = N.System(ctx, ndim, co, ln, length)
    order = None
    if ordering != 0:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, v)
    gs.assemble(kind, params, glob=(0.05, 0.6))
    return cpu, gs, order


@pytest.mark.parametrize("mesh,case", CASES, ids=[f"{m}-{c[0]}" for m, c in CASES])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
def test_gid_assembly_and_ilu_bit_exact(ctx, mesh, case, oname, ordering):
    cpu, gs, order = setup(ctx, mesh, case, ordering)
    rp, ci = gs.pattern()
    assert np.array_equal(rp, cpu.rowptr) and np.array_equal(ci, cpu.colind)
    crs, dia = gs.get_matrix()
    assert np.array_equal(crs, cpu.crs) and np.array_equal(dia, cpu.dia)
    assert np.array_equal(gs.get_vector(-1), cpu.res)
    gs.factor()
    f = cpu.ilu0(order)
    g = gs.get_factor()
    for k in ["frowptr", "fcolind", "diaind"]:
        assert np.array_equal(g[k], f[k]), k
    assert np.array_equal(g["fcrs"], f["fcrs"]) and np.array_equal(g["fdia"], f["fdia"])
    r = np.random.default_rng(3).normal(size=cpu.nnode * cpu.len)
    gs.resize_tmp(1)
    gs.set_vector(0, r)
    gs.prec_apply(0)
    assert np.array_equal(gs.get_vector(0), cpu.ilu_apply(f, r))


@pytest.mark.parametrize("mesh,case", [CASES[0], CASES[2], CASES[5]], ids=["tri-poisson", "tri-solid", "tet-solid"])
@pytest.mark.parametrize("oname,ordering", ORDERS, ids=[o[0] for o in ORDERS])
def test_gid_pcg_vs_oracle(ctx, mesh, case, oname, ordering):
    cpu, gs, order = setup(ctx, mesh, case, ordering)
    f = cpu.ilu0(order)
    ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-10, 3000)
    gs.factor()
    gok, gconv, git = gs.solve_pcg(1e-10, 3000)
    assert ok and gok and gconv < 1e-10
    assert abs(git - it) <= 2, (git, it)
    xg = gs.get_vector(-2)
    b = cpu.res
    rr = b - cpu.matvec(1.0, xg, 0.0, np.zeros_like(b))
    assert np.linalg.norm(rr) <= 1e-9 * np.linalg.norm(b)
