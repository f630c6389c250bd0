"""Load-stepped StVK Newton for many instances (dfm_newton_batch): each instance follows the Newton loop of
CEqn_Solid3D_Linear::Solve (src/femeqn/eqnsys_solid.cpp:174-209); the linear solves of all instances run as
one batch.

- An instance gives bit-identical results alone and inside the batch (same Newton and PCG iteration counts).
- Against the oracle's Newton loop (CPU: assembly, Dirichlet, ILU(0), PCG, u += du in the reference's
  arithmetic): the same Newton iteration counts per load step, and the final displacement agrees within the
  inexact-Newton tolerance (linear tolerance 1e-6, Newton tolerance 1e-6 on ||r||/||r_0||).
"""
import numpy as np
import pytest

import meshgen as M
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


NX, NY = 8, 2
LAM, MU = 0.3 / (1.3 * 0.4), 1.0 / 2.6  # E = 1, nu = 0.3
PARAMS = np.array([LAM, MU, 0.0, 0.0, 0.0, 0.0])
LOADS = [0.001, 0.006, 0.012, 0.02]
NSTEPS = 3


def _problem():
    co, tet = M.box_tet(NX, NY, NY, 4.0, 1.0, 1.0, 0x5EED0005, 0.0)
    nn = co.size // 3
    ix = np.arange(nn) % (NX + 1)
    bc = np.zeros((nn, 3), np.int32)
    bc[ix == 0] = 1
    # end face x = 4: two triangles per quad, unit traction (0, 0, -1) scaled per instance
    from delfem_amd.workloads import box_face_tris

    faces = box_face_tris(NX, NY, NY, 0, 1).ravel()
    trac = np.tile([0.0, 0.0, -1.0], faces.size // 3)
    return co, tet, bc.ravel(), faces, trac


def _gpu_systems(ctx, n):
    """One system built from the mesh, the others cloned from it (same ordering and Dirichlet flags)."""
    co, tet, bc, faces, trac = _problem()
    s = N.System(ctx, 3, co, tet, 3)
    order, _ = s.make_ordering(N.ORDER_BLOCKCOLOR)
    s.set_ordering(order)
    s.set_bc(bc)
    out = [s] + [s.clone() for _ in range(n - 1)]
    for t in out:
        t.set_params(PARAMS)
        t.set_loads(faces, trac)
    return out


def _oracle_newton(load, max_newton=40, ntol=1e-6, ltol=1e-6):
    co, tet, bc, faces, trac = _problem()
    cpu = P.CpuSystem(3, co, tet, 3)
    u = np.zeros(co.size)
    counts = []
    for step in range(1, NSTEPS + 1):
        scale = load * step / NSTEPS
        ini = 0.0
        n_it = max_newton
        for it in range(max_newton):
            cpu.zero()
            cpu.assemble(N.STVK3D, PARAMS, u)
            cpu.traction(faces, trac * scale)
            norm = cpu.set_bc(bc)
            if it == 0:
                if norm < 1e-20:
                    n_it = 0
                    break
                ini = norm
            if norm < ini * ntol:
                n_it = it
                break
            f = cpu.ilu0()
            _, _, _, x, _ = cpu.solve(f, "pcg", ltol, 1000)
            u = u + x
        counts.append(n_it)
    return u, counts


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_newton_alone_equals_batched(ctx, precision):
    prec = N.MIXED if precision == "mixed" else N.FP64
    alone = []
    for ld in LOADS:
        (s,) = _gpu_systems(ctx, 1)
        nit, lit, lun = N.newton_batch(ctx, [s], N.STVK3D, [ld], NSTEPS, precision=prec)
        alone.append((s.get_field(), nit[0].copy(), int(lit[0])))
        s.close()
    systems = _gpu_systems(ctx, len(LOADS))
    nit, lit, lun = N.newton_batch(ctx, systems, N.STVK3D, LOADS, NSTEPS, precision=prec)
    for k, s in enumerate(systems):
        u = s.get_field()
        assert np.array_equal(u, alone[k][0])
        assert np.array_equal(nit[k], alone[k][1]) and int(lit[k]) == alone[k][2]
        assert lun[k] == 0
        s.close()


def test_newton_matches_oracle(ctx):
    systems = _gpu_systems(ctx, len(LOADS))
    nit, lit, lun = N.newton_batch(ctx, systems, N.STVK3D, LOADS, NSTEPS, precision=N.FP64)
    for k, (s, ld) in enumerate(zip(systems, LOADS)):
        u_ref, counts = _oracle_newton(ld)
        u = s.get_field()
        assert np.abs(nit[k] - np.array(counts)).max() <= 1, (nit[k], counts)
        assert np.linalg.norm(u - u_ref) <= 1e-4 * np.linalg.norm(u_ref), (ld, np.linalg.norm(u - u_ref))
        # the tip deflects downwards, more for larger loads
        assert u.reshape(-1, 3)[:, 2].min() < 0
        s.close()
