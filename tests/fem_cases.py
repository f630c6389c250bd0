"""Seeded FEM systems shared by the GPU tests: the same mesh, materials, fields and Dirichlet flags assembled
by the oracle (CPU) and by the HIP library (through the C-ABI)."""
import numpy as np

import meshgen as M
from oracle import pyref as P

LAM2, MU2 = 0.3 / (1 - 0.09), 1.0 / 2.6
LAM3, MU3 = 0.3 / (1.3 * 0.4), 1.0 / 2.6


def mesh(ndim, small=True):
    if ndim == 2:
        return M.rect_tri(12, 7, 1.0, 0.4, 31) if small else M.rect_tri(40, 24, 1.0, 0.6, 31)
    return M.box_tet(4, 3, 3, 1.0, 0.8, 0.7, 32) if small else M.box_tet(8, 6, 6, 1.0, 0.8, 0.7, 32)


def flags(coords, ndim, length):
    x = coords.reshape(-1, ndim)[:, 0]
    f = np.zeros((x.size, length), np.int32)
    f[x == x.min()] = 1
    return f.ravel()


KINDS = [
    ("poisson2d", 2, 1, 0, lambda r, n: np.stack([r.uniform(0.5, 2, n), np.full(n, 1.0)], 1), True, False),
    ("poisson3d", 3, 1, 0, lambda r, n: np.stack([r.uniform(0.5, 2, n), np.full(n, 0.7)], 1), True, False),
    ("diffusion2d", 2, 1, 5, lambda r, n: np.stack([np.full(n, 1.1), r.uniform(0.5, 2, n), np.full(n, 0.3)], 1),
     True, True),
    ("solid2d", 2, 2, 1, lambda r, n: np.stack([LAM2 * r.uniform(0.5, 2, n), np.full(n, MU2), np.ones(n),
                                                np.zeros(n), np.full(n, -0.01)], 1), True, False),
    ("solid3d", 3, 3, 2, lambda r, n: np.stack([np.full(n, LAM3), MU3 * r.uniform(0.5, 2, n), np.ones(n),
                                                np.full(n, 0.01), np.zeros(n), np.full(n, -0.02)], 1), True, False),
    ("stvk2d", 2, 2, 3, lambda r, n: np.stack([np.full(n, LAM2), MU2 * r.uniform(0.5, 2, n), np.ones(n),
                                               np.zeros(n), np.full(n, -0.05)], 1), True, False),
    ("stvk3d", 3, 3, 4, lambda r, n: np.stack([LAM3 * r.uniform(0.5, 2, n), np.full(n, MU3), np.ones(n),
                                               np.zeros(n), np.zeros(n), np.full(n, -0.05)], 1), True, False),
]


def setup(N, ctx, case, ordering, small=True):
    name, ndim, length, kind, pfun, use_u, use_v = case
    co, ln = mesh(ndim, small)
    n = co.size // ndim
    nelem = ln.size // (ndim + 1)
    rng = np.random.default_rng(17)
    params = pfun(rng, nelem)
    u = rng.uniform(-0.05, 0.05, n * length) if use_u else None
    v = rng.uniform(-0.05, 0.05, n * length) if use_v else None
    fl = flags(co, ndim, length)
    glob = (0.05, 0.6)
    cpu = P.CpuSystem(ndim, co, ln, length)
    cpu.assemble(kind, params, u, v, glob)
    cpu.set_bc(fl)
    gs = N.System(ctx, ndim, co, ln, length)
    order = None
    if ordering != N.ORDER_NATURAL:
        order, _ = gs.make_ordering(ordering)
        gs.set_ordering(order)
    gs.set_bc(fl)
    gs.set_field(u, v)
    rn = gs.assemble(kind, params, glob=glob)
    return cpu, gs, order, rn


