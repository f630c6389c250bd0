"""Synthetic workloads of the benchmark configurations, generated on the host from seeded counter-based
draws (SplitMix64, include/dfm_gpu.h) so that every consumer sees bit-identical inputs.

  c1  2D plane-stress elasticity, rectangle 1.0 x 0.2, 128 x 32 quads -> P1 triangles, jitter +-10 %,
      E = 1, nu = 0.3, left edge clamped, body force (0, -0.01), ILU(0)-PCG to 1e-10.
  c2  2D Poisson/diffusion on the unit square, 1024 x 1024 quads -> P1 triangles, jitter +-10 %,
      per-element conductivity lognormal(0, 1), f = 1, u = 0 on the boundary, 1e-8.
  c3  3D elasticity, unit cube 96^3 hexes -> 6 Kuhn tets each, jitter +-10 %, per-element E lognormal(0, 0.5),
      nu = 0.3, bottom face clamped, seeded traction U(-1,1)^3 * 1e-2 per top quad, 1e-8.
  c4  serving stream: requests arriving as a seeded Poisson process, each a 2D Poisson problem (unit square,
      conductivity lognormal(0, 1), f = 1, boundary u = 0) or a 2D plane-stress elasticity problem (rectangle
      5:1, E lognormal(0, 0.25) per element, nu = 0.3, left edge clamped, body force (0, -g), g ~ U[0.005, 0.02])
      with equal probability, dofs log-uniform in [2k, 500k], jitter +-10 %, 1e-8.
  c5  3D St. Venant-Kirchhoff cantilever 4 x 1 x 1, 64 x 16 x 16 hexes -> tets, E = 1, nu = 0.3, clamped at x = 0,
      end traction (0, 0, -P) on x = 4.
`scale` shrinks the grids (tests use small versions of the same problems).
"""
import numpy as np

from . import _native as N

SEEDS = {"c1": 0x5EED0001, "c2": 0x5EED0002, "c3": 0x5EED0003, "c4": 0x5EED0004, "c5": 0x5EED0005}


def plane_stress(E, nu):
    return E * nu / (1.0 - nu * nu), E / (2.0 * (1.0 + nu))


def lame3d(E, nu):
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def box_face_tris(nx, ny, nz, axis, side):
    """Triangles of one face of the box grid, split along the tangent (0,0)-(1,1) diagonal as the Kuhn tets'
    faces are; two per quad, quads in (a fastest, then b) order over the face's tangent axes a < b."""
    n = (nx, ny, nz)
    a, b = [d for d in range(3) if d != axis]
    jb, ja = np.meshgrid(np.arange(n[b]), np.arange(n[a]), indexing="ij")
    ja, jb = ja.ravel(), jb.ravel()

    def nid(da, db):
        c = [None, None, None]
        c[axis] = np.full(ja.shape, 0 if side == 0 else n[axis])
        c[a] = ja + da
        c[b] = jb + db
        return (c[2] * (ny + 1) + c[1]) * (nx + 1) + c[0]

    c00, c10, c01, c11 = nid(0, 0), nid(1, 0), nid(0, 1), nid(1, 1)
    tri = np.empty((ja.size * 2, 3), dtype=np.int32)
    tri[0::2] = np.stack([c00, c10, c11], axis=1)
    tri[1::2] = np.stack([c00, c11, c01], axis=1)
    return tri


def c1(scale=1.0):
    nx, ny = max(2, int(round(128 * scale))), max(2, int(round(32 * scale)))
    seed = SEEDS["c1"]
    coords, tri = N.rect_tri(nx, ny, 1.0, 0.2, seed, 0.1)
    lam, mu = plane_stress(1.0, 0.3)
    ix = np.arange((nx + 1) * (ny + 1)) % (nx + 1)
    bc = np.zeros(((nx + 1) * (ny + 1), 2), np.int32)
    bc[ix == 0] = 1
    return dict(name="c1", ndim=2, len=2, coords=coords, lnods=tri, kind=N.SOLID2D,
                params=np.array([lam, mu, 1.0, 0.0, -0.01]), bc=bc.ravel(), faces=None, trac=None, tol=1e-10,
                grid=(nx, ny))


def c2(scale=1.0, n=None):
    nx = n if n is not None else max(2, int(round(1024 * scale)))
    seed = SEEDS["c2"]
    coords, tri = N.rect_tri(nx, nx, 1.0, 1.0, seed, 0.1)
    nelem = tri.size // 3
    alpha = np.exp(N.rng_normal(seed, 4, nelem))
    params = np.stack([alpha, np.ones(nelem)], axis=1)
    nn = (nx + 1) * (nx + 1)
    ix, iy = np.arange(nn) % (nx + 1), np.arange(nn) // (nx + 1)
    bc = ((ix == 0) | (ix == nx) | (iy == 0) | (iy == nx)).astype(np.int32)
    return dict(name="c2", ndim=2, len=1, coords=coords, lnods=tri, kind=N.POISSON, params=params, bc=bc,
                faces=None, trac=None, tol=1e-8, grid=(nx, nx))


def c3(scale=1.0, n=None):
    nx = n if n is not None else max(2, int(round(96 * scale)))
    seed = SEEDS["c3"]
    coords, tet = N.box_tet(nx, nx, nx, 1.0, 1.0, 1.0, seed, 0.1)
    nelem = tet.size // 4
    E = np.exp(0.5 * N.rng_normal(seed, 4, nelem))
    lam, mu = lame3d(E, 0.3)
    params = np.stack([lam, mu, np.zeros(nelem), np.zeros(nelem), np.zeros(nelem), np.zeros(nelem)], axis=1)
    nn = (nx + 1) ** 3
    iz = np.arange(nn) // ((nx + 1) * (nx + 1))
    bc = np.zeros((nn, 3), np.int32)
    bc[iz == 0] = 1
    faces = box_face_tris(nx, nx, nx, 2, 1)
    nquad = faces.shape[0] // 2
    t = (2.0 * N.rng_uniform(seed, 5, 3 * nquad) - 1.0) * 1e-2
    trac = np.repeat(t.reshape(nquad, 3), 2, axis=0)
    return dict(name="c3", ndim=3, len=3, coords=coords, lnods=tet, kind=N.SOLID3D, params=params, bc=bc.ravel(),
                faces=faces.ravel(), trac=trac.ravel(), tol=1e-8, grid=(nx, nx, nx))


def c5(scale=1.0, load=0.01, kind=None):
    nx, ny = max(2, int(round(64 * scale))), max(1, int(round(16 * scale)))
    coords, tet = N.box_tet(nx, ny, ny, 4.0, 1.0, 1.0, SEEDS["c5"], 0.0)
    lam, mu = lame3d(1.0, 0.3)
    nn = (nx + 1) * (ny + 1) * (ny + 1)
    ix = np.arange(nn) % (nx + 1)
    bc = np.zeros((nn, 3), np.int32)
    bc[ix == 0] = 1
    faces = box_face_tris(nx, ny, ny, 0, 1)
    trac = np.tile([0.0, 0.0, -load], faces.shape[0])
    return dict(name="c5", ndim=3, len=3, coords=coords, lnods=tet, kind=N.STVK3D if kind is None else kind,
                params=np.array([lam, mu, 0.0, 0.0, 0.0, 0.0]), bc=bc.ravel(), faces=faces.ravel(), trac=trac,
                tol=1e-8, grid=(nx, ny, ny))


def c4_request(i, dof_range=(2000, 500000)):
    """Request i of the serving stream (deterministic in i)."""
    base = SEEDS["c4"]
    u = N.rng_uniform(base, 1000 + i, 3)
    lo, hi = np.log(dof_range[0]), np.log(dof_range[1])
    dofs = float(np.exp(lo + u[1] * (hi - lo)))
    seed = base + 7919 * (i + 1)
    if u[0] < 0.5:
        k = max(2, int(round(np.sqrt(dofs))) - 1)
        coords, tri = N.rect_tri(k, k, 1.0, 1.0, seed, 0.1)
        nelem = tri.size // 3
        params = np.stack([np.exp(N.rng_normal(seed, 4, nelem)), np.ones(nelem)], axis=1)
        nn = (k + 1) * (k + 1)
        ix, iy = np.arange(nn) % (k + 1), np.arange(nn) // (k + 1)
        bc = ((ix == 0) | (ix == k) | (iy == 0) | (iy == k)).astype(np.int32)
        return dict(kind=N.POISSON, ndim=2, len=1, coords=coords, lnods=tri, params=params, bc=bc, ndof=nn,
                    grid=(k, k))
    ny = max(2, int(round(np.sqrt(dofs / 10.0))) - 1)
    nx = 5 * ny
    coords, tri = N.rect_tri(nx, ny, 1.0, 0.2, seed, 0.1)
    nelem = tri.size // 3
    E = np.exp(0.25 * N.rng_normal(seed, 4, nelem))
    lam, mu = plane_stress(E, 0.3)
    g = 0.005 + 0.015 * u[2]
    params = np.stack([lam, mu, np.ones(nelem), np.zeros(nelem), np.full(nelem, -g)], axis=1)
    nn = (nx + 1) * (ny + 1)
    bc = np.zeros((nn, 2), np.int32)
    bc[np.arange(nn) % (nx + 1) == 0] = 1
    return dict(kind=N.SOLID2D, ndim=2, len=2, coords=coords, lnods=tri, params=params, bc=bc.ravel(), ndof=2 * nn,
                grid=(nx, ny))


def c4_arrivals(n, rate):
    """Arrival times (s) of a seeded Poisson process with the given rate (requests/s)."""
    u = N.rng_uniform(SEEDS["c4"], 1, n)
    return np.cumsum(-np.log1p(-u) / rate)


def make(name, **kw):
    return {"c1": c1, "c2": c2, "c3": c3, "c5": c5}[name](**kw)


def build_system(ctx, w, ordering=N.ORDER_BLOCKCOLOR, tile=0):
    """System with pattern, ordering and Dirichlet flags set (not yet assembled)."""
    s = N.System(ctx, w["ndim"], w["coords"], w["lnods"], w["len"])
    if ordering != N.ORDER_NATURAL:
        order, _ = s.make_ordering(ordering, tile)
        s.set_ordering(order)
    s.set_bc(w["bc"])
    return s


def assemble(s, w, u=None):
    if u is not None:
        s.set_field(u)
    return s.assemble(w["kind"], w["params"], faces=w["faces"], trac=w["trac"])
