"""CPU-side checks of the product library: it loads, exports every entry point include/dfm_gpu.h declares,
and its host-only seeded generators reproduce the independent numpy twin in tests/meshgen.py."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import meshgen as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "dfm_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(dfm_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 35, names
    for required in ["dfm_sys_create_mesh", "dfm_sys_assemble", "dfm_prec_factor", "dfm_solve_pcg",
                     "dfm_solve_pbicgstab", "dfm_sys_matvec", "dfm_sys_dot"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from delfem_amd import _native

    lib = _native.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_rect_tri_matches_numpy_twin():
    from delfem_amd import _native as N

    for (nx, ny, lx, ly, seed, jit) in [(8, 5, 1.0, 0.2, 7, 0.1), (33, 17, 2.0, 1.0, 0x5EED0001, 0.1),
                                        (4, 4, 1.0, 1.0, 3, 0.0)]:
        c, t = N.rect_tri(nx, ny, lx, ly, seed, jit)
        c2, t2 = M.rect_tri(nx, ny, lx, ly, seed, jit)
        assert np.array_equal(c, c2) and np.array_equal(t, t2)


def test_box_tet_matches_numpy_twin_and_volumes():
    from delfem_amd import _native as N

    for (n, seed, jit) in [((3, 2, 4), 11, 0.1), ((5, 5, 5), 0x5EED0003, 0.1), ((4, 2, 2), 5, 0.0)]:
        c, t = N.box_tet(*n, 1.0, 0.7, 1.3, seed, jit)
        c2, t2 = M.box_tet(*n, 1.0, 0.7, 1.3, seed, jit)
        assert np.array_equal(c, c2) and np.array_equal(t, t2)
        p = c.reshape(-1, 3)[t.reshape(-1, 4)]
        vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]) / 6
        assert np.all(vol > 0)
        assert np.isclose(vol.sum(), 1.0 * 0.7 * 1.3, rtol=1e-12)  # the tets tile the box


def test_rng_matches_numpy_twin():
    from delfem_amd import _native as N

    u = N.rng_uniform(0x5EED0002, 4, 1000)
    assert np.array_equal(u, M.u01(0x5EED0002, 4, np.arange(1000, dtype=np.uint64)))
    g = N.rng_normal(0x5EED0002, 4, 1000)
    g2 = M.normal(0x5EED0002, 4, np.arange(1000, dtype=np.uint64))
    assert np.allclose(g, g2, rtol=1e-14, atol=1e-15)  # libm vs numpy log/cos may differ in the last bit
    assert abs(g.mean()) < 0.1 and abs(g.std() - 1) < 0.1


def test_c2_mesh_counts():
    from delfem_amd import workloads as W

    w = W.c2(n=64)
    assert w["coords"].size // 2 == 65 * 65 and w["lnods"].size // 3 == 2 * 64 * 64
    assert w["bc"].sum() == 4 * 64


@pytest.mark.parametrize("name", ["c1", "c3", "c5"])
def test_workload_shapes(name):
    from delfem_amd import workloads as W

    w = W.make(name, scale=0.125)
    n = w["coords"].size // w["ndim"]
    assert w["bc"].size == n * w["len"]
    if w["faces"] is not None:
        assert w["trac"].size == w["faces"].size


def _pattern(ndim):
    co, ln = M.rect_tri(14, 6, 1.0, 0.3, 21) if ndim == 2 else M.box_tet(4, 3, 5, 1.0, 0.7, 0.9, 22)
    from oracle import pyref as P

    cpu = P.CpuSystem(ndim, co, ln, 1)
    return cpu.rowptr, cpu.colind


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind", [3, 4, 5], ids=["rcm", "rcm2", "amd"])
def test_pattern_ordering_matches_oracle(ndim, kind):
    # host-only entry point: COrdering_Blk's orderings, against the restatement (itself pinned to the reference in
    # test_oracle_vs_reference.py); no device is touched
    from delfem_amd import _native as N
    from oracle import pyref as P

    rp, ci = _pattern(ndim)
    g = N.pattern_ordering(kind, rp, ci)
    assert np.array_equal(g, P.ordering(kind, rp, ci))
    assert np.array_equal(np.sort(g), np.arange(rp.size - 1))
    deg = np.diff(rp.astype(np.int64))
    if kind == 5:
        # taken by ascending degree, lowest index first among equals
        assert np.all(np.diff(deg[g]) >= 0)
        for d in np.unique(deg):
            assert np.all(np.diff(g[deg[g] == d]) > 0)
    if kind == 3:
        # the numbers written into the slots are the reversed breadth-first positions: block 0 is taken first, so
        # its slot holds the last number
        assert g[0] == rp.size - 2


def test_pattern_ordering_disconnected_graph():
    # two separate edges 0-1 and 2-3: RCM's search from block 0 never reaches 2 and 3, where the reference asserts
    # ("isolated island", ordering_blk.cpp:81-85); the library returns an error. RCM2 seeds every block without a
    # lower neighbour (0 and 2) and the degree ordering has no such restriction.
    from delfem_amd import _native as N
    from oracle import pyref as P

    rp = np.array([0, 1, 2, 3, 4], np.uint32)
    ci = np.array([1, 0, 3, 2], np.uint32)
    with pytest.raises(N.DfmError):
        N.pattern_ordering(3, rp, ci)
    assert P.ordering(3, rp, ci) is None
    assert np.array_equal(N.pattern_ordering(4, rp, ci), [0, 2, 1, 3])
    assert np.array_equal(N.pattern_ordering(4, rp, ci), P.ordering(4, rp, ci))
    assert np.array_equal(N.pattern_ordering(5, rp, ci), np.arange(4))


GID = os.path.join(ROOT, "tests", "golden", "gid")


def _parse_gid(path):
    # independent twin of the reader: header, numbered coordinates, numbered elements (1-based)
    lines = open(path).read().splitlines()
    head = lines[0].split()
    dim, etype, nnoel = int(head[2]), head[4], int(head[6])
    i = lines.index("Coordinates") + 1
    xyz = []
    while not lines[i].startswith("end coordinates"):
        f = lines[i].split()
        assert int(f[0]) == len(xyz) + 1
        xyz.append([float(v) for v in f[1:1 + dim]] + [0.0] * (3 - dim))
        i += 1
    i = lines.index("Elements") + 1
    con = []
    while not lines[i].startswith("end elements"):
        f = lines[i].split()
        assert int(f[0]) == len(con) + 1
        con.append([int(v) - 1 for v in f[1:1 + nnoel]])
        i += 1
    return etype, np.array(xyz), np.array(con, np.int32)


@pytest.mark.parametrize("name", ["rect_tri", "rect_tri2", "hexa_tri", "cylinder_tet"])
def test_read_gid_matches_parser(name):
    from delfem_amd import _native as N

    ndim, co, ln = N.read_gid(os.path.join(GID, name + ".msh"))
    etype, xyz, con = _parse_gid(os.path.join(GID, name + ".msh"))
    assert np.array_equal(ln, con.ravel())
    if etype == "Triangle":
        # coplanar "dimension 3" points: the constant axis is dropped, the next two kept cyclically
        v = xyz.var(axis=0)
        d = int(np.argmin(v))
        assert ndim == 2 and v[d] <= 1e-4 * v.sum()
        assert np.array_equal(co.reshape(-1, 2), xyz[:, [(d + 1) % 3, (d + 2) % 3]])
        p = co.reshape(-1, 2)[con]
        area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                      - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
        assert np.all(np.abs(area) > 0)
    else:
        assert ndim == 3 and np.array_equal(co.reshape(-1, 3), xyz)
        p = co.reshape(-1, 3)[con]
        vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6
        assert np.all(np.abs(vol) > 0)


def test_read_gid_refusals():
    from delfem_amd import _native as N

    with pytest.raises(N.DfmError):
        N.read_gid(os.path.join(GID, "rect_quad.msh"))  # quadrilaterals: not P1
    with pytest.raises(N.DfmError):
        N.read_gid(os.path.join(GID, "missing.msh"))
