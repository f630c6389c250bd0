"""Multi-segment systems: the oracle's restatement (cpu_ref.c, cr_ms_*) of CLinearSystem's segment operations and
CPreconditioner_ILU's multi-segment ILU(0) (CMatDiaFrac_BlkCrs diagonal segments, CMatFrac_BlkCrs couplings) pinned
against the reference build bit for bit, against the committed fixture (which needs no reference), and against what
the mathematics fixes."""
import ctypes as C

import numpy as np
import pytest

import msys_cases as MC
from oracle import pyref as P

needs_ref = pytest.mark.skipif(not P.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def _ref_from(nblk, lens, mats):
    R = P.RefMultiSeg.create(nblk, lens, {k: (m["rowptr"], m["colind"]) for k, m in mats.items()})
    R.set_values(mats)
    return R


def _compare_all(R, cpu, flags, res, seed):
    """FinalizeMarge, factor, SolvePrecond, MATVEC and both Krylov loops, reference vs oracle, bit for bit."""
    R.set_vector(-1, res)
    nr_ref = R.finalize(flags)
    nr_cpu, r_cpu = cpu.set_bc(flags, res)
    assert nr_ref == nr_cpu
    assert _same(R.vector(-1), r_cpu)
    for k, m in R.matrices().items():
        assert _same(m["val"], cpu.mats[k]["val"]), k
        if "dia" in m:
            assert _same(m["dia"], cpu.mats[k]["dia"]), k
    ok, fv, fd = R.factor()
    assert ok and cpu.ilu0()
    for k, v in fv.items():
        assert _same(v, cpu.fac[k]["val"]), k
    for i in range(cpu.nseg):
        assert _same(fd[i], cpu.fac[(i, i)]["dia"]), i
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(cpu.m)
    assert _same(R.precond(v), cpu.precond(v))
    y = rng.standard_normal(cpu.m)
    for alpha, beta in ((1.0, 0.0), (-0.7, 1.0), (2.5, 0.3)):
        assert _same(R.matvec(alpha, v, beta, y), cpu.matvec(alpha, v, beta, y))
    for meth in ("pbicgstab", "pcg"):
        a = R.solve(meth, r_cpu, 1e-9, 400)
        b = cpu.solve(meth, r_cpu, 1e-9, 400)
        assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2], (meth, a[:3], b[:3])
        assert _same(a[3], b[3]) and _same(a[4], b[4])


@needs_ref
@pytest.mark.parametrize("nx,ny,seed", [(6, 4, 3), (12, 6, 9)])
def test_hyper2d_reference_system(nx, ny, seed):
    """The reference's own Mooney-Rivlin P1b/P1 assembly (three segments: bubble, corner, lambda)."""
    co, ln = MC.meshgen.rect_tri(nx, ny, 1.0, 0.5, seed)[:2]
    xy = np.asarray(co).reshape(-1, 2)
    R = P.RefMultiSeg.hyper2d(co, ln, np.zeros(xy.shape[0], np.int32), 1.0, 0.4, 1.0, 0.03, -0.1)
    assert R.nseg == 3 and R.lens == [2, 2, 1] and R.nblk == [len(ln) // 3, xy.shape[0], xy.shape[0]]
    mats = R.matrices()
    res = R.vector(-1)
    cpu = P.CpuMultiSeg(R.nblk, R.lens, mats)
    flags = np.zeros(R.m, np.int32)
    left = np.nonzero(xy[:, 0] < 1e-12)[0]
    for n in left:
        flags[R.off[1] + 2 * n: R.off[1] + 2 * n + 2] = 1
    _compare_all(R, cpu, flags, res, seed)


SHAPES = [
    dict(lens=(2, 2, 1)),
    dict(lens=(3, 1, 2), bubble_pattern=True),
    dict(lens=(1, 3, 3)),
    dict(lens=(1, 1, 1), bubble_pattern=True),
    dict(lens=(2, 3, 1), absent=((0, 2), (2, 1))),       # a coupling without its transpose partner
    dict(lens=(3, 2, 1, 2), nseg=4, absent=((3, 0),)),
    dict(lens=(2, 1), nseg=2, bubble_pattern=True),
]


@needs_ref
@pytest.mark.parametrize("case", range(len(SHAPES)))
def test_synthetic_systems(case):
    kw = SHAPES[case]
    nblk, lens, mats = MC.synthetic(7, 5, seed=case, **kw)
    R = _ref_from(nblk, lens, mats)
    cpu = P.CpuMultiSeg(nblk, lens, mats)
    flags = MC.synthetic_flags(nblk, lens, case)
    res = MC.synthetic_rhs(nblk, lens, case)
    _compare_all(R, cpu, flags, res, case)


def test_fixture_against_oracle():
    """The committed fixture (the reference's outputs) reproduced by the oracle: no reference needed."""
    g = MC.hyper2d_golden()
    cpu = P.CpuMultiSeg(g["nblk"], g["lens"], g["mats"])
    nrm, r = cpu.set_bc(g["flags"], g["res"])
    assert nrm == g["res_norm"] and _same(r, g["fres"])
    for k, m in g["finalized"].items():
        assert _same(cpu.mats[k]["val"], m["val"]), k
        if "dia" in m:
            assert _same(cpu.mats[k]["dia"], m["dia"])
    assert cpu.ilu0()
    for k, v in g["factor"].items():
        assert _same(cpu.fac[k]["val"], v), k
    for i, d in enumerate(g["factor_dia"]):
        assert _same(cpu.fac[(i, i)]["dia"], d)
    assert _same(cpu.precond(g["pc_in"]), g["pc_out"])
    assert _same(cpu.matvec(1.5, g["mv_x"], 0.25, g["mv_y"]), g["mv_out"])
    for meth in ("pbicgstab", "pcg"):
        ok, conv, it, x, rr = cpu.solve(meth, g["fres"], 1e-9, 500)
        assert ok == bool(g[f"{meth}_ok"]) and conv == g[f"{meth}_conv"] and it == g[f"{meth}_iter"]
        assert _same(x, g[f"{meth}_x"]) and _same(rr, g[f"{meth}_r"])


def _dense(nblk, lens, mats):
    off = P.ms_offsets(nblk, lens)
    A = np.zeros((off[-1], off[-1]))
    for (i, j), m in mats.items():
        bi, bj = lens[i], lens[j]
        v = m["val"].reshape(-1, bi, bj)
        for r in range(nblk[i]):
            for q in range(m["rowptr"][r], m["rowptr"][r + 1]):
                c = m["colind"][q]
                A[off[i] + r * bi: off[i] + (r + 1) * bi, off[j] + c * bj: off[j] + (c + 1) * bj] += v[q]
            if i == j:
                A[off[i] + r * bi: off[i] + (r + 1) * bi, off[i] + r * bi: off[i] + (r + 1) * bi] += \
                    m["dia"].reshape(-1, bi, bi)[r]
    return A


def test_scalar_segments_equal_single_segment_ilu():
    """With 1 x 1 blocks everywhere the segment-by-segment ILU(0) is the row-by-row ILU(0) of the concatenated
    matrix (every entry receives its updates in ascending pivot order either way): the factors agree bit for bit
    with the single-segment DoILUDecomp restatement."""
    nblk, lens, mats = MC.synthetic(6, 4, lens=(1, 1, 1), seed=21, bubble_pattern=True)
    cpu = P.CpuMultiSeg(nblk, lens, mats)
    assert cpu.ilu0()
    off = P.ms_offsets(nblk, lens)
    A = _dense(nblk, lens, mats)
    n = A.shape[0]
    rows = [[c for c in np.nonzero(A[r])[0] if c != r] for r in range(n)]
    # the union pattern of the segments (no structural zeros dropped: values are nonzero almost surely)
    rp = np.zeros(n + 1, np.uint32)
    for r in range(n):
        rp[r + 1] = rp[r] + len(rows[r])
    ci = np.array([c for row in rows for c in row], np.uint32)
    crs = np.array([A[r, c] for r in range(n) for c in rows[r]])
    dia = np.diag(A).copy()
    di = np.array([rp[r] + sum(1 for c in rows[r] if c < r) for r in range(n)], np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert P.cpu_ref().cr_ilu0_factor(n, 1, p(rp), p(ci), p(di), p(crs), p(dia))
    F = np.zeros_like(A)
    for (i, j), m in mats.items():
        v = cpu.fac[(i, j)]["val"]
        for r in range(nblk[i]):
            for q in range(m["rowptr"][r], m["rowptr"][r + 1]):
                F[off[i] + r, off[j] + m["colind"][q]] = v[q]
            if i == j:
                F[off[i] + r, off[i] + r] = cpu.fac[(i, j)]["dia"][r]
    G = np.zeros_like(A)
    for r in range(n):
        for q in range(rp[r], rp[r + 1]):
            G[r, ci[q]] = crs[q]
        G[r, r] = dia[r]
    assert np.array_equal(F, G)


def test_complete_pattern_is_exact_lu():
    """Where every pattern is dense no fill is dropped: the multi-segment ILU(0) is a complete block LU and its
    SolvePrecond solves A z = v; the PBiCGSTAB solution then satisfies the dense system to its tolerance."""
    rng = np.random.default_rng(5)
    nblk, lens = [3, 4, 2], [2, 1, 3]
    mats = {}
    for i in range(3):
        for j in range(3):
            ncol = nblk[j]
            rows = [[c for c in range(ncol) if not (i == j and c == r)] for r in range(nblk[i])]
            rp, ci = MC._crs(rows, ncol)
            m = {"rowptr": rp, "colind": ci, "val": rng.uniform(-1, 1, ci.size * lens[i] * lens[j])}
            if i == j:
                m["dia"] = (rng.uniform(-1, 1, (nblk[i], lens[i], lens[i])) + 12 * np.eye(lens[i])).ravel()
            mats[(i, j)] = m
    cpu = P.CpuMultiSeg(nblk, lens, mats)
    assert cpu.ilu0()
    A = _dense(nblk, lens, mats)
    v = rng.standard_normal(A.shape[0])
    assert np.allclose(cpu.precond(v), np.linalg.solve(A, v), rtol=1e-12, atol=1e-13)
    assert np.allclose(cpu.matvec(1.0, v, 0.0, np.zeros_like(v)), A @ v, rtol=1e-13, atol=1e-13)
    ok, conv, it, x, r = cpu.solve("pbicgstab", v, 1e-12, 50)
    assert ok and it <= 3
    assert np.linalg.norm(A @ x - v) <= 1e-11 * np.linalg.norm(v)


def test_dropped_coupling_is_a_mistake_the_pins_catch():
    """A plausible mistake (the Dia LowUp update from earlier segments dropped) changes the factor: the pin above
    would fail. Planted in the oracle's inputs on the CPU: the lambda segment factored without its couplings."""
    g = MC.hyper2d_golden()
    mats = {k: v for k, v in g["finalized"].items() if not (k[0] == 2 and k[1] != 2)}
    cpu = P.CpuMultiSeg(g["nblk"], g["lens"], mats)
    ok = cpu.ilu0()
    assert not ok or not np.array_equal(cpu.fac[(2, 2)]["dia"], g["factor_dia"][2])
