"""Multi-segment systems for the tests: the reference's own 2D Mooney-Rivlin P1b/P1 system (a committed fixture,
tests/golden/msys_hyper2d.npz, written by tests/golden/make_msys_hyper2d.py from the reference build) and seeded
synthetic systems with the same segment structure (element segment, two node segments) and any block lengths.

A system is (nblk, lens, mats) with mats = {(i, j): dict(rowptr, colind, val[, dia])}; vectors and Dirichlet flags
are concatenated over the segments."""
import os

import numpy as np

import meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "msys_hyper2d.npz")


def _crs(rows, ncol):
    rp = np.zeros(len(rows) + 1, np.uint32)
    for i, r in enumerate(rows):
        rp[i + 1] = rp[i] + len(r)
    ci = np.array([c for r in rows for c in sorted(r)], np.uint32)
    assert all(0 <= c < ncol for r in rows for c in r)
    return rp, ci


def mesh_graph(nx, ny, seed):
    coords, lnods = meshgen.rect_tri(nx, ny, 1.0, ny / nx, seed)[:2]
    lnods = np.asarray(lnods).reshape(-1, 3)
    n, ne = np.asarray(coords).size // 2, lnods.shape[0]
    psup = [set() for _ in range(n)]
    n2e = [set() for _ in range(n)]
    edge = {}
    for e, tri in enumerate(lnods):
        for a in tri:
            n2e[a].add(e)
            for b in tri:
                if a != b:
                    psup[a].add(b)
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            edge.setdefault((min(a, b), max(a, b)), []).append(e)
    eadj = [set() for _ in range(ne)]
    for es in edge.values():
        if len(es) == 2:
            eadj[es[0]].add(es[1])
            eadj[es[1]].add(es[0])
    return np.asarray(coords, np.float64), lnods, psup, n2e, eadj


def synthetic(nx, ny, lens=(2, 2, 1), seed=0, bubble_pattern=False, absent=(), nseg=3):
    """Segments: 0 elements, 1 nodes, 2 nodes (as the P1b/P1 system: bubble, corner, lambda), optionally a fourth
    node segment; couplings by incidence (element-node, node-node including the node itself); pairs in `absent`
    left out. Values seeded: off-diagonal uniform(-1, 1), diagonal blocks shifted by the row's absolute sum so the
    factorisation stays away from zero pivots."""
    coords, lnods, psup, n2e, eadj = mesh_graph(nx, ny, seed)
    n, ne = len(psup), len(eadj)
    kinds = ["e", "n", "n", "n"][:nseg]
    nblk = [ne if k == "e" else n for k in kinds]
    lens = list(lens)[:nseg]
    rng = np.random.default_rng(1000 + seed)
    mats = {}
    for i in range(nseg):
        for j in range(nseg):
            if (i, j) in absent:
                continue
            ki, kj = kinds[i], kinds[j]
            if i == j:
                rows = [sorted(eadj[r]) if bubble_pattern else [] for r in range(ne)] if ki == "e" else \
                    [sorted(psup[r]) for r in range(n)]
            elif ki == "e" and kj == "n":
                rows = [sorted(lnods[r].tolist()) for r in range(ne)]
            elif ki == "n" and kj == "e":
                rows = [sorted(n2e[r]) for r in range(n)]
            else:
                rows = [sorted(psup[r] | {r}) for r in range(n)]
            rp, ci = _crs(rows, nblk[j])
            bs = lens[i] * lens[j]
            m = {"rowptr": rp, "colind": ci, "val": rng.uniform(-1.0, 1.0, ci.size * bs)}
            if i == j:
                m["dia"] = rng.uniform(-0.5, 0.5, nblk[i] * bs)
            mats[(i, j)] = m
    # diagonal dominance per dof row
    off = np.cumsum([0] + [nb * ln for nb, ln in zip(nblk, lens)])
    rowsum = np.zeros(off[-1])
    for (i, j), m in mats.items():
        bi, bj = lens[i], lens[j]
        v = np.abs(m["val"]).reshape(-1, bi, bj).sum(axis=2)
        for r in range(nblk[i]):
            for q in range(m["rowptr"][r], m["rowptr"][r + 1]):
                rowsum[off[i] + r * bi: off[i] + (r + 1) * bi] += v[q]
    for i in range(nseg):
        if (i, i) not in mats:
            continue
        b = lens[i]
        d = mats[(i, i)]["dia"].reshape(-1, b, b)
        for r in range(nblk[i]):
            for a in range(b):
                d[r, a, a] += 1.0 + rowsum[off[i] + r * b + a] + np.abs(d[r, a]).sum()
        mats[(i, i)]["dia"] = d.ravel()
    return nblk, lens, mats


def synthetic_flags(nblk, lens, seed):
    """Seeded Dirichlet flags: about 10 % of the dofs of every segment but the first."""
    rng = np.random.default_rng(2000 + seed)
    out = []
    for i, (nb, ln) in enumerate(zip(nblk, lens)):
        f = (rng.uniform(size=nb * ln) < 0.1).astype(np.int32) if i > 0 else np.zeros(nb * ln, np.int32)
        out.append(f)
    return np.concatenate(out)


def synthetic_rhs(nblk, lens, seed):
    m = sum(nb * ln for nb, ln in zip(nblk, lens))
    return np.random.default_rng(3000 + seed).standard_normal(m)


def hyper2d_golden():
    """The committed fixture: inputs (mesh, fixed nodes), the reference's assembled matrices before and after
    FinalizeMarge, its factor, preconditioner and solver outputs."""
    z = np.load(GOLDEN, allow_pickle=False)
    nseg = int(z["nseg"])
    nblk, lens = [int(x) for x in z["nblk"]], [int(x) for x in z["lens"]]
    mats, fin, fac = {}, {}, {}
    for i in range(nseg):
        for j in range(nseg):
            key = f"{i}_{j}"
            if f"rowptr_{key}" not in z:
                continue
            base = {"rowptr": z[f"rowptr_{key}"], "colind": z[f"colind_{key}"]}
            mats[(i, j)] = dict(base, val=z[f"val_{key}"])
            fin[(i, j)] = dict(base, val=z[f"fval_{key}"])
            fac[(i, j)] = z[f"factor_{key}"]
            if i == j:
                mats[(i, j)]["dia"] = z[f"dia_{i}"]
                fin[(i, j)]["dia"] = z[f"fdia_{i}"]
    g = {k: z[k] for k in z.files}
    g.update(nblk=nblk, lens=lens, mats=mats, finalized=fin, factor=fac,
             factor_dia=[z[f"factor_dia_{i}"] for i in range(nseg)])
    return g
