"""Writes tests/golden/msys_hyper2d.npz from the reference build (oracle/_ref, compiled by `make -C oracle ref`):
the 2D Mooney-Rivlin P1b/P1 system that Fem::Eqn::AddLinSys_Hyper2D_Static assembles into a three-segment
CLinearSystem_Field (displacement BUBBLE, displacement CORNER, lambda CORNER), before and after FinalizeMarge with
the left edge clamped, CPreconditioner_ILU's ILU(0) factor of it, its SolvePrecond of a seeded vector, MATVEC, and
Solve_PBiCGSTAB / Solve_PCG from its residual. Run from the repository root: python tests/golden/make_msys_hyper2d.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyref as P  # noqa: E402
from tests import meshgen  # noqa: E402

NX, NY, SEED = 8, 4, 5
C1, C2, RHO, GX, GY = 1.0, 0.5, 1.0, 0.02, -0.1


def main():
    coords, lnods = meshgen.rect_tri(NX, NY, 1.0, 0.5, SEED)[:2]
    coords, lnods = np.asarray(coords, np.float64), np.asarray(lnods, np.int32)
    xy = coords.reshape(-1, 2)
    clamped = (xy[:, 0] < 1e-12).astype(np.int32)
    # assembled with no flag set: FinalizeMarge leaves the values as the element loop made them
    R = P.RefMultiSeg.hyper2d(coords, lnods, np.zeros_like(clamped), C1, C2, RHO, GX, GY)
    out = {"coords": coords, "lnods": lnods, "clamped": clamped, "params": np.array([C1, C2, RHO, GX, GY]),
           "nseg": np.int32(R.nseg), "nblk": np.array(R.nblk, np.int32), "lens": np.array(R.lens, np.int32)}
    mats = R.matrices()
    res0 = R.vector(-1)
    for (i, j), m in mats.items():
        out[f"rowptr_{i}_{j}"] = m["rowptr"]
        out[f"colind_{i}_{j}"] = m["colind"]
        out[f"val_{i}_{j}"] = m["val"]
        if i == j:
            out[f"dia_{i}"] = m["dia"]
    out["res"] = res0
    # the corner segment is the one of mesh-node count and two dofs (AddPattern_Field adds BUBBLE first)
    icorner = [i for i in range(R.nseg) if R.nblk[i] == xy.shape[0] and R.lens[i] == 2][0]
    flags = np.zeros(R.m, np.int32)
    for n in np.nonzero(clamped)[0]:
        flags[R.off[icorner] + 2 * n: R.off[icorner] + 2 * n + 2] = 1
    out["flags"] = flags
    out["res_norm"] = np.float64(R.finalize(flags))
    out["fres"] = R.vector(-1)
    for (i, j), m in R.matrices().items():
        out[f"fval_{i}_{j}"] = m["val"]
        if i == j:
            out[f"fdia_{i}"] = m["dia"]
    ok, fv, fd = R.factor()
    assert ok
    for (i, j), v in fv.items():
        out[f"factor_{i}_{j}"] = v
    for i in range(R.nseg):
        out[f"factor_dia_{i}"] = fd[i]
    v = np.random.default_rng(7).standard_normal(R.m)
    out["pc_in"], out["pc_out"] = v, R.precond(v)
    y = np.random.default_rng(8).standard_normal(R.m)
    out["mv_x"], out["mv_y"], out["mv_out"] = v, y, R.matvec(1.5, v, 0.25, y)
    fres = out["fres"]
    for meth in ("pbicgstab", "pcg"):
        ok, conv, it, x, r = R.solve(meth, fres, 1e-9, 500)
        out[f"{meth}_ok"], out[f"{meth}_conv"], out[f"{meth}_iter"] = np.int32(ok), np.float64(conv), np.int32(it)
        out[f"{meth}_x"], out[f"{meth}_r"] = x, r
    path = os.path.join(ROOT, "tests", "golden", "msys_hyper2d.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes; segments", R.nblk, R.lens,
          "pbicgstab iterations", int(out["pbicgstab_iter"]), "pcg", int(out["pcg_iter"]))


if __name__ == "__main__":
    main()
