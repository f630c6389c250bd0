"""Runs one of the reference's matrix-level Krylov loops (MatVec::Sol::Solve_*) in a process of its own and prints
the result as JSON: those loops create their temporaries with `new double[]` and pass them to
CMatDia_BlkCrs::MatVec with beta = 0, which multiplies the uninitialised values by 0 (src/matvec/
solver_mat_iter.cpp:219-231 with src/matvec/matdia_blkcrs.cpp:940), so a heap that held a NaN makes them
diverge. The caller runs it in a fresh process with MALLOC_PERTURB_=255, under which glibc hands out every
allocation filled with zero bytes, which makes the reference's result reproducible.
Usage: python ref_mat_runner.py ndim length nx ny [nz] method tol max_iter"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)


def main():
    import test_oracle_vs_reference as T

    a = sys.argv[1:]
    ndim, length = int(a[0]), int(a[1])
    nxyz = tuple(int(v) for v in a[2:4 + (ndim == 3)])
    method, tol, maxit = a[-3], float(a[-2]), int(a[-1])
    ref, cpu = T._assembled(ndim, length, nxyz)
    ok, conv, it = ref.solve_mat(method, tol, maxit)
    print(json.dumps({"ok": int(ok), "conv": conv, "it": it, "x": ref.vector(-2).tolist(),
                      "r": ref.vector(-1).tolist()}))


if __name__ == "__main__":
    main()
