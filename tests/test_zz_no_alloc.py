"""dfm_sys_true_residual_noalloc, the entry point that allocates no device memory, against dfm_sys_true_residual
through the C-ABI: the same kernels over the same grid, so the two agree bit for bit; two calls in a row, the
second on a smaller system; no device bytes are allocated by the calls (the library's allocation count is
unchanged); the result is the true residual the oracle's matrix gives to rounding."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from delfem_amd import _native as N
    from delfem_amd import workloads as W

    ctx = N.Context(0)
    yield N, W, ctx
    ctx.close()


def _solved(N, W, ctx, w, steps):
    s = W.build_system(ctx, w, N.ORDER_BLOCKCOLOR)
    W.assemble(s, w)
    s.factor()
    s.solve_pcg(1e-12, steps, N.FP64)  # a partial solve: the residual is well above rounding
    return s


def _noalloc(N, s):
    rel = C.c_double(-1.0)
    rc = N.lib().dfm_sys_true_residual_noalloc(s.h, C.byref(rel))
    assert rc == 0, N.lib().dfm_last_error()
    return rel.value


def test_noalloc_true_residual_matches(mods):
    N, W, ctx = mods
    big = _solved(N, W, ctx, W.make("c3", n=10), 15)
    small = _solved(N, W, ctx, W.make("c2", n=20), 10)
    ref_big, ref_small = big.true_residual(), small.true_residual()
    before = N.device_bytes()
    got_big = _noalloc(N, big)
    got_small = _noalloc(N, small)
    assert N.device_bytes() == before
    assert got_big == ref_big and got_small == ref_small, (got_big, ref_big, got_small, ref_small)
    assert got_big > 1e-6 and got_small > 1e-6
    # against the oracle's matrix (host): the same quantity to rounding
    from oracle import pyref as P

    w = W.make("c2", n=20)
    cpu = P.CpuSystem(w["ndim"], w["coords"], w["lnods"], w["len"])
    cpu.assemble(w["kind"], w["params"])
    cpu.set_bc(w["bc"])
    x = small.get_vector(-2)
    r = cpu.res - cpu.matvec(1.0, x, 0.0, np.zeros_like(x))
    want = np.linalg.norm(r) / np.linalg.norm(cpu.res)
    assert abs(got_small - want) <= 1e-10 * want
    # a second pair of calls sees the scratch of the first and must not depend on it
    assert _noalloc(N, big) == ref_big and _noalloc(N, small) == ref_small
    big.close()
    small.close()
