"""Continuous-batching server (dfm_server_*): requests of both kinds and many sizes join and leave a running
batch; each result equals the same request served alone (bit for bit), meets the tolerance on the FP64 true
residual (checked with the oracle's matrix), and admission respects the slot and memory limits."""
import numpy as np
import pytest

from oracle import pyref as P

pytestmark = pytest.mark.gpu

N = W = None


@pytest.fixture(scope="module")
def ctx():
    global N, W
    from delfem_amd import _native, workloads

    N, W = _native, workloads
    c = _native.Context(0)
    yield c
    c.close()


NREQ = 10


def _requests():
    return [W.c4_request(i, dof_range=(2000, 30000)) for i in range(NREQ)]


def _submit(srv, q):
    return srv.submit(q["ndim"], q["len"], q["kind"], q["coords"], q["lnods"], q["params"], q["bc"], tol=1e-8)


def _drain(srv, tickets, limit=100000):
    for _ in range(limit):
        if all(srv.poll(t)["state"] >= N.REQ_CONVERGED for t in tickets):
            return
        srv.progress(8)
    raise AssertionError("requests did not finish")


def test_served_equals_alone(ctx):
    reqs = _requests()
    assert {q["kind"] for q in reqs} == {N.POISSON, N.SOLID2D}
    # alone: one slot, one request at a time
    alone = []
    srv = N.Server(ctx, N.MIXED, 1 << 30, max_active=1, nworkers=1)
    for q in reqs:
        t = _submit(srv, q)
        _drain(srv, [t])
        info = srv.poll(t)
        alone.append((srv.result(t, q["ndof"]), info))
        srv.release(t)
    srv.close()
    # served: requests arrive while others are solving
    srv = N.Server(ctx, N.MIXED, 1 << 30, max_active=6, nworkers=3)
    tickets = []
    for k, q in enumerate(reqs):
        tickets.append(_submit(srv, q))
        if k % 3 == 2:
            for _ in range(3):
                srv.progress(8)
    _drain(srv, tickets)
    for q, t, (ua, ia) in zip(reqs, tickets, alone):
        info = srv.poll(t)
        assert info["state"] == ia["state"] == N.REQ_CONVERGED
        assert info["iters"] == ia["iters"] and info["fell_back"] == ia["fell_back"]
        u = srv.result(t, q["ndof"])
        assert np.array_equal(u, ua)
        # FP64 true residual with the oracle's matrix
        cpu = P.CpuSystem(q["ndim"], q["coords"], q["lnods"], q["len"])
        cpu.assemble(q["kind"], q["params"])
        cpu.set_bc(q["bc"])
        b = cpu.res
        r = b - cpu.matvec(1.0, u, 0.0, np.zeros_like(b))
        bound = 2e-8 if info["fell_back"] else 1.01e-8
        assert np.linalg.norm(r) <= bound * np.linalg.norm(b), (info, np.linalg.norm(r) / np.linalg.norm(b))
        srv.release(t)
    srv.close()


def test_admission_limits(ctx):
    reqs = _requests()[:5]
    # a budget below any request: one request in flight at a time
    srv = N.Server(ctx, N.MIXED, 1, max_active=4, nworkers=2)
    tickets = [_submit(srv, q) for q in reqs]
    for _ in range(100000):
        active, queued = srv.progress(8)
        assert active <= 1
        if all(srv.poll(t)["state"] >= N.REQ_CONVERGED for t in tickets):
            break
    for t in tickets:
        assert srv.poll(t)["state"] == N.REQ_CONVERGED
        srv.release(t)
    srv.close()
    # two slots
    srv = N.Server(ctx, N.MIXED, 1 << 30, max_active=2, nworkers=2)
    tickets = [_submit(srv, q) for q in reqs]
    for _ in range(100000):
        active, queued = srv.progress(8)
        assert active <= 2
        if all(srv.poll(t)["state"] >= N.REQ_CONVERGED for t in tickets):
            break
    srv.close()


def test_served_newmark_request(ctx):
    """A Newmark request (four kind scalars) is served like the same system solved alone in FP64: the served
    update solves the same system to the tolerance."""
    import meshgen as Mg

    co, ln = Mg.box_tet(4, 3, 3, 1.0, 0.8, 0.7, 41)
    n = co.size // 3
    x = co.reshape(-1, 3)
    bc = np.zeros((n, 3), np.int32)
    bc[x[:, 0] == x[:, 0].min()] = 1
    bc = bc.ravel()
    par = np.array([0.5, 0.4, 1.0, 0.0, 0.0, -0.05])
    glob = np.array([0.05, 0.5, 0.25, 1.0])
    print("served newmark request", flush=True)
    srv = N.Server(ctx, precision=N.FP64, budget_bytes=1 << 30, max_active=4, nworkers=1)
    t = srv.submit(3, 3, N.SOLID3D_NEWMARK, co, ln, par, bc, tol=1e-10, max_iter=5000, glob=glob)
    _drain(srv, [t])
    assert srv.poll(t)["state"] == N.REQ_CONVERGED
    got = srv.result(t, 3 * n)
    srv.release(t)
    srv.close()
    cpu = P.CpuSystem(3, co, ln, 3)
    z = np.zeros(3 * n)
    cpu.assemble_newmark(par, z, z, z, glob)
    cpu.set_bc(bc)
    ok, _, _, want, _ = cpu.solve(cpu.ilu0(), "pcg", 1e-12, 5000)
    assert ok
    assert np.linalg.norm(got - want) <= 1e-7 * np.linalg.norm(want)


def test_served_pbicgstab_requests(ctx):
    """PBiCGSTAB requests beside batched PCG ones: each is solved by a setup worker on its own stream with the
    server's precision (mixed: the FP32 factor in the FP64 loop, the true residual checked), bit for bit the result
    of the same system solved alone, its true residual with the oracle's matrix within the tolerance."""
    reqs = _requests()[:6]
    srv = N.Server(ctx, N.MIXED, 1 << 30, max_active=4, nworkers=2)
    tickets = [srv.submit(q["ndim"], q["len"], q["kind"], q["coords"], q["lnods"], q["params"], q["bc"], tol=1e-8,
                          method=k % 2) for k, q in enumerate(reqs)]
    _drain(srv, tickets)
    for k, (q, t) in enumerate(zip(reqs, tickets)):
        info = srv.poll(t)
        assert info["state"] == N.REQ_CONVERGED, info
        u = srv.result(t, q["ndof"])
        if k % 2:
            s = W.build_system(ctx, dict(q, faces=None, trac=None), N.ORDER_BLOCKCOLOR)
            s.assemble(q["kind"], q["params"])
            s.factor()
            ok, conv, it = s.solve_pbicgstab(1e-8, 20000, N.MIXED)
            assert ok and it == info["iters"] and conv == info["conv"]
            assert np.array_equal(s.get_vector(-2), u)
            assert conv == srv.true_residual(t) and conv <= 1e-8
            s.close()
        cpu = P.CpuSystem(q["ndim"], q["coords"], q["lnods"], q["len"])
        cpu.assemble(q["kind"], q["params"])
        cpu.set_bc(q["bc"])
        b = cpu.res
        r = b - cpu.matvec(1.0, u, 0.0, np.zeros_like(b))
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b)
        srv.release(t)
    srv.close()
