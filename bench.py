#!/usr/bin/env python3
"""DOF-solves/s at 1e-8 relative residual, mixed precision vs FP64, with p50/p99 solve latency, on one GPU.

Default workload (the headline line): 3D linear elasticity, P1 tetrahedra from a 96^3 hex grid split 6 tets/hex
(912,673 nodes, 2,738,019 dofs, 3x3 block-CRS), +-10 % node jitter, per-element E lognormal(0, 0.5), nu = 0.3,
bottom face clamped, seeded traction on the top face; synthetic inputs generated on the host and resident in HBM
before timing. One step = one solve of the hot path: assembly -> Dirichlet rows/columns -> ILU(0) numeric
factorisation -> PCG to a 1e-8 relative residual. `value` is the mixed-precision path (FP32 matrix/factor/SpMV/ILU
inside FP64 iterative refinement, convergence declared on the FP64 true residual); the plain FP64 path is timed
beside it on the same system, and the error of the mixed solution against the FP64 one is reported.

`served` (same line): the serving stream of 2,000 requests (2D Poisson or 2D elasticity, 50/50, dofs log-uniform in
[2k, 500k]) arriving as a seeded Poisson process above capacity, continuous batching under a 64 GB budget, served
in mixed precision and in FP64, and the plain FP64 GPU path one request at a time on the first requests; per-request
error of the mixed result against the FP64 one; p50/p99 latency.

The roofline object is for the dominant kernel family of the mixed path's inner loop, the ILU(0) sweep phases as
the PCG iteration runs them (the r/x update folded into the forward phases, (r, z) into the backward ones), timed
with HIP events on the library's stream; `traffic` is the HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE
(profiles/c3_hbm_traffic.md). The CPU baseline is the reference DelFEM itself (oracle/_ref, compiled from its own
sources with its own flags) on a bounded sample of the same mesh, or the oracle's C restatement where that build
is absent (see its `kind` and `sample`).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-solves/s at 1e-8 rel. residual, mixed vs FP64; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(profile, kernel_prefixes):
    """HBM bytes per dispatch of the named rows of a committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/<profile>, last column in MB), averaged over the rows; None if the summary or a row is missing."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("| `"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                vals[cells[0].strip("`")] = float(cells[-1]) * 1e6
            except ValueError:
                continue
    got = []
    for k in kernel_prefixes:
        if k not in vals:
            return None
        got.append(vals[k])
    return sum(got) / len(got)


def _ref_params(w):
    """One uniform material for the reference's per-call parameters (its element arrays take one set each):
    E = 1, nu = 0.3 (plane stress in 2D), a body force giving a non-zero right-hand side; the work per element and
    per iteration does not depend on the values."""
    from delfem_amd import workloads as W

    if w["kind"] == 2:  # 3D elasticity
        lam, mu = W.lame3d(1.0, 0.3)
        return [lam, mu, 1.0, 0.0, 0.0, -0.01]
    if w["kind"] == 1:  # 2D elasticity
        lam, mu = W.plane_stress(1.0, 0.3)
        return [lam, mu, 1.0, 0.0, -0.01]
    return [1.0, 1.0]  # Poisson: alpha, source


class _StdoutToStderr:
    """Sends what native code writes to file descriptor 1 to descriptor 2 (the reference library prints progress
    lines), so that standard output carries only the result line."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def _cpu_reference(w, gpu_iters, budget_s):
    """The reference DelFEM (oracle/_ref: its own sources and compiler flags, driven through its public API) on
    the full mesh: CLinearSystem_Field pattern, AddLinSys_* + FinalizeMarge, then CPreconditioner_ILU (natural
    order) + Solve_PCG for a bounded number of iterations, twice, the per-iteration time from the difference; the
    solve is extrapolated to the GPU FP64 iteration count."""
    import numpy as np

    from oracle import pyref as P

    par = np.array(_ref_params(w))
    with _StdoutToStderr():
        t0 = time.perf_counter()
        ref = P.RefSystem(w["ndim"], w["coords"], w["lnods"], w["len"])
        t1 = time.perf_counter()
        ref.set_bc(w["bc"])
        ref.assemble(w["kind"], par)
        t2 = time.perf_counter()
        k1 = 2
        a = time.perf_counter()
        ref.solve("pcg", 1e-30, k1)
        t_k1 = time.perf_counter() - a
        k2 = k1 + int(max(4, min(60, budget_s / max(t_k1 / k1, 1e-6))))
        a = time.perf_counter()
        ref.solve("pcg", 1e-30, k2)
        t_k2 = time.perf_counter() - a
        t_iter = (t_k2 - t_k1) / (k2 - k1)
        t_factor = max(0.0, t_k1 - k1 * t_iter)
        total = (t1 - t0) + (t2 - t1) + t_factor + gpu_iters * t_iter
        ndof = ref.nnode * ref.len
        del ref
    return {
        "value": ndof / total, "unit": "DOF-solves/s", "cores": 1, "kind": "reference",
        "sample": (f"{w['name']} full mesh ({ndof} dofs), reference DelFEM (oracle/_ref, g++ -O2 as its Makefile), one "
                   f"thread, one uniform material (its element arrays take one parameter set each): pattern "
                   f"{t1 - t0:.2f}s, AddLinSys+FinalizeMarge {t2 - t1:.2f}s, ILU(0) natural {t_factor:.2f}s, Solve_PCG "
                   f"{t_iter * 1e3:.1f} ms/iteration from runs of {k1} and {k2} iterations; solve extrapolated to "
                   f"{gpu_iters} iterations (GPU FP64 count) -> {total:.1f}s per solve"),
        "ms_per_pcg_iter": t_iter * 1e3, "solve_s_extrapolated": total,
    }


def _cpu_port(w, gpu_iters, budget_s):
    """The oracle (plain-C restatement of the reference arithmetic, one thread) on the full mesh: setup (pattern,
    assembly, Dirichlet, ILU(0) in the reference's natural order) and a bounded number of PCG iterations; the solve
    time is extrapolated with the iteration count of the GPU FP64 solve."""
    from oracle import pyref as P

    t0 = time.perf_counter()
    cpu = P.CpuSystem(w["ndim"], w["coords"], w["lnods"], w["len"])
    cpu.assemble(w["kind"], w["params"])
    cpu.set_bc(w["bc"])
    t1 = time.perf_counter()
    f = cpu.ilu0()
    t2 = time.perf_counter()
    k1 = 5
    cpu.solve(f, "pcg", 1e-30, k1)
    t3 = time.perf_counter()
    per = (t3 - t2) / k1
    k2 = max(k1 + 5, min(200, int(budget_s / max(per, 1e-6))))
    cpu.solve(f, "pcg", 1e-30, k2)
    t4 = time.perf_counter()
    t_iter = ((t4 - t3) - (t3 - t2)) / (k2 - k1)
    t_pcg0 = (t3 - t2) - k1 * t_iter
    total = (t1 - t0) + (t2 - t1) + t_pcg0 + gpu_iters * t_iter
    ndof = cpu.nnode * cpu.len
    return {
        "value": ndof / total, "unit": "DOF-solves/s", "cores": 1, "kind": "port",
        "sample": (f"{w['name']} full mesh ({ndof} dofs), the oracle's C restatement: pattern+assembly+BC "
                   f"{t1 - t0:.2f}s, ILU(0) natural {t2 - t1:.2f}s, {k2} PCG iterations at {t_iter * 1e3:.1f} ms/iter; "
                   f"solve extrapolated to {gpu_iters} iterations (GPU FP64 count) -> {total:.1f}s per solve"),
        "ms_per_pcg_iter": t_iter * 1e3, "solve_s_extrapolated": total,
    }


def cpu_baseline(w, gpu_iters, budget_s):
    """The reference build where it is present (oracle/_ref), else the oracle's restatement."""
    from oracle import pyref as P

    note = None
    if P.ref_available():
        try:
            return _cpu_reference(w, gpu_iters, budget_s)
        except Exception as e:  # e.g. the reference library does not load on this host
            note = f"reference build unusable ({e}); "
    out = _cpu_port(w, gpu_iters, budget_s)
    if note:
        out["sample"] = note + out["sample"]
    return out


def _agg_procs():
    """Worker processes of the aggregate CPU legs: the host's CPU share (at most 16)."""
    return max(1, min(16, os.cpu_count() or 1))


def _c4_cpu_worker(job):
    """One process of the C4 CPU aggregate: the oracle's C restatement solves requests (per-element materials as
    the GPU gets them: pattern, assembly, Dirichlet, ILU(0) natural, PCG to 1e-8) until the deadline."""
    idx, deadline = job
    from delfem_amd import workloads as W
    from oracle import pyref as P

    dofs, lat = 0, []
    for i in idx:
        if time.time() > deadline:
            break
        w = W.c4_request(i)
        a = time.perf_counter()
        cpu = P.CpuSystem(w["ndim"], w["coords"], w["lnods"], w["len"])
        cpu.assemble(w["kind"], w["params"])
        cpu.set_bc(w["bc"])
        f = cpu.ilu0()
        cpu.solve(f, "pcg", 1e-8, 20000)
        lat.append(time.perf_counter() - a)
        dofs += w["ndof"]
    return dofs, lat, time.time()


def c4_cpu_aggregate(budget_s, nreq):
    """C4 on the host: _agg_procs() independent processes take the first nreq requests of the stream round-robin
    and solve them with the oracle's C restatement until budget_s has passed; DOF-solves/s over the processes."""
    import multiprocessing as mp

    P = _agg_procs()
    t0 = time.time()
    jobs = [(list(range(k, nreq, P)), t0 + budget_s) for k in range(P)]
    with mp.get_context("spawn").Pool(P) as pool:
        res = pool.map(_c4_cpu_worker, jobs)
    t1 = max(r[2] for r in res)
    dofs = sum(r[0] for r in res)
    lat = sorted(x for r in res for x in r[1])
    return {"value": dofs / (t1 - t0), "unit": "DOF-solves/s", "cores": P, "kind": "port",
            "sample": f"c4 requests 0..{nreq - 1} round-robin over {P} processes, {len(lat)} solved in {t1 - t0:.1f}s "
                      f"(budget {budget_s:.0f}s; pool start-up included), the oracle's C restatement, FP64 ILU(0)-PCG "
                      f"to 1e-8, one thread each",
            "requests_solved": len(lat), "p50_ms": pct(lat, 0.5) * 1e3 if lat else None,
            "p99_ms": pct(lat, 0.99) * 1e3 if lat else None}


def _c5_cpu_worker(k1k2):
    """One process of the C5 CPU aggregate: one StVK Newton iteration of an instance with the oracle's C
    restatement (tangent and residual, Dirichlet, ILU(0) natural) and the time per PCG iteration from runs of k1
    and k2 iterations."""
    k1, k2 = k1k2
    import numpy as np

    from delfem_amd import workloads as W
    from oracle import pyref as P

    w = W.c5()
    nelem = w["lnods"].size // 4
    a = time.perf_counter()
    cpu = P.CpuSystem(3, w["coords"], w["lnods"], 3)
    cpu.assemble(w["kind"], np.tile(w["params"], (nelem, 1)), np.zeros(w["coords"].size))
    cpu.traction(w["faces"], w["trac"])
    cpu.set_bc(w["bc"])
    f = cpu.ilu0()
    t_setup = time.perf_counter() - a
    a = time.perf_counter()
    cpu.solve(f, "pcg", 1e-30, k1)
    b = time.perf_counter()
    cpu.solve(f, "pcg", 1e-30, k2)
    c = time.perf_counter()
    return t_setup, ((c - b) - (b - a)) / (k2 - k1), w["coords"].size


def c5_cpu_aggregate(newton_per_inst, pcg_per_inst, load_steps):
    """C5 on the host: _agg_procs() processes each time one Newton iteration of an instance at once (contention
    included); an instance's solve is extrapolated with the GPU run's Newton and PCG iterations per instance."""
    import multiprocessing as mp

    P = _agg_procs()
    with mp.get_context("spawn").Pool(P) as pool:
        res = pool.map(_c5_cpu_worker, [(3, 23)] * P)
    t_setup = statistics.median(r[0] for r in res)
    t_iter = statistics.median(r[1] for r in res)
    ndof = res[0][2]
    t_inst = newton_per_inst * t_setup + pcg_per_inst * t_iter
    return {"value": P * ndof * load_steps / t_inst, "unit": "DOF-solves/s", "cores": P, "kind": "port",
            "sample": f"{P} processes at once, each one StVK Newton iteration of one instance ({ndof} dofs) with the "
                      f"oracle's C restatement: tangent+residual+ILU(0) {t_setup:.2f}s, PCG {t_iter * 1e3:.1f} ms/iteration "
                      f"(medians); an instance extrapolated to the GPU run's {newton_per_inst:.1f} Newton and "
                      f"{pcg_per_inst:.0f} PCG iterations -> {t_inst:.0f}s per instance"}


def run_c5_batch(args, N, W):
    """StVK cantilever (64x16x16 hexes x 6 tets, 56,355 dofs) as a batch of served instances with seeded load
    magnitudes U[0.001, 0.02], end traction ramped over the load steps, Newton per step to ||r|| < 1e-6 ||r_0||
    (the reference's criterion), tangent + residual re-assembled on the GPU in FP64 every Newton iteration,
    inner PCG to the reference's conv_ratio 1e-6; all instances' linear solves of a Newton iteration as one
    batch. One step of the bench = the whole load path of every instance."""
    import numpy as np

    w = W.make("c5", load=1.0)
    ctx = N.Context(0)
    t0 = time.perf_counter()
    base = W.build_system(ctx, w, N.ORDER_BLOCKCOLOR)
    if args.lev:
        base.set_fill_level(args.lev)  # clones take the fill level and its schedule
    systems = [base] + [base.clone() for _ in range(args.instances - 1)]
    scales = 0.001 + (0.02 - 0.001) * N.rng_uniform(W.SEEDS["c5"], 7, args.instances)
    log(f"[bench] c5 x{args.instances} ndof={base.ndof} setup {time.perf_counter() - t0:.1f}s")
    res = {}
    # precision per system (DFM_MIXED_AUTO): for these vector systems the FP64 loop with its true residual checked
    legs = [(N.MIXED_AUTO, "mixed"), (N.FP64, "fp64")]
    if args.c5_extra:  # a third leg in another precision, reported beside the two
        legs.append(({"fp32ilu": N.FP32_ILU, "mixed32": N.MIXED}[args.c5_extra], args.c5_extra))
    for prec, name in legs:
        for s in systems:
            s.set_field(None)
            s.set_params(w["params"])
            s.set_loads(w["faces"], w["trac"])
        ctx.synchronize()
        a = time.perf_counter()
        # one call per load step (the scale of step s is scales * s / steps, as one call over all steps forms it),
        # so that every step reports its progress
        nit = np.zeros((args.instances, args.load_steps), np.int32)
        lit = np.zeros(args.instances, np.int64)
        lun = np.zeros(args.instances, np.int64)
        for step in range(1, args.load_steps + 1):
            n1, l1, u1 = N.newton_batch(ctx, systems, N.STVK3D, scales * step / args.load_steps, 1, precision=prec,
                                        lin_tol=1e-6, lin_maxit=1000)
            nit[:, step - 1] = n1[:, 0]
            lit += l1
            lun += u1
            log(f"[bench] c5 {name} load step {step}/{args.load_steps}: newton iterations {int(n1.sum())}, "
                f"pcg iterations {int(l1.sum())}, unconverged pcg {int(u1.sum())}, "
                f"{time.perf_counter() - a:.1f}s")
        ctx.synchronize()
        dt = time.perf_counter() - a
        tip = float(np.mean([s.get_field().reshape(-1, 3)[:, 2].min() for s in systems[:4]]))
        res[name] = dict(t=dt, newton=int(nit.sum()), pcg=int(lit.sum()), unconv=int(lun.sum()), tip=tip)
        log(f"[bench] c5 {name}: {dt:.2f}s, newton iterations {nit.sum()} (max/step {nit.max()}), "
            f"pcg iterations {lit.sum()}, unconverged pcg {lun.sum()}")
    ndof = base.ndof
    work = ndof * args.instances * args.load_steps
    out = {
        "metric": "DOF-solves/s (converged Newton load steps x dofs), StVK batch",
        "value": work / res["mixed"]["t"],
        "unit": "DOF-solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": res["mixed"]["t"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly, f32+f64 PCG", "data": "synthetic",
        "config": {"workload": f"c5: 3D StVK cantilever 64x16x16 hexes x 6 tets, {ndof} dofs, "
                               f"{args.instances} instances, {args.load_steps} load steps",
                   "instances": args.instances, "load_steps": args.load_steps, "newton_tol": 1e-6,
                   "pcg_tol": 1e-6, "ilu_fill_level": args.lev,
                   "precision": "DFM_MIXED_AUTO (vector systems: the FP64 loop with its true residual checked)"},
        "fp64_value": work / res["fp64"]["t"],
        "mixed": res["mixed"], "fp64": res["fp64"],
    }
    if args.c5_extra:
        out["extra_" + args.c5_extra] = dict(res[args.c5_extra], value=work / res[args.c5_extra]["t"])
    if not args.no_cpu:
        out["cpu_aggregate"] = c5_cpu_aggregate(res["fp64"]["newton"] / args.instances,
                                                res["fp64"]["pcg"] / args.instances, args.load_steps)
        log(f"[bench] c5 cpu aggregate: {out['cpu_aggregate']['value'] / 1e6:.3f} M DOF-solves/s "
            f"({out['cpu_aggregate']['sample']})")
    for s in systems:
        s.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def c4_requests(W, n):
    """Requests 0..n-1 of the serving stream, generated on the host before any clock starts (a client holds
    them), several at a time."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=8) as ex:
        return list(ex.map(W.c4_request, range(n)))


def run_c4_plain(N, ctx, reqs, n):
    """The plain FP64 GPU path on requests 0..n-1 of the stream, one at a time as a single-system user runs it:
    pattern, block-multicolour ordering, Dirichlet flags, assembly, ILU(0) and FP64 PCG (the reference's recursive
    residual test) to 1e-8; DOF-solves/s counts the requests whose true residual met 1e-8."""
    import numpy as np

    if n <= 0:
        return None
    t_all = t_solve = 0.0
    dofs = 0
    iters, trues = [], []
    for i in range(n):
        r = reqs[i]
        ctx.synchronize()
        t0 = time.perf_counter()
        s = N.System(ctx, r["ndim"], r["coords"], r["lnods"], r["len"])
        order, _ = s.make_ordering(N.ORDER_BLOCKCOLOR)
        s.set_ordering(order)
        s.set_bc(r["bc"])
        s.assemble(r["kind"], r["params"])
        s.factor()
        ctx.synchronize()
        t1 = time.perf_counter()
        ok, conv, it = s.solve_pcg(1e-8, 20000, N.FP64)
        ctx.synchronize()
        t2 = time.perf_counter()
        rel = s.true_residual()  # outside the timed calls
        s.close()
        t_all += t2 - t0
        t_solve += t2 - t1
        trues.append(rel)
        if ok and rel <= 1e-8 * (1.0 + 1e-9):
            dofs += r["ndof"]
        iters.append(it)
    return {"requests": n, "dofs_per_s": dofs / t_all, "solve_dofs_per_s": dofs / t_solve, "wall_s": t_all,
            "solve_s": t_solve, "solved_dofs": float(dofs), "met_tolerance": int(np.sum(np.array(trues) <= 1e-8 * (1 + 1e-9))),
            "max_true_rel_residual": float(np.max(trues)), "mean_pcg_iters": float(np.mean(iters))}


def serve_stream(N, ctx, reqs, arrivals, slots, precision, ref=None, keep=False, step_iters=16):
    """The requests submitted at their arrival times (s from the start) to a continuous-batching server of the
    given precision (64 GB admission budget); every result is fetched when it finishes. Aggregate DOF-solves/s
    over the time from the first arrival to the last completion, latency from arrival to completion. ref: the
    FP64 solutions of the same requests, for the relative error of each result; keep: return the solutions."""
    import numpy as np

    n = len(reqs)
    nworkers = max(1, min(16, (os.cpu_count() or 8) - 2))
    srv = N.Server(ctx, precision, 64 << 30, max_active=slots, nworkers=nworkers)
    tol = 1e-8
    lat, iters, dofs, fb, trues, met, errs = [], [], [], [], [], [], []
    sols = {} if keep or ref is not None else None
    inflight = {}
    nxt = 0
    t0 = time.perf_counter()
    last_done = t0
    max_active = max_queue = 0
    tc = {"submit": 0.0, "progress": 0.0, "poll": 0.0, "fetch": 0.0}
    while nxt < n or inflight:
        now = time.perf_counter() - t0
        c0 = time.perf_counter()
        while nxt < n and arrivals[nxt] <= now:
            r = reqs[nxt]
            t = srv.submit(r["ndim"], r["len"], r["kind"], r["coords"], r["lnods"], r["params"], r["bc"], tol=tol)
            inflight[t] = (nxt, time.perf_counter() - t0)
            nxt += 1
        c1 = time.perf_counter()
        active, queued = srv.progress(step_iters, 1)
        c2 = time.perf_counter()
        tc["submit"] += c1 - c0
        tc["progress"] += c2 - c1
        max_active, max_queue = max(max_active, active), max(max_queue, queued)
        for t in list(inflight):
            c3 = time.perf_counter()
            info = srv.poll(t)
            tc["poll"] += time.perf_counter() - c3
            if info["state"] < N.REQ_CONVERGED:
                continue
            i, ts = inflight.pop(t)
            if info["state"] not in (N.REQ_CONVERGED, N.REQ_MAXITER):
                raise RuntimeError(f"request {i} failed: {info}")
            c3 = time.perf_counter()
            x = srv.result(t, reqs[i]["ndof"])
            rel = srv.true_residual(t)
            srv.release(t)
            tc["fetch"] += time.perf_counter() - c3
            last_done = time.perf_counter()
            lat.append(ts - arrivals[i] + info["latency_s"])
            iters.append(info["iters"])
            dofs.append(reqs[i]["ndof"])
            fb.append(info["fell_back"])
            trues.append(rel)
            met.append(rel <= tol * (1.0 + 1e-9))
            if keep or ref is not None:
                sols[i] = x
        if nxt < n and not inflight:
            time.sleep(max(0.0, min(0.001, arrivals[nxt] - (time.perf_counter() - t0))))
    wall = last_done - t0 - arrivals[0]
    stats = srv.stats()
    srv.close()
    if ref is not None:  # after the clock: the relative error of every result against its FP64 solution
        for i, x in sols.items():
            errs.append(float(np.linalg.norm(x - ref[i]) / max(np.linalg.norm(ref[i]), 1e-300)))
        if not keep:
            sols = None
    total = float(np.sum(dofs))
    solved = float(np.sum(np.array(dofs)[np.array(met)]))
    # DOF-solves/s counts the requests whose true residual met the tolerance; the others' time still counts
    out = {"dofs_per_s": solved / wall, "dofs_per_s_all_results": total / wall, "wall_s": wall, "total_dofs": total,
           "met_tolerance": int(np.sum(met)), "results": len(met),
           "max_true_rel_residual": float(np.max(trues)), "median_true_rel_residual": float(np.median(trues)),
           "p50_ms": float(np.percentile(lat, 50)) * 1e3, "p99_ms": float(np.percentile(lat, 99)) * 1e3,
           "mean_pcg_iters": float(np.mean(iters)),
           "fell_back_to_fp64": int(np.sum(fb)), "max_active": max_active, "max_queued": max_queue,
           "setup_workers": nworkers, "server_host_time": stats, "client_host_time": tc}
    if errs:
        out["error_vs_fp64"] = {"max": float(np.max(errs)), "p99": float(np.percentile(errs, 99)),
                                "median": float(np.median(errs)), "results": len(errs)}
    return out, sols


def served_object(args, N, W, ctx):
    """The serving stream, served in FP64 and in mixed precision (DFM_MIXED_AUTO per request), and the plain FP64
    path one request at a time on the first --plain requests."""
    import numpy as np

    n = args.requests
    t0 = time.perf_counter()
    reqs = c4_requests(W, n)
    arrivals = W.c4_arrivals(n, args.rate) if args.rate > 0 else np.zeros(n)
    total = float(sum(r["ndof"] for r in reqs))
    log(f"[bench] c4: {n} requests ({total / 1e6:.1f} M dofs) generated in {time.perf_counter() - t0:.1f}s, "
        f"arrivals {args.rate}/s, {args.slots} slots")
    fp64, sols = serve_stream(N, ctx, reqs, arrivals, args.slots, N.FP64, keep=True, step_iters=args.step_iters)
    log(f"[bench] c4 served fp64: {fp64['dofs_per_s'] / 1e6:.2f} M DOF-solves/s over {fp64['wall_s']:.1f}s, "
        f"p50 {fp64['p50_ms']:.0f} ms, p99 {fp64['p99_ms']:.0f} ms, {fp64['met_tolerance']}/{n} met 1e-8 (max true "
        f"residual {fp64['max_true_rel_residual']:.2e})")
    mixed, _ = serve_stream(N, ctx, reqs, arrivals, args.slots, N.MIXED_AUTO, ref=sols, step_iters=args.step_iters)
    del sols
    log(f"[bench] c4 served mixed: {mixed['dofs_per_s'] / 1e6:.2f} M DOF-solves/s over {mixed['wall_s']:.1f}s, "
        f"p50 {mixed['p50_ms']:.0f} ms, p99 {mixed['p99_ms']:.0f} ms, {mixed['met_tolerance']}/{n} met 1e-8 (max true "
        f"residual {mixed['max_true_rel_residual']:.2e}), error vs fp64 max {mixed['error_vs_fp64']['max']:.2e}")
    plain = run_c4_plain(N, ctx, reqs, min(args.plain, n) if args.plain >= 0 else n)
    out = {
        "workload": (f"c4: {n} requests, seeded Poisson arrivals at {args.rate}/s, 2D Poisson (unit square, lognormal(0,1) "
                     f"conductivity) or 2D plane-stress elasticity (5:1 strip), 50/50, dofs log-uniform in [2k, 500k], "
                     f"64 GB admission budget, {args.slots} slots, 1e-8 relative residual"),
        "requests": n, "rate_per_s": args.rate, "total_dofs": total,
        "offered_dofs_per_s": total / (arrivals[-1] - arrivals[0]) if n > 1 and args.rate > 0 else None,
        "value": mixed["dofs_per_s"], "unit": "DOF-solves/s", "p50_ms": mixed["p50_ms"], "p99_ms": mixed["p99_ms"],
        "mixed": mixed, "fp64": fp64,
        "mixed_over_served_fp64": mixed["dofs_per_s"] / fp64["dofs_per_s"],
        "precision": "DFM_MIXED_AUTO (Poisson: FP32 matrix/factor/vectors with FP64 reliable updates; elasticity: "
                     "the FP64 loop with its true residual checked, DFM_FP64_TRUE); conv_ratio is the true relative "
                     "residual",
    }
    if plain:
        out["plain_fp64"] = plain
        out["mixed_over_plain_fp64"] = mixed["dofs_per_s"] / plain["dofs_per_s"]
        log(f"[bench] c4 plain FP64, one request at a time ({plain['requests']} requests): "
            f"{plain['dofs_per_s'] / 1e6:.2f} M DOF-solves/s end to end; served mixed / plain = "
            f"{out['mixed_over_plain_fp64']:.2f}, served mixed / served fp64 = {out['mixed_over_served_fp64']:.2f}")
    return out


def run_c4_serving(args, N, W):
    """The serving stream alone: the served object as the line."""
    ctx = N.Context(0)
    sv = served_object(args, N, W, ctx)
    ctx.close()
    out = {
        "metric": "DOF-solves/s at 1e-8 rel. residual (serving, mixed precision per request); p50/p99 request latency",
        "value": sv["value"], "unit": "DOF-solves/s", "n_gpus": 1, "steps": args.requests, "warmup": 0,
        "ms_per_step": sv["mixed"]["wall_s"] / args.requests * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic", "config": {"workload": sv["workload"]},
        "p50_ms": sv["p50_ms"], "p99_ms": sv["p99_ms"], "served": sv,
    }
    if not args.no_cpu:
        out["cpu_aggregate"] = c4_cpu_aggregate(args.cpu_budget * 2.0, min(args.requests, 400))
        log(f"[bench] c4 cpu aggregate: {out['cpu_aggregate']['value'] / 1e6:.3f} M DOF-solves/s "
            f"({out['cpu_aggregate']['sample']})")
    print(json.dumps(out), flush=True)


def saved_field(w, k):
    """Field values of re-solve k of the saved-stiffness workload: new prescribed values on the Dirichlet dofs
    (C2: a travelling wave along the boundary; C3: a small rotation and shear of the clamped face), zero
    elsewhere (MakeResidual reads only the prescribed ones)."""
    import numpy as np

    x = w["coords"].reshape(-1, w["ndim"])
    fl = w["bc"].reshape(x.shape[0], w["len"])
    u = np.zeros((x.shape[0], w["len"]))
    ph = 0.7 * k
    if w["len"] == 1:
        u[:, 0] = np.sin(2.0 * np.pi * (x[:, 0] + 0.5 * x[:, 1]) + ph)
    else:
        c = x - x.mean(axis=0)
        u[:, 0] = 1e-3 * (np.cos(ph) * c[:, 1] + 0.3 * np.sin(ph) * c[:, 0])
        u[:, 1] = 1e-3 * (-np.cos(ph) * c[:, 0] + 0.3 * np.sin(ph) * c[:, 1])
    u[fl == 0] = 0.0
    return u.ravel()


def run_saved(args, N, W, order_kind, w, ctx):
    """Saved-stiffness re-solves (CLinearSystem_Save): the matrix, its ILU(0) factor, the external force and the
    boundary coupling are kept; each step takes new Dirichlet values (a host upload of the field), recomputes
    the residual (MakeResidual), solves with PCG and sets the field (UpdateValueOfField). The step of the
    assembled path (assembly, Dirichlet treatment, factorisation) is what the saved path leaves out."""
    t0 = time.perf_counter()
    s = W.build_system(ctx, w, order_kind)
    ndof = s.ndof
    ctx.synchronize()
    a = time.perf_counter()
    s.assemble_save(w["kind"], w["params"])
    s.factor()
    ctx.synchronize()
    setup_ms = (time.perf_counter() - a) * 1e3
    log(f"[bench] saved {w['name']} ndof={ndof} boundary terms={s.boundary_terms()} assemble_save+factor "
        f"{setup_ms:.1f} ms (setup {time.perf_counter() - t0:.1f}s)")
    fields = [saved_field(w, k) for k in range(args.warmup + args.steps)]
    tol = w["tol"]

    def step(k, precision):
        ctx.synchronize()
        a = time.perf_counter()
        s.set_field(fields[k])
        s.make_residual()
        ok, conv, it = s.solve_pcg(tol, args.max_iter, precision)
        s.update_field_save()
        ctx.synchronize()
        return time.perf_counter() - a, ok, it

    res = {}
    for prec, name in [(N.MIXED, "mixed"), (N.FP64, "fp64")]:
        for k in range(args.warmup):
            step(k, prec)
        lat, its, ok_all = [], [], True
        for k in range(args.steps):
            dt, ok, it = step(args.warmup + k, prec)
            lat.append(dt)
            its.append(it)
            ok_all &= ok
        # the true residual of the solve itself, read before UpdateValueOfField rewrites the fixed dofs of the
        # update with the prescribed values (one more re-solve, untimed)
        s.set_field(fields[-1])
        s.make_residual()
        s.solve_pcg(tol, args.max_iter, prec)
        true_rel = s.true_residual()
        s.update_field_save()
        res[name] = dict(lat=lat, its=its, ok=ok_all, true_rel=true_rel)
        log(f"[bench] saved {name}: {statistics.median(lat) * 1e3:.1f} ms/re-solve, iters {its}, true rel residual "
            f"{true_rel:.2e}")
    mr_ms, mr_b = s.bench_kernel(6, 50)
    achieved = mr_b / (mr_ms * 1e-3) / 1e9
    log(f"[bench] make_residual {mr_ms * 1e3:.2f} us ({achieved:.0f} GB/s, {mr_b / 1e6:.1f} MB)")
    mixed = res["mixed"]
    t = sum(mixed["lat"])
    out = {
        "metric": "saved-stiffness re-solves: DOF-solves/s at 1e-8 rel. residual (new Dirichlet values per solve)",
        "value": ndof * args.steps / t,
        "unit": "DOF-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{w['name']} geometry and materials, saved stiffness, time-varying Dirichlet values",
                   "dofs": ndof, "ordering": args.ordering,
                   "step": "field upload + MakeResidual + PCG solve + UpdateValueOfField (no assembly, no factor)"},
        "fp64_value": ndof * args.steps / sum(res["fp64"]["lat"]),
        "p50_ms": statistics.median(mixed["lat"]) * 1e3,
        "fp64_p50_ms": statistics.median(res["fp64"]["lat"]) * 1e3,
        "iters_mixed_inner": mixed["its"][-1],
        "iters_fp64": res["fp64"]["its"][-1],
        "true_rel_residual_mixed": mixed["true_rel"],
        "true_rel_residual_fp64": res["fp64"]["true_rel"],
        "converged": bool(mixed["ok"] and res["fp64"]["ok"]),
        "assemble_save_plus_factor_ms": setup_ms,
        "boundary_terms": s.boundary_terms(),
        "roofline": {
            "kernel": "k_make_residual (MakeResidual: r = f - K_b u, update = u)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": mr_ms * 1e3,
            "algorithmic_bytes": mr_b,
        },
    }
    s.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_newmark(args, N, W, order_kind, w, ctx):
    """Newmark-beta time steps of linear elastodynamics (CEqn_Solid3D_Linear / CEqnSystem_Solid2D, not
    stationary): per step the Newmark assembly (M + beta dt^2 K, predictor residual), Dirichlet treatment, ILU
    factorisation, PCG to the workload's tolerance, UpdateValueOfField_NewmarkBeta. The workload's geometry,
    materials (density 1) and Dirichlet flags, body force along -z (-y in 2D), starting at rest."""
    import numpy as np

    s = W.build_system(ctx, w, order_kind)
    ndof = s.ndof
    kind = N.SOLID3D_NEWMARK if w["ndim"] == 3 else N.SOLID2D_NEWMARK
    par = np.array(w["params"], dtype=np.float64)
    if par.ndim == 1:
        par = par[None, :].repeat(w["lnods"].size // (w["ndim"] + 1), axis=0)
    par = par.copy()
    par[:, 2] = 1.0
    par[:, 3:] = 0.0
    par[:, -1] = -1e-2
    dt, gamma, beta = 0.05, 0.5, 0.25
    glob = (dt, gamma, beta, 1.0)
    s.set_field(None, None)
    s.set_acceleration(None)
    s.set_params(par)

    def step(precision):
        ctx.synchronize()
        a = time.perf_counter()
        s.assemble_resident(kind, glob=glob, norm=False)
        s.factor()
        ok, conv, it = s.solve_pcg(w["tol"], args.max_iter, precision)
        true_rel = s.true_residual()
        s.update_field_newmark(gamma, beta, dt, True)
        ctx.synchronize()
        return time.perf_counter() - a, ok, it, true_rel

    res = {}
    for prec, name in [(N.MIXED_AUTO, "auto"), (N.FP64, "fp64")]:
        s.set_field(None, None)
        s.set_acceleration(None)
        for _ in range(args.warmup):
            step(prec)
        lat, its, rels, ok_all = [], [], [], True
        for _ in range(args.steps):
            dt_, ok, it, rel = step(prec)
            lat.append(dt_)
            its.append(it)
            rels.append(rel)
            ok_all &= ok
        res[name] = dict(lat=lat, its=its, ok=ok_all, rel=max(rels))
        log(f"[bench] newmark {name}: {statistics.median(lat) * 1e3:.1f} ms/step, iters {its}, "
            f"max true rel residual {max(rels):.2e}")
    sw_ms, sw_b = s.bench_kernel(5, 20)
    achieved = sw_b / (sw_ms * 1e-3) / 1e9
    head = res["auto"]
    t = sum(head["lat"])
    out = {
        "metric": "Newmark-beta time steps: DOF-steps/s (assembly + factor + PCG to the tolerance + field update)",
        "value": ndof * args.steps / t,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{w['name']} geometry, elastodynamics, density 1, dt {dt}, beta {beta}, "
                               f"gamma {gamma}", "dofs": ndof, "ordering": args.ordering,
                   "tol_true_rel_residual": w["tol"], "precision": "DFM_MIXED_AUTO"},
        "fp64_value": ndof * args.steps / sum(res["fp64"]["lat"]),
        "p50_ms": statistics.median(head["lat"]) * 1e3,
        "fp64_p50_ms": statistics.median(res["fp64"]["lat"]) * 1e3,
        "iters": head["its"],
        "iters_fp64": res["fp64"]["its"],
        "max_true_rel_residual": head["rel"],
        "converged": bool(head["ok"] and res["fp64"]["ok"]),
        "roofline": {
            "kernel": "k_fwd_update/k_fwd_bwd_tile/k_bwd_rz<B=%d, FP64> (ILU sweep phases of the PCG iteration)" % w["len"],
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": sw_b,
        },
    }
    s.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def solve_legs(args, N, s, w, ctx):
    """Timed steps (assembly + Dirichlet + ILU factor + PCG) in mixed precision and in FP64; the solution of each
    leg's last step is kept for the error of the mixed result against the FP64 one."""
    tol = w["tol"]

    def step(precision):
        ctx.synchronize()
        a = time.perf_counter()
        s.assemble_resident(w["kind"], norm=False)
        s.factor()
        ok, conv, it = s.solve_pcg(tol, args.max_iter, precision)
        ctx.synchronize()
        return time.perf_counter() - a, ok, conv, it

    res = {}
    for prec, name in [(N.MIXED, "mixed"), (N.FP64, "fp64")]:
        for _ in range(args.warmup):
            step(prec)
        lat, its = [], []
        ok_all = True
        conv = 0.0
        for k in range(args.steps):
            dt, ok, conv, it = step(prec)
            lat.append(dt)
            its.append(it)
            ok_all &= ok
        true_rel = s.true_residual()
        st = s.last_solve_stats()
        nrel, fell_back = s.last_solve_info()
        res[name] = dict(lat=lat, its=its, ok=ok_all, conv=conv, true_rel=true_rel, stats=st, nrel=nrel,
                         fell_back=fell_back, x=s.get_vector(-2))
        log(f"[bench] {name}: {statistics.median(lat) * 1e3:.1f} ms/solve, iters {its[-1]} (executed {st['iters']}), "
            f"conv {conv:.2e}, "
            f"true rel residual {true_rel:.2e}, solve {st['ms']:.1f} ms, {st['ms_per_iter'] * 1e3:.1f} us/iter"
            + (f", reliable updates {nrel}" if name == "mixed" else "") + (", fell back to FP64" if fell_back else ""))
    return res


def pct(v, q):
    v = sorted(v)
    if len(v) == 1:
        return v[0]
    i = q * (len(v) - 1)
    lo = int(i)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (i - lo)


WORKLOAD_NAMES = {
    "c1": "2D plane-stress elasticity P1, rectangle 1.0x0.2, 128x32 quads x 2 tris, E=1, nu=0.3, left edge clamped",
    "c2": "2D Poisson P1, unit square 1024x1024 quads x 2 tris, lognormal(0,1) conductivity, f=1, u=0 on the boundary",
    "c3": "3D linear elasticity P1 tets, unit cube 96^3 hexes x 6, E lognormal(0,0.5), nu=0.3, bottom clamped, "
          "seeded top traction",
}


def run_single(args, N, W, w, ctx, order_kind):
    """One system of a configuration: timed mixed and FP64 steps, kernel timings, roofline, CPU baseline; with
    the serving stream beside it unless --no-served."""
    import numpy as np

    t0 = time.perf_counter()
    s = W.build_system(ctx, w, order_kind)
    s.set_params(w["params"])
    s.set_loads(w["faces"], w["trac"])
    nph = s.schedule_info()
    ndof = s.ndof
    log(f"[bench] {w['name']} ndof={ndof} ncrs={s.ncrs} ordering={args.ordering} phases={nph[:3]} "
        f"setup {time.perf_counter() - t0:.1f}s")
    res = solve_legs(args, N, s, w, ctx)
    mixed, fp64 = res["mixed"], res["fp64"]
    err = float(np.linalg.norm(mixed["x"] - fp64["x"]) / np.linalg.norm(fp64["x"]))

    # kernels of the inner loops and the factorisation, timed with HIP events on the library's stream
    spmv32_ms, spmv32_b = s.bench_kernel(1, 50)
    spmv64_ms, spmv64_b = s.bench_kernel(0, 50)
    ilu32_ms, ilu32_b = s.bench_kernel(3, 20)
    ilu64_ms, ilu64_b = s.bench_kernel(2, 20)
    sw32_ms, sw32_b = s.bench_kernel(4, 20)
    sw64_ms, sw64_b = s.bench_kernel(5, 20)
    fac_ms, fac_b = s.bench_kernel(7, 5)
    log(f"[bench] spmv fp32 {spmv32_ms * 1e3:.1f} us ({spmv32_b / spmv32_ms / 1e6:.0f} GB/s), spmv fp64 "
        f"{spmv64_ms * 1e3:.1f} us ({spmv64_b / spmv64_ms / 1e6:.0f} GB/s), pcg sweep phases fp32 {sw32_ms * 1e3:.1f} us "
        f"({sw32_b / sw32_ms / 1e6:.0f} GB/s), fp64 {sw64_ms * 1e3:.1f} us ({sw64_b / sw64_ms / 1e6:.0f} GB/s), "
        f"ILU factor {fac_ms:.3f} ms ({fac_b / fac_ms / 1e6:.0f} GB/s); mixed vs fp64 solution error {err:.2e}")
    nlaunch = nph[3]  # sweep launches per PCG iteration (the last forward and first backward phase may be one)
    B = w["len"]
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated against a copy kernel) from the
    # committed summary of this workload
    prof = f"{w['name']}_hbm_traffic.md"
    blockcolor = args.ordering == "blockcolor"
    sweep_traffic = measured_traffic(prof, ["sweep_phases_fp32_per_launch"]) if blockcolor else None
    spmv_traffic = measured_traffic(prof, [f"dfm::k_pcg_spmv_sell<{B}, float, false>"]) if blockcolor else None
    K = args.steps
    head = fp64 if args.config == "c1" else mixed  # the plain-FP64 configuration reports its FP64 path
    t_head = sum(head["lat"])
    value = ndof * K / t_head
    value64 = ndof * K / sum(fp64["lat"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": t_head / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head is fp64 else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{w['name']}: {WORKLOAD_NAMES[w['name']]}",
            "dofs": ndof,
            "nnz_blocks_offdiag": s.ncrs,
            "block": B,
            "tol_rel_residual": w["tol"],
            "ordering": args.ordering,
            "ilu_phases_fwd_bwd": [nph[0], nph[1]],
            "step": "assemble + Dirichlet + ILU(0) factor + PCG solve",
        },
        "value_path": "fp64" if head is fp64 else "mixed",
        "fp64_value": value64,
        "mixed_value": ndof * K / sum(mixed["lat"]),
        "mixed_over_fp64": (ndof * K / sum(mixed["lat"])) / value64,
        "p50_ms": pct(head["lat"], 0.5) * 1e3,
        "p99_ms": pct(head["lat"], 0.99) * 1e3,
        "fp64_p50_ms": pct(fp64["lat"], 0.5) * 1e3,
        "fp64_p99_ms": pct(fp64["lat"], 0.99) * 1e3,
        "iters_mixed_inner": mixed["stats"]["iters"],
        "iters_fp64": fp64["its"][-1],
        "iters_executed_fp64": fp64["stats"]["iters"],
        "true_rel_residual_mixed": mixed["true_rel"],
        "true_rel_residual_fp64": fp64["true_rel"],
        "mixed_error_vs_fp64": err,
        "converged": bool(mixed["ok"] and fp64["ok"]),
        "mixed_reliable_updates": mixed["nrel"],
        "mixed_fell_back_to_fp64": mixed["fell_back"],
        "ms_per_pcg_iter_mixed": mixed["stats"]["ms_per_iter"],
        "ms_per_pcg_iter_fp64": fp64["stats"]["ms_per_iter"],
        "solve_ms_mixed": mixed["stats"]["ms"],
        "solve_ms_fp64": fp64["stats"]["ms"],
        # dominant kernel family of the mixed inner loop: the sweep phases (one launch per phase and direction,
        # forward phases with the r/x update, backward phases with (r, z)); per launch = one iteration's
        # algorithmic bytes of the phases and HIP-event time, over its launches
        "roofline": {
            "kernel": "k_fwd_update/k_fwd_bwd_tile/k_bwd_rz<B=%d, FP32> (ILU(0) phase sweeps of the PCG "
                      "iteration, %d launches)" % (B, nlaunch),
            "bound": "hbm",
            "achieved": sw32_b / (sw32_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sw32_b / (sw32_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": sweep_traffic,
            "avg_launch_us": sw32_ms * 1e3 / nlaunch,
            "algorithmic_bytes": sw32_b / nlaunch,
        },
        "roofline_spmv": {
            "kernel": "k_pcg_spmv_sell<B=%d, FP32> (SELL-64 SpMV + dot)" % B,
            "bound": "hbm",
            "achieved": spmv32_b / (spmv32_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmv32_b / (spmv32_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": spmv_traffic,
            "avg_launch_us": spmv32_ms * 1e3,
            "algorithmic_bytes": spmv32_b,
        },
        "kernels": {
            "spmv_fp64": {"us": spmv64_ms * 1e3, "GB/s": spmv64_b / spmv64_ms / 1e6, "bytes": spmv64_b},
            "ilu_apply_fp32": {"us": ilu32_ms * 1e3, "GB/s": ilu32_b / ilu32_ms / 1e6, "bytes": ilu32_b},
            "ilu_apply_fp64": {"us": ilu64_ms * 1e3, "GB/s": ilu64_b / ilu64_ms / 1e6, "bytes": ilu64_b},
            "pcg_sweep_phases_fp64": {"us": sw64_ms * 1e3, "GB/s": sw64_b / sw64_ms / 1e6, "bytes": sw64_b},
            "ilu_factor_fp64": {"us": fac_ms * 1e3, "GB/s": fac_b / fac_ms / 1e6, "bytes": fac_b},
        },
    }
    s.close()
    if args.config == "c3" and not args.no_served:
        out["served"] = served_object(args, N, W, ctx)
    ctx.close()
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        try:
            out["cpu_baseline"] = cpu_baseline(w, fp64["its"][-1] + 1, args.cpu_budget)
            out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
            log(f"[bench] cpu baseline: {out['cpu_baseline']['sample']}")
        except Exception as e:  # the baseline is reported, never the measured product
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--ordering", default="blockcolor", choices=["natural", "multicolor", "blockcolor"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="c3: leave out the serving stream")
    ap.add_argument("--newmark", action="store_true",
                    help="c1/c3: Newmark-beta time steps of linear elastodynamics on the workload's geometry")
    ap.add_argument("--saved", action="store_true",
                    help="c1/c2/c3: saved-stiffness re-solves with new Dirichlet values per step")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--max-iter", type=int, default=20000)
    ap.add_argument("--instances", type=int, default=256, help="c5: instances in the batch")
    ap.add_argument("--load-steps", type=int, default=10, help="c5: load steps")
    ap.add_argument("--lev", type=int, default=0, help="c5: ILU fill level of the inner solves")
    ap.add_argument("--c5-extra", default="", choices=["", "fp32ilu", "mixed32"],
                    help="c5: a third leg with DFM_FP32_ILU (fp32ilu) or DFM_MIXED (mixed32) inner solves")
    ap.add_argument("--requests", type=int, default=2000, help="c4: requests in the stream")
    ap.add_argument("--rate", type=float, default=100.0, help="c4: arrival rate (requests/s)")
    ap.add_argument("--step-iters", type=int, default=16, help="c4: batched iterations per server step")
    ap.add_argument("--slots", type=int, default=128, help="c4: systems solved at once (128: measured against 64 and 256)")
    ap.add_argument("--plain", type=int, default=200,
                    help="c4: requests (the first ones of the stream) also solved by the plain FP64 path one at a "
                         "time; -1: all, 0: none")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (--gpus {args.gpus} is not supported): requests and instance batches "
              f"shard across GPUs with no traffic between them, one large system by domain decomposition, which "
              f"this program does not run", file=sys.stderr)
        sys.exit(2)

    from delfem_amd import _native as N
    from delfem_amd import workloads as W

    if args.config == "c5":
        return run_c5_batch(args, N, W)
    if args.config == "c4":
        return run_c4_serving(args, N, W)
    w = W.make(args.config)
    ctx = N.Context(0)
    order_kind = {"natural": N.ORDER_NATURAL, "multicolor": N.ORDER_MULTICOLOR,
                  "blockcolor": N.ORDER_BLOCKCOLOR}[args.ordering]
    if args.saved:
        return run_saved(args, N, W, order_kind, w, ctx)
    if args.newmark:
        if w["len"] == 1:
            print("--newmark needs a vector workload (c1 or c3)", file=sys.stderr)
            sys.exit(2)
        return run_newmark(args, N, W, order_kind, w, ctx)
    return run_single(args, N, W, w, ctx, order_kind)


if __name__ == "__main__":
    main()
