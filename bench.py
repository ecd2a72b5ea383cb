"""Benchmark: 3D evolution steps/s of the forward shape-evolution solve on one
MI355X, with the roofline of its dominant kernel and the FP64 CPU oracle
timed beside it.

    python bench.py [--steps K (50)] [--warmup W] [--repeats R] [--config c2|c2-all|c1]
                    [--precision fp64|mixed] [--no-cpu]

A "step" is one evolution time step (velocity solve, landmark update,
reaction-diffusion update) of the workload; the timed region is R solves of
K steps each, every solve ending with the varifold objective, with the mesh
and parameters already resident in HBM (a device-resident session).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D evolution steps/s at ≤1e-6 rel err vs FP64; p50/p99 solve latency"

# algorithmic work (counted per unit, see the measurement notes of the design)
FLOP_PER_PAIR = {-1: 14, 0: 18, 1: 19, 2: 21, 3: 23, 4: 25}   # d2 (8) + kernel + 3 multiply-adds (6)
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the stored product, from rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE runs of this script (KiB per dispatch, averaged over the
# kstore_mv_k dispatches), each scaled by what the same counter reads for a
# plain device copy of known size (--copy-calibrate): FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0 on MI355X.  Keyed by (config, precision); bytes per launch,
# mean of the K_xc u and K_cx y launches (profiles/kstore_mv_k_*_c2_*.md).
TRAFFIC: dict = {("c2", "mixed"): 1.0517e9, ("c2", "fp64"): 2.0645e9}
TRAFFIC_SOURCE = ("not measured in this run: the counter runs of profiles/kstore_mv_k_float_c2_mixed.md and "
                  "profiles/kstore_mv_k_double_c2_fp64.md (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)")


def workload(name: str, steps: int):
    from shape_atrophy_amd import workloads as W
    if name == "c2":
        p = W.shell_problem(level=5, n_layers=5, time_steps=50, n_ctl=5000)
        desc = "ellipsoidal shell, icosphere-5 x 5 layers, Matern nu=3/2 on 5000 control points"
    elif name == "c2-all":
        p = W.shell_problem(level=5, n_layers=5, time_steps=50)
        desc = "ellipsoidal shell, icosphere-5 x 5 layers, Matern nu=3/2 on all landmarks"
    elif name == "c1":
        p = W.ring_problem()
        desc = "2D concentric-ring disk, 10 rings x 200 nodes"
    else:
        raise SystemExit(f"unknown config {name}")
    p["timeNum"] = steps + 1
    return p, desc


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def run_requests(api, meshes, reqs, precision, workers, mem_budget=0, rate=None, seed=5):
    """Submit requests (mesh index, tauVar, params) to a server holding the
    meshes; all at once (rate None) or as a Poisson stream of the given rate
    (requests/s).  Returns the results and the wall time from the first
    submission to the last completion."""
    srv = api.Server(workers=workers, mem_budget=mem_budget)
    ids = [srv.add_mesh(p, precision) for p in meshes]
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    t_next = t0
    rids = []
    for m, tv, prm in reqs:
        if rate:
            t_next += rng.exponential(1.0 / rate)
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        rids.append(srv.submit(ids[m], tv, meshes[m]["tauPrm"], prm))
    out = []
    for k, r in enumerate(rids):
        g = srv.wait(r)
        out.append(g)
        m = reqs[k][0]
        print(f"[bench]   {precision} req {k}: V={meshes[m]['q0'].shape[1]} d={meshes[m]['q0'].shape[0]} "
              f"steps={g['steps']} iters={g['pcg_iters']} sweeps={g['refine']} fp64_steps={g['fp64_steps']} "
              f"fail={g['fail_reason']} "
              f"run={g['seconds']:.2f}s done_at={time.perf_counter() - t0:.1f}s rc={g['rc']} status={g['status']}",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    srv.close()
    return out, wall


def mode_precision(args, api, W):
    """Mixed form against the plain FP64 form on one workload: relative
    errors of the final displacement, the concentration history, the
    objective and its gradient (config c3: the 2D ring of c1 with the kernel
    on every landmark as control points)."""
    steps = args.steps if args.steps is not None else 20
    if args.config == "c3":
        prob = W.ring_problem()
        prob["ctlIdx"] = np.arange(prob["q0"].shape[1], dtype=np.int32)
        desc = "2D concentric-ring disk, 10 rings x 200 nodes, Matern nu=4 on every landmark (control points = landmarks)"
    else:
        prob, desc = workload(args.config, steps)
    prob["params"] = dict(prob["params"], timeStp=1.0 / steps)
    prob["timeNum"] = steps + 1
    out = {}
    for prec in ("fp64", "mixed"):
        s = api.Session(prob, precision=prec)
        t0 = time.perf_counter()
        r = s.run(prob["tauVar"][:, 0], prob["tauPrm"], want_traj=True, want_grad=True)
        r["wall"] = time.perf_counter() - t0
        s.close()
        out[prec] = r
        print(f"[bench] {prec}: {r['steps']} steps {r['wall']:.3f}s obj={r['obj']:.15e} iters={r['pcg_iters']} "
              f"sweeps={r['refine']} fp64_steps={r['fp64_steps']} status={r['status']}", file=sys.stderr, flush=True)
    a, m = out["fp64"], out["mixed"]
    disp = np.linalg.norm(a["lmkStk"][-1] - a["lmkStk"][0])
    errs = {
        "displacement": float(np.linalg.norm(m["lmkStk"][-1] - a["lmkStk"][-1]) / disp),
        "concentration": float(np.linalg.norm(m["tauMat"] - a["tauMat"]) / np.linalg.norm(a["tauMat"])),
        "objective": float(abs(m["obj"] - a["obj"]) / abs(a["obj"])),
        "gradient": float(np.linalg.norm(m["grad"] - a["grad"]) / np.linalg.norm(a["grad"])),
    }
    return {
        "metric": "relative error of the mixed form vs the plain FP64 form",
        "value": max(errs.values()), "unit": "max relative error", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1e3 * m["wall"] / max(m["steps"], 1), 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": desc, "config": args.config, "landmarks": int(prob["q0"].shape[1]),
                   "control_points": int(len(prob.get("ctlIdx", [])))},
        "rel_err": errs,
        "mixed": {"steps_per_s": round(m["steps"] / m["wall"], 3), "pcg_iters": m["pcg_iters"],
                  "refine_sweeps": m["refine"], "fp64_steps": m["fp64_steps"]},
        "fp64": {"steps_per_s": round(a["steps"] / a["wall"], 3), "pcg_iters": a["pcg_iters"]},
    }


def mode_batch(args, api, W):
    """64 independent meshes of uneven size (half 2D rings, half 3D shells),
    each with its own parameters, 20 steps, run concurrently through the
    server in the mixed form, against the same solves run one at a time in
    the plain FP64 form."""
    steps = args.steps if args.steps is not None else 20
    meshes = W.batch_problems(n=args.batch_n, time_steps=steps)
    reqs = [(k, p["tauVar"][:, 0], p["params"]) for k, p in enumerate(meshes)]
    sizes = [p["q0"].shape[1] for p in meshes]
    print(f"[bench] batch: {len(meshes)} meshes, V {min(sizes)}..{max(sizes)}, {steps} steps", file=sys.stderr,
          flush=True)
    got, wall = run_requests(api, meshes, reqs, args.precision, args.workers)
    failed = [k for k, g in enumerate(got) if g["rc"] != 0 or g["status"] != 0]
    if failed:
        print(f"[bench] batch: {len(failed)} failed requests: {failed}", file=sys.stderr, flush=True)
    tot_steps = sum(g["steps"] for g in got)
    print(f"[bench] batched {args.precision}: {tot_steps} steps in {wall:.2f}s", file=sys.stderr, flush=True)
    ref, wall64 = run_requests(api, meshes, reqs, "fp64", 1)
    print(f"[bench] serial fp64: {sum(g['steps'] for g in ref)} steps in {wall64:.2f}s", file=sys.stderr, flush=True)
    ok = [k for k in range(len(got)) if k not in failed and ref[k]["status"] == 0 and ref[k]["rc"] == 0]
    err = max((abs(got[k]["obj"] - ref[k]["obj"]) / abs(ref[k]["obj"]) for k in ok), default=None)
    lat = [g["t_end"] - g["t_submit"] for g in got]
    return {
        "metric": "batched evolution steps/s (64 meshes, mixed form) vs serial plain FP64",
        "value": round(tot_steps / wall, 3), "unit": "steps/s", "n_gpus": 1, "steps": tot_steps,
        "warmup": 0, "ms_per_step": round(1e3 * wall / tot_steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": f"{len(meshes)} layered meshes, half 2D rings V 1k-20k, half 3D shells V 5k-60k, "
                               f"log-uniform sizes, per-mesh parameters +-30%, control points 10% (max 2000)",
                   "time_steps": steps, "workers": args.workers, "precision": args.precision,
                   "landmarks_total": int(sum(sizes))},
        "serial_fp64": {"value": round(sum(g["steps"] for g in ref) / wall64, 3), "unit": "steps/s"},
        "speedup_vs_serial_fp64": round(wall64 / wall, 3),
        "max_obj_rel_err_vs_fp64": err,
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "failed_requests": len(failed),
    }


def mode_serve(args, api, W):
    """Poisson stream of forward-evolution + objective requests over a pool of
    meshes (2D/3D, V log-uniform 1k-200k), each request with its own
    parameter vector; rate 0.7 x the saturated throughput measured first.
    p50/p99 latency and requests/s, against serial plain FP64 solves."""
    steps = args.steps if args.steps is not None else 10
    pool = W.batch_problems(n=args.pool, seed=5, time_steps=steps, v2d=(1000, 200000), v3d=(1000, 200000),
                            mixed_dims=True)
    rng = np.random.default_rng(55)
    reqs = []
    for k in range(args.requests):
        m = int(rng.integers(len(pool)))
        prm = W.perturb_params(pool[m]["params"], rng)
        tv = pool[m]["tauVar"][:, 0] * (1.0 + 0.05 * rng.standard_normal(pool[m]["tauVar"].shape[0]))
        reqs.append((m, tv, prm))
    sizes = [p["q0"].shape[1] for p in pool]
    print(f"[bench] serve: pool of {len(pool)} meshes V {min(sizes)}..{max(sizes)}, {len(reqs)} requests",
          file=sys.stderr, flush=True)
    sat, wsat = run_requests(api, pool, reqs, args.precision, args.workers, mem_budget=args.mem_budget)
    x_sat = len(reqs) / wsat
    rate = 0.7 * x_sat
    print(f"[bench] saturated: {x_sat:.3f} req/s; Poisson rate {rate:.3f} req/s", file=sys.stderr, flush=True)
    got, wall = run_requests(api, pool, reqs, args.precision, args.workers, mem_budget=args.mem_budget, rate=rate)
    ref, w64 = run_requests(api, pool, reqs, "fp64", 1)
    failed = [k for k, g in enumerate(sat + got + ref) if g["rc"] != 0 or g["status"] != 0]
    if failed:
        print(f"[bench] serve: {len(failed)} failed requests (indices into burst+stream+fp64): {failed}",
              file=sys.stderr, flush=True)
    lat = [g["t_end"] - g["t_submit"] for g in got]
    qlat = [g["t_start"] - g["t_submit"] for g in got]
    ok = [k for k in range(len(reqs)) if all(x[k]["status"] == 0 and x[k]["rc"] == 0 for x in (sat, got, ref))]
    err = max((abs(got[k]["obj"] - ref[k]["obj"]) / abs(ref[k]["obj"]) for k in ok), default=None)
    same = max((abs(got[k]["obj"] - sat[k]["obj"]) / abs(sat[k]["obj"]) for k in ok), default=None)
    return {
        "metric": "served requests/s and p50/p99 latency (Poisson arrivals, mixed form) vs serial plain FP64",
        "value": round(len(reqs) / wall, 4), "unit": "requests/s", "n_gpus": 1, "steps": int(sum(g["steps"] for g in got)),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": f"pool of {len(pool)} layered meshes (2D rings / 3D shells) V 1k-200k log-uniform, "
                               f"{len(reqs)} requests with per-request parameters, {steps} steps each",
                   "workers": args.workers, "precision": args.precision, "arrival_rate": round(rate, 4)},
        "saturated": {"value": round(x_sat, 4), "unit": "requests/s"},
        "serial_fp64": {"value": round(len(reqs) / w64, 4), "unit": "requests/s"},
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "queue_s": {"p50": percentile(qlat, 50), "p99": percentile(qlat, 99)},
        "max_obj_rel_err_vs_fp64": err,
        "max_rel_diff_stream_vs_burst": same,
        "failed_requests": len(failed),
    }


def mode_fit(args, api, W):
    """Fit of the four tau variables of the 2D ring (C1 mesh, 20 steps) to a
    target generated by the forward model at known values, from a start 8-12 %
    away: BFGS with strong-Wolfe line searches, each gradient a batch of nine
    candidates (central differences) run concurrently on the request server.
    Throughput in objective evaluations/s, against the same candidate batches
    run one at a time on a session and against the FP64 oracle on the host."""
    from shape_atrophy_amd import fit
    prec = "fp64"   # all-landmark kernel: the mixed form applies to control points only
    steps = args.steps if args.steps is not None else 20
    p = W.ring_problem()
    p["timeNum"] = steps + 1
    th = p["tauVar"][:, 0].copy()
    s = api.Session(p)
    r = s.run(th, p["tauPrm"], want_traj=True)
    p["target"] = fit.target_from_shape(p, r["lmkStk"][-1])
    s.close()
    x0 = th * np.array([1.08, 0.92, 1.12, 0.93])
    print(f"[bench] fit: V={p['q0'].shape[1]} steps={steps} theta*={th} x0={x0}", file=sys.stderr, flush=True)
    obj = fit.ServerObjective(p, precision=prec, workers=args.workers)
    obj(fit.fd_candidates(x0, fit.fd_steps(x0, fit.OptData())))        # warm-up: sessions created
    t0 = time.perf_counter()
    res = fit.bfgs(obj, x0, fit.OptData(itrMax=60, tolVal=1e-7, verbose=True))
    wall = time.perf_counter() - t0
    obj.close()
    err = float(np.abs(res.x - th).max())
    print(f"[bench] fit: {res.status} after {res.iterations} iterations, {res.evaluations} evaluations in "
          f"{wall:.2f}s; max |theta - theta*| {err:.3e}", file=sys.stderr, flush=True)
    # the same candidate batch served one at a time
    X = fit.fd_candidates(res.x, fit.fd_steps(res.x, fit.OptData()))
    so = fit.SessionObjective(p, precision=prec)
    so(X[:, :1])
    t1 = time.perf_counter()
    so(X)
    serial = X.shape[1] / (time.perf_counter() - t1)
    so.close()
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.build()
        n_cpu = 2
        t2 = time.perf_counter()
        for j in range(n_cpu):
            O.objfcn(dict(p, tauVar=X[:, j:j + 1].copy()), threads=os.cpu_count())
        cpu = {"value": round(n_cpu / (time.perf_counter() - t2), 4), "unit": "evaluations/s",
               "cores": os.cpu_count(), "kind": "port",
               "sample": f"{n_cpu} objective evaluations of the same workload, FP64 oracle"}
    return {
        "metric": "objective evaluations/s while fitting tauVar (BFGS, central-difference gradients)",
        "value": round(res.evaluations / wall, 3), "unit": "evaluations/s", "n_gpus": 1,
        "steps": res.evaluations * steps, "warmup": 1, "ms_per_step": round(1e3 * wall / (res.evaluations * steps), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2D concentric-ring disk, 10 rings x 200 nodes, target from the forward model at "
                               "known tau variables", "time_steps": steps, "workers": args.workers,
                   "precision": prec},
        "fit": {"status": res.status, "iterations": res.iterations, "evaluations": res.evaluations,
                "batches": res.batches, "seconds": round(wall, 3), "objective": res.f,
                "max_abs_err_theta": err},
        "serial_session": {"value": round(serial, 3), "unit": "evaluations/s"},
        "cpu_baseline": cpu,
    }


def mode_products(args, api, W):
    """The control-point kernel products of the mixed form's inner solve in
    their two implementations on one workload: stored FP32 matrices
    (kstore_mv_k) and matrix-free on the matrix cores (kx_mfma_k,
    SA_KX_MFMA=1).  Relative error of K_xc u and K_cx y against the FP64
    products, and a solve of --steps steps in each (CG iterations, the
    average product time from the kernel timer, steps/s)."""
    steps = args.steps if args.steps is not None else 10
    prob, desc = workload(args.config, steps)
    d, V = prob["q0"].shape
    nc = len(prob["ctlIdx"])
    rng = np.random.default_rng(1)
    u, y = rng.standard_normal((d, nc)), rng.standard_normal((d, V))
    tv, tp = prob["tauVar"][:, 0], prob["tauPrm"]
    s64 = api.Session(prob, precision="fp64")
    ref = (s64.probe("kxc", tv, tp, u).reshape(d, V), s64.probe("kcx", tv, tp, y).reshape(d, nc))
    r64 = s64.run(tv, tp)
    s64.close()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {}
    for name, env in (("stored_fp32", "0"), ("matrix_cores", "1")):
        os.environ["SA_KX_MFMA"] = env
        s = api.Session(prob, precision="mixed")
        kx = s.probe("kxc", tv, tp, u).reshape(d, V)
        kc = s.probe("kcx", tv, tp, y).reshape(d, nc)
        s.run(tv, tp)                      # warm-up
        s.set_profiling(True)
        r = s.run(tv, tp)
        s.close()
        out[name] = {"kxc_rel_err": rel(kx, ref[0]), "kcx_rel_err": rel(kc, ref[1]), "steps": r["steps"],
                     "seconds": round(r["seconds"], 4), "steps_per_s": round(r["steps"] / r["seconds"], 3),
                     "pcg_iters": r["pcg_iters"], "refine_sweeps": r["refine"],
                     "product_avg_us": round(1e3 * r["kmatvec_ms"] / max(r["kmatvec_count"], 1), 1),
                     "obj_rel_err_vs_fp64": abs(r["obj"] - r64["obj"]) / abs(r64["obj"])}
        print(f"[bench] {name}: {out[name]}", file=sys.stderr, flush=True)
    os.environ.pop("SA_KX_MFMA", None)
    return {"metric": "mixed-form inner kernel products: stored FP32 vs matrix cores",
            "value": out["matrix_cores"]["product_avg_us"], "unit": "us per product (matrix cores)",
            "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": None, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
            "config": {"workload": desc, "landmarks": V, "control_points": nc}, "fp64": {"pcg_iters": r64["pcg_iters"]},
            **out}


def mode_meshes(args, api, W):
    """Three meshes of the 64-mesh batch solved one at a time on their own
    sessions (no server, one stream): a 3D shell of 5k-9k landmarks, a 2D
    ring of 1k-3k and a 3D shell of 20k-40k, mixed form, --steps steps
    (default 10) after a warm-up solve.  Per mesh: wall time per CG
    iteration; under rocprofv3 --kernel-trace, the per-kernel times of
    small meshes."""
    steps = args.steps if args.steps is not None else 10
    probs = W.batch_problems(n=args.batch_n, time_steps=steps)
    sizes = [p["q0"].shape[1] for p in probs]
    out = []
    for want_d, lo, hi in ((3, 5000, 9000), (2, 1000, 3000), (3, 20000, 40000)):
        k = next((k for k, p in enumerate(probs) if p["dim"] == want_d and lo <= sizes[k] <= hi), None)
        if k is None:
            continue
        p = probs[k]
        s = api.Session(p, precision=args.precision)
        try:
            tv, tp = p["tauVar"][:, 0], p["tauPrm"]
            s.run(tv, tp)
            t0 = time.perf_counter()
            r = s.run(tv, tp)
            wall = time.perf_counter() - t0
        finally:
            s.close()
        row = {"mesh": k, "dim": p["dim"], "landmarks": sizes[k], "control_points": len(p["ctlIdx"]),
               "steps": r["steps"], "seconds": round(wall, 4), "pcg_iters": r["pcg_iters"],
               "refine_sweeps": r["refine"], "fp64_steps": r["fp64_steps"],
               "us_per_cg_iteration": round(1e6 * wall / max(r["pcg_iters"], 1), 1)}
        print(f"[bench] mesh {row}", file=sys.stderr, flush=True)
        out.append(row)
    return {"metric": "single-mesh solves from the batch: wall time per CG iteration", "value": None,
            "unit": "us per CG iteration", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": None,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
            "data": "synthetic", "config": {"workload": "three meshes of the 64-mesh batch, one at a time",
                                            "precision": args.precision}, "meshes": out}


def mode_operator(args, api, W):
    """The operator apply (E + btmWgt P) x of one workload, as the CG runs it:
    the fused launch (op_fused_k, the default) against the two-launch apply
    (elastic_elem_k + gather_op_k, SA_OP_FUSED=0).  Device time per apply
    from sa_session_time_product (mean of --tune-reps launches, best of
    three), against the algorithmic bytes 64 E + 72 V (connectivity and
    frames per element; positions, input and output per vertex, each once)."""
    prob, desc = workload(args.config, 1)
    d, V = prob["q0"].shape
    E = prob["elm"]["elmVtxMat"].shape[1]
    alg = 64 * E + 72 * V if d == 3 else 40 * E + 48 * V
    out = {}
    for name, env in (("fused", "1"), ("two_launch", "0")):
        os.environ["SA_OP_FUSED"] = env
        s = api.Session(prob)
        s.probe("elastic", prob["tauVar"][:, 0], prob["tauPrm"], np.zeros((d, V)))
        us = min(s.time_product("operator", args.tune_reps) for _ in range(3))
        s.close()
        out[name] = {"us": round(us, 2), "achieved_GBs": round(alg / us * 1e-3, 1),
                     "frac": round(alg / us * 1e-3 / PEAK_HBM_GBS, 4)}
        print(f"[bench] operator {name}: {out[name]}", file=sys.stderr, flush=True)
    os.environ.pop("SA_OP_FUSED", None)
    return {"metric": "operator apply (E + btmWgt P) x, device time per apply", "value": out["fused"]["us"],
            "unit": "us", "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": desc, "landmarks": V, "elements": E, "reps": args.tune_reps},
            "algorithmic_bytes": alg, **out}


TUNE_TARGETS = (32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384)
# FP32 products: a block needs at most 24 KiB of LDS, so the targets between
# 256 and 2048 are resolved finer (one resident round is 4-6 blocks per CU)
TUNE_TARGETS_F32 = (32, 64, 128, 256, 384, 512, 768, 1024, 1280, 1536, 2048, 4096)
# the library's built-in targets (KS_BLOCKS = 1024; 2048 for the all-landmark
# product), kept unless a target is more than TUNE_KEEP faster: two sweeps on
# different boxes differed by up to 5% at the same setting on the large
# products
TUNE_DEFAULT = {("xc", "f32"): 1024, ("cx", "f32"): 1024, ("xc", "f64"): 1024, ("cx", "f64"): 1024,
                ("cc", "f32"): 1024, ("cc", "f64"): 1024, ("kk", "f64"): 2048,
                ("op", "f64"): 1}   # the fused operator apply's launch shape (index, 0-3)
OP_SHAPES = (0, 1, 2, 3)
TUNE_KEEP = 1.05


def tune_pick(prod: str, ty: str, times: dict) -> int:
    """The target the table keeps: the built-in one unless another is more
    than TUNE_KEEP faster, then the smallest within 1% of the fastest (for
    the operator, the launch shape)."""
    best = min(times.values())
    dflt = TUNE_DEFAULT[(prod, ty)]
    if dflt in times and times[dflt] <= TUNE_KEEP * best:
        return dflt
    return min(t for t in times if times[t] <= 1.01 * best)


def tune_shapes(W):
    """Meshes of the tile sweep: control-point shells from the batch range to
    the largest served size, and all-landmark meshes."""
    ctl = [("shell 3x4, 256 control points", dict(level=3, n_layers=4, n_ctl=256)),
           ("shell 4x3, 500 control points", dict(level=4, n_layers=3, n_ctl=500)),
           ("shell 4x5, 1281 control points", dict(level=4, n_layers=5, n_ctl=1281)),
           ("shell 5x3, 2000 control points", dict(level=5, n_layers=3, n_ctl=2000)),
           ("shell 5x5, 5000 control points", dict(level=5, n_layers=5, n_ctl=5000)),
           ("shell 6x5, 2000 control points", dict(level=6, n_layers=5, n_ctl=2000)),
           # the batch's and the served pool's larger shells (2,000 control points)
           ("shell 4x8, 2000 control points", dict(level=4, n_layers=8, n_ctl=2000)),
           ("shell 5x4, 2000 control points", dict(level=5, n_layers=4, n_ctl=2000)),
           ("shell 5x8, 2000 control points", dict(level=5, n_layers=8, n_ctl=2000)),
           ("shell 6x4, 2000 control points", dict(level=6, n_layers=4, n_ctl=2000))]
    for desc, kw in ctl:
        yield desc, W.shell_problem(time_steps=2, **kw)
    yield "2D ring disk, 2000 landmarks", W.ring_problem()
    yield "2D ring disk, 6000 landmarks", W.ring_problem(n_rings=20, n_per_ring=300)
    yield "shell 4x5, all landmarks", W.shell_problem(level=4, n_layers=5, time_steps=2)
    yield "shell 5x5, all landmarks", W.shell_problem(level=5, n_layers=5, time_steps=2)


def mode_tune(args, api, W):
    """Offline sweep of the tile table: for each mesh, precision and stored
    product, the device time of one product (kernel and slice sum, the mean
    of --tune-reps launches, best of three) at each block target, and for
    each mesh the fused operator apply at each launch shape; the table keeps
    what tune_pick chooses.  Writes the table to --tune-out (copy it beside
    the library, or name it in SA_TILE_TABLE, to use it)."""
    rows, table = [], []
    names = {"xc": "kxc", "cx": "kcx", "cc": "precond", "kk": "kmatvec"}
    op_done = set()
    f32_only, f64_only = args.tune_only == "f32", args.tune_only == "f64"
    kk_only = args.tune_only == "kk"
    for desc, p in tune_shapes(W):
        d, V = p["q0"].shape
        nc = len(p["ctlIdx"]) if p.get("ctlIdx") is not None else 0
        E = p["elm"]["elmVtxMat"].shape[1]
        if ((f32_only or f64_only) and not nc) or (kk_only and nc):
            continue
        if (E, V) not in op_done and args.tune_only == "all":
            # the fused operator apply: one launch shape per session
            op_done.add((E, V))
            times = {}
            for c in OP_SHAPES:
                os.environ["SA_OPF_CFG"] = str(c)
                s = api.Session(p)
                try:
                    s.probe("elastic", p["tauVar"][:, 0], p["tauPrm"], np.zeros((d, V)))
                    times[c] = min(s.time_product("operator", args.tune_reps) for _ in range(3))
                finally:
                    s.close()
            os.environ.pop("SA_OPF_CFG", None)
            pick = tune_pick("op", "f64", times)
            table.append(f"op f64 {E} {V} {pick}")
            rows.append({"mesh": desc, "precision": "fp64", "product": "op", "rows": E, "cols": V,
                         "us": {str(t): round(v, 2) for t, v in times.items()}, "target": pick})
            print(f"[tune] {desc} operator (E {E}, V {V}): " + " ".join(f"{t}:{v:.1f}" for t, v in times.items())
                  + f" -> {pick}", file=sys.stderr, flush=True)
        precs = ("mixed",) if f32_only else ("fp64",) if f64_only else ("mixed", "fp64") if nc else ("fp64",)
        for prec in precs:
            f32 = prec == "mixed"
            prods = ("xc", "cx", "cc") if nc else ("kk",)
            for prod in prods:
                shape = {"xc": (V, nc), "cx": (nc, V), "cc": (nc, nc), "kk": (V, V)}[prod]
                times = {}
                for t in (TUNE_TARGETS_F32 if f32 and prod != "kk" else TUNE_TARGETS):
                    os.environ["SA_KS_TARGET"] = f"{prod}={t}"
                    s = api.Session(p, precision=prec)
                    try:
                        times[t] = min(s.time_product(names[prod], args.tune_reps) for _ in range(3))
                    finally:
                        s.close()
                ty = "f32" if f32 and prod != "kk" else "f64"
                pick = tune_pick(prod, ty, times)
                table.append(f"{prod} {ty} {shape[0]} {shape[1]} {pick}")
                rows.append({"mesh": desc, "precision": prec, "product": prod, "rows": shape[0], "cols": shape[1],
                             "us": {str(t): round(v, 2) for t, v in times.items()}, "target": pick})
                print(f"[tune] {desc} {prec} {prod} {shape}: " +
                      " ".join(f"{t}:{v:.1f}" for t, v in times.items()) + f" -> {pick}", file=sys.stderr, flush=True)
    os.environ.pop("SA_KS_TARGET", None)
    if os.path.dirname(args.tune_out):
        os.makedirs(os.path.dirname(args.tune_out), exist_ok=True)
    with open(args.tune_out, "w") as f:
        f.write("# product type rows cols target-blocks; op: elements landmarks launch-shape  (python bench.py --mode tune)\n")
        f.write("\n".join(table) + "\n")
    return {"metric": "tile sweep of the stored kernel products", "value": None, "unit": "us per product",
            "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
            "config": {"workload": "tile sweep", "targets": list(TUNE_TARGETS), "reps": args.tune_reps},
            "sweep": rows, "table": table}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--repeats", type=int, default=1, help="solves of --steps steps in the timed region")
    ap.add_argument("--latency-solves", type=int, default=10,
                    help="further solves (kernel timer off) whose wall times give the p50/p99 solve latency")
    ap.add_argument("--serve-requests", type=int, default=4,
                    help="mixed solves of the workload with perturbed tau variables served by 4 workers beside "
                         "each alone (0: skip)")
    ap.add_argument("--cpu-c1-solves", type=int, default=5, help="C1 oracle solves whose median the CPU leg reports")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--vel-prec", default="default", choices=["default", "jacobi"],
                    help="preconditioner of the velocity CG (jacobi: control-point form only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true",
                    help="time without the kernel timer's events (no roofline, no operator or varifold timing)")
    ap.add_argument("--cpu-iters", type=int, default=0, help="CG iterations of the CPU sample (0: sized to the workload)")
    ap.add_argument("--mode", default="single",
                    choices=["single", "batch", "serve", "precision", "fit", "products", "tune", "operator", "meshes"],
                    help="single: one solve of --config; batch: 64 uneven meshes concurrently; serve: Poisson stream")
    ap.add_argument("--tune-reps", type=int, default=30, help="product launches per timing of --mode tune")
    ap.add_argument("--tune-out", default="tiles.tuned.txt", help="where --mode tune writes its tile table")
    ap.add_argument("--tune-only", default="all", choices=["all", "f32", "f64", "kk"],
                    help="--mode tune: every product and the operator, only the FP32 products of the mixed form, "
                         "only the FP64 control-point products, or only the all-landmark FP64 products")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--pool", type=int, default=12)
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--mem-budget", type=int, default=0)
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only run 8 device-to-device copies of MIB MiB (scale for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibrate:
        import torch
        n = args.copy_calibrate << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(8):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solve runs on one MI355X (asked for --gpus {args.gpus}); "
              "the library has no multi-GPU sharding", file=sys.stderr)
        sys.exit(2)

    from shape_atrophy_amd import api
    if args.mode != "single":
        from shape_atrophy_amd import workloads as W
        res = {"batch": mode_batch, "serve": mode_serve, "precision": mode_precision,
               "fit": mode_fit, "products": mode_products, "tune": mode_tune,
               "operator": mode_operator, "meshes": mode_meshes}[args.mode](args, api, W)
        print(json.dumps(res))
        return
    steps = args.steps if args.steps is not None else 50   # the configured solve: 50 time steps
    prob, desc = workload(args.config, steps)
    prob["params"] = dict(prob["params"], velPrec=args.vel_prec)
    d, V = prob["q0"].shape
    E = prob["elm"]["elmVtxMat"].shape[1]
    F = prob["elm"]["vfdVtxMat"].shape[1]
    print(f"[bench] {args.config}: V={V} E={E} F={F} steps={steps} precision={args.precision}", file=sys.stderr,
          flush=True)

    sess = api.Session(prob, precision=args.precision)
    tv, tp = prob["tauVar"][:, 0], prob["tauPrm"]
    if args.warmup > 0:
        sess.set_time_num(min(args.warmup, steps) + 1)
        w = sess.run(tv, tp)
        print(f"[bench] warmup: {w['steps']} steps in {w['seconds']:.3f}s status={w['status']}", file=sys.stderr,
              flush=True)
    sess.set_time_num(steps + 1)
    sess.set_profiling(not args.no_profile)
    runs = []
    t0 = time.perf_counter()
    for r in range(args.repeats):
        out = sess.run(tv, tp)
        runs.append(out)
        print(f"[bench] solve {r}: {out['steps']} steps {out['seconds']:.3f}s obj={out['obj']:.12e} "
              f"pcg={out['pcg_iters']} rd={out['rd_iters']} status={out['status']}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    total_steps = sum(o["steps"] for o in runs)
    value = total_steps / wall
    # solve latency: further solves without the kernel timer, each timed alone
    sess.set_profiling(False)
    lat, lat_runs = [], []
    for r in range(args.latency_solves):
        out = sess.run(tv, tp)
        lat.append(out["seconds"])
        lat_runs.append(out)
    if lat:
        print(f"[bench] latency: {len(lat)} solves, p50 {percentile(lat, 50):.3f}s p99 {percentile(lat, 99):.3f}s",
              file=sys.stderr, flush=True)
    else:
        lat = [o["seconds"] for o in runs]

    # roofline of the dominant kernel: the kernel-matrix product
    kmv_ms = sum(o["kmatvec_ms"] for o in runs)
    kmv_n = sum(o["kmatvec_count"] for o in runs)
    op_ms = sum(o["operator_ms"] for o in runs)
    op_n = sum(o["operator_count"] for o in runs)
    order = int(prob["params"]["knlOrder"])
    nc = len(prob.get("ctlIdx", [])) or 0
    roof = None
    if kmv_n:
        avg = kmv_ms / kmv_n
        if nc:
            # each launch (K_xc u or K_cx y) reads one stored V x C kernel
            # matrix (FP64 plain form / FP32 mixed form) once, plus its
            # D-column input and output vectors
            esz = 8 if args.precision == "fp64" else 4
            B = esz * V * nc + 8 * d * (V + nc)
            ach = B / (avg * 1e-3) / 1e9
            roof = {"bound": "hbm", "kernel": f"kstore_mv_k<{d}, {'double' if esz == 8 else 'float'}>",
                    "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": TRAFFIC.get((args.config, args.precision)),
                    "traffic_source": TRAFFIC_SOURCE,
                    "avg_ms": round(avg, 4), "launches": kmv_n,
                    "work": f"{esz} B x {V} x {nc} stored kernel entries + 8 B x {d} x ({V} + {nc}) vectors per launch"}
        elif 8 * ((V + 127) // 128 * 128) * V <= (48 << 30):
            # all landmarks, K stored per step (FP64, panel-major): each launch
            # reads K once (the CG's slice sum is in the direction kernel)
            B = 8 * V * V + 8 * d * 3 * V
            ach = B / (avg * 1e-3) / 1e9
            roof = {"bound": "hbm", "kernel": f"kstore_mv_k<{d}, double> (stored K)",
                    "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": TRAFFIC.get((args.config, args.precision)),
                    "traffic_source": TRAFFIC_SOURCE,
                    "avg_ms": round(avg, 4), "launches": kmv_n,
                    "work": f"8 B x {V} x {V} stored kernel entries + 8 B x {d} x 3 x {V} vectors per product"}
        else:
            pairs = V * V
            flops = FLOP_PER_PAIR[order] * pairs
            ach = flops / (avg * 1e-3) / 1e12
            roof = {"bound": "mfma", "engine": "fp64 VALU (fp64 vector peak = fp64 matrix peak on MI355X)",
                    "kernel": "kmatvec_k", "achieved": round(ach, 3),
                    "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4),
                    "traffic": None, "avg_ms": round(avg, 4), "launches": kmv_n,
                    "work": f"{FLOP_PER_PAIR[order]} flop/pair x {pairs} pairs"}
    op = None
    if op_n:
        avg = op_ms / op_n
        B = (64 * E + 72 * V) if d == 3 else (40 * E + 48 * V)
        kern = "elastic_elem_k + gather_op_k" if os.environ.get("SA_OP_FUSED", "1")[:1] == "0" else "op_fused_k"
        op = {"bound": "hbm", "kernel": kern, "achieved": round(B / (avg * 1e-3) / 1e9, 1),
              "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(B / (avg * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
              "avg_ms": round(avg, 4), "algorithmic_bytes": B}

    # the objective's varifold row sums (one launch per solve).  In the mixed
    # form the pair contractions run on FP64 MFMA: three
    # v_mfma_f64_16x16x4f64 of 2,048 flop per 16 x 16 tile of pairs, 24 MFMA
    # flop per pair; the centre and direction kernels (an exponential per
    # pair) stay on the FP64 VALU, which bounds the launch
    vfd = None
    vfd_ms = sum(o["varifold_ms"] for o in runs)
    vfd_pairs = sum(o["varifold_pairs"] for o in runs)
    if vfd_ms > 0 and vfd_pairs:
        on_mfma = all(o["varifold_mfma"] for o in runs)
        t = vfd_ms * 1e-3
        vfd = {"kernel": f"vfd_rows_mfma_k<{d}>" if on_mfma else f"vfd_rows_k<{d}, false>", "bound": "fp64 VALU",
               "pairs": vfd_pairs, "ms": round(vfd_ms, 4), "gpairs_per_s": round(vfd_pairs / t / 1e9, 2)}
        if on_mfma:
            ach = 24 * vfd_pairs / t / 1e12
            vfd.update({"mfma_achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "mfma_frac": round(ach / PEAK_FP64_TFLOPS, 4),
                        "work": "24 MFMA flop per pair (3 x 16x16x4 FP64 per 256 pairs)"})

    # served pass: mixed solves of this workload with perturbed tau variables
    # through the request server (4 workers, one session per request in
    # flight), each also run alone on the bench session
    served = None
    if args.serve_requests > 0:
        served = serve_pass(api, prob, args, steps, sess)

    cpu = cpu_leg(prob, args, op) if not args.no_cpu else None

    # accuracy of the timed form against the plain FP64 GPU form on the same
    # solve (the FP64 form is held to the FP64 oracle on this workload's first
    # two steps by tests/test_gpu_configs.py)
    ref64 = None
    if args.precision != "fp64":
        s64 = api.Session(prob, precision="fp64")
        s64.set_time_num(steps + 1)
        t1 = time.perf_counter()
        o64 = s64.run(tv, tp)
        w64 = time.perf_counter() - t1
        s64.close()
        ref64 = {"value": round(o64["steps"] / w64, 4), "objective": o64["obj"], "status": o64["status"],
                 "obj_rel_err": abs(runs[-1]["obj"] - o64["obj"]) / abs(o64["obj"]),
                 "pcg_iters_per_step": round(o64["pcg_iters"] / max(o64["steps"], 1), 1)}
        print(f"[bench] fp64 form: {o64['steps']} steps {w64:.3f}s obj={o64['obj']:.12e}", file=sys.stderr,
              flush=True)
    sess.close()

    # the metric holds only at <= 1e-6 relative error of the objective against
    # FP64, with every solve completed: otherwise value is 0 (value_unchecked
    # keeps the throughput)
    all_runs = runs + lat_runs
    failed = sum(1 for o in all_runs if o["status"] != 0 or o["steps"] != steps)
    acc_ok = failed == 0 and (ref64 is None or (ref64["status"] == 0 and ref64["obj_rel_err"] <= 1e-6))
    res = {
        "metric": METRIC,
        "value": round(value, 4) if acc_ok else 0.0,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": total_steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / max(total_steps, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": desc, "config": args.config, "landmarks": V, "tets": E, "varifold_facets": F,
                   "control_points": nc, "time_steps_per_solve": steps, "solves": args.repeats,
                   "precision": args.precision,
                   "velocity_preconditioner": preconditioner_name(prob, args.precision),
                   "elasticity": prob["params"]["elaType"]},
        "accuracy_ok": acc_ok,
        "value_unchecked": round(value, 4),
        "failed_solves": failed,
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99), "solves": len(lat),
                      "steps_per_solve": steps},
        "steps_per_s_without_kernel_timer": (round(sum(o["steps"] for o in lat_runs) / sum(lat), 4)
                                             if lat_runs else None),
        "pcg_iters_per_step": round(sum(o["pcg_iters"] for o in runs) / max(total_steps, 1), 1),
        "refine_sweeps_per_step": round(sum(o["refine"] for o in runs) / max(total_steps, 1), 2),
        "objective": runs[-1]["obj"],
        "roofline": roof,
        "operator_roofline": op,
        "varifold": vfd,
        "cpu_baseline": cpu,
        "fp64_form": ref64,
        "served": served,
    }
    print(json.dumps(res))


def preconditioner_name(prob, precision):
    if prob["params"].get("velPrec", "default") == "jacobi":
        return "control-point CG preconditioned by the diagonal of the control-point system (Jacobi)" + (
            "; FP32 inner solves in FP64 refinement" if precision == "mixed" else "")
    if prob.get("ctlIdx") is not None and len(prob["ctlIdx"]):
        return ("control-point CG preconditioned by X^T X = (K_cc + eps I)^-1 (inverted Cholesky factor) "
                "plus a spectral term of Ritz vectors" + ("; FP32 inner solves in FP64 refinement"
                                                          if precision == "mixed" else ""))
    return "K-preconditioned CG on ldmWgt K^-1 + E + btmWgt P (ipcg.cu)"


def serve_pass(api, prob, args, steps, sess):
    """Mixed solves with perturbed tau variables: each alone on the bench
    session, then all served at once by 4 workers; served steps/s, latency,
    failures and the largest alone-vs-mix difference of the objective."""
    rng = np.random.default_rng(77)
    tp = prob["tauPrm"]
    tvs = [prob["tauVar"][:, 0] * (1.0 + 0.02 * rng.standard_normal(prob["tauVar"].shape[0]))
           for _ in range(args.serve_requests)]
    alone = [sess.run(tv, tp) for tv in tvs]
    pr = dict(prob, timeNum=steps + 1)
    srv = api.Server(workers=min(4, args.serve_requests))
    mid = srv.add_mesh(pr, args.precision)
    t0 = time.perf_counter()
    rids = [srv.submit(mid, tv, tp) for tv in tvs]
    got = [srv.wait(r) for r in rids]
    wall = time.perf_counter() - t0
    info = srv.info()
    srv.close()
    ok = [k for k, g in enumerate(got) if g["rc"] == 0 and g["status"] == 0 and alone[k]["status"] == 0]
    diff = max((abs(got[k]["obj"] - alone[k]["obj"]) / abs(alone[k]["obj"]) for k in ok), default=None)
    lat = [g["t_end"] - g["t_submit"] for g in got]
    out = {"requests": len(tvs), "workers": min(4, args.serve_requests), "precision": args.precision,
           "steps_per_s": round(sum(g["steps"] for g in got) / wall, 4),
           "serial_alone_steps_per_s": round(sum(a["steps"] for a in alone) / sum(a["seconds"] for a in alone), 4),
           "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
           "failed_requests": len(tvs) - len(ok), "max_rel_diff_alone_vs_mix": diff,
           "factorizations": sum(g["factorizations"] for g in got), "device_bytes_reserved": info["mem_used"]}
    print(f"[bench] served: {out}", file=sys.stderr, flush=True)
    return out


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": model}


def cpu_leg(prob, args, op):
    """The FP64 oracle on the host: one complete step of the workload (the
    first: geometry, kernel matrices and their factorisation, the velocity CG
    from a zero start to 1e-10, the landmark update and the
    reaction-diffusion solve) on all the cores this process may use; the C1
    full solve (20 steps and the objective), median of --cpu-c1-solves; the
    operator apply alone, median of 100 on all cores and of 20 on one."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from shape_atrophy_amd import workloads as W
    O.build()
    # every core this process may run on, within the OpenMP share the host sets
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    t0 = time.perf_counter()
    o = O.objfcn(dict(prob, timeNum=2), with_target=False, threads=threads)
    t_step = time.perf_counter() - t0
    print(f"[bench] cpu: one step {t_step:.2f}s on {threads} threads (status {o['status']}, "
          f"{int(o['pcg_iters'][0])} CG iterations)", file=sys.stderr, flush=True)
    cpu = {"value": round(1.0 / t_step, 6), "unit": "steps/s", "cores": threads, "kind": "port",
           "sample": (f"FP64 oracle: one complete first step of the same workload ({int(o['pcg_iters'][0])} CG "
                      f"iterations to 1e-10 from a zero start, reaction-diffusion included), {t_step:.2f}s "
                      f"on {threads} threads"),
           "host": cpu_info(), "step_s": round(t_step, 3), "status": o["status"]}
    c1 = W.ring_problem()
    ts = []
    for _ in range(args.cpu_c1_solves):
        t1 = time.perf_counter()
        r1 = O.objfcn(c1, threads=threads)
        ts.append(time.perf_counter() - t1)
    if ts:
        med = float(np.median(ts))
        cpu["c1_full_solve"] = {"median_s": round(med, 3), "solves": len(ts), "steps_per_s": round(20 / med, 4),
                                "cores": threads, "status": r1["status"]}
        print(f"[bench] cpu: C1 solve median {med:.2f}s of {len(ts)}", file=sys.stderr, flush=True)
    one = O.time_elastic_apply(prob, reps=20, threads=1)
    allc = O.time_elastic_apply(prob, reps=100, threads=threads)
    cpu["operator_apply"] = {"median_ms": round(1e3 * allc, 3), "cores": threads, "reps": 100,
                             "single_core_median_ms": round(1e3 * one, 3), "single_core_reps": 20,
                             "gpu_avg_ms": op["avg_ms"] if op else None}
    return cpu


if __name__ == "__main__":
    main()
