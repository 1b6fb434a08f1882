"""Benchmark: solve time to relative KKT 1e-6 on one MI355X.

Default workload: the single large LP -- m = 1,000,000 rows (half =, half <=),
n = 2,000,000 columns, ~20M nonzeros with power-law row lengths (Zipf 1.5, capped at
5,000, rescaled to 20M), N(0,1) values, 0 <= x <= 10 on 30% of the columns; plain
FP64 PDLP with Ruiz + Chambolle-Pock scaling and adaptive restarts, reltol 1e-6.

One step = one chunk of the PDLP loop: check_every (100) PDHG iterations and their
termination/restart check, on the LP already resident in HBM.  The warmup and timed
steps are the first chunks of ONE full solve, which then runs on to termination:
`value` is LPs solved per second = 1 / that solve's time (scaling, step size and the
loop; the LP resident in HBM), `ms_per_step` the mean time of the timed chunks.  A
wall-clock guard (--budget seconds from process start) never starts a slice of the
loop it cannot finish; a solve stopped by it is reported as such, with value null.

Also reported: the roofline of the PDHG iteration (the two fused step kernels, timed
live with HIP events on the solver's stream over the timed steps), per-kernel times,
the step size of the reference's power-method rule beside the converged one, and the
CPU oracle (FP64, OpenMP) on a bounded sample of the same LP on this host, projected to
the GPU solve's pass count.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|plain|batch|serve|mcf|gather]
"""
import argparse
import json
from dataclasses import dataclass
import os
import sys
import time

T_PROC0 = time.perf_counter()   # the wall-clock guard counts from here
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "coolpdlp.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
FORMATS = ["auto", "csr", "ell", "xslice", "xslice_all"]   # index = CPDLP_FMT_*


def measured_traffic(workload: str, fmt_a: str, fmt_at: str):
    """HBM bytes per PDHG iteration measured with rocprofv3 (FETCH_SIZE and WRITE_SIZE
    passes, each scaled by the same counter's reading for a plain copy of known size),
    as recorded in profiles/pdhg_steps_<workload>_hbm_traffic.md for these formats;
    None when no such measurement exists."""
    path = os.path.join(ROOT, "profiles", f"pdhg_steps_{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    rec = {}
    for line in open(path):
        if line.startswith("record:"):
            for kv in line[len("record:"):].split():
                k, _, v = kv.partition("=")
                rec[k] = v
    if rec.get("workload") != workload or rec.get("A") != fmt_a or rec.get("At") != fmt_at:
        return None, None
    return int(rec["traffic_bytes_per_iteration"]), os.path.relpath(path, ROOT)


def gather_ceiling():
    """The L2-resident gather rate of the CSR tile kernel, as recorded in
    profiles/gather_ceiling_csr_tile_kernel.md (bench.py --workload gather); None when
    that record is absent."""
    path = os.path.join(ROOT, "profiles", "gather_ceiling_csr_tile_kernel.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("record:"):
            rec = dict(kv.partition("=")[::2] for kv in line[len("record:"):].split())
            return float(rec["gathers_per_s"]), os.path.relpath(path, ROOT)
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def iteration_bytes(m, n, nnz, V=8, I=4):
    """Algorithmic HBM bytes of one PDHG iteration (both step kernels):
    primal kernel: A' (rowptr + col/val), gather y once, read x,c,lv,uv, write x+,xbar,
                   read+write x_avg   -> 4(n+1) + nnz(I+V) + mV + 8nV
    dual kernel:   A (rowptr + col/val), gather xbar once, read y,lc,uc, write y+,
                   read+write y_avg   -> 4(m+1) + nnz(I+V) + nV + 6mV"""
    primal = I * (n + 1) + nnz * (I + V) + m * V + 8 * n * V
    dual = I * (m + 1) + nnz * (I + V) + n * V + 6 * m * V
    return primal, dual


def make_workload(name):
    from coolpdlp_hip import generators as G
    if name == "large":
        g = G.power_law_lp()
        desc = dict(workload="single_large_lp_power_law", m=g.m, n=g.n, nnz=g.nnz,
                    row_length="zipf1.5_cap5000_rescaled", box_frac=0.3, reltol=1e-6,
                    algorithm="PDLP", scaling="ruiz10+chambolle_pock", precision="fp64")
        return g, 1e-6, desc
    if name == "plain":
        g = G.small_dense_lp()
        desc = dict(workload="plain_lp_m2000_n5000", m=g.m, n=g.n, nnz=g.nnz, reltol=1e-4,
                    algorithm="PDLP", precision="fp64")
        return g, 1e-4, desc
    raise SystemExit(f"unknown workload {name}")


def polish(args) -> dict:
    """The mixed precision's hand-over tolerance, when --polish-reltol sets it."""
    return {} if args.polish_reltol is None else {"polish_reltol": args.polish_reltol}


def bench_batch(args):
    """4,096 uneven LPs in one batch (one workgroup per LP, mixed precision, reltol 1e-6
    per LP) against the same LPs solved one at a time on the plain GPU path (all of them
    by default; --plain-sample N times a stratified sample and scales it, marked
    projected)."""
    import numpy as np
    import torch
    import coolpdlp_hip as cp
    from coolpdlp_hip import generators as G
    t = time.perf_counter()
    specs = G.batch_specs(args.batch_count)
    lps = [G.column_lp(m, n, (6, 12), sd, name=f"batch_{i}") for i, (m, n, sd) in enumerate(specs)]
    nnz = sum(g.nnz for g in lps)
    log(f"[bench] generated {len(lps)} LPs, nnz={nnz} in {time.perf_counter() - t:.1f}s")
    reltol = 1e-6
    algo = cp.PDLP("mixed", termination_reltol=reltol, time_limit=args.time_limit,
                   detect_infeasibility=int(args.detect_infeasibility), **polish(args))
    arena = int(sum(cp.batch.footprint(g.m, g.n, g.nnz) for g in lps) * 1.05) + (64 << 20)
    B = cp.BatchSolver(algo, hbm_bytes=arena, tail=args.batch_tail, tail_min_passes=args.batch_tail_passes,
                       tail_workers=args.batch_tail_workers)
    t = time.perf_counter()
    ids = [B.add(g) for g in lps]
    log(f"[bench] admitted to HBM ({arena / 2**30:.2f} GiB arena) in {time.perf_counter() - t:.1f}s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done_t = {}
    rounds = 0
    while True:
        running = B.run(args.checks_per_round)
        rounds += 1
        now = time.perf_counter() - t0
        for i in B.poll():
            done_t[i] = now
        if rounds % 10 == 0:
            log(f"[bench] round {rounds}: {len(done_t)}/{len(ids)} done at {now:.1f}s")
        if running == 0:
            break
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res = [B.get(i) for i in ids]
    status = [r[1].termination_status.name for r in res]
    lat = np.array([done_t[i] for i in ids])
    rel_obj = np.array([abs(float(sol.x @ g.c) - g.objective) / (1 + abs(g.objective))
                        for g, (sol, st) in zip(lps, res)])
    passes = np.array([st.kkt_passes for _, st in res])
    powm = np.array([st.powm_iters for _, st in res])
    # plain one-at-a-time GPU path: every LP (the default), or a stratified sample
    # (sorted by size, evenly spaced) whose mean is scaled to all LPs
    order = np.argsort([g.nnz for g in lps])
    n_plain = len(lps) if args.plain_sample <= 0 else min(args.plain_sample, len(lps))
    sample = order[np.linspace(0, len(lps) - 1, n_plain).astype(int)] if n_plain < len(lps) else order
    plain_t, plain_status, plain_err = [], [], []
    t_plain0 = time.perf_counter()
    for k in sample:
        g = lps[k]
        milp = cp.MILP.from_generated(g)
        eng = cp.Engine(milp, cp.PDLP("mixed", termination_reltol=reltol, time_limit=args.time_limit,
                                      **polish(args)))
        torch.cuda.synchronize()
        t = time.perf_counter()
        sol_p, st = eng.solve()
        plain_t.append(time.perf_counter() - t)
        eng.close()
        plain_status.append(st.termination_status.name)
        plain_err.append(abs(float(sol_p.x @ g.c) - g.objective) / (1 + abs(g.objective)))
        if len(plain_t) % 512 == 0:
            log(f"[bench] plain path: {len(plain_t)}/{len(sample)} LPs in {time.perf_counter() - t_plain0:.1f}s")
    measured = len(sample) == len(lps)
    plain_total = float(np.sum(plain_t)) if measured else float(np.mean(plain_t)) * len(lps)
    value = len(lps) / elapsed
    return {
        "metric": "solve time to rel-KKT 1e-6; batched LPs/s + p50/p99 vs plain PDLP, 1 GPU",
        "value": value, "unit": "LPs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * elapsed, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 bulk / f64 checks+polish",
        "data": "synthetic (seeded generator with known optima)",
        "config": {"workload": "batch_4096_uneven_lps", "count": len(lps), "nnz_total": nnz,
                   "m_range": [50, 5000], "reltol": reltol, "precision": "mixed",
                   "polish_reltol": algo.params.polish_reltol,
                   "detect_infeasibility": bool(args.detect_infeasibility),
                   "checks_per_round": args.checks_per_round,
                   "tail_continuation": ({"below": args.batch_tail, "min_passes": args.batch_tail_passes,
                                          "workers": args.batch_tail_workers}
                                         if args.batch_tail > 0 else None),
                   "timed_region": "batch (LPs admitted beforehand)"},
        "continued_on_plain_engines": int(sum(B.continued(i) for i in ids)),
        "batch_seconds": elapsed, "rounds": rounds,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "optimal": int(sum(s == "OPTIMAL" for s in status)), "count": len(lps),
        "not_optimal": [{"index": i, "status": status[i], "kkt_passes": int(passes[i]),
                         "m": lps[i].m, "n": lps[i].n, "nnz": lps[i].nnz}
                        for i in range(len(lps)) if status[i] != "OPTIMAL"],
        "max_objective_rel_err": float(rel_obj.max()), "median_kkt_passes": float(np.median(passes)),
        "mean_kkt_passes": float(np.mean(passes)), "fp32_pass_share": float(np.sum([st.fp32_passes for _, st in res]) / max(1, passes.sum())),
        "powm_iters_median": float(np.median(powm)), "powm_iters_max": int(powm.max()),
        "plain_gpu": {"lps_solved": int(len(sample)), "measured_over_all_lps": measured,
                      "mean_seconds_per_lp": float(np.mean(plain_t)),
                      ("total_seconds" if measured else "projected_total_seconds"): plain_total,
                      "lps_per_s": len(lps) / plain_total,
                      "optimal": int(sum(x == "OPTIMAL" for x in plain_status)),
                      "max_objective_rel_err": float(max(plain_err)),
                      "wall_seconds_incl_engine_creation": time.perf_counter() - t_plain0},
        "speedup_vs_plain": plain_total / elapsed,
        # the tail: the last instances to finish (one workgroup each) set the batch's time
        "last_finished": [{"m": lps[i].m, "n": lps[i].n, "nnz": lps[i].nnz, "kkt_passes": int(passes[i]),
                           "fp32_passes": int(res[i][1].fp32_passes), "finished_s": float(lat[i])}
                          for i in np.argsort(lat)[-8:]],
    }


def bench_mcf(args):
    """Mixed-precision min-cost flow: 500,000 nodes, 4,000,000 arcs; FP32 matrix,
    iterates and products, FP64 checks and FP64 polish, relative KKT 1e-6; objective
    and feasibility error against the oracle's FP64 solve of the same LP at relative
    KKT 1e-8 (CPU); the GPU's own FP64 solves at 1e-6 and 1e-8 are reported beside it."""
    import numpy as np
    import torch
    import coolpdlp_hip as cp
    from coolpdlp_hip import generators as G
    t = time.perf_counter()
    g = G.min_cost_flow_lp(nodes=args.mcf_nodes)
    log(f"[bench] min-cost flow {g.m} nodes x {g.n} arcs in {time.perf_counter() - t:.1f}s "
        f"(supply scale {g.meta['supply_scale']:.3f})")
    milp = cp.MILP.from_generated(g)
    A = g.csc()

    def feas(x):
        r = A @ x - g.lc
        viol = max(float(np.max(np.abs(r))), float(np.max(np.maximum(g.lv - x, 0))),
                   float(np.max(np.maximum(x - g.uv, 0))))
        return viol / (1 + float(np.max(np.abs(g.lc))))

    eng = cp.Engine(milp, cp.PDLP("mixed", termination_reltol=1e-6, time_limit=args.time_limit,
                                  powm_tol=args.powm_tol, show_progress=1, **polish(args)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol, st = eng.solve()
    el = time.perf_counter() - t0
    eng.close()
    log(f"[bench] mixed: {st.termination_status.name} passes={st.kkt_passes} (fp32 {st.fp32_passes}) "
        f"{el:.1f}s")
    f64_eng = cp.Engine(milp, cp.PDLP("float64", termination_reltol=1e-6, time_limit=args.time_limit,
                                      powm_tol=args.powm_tol, show_progress=1))
    t1 = time.perf_counter()
    _, f64st = f64_eng.solve()
    el_f64 = time.perf_counter() - t1
    f64_eng.close()
    log(f"[bench] fp64 at the same tolerance: {f64st.termination_status.name} passes={f64st.kkt_passes} "
        f"{el_f64:.1f}s")
    gpu_ref_eng = cp.Engine(milp, cp.PDLP("float64", termination_reltol=1e-8, time_limit=args.time_limit,
                                          powm_tol=args.powm_tol, show_progress=1))
    t1 = time.perf_counter()
    gref, grst = gpu_ref_eng.solve()
    el_gref = time.perf_counter() - t1
    gpu_ref_eng.close()
    log(f"[bench] GPU fp64 at 1e-8: {grst.termination_status.name} passes={grst.kkt_passes} {el_gref:.1f}s")
    # the stated reference: the oracle's FP64 solve at relative KKT 1e-8 on this host
    import oracle as O
    cores = len(os.sched_getaffinity(0))
    O.set_threads(max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))))
    t1 = time.perf_counter()
    ref = O.solve(g, O.default_params(termination_reltol=1e-8, powm_tol=args.powm_tol,
                                      max_kkt_passes=10**7, time_limit=args.mcf_oracle_limit), hist_cap=16)
    el_ref = time.perf_counter() - t1
    log(f"[bench] oracle fp64 at 1e-8: {ref.status} passes={ref.stats.kkt_passes} {el_ref:.1f}s")
    obj, obj_ref = float(sol.x @ g.c), float(ref.x @ g.c)
    return {
        "metric": "solve time to rel-KKT 1e-6; batched LPs/s + p50/p99 vs plain PDLP, 1 GPU",
        "value": 1.0 / el, "unit": "LPs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 bulk / f64 checks+polish", "data": "synthetic (seeded min-cost flow)",
        "config": {"workload": "min_cost_flow_mixed", "nodes": g.m, "arcs": g.n, "nnz": g.nnz,
                   "reltol": 1e-6, "reference_reltol": 1e-8, "powm_tol": args.powm_tol,
                   "polish_reltol": args.polish_reltol if args.polish_reltol is not None else 1e-5},
        "status": st.termination_status.name, "kkt_passes": st.kkt_passes,
        "fp32_passes": st.fp32_passes, "relative_kkt": st.err.relative(),
        "objective": obj, "objective_fp64_ref": obj_ref,
        "objective_rel_err": abs(obj - obj_ref) / (1 + abs(obj_ref)),
        "feasibility_err": feas(sol.x), "feasibility_err_fp64_ref": feas(ref.x),
        "fp64_ref": "oracle (CPU, FP64, OpenMP) at relative KKT 1e-8",
        "fp64_ref_status": ref.status, "fp64_ref_seconds": el_ref,
        "fp64_ref_passes": ref.stats.kkt_passes, "fp64_ref_cores": O_threads(),
        "gpu_fp64_1e8": {"status": grst.termination_status.name, "seconds": el_gref,
                         "passes": grst.kkt_passes, "objective": float(gref.x @ g.c),
                         "objective_rel_err_vs_oracle": abs(float(gref.x @ g.c) - obj_ref) / (1 + abs(obj_ref))},
        "fp64_same_tol_seconds": el_f64, "fp64_same_tol_passes": f64st.kkt_passes,
        "fp64_same_tol_status": f64st.termination_status.name,
        "speedup_vs_fp64_same_tol": el_f64 / el,
    }


def bench_gather(args):
    """The gather-rate ceiling of the CSR tile kernel the steps run: the library's own
    product (cpdlp_matrix_mul, CSR format, thread-per-row sums) of a 2,000,000-row matrix
    with 10 uniform random distinct columns per row (20M FP64 gathers), against a
    gathered vector of 16 entries (one 128-byte line: every gather hits the L1) and of
    1, 4, 8 and 16 MB.  Reports the time per product (HIP events on the launching
    stream), gathers per second and algorithmic GB/s."""
    import numpy as np
    import scipy.sparse as sp
    import torch
    import coolpdlp_hip as cp
    rows, per_row = args.gather_rows, 10
    rng = np.random.default_rng(5)
    points = []
    for n in (16, 131_072, 524_288, 1_048_576, 2_097_152):
        t = time.perf_counter()
        # distinct columns per row: a random offset plus sorted distinct steps, modulo n
        steps = np.sort(rng.choice(max(n // per_row, 1), size=(rows, per_row)), axis=1) if n > 16 else None
        if n > 16:
            cols = (rng.integers(0, n, size=(rows, 1)) + steps * per_row + np.arange(per_row)) % n
        else:
            cols = np.argsort(rng.random((rows, 16)), axis=1)[:, :per_row]
        cols = np.sort(cols, axis=1).astype(np.int32)
        indptr = np.arange(0, rows * per_row + 1, per_row, dtype=np.int64)
        A = sp.csr_matrix((rng.standard_normal(rows * per_row), cols.ravel(), indptr), shape=(rows, n)).tocsc()
        M = cp.GPUSparseMatrixCSR(A)
        b = torch.rand(n, dtype=torch.float64, device="cuda:0")
        c = torch.empty(rows, dtype=torch.float64, device="cuda:0")
        for _ in range(3):
            M.mul(c, b)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = args.gather_reps
        e0.record()
        for _ in range(reps):
            M.mul(c, b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nnz = rows * per_row
        alg = 12 * nnz + 4 * (rows + 1) + 8 * rows + 8 * n
        pt = {"vector_entries": n, "vector_mb": 8 * n / 2**20, "ms": ms, "gathers_per_s": nnz / (ms * 1e-3),
              "algorithmic_GBs": alg / (ms * 1e-3) / 1e9}
        log(f"[bench] gather n={n}: {ms * 1e3:.1f} us, {pt['gathers_per_s'] / 1e9:.1f} G gathers/s "
            f"(setup {time.perf_counter() - t:.1f}s)")
        points.append(pt)
        M.close()
        del b, c
    return {"metric": "gather-rate ceiling of the CSR tile kernel (not the headline metric)",
            "value": points[-1]["gathers_per_s"], "unit": "gathers/s", "n_gpus": 1,
            "steps": args.gather_reps, "warmup": 3, "ms_per_step": points[-1]["ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "gather_ceiling_csr_tile", "rows": rows, "nnz_per_row": per_row},
            "points": points}


def O_threads() -> int:
    """Threads the oracle runs on in this process."""
    cores = len(os.sched_getaffinity(0))
    return max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))


def _class_stats(recs, picks, classes=("small", "medium", "large")):
    """Per request class: submitted, completed, OPTIMAL, p50/p99 latency of the OPTIMAL ones."""
    import numpy as np
    out = {}
    for c in classes:
        mine = [r for r in recs if picks[r.rid][0] == c]
        lat = np.array([r.t_done - r.t_arrival for r in mine if r.status == "OPTIMAL"])
        out[c] = {"submitted": int(sum(p[0] == c for p in picks)), "completed": len(mine),
                  "optimal": int(len(lat)),
                  "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
                  "p99_s": float(np.percentile(lat, 99)) if len(lat) else None}
    return out


def _served(recs, t0_rel=0.0):
    """OPTIMAL results per second of the stream (until the last OPTIMAL completion), and
    the p50/p99 latency of those results."""
    import numpy as np
    opt = [r for r in recs if r.status == "OPTIMAL"]
    span = max((r.t_done for r in opt), default=0.0) - t0_rel
    lat = np.array([r.t_done - r.t_arrival for r in opt])
    return {"completed": len(recs), "optimal": len(opt),
            "served_lps_per_s": len(opt) / span if span > 0 else 0.0,
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None}


def bench_serve(args):
    """Serving loop: Poisson arrivals at each QPS of the sweep for `serve_duration`
    seconds (then up to `serve_drain` seconds more to finish), request mix 70% small
    (m log-uniform in [50, 2k]), 25% medium ([2k, 50k]), 5% large ([50k, 500k]), n = 2m,
    8 nonzeros per column.  Requests are drawn from a seeded pool of distinct LPs per
    class.  Latency = finish - arrival, to relative KKT 1e-6; only OPTIMAL results count
    as served.  Every request, on every route and in the baseline, runs the power method
    to powm_tol (the reference's eps*n^3 rule stops after one Rayleigh quotient for
    n >~ 3e5 and the large requests then diverge).  The baseline is a first-come-first-
    served server that solves one request at a time on a plain engine, run on the same
    arrivals (--serve-fcfs run), or modelled from each pool LP's measured solve time
    (--serve-fcfs replay)."""
    import numpy as np
    import coolpdlp_hip as cp
    from coolpdlp_hip import generators as G
    from coolpdlp_hip.serving import NativeServer, Server, fcfs_latencies, poisson_arrivals
    rng = np.random.default_rng(17)
    reltol = 1e-6
    bounds = {"small": (50, 2_000), "medium": (2_000, 50_000), "large": (50_000, 500_000)}
    sizes = {"small": args.pool_small, "medium": args.pool_medium, "large": args.pool_large}
    pool = {}
    t = time.perf_counter()
    for cls, (lo, hi) in bounds.items():
        ms = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=sizes[cls]))).astype(int)
        pool[cls] = [G.request_lp(int(m), 2 * int(m), 1000 + 97 * k + len(pool)) for k, m in enumerate(ms)]
    log(f"[bench] request pool {({k: len(v) for k, v in pool.items()})} in {time.perf_counter() - t:.1f}s")
    kw = dict(termination_reltol=reltol, time_limit=args.time_limit, powm_tol=args.powm_tol)

    def plain_solve(g):
        """The baseline server's work for one request: a plain engine, alone."""
        eng = cp.Engine(cp.MILP.from_generated(g), cp.PDLP("mixed", None, "csr", **kw))
        try:
            return eng.solve()
        finally:
            eng.close()

    service = {}
    if args.serve_fcfs == "replay":
        for cls, lps in pool.items():
            for k, g in enumerate(lps):
                t = time.perf_counter()
                plain_solve(g)
                service[(cls, k)] = time.perf_counter() - t
    out_points = []
    for qps in args.qps:
        arr = poisson_arrivals(qps, args.serve_duration, seed=int(qps))
        u = rng.random(len(arr))
        cls_of = np.where(u < 0.70, "small", np.where(u < 0.95, "medium", "large"))
        picks = [(c, int(rng.integers(len(pool[c])))) for c in cls_of]
        lanes = ((1_000_000, args.serve_medium_workers), (float("inf"), 1))
        # ---- the server (continuous batching)
        if args.serve_native:   # the library's own server (cpdlp_server_*)
            srv = NativeServer(cp.PDLP("mixed", batch_streams=1, **kw),   # per batch route
                               batch_bytes=args.serve_batch_gib << 30, checks_per_round=args.serve_checks,
                               lanes=lanes, batch_limit=args.serve_batch_limit,
                               batch2_limit=args.serve_batch2_limit, batch2_bytes=int(args.serve_batch2_gib * 2**30))
            srv.done, srv.rounds, srv.round_seconds, meta = [], 0, [], {}
            srv.start()           # the library pumps on its own thread
        else:
            srv = Server(reltol=reltol, batch_bytes=args.serve_batch_gib << 30, powm_tol=args.powm_tol,
                         checks_per_round=args.serve_checks, time_limit=args.time_limit,
                         lanes=lanes, overflow_queue=args.serve_overflow, batch_limit=args.serve_batch_limit)
        max_obj_err = 0.0
        t0 = time.perf_counter()
        k = 0
        deadline = t0 + args.serve_duration + args.serve_drain
        last_log = t0
        while time.perf_counter() < deadline:
            now = time.perf_counter() - t0
            while k < len(arr) and arr[k] <= now:
                g = pool[picks[k][0]][picks[k][1]]
                if args.serve_native:
                    meta[srv.submit(g)] = (k, arr[k], time.perf_counter() - t0, g)
                else:
                    srv.submit(g, t0 + arr[k])
                k += 1
            if args.serve_native:
                if k >= len(arr) and not meta:
                    break
                time.sleep(0.001)
                for i in srv.poll():
                    rid, t_arr, t_sub, g = meta.pop(i)
                    try:
                        sol, st, secs, _ = srv.get(i)
                        status, obj = st.termination_status.name, float(sol.x @ g.c)
                    except Exception as e:   # noqa: BLE001 -- a failed request is reported
                        status, obj, secs = f"ERROR: {e}", float("nan"), 0.0
                    srv.done.append(NativeDone(rid, t_arr, t_sub + secs, status, obj))
                    if status == "OPTIMAL":
                        max_obj_err = max(max_obj_err, abs(obj - g.objective) / (1 + abs(g.objective)))
            else:
                if k >= len(arr) and srv.pending() == 0:
                    break
                srv.pump()
            if time.perf_counter() - last_log > 10:
                last_log = time.perf_counter()
                log(f"[bench] qps={qps} t={now:.1f}s submitted={k} done={len(srv.done)}")
        srv.close()
        if args.serve_native:
            done = srv.done
        else:   # the Python scheduler keeps absolute times
            done = [NativeDone(r.rid, r.t_arrival - t0, r.t_done - t0, r.status, r.objective) for r in srv.done]
            for r in srv.done:
                if r.status == "OPTIMAL":
                    g = r.lp
                    max_obj_err = max(max_obj_err, abs(r.objective - g.objective) / (1 + abs(g.objective)))
        server = _served(done)
        server["per_class"] = _class_stats(done, picks)
        server["max_objective_rel_err"] = max_obj_err
        log(f"[bench] qps={qps} server: {server}")
        # ---- the baseline: one request at a time, first come first served
        if args.serve_fcfs == "run":
            fdone = []
            t0 = time.perf_counter()
            j = 0
            while j < len(arr):
                now = time.perf_counter() - t0
                if now > args.serve_duration + args.serve_drain:
                    break
                if arr[j] > now:
                    time.sleep(min(arr[j] - now, 0.01))
                    continue
                g = pool[picks[j][0]][picks[j][1]]
                _, st = plain_solve(g)
                fdone.append(NativeDone(j, arr[j], time.perf_counter() - t0, st.termination_status.name, 0.0))
                j += 1
            fcfs = _served(fdone)
            fcfs["per_class"] = _class_stats(fdone, picks)
            fcfs["mode"] = "run"
        elif args.serve_fcfs == "replay":
            fl = fcfs_latencies(arr, np.array([service[p] for p in picks]))
            fdone = [NativeDone(j, arr[j], arr[j] + fl[j], "OPTIMAL", 0.0) for j in range(len(arr))
                     if arr[j] + fl[j] <= args.serve_duration + args.serve_drain]
            fcfs = _served(fdone)
            fcfs["per_class"] = _class_stats(fdone, picks)
            fcfs["mode"] = "replay of measured solve times"
        else:
            fcfs = None
        if fcfs is not None:
            log(f"[bench] qps={qps} fcfs: {fcfs}")
        pt = {"qps": qps, "offered": int(len(arr)), "server": server, "fcfs": fcfs,
              "served_over_offered": server["served_lps_per_s"] / qps,
              "speedup_over_fcfs": (server["served_lps_per_s"] / fcfs["served_lps_per_s"])
              if fcfs and fcfs["served_lps_per_s"] > 0 else None}
        out_points.append(pt)
    best = max(out_points, key=lambda p: p["server"]["served_lps_per_s"])
    tracking = [p["qps"] for p in out_points if p["served_over_offered"] >= 0.9]
    return {
        "metric": "solve time to rel-KKT 1e-6; batched LPs/s + p50/p99 vs plain PDLP, 1 GPU",
        "value": best["server"]["served_lps_per_s"], "unit": "LPs/s (OPTIMAL results)", "n_gpus": 1,
        "steps": len(out_points), "warmup": 0, "ms_per_step": 1e3 * args.serve_duration,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 bulk / f64 checks+polish",
        "data": "synthetic (seeded request pool with known optima)",
        "config": {"workload": "serving_poisson_mix", "qps": args.qps, "duration_s": args.serve_duration,
                   "drain_s": args.serve_drain, "medium_workers": args.serve_medium_workers,
                   "overflow_queue": args.serve_overflow, "batch_limit": args.serve_batch_limit,
                   "batch2_limit": args.serve_batch2_limit if args.serve_native else 0,
                   "batch2_gib": args.serve_batch2_gib,
                   "server": "library (cpdlp_server_*)" if args.serve_native else "python scheduler",
                   "fcfs": args.serve_fcfs, "powm_tol": args.powm_tol,
                   "mix": "70/25/5 small/medium/large", "reltol": reltol,
                   "pool": {k: len(v) for k, v in pool.items()}},
        "knee_qps": max(tracking) if tracking else None,
        "points": out_points,
    }


@dataclass
class NativeDone:
    """A request finished by the library's server, in the fields the serving report reads."""
    rid: int
    t_arrival: float
    t_done: float
    status: str
    objective: float


def cpu_baseline(g, gpu_passes, reltol, budget_passes, powm_tol, seconds_left):
    """Oracle FP64 PDLP on this host: the whole solve when it fits in about 30 s,
    else a bounded sample of `budget_passes` passes (plus their checks) projected to
    the GPU solve's pass count (setup + passes x the sample's time per pass)."""
    import oracle as O
    cores = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", cores))
    threads = max(1, min(cores, omp))
    O.set_threads(threads)
    p = O.default_params(termination_reltol=reltol, max_kkt_passes=budget_passes, time_limit=1e9,
                         powm_tol=powm_tol)
    t = time.perf_counter()
    r = O.solve(g, p, hist_cap=16)
    wall = time.perf_counter() - t
    loop = max(wall - r.stats.setup_seconds, 1e-9)
    per_pass = loop / max(r.stats.kkt_passes, 1)
    projected = r.stats.setup_seconds + per_pass * gpu_passes
    if projected <= min(30.0, seconds_left - wall):   # small enough: time the whole solve
        p = O.default_params(termination_reltol=reltol, time_limit=1e9, powm_tol=powm_tol)
        t = time.perf_counter()
        r = O.solve(g, p, hist_cap=16)
        wall = time.perf_counter() - t
        return {
            "value": 1.0 / wall, "unit": "LPs/s", "cores": threads, "kind": "port",
            "projected": False,
            "sample": (f"oracle FP64 PDLP, the whole solve: {r.status} after {r.stats.kkt_passes} "
                       f"passes in {wall:.2f}s"),
            "status_at_sample": r.status,
        }
    return {
        "value": 1.0 / projected, "unit": "LPs/s", "cores": threads, "kind": "port",
        "projected": True,
        "sample": (f"PROJECTED: oracle FP64 PDLP, {r.stats.kkt_passes} passes incl. their checks "
                   f"({wall:.1f}s wall, setup {r.stats.setup_seconds:.1f}s, "
                   f"{1e3 * per_pass:.1f} ms/pass) projected to the GPU solve's "
                   f"{gpu_passes} passes: {projected:.0f}s per LP"),
        "status_at_sample": r.status,
    }


def budget_left(args) -> float:
    """Seconds left of the wall-clock budget (counted from process start)."""
    return args.budget - (time.perf_counter() - T_PROC0)


def bench_single(args):
    """One LP (large: the power-law LP; plain: the m=2,000 LP).  Large: a step is one
    chunk of the loop inside one timed full solve.  Plain: a step is one full solve
    (the whole solve takes ~9 chunks)."""
    import torch
    import coolpdlp_hip as cp
    t = time.perf_counter()
    g, reltol, desc = make_workload(args.workload)
    log(f"[bench] generated {desc['workload']} m={g.m} n={g.n} nnz={g.nnz} in "
        f"{time.perf_counter() - t:.1f}s")
    milp = cp.MILP.from_generated(g)
    # powm_tol: the reference's power-method rule (tol = eps*n^3) stops after one
    # Rayleigh quotient of a random vector at n = 2e6, which underestimates ||A|| and
    # makes eta*||A|| > 1; the benchmark converges the power method to 1e-4 instead and
    # reports the reference rule's eta beside it
    algo = cp.PDLP("float64", None, args.format, termination_reltol=reltol, time_limit=args.time_limit,
                   max_kkt_passes=args.max_passes, device=0, powm_tol=args.powm_tol,
                   show_progress=0, detect_infeasibility=int(args.detect_infeasibility),
                   long_row=args.long_row)
    t = time.perf_counter()
    eng = cp.Engine(milp, algo)
    torch.cuda.synchronize()
    log(f"[bench] uploaded to HBM in {time.perf_counter() - t:.1f}s")
    chunked = args.workload == "large"
    guard = None
    if chunked:
        t_solve0 = time.perf_counter()
        eng.initialize()
        status = cp.STILL_RUNNING
        for w in range(args.warmup):
            status = eng.run_checks(1)
        s0 = eng.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        timed = 0
        for k in range(args.steps):
            if status != cp.STILL_RUNNING:
                break
            status = eng.run_checks(1)
            timed += 1
        torch.cuda.synchronize()
        steps_elapsed = time.perf_counter() - t0
        s1 = eng.stats()
        per_check = steps_elapsed / max(timed, 1)
        log(f"[bench] {timed} timed steps: {1e3 * per_check:.2f} ms per chunk of "
            f"{algo.params.check_every} iterations")
        # the rest of the same solve, in slices the guard can account for
        last_log = time.perf_counter()
        while status == cp.STILL_RUNNING:
            left = budget_left(args) - args.reserve
            slice_checks = max(1, min(args.slice_checks, int(left / max(per_check, 1e-6) / 1.25)))
            if left <= 0 or slice_checks * per_check * 1.25 > left:
                guard = f"stopped by the wall-clock guard after {eng.stats().kkt_passes} passes"
                log(f"[bench] {guard}")
                break
            status = eng.run_checks(slice_checks)
            if time.perf_counter() - last_log > 20:
                last_log = time.perf_counter()
                st = eng.stats()
                log(f"[bench] solve t={time.perf_counter() - t_solve0:.1f}s passes={st.kkt_passes} "
                    f"rel={st.err.relative():.3e}")
        torch.cuda.synchronize()
        solve_seconds = time.perf_counter() - t_solve0
        st = eng.stats()
        steps_done, warm_done = timed, args.warmup
        ms_per_step = 1e3 * per_check
        # roofline over the timed steps: HIP events on the solver's stream around each chunk
        d_ms = s1.step_kernel_ms - s0.step_kernel_ms
        d_launch = s1.step_kernel_launches - s0.step_kernel_launches
    else:
        for w in range(args.warmup):
            _, st = eng.solve()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            _, st = eng.solve()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        solve_seconds = elapsed / max(args.steps, 1)
        steps_done, warm_done = args.steps, args.warmup
        ms_per_step = 1e3 * solve_seconds
        d_ms, d_launch = st.step_kernel_ms, st.step_kernel_launches
    log(f"[bench] solve: {st.termination_status.name} passes={st.kkt_passes} "
        f"{solve_seconds:.2f}s rel={st.err.relative():.3e}")
    sol = eng.solution()
    obj = cp.objective_value(sol.x, milp)

    bp, bd = iteration_bytes(g.m, g.n, g.nnz)
    t_iter_ms = d_ms / max(d_launch // 2, 1)
    achieved = (bp + bd) / (t_iter_ms * 1e-3) / 1e9
    t_iter_all = st.step_kernel_ms / max(st.step_kernel_launches // 2, 1)
    ms_p, ms_d = eng.profile_steps(20)
    kernels = {
        "primal_step": {"ms": ms_p, "bytes": bp, "GBs": bp / (ms_p * 1e-3) / 1e9,
                        "frac": bp / (ms_p * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "dual_step": {"ms": ms_d, "bytes": bd, "GBs": bd / (ms_d * 1e-3) / 1e9,
                      "frac": bd / (ms_d * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "timed": "each half-step alone after the solve, 20 launches between HIP events on the "
                 "solver's stream (launch gaps included, so the two sum to a little more than "
                 "ms_per_iteration)",
    }
    traffic, traffic_src = measured_traffic(desc["workload"], FORMATS[st.format_A], FORMATS[st.format_At])
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "pdhg_iteration (primal_step + dual_step)",
                "timed_over": "the timed steps" if chunked else "the last solve",
                "bytes_per_iteration": bp + bd, "ms_per_iteration": t_iter_ms,
                "ms_per_iteration_whole_solve": t_iter_all, "kernels": kernels}
    # the bound these kernels meet first: L2 request latency of the random gathers (one per
    # nonzero and product), against the tile kernel's rate with an L2-resident vector
    ceil_rate, ceil_src = gather_ceiling()
    gathers = 2 * g.nnz
    g_rate = gathers / (t_iter_ms * 1e-3)
    gather_roofline = {"bound": "l2_gather_latency", "achieved": g_rate, "peak": ceil_rate,
                       "unit": "gathers/s", "frac": (g_rate / ceil_rate) if ceil_rate else None,
                       "gathers_per_iteration": gathers, "peak_source": ceil_src}
    eng.close()

    if args.calibrate_copy:
        a = torch.empty(64 * 2**20, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 5 x 512 MiB read + 512 MiB written")
        del a, b

    # the reference's own power-method rule (powm! with tol = eps*n^3, maxiter = n)
    eta_ref = None
    if args.workload == "large" and not args.no_eta_ref and budget_left(args) > args.reserve + 30:
        ref_algo = cp.PDLP("float64", None, args.format, termination_reltol=reltol, device=0,
                           powm_tol=float("nan"))
        with cp.Engine(milp, ref_algo) as e2:
            e2.prepare()
            eta_r, _, it_r = e2.step_size()
        eta_ref = {"eta": eta_r, "powm_iters": it_r, "eta_over_converged": eta_r / st.eta,
                   "rule": "powm! tol = eps(Float64)*n^3, maxiter = n (linalg.jl:86-96)"}
        log(f"[bench] reference-rule eta {eta_r:.6e} after {it_r} power iterations "
            f"(converged eta {st.eta:.6e})")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(g, st.kkt_passes, reltol, args.cpu_passes, args.powm_tol,
                           budget_left(args) - 15)
        log(f"[bench] cpu baseline: {cpu['sample']}")

    # a solve capped by --max-passes (profiling runs) or stopped by the guard has no solve time
    finished = st.termination_status == cp.OPTIMAL and guard is None
    return {
        "metric": "solve time to rel-KKT 1e-6; batched LPs/s + p50/p99 vs plain PDLP, 1 GPU",
        "value": (1.0 / solve_seconds) if finished else None, "unit": "LPs/s", "n_gpus": 1,
        "steps": steps_done, "warmup": warm_done, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator with a known optimum)",
        "config": desc,
        "step": ("one chunk of check_every PDHG iterations + its check, inside the timed solve"
                 if chunked else "one full solve"),
        "solve_seconds": solve_seconds, "loop_seconds": st.loop_seconds,
        "wall_clock_guard": guard,
        "detect_infeasibility": bool(args.detect_infeasibility),
        "termination_status": st.termination_status.name, "kkt_passes": st.kkt_passes,
        "relative_kkt": st.err.relative(), "restarts": st.restarts,
        "objective": obj, "objective_closed_form": g.objective,
        "objective_rel_err": abs(obj - g.objective) / (1 + abs(g.objective)),
        "formats": {"A": FORMATS[st.format_A], "At": FORMATS[st.format_At]},
        "long_row": args.long_row,
        "line_grouped_tiles": {"A": bool(st.line_tiles & 1), "At": bool(st.line_tiles & 2)},
        "row_order": os.environ.get("CPDLP_ORDER", "first_touch"),
        "setup_seconds": st.setup_seconds, "eta": st.eta, "powm_iters": st.powm_iters,
        "powm_tol": args.powm_tol, "eta_reference_rule": eta_ref,
        "roofline": roofline, "gather_roofline": gather_roofline, "cpu_baseline": cpu,
        "process_seconds": time.perf_counter() - T_PROC0,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="large")
    ap.add_argument("--cpu-passes", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--time-limit", type=float, default=600.0)
    ap.add_argument("--budget", type=float, default=560.0,
                    help="wall-clock seconds from process start the whole run must fit in")
    ap.add_argument("--reserve", type=float, default=75.0,
                    help="seconds of the budget kept for the report after the solve")
    ap.add_argument("--slice-checks", type=int, default=100,
                    help="checks per loop call after the timed steps (guard granularity)")
    ap.add_argument("--no-eta-ref", action="store_true",
                    help="skip the reference power-method rule's step size")
    ap.add_argument("--powm-tol", type=float, default=1e-4)
    ap.add_argument("--detect-infeasibility", action="store_true",
                    help="opt-in ray certificates at every check (timed with the solve)")
    ap.add_argument("--polish-reltol", type=float, default=None,
                    help="mixed precision: FP32 phase stops at max(this, reltol) (engine default 1e-5)")
    # profiling runs only: cap the solve at this many passes (same kernels, fewer of
    # them); the default solves to the tolerance
    ap.add_argument("--max-passes", type=int, default=10**7)
    ap.add_argument("--long-row", type=int, default=64,
                    help="rows longer than this are long rows (wavefront rows, XCD slices)")
    # PMC calibration: also run a plain 512 MiB device-to-device copy kernel 5 times
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--format", default="auto", choices=FORMATS)
    ap.add_argument("--formats", default=None,
                    help="A,At formats for an auto run (sets CPDLP_FORMATS; for counter runs)")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--checks-per-round", type=int, default=20)
    ap.add_argument("--batch-tail", type=int, default=16,
                    help="continue the last N unfinished batch LPs on plain engines (0: off)")
    ap.add_argument("--batch-tail-passes", type=int, default=2000)
    ap.add_argument("--batch-tail-workers", type=int, default=8)
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="batch: LPs solved one at a time for the comparison (0: all of them)")
    ap.add_argument("--qps", type=float, nargs="+", default=[100.0, 300.0, 1000.0, 3000.0, 5000.0])
    ap.add_argument("--serve-duration", type=float, default=60.0)
    ap.add_argument("--serve-drain", type=float, default=30.0)
    ap.add_argument("--serve-fcfs", default="run", choices=["run", "replay", "none"],
                    help="the one-at-a-time baseline: run it on the same arrivals, model it, or skip it")
    ap.add_argument("--serve-checks", type=int, default=20)
    ap.add_argument("--serve-batch-gib", type=int, default=16)
    ap.add_argument("--serve-medium-workers", type=int, default=3)
    ap.add_argument("--serve-batch2-limit", type=int, default=0,
                    help="m + n limit of the library server's second batch route (0: off)")
    ap.add_argument("--serve-batch2-gib", type=float, default=2.0)
    ap.add_argument("--serve-overflow", type=int, default=0)
    ap.add_argument("--serve-batch-limit", type=int, default=8192)
    ap.add_argument("--serve-python", dest="serve_native", action="store_false",
                    help="serve with the Python scheduler instead of the library's own server (cpdlp_server_*)")
    ap.add_argument("--pool-small", type=int, default=48)
    ap.add_argument("--pool-medium", type=int, default=16)
    ap.add_argument("--pool-large", type=int, default=3)
    ap.add_argument("--mcf-nodes", type=int, default=500_000)
    ap.add_argument("--gather-rows", type=int, default=2_000_000)
    ap.add_argument("--gather-reps", type=int, default=20)
    ap.add_argument("--mcf-oracle-limit", type=float, default=900.0,
                    help="time limit (s) of the oracle's FP64 reference solve of the min-cost flow")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (got --gpus {args.gpus}): the solver has no multi-GPU "
              "path; a multi-GPU run would shard A by row blocks with an all-gather of the "
              "iterates per half-step.", file=sys.stderr)
        sys.exit(2)

    if args.formats:
        os.environ["CPDLP_FORMATS"] = args.formats
    import coolpdlp_hip as cp
    import torch

    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "mcf":
        print(json.dumps(bench_mcf(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.workload == "gather":
        print(json.dumps(bench_gather(args)), flush=True)
        return

    print(json.dumps(bench_single(args)), flush=True)


if __name__ == "__main__":
    main()
