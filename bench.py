"""Benchmark: mini-KBA sweep grind time and solve time on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (one "step" = one full steady-state solve, outer/inner iterations to
convergence): 64^3 cells on the unit cube, 32 groups, 36 angles per octant
(288 directions), P3 scattering (16 flux moments), two materials with the
unit source in the central cube, vacuum boundaries, epsi 1e-4.  Inputs are
resident in HBM before the timed region.

Prints one JSON line.  `value` is the FP64 sweep grind time
t_sweep / (cells * 8*nang * ng * inner_iterations) in ns (lower is better),
where t_sweep covers the sweep launches and the sum of the octant partials
into the flux moments (which the reference does inside its sweep kernel).
`roofline` prices the algorithmic bytes per unit of the reference data flow,
B64 = 8 + (24*cmom + 8)/nang, against 8 TB/s over the same t_sweep, timed
with HIP events on the solver's stream.  `cpu_baseline` times the CPU oracle
in the operation order of the reference's AVX sweep variant (the one it
registers on x86 with AVX) over one inner iteration's sweeps of the same
problem on every host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "legion-snap_amd"))
sys.path.insert(0, REPO)

METRIC = "sweep grind time ns/(cell·angle·group) + solve s, FP64 vs mixed @1e-6 rel"
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "sweep_config1_fp64_traffic.json")


def measured_traffic(size: int):
    """HBM bytes of one inner sweep of the FP64 sweep kernel (all its
    launches, the unit `achieved` is priced over) from the committed
    FETCH_SIZE / WRITE_SIZE profile of this workload (calibrated against a
    plain copy), or None when no profile covers it."""
    if size != 64 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        return float(json.load(f)["bytes_per_sweep"])


def bytes_per_unit(nang: int, cmom: int, real_bytes: int) -> float:
    """Algorithmic HBM bytes per cell x direction x group of the reference
    data flow: dinv streamed once per octant; source moments, t_xs and the
    flux-moment read-modify-write amortised over the octant's angles."""
    return real_bytes + (24.0 * cmom + 8.0) / nang


CPU_INFO = {}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_bandwidth_gbs(torch, gib: int = 2, reps: int = 5) -> float:
    """Device-to-device copy rate (bytes read + written per second, best of
    reps) of a gib-GiB buffer: the HBM rate a streaming kernel reaches here."""
    n = (gib << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b) * 1e-3)
    del src, dst
    return 2 * n * 8 / best / 1e9


def cpu_baseline(problem, budget_s: float = 25.0, reps: int = 5):
    """Time the CPU oracle on one inner iteration's sweeps (all octants and
    groups) of `problem` using every host core (best of `reps` after a
    warm-up), or of a reduced group count when that would exceed the budget
    (the grind time is per unit).  The sweeps run in the reference's AVX
    variant's operation order when nang % 4 == 0, its generic order else."""
    from oracle import oracle
    import dataclasses
    order = "avx" if problem.nang % 4 == 0 else "generic"
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores) or cores)))
    # probe on 2 groups, then size the sample to the budget
    probe = dataclasses.replace(problem, ng=2, sigt=problem.sigt[:, :2], siga=problem.siga[:, :2],
                                sigs=problem.sigs[:, :2], slgg=problem.slgg[:2, :, :2, :],
                                qi=problem.qi[:2])
    t = oracle.time_sweeps(probe, 1, cores, order=order)
    per_group = t / 2.0
    ng = int(max(2, min(problem.ng, budget_s / (reps + 1) / max(per_group, 1e-9))))
    sample = dataclasses.replace(problem, ng=ng, sigt=problem.sigt[:, :ng], siga=problem.siga[:, :ng],
                                 sigs=problem.sigs[:, :ng], slgg=problem.slgg[:ng, :, :ng, :],
                                 qi=problem.qi[:ng])
    oracle.time_sweeps(sample, 1, cores, order=order)  # warm-up
    t = min(oracle.time_sweeps(sample, 1, cores, order=order) for _ in range(reps))
    units = sample.ncells * 8 * sample.nang * ng
    # the same sweeps on one core (the two-group probe), and the host CPU
    t1 = oracle.time_sweeps(probe, 1, 1, order=order)
    CPU_INFO["one_core_grind_ns"] = t1 * 1e9 / (probe.ncells * 8 * probe.nang * probe.ng)
    CPU_INFO["threads"] = cores
    CPU_INFO["model"] = cpu_model()
    return {
        "value": t * 1e9 / units,
        "unit": "ns/(cell·angle·group)",
        "cores": cores,
        "kind": "port",
        "sample": (f"one inner iteration's sweeps (8 octants x {ng} of {problem.ng} groups) of the "
                   f"{problem.nx}^3-cell, {problem.nang}-angle, P{problem.nmom - 1} workload in the "
                   f"reference's {order.upper() if order == 'avx' else order} variant's operation "
                   f"order, {units:.3g} cell-angle-groups in {t:.2f} s (best of {reps} after a "
                   f"warm-up, {cores} threads)"),
    }


def rocprof_summary(dirs, kernel_filter: str = "", sweeps_of: str = "",
                    launches_per_sweep: int = 0) -> str:
    """Markdown tables from rocprofv3 CSV output under `dirs`: per-kernel
    call counts and durations (kernel_stats.csv of --stats), and per-kernel
    counter values (counter_collection.csv of --pmc) as the mean per
    dispatch.  With sweeps_of (a kernel-name prefix) and launches_per_sweep,
    also each counter summed over one inner sweep of that kernel (dispatches
    / launches_per_sweep sweeps)."""
    import csv
    import glob
    out = []
    for d in dirs:
        for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)):
            rows = list(csv.DictReader(open(f)))
            out.append(f"kernel statistics ({os.path.relpath(f, d)})\n")
            out.append("| kernel | calls | mean (ms) | min (ms) | max (ms) | share |")
            out.append("|---|---|---|---|---|---|")
            for r in rows:
                if kernel_filter and kernel_filter not in r["Name"]:
                    continue
                out.append(f"| `{r['Name']}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4g} | "
                           f"{float(r['MinNs']) / 1e6:.4g} | {float(r['MaxNs']) / 1e6:.4g} | "
                           f"{float(r['Percentage']):.1f}% |")
            out.append("")
        per = {}
        for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
            for r in csv.DictReader(open(f)):
                name = r["Kernel_Name"]
                if kernel_filter and kernel_filter not in name:
                    continue
                k = per.setdefault(name, {"res": (r["VGPR_Count"], r["Accum_VGPR_Count"],
                                                  r["SGPR_Count"], r["LDS_Block_Size"],
                                                  r["Scratch_Size"]), "c": {}})
                c = k["c"].setdefault(r["Counter_Name"], [0.0, set()])
                c[0] += float(r["Counter_Value"])
                c[1].add(r["Dispatch_Id"])
        if per:
            out.append(f"counters ({d})\n")
            out.append("| kernel | VGPR/AGPR/SGPR/LDS/scratch | counter | dispatches | mean per dispatch"
                       + (" | per inner sweep |" if sweeps_of else " |"))
            out.append("|---|---|---|---|---|" + ("---|" if sweeps_of else ""))
            for name, k in sorted(per.items()):
                for cn, (tot, ids) in sorted(k["c"].items()):
                    line = (f"| `{name}` | {'/'.join(k['res'])} | {cn} | {len(ids)} | "
                            f"{tot / max(1, len(ids)):.5g} |")
                    if sweeps_of:
                        if name.startswith(sweeps_of) and launches_per_sweep:
                            line += f" {tot / (len(ids) / launches_per_sweep):.5g} |"
                        else:
                            line += " |"
                    out.append(line)
            out.append("")
    return "\n".join(out)


def dropin_timing(torch, n: int = 64, ng: int = 32, nang: int = 36, nmom: int = 4,
                  streams: int = 4, reps: int = 3) -> dict:
    """The drop-in sweep driven as Legion's task bodies drive it: one inner
    iteration is 8 MiniKBA corner tasks over all groups, the tasks spread over
    `streams` streams, every corner accumulating into the same flux fields.
    Three forms: run_gpu_sweep once per group (the reference's task-body
    loop, exact kernel), one group-range call per corner (exact), and one
    group-range call per corner with the fused kernel.  Synthetic inputs on
    the device; wall time per inner iteration (best of `reps`, host enqueue
    included) and the grind time over cells x 8 nang x ng."""
    from snap_amd import tasks
    cmom = nmom * nmom
    quad = tasks.quadrature(nang, nmom)
    tasks.initialize_gpu_context(quad, nang, nmom, (n, n, n))
    h = (2.0 * n, 2.0 * n, 2.0 * n)
    gen = torch.Generator(device="cuda").manual_seed(5)
    txs = 0.5 + torch.rand((ng, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    dinv = torch.zeros((ng, n, n, n, nang), dtype=torch.float64, device="cuda")
    tasks.calculate_geometry_param(txs, dinv, h)
    qtot = torch.rand((ng, n, n, n, cmom), dtype=torch.float64, device="cuda", generator=gen)
    qtot[..., 1:] *= 0.1
    flux = torch.zeros((ng, n, n, n), dtype=torch.float64, device="cuda")
    fluxm = torch.zeros((ng, n, n, n, cmom - 1), dtype=torch.float64, device="cuda")
    ss = [torch.cuda.Stream() for _ in range(streams)]
    units = n ** 3 * 8 * nang * ng
    out = {"streams": streams, "tasks": 8, "groups_per_task": ng}
    for key, form in (("per_group_calls_exact", None), ("group_range_exact", "exact"),
                      ("group_range_fused", "fused")):
        best = float("inf")
        for r in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for corner in range(8):
                with torch.cuda.stream(ss[corner % streams]):
                    tasks.mini_kba(corner, range(ng), qtot, flux, fluxm, dinv, txs, None, h, nang,
                                   nmom, group_range=form)
            torch.cuda.synchronize()
            if r:  # the first pass sizes the workspaces and loads the kernels
                best = min(best, time.perf_counter() - t0)
        out[key] = {"ms_per_inner": best * 1e3, "grind_ns": best * 1e9 / units}
    return out


def deck_suite(names, cpu: bool) -> dict:
    """Every reference input deck through the GPU solver (FP64, fused): the
    decks' own sizes, time steps and iteration control.  The CPU column times
    the oracle's steady sweeps of the same deck (one inner iteration, reduced
    groups for the large ones) as the per-unit CPU cost."""
    from oracle import oracle
    from snap_amd import deck
    from snap_amd.solver import Solver
    import dataclasses
    out = {}
    for name in names:
        path = os.path.join(REPO, "tests", "golden", "decks", name)
        p, _ = deck.read_deck(path)
        with Solver(p, precision="fp64") as s:
            s.run()  # warm-up
            st = s.run()
            plan = s.plan()
            mms = s.mms_compare() if p.mms else None
        row = {"cells": p.ncells, "groups": p.ng, "angles": p.nang, "steps": p.nsteps,
               "inner": st["inner_total"], "outer": st["outer_total"],
               "converged": st["converged"], "grind_ns": st["grind_ns"],
               "solve_s": st["solve_ms"] / 1e3, "sweep_s": st["sweep_ms"] / 1e3,
               "sweep": ("one cell per workgroup" if plan["cells_per_wave"] == 0 else
                         "one wavefront per tile" if plan["waves"] == 1 and
                         plan["cells_per_wave"] == 16 else "multi-wave tiles")}
        if mms is not None:
            row["mms_max_diff"] = mms[0]
        if cpu:
            q = dataclasses.replace(p, timedep=False, mms=False,
                                    qi=np.ones((p.ng, p.nz, p.ny, p.nx)))
            units = p.ncells * 8 * p.nang * p.ng
            if units > 4e8:  # fewer groups keep the sample near a second
                g = max(1, int(p.ng * 4e8 / units))
                q = dataclasses.replace(q, ng=g, sigt=q.sigt[:, :g], siga=q.siga[:, :g],
                                        sigs=q.sigs[:, :g], slgg=q.slgg[:g, :, :g],
                                        qi=q.qi[:g])
            t = oracle.time_sweeps(q, 1)
            row["cpu_grind_ns"] = t * 1e9 / (q.ncells * 8 * q.nang * q.ng)
            row["speedup_vs_cpu"] = row["cpu_grind_ns"] / row["grind_ns"]
        out[name] = row
        print(name, json.dumps(row), file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the adjoint (config 3) and batch (config 4) measurements")
    ap.add_argument("--size", type=int, default=64, help="grid edge (default: the 64^3 workload)")
    ap.add_argument("--decks", nargs="*", default=None,
                    help="run the reference input decks (all when no names are given) and "
                         "print their suite line instead of the default workload")
    ap.add_argument("--sweeps", nargs=2, metavar=("KERNEL", "PRECISION"), default=None,
                    help="profiling harness: one solve of the workload, then --count timed inner "
                         "sweeps with the given sweep kernel (auto, tiled, wave_tile) and precision "
                         "(fp64, fp64_exact, mixed); prints one JSON line")
    ap.add_argument("--count", type=int, default=3, help="timed inner sweeps for --sweeps")
    ap.add_argument("--dinv", default="default", help="dinv mode for --sweeps (stream, recompute)")
    ap.add_argument("--tz", type=int, default=0, help="tile depth for --sweeps (0 = default)")
    ap.add_argument("--summarize", nargs="+", metavar="DIR", default=None,
                    help="print markdown tables of the rocprofv3 CSV output under DIR(s) "
                         "(no GPU needed) and exit")
    ap.add_argument("--kernel", default="", help="with --summarize: only kernels containing this")
    ap.add_argument("--sweeps-of", default="", help="with --summarize: kernel-name prefix whose "
                    "counters are also summed per inner sweep")
    ap.add_argument("--launches-per-sweep", type=int, default=0,
                    help="with --sweeps-of: launches of that kernel per inner sweep")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device-to-device copy and exit (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE traffic counters)")
    args = ap.parse_args()
    if args.summarize:
        print(rocprof_summary(args.summarize, args.kernel, args.sweeps_of, args.launches_per_sweep))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        return 2

    import torch
    from snap_amd import problems
    from snap_amd.solver import Solver

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if args.decks is not None:
        names = args.decks or sorted(os.listdir(os.path.join(REPO, "tests", "golden", "decks")))
        names = [n for n in names if n.endswith(".in")]
        rows = deck_suite(names, cpu=not args.no_cpu)
        tot_units = sum(r["cells"] * 8 * r["angles"] * r["groups"] * r["inner"] for r in rows.values())
        tot_sweep = sum(r["sweep_s"] for r in rows.values())
        print(json.dumps({"metric": "reference input decks: FP64 sweep grind time and solve time",
                          "value": tot_sweep * 1e9 / tot_units, "unit": "ns/(cell·angle·group)",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
                          "data": "the reference's input decks (input/*.in)", "decks": rows}))
        return 0
    if args.sweeps is not None:
        kern, prec = args.sweeps
        p = problems.config1(n=args.size)
        with Solver(p, precision=prec, sweep_kernel=kern, dinv=args.dinv,
                    tile=(0, 0, args.tz)) as s:
            st = s.run()
            ts = s.time_sweeps(args.count, prec)
            plan = s.plan()
        print(json.dumps({"kernel": kern, "precision": prec, "dinv": args.dinv, "plan": plan,
                          "inner": st["inner_total"], "grind_ns": st["grind_ns"], **ts}))
        return 0
    if args.copy_probe:
        n = 1 << 27  # 1 GiB of doubles
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))
        return 0
    p = problems.config1(n=args.size)
    units = p.units_per_sweep
    details = {}

    # ---- FP64, fused multiply-adds (the default plan: the one-wavefront tile
    # sweep for 33..64 angles)
    def fp64_run(kernel):
        with Solver(p, precision="fp64", sweep_kernel=kernel) as s:
            for _ in range(args.warmup):
                s.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            stats = [s.run() for _ in range(args.steps)]
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            f, _ = s.flux()
            ts = s.time_sweeps(3)
            plan = s.plan()
        return stats, wall, f, ts, plan

    stats, wall, f64, ts, plan64 = fp64_run("auto")
    inner = stats[-1]["inner_total"]
    sweep_ms = float(np.mean([st["sweep_ms"] for st in stats]))
    solve_ms = float(np.mean([st["solve_ms"] for st in stats]))
    grind = sweep_ms * 1e6 / (units * inner)
    b64 = bytes_per_unit(p.nang, p.cmom, 8)
    # algorithmic bytes of one inner sweep over its measured time (sweep
    # launches plus the octant sum, HIP events on the solver's stream)
    achieved = units * b64 / (ts["ms_per_sweep"] * 1e-3) / 1e9
    details["fp64"] = {"grind_ns": grind, "sweep_ms_per_inner": sweep_ms / inner,
                       "solve_s": solve_ms / 1e3, "inner": inner, "outer": stats[-1]["outer_total"],
                       "wall_s_per_step": wall / args.steps,
                       "timed_sweep_ms": ts["ms_per_sweep"], "launch_ms": ts["ms_per_launch"],
                       "launches_per_sweep": ts["launches_per_sweep"], "bytes_per_unit": b64,
                       "plan": plan64}
    if not args.no_extra:
        st_t, _, f_t, ts_t, plan_t = fp64_run("tiled")
        g_t = float(np.mean([st["sweep_ms"] for st in st_t])) * 1e6 / (units * st_t[-1]["inner_total"])
        details["fp64_multiwave_tiles"] = {
            "grind_ns": g_t, "timed_sweep_ms": ts_t["ms_per_sweep"], "plan": plan_t,
            "inner": st_t[-1]["inner_total"],
            "maxrel_vs_default": float(np.max(np.abs(f_t - f64) / np.abs(f64)))}

    # ---- the drop-in sweep as Legion's task bodies would drive it
    if not args.no_extra:
        details["dropin_config1"] = dropin_timing(torch)
        details["dropin_config1"]["solver_grind_ns"] = grind

    # ---- FP64 in the reference's operation order (bit-equal to the oracle)
    with Solver(p, precision="fp64_exact") as s:
        st = s.run()
        fex, _ = s.flux()
    details["fp64_exact"] = {"grind_ns": st["sweep_ms"] * 1e6 / (units * st["inner_total"]),
                             "solve_s": st["solve_ms"] / 1e3, "inner": st["inner_total"],
                             "maxrel_vs_fp64_fused": float(np.max(np.abs(fex - f64) / np.abs(fex)))}

    # ---- mixed precision (FP64 moments / sources / tests): the default
    # (FP32 sweeps, FP64 polish sweeps at the end of each outer), per-outer
    # defect correction with FP32 correction sweeps and FP64 polish, and the
    # two without polish
    for key, kw in (("mixed", {}), ("mixed_refine", {"refine": True}),
                    ("mixed_fp32_only", {"fp64_polish": 0.0}),
                    ("mixed_refine_no_polish", {"refine": True, "fp64_polish": 0.0})):
        with Solver(p, precision="mixed", **kw) as s:
            for _ in range(args.warmup):
                s.run()
            st = s.run()
            fmx, _ = s.flux()
            plan = s.plan()
        d = np.abs(fmx - fex)
        g = st["sweep_ms"] * 1e6 / (units * st["inner_total"])
        details[key] = {"grind_ns": g, "solve_s": st["solve_ms"] / 1e3, "inner": st["inner_total"],
                        "outer": st["outer_total"], "sweeps_fp64": st["sweeps_fp64"],
                        "sweeps_fp32": st["sweeps_fp32"],
                        "maxrel_vs_fp64": float(np.max(d / np.abs(fex))),
                        "l2rel_vs_fp64": float(np.linalg.norm(d) / np.linalg.norm(fex)),
                        "speedup_vs_fp64_grind": grind / g,
                        "speedup_vs_fp64_solve": solve_ms / st["solve_ms"], "plan": plan}

    # ---- adjoint gradient (config 3: 48^3, 16 groups, 24 angles, P0): the
    # FP64 forward solve, then dR/dsiga for all cells by reverse sweeps in
    # FP64 and in FP32 (on the FP64 solve's fixup decisions)
    if not args.no_extra:
        c3 = problems.config3()
        resp = np.zeros((c3.ng, c3.nz, c3.ny, c3.nx))
        resp[:, -8:, -8:, -8:] = 1.0
        with Solver(c3, precision="fp64") as s:
            st = s.run()
            for prec in ("fp64", "mixed"):
                s.gradient(resp, precision=prec)  # warm-up
                t0 = time.perf_counter()
                g, R, it = s.gradient(resp, precision=prec)
                tg = time.perf_counter() - t0
                details[f"adjoint_config3_{prec}"] = {
                    "forward_solve_s": st["solve_ms"] / 1e3, "gradient_s": tg,
                    "adjoint_iterations": it, "response": R, "cells": c3.ncells,
                    "grad_siga_max_abs": float(np.max(np.abs(g.sum(axis=0))))}

    # ---- batch of uneven problems (config 4: 48 problems, edges 24..64,
    # 16 groups, 24 angles, P3), one launch per sweep stage, graph-captured
    # outer iterations, autotuned tile shape (cached for the second batch)
    if not args.no_extra:
        from snap_amd.batch import Batch
        c4 = problems.config4()
        for prec in ("fp64", "mixed"):
            with Batch(c4, precision=prec, use_graph=True) as b:
                st0, _ = b.run()  # warm-up (captures the graph)
                st, per = b.run()
            with Batch(c4, precision=prec, use_graph=True) as b:
                st2, _ = b.run()
            times = np.array([x["time_ms"] for x in per])
            inners = [x["inner_total"] for x in per]
            outers = [x["outer_total"] for x in per]
            details[f"batch_config4_{prec}"] = {
                "problems": len(c4), "cells": int(sum(q.ncells for q in c4)),
                "grind_ns": st["grind_ns"], "solve_s": st["solve_ms"] / 1e3,
                "converged": st["converged"], "outer_iterations": st["outer_iterations"],
                "per_problem_inner_min_max": [min(inners), max(inners)],
                "per_problem_outer_min_max": [min(outers), max(outers)],
                "sweeps_fp64": int(sum(x["sweeps_fp64"] for x in per)),
                "sweeps_total": int(sum(inners)),
                # device clock at the end of the (lockstep) outer iteration in
                # which the problem's outer test passed
                "time_to_convergence_p50_ms": float(np.percentile(times, 50)),
                "time_to_convergence_p99_ms": float(np.percentile(times, 99)),
                "tile": list(st["tile"]), "angles_per_lane": st["angles_per_lane"],
                "groups_per_wave": st["groups_per_wave"],
                "autotune_s": st0["autotune_ms"] / 1e3,
                "second_batch_autotune_s": st2["autotune_ms"] / 1e3,
                "second_batch_plan_cached": st2["plan_cached"]}
        details["batch_config4_mixed_speedup"] = (details["batch_config4_fp64"]["solve_s"] /
                                                  details["batch_config4_mixed"]["solve_s"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p)
        details["fp64_speedup_vs_cpu"] = cpu["value"] / grind
        details["cpu"] = dict(CPU_INFO)
    # the sweep's algorithmic rate against a measured copy as well as the
    # published peak
    if not args.no_extra:
        copy_gbs = copy_bandwidth_gbs(torch)
        details["hbm_copy_GBps"] = copy_gbs
        details["roofline_frac_of_copy"] = achieved / copy_gbs

    line = {
        "metric": METRIC,
        "value": grind,
        "unit": "ns/(cell·angle·group)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": solve_ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"config 1: {p.nx}^3 cells, {p.ng} groups, {p.nang} angles/octant, "
                                f"P{p.nmom - 1} ({p.cmom} moments), two materials, FP64 solve to "
                                f"convergence (one step = one solve)"),
                   "cells": p.ncells, "groups": p.ng, "angles_per_octant": p.nang, "nmom": p.nmom},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.size)},
        "cpu_baseline": cpu,
        "details": details,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
