"""Run one of the reference's input decks on the GPU, as the reference's
executable does with `snap <deck>`:

    python -m snap_amd.run_deck input.in [--precision fp64|fp64_exact|mixed] [--quiet]

Prints the deck settings (Snap::report_arguments, snap.cc:1344-1380), the
mapper tunables the reference would pick on one GPU (mapper.cc:101-172), one
line per convergence test and the execution summary in the words of the
reference's ConvergenceMonad (convergence.cc:153-263), and the MMS comparison
for manufactured-solution decks (snap.cc:542-559).  Exits non-zero when the
MMS acceptance fails, as the reference asserts.
"""
from __future__ import annotations

import argparse
import ctypes
import sys
from typing import List

from . import _lib, deck
from .solver import Solver


def report_arguments(p, extra) -> List[str]:
    mat = {0: "Homogenous Layout", 1: "Center Layout", 2: "Corner Layout"}[int(extra.get("mat_opt", 0))]
    src = {0: "Everywhere", 1: "Center", 2: "Corner", 3: "MMS"}[int(extra.get("src_opt", 0))]
    return [
        f"Dimensions: {int(extra.get('ndimen', 3))}",
        f"X-Chunks: {int(extra.get('ichunk', 1))}",
        f"Y-Chunks: {int(extra.get('npey', 1))}",
        f"Z-Chunks: {int(extra.get('npez', 1))}",
        f"nx,ny,nz: {p.nx},{p.ny},{p.nz}",
        f"lx,ly,lz: {p.lx:.8g},{p.ly:.8g},{p.lz:.8g}",
        f"Moments: {p.nmom}",
        f"Angles: {p.nang}",
        f"Groups: {p.ng}",
        f"Convergence: {p.epsi:.8g}",
        f"Max Inner Iterations: {p.iitm}",
        f"Max Outer Iterations: {p.oitm}",
        f"Time Dependent: {'Yes' if p.timedep else 'No'}",
        f"Total Simulation Time: {p.tf:.8g}",
        f"Total Steps: {p.nsteps}",
        f"Material Layout: {mat}",
        f"Source Layout: {src}",
        f"Fixup Flux: {'Yes' if p.fixup else 'No'}",
    ]


def tunables(ng: int, iitm: int, processors: int = 1, sms: int = 256) -> dict:
    out = (ctypes.c_int32 * 4)()
    _lib.check(_lib.load().snap_hip_mapper_tunables(ng, iitm, processors, sms, out),
               "snap_hip_mapper_tunables")
    return {"outer_runahead": out[0], "inner_runahead": out[1], "energy_chunk": out[2],
            "sms_per_sweep": out[3]}


def monad_report(log: List[dict]) -> List[str]:
    """The ConvergenceMonad's lines for the recorded tests: its loop counters
    (the inner counter restarts after a converged inner and at every outer
    test, the outer counter after a converged outer) and the time since the
    previous result of the same kind."""
    lines = []
    inner_no = outer_no = step_no = 0
    inner_start = outer_start = step_start = 0.0
    tot_inner = tot_outer = 0
    t_inner = t_outer = t_step = 0.0
    for r in log:
        t = r["t_us"]
        if r["kind"] == "inner":
            dt = t - inner_start
            word = "CONVERGED" if r["converged"] else "did not converge"
            lines.append(f"Inner loop {inner_no} of outer loop {outer_no} of time step {step_no} "
                         f"{word} in {int(dt)} microseconds")
            inner_no = 0 if r["converged"] else inner_no + 1
            tot_inner += 1
            t_inner += dt
            inner_start = t
        else:
            dt = t - outer_start
            if r["converged"]:
                lines.append(f"Outer loop {outer_no} of time step {step_no} CONVERGED in "
                             f"{int(dt)} microseconds")
                st = t - step_start
                lines.append(f"Time step {step_no} took {int(st)} microseconds")
                step_no += 1
                outer_no = 0
                step_start = t
                t_step += st
            else:
                lines.append(f"Outer loop {outer_no} of time step {step_no} did not converge "
                             f"in {int(dt)} microsecond")
                outer_no += 1
            tot_outer += 1
            t_outer += dt
            outer_start = t
            inner_no = 0
    bar = "-" * 57
    lines += [bar, "SNAP Execution Summary", bar,
              f"  Execution Time: {int(t_step)} us",
              f"  Total Time Steps: {step_no} (avg {t_step / max(step_no, 1):.8g} us / iter)",
              f"  Total Outer Loops: {tot_outer} (avg {t_outer / max(tot_outer, 1):.8g} us / iter)",
              f"  Total Inner Loops: {tot_inner} (avg {t_inner / max(tot_inner, 1):.8g} us / iter)",
              bar]
    return lines


def run(path: str, precision: str = "fp64", quiet: bool = False) -> dict:
    p, extra = deck.read_deck(path)
    out = report_arguments(p, extra)
    tun = tunables(p.ng, p.iitm)
    out.append("Mapper tunables (1 GPU): " + ", ".join(f"{k}={v}" for k, v in tun.items()))
    with Solver(p, precision=precision) as s:
        st = s.run()
        log = s.iteration_log()
        mms = s.mms_compare() if p.mms else None
    out += monad_report(log)
    ok = True
    if mms is not None:
        counted = (p.nx * int(extra.get("ichunk", 1))) * (p.ny * int(extra.get("npey", 1))) * \
            (p.nz * int(extra.get("npez", 1)))
        ref_avg = mms[2] * p.ncells / counted
        out += [f"MMS Max Diff: {mms[0]:.8g}", f"MMS Min Diff: {mms[1]:.8g}",
                f"MMS Avg Diff: {ref_avg:.8g} (over {counted} cells as snap.cc:546 counts; "
                f"{mms[2]:.8g} over the {p.ncells} cells)"]
        if mms[0] > 0.1:
            out.append("MMS FAILURE! MAX IS LARGER THAN 0.1!")
            ok = False
        if ref_avg > 0.001:
            out.append("MMS FAILURE! AVG IS LARGER THAN 0.001!")
            ok = False
    out.append(f"grind time {st['grind_ns']:.6g} ns per cell-angle-group over "
               f"{st['inner_total']} inner iterations; solve {st['solve_ms'] / 1e3:.6g} s")
    if not quiet:
        print("\n".join(out))
    return {"stats": st, "mms": mms, "ok": ok, "lines": out, "tunables": tun}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("deck")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp64_exact", "mixed"])
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    r = run(a.deck, a.precision, a.quiet)
    return 0 if r["ok"] else 1


if __name__ == "__main__":
    sys.exit(main())
