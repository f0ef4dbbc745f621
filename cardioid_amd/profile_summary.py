"""Summarises rocprofv3 results of one kernel into a short markdown file and
a traffic JSON (used by bench.py for roofline.traffic).

    python -m cardioid_amd.profile_summary --kernel NAME --cells N --bytes B \
        --trace DIR --fetch DIR --write DIR --out profiles/STEM \
        --title "..." --cmd-trace "..." --cmd-fetch "..." --cmd-write "..."

DIR are rocprofv3 output directories (SQLite results).  FETCH_SIZE and
WRITE_SIZE (KiB per dispatch) are scaled by what the same counters read for
the 512 MiB device-to-device copy that the profiled program runs first
(`__amd_rocclr_copyBuffer`: 512 MiB read and 512 MiB written).
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sqlite3

COPY_BYTES = 512 * 2 ** 20     # per copy; profile_step makes 3 (the runtime may split each)
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _db(d):
    f = sorted(glob.glob(os.path.join(d, "**", "*.db"), recursive=True))
    if not f:
        raise FileNotFoundError(f"no rocprofv3 database under {d}")
    return sqlite3.connect(f[0])


def trace_rows(d):
    c = _db(d)
    return [tuple(r) for r in c.execute(
        "select name, total_calls, total_duration, average, percentage from top_kernels")]


def counter_means(d, counter):
    """kernel name -> (mean per dispatch, dispatches, sum over dispatches)"""
    c = _db(d)
    rows = c.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                     "where counter_name = ? group by kernel_name, dispatch_id", (counter,))
    acc = {}
    for name, _, v in rows:
        acc.setdefault(name, []).append(v)
    return {k: (sum(v) / len(v), len(v), sum(v)) for k, v in acc.items()}


def _pick(d, kernel):
    keys = [k for k in d if kernel in k]
    if not keys:
        raise KeyError(f"{kernel} not in {list(d)}")
    return keys[0]


# counters kept per CU (one instance each), whose _sum over the CUs divided
# by the CU count and the kernel's cycles is a busy or stall fraction
_UNIT_FRACTIONS = ("TA_TA_BUSY_sum", "TD_TD_BUSY_sum", "TD_TC_STALL_sum",
                   "TCP_PENDING_STALL_CYCLES_sum", "TA_DATA_STALLED_BY_TC_CYCLES_sum",
                   "TA_ADDR_STALLED_BY_TC_CYCLES_sum")
# counters reported per 64-cell tile
_UNIT_PER_TILE = ("TCP_TOTAL_CACHE_ACCESSES_sum", "TD_LOAD_WAVEFRONT_sum", "TCP_TCC_READ_REQ_sum",
                  "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR")


def units_section(a, per):
    """Texture address (TA), data (TD) and L1 (TCP) counters of the kernel
    and of the calibration copy: busy and stall fractions of the kernel's
    cycles (GRBM_GUI_ACTIVE counts every XCD, so one XCD's cycles are it
    divided by the XCD count), and accesses per 64-cell tile."""
    cols = (a.kernel, COPY_KERNEL)
    vals = {}
    for d in a.units:
        c = _db(d)
        for (name,) in c.execute("select distinct counter_name from counters_collection"):
            m = counter_means(d, name)
            vals[name] = {k: m[_pick(m, k)][0] for k in cols if any(k in x for x in m)}
    lines = ["## Memory units (TA, TD, TCP)", "", "Commands:", ""]
    lines += [f"    {c}" for c in a.cmd_units]
    lines += ["", f"Per dispatch. Fractions: counter / ({a.cus} CUs x GRBM_GUI_ACTIVE / {a.xcds} "
              "XCDs); per tile: counter / (tissue cells / 64) per launch.", "",
              f"| counter | `{a.kernel}` | `{COPY_KERNEL}` |", "|---|---|---|"]
    grbm = vals.get("GRBM_GUI_ACTIVE", {})
    tiles = a.cells / 64.0 / per
    for name, v in vals.items():
        row = []
        for k in cols:
            x = v.get(k)
            if x is None:
                row.append("-")
            elif name in _UNIT_FRACTIONS and grbm.get(k):
                row.append(f"{x:.4g} (**{x / (a.cus * grbm[k] / a.xcds):.2f}** busy)"
                           if "BUSY" in name else
                           f"{x:.4g} ({x / (a.cus * grbm[k] / a.xcds):.2f} of the cycles)")
            elif name in _UNIT_PER_TILE and k == a.kernel:
                row.append(f"{x:.4g} ({x / tiles:.1f} per tile)")
            else:
                row.append(f"{x:.4g}")
        lines.append(f"| {name} | {row[0]} | {row[1]} |")
    return lines + [""]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--cells", type=int, required=True)
    ap.add_argument("--bytes", type=float, required=True, help="algorithmic bytes per cell-step")
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--out", required=True)
    ap.add_argument("--title", required=True)
    ap.add_argument("--workload", default="")
    ap.add_argument("--cmd-trace", default="")
    ap.add_argument("--cmd-fetch", default="")
    ap.add_argument("--cmd-write", default="")
    ap.add_argument("--note", default="")
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--launches-per-step", type=int, default=1,
                    help="launches of the kernel that together make one step (the drop-in "
                         "reaction: one per cell type); per-step figures sum them")
    ap.add_argument("--valu", help="rocprofv3 --pmc run of SQ_WAVES SQ_INSTS_VALU "
                    "SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES GRBM_GUI_ACTIVE GRBM_COUNT")
    ap.add_argument("--cmd-valu", default="")
    ap.add_argument("--flops", type=float, default=1060.0,
                    help="algorithmic floating-point operations per cell-step (FMA = 2); "
                         "0 leaves the flop fraction out")
    ap.add_argument("--units", action="append", default=[],
                    help="a rocprofv3 --pmc run of texture-unit / L1 counters (repeatable); "
                         "busy fractions need GRBM_GUI_ACTIVE in the same run")
    ap.add_argument("--cmd-units", action="append", default=[])
    ap.add_argument("--cus", type=int, default=256)
    ap.add_argument("--xcds", type=int, default=8)
    ap.add_argument("--peak-tflops", type=float, default=78.6)
    ap.add_argument("--simds", type=int, default=1024)
    ap.add_argument("--clock-ghz", type=float, default=2.4)
    a = ap.parse_args(argv)

    rows = trace_rows(a.trace)
    k = [r for r in rows if a.kernel in r[0]][0]
    per = a.launches_per_step
    avg_us = k[3] * per          # one step's launches
    lines = [f"# {a.title}", ""]
    if a.note:
        lines += [a.note, ""]
    lines += ["## Kernel trace", "", "Command (from the repo root, on the MI355X):", "",
              f"    {a.cmd_trace}", "", "| kernel | calls | total (us) | average (us) | % |",
              "|---|---|---|---|---|"]
    for r in rows[:6]:
        lines.append(f"| `{r[0][:90]}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} | {r[4]:.2f} |")
    achieved = a.bytes * a.cells / (avg_us * 1e-6) / 1e9
    what = ("per launch (one step of" if per == 1 else
            f"per step ({per} launches of {k[3]:.1f} us on average, together one step of")
    lines += ["", f"Average duration of `{a.kernel}`: {avg_us:.1f} us {what} "
              f"{a.cells} tissue cells) = {a.cells / avg_us / 1e3:.3f} Gcell-steps/s; at "
              f"{a.bytes:g} algorithmic bytes per cell-step that is {achieved:.0f} GB/s "
              f"({achieved / 8000:.3f} of 8 TB/s).", ""]
    out = {"kernel": k[0], "workload": a.workload, "tissue_cells": a.cells,
           "average_us": avg_us, "algorithmic_bytes_per_cell_step": a.bytes,
           "launches_per_step": per}
    if a.fetch and a.write:
        f = counter_means(a.fetch, "FETCH_SIZE")
        w = counter_means(a.write, "WRITE_SIZE")
        cf, cw = f[_pick(f, COPY_KERNEL)], w[_pick(w, COPY_KERNEL)]
        fs = a.copies * COPY_BYTES / (cf[2] * 1024)
        ws = a.copies * COPY_BYTES / (cw[2] * 1024)
        fk, nf, _ = f[_pick(f, a.kernel)]
        wk, nw, _ = w[_pick(w, a.kernel)]
        fb, wb = fk * 1024 * fs * per, wk * 1024 * ws * per
        lines += ["## HBM traffic per step (FETCH_SIZE, WRITE_SIZE)", "",
                  "Commands (one counter per run):", "", f"    {a.cmd_fetch}", f"    {a.cmd_write}",
                  "", f"Calibration: the {a.copies} device-to-device copies of 512 MiB at the start "
                  f"of the same runs ({cf[1]} dispatches of `{COPY_KERNEL}`) read FETCH_SIZE = "
                  f"{cf[2] / a.copies:.1f} KiB and WRITE_SIZE = {cw[2] / a.copies:.1f} KiB per copy, "
                  f"so bytes read = FETCH_SIZE x 1024 x {fs:.4f} and bytes written = WRITE_SIZE x "
                  f"1024 x {ws:.4f}.", "",
                  "| | raw (KiB/launch) | launches | scaled bytes/step | bytes per tissue cell-step |",
                  "|---|---|---|---|---|",
                  f"| FETCH_SIZE | {fk:.1f} | {nf} | {fb:.4e} | {fb / a.cells:.1f} |",
                  f"| WRITE_SIZE | {wk:.1f} | {nw} | {wb:.4e} | {wb / a.cells:.1f} |",
                  f"| total | | | {fb + wb:.4e} | {(fb + wb) / a.cells:.1f} "
                  f"(algorithmic {a.bytes:g}) |", ""]
        out.update(fetch_bytes_per_launch=fb, write_bytes_per_launch=wb,
                   bytes_per_launch=fb + wb, fetch_scale=fs, write_scale=ws)
    if a.valu:
        cnt = {c: counter_means(a.valu, c) for c in ("SQ_WAVES", "SQ_INSTS_VALU",
                                                     "SQ_ACTIVE_INST_VALU")}
        waves = cnt["SQ_WAVES"][_pick(cnt["SQ_WAVES"], a.kernel)][0] * per
        insts = cnt["SQ_INSTS_VALU"][_pick(cnt["SQ_INSTS_VALU"], a.kernel)][0] * per
        active = cnt["SQ_ACTIVE_INST_VALU"][_pick(cnt["SQ_ACTIVE_INST_VALU"], a.kernel)][0] * per
        tiles = a.cells / 64.0
        # a wave64 VALU instruction occupies its SIMD for 4 cycles
        avail = a.simds * a.clock_ghz * 1e9 * avg_us * 1e-6
        issue = 4 * insts / avail
        flop_frac = a.flops * a.cells / (avg_us * 1e-6) / (a.peak_tflops * 1e12)
        flop_line = (f"  Algorithmic arithmetic: {a.flops:g} flop per cell-step = "
                     f"{flop_frac * a.peak_tflops:.1f} TFLOP/s = {flop_frac:.3f} of "
                     f"{a.peak_tflops} TFLOP/s." if a.flops > 0 else "")
        lines += ["## VALU", "", "Command:", "", f"    {a.cmd_valu}", "",
                  "| | per step | per wave | per 64-cell tile |", "|---|---|---|---|",
                  f"| SQ_WAVES | {waves:.4g} | | |",
                  f"| SQ_INSTS_VALU | {insts:.4g} | {insts / waves:.0f} | "
                  f"{insts / tiles:.0f} |",
                  f"| SQ_ACTIVE_INST_VALU | {active:.4g} | {active / waves:.0f} | "
                  f"{active / tiles:.0f} |", "",
                  "(A wave of a kernel on the resident grid walks many tiles; the per-tile "
                  "figure compares kernels of either kind.)", "",
                  f"VALU issue: 4 cycles x {insts:.4g} instructions over {a.simds} SIMDs x "
                  f"{a.clock_ghz} GHz x {avg_us:.1f} us = **{issue:.2f}** of the SIMDs' cycles "
                  f"(at the peak clock; a lower clock raises it).{flop_line}", ""]
        out.update(valu_instructions_per_wave=insts / waves,
                   valu_instructions_per_tile=insts / tiles, valu_issue_frac=issue,
                   flop_per_cell_step=a.flops, flop_frac=flop_frac,
                   peak_tflops=a.peak_tflops, clock_ghz=a.clock_ghz)
    if a.units:
        lines += units_section(a, per)
    with open(a.out + ".md", "w") as fh:
        fh.write("\n".join(lines))
    with open(a.out + "_traffic.json", "w") as fh:
        json.dump(out, fh, indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
