"""Summaries of rocprofv3 CSV output for the profiles/ directory.

    python 3d-channel-flow-code_amd/nsx/profsum.py stats DIR
        per-kernel time table of DIR/**/*_kernel_stats.csv (--kernel-trace --stats)
    python 3d-channel-flow-code_amd/nsx/profsum.py traffic FETCH_DIR WRITE_DIR N_COPY
        per-kernel HBM traffic of two --pmc runs (FETCH_SIZE, then WRITE_SIZE,
        KiB per dispatch), each scaled by the factor that makes the k_copy
        dispatch of N_COPY doubles read and write 8 N_COPY bytes; the mean
        over the kernel's dispatches that did their work (a solver launch
        queued after the stop returns at once, moving under a tenth of the
        kernel's largest fetch: those are left out), the same launches the
        bench's event timing averages over

The traffic summary ends with one machine-readable line per kernel,
`traffic: <kernel> <fetch bytes> <write bytes> <dispatches>`, which bench.py
reads for the `traffic` field of its roofline.
"""
from __future__ import annotations

import csv
import glob
import os
import sys
from collections import defaultdict

import numpy as np


def _csvs(d: str, suffix: str):
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def short(name: str) -> str:
    n = name.replace("(anonymous namespace)::", "")
    n = n.split("(")[0].strip()
    n = n[5:] if n.startswith("void ") else n
    # template arguments without spaces, so a name is one word: k_cgs<float,1>
    return n.replace(", ", ",")


def stats(d: str) -> str:
    files = _csvs(d, "_kernel_stats.csv")
    if not files:
        raise SystemExit(f"no *_kernel_stats.csv under {d}")
    rows = list(csv.DictReader(open(files[0])))
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:20]:
        out.append(f"| {short(r['Name'])} | {r['Calls']} | {int(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {int(r['MinNs']) / 1e3:.2f} | {int(r['MaxNs']) / 1e3:.2f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out)


def _pmc(d: str, counter: str):
    files = _csvs(d, "_counter_collection.csv")
    if not files:
        raise SystemExit(f"no *_counter_collection.csv under {d}")
    per = defaultdict(list)
    for r in csv.DictReader(open(files[0])):
        if r.get("Counter_Name") != counter:
            continue
        per[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return per


def traffic(fetch_dir: str, write_dir: str, n_copy: int) -> str:
    fe, wr = _pmc(fetch_dir, "FETCH_SIZE"), _pmc(write_dir, "WRITE_SIZE")
    if "k_copy" not in fe or "k_copy" not in wr:
        raise SystemExit("the runs hold no k_copy dispatch to calibrate against")
    copy_bytes = 8.0 * n_copy
    f_scale = copy_bytes / (sum(fe["k_copy"]) / len(fe["k_copy"]) * 1024.0)
    w_scale = copy_bytes / (sum(wr["k_copy"]) / len(wr["k_copy"]) * 1024.0)
    out = [f"calibration: k_copy of {n_copy} doubles ({copy_bytes / 1e6:.1f} MB each way): "
           f"FETCH_SIZE x {f_scale:.4f}, WRITE_SIZE x {w_scale:.4f}", "",
           "| kernel | dispatches (that worked) | fetch MB/dispatch (mean) | write MB/dispatch (mean) |",
           "|---|---|---|---|"]
    lines = []

    def worked(fv, wv):
        f, w = np.asarray(fv, dtype=float), np.asarray(wv, dtype=float)
        n = min(len(f), len(w))
        f, w = f[:n], w[:n]
        # a launch queued after its solve stopped returns at once and moves
        # a few KB; every working launch moves at least one vector
        keep = f >= 0.01 * f.max() if n else np.zeros(0, bool)
        return f[keep], w[keep]

    tab = {k: worked(fe[k], wr[k]) for k in set(fe) & set(wr)}
    for k in sorted(tab, key=lambda k: -float(tab[k][0].mean()) if len(tab[k][0]) else 0.0):
        fk, wk = tab[k]
        if not len(fk):
            continue
        f = float(fk.mean()) * 1024.0 * f_scale
        w = float(wk.mean()) * 1024.0 * w_scale
        out.append(f"| {k} | {len(fe[k])} ({len(fk)}) | {f / 1e6:.1f} | {w / 1e6:.1f} |")
        lines.append(f"traffic: {k} {f:.0f} {w:.0f} {len(fk)}")
    return "\n".join(out + [""] + lines)


def read_traffic(path: str) -> dict:
    """{kernel: (fetch bytes, write bytes)} from a traffic summary, or {}"""
    res = {}
    if not os.path.exists(path):
        return res
    for line in open(path):
        if line.startswith("traffic: "):
            _, k, f, w, _n = line.split()
            res[k] = (float(f), float(w))
    return res


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        print(stats(sys.argv[2]))
    elif len(sys.argv) >= 5 and sys.argv[1] == "traffic":
        print(traffic(sys.argv[2], sys.argv[3], int(sys.argv[4])))
    else:
        raise SystemExit(__doc__)
