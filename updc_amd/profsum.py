"""Summaries of rocprofv3 CSV output for the iteration kernels.

Two summaries, each a short text file:

* ``trace``: from ``rocprofv3 --kernel-trace --stats -f csv`` output, the
  per-stage dispatch count and duration statistics of the iteration kernels
  (dispatches from the first iteration on; the set-up kernels are skipped).
* ``traffic``: from two ``rocprofv3 --pmc`` runs (FETCH_SIZE, then WRITE_SIZE,
  each reported in KiB per dispatch), the HBM bytes per dispatch of each
  iteration stage, each counter scaled by what it reads for the copy kernel
  ``k_copy_ref`` of known size (``--copy-bytes`` read and written).

Usage:
  python -m updc_amd.profsum trace DIR --cmd "<rocprofv3 command>"
  python -m updc_amd.profsum traffic FETCH_DIR WRITE_DIR --copy-bytes N \\
      --cmd "<fetch command>" --cmd "<write command>"

The traffic summary carries one machine-readable line per stage,
``stage <name> fetch_bytes <F> write_bytes <W> traffic_bytes <F+W>``, which
``bench.py`` reads to fill ``roofline.traffic``.
"""
from __future__ import annotations

import argparse
import csv
import re
import statistics
import sys
from collections import defaultdict
from pathlib import Path

# iteration stage names, in launch order (solver.hip launch_iteration)
STAGES = ("row_r2c_x", "col_fwd_bcast", "row_c2r_blur", "gather_freeze", "row_r2c_gather", "adjoint_image",
          "row_r2c_adj", "col_adj_accum", "col_inv_corr", "row_c2r_corr", "update", "row_c2r_update")


def _template_args(name: str) -> list[str]:
    m = re.search(r"<(.*?)>\(", name)
    return [a.strip() for a in m.group(1).split(",")] if m else []


def _row_r2c_gather(name: str) -> bool:
    # k_row_r2c<T, NT, VMAX, RB, MODE>: MODE 1 is the fused CSC gather; the
    # persistent k_row_r2c_p is the plain mode only
    return "k_row_r2c<" in name and _template_args(name)[-1] == "1"


def classify(name: str, prev: str | None) -> str | None:
    """Iteration stage of a dispatch, from the kernel name and the stage of
    the iteration dispatch before it, or None for other kernels."""
    if "k_row_r2c" in name:
        if _row_r2c_gather(name):
            return "row_r2c_gather"
        return "row_r2c_adj" if prev == "adjoint_image" else "row_r2c_x"
    if "k_row_c2r_upd" in name:
        return "row_c2r_update"
    if "k_row_c2r" in name:
        return "row_c2r_blur" if prev == "col_fwd_bcast" else "row_c2r_corr"
    if "k_col_e<" in name or "k_col<" in name:
        args = _template_args(name)
        mode = args[5] if "k_col_e<" in name else args[-1]
        # COL_INV (1) appears in the batched iteration only as the inverse of
        # the summed adjoint spectrum kept in FP64
        return {"1": "col_inv_corr", "2": "col_fwd_bcast", "3": "col_adj_accum", "4": "col_adj_accum"}.get(mode)
    if "k_gather<" in name:
        return "gather_freeze"
    if "k_csc_image<" in name or "k_csc_image16<" in name:
        return "adjoint_image"
    if "k_update<" in name:
        return "update"
    return None


def _find(d: Path, suffix: str) -> Path:
    hits = sorted(d.rglob(f"*{suffix}"))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits[0]


def _iteration_stages(rows: list[dict], key_name: str) -> list[tuple[str, dict]]:
    """(stage, row) of the dispatches from the first iteration on, in
    dispatch order: the plain row R2C that precedes the first gather opens
    the window."""
    rows = sorted(rows, key=lambda r: int(r["Dispatch_Id"]))
    first = next((i for i, r in enumerate(rows) if "k_gather<" in r[key_name]), None)
    if first is None:
        return []
    start = first
    for i in range(first, -1, -1):
        nm = rows[i][key_name]
        if "k_row_r2c" in nm and not _row_r2c_gather(nm):
            start = i
            break
    out, prev = [], None
    for r in rows[start:]:
        st = classify(r[key_name], prev)
        if st is not None:
            out.append((st, r))
            prev = st
    return out


def trace_summary(d: Path, cmds: list[str]) -> str:
    path = _find(d, "kernel_trace.csv")
    rows = list(csv.DictReader(open(path)))
    per = defaultdict(list)
    names = {}
    for st, r in _iteration_stages(rows, "Kernel_Name"):
        per[st].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        names[st] = r["Kernel_Name"].split("(")[0]
    total = sum(sum(v) for v in per.values())
    out = ["# rocprofv3 kernel trace: iteration kernels", ""]
    out += [f"command: {c}" for c in cmds]
    out += [f"source: {path.name} (dispatches from the first iteration on)", "",
            f"{'stage':16s} {'calls':>5s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'share':>6s}  kernel"]
    for st in STAGES:
        v = per.get(st)
        if not v:
            continue
        out.append(f"{st:16s} {len(v):5d} {statistics.fmean(v):9.1f} {min(v):9.1f} {max(v):9.1f} "
                   f"{100 * sum(v) / total:5.1f}%  {names[st]}")
    # every iteration opens with the row R2C of x
    n_it = len(per.get("row_r2c_x", [])) or 1
    out += ["", f"iterations: {n_it}; iteration kernels: {total / n_it / 1e3:.3f} ms per iteration"]
    return "\n".join(out) + "\n"


def _counter_means(d: Path, counter: str) -> tuple[dict, float]:
    """Per-stage mean of `counter` (KiB per dispatch) and the copy kernel's."""
    path = _find(d, "counter_collection.csv")
    rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
    copy = [float(r["Counter_Value"]) for r in rows if "k_copy_ref" in r["Kernel_Name"]]
    if not copy:
        raise SystemExit(f"{path}: no k_copy_ref dispatch to calibrate {counter}")
    per = defaultdict(list)
    for st, r in _iteration_stages(rows, "Kernel_Name"):
        per[st].append(float(r["Counter_Value"]))
    return {k: statistics.fmean(v) for k, v in per.items()}, statistics.fmean(copy)


def traffic_summary(fetch_dir: Path, write_dir: Path, copy_bytes: int, cmds: list[str]) -> str:
    f, fc = _counter_means(fetch_dir, "FETCH_SIZE")
    w, wc = _counter_means(write_dir, "WRITE_SIZE")
    kf = copy_bytes / (fc * 1024.0)
    kw = copy_bytes / (wc * 1024.0)
    out = ["# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: HBM bytes per dispatch of the iteration kernels", ""]
    out += [f"command: {c}" for c in cmds]
    out += ["", f"calibration: k_copy_ref reads and writes {copy_bytes} B per dispatch; FETCH_SIZE reads "
                f"{fc:.1f} KiB -> scale {kf:.4f}, WRITE_SIZE reads {wc:.1f} KiB -> scale {kw:.4f}", ""]
    for st in STAGES:
        if st not in f or st not in w:
            continue
        fb, wb = f[st] * 1024.0 * kf, w[st] * 1024.0 * kw
        out.append(f"stage {st} fetch_bytes {fb:.0f} write_bytes {wb:.0f} traffic_bytes {fb + wb:.0f}")
    return "\n".join(out) + "\n"


def read_traffic(path: Path) -> dict:
    """stage -> HBM bytes per dispatch, from a traffic summary."""
    res = {}
    for line in Path(path).read_text().splitlines():
        p = line.split()
        if len(p) == 8 and p[0] == "stage" and p[6] == "traffic_bytes":
            res[p[1]] = dict(fetch=int(p[3]), write=int(p[5]), total=int(p[7]))
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir", type=Path)
    t.add_argument("--cmd", action="append", default=[])
    p = sub.add_parser("traffic")
    p.add_argument("fetch_dir", type=Path)
    p.add_argument("write_dir", type=Path)
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--cmd", action="append", default=[])
    a = ap.parse_args(argv)
    if a.what == "trace":
        sys.stdout.write(trace_summary(a.dir, a.cmd))
    else:
        sys.stdout.write(traffic_summary(a.fetch_dir, a.write_dir, a.copy_bytes, a.cmd))
    return 0


if __name__ == "__main__":
    sys.exit(main())
