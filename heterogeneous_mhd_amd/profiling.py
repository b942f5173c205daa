"""Summaries of rocprofv3 output for the MHD kernels, and the copy kernel
that calibrates the HBM byte counters.

    python -m heterogeneous_mhd_amd.profiling calib [--gib 4]
        copies a buffer of known size a few times (run under rocprofv3 --pmc
        to read what FETCH_SIZE / WRITE_SIZE report for known bytes)
    python -m heterogeneous_mhd_amd.profiling stats run_kernel_stats.csv
        per-kernel calls / mean / min / max duration table
    python -m heterogeneous_mhd_amd.profiling traffic --fetch F.csv --write W.csv
            --calib-fetch CF.csv --calib-write CW.csv --gib 4 --cells N
        calibrated HBM bytes per dispatch and per cell of each kernel; the last
        line is 'json: {...}' with bytes per launch keyed by operator

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch.  Each is scaled by
known_bytes / (counter * 1024) measured on the calibration copy, so the
figures are HBM bytes as a plain copy of known size defines them.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys

# kernel-name prefix -> operator whose C call launches it
_OPS = [
    (re.compile(r"\badvect_fluid_kernel<\d+, \d+, false, false, \d+, \w+, 2>"), "advect2_fluidx"),
    (re.compile(r"\badvect_fluid_kernel<\d+, \d+, false, true, \d+, \w+, 2>"), "advect2_fluidx fast"),
    (re.compile(r"\badvect_fluid_kernel<\d+, \d+, false, false[,>]"), "advect_fluidx"),
    (re.compile(r"\badvect_fluid_kernel<\d+, \d+, false, true[,>]"), "advect_fluidx fast"),
    (re.compile(r"\badvect_fluid_kernel<\d+, \d+, true, false[,>]"), "advect_fluidx+calcfl"),
    (re.compile(r"\badvect_fluid_kernel<\d+, \d+, true, true[,>]"), "advect_fluidx+calcfl fast"),
    (re.compile(r"\bfluidx_kernel<\d+, \d+, false, false[,>]"), "fluidx"),
    (re.compile(r"\bfluidx_kernel<\d+, \d+, false, true[,>]"), "fluidx fast"),
    (re.compile(r"\bfluidx_kernel<\d+, \d+, true, false[,>]"), "fluidx+calcfl"),
    (re.compile(r"\bfluidx_kernel<\d+, \d+, true, true[,>]"), "fluidx+calcfl fast"),
    (re.compile(r"\badvect_kernel<\d+, \d+, false[,>]"), "advectbyzx"),
    (re.compile(r"\badvect_kernel<\d+, \d+, true[,>]"), "advectbyzx fast"),
    (re.compile(r"\badvect_transpose_kernel<\d+, false, 12>"), "advect_transpose12"),
    (re.compile(r"\badvect_transpose_kernel<\d+, true, 12>"), "advect_transpose12 fast"),
    (re.compile(r"\badvect_transpose_kernel<\d+, false, 13>"), "advect_transpose13"),
    (re.compile(r"\badvect_transpose_kernel<\d+, true, 13>"), "advect_transpose13 fast"),
    (re.compile(r"\btranspose_advect_kernel<\d+, false, 12>"), "transpose12_advect"),
    (re.compile(r"\btranspose_advect_kernel<\d+, true, 12>"), "transpose12_advect fast"),
    (re.compile(r"\btranspose_advect_kernel<\d+, false, 13>"), "transpose13_advect"),
    (re.compile(r"\btranspose_advect_kernel<\d+, true, 13>"), "transpose13_advect fast"),
    (re.compile(r"\bstream_copy_kernel"), "copy_matrix (streaming copy)"),
    (re.compile(r"\btranspose_kernel(?!<)"), "transpose (u and b)"),
    (re.compile(r"\btranspose_kernel<5>"), "transpose (u, 5 comps)"),
    (re.compile(r"\btranspose_kernel<3>"), "transpose (b, 3 comps)"),
    (re.compile(r"\bcfl_kernel"), "calcfl"),
    (re.compile(r"\bcfl_fold_kernel"), "calcfl fold"),
    (re.compile(r"\bbegin_step_kernel"), "begin_step"),
    (re.compile(r"copyBuffer"), "copy"),
    (re.compile(r"ens_fluid_kernel<\d+, \d+, \d+, true"), "ensemble fluidx+calcfl"),
    (re.compile(r"ens_fluid_kernel"), "ensemble fluidx"),
    (re.compile(r"ens_advect_kernel"), "ensemble advectbyzx"),
    (re.compile(r"ens_fused_kernel<\d+, \d+, \d+, true"), "ensemble advect_fluidx+calcfl"),
    (re.compile(r"ens_fused_kernel"), "ensemble advect_fluidx"),
    (re.compile(r"ens_advect_transpose_kernel<\w+, 12>"), "ensemble advect_transpose12"),
    (re.compile(r"ens_advect_transpose_kernel<\w+, 13>"), "ensemble advect_transpose13"),
    (re.compile(r"ens_transpose_advect_kernel<\w+, 12>"), "ensemble transpose12_advect"),
    (re.compile(r"ens_transpose_advect_kernel<\w+, 13>"), "ensemble transpose13_advect"),
    (re.compile(r"ens_transpose_kernel<5>"), "ensemble transpose (u)"),
    (re.compile(r"ens_transpose_kernel<3>"), "ensemble transpose (b)"),
    (re.compile(r"ens_cfl_fold_kernel"), "ensemble calcfl fold"),
    (re.compile(r"ens_begin_kernel"), "ensemble begin_step"),
    (re.compile(r"\bconv_first_kernel"), "corr_conv first layer (GELU)"),
    (re.compile(r"\bconv_kernel<0"), "corr_conv hidden (GELU)"),
    (re.compile(r"\bconv_kernel<1"), "corr_conv output"),
    (re.compile(r"\bconv_kernel<2"), "corr_conv input gradient"),
    (re.compile(r"\bconv_dw_kernel"), "corr_conv_dw"),
    (re.compile(r"\bdw_reduce_kernel"), "corr_dw_reduce"),
    (re.compile(r"\bpack_kernel"), "corr_pack"),
    (re.compile(r"\bpack_features_kernel"), "corr_pack_features"),
    (re.compile(r"\b(ysum|ymean|apply|check|decide|commit)_kernel"), "corr_apply"),
    (re.compile(r"\bweight_prep_kernel"), "corr_weight_prep"),
    (re.compile(r"\bic_pass\d"), "ic"),
]


def short_name(name: str) -> str:
    s = name.replace("(anonymous namespace)::", "").replace("mhd::", "").replace("void ", "")
    s = s.replace("at::native::", "")
    depth, out = 0, []
    for ch in s:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0 and out:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out)[:60]


def op_of(name: str) -> str | None:
    for rx, op in _OPS:
        if rx.search(name):
            return op
    return None


def calib_copy(gib: float, reps: int = 3) -> int:
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return 4 * n


def read_counters(path: str, counter: str):
    """kernel name -> list of per-dispatch values of one counter."""
    d = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def _largest_copy(d):
    vals = [v for k, vs in d.items() if "copyBuffer" in k for v in vs]
    if not vals:
        raise SystemExit("no copyBuffer dispatch in the calibration run")
    return max(vals)


def stats(path: str) -> None:
    rows = list(csv.DictReader(open(path)))
    print(f"{'kernel':60s} {'operator':28s} {'calls':>6s} {'mean_ms':>9s} {'min_ms':>9s} "
          f"{'max_ms':>9s} {'total_%':>7s}")
    for r in rows:
        print(f"{short_name(r['Name']):60s} {str(op_of(r['Name'])):28s} {int(r['Calls']):6d} "
              f"{float(r['AverageNs']) / 1e6:9.3f} {float(r['MinNs']) / 1e6:9.3f} "
              f"{float(r['MaxNs']) / 1e6:9.3f} {float(r['Percentage']):7.2f}")


def traffic(a) -> None:
    known = int(a.gib * (1 << 30)) // 4 * 4
    f_cal = known / (_largest_copy(read_counters(a.calib_fetch, "FETCH_SIZE")) * 1024)
    w_cal = known / (_largest_copy(read_counters(a.calib_write, "WRITE_SIZE")) * 1024)
    fetch = read_counters(a.fetch, "FETCH_SIZE")
    write = read_counters(a.write, "WRITE_SIZE")
    print(f"calibration copy of {known} bytes: FETCH_SIZE x {f_cal:.4f}, WRITE_SIZE x {w_cal:.4f}")
    print(f"cells per dispatch: {a.cells}")
    print(f"{'kernel':60s} {'n':>3s} {'fetch_KiB':>12s} {'write_KiB':>12s} {'read_B/cell':>11s} "
          f"{'write_B/cell':>12s} {'total_B/cell':>12s}")
    out = {}
    for k in sorted(set(fetch) | set(write)):
        op = op_of(k)
        if op is None or op in ("copy", "begin_step", "calcfl fold"):
            continue
        fk = sum(fetch.get(k, [0])) / max(1, len(fetch.get(k, [])))
        wk = sum(write.get(k, [0])) / max(1, len(write.get(k, [])))
        rb, wb = fk * 1024 * f_cal, wk * 1024 * w_cal
        print(f"{short_name(k):60s} {len(fetch.get(k, [])):3d} {fk:12.5g} {wk:12.5g} "
              f"{rb / a.cells:11.2f} {wb / a.cells:12.2f} {(rb + wb) / a.cells:12.2f}")
        out[op] = {"kernel": short_name(k), "bytes_per_launch": rb + wb, "read": rb, "write": wb}
    print("json: " + json.dumps({"cells": a.cells, "fetch_factor": f_cal, "write_factor": w_cal,
                                 "ops": out}))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("calib")
    c.add_argument("--gib", type=float, default=4.0)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    t = sub.add_parser("traffic")
    for o in ("--fetch", "--write", "--calib-fetch", "--calib-write"):
        t.add_argument(o, required=True)
    t.add_argument("--gib", type=float, default=4.0)
    t.add_argument("--cells", type=int, required=True)
    a = ap.parse_args(argv)
    if a.cmd == "calib":
        print(f"copied {calib_copy(a.gib)} bytes x 3", flush=True)
    elif a.cmd == "stats":
        stats(a.csv)
    else:
        traffic(a)


if __name__ == "__main__":
    sys.exit(main())
