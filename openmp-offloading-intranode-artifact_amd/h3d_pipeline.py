#!/usr/bin/env python3
"""Timing rows and summaries in the artifact's benchmark-pipeline formats.

The artifact's HIP tree times every binary with run.sh: one timings.csv row
per run (3D_Heat_Benchmark/AMD-MI250X/Source-Results/HIP/run.sh:70,196-232)
and the solver's stdout logged under a '>>> Running:' header (run.sh:128,
200).  aggregate.sh (:31-113) then merges the run logs into summary.csv,
one row per (variant, gpus, N) with the min / median / max of the solver's
own 'Time = S seconds' and the step count it reported.  This module writes
the same three files for the h3d program, so MI355X rows sit beside the
artifact's MI250X variants, and aggregates run logs from either side.

  python h3d_pipeline.py run [--sizes 512 640 ...] [--reps 3] [--out DIR]
  python h3d_pipeline.py aggregate results_*/run.log [-o summary.csv]
  python h3d_pipeline.py gamma summary.csv [--exe 1-hip-h3d] [--output F]

`gamma` is the machine constant of the artifact's performance model
(3D_Heat_Benchmark/*/Source-Results/model/compute_gamma.py): per grid of one
variant, gamma = T1 / (steps (N-2)^3) seconds per point-update, and its
plateau, the mean over the largest grids.

The run case of size N is the artifact's: `h3d N T` with T = STEPS * dt
printed %.7g, dt = h^2/(12 kappa_max), h = 1/(N-1), kappa_max = 0.95 and
STEPS = 500 (submit.sh:10-17,72-75), WRITE_TXT=0 (run.sh:52).
"""
import argparse
import csv
import io
import math
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
H3D_BIN = os.path.join(HERE, "bin", "h3d")

# run.sh:70 and aggregate.sh:91
TIMINGS_HEADER = ("rep,job_id,variant,family,gpus,rocr_mask,N,dt,steps,"
                  "wall_sec,reported_sec,status")
SUMMARY_HEADER = ("variant,gpus,N,samples,failed,min_sec,median_sec,max_sec,"
                  "rel_spread_pct,steps")
N_LIST = (512, 640, 768, 896, 1024, 1152, 1280)  # submit.sh:12

_RUNNING = re.compile(r"^>>> Running:")
_ESTIMATED = re.compile(r"Estimated\s+timesteps:")
_COMPLETE = re.compile(r"[0-9]+\s+timesteps?\s+complete")
_TIME = re.compile(r"^\s*Time\s*=")
_REPORTED = re.compile(r"^\s*Time\s*=\s*([0-9.eE+-]+)\s*seconds")


def sim_time(n, steps=500, kappa_max=0.95):
    """The T argument of size n: STEPS * h^2/(12 kappa_max), %.7g
    (submit.sh:72-75)."""
    h = 1.0 / (n - 1)
    dt = h * h / (12.0 * kappa_max)
    return "%.7g" % (steps * dt)


def _num(tok):
    """awk's string-to-number conversion of a field: the longest leading
    decimal number, 0 if there is none."""
    m = re.match(r"\s*[-+]?(\d+\.?\d*([eE][-+]?\d+)?|\.\d+([eE][-+]?\d+)?)", tok)
    return float(m.group(0)) if m else 0.0


def _fmt_num(x):
    """awk's output of a number through %s: integers print as integers,
    other values as %.6g."""
    return str(int(x)) if x == int(x) and abs(x) < 1e16 else "%.6g" % x


def parse_run_log(lines):
    """Timed runs in run-log lines, as aggregate.sh's first pass reads them
    (aggregate.sh:31-72): a '>>> Running:' line names the variant (a token
    matching ^[0-9]+-hip once a leading './' is dropped), gpus= and N=; the
    step count is the solver's 'N timesteps complete', else its 'Estimated
    timesteps: ~N'; the 'Time = S seconds' line closes the run.  Returns
    (variant, gpus, N, seconds, steps) tuples of strings, steps '' when the
    log gave none."""
    out = []
    variant = gpus = size = steps_done = steps_est = ""
    armed = False
    for line in lines:
        line = line.rstrip("\n")
        f = line.split()
        if _RUNNING.search(line):
            variant = gpus = size = steps_done = steps_est = ""
            armed = False
            for tok in f:
                t = tok[2:] if tok.startswith("./") else tok
                if re.match(r"^[0-9]+-hip", t):
                    variant = t
                if tok.startswith("gpus="):
                    gpus = tok.split("=")[1]
                if tok.startswith("N="):
                    size = tok.split("=")[1]
            armed = variant != "" and gpus != "" and size != ""
            continue
        if not armed:
            continue
        if _ESTIMATED.search(line):
            s = f[-1]
            steps_est = _fmt_num(_num(s[1:] if s.startswith("~") else s))
            continue
        if _COMPLETE.search(line):
            steps_done = _fmt_num(_num(f[0]))
            continue
        if _TIME.search(line):
            for i, tok in enumerate(f):
                if tok == "=" and i + 1 < len(f):
                    steps = steps_done if steps_done != "" else steps_est
                    out.append((variant, gpus, size, f[i + 1], steps))
                    armed = False
                    break
    return out


def _median(a):
    a = sorted(a)
    n = len(a)
    m = (n + 1) // 2
    return a[m - 1] if n % 2 else (a[m - 1] + a[m]) / 2.0


def summarize(records):
    """summary.csv lines (header first) from parse_run_log records, as
    aggregate.sh's second pass (:80-113): per (variant, gpus, N) the sample
    count, min / median / max seconds (%.6f), the spread 100 (max - min) /
    median (%.1f) and the last step count seen; rows ordered by variant,
    then N numerically (sort -t, -k1,1 -k3,3n; variants compared bytewise)."""
    groups, order = {}, []
    for variant, gpus, size, sec, steps in records:
        key = (variant, gpus, size)
        if key not in groups:
            groups[key] = {"v": [], "steps": ""}
            order.append(key)
        groups[key]["v"].append(_num(sec))
        if steps != "":
            groups[key]["steps"] = steps
    rows = []
    for key in order:
        g = groups[key]
        v = sorted(g["v"])
        med = _median(v)
        spread = 100.0 * (v[-1] - v[0]) / med if med > 0 else 0.0
        rows.append("%s,%s,%s,%d,0,%.6f,%.6f,%.6f,%.1f,%s" % (
            key[0], key[1], key[2], len(v), v[0], med, v[-1], spread, g["steps"]))
    rows.sort(key=lambda r: (r.split(",")[0].encode(), _num(r.split(",")[2]), r.encode()))
    return [SUMMARY_HEADER] + rows


GAMMA_FIELDS = ("N", "interior_points", "timesteps", "T1_seconds", "gamma_s_per_point")


def gamma_rows(summary_lines, variant="1-omp", time_column="median_sec"):
    """Machine constant per grid from summary.csv lines, as
    compute_gamma.py's read_summary (:124-186): rows of `variant` with an
    integer N > 2, steps > 0 and a finite positive `time_column`; gamma =
    T1 / (steps (N-2)^3); sorted by N.  Other rows are skipped."""
    out = []
    for row in csv.DictReader(summary_lines):
        if (row.get("variant") or "").strip() != variant:
            continue
        try:
            n, steps = int(row["N"]), int(row["steps"])
            t1 = float(row[time_column])
        except (KeyError, TypeError, ValueError):
            continue
        if n <= 2 or steps <= 0 or not math.isfinite(t1) or t1 <= 0:
            continue
        interior = (n - 2) ** 3
        out.append({"N": n, "interior_points": interior, "timesteps": steps,
                    "T1_seconds": t1, "gamma_s_per_point": t1 / (steps * interior)})
    out.sort(key=lambda r: r["N"])
    return out


def gamma_csv(rows):
    """gamma_single_gpu.csv text of gamma rows (compute_gamma.py:203-219:
    the csv module's writer, CRLF lines, numbers as Python prints them)."""
    buf = io.StringIO(newline="")
    w = csv.DictWriter(buf, fieldnames=GAMMA_FIELDS)
    w.writeheader()
    w.writerows(rows)
    return buf.getvalue()


def gamma_plateau(rows, tail=3):
    """Mean gamma of the `tail` largest grids (compute_gamma.py:314-318)."""
    last = rows[-min(tail, len(rows)):]
    return sum(r["gamma_s_per_point"] for r in last) / len(last)


def run_case(binary, variant, n, steps=500, rep=1, job_id="local", timeout=1800,
             extra_args=(), gpus=1, mask="0"):
    """One timed run of `binary n T`, as run.sh's run_bench (:196-232).
    Returns (timings_row, run_log_text): the row in timings.csv's schema
    (wall seconds %.4f around the process, the solver's reported seconds or
    N/A, status ok / timeout / runtime_error) and the log block, a
    '>>> Running:' header followed by the program's stdout and stderr."""
    T = sim_time(n, steps)
    env = dict(os.environ, WRITE_TXT="0")
    cmd = [binary, str(n), T, *extra_args]
    t0 = time.monotonic()
    try:
        r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           timeout=timeout)
        out = r.stdout.decode(errors="replace")
        status = "ok" if r.returncode == 0 else "runtime_error"
    except subprocess.TimeoutExpired as e:
        out = (e.stdout or b"").decode(errors="replace")
        status = "timeout"
    wall = "%.4f" % (time.monotonic() - t0)
    found = [m.group(1) for m in map(_REPORTED.match, out.splitlines()) if m]
    reported = found[-1] if found else "N/A"
    row = ",".join(str(x) for x in (rep, job_id, variant, f"{gpus}gpu", gpus, mask, n, T,
                                    steps, wall, reported, status))
    log = (f">>> Running: {variant}  gpus={gpus}  mask={mask}  N={n}\n" + out
           + ("" if out.endswith("\n") or not out else "\n"))
    return row, log


def cmd_run(a):
    if not os.access(a.bin, os.X_OK):
        sys.exit(f"{a.bin} is not an executable: build it first (make -C {HERE})")
    out = a.out or time.strftime(f"results_{a.job_id}_%Y%m%d_%H%M%S")
    os.makedirs(out, exist_ok=True)
    tpath, lpath = os.path.join(out, "timings.csv"), os.path.join(out, "run.log")
    with open(tpath, "w") as ft, open(lpath, "w") as fl:
        ft.write(TIMINGS_HEADER + "\n")
        for n in a.sizes:
            for rep in range(1, a.reps + 1):
                row, log = run_case(a.bin, a.variant, n, a.steps, rep, a.job_id, a.timeout,
                                    a.arg)
                ft.write(row + "\n")
                fl.write(log)
                ft.flush()
                fl.flush()
                print(row, flush=True)
    with open(lpath) as f:
        summary = summarize(parse_run_log(f))
    with open(os.path.join(out, "summary.csv"), "w") as f:
        f.write("\n".join(summary) + "\n")
    print("\n".join(summary))


def cmd_aggregate(a):
    lines = []
    for path in a.logs:
        with open(path, errors="replace") as f:
            lines.extend(f)
    summary = "\n".join(summarize(parse_run_log(lines))) + "\n"
    if a.o:
        with open(a.o, "w") as f:
            f.write(summary)
    else:
        sys.stdout.write(summary)


def cmd_gamma(a):
    with open(a.summary, newline="", encoding="utf-8") as f:
        rows = gamma_rows(f, a.exe, a.time_column)
    if not rows:
        sys.exit(f"No usable rows found for variant {a.exe!r} in {a.summary}")
    print(f"{'N':>6} {'interior=(n-2)^3':>18} {'steps':>7} {'T1 [s]':>12} "
          f"{'gamma [s/pt]':>15} {'gamma [ps/pt]':>15}")
    for r in rows:
        g = r["gamma_s_per_point"]
        print(f"{r['N']:>6} {r['interior_points']:>18} {r['timesteps']:>7} "
              f"{r['T1_seconds']:>12.6f} {g:>15.4e} {g * 1e12:>15.3f}")
    last = rows[-min(a.tail, len(rows)):]
    p = gamma_plateau(rows, a.tail)
    print(f"Plateau gamma (mean of N = {', '.join(str(r['N']) for r in last)}): "
          f"{p:.4e} s/point-update ({p * 1e12:.2f} ps/point-update)")
    out = a.output or os.path.join(os.path.dirname(os.path.abspath(a.summary)),
                                   "gamma_single_gpu.csv")
    with open(out, "w", newline="", encoding="utf-8") as f:
        f.write(gamma_csv(rows))
    print(f"Wrote {out}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run", help="time the h3d program at the artifact's sizes")
    r.add_argument("--bin", default=H3D_BIN)
    r.add_argument("--variant", default="1-hip-h3d",
                   help="name in the rows; aggregate.sh keys on ^[0-9]+-hip")
    r.add_argument("--sizes", type=int, nargs="+", default=list(N_LIST))
    r.add_argument("--reps", type=int, default=3)
    r.add_argument("--steps", type=int, default=500)
    r.add_argument("--job-id", default="local")
    r.add_argument("--timeout", type=float, default=1800)
    r.add_argument("--arg", action="append", default=[],
                   help="extra h3d option, repeatable (e.g. --arg=--graph)")
    r.add_argument("--out", default="")
    g = sub.add_parser("aggregate", help="summary.csv from run logs")
    g.add_argument("logs", nargs="+")
    g.add_argument("-o", default="")
    m = sub.add_parser("gamma", help="machine constant from a summary.csv")
    m.add_argument("summary")
    m.add_argument("--exe", default="1-omp", help="variant to select (default: 1-omp)")
    m.add_argument("--time-column", default="median_sec",
                   choices=("min_sec", "median_sec", "max_sec"))
    m.add_argument("--tail", type=int, default=3,
                   help="number of largest grids in the plateau mean")
    m.add_argument("--output", default="",
                   help="default: gamma_single_gpu.csv beside the summary")
    a = ap.parse_args(argv)
    if a.cmd == "gamma" and a.tail <= 0:
        ap.error("--tail must be positive")
    {"run": cmd_run, "aggregate": cmd_aggregate, "gamma": cmd_gamma}[a.cmd](a)


if __name__ == "__main__":
    main()
