#!/usr/bin/env python3
"""Writes the benchmark-pipeline fixtures from the reference's recorded MI250X
HIP results (run here, where the reference tree is present; the fixtures are
committed and the tests never read the reference):

  hip_mi250x_runlog.txt   the lines of every results_*/run.log that
                          aggregate.sh parses ('>>> Running:' headers,
                          'Estimated timesteps', 'N timesteps complete',
                          'Time = S seconds'), files in sorted order, each
                          introduced by a '# <dir>/run.log' line
  hip_mi250x_summary.csv  the summary.csv aggregate.sh wrote from them
  hip_mi250x_timings_header.txt  the header line of run.sh's timings.csv
  gamma_<machine>_summary.csv    the OpenMP summary.csv of each machine
  gamma_<machine>_single_gpu.csv the gamma_single_gpu.csv compute_gamma.py
                                 wrote from it (machine: mi250x, h100,
                                 max1550), byte for byte (CRLF lines)

Source: 3D_Heat_Benchmark/AMD-MI250X/Source-Results/HIP/{results_*/run.log,
summary.csv, results_*/timings.csv} and 3D_Heat_Benchmark/<machine dir>/
Source-Results/OpenMP/{summary.csv, gamma_single_gpu.csv}.
"""
import glob
import os
import re
import sys

SRC = "/root/reference/3D_Heat_Benchmark/AMD-MI250X/Source-Results/HIP"
GAMMA = {"mi250x": "AMD-MI250X", "h100": "NVIDIA-H100", "max1550": "Intel-1550"}
HERE = os.path.dirname(os.path.abspath(__file__))
KEEP = re.compile(r"^>>> Running:|Estimated\s+timesteps:|[0-9]+\s+timesteps?\s+complete"
                  r"|^\s*Time\s*=")


def main():
    logs = sorted(glob.glob(os.path.join(SRC, "results_*", "run.log")))
    if not logs:
        sys.exit(f"no run logs under {SRC}")
    out = []
    for path in logs:
        out.append("# " + os.path.relpath(path, SRC))
        with open(path, errors="replace") as f:
            out.extend(line.rstrip("\n") for line in f if KEEP.search(line))
    with open(os.path.join(HERE, "hip_mi250x_runlog.txt"), "w") as f:
        f.write("\n".join(out) + "\n")
    with open(os.path.join(SRC, "summary.csv")) as f:
        summary = f.read()
    with open(os.path.join(HERE, "hip_mi250x_summary.csv"), "w") as f:
        f.write(summary)
    timings = sorted(glob.glob(os.path.join(SRC, "results_*", "timings.csv")))
    with open(timings[0]) as f:
        header = f.readline()
    with open(os.path.join(HERE, "hip_mi250x_timings_header.txt"), "w") as f:
        f.write(header)
    print(f"{len(logs)} run logs, {len(out)} lines; summary {summary.count(chr(10))} lines")
    for tag, machine in GAMMA.items():
        d = f"/root/reference/3D_Heat_Benchmark/{machine}/Source-Results/OpenMP"
        for src, dst in (("summary.csv", f"gamma_{tag}_summary.csv"),
                         ("gamma_single_gpu.csv", f"gamma_{tag}_single_gpu.csv")):
            with open(os.path.join(d, src), "rb") as f:
                data = f.read()
            with open(os.path.join(HERE, dst), "wb") as f:
                f.write(data)
        print(f"gamma fixtures for {machine}")


if __name__ == "__main__":
    main()
