#!/usr/bin/env python3
"""Write the golden fixtures of the reference-mode heat update.

Runs the reference's own serial program (OpenMP_Single_Source/
heat3D_variable_coeff.c, compiled from source into oracle/_ref/serial by
`make -C oracle ref`) in a scratch directory, with the same `n T` command
line the reference's pipeline uses (T = STEPS*dt printed with %.7g, as in
AMD-MI250X/Source-Results/OpenMP/submit.sh:61-62), and keeps:

  ref_n{n}_s{STEPS}.txt.gz   the u_serial.txt it writes (%.17g, k,j,i order)
  index.json                 n, T string, recorded step count, stdout lines

Only data leaves this script; it is run in the build container, where the
reference tree is mounted, never at test time.
"""
import gzip
import json
import os
import pathlib
import shutil
import subprocess
import sys
import tempfile

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent.parent
SERIAL = REPO / "oracle" / "_ref" / "serial"

CASES = [(5, 1), (5, 3), (5, 10), (9, 25), (17, 20), (33, 8)]


def t_arg(n, steps, kappa_max=0.95):
    h = 1.0 / (n - 1)
    dt = h * h / (12.0 * kappa_max)
    return "%.7g" % (steps * dt)


def main():
    if not SERIAL.exists():
        subprocess.run(["make", "-C", str(REPO / "oracle"), "ref"], check=True)
    index = []
    for n, steps in CASES:
        T = t_arg(n, steps)
        with tempfile.TemporaryDirectory() as td:
            env = dict(os.environ, WRITE_TXT="1", OMP_NUM_THREADS="2")
            out = subprocess.run([str(SERIAL), str(n), T], cwd=td, env=env,
                                 check=True, capture_output=True, text=True)
            name = f"ref_n{n}_s{steps}.txt.gz"
            with open(os.path.join(td, "u_serial.txt"), "rb") as src, \
                    gzip.GzipFile(HERE / name, "wb", mtime=0) as dst:
                shutil.copyfileobj(src, dst)
        lines = [ln for ln in out.stdout.splitlines()
                 if ln and not ln.startswith("Time =")]
        recorded = int(next(ln for ln in lines
                            if ln.endswith("timesteps complete")).split()[0])
        index.append({"n": n, "steps_arg": steps, "T": T,
                      "steps": recorded, "file": name, "stdout": lines})
        print(f"n={n} T={T} -> {recorded} steps -> {name}")
    (HERE / "index.json").write_text(json.dumps(index, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
