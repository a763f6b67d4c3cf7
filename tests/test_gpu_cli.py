"""The h3d program as a drop-in for the reference's `./bin n T` binaries:
same command line, the stdout lines the reference's pipeline parses
(aggregate.sh:31-40) and the %.17g solution file it verifies
(verify.sh:239-286)."""
import gzip
import json
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "openmp-offloading-intranode-artifact_amd", "bin", "h3d")
GOLDEN = os.path.join(REPO, "tests", "golden")
INDEX = json.load(open(os.path.join(GOLDEN, "index.json")))


def run(args, cwd, write=True, timeout=120):
    env = dict(os.environ, WRITE_TXT="1" if write else "0")
    r = subprocess.run([EXE] + args, cwd=cwd, env=env, capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("entry", [e for e in INDEX if e["n"] in (9, 33)],
                         ids=lambda e: f"n{e['n']}")
def test_cli_reference_mode_matches_golden(entry, tmp_path):
    out = run([str(entry["n"]), entry["T"]], str(tmp_path))
    assert re.search(rf"^{entry['steps']} timesteps complete$", out, re.M)
    assert re.search(r"^Time = [0-9.]+ seconds$", out, re.M)
    assert re.search(r"^Estimated timesteps:\s+~\d+$", out, re.M)
    with gzip.open(os.path.join(GOLDEN, entry["file"]), "rt") as f:
        want = f.read()
    assert (tmp_path / "u_h3d_1gpu.txt").read_text() == want


def test_cli_graph_equals_eager(tmp_path):
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run(["40", "--steps", "9", "--mode", "seeded", "--kappa", "lognormal"], str(a))
    run(["40", "--steps", "9", "--mode", "seeded", "--kappa", "lognormal", "--graph"], str(b))
    assert (a / "u_h3d_1gpu.txt").read_text() == (b / "u_h3d_1gpu.txt").read_text()


@pytest.mark.parametrize("prec", ["f32c", "f32c16"])
def test_cli_fp32_check_passes(prec, tmp_path):
    out = run(["66", "--steps", "200", "--mode", "seeded", "--precision", prec,
               "--kblock", "2", "--check"], str(tmp_path), write=False)
    m = re.search(r"vs FP64: rel_l2 = ([0-9.e+-]+)", out)
    assert m and float(m.group(1)) < 1e-5, out


def test_cli_bf16_kappa_runs(tmp_path):
    out = run(["50", "--steps", "30", "--mode", "seeded", "--precision", "f32c",
               "--kappa-bf16", "--check"], str(tmp_path), write=False)
    m = re.search(r"vs FP64: rel_l2 = ([0-9.e+-]+)", out)
    assert m and float(m.group(1)) < 1e-5, out


def test_cli_slab_runs_with_face_exchange(tmp_path):
    out = run(["34", "--steps", "5", "--slab", "4:1"], str(tmp_path))
    assert "Block sizes: nz0=10 nz1=10 nz2=10 nz3=10" in out
    vals = np.loadtxt(tmp_path / "u_h3d_slab1of4.txt")
    assert vals.size == 10 * 34 * 34 and np.all(np.isfinite(vals))
    v = vals.reshape(10, 34, 34)
    # ghost planes hold the z-periodic stand-in copies of the edge planes
    assert np.array_equal(v[0], v[8]) and np.array_equal(v[9], v[1])


def test_cli_autotune(tmp_path):
    out = run(["70", "--steps", "4", "--mode", "seeded", "--autotune"], str(tmp_path),
              write=False)
    assert re.search(r"^Autotuned: rows=\d+ waves=\d+ zchunk=\d+", out, re.M)
