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


@pytest.mark.parametrize("kblock", [2, 3])
@pytest.mark.parametrize("entry", [e for e in INDEX if e["n"] in (9, 17, 33)],
                         ids=lambda e: f"n{e['n']}")
def test_cli_fused_fp64_passes_match_golden(entry, kblock, tmp_path):
    """`--kblock` fuses FP64 steps into passes (the reference fields carry
    rhs): the dump is still the reference's serial program's, byte for
    byte, including step counts that are not multiples of kblock."""
    out = run([str(entry["n"]), entry["T"], "--kblock", str(kblock), "--no-autotune"],
              str(tmp_path))
    assert re.search(rf"^{entry['steps']} timesteps complete$", out, re.M)
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


@pytest.mark.parametrize("mode,part", [("reference", 1), ("seeded", 0)])
def test_cli_deep_slab_runs_with_band_exchange(tmp_path, mode, part):
    """--slab with --kblock 2: two ghost planes per side, 2-step passes and
    one band exchange per pass; the first slab of a seeded run starts one
    plane below the domain (the stand-in repeats the boundary plane)."""
    out = run(["34", "--steps", "5", "--mode", mode, "--slab", f"4:{part}", "--kblock", "2"],
              str(tmp_path))
    assert "Ghost planes per side: 2" in out and "5 timesteps complete" in out
    vals = np.loadtxt(tmp_path / f"u_h3d_slab{part}of4.txt")
    assert vals.size == 12 * 34 * 34 and np.all(np.isfinite(vals))
    v = vals.reshape(12, 34, 34)
    # ghosts hold the z-periodic stand-in bands of owned planes
    assert np.array_equal(v[0:2], v[8:10]) and np.array_equal(v[10:12], v[2:4])


@pytest.mark.parametrize("prec,kblock", [("f64", 4), ("f32c", 5)])
def test_cli_fixed_depth_is_autotuned_on_large_grids(prec, kblock, tmp_path):
    """A --kblock the library accepts is searched at that depth when the
    program autotunes by default (from 128^3 interior points), and every
    shape it may pick gives the bits of the default shape."""
    args = ["130", "--steps", "11", "--mode", "seeded", "--precision", prec,
            "--kblock", str(kblock)]
    tuned = run(args, str(tmp_path), write=False)
    assert re.search(rf"^Autotuned: .*kblock={kblock}$", tuned, re.M), tuned
    plain = run(args + ["--no-autotune"], str(tmp_path), write=False)
    line = lambda out: next(ln for ln in out.splitlines() if ln.startswith("checksum ="))
    assert line(tuned) == line(plain)


def cli_deep_slab_initial(gpu, n, parts, part, ghost, seeded):
    """The initial slab the program builds for --slab parts:part with
    --kblock ghost (host/h3d_main.c): the reference's partition init with
    `ghost` ghost planes, or the seeded generators on the planes inside the
    domain with the nearest domain plane repeated beyond it; and the
    step's factor (seeded: from the global kappa bound)."""
    rc, nzd, gks = gpu.slab_decompose(n, parts)
    assert rc == 0
    nz = nzd[part] - 2 + 2 * ghost
    gk0 = gks[part] - ghost
    g = gpu.Grid.dense(n, n, nz)
    P = n * n
    if not seeded:
        h, dt, factor = gpu.params_reference(n)
        u, rhs, k = (np.zeros(nz * P) for _ in range(3))
        gpu.call("h3d_init_partition", gpu.ptr(u), gpu.ptr(rhs), gpu.ptr(k), nz, gk0 + ghost,
                 ghost, n, n, n, h, dt)
        return g, u, rhs, k, factor
    lo, hi = max(gk0, 0), min(gk0 + nz, n)
    gin = gpu.Grid.dense(n, n, hi - lo)
    ui, ki = gpu.gen_seeded(gin, lo, n, n, n, 0, 20260101)
    u, k = np.zeros(nz * P), np.zeros(nz * P)
    for z in range(nz):
        src = min(max(z + gk0, lo), hi - 1) - lo
        u[z * P:(z + 1) * P] = ui[src * P:(src + 1) * P]
        k[z * P:(z + 1) * P] = ki[src * P:(src + 1) * P]
    _, kg = gpu.gen_seeded(gpu.Grid.dense(n), 0, n, n, n, 0, 20260101)
    h = 1.0 / (n - 1)
    fmax = gpu.lib().h3d_face_sum_max_host(gpu.ptr(kg), gpu.C.byref(gpu.Grid.dense(n)))
    dt = 0.9 * 2.0 * h * h / fmax
    return g, u, None, k, dt / h / h / 2.0


@pytest.mark.parametrize("mode,part", [("reference", 1), ("seeded", 0), ("seeded", 3)])
def test_cli_deep_slab_equals_oracle(gpu, oracle, tmp_path, mode, part):
    """The program's deep-ghost slab (--slab 4:p --kblock 2) ends on the
    bits of the oracle stepping the same initial slab -- built on the host
    as the program builds it -- with the same band copies after every
    2-step pass."""
    from test_gpu_face_reduce import deep_standin_oracle
    n, ghost = 34, 2
    out = run([str(n), "--steps", "5", "--mode", mode, "--slab", f"4:{part}", "--kblock",
               str(ghost)], str(tmp_path))
    steps = int(re.search(r"^(\d+) timesteps complete$", out, re.M).group(1))
    g, u, rhs, k, factor = cli_deep_slab_initial(gpu, n, 4, part, ghost, mode == "seeded")
    want = deep_standin_oracle(oracle, u, rhs, k, g, factor, ghost, steps)
    got = np.loadtxt(tmp_path / f"u_h3d_slab{part}of4.txt")
    assert got.size == want.size and np.array_equal(got, want)


@pytest.mark.parametrize("prec,ghost,part", [("f32c", 1, 1), ("f32c", 3, 0), ("f32c16", 2, 3),
                                             ("f32", 2, 2)])
def test_cli_fp32_slab_equals_oracle(gpu, oracle, tmp_path, prec, ghost, part):
    """An FP32 slab (--slab 4:p --precision ... --kblock K): passes of K
    steps with bands of K FP32 planes (and FP16 residual planes in f32c16)
    exchanged after each; the dump equals the oracle stepping the same
    initial slab, rounded to FP32, with the same band copies, and --check
    compares against the FP64 slab with the same exchange."""
    from test_gpu_f32_slab import standin_oracle_f32
    n = 34
    args = [str(n), "--steps", "5", "--mode", "seeded", "--precision", prec,
            "--slab", f"4:{part}", "--check"]
    if ghost > 1:
        args += ["--kblock", str(ghost)]
    out = run(args, str(tmp_path))
    assert re.search(r"^vs FP64: .* PASS$", out, re.M), out
    steps = int(re.search(r"^(\d+) timesteps complete$", out, re.M).group(1))
    g, u, _, k, factor = cli_deep_slab_initial(gpu, n, 4, part, ghost, True)
    F = dict(g=g, u=u.astype(np.float32), k=k.astype(np.float32), bits=None,
             factor=float(np.float32(factor)))
    want, _ = standin_oracle_f32(oracle, F, g, ghost, steps, {"f32": 0, "f32c": 1,
                                                             "f32c16": 2}[prec])
    got = np.loadtxt(tmp_path / f"u_h3d_slab{part}of4.txt")
    assert got.size == want.size and np.array_equal(got, want.astype(np.float64))


@pytest.mark.parametrize("prec,part", [("f32c", 0), ("f32c16", 2)])
def test_cli_fp32_slab_bf16_kappa_takes_the_field_scale(gpu, oracle, tmp_path, prec, part):
    """--kappa-bf16 on a slab: kappa is stored with the whole field's BF16
    scale (every slab of the split the same one, as the whole field run
    has), and the dump equals the oracle's stand-in run on those bits."""
    from test_gpu_f32_slab import standin_oracle_f32
    n, ghost = 34, 2
    out = run([str(n), "--steps", "5", "--mode", "seeded", "--precision", prec, "--kappa-bf16",
               "--slab", f"4:{part}", "--kblock", str(ghost)], str(tmp_path))
    steps = int(re.search(r"^(\d+) timesteps complete$", out, re.M).group(1))
    g, u, _, k, factor = cli_deep_slab_initial(gpu, n, 4, part, ghost, True)
    _, kg = gpu.gen_seeded(gpu.Grid.dense(n), 0, n, n, n, 0, 20260101)
    _, scale = gpu.quantize_bf16(kg)
    F = dict(g=g, u=u.astype(np.float32), k=None, bits=gpu.quantize_bf16_scaled(k, scale),
             factor=float(np.float32(factor * scale)))
    want, _ = standin_oracle_f32(oracle, F, g, ghost, steps, {"f32c": 1, "f32c16": 2}[prec])
    got = np.loadtxt(tmp_path / f"u_h3d_slab{part}of4.txt")
    assert got.size == want.size and np.array_equal(got, want.astype(np.float64))


@pytest.mark.parametrize("prec", ["f64", "f32c16"])
def test_cli_slab_autotune_keeps_bits(tmp_path, prec):
    """--autotune on a slab tunes the interior on the slab pass itself
    (h3d_autotune_slab) and changes no bit of the result."""
    args = ["34", "--steps", "5", "--mode", "seeded", "--precision", prec, "--slab", "4:1",
            "--kblock", "2"]
    tuned = run(args + ["--autotune"], str(tmp_path), write=False)
    assert re.search(r"^Autotuned: .*kblock=2$", tuned, re.M), tuned
    plain = run(args + ["--no-autotune"], str(tmp_path), write=False)
    line = lambda out: next(ln for ln in out.splitlines() if ln.startswith("checksum ="))
    assert line(tuned) == line(plain)


def test_cli_autotune(tmp_path):
    out = run(["70", "--steps", "4", "--mode", "seeded", "--autotune"], str(tmp_path),
              write=False)
    assert re.search(r"^Autotuned: rows=\d+ waves=\d+ zchunk=\d+", out, re.M)


VERIFY = os.path.join(REPO, "openmp-offloading-intranode-artifact_amd", "bin", "h3d-verify")


def verify(args, cwd):
    return subprocess.run([VERIFY] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=120)


def compare_line(label, name, m):
    """The metrics line of compare.py:101-103 for oracle metrics m."""
    ok = m[6] == 0 and m[7] == 0 and m[8] == 0
    return (f"  [{'PASS' if ok else 'FAIL'} ] {label:<22} {name:<34} "
            f"maxabs={m[0]:.3e}  rms={m[2]:.3e}  relL2={m[3]:.3e}  maxrel={m[4]:.3e}")


def test_verify_passes_cli_dump_against_reference_dump(tmp_path):
    """The pipeline's verify step: the reference serial program's dump as
    u_serial.txt, h3d's own dump found by name, bit-identical -> PASS."""
    entry = next(e for e in INDEX if e["n"] == 33)
    with gzip.open(os.path.join(GOLDEN, entry["file"]), "rt") as f:
        (tmp_path / "u_serial.txt").write_text(f.read())
    run([str(entry["n"]), entry["T"]], str(tmp_path))
    r = verify([], str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Reference: u_serial.txt  ({33 ** 3} values)" in r.stdout
    assert "Tolerance: atol=1e-06, rtol=1e-05" in r.stdout
    assert ("  [PASS ] 1-GPU h3d              u_h3d_1gpu.txt                     "
            "maxabs=0.000e+00  rms=0.000e+00  relL2=0.000e+00  maxrel=0.000e+00") in r.stdout
    assert "Summary: 1/1 passed (atol=1e-06, rtol=1e-05)" in r.stdout


def test_verify_metrics_and_failures_match_oracle(oracle, tmp_path):
    n = 17
    rng = np.random.default_rng(5)
    ref = rng.standard_normal(n ** 3)
    ref[::40] = 0.0
    good = ref * (1 + 1e-7 * rng.standard_normal(ref.size))
    bad = ref + 1e-3 * rng.standard_normal(ref.size)
    nan = good.copy()
    nan[100] = np.nan
    for name, a in (("ref.txt", ref), ("good.txt", good), ("bad.txt", bad), ("nan.txt", nan),
                    ("short.txt", ref[:-1])):
        (tmp_path / name).write_text("".join("%.17g\n" % v for v in a))
    r = verify(["--ref", "ref.txt", "--atol", "1e-6", "--rtol", "1e-5", "good.txt",
                "bad.txt", "nan.txt", "short.txt", "missing.txt"], str(tmp_path))
    assert r.returncode == 1, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    for name, a in (("good.txt", good), ("bad.txt", bad), ("nan.txt", nan)):
        m = oracle.metrics(ref, a, 1e-6, 1e-5)
        assert compare_line(name, name, m) in lines, (name, r.stdout)
    m = oracle.metrics(ref, bad, 1e-6, 1e-5)
    i = int(m[5])
    assert f"           worst @ idx {i}: ref={ref[i]:.6e}  cand={bad[i]:.6e}" in lines
    assert any(ln.endswith("NaN=1 Inf=0") for ln in lines)
    assert (f"  [FAIL ] {'short.txt':<22} {'short.txt':<34} size mismatch: "
            f"ref={n ** 3} cand={n ** 3 - 1}") in lines
    assert "Summary: 1/4 passed (atol=1e-06, rtol=1e-05)" in r.stdout
    # nothing to compare -> 2
    r = verify(["--ref", "ref.txt", "missing.txt"], str(tmp_path))
    assert r.returncode == 2 and "no candidate GPU output files found" in r.stdout


def test_cli_autotunes_large_grids_with_identical_results(tmp_path):
    """From 128^3 interior points the program tunes its launch shapes by
    default; every shape computes the same bits, so the checksum line is
    the same as with --no-autotune."""
    args = ["130", "--steps", "20", "--mode", "seeded"]
    tuned = run(args, str(tmp_path), write=False)
    plain = run(args + ["--no-autotune"], str(tmp_path), write=False)
    assert "Autotuned:" in tuned and "Autotuned:" not in plain
    line = lambda out: next(ln for ln in out.splitlines() if ln.startswith("checksum ="))
    assert line(tuned) == line(plain)


def test_pipeline_rows_from_h3d(tmp_path):
    """The program timed the way the artifact's run.sh times its binaries:
    timings.csv rows with status ok and the solver's reported seconds, and
    a summary whose step counts are the step rule's for each (n, T)."""
    import h3d_pipeline as P
    import h3d_py as h3d
    out = tmp_path / "res"
    P.main(["run", "--sizes", "34", "66", "--reps", "2", "--out", str(out)])
    rows = [r.split(",") for r in (out / "timings.csv").read_text().splitlines()[1:]]
    assert len(rows) == 4
    assert all(r[11] == "ok" and float(r[10]) > 0 for r in rows), rows
    summary = (out / "summary.csv").read_text().splitlines()
    assert len(summary) == 3
    for line, n in zip(summary[1:], (34, 66)):
        f = line.split(",")
        want = h3d.count_steps(float(P.sim_time(n)), h3d.params_reference(n)[1])
        assert f[:4] == ["1-hip-h3d", "1", str(n), "2"] and f[9] == str(want), line
