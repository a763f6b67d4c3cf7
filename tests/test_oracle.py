"""Pins the CPU oracle against evidence other than itself.

- golden fixtures written by the reference's own serial program
  (OpenMP_Single_Source/heat3D_variable_coeff.c compiled from source, see
  tests/golden/make_golden.py), compared byte for byte;
- values recorded in the reference's logs (HIP/verify-20370244.out:115-150,
  HIP/summary.csv steps column);
- closed forms (the discrete eigenmode of the constant-kappa operator);
- brute-force numpy restatements, and mutants of them that must fail the
  golden fixtures (a dropped term, a flipped sign, a transposed index).
"""
import gzip
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INDEX = json.load(open(os.path.join(GOLDEN, "index.json")))


def golden_lines(entry):
    with gzip.open(os.path.join(GOLDEN, entry["file"]), "rt") as f:
        return f.read().splitlines()


def dump_lines(u):
    return ["%.17g" % v for v in u]


@pytest.mark.parametrize("entry", INDEX, ids=lambda e: f"n{e['n']}_T{e['T']}")
def test_oracle_matches_reference_binary_bitwise(oracle, entry):
    n = entry["n"]
    h, dt, factor = oracle.params_reference(n)
    steps = oracle.count_steps(float(entry["T"]), dt)
    assert steps == entry["steps"]
    u, un, rhs, k = oracle.init_reference(n)
    res = oracle.run_f64(u, un, rhs, k, n, n, n, steps, factor=factor)
    assert dump_lines(res) == golden_lines(entry)


@pytest.mark.parametrize("entry", INDEX, ids=lambda e: f"n{e['n']}_T{e['T']}")
def test_oracle_openmp_equals_serial(oracle, entry):
    n = entry["n"]
    h, dt, factor = oracle.params_reference(n)
    steps = entry["steps"]
    u, un, rhs, k = oracle.init_reference(n)
    a = oracle.run_f64(u.copy(), un.copy(), rhs, k, n, n, n, steps, factor=factor)
    b = oracle.run_f64(u.copy(), un.copy(), rhs, k, n, n, n, steps, factor=factor,
                       nthreads=4)
    assert np.array_equal(a, b)


REF_SERIAL = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "serial")


@pytest.mark.skipif(not os.path.exists(REF_SERIAL),
                    reason="reference serial program not built (needs the reference tree)")
@pytest.mark.parametrize("n,steps", [(24, 13), (41, 6), (50, 3)])
def test_oracle_matches_reference_binary_live(oracle, tmp_path, n, steps):
    """Beyond the committed fixtures: the reference's serial program built
    from source (oracle/Makefile `ref`), at sizes with odd and even n and a
    step count from the %.7g T rule, against the oracle, byte for byte."""
    import subprocess
    h, dt, factor = oracle.params_reference(n)
    T = "%.7g" % (steps * dt)
    env = dict(os.environ, WRITE_TXT="1", OMP_NUM_THREADS="2")
    out = subprocess.run([os.path.abspath(REF_SERIAL), str(n), T], cwd=tmp_path, env=env,
                         check=True, capture_output=True, text=True, timeout=120)
    recorded = int(next(ln for ln in out.stdout.splitlines()
                        if ln.endswith("timesteps complete")).split()[0])
    assert recorded == oracle.count_steps(float(T), dt)
    u, un, rhs, k = oracle.init_reference(n)
    res = oracle.run_f64(u, un, rhs, k, n, n, n, recorded, factor=factor)
    assert open(tmp_path / "u_serial.txt").read().splitlines() == dump_lines(res)


def test_dump_format_roundtrip(oracle, tmp_path):
    entry = INDEX[-1]
    n = entry["n"]
    h, dt, factor = oracle.params_reference(n)
    u, un, rhs, k = oracle.init_reference(n)
    res = oracle.run_f64(u, un, rhs, k, n, n, n, entry["steps"], factor=factor)
    path = str(tmp_path / "u.txt")
    assert oracle.lib().orc_write_dump(path.encode(), oracle.p(res), n) == 0
    text = open(path).read().splitlines()
    assert text == golden_lines(entry)
    assert np.array_equal(np.array([float(x) for x in text]), res)  # %.17g round-trips


def test_step_counts_recorded_by_reference(oracle):
    # T = STEPS*dt printed %.7g (submit.sh:61-62); the loop test
    # `while (t<T)` then runs STEPS or STEPS+1 times.  Recorded counts:
    # HIP/verify-20370244.out:116 (N=256, 100 -> 101) and the steps column of
    # AMD-MI250X/Source-Results/HIP/summary.csv:2-8 (STEPS=500).
    recorded = {(256, 100): 101, (512, 500): 501, (640, 500): 500, (768, 500): 501,
                (896, 500): 501, (1024, 500): 501, (1152, 500): 500, (1280, 500): 500}
    for (n, s), want in recorded.items():
        h, dt, f = oracle.params_reference(n)
        T = float("%.7g" % (s * dt))
        assert oracle.count_steps(T, dt) == want, (n, s)


def test_params_match_recorded_banner(oracle):
    # HIP/verify-20370244.out:141-150 (N=256)
    h, dt, factor = oracle.params_reference(256)
    assert "%.6e" % h == "3.921569e-03"
    assert "%.6e" % dt == "1.349009e-06"
    assert "%.6e" % factor == "4.385965e-02"
    assert "%.6f" % (dt / (h * h)) == "0.087719"
    # factor is dt/h/h/2 evaluated left to right, not dt/(2h^2)
    assert factor == dt / h / h / 2.0


def numpy_step(u, k, rhs, factor, mutate=None):
    """Brute-force restatement of heat3D_variable_coeff.c:174-181."""
    c = u[1:-1, 1:-1, 1:-1]
    kc = k[1:-1, 1:-1, 1:-1]

    def nb(a, dz, dy, dx):
        n0, n1, n2 = a.shape
        return a[1 + dz:n0 - 1 + dz, 1 + dy:n1 - 1 + dy, 1 + dx:n2 - 1 + dx]

    A = (nb(k, 0, 0, 1) + kc) * (nb(u, 0, 0, 1) - c)
    B = (kc + nb(k, 0, 0, -1)) * (c - nb(u, 0, 0, -1))
    Cc = (nb(k, 0, 1, 0) + kc) * (nb(u, 0, 1, 0) - c)
    D = (kc + nb(k, 0, -1, 0)) * (c - nb(u, 0, -1, 0))
    E = (nb(k, 1, 0, 0) + kc) * (nb(u, 1, 0, 0) - c)
    F = (kc + nb(k, -1, 0, 0)) * (c - nb(u, -1, 0, 0))
    if mutate == "drop_E":
        E = 0.0 * E
    elif mutate == "sign_D":
        D = -D
    elif mutate == "transpose_C":  # j+1 neighbour taken along x
        Cc = (nb(k, 0, 0, 1) + kc) * (nb(u, 0, 1, 0) - c)
    S = A - B + Cc - D + E - F
    out = u.copy()
    out[1:-1, 1:-1, 1:-1] = c + nb(rhs, 0, 0, 0) + factor * S
    return out


@pytest.mark.parametrize("mutate", [None, "drop_E", "sign_D", "transpose_C"])
def test_golden_pins_every_term(oracle, mutate):
    entry = next(e for e in INDEX if e["n"] == 9)
    n, steps = entry["n"], entry["steps"]
    h, dt, factor = oracle.params_reference(n)
    u, un, rhs, k = oracle.init_reference(n)
    u, rhs, k = (a.reshape(n, n, n) for a in (u, rhs, k))
    for _ in range(steps):
        u = numpy_step(u, k, rhs, factor, mutate)
    same = dump_lines(u.ravel()) == golden_lines(entry)
    assert same == (mutate is None)


def test_constant_kappa_eigenmode_decay(oracle):
    # kappa = c, rhs = 0, u0 = sin(pi x) sin(pi y) sin(pi z): each step
    # multiplies u by g = 1 - 12*factor*c*(1 - cos(pi h)).
    n, c, steps = 21, 0.7, 40
    h, dt, factor = oracle.params_reference(n)
    x = np.arange(n) * h
    s = np.sin(np.pi * x)
    u0 = (s[:, None, None] * s[None, :, None] * s[None, None, :]).ravel()
    u0 = u0.reshape(n, n, n)
    u0[0], u0[-1], u0[:, 0], u0[:, -1], u0[:, :, 0], u0[:, :, -1] = 0, 0, 0, 0, 0, 0
    u0 = u0.ravel().copy()
    k = np.full(n ** 3, c)
    res = oracle.run_f64(u0.copy(), u0.copy(), None, k, n, n, n, steps, factor=factor)
    g = 1.0 - 12.0 * factor * c * (1.0 - math.cos(math.pi * h))
    assert np.max(np.abs(res - g ** steps * u0)) < 1e-13


def test_interior_boundary_split_equals_full(oracle):
    n = 11
    h, dt, factor = oracle.params_reference(n)
    u, un, rhs, k = oracle.init_reference(n)
    full = un.copy()
    oracle.step_f64(full, u, rhs, k, n, n, n, factor, 1, n - 2)
    split = un.copy()
    oracle.step_f64(split, u, rhs, k, n, n, n, factor, 1, 1)
    oracle.step_f64(split, u, rhs, k, n, n, n, factor, n - 2, n - 2)
    oracle.step_f64(split, u, rhs, k, n, n, n, factor, 2, n - 3)
    assert np.array_equal(full, split)


def test_slab_decompose_closed_forms(oracle):
    # 4-openmp-stream-omp-p2p.c:391-415: remainder to the lower devices
    assert oracle.slab_decompose(1026, 4) == (0, [258] * 4, [1, 257, 513, 769])
    assert oracle.slab_decompose(1024, 4) == (0, [258, 258, 257, 257], [1, 257, 513, 768])
    # 2-openmp.c:294-316: remainder to the higher device
    assert oracle.slab_decompose(1025, 2) == (0, [513, 514], [1, 512])
    assert oracle.slab_decompose(9, 4)[0] == -1  # slabs of 3 planes are refused
    assert oracle.slab_decompose(10, 4)[0] == 0


def test_slab_partition_init_matches_global(oracle):
    n, parts = 14, 4
    h, dt, f = oracle.params_reference(n)
    u, _, rhs, k = oracle.init_reference(n)
    rc, nzd, gks = oracle.slab_decompose(n, parts)
    for nz_l, gk in zip(nzd, gks):
        su, sr, sk = (np.empty(nz_l * n * n) for _ in range(3))
        oracle.lib().orc_init_partition(oracle.p(su), oracle.p(sr), oracle.p(sk),
                                        nz_l, gk, 1, n, n, n, h, dt)
        lo = (gk - 1) * n * n
        assert np.array_equal(sr, rhs[lo:lo + nz_l * n * n])
        assert np.array_equal(sk, k[lo:lo + nz_l * n * n])
        assert np.array_equal(su, u[lo:lo + nz_l * n * n])


def deep_slab_run(oracle, n, parts, ghost, steps, shift=0):
    """The deep-ghost exchange on the CPU: the reference's split widened to
    `ghost` ghost planes per side (ending at the global Dirichlet planes),
    passes of up to `ghost` oracle steps over each slab array, then the
    top owned band of slab p into the lower ghosts of p+1 and the bottom
    owned band of p+1 into the upper ghosts of p; `shift` moves the first
    band down by that many planes (a planted mistake).  Returns the owned
    planes reassembled into the global array."""
    h, dt, factor = oracle.params_reference(n)
    u, _, rhs, k = oracle.init_reference(n)
    rc, nzd, gks = oracle.slab_decompose(n, parts)
    assert rc == 0
    P = n * n
    slabs = []
    for nz_l, gk in zip(nzd, gks):
        own_lo, own_hi = gk, gk + nz_l - 3
        start, end = max(0, own_lo - ghost), min(n - 1, own_hi + ghost)
        sl = slice(start * P, (end + 1) * P)
        slabs.append(dict(nz=end - start + 1, start=start, lo=own_lo - start, hi=own_hi - start,
                          old=u[sl].copy(), new=u[sl].copy(), rhs=rhs[sl], kap=k[sl]))
    left = steps
    while left:
        r = min(ghost, left)
        for s in slabs:
            full = oracle.run_f64(s["old"].copy(), s["old"].copy(), s["rhs"], s["kap"], n, n,
                                  s["nz"], r, factor=factor)
            s["new"][s["lo"] * P:(s["hi"] + 1) * P] = full[s["lo"] * P:(s["hi"] + 1) * P]
        for a, b in zip(slabs, slabs[1:]):
            top = a["hi"] - ghost + 1 - shift
            b["new"][:ghost * P] = a["new"][top * P:(top + ghost) * P]
            a["new"][(a["hi"] + 1) * P:(a["hi"] + 1 + ghost) * P] = \
                b["new"][b["lo"] * P:(b["lo"] + ghost) * P]
        for s in slabs:
            s["old"], s["new"] = s["new"], s["old"]
        left -= r
    out = u.copy()
    for s in slabs:
        g0 = s["start"]
        out[(g0 + s["lo"]) * P:(g0 + s["hi"] + 1) * P] = s["old"][s["lo"] * P:(s["hi"] + 1) * P]
    return out, oracle.run_f64(u.copy(), u.copy(), rhs, k, n, n, n, steps, factor=factor)


@pytest.mark.parametrize("parts,ghost", [(2, 2), (4, 2), (4, 3), (2, 4)], ids=str)
def test_deep_ghost_exchange_equals_global_run(oracle, parts, ghost):
    """Slabs with `ghost` ghost planes, advanced `ghost` steps per pass with
    one band exchange per pass, reassemble to the global run bit for bit;
    a band taken one plane too low breaks it (the check the GPU test of the
    same procedure relies on)."""
    got, want = deep_slab_run(oracle, 38, parts, ghost, 2 * ghost + 1)
    assert np.array_equal(got, want)
    bad, _ = deep_slab_run(oracle, 38, parts, ghost, 2 * ghost + 1, shift=1)
    assert not np.array_equal(bad, want)


def test_metrics_hand_vectors(oracle):
    ref = np.array([1.0, -2.0, 0.0, 4.0, 1e-13])
    cand = np.array([1.0, -2.5, 1e-7, 4.0, 2e-13])
    m = oracle.metrics(ref, cand, atol=1e-6, rtol=1e-5)
    d = cand - ref
    assert m[0] == 0.5 and m[5] == 1
    assert m[1] == pytest.approx(np.abs(d).mean(), rel=1e-15)
    assert m[2] == pytest.approx(math.sqrt((d ** 2).mean()), rel=1e-15)
    assert m[3] == pytest.approx(np.linalg.norm(d) / np.linalg.norm(ref), rel=1e-15)
    assert m[4] == pytest.approx(0.25, rel=1e-15)  # |ref| > 1e-12 only
    assert m[6] == 0 and m[7] == 0
    assert m[8] == 1  # only the -2 -> -2.5 cell fails atol + rtol*|ref|
    bad = cand.copy()
    bad[2] = np.nan
    bad[3] = np.inf
    m = oracle.metrics(ref, bad)
    assert m[6] == 1 and m[7] == 1 and m[8] == 3
    assert math.isnan(m[0]) and m[5] == 2  # numpy: argmax returns the first NaN


def test_stats_against_fsum(oracle):
    rng = np.random.default_rng(5)
    u = rng.standard_normal(7 * 6 * 5) * 10 ** rng.uniform(-8, 8, 7 * 6 * 5)
    s = oracle.stats(u, 7, 6, 5)
    assert s[0] == pytest.approx(math.fsum(u), rel=1e-15, abs=1e-300)
    assert s[1] == pytest.approx(math.fsum(u * u), rel=1e-15)
    assert s[2] == np.abs(u).max()
    u[3] = np.nan
    assert oracle.stats(u, 7, 6, 5)[3] == 1


def test_face_sum_max_brute_force(oracle):
    rng = np.random.default_rng(1)
    nz, ny, nx = 5, 6, 7
    k = rng.uniform(0.5, 2.0, nz * ny * nx)
    K = k.reshape(nz, ny, nx)
    best = 0.0
    for z in range(1, nz - 1):
        for y in range(1, ny - 1):
            for x in range(1, nx - 1):
                c = K[z, y, x]
                s = sum(K[z + dz, y + dy, x + dx] + c for dz, dy, dx in
                        [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)])
                best = max(best, s)
    assert oracle.face_sum_max(k, nx, ny, nz) == pytest.approx(best, rel=1e-15)
    # constant kappa: 12 c, so dt <= 2h^2/(12c) = h^2/(6c)
    assert oracle.face_sum_max(np.full(nz * ny * nx, 0.75), nx, ny, nz) == 9.0


# ---------------------------------------------------------------------------
# FP32 paths (no reference counterpart).

def f32_problem(nx, ny, nz, seed=3):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, (nz, ny, nx)).astype(np.float32)
    u[0], u[-1], u[:, 0], u[:, -1], u[:, :, 0], u[:, :, -1] = 0, 0, 0, 0, 0, 0
    k = rng.uniform(0.5, 1.5, (nz, ny, nx)).astype(np.float32)
    return u.ravel(), k.ravel()


def np_fmaf(a, b, c):
    """float32 fused multiply-add a*b + c with one rounding, emulated in
    float64: the product of two float32 values is exact in float64; the
    sum is rounded to float64 and its error recovered exactly (TwoSum); a
    float64 sum that sits on a float32 rounding midpoint is resolved by the
    sign of that error (the only case where rounding twice differs from
    rounding once)."""
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    c = np.asarray(c, np.float32).astype(np.float64)
    p = a * b
    s = p + c
    bb = s - p
    err = (p - (s - bb)) + (c - bb)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    away = np.where(r64 < s, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
    o = np.nextafter(r, away)
    o64 = o.astype(np.float64)
    tie = (r64 != s) & ((s - r64) == (o64 - s))
    take_o = tie & (err != 0) & ((err > 0) == (o64 > s))
    return np.where(take_o, o, r).astype(np.float32)


def f32_S(u, k):
    """S of the FP32 update in the reference's term order, rounded per
    operation in float32."""
    c = u[1:-1, 1:-1, 1:-1]
    kc = k[1:-1, 1:-1, 1:-1]

    def nb(a, dz, dy, dx):
        n0, n1, n2 = a.shape
        return a[1 + dz:n0 - 1 + dz, 1 + dy:n1 - 1 + dy, 1 + dx:n2 - 1 + dx]

    return ((nb(k, 0, 0, 1) + kc) * (nb(u, 0, 0, 1) - c)
            - (kc + nb(k, 0, 0, -1)) * (c - nb(u, 0, 0, -1))
            + (nb(k, 0, 1, 0) + kc) * (nb(u, 0, 1, 0) - c)
            - (kc + nb(k, 0, -1, 0)) * (c - nb(u, 0, -1, 0))
            + (nb(k, 1, 0, 0) + kc) * (nb(u, 1, 0, 0) - c)
            - (kc + nb(k, -1, 0, 0)) * (c - nb(u, -1, 0, 0)))


def numpy_f32_plain(u, k, factor, steps, shape):
    u = u.reshape(shape).astype(np.float32).copy()
    k = k.reshape(shape).astype(np.float32)
    f = np.float32(factor)
    for _ in range(steps):
        S = f32_S(u, k)
        u = u.copy()
        u[1:-1, 1:-1, 1:-1] = np_fmaf(f, S, u[1:-1, 1:-1, 1:-1])
    return u.ravel()


def _fma_exact(a, b, c):
    """a*b + c rounded once to float32, from exact rational arithmetic."""
    from fractions import Fraction
    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    r = np.float32(float(v))  # candidate; fix by comparing exact distances
    best = r
    for cand in (np.nextafter(r, np.float32(-np.inf)), r, np.nextafter(r, np.float32(np.inf))):
        d, db = abs(Fraction(float(cand)) - v), abs(Fraction(float(best)) - v)
        if d < db or (d == db and (cand.view(np.uint32) & 1) == 0 and cand != best):
            best = cand
    return best


def test_np_fmaf_rounds_once():
    """The float64 emulation of the FP32 fused multiply-add against exact
    rational arithmetic and against the C library's fmaf, on random
    operands and on a constructed case where rounding twice (to float64,
    then to float32) lands on a tie and goes the wrong way."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.fmaf.restype = ctypes.c_float
    libm.fmaf.argtypes = [ctypes.c_float] * 3
    rng = np.random.default_rng(11)
    a = (rng.standard_normal(400) * 10.0 ** rng.integers(-3, 3, 400)).astype(np.float32)
    b = (rng.standard_normal(400) * 10.0 ** rng.integers(-3, 3, 400)).astype(np.float32)
    c = (-a.astype(np.float64) * b * (1 + rng.standard_normal(400) * 1e-6)).astype(np.float32)
    # the double-rounding trap: a*b + c = 1 + 2^-23 + 2^-24 - 2^-70 exactly
    trap = (np.float32(2.0 ** -12 * (1 + 2.0 ** -23)), np.float32(2.0 ** -12 * (1 - 2.0 ** -23)),
            np.float32(1 + 2.0 ** -23))
    a, b, c = (np.append(x, t) for x, t in zip((a, b, c), trap))
    got = np_fmaf(a, b, c)
    for i in range(a.size):
        want = _fma_exact(a[i], b[i], c[i])
        assert got[i].view(np.uint32) == want.view(np.uint32), (a[i], b[i], c[i])
        assert np.float32(libm.fmaf(a[i], b[i], c[i])).view(np.uint32) == want.view(np.uint32)
    assert got[-1] == np.float32(1 + 2.0 ** -23)  # not the tie's even neighbour
    naive = (a[-1].astype(np.float64) * b[-1] + c[-1]).astype(np.float32)
    assert naive != got[-1]


def test_f32_plain_matches_numpy(oracle):
    nx, ny, nz = 9, 8, 7
    u, k = f32_problem(nx, ny, nz)
    factor = 0.05
    got, _ = oracle.run_f32(u, k, None, nx, ny, nz, factor, 6, 2, 0)
    with np.errstate(all="ignore"):
        # numpy float32 arithmetic rounds each operation like the C code
        want = numpy_f32_plain(u, k, factor, 6, (nz, ny, nx))
    assert np.array_equal(got, want)


def np_ulp_exp(u):
    """max(ilogb(u), -126), -126 for 0 (numpy float32 restatement)."""
    m, e = np.frexp(u.astype(np.float32))
    e = (e - 1).astype(np.int32)
    e[(u == 0) | ~np.isfinite(u)] = -126
    return np.maximum(e, -126)


def numpy_f32_compensated(u, k, factor, steps, kb, mode, shape):
    """Kahan-compensated FP32 steps with the residual reset per pass of kb
    steps (mode 1) or kept across passes as FP16 scaled by the ulp (mode 2)."""
    u = u.reshape(shape).astype(np.float32).copy()
    k = k.reshape(shape)
    f = np.float32(factor)
    comp = np.zeros(shape, np.float16)
    done = 0
    inner = (slice(1, -1),) * 3
    while done < steps:
        ns = min(kb, steps - done)
        if mode == 2:
            e = np.ldexp(comp.astype(np.float32), np_ulp_exp(u) - 23).astype(np.float32)
        else:
            e = np.zeros(shape, np.float32)
        def nb(a, dz, dy, dx):
            n0, n1, n2 = a.shape
            return a[1 + dz:n0 - 1 + dz, 1 + dy:n1 - 1 + dy, 1 + dx:n2 - 1 + dx]

        for _ in range(ns):
            c = u[inner]
            kc = k[inner]  # S in float32, in the reference's term order
            S = ((nb(k, 0, 0, 1) + kc) * (nb(u, 0, 0, 1) - c)
                 - (kc + nb(k, 0, 0, -1)) * (c - nb(u, 0, 0, -1))
                 + (nb(k, 0, 1, 0) + kc) * (nb(u, 0, 1, 0) - c)
                 - (kc + nb(k, 0, -1, 0)) * (c - nb(u, 0, -1, 0))
                 + (nb(k, 1, 0, 0) + kc) * (nb(u, 1, 0, 0) - c)
                 - (kc + nb(k, -1, 0, 0)) * (c - nb(u, -1, 0, 0)))
            y = np_fmaf(f, S, e[inner])
            t = c + y
            e[inner] = y - (t - c)
            u = u.copy()
            u[inner] = t
        if mode == 2:
            sc = np.ldexp(e, 23 - np_ulp_exp(u)).astype(np.float32)
            comp = sc.astype(np.float16)
        done += ns
    return u.ravel()


@pytest.mark.parametrize("mode,kb", [(1, 1), (1, 3), (2, 1), (2, 4)])
def test_f32_compensated_matches_numpy(oracle, mode, kb):
    nx, ny, nz = 10, 9, 8
    u, k = f32_problem(nx, ny, nz, seed=6)
    factor = 0.05
    steps = 2 * kb + 1
    got, _ = oracle.run_f32(u, k, None, nx, ny, nz, factor, steps, kb, mode)
    with np.errstate(all="ignore"):
        want = numpy_f32_compensated(u, k, factor, steps, kb, mode, (nz, ny, nx))
    assert np.array_equal(got, want)


def test_f32_compensated_is_closer_to_f64(oracle):
    # small increments (factor*S << u), the regime compensation is for: the
    # rounding of u + d then dominates the plain update's error
    nx, ny, nz = 12, 11, 10
    u, k = f32_problem(nx, ny, nz, seed=9)
    u64 = u.astype(np.float64)
    k64 = k.astype(np.float64)
    factor = np.float32(0.004)
    steps = 2000
    ref = oracle.run_f64(u64.copy(), u64.copy(), None, k64, nx, ny, nz, steps,
                         factor=float(factor))
    err = {}
    for mode, kb in [(0, 4), (1, 4), (2, 4), (2, 1)]:
        got, _ = oracle.run_f32(u, k, None, nx, ny, nz, float(factor), steps, kb, mode)
        err[(mode, kb)] = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err[(1, 4)] < err[(0, 4)]
    assert err[(2, 4)] < err[(1, 4)]
    assert err[(2, 1)] < 1e-6


def test_comp16_encode_decode(oracle):
    L = oracle.lib()
    for u in [1.0, 3.7, 1e-5, -2.5e3, 0.0]:
        u32 = np.float32(u)
        ulp = np.spacing(np.abs(u32)) if u != 0 else np.float32(2.0 ** -149)
        for frac in [0.0, 0.5, -0.5, 0.123, -0.377]:
            e = np.float32(frac * ulp)
            back = L.orc_comp_decode(L.orc_comp_encode(e, u32), u32)
            assert abs(back - e) <= abs(float(ulp)) * 2.0 ** -11
    # exact for values representable in FP16 relative to the ulp
    assert L.orc_comp_decode(L.orc_comp_encode(np.float32(2.0 ** -24), 1.0), 1.0) == 2.0 ** -24


def test_bf16_conversion_matches_numpy(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(0)
    x = rng.uniform(-4, 4, 2000).astype(np.float32)
    w = x.view(np.uint32).astype(np.uint64)
    want = ((w + 0x7FFF + ((w >> 16) & 1)) >> 16).astype(np.uint16)
    got = np.array([L.orc_f32_to_bf16(float(v)) for v in x], dtype=np.uint16)
    assert np.array_equal(got, want)
    back = np.array([L.orc_bf16_to_f32(int(b)) for b in got], dtype=np.float32)
    assert np.all(np.abs(back - x) <= np.abs(x) * 2.0 ** -8)
