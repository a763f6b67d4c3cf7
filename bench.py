#!/usr/bin/env python3
"""Benchmark of the 3-D variable-coefficient heat update on one MI355X.

Workload (the FP64 GPU configuration the headline metric is quoted on):
512^3 interior + 1-cell Dirichlet ring (514^3 points), seeded sine kappa
(1 + 0.5 sin sin sin) x U[0.9,1.1], Gaussian hot spot u0 + U[0,1e-3],
dt = 0.9 x the FP64 positivity bound, FP64 storage and arithmetic.  One
"step" is one time step of the whole field; --steps of them are timed
(default 1000, as in the configuration), after --warmup untimed steps.

Reported beside it, in the same run:
  - the measured HBM stream bandwidth (copy / triad probe on 4 GiB arrays)
    and host <-> device copy rates,
  - the FP32-compensated, temporally blocked kernel on the same field and
    step count (the autotuner's fastest fused depth, and single-step
    passes), with GLUP/s, % of stream bandwidth, speed-up over the FP64
    kernel and rel-L2 / max error against an FP64 run on the same
    FP32-rounded inputs,
  - the other configurations: the reference's own GPU case (n = 512,
    reference fields with rhs, 501 steps), one slab of a 4-way split of
    1024^3 with in-HBM face exchange, 1024^3 FP32 for 2000 steps in four
    storage variants, and a 64-member mixed ensemble (--no-configs skips),
  - the CPU baseline: the reference's own serial program (built from its
    source into oracle/_ref) with OpenMP on this host's cores and on one
    thread, on the reference's recorded CPU case (n = 256, 101 steps), and
    the plain C oracle on one core on 128^3 x 100 steps.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "openmp-offloading-intranode-artifact_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(REPO, "tests"))

import numpy as np  # noqa: E402

import h3d_py as h3d  # noqa: E402

METRIC = "GLUP/s + % HBM roofline; FP32-compensated vs FP64 at stated rel-L2, 1 MI355X"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--m", type=int, default=512, help="interior points per axis")
    ap.add_argument("--kblock", type=int, default=0,
                    help="FP32 time steps per pass (0: autotune)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 leg")
    ap.add_argument("--dense", action="store_true",
                    help="reference's dense layout instead of 128-byte aligned rows")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the slab, 1024^3 FP32 and ensemble configurations")
    ap.add_argument("--tune64", default="",
                    help="kblock,rows,waves,zchunk: skip FP64 autotuning")
    ap.add_argument("--tune64s", default="",
                    help="rows,waves,zchunk of the single-step FP64 leg: skip its autotuning")
    ap.add_argument("--tune32", default="",
                    help="kblock,rows,waves,zchunk: skip FP32 autotuning")
    return ap.parse_args()


def traffic_from_profiles(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE, scaled by the copy-kernel calibration), committed under
    profiles/; None when not measured for this kernel and shape."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(name, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline():
    """The reference's own program (heat3D_variable_coeff.c, built from
    source) on its recorded CPU case, timed on this host: with OpenMP on the
    host cores (the headline baseline) and on one thread (the serial C
    reference); plus the oracle on one core.  Each is bounded to seconds."""
    out = {}
    ncores = len(os.sched_getaffinity(0))
    threads = min(ncores, 16)
    ref = os.path.join(REPO, "oracle", "_ref", "serial")
    n, s = 256, 100
    h = 1.0 / (n - 1)
    dt = h * h / (12.0 * 0.95)
    T = "%.7g" % (s * dt)  # submit.sh:61-62 -> 101 steps

    def run_ref(nthreads):
        env = dict(os.environ, WRITE_TXT="0", OMP_NUM_THREADS=str(nthreads),
                   OMP_PROC_BIND="spread", OMP_PLACES="cores")
        r = subprocess.run([ref, str(n), T], env=env, capture_output=True, text=True,
                           timeout=300, cwd="/tmp")
        steps = int(re.search(r"^(\d+) timesteps complete", r.stdout, re.M).group(1))
        sec = float(re.search(r"^Time = ([0-9.]+) seconds", r.stdout, re.M).group(1))
        return {"value": (n - 2) ** 3 * steps / sec / 1e9, "unit": "GLUP/s",
                "cores": nthreads, "kind": "reference",
                "sample": f"reference heat3D_variable_coeff.c built from source, "
                          f"./serial {n} {T} ({steps} steps, {n - 2}^3 interior), "
                          f"OMP_NUM_THREADS={nthreads}, {sec:.3f} s"}

    if os.path.exists(ref):
        out = run_ref(threads)
        out["host_cpus"] = ncores
        out["serial_reference"] = run_ref(1)
    # one-core C oracle on the 128^3 configuration (port of the same loop)
    import oracle_py as oracle
    m = 128
    nn = m + 2
    g = h3d.Grid.dense(nn)
    u, k = h3d.gen_seeded(g, 0, nn, nn, nn, 0, 20260101)
    fac = 0.9 * 2.0 / oracle.face_sum_max(k, nn, nn, nn) / 2.0
    t0 = time.perf_counter()
    oracle.run_f64(u.copy(), u.copy(), None, k, nn, nn, nn, 100, factor=fac, nthreads=1)
    sec = time.perf_counter() - t0
    serial = {"value": m ** 3 * 100 / sec / 1e9, "unit": "GLUP/s", "cores": 1,
              "kind": "port", "sample": f"C oracle, 128^3 interior x 100 steps, {sec:.3f} s"}
    if not out:
        out = dict(serial)
    out["serial"] = serial
    return out


def seeded_problem(gx, gy, gz, kind, seed, g=None, gk0=0, member=0):
    """Seeded fields of a (gx, gy, gz)-point problem (or of the planes of
    grid g starting at global plane gk0) and factor = dt/h/h/2 with
    dt = 0.9 x the positivity bound of the generated kappa."""
    g = g or h3d.pitched_grid(gx, gy, gz)
    u, k = h3d.gen_seeded(g, gk0, gx, gy, gz, kind, seed, member)
    h = 1.0 / (gx - 1)
    fmax = h3d.lib().h3d_face_sum_max_host(h3d.ptr(k), h3d.C.byref(g))
    dt = 0.9 * 2.0 * h * h / fmax
    return g, u, k, dt / h / h / 2.0


def leg_reference_workload(bw, stream):
    """The reference's own GPU benchmark case, for comparison with its
    published numbers: reference-mode fields (sin/cos closed forms, rhs),
    n = 512 points per axis (510^3 interior) and T = 500 dt printed %.7g,
    which the step rule turns into 501 steps
    (AMD-MI250X/Source-Results/HIP/summary.csv:2; submit.sh:61-62).  Device
    fields in the h3d program's layout (rows padded to 32 elements,
    interior rows 128-B aligned).  32 B per update (u, rhs, kappa read; u
    written)."""
    n = 512
    h, dt, factor = h3d.params_reference(n)
    T = float("%.7g" % (500 * dt))
    steps = h3d.count_steps(T, dt)
    g = h3d.pitched_grid(n, n, n)
    u, r, k = (np.zeros(g.elems) for _ in range(3))
    h3d.call("h3d_init_reference", h3d.ptr(u), h3d.ptr(r), h3d.ptr(k), h3d.C.byref(g))
    da, db, dr, dk = (h3d.AlignedField(g, np.float64) for _ in range(4))
    for d, x in ((da, u), (db, u), (dr, r), (dk, k)):
        d.upload(x)
    t = h3d.Tune()
    h3d.call("h3d_autotune", 0, da.ptr, db.ptr, dk.ptr, 0, h3d.C.byref(g), 1, 0, stream.ptr,
             h3d.C.byref(t))  # mode 1: shapes that hold rhs
    ms, fb = h3d.C.c_float(), h3d.C.c_int()
    h3d.call("h3d_run_f64", da.ptr, db.ptr, dr.ptr, dk.ptr, h3d.C.byref(g), factor, 10,
             h3d.C.byref(t), 0, stream.ptr, h3d.C.byref(ms), h3d.C.byref(fb))
    da.upload(u)
    db.upload(u)
    h3d.call("h3d_run_f64", da.ptr, db.ptr, dr.ptr, dk.ptr, h3d.C.byref(g), factor, steps,
             h3d.C.byref(t), 0, stream.ptr, h3d.C.byref(ms), h3d.C.byref(fb))
    sec = ms.value / 1e3
    lups = (n - 2) ** 3 * steps
    glups = lups / sec / 1e9
    st = h3d.stats(db if fb.value else da, g)
    for d in (da, db, dr, dk):
        d.close()
    mi250x_hip = 13.06  # GLUP/s, 1 GCD, HIP/summary.csv:2 (5.0899 s for 501 steps)
    bpl = 32.0 / max(1, t.kblock)  # u, rhs, kappa read and u written once per pass
    return {"value": round(glups, 3), "unit": "GLUP/s", "dtype": "f64", "n": n,
            "steps": steps, "T": "%.7g" % T, "seconds": round(sec, 5),
            "algorithmic_bytes_per_lup": bpl,
            "achieved_gbs": round(bpl * lups / sec / 1e9, 1),
            "frac_of_stream": round(bpl * lups / sec / 1e9 / bw, 4),
            "tune": [t.rows, t.waves, t.zchunk], "kblock": t.kblock, "checksum": float(st[0]),
            "reference_mi250x_hip_glups": mi250x_hip,
            "ratio_to_reference_mi250x_hip": round(glups / mi250x_hip, 2)}


def leg_cfg3(bw, stream):
    """One GPU's slab of a 4-way z split of 1024^3: 1024 x 1024 x 256 owned
    planes + 2 ghosts (1026 x 1026 x 258 points), lognormal kappa, 500 FP64
    steps, boundary planes and in-HBM face pack/unpack overlapped with the
    interior, replayed from a hipGraph.  dt from this slab's kappa."""
    n, planes, steps = 1026, 258, 500
    rc, nzd, gks = h3d.slab_decompose(n, 4)
    assert rc == 0 and nzd[1] == planes
    g = h3d.pitched_grid(n, n, planes)
    g, u, k, factor = seeded_problem(n, n, n, 1, 20260101, g=g, gk0=gks[1] - 1)
    fields = []
    for x in (u, u, k):
        f = h3d.AlignedField(g, np.float64)
        f.upload(x)
        fields.append(f)
    da, db, dk = fields
    lo, hi = h3d.DeviceArray(n * n * 8), h3d.DeviceArray(n * n * 8)
    ms, fb = h3d.C.c_float(), h3d.C.c_int()
    args = (da.ptr, db.ptr, None, dk.ptr, h3d.C.byref(g), factor, lo.ptr, hi.ptr)
    t = h3d.Tune()
    h3d.call("h3d_autotune", 0, da.ptr, db.ptr, dk.ptr, 0, h3d.C.byref(g), 0, 1, stream.ptr,
             h3d.C.byref(t))  # the slab step runs single steps
    h3d.call("h3d_run_slab_f64", *args, 10, h3d.C.byref(t), 1, stream.ptr, h3d.C.byref(ms),
             h3d.C.byref(fb))
    h3d.call("h3d_run_slab_f64", *args, steps, h3d.C.byref(t), 1, stream.ptr,
             h3d.C.byref(ms), h3d.C.byref(fb))
    lups = (n - 2) * (n - 2) * (planes - 2) * steps
    # per step: 2 packs + 2 unpacks, each reading and writing one plane
    face_bytes = 4 * 2 * n * n * 8 * steps
    gbs = (24.0 * lups + face_bytes) / (ms.value / 1e3) / 1e9
    st = h3d.stats(db if fb.value else da, g)
    for f in fields + [lo, hi]:
        f.close()
    deep = {f"ghost{K}": slab_deep(n, planes - 2, gks[1], K, factor, steps, bw, stream)
            for K in (2, 3)}
    return {"value": round(lups / (ms.value / 1e3) / 1e9, 3), "unit": "GLUP/s",
            "dtype": "f64", "steps": steps, "ms_per_step": round(ms.value / steps, 5),
            "achieved_gbs": round(gbs, 1), "frac_of_stream": round(gbs / bw, 4),
            "slab": [n, n, planes], "nz_split": nzd, "face_bytes": n * n * 8,
            "tune": [t.rows, t.waves, t.zchunk],
            "nonfinite": float(st[3]),
            "fused_passes_deep_ghosts": deep}


def slab_deep(n, owned, gk_owned, K, factor, steps, bw, stream):
    """The same slab held with K ghost planes per side (n x n x (owned +
    2K) points): fused passes of K steps, bands of K planes exchanged in
    HBM once per pass, the interior overlapped with the bands.  Same update
    count and factor as the one-ghost slab; the interior pass shape comes
    from the autotuner at pass depth K."""
    planes = owned + 2 * K
    g = h3d.pitched_grid(n, n, planes)
    u, k = h3d.gen_seeded(g, gk_owned - K, n, n, n, 1, 20260101)
    fields = []
    for x in (u, u, k):
        f = h3d.AlignedField(g, np.float64)
        f.upload(x)
        fields.append(f)
    da, db, dk = fields
    lo, hi = h3d.DeviceArray(K * n * n * 8), h3d.DeviceArray(K * n * n * 8)
    ms, fb = h3d.C.c_float(), h3d.C.c_int()
    t = h3d.Tune()
    h3d.call("h3d_autotune", 0, da.ptr, db.ptr, dk.ptr, 0, h3d.C.byref(g), 0, K, stream.ptr,
             h3d.C.byref(t))
    args = (da.ptr, db.ptr, None, dk.ptr, h3d.C.byref(g), factor, K, lo.ptr, hi.ptr)
    h3d.call("h3d_run_slab_pass_f64", *args, 4 * K, h3d.C.byref(t), 1, stream.ptr,
             h3d.C.byref(ms), h3d.C.byref(fb))
    h3d.call("h3d_run_slab_pass_f64", *args, steps, h3d.C.byref(t), 1, stream.ptr,
             h3d.C.byref(ms), h3d.C.byref(fb))
    lups = (n - 2) * (n - 2) * owned * steps
    passes = -(-steps // K)
    # per pass: 2 band packs + 2 band unpacks of K planes, read and written
    face_bytes = 4 * 2 * K * n * n * 8 * passes
    gbs = (24.0 * lups / K + face_bytes) / (ms.value / 1e3) / 1e9
    st = h3d.stats(db if fb.value else da, g)
    for f in fields + [lo, hi]:
        f.close()
    return {"glups": round(lups / (ms.value / 1e3) / 1e9, 3),
            "ms_per_step": round(ms.value / steps, 5), "achieved_gbs": round(gbs, 1),
            "frac_of_stream": round(gbs / bw, 4), "slab": [n, n, planes],
            "band_bytes": K * n * n * 8, "tune": [t.rows, t.waves, t.zchunk, t.kblock],
            "nonfinite": float(st[3])}


def leg_cfg4(bw, stream):
    """1024^3 interior, FP32 storage, compensated, temporally blocked
    (kblock from the autotuner), 2000 steps: kappa in FP32 and in BF16 with a
    per-field scale, residual in registers and in an FP16 field; each against
    the FP64 kernel run on the same (rounded / dequantised) inputs."""
    m, steps = 1024, 2000
    n = m + 2
    g, u, k, factor = seeded_problem(n, n, n, 0, 20260101)
    u32 = u.astype(np.float32)
    del u
    k32 = k.astype(np.float32)
    bits, scale = h3d.quantize_bf16(k)
    del k
    out = {"interior": [m, m, m], "steps": steps}
    u64 = u32.astype(np.float64)

    def f64_reference(kappa64):
        fs = []
        for x in (u64, u64, kappa64):
            f = h3d.AlignedField(g, np.float64)
            f.upload(x)
            fs.append(f)
        ms, fb = h3d.C.c_float(), h3d.C.c_int()
        t64 = h3d.Tune()
        h3d.call("h3d_autotune", 0, fs[0].ptr, fs[1].ptr, fs[2].ptr, 0, h3d.C.byref(g), 0, 0,
                 stream.ptr, h3d.C.byref(t64))
        h3d.call("h3d_run_f64", fs[0].ptr, fs[1].ptr, None, fs[2].ptr, h3d.C.byref(g),
                 factor, steps, h3d.C.byref(t64), 0, stream.ptr, h3d.C.byref(ms),
                 h3d.C.byref(fb))
        keep = fs[1] if fb.value else fs[0]
        for f in fs:
            if f is not keep:
                f.close()
        return keep, ms.value

    fa, fbuf = h3d.AlignedField(g, np.float32), h3d.AlignedField(g, np.float32)
    ca, cb = h3d.AlignedField(g, np.uint16), h3d.AlignedField(g, np.uint16)
    for label, kap_host, kap_dtype, s in (("f32_kappa", k32, np.float32, 1.0),
                                          ("bf16_kappa", bits, np.uint16, scale)):
        ref, ms64 = f64_reference(k32.astype(np.float64) if s == 1.0 and kap_dtype == np.float32
                                  else h3d.bf16_to_f64(bits, scale))
        fk = h3d.AlignedField(g, kap_dtype)
        fk.upload(kap_host)
        kf, kb = (fk.ptr, None) if kap_dtype == np.float32 else (None, fk.ptr)
        f32f = h3d.C.c_float(np.float32(factor * s))
        for mode, mname in ((1, "carry"), (2, "fp16_residual")):
            t = h3d.Tune()
            h3d.call("h3d_autotune", 1, fa.ptr, fbuf.ptr, fk.ptr,
                     1 if kb is not None else 0, h3d.C.byref(g), mode, 0, stream.ptr,
                     h3d.C.byref(t))
            fa.upload(u32)
            fbuf.upload(u32)
            zero = np.zeros(g.elems, np.uint16)
            ca.upload(zero)
            cb.upload(zero)
            ms, fb = h3d.C.c_float(), h3d.C.c_int()
            h3d.call("h3d_run_f32", fa.ptr, fbuf.ptr, kf, kb, ca.ptr, cb.ptr, h3d.C.byref(g),
                     f32f, steps, mode, h3d.C.byref(t), 1, stream.ptr, h3d.C.byref(ms),
                     h3d.C.byref(fb))
            mt = h3d.compare(ref, fbuf if fb.value else fa, g, f32=True)
            passes = -(-steps // t.kblock)
            bpl = (8.0 if kb is None else 6.0) + 4.0 + (4.0 if mode == 2 else 0.0)
            gbs = bpl * m ** 3 * passes / (ms.value / 1e3) / 1e9
            out[f"{label}_{mname}"] = {
                "value": round(m ** 3 * steps / (ms.value / 1e3) / 1e9, 3), "unit": "GLUP/s",
                "kblock": t.kblock, "tune": [t.rows, t.waves, t.zchunk],
                "algorithmic_bytes_per_lup": round(bpl / t.kblock, 3),
                "achieved_gbs": round(gbs, 1), "frac_of_stream": round(gbs / bw, 4),
                "rel_l2_vs_fp64": float(mt[3]), "max_abs_vs_fp64": float(mt[0]),
                "max_rel_vs_fp64": float(mt[4])}
            log(f"cfg4 {label} {mname}: {out[f'{label}_{mname}']}")
        out[f"{label}_fp64_reference_glups"] = round(m ** 3 * steps / (ms64 / 1e3) / 1e9, 3)
        fk.close()
        ref.close()
    for f in (fa, fbuf, ca, cb):
        f.close()
    out["kappa_bf16_scale"] = scale
    return out


def leg_cfg5(bw, stream):
    """64 independent problems, each axis uniform in [48, 256] interior
    points, lognormal kappa, Gaussian u0, alternating FP64 / FP32-compensated
    members, 500 steps captured in one hipGraph; FP32 members against their
    FP64 twins."""
    rng = np.random.default_rng(20260101)
    dims = [tuple(int(x) + 2 for x in rng.integers(48, 257, 3)) for _ in range(64)]

    def build(specs):
        mems, keep = [], []
        for s in specs:
            dt = np.float32 if s["precision"] else np.float64
            da, db, dk = (h3d.DeviceArray.from_host(np.ascontiguousarray(x, dtype=dt))
                          for x in (s["u"], s["u"], s["k"]))
            keep.append((da, db, dk))
            mems.append(h3d.Member(da.ptr, db.ptr, dk.ptr, s["g"], s["factor"], s["precision"]))
        arr = (h3d.Member * len(mems))(*mems)
        e = h3d.C.c_void_p()
        h3d.call("h3d_ensemble_create", h3d.C.byref(e), arr, len(mems), None, 0)
        return e, keep

    specs = []
    for mi, (nx, ny, nz) in enumerate(dims):
        g = h3d.Grid.dense(nx, ny, nz)
        u, k = h3d.gen_seeded(g, 0, nx, ny, nz, 1, 20260101, mi)
        prec = mi % 2
        if prec:
            u, k = u.astype(np.float32).astype(np.float64), k.astype(np.float32).astype(np.float64)
        h = 1.0 / (nx - 1)
        fmax = h3d.lib().h3d_face_sum_max_host(h3d.ptr(k), h3d.C.byref(g))
        specs.append(dict(g=g, u=u, k=k, factor=0.9 * 2 * h * h / fmax / h / h / 2.0,
                          precision=prec))
    e, keep = build(specs)
    ms = h3d.C.c_float()
    et = h3d.Tune()
    h3d.call("h3d_ensemble_autotune", e, stream.ptr, h3d.C.byref(et))  # z-march depth
    h3d.call("h3d_ensemble_run", e, 500, 1, stream.ptr, h3d.C.byref(ms))  # builds the graph
    h3d.call("h3d_ensemble_run", e, 500, 1, stream.ptr, h3d.C.byref(ms))  # timed: 1000 total
    nwork, nblk, done = h3d.C.c_int(), h3d.C.c_int(), h3d.C.c_longlong()
    h3d.call("h3d_ensemble_info", e, h3d.C.byref(nwork), h3d.C.byref(nblk), h3d.C.byref(done))
    cells = sum((x - 2) * (y - 2) * (z - 2) for x, y, z in dims)
    sec = ms.value / 1e3
    glups = cells * 500 / sec / 1e9
    twins = [dict(s, precision=0) for s in specs[1::2]]
    e2, keep2 = build(twins)
    h3d.call("h3d_ensemble_run", e2, 1000, 1, stream.ptr, h3d.C.byref(ms))
    per_member = []
    in_b = h3d.C.c_int()
    for mi, ((da, db, dk), (ta, tb, tk)) in enumerate(zip(keep[1::2], keep2)):
        h3d.call("h3d_ensemble_where", e, 2 * mi + 1, h3d.C.byref(in_b))
        got = (db if in_b.value else da).download().astype(np.float64)
        h3d.call("h3d_ensemble_where", e2, mi, h3d.C.byref(in_b))
        ref = (tb if in_b.value else ta).download()  # FP64 twin, 1000 steps
        per_member.append({"member": 2 * mi + 1, "dims": list(dims[2 * mi + 1]),
                           "rel_l2": float("%.3g" % (np.linalg.norm(got - ref)
                                                     / np.linalg.norm(ref))),
                           "max_abs": float("%.3g" % np.max(np.abs(got - ref)))})
    worst = max(m["rel_l2"] for m in per_member)
    h3d.call("h3d_ensemble_destroy", e)
    h3d.call("h3d_ensemble_destroy", e2)
    for trip in keep + keep2:
        for d in trip:
            d.close()
    # FP64 members 24 B and FP32 members 4+4+4 B plus 2+2 B residual per
    # cell and pass of kblock steps
    k64 = et.kblock
    b64 = sum(24.0 / k64 * (x - 2) * (y - 2) * (z - 2) for x, y, z in dims[0::2])
    b32 = sum(16.0 / et.kblock * (x - 2) * (y - 2) * (z - 2) for x, y, z in dims[1::2])
    gbs = (b64 + b32) * 500 / sec / 1e9
    return {"value": round(glups, 3), "unit": "GLUP/s (aggregate)", "members": 64,
            "cells": int(cells), "steps_timed": 500, "graph": True, "zchunk": et.zchunk,
            "kblock": et.kblock,
            "ms_per_step": round(1e3 * sec / 500, 5),
            "work_items": nwork.value, "workgroups": nblk.value,
            "achieved_gbs": round(gbs, 1), "frac_of_stream": round(gbs / bw, 4),
            "worst_fp32_member_rel_l2_vs_fp64": worst,
            "fp32_members_vs_fp64": per_member}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: one GPU only (--gpus {a.gpus} requested): the update would "
              "shard as z-slabs, one per GPU, exchanging one face per neighbour per "
              "step, but no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    if h3d.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    arch = (h3d.C.c_char * 64)()
    h3d.call("h3d_device_arch", 0, arch, 64)
    log("device", arch.value.decode())

    # ---- HBM stream bandwidth (the roofline) ----
    copy_gbs, triad_gbs = h3d.stream_probe(4 << 30, reps=5)
    bw = max(copy_gbs, triad_gbs)
    log(f"stream: copy {copy_gbs:.0f} GB/s, triad {triad_gbs:.0f} GB/s")
    # host <-> device rates (the boundary takes device buffers, so these are
    # reported beside the headline, never folded into it)
    transfer = {}
    for kind, pinned in (("pinned", True), ("pageable", False)):
        h2d, d2h, lh, ld = h3d.transfer_probe(64 << 20, reps=3, pinned=pinned)
        transfer[kind] = {"h2d_gbs": round(h2d, 2), "d2h_gbs": round(d2h, 2),
                          "h2d_4k_us": round(lh, 2), "d2h_4k_us": round(ld, 2)}
    log(f"transfer: {transfer}")

    # ---- problem ----
    m = a.m
    n = m + 2
    # rows padded to 32 elements and offset so each row's first interior
    # point starts a 128-byte line (the C ABI takes any row/plane pitch)
    g = h3d.pitched_grid(n, n, n) if not a.dense else h3d.Grid.dense(n)
    t0 = time.perf_counter()
    u0, kap = h3d.gen_seeded(g, 0, n, n, n, 0, 20260101)
    hh = 1.0 / (n - 1)
    fmax = h3d.lib().h3d_face_sum_max_host(h3d.ptr(kap), h3d.C.byref(g))
    dt = 0.9 * 2.0 * hh * hh / fmax
    factor = dt / hh / hh / 2.0
    log(f"generated {n}^3 fields in {time.perf_counter() - t0:.1f} s; dt={dt:.6e}")

    stream = h3d.Stream()
    def field(host, dtype):
        f = h3d.AlignedField(g, dtype)
        f.upload(host)
        return f

    da, db, dk = (field(x, np.float64) for x in (u0, u0, kap))
    cells = m ** 3
    ms, fb = h3d.C.c_float(), h3d.C.c_int()

    def fp64_run(tune):
        """Timed FP64 run of a.steps steps from u0 (after a.warmup untimed
        steps): GLUP/s by the host clock, device seconds per launch (one
        launch = one pass of tune.kblock steps), final buffer."""
        if a.warmup:
            h3d.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, h3d.C.byref(g), factor,
                     a.warmup, h3d.C.byref(tune), 0, stream.ptr, h3d.C.byref(ms),
                     h3d.C.byref(fb))
        # fresh initial state for the timed run
        da.upload(u0)
        db.upload(u0)
        h3d.call("h3d_device_synchronize", 0)
        t0 = time.perf_counter()
        h3d.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, h3d.C.byref(g), factor,
                 a.steps, h3d.C.byref(tune), 0, stream.ptr, h3d.C.byref(ms), h3d.C.byref(fb))
        h3d.call("h3d_device_synchronize", 0)
        wall = time.perf_counter() - t0
        launches = -(-a.steps // max(1, tune.kblock))
        return cells * a.steps / wall / 1e9, wall, ms.value / 1e3 / launches

    def tuned(fixed_kblock):
        t = h3d.Tune()
        h3d.call("h3d_autotune", 0, da.ptr, db.ptr, dk.ptr, 0, h3d.C.byref(g), 0,
                 fixed_kblock, stream.ptr, h3d.C.byref(t))
        return t

    if a.tune64:
        k_, r_, w_, z_ = (int(x) for x in a.tune64.split(","))
        tune = h3d.Tune(r_, w_, z_, k_)
    else:
        tune = tuned(0)  # fused passes allowed: the fastest time per step
    kb64 = max(1, tune.kblock)
    log(f"FP64 tune: kblock={kb64} rows={tune.rows} waves={tune.waves} zchunk={tune.zchunk}")
    glups, wall, per_launch_s = fp64_run(tune)
    alg_bytes = 24.0 * cells  # per launch: u_old + kappa read, u_new written once
    achieved = alg_bytes / per_launch_s / 1e9
    res64 = db if fb.value else da
    st = h3d.stats(res64, g)
    log(f"FP64: {a.steps} steps in {wall:.3f} s -> {glups:.1f} GLUP/s, "
        f"{achieved:.0f} GB/s ({100 * achieved / bw:.1f}% of stream)")
    # the single-step kernel (one launch per step, as the reference runs)
    if a.tune64s:
        r_, w_, z_ = (int(x) for x in a.tune64s.split(","))
        tune1 = h3d.Tune(r_, w_, z_, 1)
    else:
        tune1 = tune if kb64 == 1 else tuned(1)
    glups1, wall1, per_launch1 = fp64_run(tune1)
    single = {"value": round(glups1, 3), "unit": "GLUP/s", "kblock": 1,
              "tune": {"rows": tune1.rows, "waves": tune1.waves, "zchunk": tune1.zchunk},
              "ms_per_step": round(1e3 * wall1 / a.steps, 5),
              "achieved_gbs": round(alg_bytes / per_launch1 / 1e9, 1),
              "frac_of_stream": round(alg_bytes / per_launch1 / 1e9 / bw, 4),
              "algorithmic_bytes_per_lup": 24,
              "traffic": traffic_from_profiles(f"h3d_march_f64:{m}"),
              "kernel_ms_per_launch": round(1e3 * per_launch1, 5)}
    log(f"FP64 single steps: {glups1:.1f} GLUP/s ({100 * single['frac_of_stream']:.1f}% of stream)")
    kname = f"h3d_pass_f64_kernel:{m}:kb{kb64}" if kb64 > 1 else f"h3d_march_f64:{m}"

    result = {
        "metric": METRIC, "value": round(glups, 3), "unit": "GLUP/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * wall / a.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generators)",
        "config": {"workload": f"heat3d_variable_coeff FP64, {m}^3 interior + Dirichlet ring, "
                               f"seeded sine kappa, Gaussian u0, dt=0.9*bound, rhs=0",
                   "interior": [m, m, m], "points": [n, n, n], "seed": 20260101,
                   "layout": "dense" if a.dense else f"rows padded to {g.sy} elements, "
                                                     f"interior rows 128-B aligned",
                   "tune": {"rows": tune.rows, "waves": tune.waves, "zchunk": tune.zchunk,
                            "kblock": kb64}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(kname),
                     "kernel": kname.split(":")[0], "steps_per_launch": kb64,
                     "frac_of_stream": round(achieved / bw, 4),
                     "algorithmic_bytes_per_lup": round(24.0 / kb64, 3),
                     "kernel_ms_per_launch": round(1e3 * per_launch_s, 5)},
        "fp64_single_step": single,
        "stream": {"copy_gbs": round(copy_gbs, 1), "triad_gbs": round(triad_gbs, 1),
                   "array_bytes": 4 << 30},
        "transfer": dict(transfer, bytes=64 << 20),
        "checksum": {"sum": float(st[0]), "sumsq": float(st[1]), "max": float(st[2])},
    }

    # ---- FP32 compensated, temporally blocked ----
    if not a.no_fp32:
        u32 = u0.astype(np.float32)
        k32 = kap.astype(np.float32)
        # FP64 reference on the FP32-rounded inputs (untimed)
        da.upload(u32.astype(np.float64))
        db.upload(u32.astype(np.float64))
        dk.upload(k32.astype(np.float64))
        h3d.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, h3d.C.byref(g), factor,
                 a.steps, h3d.C.byref(tune), 0, stream.ptr, h3d.C.byref(ms), h3d.C.byref(fb))
        ref = db if fb.value else da
        fa, fbuf, fk = (field(x, np.float32) for x in (u32, u32, k32))
        f32f = h3d.C.c_float(np.float32(factor))

        def fp32_leg(fixed_kblock, fixed=""):
            if fixed:
                k_, r_, w_, z_ = (int(x) for x in fixed.split(","))
                t32 = h3d.Tune(r_, w_, z_, k_)
            else:
                t32 = h3d.Tune()
                h3d.call("h3d_autotune", 1, fa.ptr, fbuf.ptr, fk.ptr, 0, h3d.C.byref(g), 1,
                         fixed_kblock, stream.ptr, h3d.C.byref(t32))
            log(f"FP32 tune: kblock={t32.kblock} rows={t32.rows} waves={t32.waves} "
                f"zchunk={t32.zchunk}")
            h3d.call("h3d_run_f32", fa.ptr, fbuf.ptr, fk.ptr, None, None, None,
                     h3d.C.byref(g), f32f, max(a.warmup, t32.kblock), 1, h3d.C.byref(t32), 0,
                     stream.ptr, h3d.C.byref(ms), h3d.C.byref(fb))
            fa.upload(u32)
            fbuf.upload(u32)
            h3d.call("h3d_device_synchronize", 0)
            t0 = time.perf_counter()
            h3d.call("h3d_run_f32", fa.ptr, fbuf.ptr, fk.ptr, None, None, None,
                     h3d.C.byref(g), f32f, a.steps, 1, h3d.C.byref(t32), 0, stream.ptr,
                     h3d.C.byref(ms), h3d.C.byref(fb))
            h3d.call("h3d_device_synchronize", 0)
            wall32 = time.perf_counter() - t0
            glups32 = cells * a.steps / wall32 / 1e9
            passes = -(-a.steps // t32.kblock)
            bytes32 = 12.0 * cells * passes  # u + kappa read, u written, per pass
            gbs32 = bytes32 / (ms.value / 1e3) / 1e9
            m32 = h3d.compare(ref, fbuf if fb.value else fa, g, f32=True)
            log(f"FP32c k={t32.kblock}: {glups32:.1f} GLUP/s, {gbs32:.0f} GB/s "
                f"({100 * gbs32 / bw:.1f}% of stream), rel-L2 {m32[3]:.2e}, "
                f"max-abs {m32[0]:.2e}")
            return {
                "value": round(glups32, 3), "unit": "GLUP/s",
                "dtype": "f32 (compensated, register carry)"
                         if t32.kblock > 1 else "f32 (single-step passes; the carried "
                                                "residual spans one step)",
                "kblock": t32.kblock,
                "tune": {"rows": t32.rows, "waves": t32.waves, "zchunk": t32.zchunk},
                "ms_per_step": round(1e3 * wall32 / a.steps, 5),
                "achieved_gbs": round(gbs32, 1), "frac_of_stream": round(gbs32 / bw, 4),
                "algorithmic_bytes_per_lup": round(12.0 / t32.kblock, 3),
                "speedup_vs_fp64": round(glups32 / glups, 3),
                "speedup_vs_fp64_single_step": round(glups32 / glups1, 3),
                "rel_l2_vs_fp64": float(m32[3]), "max_abs_vs_fp64": float(m32[0]),
                "max_rel_vs_fp64": float(m32[4]),
                # the speed target is against the plain (single-step) FP64 kernel
                "targets": {"frac_of_stream>=0.8": bool(gbs32 / bw >= 0.8),
                            "speedup>=1.8": bool(glups32 / glups1 >= 1.8),
                            "rel_l2<=1e-5": bool(m32[3] <= 1e-5)},
            }

        # fastest kblock (the autotuner's choice, or --tune32 / --kblock)
        result["fp32c"] = fp32_leg(a.kblock, a.tune32)
        # two fused steps: the temporally blocked depth closest to the stream
        # rate; single-step passes: the stream-bound end of the family
        if result["fp32c"]["kblock"] != 2:
            result["fp32c_k2"] = fp32_leg(2)
        if result["fp32c"]["kblock"] != 1:
            result["fp32c_k1"] = fp32_leg(1)

    # ---- the other configurations, each reported beside the headline ----
    for d in (da, db, dk):
        d.close()
    if not a.no_fp32:
        for d in (fa, fbuf, fk):
            d.close()
    if not a.no_configs:
        result["configs"] = {}
        for name, fn in (("reference_workload_n512_501steps", leg_reference_workload),
                         ("slab_1024x1024x256_fp64", leg_cfg3),
                         ("fp32c_1024cubed_2000steps", leg_cfg4),
                         ("ensemble_64_members_500steps", leg_cfg5)):
            log(f"{name} ...")
            try:
                result["configs"][name] = fn(bw, stream)
            except Exception as exc:  # an extra leg never hides the headline
                result["configs"][name] = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"{name}: {json.dumps(result['configs'][name])[:300]}")

    # ---- CPU baseline ----
    if not a.no_cpu:
        log("CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline()
        log(f"CPU: {result['cpu_baseline']['value']:.3f} GLUP/s "
            f"({result['cpu_baseline']['cores']} threads)")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
