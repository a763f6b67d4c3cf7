"""Ensembles on the GPU: every member of a mixed FP64 / FP32-compensated
ensemble advanced by the persistent work-list kernel equals the same
member advanced alone (FP64: h3d_run_f64; FP32: the CPU oracle replaying
the same passes, with the FP16 residual field carried across them
(precision 1) or the residual carried in registers within each pass
(precision 2)), bit for bit, eagerly and from a hipGraph; and the
64-member configuration keeps every FP32 member within rel-L2 1e-5 of its
FP64 twin after 500 steps."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_member(h3d, dims, m, seed, precision, kind=1):
    nx, ny, nz = dims
    g = h3d.Grid.dense(nx, ny, nz)
    u, k = h3d.gen_seeded(g, 0, nx, ny, nz, kind, seed, member=m)
    h = 1.0 / (nx - 1)
    if precision:
        u, k = u.astype(np.float32), k.astype(np.float32)
    k64 = k.astype(np.float64)
    fmax = h3d.lib().h3d_face_sum_max_host(h3d.ptr(k64), C.byref(g))
    factor = 0.9 * 2 * h * h / fmax / h / h / 2.0
    return dict(g=g, u=u, k=k, factor=factor, precision=precision)


def build(h3d, specs):
    mems, keep = [], []
    for s in specs:
        da, db, dk = (h3d.DeviceArray.from_host(x) for x in (s["u"], s["u"], s["k"]))
        keep.append((da, db, dk))
        mems.append(h3d.Member(da.ptr, db.ptr, dk.ptr, s["g"], s["factor"], s["precision"]))
    arr = (h3d.Member * len(mems))(*mems)
    e = C.c_void_p()
    h3d.call("h3d_ensemble_create", C.byref(e), arr, len(mems), None, 0)
    return e, keep


def where(h3d, e, m):
    in_b = C.c_int()
    h3d.call("h3d_ensemble_where", e, m, C.byref(in_b))
    return in_b.value


def pass_depths(steps, kb):
    """FP32 pass depths of one run of `steps` steps at temporal-blocking
    depth kb: full groups of kb, then single steps."""
    full = steps // kb if kb > 1 else 0
    return [kb] * full + [1] * (steps - full * kb)


def run(h3d, e, steps, graph, stream):
    ms = C.c_float()
    h3d.call("h3d_ensemble_run", e, steps, graph, stream.ptr, C.byref(ms))
    return ms.value


SPECS = [((5, 5, 5), 0), ((70, 9, 40), 1), ((66, 66, 66), 0), ((48, 131, 57), 1),
         ((130, 50, 3 + 2), 0), ((3, 80, 70), 1), ((100, 33, 150), 1), ((40, 70, 33), 2),
         ((81, 5, 66), 2)]


def f32_mode(precision):
    """The FP32 pass mode of an ensemble member's precision."""
    return 2 if precision == 1 else 1


def oracle_f32_passes(oracle, sp, depths):
    """The oracle replaying an FP32 member's passes: the FP16 residual
    carried across passes (precision 1), or reset at each pass (2)."""
    g = sp["g"]
    f32 = float(np.float32(sp["factor"]))
    u, comp = sp["u"], None
    for d in depths:
        u, comp = oracle.run_f32(u, sp["k"], None, g.nx, g.ny, g.nz, f32, d, d,
                                 f32_mode(sp["precision"]), comp0=comp)
    return u


@pytest.mark.parametrize("graph", [0, 1])
def test_members_equal_standalone_runs(gpu, oracle, graph):
    specs = [make_member(gpu, d, m, 77, p) for m, (d, p) in enumerate(SPECS)]
    e, keep = build(gpu, specs)
    s = gpu.Stream()
    steps = 7
    run(gpu, e, 3, graph, s)
    run(gpu, e, 4, graph, s)  # parity carries over between runs
    nwork, nblk, done = C.c_int(), C.c_int(), C.c_longlong()
    gpu.call("h3d_ensemble_info", e, C.byref(nwork), C.byref(nblk), C.byref(done))
    assert done.value == steps and nwork.value > len(SPECS) and nblk.value >= 1
    for sp, (da, db, dk) in zip(specs, keep):
        got = db.download()  # odd total: results in u_b
        g = sp["g"]
        if sp["precision"] == 0:
            want = oracle.run_f64(sp["u"].copy(), sp["u"].copy(), None, sp["k"], g.nx, g.ny,
                                  g.nz, steps, factor=sp["factor"])
            assert np.array_equal(got, want), g.nx
        else:
            want = oracle_f32_passes(oracle, sp, [1] * steps)
            assert got.tobytes() == want.tobytes(), g.nx
    gpu.call("h3d_ensemble_destroy", e)


@pytest.mark.parametrize("zchunk", [0, 3, 32])
def test_zchunk_and_autotune_keep_results(gpu, oracle, zchunk):
    """The z-march depth changes the work list, not the arithmetic: members
    stay bit-identical to standalone runs for a given depth, and the
    autotuner's timed launches (between two runs, at an odd step count)
    leave the members' state untouched."""
    specs = [make_member(gpu, d, m, 91, p) for m, (d, p) in enumerate(SPECS)]
    mems, keep = [], []
    for sp in specs:
        da, db, dk = (gpu.DeviceArray.from_host(x) for x in (sp["u"], sp["u"], sp["k"]))
        keep.append((da, db, dk))
        mems.append(gpu.Member(da.ptr, db.ptr, dk.ptr, sp["g"], sp["factor"], sp["precision"]))
    arr = (gpu.Member * len(mems))(*mems)
    e = C.c_void_p()
    t = gpu.Tune(0, 0, zchunk, 0)
    gpu.call("h3d_ensemble_create", C.byref(e), arr, len(mems), C.byref(t), 0)
    s = gpu.Stream()
    run(gpu, e, 3, 1, s)
    out = gpu.Tune()
    gpu.call("h3d_ensemble_autotune", e, s.ptr, C.byref(out))
    assert out.zchunk in (32, 64, 128, 256) and out.kblock in (1, 2, 3)
    run(gpu, e, 2, 1, s)
    steps = 5
    # passes: three single steps, then two steps at the tuner's depth
    depths = [1, 1, 1] + pass_depths(2, out.kblock)
    for m, (sp, (da, db, dk)) in enumerate(zip(specs, keep)):
        got = (db if where(gpu, e, m) else da).download()
        g = sp["g"]
        if sp["precision"] == 0:
            want = oracle.run_f64(sp["u"].copy(), sp["u"].copy(), None, sp["k"], g.nx, g.ny,
                                  g.nz, steps, factor=sp["factor"])
            assert np.array_equal(got, want), g.nx
        else:
            u = oracle_f32_passes(oracle, sp, depths)
            assert got.tobytes() == u.tobytes(), (g.nx, out.kblock)
    gpu.call("h3d_ensemble_destroy", e)


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("kb", [2, 3])
def test_fp32_pass_depth_members_equal_standalone(gpu, oracle, kb, graph):
    """Every member fused kb steps per pass: each FP32 member equals the
    oracle replaying the same passes with the FP16 residual carried across
    them, and each FP64 member its plain run, over runs whose lengths are
    not multiples of kb."""
    specs = [make_member(gpu, d, m, 13, p) for m, (d, p) in enumerate(SPECS)]
    check_members_over_runs(gpu, oracle, specs, kb, graph)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_single_precision_ensembles_equal_standalone(gpu, oracle, prec):
    """One persistent kernel serves FP64 and FP32 tiles; an ensemble whose
    members all have one precision takes only that path, and each member
    still equals its standalone run (3-step passes and single steps, from
    a graph)."""
    specs = [make_member(gpu, d, m, 29, prec) for m, (d, _) in enumerate(SPECS)]
    check_members_over_runs(gpu, oracle, specs, 3, 1)


def check_members_over_runs(gpu, oracle, specs, kb, graph):
    """Runs of 3, 4 and kb steps at pass depth kb, then every member
    against the oracle."""
    mems, keep = [], []
    for sp in specs:
        da, db, dk = (gpu.DeviceArray.from_host(x) for x in (sp["u"], sp["u"], sp["k"]))
        keep.append((da, db, dk))
        mems.append(gpu.Member(da.ptr, db.ptr, dk.ptr, sp["g"], sp["factor"], sp["precision"]))
    arr = (gpu.Member * len(mems))(*mems)
    e = C.c_void_p()
    t = gpu.Tune(0, 0, 0, kb)
    gpu.call("h3d_ensemble_create", C.byref(e), arr, len(mems), C.byref(t), 0)
    s = gpu.Stream()
    runs = [3, 4, kb]
    for n in runs:
        run(gpu, e, n, graph, s)
    depths = [d for n in runs for d in pass_depths(n, kb)]
    steps = sum(runs)
    for m, (sp, (da, db, dk)) in enumerate(zip(specs, keep)):
        got = (db if where(gpu, e, m) else da).download()
        g = sp["g"]
        if sp["precision"] == 0:
            # one flip per pass, as for FP32 members
            assert where(gpu, e, m) == len(depths) % 2
            want = oracle.run_f64(sp["u"].copy(), sp["u"].copy(), None, sp["k"], g.nx, g.ny,
                                  g.nz, steps, factor=sp["factor"])
            assert np.array_equal(got, want), g.nx
        else:
            assert where(gpu, e, m) == len(depths) % 2
            u = oracle_f32_passes(oracle, sp, depths)
            assert got.tobytes() == u.tobytes(), (g.nx, kb)
    gpu.call("h3d_ensemble_destroy", e)


def test_graph_equals_eager_over_many_steps(gpu):
    specs = [make_member(gpu, d, m, 5, p) for m, (d, p) in enumerate(SPECS[1:5])]
    outs = []
    for graph in (0, 1):
        e, keep = build(gpu, specs)
        s = gpu.Stream()
        run(gpu, e, 20, graph, s)
        run(gpu, e, 20, graph, s)  # a cached graph replayed
        outs.append([d[0].download() for d in keep])
        gpu.call("h3d_ensemble_destroy", e)
    for a, b in zip(*outs):
        assert a.tobytes() == b.tobytes()


def test_ensemble_rejects_bad_members(gpu):
    sp = make_member(gpu, (20, 20, 20), 0, 1, 0)
    d = gpu.DeviceArray.from_host(sp["u"])
    bad = (gpu.Member * 1)(gpu.Member(d.ptr, d.ptr, d.ptr, sp["g"], 0.1, 0))
    e = C.c_void_p()
    assert gpu.lib().h3d_ensemble_create(C.byref(e), bad, 1, None, 0) == -1
    bad2 = (gpu.Member * 1)(gpu.Member(d.ptr, d.ptr, d.ptr, sp["g"], 0.1, 3))
    assert gpu.lib().h3d_ensemble_create(C.byref(e), bad2, 1, None, 0) == -1


def ensemble64_dims(seed=20260101, count=64):
    rng = np.random.default_rng(seed)
    return [tuple(int(x) + 2 for x in rng.integers(48, 257, 3)) for _ in range(count)]


@pytest.mark.parametrize("prec32", [1, 2])
def test_ensemble64_500_steps_accuracy(gpu, prec32):
    """64 members, each axis uniform in [48, 256] interior points, lognormal
    kappa, Gaussian u0, alternating FP64 / FP32-compensated (FP16 residual
    field, or register carry), 500 steps from one hipGraph; each FP32 member
    against the same member run in FP64 on its FP32-rounded inputs."""
    dims = ensemble64_dims()
    specs = [make_member(gpu, d, m, 20260101, prec32 * (m % 2)) for m, d in enumerate(dims)]
    e, keep = build(gpu, specs)
    s = gpu.Stream()
    ms = run(gpu, e, 500, 1, s)
    cells = sum((d[0] - 2) * (d[1] - 2) * (d[2] - 2) for d in dims)
    print(f"ensemble: {cells * 500 / (ms / 1e3) / 1e9:.1f} GLUP/s aggregate")
    # FP64 twins of the FP32 members
    twins = []
    for sp in specs[1::2]:
        t = dict(sp)
        t["u"] = sp["u"].astype(np.float64)
        t["k"] = sp["k"].astype(np.float64)
        t["precision"] = 0
        twins.append(t)
    e2, keep2 = build(gpu, twins)
    run(gpu, e2, 500, 1, s)
    worst = 0.0
    for sp, (da, db, dk), (ta, tb, tk) in zip(specs[1::2], keep[1::2], keep2):
        got = da.download().astype(np.float64)  # even step count: results in u_a
        ref = ta.download()
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        worst = max(worst, rel)
    print("worst FP32 member rel-L2 vs FP64:", worst)
    assert worst <= 1e-5
    for sp, (da, db, dk) in zip(specs[0::2], keep[0::2]):
        assert np.all(np.isfinite(da.download()))
    gpu.call("h3d_ensemble_destroy", e)
    gpu.call("h3d_ensemble_destroy", e2)


@pytest.mark.parametrize("prec32", [1, 2])
def test_ensemble64_members_bitwise_first_steps(gpu, oracle, prec32):
    """The 64-member configuration at full size with every member fused 3
    steps per pass (the depth the benchmark's tuner picks): after 7 steps
    (two 3-step passes and a single step) every FP64 member equals the
    oracle's run bit for bit, and a sample of FP32 members equals the
    oracle replaying the same passes, with the FP16 residual carried across
    them (precision 1) or reset at each pass (precision 2)."""
    dims = ensemble64_dims()
    specs = [make_member(gpu, d, m, 20260101, prec32 * (m % 2)) for m, d in enumerate(dims)]
    mems, keep = [], []
    for sp in specs:
        da, db, dk = (gpu.DeviceArray.from_host(x) for x in (sp["u"], sp["u"], sp["k"]))
        keep.append((da, db, dk))
        mems.append(gpu.Member(da.ptr, db.ptr, dk.ptr, sp["g"], sp["factor"], sp["precision"]))
    arr = (gpu.Member * len(mems))(*mems)
    e = C.c_void_p()
    t = gpu.Tune(0, 0, 0, 3)
    gpu.call("h3d_ensemble_create", C.byref(e), arr, len(mems), C.byref(t), 0)
    s = gpu.Stream()
    steps = 7
    run(gpu, e, steps, 1, s)
    depths = pass_depths(steps, 3)
    checked32 = 0
    for m, (sp, (da, db, dk)) in enumerate(zip(specs, keep)):
        g = sp["g"]
        if sp["precision"] != 0 and m % 8 != 1:
            continue  # FP32 sample: members 1, 9, 17, ... (the oracle's FP32 path is serial)
        got = (db if where(gpu, e, m) else da).download()
        if sp["precision"] == 0:
            want = oracle.run_f64(sp["u"].copy(), sp["u"].copy(), None, sp["k"], g.nx, g.ny,
                                  g.nz, steps, factor=sp["factor"], nthreads=16)
            assert np.array_equal(got, want), (m, dims[m])
        else:
            u = oracle_f32_passes(oracle, sp, depths)
            assert got.tobytes() == u.tobytes(), (m, dims[m])
            checked32 += 1
    assert checked32 == 8
    gpu.call("h3d_ensemble_destroy", e)
    for trip in keep:
        for d in trip:
            d.close()
