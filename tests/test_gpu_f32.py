"""FP32-storage kernels on the GPU: every temporally blocked pass is checked
bit for bit against the CPU oracle stepping the same arithmetic one step at
a time, and the compensated paths against the FP64 kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(gpu, nx, ny, nz, seed=5, kind=1, bf16=False):
    g = gpu.Grid.dense(nx, ny, nz)
    u, k = gpu.gen_seeded(g, 0, nx, ny, nz, kind, seed)
    h = 1.0 / (nx - 1)
    if bf16:
        bits, s = gpu.quantize_bf16(k)
        k64 = gpu.bf16_to_f64(bits, s)
    else:
        bits, s = None, 1.0
        k64 = k.astype(np.float32).astype(np.float64)
    fmax = gpu.lib().h3d_face_sum_max_host(gpu.ptr(k64), gpu.C.byref(g))
    dt = 0.9 * 2 * h * h / fmax
    factor = dt / h / h / 2.0
    return dict(g=g, u32=u.astype(np.float32), k32=k.astype(np.float32), bits=bits,
                scale=s, k64=k64, factor=factor)


def gpu_run(gpu, P, steps, mode, tune, graph=0):
    g = P["g"]
    da = gpu.DeviceArray.from_host(P["u32"])
    db = gpu.DeviceArray.from_host(P["u32"])
    if P["bits"] is not None:
        dk = gpu.DeviceArray.from_host(P["bits"])
        kf, kb = None, dk.ptr
    else:
        dk = gpu.DeviceArray.from_host(P["k32"])
        kf, kb = dk.ptr, None
    ca = gpu.DeviceArray.from_host(np.zeros(g.elems, np.uint16))
    cb = gpu.DeviceArray.from_host(np.zeros(g.elems, np.uint16))
    f32f = np.float32(P["factor"] * P["scale"])
    s = gpu.Stream()
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    gpu.call("h3d_run_f32", da.ptr, db.ptr, kf, kb, ca.ptr, cb.ptr, gpu.C.byref(g),
             gpu.C.c_float(f32f), steps, mode, gpu.C.byref(tune), graph, s.ptr,
             gpu.C.byref(ms), gpu.C.byref(fb))
    u = (db if fb.value else da).download()
    c = (cb if fb.value else ca).download()
    return u, c


def oracle_run(oracle, P, steps, kb, mode):
    g = P["g"]
    f32f = float(np.float32(P["factor"] * P["scale"]))
    if P["bits"] is not None:
        return oracle.run_f32(P["u32"], None, P["bits"], g.nx, g.ny, g.nz, f32f, steps, kb, mode)
    return oracle.run_f32(P["u32"], P["k32"], None, g.nx, g.ny, g.nz, f32f, steps, kb, mode)


CASES = [
    # shape, kblock, rows, waves, zchunk
    ((20, 18, 16), 1, 2, 4, 0),
    ((70, 40, 30), 2, 2, 8, 0),
    ((70, 40, 30), 2, 4, 4, 7),
    ((65, 33, 21), 3, 2, 8, 5),
    ((130, 35, 19), 4, 4, 8, 0),
    ((9, 75, 12), 4, 2, 8, 3),
    ((61, 61, 13), 6, 4, 4, 4),
    ((5, 5, 5), 2, 2, 4, 1),
    ((70, 40, 30), 2, 4, 8, 29),
    ((9, 75, 12), 1, 4, 4, 11),
    ((70, 40, 30), 1, 1, 2, 0),
    ((130, 35, 19), 1, 4, 8, 7),
    # 16-wave tiles (one 1024-thread workgroup per CU) of 2..6 rows per lane
    ((70, 40, 30), 2, 2, 16, 0),
    ((70, 100, 30), 2, 4, 16, 13),
    ((130, 90, 21), 2, 5, 16, 0),
    ((70, 110, 17), 2, 6, 16, 9),
    ((130, 70, 21), 3, 4, 16, 0),
    ((65, 50, 21), 3, 3, 16, 7),
    # deep passes: 7 and 8 steps on 8-wave tiles of 3 and 4 rows (the
    # remainder pass of a run on the depth's default shape where the
    # single-step tile has no 3-row form)
    ((130, 35, 40), 8, 3, 8, 0),
    ((70, 60, 30), 8, 4, 8, 11),
    ((65, 40, 25), 7, 3, 8, 0),
    ((65, 40, 25), 5, 4, 8, 0),
]


@pytest.mark.parametrize("case", CASES, ids=str)
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("bf16", [False, True])
def test_pass_matches_oracle_bitwise(gpu, oracle, case, mode, bf16):
    shape, kb, rows, waves, zc = case
    P = problem(gpu, *shape, bf16=bf16)
    steps = 2 * kb + 1  # two full passes and a remainder pass
    t = gpu.Tune(rows, waves, zc, kb)
    got_u, got_c = gpu_run(gpu, P, steps, mode, t)
    want_u, want_c = oracle_run(oracle, P, steps, kb, mode)
    assert got_u.tobytes() == want_u.tobytes()
    if mode == 2:
        g = P["g"]
        inner = np.zeros((g.nz, g.ny, g.nx), bool)
        inner[1:-1, 1:-1, 1:-1] = True
        inner = inner.ravel()
        assert np.array_equal(got_c[inner], want_c[inner])


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("kb", [1, 2, 3, 5])
@pytest.mark.parametrize("shape", [(3, 3, 3), (3, 5, 4), (4, 3, 6)], ids=str)
def test_pass_minimum_grids(gpu, oracle, shape, kb, mode):
    """Grids of 3 and 4 points on an axis: single steps and passes of 2, 3
    and 5 steps on each depth's default shape give the oracle's bits."""
    P = problem(gpu, *shape)
    steps = 2 * kb + 1
    got_u, _ = gpu_run(gpu, P, steps, mode, gpu.Tune(0, 0, 0, kb))
    want_u, _ = oracle_run(oracle, P, steps, kb, mode)
    assert got_u.tobytes() == want_u.tobytes()


def test_plain_fused_pass_equals_single_steps(gpu):
    P = problem(gpu, 100, 60, 40, seed=2)
    a, _ = gpu_run(gpu, P, 8, 0, gpu.Tune(2, 8, 0, 4))
    b, _ = gpu_run(gpu, P, 8, 0, gpu.Tune(2, 8, 0, 1))
    assert a.tobytes() == b.tobytes()


def test_graph_replay_equals_eager(gpu):
    P = problem(gpu, 90, 50, 35, seed=4)
    t = gpu.Tune(2, 8, 0, 2)
    a, ca = gpu_run(gpu, P, 21, 2, t, graph=0)
    b, cb = gpu_run(gpu, P, 21, 2, t, graph=1)
    assert a.tobytes() == b.tobytes() and ca.tobytes() == cb.tobytes()


def test_compensation_accuracy_vs_f64(gpu):
    """1000 steps at 98^3: compensated FP32 (register carry and FP16
    residual) stays within rel-L2 1e-5 of the FP64 kernel run on the same
    rounded inputs; plain FP32 is reported alongside."""
    n = 98
    P = problem(gpu, n, n, n, seed=20260101, kind=0)
    g = P["g"]
    u64 = P["u32"].astype(np.float64)
    da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u64, u64, P["k64"]))
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    steps = 1000
    gpu.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, gpu.C.byref(g), P["factor"],
             steps, None, 0, None, gpu.C.byref(ms), gpu.C.byref(fb))
    ref = db if fb.value else da
    err = {}
    for mode, kb in [(0, 2), (1, 2), (1, 4), (2, 2)]:
        u, _ = gpu_run(gpu, P, steps, mode, gpu.Tune(2, 8, 0, kb))
        m = gpu.compare(ref, gpu.DeviceArray.from_host(u), g, f32=True)
        err[(mode, kb)] = m[3]
    print("rel-L2 vs FP64:", err)
    assert err[(1, 2)] <= 1e-5 and err[(1, 4)] <= 1e-5 and err[(2, 2)] <= 1e-5
    assert err[(2, 2)] <= err[(0, 2)]


def test_pass_argument_errors(gpu):
    g = gpu.Grid.dense(20)
    d1, d2 = gpu.DeviceArray(g.elems * 4), gpu.DeviceArray(g.elems * 4)
    L = gpu.lib()
    f = gpu.C.c_float(0.01)
    assert L.h3d_pass_f32(d1.ptr, d2.ptr, None, None, None, None, gpu.C.byref(g), f, 1, 0, None, None) == -1
    assert L.h3d_pass_f32(d1.ptr, d2.ptr, d2.ptr, None, None, None, gpu.C.byref(g), f, 9, 0, None, None) == -2
    assert L.h3d_pass_f32(d1.ptr, d2.ptr, d2.ptr, None, None, None, gpu.C.byref(g), f, 1, 2, None, None) == -1
    t = gpu.Tune(2, 2, 0, 4)  # 4 rows per workgroup cannot hold a 2*4 halo
    assert L.h3d_pass_f32(d1.ptr, d2.ptr, d2.ptr, None, None, None, gpu.C.byref(g), f, 4, 0,
                          gpu.C.byref(t), None) == -1


@pytest.mark.parametrize("kb", [1, 2])
def test_largest_plane_stride_bitwise(gpu, oracle, kb):
    """Planes of 2^28 elements (1 GiB of FP32), the largest the 32-bit
    in-plane byte offsets allow, for the single-step and the fused tile:
    bit-identical to the same values stored densely."""
    from test_gpu_f64 import get_rows, put_rows
    nx, ny, nz = 70, 9, 5
    sy, sz = 72, 1 << 28
    P = problem(gpu, nx, ny, nz)
    steps = 2 * kb
    want, _ = oracle.run_f32(P["u32"], P["k32"], None, nx, ny, nz,
                             float(np.float32(P["factor"])), steps, kb, 1)
    g = gpu.Grid(nx, ny, nz, sy, sz)
    bufs = [gpu.DeviceArray(sz * nz * 4, dtype=np.float32) for _ in range(3)]
    for b, a in zip(bufs, (P["u32"], P["u32"], P["k32"])):
        put_rows(gpu, b, a, nx, ny, nz, sy, sz)
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    t = gpu.Tune(0, 0, 0, kb)
    gpu.call("h3d_run_f32", bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, None, None, None,
             gpu.C.byref(g), gpu.C.c_float(np.float32(P["factor"])), steps, 1, gpu.C.byref(t),
             0, None, gpu.C.byref(ms), gpu.C.byref(fb))
    got = get_rows(gpu, bufs[1] if fb.value else bufs[0], np.float32, nx, ny, nz, sy, sz)
    assert got.tobytes() == want.tobytes()
    for b in bufs:
        b.close()
