"""FP64 kernel parity on the GPU: bit-exact against the reference's own
serial program (golden fixtures) and against the CPU oracle."""
import gzip
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INDEX = json.load(open(os.path.join(GOLDEN, "index.json")))


@pytest.mark.parametrize("entry", INDEX, ids=lambda e: f"n{e['n']}_T{e['T']}")
def test_reference_mirror_matches_golden(gpu, entry):
    """h3d_heat_update (the reference's heat_update signature) in the
    reference's own time loop reproduces u_serial.txt byte for byte."""
    n = entry["n"]
    h, dt, factor = gpu.params_reference(n)
    steps = gpu.count_steps(float(entry["T"]), dt)
    assert steps == entry["steps"]
    u, rhs, k = gpu.init_reference(n)
    du_old, du_new = gpu.DeviceArray.from_host(u), gpu.DeviceArray.from_host(u)
    dr, dk = gpu.DeviceArray.from_host(rhs), gpu.DeviceArray.from_host(k)
    for _ in range(steps):
        gpu.call("h3d_heat_update", du_new.ptr, du_old.ptr, dr.ptr, dk.ptr, n, n, n, factor, 0)
        du_old, du_new = du_new, du_old
    gpu.call("h3d_device_synchronize", 0)
    res = du_old.download()
    with gzip.open(os.path.join(GOLDEN, entry["file"]), "rt") as f:
        want = f.read().splitlines()
    assert ["%.17g" % v for v in res] == want


SHAPES = [
    (5, 5, 5), (66, 66, 66), (67, 45, 33), (130, 9, 12), (3, 70, 8), (200, 3, 4),
]
TUNES = [None, (1, 1, 1), (2, 4, 7), (4, 2, 0), (8, 8, 3), (1, 16, 64)]


def seeded(gpu, nx, ny, nz, seed=11, kind=0, pitch=0, rows_pad=0):
    g = gpu.Grid(nx, ny, nz, nx + pitch, (nx + pitch) * (ny + rows_pad))
    u, k = gpu.gen_seeded(g, 0, nx, ny, nz, kind, seed)
    return g, u, k


def oracle_steps(oracle, g, u, k, factor, steps, rhs=None):
    return oracle.run_f64(u.copy(), u.copy(), rhs, k, g.nx, g.ny, g.nz, steps,
                          sy=g.sy, sz=g.sz, factor=factor)


@pytest.mark.parametrize("shape", SHAPES, ids=str)
@pytest.mark.parametrize("tune", TUNES, ids=str)
def test_step_matches_oracle_bitwise(gpu, oracle, shape, tune):
    nx, ny, nz = shape
    g, u, k = seeded(gpu, nx, ny, nz, kind=1)
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz) / 2.0
    steps = 3
    want = oracle_steps(oracle, g, u, k, factor, steps)
    t = gpu.Tune(*tune, 1) if tune else None
    da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u, u, k))
    s = gpu.Stream()
    for i in range(steps):
        src, dst = (da, db) if i % 2 == 0 else (db, da)
        gpu.call("h3d_step_f64", dst.ptr, src.ptr, None, dk.ptr, gpu.C.byref(g), factor,
                 1, nz - 2, gpu.C.byref(t) if t else None, s.ptr)
    s.sync()
    got = (db if steps % 2 else da).download()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(64, 20, 12), (80, 33, 9), (144, 17, 10), (16, 16, 5),
                                   (48, 70, 6)], ids=str)
@pytest.mark.parametrize("shift", [0, 1, 2, 7, 15], ids=lambda s: f"shift{s}")
@pytest.mark.parametrize("with_rhs", [False, True], ids=["norhs", "rhs"])
def test_line_aligned_tile_origins(gpu, oracle, shape, shift, with_rhs):
    """Dense rows a whole number of 128-byte lines long (nx a multiple of
    16): the tiles' x origin follows the line offset of u_new, here a field
    placed `shift` elements into its buffer, so lane 0 of a tile sits at
    x = 0, at x = 1 or left of the array (x0 = 0, 1, -49, -50, -55 on a
    line-aligned allocation).  Bit-identical to the oracle either way."""
    C = gpu.C
    nx, ny, nz = shape
    g, u, k = seeded(gpu, nx, ny, nz, kind=1)
    rhs = np.random.default_rng(9).uniform(-1e-4, 1e-4, g.elems) if with_rhs else None
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz) / 2.0
    steps = 3
    want = oracle_steps(oracle, g, u, k, factor, steps, rhs=rhs)
    nbytes = g.elems * 8
    bufs = [gpu.DeviceArray(nbytes + 256) for _ in range(4)]
    pa, pb, pk, pr = (C.c_void_p(b.ptr.value + 8 * shift) for b in bufs)
    for p, x in ((pa, u), (pb, u), (pk, k)) + (((pr, rhs),) if with_rhs else ()):
        gpu.call("h3d_target_memcpy", p, gpu.ptr(x), C.c_size_t(nbytes), 0, -1)
    s = gpu.Stream()
    for i in range(steps):
        src, dst = (pa, pb) if i % 2 == 0 else (pb, pa)
        gpu.call("h3d_step_f64", dst, src, pr if with_rhs else None, pk, C.byref(g), factor,
                 1, nz - 2, None, s.ptr)
    s.sync()
    got = np.empty(g.elems)
    gpu.call("h3d_target_memcpy", gpu.ptr(got), pb if steps % 2 else pa, C.c_size_t(nbytes),
             -1, 0)
    assert np.array_equal(got, want)


def test_pitched_layout_and_rhs(gpu, oracle):
    nx, ny, nz = 70, 21, 17
    g, u, k = seeded(gpu, nx, ny, nz, pitch=10, rows_pad=3)
    rhs = np.random.default_rng(4).uniform(-1e-4, 1e-4, g.elems)
    factor = 0.02
    want = oracle_steps(oracle, g, u, k, factor, 4, rhs=rhs)
    da, db, dk, dr = (gpu.DeviceArray.from_host(x) for x in (u, u, k, rhs))
    ms = gpu.C.c_float()
    fb = gpu.C.c_int()
    gpu.call("h3d_run_f64", da.ptr, db.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor, 4,
             None, 0, None, gpu.C.byref(ms), gpu.C.byref(fb))
    got = (db if fb.value else da).download()
    assert fb.value == 0
    assert np.array_equal(got, want)
    # padding between rows and planes is never written
    pad = got.reshape(nz, ny + 3, nx + 10)
    assert np.array_equal(pad[:, :, nx:], u.reshape(nz, ny + 3, nx + 10)[:, :, nx:])


def test_interior_and_boundary_mirrors_equal_full_step(gpu, oracle):
    n = 40
    h, dt, factor = gpu.params_reference(n)
    u, rhs, k = gpu.init_reference(n)
    da, dr, dk = (gpu.DeviceArray.from_host(x) for x in (u, rhs, k))
    full, split = gpu.DeviceArray.from_host(u), gpu.DeviceArray.from_host(u)
    gpu.call("h3d_heat_update", full.ptr, da.ptr, dr.ptr, dk.ptr, n, n, n, factor, 0)
    gpu.call("h3d_heat_update_boundary_bottom", split.ptr, da.ptr, dr.ptr, dk.ptr, n, n, n, factor, 0)
    gpu.call("h3d_heat_update_boundary_top", split.ptr, da.ptr, dr.ptr, dk.ptr, n, n, n, factor, 0)
    gpu.call("h3d_heat_update_interior", split.ptr, da.ptr, dr.ptr, dk.ptr, n, n, n, factor,
             2, n - 3, 0)
    gpu.call("h3d_device_synchronize", 0)
    a, b = full.download(), split.download()
    assert np.array_equal(a, b)
    want = u.copy()
    oracle.step_f64(want, u, rhs, k, n, n, n, factor, 1, n - 2)
    assert np.array_equal(a, want)


def test_graph_replay_equals_eager(gpu):
    nx, ny, nz = 80, 50, 40
    g, u, k = seeded(gpu, nx, ny, nz, kind=1, seed=3)
    factor = 0.01
    res = []
    s = gpu.Stream()
    for graph in (0, 1):
        da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u, u, k))
        ms, fb = gpu.C.c_float(), gpu.C.c_int()
        gpu.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, gpu.C.byref(g), factor, 11,
                 None, graph, s.ptr, gpu.C.byref(ms), gpu.C.byref(fb))
        res.append((db if fb.value else da).download())
        # replay again on the same buffers: graph replays after eager work
        gpu.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, gpu.C.byref(g), factor, 6,
                 None, graph, s.ptr, gpu.C.byref(ms), gpu.C.byref(fb))
        res.append(da.download())
    assert np.array_equal(res[0], res[2])
    assert np.array_equal(res[1], res[3])


def test_128_cubed_100_steps(gpu, oracle):
    """128^3 interior, seeded sine kappa, dt = 0.9 * bound, 100 steps: the
    GPU must agree with the CPU oracle to <= 1e-12 max relative error; with
    contraction off and the same association order it is bit-exact."""
    m = 128
    n = m + 2
    g, u, k = seeded(gpu, n, n, n, seed=20260101)
    h = 1.0 / (n - 1)
    dt = 0.9 * 2 * h * h / gpu.lib().h3d_face_sum_max_host(gpu.ptr(k), gpu.C.byref(g))
    factor = dt / h / h / 2.0
    want = oracle.run_f64(u.copy(), u.copy(), None, k, n, n, n, 100, factor=factor, nthreads=8)
    da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u, u, k))
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    gpu.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, gpu.C.byref(g), factor, 100,
             None, 0, None, gpu.C.byref(ms), gpu.C.byref(fb))
    got = (db if fb.value else da).download()
    m_ = oracle.metrics(want, got)
    assert m_[4] <= 1e-12
    assert np.array_equal(got, want)


def test_argument_errors(gpu):
    g = gpu.Grid.dense(10)
    d = gpu.DeviceArray(g.elems * 8)
    e = gpu.DeviceArray(g.elems * 8)
    L = gpu.lib()
    assert L.h3d_step_f64(d.ptr, d.ptr, None, d.ptr, gpu.C.byref(g), 0.1, 1, 8, None, None) == -1
    assert L.h3d_step_f64(e.ptr, d.ptr, None, d.ptr, gpu.C.byref(g), 0.1, 0, 8, None, None) == -1
    assert L.h3d_step_f64(e.ptr, d.ptr, None, d.ptr, gpu.C.byref(g), 0.1, 1, 9, None, None) == -1
    bad = gpu.Tune(3, 1, 1, 1)
    assert L.h3d_step_f64(e.ptr, d.ptr, None, d.ptr, gpu.C.byref(g), 0.1, 1, 8,
                          gpu.C.byref(bad), None) == -1
    assert L.h3d_heat_update(e.ptr, d.ptr, None, d.ptr, 10, 10, 10, 0.1, 7) == -3
    # an empty range is a no-op, as the reference's loops are
    assert L.h3d_step_f64(e.ptr, d.ptr, None, d.ptr, gpu.C.byref(g), 0.1, 5, 4, None, None) == 0


def put_rows(gpu, buf, a, nx, ny, nz, sy, sz):
    """Copy dense [k][j][i] rows of `a` into a strided device buffer."""
    item = a.itemsize
    for kk in range(nz):
        for j in range(ny):
            row = np.ascontiguousarray(a[(kk * ny + j) * nx:(kk * ny + j + 1) * nx])
            dst = gpu.C.c_void_p(buf.ptr.value + (kk * sz + j * sy) * item)
            gpu.call("h3d_target_memcpy", dst, gpu.ptr(row), gpu.C.c_size_t(row.nbytes), 0, -1)


def get_rows(gpu, buf, dtype, nx, ny, nz, sy, sz):
    out = np.empty(nx * ny * nz, dtype)
    item = out.itemsize
    for kk in range(nz):
        for j in range(ny):
            row = np.empty(nx, dtype)
            src = gpu.C.c_void_p(buf.ptr.value + (kk * sz + j * sy) * item)
            gpu.call("h3d_target_memcpy", gpu.ptr(row), src, gpu.C.c_size_t(row.nbytes), -1, 0)
            out[(kk * ny + j) * nx:(kk * ny + j + 1) * nx] = row
    return out


@pytest.mark.parametrize("kb", [1, 2], ids=["step", "fused"])
def test_largest_plane_stride_bitwise(gpu, oracle, kb):
    """Planes of 2^28 elements (2 GiB of FP64), the largest the kernels'
    32-bit in-plane byte offsets allow: a thin problem laid out with that
    plane stride steps bit-identically to the same values stored densely
    (the last planes start 6 GiB into the buffers), with single steps and
    with a 2-step fused pass plus a single step (whose out-of-range store
    offset, 2^31 bytes, equals such a plane's size)."""
    nx, ny, nz = 70, 5, 4
    sy, sz = 72, 1 << 28
    gd = gpu.Grid.dense(nx, ny, nz)
    u, k = gpu.gen_seeded(gd, 0, nx, ny, nz, 1, 5)
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz) / 2.0
    want = oracle.run_f64(u.copy(), u.copy(), None, k, nx, ny, nz, 3, factor=factor)
    g = gpu.Grid(nx, ny, nz, sy, sz)
    bufs = [gpu.DeviceArray(sz * nz * 8) for _ in range(3)]
    for b, a in zip(bufs, (u, u, k)):
        put_rows(gpu, b, a, nx, ny, nz, sy, sz)
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    t = gpu.Tune(2, 8, 0, 2) if kb == 2 else None
    gpu.call("h3d_run_f64", bufs[0].ptr, bufs[1].ptr, None, bufs[2].ptr, gpu.C.byref(g),
             factor, 3, gpu.C.byref(t) if t else None, 0, None, gpu.C.byref(ms),
             gpu.C.byref(fb))
    got = get_rows(gpu, bufs[1] if fb.value else bufs[0], np.float64, nx, ny, nz, sy, sz)
    assert np.array_equal(got, want)
    for b in bufs:
        b.close()


PASS_CASES = [
    # shape, kblock, rows, waves, zchunk, rhs (shapes whose LDS fits 160 KB)
    ((70, 40, 30), 2, 2, 8, 0, False),
    ((70, 40, 30), 2, 2, 8, 0, True),
    ((70, 40, 30), 2, 4, 4, 7, True),
    ((65, 33, 21), 3, 2, 8, 5, False),
    ((65, 33, 21), 3, 2, 8, 5, True),
    ((70, 40, 30), 3, 4, 8, 6, True),   # 32-row tiles at 3 and 4 steps
    ((70, 70, 17), 4, 4, 8, 0, False),
    ((70, 40, 30), 3, 2, 16, 7, True),  # 32 rows as 16 waves of 2 (160 KB)
    ((66, 66, 40), 2, 2, 16, 0, False),
    ((130, 35, 19), 4, 4, 4, 0, False),
    ((9, 75, 12), 2, 4, 8, 3, False),
    ((9, 75, 12), 3, 4, 4, 3, True),
    ((5, 5, 5), 2, 2, 4, 1, False),
    ((5, 5, 5), 2, 2, 4, 1, True),
    ((61, 61, 13), 3, 4, 4, 4, True),
    ((66, 66, 66), 2, 4, 8, 16, False),
    ((66, 66, 66), 2, 4, 4, 16, True),
    ((40, 20, 50), 3, 2, 4, 6, False),  # two output rows per tile
]


@pytest.mark.parametrize("kb", [2, 3, 4])
@pytest.mark.parametrize("shape,with_rhs", [((3, 3, 3), False), ((3, 5, 4), True),
                                            ((4, 3, 6), False), ((6, 4, 3), True)], ids=str)
def test_fused_pass_minimum_grids(gpu, oracle, shape, with_rhs, kb):
    """Grids of 3 and 4 points on an axis (one or two interior points): a
    tile and its warm-up planes reach far past the array on every side, and
    passes on each depth's default shape still give the oracle's bits with
    the ring untouched."""
    nx, ny, nz = shape
    g, u, k = seeded(gpu, nx, ny, nz, kind=1)
    rhs = np.random.default_rng(5).uniform(-1e-4, 1e-4, g.elems) if with_rhs else None
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz) / 2.0
    steps = 2 * kb + 1
    want = oracle_steps(oracle, g, u, k, factor, steps, rhs=rhs)
    da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u, u, k))
    dr = gpu.DeviceArray.from_host(rhs) if with_rhs else None
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    t = gpu.Tune(0, 0, 0, kb)
    gpu.call("h3d_run_f64", da.ptr, db.ptr, dr.ptr if dr else None, dk.ptr, gpu.C.byref(g),
             factor, steps, gpu.C.byref(t), 0, None, gpu.C.byref(ms), gpu.C.byref(fb))
    got = (db if fb.value else da).download()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("extra", [0, 1], ids=["2k+1", "3k+k-1"])
@pytest.mark.parametrize("case", PASS_CASES, ids=str)
def test_fused_pass_matches_oracle_bitwise(gpu, oracle, case, extra):
    """Temporally blocked FP64 passes (h3d_run_f64 with kblock > 1): full
    passes and a remainder of 1 or kblock-1 steps, bit-identical to the
    oracle stepping the reference's update one step at a time, in the
    buffer the pass count (h3d_run_f64_passes) implies."""
    shape, kb, rows, waves, zc, with_rhs = case
    nx, ny, nz = shape
    g, u, k = seeded(gpu, nx, ny, nz, kind=1)
    rhs = np.random.default_rng(3).uniform(-1e-4, 1e-4, g.elems) if with_rhs else None
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz) / 2.0
    steps = 2 * kb + 1 if extra == 0 else 4 * kb - 1
    want = oracle_steps(oracle, g, u, k, factor, steps, rhs=rhs)
    da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u, u, k))
    dr = gpu.DeviceArray.from_host(rhs) if with_rhs else None
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    t = gpu.Tune(rows, waves, zc, kb)
    gpu.call("h3d_run_f64", da.ptr, db.ptr, dr.ptr if dr else None, dk.ptr, gpu.C.byref(g),
             factor, steps, gpu.C.byref(t), 0, None, gpu.C.byref(ms), gpu.C.byref(fb))
    got = (db if fb.value else da).download()
    npass = gpu.C.c_int()
    gpu.call("h3d_run_f64_passes", gpu.C.byref(g), 1 if with_rhs else 0, steps, gpu.C.byref(t),
             gpu.C.byref(npass), None)
    assert npass.value in (steps // kb, -(-steps // kb))
    assert fb.value == npass.value % 2
    assert np.array_equal(got, want)


RANGE_CASES = [
    # shape, nsteps, k_start, k_end, rows, waves, zchunk, rhs, row pitch
    ((40, 30, 24), 2, 5, 11, 2, 8, 0, False, 0),
    ((40, 30, 24), 3, 1, 1, 2, 8, 0, True, 0),     # the bottom plane alone
    ((40, 30, 24), 3, 22, 22, 2, 4, 0, False, 0),  # the top plane alone
    ((70, 21, 30), 2, 3, 26, 4, 4, 5, True, 10),   # chunked, pitched
    ((33, 40, 20), 4, 4, 15, 4, 4, 3, False, 0),
    ((40, 30, 24), 1, 6, 9, 2, 2, 0, True, 0),     # one step: h3d_step_f64
    ((33, 40, 20), 4, 1, 18, 0, 0, 0, True, 0),    # default shape at 4 steps
    ((40, 30, 24), 3, 5, 11, 2, 16, 5, True, 0),   # 16 waves of 2 rows
]


@pytest.mark.parametrize("case", RANGE_CASES, ids=str)
def test_fused_pass_plane_range(gpu, oracle, case):
    """A pass over planes k_start..k_end writes exactly those planes'
    interior cells, with the bits of nsteps oracle steps over the whole
    array; every other value of u_out (ring, padding, other planes) keeps
    what it held."""
    shape, ns, ks, ke, rows, waves, zc, with_rhs, pitch = case
    nx, ny, nz = shape
    g, u, k = seeded(gpu, nx, ny, nz, kind=1, pitch=pitch)
    rhs = np.random.default_rng(4).uniform(-1e-4, 1e-4, g.elems) if with_rhs else None
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz, g.sy, g.sz) / 2.0
    full = oracle_steps(oracle, g, u, k, factor, ns, rhs=rhs)
    pre = np.random.default_rng(9).standard_normal(g.elems)
    want = pre.copy()
    w3, f3 = (a.reshape(nz, ny, g.sy) for a in (want, full))
    w3[ks:ke + 1, 1:ny - 1, 1:nx - 1] = f3[ks:ke + 1, 1:ny - 1, 1:nx - 1]
    di, do, dk = (gpu.DeviceArray.from_host(x) for x in (u, pre, k))
    dr = gpu.DeviceArray.from_host(rhs) if with_rhs else None
    t = gpu.Tune(rows, waves, zc, ns)
    s = gpu.Stream()
    gpu.call("h3d_pass_f64", do.ptr, di.ptr, dr.ptr if dr else None, dk.ptr, gpu.C.byref(g),
             factor, ns, ks, ke, gpu.C.byref(t), s.ptr)
    s.sync()
    assert np.array_equal(do.download(), want)


def test_fused_pass_pitched_and_graph(gpu, oracle):
    """Padded rows and planes are never written; the hipGraph replay of
    fused passes equals the eager loop."""
    nx, ny, nz = 70, 21, 17
    g, u, k = seeded(gpu, nx, ny, nz, pitch=10, rows_pad=3)
    rhs = np.random.default_rng(5).uniform(-1e-4, 1e-4, g.elems)
    factor = 0.02
    steps = 9
    want = oracle_steps(oracle, g, u, k, factor, steps, rhs=rhs)
    t = gpu.Tune(2, 8, 5, 2)
    s = gpu.Stream()
    outs = []
    for graph in (0, 1):
        da, db, dk, dr = (gpu.DeviceArray.from_host(x) for x in (u, u, k, rhs))
        ms, fb = gpu.C.c_float(), gpu.C.c_int()
        gpu.call("h3d_run_f64", da.ptr, db.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor, steps,
                 gpu.C.byref(t), graph, s.ptr, gpu.C.byref(ms), gpu.C.byref(fb))
        outs.append((db if fb.value else da).download())
    assert np.array_equal(outs[0], want) and np.array_equal(outs[1], want)
    pad = outs[0].reshape(nz, ny + 3, nx + 10)
    assert np.array_equal(pad[:, :, nx:], u.reshape(nz, ny + 3, nx + 10)[:, :, nx:])
    assert np.array_equal(pad[:, ny:, :], u.reshape(nz, ny + 3, nx + 10)[:, ny:, :])


@pytest.mark.parametrize("kb", [2, 3, 4])
def test_fused_pass_equals_single_steps_258(gpu, kb):
    """At 256^3 interior points (every z chunk, tile and XCD remap in use):
    kb-step passes give the bits of single steps."""
    n = 258
    g = gpu.pitched_grid(n, n, n)
    u, k = gpu.gen_seeded(g, 0, n, n, n, 0, 20260101)
    factor = 0.9 * 2.0 / gpu.lib().h3d_face_sum_max_host(gpu.ptr(k), gpu.C.byref(g)) / 2.0
    steps = 2 * kb
    res = []
    rows, waves = (4, 4) if kb == 4 else (2, 8)  # shapes whose LDS fits
    for t in (gpu.Tune(2, 4, 0, 1), gpu.Tune(rows, waves, 32, kb)):
        fs = []
        for x in (u, u, k):
            f = gpu.AlignedField(g, np.float64)
            f.upload(x)
            fs.append(f)
        ms, fb = gpu.C.c_float(), gpu.C.c_int()
        gpu.call("h3d_run_f64", fs[0].ptr, fs[1].ptr, None, fs[2].ptr, gpu.C.byref(g), factor,
                 steps, gpu.C.byref(t), 0, None, gpu.C.byref(ms), gpu.C.byref(fb))
        res.append((fs[1] if fb.value else fs[0]).download())
    assert np.array_equal(res[0], res[1])


def test_fused_pass_argument_errors(gpu):
    g = gpu.Grid.dense(20)
    d1, d2 = gpu.DeviceArray(g.elems * 8), gpu.DeviceArray(g.elems * 8)
    L = gpu.lib()
    f = 0.01
    assert L.h3d_pass_f64(d1.ptr, d1.ptr, None, d2.ptr, gpu.C.byref(g), f, 2, 1, 18, None, None) == -1
    assert L.h3d_pass_f64(d1.ptr, d2.ptr, None, None, gpu.C.byref(g), f, 2, 1, 18, None, None) == -1
    assert L.h3d_pass_f64(d1.ptr, d2.ptr, None, d2.ptr, gpu.C.byref(g), f, 5, 1, 18, None, None) == -2
    assert L.h3d_pass_f64(d1.ptr, d2.ptr, None, d2.ptr, gpu.C.byref(g), f, 0, 1, 18, None, None) == -2
    t = gpu.Tune(2, 4, 0, 4)  # 8 rows cannot hold a 2*4 halo
    assert L.h3d_pass_f64(d1.ptr, d2.ptr, None, d2.ptr, gpu.C.byref(g), f, 4, 1, 18, gpu.C.byref(t),
                          None) == -1
    t = gpu.Tune(2, 8, 0, 5)
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    assert L.h3d_run_f64(d1.ptr, d2.ptr, None, d2.ptr, gpu.C.byref(g), f, 5, gpu.C.byref(t), 0,
                         None, gpu.C.byref(ms), gpu.C.byref(fb)) == -2


def test_tuned_run_with_deeper_remainder_passes(gpu, oracle):
    """The autotuner's free FP64 search also times the next pass depth; a
    run whose step count leaves a remainder may then carry it on deeper
    passes (h3d_run_f64_passes says how many).  Either way the result is
    the bits of single steps, in the buffer the pass count implies."""
    C = gpu.C
    nx, ny, nz = 134, 134, 134
    g, u, k = seeded(gpu, nx, ny, nz, kind=1)
    factor = 0.9 * 2.0 / oracle.face_sum_max(k, nx, ny, nz) / 2.0
    da, db, dk = (gpu.DeviceArray.from_host(x) for x in (u, u, k))
    t = gpu.Tune()
    gpu.call("h3d_autotune", 0, da.ptr, db.ptr, dk.ptr, 0, C.byref(g), 0, 0, None, C.byref(t))
    kb = max(1, t.kblock)
    for steps in (2 * kb + 1, 3 * kb + kb - 1):
        n, d = C.c_int(), C.c_int()
        gpu.call("h3d_run_f64_passes", C.byref(g), 0, steps, C.byref(t), C.byref(n), C.byref(d))
        full, rem = divmod(steps, kb)
        assert n.value in (full, full + (1 if rem else 0))
        assert d.value == (rem if n.value == full and rem else 0)
        want = oracle_steps(oracle, g, u, k, factor, steps)
        da.upload(u)
        db.upload(u)
        ms, fb = C.c_float(), C.c_int()
        gpu.call("h3d_run_f64", da.ptr, db.ptr, None, dk.ptr, C.byref(g), factor, steps,
                 C.byref(t), 0, None, C.byref(ms), C.byref(fb))
        assert fb.value == n.value % 2
        assert np.array_equal((db if fb.value else da).download(), want)
