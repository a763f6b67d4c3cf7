"""The entry point that allocates nothing (h3d_stats_f64_noalloc): its
device-side results equal h3d_stats_f64's bit for bit, on two calls in a
row (the second on a smaller field, so fewer partial records than the
first left in the library's array), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def field(gpu, nx, ny, nz, seed, pitch=0):
    g = gpu.Grid(nx, ny, nz, nx + pitch, (nx + pitch) * ny)
    u, _ = gpu.gen_seeded(gpu.Grid(nx, ny, nz, nx + pitch, (nx + pitch) * ny), 0, nx, ny, nz,
                          1, seed)
    rng = np.random.default_rng(seed)
    u += rng.standard_normal(u.size) * 1e-3  # signed values, padding included
    return g, u


def test_noalloc_stats_equal_allocating_stats(gpu):
    s = gpu.Stream()
    out_dev = gpu.DeviceArray(4 * 8)
    # large first (1024 partial records), then smaller (fewer records), then
    # a field with non-finite values
    cases = [field(gpu, 300, 260, 200, 1), field(gpu, 37, 20, 11, 2, pitch=3),
             field(gpu, 65, 9, 5, 3)]
    cases[2][1][7] = np.nan
    cases[2][1][100] = np.inf
    for g, u in cases:
        du = gpu.DeviceArray.from_host(u)
        want = np.zeros(4)
        gpu.call("h3d_stats_f64", du.ptr, gpu.C.byref(g), gpu.ptr(want), s.ptr)
        out_dev.upload(np.full(4, -1.0))
        gpu.call("h3d_stats_f64_noalloc", du.ptr, gpu.C.byref(g), out_dev.ptr, s.ptr)
        s.sync()
        got = out_dev.download()
        assert got.tobytes() == want.tobytes(), (g.nx, g.ny, g.nz, got, want)
        du.close()


def test_noalloc_stats_argument_errors(gpu):
    L = gpu.lib()
    g = gpu.Grid.dense(8)
    d = gpu.DeviceArray(g.elems * 8)
    assert L.h3d_stats_f64_noalloc(None, gpu.C.byref(g), d.ptr, None) == -1
    assert L.h3d_stats_f64_noalloc(d.ptr, gpu.C.byref(g), None, None) == -1
    bad = gpu.Grid(2, 8, 8, 2, 16)
    assert L.h3d_stats_f64_noalloc(d.ptr, gpu.C.byref(bad), d.ptr, None) == -1
