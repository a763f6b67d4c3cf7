"""Argument checks of the C ABI, on the CPU.

Every compute entry point validates its arguments before touching a device
and returns H3D_EARG (-1) or H3D_ELIMIT (-2), never exits or aborts (the
reference's HIP programs exit on error, 1-hip.cpp:43-51).  The pointers
below are never dereferenced: each call is rejected by a check that
precedes any device work, which is what these tests pin."""
import ctypes as C

import pytest

EARG, ELIMIT = -1, -2
FAKE = [C.c_void_p(0x1000 * (i + 1)) for i in range(8)]  # distinct, never read


@pytest.fixture(scope="module")
def L(h3d):
    return h3d.lib()


def grid(h3d, nx=20, ny=18, nz=16, sy=None, sz=None):
    sy = nx if sy is None else sy
    sz = sy * ny if sz is None else sz
    return h3d.Grid(nx, ny, nz, sy, sz)


BAD_GRIDS = [
    dict(nx=2), dict(ny=2), dict(nz=2),
    dict(nx=20, sy=19),                      # row stride below the row
    dict(nx=20, ny=18, sy=20, sz=359),       # plane stride below the plane
    dict(nx=20, ny=18, sy=20, sz=(1 << 28) + 1),  # past 32-bit in-plane offsets
]


@pytest.mark.parametrize("bad", BAD_GRIDS, ids=str)
def test_every_entry_point_rejects_bad_grids(h3d, L, bad):
    g = C.byref(grid(h3d, **bad))
    u, v, k, r, c1, c2, f1, f2 = FAKE
    out = (C.c_double * 9)()
    ms, fb = C.c_float(), C.c_int()
    assert L.h3d_step_f64(u, v, None, k, g, 0.1, 1, 1, None, None) == EARG
    assert L.h3d_run_f64(u, v, None, k, g, 0.1, 3, None, 0, None, C.byref(ms), C.byref(fb)) == EARG
    assert L.h3d_pass_f64(u, v, None, k, g, 0.1, 2, 1, 1, None, None) == EARG
    assert L.h3d_pass_f32(u, v, k, None, None, None, g, C.c_float(0.1), 1, 0, None, None) == EARG
    assert L.h3d_run_f32(u, v, k, None, None, None, g, C.c_float(0.1), 3, 0, None, 0, None,
                         C.byref(ms), C.byref(fb)) == EARG
    assert L.h3d_pack_face_f64(f1, u, g, 1, None) == EARG
    assert L.h3d_unpack_face_f32(u, f1, g, 1, None) == EARG
    assert L.h3d_stats_f64(u, g, out, None) == EARG
    assert L.h3d_compare_f64(u, v, g, 1e-6, 1e-5, out, None) == EARG
    assert L.h3d_face_sum_max_f64(k, g, out, None) == EARG
    assert L.h3d_slab_step_f64(u, v, None, k, g, 0.1, f1, f2, None, FAKE[7]) == EARG
    assert L.h3d_slab_pass_f64(u, v, None, k, g, 0.1, 1, 1, f1, f2, None, FAKE[7]) == EARG
    assert L.h3d_pack_band_f64(f1, u, g, 0, 1, None) == EARG
    assert L.h3d_unpack_band_f64(u, f1, g, 0, 1, None) == EARG


def test_fp64_step_checks(h3d, L):
    g = grid(h3d)
    u, v, k = FAKE[:3]
    gp = C.byref(g)
    assert L.h3d_step_f64(None, v, None, k, gp, 0.1, 1, 14, None, None) == EARG
    assert L.h3d_step_f64(u, None, None, k, gp, 0.1, 1, 14, None, None) == EARG
    assert L.h3d_step_f64(u, v, None, None, gp, 0.1, 1, 14, None, None) == EARG
    assert L.h3d_step_f64(u, u, None, k, gp, 0.1, 1, 14, None, None) == EARG  # in place
    assert L.h3d_step_f64(u, v, None, k, gp, 0.1, 0, 14, None, None) == EARG  # ring plane
    assert L.h3d_step_f64(u, v, None, k, gp, 0.1, 1, 15, None, None) == EARG
    for rows, waves in ((3, 2), (16, 2), (2, 0 - 1), (2, 17)):
        t = h3d.Tune(rows, waves, 8, 0)
        assert L.h3d_step_f64(u, v, None, k, gp, 0.1, 1, 14, C.byref(t), None) == EARG
    ms, fb = C.c_float(), C.c_int()
    assert L.h3d_run_f64(u, v, None, k, gp, 0.1, -1, None, 0, None, C.byref(ms),
                         C.byref(fb)) == EARG
    assert L.h3d_run_f64(u, v, None, k, gp, 0.1, 4, None, 1, None, C.byref(ms),
                         C.byref(fb)) == EARG  # a graph needs a created stream


def test_fp64_pass_checks(h3d, L):
    g = grid(h3d)
    gp = C.byref(g)
    u, v, k, r = FAKE[:4]
    P = L.h3d_pass_f64
    assert P(u, u, None, k, gp, 0.1, 2, 1, 14, None, None) == EARG  # in place
    assert P(None, v, None, k, gp, 0.1, 2, 1, 14, None, None) == EARG
    assert P(u, v, None, None, gp, 0.1, 2, 1, 14, None, None) == EARG
    assert P(u, v, None, k, gp, 0.1, 0, 1, 14, None, None) == ELIMIT
    assert P(u, v, None, k, gp, 0.1, 5, 1, 14, None, None) == ELIMIT
    for kblock, rows, waves in ((2, 1, 4), (2, 2, 2), (3, 8, 8), (4, 2, 4), (4, 2, 16),
                                (2, 4, 16)):
        t = h3d.Tune(rows, waves, 8, kblock)
        assert P(u, v, None, k, gp, 0.1, kblock, 1, 14, C.byref(t), None) == EARG
    t = h3d.Tune(2, 8, -1, 2)
    assert P(u, v, None, k, gp, 0.1, 2, 1, 14, C.byref(t), None) == EARG  # z chunk
    for ks, ke in ((0, 14), (1, 15), (-3, 2)):  # ring planes or outside the array
        assert P(u, v, None, k, gp, 0.1, 2, ks, ke, None, None) == EARG
    ms, fb = C.c_float(), C.c_int()
    t = h3d.Tune(2, 8, 8, 5)
    assert L.h3d_run_f64(u, v, None, k, gp, 0.1, 10, C.byref(t), 0, None, C.byref(ms),
                         C.byref(fb)) == ELIMIT


def test_fp32_pass_checks(h3d, L):
    g = grid(h3d)
    gp = C.byref(g)
    u, v, kf, kb, c1, c2 = FAKE[:6]
    f = C.c_float(0.1)
    P = L.h3d_pass_f32
    assert P(u, u, kf, None, None, None, gp, f, 1, 0, None, None) == EARG
    assert P(u, v, None, None, None, None, gp, f, 1, 0, None, None) == EARG  # no kappa
    assert P(u, v, kf, kb, None, None, gp, f, 1, 0, None, None) == EARG      # both kappas
    assert P(u, v, kf, None, None, None, gp, f, 1, 3, None, None) == EARG    # no mode 3
    assert P(u, v, kf, None, None, None, gp, f, 1, 2, None, None) == EARG    # mode 2: residuals
    assert P(u, v, kf, None, c1, c1, gp, f, 1, 2, None, None) == EARG        # in place
    assert P(u, v, kf, None, None, None, gp, f, 0, 0, None, None) == ELIMIT
    assert P(u, v, kf, None, None, None, gp, f, 9, 0, None, None) == ELIMIT
    for kblock, rows, waves in ((2, 1, 4), (2, 2, 2), (3, 8, 8), (1, 3, 4), (1, 2, 9)):
        t = h3d.Tune(rows, waves, 8, kblock)
        assert P(u, v, kf, None, None, None, gp, f, kblock, 0, C.byref(t), None) == EARG
    ms, fb = C.c_float(), C.c_int()
    t = h3d.Tune(0, 0, 0, 9)
    assert L.h3d_run_f32(u, v, kf, None, None, None, gp, f, 10, 0, C.byref(t), 0, None,
                         C.byref(ms), C.byref(fb)) == ELIMIT


def test_face_and_slab_checks(h3d, L):
    g = grid(h3d)
    gp = C.byref(g)
    u, v, k, f1, f2, s = FAKE[:6]
    for plane in (-1, g.nz):
        assert L.h3d_pack_face_f64(f1, u, gp, plane, None) == EARG
        assert L.h3d_unpack_face_f64(u, f1, gp, plane, None) == EARG
    assert L.h3d_pack_face_f64(None, u, gp, 1, None) == EARG
    thin = grid(h3d, nz=3)
    assert L.h3d_slab_step_f64(u, v, None, k, C.byref(thin), 0.1, f1, f2, None, s) == EARG
    assert L.h3d_slab_step_f64(u, v, None, k, gp, 0.1, None, f2, None, s) == EARG
    ms, fb = C.c_float(), C.c_int()
    assert L.h3d_run_slab_f64(u, v, None, k, gp, 0.1, f1, f2, 4, None, 0, None,
                              C.byref(ms), C.byref(fb)) == EARG  # needs a created stream
    # bands: planes k .. k+count-1 must lie in the array
    for plane, count in ((-1, 2), (15, 2), (14, 3), (1, 0)):
        assert L.h3d_pack_band_f64(f1, u, gp, plane, count, None) == EARG
        assert L.h3d_unpack_band_f64(u, f1, gp, plane, count, None) == EARG
    # deep slabs: ghost 1..4, nz >= 3*ghost, nsteps 1..ghost
    SP, RP = L.h3d_slab_pass_f64, L.h3d_run_slab_pass_f64
    for ghost, nsteps in ((0, 1), (5, 1), (2, 0), (2, 3)):
        assert SP(u, v, None, k, gp, 0.1, ghost, nsteps, f1, f2, None, s) == EARG
    g11 = grid(h3d, nz=11)
    assert SP(u, v, None, k, C.byref(g11), 0.1, 4, 4, f1, f2, None, s) == EARG
    assert SP(u, u, None, k, gp, 0.1, 2, 2, f1, f2, None, s) == EARG
    assert SP(u, v, None, k, gp, 0.1, 2, 2, f1, f1, None, s) == EARG
    assert RP(u, v, None, k, gp, 0.1, 2, f1, f2, -1, None, 0, s, C.byref(ms),
              C.byref(fb)) == EARG
    assert RP(u, v, None, k, gp, 0.1, 2, f1, f2, 4, None, 0, None, C.byref(ms),
              C.byref(fb)) == EARG  # needs a created stream


def test_ensemble_and_tuner_checks(h3d, L):
    e = C.c_void_p()
    assert L.h3d_ensemble_create(C.byref(e), None, 1, None, 0) == EARG
    mem = (h3d.Member * 1)()
    assert L.h3d_ensemble_create(C.byref(e), mem, 0, None, 0) == EARG
    assert L.h3d_ensemble_run(None, 1, 0, None, None) == EARG
    assert L.h3d_ensemble_autotune(None, None, None) == EARG
    assert L.h3d_ensemble_destroy(None) == EARG
    g = grid(h3d)
    t = h3d.Tune()
    u, v, k = FAKE[:3]
    assert L.h3d_autotune(2, u, v, k, 0, C.byref(g), 0, 0, None, C.byref(t)) == EARG
    assert L.h3d_autotune(1, u, v, k, 0, C.byref(g), 3, 0, None, C.byref(t)) == EARG
    assert L.h3d_autotune(1, u, v, k, 0, C.byref(g), 1, 9, None, C.byref(t)) == ELIMIT
    assert L.h3d_autotune(0, u, v, k, 0, C.byref(g), 2, 0, None, C.byref(t)) == EARG  # FP64: rhs flag
    assert L.h3d_autotune(0, u, v, k, 0, C.byref(g), 0, 5, None, C.byref(t)) == ELIMIT
    bad = grid(h3d, nx=2)
    assert L.h3d_autotune(0, u, v, k, 0, C.byref(bad), 0, 0, None, C.byref(t)) == EARG


def test_run_plan_query(h3d, L):
    """h3d_run_f64_passes: the launches of a run, on the CPU (it reads only
    the autotuner's cache, empty for this grid)."""
    g = grid(h3d, nx=23, ny=19, nz=17)
    n, d = C.c_int(), C.c_int()
    for kb, steps, want in ((3, 20, 7), (3, 21, 7), (2, 5, 3), (1, 9, 9), (4, 9, 3), (0, 4, 4)):
        t = h3d.Tune(2, 8, 0, kb)
        assert L.h3d_run_f64_passes(C.byref(g), 0, steps, C.byref(t), C.byref(n),
                                    C.byref(d)) == 0
        assert (n.value, d.value) == (want, 0)
    assert L.h3d_run_f64_passes(C.byref(g), 1, 20, None, C.byref(n), None) == 0
    assert n.value == 20
    t = h3d.Tune(2, 8, 0, 5)
    assert L.h3d_run_f64_passes(C.byref(g), 0, 20, C.byref(t), C.byref(n), None) == ELIMIT
    assert L.h3d_run_f64_passes(C.byref(g), 0, -1, None, C.byref(n), None) == EARG
    assert L.h3d_run_f64_passes(C.byref(g), 0, 4, None, None, None) == EARG
    bad = grid(h3d, nx=2)
    assert L.h3d_run_f64_passes(C.byref(bad), 0, 4, None, C.byref(n), None) == EARG
