"""Face pack/unpack, the z-slab decomposition and the FP64 reductions on
the GPU."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,pitch", [((7, 5, 4), 0), ((130, 66, 5), 0),
                                        ((33, 17, 6), 5), ((1026, 3, 3), 0),
                                        ((3, 1026, 3), 0)], ids=str)
@pytest.mark.parametrize("f32", [False, True])
def test_pack_unpack_bitwise(gpu, oracle, dims, pitch, f32):
    nx, ny, nz = dims
    g = gpu.Grid(nx, ny, nz, nx + pitch, (nx + pitch) * ny)
    dt = np.float32 if f32 else np.float64
    rng = np.random.default_rng(nx * ny)
    u = rng.standard_normal(g.elems).astype(dt)
    u.view(np.uint8)[::97] = 0xFF  # odd bit patterns (NaNs) travel unchanged
    du = gpu.DeviceArray.from_host(u)
    buf = gpu.DeviceArray(nx * ny * u.itemsize, dtype=dt)
    sfx = "f32" if f32 else "f64"
    s = gpu.Stream()
    for k in (0, nz // 2, nz - 1):
        gpu.call(f"h3d_pack_face_{sfx}", buf.ptr, du.ptr, gpu.C.byref(g), k, s.ptr)
        s.sync()
        got = buf.download()
        want = u.reshape(nz, ny, nx + pitch)[k, :, :nx].ravel()
        assert got.tobytes() == want.tobytes()
    # unpack a different plane into k = 1 and compare whole arrays
    face = rng.standard_normal(nx * ny).astype(dt)
    buf.upload(face)
    gpu.call(f"h3d_unpack_face_{sfx}", du.ptr, buf.ptr, gpu.C.byref(g), 1, s.ptr)
    s.sync()
    want = u.copy().reshape(nz, ny, nx + pitch)
    want[1, :, :nx] = face.reshape(ny, nx)
    assert du.download().tobytes() == want.ravel().tobytes()
    L = gpu.lib()
    assert L.h3d_pack_face_f64(buf.ptr, du.ptr, gpu.C.byref(g), nz, None) == -1


@pytest.mark.parametrize("parts", [2, 4])
def test_slabs_with_face_exchange_equal_global_run(gpu, oracle, parts):
    """Split a reference-mode problem into z-slabs (the reference's
    decomposition), step each slab with the boundary/interior kernel split
    and exchange faces through pack -> unpack between slab buffers with the
    reference's transfer map; the reconstructed field equals the global run
    bit for bit (as the reference's multi-GPU variants equal its serial code,
    HIP/verify-20370244.out:167-586)."""
    n, steps = 38, 7
    h, dt, factor = gpu.params_reference(n)
    u, rhs, k = gpu.init_reference(n)
    glob = oracle.run_f64(u.copy(), u.copy(), rhs, k, n, n, n, steps, factor=factor)
    rc, nzd, gks = gpu.slab_decompose(n, parts)
    assert rc == 0
    L = gpu.lib()
    slabs = []
    for nz_l, gk in zip(nzd, gks):
        a = [np.zeros(nz_l * n * n) for _ in range(3)]
        assert L.h3d_init_partition(*(gpu.ptr(x) for x in a), nz_l, gk, 1, n, n, n, h, dt) == 0
        su, sr, sk = a
        slabs.append(dict(g=gpu.Grid.dense(n, n, nz_l),
                          old=gpu.DeviceArray.from_host(su), new=gpu.DeviceArray.from_host(su),
                          rhs=gpu.DeviceArray.from_host(sr), kap=gpu.DeviceArray.from_host(sk)))
    face = gpu.DeviceArray(n * n * 8)
    s = gpu.Stream()
    for _ in range(steps):
        for p, sl in enumerate(slabs):
            g = sl["g"]
            nz_l = g.nz
            args = (sl["new"].ptr, sl["old"].ptr, sl["rhs"].ptr, sl["kap"].ptr, gpu.C.byref(g), factor)
            gpu.call("h3d_step_f64", *args, 1, 1, None, s.ptr)
            gpu.call("h3d_step_f64", *args, nz_l - 2, nz_l - 2, None, s.ptr)
            if nz_l > 4:
                gpu.call("h3d_step_f64", *args, 2, nz_l - 3, None, s.ptr)
        # transfer map (4-openmp-stream-omp-p2p.c:609-616): plane k=nz-2 of
        # slab p -> ghost k=0 of slab p+1; plane k=1 of slab p+1 -> ghost
        # k=nz-1 of slab p.  Faces land in u_new, read as u_old next step.
        for p in range(parts - 1):
            lo, hi = slabs[p], slabs[p + 1]
            gpu.call("h3d_pack_face_f64", face.ptr, lo["new"].ptr, gpu.C.byref(lo["g"]), lo["g"].nz - 2, s.ptr)
            gpu.call("h3d_unpack_face_f64", hi["new"].ptr, face.ptr, gpu.C.byref(hi["g"]), 0, s.ptr)
            gpu.call("h3d_pack_face_f64", face.ptr, hi["new"].ptr, gpu.C.byref(hi["g"]), 1, s.ptr)
            gpu.call("h3d_unpack_face_f64", lo["new"].ptr, face.ptr, gpu.C.byref(lo["g"]), lo["g"].nz - 1, s.ptr)
        for sl in slabs:
            sl["old"], sl["new"] = sl["new"], sl["old"]
    s.sync()
    # reconstruction (4-openmp-stream-omp-p2p.c:776-820): owned planes only,
    # plus the physical ghost planes of the first and last slab
    out = np.zeros(n ** 3)
    for p, sl in enumerate(slabs):
        a = sl["old"].download()
        assert L.h3d_slab_reconstruct(gpu.ptr(out), gpu.ptr(a), n, parts, p) == 0
    assert np.array_equal(out, glob)


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("nz", [4, 5, 19])
def test_slab_step_with_inhbm_exchange(gpu, oracle, graph, nz):
    """h3d_run_slab_f64 (boundary planes, interior on a second stream, face
    pack/unpack into the slab's own ghosts) against the oracle doing the
    same update followed by the same plane copies."""
    n, steps = 37, 9
    g = gpu.Grid.dense(n, n, nz)
    u, k = gpu.gen_seeded(g, 0, n, n, nz, 1, 3)  # a slab of a taller problem
    u[:] = np.where(u == 0.0, 0.5, u)  # ghosts hold data, as in a real slab
    rhs = np.random.default_rng(2).uniform(0, 1e-4, g.elems)
    factor = 0.02
    a, b = u.copy(), u.copy()
    P = n * n
    for _ in range(steps):
        oracle.step_f64(b, a, rhs, k, n, n, nz, factor, 1, nz - 2)
        b[0:P] = b[(nz - 2) * P:(nz - 1) * P]
        b[(nz - 1) * P:nz * P] = b[P:2 * P]
        a, b = b, a
    da, db, dk, dr = (gpu.DeviceArray.from_host(x) for x in (u, u, k, rhs))
    lo, hi = gpu.DeviceArray(P * 8), gpu.DeviceArray(P * 8)
    s = gpu.Stream()
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    gpu.call("h3d_run_slab_f64", da.ptr, db.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor,
             lo.ptr, hi.ptr, steps, None, graph, s.ptr, gpu.C.byref(ms), gpu.C.byref(fb))
    got = (db if fb.value else da).download()
    assert np.array_equal(got, a)
    # one step through the single-step entry point
    one, src = gpu.DeviceArray.from_host(u), gpu.DeviceArray.from_host(u)
    gpu.call("h3d_slab_step_f64", one.ptr, src.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor,
             lo.ptr, hi.ptr, None, s.ptr)
    s.sync()
    want = u.copy()
    oracle.step_f64(want, u, rhs, k, n, n, nz, factor, 1, nz - 2)
    want[0:P] = want[(nz - 2) * P:(nz - 1) * P]
    want[(nz - 1) * P:] = want[P:2 * P]
    assert np.array_equal(one.download(), want)


@pytest.mark.parametrize("dims,pitch,k,count", [((7, 5, 9), 0, 2, 3), ((33, 17, 8), 5, 0, 8),
                                                ((130, 66, 6), 0, 4, 2), ((40, 9, 7), 3, 6, 1)],
                         ids=str)
def test_band_pack_unpack_bitwise(gpu, dims, pitch, k, count):
    """Planes k..k+count-1 to a dense [k][j][i] buffer and back, dense
    (one flat copy) and pitched (row copies)."""
    nx, ny, nz = dims
    g = gpu.Grid(nx, ny, nz, nx + pitch, (nx + pitch) * ny)
    rng = np.random.default_rng(nz * count)
    u = rng.standard_normal(g.elems)
    u.view(np.uint8)[::89] = 0xFF
    du = gpu.DeviceArray.from_host(u)
    buf = gpu.DeviceArray(count * nx * ny * 8)
    s = gpu.Stream()
    gpu.call("h3d_pack_band_f64", buf.ptr, du.ptr, gpu.C.byref(g), k, count, s.ptr)
    s.sync()
    u3 = u.reshape(nz, ny, nx + pitch)
    assert buf.download().tobytes() == u3[k:k + count, :, :nx].ravel().tobytes()
    band = rng.standard_normal(count * nx * ny)
    buf.upload(band)
    k2 = nz - count
    gpu.call("h3d_unpack_band_f64", du.ptr, buf.ptr, gpu.C.byref(g), k2, count, s.ptr)
    s.sync()
    want = u.copy().reshape(nz, ny, nx + pitch)
    want[k2:, :, :nx] = band.reshape(count, ny, nx)
    assert du.download().tobytes() == want.ravel().tobytes()


def deep_standin_oracle(oracle, u, rhs, k, g, factor, ghost, steps):
    """Passes of `ghost` steps (the last one shorter) of the whole slab
    array, each followed by the periodic band copies: owned planes
    nz-2*ghost..nz-1-ghost into ghosts 0..ghost-1, owned planes
    ghost..2*ghost-1 into ghosts nz-ghost..nz-1."""
    P, nz = g.sz, g.nz
    a = u.copy()
    left = steps
    while left:
        r = min(ghost, left)
        a = oracle.run_f64(a.copy(), a.copy(), rhs, k, g.nx, g.ny, nz, r, sy=g.sy, sz=g.sz,
                           factor=factor).copy()
        a[0:ghost * P] = a[(nz - 2 * ghost) * P:(nz - ghost) * P]
        a[(nz - ghost) * P:] = a[ghost * P:2 * ghost * P]
        left -= r
    return a


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("ghost,nz,pitch", [(1, 5, 0), (2, 6, 0), (2, 9, 0), (2, 23, 0),
                                            (3, 9, 0), (3, 16, 0), (4, 30, 0), (2, 12, 11),
                                            (3, 16, 27)], ids=str)
def test_deep_slab_standin_matches_oracle(gpu, oracle, graph, ghost, nz, pitch):
    """h3d_run_slab_pass_f64 (fused passes of `ghost` steps, boundary bands
    first, interior on a second stream, band exchange into the slab's own
    ghosts) against the oracle stepping the whole slab array and copying
    the bands, on dense and padded rows; with ghost 1 it is
    h3d_run_slab_f64."""
    n, steps = 37, 2 * ghost + 3
    g = gpu.Grid(n, n, nz, n + pitch, (n + pitch) * n)
    u, k = gpu.gen_seeded(g, 0, n, n, nz, 1, 5)
    u[:] = np.where(u == 0.0, 0.5, u)  # ghosts hold data, as in a real slab
    rhs = np.random.default_rng(6).uniform(0, 1e-4, g.elems)
    factor = 0.02
    want = deep_standin_oracle(oracle, u, rhs, k, g, factor, ghost, steps)
    da, db, dk, dr = (gpu.DeviceArray.from_host(x) for x in (u, u, k, rhs))
    lo, hi = gpu.DeviceArray(ghost * n * n * 8), gpu.DeviceArray(ghost * n * n * 8)
    s = gpu.Stream()
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    t = gpu.Tune(4, 4, 0, ghost) if ghost > 1 else None
    gpu.call("h3d_run_slab_pass_f64", da.ptr, db.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor,
             ghost, lo.ptr, hi.ptr, steps, gpu.C.byref(t) if t else None, graph, s.ptr,
             gpu.C.byref(ms), gpu.C.byref(fb))
    npass = -(-steps // ghost)
    assert fb.value == npass % 2
    assert np.array_equal((db if fb.value else da).download(), want)
    if ghost == 1:
        ea, eb = gpu.DeviceArray.from_host(u), gpu.DeviceArray.from_host(u)
        gpu.call("h3d_run_slab_f64", ea.ptr, eb.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor,
                 lo.ptr, hi.ptr, steps, None, graph, s.ptr, gpu.C.byref(ms), gpu.C.byref(fb))
        assert np.array_equal((eb if fb.value else ea).download(), want)
    # one shorter pass through the single-pass entry point
    one, src = gpu.DeviceArray.from_host(u), gpu.DeviceArray.from_host(u)
    r = max(1, ghost - 1)
    gpu.call("h3d_slab_pass_f64", one.ptr, src.ptr, dr.ptr, dk.ptr, gpu.C.byref(g), factor,
             ghost, r, lo.ptr, hi.ptr, None, s.ptr)
    s.sync()
    assert np.array_equal(one.download(),
                          deep_standin_oracle(oracle, u, rhs, k, g, factor, ghost, r))


@pytest.mark.parametrize("parts,ghost", [(2, 2), (4, 2), (4, 3), (2, 4), (4, 4)], ids=str)
def test_deep_slabs_with_band_exchange_equal_global_run(gpu, oracle, parts, ghost):
    """The deep-ghost algorithm across slabs: split a reference-mode
    problem into the reference's z-slabs, widened to `ghost` ghost planes
    per side (the first and last slab end at the global Dirichlet planes);
    advance each slab by fused passes over its owned planes
    (h3d_pass_f64 with a plane range) and swap bands of `ghost` planes
    between neighbours (h3d_pack_band_f64 -> h3d_unpack_band_f64) once per
    pass.  The owned planes reassembled equal the global run bit for bit."""
    n, steps = 38, 2 * ghost + 1
    h, dt, factor = gpu.params_reference(n)
    u, rhs, k = gpu.init_reference(n)
    glob = oracle.run_f64(u.copy(), u.copy(), rhs, k, n, n, n, steps, factor=factor)
    rc, nzd, gks = gpu.slab_decompose(n, parts)
    assert rc == 0
    P = n * n
    slabs = []
    for nz_l, gk in zip(nzd, gks):
        own_lo, own_hi = gk, gk + nz_l - 3  # owned global planes (2 ghosts in nz_l)
        assert own_hi - own_lo + 1 >= ghost
        start, end = max(0, own_lo - ghost), min(n - 1, own_hi + ghost)
        sl = slice(start * P, (end + 1) * P)
        slabs.append(dict(g=gpu.Grid.dense(n, n, end - start + 1), start=start,
                          lo=own_lo - start, hi=own_hi - start,
                          old=gpu.DeviceArray.from_host(u[sl]),
                          new=gpu.DeviceArray.from_host(u[sl]),
                          rhs=gpu.DeviceArray.from_host(rhs[sl]),
                          kap=gpu.DeviceArray.from_host(k[sl])))
    band = gpu.DeviceArray(ghost * P * 8)
    s = gpu.Stream()
    left = steps
    while left:
        r = min(ghost, left)
        for sl in slabs:
            gpu.call("h3d_pass_f64", sl["new"].ptr, sl["old"].ptr, sl["rhs"].ptr,
                     sl["kap"].ptr, gpu.C.byref(sl["g"]), factor, r, sl["lo"], sl["hi"],
                     None, s.ptr)
        for p in range(parts - 1):
            a, b = slabs[p], slabs[p + 1]
            # top owned band of a -> lower ghosts of b; bottom owned band of
            # b -> upper ghosts of a
            gpu.call("h3d_pack_band_f64", band.ptr, a["new"].ptr, gpu.C.byref(a["g"]),
                     a["hi"] - ghost + 1, ghost, s.ptr)
            gpu.call("h3d_unpack_band_f64", b["new"].ptr, band.ptr, gpu.C.byref(b["g"]), 0,
                     ghost, s.ptr)
            gpu.call("h3d_pack_band_f64", band.ptr, b["new"].ptr, gpu.C.byref(b["g"]), b["lo"],
                     ghost, s.ptr)
            gpu.call("h3d_unpack_band_f64", a["new"].ptr, band.ptr, gpu.C.byref(a["g"]),
                     a["hi"] + 1, ghost, s.ptr)
        for sl in slabs:
            sl["old"], sl["new"] = sl["new"], sl["old"]
        left -= r
    s.sync()
    out = u.copy()  # global Dirichlet planes 0 and n-1 never change
    for sl in slabs:
        a = sl["old"].download()
        g0 = sl["start"]
        out[(g0 + sl["lo"]) * P:(g0 + sl["hi"] + 1) * P] = a[sl["lo"] * P:(sl["hi"] + 1) * P]
    assert np.array_equal(out, glob)


def test_stats_matches_oracle(gpu, oracle):
    n = 45
    g = gpu.Grid(n, n - 3, n - 5, n + 3, (n + 3) * (n - 3))
    u, k = gpu.gen_seeded(g, 0, n, n - 3, n - 5, 1, 8)
    rng = np.random.default_rng(0)
    u = u * 10 ** rng.uniform(-3, 3, u.size)
    du = gpu.DeviceArray.from_host(u)
    got = gpu.stats(du, g)
    want = oracle.stats(u, g.nx, g.ny, g.nz, g.sy, g.sz)
    assert got[0] == pytest.approx(want[0], rel=1e-14)
    assert got[1] == pytest.approx(want[1], rel=1e-14)
    assert got[2] == want[2] and got[3] == 0
    again = gpu.stats(du, g)
    assert np.array_equal(got, again)  # deterministic
    u32 = u.astype(np.float32)
    got32 = gpu.stats(gpu.DeviceArray.from_host(u32), g, f32=True)
    want32 = oracle.stats(u32.astype(np.float64), g.nx, g.ny, g.nz, g.sy, g.sz)
    assert got32[0] == pytest.approx(want32[0], rel=1e-14)
    u[17] = np.nan
    u[40] = np.inf
    assert gpu.stats(gpu.DeviceArray.from_host(u), g)[3] == 2


def test_compare_matches_oracle_metrics(gpu, oracle):
    nx, ny, nz = 50, 40, 30
    g = gpu.Grid.dense(nx, ny, nz)
    rng = np.random.default_rng(3)
    ref = rng.standard_normal(g.elems)
    ref[::50] = 0.0
    cand = ref + rng.standard_normal(g.elems) * 1e-6 * np.abs(ref)
    cand[123] += 1e-3
    cand[777] = np.inf
    dr, dc = gpu.DeviceArray.from_host(ref), gpu.DeviceArray.from_host(cand)
    got = gpu.compare(dr, dc, g, 1e-6, 1e-5)
    want = oracle.metrics(ref, cand, 1e-6, 1e-5)
    assert got[5] == want[5] == 777
    for i in (6, 7, 8):
        assert got[i] == want[i]
    fin = cand.copy()
    fin[777] = ref[777]
    dc.upload(fin)
    got = gpu.compare(dr, dc, g, 1e-6, 1e-5)
    want = oracle.metrics(ref, fin, 1e-6, 1e-5)
    assert got[0] == want[0] and got[5] == want[5] == 123
    for i in (1, 2, 3, 4):
        assert got[i] == pytest.approx(want[i], rel=1e-12)
    assert got[8] == want[8]
    c32 = fin.astype(np.float32)
    got = gpu.compare(dr, gpu.DeviceArray.from_host(c32), g, 1e-6, 1e-5, f32=True)
    want = oracle.metrics(ref, c32.astype(np.float64), 1e-6, 1e-5)
    assert got[3] == pytest.approx(want[3], rel=1e-12) and got[8] == want[8]


def test_face_sum_max_device_matches_host(gpu, oracle):
    n = 61
    g = gpu.Grid.dense(n, n - 7, n - 11)
    _, k = gpu.gen_seeded(g, 0, g.nx, g.ny, g.nz, 1, 21)
    dk = gpu.DeviceArray.from_host(k)
    out = np.zeros(1)
    gpu.call("h3d_face_sum_max_f64", dk.ptr, gpu.C.byref(g), gpu.ptr(out), None)
    assert out[0] == oracle.face_sum_max(k, g.nx, g.ny, g.nz)


def test_stream_probe_runs(gpu):
    c, t = gpu.stream_probe(1 << 28, reps=2)
    assert c > 100.0 and t > 100.0 and math.isfinite(c)


@pytest.mark.parametrize("pinned", [True, False])
def test_transfer_probe_runs(gpu, pinned):
    h2d, d2h, lat_h2d, lat_d2h = gpu.transfer_probe(1 << 26, reps=2, pinned=pinned)
    assert h2d > 0.5 and d2h > 0.5 and 0 < lat_h2d < 1e4 and 0 < lat_d2h < 1e4
    L = gpu.lib()
    out = np.zeros(4)
    assert L.h3d_transfer_probe(100, 1, 1, 0, gpu.ptr(out)) == -1


@pytest.mark.parametrize("nbytes", [8, 4096, (1 << 20) + 8, 1 << 26], ids=str)
@pytest.mark.parametrize("pinned", [True, False])
@pytest.mark.parametrize("asynchronous", [False, True])
@pytest.mark.parametrize("d2h", [False, True])
def test_transfer_bench_delivers_every_word(gpu, nbytes, pinned, asynchronous, d2h):
    """The verified transfer measurement (the reference's synchronous and
    asynchronous H2D / D2H benchmarks): every 64-bit word arrives with its
    own value, in both directions, from pinned and pageable memory, and
    the rate is a positive finite number."""
    gbs, sec, bad, checked = gpu.transfer_bench(nbytes, pinned, asynchronous, d2h)
    assert bad == 0 and checked == nbytes // 8
    assert sec > 0 and math.isfinite(gbs) and gbs > 0


def test_transfer_bench_argument_errors(gpu):
    L = gpu.lib()
    out = np.zeros(4)
    assert L.h3d_transfer_bench(12, 1, 0, 0, 0, gpu.ptr(out)) == -1   # not whole words
    assert L.h3d_transfer_bench(0, 1, 0, 0, 0, gpu.ptr(out)) == -1
    assert L.h3d_transfer_bench(64, 1, 2, 0, 0, gpu.ptr(out)) == -1   # async flag
    assert L.h3d_transfer_bench(64, 1, 0, 2, 0, gpu.ptr(out)) == -1   # direction
    assert L.h3d_transfer_bench(64, 1, 0, 0, 0, None) == -1
    assert L.h3d_transfer_bench(64, 1, 0, 0, 99, gpu.ptr(out)) == -3  # no such device


def test_device_copy_completes_before_return(gpu):
    """h3d_target_memcpy is synchronous for device-to-device copies too:
    a pass queued right after it on a non-blocking stream reads the copied
    field, so the result equals the same pass on a field uploaded from the
    host (the copy of 1 GiB takes longer than the launch of the pass)."""
    n = 514
    g = gpu.Grid.dense(n)
    u, k = gpu.gen_seeded(g, 0, n, n, n, 0, 11)
    u32, k32 = u.astype(np.float32), k.astype(np.float32)
    src = gpu.DeviceArray.from_host(u32)
    dk = gpu.DeviceArray.from_host(k32)
    s = gpu.Stream()
    outs = []
    for via_copy in (False, True):
        a = gpu.DeviceArray.from_host(np.zeros_like(u32))
        b = gpu.DeviceArray.from_host(u32)
        if via_copy:
            gpu.call("h3d_target_memcpy", a.ptr, src.ptr, gpu.C.c_size_t(src.nbytes), 0, 0)
        else:
            a.upload(u32)
        t = gpu.Tune(0, 0, 0, 2)
        gpu.call("h3d_pass_f32", b.ptr, a.ptr, dk.ptr, None, None, None, gpu.C.byref(g),
                 gpu.C.c_float(0.01), 2, 1, gpu.C.byref(t), s.ptr)
        s.sync()
        outs.append(b.download())
    assert outs[0].tobytes() == outs[1].tobytes()


def test_compare_pass_rule_at_the_tolerance_boundary(gpu, oracle):
    """The PASS rule |c - r| <= atol + rtol*|r| decided on the device as
    numpy decides it (compare.py:65-76, unfused arithmetic), on candidates
    one step below, at and one step above the rounded r + bound.  With
    rtol = 0.3 the bound's spacing is close to that of c - r (about a
    tenth of the cells sit exactly on the bound), so a fused multiply-add
    in the bound (one rounding instead of two) changes the decision on
    thousands of these cells; the failure count must equal numpy's and
    the oracle's."""
    rng = np.random.default_rng(17)
    atol, rtol = 1e-6, 0.3
    r = rng.uniform(1.0, 2.0, 30_000) * rng.choice([-1.0, 1.0], 30_000)
    c = r + (atol + rtol * np.abs(r))
    ref = np.concatenate([r, r, r])
    cand = np.concatenate([np.nextafter(c, -np.inf), c, np.nextafter(c, np.inf)])
    ad, bound = np.abs(cand - ref), atol + rtol * np.abs(ref)
    want = int(np.sum(~(ad <= bound)))
    assert int(np.sum(ad == bound)) > 1000 and 0 < want < ref.size
    # a 3 x 3 x nx block holds them (every point of the array is compared)
    nx = ref.size // 9
    assert 9 * nx == ref.size
    g = gpu.Grid(nx, 3, 3, nx, 3 * nx)
    dr, dc = gpu.DeviceArray.from_host(ref), gpu.DeviceArray.from_host(cand)
    got = gpu.compare(dr, dc, g, atol=atol, rtol=rtol)
    assert int(got[8]) == want
    assert int(oracle.metrics(ref, cand, atol, rtol)[8]) == want
