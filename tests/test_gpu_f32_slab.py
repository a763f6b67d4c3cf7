"""FP32 fields in slabs with deep ghosts on the GPU: the plane-range FP32
pass (h3d_pass_f32_range) writes exactly the planes it is given, with the
bits of the whole-field pass, and slabs advanced by range passes with band
exchange -- of u, and of the FP16 residual field in the COMP16 mode --
reassemble to the global run bit for bit, as the FP64 slabs do."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fields(gpu, nx, ny, nz, pitch=0, seed=7, bf16=False):
    """Seeded FP32 fields (kappa as FP32, or BF16 bits and scale) and the
    FP32 factor the kernels take."""
    g = gpu.Grid(nx, ny, nz, nx + pitch, (nx + pitch) * ny)
    u, k = gpu.gen_seeded(g, 0, nx, ny, nz, 1, seed)
    h = 1.0 / (nx - 1)
    if bf16:
        bits, scale = gpu.quantize_bf16(k)
        k64 = gpu.bf16_to_f64(bits, scale)
    else:
        bits, scale = None, 1.0
        k64 = k.astype(np.float32).astype(np.float64)
    fmax = gpu.lib().h3d_face_sum_max_host(gpu.ptr(k64), gpu.C.byref(g))
    factor = float(np.float32(0.9 * 2 * h * h / fmax / h / h / 2.0 * scale))
    return dict(g=g, u=u.astype(np.float32), k=k.astype(np.float32), bits=bits,
                factor=factor)


def oracle_pass(oracle, F, u, comp, g, ns, mode):
    """One pass of ns steps over the whole array of grid g."""
    kf, kb = (None, F["bits"]) if F["bits"] is not None else (F["k"], None)
    return oracle.run_f32(u, kf, kb, g.nx, g.ny, g.nz, F["factor"], ns, ns, mode, comp0=comp,
                          sy=g.sy, sz=g.sz)


def kappa_args(gpu, F, part=slice(None)):
    if F["bits"] is not None:
        dk = gpu.DeviceArray.from_host(np.ascontiguousarray(F["bits"][part]))
        return dk, None, dk.ptr
    dk = gpu.DeviceArray.from_host(np.ascontiguousarray(F["k"][part]))
    return dk, dk.ptr, None


RANGE_CASES = [
    # shape, pitch, nsteps, k_start, k_end, mode, bf16, tune (rows, waves, zchunk)
    ((70, 40, 30), 0, 3, 1, 28, 1, False, (4, 16, 0)),
    ((70, 40, 30), 0, 3, 5, 11, 2, False, (4, 8, 4)),
    ((65, 33, 21), 7, 2, 2, 2, 1, True, (2, 8, 0)),
    ((65, 33, 21), 0, 1, 3, 17, 2, False, (2, 4, 5)),
    ((40, 70, 25), 0, 4, 9, 23, 2, True, (4, 8, 6)),
    ((40, 70, 25), 5, 1, 1, 1, 0, False, (1, 2, 0)),
]


@pytest.mark.parametrize("case", RANGE_CASES, ids=str)
def test_f32_pass_plane_range(gpu, oracle, case):
    """A pass over planes k_start..k_end writes those planes' interior
    cells (u, and the residual in COMP16 mode) with the bits of the same
    pass over the whole field; every other value of u_out and comp_out
    (rings, padding, other planes) keeps what it held."""
    shape, pitch, ns, ks, ke, mode, bf16, tune = case
    nx, ny, nz = shape
    F = fields(gpu, nx, ny, nz, pitch, bf16=bf16)
    g = F["g"]
    # a residual field as one earlier pass leaves it
    _, comp0 = oracle_pass(oracle, F, F["u"], None, g, 2, 2)
    comp0 = comp0 if mode == 2 else np.zeros(g.elems, np.uint16)
    full_u, full_c = oracle_pass(oracle, F, F["u"], comp0, g, ns, mode)
    rng = np.random.default_rng(3)
    pre_u = rng.standard_normal(g.elems).astype(np.float32)
    pre_c = rng.integers(0, 65535, g.elems, dtype=np.uint16)
    want_u, want_c = pre_u.copy(), pre_c.copy()
    sl = (slice(ks, ke + 1), slice(1, ny - 1), slice(1, nx - 1))
    want_u.reshape(nz, ny, g.sy)[sl] = full_u.reshape(nz, ny, g.sy)[sl]
    if mode == 2:
        want_c.reshape(nz, ny, g.sy)[sl] = full_c.reshape(nz, ny, g.sy)[sl]
    di, do = gpu.DeviceArray.from_host(F["u"]), gpu.DeviceArray.from_host(pre_u)
    ci, co = gpu.DeviceArray.from_host(comp0), gpu.DeviceArray.from_host(pre_c)
    dk, kf, kb = kappa_args(gpu, F)
    t = gpu.Tune(*tune, ns)
    s = gpu.Stream()
    gpu.call("h3d_pass_f32_range", do.ptr, di.ptr, kf, kb, co.ptr, ci.ptr, gpu.C.byref(g),
             gpu.C.c_float(F["factor"]), ns, mode, ks, ke, gpu.C.byref(t), s.ptr)
    s.sync()
    assert do.download().tobytes() == want_u.tobytes()
    assert co.download().tobytes() == want_c.tobytes()


def test_f32_pass_range_arguments(gpu):
    g = gpu.Grid.dense(20, 20, 12)
    a, b = gpu.DeviceArray(g.elems * 4), gpu.DeviceArray(g.elems * 4)
    L = gpu.lib()
    EARG = gpu.H3D_EARG
    f = gpu.C.c_float(0.01)
    assert L.h3d_pass_f32_range(a.ptr, b.ptr, a.ptr, None, None, None, gpu.C.byref(g), f, 2,
                                1, 0, 5, None, None) == EARG  # k_start < 1
    assert L.h3d_pass_f32_range(a.ptr, b.ptr, a.ptr, None, None, None, gpu.C.byref(g), f, 2,
                                1, 1, 11, None, None) == EARG  # k_end > nz-2
    # an empty range writes nothing and succeeds
    assert L.h3d_pass_f32_range(a.ptr, b.ptr, b.ptr, None, None, None, gpu.C.byref(g), f, 2,
                                1, 6, 5, None, None) == 0


@pytest.mark.parametrize("parts,ghost,mode,bf16", [(2, 2, 1, False), (4, 3, 2, False),
                                                   (2, 4, 2, True), (4, 2, 0, False),
                                                   (2, 6, 1, False)], ids=str)
def test_f32_deep_slabs_with_band_exchange_equal_global_run(gpu, oracle, parts, ghost, mode,
                                                            bf16):
    """The reference's z-split of a seeded FP32 problem, each slab widened
    to `ghost` ghost planes per side, advanced by range passes over its
    owned planes (h3d_pass_f32_range) with bands of `ghost` planes swapped
    between neighbours once per pass -- u, and the FP16 residual field in
    COMP16 mode.  The owned planes reassembled equal the global run of the
    same passes bit for bit."""
    n = 34
    steps = 2 * ghost + 1
    F = fields(gpu, n, n, n, bf16=bf16)
    g = F["g"]
    glob_u, glob_c = oracle.run_f32(F["u"], None if bf16 else F["k"], F["bits"], n, n, n,
                                    F["factor"], steps, ghost, mode)
    rc, nzd, gks = gpu.slab_decompose(n, parts)
    assert rc == 0
    P = n * n
    slabs = []
    zero_c = np.zeros(g.elems, np.uint16)
    for nz_l, gk in zip(nzd, gks):
        own_lo, own_hi = gk, gk + nz_l - 3
        assert own_hi - own_lo + 1 >= ghost
        start, end = max(0, own_lo - ghost), min(n - 1, own_hi + ghost)
        part = slice(start * P, (end + 1) * P)
        dk, kf, kb = kappa_args(gpu, F, part)
        slabs.append(dict(g=gpu.Grid.dense(n, n, end - start + 1), start=start,
                          lo=own_lo - start, hi=own_hi - start, dk=dk, kf=kf, kb=kb,
                          old=gpu.DeviceArray.from_host(F["u"][part]),
                          new=gpu.DeviceArray.from_host(F["u"][part]),
                          cold=gpu.DeviceArray.from_host(zero_c[part]),
                          cnew=gpu.DeviceArray.from_host(zero_c[part])))
    band = gpu.DeviceArray(ghost * P * 4)
    cband = gpu.DeviceArray(ghost * P * 2)
    s = gpu.Stream()
    left = steps
    while left:
        r = min(ghost, left)
        for sl in slabs:
            gpu.call("h3d_pass_f32_range", sl["new"].ptr, sl["old"].ptr, sl["kf"], sl["kb"],
                     sl["cnew"].ptr, sl["cold"].ptr, gpu.C.byref(sl["g"]),
                     gpu.C.c_float(F["factor"]), r, mode, sl["lo"], sl["hi"], None, s.ptr)
        for p in range(parts - 1):
            a, b = slabs[p], slabs[p + 1]
            pairs = [("new", band, "h3d_pack_band_f32", "h3d_unpack_band_f32")]
            if mode == 2:
                pairs.append(("cnew", cband, "h3d_pack_band_u16", "h3d_unpack_band_u16"))
            for field, buf, pack, unpack in pairs:
                # top owned band of a -> lower ghosts of b; bottom owned band
                # of b -> upper ghosts of a
                gpu.call(pack, buf.ptr, a[field].ptr, gpu.C.byref(a["g"]), a["hi"] - ghost + 1,
                         ghost, s.ptr)
                gpu.call(unpack, b[field].ptr, buf.ptr, gpu.C.byref(b["g"]), 0, ghost, s.ptr)
                gpu.call(pack, buf.ptr, b[field].ptr, gpu.C.byref(b["g"]), b["lo"], ghost,
                         s.ptr)
                gpu.call(unpack, a[field].ptr, buf.ptr, gpu.C.byref(a["g"]), a["hi"] + 1,
                         ghost, s.ptr)
        for sl in slabs:
            sl["old"], sl["new"] = sl["new"], sl["old"]
            sl["cold"], sl["cnew"] = sl["cnew"], sl["cold"]
        left -= r
    s.sync()
    out_u, out_c = F["u"].copy(), np.zeros(g.elems, np.uint16)
    for sl in slabs:
        a, c = sl["old"].download(), sl["cold"].download()
        g0, lo, hi = sl["start"], sl["lo"], sl["hi"]
        out_u[(g0 + lo) * P:(g0 + hi + 1) * P] = a[lo * P:(hi + 1) * P]
        out_c[(g0 + lo) * P:(g0 + hi + 1) * P] = c[lo * P:(hi + 1) * P]
    assert out_u.tobytes() == glob_u.tobytes()
    if mode == 2:
        assert out_c.tobytes() == glob_c.tobytes()


def standin_oracle_f32(oracle, F, g, ghost, steps, mode):
    """Passes of `ghost` steps (the last one shorter) of the whole slab
    array, each followed by the periodic band copies of u (and of the
    residual in COMP16 mode): owned planes nz-2*ghost..nz-1-ghost into
    ghosts 0..ghost-1, owned planes ghost..2*ghost-1 into ghosts
    nz-ghost..nz-1."""
    P, nz = g.sz, g.nz
    u, c = F["u"].copy(), np.zeros(g.elems, np.uint16)
    left = steps
    while left:
        r = min(ghost, left)
        u, c2 = oracle_pass(oracle, F, u, c, g, r, mode)
        u = u.copy()
        c = c2.copy() if mode == 2 else c
        for a in ((u, c) if mode == 2 else (u,)):
            a[0:ghost * P] = a[(nz - 2 * ghost) * P:(nz - ghost) * P]
            a[(nz - ghost) * P:] = a[ghost * P:2 * ghost * P]
        left -= r
    return u, c


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("ghost,nz,pitch,mode,bf16", [(1, 5, 0, 1, False), (2, 9, 0, 2, False),
                                                      (3, 16, 5, 2, True),
                                                      (4, 30, 0, 1, False),
                                                      (3, 12, 0, 0, False)], ids=str)
def test_f32_slab_standin_matches_oracle(gpu, oracle, graph, ghost, nz, pitch, mode, bf16):
    """h3d_run_slab_pass_f32 (passes of `ghost` steps, bands first, the
    owned planes between them on a second stream, band exchange into the
    slab's own ghosts) against the oracle stepping the whole slab array and
    copying the bands; and one shorter pass through h3d_slab_pass_f32."""
    n, steps = 37, 2 * ghost + 3
    F = fields(gpu, n, n, nz, pitch, seed=9, bf16=bf16)
    g = F["g"]
    F["u"] = np.where(F["u"] == 0.0, np.float32(0.5), F["u"])  # ghosts hold data
    want_u, want_c = standin_oracle_f32(oracle, F, g, ghost, steps, mode)
    da, db = gpu.DeviceArray.from_host(F["u"]), gpu.DeviceArray.from_host(F["u"])
    zero_c = np.zeros(g.elems, np.uint16)
    ca, cb = gpu.DeviceArray.from_host(zero_c), gpu.DeviceArray.from_host(zero_c)
    dk, kf, kb = kappa_args(gpu, F)
    lo, hi = gpu.DeviceArray(ghost * n * n * 4), gpu.DeviceArray(ghost * n * n * 4)
    clo, chi = gpu.DeviceArray(ghost * n * n * 2), gpu.DeviceArray(ghost * n * n * 2)
    cargs = (clo.ptr, chi.ptr) if mode == 2 else (None, None)
    s = gpu.Stream()
    ms, fb = gpu.C.c_float(), gpu.C.c_int()
    t = gpu.Tune(4, 4, 0, ghost) if ghost > 1 else None
    gpu.call("h3d_run_slab_pass_f32", da.ptr, db.ptr, kf, kb, ca.ptr, cb.ptr, gpu.C.byref(g),
             gpu.C.c_float(F["factor"]), mode, ghost, lo.ptr, hi.ptr, *cargs, steps,
             gpu.C.byref(t) if t else None, graph, s.ptr, gpu.C.byref(ms), gpu.C.byref(fb))
    npass = -(-steps // ghost)
    assert fb.value == npass % 2
    assert (db if fb.value else da).download().tobytes() == want_u.tobytes()
    if mode == 2:
        assert (cb if fb.value else ca).download().tobytes() == want_c.tobytes()
    # one shorter pass through the single-pass entry point
    r = max(1, ghost - 1)
    one, src = gpu.DeviceArray.from_host(F["u"]), gpu.DeviceArray.from_host(F["u"])
    c1, c0 = gpu.DeviceArray.from_host(zero_c), gpu.DeviceArray.from_host(zero_c)
    gpu.call("h3d_slab_pass_f32", one.ptr, src.ptr, kf, kb, c1.ptr, c0.ptr, gpu.C.byref(g),
             gpu.C.c_float(F["factor"]), mode, ghost, r, lo.ptr, hi.ptr, *cargs, None, s.ptr)
    s.sync()
    w1, wc1 = standin_oracle_f32(oracle, F, g, ghost, r, mode)
    assert one.download().tobytes() == w1.tobytes()
    if mode == 2:
        assert c1.download().tobytes() == wc1.tobytes()


def test_f32_slab_arguments(gpu):
    n, nz = 20, 12
    g = gpu.Grid.dense(n, n, nz)
    a, b = gpu.DeviceArray(g.elems * 4), gpu.DeviceArray(g.elems * 4)
    k = gpu.DeviceArray(g.elems * 4)
    lo, hi = gpu.DeviceArray(4 * n * n * 4), gpu.DeviceArray(4 * n * n * 4)
    L = gpu.lib()
    EARG = gpu.H3D_EARG
    f = gpu.C.c_float(0.01)
    # COMP16 needs residual fields and their band buffers
    assert L.h3d_slab_pass_f32(a.ptr, b.ptr, k.ptr, None, None, None, gpu.C.byref(g), f, 2, 2,
                               2, lo.ptr, hi.ptr, None, None, None, None) == EARG
    # ghost depth beyond the fused passes, nsteps beyond the ghost depth,
    # and a slab too thin for its ghosts
    assert L.h3d_slab_pass_f32(a.ptr, b.ptr, k.ptr, None, None, None, gpu.C.byref(g), f, 1, 9,
                               1, lo.ptr, hi.ptr, None, None, None, None) == EARG
    assert L.h3d_slab_pass_f32(a.ptr, b.ptr, k.ptr, None, None, None, gpu.C.byref(g), f, 1, 2,
                               3, lo.ptr, hi.ptr, None, None, None, None) == EARG
    assert L.h3d_slab_pass_f32(a.ptr, b.ptr, k.ptr, None, None, None, gpu.C.byref(g), f, 1, 5,
                               1, lo.ptr, hi.ptr, None, None, None, None) == EARG
    # both kappa forms at once
    assert L.h3d_slab_pass_f32(a.ptr, b.ptr, k.ptr, k.ptr, None, None, gpu.C.byref(g), f, 1,
                               2, 2, lo.ptr, hi.ptr, None, None, None, None) == EARG


SLAB_TUNE_CASES = [
    # kind (0 FP64, 1 FP32), mode, ghost, bf16 kappa
    (1, 2, 3, False), (1, 1, 2, True), (1, 0, 1, False), (0, 0, 2, False), (0, 1, 1, False),
    (0, 0, 3, False),
]


@pytest.mark.parametrize("kind,mode,ghost,bf16", SLAB_TUNE_CASES, ids=str)
def test_autotune_slab_reads_only_and_keeps_bits(gpu, kind, mode, ghost, bf16):
    """h3d_autotune_slab times whole slab passes on scratch copies: the
    caller's u, kappa and residual field are unchanged, the tune has the
    slab's depth, and a run with it equals a run on the default shape bit
    for bit."""
    n, nz, steps = 70, 40, 2 * ghost + 1
    g = gpu.Grid(n, n, nz, n + 3, (n + 3) * n)
    u, k = gpu.gen_seeded(g, 0, n, n, nz, 1, 21)
    u = np.where(u == 0.0, 0.5, u)  # ghosts hold data
    esz = 8 if kind == 0 else 4
    uh = u.astype(np.float64 if kind == 0 else np.float32)
    if kind == 1 and bf16:
        kh, _ = gpu.quantize_bf16(k)
    else:
        kh = k.astype(uh.dtype)
    ch = np.random.default_rng(4).integers(0, 1 << 10, g.elems).astype(np.uint16)
    rhs = np.random.default_rng(6).uniform(-1e-4, 1e-4, g.elems) if kind == 0 and mode else None
    da, dk, dc = (gpu.DeviceArray.from_host(x) for x in (uh, kh, ch))
    lo, hi = gpu.DeviceArray(ghost * n * n * esz), gpu.DeviceArray(ghost * n * n * esz)
    clo, chi = gpu.DeviceArray(ghost * n * n * 2), gpu.DeviceArray(ghost * n * n * 2)
    s = gpu.Stream()
    t = gpu.Tune()
    c16 = kind == 1 and mode == 2
    gpu.call("h3d_autotune_slab", kind, da.ptr, dk.ptr, int(bf16), dc.ptr if c16 else None,
             gpu.C.byref(g), mode, ghost, lo.ptr, hi.ptr, clo.ptr if c16 else None,
             chi.ptr if c16 else None, s.ptr, gpu.C.byref(t))
    assert da.download().tobytes() == uh.tobytes()
    assert dk.download().tobytes() == kh.tobytes() and dc.download().tobytes() == ch.tobytes()
    assert t.kblock == ghost and t.rows > 0 and t.waves > 0 and t.zchunk > 0

    def run(tune):
        a, b = gpu.DeviceArray.from_host(uh), gpu.DeviceArray.from_host(uh)
        ca, cb = gpu.DeviceArray.from_host(ch), gpu.DeviceArray.from_host(ch)
        ms, fb = gpu.C.c_float(), gpu.C.c_int()
        tp = gpu.C.byref(tune) if tune is not None else None
        if kind == 1:
            kf, kb = (None, dk.ptr) if bf16 else (dk.ptr, None)
            gpu.call("h3d_run_slab_pass_f32", a.ptr, b.ptr, kf, kb, ca.ptr, cb.ptr,
                     gpu.C.byref(g), gpu.C.c_float(1e-3), mode, ghost, lo.ptr, hi.ptr,
                     clo.ptr if c16 else None, chi.ptr if c16 else None, steps, tp, 0, s.ptr,
                     gpu.C.byref(ms), gpu.C.byref(fb))
        else:
            dr = gpu.DeviceArray.from_host(rhs) if rhs is not None else None
            gpu.call("h3d_run_slab_pass_f64", a.ptr, b.ptr, dr.ptr if dr else None, dk.ptr,
                     gpu.C.byref(g), 1e-3, ghost, lo.ptr, hi.ptr, steps, tp, 0, s.ptr,
                     gpu.C.byref(ms), gpu.C.byref(fb))
        out = (b if fb.value else a).download().tobytes()
        return out + ((cb if fb.value else ca).download().tobytes() if c16 else b"")

    assert run(t) == run(None)


def test_autotune_slab_arguments(gpu):
    n, nz = 20, 12
    g = gpu.Grid.dense(n, n, nz)
    a = gpu.DeviceArray(g.elems * 8)
    k = gpu.DeviceArray(g.elems * 8)
    lo, hi = gpu.DeviceArray(4 * n * n * 8), gpu.DeviceArray(4 * n * n * 8)
    L = gpu.lib()
    EARG = gpu.H3D_EARG
    s = gpu.Stream()
    t = gpu.Tune()
    args = lambda kind, mode, ghost, stream=s.ptr, comp=None: (
        kind, a.ptr, k.ptr, 0, comp, gpu.C.byref(g), mode, ghost, lo.ptr, hi.ptr, None, None,
        stream, gpu.C.byref(t))
    assert L.h3d_autotune_slab(*args(0, 0, 2, stream=None)) == EARG  # needs a created stream
    assert L.h3d_autotune_slab(*args(0, 0, 0)) == EARG               # ghost depth 0
    assert L.h3d_autotune_slab(*args(0, 0, 5)) == EARG               # FP64 ghosts 1..4
    assert L.h3d_autotune_slab(*args(0, 2, 2)) == EARG               # FP64 modes 0, 1
    assert L.h3d_autotune_slab(*args(1, 2, 2)) == EARG               # COMP16 needs residuals
    assert L.h3d_autotune_slab(*args(1, 1, 5)) == EARG               # slab too thin
    assert L.h3d_autotune_slab(*args(2, 0, 2)) == EARG               # no such kind
