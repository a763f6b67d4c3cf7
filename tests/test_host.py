"""CPU-only checks of the product library (no compute on a device): every
symbol of include/h3d.h is exported, and the host-side set-up functions
agree with the oracle (reference mode) or with an independent numpy
restatement (seeded generators, which the reference does not have)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "h3d.h")
PKG = os.path.join(REPO, "openmp-offloading-intranode-artifact_amd")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(h3d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(h3d):
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(h3d.lib(), n)]
    assert not missing
    # and the ctypes binding knows every one of them
    assert not [n for n in names if n not in h3d.SIGNATURES]


def test_error_strings(h3d):
    L = h3d.lib()
    assert L.h3d_error_string(0) == b"success"
    assert L.h3d_error_string(-1) == b"invalid argument"
    assert L.h3d_error_string(-3) == b"no such device"


def test_params_and_steps_match_oracle(h3d, oracle):
    for n in (3, 5, 51, 256, 514, 1026):
        assert h3d.params_reference(n) == oracle.params_reference(n)
        h, dt, f = oracle.params_reference(n)
        for s in (1, 3, 100, 500):
            T = float("%.7g" % (s * dt))
            assert h3d.count_steps(T, dt) == oracle.count_steps(T, dt)


@pytest.mark.parametrize("n", [5, 17, 33])
def test_init_reference_bitwise(h3d, oracle, n):
    u, r, k = h3d.init_reference(n)
    ou, _, orr, ok = oracle.init_reference(n)
    assert np.array_equal(u, ou) and np.array_equal(r, orr) and np.array_equal(k, ok)


def test_init_partition_matches_oracle(h3d, oracle):
    n = 18
    h, dt, f = oracle.params_reference(n)
    rc, nzd, gks = h3d.slab_decompose(n, 4)
    assert (rc, nzd, gks) == oracle.slab_decompose(n, 4)
    L = h3d.lib()
    for nz_l, gk in zip(nzd, gks):
        a = [np.empty(nz_l * n * n) for _ in range(3)]
        b = [np.empty(nz_l * n * n) for _ in range(3)]
        assert L.h3d_init_partition(*(h3d.ptr(x) for x in a), nz_l, gk, 1, n, n, n, h, dt) == 0
        oracle.lib().orc_init_partition(*(oracle.p(x) for x in b), nz_l, gk, 1, n, n, n, h, dt)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_slab_decompose_matches_oracle(h3d, oracle):
    for n in range(3, 40):
        for parts in (1, 2, 4):
            a, b = h3d.slab_decompose(n, parts), oracle.slab_decompose(n, parts)
            assert (a[0] == 0) == (b[0] == 0), (n, parts)
            if a[0] == 0:
                assert a[1:] == b[1:]


def splitmix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def np_uniform(seed, member, stream, idx):
    with np.errstate(over="ignore"):
        key = splitmix(np.uint64(seed) ^ splitmix(np.uint64((member << 8) | stream)))
        z = splitmix(key + np.uint64(0x9E3779B97F4A7C15) * (idx.astype(np.uint64) + np.uint64(1)))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def test_rng_matches_numpy_restatement(h3d):
    idx = np.arange(0, 5000, 7, dtype=np.uint64)
    for seed, member, stream in [(20260101, 0, 1), (1, 3, 2), (2 ** 63 + 5, 63, 4)]:
        want = np_uniform(seed, member, stream, idx)
        got = np.array([h3d.lib().h3d_rng_uniform(seed, member, stream, int(i)) for i in idx])
        assert np.array_equal(got, want)
        assert got.min() >= 0.0 and got.max() < 1.0


@pytest.mark.parametrize("kind", [0, 1])
def test_seeded_fields_match_numpy(h3d, kind):
    nx, ny, nz = 9, 7, 6
    seed = 77
    g = h3d.Grid.dense(nx, ny, nz)
    u, k = h3d.gen_seeded(g, 0, nx, ny, nz, kind, seed, member=2)
    h = 1.0 / (nx - 1)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    gid = ((K * ny + J) * nx + I).ravel().astype(np.uint64)
    x, y, z = (I * h).ravel(), (J * h).ravel(), (K * h).ravel()
    if kind == 0:
        base = 1 + 0.5 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
        want_k = base * (0.9 + 0.2 * np_uniform(seed, 2, 1, gid))
    else:
        a = 1.0 - np_uniform(seed, 2, 3, gid)
        b = np_uniform(seed, 2, 4, gid)
        want_k = np.exp(0.5 * np.sqrt(-2 * np.log(a)) * np.cos(2 * np.pi * b))
    assert np.allclose(k, want_k, rtol=4e-16, atol=0)
    c = (np.array([nx, ny, nz]) - 1) * h * 0.5
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    want_u = np.exp(-r2 / (2 * 0.1 ** 2)) + 1e-3 * np_uniform(seed, 2, 2, gid)
    ring = ((I == 0) | (I == nx - 1) | (J == 0) | (J == ny - 1) | (K == 0) | (K == nz - 1)).ravel()
    want_u[ring] = 0.0
    assert np.allclose(u, want_u, rtol=4e-16, atol=0)
    assert np.all(u[ring] == 0.0)
    if kind == 0:
        assert k.min() > 0.5 * 0.9 - 1e-12 and k.max() < 1.5 * 1.1 + 1e-12


def test_seeded_slab_equals_global_slice(h3d):
    n = 16
    g = h3d.Grid.dense(n)
    u, k = h3d.gen_seeded(g, 0, n, n, n, 1, 99)
    gs = h3d.Grid.dense(n, n, 6)
    su, sk = h3d.gen_seeded(gs, 5, n, n, n, 1, 99)
    assert np.array_equal(su, u[5 * n * n:11 * n * n])
    assert np.array_equal(sk, k[5 * n * n:11 * n * n])
    # pitched output holds the same values
    gp = h3d.Grid(n, n, n, n + 3, (n + 3) * (n + 1))
    pu, pk = h3d.gen_seeded(gp, 0, n, n, n, 1, 99)
    P = pk.reshape(n, n + 1, n + 3)[:, :n, :n].ravel()
    assert np.array_equal(P, k)


def test_face_sum_max_host_matches_oracle(h3d, oracle):
    n = 12
    g = h3d.Grid.dense(n)
    _, k = h3d.gen_seeded(g, 0, n, n, n, 1, 4)
    assert h3d.lib().h3d_face_sum_max_host(h3d.ptr(k), g) == oracle.face_sum_max(k, n, n, n)


def test_quantize_bf16_power_of_two_scale(h3d):
    k = np.random.default_rng(2).lognormal(0, 0.5, 4000)
    bits, s = h3d.quantize_bf16(k)
    assert s == 2.0 ** np.ceil(np.log2(k.max())) or s == 2.0 ** (np.floor(np.log2(k.max())) + 1)
    deq = h3d.bf16_to_f64(bits, s)
    assert np.all(np.abs(deq - k) <= k * 2.0 ** -8)
    # bf16(x) * s is exact: the quantised values times the scale are the
    # values the FP64 reference run uses
    assert np.array_equal(deq / s, (bits.astype(np.uint32) << 16).view(np.float32))


def test_bf16_slab_of_a_field_takes_the_field_scale(h3d):
    """A slab's kappa quantized with the whole field's scale has the bits of
    the field's quantization there; quantized with its own maximum's scale
    it can differ (a slab whose maximum is below half the field's gets a
    smaller power of two).  The scale is the power of two above max|kappa|."""
    rng = np.random.default_rng(8)
    k = rng.uniform(0.1, 3.0, 4000)
    k[:1000] *= 0.2  # the first "slab" peaks below half the field's maximum
    bits, s = h3d.quantize_bf16(k)
    sc = h3d.C.c_double()
    assert h3d.lib().h3d_bf16_scale(h3d.ptr(k), h3d.C.c_size_t(k.size), h3d.C.byref(sc)) == 0
    assert sc.value == s == 4.0  # max in [2, 4)
    slab = k[:1000]
    assert np.array_equal(h3d.quantize_bf16_scaled(slab, s), bits[:1000])
    own, s_own = h3d.quantize_bf16(slab)
    assert s_own == 1.0 and not np.array_equal(own, bits[:1000])
    np.testing.assert_array_equal(h3d.bf16_to_f64(h3d.quantize_bf16_scaled(slab, s), s),
                                  h3d.bf16_to_f64(bits[:1000], s))
    L = h3d.lib()
    out = np.empty(4, np.uint16)
    for bad in (3.0, 0.0, -2.0, float("inf")):  # not a positive power of two
        assert L.h3d_quantize_bf16_scaled(h3d.ptr(k), h3d.ptr(out), h3d.C.c_size_t(4),
                                          h3d.C.c_double(bad)) == h3d.H3D_EARG
    assert L.h3d_quantize_bf16_scaled(h3d.ptr(np.array([1e300, 0, 0, 0])), h3d.ptr(out),
                                      h3d.C.c_size_t(4), h3d.C.c_double(1.0)) == h3d.H3D_EARG


def test_write_dump_format(h3d, tmp_path):
    n = 4
    g = h3d.Grid.dense(n)
    u = np.random.default_rng(0).standard_normal(n ** 3)
    path = str(tmp_path / "d.txt")
    assert h3d.lib().h3d_write_dump(path.encode(), h3d.ptr(u), g) == 0
    lines = open(path).read().splitlines()
    assert lines == ["%.17g" % v for v in u]


def test_argument_errors_need_no_device(h3d):
    L = h3d.lib()
    g = h3d.Grid(2, 5, 5, 2, 10)  # nx < 3
    assert L.h3d_step_f64(None, None, None, None, g, 0.1, 1, 3, None, None) == -1
    g = h3d.Grid(5, 5, 5, 4, 25)  # sy < nx
    assert L.h3d_stats_f64(None, g, None, None) == -1
    nz = (h3d.C.c_int * 4)()
    assert L.h3d_slab_decompose(9, 4, nz, nz) == -1
    assert L.h3d_slab_decompose(9, 3, nz, nz) == -1


def test_cli_rejects_bad_arguments():
    exe = os.path.join(PKG, "bin", "h3d")
    r = subprocess.run([exe, "2", "1.0"], capture_output=True, text=True)
    assert r.returncode == 1 and "at least 3 points" in r.stderr
    r = subprocess.run([exe, "10", "--precision", "f16"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([exe, "9", "--slab", "4:0", "--steps", "1"], capture_output=True,
                       text=True, env=dict(os.environ, WRITE_TXT="0"))
    assert r.returncode == 1 and "cannot be split" in r.stderr
    # deep slabs: 1..4 ghost planes per side, and at least as many owned planes
    r = subprocess.run([exe, "34", "--slab", "4:1", "--kblock", "5", "--steps", "1"],
                       capture_output=True, text=True, env=dict(os.environ, WRITE_TXT="0"))
    assert r.returncode == 1 and "1..4" in r.stderr
    r = subprocess.run([exe, "14", "--slab", "4:1", "--kblock", "4", "--steps", "1"],
                       capture_output=True, text=True, env=dict(os.environ, WRITE_TXT="0"))
    assert r.returncode == 1 and "cannot hold 4 ghost planes" in r.stderr


def test_verify_tool_exit_codes_without_device(tmp_path):
    """compare.py's exit codes (compare.py:120-151): 2 without a reference
    file, 2 on a bad option; nothing here touches a GPU."""
    exe = os.path.join(PKG, "bin", "h3d-verify")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2
    assert "ERROR: reference file 'u_serial.txt' not found." in r.stdout
    r = subprocess.run([exe, "--tolerance", "1"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    (tmp_path / "r.txt").write_text("1\n2\n")
    r = subprocess.run([exe, "--ref", "r.txt"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "not n^3" in r.stdout


@pytest.mark.parametrize("n,parts", [(9, 2), (10, 2), (11, 4), (22, 4)])
def test_slab_reconstruct_is_the_reference_assembly(h3d, n, parts):
    """Each slab holds global planes gk_start-1 .. gk_start+nz_d-2; the
    reconstruction takes the owned ones plus the physical boundary planes
    (4-openmp-stream-omp-p2p.c:776-820) and so rebuilds the global array
    exactly, whichever slab writes first; interface ghosts come from the
    slab that owns them."""
    L = h3d.lib()
    nz_d, gks = (C.c_int * 4)(), (C.c_int * 4)()
    assert L.h3d_slab_decompose(n, parts, nz_d, gks) == 0
    nn = n * n
    glob = np.arange(n * nn, dtype=np.float64)
    out = np.full(n * nn, -1.0)
    for p in reversed(range(parts)):
        planes = [gks[p] - 1 + k for k in range(nz_d[p])]
        slab = np.concatenate([glob[g * nn:(g + 1) * nn] for g in planes])
        # ghost planes hold stale data in a real run: make them wrong here
        if p > 0:
            slab[:nn] = -7.0
        if p < parts - 1:
            slab[-nn:] = -7.0
        assert L.h3d_slab_reconstruct(h3d.ptr(out), h3d.ptr(slab), n, parts, p) == 0
    assert np.array_equal(out, glob)
    assert L.h3d_slab_reconstruct(h3d.ptr(out), h3d.ptr(glob), n, parts, parts) == -1
    assert L.h3d_slab_reconstruct(h3d.ptr(out), h3d.ptr(glob), n, 3, 0) == -1


def test_slab_reconstruct_plane_ranges_n1026(h3d):
    """At the slab configuration's size (n = 1026, 4 slabs of 258 planes)
    each slab contributes 256 owned planes, the outer ones one boundary
    plane more: ranges 0..256, 257..512, 513..768, 769..1025."""
    L = h3d.lib()
    nz_d, gks = (C.c_int * 4)(), (C.c_int * 4)()
    assert L.h3d_slab_decompose(1026, 4, nz_d, gks) == 0
    assert list(nz_d) == [258] * 4 and list(gks) == [1, 257, 513, 769]

