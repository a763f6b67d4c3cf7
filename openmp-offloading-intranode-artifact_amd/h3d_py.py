"""ctypes binding of libh3d.so (C ABI declared in include/h3d.h).

Used by the tests and by bench.py.  It loads the in-tree library built by
`make -C openmp-offloading-intranode-artifact_amd` and raises if it is
missing: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libh3d.so")

H3D_OK, H3D_EARG, H3D_ELIMIT, H3D_ENODEV, H3D_EHOST = 0, -1, -2, -3, -4
F32_PLAIN, F32_CARRY, F32_COMP16 = 0, 1, 2


class H3DError(RuntimeError):
    def __init__(self, call, code, msg):
        super().__init__(f"{call} failed: {msg} ({code})")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("sy", C.c_int), ("sz", C.c_longlong)]

    @classmethod
    def dense(cls, nx, ny=None, nz=None):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        return cls(nx, ny, nz, nx, nx * ny)

    @property
    def elems(self):
        return self.sz * self.nz


class Tune(C.Structure):
    _fields_ = [("rows", C.c_int), ("waves", C.c_int), ("zchunk", C.c_int),
                ("kblock", C.c_int)]


class Member(C.Structure):
    _fields_ = [("u_a", C.c_void_p), ("u_b", C.c_void_p),
                ("kappa", C.c_void_p), ("g", Grid), ("factor", C.c_double),
                ("precision", C.c_int)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_GP = C.POINTER(Grid)
_TP = C.POINTER(Tune)

# name -> (restype, argtypes)
SIGNATURES = {
    "h3d_heat_update": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _I]),
    "h3d_heat_update_interior": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _I, _I, _I]),
    "h3d_heat_update_boundary_bottom": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _I]),
    "h3d_heat_update_boundary_top": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _I]),
    "h3d_step_f64": (_I, [_P, _P, _P, _P, _GP, _D, _I, _I, _TP, _P]),
    "h3d_pass_f64": (_I, [_P, _P, _P, _P, _GP, _D, _I, _I, _I, _TP, _P]),
    "h3d_run_f64": (_I, [_P, _P, _P, _P, _GP, _D, _I, _TP, _I, _P,
                         C.POINTER(C.c_float), C.POINTER(_I)]),
    "h3d_run_f64_passes": (_I, [_GP, _I, _I, _TP, C.POINTER(_I), C.POINTER(_I)]),
    "h3d_slab_step_f64": (_I, [_P, _P, _P, _P, _GP, _D, _P, _P, _TP, _P]),
    "h3d_run_slab_f64": (_I, [_P, _P, _P, _P, _GP, _D, _P, _P, _I, _TP, _I, _P,
                              C.POINTER(C.c_float), C.POINTER(_I)]),
    "h3d_slab_pass_f64": (_I, [_P, _P, _P, _P, _GP, _D, _I, _I, _P, _P, _TP, _P]),
    "h3d_run_slab_pass_f64": (_I, [_P, _P, _P, _P, _GP, _D, _I, _P, _P, _I, _TP, _I,
                                   _P, C.POINTER(C.c_float), C.POINTER(_I)]),
    "h3d_pack_band_f64": (_I, [_P, _P, _GP, _I, _I, _P]),
    "h3d_unpack_band_f64": (_I, [_P, _P, _GP, _I, _I, _P]),
    "h3d_pack_band_f32": (_I, [_P, _P, _GP, _I, _I, _P]),
    "h3d_unpack_band_f32": (_I, [_P, _P, _GP, _I, _I, _P]),
    "h3d_pack_band_u16": (_I, [_P, _P, _GP, _I, _I, _P]),
    "h3d_unpack_band_u16": (_I, [_P, _P, _GP, _I, _I, _P]),
    "h3d_slab_pass_f32": (_I, [_P, _P, _P, _P, _P, _P, _GP, C.c_float, _I, _I, _I, _P, _P,
                               _P, _P, _TP, _P]),
    "h3d_run_slab_pass_f32": (_I, [_P, _P, _P, _P, _P, _P, _GP, C.c_float, _I, _I, _P, _P,
                                   _P, _P, _I, _TP, _I, _P, C.POINTER(C.c_float),
                                   C.POINTER(C.c_int)]),
    "h3d_pack_face_f64": (_I, [_P, _P, _GP, _I, _P]),
    "h3d_unpack_face_f64": (_I, [_P, _P, _GP, _I, _P]),
    "h3d_pack_face_f32": (_I, [_P, _P, _GP, _I, _P]),
    "h3d_unpack_face_f32": (_I, [_P, _P, _GP, _I, _P]),
    "h3d_stats_f64": (_I, [_P, _GP, _P, _P]),
    "h3d_stats_f32": (_I, [_P, _GP, _P, _P]),
    "h3d_stats_f64_noalloc": (_I, [_P, _GP, _P, _P]),
    "h3d_compare_f64": (_I, [_P, _P, _GP, _D, _D, _P, _P]),
    "h3d_compare_f32": (_I, [_P, _P, _GP, _D, _D, _P, _P]),
    "h3d_face_sum_max_f64": (_I, [_P, _GP, _P, _P]),
    "h3d_pass_f32": (_I, [_P, _P, _P, _P, _P, _P, _GP, C.c_float, _I, _I, _TP, _P]),
    "h3d_pass_f32_range": (_I, [_P, _P, _P, _P, _P, _P, _GP, C.c_float, _I, _I, _I, _I, _TP,
                                _P]),
    "h3d_run_f32": (_I, [_P, _P, _P, _P, _P, _P, _GP, C.c_float, _I, _I, _TP,
                         _I, _P, C.POINTER(C.c_float), C.POINTER(_I)]),
    "h3d_autotune": (_I, [_I, _P, _P, _P, _I, _GP, _I, _I, _P, _TP]),
    "h3d_autotune_slab": (_I, [_I, _P, _P, _I, _P, _GP, _I, _I, _P, _P, _P, _P, _P, _TP]),
    "h3d_device_count": (_I, [C.POINTER(_I)]),
    "h3d_target_alloc": (_I, [C.POINTER(_P), C.c_size_t, _I]),
    "h3d_target_free": (_I, [_P, _I]),
    "h3d_target_memcpy": (_I, [_P, _P, C.c_size_t, _I, _I]),
    "h3d_device_synchronize": (_I, [_I]),
    "h3d_stream_create": (_I, [C.POINTER(_P), _I]),
    "h3d_stream_destroy": (_I, [_P]),
    "h3d_stream_synchronize": (_I, [_P]),
    "h3d_error_string": (C.c_char_p, [_I]),
    "h3d_device_arch": (_I, [_I, C.c_char_p, _I]),
    "h3d_stream_probe": (_I, [C.c_size_t, _I, _I, C.POINTER(_D), C.POINTER(_D)]),
    "h3d_transfer_probe": (_I, [C.c_size_t, _I, _I, _I, _P]),
    "h3d_transfer_bench": (_I, [C.c_size_t, _I, _I, _I, _I, _P]),
    "h3d_params_reference": (_I, [_I, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
    "h3d_count_steps": (_I, [_D, _D]),
    "h3d_init_reference": (_I, [_P, _P, _P, _GP]),
    "h3d_init_partition": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _D]),
    "h3d_slab_decompose": (_I, [_I, _I, C.POINTER(_I), C.POINTER(_I)]),
    "h3d_slab_reconstruct": (_I, [_P, _P, _I, _I, _I]),
    "h3d_gen_seeded": (_I, [_P, _P, _GP, _I, _I, _I, _I, _I, C.c_uint64, _I]),
    "h3d_rng_uniform": (_D, [C.c_uint64, _I, _I, C.c_uint64]),
    "h3d_face_sum_max_host": (_D, [_P, _GP]),
    "h3d_quantize_bf16": (_I, [_P, _P, C.c_size_t, C.POINTER(_D)]),
    "h3d_bf16_scale": (_I, [_P, C.c_size_t, C.POINTER(_D)]),
    "h3d_quantize_bf16_scaled": (_I, [_P, _P, C.c_size_t, _D]),
    "h3d_write_dump": (_I, [C.c_char_p, _P, _GP]),
    "h3d_ensemble_create": (_I, [C.POINTER(_P), C.POINTER(Member), _I, _TP, _I]),
    "h3d_ensemble_run": (_I, [_P, _I, _I, _P, C.POINTER(C.c_float)]),
    "h3d_ensemble_autotune": (_I, [_P, _P, _TP]),
    "h3d_ensemble_where": (_I, [_P, _I, C.POINTER(_I)]),
    "h3d_ensemble_destroy": (_I, [_P]),
    "h3d_ensemble_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_longlong)]),
}

_lib = None


def lib():
    """The loaded libh3d.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `make -C "
                               f"{os.path.basename(HERE)}` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def check(name, rc):
    if rc != H3D_OK:
        msg = lib().h3d_error_string(rc)
        raise H3DError(name, rc, msg.decode() if msg else "?")
    return rc


def call(name, *args):
    return check(name, getattr(lib(), name)(*args))


def ptr(a):
    """Host pointer of a numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def device_count():
    n = C.c_int(0)
    call("h3d_device_count", C.byref(n))
    return n.value


class DeviceArray:
    """Device buffer owned by this object (freed on close/GC)."""

    def __init__(self, nbytes, device=0, dtype=np.float64):
        self.nbytes = int(nbytes)
        self.device = device
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        call("h3d_target_alloc", C.byref(p), C.c_size_t(self.nbytes), device)
        self.ptr = p

    @classmethod
    def from_host(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, device, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        call("h3d_target_memcpy", self.ptr, ptr(a), C.c_size_t(a.nbytes), self.device, -1)

    def download(self, shape=None):
        out = np.empty(self.nbytes // self.dtype.itemsize, dtype=self.dtype)
        call("h3d_target_memcpy", ptr(out), self.ptr, C.c_size_t(self.nbytes), -1, self.device)
        return out if shape is None else out.reshape(shape)

    def close(self):
        if self.ptr is not None and _lib is not None:
            lib().h3d_target_free(self.ptr, self.device)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pitched_grid(nx, ny, nz, align_elems=32):
    """Grid whose row stride is a multiple of `align_elems` (32 elements =
    128 B of FP32, 256 B of FP64), for use with AlignedField."""
    sy = -(-nx // align_elems) * align_elems
    return Grid(nx, ny, nz, sy, sy * ny)


class AlignedField:
    """Device buffer for a pitched grid, offset so that element 1 of every
    row (the first interior point) starts a 128-byte line."""

    def __init__(self, g, dtype, device=0):
        self.dtype = np.dtype(dtype)
        self.nbytes = g.elems * self.dtype.itemsize
        self.buf = DeviceArray(self.nbytes + 256, device, self.dtype)
        base = self.buf.ptr.value
        off = (-(base + self.dtype.itemsize)) % 128
        self.ptr = C.c_void_p(base + off)
        self.device = device

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        call("h3d_target_memcpy", self.ptr, ptr(a), C.c_size_t(a.nbytes), self.device, -1)

    def download(self):
        out = np.empty(self.nbytes // self.dtype.itemsize, dtype=self.dtype)
        call("h3d_target_memcpy", ptr(out), self.ptr, C.c_size_t(self.nbytes), -1, self.device)
        return out

    def close(self):
        self.buf.close()


class Stream:
    def __init__(self, device=0):
        p = C.c_void_p()
        call("h3d_stream_create", C.byref(p), device)
        self.ptr = p

    def sync(self):
        call("h3d_stream_synchronize", self.ptr)

    def close(self):
        if self.ptr is not None and _lib is not None:
            lib().h3d_stream_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- host-side helpers ----------------------------------------------------

def params_reference(n):
    h, dt, f = C.c_double(), C.c_double(), C.c_double()
    call("h3d_params_reference", n, C.byref(h), C.byref(dt), C.byref(f))
    return h.value, dt.value, f.value


def count_steps(T, dt):
    return lib().h3d_count_steps(T, dt)


def init_reference(n, with_rhs=True):
    g = Grid.dense(n)
    u = np.empty(g.elems)
    k = np.empty(g.elems)
    r = np.empty(g.elems) if with_rhs else None
    call("h3d_init_reference", ptr(u), ptr(r), ptr(k), C.byref(g))
    return u, r, k


def gen_seeded(g, gk0, gnx, gny, gnz, kappa_kind, seed, member=0):
    u = np.zeros(g.elems)  # row/plane padding of a pitched grid stays 0
    k = np.zeros(g.elems)
    call("h3d_gen_seeded", ptr(u), ptr(k), C.byref(g), gk0, gnx, gny, gnz,
         kappa_kind, C.c_uint64(seed), member)
    return u, k


def slab_decompose(n, parts):
    nz = (C.c_int * 4)()
    gk = (C.c_int * 4)()
    rc = lib().h3d_slab_decompose(n, parts, nz, gk)
    return rc, list(nz)[:parts], list(gk)[:parts]


def quantize_bf16(kappa):
    kappa = np.ascontiguousarray(kappa, dtype=np.float64)
    bits = np.empty(kappa.size, dtype=np.uint16)
    s = C.c_double()
    call("h3d_quantize_bf16", ptr(kappa), ptr(bits), C.c_size_t(kappa.size), C.byref(s))
    return bits, s.value


def quantize_bf16_scaled(kappa, scale):
    """bf16(kappa/scale) with a given power-of-two scale (a slab of a field
    takes the whole field's scale)."""
    kappa = np.ascontiguousarray(kappa, dtype=np.float64)
    bits = np.empty(kappa.size, dtype=np.uint16)
    call("h3d_quantize_bf16_scaled", ptr(kappa), ptr(bits), C.c_size_t(kappa.size),
         C.c_double(scale))
    return bits


def bf16_to_f64(bits, scale):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64) * scale


def stats(d, g, f32=False, stream=None):
    out = np.zeros(4)
    call("h3d_stats_f32" if f32 else "h3d_stats_f64", d.ptr, C.byref(g), ptr(out),
         stream.ptr if stream else None)
    return out


def compare(ref, cand, g, atol=1e-6, rtol=1e-5, f32=False, stream=None):
    out = np.zeros(9)
    call("h3d_compare_f32" if f32 else "h3d_compare_f64", ref.ptr, cand.ptr,
         C.byref(g), atol, rtol, ptr(out), stream.ptr if stream else None)
    return out


def stream_probe(nbytes, reps=5, device=0):
    c, t = C.c_double(), C.c_double()
    call("h3d_stream_probe", C.c_size_t(nbytes), reps, device, C.byref(c), C.byref(t))
    return c.value, t.value


def transfer_bench(nbytes, pinned=True, asynchronous=False, d2h=False, device=0):
    """(GB/s, seconds, words wrong, words checked) of one verified copy."""
    out = np.zeros(4)
    call("h3d_transfer_bench", C.c_size_t(nbytes), int(pinned), int(asynchronous),
         int(d2h), device, ptr(out))
    return float(out[0]), float(out[1]), int(out[2]), int(out[3])


def transfer_probe(nbytes, reps=3, pinned=True, device=0):
    """(H2D GB/s, D2H GB/s, H2D us, D2H us) of host<->device copies."""
    out = np.zeros(4)
    call("h3d_transfer_probe", C.c_size_t(nbytes), reps, int(pinned), device, ptr(out))
    return tuple(float(x) for x in out)
