"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported by tests/, smoke() and bench.py's cpu_baseline
leg only, always as the checker, never as the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_P, _I, _D, _LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong
SIGS = {
    "orc_params_reference": (None, [_I, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
    "orc_count_steps": (_I, [_D, _D]),
    "orc_init_reference": (None, [_P, _P, _P, _P, _I]),
    "orc_init_partition": (None, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _D]),
    "orc_step_f64": (None, [_P, _P, _P, _P, _I, _I, _I, _I, _LL, _D, _I, _I]),
    "orc_run_f64": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _LL, _D, _I, _I]),
    "orc_slab_decompose": (_I, [_I, _I, C.POINTER(_I), C.POINTER(_I)]),
    "orc_pack_face": (None, [_P, _P, _I, _I, _I, _LL, _I]),
    "orc_unpack_face": (None, [_P, _P, _I, _I, _I, _LL, _I]),
    "orc_metrics": (None, [_P, _P, C.c_size_t, _D, _D, _P]),
    "orc_stats": (None, [_P, _I, _I, _I, _I, _LL, _P]),
    "orc_face_sum_max": (_D, [_P, _I, _I, _I, _I, _LL]),
    "orc_write_dump": (_I, [C.c_char_p, _P, _I]),
    "orc_run_f32": (None, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _LL,
                           C.c_float, _I, _I, _I, C.POINTER(_I)]),
    "orc_comp_encode": (C.c_uint16, [C.c_float, C.c_float]),
    "orc_comp_decode": (C.c_float, [C.c_uint16, C.c_float]),
    "orc_bf16_to_f32": (C.c_float, [C.c_uint16]),
    "orc_f32_to_bf16": (C.c_uint16, [C.c_float]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                           capture_output=True)
        L = C.CDLL(LIB_PATH)
        for n, (r, a) in SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def params_reference(n):
    h, dt, f = C.c_double(), C.c_double(), C.c_double()
    lib().orc_params_reference(n, C.byref(h), C.byref(dt), C.byref(f))
    return h.value, dt.value, f.value


def count_steps(T, dt):
    return lib().orc_count_steps(T, dt)


def init_reference(n):
    u = np.empty(n ** 3)
    un = np.empty(n ** 3)
    r = np.empty(n ** 3)
    k = np.empty(n ** 3)
    lib().orc_init_reference(p(u), p(un), p(r), p(k), n)
    return u, un, r, k


def run_f64(ua, ub, rhs, kappa, nx, ny, nz, steps, sy=None, sz=None, factor=None,
            nthreads=1):
    """Runs in place; returns the array holding the final field."""
    sy = nx if sy is None else sy
    sz = sy * ny if sz is None else sz
    w = lib().orc_run_f64(p(ua), p(ub), p(rhs), p(kappa), nx, ny, nz, sy, sz,
                          factor, steps, nthreads)
    return ub if w else ua


def step_f64(un, uo, rhs, kappa, nx, ny, nz, factor, k_start, k_end, sy=None, sz=None):
    sy = nx if sy is None else sy
    sz = sy * ny if sz is None else sz
    lib().orc_step_f64(p(un), p(uo), p(rhs), p(kappa), nx, ny, nz, sy, sz, factor,
                       k_start, k_end)


def slab_decompose(n, parts):
    nz = (C.c_int * 4)()
    gk = (C.c_int * 4)()
    rc = lib().orc_slab_decompose(n, parts, nz, gk)
    return rc, list(nz)[:parts], list(gk)[:parts]


def metrics(ref, cand, atol=1e-6, rtol=1e-5):
    out = np.zeros(9)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    cand = np.ascontiguousarray(cand, dtype=np.float64)
    lib().orc_metrics(p(ref), p(cand), ref.size, atol, rtol, p(out))
    return out


def stats(u, nx, ny, nz, sy=None, sz=None):
    sy = nx if sy is None else sy
    sz = sy * ny if sz is None else sz
    out = np.zeros(4)
    lib().orc_stats(p(np.ascontiguousarray(u, dtype=np.float64)), nx, ny, nz, sy, sz, p(out))
    return out


def face_sum_max(kappa, nx, ny, nz, sy=None, sz=None):
    sy = nx if sy is None else sy
    sz = sy * ny if sz is None else sz
    return lib().orc_face_sum_max(p(kappa), nx, ny, nz, sy, sz)


def run_f32(u0, kappa_f32, kappa_bf16, nx, ny, nz, factor, steps, kb, mode,
            comp0=None, sy=None, sz=None):
    """Returns (u_final float32, comp_final uint16 or None)."""
    sy = nx if sy is None else sy
    sz = sy * ny if sz is None else sz
    ua = np.array(u0, dtype=np.float32, copy=True)
    ub = ua.copy()
    n = ua.size
    ca = (np.zeros(n, np.uint16) if comp0 is None else np.array(comp0, np.uint16))
    cb = ca.copy()
    fi = C.c_int(0)
    lib().orc_run_f32(p(ua), p(ub), p(kappa_f32), p(kappa_bf16), p(ca), p(cb),
                      nx, ny, nz, sy, sz, C.c_float(factor), steps, kb, mode,
                      C.byref(fi))
    u = ub if (fi.value & 1) else ua
    c = cb if (fi.value & 2) else ca
    return u, (c if mode == 2 else None)
