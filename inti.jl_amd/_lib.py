"""ctypes binding of libinti_amd.so (the C-ABI in include/inti_amd.h).

The library is built in-tree next to this file. There is no fallback: if it
is missing or has no GPU to run on, calls raise.
"""
import atexit
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libinti_amd.so")

INTI_OK = 0
INTI_ERR_ARG = 1
INTI_ERR_HIP = 2
INTI_ERR_UNSUPPORTED = 3
INTI_ERR_NOMEM = 4
INTI_ERR_SIZE = 5

LAPLACE, HELMHOLTZ = 0, 1
SL, DL, ADL, HS, COMB_SD, COMB_AH = 0, 1, 2, 3, 4, 5
FF_SL, FF_DL, FF_COMB = 6, 7, 8
FP64, MIXED = 0, 1

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "inti_version": (C.c_int, []),
    "inti_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "inti_ctx_destroy": (C.c_int, [_vp]),
    "inti_last_error": (C.c_char_p, [_vp]),
    "inti_ctx_synchronize": (C.c_int, [_vp]),
    "inti_ctx_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "inti_ctx_set_profiling": (C.c_int, [_vp, C.c_int]),
    "inti_ctx_last_kernel_ms": (C.c_int, [_vp, _dp]),
    "inti_device_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "inti_device_free": (C.c_int, [_vp, _vp]),
    "inti_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "inti_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "inti_device_copy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "inti_quad_upload": (C.c_int, [_vp, C.c_int, C.c_int64, _dp, _dp, _dp, C.POINTER(_vp)]),
    "inti_quad_destroy": (C.c_int, [_vp]),
    "inti_op_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _vp, _vp,
                                 C.c_int, C.POINTER(_vp)]),
    "inti_op_destroy": (C.c_int, [_vp]),
    "inti_op_set_identity": (C.c_int, [_vp, C.c_double, C.c_double]),
    "inti_op_set_wavenumber_imag": (C.c_int, [_vp, C.c_double]),
    "inti_op_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "inti_op_set_symmetric": (C.c_int, [_vp, C.c_int, C.c_int]),
    "inti_op_apply": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "inti_op_apply_device": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64, _dp, _dp]),
    "inti_microbench": (C.c_int, [_vp, C.c_int, _dp]),
    "inti_op_apply_device_noalloc": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64, _dp, _dp]),
    "inti_op_assemble_dense": (C.c_int, [_vp, _dp]),
    "inti_op_adjoint": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "inti_op_assemble_dense_device": (C.c_int, [_vp, _vp, C.c_int64]),
    "inti_dense_add_csr": (C.c_int, [_vp, _vp, C.c_double, C.c_double, _vp, C.c_int64]),
    "inti_dense_gemm": (C.c_int, [_vp, C.c_int, C.c_int64, C.c_int64, C.c_int, _vp, C.c_int64, _vp,
                                  C.c_int64, _vp, C.c_int64, _dp, _dp]),
    "inti_batch_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp), _dp, _dp,
                                    C.c_int, C.POINTER(_vp)]),
    "inti_batch_destroy": (C.c_int, [_vp]),
    "inti_batch_info": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "inti_batch_set_identity": (C.c_int, [_vp, _dp]),
    "inti_batch_attach_correction": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_double]),
    "inti_batch_bdim_build": (C.c_int, [_vp, C.c_int, _i64p, _i64p, _dp, _dp, C.POINTER(_vp), C.POINTER(_vp)]),
    "inti_batch_set_launch": (C.c_int, [_vp, C.c_int64, C.c_int]),
    "inti_batch_apply_device": (C.c_int, [_vp, C.c_int, _vp, C.c_int64, _vp, C.c_int64, _dp, _dp]),
    "inti_batch_apply": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "inti_batch_autotune": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double)]),
    "inti_batch_gmres": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.POINTER(C.c_int), _dp]),
    "inti_op_attach_correction": (C.c_int, [_vp, _vp, C.c_double, C.c_double]),
    "inti_op_detach_correction": (C.c_int, [_vp]),
    "inti_bdim_build": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int64, _i64p, _i64p, _i64p,
                                  _dp, _dp, C.POINTER(_vp), C.POINTER(_vp), _dp]),
    "inti_bdim_build_types": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_int), _i64p, _i64p, _i64p,
                                        _i64p, _dp, _dp, C.POINTER(_vp), C.POINTER(_vp), _dp]),
    "inti_lsq_build": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int64, _dp, _i64p, C.c_int64,
                                 C.c_int64, _i64p, _i64p, _dp, C.POINTER(_vp), _dp]),
    "inti_lsq_build_types": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _i64p, _dp, _i64p,
                                       C.c_int64, C.c_int64, _i64p, _i64p, _dp, C.POINTER(_vp), _dp]),
    "inti_gmres": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_int, C.POINTER(C.c_int), _dp]),
    "inti_gmres_prec": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.c_int, C.POINTER(C.c_int), _dp]),
    "inti_refine_solve_prec": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_double,
                                         C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), _dp]),
    "inti_refine_solve": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_double,
                                    C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), _dp]),
    "inti_csr_info": (C.c_int, [_vp, _i64p, _i64p, _i64p, C.POINTER(C.c_int)]),
    "inti_csr_export": (C.c_int, [_vp, _i64p, _i64p, _dp]),
    "inti_csr_destroy": (C.c_int, [_vp]),
}

_lib = None
_lock = threading.RLock()
_exiting = False


@atexit.register
def _mark_exit():
    """At interpreter exit no handle is released any more: the process
    teardown frees the device memory, and no library call races the
    finalisation order of Python objects."""
    global _exiting
    _exiting = True


def exiting():
    return _exiting


class IntiError(RuntimeError):
    """Raised when a libinti_amd call returns a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libinti_amd error {code}: {msg}")
        self.code = code


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBPATH):
                raise ImportError(f"{LIBPATH} not built; run __graft_entry__.build()")
            lib = C.CDLL(LIBPATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            _lib = lib
    return _lib


def check(ctx, code):
    if code != INTI_OK:
        msg = load().inti_last_error(ctx).decode()
        raise IntiError(code, msg)


def dptr(a):
    """double* of a C-contiguous float64/complex128 numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.ndim == 1
    if a.dtype == np.complex128:
        return a.ctypes.data_as(_dp)
    assert a.dtype == np.float64, a.dtype
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64
    return a.ctypes.data_as(_i64p)


class Context:
    """One device context (HIP stream + scratch) of the library."""

    def __init__(self, device=0):
        lib = load()
        h = _vp()
        code = lib.inti_ctx_create(device, C.byref(h))
        if code != INTI_OK:
            raise IntiError(code, lib.inti_last_error(None).decode())
        self.handle = h
        self.lib = lib

    def close(self):
        if self.handle and not _exiting:
            self.lib.inti_ctx_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code):
        check(self.handle, code)

    def synchronize(self):
        self.check(self.lib.inti_ctx_synchronize(self.handle))

    def stream(self):
        s = _vp()
        self.check(self.lib.inti_ctx_stream(self.handle, C.byref(s)))
        return s.value

    def set_profiling(self, on):
        self.check(self.lib.inti_ctx_set_profiling(self.handle, 1 if on else 0))

    def last_kernel_ms(self):
        v = C.c_double()
        self.check(self.lib.inti_ctx_last_kernel_ms(self.handle, C.byref(v)))
        return v.value


_default_ctx = None


def default_context():
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
    return _default_ctx
