"""Assembled operators kept in HBM and applied to many right-hand sides on
the FP64 matrix cores.

`DenseOperator(A)` is `assemble_matrix(A)` (src/nystrom.jl:103-124) plus the
`axpy!(true, dA, Amat)` of every attached correction (src/api.jl:171-173),
done on the device; `apply` is the dense product with many densities at once
(a real GEMM, inti_dense_gemm).
"""
import ctypes as C

import numpy as np

from . import _lib


class DenseOperator:
    def __init__(self, iop, corrections=()):
        """iop: an IntegralOperator; corrections: (SparseCorrection, coef)
        pairs added to the matrix."""
        self.ctx = iop.ctx
        self.shape = iop.shape
        self.dtype = iop.dtype
        m, n = self.shape
        cw = 2 if self.dtype == np.complex128 else 1
        self._ptr = C.c_void_p()
        self.ctx.check(self.ctx.lib.inti_device_alloc(self.ctx.handle, 8 * cw * m * n, C.byref(self._ptr)))
        self.ctx.check(self.ctx.lib.inti_op_assemble_dense_device(iop.handle, self._ptr, m))
        for corr, coef in corrections:
            coef = complex(coef)
            self.ctx.check(self.ctx.lib.inti_dense_add_csr(self.ctx.handle, corr.handle, coef.real, coef.imag,
                                                           self._ptr, m))

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p is not None and p.value and not _lib._exiting:
            self.ctx.lib.inti_device_free(self.ctx.handle, p)
            self._ptr = None

    @property
    def device_pointer(self):
        return self._ptr

    def to_numpy(self):
        m, n = self.shape
        out = np.empty((n, m), dtype=self.dtype)  # every byte is copied over
        self.ctx.check(self.ctx.lib.inti_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self._ptr,
                                                    out.nbytes))
        return out.T

    def apply(self, X, Y=None, alpha=1.0, beta=0.0):
        """alpha A X + beta Y for X of shape (n,) or (n, nrhs)."""
        m, n = self.shape
        X = np.asarray(X)
        vec = X.ndim == 1
        Xc = np.ascontiguousarray(X.reshape(n, -1).T, dtype=self.dtype)  # column-major n x R
        R = Xc.shape[0]
        Yc = (np.zeros((R, m), dtype=self.dtype) if Y is None
              else np.ascontiguousarray(np.asarray(Y).reshape(m, -1).T, dtype=self.dtype))
        dx, dy = C.c_void_p(), C.c_void_p()
        lib, h = self.ctx.lib, self.ctx.handle
        self.ctx.check(lib.inti_device_alloc(h, max(Xc.nbytes, 8), C.byref(dx)))
        try:
            self.ctx.check(lib.inti_device_alloc(h, max(Yc.nbytes, 8), C.byref(dy)))
            try:
                self.ctx.check(lib.inti_memcpy_h2d(h, dx, Xc.ctypes.data_as(C.c_void_p), Xc.nbytes))
                self.ctx.check(lib.inti_memcpy_h2d(h, dy, Yc.ctypes.data_as(C.c_void_p), Yc.nbytes))
                a = np.array([complex(alpha).real, complex(alpha).imag])
                b = np.array([complex(beta).real, complex(beta).imag])
                self.ctx.check(lib.inti_dense_gemm(h, 1 if self.dtype == np.complex128 else 0, m, n, R, self._ptr,
                                                   m, dx, n, dy, m, _lib.dptr(a), _lib.dptr(b)))
                self.ctx.check(lib.inti_memcpy_d2h(h, Yc.ctypes.data_as(C.c_void_p), dy, Yc.nbytes))
            finally:
                lib.inti_device_free(h, dy)
        finally:
            lib.inti_device_free(h, dx)
        out = Yc.T
        return out.reshape(-1) if vec else np.ascontiguousarray(out)

    def __matmul__(self, X):
        return self.apply(X)
