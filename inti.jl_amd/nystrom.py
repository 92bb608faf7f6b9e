"""IntegralOperator / assemble_matrix on the GPU (Inti.jl src/nystrom.jl).

An IntegralOperator is matrix-free: entries K(x_i, y_j) * w_j
(nystrom.jl:93-96) are recomputed by the direct-sum kernel at every apply.
A sparse correction (from bdim_correction) can be attached, which fuses
y += coef * dA x into the same apply.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from .kernels import CombinedFieldKernel, DoubleLayerKernel, Helmholtz, SingleLayerKernel


def _green_multiplier(s):
    """-1 inside, 0 outside, -1/2 on (nystrom.jl:196-207)."""
    if s == "inside":
        return -1.0
    if s == "outside":
        return 0.0
    if s == "on":
        return -0.5
    raise ValueError(f"Unknown target location {s}. Expected :inside, :outside, or :on")


class SparseCorrection:
    """A CSR correction held on the device (rows = targets)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        m, n, nnz, cp = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        ctx.check(ctx.lib.inti_csr_info(handle, C.byref(m), C.byref(n), C.byref(nnz), C.byref(cp)))
        self.shape = (m.value, n.value)
        self.nnz = nnz.value
        self.is_complex = bool(cp.value)

    def __del__(self):
        try:
            if self.handle and not _lib.exiting():
                self.ctx.lib.inti_csr_destroy(self.handle)
        except Exception:
            pass

    def export(self):
        """(rowptr, colind, values) as numpy arrays, 0-based CSR."""
        m, n = self.shape
        rowptr = np.zeros(m + 1, dtype=np.int64)
        colind = np.zeros(self.nnz, dtype=np.int64)
        val = np.zeros(self.nnz, dtype=np.complex128 if self.is_complex else np.float64)
        self.ctx.check(self.ctx.lib.inti_csr_export(
            self.handle, _lib.i64ptr(rowptr), _lib.i64ptr(colind), _lib.dptr(val)))
        return rowptr, colind, val

    def to_scipy(self):
        rowptr, colind, val = self.export()
        return sp.csr_matrix((val, colind, rowptr), shape=self.shape)


class IntegralPotential:
    """IntegralPotential(kernel, quadrature) (nystrom.jl:20-38): pot[sigma] is
    a function of evaluation points x, (dim,) or (dim, m), returning
    sum_j K(x, y_j) sigma_j w_j; the sum runs as a GPU direct sum over all
    points of one call. No singular correction: x must lie off the surface."""

    def __init__(self, kernel, quadrature, precision="fp64", ctx=None):
        self.kernel = kernel
        self.quadrature = quadrature
        self.precision = precision
        self.ctx = ctx

    def __getitem__(self, sigma):
        sigma = np.asarray(sigma)
        if sigma.shape[0] != len(self.quadrature):
            raise ValueError("the density needs one value per quadrature node")

        def evaluate(x):
            from .quadrature import points
            x = np.asarray(x, dtype=np.float64)
            single = x.ndim == 1
            X = x.reshape(self.quadrature.dim, -1)
            op = IntegralOperator(self.kernel, points(X), self.quadrature, precision=self.precision,
                                  ctx=self.ctx)
            out = op.apply(sigma)
            return out[0] if single and out.ndim == 1 else out
        return evaluate


def SingleLayerPotential(op, source):
    """nystrom.jl:40-47."""
    return IntegralPotential(SingleLayerKernel(op), source)


def DoubleLayerPotential(op, source):
    """nystrom.jl:49-56."""
    return IntegralPotential(DoubleLayerKernel(op), source)


_SYMMETRIC_DEFAULT = None


def set_default_symmetric(mode):
    """Symmetric-evaluation mode given to every IntegralOperator created
    afterwards (see IntegralOperator.set_symmetric); None keeps the
    library's default (1)."""
    global _SYMMETRIC_DEFAULT
    _SYMMETRIC_DEFAULT = None if mode is None else int(mode)


class IntegralOperator:
    """IntegralOperator(kernel, target, source=target) (nystrom.jl:67-89).

    precision: "fp64" (default) or "mixed" (FP32 transcendental unit, FP64
    geometry and accumulation)."""

    def __init__(self, kernel, target, source=None, precision="fp64", ctx=None):
        source = target if source is None else source
        if target.dim != source.dim:
            raise ValueError("target and source must have the same ambient dimension")
        if kernel.op.dim != source.dim:
            raise ValueError("kernel dimension differs from the quadrature's")
        self.kernel = kernel
        self.target = target
        self.source = source
        self.precision = precision
        self.ctx = ctx or _lib.default_context()
        self._h = None
        self._corr = []  # keep attached corrections alive
        self.identity = 0.0

    @property
    def shape(self):
        return (len(self.target), len(self.source))

    @property
    def dtype(self):
        return np.complex128 if self.kernel.is_complex else np.float64

    @property
    def handle(self):
        if self._h is None:
            k = self.kernel
            coefs = None
            if isinstance(k, CombinedFieldKernel):
                coefs = np.array([k.c0.real, k.c0.imag, k.c1.real, k.c1.imag])
            h = C.c_void_p()
            prec = _lib.MIXED if self.precision == "mixed" else _lib.FP64
            kr, ki = (k.op.kr, k.op.ki) if hasattr(k.op, "kr") else (k.op.k, 0.0)
            self.ctx.check(self.ctx.lib.inti_op_create(
                self.ctx.handle, k.op.pde, k.kind, k.op.dim, kr,
                _lib.dptr(coefs), self.target.device(self.ctx), self.source.device(self.ctx),
                prec, C.byref(h)))
            if ki != 0:
                st = self.ctx.lib.inti_op_set_wavenumber_imag(h, ki)
                if st != 0:
                    self.ctx.lib.inti_op_destroy(h)
                    self.ctx.check(st)
            self._h = h
            if _SYMMETRIC_DEFAULT is not None:
                self.ctx.check(self.ctx.lib.inti_op_set_symmetric(h, _SYMMETRIC_DEFAULT, 0))
        return self._h

    def __del__(self):
        try:
            if self._h is not None and not _lib.exiting():
                self.ctx.lib.inti_op_destroy(self._h)
        except Exception:
            pass

    def set_identity(self, c):
        c = complex(c)
        self.ctx.check(self.ctx.lib.inti_op_set_identity(self.handle, c.real, c.imag))
        self.identity = c

    def set_symmetric(self, mode, group=0):
        """Symmetric evaluation when target is source (single right-hand
        side): mode 0 off, 1 automatic (default: where the work fills the
        GPU), 2 wherever it applies; group = tiles of 256 nodes per work group
        (0 automatic)."""
        self.ctx.check(self.ctx.lib.inti_op_set_symmetric(self.handle, int(mode), int(group)))

    def attach(self, corr, coef=1.0):
        """Fuse coef * corr (a SparseCorrection) into every apply."""
        coef = complex(coef)
        self.ctx.check(self.ctx.lib.inti_op_attach_correction(self.handle, corr.handle,
                                                              coef.real, coef.imag))
        self._corr.append(corr)

    def detach(self):
        self.ctx.check(self.ctx.lib.inti_op_detach_correction(self.handle))
        self._corr = []

    def apply(self, x, y=None, alpha=1.0, beta=0.0):
        """y = alpha * A x + beta * y; x is (n,) or (n, nrhs)."""
        x = np.asarray(x)
        vec = x.ndim == 1
        X = x.reshape(len(self.source), -1)
        nrhs = X.shape[1]
        Xc = np.ascontiguousarray(X.T, dtype=self.dtype)  # column-major n x nrhs
        if y is None:
            Yc = np.zeros((nrhs, len(self.target)), dtype=self.dtype)
        else:
            Yc = np.ascontiguousarray(np.asarray(y).reshape(len(self.target), -1).T, dtype=self.dtype)
        a = np.array([complex(alpha).real, complex(alpha).imag])
        b = np.array([complex(beta).real, complex(beta).imag])
        self.ctx.check(self.ctx.lib.inti_op_apply(self.handle, nrhs, _lib.dptr(Xc), _lib.dptr(Yc),
                                                  _lib.dptr(a), _lib.dptr(b)))
        out = Yc.T
        return out.reshape(-1) if vec else np.ascontiguousarray(out)

    def __matmul__(self, x):
        return self.apply(x)

    def pullback(self, ubar, sigma, *, wrt=("sigma",), dcoefs=None):
        """Reverse mode of y = A sigma on the GPU (kernel values recomputed):
        the gradients of Re<ubar, y> requested in `wrt` among "sigma"
        (A^H ubar), "source_nodes", "source_normals", "source_weights",
        "target_nodes", "nodes" (operators whose target is their source: the
        gradient of the nodes, the sum of the source and target parts, in one
        output) and "k". dcoefs = (dc0/dk, dc1/dk) for combined kernels with
        k-dependent coefficients. Returns a dict."""
        m, n = self.shape
        dim = self.source.dim
        u = np.ascontiguousarray(ubar, dtype=self.dtype).reshape(m)
        s = np.ascontiguousarray(sigma, dtype=self.dtype).reshape(n)
        out = {}
        bufs = {
            "sigma": np.zeros(n, dtype=self.dtype),
            "source_nodes": np.zeros((dim, n)),
            "source_normals": np.zeros((dim, n)),
            "source_weights": np.zeros(n),
            "target_nodes": np.zeros((dim, m)),
            "nodes": np.zeros((dim, n)),
            "k": np.zeros(1),
        }
        for w in wrt:
            if w not in bufs:
                raise ValueError(f"unknown gradient {w!r}")
        if "nodes" in wrt:
            if self.target is not self.source:
                raise ValueError("'nodes' needs an operator whose target is its source")
            if "source_nodes" in wrt or "target_nodes" in wrt:
                raise ValueError("'nodes' is the sum of 'source_nodes' and 'target_nodes': ask for one or the parts")
        ptr = {w: (_lib.dptr(bufs[w]) if w in wrt else None) for w in bufs}
        if "nodes" in wrt:  # one output for both parts: the library sums them
            ptr["source_nodes"] = ptr["target_nodes"] = ptr["nodes"]
        dc = None
        if dcoefs is not None:
            d0, d1 = complex(dcoefs[0]), complex(dcoefs[1])
            dc = np.array([d0.real, d0.imag, d1.real, d1.imag])
        self.ctx.check(self.ctx.lib.inti_op_adjoint(
            self.handle, _lib.dptr(u), _lib.dptr(s), ptr["sigma"], ptr["source_nodes"],
            ptr["source_normals"], ptr["source_weights"], ptr["target_nodes"], _lib.dptr(dc), ptr["k"]))
        for w in wrt:
            out[w] = bufs[w][0] if w == "k" else bufs[w]
        return out


def rrule(A, sigma):
    """ChainRulesCore-style rrule of y = A * sigma: returns (y, pullback),
    pullback(ybar, wrt=...) -> dict of gradients (IntegralOperator.pullback)."""
    iop = A.iop if hasattr(A, "iop") else A
    y = A.apply(sigma)

    def pullback(ybar, wrt=("sigma",), dcoefs=None):
        return iop.pullback(ybar, sigma, wrt=wrt, dcoefs=dcoefs)
    return y, pullback


def assemble_matrix(iop):
    """Dense matrix of an IntegralOperator (nystrom.jl:103-124); the entries
    are computed on the GPU (FP64)."""
    m, n = iop.shape
    out = np.zeros((n, m), dtype=iop.dtype)  # column-major m x n
    iop.ctx.check(iop.ctx.lib.inti_op_assemble_dense(iop.handle, _lib.dptr(out)))
    return out.T


def is_helmholtz(op):
    return isinstance(op, Helmholtz)


def _iop(A):
    return A.iop if hasattr(A, "iop") else A


def gmres(A, b, x0=None, *, restart=100, maxiter=100, rtol=1e-6, atol=0.0, use_graph=True, precond=None):
    """Restarted GMRES on the GPU for the square operator A (IntegralOperator
    or corrected operator): the `gmres(L, rhs; restart, maxiter)` of
    IterativeSolvers used on Inti's operators
    (docs/src/examples/helmholtz_scattering.jl:494-495). Stops when
    ||b - A x|| <= max(rtol ||b||, atol). precond: a right preconditioner P
    (an operator of the same size and value type): GMRES on A P y = b,
    x = P y. Returns (x, info)."""
    A = _iop(A)
    P = None if precond is None else _iop(precond)
    b = np.ascontiguousarray(b, dtype=A.dtype)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=A.dtype).copy()
    it = C.c_int()
    rr = C.c_double()
    A.ctx.check(A.ctx.lib.inti_gmres_prec(A.handle, None if P is None else P.handle, 0,
                                          b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                                          restart, maxiter, rtol, atol, 1 if use_graph else 0, C.byref(it),
                                          C.byref(rr)))
    return x, {"iterations": it.value, "relres": rr.value}


def gmres_adjoint(A, g, x0=None, *, restart=100, maxiter=300, rtol=1e-10):
    """Solves A^H lam = g, the adjoint solve of an objective that depends on
    x = A^{-1} b (dJ/db = lam, dJ/dA = -lam x^H). A^H is applied on the GPU
    as A's pullback with sigma-bar alone (the transposed direct sum at the
    forward's cost, the correction's adjoint and the conjugated identity
    included); the GMRES iteration runs on the host (scipy). Returns
    (lam, info)."""
    import scipy.sparse.linalg as sla
    A = _iop(A)
    m, n = A.shape
    if m != n:
        raise ValueError("gmres_adjoint needs a square operator")
    dummy = np.zeros(n, dtype=A.dtype)
    count = [0]

    def mv(u):
        count[0] += 1
        return A.pullback(np.asarray(u, dtype=A.dtype), dummy, wrt=("sigma",))["sigma"]

    op = sla.LinearOperator((n, n), matvec=mv, dtype=A.dtype)
    g = np.ascontiguousarray(g, dtype=A.dtype)
    lam, flag = sla.gmres(op, g, x0=x0, rtol=rtol, atol=0.0, restart=restart, maxiter=max(1, maxiter // restart))
    res = float(np.linalg.norm(g - mv(lam)) / np.linalg.norm(g))
    count[0] -= 1
    return lam, {"iterations": count[0], "relres": res, "converged": flag == 0}


def refine_solve(A_hi, A_lo, b, x0=None, *, restart=100, maxiter_inner=100, inner_rtol=1e-6,
                 max_outer=10, rtol=1e-10, use_graph=True, precond=None):
    """Iterative refinement: FP64 residuals with A_hi, GMRES corrections with
    A_lo (e.g. the MIXED operator), right-preconditioned by precond if given.
    Returns (x, info)."""
    A_hi, A_lo = _iop(A_hi), _iop(A_lo)
    P = None if precond is None else _iop(precond)
    b = np.ascontiguousarray(b, dtype=A_hi.dtype)
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=A_hi.dtype).copy()
    o, tot = C.c_int(), C.c_int()
    rr = C.c_double()
    A_hi.ctx.check(A_hi.ctx.lib.inti_refine_solve_prec(
        A_hi.handle, A_lo.handle, None if P is None else P.handle, 0, b.ctypes.data_as(C.c_void_p),
        x.ctypes.data_as(C.c_void_p), restart, maxiter_inner, inner_rtol, max_outer, rtol,
        1 if use_graph else 0, C.byref(o), C.byref(tot), C.byref(rr)))
    return x, {"outer": o.value, "inner_iterations": tot.value, "relres": rr.value}
