"""Batches of independent problems on the GPU: one block-diagonal operator
over P quadratures (each with its own wavenumber), applied in one persistent
launch, with a batched restarted GMRES over every (problem, right-hand side)
system whose Arnoldi step is captured in a hipGraph.

The reference has no batch; this replaces a loop over problems of
IntegralOperator(K, Q_p, Q_p) (src/nystrom.jl:79-89) and
gmres(L_p, b_p) (docs/src/examples/helmholtz_scattering.jl:494-495).

Block vectors are (ntot, nrhs) arrays: problem p owns rows
offsets[p]:offsets[p+1].
"""
import ctypes as C

import numpy as np

from . import _lib
from .kernels import CombinedFieldKernel, Helmholtz


class BatchOperator:
    """BatchOperator(kernels, quads, precision="fp64"): kernels[p] over
    quads[p] (target = source). All kernels share their kind and PDE
    dimension; each carries its own wavenumber (and combined-field
    coefficients)."""

    def __init__(self, kernels, quads, precision="fp64", ctx=None):
        if len(kernels) != len(quads) or not kernels:
            raise ValueError("one kernel per quadrature, at least one problem")
        k0 = kernels[0]
        if any(getattr(K.op, "ki", 0.0) != 0 for K in kernels):
            raise NotImplementedError("batches take real wavenumbers")
        kind, pde, dim = k0.kind, k0.op.pde, k0.op.dim
        for K in kernels:
            if K.kind != kind or K.op.pde != pde or K.op.dim != dim:
                raise ValueError("all kernels of a batch share kind, PDE and dimension")
        self.kernels = list(kernels)
        self.quads = list(quads)
        self.ctx = ctx or _lib.default_context()
        self.dtype = np.complex128 if isinstance(k0.op, Helmholtz) else np.float64
        self.nprob = len(quads)
        handles = (C.c_void_p * self.nprob)(*[q.device(self.ctx) for q in quads])
        ks = np.array([K.op.k for K in kernels], dtype=np.float64)
        coefs = None
        if isinstance(k0, CombinedFieldKernel):
            coefs = np.array([[K.c0.real, K.c0.imag, K.c1.real, K.c1.imag] for K in kernels]).reshape(-1)
        prec = _lib.MIXED if precision == "mixed" else _lib.FP64
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.inti_batch_create(
            self.ctx.handle, pde, kind, dim, self.nprob, C.cast(handles, C.POINTER(C.c_void_p)),
            _lib.dptr(ks), _lib.dptr(coefs), prec, C.byref(h)))
        self.handle = h
        self._corr = []
        nt = C.c_int64()
        off = np.zeros(self.nprob + 1, dtype=np.int64)
        self.ctx.check(self.ctx.lib.inti_batch_info(h, C.byref(nt), off.ctypes.data_as(C.POINTER(C.c_int64))))
        self.ntot = nt.value
        self.offsets = off

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and h.value and not _lib._exiting:
            self.ctx.lib.inti_batch_destroy(h)
            self.handle = None

    def split(self, X):
        """Per-problem row blocks of a block vector."""
        return [X[self.offsets[p]:self.offsets[p + 1]] for p in range(self.nprob)]

    def join(self, blocks):
        return np.concatenate([np.asarray(b) for b in blocks], axis=0)

    def set_identity(self, c):
        """c * I added to every problem (scalar) or c[p] to problem p."""
        c = np.broadcast_to(np.asarray(c, dtype=np.complex128), (self.nprob,))
        cid = np.ascontiguousarray(np.stack([c.real, c.imag], axis=1).reshape(-1))
        self.ctx.check(self.ctx.lib.inti_batch_set_identity(self.handle, _lib.dptr(cid)))

    def bdim_correction(self, *, green_multiplier=None, parameters=None):
        """The DIM corrections of every problem's quadrature with itself
        (bdim_correction, bdim.jl:50-196, per problem: self pattern, one
        element type shared by the problems) in one pass over the batch:
        traces in one launch, Theta in one launch over (problem, target
        tile) work items, the least squares of all elements in one launch. green_multiplier: one value
        for all nodes (default -1/2) or one per node (ntot). Returns [(dS,
        dD)] per problem, SparseCorrection objects."""
        from .bdim import DimParameters
        from .nystrom import SparseCorrection
        nq = {q.qtags.shape[0] for q in self.quads}
        if len(nq) != 1 or any(len(q.etype_qtags) != 1 for q in self.quads):
            raise NotImplementedError("bdim_correction of a batch: one element type, the same nq in every problem")
        nq = nq.pop()
        ne = np.ascontiguousarray([q.qtags.shape[1] for q in self.quads], dtype=np.int64)
        tags = np.ascontiguousarray(np.concatenate([q.qtags.T.reshape(-1) for q in self.quads]), dtype=np.int64)
        mu = np.full(self.ntot, -0.5) if green_multiplier is None else np.broadcast_to(
            np.asarray(green_multiplier, dtype=np.float64), (self.ntot,))
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        params = parameters or DimParameters()
        pv = np.array([params.sources_oversample_factor, params.sources_radius_multiplier])
        hS = (C.c_void_p * self.nprob)()
        hD = (C.c_void_p * self.nprob)()
        self.ctx.check(self.ctx.lib.inti_batch_bdim_build(self.handle, nq, _lib.i64ptr(ne), _lib.i64ptr(tags),
                                                          _lib.dptr(mu), _lib.dptr(pv), hS, hD))
        return [(SparseCorrection(self.ctx, C.c_void_p(hS[p])), SparseCorrection(self.ctx, C.c_void_p(hD[p])))
                for p in range(self.nprob)]

    def attach(self, p, corr, coef=1.0):
        """Add coef * corr (a SparseCorrection of problem p) to problem p."""
        coef = complex(coef)
        self.ctx.check(self.ctx.lib.inti_batch_attach_correction(self.handle, p, corr.handle, coef.real,
                                                                 coef.imag))
        self._corr.append(corr)

    def set_launch(self, chunk_len=0, grid=0):
        self.ctx.check(self.ctx.lib.inti_batch_set_launch(self.handle, int(chunk_len), int(grid)))

    def apply(self, X, Y=None, alpha=1.0, beta=0.0):
        """alpha A X + beta Y for a block vector X (ntot,) or (ntot, nrhs)."""
        X = np.asarray(X)
        vec = X.ndim == 1
        Xc = np.ascontiguousarray(X.reshape(self.ntot, -1).T, dtype=self.dtype)
        nrhs = Xc.shape[0]
        Yc = (np.zeros_like(Xc) if Y is None
              else np.ascontiguousarray(np.asarray(Y).reshape(self.ntot, -1).T, dtype=self.dtype))
        a = np.array([complex(alpha).real, complex(alpha).imag])
        b = np.array([complex(beta).real, complex(beta).imag])
        self.ctx.check(self.ctx.lib.inti_batch_apply(self.handle, nrhs, _lib.dptr(Xc), _lib.dptr(Yc),
                                                     _lib.dptr(a), _lib.dptr(b)))
        out = Yc.T
        return out.reshape(-1) if vec else np.ascontiguousarray(out)

    def __matmul__(self, X):
        return self.apply(X)

    def autotune(self, nrhs=16, reps=3):
        """Pick the work-item chunk length and persistent grid by timing."""
        c, g, ms = C.c_int64(), C.c_int(), C.c_double()
        self.ctx.check(self.ctx.lib.inti_batch_autotune(self.handle, nrhs, reps, C.byref(c), C.byref(g),
                                                        C.byref(ms)))
        return {"chunk_len": c.value, "grid": g.value, "ms": ms.value}

    def _system_norms(self, X):
        """2-norm of every system's block (problem p, column r): (nprob, nrhs)."""
        sq = np.abs(X) ** 2
        cs = np.concatenate([np.zeros((1, sq.shape[1])), np.cumsum(sq, axis=0)])
        return np.sqrt(np.maximum(cs[self.offsets[1:]] - cs[self.offsets[:-1]], 0.0))

    def refine_solve(self, lo, B, X0=None, *, restart=100, maxiter_inner=100, inner_rtol=1e-6,
                     max_outer=10, rtol=1e-10, use_graph=True):
        """Iterative refinement of every system of A X = B: FP64 residuals
        with this operator, corrections by the batched GMRES of `lo` (the
        same problems, e.g. in mixed precision). Systems whose residual is
        at rtol get a zero right-hand side (no further corrections).
        Returns (X, info) with info["relres"] of shape (nprob, nrhs)."""
        if lo.nprob != self.nprob or lo.ntot != self.ntot:
            raise ValueError("refine_solve: the two batches must hold the same problems")
        B = np.asarray(B, dtype=self.dtype)
        vec = B.ndim == 1
        B2 = B.reshape(self.ntot, -1)
        X = np.zeros_like(B2) if X0 is None else np.array(np.asarray(X0, dtype=self.dtype).reshape(self.ntot, -1))
        bn = self._system_norms(B2)
        bn[bn == 0] = 1.0
        inner, outer = 0, 0
        while True:
            R = B2 - self.apply(X).reshape(self.ntot, -1)
            rel = self._system_norms(R) / bn
            if outer >= max_outer or np.all(rel <= rtol):
                break
            done = rel <= rtol
            for p in range(self.nprob):  # converged systems: zero right-hand side
                if done[p].any():
                    R[self.offsets[p]:self.offsets[p + 1], done[p]] = 0
            D, info = lo.gmres(R, restart=restart, maxiter=maxiter_inner, rtol=inner_rtol, use_graph=use_graph)
            X += D.reshape(self.ntot, -1)
            inner += info["iterations"]
            outer += 1
        X = X.reshape(-1) if vec else X
        return X, {"outer": outer, "inner_iterations": inner, "relres": rel}

    def gmres(self, B, X0=None, *, restart=30, maxiter=100, rtol=1e-6, atol=0.0, use_graph=True):
        """Solve every system (p, r) of A X = B; returns (X, info) with
        info["relres"] of shape (nprob, nrhs)."""
        B = np.asarray(B)
        vec = B.ndim == 1
        Bc = np.ascontiguousarray(B.reshape(self.ntot, -1).T, dtype=self.dtype)
        nrhs = Bc.shape[0]
        Xc = (np.zeros_like(Bc) if X0 is None
              else np.ascontiguousarray(np.asarray(X0).reshape(self.ntot, -1).T, dtype=self.dtype))
        it = C.c_int()
        rel = np.zeros(self.nprob * nrhs)
        self.ctx.check(self.ctx.lib.inti_batch_gmres(
            self.handle, 0, nrhs, _lib.dptr(Bc), _lib.dptr(Xc), restart, maxiter, rtol, atol,
            1 if use_graph else 0, C.byref(it), _lib.dptr(rel)))
        X = Xc.T
        X = X.reshape(-1) if vec else np.ascontiguousarray(X)
        return X, {"iterations": it.value, "relres": rel.reshape(self.nprob, nrhs)}
