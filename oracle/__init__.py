"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
baseline leg of bench.py as the CPU reference. The product never imports it.

All values are returned as complex128 (Laplace values have zero imaginary
part). Point sets are passed as objects with .coords (dim, n), .normals
(dim, n) or None, .weights (n,), .qtags (nq, ne).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LIB_FMA = os.path.join(HERE, "liboracle_fma.so")
LIB_REV = os.path.join(HERE, "liboracle_rev.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_lib = None
_lib_fma = None
_lib_rev = None


def _load(variant=None):
    """The oracle library; variant="fma" gives the FMA-contracted build,
    variant="rev" the build whose Theta sums run over the sources in
    reverse order."""
    global _lib, _lib_fma, _lib_rev
    if variant == "fma":
        if _lib_fma is None:
            _lib_fma = _bind(LIB_FMA)
        return _lib_fma
    if variant == "rev":
        if _lib_rev is None:
            _lib_rev = _bind(LIB_REV)
        return _lib_rev
    if _lib is None:
        _lib = _bind(LIB)
    return _lib


def _bind(path):
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(path)
    lib.ork_kernel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.ork_hankel1.argtypes = [C.c_int, C.c_double, _dp]
    lib.ork_assemble.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int64, _dp, _dp,
                                 C.c_int64, _dp, _dp, _dp, _dp]
    lib.ork_matvec.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int64, _dp, _dp,
                               C.c_int64, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_int64,
                               C.c_int64, C.c_int]
    lib.ork_dim_monopoles.restype = C.c_int
    lib.ork_dim_monopoles.argtypes = [C.c_int, C.c_int64, _dp, C.c_int, C.c_double, C.c_double, _dp]
    lib.ork_bdim.restype = C.c_int64
    lib.ork_bdim.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int64, _dp, _dp,
                             C.c_int64, _dp, _dp, _dp, C.c_int, C.c_int64, _i64p, _i64p, _i64p,
                             _dp, C.c_double, C.c_double, _i64p, _i64p, _dp, _dp, _dp]
    lib.ork_bdim_types.restype = C.c_int64
    lib.ork_bdim_types.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int64, _dp, _dp,
                                   C.c_int64, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_int), _i64p, _i64p,
                                   _i64p, _i64p, _dp, C.c_double, C.c_double, _i64p, _i64p, _dp, _dp, _dp]
    lib.ork_nearest_lists.restype = C.c_int64
    lib.ork_nearest_lists.argtypes = [C.c_int, C.c_int64, _dp, C.c_int64, _dp, C.c_int, C.c_int64,
                                      _i64p, C.c_double, _i64p, _i64p]
    lib.ork_csr_matvec.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p, _dp, C.c_int, _dp, _dp,
                                   _dp, _dp]
    lib.ork_adjoint.restype = C.c_int
    lib.ork_adjoint.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, C.c_int64, _dp,
                                C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    return lib


def _d(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _i(a):
    if a is None:
        return None
    return a.ctypes.data_as(_i64p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _coefs(coefs):
    c = np.zeros(4) if coefs is None else np.asarray(
        [complex(coefs[0]).real, complex(coefs[0]).imag, complex(coefs[1]).real,
         complex(coefs[1]).imag], dtype=np.float64)
    return c


def hankel1(n, x):
    out = np.zeros(2)
    _load().ork_hankel1(n, float(x), _d(out))
    return complex(out[0], out[1])


def kernel(pde, dim, kind, k, x, nx, y, ny, coefs=None):
    out = np.zeros(2)
    c = _coefs(coefs)
    _load().ork_kernel(pde, dim, kind, k, _d(c), _d(_f64(x)), _d(_f64(nx)), _d(_f64(y)),
                       _d(_f64(ny)), _d(out))
    return complex(out[0], out[1])


def assemble(pde, dim, kind, k, trg, src, coefs=None):
    m, n = trg.coords.shape[1], src.coords.shape[1]
    out = np.zeros((n, m), dtype=np.complex128)
    c = _coefs(coefs)
    _load().ork_assemble(pde, dim, kind, k, _d(c), m, _d(_f64(trg.coords)), _d(_f64(trg.normals)),
                         n, _d(_f64(src.coords)), _d(_f64(src.normals)), _d(_f64(src.weights)),
                         _d(out.view(np.float64)))
    return out.T


def matvec(pde, dim, kind, k, trg, src, sigma, coefs=None, rows=None, nthreads=0):
    """Direct sum y = K sigma for rows [lo, hi) (others are zero)."""
    m, n = trg.coords.shape[1], src.coords.shape[1]
    s = np.asarray(sigma, dtype=np.complex128)
    vec = s.ndim == 1
    S = s.reshape(n, -1)
    nrhs = S.shape[1]
    X = np.ascontiguousarray(S.T)
    Y = np.zeros((nrhs, m), dtype=np.complex128)
    lo, hi = (0, m) if rows is None else rows
    c = _coefs(coefs)
    _load().ork_matvec(pde, dim, kind, k, _d(c), m, _d(_f64(trg.coords)), _d(_f64(trg.normals)),
                       n, _d(_f64(src.coords)), _d(_f64(src.normals)), _d(_f64(src.weights)),
                       nrhs, _d(X.view(np.float64)), _d(Y.view(np.float64)), lo, hi, nthreads)
    out = Y.T
    return out.reshape(-1) if vec else out


def monopoles(dim, src_coords, nq, oversample=3.0, radius_mult=1.5):
    ns = int(np.ceil(oversample * nq))
    zs = np.zeros((dim, ns))
    n = src_coords.shape[1]
    _load().ork_dim_monopoles(dim, n, _d(_f64(src_coords)), nq, oversample, radius_mult, _d(zs))
    return zs


def _etypes(src):
    return getattr(src, "etype_qtags", None) or [src.qtags]


def bdim(pde, dim, k, trg, src, mu, derivative=False, near=None, oversample=3.0,
         radius_mult=1.5, variant=None):
    """CSR (rowptr, colind, valS, valD) and diagnostics of bdim_correction;
    a source with several element types takes one near list per type."""
    m, n = trg.coords.shape[1], src.coords.shape[1]
    types = _etypes(src)
    nqs = np.array([q.shape[0] for q in types], dtype=np.int32)
    nes = np.array([q.shape[1] for q in types], dtype=np.int64)
    qt = np.ascontiguousarray(np.concatenate([q.T.reshape(-1) for q in types]), dtype=np.int64)
    if near is None:
        cnt_tri = int(np.sum(nqs.astype(np.int64) ** 2 * nes))
        near_ptr = near_idx = None
    else:
        lists = [near] if len(types) == 1 else list(near)
        if len(lists) != len(types):
            raise ValueError("one near list per element type")
        near_ptr = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.int64) for p, _ in lists]))
        near_idx = np.ascontiguousarray(np.concatenate([np.asarray(i, dtype=np.int64) for _, i in lists]))
        cnt_tri = sum(int(p[-1]) * int(nq) for (p, _), nq in zip(lists, nqs))
    nnz_max = max(cnt_tri, 1)
    rowptr = np.zeros(m + 1, dtype=np.int64)
    colind = np.zeros(nnz_max, dtype=np.int64)
    valS = np.zeros(nnz_max, dtype=np.complex128)
    valD = np.zeros(nnz_max, dtype=np.complex128)
    diag = np.zeros(4)
    nnz = _load(variant).ork_bdim_types(
        pde, dim, k, 1 if derivative else 0, m, _d(_f64(trg.coords)), _d(_f64(trg.normals)), n,
        _d(_f64(src.coords)), _d(_f64(src.normals)), _d(_f64(src.weights)), len(types),
        nqs.ctypes.data_as(C.POINTER(C.c_int)), _i(nes), _i(qt), _i(near_ptr), _i(near_idx),
        _d(_f64(mu)), oversample, radius_mult, _i(rowptr), _i(colind),
        _d(valS.view(np.float64)), _d(valD.view(np.float64)), _d(diag))
    if nnz < 0:
        raise ValueError("oracle bdim rejected its arguments")
    return rowptr, colind[:nnz], valS[:nnz], valD[:nnz], diag


def nearest_lists(trg, src, maxdist=np.inf):
    """Element -> target lists; one (near_ptr, near_idx) pair per element type
    when the source has several (quadrature.jl:264-286 runs per type)."""
    m, n = trg.coords.shape[1], src.coords.shape[1]
    out = []
    for qtags in _etypes(src):
        nq, ne = qtags.shape
        qt = np.ascontiguousarray(qtags.T.reshape(-1), dtype=np.int64)
        near_ptr = np.zeros(ne + 1, dtype=np.int64)
        near_idx = np.zeros(max(m, 1), dtype=np.int64)
        cnt = _load().ork_nearest_lists(trg.coords.shape[0], m, _d(_f64(trg.coords)), n,
                                        _d(_f64(src.coords)), nq, ne, _i(qt), float(maxdist),
                                        _i(near_ptr), _i(near_idx))
        out.append((near_ptr, near_idx[:cnt]))
    return out[0] if len(out) == 1 else out


def csr_matvec(m, n, rowptr, colind, val, x):
    x = np.asarray(x, dtype=np.complex128)
    vec = x.ndim == 1
    X = np.ascontiguousarray(x.reshape(n, -1).T)
    Y = np.zeros((X.shape[0], m), dtype=np.complex128)
    a = np.array([1.0, 0.0])
    b = np.array([0.0, 0.0])
    vv = np.ascontiguousarray(val, dtype=np.complex128)
    _load().ork_csr_matvec(m, n, _i(np.ascontiguousarray(rowptr, dtype=np.int64)),
                           _i(np.ascontiguousarray(colind, dtype=np.int64)), _d(vv.view(np.float64)),
                           X.shape[0], _d(X.view(np.float64)), _d(a), _d(b), _d(Y.view(np.float64)))
    out = Y.T
    return out.reshape(-1) if vec else out


def corrected_matvec(op_pde, dim, kind, k, quad, sigma, mu, which="D", coefs=None, nthreads=0):
    """(A + dA) sigma on one surface, A the SL (which='S') or DL ('D')
    operator, dA from the oracle's bdim_correction (self pattern)."""
    rowptr, colind, vS, vD, _ = bdim(op_pde, dim, k, quad, quad, mu)
    y = matvec(op_pde, dim, kind, k, quad, quad, sigma, coefs=coefs, nthreads=nthreads)
    val = vD if which == "D" else vS
    return y + csr_matvec(len(quad.weights), len(quad.weights), rowptr, colind, val, sigma)


def adjoint(pde, dim, kind, k, trg, src, ubar, sigma, coefs=None, dcoefs=None, want_xbar=True):
    """Pullback of y = K sigma: dict with sigma_bar (n,), ybar/nbar (dim, n),
    wbar (n,), xbar (dim, m) and kbar (see ork_adjoint). coefs = (c0, c1),
    dcoefs = (dc0/dk, dc1/dk), complex."""
    m, n = trg.coords.shape[1], src.coords.shape[1]
    u = np.ascontiguousarray(np.asarray(ubar, dtype=np.complex128))
    sg = np.ascontiguousarray(np.asarray(sigma, dtype=np.complex128))
    sb = np.zeros(n, dtype=np.complex128)
    yb = np.zeros((dim, n))
    nb = np.zeros((dim, n))
    wb = np.zeros(n)
    xb = np.zeros((dim, m)) if want_xbar else None
    kb = np.zeros(1)
    c = _coefs(coefs)
    dc = None if dcoefs is None else _coefs(dcoefs)
    rc = _load().ork_adjoint(pde, dim, kind, k, _d(c), _d(dc), m, _d(_f64(trg.coords)), n,
                             _d(_f64(src.coords)), _d(_f64(src.normals)), _d(_f64(src.weights)),
                             _d(u.view(np.float64)), _d(sg.view(np.float64)), _d(sb.view(np.float64)),
                             _d(yb), _d(nb), _d(wb), _d(xb), _d(kb))
    if rc != 0:
        raise ValueError("oracle adjoint: unsupported kind")
    return {"sigma_bar": sb, "ybar": yb, "nbar": nb, "wbar": wb, "xbar": xb, "kbar": kb[0]}
