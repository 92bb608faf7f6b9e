"""The call sequences of the Julia shim (inti.jl_amd/julia/IntiMI355X.jl),
replayed through the same C-ABI entry points with ctypes (no Julia
toolchain in the build image):

  * one device operator per (IntegralOperator, precision): an FP64 handle
    created after a mixed one on the same quadratures is its own operator,
    its pullback works, the mixed one's is refused;
  * upload of nodes without normals (volume quadratures, point targets) with
    a NULL normals pointer, used as targets and as single-layer sources;
  * single_double_layer with compression :gpu: DIM build, attach, destroy
    the build's CSR, apply (equal to the Python mirror's corrected operator);
  * the rrule of a corrected operator: sigma-bar only (NULL geometry
    outputs); geometry outputs are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64, MIXED = 0, 1
LAP, HEL = 0, 1
SL, DL = 0, 1
ERR_ARG, ERR_UNSUPPORTED = 1, 3


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _upload(ctx, coords, normals, weights):
    h = C.c_void_p()
    dim, n = coords.shape
    ctx.check(ctx.lib.inti_quad_upload(ctx.handle, dim, n, _dp(np.ascontiguousarray(coords)),
                                       _dp(None if normals is None else np.ascontiguousarray(normals)),
                                       _dp(np.ascontiguousarray(weights)), C.byref(h)))
    return h


def _op(ctx, pde, kind, dim, k, trg, src, precision):
    h = C.c_void_p()
    rc = ctx.lib.inti_op_create(ctx.handle, pde, kind, dim, k, None, trg, src, precision, C.byref(h))
    return rc, h


def _apply(ctx, h, x, m):
    y = np.zeros(m, dtype=x.dtype)
    one = (C.c_double * 2)(1.0, 0.0)
    zero = (C.c_double * 2)(0.0, 0.0)
    ctx.check(ctx.lib.inti_op_apply(h, 1, _dp(x.view(np.float64)), _dp(y.view(np.float64)), one, zero))
    return y


def test_handles_per_precision_and_adjoint(inti):
    ctx = inti.default_context()
    Q = inti.sphere_quadrature(4)
    n = len(Q)
    q = _upload(ctx, Q.coords, Q.normals, Q.weights)
    try:
        rc, hm = _op(ctx, HEL, DL, 3, 2.0, q, q, MIXED)
        ctx.check(rc)
        rc, hf = _op(ctx, HEL, DL, 3, 2.0, q, q, FP64)
        ctx.check(rc)
        assert hm.value != hf.value
        rng = np.random.default_rng(1)
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        ym, yf = _apply(ctx, hm, x, n), _apply(ctx, hf, x, n)
        ref = inti.IntegralOperator(inti.DoubleLayerKernel(inti.Helmholtz(k=2.0, dim=3)), Q).apply(x)
        assert np.linalg.norm(yf - ref) <= 1e-13 * np.linalg.norm(ref)
        assert 0 < np.linalg.norm(ym - yf) <= 1e-6 * np.linalg.norm(yf)
        sb = np.zeros(n, dtype=np.complex128)
        yb = np.zeros((3, n))
        kb = np.zeros(1)
        ctx.check(ctx.lib.inti_op_adjoint(hf, _dp(x.view(np.float64)), _dp(x.view(np.float64)),
                                          _dp(sb.view(np.float64)), _dp(yb), None, None, None, None, _dp(kb)))
        rc = ctx.lib.inti_op_adjoint(hm, _dp(x.view(np.float64)), _dp(x.view(np.float64)),
                                     _dp(sb.view(np.float64)), _dp(yb), None, None, None, None, _dp(kb))
        assert rc == ERR_UNSUPPORTED
        ctx.lib.inti_op_destroy(hm)
        ctx.lib.inti_op_destroy(hf)
    finally:
        ctx.lib.inti_quad_destroy(q)


def test_upload_without_normals(inti, oracle):
    """Volume-like nodes (no normals) as targets and as single-layer sources;
    a double layer on them is refused with an argument error."""
    ctx = inti.default_context()
    rng = np.random.default_rng(2)
    V = rng.uniform(-0.5, 0.5, size=(3, 300))
    w = rng.uniform(0.1, 1.0, 300)
    Q = inti.sphere_quadrature(3)
    qv = _upload(ctx, V, None, w)
    qs = _upload(ctx, Q.coords, Q.normals, Q.weights)
    try:
        rc, h = _op(ctx, LAP, DL, 3, 0.0, qv, qs, FP64)  # targets without normals
        ctx.check(rc)
        x = rng.normal(size=len(Q))
        y = _apply(ctx, h, x, 300)
        ref = oracle.matvec(0, 3, 1, 0.0, inti.points(V), Q, x).real
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
        ctx.lib.inti_op_destroy(h)
        rc, h = _op(ctx, LAP, SL, 3, 0.0, qs, qv, FP64)  # sources without normals
        ctx.check(rc)
        s = rng.normal(size=300)
        src = inti.Quadrature(V, None, w, np.zeros((1, 0), dtype=np.int64))
        y = _apply(ctx, h, s, len(Q))
        ref = oracle.matvec(0, 3, 0, 0.0, Q, src, s).real
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
        ctx.lib.inti_op_destroy(h)
        rc, h = _op(ctx, LAP, DL, 3, 0.0, qs, qv, FP64)
        if rc == 0:  # refused at creation or at the first apply
            y = np.zeros(len(Q))
            one = (C.c_double * 2)(1.0, 0.0)
            rc = ctx.lib.inti_op_apply(h, 1, _dp(s), _dp(y), one, one)
            ctx.lib.inti_op_destroy(h)
        assert rc == ERR_ARG
    finally:
        ctx.lib.inti_quad_destroy(qv)
        ctx.lib.inti_quad_destroy(qs)


def test_single_double_layer_and_rrule_sequence(inti):
    ctx = inti.default_context()
    Q = inti.circle_quadrature(50, qorder=8)
    n = len(Q)
    k = 3.0
    q = _upload(ctx, Q.coords, Q.normals, Q.weights)
    try:
        ctx.check((rc := _op(ctx, HEL, SL, 2, k, q, q, FP64))[0])
        hS = rc[1]
        ctx.check((rc := _op(ctx, HEL, DL, 2, k, q, q, FP64))[0])
        hD = rc[1]
        qt = np.ascontiguousarray(Q.qtags.T.reshape(-1), dtype=np.int64)
        mu = np.full(n, -0.5)
        dS, dD = C.c_void_p(), C.c_void_p()
        ctx.check(ctx.lib.inti_bdim_build(hS, hD, 0, Q.nq, Q.ne, qt.ctypes.data_as(C.POINTER(C.c_int64)),
                                          None, None, _dp(mu), None, C.byref(dS), C.byref(dD), None))
        ctx.check(ctx.lib.inti_op_attach_correction(hS, dS, 1.0, 0.0))
        ctx.check(ctx.lib.inti_op_attach_correction(hD, dD, 1.0, 0.0))
        ctx.check(ctx.lib.inti_csr_destroy(dS))
        ctx.check(ctx.lib.inti_csr_destroy(dD))
        rng = np.random.default_rng(3)
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        yS, yD = _apply(ctx, hS, x, n), _apply(ctx, hD, x, n)
        S, D = inti.single_double_layer(op=inti.Helmholtz(k=k, dim=2), target=Q, source=Q,
                                        compression={"method": "gpu"})
        assert np.linalg.norm(yS - S @ x) <= 1e-13 * np.linalg.norm(yS)
        assert np.linalg.norm(yD - D @ x) <= 1e-13 * np.linalg.norm(yD)
        # rrule of the corrected operator: sigma-bar = D^H ubar only
        sb = np.zeros(n, dtype=np.complex128)
        ctx.check(ctx.lib.inti_op_adjoint(hD, _dp(x.view(np.float64)), _dp(x.view(np.float64)),
                                          _dp(sb.view(np.float64)), None, None, None, None, None, None))
        Dm = inti.DenseOperator(inti.IntegralOperator(inti.DoubleLayerKernel(inti.Helmholtz(k=k, dim=2)), Q),
                                corrections=[(inti.bdim_correction(
                                    inti.Helmholtz(k=k, dim=2), Q, Q,
                                    inti.IntegralOperator(inti.SingleLayerKernel(inti.Helmholtz(k=k, dim=2)), Q),
                                    inti.IntegralOperator(inti.DoubleLayerKernel(inti.Helmholtz(k=k, dim=2)), Q),
                                    green_multiplier=mu)[1], 1.0)]).to_numpy()
        ref = Dm.conj().T @ x
        assert np.linalg.norm(sb - ref) <= 1e-12 * np.linalg.norm(ref)
        yb = np.zeros((2, n))
        rc = ctx.lib.inti_op_adjoint(hD, _dp(x.view(np.float64)), _dp(x.view(np.float64)),
                                     _dp(sb.view(np.float64)), _dp(yb), None, None, None, None, None)
        assert rc == ERR_UNSUPPORTED
        ctx.lib.inti_op_destroy(hS)
        ctx.lib.inti_op_destroy(hD)
    finally:
        ctx.lib.inti_quad_destroy(q)
