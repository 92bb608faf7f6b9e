"""The allocation-free apply (inti_op_apply_device_noalloc) through the C-ABI
against inti_op_apply_device on the same operator and device operands: two
calls in a row, the second with a smaller input, each within 1e-13
relative (the ordinary apply may take the symmetric sum or another source
split; the FP64 bar of the direct sum is 1e-12), and bit for bit where both
run the same plan. A shape whose partial sums exceed the static buffer is
refused with INTI_ERR_SIZE before anything is launched."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INTI_ERR_SIZE = 5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(ctx, lib, host):
    p = C.c_void_p()
    ctx.check(lib.inti_device_alloc(ctx.handle, host.nbytes, C.byref(p)))
    ctx.check(lib.inti_memcpy_h2d(ctx.handle, p, host.ctypes.data_as(C.c_void_p), host.nbytes))
    return p


def _get(ctx, lib, p, like):
    out = np.zeros_like(like)
    ctx.check(lib.inti_memcpy_d2h(ctx.handle, out.ctypes.data_as(C.c_void_p), p, out.nbytes))
    return out


def _both(inti, op, X, alpha, beta, Y0):
    """(ordinary apply, allocation-free apply) of op on the columns of X."""
    ctx = inti.default_context()
    lib = ctx.lib
    n, r = X.shape
    m = op.shape[0]
    Xc = np.ascontiguousarray(X.T)   # column-major n x r
    Yc = np.ascontiguousarray(Y0.T)
    a = (C.c_double * 2)(alpha.real, alpha.imag)
    b = (C.c_double * 2)(beta.real, beta.imag)
    dx, dy1, dy2 = _dev(ctx, lib, Xc), _dev(ctx, lib, Yc), _dev(ctx, lib, Yc)
    try:
        ctx.check(lib.inti_op_apply_device(op.handle, r, dx, n, dy1, m, a, b))
        rc = lib.inti_op_apply_device_noalloc(op.handle, r, dx, n, dy2, m, a, b)
        ctx.check(rc)
        ctx.synchronize()
        return _get(ctx, lib, dy1, Yc).T, _get(ctx, lib, dy2, Yc).T
    finally:
        for p in (dx, dy1, dy2):
            lib.inti_device_free(ctx.handle, p)


def test_noalloc_apply_matches_apply_twice(inti):
    rng = np.random.default_rng(3)
    # call 1: corrected 3D Helmholtz combined field + identity, 17,280 nodes,
    # 17 columns (a block of 16 on the matrix cores, then one; the 16-column
    # block's source split is capped by the buffer)
    say("[gpu] noalloc call 1: 3D Helmholtz CFIE, 17,280 nodes, 17 columns")
    Q = inti.sphere_quadrature(12)
    k = 2 * np.pi
    L, _ = inti.combined_field(op=inti.Helmholtz(k=k, dim=3), quad=Q, c_single=-1j * k, c_double=1.0,
                               c_identity=0.5)
    n = len(Q)
    X = rng.normal(size=(n, 17)) + 1j * rng.normal(size=(n, 17))
    Y0 = rng.normal(size=(n, 17)) + 1j * rng.normal(size=(n, 17))
    y_ref, y_na = _both(inti, L.iop, X, 1.5 - 0.5j, 0.25 + 1j, Y0)
    e1 = np.linalg.norm(y_na - y_ref) / np.linalg.norm(y_ref)
    say(f"  rel diff {e1:.2e}")
    assert e1 <= 1e-13
    # call 2: smaller, corrected 2D Laplace double layer, 200 nodes, 1 column
    say("[gpu] noalloc call 2: 2D Laplace DL, 200 nodes, 1 column")
    Q2 = inti.circle_quadrature(40, qorder=8)
    op = inti.Laplace(dim=2)
    D = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q2)
    S = inti.IntegralOperator(inti.SingleLayerKernel(op), Q2)
    D.attach(inti.bdim_correction(op, Q2, Q2, S, D, green_multiplier=np.full(len(Q2), -0.5))[1])
    x2 = rng.normal(size=(len(Q2), 1))
    y_ref2, y_na2 = _both(inti, D, x2, 1.0 + 0j, 0j, np.zeros_like(x2))
    e2 = np.linalg.norm(y_na2 - y_ref2) / np.linalg.norm(y_ref2)
    say(f"  rel diff {e2:.2e}, identical bits: {np.array_equal(y_na2, y_ref2)}")
    assert np.array_equal(y_na2, y_ref2)  # same one-direction plan, no cap


def test_noalloc_apply_refuses_oversized_partials(inti):
    """m * 16 columns * 16 bytes beyond the 64 MiB buffer: 300,000 targets,
    16 complex columns. Refused with INTI_ERR_SIZE (no kernel launched)."""
    say("[gpu] noalloc oversized shape")
    ctx = inti.default_context()
    lib = ctx.lib
    m = 300_000
    trg = inti.points(np.stack([np.linspace(2, 3, m), np.zeros(m), np.zeros(m)]))
    src = inti.sphere_quadrature(2)
    A = inti.IntegralOperator(inti.SingleLayerKernel(inti.Helmholtz(k=1.0, dim=3)), trg, src)
    n = len(src)
    X = np.zeros((16, n), dtype=np.complex128)
    Y = np.zeros((16, m), dtype=np.complex128)
    dx, dy = _dev(ctx, lib, X), _dev(ctx, lib, Y)
    try:
        one = (C.c_double * 2)(1.0, 0.0)
        zero = (C.c_double * 2)(0.0, 0.0)
        rc = lib.inti_op_apply_device_noalloc(A.handle, 16, dx, n, dy, m, one, zero)
        assert rc == INTI_ERR_SIZE
        assert b"exceed" in lib.inti_last_error(ctx.handle)
        # one column of the same shape fits (300,000 x 16 bytes per split)
        rc = lib.inti_op_apply_device_noalloc(A.handle, 1, dx, n, dy, m, one, zero)
        assert rc == 0
        ctx.synchronize()
    finally:
        lib.inti_device_free(ctx.handle, dx)
        lib.inti_device_free(ctx.handle, dy)
