"""Dense multi-right-hand-side products on the FP64 matrix cores
(inti_dense_gemm) against numpy, and the assembled corrected operator
against the matrix-free one."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gemm(inti, A, X, Y, alpha, beta):
    ctx = inti.default_context()
    lib, h = ctx.lib, ctx.handle
    cplx = np.iscomplexobj(A)
    m, n = A.shape
    R = X.shape[1]
    dt = np.complex128 if cplx else np.float64
    Ac = np.asfortranarray(A.astype(dt))
    Xc = np.asfortranarray(X.astype(dt))
    Yc = np.asfortranarray(Y.astype(dt))
    ptr = {}
    for name, arr in (("A", Ac), ("X", Xc), ("Y", Yc)):
        p = C.c_void_p()
        ctx.check(lib.inti_device_alloc(h, max(arr.nbytes, 8), C.byref(p)))
        ctx.check(lib.inti_memcpy_h2d(h, p, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
        ptr[name] = p
    a = np.array([complex(alpha).real, complex(alpha).imag])
    b = np.array([complex(beta).real, complex(beta).imag])
    ctx.check(lib.inti_dense_gemm(h, 1 if cplx else 0, m, n, R, ptr["A"], m, ptr["X"], n, ptr["Y"], m,
                                  a.ctypes.data_as(C.POINTER(C.c_double)), b.ctypes.data_as(C.POINTER(C.c_double))))
    out = np.zeros_like(Yc)
    ctx.check(lib.inti_memcpy_d2h(h, out.ctypes.data_as(C.c_void_p), ptr["Y"], out.nbytes))
    for p in ptr.values():
        ctx.check(lib.inti_device_free(h, p))
    return out


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n,R", [(37, 53, 1), (200, 130, 5), (300, 257, 16), (129, 4, 24), (16, 300, 3),
                                   (90, 70, 40), (65, 333, 100)])
def test_dense_gemm_matches_numpy(inti, cplx, m, n, R):
    """Every (m, n, R) shape, ragged against the 128 x 32 tiles and the
    k-chunks, real and complex, with alpha and beta: 1e-14 relative."""
    print(f"dense gemm cplx={cplx} m={m} n={n} R={R}", flush=True)
    rng = np.random.default_rng(m * 1000 + n + R)

    def rnd(*s):
        x = rng.normal(size=s)
        return x + 1j * rng.normal(size=s) if cplx else x
    A, X, Y = rnd(m, n), rnd(n, R), rnd(m, R)
    alpha, beta = (1.5 - 0.5j, 0.25 + 2j) if cplx else (1.5, -0.75)
    got = _gemm(inti, A, X, Y, alpha, beta)
    ref = alpha * (A @ X) + beta * Y
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref)


def test_dense_gemm_lane_layout_unit_matrices(inti):
    """Structured inputs that expose a transposed or permuted fragment:
    A = E_ij (one entry) against X = identity picks exactly entry (i, j)."""
    print("dense gemm layout", flush=True)
    m = n = 48
    for (i, j) in ((0, 0), (5, 17), (16, 3), (31, 47), (47, 30)):
        A = np.zeros((m, n))
        A[i, j] = 1.0
        got = _gemm(inti, A, np.eye(n), np.zeros((m, n)), 1.0, 0.0)
        assert got[i, j] == 1.0 and np.count_nonzero(got) == 1


@pytest.mark.parametrize("pde,k", [(0, 0.0), (1, 3.0)])
def test_dense_operator_matches_matrix_free(inti, pde, k):
    """The assembled corrected double layer (GPU assembly + on-device
    scatter-add of the DIM correction) applied to 16 right-hand sides by
    the matrix-core GEMM equals the matrix-free corrected apply (1e-12)."""
    print(f"dense operator pde={pde}", flush=True)
    Q = inti.sphere_quadrature(4)
    op = inti.Laplace(dim=3) if pde == 0 else inti.Helmholtz(k=k, dim=3)
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), Q)
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)
    dS, dD = inti.bdim_correction(op, Q, Q, Sop, Dop, green_multiplier=np.full(len(Q), -0.5))
    Dd = inti.DenseOperator(Dop, corrections=[(dD, 1.0)])
    Dop.attach(dD)
    rng = np.random.default_rng(3)
    X = rng.normal(size=(len(Q), 16))
    if pde == 1:
        X = X + 1j * rng.normal(size=X.shape)
    ref = Dop.apply(X)
    got = Dd.apply(X)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    M = Dd.to_numpy()
    ref_dense = inti.assemble_matrix(inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)) + dD.to_scipy().toarray()
    assert np.abs(M - ref_dense).max() <= 1e-14 * np.abs(ref_dense).max()
