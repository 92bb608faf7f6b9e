"""The corrected 2D Laplace double layer (bdim_correction + assemble_matrix +
matvec, src/bdim.jl:50-196, src/nystrom.jl:93-124) restated in numpy
extended precision (np.longdouble: the x87 80-bit format, 64-bit mantissa,
unit roundoff 5.4e-20, about 2,000 times finer than FP64).

Its inputs are the FP64 ones every implementation shares: nodes, normals,
weights, element tags and the FP64 monopole centres. On inputs whose DIM
least squares have condition numbers kappa ~ 5e8, an FP64 evaluation of the
algorithm is fixed only to ~kappa * 1e-16 in the weights; this evaluation
is fixed to ~kappa * 5e-20, so it stands in for the exact result of the
algorithm on those inputs, and the FP64 implementations (the oracle, the
GPU) can each be measured against it.

The least squares use Householder QR with the reflector convention of the
generic unblocked QR (bdim.jl:163); any backward-stable method gives the
same solution to the precision that matters here.

TEST INFRASTRUCTURE ONLY."""
import numpy as np

LD = np.longdouble
TWO_PI = LD(2) * LD("3.14159265358979323846264338327950288")
SAME_POINT_TOLERANCE = 1e-14  # src/utils.jl:235


def _ld(a):
    return np.asarray(a, dtype=np.float64).astype(LD)


def _laplace2d(kind, X, NX, Y, NY):
    """Kernel matrix K(x_i, y_j), src/kernels.jl:137-204 (2D)."""
    r0 = X[0][:, None] - Y[0][None, :]
    r1 = X[1][:, None] - Y[1][None, :]
    d2 = r0 * r0 + r1 * r1
    near = np.sqrt(d2) <= SAME_POINT_TOLERANCE
    d2s = np.where(near, LD(1), d2)
    if kind == "SL":
        K = -np.log(d2s) / (LD(2) * TWO_PI)
    elif kind == "DL":
        K = (r0 * NY[0][None, :] + r1 * NY[1][None, :]) / (TWO_PI * d2s)
    elif kind == "ADL":
        K = -(r0 * NX[0][:, None] + r1 * NX[1][:, None]) / (TWO_PI * d2s)
    else:
        raise ValueError(kind)
    return np.where(near, LD(0), K)


def _householder_ls(A, B):
    """Least-squares solutions of A x = b for every column b of B (A tall,
    full rank), Householder QR with v = x + sign(x1) ||x|| e1."""
    A = A.copy()
    B = B.copy()
    m, n = A.shape
    for j in range(n):
        x = A[j:, j]
        nu = np.sqrt(np.sum(x * x))
        s = LD(1) if x[0] >= 0 else LD(-1)
        v = x.copy()
        v[0] += s * nu
        vv = np.sum(v * v)
        if vv == 0:
            continue
        A[j:, j:] -= np.outer(v, (LD(2) / vv) * (v @ A[j:, j:]))
        B[j:, :] -= np.outer(v, (LD(2) / vv) * (v @ B[j:, :]))
    R = A[:n, :n]
    Xs = np.zeros((n, B.shape[1]), dtype=LD)
    for i in range(n - 1, -1, -1):
        Xs[i] = (B[i] - R[i, i + 1:] @ Xs[i + 1:]) / R[i, i]
    return Xs


def corrected_dl_laplace2d(Q, zs, mu, sigma):
    """(D + dD) sigma in extended precision, self pattern. Q: coords (2, n),
    normals, weights, qtags (nq, ne); zs: FP64 monopole centres (2, ns)."""
    X, N, w = _ld(Q.coords), _ld(Q.normals), _ld(Q.weights)
    Z = _ld(zs)
    s = _ld(sigma)
    n = X.shape[1]
    nq, ne = Q.qtags.shape
    S = _laplace2d("SL", X, N, X, N) * w[None, :]
    D = _laplace2d("DL", X, N, X, N) * w[None, :]
    B0 = _laplace2d("SL", X, N, Z, None)    # G(y_j, z_k)
    B1 = _laplace2d("ADL", X, N, Z, None)   # ADL(y_j, z_k), target normal n_j
    theta = _ld(mu)[:, None] * B0 + S @ B1 - D @ B0
    y = D @ s
    for e in range(ne):
        jg = Q.qtags[:, e]
        A = np.concatenate([B0[jg], B1[jg]]).T          # ns x 2nq
        W = _householder_ls(A, theta[jg].T)            # 2nq x nq (one column per target)
        # dD[i, jg] = W[:nq, i]: the correction of row i = jg[t]
        y[jg] += W[:nq].T @ s[jg]
    return y.astype(np.float64)
