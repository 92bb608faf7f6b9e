"""The extended-precision restatement of the corrected 2D Laplace double layer
(tests/_extended.py) on the 640-node unit-circle workload (128 P2 elements x
5 Gauss nodes, sigma = cos(3 theta) + 0.1 N(0,1)), against the oracle.

Where the algorithm is well conditioned in its output (the smooth density
exp(cos theta), the uncorrected sum) the two agree to FP64 rounding, which
pins the restatement. On the seeded noisy density the oracle itself (built
with or without FMA contraction) is about 6e-5 from the extended result:
the DIM least squares have R-diagonal ratios of 5.3e8, so every FP64
evaluation of this pipeline is fixed only to that level on this input."""
import numpy as np

from tests import _extended


def _setup(inti, oracle):
    import bench
    Q, sig = bench.circle_workload(inti)
    n = len(Q)
    mu = np.full(n, -0.5)
    zs = oracle.monopoles(2, Q.coords, Q.nq)
    return Q, sig, mu, zs


def _oracle_corrected(oracle, Q, mu, s, variant=None):
    n = len(Q)
    rp, ci, vS, vD, _ = oracle.bdim(0, 2, 0.0, Q, Q, mu, variant=variant)
    A = oracle.assemble(0, 2, 1, 0.0, Q, Q).real
    return A @ s + oracle.csr_matvec(n, n, rp, ci, vD, s).real


def test_extended_kernel_matrices_match_oracle(inti, oracle):
    """Entry-wise to 1e-13 of the largest entry: for neighbouring nodes on
    the circle r.n ~ d^2 / 2 cancels, and the FP64 oracle entries there carry
    relative errors ~1e-16 / d (1.6e-14 of the largest entry measured)."""
    Q, _, _, _ = _setup(inti, oracle)
    X, N = _extended._ld(Q.coords), _extended._ld(Q.normals)
    for kind, kid in (("SL", 0), ("DL", 1), ("ADL", 2)):
        K = (_extended._laplace2d(kind, X, N, X, N) * _extended._ld(Q.weights)[None, :]).astype(np.float64)
        R = oracle.assemble(0, 2, kid, 0.0, Q, Q).real
        assert np.abs(K - R).max() <= 1e-13 * np.abs(R).max(), kind


def test_extended_matches_oracle_on_smooth_density(inti, oracle):
    Q, _, mu, zs = _setup(inti, oracle)
    th = np.arctan2(Q.coords[1], Q.coords[0])
    s = np.exp(np.cos(th))
    ye = _extended.corrected_dl_laplace2d(Q, zs, mu, s)
    yo = _oracle_corrected(oracle, Q, mu, s)
    err = np.linalg.norm(yo - ye) / np.linalg.norm(ye)
    assert err <= 1e-12, err


def test_fp64_error_scale_on_seeded_density(inti, oracle):
    """Both oracle builds sit 1e-5..1e-4 from the extended result: the
    size of the FP64 error of this pipeline on this density."""
    Q, sig, mu, zs = _setup(inti, oracle)
    ye = _extended.corrected_dl_laplace2d(Q, zs, mu, sig)
    for variant in (None, "fma"):
        err = np.linalg.norm(_oracle_corrected(oracle, Q, mu, sig, variant) - ye) / np.linalg.norm(ye)
        assert 1e-6 <= err <= 1e-4, (variant, err)


def test_planted_sign_error_in_extended_is_caught(inti, oracle):
    """A wrong Green multiplier in the extended pipeline breaks agreement on
    the smooth density."""
    Q, _, mu, zs = _setup(inti, oracle)
    th = np.arctan2(Q.coords[1], Q.coords[0])
    s = np.exp(np.cos(th))
    ye = _extended.corrected_dl_laplace2d(Q, zs, -mu, s)
    yo = _oracle_corrected(oracle, Q, mu, s)
    assert np.linalg.norm(yo - ye) / np.linalg.norm(yo) > 1e-6
