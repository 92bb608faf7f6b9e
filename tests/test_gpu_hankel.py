"""The device Hankel functions, pinned through the 2D Helmholtz kernels on the
GPU against 50-digit mpmath values (tests/golden/golden.npz) over
x = kd in [1e-6, 1e3], which covers the 2D batch workload (kd up to ~240).

Targets x_i = (x_i, 0), one source at the origin with unit weight and
normal (1, 0), k = 1: the single layer gives (i/4) H0(x_i) and the double
layer (i/4) H1(x_i) (src/kernels.jl:360-371, hankelh1 at :344-345).

Bars (relative error of the complex value):
  FP64:  1e-13 + 2.2e-16 x. The second term is the conditioning of H_n to
         its argument: forming the distance and the phase k d / (2 pi) costs
         an ulp or two of x in any FP64 implementation, which moves H_n by
         about x * eps relative.
  MIXED: 1e-6."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _values(inti, x, kind, precision):
    n = len(x)
    trg = inti.points(np.stack([x, np.zeros(n)]))
    src = inti.Quadrature(np.zeros((2, 1)), np.array([[1.0], [0.0]]), np.ones(1),
                          np.zeros((1, 1), dtype=np.int64))
    op = inti.Helmholtz(k=1.0, dim=2)
    K = inti.SingleLayerKernel(op) if kind == "SL" else inti.DoubleLayerKernel(op)
    return inti.IntegralOperator(K, trg, src, precision=precision).apply(np.ones(1, dtype=np.complex128))


@pytest.mark.parametrize("kind", ["SL", "DL"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_hankel_against_mpmath(inti, golden, kind, precision):
    say(f"[gpu] hankel {kind} {precision}")
    x = golden["hankel_x"]
    ref = 0.25j * (golden["hankel_h0"] if kind == "SL" else golden["hankel_h1"])
    got = _values(inti, x, kind, precision)
    err = np.abs(got - ref) / np.abs(ref)
    bar = 1e-13 + 2.2e-16 * x if precision == "fp64" else np.full_like(x, 1e-6)
    worst = int(np.argmax(err / bar))
    say(f"  max rel err {err.max():.3e} (at x = {x[np.argmax(err)]:.4g}); worst vs bar at x = {x[worst]:.4g}")
    assert np.all(err <= bar), (float(err[worst]), float(x[worst]))
