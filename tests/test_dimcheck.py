"""The backward-error check of DIM weights (tests/_dimcheck.py) on the CPU:
both builds of the oracle (with and without FMA contraction) meet the bar,
weights perturbed by 1e-8 relative fail it, and so do weights of the wrong
element order or sign convention."""
import numpy as np
import pytest

from tests import _dimcheck as dc


def _cases(inti):
    return {"circle_laplace": (inti.circle_quadrature(128, qorder=9), 0, 2, 0.0),
            "ellipse_helmholtz": (inti.ellipse_quadrature(2.9, 0.3, 400, qorder=11), 1, 2, 35.0),
            "sphere_helmholtz": (inti.sphere_quadrature(6), 1, 3, 3.0)}


@pytest.mark.parametrize("case", ["circle_laplace", "ellipse_helmholtz", "sphere_helmholtz"])
def test_backward_error_separates_correct_and_perturbed_weights(inti, oracle, case):
    Q, pde, dim, k = _cases(inti)[case]
    mu = np.full(len(Q), -0.5)
    elems = dc.sample_elements(Q.ne, 12, 1, must=(0,))
    rows, A, th, tau = dc.ls_data(oracle, pde, dim, k, Q, mu, elems)
    near = dc.near_lists(Q, elems)
    for variant in (None, "fma"):
        rp, ci, vS, vD, _ = oracle.bdim(pde, dim, k, Q, Q, mu, near=near, variant=variant)
        eta = dc.weights_backward_error(Q, elems, A, th, tau, rows, rp, ci, vS, rp, ci, vD)
        assert eta.max() <= dc.ETA_BAR / 10, (variant, eta.max())
    rng = np.random.default_rng(0)
    bad = vD * (1 + 1e-8 * rng.normal(size=vD.shape))
    eta = dc.weights_backward_error(Q, elems, A, th, tau, rows, rp, ci, vS, rp, ci, bad)
    assert eta.max() > dc.ETA_BAR, eta.max()
    # the dS sign convention (dS = -W[nq+1:2nq]) flipped
    eta = dc.weights_backward_error(Q, elems, A, th, tau, rows, rp, ci, -vS, rp, ci, vD)
    assert eta.max() > 10 * dc.ETA_BAR
