"""C-ABI surface without a GPU: the library loads, exports exactly the
declared entry points, and rejects bad arguments with status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    hdr = open(os.path.join(ROOT, "include", "inti_amd.h")).read()
    return set(re.findall(r"^(?:int|const char \*)\s*(inti_\w+)\(", hdr, flags=re.M))


def test_library_exports_every_declared_symbol(inti):
    lib = inti._lib.load()
    names = _declared()
    assert len(names) >= 27
    for name in names:
        assert hasattr(lib, name), name
    # the ctypes table covers the header, and nothing more
    assert names == set(inti._lib.SIGNATURES)


def test_header_constants_match_binding(inti):
    hdr = open(os.path.join(ROOT, "include", "inti_amd.h")).read()
    defs = dict(re.findall(r"#define (INTI_\w+) (\d+)", hdr))
    L = inti._lib
    assert int(defs["INTI_OK"]) == L.INTI_OK
    assert int(defs["INTI_ERR_ARG"]) == L.INTI_ERR_ARG
    assert int(defs["INTI_HELMHOLTZ"]) == L.HELMHOLTZ
    assert int(defs["INTI_COMB_AH"]) == L.COMB_AH
    assert int(defs["INTI_MIXED"]) == L.MIXED


def test_null_handles_are_rejected(inti):
    lib = inti._lib.load()
    assert lib.inti_version() >= 100
    assert lib.inti_ctx_destroy(None) == inti._lib.INTI_ERR_ARG
    assert lib.inti_quad_destroy(None) == inti._lib.INTI_ERR_ARG
    assert lib.inti_op_apply(None, 1, None, None, None, None) == inti._lib.INTI_ERR_ARG
    assert lib.inti_op_assemble_dense(None, None) == inti._lib.INTI_ERR_ARG
    assert lib.inti_csr_export(None, None, None, None) == inti._lib.INTI_ERR_ARG
    assert lib.inti_quad_upload(None, 2, 0, None, None, None, None) == inti._lib.INTI_ERR_ARG


def test_context_without_device_fails_loudly(inti):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present: covered by the GPU tests")
    with pytest.raises(inti.IntiError) as e:
        inti.Context(0)
    assert "device" in str(e.value).lower()


def test_host_option_validation(inti):
    from inti_jl_amd import api
    Q = inti.circle_quadrature(8)
    with pytest.raises(ValueError):
        api._normalize_compression({"method": "hmatrix"})
    with pytest.raises(ValueError):
        api._normalize_correction({"method": "adaptive"}, Q, Q)
    P = inti.points(np.zeros((2, 3)))
    with pytest.raises(ValueError):
        api._normalize_correction({"method": "dim"}, P, Q)
    with pytest.raises(ValueError):
        api._normalize_correction({"method": "dim", "target_location": "inside",
                                   "green_multiplier": [0, 0, 0]}, P, Q)
    assert inti._green_multiplier("on") == -0.5
    assert inti._green_multiplier("inside") == -1.0
    assert inti._green_multiplier("outside") == 0.0
    with pytest.raises(ValueError):
        inti._green_multiplier("above")
    with pytest.raises(NotImplementedError):
        inti.Helmholtz(k=1 + 1j, dim=2)


def test_nearest_lists_match_oracle(inti, oracle):
    """etype_to_nearest_points (quadrature.jl:264-286) vs brute force."""
    Q = inti.circle_quadrature(16, qorder=5)
    rng = np.random.default_rng(5)
    P = inti.points(rng.uniform(-1.4, 1.4, size=(2, 200)))
    for maxdist in (np.inf, 0.2):
        ptr, idx = inti.etype_to_nearest_points(P, Q, maxdist)
        optr, oidx = oracle.nearest_lists(P, Q, maxdist)
        assert np.array_equal(ptr, optr) and np.array_equal(idx, oidx)
    assert inti.etype_to_nearest_points(Q, Q) is None
