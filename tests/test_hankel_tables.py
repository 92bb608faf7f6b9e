"""The device Hankel functions H0^(1), H1^(1) (devmath.hpp: hankel01_f64,
hankel01_mixed) evaluated here in numpy from the coefficient tables of
hankel_coeffs.h, in the device's branch structure and Horner order, against
the 50-digit mpmath values in tests/golden/golden.npz over x in [1e-6, 1e3]
(the kd range of the workloads, up to kd ~ 240 for the 2D batch).

FP64 tables in float64 arithmetic: relative error of the complex value
<= 1e-13. FP32 tables in float32 arithmetic (the mixed kernels): <= 1e-6.
sin/cos/log/rsqrt are numpy's here; the device's own transcendentals are
pinned by the GPU test of the same goldens (test_gpu_hankel.py). A wrong
digit, a dropped term, a swapped table or a wrong phase (x - pi/4) fails
these bounds."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "inti.jl_amd", "csrc", "hankel_coeffs.h")
GAMMA = 0.57721566490153286061


def _tables():
    text = open(HDR).read()
    out = {}
    for m in re.finditer(r"(HK_F(?:64|32)_[A-Z0-9_]+)(\[\d+\])?\s*=\s*(\{[^}]*\}|[^;]+);", text):
        name, body = m.group(1), m.group(3).strip()
        if body.startswith("{"):
            vals = [t.strip().rstrip("f") for t in body[1:-1].split(",") if t.strip()]
            out[name] = [float(v) for v in vals]
        else:
            out[name] = float(body.rstrip("f"))
    return out


def _horner(c, u, dt):
    s = np.full_like(u, dt(c[-1]))
    for v in c[-2::-1]:
        s = (s * u + dt(v)).astype(dt)
    return s


def _hankel01(x, prec, T=None):
    T = _tables() if T is None else T
    dt = np.float64 if prec == "F64" else np.float32
    P = "HK_" + prec + "_"
    x64 = np.asarray(x, dtype=np.float64)
    x = x64.astype(dt)
    xs = T[P + "XS"]
    h0 = np.zeros(x.shape, np.complex128)
    h1 = np.zeros(x.shape, np.complex128)
    lo = x < xs
    xl = x[lo]
    u = (dt(0.25) * xl * xl).astype(dt)
    a0, a1 = _horner(T[P + "A0"], u, dt), _horner(T[P + "A1"], u, dt)
    b0, b1 = _horner(T[P + "B0"], u, dt), _horner(T[P + "B1"], u, dt)
    L = np.log(dt(0.5) * xl).astype(dt)
    J0, J1 = a0, (dt(0.5) * xl * a1).astype(dt)
    tp = dt(2 / np.pi)
    h0[lo] = J0 + 1j * (tp * ((L + dt(GAMMA)) * J0 + b0))
    h1[lo] = J1 + 1j * (-tp / xl + tp * L * J1 - dt(0.5 / np.pi) * xl * b1)
    xh, xh64 = x[~lo], x64[~lo]
    ix = (dt(1) / xh).astype(dt)
    t = (dt(xs * xs) * ix * ix).astype(dt)

    def rat(name):
        return (_horner(T[P + name + "_NUM"], t, dt) / _horner(T[P + name + "_DEN"], t, dt)).astype(dt)

    P0, Q0, P1, Q1 = rat("P0"), rat("Q0") * ix, rat("P1"), rat("Q1") * ix
    amp = (dt(np.sqrt(2 / np.pi)) / np.sqrt(xh)).astype(dt)
    # the phase x - pi/4 in revolutions, u = fma(x, 1/(2 pi), -1/8), formed
    # from the FP64 argument in both variants and reduced exactly to [-1/2, 1/2]
    u = (xh64.astype(np.longdouble) * np.longdouble(1 / (2 * np.pi)) - np.longdouble(0.125)).astype(np.float64)
    f = u - np.rint(u)
    s, c = np.sin(2 * np.pi * f).astype(dt), np.cos(2 * np.pi * f).astype(dt)
    h0[~lo] = amp * (P0 * c - Q0 * s) + 1j * amp * (P0 * s + Q0 * c)
    h1[~lo] = amp * (P1 * s + Q1 * c) + 1j * amp * (Q1 * s - P1 * c)
    return h0, h1


def test_tables_parse():
    T = _tables()
    for prec, ns, nr in (("F64", 18, 7), ("F32", 10, 4)):
        for t in ("A0", "A1", "B0", "B1"):
            assert len(T[f"HK_{prec}_{t}"]) == ns
        for t in ("P0", "Q0", "P1", "Q1"):
            assert len(T[f"HK_{prec}_{t}_NUM"]) == nr and len(T[f"HK_{prec}_{t}_DEN"]) == nr
    # leading series terms are exact: J0 = 1 - u + u^2/4, J1/(x/2) = 1 - u/2
    assert T["HK_F64_A0"][:3] == [1.0, -1.0, 0.25] and T["HK_F64_A1"][:2] == [1.0, -0.5]


def test_fp64_tables_against_mpmath(golden):
    x = golden["hankel_x"]
    h0, h1 = _hankel01(x, "F64")
    e0 = np.abs(h0 - golden["hankel_h0"]) / np.abs(golden["hankel_h0"])
    e1 = np.abs(h1 - golden["hankel_h1"]) / np.abs(golden["hankel_h1"])
    assert x.min() <= 1e-6 and x.max() >= 1e3
    assert e0.max() <= 1e-13 and e1.max() <= 1e-13, (e0.max(), e1.max())


def test_fp32_tables_against_mpmath(golden):
    x = golden["hankel_x"]
    h0, h1 = _hankel01(x, "F32")
    e0 = np.abs(h0 - golden["hankel_h0"]) / np.abs(golden["hankel_h0"])
    e1 = np.abs(h1 - golden["hankel_h1"]) / np.abs(golden["hankel_h1"])
    assert e0.max() <= 1e-6 and e1.max() <= 1e-6, (e0.max(), e1.max())


def test_planted_table_error_is_caught(golden):
    """A coefficient off in its 11th digit must fail the FP64 bound."""
    T = _tables()
    T["HK_F64_P0_NUM"] = list(T["HK_F64_P0_NUM"])
    T["HK_F64_P0_NUM"][2] *= 1 + 1e-11
    x = golden["hankel_x"]
    h0, _ = _hankel01(x, "F64", T)
    e0 = np.abs(h0 - golden["hankel_h0"]) / np.abs(golden["hankel_h0"])
    assert e0.max() > 1e-13
