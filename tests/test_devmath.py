"""The FP64 sin/cos of a phase in revolutions used by the device kernels
(devmath.hpp: sincos_rev64): its coefficient tables, read from the header,
and its reduction q = rint(4u), f = u - q/4, evaluated here in numpy with
the same Horner order, against 50-digit mpmath sin(2 pi u), cos(2 pi u) of
the same double u. A wrong coefficient digit, degree or quadrant rule
breaks the 4-ulp bound."""
import os
import re

import mpmath as mp
import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "inti.jl_amd", "csrc", "devmath.hpp")


def _table(name):
    text = open(HDR).read()
    body = re.search(name + r"\[\d+\]\s*=\s*\{([^}]*)\}", text).group(1)
    return [float.fromhex(t.strip()) for t in body.split(",") if t.strip()]


def _sincos_rev(u):
    sinc, cosc = _table("kSinRev"), _table("kCosRev")
    q = np.rint(4.0 * u)
    f = u - 0.25 * q
    z = f * f
    ps = np.full_like(u, sinc[-1])
    for c in sinc[-2::-1]:
        ps = ps * z + c
    pc = np.full_like(u, cosc[-1])
    for c in cosc[-2::-1]:
        pc = pc * z + c
    S, C = f * ps, pc
    qi = q.astype(np.int64) & 3
    s0 = np.where(qi & 1, C, S)
    c0 = np.where(qi & 1, S, C)
    s = np.where(qi & 2, -s0, s0)
    c = np.where((qi ^ (qi >> 1)) & 1, -c0, c0)
    return s, c


def test_tables_are_the_taylor_coefficients():
    mp.mp.dps = 50
    tp = 2 * mp.pi
    sinc, cosc = _table("kSinRev"), _table("kCosRev")
    assert len(sinc) == 8 and len(cosc) == 9
    for k, v in enumerate(sinc):
        assert v == float((-1) ** k * tp ** (2 * k + 1) / mp.factorial(2 * k + 1))
    for k, v in enumerate(cosc):
        assert v == float((-1) ** k * tp ** (2 * k) / mp.factorial(2 * k))


def test_sincos_rev_against_mpmath():
    mp.mp.dps = 40
    rng = np.random.default_rng(5)
    u = np.concatenate([rng.uniform(-1, 1, 3000), rng.uniform(-300, 300, 3000),
                        np.arange(-8, 9) / 8.0, np.arange(-8, 9) / 8.0 + 1e-12,
                        np.arange(-8, 9) / 8.0 - 1e-12, [0.0, 1e-300, -1e-300, 1e6 + 0.3]])
    s, c = _sincos_rev(u)
    es = max(abs(float(mp.sin(2 * mp.pi * mp.mpf(x))) - y) for x, y in zip(u, s))
    ec = max(abs(float(mp.cos(2 * mp.pi * mp.mpf(x))) - y) for x, y in zip(u, c))
    assert es <= 4.5e-16 and ec <= 4.5e-16, (es, ec)
