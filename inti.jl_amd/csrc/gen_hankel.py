"""Generate hankel_coeffs.h: coefficient tables for the device Hankel
functions H0^(1), H1^(1) of real argument used by the 2D Helmholtz kernels
(Inti.jl src/kernels.jl:344 calls Bessels.hankelh1 for real k*d).

Method (DLMF / Abramowitz-Stegun, own fits):
  * x < XS: ascending series (A&S 9.1.10, 9.1.11)
        J0 = A0(u),  J1 = (x/2) A1(u),                     u = x^2/4
        Y0 = (2/pi) [(ln(x/2) + gamma) J0 + B0(u)]
        Y1 = -2/(pi x) + (2/pi) ln(x/2) J1 - (x/(2 pi)) B1(u)
    with A0, A1, B0, B1 entire power series in u, truncated.
  * x >= XS: modulus/phase form (A&S 9.2.5-9.2.6)
        H_n(x) = sqrt(2/(pi x)) (P_n + i Q_n) exp(i (x - (2n+1) pi/4))
    with P_n = NP(t)/DP(t) and Q_n = (1/x) NQ(t)/DQ(t), t = XS^2/x^2 in (0,1],
    rational least-squares fits (Sanathanan-Koerner iteration, relative
    weights) to mpmath values; the t -> 0 end uses the asymptotic series.

Two tables: FP64 (relative error ~1e-16) and FP32 (~1e-8).
Run:  python gen_hankel.py > hankel_coeffs.h     (needs mpmath)
"""
import sys

import mpmath as mp

mp.mp.dps = 50


def ak(n, k):
    nu2 = 4 * n * n
    p = mp.mpf(1)
    for j in range(1, k + 1):
        p *= nu2 - (2 * j - 1) ** 2
    return p / (mp.factorial(k) * mp.mpf(8) ** k)


def pq_exact(n, x):
    x = mp.mpf(x)
    if x > 2000:
        P = mp.mpf(0)
        Q = mp.mpf(0)
        for k in range(31):
            t = ak(n, k) / x ** k
            r = k % 4
            if r == 0:
                P += t
            elif r == 1:
                Q += t
            elif r == 2:
                P -= t
            else:
                Q -= t
        return P, Q
    H = mp.besselj(n, x) + 1j * mp.bessely(n, x)
    v = H * mp.sqrt(mp.pi * x / 2) * mp.exp(-1j * (x - (2 * n + 1) * mp.pi / 4))
    return v.real, v.imag


def target(n, which, XS, t):
    if t == 0:
        return mp.mpf(1) if which == 0 else ak(n, 1)
    x = mp.mpf(XS) / mp.sqrt(t)
    P, Q = pq_exact(n, x)
    return P if which == 0 else Q * x


def ratfit(f, L, M, npts=120, iters=8):
    ts = [(1 - mp.cos(mp.pi * (i + mp.mpf(1) / 2) / npts)) / 2 for i in range(npts)]
    fv = [f(t) for t in ts]
    dprev = [mp.mpf(1)] * npts
    for _ in range(iters):
        rows, rhs = [], []
        for t, fx, dp in zip(ts, fv, dprev):
            w = 1 / (dp * fx)
            rows.append([w * t ** i for i in range(L + 1)] + [-w * fx * t ** j for j in range(1, M + 1)])
            rhs.append(w * fx)
        sol, _ = mp.qr_solve(mp.matrix(rows), mp.matrix(rhs))
        a = [sol[i] for i in range(L + 1)]
        b = [mp.mpf(1)] + [sol[L + j] for j in range(1, M + 1)]
        dprev = [sum(b[j] * t ** j for j in range(M + 1)) for t in ts]
    return a, b


def series_coeffs(K):
    A0 = [(-1) ** k / mp.factorial(k) ** 2 for k in range(K)]
    A1 = [(-1) ** k / (mp.factorial(k) * mp.factorial(k + 1)) for k in range(K)]
    H = [mp.mpf(0)]
    for k in range(1, K + 1):
        H.append(H[-1] + mp.mpf(1) / k)
    B0 = [mp.mpf(0)] + [(-1) ** (k + 1) * H[k] / mp.factorial(k) ** 2 for k in range(1, K)]
    # psi(k+1) + psi(k+2) = -2 gamma + H_k + H_{k+1}
    B1 = [(-1) ** k * (-2 * mp.euler + H[k] + H[k + 1]) / (mp.factorial(k) * mp.factorial(k + 1))
          for k in range(K)]
    return A0, A1, B0, B1


def horner(c, u):
    s = mp.mpf(0)
    for v in reversed(c):
        s = s * u + v
    return s


def check_series(K, XS):
    """Max relative error of the small-x formulas vs mpmath for x < XS."""
    A0, A1, B0, B1 = series_coeffs(K)
    worst = 0
    for i in range(1, 200):
        x = mp.mpf(XS) * i / 200
        u = x * x / 4
        J0 = horner(A0, u)
        J1 = x / 2 * horner(A1, u)
        L = mp.log(x / 2)
        Y0 = 2 / mp.pi * ((L + mp.euler) * J0 + horner(B0, u))
        Y1 = -2 / (mp.pi * x) + 2 / mp.pi * L * J1 - x / (2 * mp.pi) * horner(B1, u)
        for n, (J, Y) in enumerate(((J0, Y0), (J1, Y1))):
            ex = mp.besselj(n, x) + 1j * mp.bessely(n, x)
            worst = max(worst, abs((J + 1j * Y) - ex) / abs(ex))
    return worst


def check_rational(fits, XS):
    worst = 0
    for i in range(1, 400):
        x = mp.mpf(XS) + (mp.mpf(i) / 400) ** 2 * 200
        t = (mp.mpf(XS) / x) ** 2
        for n in (0, 1):
            aP, bP = fits[(n, 0)]
            aQ, bQ = fits[(n, 1)]
            P = horner(aP, t) / horner(bP, t)
            Q = horner(aQ, t) / horner(bQ, t) / x
            Pe, Qe = pq_exact(n, x)
            worst = max(worst, abs(mp.mpc(P, Q) - mp.mpc(Pe, Qe)) / abs(mp.mpc(Pe, Qe)))
    return worst


def emit(name, vals, ctype, suffix):
    body = ", ".join(mp.nstr(v, 20, min_fixed=-1, max_fixed=-1) + suffix for v in vals)
    return f"static constexpr {ctype} {name}[{len(vals)}] = {{{body}}};\n"


def build(prec, XS, K, L, M):
    ctype = "double" if prec == "f64" else "float"
    suffix = "" if prec == "f64" else "f"
    fits = {}
    for n in (0, 1):
        for which in (0, 1):
            fits[(n, which)] = ratfit(lambda t: target(n, which, XS, t), L, M)
    es = check_series(K, XS)
    er = check_rational(fits, XS)
    out = [f"// {prec}: x < {XS}: {K}-term series (max rel err {mp.nstr(es, 3)});"
           f" x >= {XS}: ({L},{M}) rationals (max rel err {mp.nstr(er, 3)})\n"]
    out.append(f"static constexpr {ctype} HK_{prec.upper()}_XS = {XS}{'.0' if prec == 'f64' else '.0f'};\n")
    A0, A1, B0, B1 = series_coeffs(K)
    for nm, c in (("A0", A0), ("A1", A1), ("B0", B0), ("B1", B1)):
        out.append(emit(f"HK_{prec.upper()}_{nm}", c, ctype, suffix))
    for n in (0, 1):
        for which, nm in ((0, "P"), (1, "Q")):
            a, b = fits[(n, which)]
            out.append(emit(f"HK_{prec.upper()}_{nm}{n}_NUM", a, ctype, suffix))
            out.append(emit(f"HK_{prec.upper()}_{nm}{n}_DEN", b, ctype, suffix))
    return "".join(out), es, er


def main():
    hdr = ["// Generated by gen_hankel.py -- do not edit by hand.\n",
           "// Coefficients for H0^(1), H1^(1) of real argument; see gen_hankel.py.\n",
           "#pragma once\n"]
    s64, es64, er64 = build("f64", 5, 18, 6, 6)
    s32, es32, er32 = build("f32", 3, 10, 3, 3)
    assert es64 < 2e-14 and er64 < 1e-15, (es64, er64)
    assert es32 < 2e-8 and er32 < 2e-8, (es32, er32)
    sys.stdout.write("".join(hdr) + s64 + s32)


if __name__ == "__main__":
    main()
