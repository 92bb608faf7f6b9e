"""Writes golden.npz: reference values computed with mpmath (independent of
both the oracle and the product) at 40 significant digits, rounded to FP64.

  hankel_x, hankel_h{0,1,2}: H_n^(1)(x) = J_n(x) + i Y_n(x) on a log grid of
      x in [1e-6, 1e3] plus points near zeros of J0, J1, Y0, Y1 and the
      region boundaries of the device evaluation (3, 5);
  gl_{n}_x, gl_{n}_w: n-point Gauss-Legendre rule on [0, 1], n = 1..12.

Run:  python tests/golden/make_golden.py   (needs mpmath)
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def hankel_points():
    xs = list(np.logspace(-6, 3, 181))
    for z in (0.8935769662791675, 2.404825557695773, 3.8317059702075125, 3.957678419314858,
              5.520078110286311, 7.015586669815619, 2.197141326031017):
        xs += [z * (1 - 1e-9), z, z * (1 + 1e-9)]
    for b in (3.0, 5.0, 25.0):
        xs += [b * (1 - 1e-12), b, b * (1 + 1e-12)]
    return np.array(sorted(set(float(x) for x in xs)))


def gl(n):
    # roots of P_n by Newton in mp precision from Chebyshev-like guesses
    xs, ws = [], []
    for i in range(1, n + 1):
        x = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (n + mp.mpf(1) / 2))
        for _ in range(100):
            p = mp.legendre(n, x)
            dp = n * (x * p - mp.legendre(n - 1, x)) / (x * x - 1)
            dx = p / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** (-38):
                break
        dp = n * (x * mp.legendre(n, x) - mp.legendre(n - 1, x)) / (x * x - 1)
        xs.append((x + 1) / 2)
        ws.append(1 / ((1 - x * x) * dp * dp))
    order = sorted(range(n), key=lambda i: xs[i])
    return np.array([float(xs[i]) for i in order]), np.array([float(ws[i]) for i in order])


def main():
    out = {}
    x = hankel_points()
    out["hankel_x"] = x
    for n in (0, 1, 2):
        h = np.array([complex(mp.besselj(n, mp.mpf(v)) + 1j * mp.bessely(n, mp.mpf(v))) for v in x])
        out[f"hankel_h{n}"] = h
    for n in range(1, 13):
        gx, gw = gl(n) if n > 1 else (np.array([0.5]), np.array([1.0]))
        out[f"gl_{n}_x"] = gx
        out[f"gl_{n}_w"] = gw
    np.savez(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
