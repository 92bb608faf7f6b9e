"""Helpers of the workload-size parity tests: row samples of the corrected
operators from the oracle, and the backward error of DIM weights.

The DIM least-squares problems (src/bdim.jl:137-185) are min || A w - t ||
with A = M^T (ns x 2 nq; M = [gamma0 B; gamma1 B] at the element's nodes)
and t = Theta[i, :] = mu G(x_i, z) + S_i B1 - D_i B0 (bdim.jl:106-134). Their
condition numbers reach 1e8-1e12 on the workloads (R-diagonal ratios), so
two valid FP64 roundings of the same algorithm disagree in the weights far
above 1e-12 (the oracle built with and without FMA contraction differs by up
to 50 % on elongated ellipses). What a correct implementation must meet is
a small backward error against the ORACLE's A and t:

  eta = || A^H (A w - t) || / (||A||_F (||A||_F ||w|| + tau)),

tau = || |mu| |G(x_i, z)| + |S_i| |B1| + |D_i| |B0| || (the sum of the
magnitudes of Theta's terms). Theta is a long sum with cancellation (on the
491,520-node sphere ||t|| ~ 3e-5 against tau ~ 0.16), and any FP64
summation order is exact only for terms perturbed at rounding level: tau is
the scale of that perturbation, as ||A|| ||w|| is the scale of the solve's.
A weight vector perturbed by 1e-8 relative gives eta > 2e-13 on these
problems; the oracle's own two builds give eta <= 2e-16.

TEST INFRASTRUCTURE ONLY (calls the oracle)."""
from types import SimpleNamespace as NS

import numpy as np

ETA_BAR = 1e-13


def sample_elements(ne, count, seed, must=()):
    rng = np.random.default_rng(seed)
    e = set(int(v) for v in rng.choice(ne, size=min(count, ne), replace=False))
    e.update(int(v) for v in must)
    return np.array(sorted(e), dtype=np.int64)


def near_lists(Q, elems):
    """Element -> target lists correcting only the sampled elements' own nodes."""
    ptr = np.zeros(Q.ne + 1, dtype=np.int64)
    cnt = np.zeros(Q.ne, dtype=np.int64)
    cnt[elems] = Q.nq
    ptr[1:] = np.cumsum(cnt)
    idx = Q.qtags[:, elems].T.reshape(-1).astype(np.int64)
    return ptr, idx


def rows_of(Q, elems):
    return Q.qtags[:, elems].T.reshape(-1).astype(np.int64)


def subset(Q, rows):
    return NS(coords=np.ascontiguousarray(Q.coords[:, rows]),
              normals=None if Q.normals is None else np.ascontiguousarray(Q.normals[:, rows]),
              weights=np.ascontiguousarray(Q.weights[rows]))


def csr_rows_apply(rp, ci, val, rows, x):
    """(CSR restricted to rows) @ x for x of shape (n,) or (n, r)."""
    out = []
    for i in rows:
        s = slice(rp[i], rp[i + 1])
        out.append(val[s] @ x[ci[s]])
    return np.array(out)


def same_pattern(rp_a, ci_a, rp_b, ci_b, rows):
    for i in rows:
        a = ci_a[rp_a[i]:rp_a[i + 1]]
        b = ci_b[rp_b[i]:rp_b[i + 1]]
        if a.shape != b.shape or np.any(a != b):
            return False
    return True


def ls_data(oracle, pde, dim, k, Q, mu, elems, derivative=False, chunk=48):
    """The oracle's least-squares data for the sampled elements: per element
    A_e = M_e^T (ns x 2nq), the Theta rows of its nodes (bdim.jl:94-134)
    and tau per row (the magnitude scale of Theta's terms)."""
    nq = Q.nq
    zs = oracle.monopoles(dim, Q.coords, nq)
    ns = zs.shape[1]
    Z = NS(coords=zs, normals=None, weights=np.ones(ns))
    rows = rows_of(Q, elems)
    T = subset(Q, rows)
    # traces B0 = G(y_j, z_k), B1 = ADL(y_j, z_k) at every source node
    B0 = oracle.assemble(pde, dim, 0, k, Q, Z)
    B1 = oracle.assemble(pde, dim, 2, k, Q, Z)
    kG = 2 if derivative else 0
    kS, kD = (2, 3) if derivative else (0, 1)
    th = oracle.matvec(pde, dim, kS, k, T, Q, B1) - oracle.matvec(pde, dim, kD, k, T, Q, B0)
    G = oracle.assemble(pde, dim, kG, k, T, Z)
    th = th + mu[rows][:, None] * G
    aB0, aB1 = np.abs(B0), np.abs(B1)
    tau = np.zeros(len(rows))
    for c0 in range(0, len(rows), chunk):
        Tc = subset(Q, rows[c0:c0 + chunk])
        acc = np.abs(mu[rows[c0:c0 + chunk]])[:, None] * np.abs(G[c0:c0 + chunk])
        acc += np.abs(oracle.assemble(pde, dim, kS, k, Tc, Q)) @ aB1
        acc += np.abs(oracle.assemble(pde, dim, kD, k, Tc, Q)) @ aB0
        tau[c0:c0 + chunk] = np.linalg.norm(acc, axis=1)
    A = []
    for e in elems:
        jg = Q.qtags[:, e]
        A.append(np.concatenate([B0[jg], B1[jg]]).T)
    return rows, A, th, tau


def weights_backward_error(Q, elems, A, th, tau, rows, rpS, ciS, vS, rpD, ciD, vD):
    """eta per corrected target (see module docstring); weights taken from
    CSR corrections dS, dD (dD = W[1:nq], dS = -W[nq+1:2nq], bdim.jl:175-182)."""
    nq = Q.nq
    etas = np.zeros(len(rows))
    for t, i in enumerate(rows):
        e = t // nq
        cols = Q.qtags[:, elems[e]]
        dD = dict(zip(ciD[rpD[i]:rpD[i + 1]], vD[rpD[i]:rpD[i + 1]]))
        dS = dict(zip(ciS[rpS[i]:rpS[i + 1]], vS[rpS[i]:rpS[i + 1]]))
        w = np.array([dD[j] for j in cols] + [-dS[j] for j in cols])
        Ae = A[e]
        r = Ae @ w - th[t]
        nA = np.linalg.norm(Ae)
        etas[t] = np.linalg.norm(Ae.conj().T @ r) / (nA * (nA * np.linalg.norm(w) + tau[t]))
    return etas


def theta_spread(Q, elems, A, tau, rows, sigma, cS, cD, y_ref, draws=3, seed=0):
    """How far the corrected rows move when Theta is perturbed at the level
    an FP64 summation may leave in it. Theta's row t is a sum over all n
    sources whose terms have magnitudes adding up to tau_t; the rounding
    errors of such a sum, in any order, grow like a random walk, about
    sqrt(n) * 2.2e-16 * tau_t (the oracle sums sequentially, the GPU in
    tiles). Each row gets a random perturbation of that norm, the
    per-element least squares are re-solved (numpy lstsq) and the change of
    the rows of cS dS + cD dD applied to sigma (dD = W[1:nq],
    dS = -W[nq+1:2nq]) is measured. Returns the largest relative change
    over `draws` draws."""
    rng = np.random.default_rng(seed)
    nq = Q.nq
    worst = 0.0
    for _ in range(draws):
        dy = np.zeros((len(rows),) + np.shape(sigma)[1:], dtype=np.complex128)
        for e, el in enumerate(elems):
            cols = Q.qtags[:, el]
            ns = A[e].shape[0]
            xi = rng.normal(size=(ns, nq)) + (1j * rng.normal(size=(ns, nq)) if np.iscomplexobj(A[e]) else 0)
            xi *= np.sqrt(len(Q)) * 2.2e-16 * tau[e * nq:(e + 1) * nq][None, :] / np.linalg.norm(xi, axis=0)[None, :]
            dW = np.linalg.lstsq(A[e], xi, rcond=None)[0]           # 2nq x nq: one column per target
            comb = cD * dW[:nq] - cS * dW[nq:]                      # weights on the element's nodes
            dy[e * nq:(e + 1) * nq] = comb.T @ sigma[cols]
        worst = max(worst, float(np.linalg.norm(dy) / np.linalg.norm(y_ref)))
    return worst


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))
