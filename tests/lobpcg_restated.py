"""The reference's LOBPCG restated in numpy on top of an H·v callable (test
helper): src/CalcByLOBPCG.c:328-632 with diag_ovrp (:43-123) and
calc_preshift (:130-147); numpy's eigh stands in for zheevd.  Vectors are
kept normalised explicitly, as in the reference."""

import math

import numpy as np


def _preshift(eig, res, eps):
    if res >= 1.0:
        return 0.0
    k = math.trunc(math.log10(abs(eig)))
    i = math.ceil(math.log10(eps)) if eps > res else math.ceil(math.log10(res))
    return math.trunc(eig / 10.0 ** (k + i)) * 10.0 ** (k + i)


def _diag_ovrp(hsub, ovlp):
    o, U = np.linalg.eigh(ovlp)
    keep = o > 1.0e-10
    T = U[:, keep] / np.sqrt(o[keep])
    e, Y = np.linalg.eigh(T.conj().T @ hsub @ T)
    return e, T @ Y


def lobpcg_restated(hv, x0, max_iter, eps0, diag=None, precg=False, eig_pos_shift=0.0):
    """x0: (dim, k) initial block.  Returns (eigenvalues = final Rayleigh
    quotients, history rows (max residual, mu_0..mu_{k-1}) per step)."""
    dim, k = x0.shape
    x = x0 / np.linalg.norm(x0, axis=0)
    X = np.stack([hv(x[:, c]) for c in range(k)], axis=1)
    p = np.zeros_like(x)
    P = np.zeros_like(x)
    mu = np.array([np.vdot(x[:, c], X[:, c]).real for c in range(k)])
    hist = []
    for stp in range(1, max_iter + 1):
        eps = eps0 * max(1.0, np.max(np.abs(mu)))
        w = X - mu * x
        dn = np.linalg.norm(w, axis=0)
        hist.append([dn.max(), *mu])
        if dn.max() < eps:
            break
        if stp != 1:
            if precg:
                for c in range(k):
                    shift = _preshift(mu[c] + eig_pos_shift, dn[c], eps) - eig_pos_shift
                    pc = diag - shift
                    m = np.abs(pc) > eps
                    w[m, c] /= pc[m]
            w = w / np.linalg.norm(w, axis=0)
        W = np.stack([hv(w[:, c]) for c in range(k)], axis=1)
        B = np.concatenate([w, x, p], axis=1)
        HB = np.concatenate([W, X, P], axis=1)
        hsub = B.conj().T @ HB
        ovlp = B.conj().T @ B
        mu = np.array([hsub[k + c, k + c].real for c in range(k)])
        e, Cf = _diag_ovrp(hsub, ovlp)
        mu = 0.5 * (mu + e[:k])
        C = Cf[:, :k]
        x, X = B @ C, HB @ C
        Cp = C.copy()
        Cp[k:2 * k] = 0.0
        p, P = B @ Cp, HB @ Cp
        nx, npn = np.linalg.norm(x, axis=0), np.linalg.norm(p, axis=0)
        x, X, p, P = x / nx, X / nx, p / npn, P / npn
    ev = np.array([np.vdot(x[:, c], X[:, c]).real for c in range(k)])
    return ev, np.array(hist)
