"""Brute-force time-convolutionless generators for a Gaussian bath (test helper).

This evaluates the TCL expansion directly from its definition, sharing
nothing with the F/C/R -> MIKX -> NAKZWAN route of the reference
(tcl4_kernels.cpp, tcl4_mikx.cpp, tcl4_assemble.cpp):

* interaction V = A (x) B, zero-mean Gaussian bath with correlation
  C(tau) = <B(tau) B(0)>, C(-tau) = conj(C(tau));
* Liouvillian L(s) X = -i [A(s - t) (x) B(s), X] in the Schroedinger frame of
  the generator at time t (A(s) = e^{iHs} A e^{-iHs});
* second and fourth order of the TCL generator as ordered cumulants
  (Breuer & Petruccione, The Theory of Open Quantum Systems, 2002, sec. 9.2):
    K2(t) = int_0^t dt1 <L(t) L(t1)>,
    K4(t) = int_0^t dt1 int_0^t1 dt2 int_0^t2 dt3 [<L L1 L2 L3> - <L L1><L2 L3>
            - <L L2><L1 L3> - <L L3><L1 L2>],
  where <...> = P ... P is the bath average; the bath moments follow from
  Wick's theorem on the operator string each left/right placement produces;
* nested Gauss-Legendre quadrature on the simplex t3 < t2 < t1 < t.

The bath correlation is a sum of complex exponentials, so Gamma(w, t) =
int_0^t e^{iw tau} C(tau) d tau has a closed form, and the oracle can be fed
exact Gamma tables.  Superoperators act on column-major vec(rho):
vec(X rho Y) = (Y^T (x) X) vec(rho).
"""
from __future__ import annotations

import itertools

import numpy as np


class ExpBath:
    """C(tau) = sum_k a_k exp(-z_k tau) for tau >= 0, C(-tau) = conj C(tau)."""

    def __init__(self, a, z):
        self.a = np.asarray(a, np.complex128)
        self.z = np.asarray(z, np.complex128)

    def C(self, tau):
        tau = np.asarray(tau, np.float64)
        v = (self.a[None, :] * np.exp(-self.z[None, :] * np.abs(tau).reshape(-1, 1))).sum(1).reshape(tau.shape)
        return np.where(tau >= 0, v, np.conj(v))

    def gamma(self, omega: float, t):
        """Gamma(omega, t) = int_0^t e^{i omega tau} C(tau) d tau (closed form)."""
        t = np.asarray(t, np.float64)
        out = np.zeros(t.shape, np.complex128)
        for a, z in zip(self.a, self.z):
            k = 1j * omega - z
            out = out + a * np.expm1(k * t) / k
        return out

    def gamma_table(self, omegas, Nt: int, dt: float) -> np.ndarray:
        """[Nt, nf] table of Gamma(omega_b, k dt)."""
        tg = np.arange(Nt) * dt
        return np.stack([self.gamma(w, tg) for w in omegas], axis=1)


def _gl(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1) / 2, w / 2


def _kron_batch(Y, X):
    P, n, _ = Y.shape
    m = X.shape[1]
    return np.einsum("pij,pkl->pikjl", Y, X).reshape(P, n * m, n * m)


def _left(X):
    I = np.broadcast_to(np.eye(X.shape[1]), X.shape)
    return _kron_batch(I, X)


def _right(X):
    I = np.broadcast_to(np.eye(X.shape[1]), X.shape)
    return _kron_batch(np.transpose(X, (0, 2, 1)), I)


def _rotated(eps, A, s):
    """e^{iHs} A e^{-iHs} for H = diag(eps), at every s."""
    ph = np.exp(1j * np.outer(s, eps))
    return ph[:, :, None] * A[None] * np.conj(ph)[:, None, :]


def _wick(bath, times, string):
    """<B(t_a) B(t_b) ...> of a zero-mean Gaussian bath for the ordered string."""
    c = lambda a, b: bath.C(times[a] - times[b])  # noqa: E731
    if len(string) == 2:
        return c(*string)
    x0, x1, x2, x3 = string
    return c(x0, x1) * c(x2, x3) + c(x0, x2) * c(x1, x3) + c(x0, x3) * c(x1, x2)


def _moment(bath, As, times, idx):
    """P L(t_idx0) L(t_idx1) ... P (idx0 outermost), summed over the left/right
    placements of each commutator."""
    P, N = As[0].shape[0], As[0].shape[1]
    D = N * N
    out = np.zeros((P, D, D), np.complex128)
    for sides in itertools.product("LR", repeat=len(idx)):
        S = np.broadcast_to(np.eye(D), (P, D, D)).astype(np.complex128)
        coef = 1.0 + 0j
        for k, side in zip(idx, sides):
            S = S @ (_left(As[k]) if side == "L" else _right(As[k]))
            coef *= -1j if side == "L" else 1j
        # tr_B(B_l1 ... rho_B ... B_r1): right factors in application order
        # (innermost first), then the left factors outermost first
        rights = [k for k, side in reversed(list(zip(idx, sides))) if side == "R"]
        lefts = [k for k, side in zip(idx, sides) if side == "L"]
        out += coef * _wick(bath, times, rights + lefts)[:, None, None] * S
    return out


def unitary(eps) -> np.ndarray:
    """L0 = -i [diag(eps), .]."""
    N = len(eps)
    H = np.diag(eps).astype(np.complex128)
    return -1j * (np.kron(np.eye(N), H) - np.kron(H.T, np.eye(N)))


def k2(bath, eps, A, t: float, n: int = 48) -> np.ndarray:
    x, w = _gl(n)
    t1 = t * x
    times = [np.full(n, t), t1]
    As = [_rotated(eps, A, s - t) for s in times]
    return ((t * w)[:, None, None] * _moment(bath, As, times, [0, 1])).sum(0)


def k4(bath, eps, A, t: float, n: int = 14) -> np.ndarray:
    x, w = _gl(n)
    t1 = np.repeat(t * x, n * n)
    w1 = np.repeat(t * w, n * n)
    t2 = np.tile(np.repeat(x, n), n) * t1
    w2 = np.tile(np.repeat(w, n), n) * t1
    t3 = np.tile(x, n * n) * t2
    w3 = np.tile(w, n * n) * t2
    W = w1 * w2 * w3
    times = [np.full(W.size, t), t1, t2, t3]
    As = [_rotated(eps, A, s - t) for s in times]
    m = lambda *i: _moment(bath, As, times, list(i))  # noqa: E731
    integrand = m(0, 1, 2, 3) - m(0, 1) @ m(2, 3) - m(0, 2) @ m(1, 3) - m(0, 3) @ m(1, 2)
    return (W[:, None, None] * integrand).sum(0)


def secular(L, eps, tol: float = 1e-9) -> np.ndarray:
    """Entries L[(n,m),(i,j)] with eps_n - eps_m == eps_i - eps_j (vec index i + N j)."""
    N = len(eps)
    fr = np.array([eps[i] - eps[j] for j in range(N) for i in range(N)])
    return np.where(np.abs(fr[:, None] - fr[None, :]) < tol, L, 0)
