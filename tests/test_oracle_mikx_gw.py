"""Pins of the oracle's MIKX gather, NAKZWAN GW assembly and L4 reshuffle (CPU only).

Fixtures are the reference's own synthetic patterns:
cpp/tests/tcl4_mikx_tests.cpp:49-160 and tests/tcl4_tests.cpp:36-90.
"""
import itertools

import numpy as np
import pytest

P2F_FIXTURE = np.array([[0, 1], [2, 3]], np.int32)  # pair_to_freq(0,0)=0, (0,1)=1, (1,0)=2, (1,1)=3


def make_val(base, a, b, c, t=None):
    if t is None:
        return complex(base + 100.0 * a + 10.0 * b + c, 0.1 * base + 0.01 * a - 0.001 * b + 0.0001 * c)
    return complex(base + 100.0 * a + 10.0 * b + c + 0.001 * t,
                   0.1 * base + 0.01 * a - 0.001 * b + 0.0001 * c + 0.00001 * t)


def fixture_slices(nf=4, t=None):
    F = np.zeros((nf, nf, nf), complex)
    C = np.zeros_like(F)
    R = np.zeros_like(F)
    for a, b, c in itertools.product(range(nf), repeat=3):
        F[a, b, c] = make_val(1.0, a, b, c, t)
        C[a, b, c] = make_val(2.0, a, b, c, t)
        R[a, b, c] = make_val(3.0, a, b, c, t)
    return F, C, R


def test_mikx_make_val_fixture_code_order(orc):
    N = 2
    F, C, R = fixture_slices()
    M, I, K, X = orc.mikx(F, C, R, P2F_FIXTURE, N)
    f = lambda a, b: P2F_FIXTURE[a, b]  # noqa: E731
    X6 = X.reshape((N,) * 6, order="F")
    stale_differs = False
    for j, k, p, q in itertools.product(range(N), repeat=4):
        row, col = j + k * N, p + q * N
        # tcl4_mikx.cpp:111-113 — M1 = F[f(j,q)][f(j,k)][f(p,j)]
        assert M[row, col] == F[f(j, q), f(j, k), f(p, j)] - R[f(j, q), f(p, q), f(q, k)]
        assert I[row, col] == F[f(j, k), f(q, p), f(k, q)]
        assert K[row, col] == R[f(j, k), f(p, q), f(q, j)]
        stale = F[f(j, k), f(j, q), f(p, j)] - R[f(j, q), f(p, q), f(q, k)]
        stale_differs |= stale != M[row, col]
        for r, s in itertools.product(range(N), repeat=2):
            assert X6[j, k, p, q, r, s] == C[f(j, k), f(p, q), f(r, s)] + R[f(j, k), f(p, q), f(r, s)]
    # the unit test's expectation (tcl4_mikx_tests.cpp:133-134) is stale: it
    # swaps the first two F indices, which differs whenever k != q
    assert stale_differs


def test_mikx_rejects_missing_bucket(orc):
    F, C, R = fixture_slices()
    p2f = P2F_FIXTURE.copy()
    p2f[1, 0] = -1
    with pytest.raises(orc.OracleError, match="pair_to_freq"):
        orc.mikx(F, C, R, p2f, 2)


def nakzwan_numpy(M, I, K, X, ops, N):
    """Independent restatement of matlab/NAKZWAN_v9.m (1-based -> 0-based)."""
    M4 = M.reshape(N, N, N, N, order="F")  # M4[j,k,p,q] = M[(j,k),(p,q)]
    I4 = I.reshape(N, N, N, N, order="F")
    K4 = K.reshape(N, N, N, N, order="F")
    X6 = X.reshape((N,) * 6, order="F")
    T = np.zeros((N, N, N, N), complex)
    for n, i, m, j in itertools.product(range(N), repeat=4):
        res = 0j
        for A in ops:
            for B in ops:
                for a in range(N):
                    for b in range(N):
                        res = res - (B[n, a] * A[a, b] * B[b, i] * A[j, m] * M4[b, i, n, a]
                                     - A[n, a] * B[a, b] * B[b, i] * A[j, m] * I4[a, n, i, b]
                                     + B[n, a] * B[a, b] * A[b, i] * A[j, m] * K4[i, b, n, a]
                                     + A[n, a] * B[a, i] * B[j, b] * A[b, m] * X6[a, n, j, b, n, i]
                                     - B[n, a] * A[a, i] * B[j, b] * A[b, m] * X6[i, a, j, b, n, i]
                                     - A[n, a] * A[a, b] * B[b, i] * B[j, m] * X6[b, a, j, m, a, i]
                                     + A[n, a] * B[a, b] * A[b, i] * B[j, m] * X6[i, b, j, m, a, i])
                        if j == m:
                            for c in range(N):
                                res = res + (A[n, a] * B[a, b] * A[b, c] * B[c, i] * M4[c, i, a, b]
                                             + A[n, a] * B[a, b] * B[b, c] * A[c, i] * K4[i, c, a, b]
                                             - A[n, a] * A[a, b] * B[b, c] * B[c, i] * I4[b, a, i, c])
        T[n, i, m, j] = res
    T2 = T.reshape(N * N, N * N, order="F")
    return T2 + T2.conj().T


@pytest.mark.parametrize("t", [0, 2, 3])
def test_gw_matches_nakzwan_on_synthetic_fixture(orc, t):
    # tests/tcl4_tests.cpp:36-90: N=2, nf=4, ops {I, sx}, make_val with t
    N = 2
    F, C, R = fixture_slices(t=t)
    ops = [np.eye(2, dtype=complex), np.array([[0, 1], [1, 0]], complex)]
    M, I, K, X = orc.mikx(F, C, R, P2F_FIXTURE, N)
    GW = orc.assemble_gw(M, I, K, X, ops, N)
    ref = nakzwan_numpy(M, I, K, X, ops, N)
    assert np.abs(GW - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(GW, GW.conj().T)  # T + T^H, exactly Hermitian


def test_gw_random_system_and_l4_reshuffle(orc):
    rng = np.random.default_rng(5)
    N, nf = 3, 7
    shape = (nf, nf, nf)
    F, C, R = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3))
    p2f = rng.integers(0, nf, size=(N, N)).astype(np.int32)
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    A = A + A.conj().T
    M, I, K, X = orc.mikx(F, C, R, p2f, N)
    GW = orc.assemble_gw(M, I, K, X, [A], N)
    assert np.abs(GW - nakzwan_numpy(M, I, K, X, [A], N)).max() <= 1e-12 * np.abs(GW).max()
    L4 = orc.gw_to_l4(GW, N)
    # L4(n,m;i,j) = GW(n,i;m,j)
    G4 = GW.reshape(N, N, N, N, order="F")  # [n,i,m,j]
    ref = G4.transpose(0, 2, 1, 3).reshape(N * N, N * N, order="F")
    assert np.array_equal(L4, ref)
    # Hermiticity preservation: L4[(m,n),(j,i)] = conj(L4[(n,m),(i,j)])
    L44 = L4.reshape(N, N, N, N, order="F")  # [n,m,i,j]
    assert np.array_equal(L44.transpose(1, 0, 3, 2), L44.conj())
    # l4_from_slices == mikx -> assemble -> reshuffle
    L4b = orc.lib().orc_l4_from_slices
    assert L4b is not None
