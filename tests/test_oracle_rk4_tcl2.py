"""Pins of the oracle's RK4 propagation and TCL2 generator (CPU only).

RK4: tests/rk4_dense_tests.cpp:43-102 closed forms; TCL2: generator.cpp:81-177.
"""
import numpy as np
import pytest
from scipy.linalg import expm


def test_constant_L_matches_expm(orc):
    # rk4_dense_tests.cpp:43-70
    D = 4
    L = np.zeros((D, D), complex)
    L[0, 0] = 0.10 - 0.03j
    L[1, 1] = -0.07 + 0.05j
    L[2, 2] = 0.02
    L[3, 3] = -0.04 - 0.02j
    L[0, 2] = 0.03 + 0.01j
    L[2, 0] = -0.02 + 0.02j
    L[1, 3] = 0.01 - 0.04j
    r0 = np.array([1.0, 0.5 - 0.2j, -0.7 + 0.1j, 0.1 + 0.3j])
    tf, dt = 1.25, 0.0025
    steps = int(round(tf / dt))
    r, _ = orc.propagate(np.repeat(L[None], steps + 1, axis=0), r0, dt)
    expect = expm(tf * L) @ r0
    err = np.linalg.norm(r - expect) / max(1.0, np.linalg.norm(expect))
    assert err <= 1e-6
    assert err <= 1e-12  # RK4 error at this step is far below the reference's tolerance


def test_time_dependent_L_midpoints(orc):
    # rk4_dense_tests.cpp:72-102: L(t) = t I, exact r(t) = exp(t^2/2) r0
    D, steps, dt = 3, 100, 0.01
    ts = np.arange(steps + 1) * dt
    Ls = np.array([t * np.eye(D) for t in ts], complex)
    Lh = np.array([(t + 0.5 * dt) * np.eye(D) for t in ts[:-1]], complex)
    r0 = np.array([0.3 + 0.1j, -0.2 + 0.4j, 0.9 - 0.3j])
    r, samples = orc.propagate(Ls, r0, dt, L_half=Lh, every=10)
    expect = np.exp(0.5 * ts[-1] ** 2) * r0
    assert np.linalg.norm(r - expect) / max(1.0, np.linalg.norm(expect)) <= 1e-6
    assert samples.shape[0] == 10 + 1  # steps 0,10,...,90 plus the final state
    # endpoint-average overload differs from the midpoint one only at O(dt^2)
    r2, _ = orc.propagate(Ls, r0, dt)
    assert np.linalg.norm(r2 - r) < 1e-3


def test_one_step_is_fourth_order_taylor(orc):
    rng = np.random.default_rng(1)
    D = 6
    L = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
    r0 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    h = 0.01
    r1 = orc.rk4_update(L, L, L, r0, h)
    P = np.eye(D) + h * L + (h * L) @ (h * L) / 2 + np.linalg.matrix_power(h * L, 3) / 6 + np.linalg.matrix_power(h * L, 4) / 24
    assert np.abs(r1 - P @ r0).max() < 1e-14 * np.abs(r0).max() * 10
    with pytest.raises(orc.OracleError):
        orc.rk4_update(L, L, L, r0, 0.0)


def tcl2_numpy(eps, A, p2f, grow):
    N = len(eps)
    I = np.eye(N)
    L = np.zeros((N * N, N * N), complex)
    Hls = np.zeros((N, N), complex)
    for b in range(len(grow)):
        Ab = np.where(p2f == b, A, 0)
        if not np.any(Ab):
            continue
        AdA = Ab.conj().T @ Ab
        g, S = 2 * grow[b].real, grow[b].imag
        Hls += S * AdA
        L += g * (np.kron(Ab.conj(), Ab) - 0.5 * np.kron(I, AdA) - 0.5 * np.kron(AdA.T, I))
    H = np.diag(eps) + 0.5 * (Hls + Hls.conj().T)
    return -1j * (np.kron(I, H) - np.kron(H.T, I)) + L


def test_tcl2_matches_kron_restatement_and_preserves_trace(orc):
    sx = np.array([[0, 1], [1, 0]], complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    S = orc.system_build(0.5 * sz + 0.5 * sx, [sz])
    rng = np.random.default_rng(0)
    grow = rng.standard_normal(S["nf"]) + 1j * rng.standard_normal(S["nf"])
    L = orc.tcl2_at(S["eps"], S["A_eig"], S["pair_to_freq"], grow)
    ref = tcl2_numpy(S["eps"], S["A_eig"][0], S["pair_to_freq"], grow)
    assert np.abs(L - ref).max() < 1e-14 * np.abs(ref).max() * 10
    N = 2
    diag_rows = [n + n * N for n in range(N)]
    assert np.abs(L[diag_rows].sum(axis=0)).max() < 1e-14 * np.abs(L).max() * 10
