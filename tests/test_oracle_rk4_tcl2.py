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


def _two_channel_system(orc, seed=7):
    rng = np.random.default_rng(seed)
    N = 3
    X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    H = (X + X.conj().T) / 2
    ops = []
    for _ in range(2):
        Y = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        ops.append((Y + Y.conj().T) / 4)
    return orc.system_build(H, ops), rng


def test_tcl2_channels_single_entry_is_the_lindblad_form(orc):
    """With one channel entry Gamma_b(alpha, beta) = g (real) and all others
    zero, the generator minus L0 acts on rho as
    g (A_beta,b rho A_alpha,b^dag - 1/2 {A_alpha,b^dag A_beta,b, rho}) (the
    master-equation form of generator.cpp:92-162, evaluated with matrix
    products, not Kronecker products): pins the channel indices, the operand
    order and the conjugation."""
    s, rng = _two_channel_system(orc)
    eps, N, nf = np.asarray(s["eps"]), 3, s["nf"]
    ops = [np.asarray(a) for a in s["A_eig"]]
    p2f = np.asarray(s["pair_to_freq"])
    L0 = orc.tcl2_general_at(eps, ops, p2f, np.zeros((nf, 2, 2), complex))
    rho = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    for b in range(nf):
        parts = [np.where(p2f == b, a, 0) for a in ops]
        for al in range(2):
            for be in range(2):
                if not np.any(parts[al].conj().T @ parts[be]):
                    continue
                G = np.zeros((nf, 2, 2), complex)
                G[b, al, be] = 0.5
                L = orc.tcl2_general_at(eps, ops, p2f, G) - L0
                got = (L @ rho.reshape(-1, order="F")).reshape(N, N, order="F")
                Aa, Ab = parts[al], parts[be]
                AdA = Aa.conj().T @ Ab
                ref = 1.0 * (Ab @ rho @ Aa.conj().T - 0.5 * (AdA @ rho + rho @ AdA))
                assert np.abs(got - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max()), (b, al, be)


def test_tcl2_channels_of_a_correlated_pair_equal_one_channel(orc):
    """B_2 = c B_1 (real c): Gamma_b(alpha, beta) = Gamma_b c_alpha c_beta with
    operators A_1, A_2 equals one channel with A_1 + c A_2 and Gamma_b."""
    s, rng = _two_channel_system(orc, seed=8)
    eps, nf = np.asarray(s["eps"]), s["nf"]
    ops = [np.asarray(a) for a in s["A_eig"]]
    p2f = np.asarray(s["pair_to_freq"])
    c = np.array([1.0, -0.7])
    gam = rng.normal(size=nf) + 1j * rng.normal(size=nf)
    G = gam[:, None, None] * np.outer(c, c)[None]
    two = orc.tcl2_general_at(eps, ops, p2f, G)
    one = orc.tcl2_at(eps, [ops[0] + c[1] * ops[1]], p2f, gam)
    assert np.abs(two - one).max() <= 1e-13 * np.abs(one).max()
