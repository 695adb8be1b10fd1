"""Pins of the oracle's F/C/R series and triple kernels (CPU only).

Reference: cpp/src/tcl/tcl4_kernels.cpp:105-281 and cpp/src/tcl/tcl4.cpp:103-183.
"""
import numpy as np
import pytest


def phases(n, omega, dt):
    tk = np.arange(n) * dt
    return np.exp(1j * omega * tk), np.exp(-1j * omega * tk)


def direct_numpy(g1, g2, omega, dt, which):
    """tcl4_kernels.m discretised as the reference's Direct method (O(N^2))."""
    n = len(g1)
    pp, pm = phases(n, omega, dt)
    out = np.zeros(n, complex)
    if which == "R":
        for k in range(n):
            out[k] = g1[k] * dt * np.sum(g2[: k + 1] * pm[: k + 1]) - dt * np.sum(g1[: k + 1] * g2[: k + 1] * pm[: k + 1])
        return out
    b = g2 if which == "F" else np.conj(g2)
    for k in range(n):
        m = np.arange(k + 1)
        out[k] = g1[k] * pm[k] * dt * np.sum(b[m] * pp[m]) - dt * np.sum(g1[k - m] * pm[k - m] * b[m])
    return out


@pytest.fixture(scope="module")
def sb_gamma(orc):
    # spin-boson bath of the reference tests: J = w exp(-w/10), beta 0.5
    dt = 0.000625
    C = orc.bcf_fft(2047, dt, 0.5, 0, 1.0, 10.0)
    omegas = np.array([-1.0, 0.0, 1.0])
    return orc.gamma_trapz(C, dt, omegas), omegas, dt


@pytest.mark.parametrize("which", ["F", "C", "R"])
def test_direct_matches_numpy_quadrature(orc, sb_gamma, which):
    G, om, dt = sb_gamma
    g1, g2 = G[:300, 0], G[:300, 2]
    omega = om[0] + om[2] + om[1]
    ref = direct_numpy(g1, g2, omega, dt, which)
    got = orc.fcr_series(g1, g2, omega, dt, which, "direct")
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("which", ["F", "C", "R"])
def test_convolution_matches_direct(orc, sb_gamma, which):
    # the reference's own cross-check (tests/tcl4_tests.cpp:188-212)
    G, om, dt = sb_gamma
    for (i, j, k) in [(0, 1, 2), (2, 2, 0), (1, 1, 1)]:
        g1, g2 = G[:, i], G[:, j]
        omega = om[i] + om[j] + om[k]
        conv = orc.fcr_series(g1, g2, omega, dt, which, "conv")
        dire = orc.fcr_series(g1, g2, omega, dt, which, "direct")
        rel = np.abs(conv - dire).max() / np.abs(dire).max()
        assert rel < 1e-9, (which, i, j, k, rel)


def test_constant_gamma_gives_zero_kernels(orc):
    # [Gamma(t) - Gamma(t - s)] vanishes for constant Gamma, so F = C = R = 0.
    n = 257
    g = np.full(n, 0.3 - 0.2j)
    for which in "FCR":
        for method in ("conv", "direct"):
            out = orc.fcr_series(g, g, 1.7, 0.01, which, method)
            assert np.abs(out).max() < 1e-13


def test_r_at_zero_is_exactly_zero(orc, sb_gamma):
    G, om, dt = sb_gamma
    assert G[0, 0] == 0
    out = orc.fcr_series(G[:, 0], G[:, 2], 0.5, dt, "R", "conv")
    assert out[0] == 0


def test_triple_kernels_per_triple_and_slicing(orc, sb_gamma):
    G, om, dt = sb_gamma
    G = G[:600]
    mirror = orc.mirror(om)
    assert list(mirror) == [2, 1, 0]
    F, C, R = orc.triple_kernels(G, om, mirror, dt, nthreads=2)
    nf = len(om)
    for (i, j, k) in [(0, 0, 0), (0, 2, 1), (2, 1, 0)]:
        omega = (om[i] + om[j]) + om[k]
        assert np.array_equal(F[i, j, k], orc.fcr_series(G[:, i], G[:, mirror[j]], omega, dt, "F"))
        assert np.array_equal(C[i, j, k], orc.fcr_series(G[:, i], G[:, j], omega, dt, "C"))
        assert np.array_equal(R[i, j, k], orc.fcr_series(G[:, i], G[:, j], omega, dt, "R"))
    # keep < Nt: same arithmetic (FFT length from Nt), so identical heads
    Fk, Ck, Rk = orc.triple_kernels(G, om, mirror, dt, keep=123, nthreads=3)
    assert np.array_equal(Fk, F[..., :123]) and np.array_equal(Ck, C[..., :123]) and np.array_equal(Rk, R[..., :123])
    assert F.shape == (nf, nf, nf, 600)
