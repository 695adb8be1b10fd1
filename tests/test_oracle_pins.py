"""Pins of the oracle by mathematics rather than by a second transcription.

* TCL4: the oracle's L4(t) (F/C/R -> MIKX -> NAKZWAN, tcl4_assemble.cpp:29-178)
  converges at first order in dt to the fourth-order TCL generator evaluated by
  brute force from its definition (tests/tcl4_bruteforce.py: ordered cumulants
  and Wick's theorem, Breuer & Petruccione sec. 9.2), which is itself checked
  against the exact dynamics of a spin coupled to one bosonic mode.  A sign
  planted in any one of the ten NAKZWAN terms breaks the convergence.
* TCL2: the drivers' generator (generator.cpp:81-177) equals L0 plus the
  secular part of the brute-force K2 when its Gamma table is read at the
  opposite frequency sign (the reference labels A(w) by eps_m - eps_n and
  reads Gamma(w) at that label; the L4 assembly reads the same table in the
  other convention, see test_reference_tcl2_relaxes_to_the_inverted_gibbs_state).
* Pure dephasing ([H, A] = 0): the Gaussian TCL expansion stops at second
  order, so L4 == 0 exactly.
* Bath correlation (bcf_fft_fun, correlation_fft.hpp:154-214): Ohmic closed
  forms at zero and finite temperature, to second order in the frequency step.
"""
import numpy as np
import pytest

from . import tcl4_bruteforce as BF

SZ = np.diag([1.0, -1.0]).astype(np.complex128)
SX = np.array([[0, 1], [1, 0]], np.complex128)
BATH = BF.ExpBath([0.3 + 0.1j, 0.2 - 0.05j], [1.0 + 2.0j, 2.5 - 0.7j])


def system(orc, N, seed=0):
    if N == 2:
        H, A = 0.5 * SZ + 0.5 * SX, SZ
    else:
        rng = np.random.default_rng(seed)
        X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        Y = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        H, A = (X + X.conj().T) / 2, (Y + Y.conj().T) / 4
    s = orc.system_build(H, [A])
    return s, np.asarray(s["eps"]), np.asarray(s["A_eig"][0])


def oracle_l4(orc, s, A, bath, dt, ts):
    tidx = [int(round(t / dt)) for t in ts]
    G = bath.gamma_table(s["omegas"], max(tidx) + 1, dt)
    return orc.l4_series(G, s["omegas"], s["mirror"], s["pair_to_freq"], [A], dt, tidx)


def test_bruteforce_tcl_matches_exact_spin_mode_dynamics():
    """K2 and K4 of the brute force equal the second and fourth order of the
    exact TCL generator Lambda'(t) Lambda(t)^-1 of a spin coupled to one
    bosonic mode in its ground state (a Gaussian bath with C(tau) = e^{-iW tau}),
    extracted from the exact dynamics at three coupling strengths."""
    from scipy.linalg import expm

    eps = np.array([-0.6, 0.6])
    th = 0.7
    A = np.array([[np.cos(th), np.sin(th)], [np.sin(th), -np.cos(th)]], np.complex128)
    W, nmax, t = 1.3, 14, 0.9
    b = np.diag(np.sqrt(np.arange(1, nmax)), 1).astype(np.complex128)
    HS, HB = np.diag(eps).astype(np.complex128), W * (b.conj().T @ b)
    d = 2 * nmax
    rhoB = np.zeros((nmax, nmax), np.complex128)
    rhoB[0, 0] = 1.0

    def generator(lam):
        Htot = np.kron(HS, np.eye(nmax)) + np.kron(np.eye(2), HB) + lam * np.kron(A, b + b.conj().T)
        Lf = -1j * (np.kron(np.eye(d), Htot) - np.kron(Htot.T, np.eye(d)))
        E = expm(Lf * t)
        Lam, dLam = np.zeros((4, 4), complex), np.zeros((4, 4), complex)
        trB = lambda R: np.einsum("iaja->ij", R.reshape(2, nmax, 2, nmax))  # noqa: E731
        for k in range(4):
            e = np.zeros(4, complex)
            e[k] = 1
            v = np.kron(e.reshape(2, 2, order="F"), rhoB).reshape(-1, order="F")
            x = E @ v
            Lam[:, k] = trB(x.reshape(d, d, order="F")).reshape(-1, order="F")
            dLam[:, k] = trB((Lf @ x).reshape(d, d, order="F")).reshape(-1, order="F")
        return dLam @ np.linalg.inv(Lam) - BF.unitary(eps)

    lams = [0.01, 0.02, 0.03]
    coef = np.linalg.solve(np.array([[l ** 2, l ** 4, l ** 6] for l in lams]),
                           np.stack([generator(l).ravel() for l in lams]))
    bath = BF.ExpBath([1.0], [1j * W])
    K2, K4 = BF.k2(bath, eps, A, t, n=60), BF.k4(bath, eps, A, t, n=24)
    assert np.linalg.norm(coef[0].reshape(4, 4) - K2) <= 1e-9 * np.linalg.norm(K2)
    assert np.linalg.norm(coef[1].reshape(4, 4) - K4) <= 1e-6 * np.linalg.norm(K4)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_oracle_l4_converges_to_bruteforce_tcl4(orc, N):
    """At dt = 0.01, 0.005, 0.0025 the gap to the brute-force K4 halves with dt
    (the F/C/R prefix sums and convolutions are first-order quadratures), and
    the Richardson combination 2 L4(dt/2) - L4(dt) is second-order close."""
    s, eps, A = system(orc, N, seed=N)
    ts = [0.4, 0.9]
    ref = [BF.k4(BATH, eps, A, t) for t in ts]
    L = {dt: oracle_l4(orc, s, A, BATH, dt, ts) for dt in (0.01, 0.005, 0.0025)}
    for q in range(len(ts)):
        nrm = np.linalg.norm(ref[q])
        e = [np.linalg.norm(L[dt][q] - ref[q]) / nrm for dt in (0.01, 0.005, 0.0025)]
        assert 1.9 < e[0] / e[1] < 2.1 and 1.9 < e[1] / e[2] < 2.1, e
        assert e[2] < 1e-2, e
        rich = np.linalg.norm(2 * L[0.0025][q] - L[0.005][q] - ref[q]) / nrm
        assert rich < 0.05 * e[2], (rich, e)


@pytest.mark.parametrize("term", range(1, 11))
def test_planted_nakzwan_sign_breaks_convergence(orc, term):
    """A sign planted in one NAKZWAN term leaves an O(1) gap to the brute force
    that does not shrink with dt."""
    s, eps, A = system(orc, 3, seed=3)
    ts = [0.9]
    ref = BF.k4(BATH, eps, A, ts[0])
    nrm = np.linalg.norm(ref)
    with orc.planted_mistake(term):
        e = [np.linalg.norm(oracle_l4(orc, s, A, BATH, dt, ts)[0] - ref) / nrm for dt in (0.005, 0.0025)]
    assert min(e) > 1e-2 and e[1] > 0.8 * e[0], (term, e)


@pytest.mark.parametrize("N", [2, 3])
def test_reference_tcl2_is_the_secular_bruteforce_tcl2(orc, N):
    """generator.cpp:81-177 with Gamma read at -omega_b equals L0 plus the
    secular part of the brute-force K2 to rounding; read at +omega_b (what the
    drivers pass) it does not."""
    s, eps, A = system(orc, N, seed=N)
    for t in (0.5, 1.2):
        ref = BF.unitary(eps) + BF.secular(BF.k2(BATH, eps, A, t), eps)
        flipped = np.array([BATH.gamma(-w, t) for w in s["omegas"]])
        got = orc.tcl2_at(eps, [A], s["pair_to_freq"], flipped)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
        natural = np.array([BATH.gamma(w, t) for w in s["omegas"]])
        assert np.linalg.norm(orc.tcl2_at(eps, [A], s["pair_to_freq"], natural) - ref) > 1e-3 * np.linalg.norm(ref)


def test_reference_tcl2_relaxes_to_the_inverted_gibbs_state(orc):
    """Consequence of the TCL2 frequency convention: the drivers' TCL2
    propagation (TCL4_spin_boson_example.cpp:284-384) of the spin-boson system
    relaxes the ground-state population to 1 - p_Gibbs; with the table read at
    -omega_b it relaxes to p_Gibbs.  The engine reproduces the reference."""
    s = orc.system_build(0.5 * SZ + 0.5 * SX, [SZ])
    eps = np.asarray(s["eps"])
    dt, steps = 0.01, 20000
    C = orc.bcf_fft(2 ** 15 - 1, dt, 1.0, 0, 0.05, 5.0)
    r0 = np.array([0, 0, 0, 1], np.complex128)
    p_gibbs = np.exp(-eps[0]) / np.exp(-eps).sum()
    for sign, expect in ((1.0, 1.0 - p_gibbs), (-1.0, p_gibbs)):
        G = orc.gamma_trapz(C, dt, sign * np.asarray(s["omegas"]))
        S = orc.tcl4_propagate(G, s["omegas"], s["mirror"], s["pair_to_freq"], eps, s["A_eig"], dt, steps, r0,
                               order=2)
        assert abs(S[-1, 0].real - expect) < 1e-6, (sign, S[-1, 0].real, expect)


def test_pure_dephasing_l4_vanishes(orc):
    """[H, A] = 0: TCL2 is exact for a Gaussian bath, so L4 == 0 at every t;
    a planted sign in any NAKZWAN term makes it nonzero."""
    s = orc.system_build(0.5 * 1.3 * SZ, [SZ])
    dt = 0.01
    C = orc.bcf_fft(4095, dt, 1.0, 0, 0.05, 5.0)
    G = orc.gamma_trapz(C, dt, s["omegas"])
    tidx = list(range(0, 2000, 37))
    args = (G, s["omegas"], s["mirror"], s["pair_to_freq"], s["A_eig"], dt, tidx)
    assert np.abs(orc.l4_series(*args)).max() == 0.0
    scale = np.abs(orc.l4_series(G, *orc_spin_boson_args(orc, C, dt, tidx))).max()
    for term in range(1, 11):
        with orc.planted_mistake(term):
            bad = np.abs(orc.l4_series(*args)).max()
        assert bad > 1e-3 * scale, term


def orc_spin_boson_args(orc, C, dt, tidx):
    s = orc.system_build(0.5 * SZ + 0.5 * SX, [SZ])
    return s["omegas"], s["mirror"], s["pair_to_freq"], s["A_eig"], dt, tidx


def ohmic_closed_form(t, alpha, wc, beta, nterms=200000):
    """C(t) = (1/2) int_0^inf alpha w e^{-w/wc} [coth(beta w/2) cos wt - i sin wt] dw
    = (alpha/2) [(1/wc + it)^-2 + 2 sum_{n>=1} Re (1/wc + n beta + it)^-2]
    (coth(x/2) = 1 + 2 sum e^{-nx}); the series tail by its integral."""
    z0 = 1.0 / wc + 1j * t
    out = 1.0 / z0 ** 2
    if np.isfinite(beta):
        n = np.arange(1, nterms + 1)[:, None]
        tail = 1.0 / (beta * (1.0 / wc + (nterms + 0.5) * beta + 1j * t))
        out = out + 2 * ((1.0 / (1.0 / wc + n * beta + 1j * t) ** 2).sum(0) + tail).real
    return 0.5 * alpha * out


@pytest.mark.parametrize("beta", [np.inf, 1.0, 0.3])
def test_bcf_matches_ohmic_closed_form(orc, beta):
    """The sampled KMS spectrum, its forward FFT and the dw/pi scale reproduce
    the Ohmic correlation function; the error falls by 4x per halving of the
    frequency step (trapezoid sum of a function that vanishes at w = 0)."""
    alpha, wc, dt = 0.05, 5.0, 0.005
    t = np.arange(400) * dt * 5
    exact = ohmic_closed_form(t, alpha, wc, beta)
    errs = []
    for N in (2 ** 14 - 1, 2 ** 15 - 1, 2 ** 16 - 1):
        C = orc.bcf_fft(N, dt, beta, 0, alpha, wc)
        errs.append(np.abs(C[: 400 * 5: 5] - exact).max() / np.abs(exact).max())
    assert errs[-1] < 3e-6, errs
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5, errs
