"""GPU parity of RK4 propagation and the TCL2+TCL4 driver composition."""
import numpy as np
import pytest
from scipy.linalg import expm

import taco_amd
from taco_amd import workloads as W

from .metrics import max_rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,steps", [(4, 50), (16, 200), (64, 30), (100, 12), (256, 6)])
def test_rk4_matches_oracle(orc, D, steps):
    rng = np.random.default_rng(D)
    base = (rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))) * 0.3
    Ls = np.array([base * np.cos(0.01 * s) for s in range(steps + 1)])
    r0 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    got_samples = []
    # D = 100 and 256 take the general-size path (generators staged in LDS, or
    # read from HBM in column blocks)
    every = 7 if steps > 7 else 2
    r = taco_amd.propagate_rk4_dense(Ls, r0, 0.0, 0.01, on_sample=lambda t, v: got_samples.append(v.copy()),
                                     sample_every=every)
    rr, ref_samples = orc.propagate(Ls, r0, 0.01, every=every)
    assert np.abs(r - rr).max() <= 1e-12 * np.abs(rr).max()
    assert len(got_samples) == len(ref_samples)
    assert max_rel_err(np.array(got_samples), ref_samples)[0] <= 1e-12


def test_rk4_closed_forms():
    D = 4
    L = np.diag([0.1 - 0.03j, -0.07 + 0.05j, 0.02, -0.04 - 0.02j])
    L[0, 2], L[2, 0], L[1, 3] = 0.03 + 0.01j, -0.02 + 0.02j, 0.01 - 0.04j
    r0 = np.array([1.0, 0.5 - 0.2j, -0.7 + 0.1j, 0.1 + 0.3j])
    r = taco_amd.propagate_rk4_dense(np.repeat(L[None], 501, 0), r0, 0.0, 0.0025)
    expect = expm(1.25 * L) @ r0
    assert np.linalg.norm(r - expect) / max(1, np.linalg.norm(expect)) <= 1e-6
    ts = np.arange(101) * 0.01
    Ls = np.array([t * np.eye(3) for t in ts], complex)
    Lh = np.array([(t + 0.005) * np.eye(3) for t in ts[:-1]], complex)
    r0 = np.array([0.3 + 0.1j, -0.2 + 0.4j, 0.9 - 0.3j])
    r = taco_amd.propagate_rk4_dense(Ls, r0, 0.0, 0.01, L_half_series=Lh)
    assert np.linalg.norm(r - np.exp(0.5) * r0) / np.linalg.norm(r0) <= 1e-6


@pytest.mark.parametrize("order", [0, 2, 4])
def test_tcl4_propagate_spin_boson(orc, order):
    w = W.spin_boson()
    S = w.system
    r0 = np.zeros(4, complex)
    r0[0] = 1.0
    got = taco_amd.tcl4_propagate(S, w.gamma, w.dt, w.steps, r0, order=order)
    ref = orc.tcl4_propagate(w.gamma, S.omegas, S.mirror, S.pair_to_freq, S.eps, S.A_eig, w.dt, w.steps, r0,
                             order=order)
    e, _ = max_rel_err(got, ref)
    print(f"order {order}: max per-t rho rel err {e:.3e}")
    assert e <= 1e-10
    tr = got[:, 0] + got[:, 3]
    assert np.abs(tr - 1).max() < 1e-12


def test_tcl4_propagate_spin_pair_degenerate(orc):
    w = W.spin_pair(Nt=4096, steps=400)
    S = w.system
    r0 = np.zeros(16, complex)
    r0[0] = 1.0
    got = taco_amd.tcl4_propagate(S, w.gamma, w.dt, w.steps, r0)
    ref = orc.tcl4_propagate(w.gamma, S.omegas, S.mirror, S.pair_to_freq, S.eps, S.A_eig, w.dt, w.steps, r0)
    assert max_rel_err(got, ref)[0] <= 1e-10


@pytest.mark.parametrize("case", ["gue3_two_ops", "spin_pair", "d8", "tfim_d16"])
def test_tcl2_bucket_lists_equal_all_bucket_walk(monkeypatch, case):
    """The TCL2 kernel's bucket-list path (each superoperator entry visits only
    the buckets with a nonzero term) against the all-bucket walk
    (TACO_TCL2_LISTS=0): the same values (exact zeros may differ in sign)."""
    if case == "gue3_two_ops":
        rng = W.SplitMix64(11)
        H = W.gue(3, rng)
        A1, A2 = W.random_hermitian_unit(3, rng), W.random_hermitian_unit(3, rng)
        w = W.Workload("g3", H, [A1, A2], taco_amd.DrudeLorentz(0.08, 4.0), 0.7, 0.01, 511, list(range(41)),
                       steps=40)
        steps = 40
    elif case == "spin_pair":
        w, steps = W.spin_pair(Nt=1024, steps=60), 60
    elif case == "d8":
        rng = W.SplitMix64(5)
        H = W.gue(8, rng)
        A = W.random_hermitian_unit(8, rng)
        w = W.Workload("g8", H, [A], taco_amd.Ohmic(0.1, 5.0), 1.0, 0.01, 255, list(range(21)), steps=20)
        steps = 20
    else:
        w, steps = W.tfim_workload(0.7, lam=0.2, dt=0.05, Nt=64, tmax=6), 6
    S, G = w.system, w.gamma
    r0 = np.zeros(S.N ** 2, complex)
    r0[0] = 1.0
    out = []
    for lists in ("1", "0"):
        monkeypatch.setenv("TACO_TCL2_LISTS", lists)
        out.append(taco_amd.tcl4_propagate(S, G, w.dt, steps, r0, order=2))
    assert np.array_equal(out[0], out[1])


def _two_channel_system(seed=7):
    rng = np.random.default_rng(seed)
    N = 3
    X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    ops = []
    for _ in range(2):
        Y = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        ops.append((Y + Y.conj().T) / 4)
    return taco_amd.System.build((X + X.conj().T) / 2, ops), rng


def test_tcl2_channel_matrices_match_oracle(orc):
    """build_tcl2_components in its general form (generator.cpp:81-177): two
    non-commuting coupling operators and full complex channel matrices
    Gamma_b(alpha, beta) with off-diagonal entries, at several times."""
    S, rng = _two_channel_system()
    Nt, nf = 40, S.nf
    G = rng.normal(size=(Nt, nf, 2, 2)) + 1j * rng.normal(size=(Nt, nf, 2, 2))
    G[:, 1, 0, 1] = 0.0          # exact zeros are skipped, as in the reference
    G[:, 2, 1, 1] = 0.3          # real: no Lamb-shift part
    tidx = [0, 3, 17, 39]
    got = taco_amd.build_tcl2_generator(S, G, tidx)
    for q, t in enumerate(tidx):
        ref = orc.tcl2_general_at(S.eps, S.A_eig, S.pair_to_freq, G[t])
        assert np.abs(got[q] - ref).max() <= 1e-13 * np.abs(ref).max(), t


def test_tcl2_diagonal_channels_match_the_drivers_generator(orc):
    """Diagonal channel matrices Gamma_b(alpha, beta) = delta_ab gamma_b are the
    drivers' case: equal to the oracle's tcl2_at on the spin-boson Gamma table."""
    w = W.spin_boson(Nt=256, tmax=20)
    S = w.system
    G = np.zeros((256, S.nf, 1, 1), complex)
    G[:, :, 0, 0] = w.gamma
    got = taco_amd.build_tcl2_generator(S, G, [5, 100])
    for q, t in enumerate([5, 100]):
        ref = orc.tcl2_at(S.eps, S.A_eig, S.pair_to_freq, w.gamma[t])
        assert np.abs(got[q] - ref).max() <= 1e-14 * np.abs(ref).max()


def test_tcl2_skips_channels_whose_adag_a_vanishes(orc):
    """The reference skips a (b, alpha, beta) whose A_alpha,b^dag A_beta,b is
    the zero matrix, even where kron(conj(A_alpha,b), A_beta,b) is not: with
    equally spaced levels the pairs (0,1) and (1,2) share the bucket -1, and
    operators on either pair meet no common row q."""
    H = np.diag([0.0, 1.0, 2.0]).astype(complex)
    A1 = np.zeros((3, 3), complex)
    A2 = np.zeros((3, 3), complex)
    A1[0, 1] = A1[1, 0] = 1.0
    A2[1, 2] = A2[2, 1] = 0.5
    S = taco_amd.System.build(H, [A1, A2])
    p2f = np.asarray(S.pair_to_freq)
    b = p2f[0, 1]
    assert p2f[1, 2] == b
    P1, P2 = (np.where(p2f == b, np.asarray(a), 0) for a in S.A_eig)
    assert not np.any(P1.conj().T @ P2) and np.any(np.kron(P1.conj(), P2))
    rng = np.random.default_rng(3)
    G = rng.normal(size=(8, S.nf, 2, 2)) + 1j * rng.normal(size=(8, S.nf, 2, 2))
    got = taco_amd.build_tcl2_generator(S, G, [2, 7])
    for q, t in enumerate([2, 7]):
        ref = orc.tcl2_general_at(S.eps, S.A_eig, S.pair_to_freq, G[t])
        assert np.abs(got[q] - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("D", [100, 256])
def test_large_d_midpoints_equal_explicit_half_series(D):
    """At large D the forward RK4 reads midpoint generators formed once on
    the device; they are the values (L_n + L_{n+1})/2 the products form on
    the fly, so the run equals the L_half_series overload given the same
    midpoints, bit for bit."""
    rng = np.random.default_rng(D)
    steps = 12
    L = (rng.standard_normal((steps + 1, D, D)) + 1j * rng.standard_normal((steps + 1, D, D))) * (0.5 / np.sqrt(D))
    r0 = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    a = taco_amd.propagate_rk4_dense(L, r0, 0.0, 0.01)
    b = taco_amd.propagate_rk4_dense(L, r0, 0.0, 0.01, L_half_series=(L[:-1] + L[1:]) * 0.5)
    assert np.array_equal(a, b)
