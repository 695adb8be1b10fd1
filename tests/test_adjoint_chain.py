"""CPU checks of the differentiable parameter chain in front of the engine
(taco_amd/adjoint.py): bath correlation C(t), trapz Gamma, eigensystem and
bucket frequencies.  Each backward is compared with central finite
differences of the oracle's own forward (bcf_fft, gamma_trapz, system_build),
on random linear functionals, so a dropped term, a wrong sign or a transposed
index fails."""
import numpy as np
import pytest
import torch

from taco_amd import adjoint as AD
from taco_amd import workloads as W


def fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


@pytest.mark.parametrize("kind,p0,p1,beta", [(0, 0.05, 5.0, 1.0), (0, 0.1, 10.0, 0.5), (1, 0.1, 7.0, 2.0)])
def test_bcf_backward_matches_oracle_fd(orc, kind, p0, p1, beta):
    N, dt = 1023, 0.01
    rng = np.random.default_rng(1)
    Wt = rng.standard_normal(N + 1) + 1j * rng.standard_normal(N + 1)

    def lin(C):
        return float(np.real(np.vdot(Wt, C)))  # Re sum conj(W) C

    p = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (p0, p1, beta)]
    grads = torch.autograd.grad(AD.BcfFun.apply(p[0], p[1], p[2], (kind, N, dt)), p,
                                grad_outputs=torch.from_numpy(Wt))
    base = [p0, p1, beta]
    for q in range(3):
        def f(v, q=q):
            args = list(base)
            args[q] = v
            return lin(orc.bcf_fft(N, dt, args[2], kind, args[0], args[1]))
        ref = fd(f, base[q], 1e-6 * abs(base[q]))
        got = float(grads[q])
        assert abs(got - ref) <= 1e-6 * max(1.0, abs(ref)), (q, got, ref)


def test_gamma_backward_matches_oracle_fd(orc):
    rng = np.random.default_rng(2)
    Nt, dt = 700, 0.01
    C = rng.standard_normal(Nt) + 1j * rng.standard_normal(Nt)
    om = np.array([-1.3, 0.0, 0.4, 2.2])
    Wg = rng.standard_normal((Nt, 4)) + 1j * rng.standard_normal((Nt, 4))
    Ct = torch.from_numpy(C).requires_grad_(True)
    omt = torch.from_numpy(om).requires_grad_(True)
    G = AD.GammaFun.apply(Ct, omt, dt)
    gC, gw = torch.autograd.grad(G, (Ct, omt), grad_outputs=torch.from_numpy(Wg))
    lin = lambda G_: float(np.real(np.vdot(Wg, G_)))
    # directional derivative along a random complex direction of C
    dC = rng.standard_normal(Nt) + 1j * rng.standard_normal(Nt)
    ref = fd(lambda s: lin(orc.gamma_trapz(C + s * dC, dt, om)), 0.0, 1e-6)
    got = float(np.real(np.vdot(gC.numpy(), dC)))
    assert abs(got - ref) <= 1e-7 * max(1.0, abs(ref))
    for b in range(4):
        def f(v, b=b):
            o = om.copy()
            o[b] = v
            return lin(orc.gamma_trapz(C, dt, o))
        ref = fd(f, om[b], 1e-6)
        assert abs(float(gw[b]) - ref) <= 1e-6 * max(1.0, abs(ref)), b


@pytest.mark.parametrize("degenerate", [False, True])
def test_eigh_backward_matches_fd(orc, degenerate):
    """A basis-independent function of the eigensystem (sum_k f(e_k) U P_k U^H
    with a fixed lab matrix) differentiated through EighFun; degenerate = the
    spin pair (eigenvalue 0 twice), where only such functions are defined."""
    rng = np.random.default_rng(3)
    if degenerate:
        w = W.spin_pair(Nt=64, steps=10)
        H0 = w.H
    else:
        H0 = W.gue(4, W.SplitMix64(5))
    dH = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    dH = 0.5 * (dH + dH.conj().T)
    if degenerate:  # a perturbation inside the model family keeps the degeneracy
        sz, sx, i2 = AD.SZ, AD.SX, AD.I2
        h1 = 0.3 * sz + 0.7 * sx
        dH = np.kron(h1, i2) + np.kron(i2, h1)
    M = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    A = W.random_hermitian_unit(4, W.SplitMix64(9))

    def value(e, U):
        # sum_k sin(e_k) (U e_k e_k^H U^H) : M  plus a lab-frame operator U diag(e) U^H : M
        P = U @ torch.diag(torch.sin(e).to(torch.complex128)) @ U.conj().T
        Q = U @ torch.diag(e.to(torch.complex128)) @ U.conj().T
        Mt = torch.from_numpy(M)
        return torch.real(torch.sum(torch.conj(Mt) * (P + Q)))

    H = torch.from_numpy(H0.copy()).requires_grad_(True)
    e, U = AD.EighFun.apply(H, [A], 1e-9)
    (gH,) = torch.autograd.grad(value(e, U), H)
    got = float(np.real(np.vdot(gH.numpy(), dH)))

    def f(s):
        ev, Uv = AD.EighFun.apply(torch.from_numpy(H0 + s * dH), [A], 1e-9)
        return float(value(ev, Uv))

    ref = fd(f, 0.0, 1e-6)
    assert abs(got - ref) <= 1e-7 * max(1.0, abs(ref)), (got, ref)


def test_bucket_frequencies_follow_eigenvalue_differences():
    """omegas[p2f[m, n]] = e_m - e_n, and the chain H -> (e, U) -> bucket
    frequencies differentiates like finite differences of System.build."""
    from taco_amd import System
    w = W.gue_pair(N_bath=1000, tmax=10)
    S = w.system
    for m in range(S.N):
        for n in range(S.N):
            assert abs(S.omegas[S.pair_to_freq[m, n]] - (S.eps[m] - S.eps[n])) < 1e-8
    rng = np.random.default_rng(4)
    dH = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    dH = 0.5 * (dH + dH.conj().T)
    wb = np.arange(S.nf, dtype=float) + 1.0
    H = torch.from_numpy(w.H.copy()).requires_grad_(True)
    e, U = AD.EighFun.apply(H, w.ops, 1e-9)
    om = AD.BucketOmegas.apply(e, S.omegas, S.pair_to_freq)
    (gH,) = torch.autograd.grad((om * torch.from_numpy(wb)).sum(), H)
    got = float(np.real(np.vdot(gH.numpy(), dH)))
    f = lambda s: float(wb @ System.build(w.H + s * dH, w.ops, 1e-9).omegas)
    ref = fd(f, 0.0, 1e-6)
    assert abs(got - ref) <= 1e-7 * max(1.0, abs(ref)), (got, ref)
