"""GPU checks of the hand-written reverse pass (taco_tape_* of the C ABI) and
of the parameter gradients built on it.

* The tape's forward equals tcl4_propagate bit for bit.
* The vector-Jacobian product of r_0..r_steps with respect to every engine
  input (Gamma table, bucket frequencies, A_eig, eigenvalues, r0) matches
  central finite differences of the oracle's forward along random
  directions: a dropped term, a sign or an index error in any of the RK4,
  L4, F/C/R or TCL2 adjoints shows up here.
* The gradient of the spin-pair loss with respect to (alpha, omega_c, beta,
  eps, Delta) matches central finite differences (h = 1e-6 relative) of the
  oracle pipeline (system build, bath correlation, trapz Gamma, TCL4 run).

Finite differences limit the comparison: the bar is 1e-6 relative (the
printed values show the actual agreement)."""
import numpy as np
import pytest
import torch

import taco_amd
from taco_amd import _lib
from taco_amd import adjoint as AD
from taco_amd import workloads as W

pytestmark = pytest.mark.gpu
FD_TOL = 1e-6


def r0_for(N):
    r = np.zeros(N * N, complex)
    r[0] = 1.0
    return r


def cases():
    out = []
    sb = W.spin_boson(Nt=1024, tmax=200, tstep=1)
    out.append(("spin_boson", sb.system, sb.gamma, sb.dt, 200, 64))
    sp = W.spin_pair(Nt=2048, steps=300)
    out.append(("spin_pair", sp.system, sp.gamma, sp.dt, 300, 64))
    rng = W.SplitMix64(11)
    H = W.gue(3, rng)
    A1, A2 = W.random_hermitian_unit(3, rng), W.random_hermitian_unit(3, rng)
    g2 = W.Workload("gue3_two_ops", H, [A1, A2], taco_amd.DrudeLorentz(0.08, 4.0), 0.7, 0.01, 511,
                    list(range(151)), steps=150)
    out.append(("gue3_two_ops", g2.system, g2.gamma, g2.dt, 150, 32))
    return out


CASES = cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_tape_forward_equals_tcl4_propagate(case):
    name, S, G, dt, steps, K = case
    tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(S.N), 4, K)
    ref = taco_amd.tcl4_propagate(S, G, dt, steps, r0_for(S.N))
    assert np.array_equal(tp.samples, ref)
    tp.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_vjp_matches_oracle_finite_differences(orc, case):
    name, S, G, dt, steps, K = case
    N = S.N
    rng = np.random.default_rng(5)
    Wt = rng.standard_normal((steps + 1, N * N)) + 1j * rng.standard_normal((steps + 1, N * N))
    tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
    g = tp.adjoint(Wt)
    tp.close()

    def loss(Gx=G, om=S.omegas, ops=S.A_eig, eps=S.eps, r0=r0_for(N)):
        s = orc.tcl4_propagate(Gx, om, S.mirror, S.pair_to_freq, eps, ops, dt, steps, r0)
        return float(np.real(np.vdot(Wt, s)))

    l0 = abs(loss())

    def check(label, got, f, h):
        # central differences: truncation O(h^2), rounding ~1e-14 |loss| / h
        ref = (f(h) - f(-h)) / (2 * h)
        tol = FD_TOL * abs(ref) + 1e-14 * l0 / h
        print(f"{name} {label}: vjp {got:.12e} fd {ref:.12e} diff {abs(got - ref):.2e} tol {tol:.2e}")
        assert abs(got - ref) <= tol, label

    dG = (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)) * np.abs(G).max()
    check("gamma", float(np.real(np.vdot(g["gamma"], dG))), lambda s: loss(Gx=G + s * dG), 1e-5)
    dw = rng.standard_normal(S.nf)
    check("omegas", float(g["omegas"] @ dw), lambda s: loss(om=S.omegas + s * dw), 1e-4)
    dA = [rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) for _ in S.A_eig]
    got = sum(float(np.real(np.vdot(a, d))) for a, d in zip(g["ops"], dA))
    check("ops", got, lambda s: loss(ops=[A + s * d for A, d in zip(S.A_eig, dA)]), 1e-5)
    de = rng.standard_normal(N)
    check("eps", float(g["eps"] @ de), lambda s: loss(eps=S.eps + s * de), 1e-4)
    dr = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    check("r0", float(np.real(np.vdot(g["r0"], dr))), lambda s: loss(r0=r0_for(N) + s * dr), 1e-2)


@pytest.mark.parametrize("case", [1, 2], ids=["spin_pair", "d3_two_ops"])
def test_tape_adjoint_repeatable(case):
    """The reverse pass has no atomics: every cotangent entry is summed by one
    thread in a fixed order, so two reverses (and a fresh tape) agree bit for
    bit, in every input."""
    name, S, G, dt, steps, K = CASES[case]
    N = S.N
    Wt = np.ones((steps + 1, N * N), complex)
    tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
    a, b = tp.adjoint(Wt), tp.adjoint(Wt)
    tp.close()
    tp2 = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
    d = tp2.adjoint(Wt)
    tp2.close()
    for key in a:
        assert np.array_equal(a[key], b[key]) and np.array_equal(a[key], d[key]), key
    if case != 1:
        return
    # the reverse of the checkpointed run equals the reverse with every state kept
    tp1 = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, 1)
    c = tp1.adjoint(Wt)
    tp1.close()
    for key in ("gamma", "omegas", "eps", "r0"):
        assert np.allclose(a[key], c[key], rtol=1e-12, atol=1e-14 * np.abs(a[key]).max()), key


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_factorised_l4_adjoint_matches_term_walk(monkeypatch, case):
    # the default (factorised) L4 reverse against the O(N^6) term walk kept
    # behind TACO_L4_ADJ_KERNEL=direct: same cotangents up to summation order
    name, S, G, dt, steps, K = case
    N = S.N
    rng = np.random.default_rng(9)
    Wt = rng.standard_normal((steps + 1, N * N)) + 1j * rng.standard_normal((steps + 1, N * N))
    out = []
    for kernel in ("factorised", "direct"):
        if kernel == "direct":
            monkeypatch.setenv("TACO_L4_ADJ_KERNEL", "direct")
        tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
        out.append(tp.adjoint(Wt))
        tp.close()
    a, b = out
    for key in ("gamma", "omegas", "eps", "r0"):
        scale = np.abs(b[key]).max()
        err = np.abs(a[key] - b[key]).max() / scale
        print(f"{name} {key}: max diff / max {err:.2e}")
        assert err <= 1e-12, key
    for x, y in zip(a["ops"], b["ops"]):
        assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()


def oracle_spin_pair_rho_lab(orc, prob, p):
    alpha, wc, beta, eps, delta = p
    H1 = 0.5 * eps * AD.SZ + 0.5 * delta * AD.SX
    H = np.kron(H1, AD.I2) + np.kron(AD.I2, H1)
    A = prob.coupling()
    sysd = orc.system_build(H, [A], prob.freq_tol)
    C = orc.bcf_fft(prob.Nt - 1, prob.dt, beta, 0, alpha, wc)
    G = orc.gamma_trapz(C, prob.dt, sysd["omegas"])
    r0 = r0_for(sysd["N"])
    s = orc.tcl4_propagate(G, sysd["omegas"], sysd["mirror"], sysd["pair_to_freq"], sysd["eps"], sysd["A_eig"],
                           prob.dt, prob.steps, r0, order=prob.order)
    N = sysd["N"]
    rho = s.reshape(-1, N, N).transpose(0, 2, 1)
    U = sysd["U"]
    return U @ rho @ U.conj().T


def test_spin_pair_parameter_gradient_matches_oracle_fd(orc):
    prob = AD.SpinPairProblem(Nt=2 ** 12, steps=400, checkpoint_every=64)
    p0 = np.array([0.05, 5.0, 1.0, 1.0, 1.0])
    u = np.where(np.random.default_rng(4).uniform(size=5) < 0.5, -1.0, 1.0)
    target = oracle_spin_pair_rho_lab(orc, prob, p0 * (1 + 0.05 * u))
    params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    loss = prob.loss(params, torch.from_numpy(target))
    (grad,) = torch.autograd.grad(loss, params)

    def oloss(p):
        d = oracle_spin_pair_rho_lab(orc, prob, p) - target
        return float(np.sum(d.real ** 2 + d.imag ** 2))

    l0 = oloss(p0)
    assert abs(float(loss) - l0) <= 1e-10 * l0
    for q, name in enumerate(prob.names):
        h = 1e-6 * abs(p0[q])
        pp, pm = p0.copy(), p0.copy()
        pp[q] += h
        pm[q] -= h
        ref = (oloss(pp) - oloss(pm)) / (2 * h)
        got = float(grad[q])
        print(f"d loss / d {name}: adjoint {got:.12e} fd {ref:.12e} rel {abs(got - ref) / abs(ref):.2e}")
        assert abs(got - ref) <= FD_TOL * abs(ref), name


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_l4_adjoint_tbar_in_scratch_matches_lds(monkeypatch, case):
    # the placement used above N = 8 (the N^4 Tbar tile in global scratch,
    # forced here with TACO_L4_ADJ_TB=global) against the LDS placement
    name, S, G, dt, steps, K = case
    N = S.N
    rng = np.random.default_rng(13)
    Wt = rng.standard_normal((steps + 1, N * N)) + 1j * rng.standard_normal((steps + 1, N * N))
    out = []
    for where in ("lds", "global"):
        if where == "global":
            monkeypatch.setenv("TACO_L4_ADJ_TB", "global")
        tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
        out.append(tp.adjoint(Wt))
        tp.close()
    a, b = out
    for key in ("gamma", "omegas", "eps", "r0"):
        err = np.abs(a[key] - b[key]).max() / np.abs(b[key]).max()
        print(f"{name} {key}: max diff / max {err:.2e}")
        assert err <= 1e-12, key
    for x, y in zip(a["ops"], b["ops"]):
        assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max()


# d = 16: the 4-spin transverse-field Ising chain (nf = 81), a short window
# with a coarse step and a strong bath so that rho moves appreciably
D16 = W.tfim_workload(0.7, lam=0.2, dt=0.05, Nt=64, tmax=6)
D16_STEPS, D16_K = 6, 4


def test_d16_tape_forward_matches_oracle(orc):
    S, G, dt = D16.system, D16.gamma, D16.dt
    assert S.N == 16 and S.nf == 81
    tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, D16_STEPS, r0_for(16), 4, D16_K)
    got = tp.samples.copy()
    tp.close()
    ref = orc.tcl4_propagate(G, S.omegas, S.mirror, S.pair_to_freq, S.eps, S.A_eig, dt, D16_STEPS, r0_for(16))
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    moved = np.linalg.norm(ref - ref[0], axis=1).max() / np.linalg.norm(ref[0])
    print(f"d16 rho per-step rel err max {err.max():.2e}; rho moves by {moved:.2e}")
    assert moved > 1e-3
    assert err.max() <= 1e-10


def test_d16_vjp_matches_fd_of_forward():
    """The d = 16 reverse (Tbar in scratch, slice cotangents in global memory,
    the general RK4 reverse for D = 256) against central differences of the
    GPU forward, which the test above pins to the oracle."""
    S, G, dt = D16.system, D16.gamma, D16.dt
    N, steps, K = 16, D16_STEPS, D16_K
    rng = np.random.default_rng(21)
    Wt = rng.standard_normal((steps + 1, N * N)) + 1j * rng.standard_normal((steps + 1, N * N))
    tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
    g = tp.adjoint(Wt)
    tp.close()

    def loss(Gx=G, om=S.omegas, ops=S.A_eig, eps=S.eps, r0=r0_for(N)):
        t = AD.Tape(Gx, om, S.mirror, S.pair_to_freq, ops, eps, dt, steps, r0, 4, K)
        s = t.samples.copy()
        t.close()
        return float(np.real(np.vdot(Wt, s)))

    l0 = abs(loss())

    def check(label, got, f, h):
        ref = (f(h) - f(-h)) / (2 * h)
        tol = FD_TOL * abs(ref) + 1e-14 * l0 / h
        print(f"d16 {label}: vjp {got:.12e} fd {ref:.12e} diff {abs(got - ref):.2e} tol {tol:.2e}")
        assert abs(got - ref) <= tol, label

    dG = (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)) * np.abs(G).max()
    check("gamma", float(np.real(np.vdot(g["gamma"], dG))), lambda s: loss(Gx=G + s * dG), 1e-5)
    dw = rng.standard_normal(S.nf)
    check("omegas", float(g["omegas"] @ dw), lambda s: loss(om=S.omegas + s * dw), 1e-4)
    dA = [rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) for _ in S.A_eig]
    got = sum(float(np.real(np.vdot(a, d))) for a, d in zip(g["ops"], dA))
    check("ops", got, lambda s: loss(ops=[A + s * d for A, d in zip(S.A_eig, dA)]), 1e-5)
    de = rng.standard_normal(N)
    check("eps", float(g["eps"] @ de), lambda s: loss(eps=S.eps + s * de), 1e-4)
    dr = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    check("r0", float(np.real(np.vdot(g["r0"], dr))), lambda s: loss(r0=r0_for(N) + s * dr), 1e-2)


def test_tfim_parameter_gradient_matches_fd():
    """Gradient of the d = 16 chain loss with respect to (lambda, gamma, beta,
    J, h) through the whole chain (Drude-Lorentz bath correlation, trapz
    Gamma, eigensystem, TCL4 run and its reverse) against central differences
    (h = 1e-6 relative) of the same pipeline's forward."""
    prob = AD.TfimChainProblem(Nt=256, dt=0.02, steps=12, checkpoint_every=4)
    p0 = np.array([0.1, 5.0, 1.0, 1.0, 0.7])
    u = np.where(np.random.default_rng(4).uniform(size=5) < 0.5, -1.0, 1.0)
    with torch.no_grad():
        target = prob.rho_lab(torch.tensor(p0 * (1 + 0.05 * u), dtype=torch.float64)).detach()
    params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    loss = prob.loss(params, target)
    (grad,) = torch.autograd.grad(loss, params)

    def floss(p):
        with torch.no_grad():
            return float(prob.loss(torch.tensor(p, dtype=torch.float64), target))

    l0 = float(loss)
    for q, name in enumerate(prob.names):
        h = 1e-6 * abs(p0[q])
        pp, pm = p0.copy(), p0.copy()
        pp[q] += h
        pm[q] -= h
        ref = (floss(pp) - floss(pm)) / (2 * h)
        got = float(grad[q])
        tol = FD_TOL * abs(ref) + 1e-14 * l0 / h
        print(f"d loss / d {name}: adjoint {got:.12e} fd {ref:.12e} rel {abs(got - ref) / abs(ref):.2e}")
        assert abs(got - ref) <= tol, name


def _mt_sets(kind):
    """Parameter sets of one model size for the multi-set tape."""
    if kind == "spin_pair_d4":
        ws = [W.spin_pair(eps=e, delta=d, alpha=a, Nt=1024, steps=120)
              for e, d, a in ((1.0, 1.0, 0.05), (0.8, 1.3, 0.08), (1.2, 0.6, 0.03))]
        return [dict(gamma=w.gamma, omegas=w.system.omegas, mirror=w.system.mirror,
                     pair_to_freq=w.system.pair_to_freq, ops=w.system.A_eig, eps=w.system.eps, dt=w.dt,
                     steps=120, r0=r0_for(4)) for w in ws], 32
    if kind == "gue3_mixed_order":
        out = []
        for k, order in enumerate((4, 2, 4)):
            rng = W.SplitMix64(40 + k)
            H = W.gue(3, rng)
            A1, A2 = W.random_hermitian_unit(3, rng), W.random_hermitian_unit(3, rng)
            w = W.Workload("g3", H, [A1, A2], taco_amd.DrudeLorentz(0.08, 4.0), 0.7, 0.01, 300 + 50 * k,
                           list(range(81)), steps=80)
            r0 = np.zeros(9, complex)
            r0[0], r0[4] = 0.6, 0.4
            out.append(dict(gamma=w.gamma, omegas=w.system.omegas, mirror=w.system.mirror,
                            pair_to_freq=w.system.pair_to_freq, ops=w.system.A_eig, eps=w.system.eps, dt=w.dt,
                            steps=80, r0=r0, order=order))
        return out, 16
    # d = 16 TFIM chain, two transverse fields (nf = 81 and 57)
    ws = [W.tfim_workload(h, lam=0.2, dt=0.05, Nt=64, tmax=6) for h in (0.7, 1.0)]
    return [dict(gamma=w.gamma, omegas=w.system.omegas, mirror=w.system.mirror, pair_to_freq=w.system.pair_to_freq,
                 ops=w.system.A_eig, eps=w.system.eps, dt=w.dt, steps=6, r0=r0_for(16)) for w in ws], 4


def _whole_slice_bytes(st):
    return 3 * len(st["omegas"]) ** 3 * (st["steps"] + 1) * 16 if st.get("order", 4) == 4 else 0


@pytest.mark.parametrize("kind", ["spin_pair_d4", "gue3_mixed_order", "tfim_d16"])
@pytest.mark.parametrize("budget", ["keep", "packed", "partial", "recompute"])
def test_multi_set_tape_equals_single_tapes(kind, budget):
    """Every set of a multi-set tape (one RK4 launch forward, one recompute
    and one walk launch in reverse for all sets) has the samples and the
    cotangents of a single tape of that set, bit for bit -- with the slices
    kept whole, kept packed (a budget just short of the whole slices), with a
    budget that packs only the first set's, and with one that forces the
    recomputation of all."""
    sets, K = _mt_sets(kind)
    rng = np.random.default_rng(31)
    bars = [rng.standard_normal((s["steps"] + 1, len(s["r0"]))) + 1j * rng.standard_normal(
        (s["steps"] + 1, len(s["r0"]))) for s in sets]
    sliced = [k for k, st in enumerate(sets) if st.get("order", 4) == 4]
    packed_budget = sum(_whole_slice_bytes(st) for st in sets) - 16
    if budget == "recompute":
        taco_amd.set_slice_budget(16)
    elif budget == "packed":
        taco_amd.set_slice_budget(packed_budget)
    elif budget == "partial":  # room for the first set's packed slices only
        taco_amd.set_slice_budget(packed_budget)
        try:
            probe = AD.MultiTape(sets, checkpoint_every=K)
            first = probe.slice_modes[sliced[0]]
            probe.close()
        finally:
            taco_amd.set_slice_budget(0)
        assert first[0] == "packed" and 0 < first[1] < _whole_slice_bytes(sets[sliced[0]])
        taco_amd.set_slice_budget(first[1] + 16)
    try:
        mt = AD.MultiTape(sets, checkpoint_every=K)
        modes = mt.slice_modes
        want = {"keep": "whole", "packed": "packed", "recompute": "recompute"}
        for k in sliced:
            if budget == "partial":
                assert modes[k][0] == ("packed" if k == sliced[0] else "recompute"), (kind, modes)
            else:
                assert modes[k][0] == want[budget], (kind, modes)
        if budget == "keep":
            assert [m[1] for m in modes] == [_whole_slice_bytes(st) for st in sets]
        assert mt.keeps_slices == (budget in ("keep", "packed"))
        got = mt.adjoint(bars)
        again = mt.adjoint(bars)
        samples = [s.copy() for s in mt.samples]
        mt.close()
    finally:
        taco_amd.set_slice_budget(0)
    for k, st in enumerate(sets):
        tp = AD.Tape(st["gamma"], st["omegas"], st["mirror"], st["pair_to_freq"], st["ops"], st["eps"], st["dt"],
                     st["steps"], st["r0"], st.get("order", 4), K)
        ref = tp.adjoint(bars[k])
        assert np.array_equal(samples[k], tp.samples), (kind, k)
        tp.close()
        for key in ("gamma", "omegas", "eps", "r0"):
            assert np.array_equal(got[k][key], ref[key]), (kind, k, key)
            assert np.array_equal(again[k][key], ref[key]), (kind, k, key)
        for x, y in zip(got[k]["ops"], ref["ops"]):
            assert np.array_equal(x, y), (kind, k)


def test_two_live_multi_set_tapes_keep_their_own_state():
    """The context lends its multi-set buffers to one live multi-set tape at a
    time: a second tape created while the first is alive runs on buffers of
    its own, so reversing the first afterwards still gives its single-tape
    cotangents, and a third tape after both are closed reuses the context's."""
    sets, K = _mt_sets("spin_pair_d4")
    other = [dict(st, r0=np.roll(st["r0"], 5)) for st in sets]
    rng = np.random.default_rng(33)
    bars = [rng.standard_normal((s_["steps"] + 1, 16)) + 0j for s_ in sets]
    a = AD.MultiTape(sets, checkpoint_every=K)
    b = AD.MultiTape(other, checkpoint_every=K)
    got_b = b.adjoint(bars)
    got_a = a.adjoint(bars)
    a.close()
    b.close()
    c = AD.MultiTape(sets, checkpoint_every=K)
    got_c = c.adjoint(bars)
    c.close()
    for k, (st, ot) in enumerate(zip(sets, other)):
        for src, res in ((st, got_a[k]), (ot, got_b[k]), (st, got_c[k])):
            tp = AD.Tape(src["gamma"], src["omegas"], src["mirror"], src["pair_to_freq"], src["ops"], src["eps"],
                         src["dt"], src["steps"], src["r0"], 4, K)
            ref = tp.adjoint(bars[k])
            tp.close()
            for key in ("gamma", "omegas", "eps", "r0"):
                assert np.array_equal(res[key], ref[key]), (k, key)


@pytest.mark.parametrize("budget", ["keep", "packed", "recompute"])
def test_multi_set_bath_tape_equals_single_bath_tapes(budget):
    """Multi-set tape with device-made Gamma tables (taco_mtape_*_bath): every
    set's samples and cotangents, the bath-parameter ones included, equal a
    BathTape of that set alone, bit for bit (slices kept or recomputed)."""
    ws = [W.spin_pair(eps=e, delta=d, alpha=a, Nt=1024, steps=120)
          for e, d, a in ((1.0, 1.0, 0.05), (0.8, 1.3, 0.08), (1.2, 0.6, 0.03))]
    baths = [(0, 0.05, 5.0, 1.0), (0, 0.08, 4.0, 0.7), (1, 0.03, 3.0, 1.5)]
    sets = [dict(Nt=1024, omegas=w.system.omegas, mirror=w.system.mirror, pair_to_freq=w.system.pair_to_freq,
                 ops=w.system.A_eig, eps=w.system.eps, dt=w.dt, steps=120, r0=r0_for(4)) for w in ws]
    rng = np.random.default_rng(32)
    bars = [rng.standard_normal((121, 16)) + 1j * rng.standard_normal((121, 16)) for _ in sets]
    if budget == "recompute":
        taco_amd.set_slice_budget(16)
    elif budget == "packed":
        taco_amd.set_slice_budget(sum(_whole_slice_bytes(st) for st in sets) - 16)
    try:
        mt = AD.MultiTape(sets, checkpoint_every=32, baths=baths)
        assert mt.keeps_slices == (budget != "recompute")
        assert {m[0] for m in mt.slice_modes} == {"keep": {"whole"}, "packed": {"packed"},
                                                  "recompute": {"recompute"}}[budget]
        got = mt.adjoint(bars)
        samples = [x.copy() for x in mt.samples]
        mt.close()
    finally:
        taco_amd.set_slice_budget(0)
    for k, (st, b) in enumerate(zip(sets, baths)):
        tp = AD.BathTape(*b, st["Nt"], st["omegas"], st["mirror"], st["pair_to_freq"], st["ops"], st["eps"],
                         st["dt"], st["steps"], st["r0"], 4, 32)
        ref = tp.adjoint(bars[k])
        assert np.array_equal(samples[k], tp.samples), k
        tp.close()
        for key in ("bath", "omegas", "eps", "r0"):
            assert np.array_equal(got[k][key], ref[key]), (k, key)
        for x, y in zip(got[k]["ops"], ref["ops"]):
            assert np.array_equal(x, y), k


@pytest.mark.parametrize("chain", ["device", "split"])
def test_tfim_losses_many_equal_set_by_set(chain):
    """The d = 16 chain's multi-set loss path (one tape for all parameter
    sets) gives every set the loss and parameter gradient of the set-by-set
    path, bit for bit, with the bath chain on the device (BathTCL4Propagate
    and its multi-set form) or in separate steps."""
    prob = AD.TfimChainProblem(Nt=256, dt=0.02, steps=12, checkpoint_every=4, bath_chain=chain)
    P = [np.array([0.1, 5.0, 1.0, 1.0, h]) for h in (0.7, 1.0, 1.3)]
    with torch.no_grad():
        targets = [prob.rho_lab(torch.tensor(p * 1.03, dtype=torch.float64)).detach() for p in P]
    params = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in P]
    losses = prob.losses_many(params, targets)
    grads = torch.autograd.grad(sum(losses), params)
    for q, p in enumerate(P):
        x = torch.tensor(p, dtype=torch.float64, requires_grad=True)
        l1 = prob.loss(x, targets[q])
        (g1,) = torch.autograd.grad(l1, x)
        assert float(l1) == float(losses[q])
        assert np.array_equal(g1.numpy(), grads[q].numpy()), q


@pytest.mark.parametrize("kind", ["d16", "d10", "d4_rows", "d3_rows"])
def test_deferred_reverse_walk_equals_in_place_walk(monkeypatch, kind):
    """The reverse RK4 walk with the generator cotangents formed afterwards by
    k_rk4_lbar (large D: over conjugate-transposed generators; D = 4, 9, 16:
    the row-group walk storing kb1..kb4) gives the cotangents of the walk that
    forms them in place, bit for bit (single and multi-set)."""
    if kind == "d16":
        S, G, dt = D16.system, D16.gamma, D16.dt
        steps, K = D16_STEPS, D16_K
    elif kind == "d4_rows":
        name, S, G, dt, steps, K = CASES[1]
    elif kind == "d3_rows":
        name, S, G, dt, steps, K = CASES[2]
    else:
        rng = W.SplitMix64(17)
        H = W.gue(10, rng)
        w = W.Workload("g10", H, [W.random_hermitian_unit(10, rng)], taco_amd.DrudeLorentz(0.1, 3.0), 0.8, 0.02, 127,
                       list(range(21)), steps=20)
        S, G, dt, steps, K = w.system, w.gamma, w.dt, 20, 8
    N = S.N
    rng = np.random.default_rng(5)
    bar = rng.standard_normal((steps + 1, N * N)) + 1j * rng.standard_normal((steps + 1, N * N))
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("TACO_RK4_DEFER", mode)
        tp = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, dt, steps, r0_for(N), 4, K)
        out[mode] = tp.adjoint(bar)
        tp.close()
        st = dict(gamma=G, omegas=S.omegas, mirror=S.mirror, pair_to_freq=S.pair_to_freq, ops=S.A_eig, eps=S.eps,
                  dt=dt, steps=steps, r0=r0_for(N))
        mt = AD.MultiTape([st, st], checkpoint_every=K)
        out[mode + "m"] = mt.adjoint([bar, bar])
        mt.close()
    for key in ("gamma", "omegas", "eps", "r0"):
        assert np.array_equal(out["0"][key], out["1"][key]), key
        for q in range(2):
            assert np.array_equal(out["0m"][q][key], out["1"][key]), key
    for x, y in zip(out["0"]["ops"], out["1"]["ops"]):
        assert np.array_equal(x, y)


def _chain_grad(prob, p, target):
    params = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    rho = prob.rho_lab(params)
    d = rho - target
    loss = (d.real ** 2 + d.imag ** 2).sum()
    (g,) = torch.autograd.grad(loss, params)
    return rho.detach().numpy(), float(loss), g.numpy()


@pytest.mark.parametrize("kind", ["spin_pair_ohmic", "spin_pair_order2", "tfim_drude", "tiny_grid"])
def test_device_bath_chain_equals_split_chain(kind):
    """The fused device chain (BathTape: C(t) and Gamma made, kept and
    reversed in device memory) against the split chain (BcfFun -> GammaFun ->
    TCL4Propagate, whose chain reverses run on the host in numpy): the same
    rho(t) bit for bit (the device producers are bit-identical), and the same
    parameter gradient up to the rounding of the two reverse implementations
    (different FFTs and summation orders: bar 1e-10 of the largest entry)."""
    if kind == "tfim_drude":
        mk = lambda chain: AD.TfimChainProblem(Nt=256, dt=0.02, steps=12, checkpoint_every=4, bath_chain=chain)
        p0 = np.array([0.1, 5.0, 1.0, 1.0, 0.7])
    elif kind == "tiny_grid":  # Nt = 2: a length-4 FFT on one thread
        mk = lambda chain: AD.SpinPairProblem(Nt=2, dt=0.01, steps=1, checkpoint_every=1, bath_chain=chain)
        p0 = np.array([0.05, 5.0, 1.0, 1.0, 1.0])
    else:
        order = 2 if kind == "spin_pair_order2" else 4
        mk = lambda chain: AD.SpinPairProblem(Nt=2 ** 12, steps=400, checkpoint_every=64, order=order,
                                              bath_chain=chain)
        p0 = np.array([0.05, 5.0, 1.0, 1.0, 1.0])
    dev, split = mk("device"), mk("split")
    u = np.where(np.random.default_rng(4).uniform(size=5) < 0.5, -1.0, 1.0)
    with torch.no_grad():
        target = split.rho_lab(torch.tensor(p0 * (1 + 0.05 * u), dtype=torch.float64)).detach()
    rho_d, loss_d, g_d = _chain_grad(dev, p0, target)
    rho_s, loss_s, g_s = _chain_grad(split, p0, target)
    assert np.array_equal(rho_d, rho_s)
    assert loss_d == loss_s
    scale = np.abs(g_s).max()
    for q, name in enumerate(dev.names):
        print(f"{kind} d loss / d {name}: device {g_d[q]:.15e} split {g_s[q]:.15e} "
              f"diff/max {abs(g_d[q] - g_s[q]) / scale:.2e}")
    assert np.abs(g_d - g_s).max() <= 1e-10 * scale
    # the device reverse repeats bit for bit
    _, _, g_d2 = _chain_grad(dev, p0, target)
    assert np.array_equal(g_d, g_d2)


def test_bath_tape_rejects_infinite_beta_cotangents():
    sp = W.spin_pair(Nt=256, steps=20)
    S = sp.system
    tp = AD.BathTape(0, 0.05, 5.0, float("inf"), 256, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, sp.dt, 20,
                     r0_for(4), 4, 8)
    # the forward runs at zero temperature (the producers support it) ...
    _, C = taco_amd.bcf_fft_fun(255, sp.dt, taco_amd.Ohmic(0.05, 5.0), float("inf"))
    G = taco_amd.compute_trapz_prefix_multi_matrix(C, sp.dt, S.omegas)
    ref = AD.Tape(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, S.eps, sp.dt, 20, r0_for(4), 4, 8)
    assert np.array_equal(tp.samples, ref.samples)
    ref.close()
    # ... and its bath cotangents are refused
    with pytest.raises(_lib.TacoInvalidArgument, match="finite beta"):
        tp.adjoint(np.ones((21, 16), complex))
    tp.close()
