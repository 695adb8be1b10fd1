"""GPU parity of the F/C/R kernels and the fused L4 builders against the
oracle, through the C ABI.  Bar: per-t relative Frobenius error <= 1e-10
(floored at 1e-6 of the series maximum, see tests/metrics.py)."""
import numpy as np
import pytest

import taco_amd
from taco_amd import workloads as W
from taco_amd._lib import TacoError, TacoInvalidArgument, TacoOutOfRange

from .metrics import max_rel_err, per_t_rel_err, ref_style_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def series_t_major(X):
    nf = X.shape[0]
    return X.reshape(nf ** 3, -1).T  # [t, nf^3]


def smooth_gamma(Nt, nf, seed):
    """Gamma-like columns: zero at t = 0, saturating, oscillating."""
    rng = np.random.default_rng(seed)
    t = np.arange(Nt) * 0.01
    cols = []
    for b in range(nf):
        a, g, w = rng.uniform(0.5, 2), rng.uniform(0.5, 3), rng.uniform(-3, 3)
        cols.append(a * (1 - np.exp(-(g - 1j * w) * t)) / (g - 1j * w))
    return np.stack(cols, axis=1)


def check_fcr(orc, S, G, dt, keep):
    """Per-t error of the F/C/R slices.  F(0) and C(0) vanish in exact
    arithmetic (Gamma(., 0) = 0 kills both terms); the computed values are FFT
    rounding noise, so t = 0 is checked in absolute terms against the series
    maximum and t >= 1 with the relative metric."""
    got = taco_amd.compute_triple_kernels(S, G, dt, keep=keep)
    ref = orc.triple_kernels(G, S.omegas, S.mirror, dt, keep=keep)
    errs = {}
    for name, g, r in zip("FCR", (got.F, got.C, got.R), ref):
        gt, rt = series_t_major(g), series_t_major(r)
        scale = np.abs(rt).max() if rt.size else 0.0
        assert np.abs(gt[0] - rt[0]).max() <= 1e-14 * max(scale, 1e-300), name
        e = 0.0
        if gt.shape[0] > 1:
            per_t, _ = per_t_rel_err(gt[1:], rt[1:])
            e = float(per_t.max())
            print(name, "per-t rel err t=1..8:", np.array2string(per_t[:8], precision=2))
        errs[name] = e
    return errs


def test_triple_kernels_spin_boson_full_grid(orc):
    w = W.spin_boson()
    errs = check_fcr(orc, w.system, w.gamma, w.dt, 2001)
    print("spin_boson F/C/R max per-t rel err", errs)
    assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize("Nt", [1, 2, 3, 5, 64, 1000, 4097])
def test_triple_kernels_edge_lengths(orc, Nt):
    w = W.spin_boson()
    S = w.system
    G = smooth_gamma(Nt, S.nf, Nt)
    errs = check_fcr(orc, S, G, 0.01, Nt)
    assert max(errs.values()) <= TOL, (Nt, errs)


def test_triple_kernels_pad_factor(orc):
    w = W.spin_boson(Nt=3000)
    S, G = w.system, w.gamma
    taco_amd.set_fcr_fft_pad_factor(4)
    try:
        got = taco_amd.compute_triple_kernels(S, G, w.dt, keep=500)
    finally:
        taco_amd.set_fcr_fft_pad_factor(0)
    ref = orc.triple_kernels(G, S.omegas, S.mirror, w.dt, keep=500, pad=4)
    for g, r in zip((got.F, got.C, got.R), ref):
        assert max_rel_err(series_t_major(g), series_t_major(r))[0] <= TOL


def l4_check(orc, w, tidx, tol=TOL):
    S = w.system
    G = w.gamma
    got = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, tidx)
    ref = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, tidx)
    e, nfloor = max_rel_err(got, ref)
    per_t, _ = per_t_rel_err(got, ref)
    print(f"{w.name}: ntau={len(tidx)} max per-t rel err {e:.3e} (floored t: {nfloor}), "
          f"ref-style {ref_style_err(got, ref):.3e}, first t: {np.array2string(per_t[:6], precision=2)}")
    assert e <= tol
    return got, ref


def test_l4_spin_boson(orc):
    w = W.spin_boson()
    l4_check(orc, w, w.tidx)


def test_l4_gue_pair_reduced_grid(orc):
    w = W.gue_pair(N_bath=20000, tmax=1000)
    l4_check(orc, w, w.tidx)


def test_l4_spin_boson_cr_phases(orc_cr):
    """TACO_PHASE_CR (correctly rounded phases on the device) against the
    oracle in the same convention."""
    w = W.spin_boson()
    l4_check(orc_cr, w, w.tidx)


def test_l4_gue_pair_reduced_grid_cr_phases(orc_cr):
    w = W.gue_pair(N_bath=20000, tmax=1000)
    l4_check(orc_cr, w, w.tidx)


def test_libm_phase_early_t_matches_reference_arithmetic(orc):
    """In the libm convention the first steps (pure FFT rounding noise, where
    the two conventions differ by ~1e-9) match the oracle as closely as the
    later ones."""
    w = W.gue_pair(N_bath=20000, tmax=40)
    got, ref = l4_check(orc, w, w.tidx)
    per_t, _ = per_t_rel_err(got, ref)
    assert per_t[1:20].max() <= 1e-11, per_t[:20]


def test_pure_dephasing_l4_vanishes_on_the_gpu(orc):
    """[H, A] = 0: the fourth-order TCL generator of a Gaussian bath is zero
    (the oracle gives exact zeros); the engine's sums leave at most rounding."""
    H = 0.5 * 1.3 * np.diag([1.0, -1.0]).astype(complex)
    A = np.diag([1.0, -1.0]).astype(complex)
    w = W.Workload("dephasing", H, [A], taco_amd.Ohmic(0.05, 5.0), 1.0, 0.01, 4095, list(range(0, 2000, 7)))
    got = taco_amd.build_TCL4_generator_hip_fused_batch(w.system, w.gamma, w.dt, w.tidx)
    sb = W.spin_boson(Nt=4096, tmax=2000, tstep=7)
    scale = np.abs(taco_amd.build_TCL4_generator_hip_fused_batch(sb.system, sb.gamma, sb.dt, sb.tidx)).max()
    assert np.abs(got).max() <= 1e-14 * scale, np.abs(got).max() / scale


def test_gpu_l4_converges_to_bruteforce_tcl4():
    """The engine's L4 approaches the fourth-order TCL generator evaluated by
    brute force from its definition (tests/tcl4_bruteforce.py) at first order
    in dt, like the oracle it matches (exact Gamma tables of an exponential
    bath, N = 3)."""
    from . import tcl4_bruteforce as BF

    rng = np.random.default_rng(3)
    X = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    Y = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    S = taco_amd.System.build((X + X.conj().T) / 2, [(Y + Y.conj().T) / 4])
    bath = BF.ExpBath([0.3 + 0.1j, 0.2 - 0.05j], [1.0 + 2.0j, 2.5 - 0.7j])
    t = 0.9
    ref = BF.k4(bath, np.asarray(S.eps), np.asarray(S.A_eig[0]), t)
    e = []
    for dt in (0.005, 0.0025):
        k = int(round(t / dt))
        got = taco_amd.build_TCL4_generator_hip_fused_batch(S, bath.gamma_table(S.omegas, k + 1, dt), dt, [k])
        e.append(np.linalg.norm(got[0] - ref) / np.linalg.norm(ref))
    assert e[1] < 1e-2 and 1.9 < e[0] / e[1] < 2.1, e


@pytest.mark.parametrize("wl", ["spin_boson", "d3"])
def test_device_mikx_and_gw_helpers_match_oracle(orc, wl):
    """taco_mikx_device (build_mikx_device) gathers M, I, K, X exactly as the
    oracle's mikx, and taco_assemble_gw_device (assemble_liouvillian_cuda_device
    [_raw]) sums the NAKZWAN terms in the oracle's order: equal bit for bit."""
    import torch

    from taco_amd import _lib

    if wl == "spin_boson":
        w = W.spin_boson(Nt=1024, tmax=300)
    else:
        w = W.heterogeneous_system(1, dims=(3,), n_time=40, log2_range=(10, 10))
    S, G = w.system, w.gamma
    N, nf, t = S.N, S.nf, 37
    F, C, R = orc.triple_kernels(G, S.omegas, S.mirror, w.dt, keep=t + 1)
    Fs, Cs, Rs = (np.ascontiguousarray(x[..., t]).reshape(-1) for x in (F, C, R))
    M, I, K, X = orc.mikx(Fs, Cs, Rs, S.pair_to_freq, N)
    ops = np.stack([np.asarray(a) for a in S.A_eig])
    GW = orc.assemble_gw(M, I, K, X, list(ops), N)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    dF, dC, dR = T(Fs), T(Cs), T(Rs)
    dpf = T(np.asarray(S.pair_to_freq, np.int32).T.reshape(-1))
    dops = T(np.stack([a.T for a in ops]).reshape(-1))
    N4, N6 = N ** 4, N ** 6
    out = [torch.zeros(n, dtype=torch.complex128, device=dev) for n in (N4, N4, N4, N6, N4, N4)]
    dM, dI, dK, dX, dGW, dGWr = out
    torch.cuda.synchronize()
    L = _lib.lib()
    _lib.check(L.taco_mikx_device(dF.data_ptr(), dC.data_ptr(), dR.data_ptr(), dpf.data_ptr(), N, nf, dM.data_ptr(),
                                  dI.data_ptr(), dK.data_ptr(), dX.data_ptr(), None))
    _lib.check(L.taco_assemble_gw_device(dM.data_ptr(), dI.data_ptr(), dK.data_ptr(), dX.data_ptr(), dops.data_ptr(),
                                         N, len(ops), dGW.data_ptr(), 0, None))
    _lib.check(L.taco_assemble_gw_device(dM.data_ptr(), dI.data_ptr(), dK.data_ptr(), dX.data_ptr(), dops.data_ptr(),
                                         N, len(ops), dGWr.data_ptr(), 1, None))
    torch.cuda.synchronize()
    col = lambda d: d.cpu().numpy().reshape(N * N, N * N).T  # noqa: E731
    assert np.array_equal(col(dM), M) and np.array_equal(col(dI), I) and np.array_equal(col(dK), K)
    assert np.array_equal(dX.cpu().numpy(), X)
    assert np.array_equal(col(dGW), GW)
    raw = col(dGWr)
    assert np.array_equal(raw + raw.conj().T, GW)


def test_libm_phase_table_budget(orc, monkeypatch):
    """A libm phase table above TACO_LIBM_TABLE_BYTES is refused with an
    invalid-argument error; the CR convention needs no table."""
    w = W.spin_boson(Nt=500)
    monkeypatch.setenv("TACO_LIBM_TABLE_BYTES", "4096")
    with pytest.raises(TacoInvalidArgument):
        taco_amd.build_TCL4_generator_hip_fused_batch(w.system, w.gamma, w.dt, [0, 10])
    with taco_amd.phase_mode(taco_amd.PHASE_CR):
        taco_amd.build_TCL4_generator_hip_fused_batch(w.system, w.gamma, w.dt, [0, 10])


def test_l4_published_setting_reduced(orc):
    w = W.published(N_bath=20000, tmax=2000)
    l4_check(orc, w, w.tidx)


@pytest.mark.parametrize("b", [0, 5])
def test_l4_heterogeneous_small(orc, b):
    w = W.heterogeneous_system(b, n_time=40, log2_range=(10, 11))
    if w.meta["d"] > 6:
        pytest.skip("d = 8 covered by test_l4_d8_small")
    l4_check(orc, w, w.tidx[:40])


def test_l4_d8_small(orc):
    rng = W.SplitMix64(77)
    H = W.gue(8, rng)
    A = W.random_hermitian_unit(8, rng)
    w = W.Workload("d8", H, [A], taco_amd.DrudeLorentz(0.1, 5.0), 1.0, 0.01, 299, [0, 1, 7, 50, 299])
    assert w.system.nf == 57
    l4_check(orc, w, w.tidx)


def test_l4_time_list_semantics(orc):
    w = W.spin_boson(Nt=500)
    S, G = w.system, w.gamma
    tidx = [37, 5, 37, 0, 499, 5]
    got = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, tidx)
    ref = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, tidx)
    assert max_rel_err(got, ref)[0] <= TOL
    one = taco_amd.build_TCL4_generator_hip_fused(S, G, w.dt, 37)
    assert np.array_equal(one, got[0])
    allt = taco_amd.build_correction_series_hip_fused(S, G[:200], w.dt)
    assert allt.shape == (200, 4, 4)
    refall = orc.l4_series(G[:200], S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, list(range(200)))
    assert max_rel_err(allt, refall)[0] <= TOL
    # deterministic: a second run is bitwise identical
    again = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, tidx)
    assert np.array_equal(again, got)


def test_l4_from_slices_fixture(orc):
    # tests/tcl4_tests.cpp:36-90 synthetic kernels, ops {I, sx}
    from .test_oracle_mikx_gw import P2F_FIXTURE, fixture_slices
    ops = [np.eye(2, dtype=complex), np.array([[0, 1], [1, 0]], complex)]
    for t in (0, 2, 3):
        F, C, R = fixture_slices(t=t)
        L4, GW = taco_amd.assemble_l4_from_slices(F, C, R, P2F_FIXTURE, ops)
        M, I, K, X = orc.mikx(F, C, R, P2F_FIXTURE, 2)
        GWr = orc.assemble_gw(M, I, K, X, ops, 2)
        assert np.abs(GW - GWr).max() <= 1e-12 * np.abs(GWr).max()
        assert np.abs(L4 - orc.gw_to_l4(GWr, 2)).max() <= 1e-12 * np.abs(GWr).max()


@pytest.mark.parametrize("case", ["gue_pair", "d8"])
def test_l4_pruned_triples_match_full_slices(orc, case, monkeypatch):
    """The L4 builders compute only the triples the GW sums read (F, C and R
    separately).  With the slice buffer NaN-filled first, an entry read but
    never written would show up; the result equals L4 assembled from the full
    nf^3 slices of compute_triple_kernels bit for bit."""
    if case == "gue_pair":
        w = W.gue_pair(N_bath=3000, tmax=300)
    else:
        rng = W.SplitMix64(77)
        w = W.Workload("d8", W.gue(8, rng), [W.random_hermitian_unit(8, rng)], taco_amd.DrudeLorentz(0.1, 5.0),
                       1.0, 0.01, 199, [0, 3, 199])
    S, G = w.system, w.gamma
    tidx = list(w.tidx)[:6]
    monkeypatch.setenv("TACO_DEBUG_POISON", "1")
    got = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, tidx)
    assert np.isfinite(got).all()
    full = taco_amd.compute_triple_kernels(S, G, w.dt, keep=max(tidx) + 1)
    for q, t in enumerate(tidx):
        L4, _ = taco_amd.assemble_l4_from_slices(full.F[..., t], full.C[..., t], full.R[..., t], S.pair_to_freq,
                                                 S.A_eig)
        assert np.array_equal(got[q], L4), (case, t)


def test_l4_error_paths():
    w = W.spin_boson(Nt=64)
    S, G = w.system, w.gamma
    with pytest.raises(TacoOutOfRange):
        taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, [3, 64])
    with pytest.raises(TacoOutOfRange):
        taco_amd.build_TCL4_generator_hip_fused(S, G, w.dt, 64)
    with pytest.raises(TacoInvalidArgument):
        taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, [1], method=taco_amd.FCRMethod.Direct)
    with pytest.raises(TacoInvalidArgument):
        taco_amd.build_TCL4_generator_hip_fused_batch(S, G, -1.0, [1])
    with pytest.raises(TacoInvalidArgument):
        taco_amd.build_TCL4_generator_hip_fused_batch(S, G[:, :2], w.dt, [1])
    import copy
    S2 = copy.deepcopy(S)
    S2.pair_to_freq[0, 1] = -1
    with pytest.raises(TacoError, match="pair_to_freq"):
        taco_amd.build_TCL4_generator_hip_fused_batch(S2, G, w.dt, [1])
    S3 = copy.deepcopy(S)
    S3.A_eig = []
    with pytest.raises(TacoInvalidArgument):
        taco_amd.build_TCL4_generator_hip_fused_batch(S3, G, w.dt, [1])
    assert taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, [5]).shape == (1, 4, 4)


def test_device_cr_sincos_bitwise_equals_host():
    import ctypes
    from taco_amd import _lib
    ctx = taco_amd.context(0)
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-1e5, 1e5, 500000), rng.uniform(-3, 3, 200000), [0.0, -0.0, np.pi]])
    out = []
    for h in (ctx.handle, None):
        s = np.zeros_like(x)
        c = np.zeros_like(x)
        _lib.check(_lib.lib().taco_cr_sincos(h, _lib.dptr(x), x.size, _lib.dptr(s), _lib.dptr(c)))
        out.append((s, c))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_fft_path_is_bitwise_at_t0(orc):
    """With identical phases, F(0) and C(0) (pure FFT rounding noise, since
    Gamma(., 0) = 0 zeroes the prefix term) equal the oracle's bit for bit:
    the three-pass FFT performs the reference's butterflies exactly."""
    w = W.spin_boson()
    S, G = w.system, w.gamma
    got = taco_amd.compute_triple_kernels(S, G, w.dt, keep=3)
    ref = orc.triple_kernels(G, S.omegas, S.mirror, w.dt, keep=3)
    assert np.array_equal(got.F[..., 0], ref[0][..., 0])
    assert np.array_equal(got.C[..., 0], ref[1][..., 0])
    assert np.array_equal(got.R[..., 0], ref[2][..., 0])


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5, 6])
def test_fft_tile_configs_bitwise_identical(orc, cfg):
    """Every FFT tile configuration the autotuner may pick performs the same
    butterflies: results are bit-identical across configurations."""
    import os
    w = W.gue_pair(N_bath=3000, tmax=300)
    S, G = w.system, w.gamma
    os.environ["TACO_FFT_TILE"] = str(cfg)
    try:
        got = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)
    finally:
        del os.environ["TACO_FFT_TILE"]
    os.environ["TACO_FFT_TILE"] = "0"
    try:
        ref = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)
    finally:
        del os.environ["TACO_FFT_TILE"]
    assert np.array_equal(got, ref)
    oref = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, w.tidx)
    assert max_rel_err(got, oref)[0] <= TOL


@pytest.mark.parametrize("case", ["gue_pair", "d8"])
def test_l4_factorised_matches_oracle(orc, case, monkeypatch):
    """The factorised L4 (O(N^5 + N^3 nf) per time; the kernel for N > 8)
    forced on systems the LDS kernel also handles: same values to rounding."""
    if case == "gue_pair":
        w = W.gue_pair(N_bath=3000, tmax=300)
        tidx = w.tidx
    else:
        rng = W.SplitMix64(77)
        w = W.Workload("d8", W.gue(8, rng), [W.random_hermitian_unit(8, rng)], taco_amd.DrudeLorentz(0.1, 5.0),
                       1.0, 0.01, 299, [0, 1, 7, 50, 299])
        tidx = w.tidx
    monkeypatch.setenv("TACO_L4_FACTORISED", "1")
    l4_check(orc, w, tidx)


@pytest.mark.parametrize("h", [0.7, 1.0])
def test_l4_tfim_d16_matches_oracle(orc, h):
    """d = 16 (4-spin transverse-field Ising chain, nf = 81 at generic h, 57
    at h = 1): the factorised kernel against the oracle's O(N^6) NAKZWAN sum."""
    w = W.tfim_workload(h, Nt=64, tmax=19)
    S = w.system
    assert S.N == 16 and S.nf == (57 if h == 1.0 else 81)
    l4_check(orc, w, [0, 5, 19])


@pytest.mark.parametrize("wl", ["spin_boson", "gue_pair"])
def test_window_mode_matches_direct_method(orc, wl):
    """TACO_FFT_WINDOW (rows t < T only, Nfft = next_pow2(2T - 1), exact
    twiddles) against the reference's own Direct O(T^2) method
    (tcl4_kernels.cpp:188-281, tcl4_tests.cpp:188-212), which has no FFT
    rounding: per-t relative error <= 1e-10 from t = 20 on and <= 1e-8 at the
    first steps (where L4 is a small difference of FFT-rounded terms); the
    parity mode reproduces the reference's recurrence-twiddle FFT, whose
    early-t error against Direct is orders of magnitude larger (1e-4 for the
    spin-boson window)."""
    w = W.spin_boson(Nt=2 ** 14, tmax=300, tstep=1) if wl == "spin_boson" else W.gue_pair(N_bath=20000, tmax=150)
    S, G = w.system, w.gamma
    ref = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, w.tidx, method="direct")
    with taco_amd.fft_mode(taco_amd.FFT_WINDOW):
        win = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)
    par = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)
    e_win, _ = max_rel_err(win, ref)
    e_par, _ = max_rel_err(par, ref)
    pt_win, _ = per_t_rel_err(win, ref)
    print(f"{wl}: vs Direct: window {e_win:.3e} (t >= 20: {pt_win[20:].max():.3e}), parity {e_par:.3e}")
    assert e_win <= 1e-8 and pt_win[20:].max() <= 1e-10
    assert e_win <= e_par
    assert taco_amd.get_fft_mode() == taco_amd.FFT_PARITY


def test_window_mode_batch_and_tape_equal_plain(orc):
    """The window mode applies to every F/C/R pass: the batch equals the plain
    builder bit for bit in it, and TCL4 propagation through it stays within
    the bar of the parity-mode oracle."""
    sb = W.spin_boson(Nt=2048, tmax=300, tstep=3)
    r0 = np.zeros(4, complex)
    r0[0] = 1.0
    with taco_amd.fft_mode(taco_amd.FFT_WINDOW):
        b = taco_amd.Batch([taco_amd.BatchItem(sb.system, sb.gamma, sb.dt, sb.tidx, steps=300, r0=r0)]).run()
        plain = taco_amd.build_TCL4_generator_hip_fused_batch(sb.system, sb.gamma, sb.dt, sb.tidx)
        rho = taco_amd.tcl4_propagate(sb.system, sb.gamma, sb.dt, 300, r0)
        assert np.array_equal(b.l4(0), plain) and np.array_equal(b.rho(0), rho)
        b.close()
    S, G = sb.system, sb.gamma
    rr = orc.tcl4_propagate(G, S.omegas, S.mirror, S.pair_to_freq, S.eps, S.A_eig, sb.dt, 300, r0)
    assert max_rel_err(rho, rr)[0] <= 1e-10


@pytest.mark.parametrize("wl", ["spin_boson", "gue_pair"])
def test_combine_split_equals_single_launch(monkeypatch, wl):
    """The F/C/R assembly split over 256-sample blocks (block totals, carries,
    blocks) forms every value by the same operations in the same order as the
    one-workgroup-per-triple walk: bit-identical slices, series and L4."""
    w = W.spin_boson(Nt=3000, tmax=1500) if wl == "spin_boson" else W.gue_pair(N_bath=5000, tmax=700)
    S, G = w.system, w.gamma
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("TACO_COMBINE_SPLIT", mode)
        ser = taco_amd.compute_triple_kernels(S, G, w.dt, keep=1200)
        L4 = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)
        res[mode] = (ser.F, ser.C, ser.R, L4)
    for a, b in zip(res["0"], res["1"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("fft", ["parity", "window"])
def test_time_windows_equal_one_pass(orc, fft):
    """A slice budget below the call's kept times runs the builder as windows
    of consecutive kept times (an F/C/R pass and the L4 per window, the
    reference's time decomposition, tcl4_mpi_omp.cpp:62-156): L4 is
    bit-identical to the one-pass build for an unsorted time list with
    repeats, in both FFT modes, and the budget is per context."""
    w = W.gue_pair(N_bath=3000, tmax=400)
    S, G = w.system, w.gamma
    nf = len(S.omegas)
    rng = np.random.default_rng(11)
    tidx = [int(t) for t in rng.integers(0, 1500, 60)] + [0, 1499, 7, 7]
    mode = taco_amd.FFT_PARITY if fft == "parity" else taco_amd.FFT_WINDOW
    default = taco_amd.get_slice_budget()
    with taco_amd.fft_mode(mode):
        one = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, tidx)
        try:
            for per in (1, 7, 25):
                taco_amd.set_slice_budget(per * 3 * nf ** 3 * 16)
                assert taco_amd.get_slice_budget() == per * 3 * nf ** 3 * 16
                got = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, tidx)
                assert np.array_equal(got, one), per
        finally:
            taco_amd.set_slice_budget(0)
    assert taco_amd.get_slice_budget() == default
    if fft == "parity":
        ref = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, tidx)
        assert max_rel_err(one, ref)[0] <= TOL


@pytest.mark.parametrize("wl", ["spin_boson", "gue_pair"])
def test_direct_method_matches_oracle_direct(orc, wl):
    """FCRMethod.Direct on the GPU (tcl4_kernels.cpp:188-281: each sum in the
    reference's order) equals the oracle's Direct series bit for bit, and the
    L4 built from it matches the oracle's Direct L4; the fused CUDA names keep
    rejecting Direct as the reference's do (tcl4_fused_cuda.cu:820-822)."""
    w = W.spin_boson(Nt=600) if wl == "spin_boson" else W.gue_pair(N_bath=400, tmax=400)
    S, G, dt = w.system, w.gamma, w.dt
    got = taco_amd.compute_triple_kernels(S, G, dt, method=taco_amd.FCRMethod.Direct)
    ref = orc.triple_kernels(G, S.omegas, S.mirror, dt, method="direct")
    for name, a, b in zip("FCR", (got.F, got.C, got.R), ref):
        assert np.array_equal(a, b), name
    tidx = [0, 1, 7, 150, len(G) - 1]
    L = taco_amd.build_TCL4_generator(S, G, dt, tidx, method=taco_amd.FCRMethod.Direct)
    Lr = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, dt, tidx, method="direct")
    assert max_rel_err(L, Lr)[0] <= TOL
    one = taco_amd.build_TCL4_generator(S, G, dt, 150, method=taco_amd.FCRMethod.Direct)
    assert np.array_equal(one, L[3])
    with pytest.raises(TacoInvalidArgument):
        taco_amd.build_TCL4_generator_cuda_fused(S, G, dt, 3, method=taco_amd.FCRMethod.Direct)


def _ref_rel_err(a, b):
    """tcl4_tests.cpp:25-28: |a - b| / max(1, |b|)."""
    return np.abs(a - b) / np.maximum(1.0, np.abs(b))


@pytest.mark.parametrize("N,dt", [(2048, 0.000625), (8192, 0.000625), (16384, 0.000625),
                                  (4000, 0.0005), (8000, 0.0005), (1000, 0.002), (2000, 0.002), (4000, 0.002),
                                  (20000, 0.0001)])
def test_convolution_vs_direct_reference_sweep(N, dt):
    """The reference's own Convolution-vs-Direct check (tcl4_tests.cpp:166-236:
    H = sx/2, A = sz/2, J = w exp(-w/10), beta = 0.5, trapz Gamma; F/C/R at
    t = 0, Nt/4, Nt/2, 3Nt/4, Nt-1 over every triple), both methods on the GPU.
    The reference prints the errors; here they are bounded (the FFT route's
    recurrence twiddles and rounding leave ~1e-12)."""
    S = taco_amd.System.build(0.5 * W.SX, [0.5 * W.SZ])
    _, C = taco_amd.bcf_fft_fun(N, dt, taco_amd.Ohmic(1.0, 10.0), 0.5)
    G = taco_amd.compute_trapz_prefix_multi_matrix(C, dt, S.omegas)
    conv = taco_amd.compute_triple_kernels(S, G, dt, method=taco_amd.FCRMethod.Convolution)
    dirk = taco_amd.compute_triple_kernels(S, G, dt, method=taco_amd.FCRMethod.Direct)
    Nt = G.shape[0]
    tids = [0, Nt // 4, Nt // 2, min(Nt - 1, 3 * Nt // 4), Nt - 1]
    errs = [float(_ref_rel_err(a[..., tids], b[..., tids]).max())
            for a, b in ((conv.F, dirk.F), (conv.C, dirk.C), (conv.R, dirk.R))]
    print(f"Case N={N}, dt={dt}: errF={errs[0]:.3e}, errC={errs[1]:.3e}, errR={errs[2]:.3e}")
    assert max(errs) <= 1e-9
