"""Phase convention: correctly rounded cos/sin of omega*(k*dt).

The oracle's correctly rounded routine (x87 long double with a quad-precision
fallback near rounding boundaries) is pinned against a 60-digit decimal
evaluation; the engine's double-double routine (host build of the device
code) must agree with it bit for bit."""
import ctypes
from decimal import Decimal, getcontext

import numpy as np
import pytest

from taco_amd import _lib, workloads as W


def decimal_sincos(x: float):
    getcontext().prec = 70
    # pi to 70 digits by the decimal-module recipe
    getcontext().prec += 2
    three = Decimal(3)
    lasts, t, s, n, na, d, da = 0, three, 3, 1, 0, 0, 24
    while s != lasts:
        lasts = s
        n, na = n + na, na + 8
        d, da = d + da, da + 32
        t = (t * n) / d
        s += t
    getcontext().prec -= 2
    pi = +s
    X = Decimal(x)
    k = (X / (pi / 2)).to_integral_value()
    r = X - k * (pi / 2)
    sr = cr = Decimal(0)
    term = r
    i = 1
    while abs(term) > Decimal(10) ** -68:
        sr += term
        term = -term * r * r / ((i + 1) * (i + 2))
        i += 2
    term = Decimal(1)
    i = 0
    while abs(term) > Decimal(10) ** -68:
        cr += term
        term = -term * r * r / ((i + 1) * (i + 2))
        i += 2
    q = int(k) % 4
    s_, c_ = [(sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr)][q]
    return float(s_), float(c_)  # Decimal -> float rounds correctly


def host_cr(x):
    x = np.ascontiguousarray(x, np.float64)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    _lib.check(_lib.lib().taco_cr_sincos(None, _lib.dptr(x), x.size, _lib.dptr(s), _lib.dptr(c)))
    return s, c


def test_oracle_cr_matches_decimal(orc):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-4e3, 4e3, 400), rng.uniform(-2, 2, 200), [0.5, 1e-8, 100.0, 3000.0]])
    s, c = orc.cr_sincos(xs)
    for x, si, ci in zip(xs, s, c):
        ds, dc = decimal_sincos(float(x))
        assert si == ds and ci == dc, x


def test_engine_cr_bitwise_equals_oracle(orc):
    rng = np.random.default_rng(2)
    w = W.gue_pair()
    om = w.system.omegas
    oms = sorted({(om[i] + om[j]) + om[k] for i in range(13) for j in range(13) for k in range(13)})
    grid = np.concatenate([o * (np.arange(0, 200001, 101) * 0.002) for o in oms[::9]])
    xs = np.concatenate([rng.uniform(-1e5, 1e5, 200000), rng.uniform(-3, 3, 100000), grid,
                         [0.0, -0.0, np.pi / 2, np.pi, 1e5 * np.pi, 0.785398163397448]])
    s1, c1 = host_cr(xs)
    s2, c2 = orc.cr_sincos(xs)
    assert np.array_equal(s1, s2) and np.array_equal(c1, c2)
    assert np.signbit(s1[np.where(xs == 0)[0]]).tolist() == np.signbit(xs[xs == 0]).tolist()


def test_phase_symmetry(orc):
    x = np.linspace(-500, 500, 10001)
    s, c = orc.cr_sincos(x)
    s2, c2 = orc.cr_sincos(-x)
    assert np.array_equal(s, -s2) and np.array_equal(c, c2)


def test_reference_libm_sensitivity_at_early_t(orc):
    """The reference's early-t L4 depends on the libm's last bits: the oracle
    with glibc phases and with correctly rounded phases differ by ~1e-10..1e-9
    per-t at the first steps (the FFT's rounding noise), then agree to ~1e-15."""
    from .metrics import per_t_rel_err
    w = W.spin_boson()
    S, G = w.system, w.gamma
    with orc.phase_mode(orc.PHASE_LIBM):
        a = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, w.tidx)
    with orc.phase_mode(orc.PHASE_CR):
        b = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, w.tidx)
    e, _ = per_t_rel_err(a, b)
    assert e[0] > 1e-12          # t = 0: pure rounding noise, libm-dependent
    assert e[5:].max() < 1e-13   # later t: libm-independent


def decimal_exp_rounded(x: float) -> float:
    getcontext().prec = 80
    return float(Decimal(x).exp())  # Decimal -> float rounds correctly (also to subnormals)


def exp_args(rng, n_random):
    return np.concatenate([
        rng.uniform(-745.2, 709.8, n_random), rng.uniform(-3, 3, n_random // 2),
        rng.uniform(-745.2, -707.0, n_random // 8),      # subnormal results
        -rng.uniform(0, 1, n_random // 4) * 200.0,       # -w / omega_c of Ohmic spectra
        [0.0, -0.0, 1.0, -1.0, 709.782712893384, 709.7827128933841, -745.1332191019412,
         -745.1332191019411, -744.44007192138126, -708.3964185322641, 1e-300, -1e-300, 5e-324]])


def test_oracle_cr_exp_matches_decimal(orc):
    """The bath correlation's exponentials (the Ohmic factor exp(-w/w_c) and
    the KMS factors) use the correctly rounded exp; the oracle's (quad
    precision rounded once) against a 80-digit decimal evaluation."""
    rng = np.random.default_rng(3)
    xs = exp_args(rng, 400)
    got = orc.cr_exp(xs)
    for x, g in zip(xs, got):
        assert g == decimal_exp_rounded(float(x)), x


def host_exp(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    _lib.check(_lib.lib().taco_cr_exp(None, _lib.dptr(x), x.size, _lib.dptr(out)))
    return out


def test_engine_cr_exp_bitwise_equals_oracle(orc):
    """The engine's double-double exp (host build of the device code) equals
    the oracle's bit for bit, including the subnormal range, overflow and the
    arguments of the bench bath (w = k dw up to the Nyquist frequency)."""
    rng = np.random.default_rng(4)
    dw = 2 * np.pi / (0.002 * 2 ** 19)
    grid = -(np.arange(1, 2 ** 18) * dw) / 10.0
    xs = np.concatenate([exp_args(rng, 400000), grid])
    a, b = host_exp(xs), orc.cr_exp(xs)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, (xs[bad[:5]], a[bad[:5]], b[bad[:5]])
    assert host_exp(np.array([710.0]))[0] == np.inf and host_exp(np.array([-746.0]))[0] == 0.0
