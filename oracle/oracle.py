"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle (``tcl4_oracle.cpp``) restates the reference serial/OpenMP TCL4
path of jcarrots/taco.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only to check or
to time the reference algorithm; the product package ``taco_amd`` never does.

Parity status of each oracle function (what pins it; see tests/test_oracle_*):

==========================  ==========================================================
function                    pinned by
==========================  ==========================================================
fft                         brute-force O(n^2) DFT (n <= 64); bitwise agreement with an
                            independent numpy restatement of the recurrence FFT
                            (correlation_fft.hpp:38-73); inverse(forward(x)) == x
causal_conv_fft             direct O(N^2) causal sum
phases (mode CR)            quad-precision sinq/cosq (libquadmath) on random and grid
                            arguments; symmetry exp(-ix) = conj(exp(ix))
fcr_series (Convolution)    the reference's own Direct method cross-check
                            (tests/tcl4_tests.cpp:188-212); constant-Gamma => 0;
                            R(0) == 0 exactly when Gamma(.,0) == 0
fcr_series (Direct)         brute-force quadrature sums written from tcl4_kernels.m
triple_kernels              per-triple equality with fcr_series; slicing (keep < Nt)
                            equals the head of the full series bitwise
bath exp (mode CR)          70-digit decimal exp on random, large and subnormal-range
                            arguments (correctly rounded exp of J and the KMS factors)
bcf_fft                     Ohmic closed forms C(t) at zero and finite temperature
                            (coth series), converging at second order in the frequency
                            step (tests/test_oracle_pins.py); brute-force DFT of the KMS
                            spectrum (bcf_fft_ohmic_simple.m)
gamma_trapz / gamma_rect    geometric-series closed form for C(t) = exp(-g t)
mikx                        make_val fixture of cpp/tests/tcl4_mikx_tests.cpp:49-160
                            with the code's M1 order (tcl4_mikx.cpp:111-113)
assemble_gw / gw_to_l4      with mikx and fcr_series: first-order convergence in dt
                            to the fourth-order TCL generator evaluated by brute force
                            (ordered cumulants + Wick, tests/tcl4_bruteforce.py, itself
                            checked against exact spin + bosonic-mode dynamics) at
                            N = 2, 3, 4; a sign planted in any of the ten NAKZWAN terms
                            breaks it; pure dephasing gives L4 == 0 exactly (the
                            planted signs do not); GW == GW^H; L4 preserves Hermiticity
tcl2_at                     equals L0 + the secular part of the brute-force K2 for the
                            Gamma table read at -omega_b (the reference's frequency
                            convention for A(w)); trace preservation of L2
rk4_update / propagate      rk4_dense_tests.cpp closed forms (expm, exp(t^2/2));
                            one step == 4th-order Taylor polynomial
system_build                residual ||HU - U diag(eps)||, unitarity, ascending
                            order, bucket counts of the configs; eigenvector phase
                            and degenerate gauge: parity unpinned (Eigen absent)
==========================  ==========================================================
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_size_t

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = POINTER(c_double)
_ip = POINTER(c_int)
_sp = POINTER(c_size_t)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "liboracle.so", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_last_error": (ctypes.c_char_p, []),
            "orc_next_pow2": (c_size_t, [c_size_t]),
            "orc_fcr_nfft": (c_size_t, [c_size_t, c_size_t]),
            "orc_fft": (c_int, [_dp, c_size_t, c_int]),
            "orc_bcf_fft": (c_int, [c_size_t, c_double, c_int, c_double, c_double, c_double, _dp]),
            "orc_gamma_trapz": (c_int, [_dp, c_size_t, c_double, _dp, c_int, _dp]),
            "orc_gamma_rect": (c_int, [_dp, c_size_t, c_double, _dp, c_int, _dp]),
            "orc_fcr_series": (c_int, [_dp, _dp, c_size_t, c_double, c_double, c_int, c_int, c_size_t, _dp]),
            "orc_causal_conv_fft": (c_int, [_dp, _dp, c_size_t, c_double, c_size_t, _dp]),
            "orc_mirror": (c_int, [_dp, c_int, c_double, _ip]),
            "orc_triple_kernels": (c_int, [_dp, c_size_t, c_int, _dp, _ip, c_double, c_int, c_size_t,
                                           c_size_t, _dp, _dp, _dp, c_int]),
            "orc_mikx": (c_int, [_dp, _dp, _dp, _ip, c_int, c_int, _dp, _dp, _dp, _dp]),
            "orc_assemble_gw": (c_int, [_dp, _dp, _dp, _dp, _dp, c_int, c_int, _dp]),
            "orc_gw_to_l4": (c_int, [_dp, c_int, _dp]),
            "orc_l4_from_slices": (c_int, [_dp, _dp, _dp, _ip, c_int, c_int, _dp, c_int, _dp]),
            "orc_l4_series": (c_int, [_dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, c_int, c_double,
                                      c_int, c_size_t, _sp, c_size_t, _dp, c_int]),
            "orc_tcl2_at": (c_int, [_dp, _dp, c_int, c_int, _ip, c_int, _dp, _dp]),
            "orc_tcl2_general_at": (c_int, [_dp, _dp, c_int, c_int, _ip, c_int, _dp, _dp]),
            "orc_rk4_update": (c_int, [_dp, _dp, _dp, _dp, c_size_t, c_double]),
            "orc_propagate": (c_int, [_dp, _dp, c_size_t, c_size_t, _dp, c_double, c_size_t, _dp, _sp]),
            "orc_tcl4_propagate": (c_int, [_dp, c_size_t, c_int, _dp, _ip, _ip, c_int, _dp, _dp, c_int,
                                           c_double, c_int, c_size_t, c_size_t, c_int, _dp, _dp, c_int]),
            "orc_system_build": (c_int, [_dp, c_int, _dp, c_int, c_double, _dp, _dp, _dp, _ip, _dp, _ip,
                                         _ip, _ip]),
            "orc_set_phase_mode": (None, [c_int]),
            "orc_get_phase_mode": (c_int, []),
            "orc_set_mutation": (None, [c_int]),
            "orc_phases": (None, [c_double, c_double, c_size_t, _dp]),
            "orc_cr_sincos": (None, [_dp, c_size_t, _dp, _dp]),
            "orc_cr_exp": (None, [_dp, c_size_t, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _chk(rc: int) -> None:
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _c(a, dtype=np.complex128):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def _fortran_c128(a):
    """Column-major complex128 copy (the reference's Eigen layout)."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def default_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# --- phase convention ---------------------------------------------------------
PHASE_LIBM = 0  # std::exp(i*x) of the platform libm (what the reference calls)
PHASE_CR = 1    # correctly rounded cos/sin of the same argument


def set_phase_mode(mode: int) -> None:
    lib().orc_set_phase_mode(int(mode))


class planted_mistake:
    """Context manager: NAKZWAN term q (1..10) enters with the wrong sign
    (tests only: shows that a check catches such a mistake)."""

    def __init__(self, term: int):
        self.term = term

    def __enter__(self):
        lib().orc_set_mutation(int(self.term))
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)


def get_phase_mode() -> int:
    return int(lib().orc_get_phase_mode())


class phase_mode:
    """Context manager selecting the oracle's phase convention."""

    def __init__(self, mode: int):
        self.mode = mode

    def __enter__(self):
        self.prev = get_phase_mode()
        set_phase_mode(self.mode)

    def __exit__(self, *exc):
        set_phase_mode(self.prev)


def phases(omega: float, dt: float, n: int) -> np.ndarray:
    out = np.zeros(n, np.complex128)
    lib().orc_phases(omega, dt, n, _p(out))
    return out


def cr_exp(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().orc_cr_exp(_p(x), x.size, _p(out))
    return out


def cr_sincos(x):
    x = np.ascontiguousarray(x, np.float64)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().orc_cr_sincos(_p(x), x.size, _p(s), _p(c))
    return s, c


# --- primitives -------------------------------------------------------------

def next_pow2(n: int) -> int:
    return int(lib().orc_next_pow2(n))


def fcr_nfft(Nt: int, pad: int = 0) -> int:
    return int(lib().orc_fcr_nfft(Nt, pad))


def fft(x, forward: bool = True) -> np.ndarray:
    a = _c(x).copy()
    _chk(lib().orc_fft(_p(a), a.size, 1 if forward else 0))
    return a


def bcf_fft(N: int, dt: float, beta: float, kind: int = 0, p0: float = 1.0, p1: float = 1.0) -> np.ndarray:
    """C(t) at t = 0..N*dt (N+1 samples); kind 0 Ohmic (alpha, wc), 1 Drude-Lorentz (lambda, gamma)."""
    out = np.zeros(N + 1, np.complex128)
    _chk(lib().orc_bcf_fft(N, dt, kind, p0, p1, beta, _p(out)))
    return out


def gamma_trapz(C, dt: float, omegas) -> np.ndarray:
    """Gamma table [Nt, nf] (column-major storage, returned as a numpy array)."""
    C = _c(C)
    om = np.ascontiguousarray(omegas, np.float64)
    G = np.zeros((len(om), C.size), np.complex128)
    _chk(lib().orc_gamma_trapz(_p(C), C.size, dt, _p(om), len(om), _p(G)))
    return G.T


def gamma_rect(C, dt: float, omegas) -> np.ndarray:
    C = _c(C)
    om = np.ascontiguousarray(omegas, np.float64)
    G = np.zeros((len(om), C.size), np.complex128)
    _chk(lib().orc_gamma_rect(_p(C), C.size, dt, _p(om), len(om), _p(G)))
    return G.T


def fcr_series(g1, g2, omega: float, dt: float, which: str, method: str = "conv", pad: int = 0) -> np.ndarray:
    g1 = _c(g1)
    g2 = _c(g2)
    out = np.zeros(g1.size, np.complex128)
    w = {"F": 0, "C": 1, "R": 2}[which]
    m = {"conv": 0, "direct": 1}[method]
    _chk(lib().orc_fcr_series(_p(g1), _p(g2), g1.size, omega, dt, w, m, pad, _p(out)))
    return out


def causal_conv_fft(f, g, dt: float, pad: int = 0) -> np.ndarray:
    f = _c(f)
    g = _c(g)
    out = np.zeros(f.size, np.complex128)
    _chk(lib().orc_causal_conv_fft(_p(f), _p(g), f.size, dt, pad, _p(out)))
    return out


def mirror(omegas, tol: float = 1e-9) -> np.ndarray:
    om = np.ascontiguousarray(omegas, np.float64)
    out = np.zeros(len(om), np.int32)
    _chk(lib().orc_mirror(_p(om), len(om), tol, _pi(out)))
    return out


def gamma_colmajor(gamma) -> np.ndarray:
    """Flat column-major buffer of a [Nt, nf] Gamma table."""
    return np.ascontiguousarray(np.asarray(gamma, np.complex128).T)


def triple_kernels(gamma, omegas, mirror_idx, dt: float, keep: int | None = None, method: str = "conv",
                   pad: int = 0, nthreads: int | None = None):
    """F, C, R arrays of shape [nf, nf, nf, keep] (series of the first `keep` samples)."""
    Nt, nf = np.asarray(gamma).shape
    keep = Nt if keep is None else keep
    g = gamma_colmajor(gamma)
    om = np.ascontiguousarray(omegas, np.float64)
    mi = np.ascontiguousarray(mirror_idx, np.int32)
    F = np.zeros((nf, nf, nf, keep), np.complex128)
    C = np.zeros_like(F)
    R = np.zeros_like(F)
    _chk(lib().orc_triple_kernels(_p(g), Nt, nf, _p(om), _pi(mi), dt, 0 if method == "conv" else 1, pad, keep,
                                  _p(F), _p(C), _p(R), nthreads or default_threads()))
    return F, C, R


def mikx(Fs, Cs, Rs, p2f, N: int):
    """M, I, K as (N^2 x N^2) column-major arrays (returned as numpy [row, col]) and X flat N^6."""
    nf = int(round(np.asarray(Fs).size ** (1 / 3)))
    Fs, Cs, Rs = _c(Fs).ravel(), _c(Cs).ravel(), _c(Rs).ravel()
    p = np.ascontiguousarray(np.asarray(p2f, np.int32).T.ravel())  # column-major
    N2 = N * N
    M = np.zeros(N2 * N2, np.complex128)
    I = np.zeros_like(M)
    K = np.zeros_like(M)
    X = np.zeros(N2 * N2 * N2, np.complex128)
    _chk(lib().orc_mikx(_p(Fs), _p(Cs), _p(Rs), _pi(p), N, nf, _p(M), _p(I), _p(K), _p(X)))
    col = lambda a: a.reshape(N2, N2).T  # noqa: E731
    return col(M), col(I), col(K), X


def assemble_gw(M, I, K, X, ops, N: int) -> np.ndarray:
    f = lambda a: np.ascontiguousarray(np.asarray(a, np.complex128).T).ravel()  # noqa: E731
    ops_flat = np.concatenate([f(o) for o in ops])
    Xf = _c(X).ravel()
    GW = np.zeros(N ** 4, np.complex128)
    Mf, If, Kf = f(M), f(I), f(K)
    _chk(lib().orc_assemble_gw(_p(Mf), _p(If), _p(Kf), _p(Xf), _p(ops_flat), len(ops), N, _p(GW)))
    return GW.reshape(N * N, N * N).T


def gw_to_l4(GW, N: int) -> np.ndarray:
    g = np.ascontiguousarray(np.asarray(GW, np.complex128).T).ravel()
    L = np.zeros(N ** 4, np.complex128)
    _chk(lib().orc_gw_to_l4(_p(g), N, _p(L)))
    return L.reshape(N * N, N * N).T


def _ops_flat(ops):
    return np.concatenate([np.ascontiguousarray(np.asarray(o, np.complex128).T).ravel() for o in ops])


def _p2f_flat(p2f):
    return np.ascontiguousarray(np.asarray(p2f, np.int32).T.ravel())


def l4_series(gamma, omegas, mirror_idx, p2f, ops, dt: float, tidx, method: str = "conv", pad: int = 0,
              nthreads: int | None = None) -> np.ndarray:
    """L4(t) for t in tidx, shape [ntau, N^2, N^2] (each block a normal numpy matrix)."""
    Nt, nf = np.asarray(gamma).shape
    N = np.asarray(p2f).shape[0]
    g = gamma_colmajor(gamma)
    om = np.ascontiguousarray(omegas, np.float64)
    mi = np.ascontiguousarray(mirror_idx, np.int32)
    pf = _p2f_flat(p2f)
    of = _ops_flat(ops)
    td = np.ascontiguousarray(tidx, np.uint64)
    out = np.zeros((len(td), N ** 4), np.complex128)
    _chk(lib().orc_l4_series(_p(g), Nt, nf, _p(om), _pi(mi), _pi(pf), N, _p(of), len(ops), dt,
                             0 if method == "conv" else 1, pad, td.ctypes.data_as(_sp), len(td), _p(out),
                             nthreads or default_threads()))
    return out.reshape(len(td), N * N, N * N).transpose(0, 2, 1)


def tcl2_at(eps, ops, p2f, gamma_row) -> np.ndarray:
    N = len(eps)
    e = np.ascontiguousarray(eps, np.float64)
    of = _ops_flat(ops)
    pf = _p2f_flat(p2f)
    gr = _c(gamma_row)
    L = np.zeros(N ** 4, np.complex128)
    _chk(lib().orc_tcl2_at(_p(e), _p(of), len(ops), N, _pi(pf), gr.size, _p(gr), _p(L)))
    return L.reshape(N * N, N * N).T


def tcl2_general_at(eps, ops, p2f, gamma_ab) -> np.ndarray:
    """TCL2 generator from channel matrices gamma_ab [nf, num_ops, num_ops]
    (gamma_ab[b, alpha, beta] = Gamma_b(alpha, beta) at one t)."""
    N = len(eps)
    e = np.ascontiguousarray(eps, np.float64)
    of = _ops_flat(ops)
    pf = _p2f_flat(p2f)
    g = np.asarray(gamma_ab, np.complex128)
    nf = g.shape[0]
    gf = np.ascontiguousarray(g.transpose(0, 2, 1)).ravel()  # [b][beta][alpha]
    L = np.zeros(N ** 4, np.complex128)
    _chk(lib().orc_tcl2_general_at(_p(e), _p(of), len(ops), N, _pi(pf), nf, _p(gf), _p(L)))
    return L.reshape(N * N, N * N).T


def rk4_update(L0, Lh, L1, r, dt: float) -> np.ndarray:
    f = lambda a: np.ascontiguousarray(np.asarray(a, np.complex128).T).ravel()  # noqa: E731
    a0, ah, a1 = f(L0), f(Lh), f(L1)
    rv = _c(r).copy()
    _chk(lib().orc_rk4_update(_p(a0), _p(ah), _p(a1), _p(rv), rv.size, dt))
    return rv


def propagate(L_series, r0, dt: float, L_half=None, every: int = 1):
    """propagate_rk4_dense_serial over L_series [steps+1, D, D]; returns (r_final, samples)."""
    Ls = np.asarray(L_series, np.complex128)
    steps = Ls.shape[0] - 1
    D = Ls.shape[1]
    flat = np.ascontiguousarray(Ls.transpose(0, 2, 1)).ravel()
    hp = None
    if L_half is not None:
        Lh = np.ascontiguousarray(np.asarray(L_half, np.complex128).transpose(0, 2, 1)).ravel()
        hp = _p(Lh)
    rv = _c(r0).copy()
    samples = np.zeros((steps // every + 2, D), np.complex128)
    ns = c_size_t(0)
    _chk(lib().orc_propagate(_p(flat), hp, steps, D, _p(rv), dt, every, _p(samples), ctypes.byref(ns)))
    return rv, samples[: ns.value]


def tcl4_propagate(gamma, omegas, mirror_idx, p2f, eps, ops, dt: float, steps: int, r0, order: int = 4,
                   method: str = "conv", pad: int = 0, nthreads: int | None = None) -> np.ndarray:
    """Driver composition of TCL4_spin_boson_example: states r(t_0..t_steps), shape [steps+1, N^2]."""
    Nt, nf = np.asarray(gamma).shape
    N = len(eps)
    g = gamma_colmajor(gamma)
    om = np.ascontiguousarray(omegas, np.float64)
    mi = np.ascontiguousarray(mirror_idx, np.int32)
    pf = _p2f_flat(p2f)
    of = _ops_flat(ops)
    e = np.ascontiguousarray(eps, np.float64)
    r = _c(r0).copy()
    samples = np.zeros((steps + 1, N * N), np.complex128)
    _chk(lib().orc_tcl4_propagate(_p(g), Nt, nf, _p(om), _pi(mi), _pi(pf), N, _p(e), _p(of), len(ops), dt,
                                  0 if method == "conv" else 1, pad, steps, order, _p(r), _p(samples),
                                  nthreads or default_threads()))
    return samples


def system_build(H, ops_lab, freq_tol: float = 1e-9) -> dict:
    H = np.asarray(H, np.complex128)
    N = H.shape[0]
    Hf = np.ascontiguousarray(H.T).ravel()
    of = _ops_flat(ops_lab)
    eps = np.zeros(N)
    U = np.zeros(N * N, np.complex128)
    Ae = np.zeros(len(ops_lab) * N * N, np.complex128)
    nf = c_int(0)
    om = np.zeros(N * N)
    mi = np.zeros(N * N, np.int32)
    pf = np.zeros(N * N, np.int32)
    zi = c_int(0)
    _chk(lib().orc_system_build(_p(Hf), N, _p(of), len(ops_lab), freq_tol, _p(eps), _p(U), _p(Ae),
                                ctypes.byref(nf), _p(om), _pi(mi), _pi(pf), ctypes.byref(zi)))
    n = nf.value
    return {
        "N": N,
        "eps": eps,
        "U": U.reshape(N, N).T.copy(),
        "A_eig": [Ae[o * N * N:(o + 1) * N * N].reshape(N, N).T.copy() for o in range(len(ops_lab))],
        "nf": n,
        "omegas": om[:n].copy(),
        "mirror": mi[:n].copy(),
        "pair_to_freq": pf.reshape(N, N).T.copy(),
        "zero_index": zi.value,
    }
