"""Host-side mirror of the reference TCL4 interface, served by the HIP engine.

Names, argument meaning and error behaviour follow jcarrots/taco:

* ``Exec`` / ``Backend``              cpp/include/taco/exec.hpp:11-19
* ``FCRMethod``                        cpp/include/taco/tcl4_kernels.hpp:9-12
* ``System`` (``sys::System``)         cpp/include/taco/system.hpp:201-226
* ``build_map`` / ``Tcl4Map``          cpp/include/taco/tcl4.hpp:13-28, cpp/src/tcl/tcl4.cpp:40-101
* ``compute_triple_kernels``           tcl4.hpp:37-42, tcl4.cpp:103-183
* ``build_TCL4_generator_hip_fused``   tcl4_fused_cuda.hpp:16-21 (the CUDA name is an alias)
* ``build_TCL4_generator_hip_fused_batch``  tcl4_fused_cuda.hpp:25-41
* ``build_correction_series_hip_fused``      tcl4_fused_cuda.hpp:44-48
* ``build_TCL4_generator`` / ``build_correction_series``  tcl4.hpp:83-95
* ``propagate_rk4_dense``              cpp/include/taco/rk4_dense.hpp:85-129

Matrices are numpy arrays indexed ``[row, col]``; the engine receives them
column-major (Eigen's layout).  Exceptions map the reference's types:
``TacoInvalidArgument`` (ValueError) for std::invalid_argument,
``TacoOutOfRange`` (IndexError) for std::out_of_range, ``TacoError``
(RuntimeError) for std::runtime_error, ``TacoOverflow`` for std::overflow_error.
"""
from __future__ import annotations

import ctypes
import enum
import threading
from ctypes import c_double, c_int, c_size_t, c_void_p
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from ._lib import TacoInvalidArgument, check, dptr, iptr, lib, sptr

__all__ = [
    "Backend", "Exec", "FCRMethod", "System", "Tcl4Map", "TripleKernelSeries", "Ohmic", "DrudeLorentz",
    "build_map", "bcf_fft_fun", "compute_trapz_prefix_multi_matrix", "bath_gamma_hip", "SpectralExpr",
    "compute_rect_prefix_multi_matrix",
    "compute_trapz_prefix_multi_matrix_hip", "compute_triple_kernels",
    "build_TCL4_generator_hip_fused", "build_TCL4_generator_hip_fused_batch", "build_correction_series_hip_fused",
    "build_TCL4_generator_cuda_fused", "build_TCL4_generator_cuda_fused_batch", "build_correction_series_cuda_fused",
    "build_TCL4_generator", "build_correction_series", "assemble_l4_from_slices", "build_tcl2_generator",
    "propagate_rk4_dense",
    "propagate_rk4_dense_serial", "propagate_rk4_dense_omp", "tcl4_propagate", "set_fcr_fft_pad_factor",
    "get_fcr_fft_pad_factor", "set_phase_mode", "get_phase_mode", "phase_mode", "PHASE_LIBM", "PHASE_CR",
    "set_fft_mode", "get_fft_mode", "fft_mode", "FFT_PARITY", "FFT_WINDOW", "set_slice_budget", "get_slice_budget",
    "context", "version", "BatchItem", "Batch", "batched_run",
]


class Backend(enum.Enum):
    Serial = 0
    Omp = 1
    Cuda = 2
    MpiOmp = 3
    MpiCuda = 4
    Hip = 5


@dataclass
class Exec:
    """exec.hpp:13-19.  Serial/Omp/Cuda/Hip requests are all served by the
    HIP engine on ``gpu_id``; the MPI backends are not part of this engine."""
    backend: Backend = Backend.Hip
    threads: int = 0
    gpu_id: int = 0
    streams: int = 1
    pinned: bool = True


class FCRMethod(enum.IntEnum):
    Convolution = 0
    Direct = 1


def version() -> str:
    return lib().taco_version().decode()


def set_fcr_fft_pad_factor(factor: int) -> None:
    """tcl4_kernels.cpp:179-181 (process-global)."""
    lib().taco_set_fcr_fft_pad_factor(int(factor))


def get_fcr_fft_pad_factor() -> int:
    return int(lib().taco_get_fcr_fft_pad_factor())


PHASE_LIBM = 0  # the platform libm's std::exp values (the reference's arithmetic); default
PHASE_CR = 1    # correctly rounded cos/sin/exp of the same arguments, evaluated on the device


def set_phase_mode(mode: int) -> None:
    """Process-global phase/exponential convention (include/taco_tcl4.h):
    PHASE_LIBM reproduces the reference's std::exp calls
    (tcl4_kernels.cpp:116-120, gamma.hpp:185, correlation_fft.hpp:154-214)
    through host-evaluated tables; PHASE_CR uses correctly rounded values."""
    check(lib().taco_set_phase_mode(int(mode)))


def get_phase_mode() -> int:
    return int(lib().taco_get_phase_mode())


FFT_PARITY = 0  # the reference's recurrence-twiddle FFT over the whole Gamma table (default)
FFT_WINDOW = 1  # only the rows t < T enter; Nfft from T; exact twiddles (no reference counterpart)


def set_fft_mode(mode: int) -> None:
    """Process-global F/C/R FFT mode (include/taco_tcl4.h): FFT_PARITY
    reproduces the reference's spectra bit for bit; FFT_WINDOW computes the
    same causal convolutions at t < T from the rows t < T only, with an FFT of
    next_pow2(2T - 1) and exact twiddles."""
    check(lib().taco_set_fft_mode(int(mode)))


def get_fft_mode() -> int:
    return int(lib().taco_get_fft_mode())


class fft_mode:
    """Context manager that switches the FFT mode and restores it."""

    def __init__(self, mode: int):
        self.mode = mode

    def __enter__(self):
        self.prev = get_fft_mode()
        set_fft_mode(self.mode)
        return self

    def __exit__(self, *exc):
        set_fft_mode(self.prev)


class phase_mode:
    """Context manager that switches the phase mode and restores it."""

    def __init__(self, mode: int):
        self.mode = mode

    def __enter__(self):
        self.prev = get_phase_mode()
        set_phase_mode(self.mode)
        return self

    def __exit__(self, *exc):
        set_phase_mode(self.prev)


def set_slice_budget(nbytes: int, device: int = 0) -> None:
    """Bytes of F/C/R slices one pass of the L4 builders may hold on this
    thread's context (0: default, 32 GiB).  Longer time sets run as windows
    of consecutive kept times, bit-identical to one pass
    (backend/cpu/tcl4_mpi_omp.cpp:62-156 splits the times the same way)."""
    check(lib().taco_set_slice_budget(context(device).handle, int(nbytes)))


def get_slice_budget(device: int = 0) -> int:
    return int(lib().taco_get_slice_budget(context(device).handle))


# --- contexts: one per (host thread, device), like the reference's thread_local workspace
_tls = threading.local()


class _Ctx:
    def __init__(self, device: int):
        h = c_void_p()
        check(lib().taco_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def __del__(self):
        try:
            if self.handle:
                lib().taco_ctx_destroy(self.handle)
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return int(lib().taco_ctx_stream(self.handle) or 0)


def context(device: int = 0) -> _Ctx:
    cache = getattr(_tls, "ctx", None)
    if cache is None:
        cache = _tls.ctx = {}
    if device not in cache:
        cache[device] = _Ctx(device)
    return cache[device]


def _ctx_for(exec_: Optional[Exec]) -> _Ctx:
    e = exec_ or Exec()
    if e.backend in (Backend.MpiOmp, Backend.MpiCuda):
        raise TacoInvalidArgument(1, "MPI backends are not provided by the single-GPU HIP engine")
    return context(e.gpu_id)


# --- inputs ------------------------------------------------------------------

def _colmajor(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, np.complex128).T)


def _ops_flat(ops) -> np.ndarray:
    return np.concatenate([_colmajor(o).ravel() for o in ops]) if len(ops) else np.zeros(0, np.complex128)


@dataclass
class Tcl4Map:
    N: int
    nf: int
    omegas: np.ndarray
    pair_to_freq: np.ndarray
    mirror_index: np.ndarray
    zero_index: int
    time_grid: Sequence[float] = ()


@dataclass
class System:
    """sys::System: eigensystem of H, Bohr-frequency buckets and A_eig."""
    H: np.ndarray
    A_lab: list
    eps: np.ndarray
    U: np.ndarray
    A_eig: list
    omegas: np.ndarray
    mirror: np.ndarray
    pair_to_freq: np.ndarray
    zero_index: int
    freq_tol: float = 1e-9

    @property
    def N(self) -> int:
        return int(self.H.shape[0])

    @property
    def nf(self) -> int:
        return int(len(self.omegas))

    @classmethod
    def build(cls, H, jump_ops_lab, freq_tol: float = 1e-9) -> "System":
        H = np.asarray(H, np.complex128)
        if H.ndim != 2 or H.shape[0] != H.shape[1]:
            raise TacoInvalidArgument(1, "Eigensystem: Hamiltonian must be square")
        N = H.shape[0]
        ops = [np.asarray(o, np.complex128) for o in jump_ops_lab]
        for o in ops:
            if o.shape != (N, N):
                raise TacoInvalidArgument(1, "to_eigen: dimension mismatch")
        Hf = _colmajor(H).ravel()
        of = _ops_flat(ops)
        eps = np.zeros(N)
        U = np.zeros(N * N, np.complex128)
        Ae = np.zeros(max(len(ops), 1) * N * N, np.complex128)
        nf = c_int(0)
        om = np.zeros(N * N)
        mi = np.zeros(N * N, np.int32)
        pf = np.zeros(N * N, np.int32)
        zi = c_int(0)
        check(lib().taco_system_build(dptr(Hf), N, dptr(of) if of.size else None, len(ops), freq_tol, dptr(eps),
                                      dptr(U), dptr(Ae), ctypes.byref(nf), dptr(om), iptr(mi), iptr(pf),
                                      ctypes.byref(zi)))
        n = nf.value
        return cls(H=H, A_lab=ops, eps=eps, U=U.reshape(N, N).T.copy(),
                   A_eig=[Ae[k * N * N:(k + 1) * N * N].reshape(N, N).T.copy() for k in range(len(ops))],
                   omegas=om[:n].copy(), mirror=mi[:n].copy(), pair_to_freq=pf.reshape(N, N).T.copy(),
                   zero_index=zi.value, freq_tol=freq_tol)

    @classmethod
    def build_many(cls, Hs, jump_ops_lab, freq_tol: float = 1e-9, exec: Optional[Exec] = None) -> list:
        """System.build of several Hamiltonians of one size on the GPU (a
        workgroup each); every System equals System.build's bit for bit.
        jump_ops_lab: one list of operators shared by all, or one list per H."""
        Hs = [np.asarray(H, np.complex128) for H in Hs]
        if not Hs:
            return []
        N = Hs[0].shape[0]
        for H in Hs:
            if H.shape != (N, N):
                raise TacoInvalidArgument(1, "Eigensystem: Hamiltonians must be square and of one size")
        per = jump_ops_lab if (len(jump_ops_lab) and isinstance(jump_ops_lab[0], (list, tuple))) else \
            [jump_ops_lab] * len(Hs)
        if len(per) != len(Hs) or len({len(o) for o in per}) != 1:
            raise TacoInvalidArgument(1, "one operator list per Hamiltonian, all of one length")
        no = len(per[0])
        S = len(Hs)
        ops = [[np.asarray(o, np.complex128) for o in ol] for ol in per]
        for ol in ops:
            for o in ol:
                if o.shape != (N, N):
                    raise TacoInvalidArgument(1, "to_eigen: dimension mismatch")
        Hf = np.concatenate([_colmajor(H).ravel() for H in Hs])
        of = np.concatenate([_ops_flat(ol) for ol in ops]) if no else np.zeros(1, np.complex128)
        eps = np.zeros(S * N)
        U = np.zeros(S * N * N, np.complex128)
        Ae = np.zeros(max(no, 1) * S * N * N, np.complex128)
        nf = np.zeros(S, np.int32)
        om = np.zeros(S * N * N)
        mi = np.zeros(S * N * N, np.int32)
        pf = np.zeros(S * N * N, np.int32)
        zi = np.zeros(S, np.int32)
        ctx = _ctx_for(exec)
        check(lib().taco_system_build_hip(ctx.handle, S, dptr(Hf), N, dptr(of), no, freq_tol, dptr(eps), dptr(U),
                                          dptr(Ae), iptr(nf), dptr(om), iptr(mi), iptr(pf), iptr(zi)))
        out = []
        NN = N * N
        for s in range(S):
            n = int(nf[s])
            out.append(cls(H=Hs[s], A_lab=ops[s], eps=eps[s * N:(s + 1) * N].copy(),
                           U=U[s * NN:(s + 1) * NN].reshape(N, N).T.copy(),
                           A_eig=[Ae[(s * no + k) * NN:(s * no + k + 1) * NN].reshape(N, N).T.copy()
                                  for k in range(no)],
                           omegas=om[s * NN:s * NN + n].copy(), mirror=mi[s * NN:s * NN + n].copy(),
                           pair_to_freq=pf[s * NN:(s + 1) * NN].reshape(N, N).T.copy(),
                           zero_index=int(zi[s]), freq_tol=freq_tol))
        return out

    def to_lab(self, A_eig) -> np.ndarray:
        return self.U @ np.asarray(A_eig) @ self.U.conj().T

    def to_eigen(self, A_lab) -> np.ndarray:
        return self.U.conj().T @ np.asarray(A_lab) @ self.U


def build_map(system: System, time_grid: Sequence[float] = ()) -> Tcl4Map:
    return Tcl4Map(N=system.N, nf=system.nf, omegas=system.omegas.copy(), pair_to_freq=system.pair_to_freq.copy(),
                   mirror_index=system.mirror.copy(), zero_index=system.zero_index, time_grid=list(time_grid))


@dataclass(frozen=True)
class Ohmic:
    """J(w) = alpha * w * exp(-w / omega_c) for w > 0 (TCL4_spin_boson_example.cpp:210-213)."""
    alpha: float
    omega_c: float
    kind: int = field(default=0, init=False)

    @property
    def params(self):
        return self.alpha, self.omega_c


@dataclass(frozen=True)
class DrudeLorentz:
    """J(w) = 2 lambda gamma w / (w^2 + gamma^2) for w > 0."""
    lam: float
    gamma: float
    kind: int = field(default=1, init=False)

    @property
    def params(self):
        return self.lam, self.gamma


class SpectralExpr:
    """A spectral density J(w) written as an expression, as the reference's
    configuration files give it (bath.J_expr with bath.params;
    expression.hpp:14-26), compiled once by the engine's host code:
    SpectralExpr("w * exp(-w/omega_c) * step(w)", omega_c=10.0)(w)."""

    def __init__(self, source: str, **params):
        names = list(params)
        arr = (ctypes.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
        vals = np.ascontiguousarray([float(params[n]) for n in names] or [0.0], np.float64)
        h = c_void_p()
        check(lib().taco_spectral_create(source.encode(), arr, dptr(vals), len(names), ctypes.byref(h)))
        self.handle = h
        self.source = source
        self.params = dict(params)

    def __call__(self, w):
        w = np.ascontiguousarray(np.atleast_1d(np.asarray(w, np.float64)))
        out = np.zeros_like(w)
        check(lib().taco_spectral_eval(self.handle, dptr(w), w.size, dptr(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().taco_spectral_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def bcf_fft_fun(N: int, dt: float, J, beta: float):
    """correlation_fft.hpp:154-214: returns (t[0..N], C[0..N]).  J is an
    Ohmic, a DrudeLorentz or a SpectralExpr."""
    C = np.zeros(N + 1, np.complex128)
    if isinstance(J, SpectralExpr):
        check(lib().taco_bcf_fft_spectral(N, dt, J.handle, beta, dptr(C)))
        return np.arange(N + 1) * dt, C
    p0, p1 = J.params
    check(lib().taco_bcf_fft(N, dt, J.kind, p0, p1, beta, dptr(C)))
    return np.arange(N + 1) * dt, C


def compute_trapz_prefix_multi_matrix(C, dt: float, omegas) -> np.ndarray:
    """gamma.hpp:166-213: Gamma table [Nt, nf]."""
    C = np.ascontiguousarray(C, np.complex128)
    om = np.ascontiguousarray(omegas, np.float64)
    G = np.zeros((len(om), C.size), np.complex128)
    check(lib().taco_gamma_trapz(dptr(C), C.size, dt, dptr(om), len(om), dptr(G)))
    return G.T


def compute_rect_prefix_multi_matrix(C, dt: float, omegas) -> np.ndarray:
    """gamma.hpp:127-160 (rectangle rule, the tcl_driver default): Gamma [Nt, nf]."""
    C = np.ascontiguousarray(C, np.complex128)
    om = np.ascontiguousarray(omegas, np.float64)
    G = np.zeros((len(om), C.size), np.complex128)
    check(lib().taco_gamma_rect(dptr(C), C.size, dt, dptr(om), len(om), dptr(G)))
    return G.T


def bath_gamma_hip(N: int, dt: float, J, beta: float, omegas=None, exec: Optional[Exec] = None):
    """GPU twin of bcf_fft_fun + compute_trapz_prefix_multi_matrix, bit-identical
    to them: returns (C[0..N], Gamma [N+1, nf]) (Gamma None without omegas)."""
    ctx = _ctx_for(exec)
    C = np.zeros(N + 1, np.complex128)
    p0, p1 = J.params
    if omegas is None:
        check(lib().taco_bath_gamma_hip(ctx.handle, N, dt, J.kind, p0, p1, beta, None, 0, dptr(C), None))
        return C, None
    om = np.ascontiguousarray(omegas, np.float64)
    G = np.zeros((len(om), N + 1), np.complex128)
    check(lib().taco_bath_gamma_hip(ctx.handle, N, dt, J.kind, p0, p1, beta, dptr(om), len(om), dptr(C), dptr(G)))
    return C, G.T


def compute_trapz_prefix_multi_matrix_hip(C, dt: float, omegas, exec: Optional[Exec] = None) -> np.ndarray:
    """GPU twin of compute_trapz_prefix_multi_matrix (bit-identical): Gamma [Nt, nf]."""
    ctx = _ctx_for(exec)
    C = np.ascontiguousarray(C, np.complex128)
    om = np.ascontiguousarray(omegas, np.float64)
    G = np.zeros((len(om), C.size), np.complex128)
    check(lib().taco_gamma_trapz_hip(ctx.handle, dptr(C), C.size, dt, dptr(om), len(om), dptr(G)))
    return G.T


# --- hot path ----------------------------------------------------------------

@dataclass
class TripleKernelSeries:
    """tcl4.hpp:30-35: F[i][j][k](t) as arrays of shape [nf, nf, nf, Nt]."""
    F: np.ndarray
    C: np.ndarray
    R: np.ndarray


def _gamma_checked(system: System, gamma_series) -> np.ndarray:
    g = np.asarray(gamma_series, np.complex128)
    if g.ndim != 2:
        raise TacoInvalidArgument(1, "gamma_series must be a [Nt, nf] matrix")
    if g.shape[1] != system.nf:
        raise TacoInvalidArgument(1, "gamma_series column count does not match frequency buckets")
    return g


def compute_triple_kernels(system: System, gamma_series, dt: float, nmax: int = 2,
                           method: FCRMethod = FCRMethod.Convolution, exec: Optional[Exec] = None,
                           keep: Optional[int] = None) -> TripleKernelSeries:
    """All nf^3 F/C/R series on the GPU.  ``keep`` (extension) limits the
    returned samples to t < keep; the values equal the full series' head."""
    del nmax  # ignored, as in the reference (tcl4.cpp:117)
    ctx = _ctx_for(exec)
    g = _gamma_checked(system, gamma_series)
    Nt, nf = g.shape
    keep = Nt if keep is None else int(keep)
    gc = _colmajor(g).ravel()
    F = np.zeros((nf, nf, nf, keep), np.complex128)
    C = np.zeros_like(F)
    R = np.zeros_like(F)
    om = np.ascontiguousarray(system.omegas, np.float64)
    mi = np.ascontiguousarray(system.mirror, np.int32)
    check(lib().taco_triple_kernels(ctx.handle, dptr(gc), Nt, nf, dptr(om), iptr(mi), dt, int(method), keep,
                                    dptr(F), dptr(C), dptr(R)))
    return TripleKernelSeries(F, C, R)


def _system_arrays(system: System):
    om = np.ascontiguousarray(system.omegas, np.float64)
    mi = np.ascontiguousarray(system.mirror, np.int32)
    pf = np.ascontiguousarray(np.asarray(system.pair_to_freq, np.int32).T.ravel())
    of = _ops_flat(system.A_eig)
    return om, mi, pf, of


def build_TCL4_generator_hip_fused_batch(system: System, gamma_series, dt: float,
                                         time_indices: Sequence[int] = (),
                                         method: FCRMethod = FCRMethod.Convolution, exec: Optional[Exec] = None,
                                         fcr_ms: Optional[list] = None) -> np.ndarray:
    """L4(t) for each t in time_indices (empty = all t < Nt): [ntau, N^2, N^2].
    As the reference's fused CUDA builders (tcl4_fused_cuda.cu:820-822), only
    FCRMethod.Convolution is accepted; build_TCL4_generator takes Direct too."""
    if FCRMethod(method) != FCRMethod.Convolution:
        raise TacoInvalidArgument(1, "build_TCL4_generator_cuda_fused: only FCRMethod::Convolution is supported")
    return _l4_batch(system, gamma_series, dt, time_indices, method, exec, fcr_ms)


def _l4_batch(system: System, gamma_series, dt: float, time_indices, method, exec, fcr_ms=None) -> np.ndarray:
    ctx = _ctx_for(exec)
    g = _gamma_checked(system, gamma_series)
    Nt, nf = g.shape
    N = system.N
    tid = np.ascontiguousarray(np.asarray(list(time_indices), dtype=np.int64), np.uint64)
    if tid.size and (np.asarray(list(time_indices)) < 0).any():
        from ._lib import TacoOutOfRange
        raise TacoOutOfRange(2, "time_index out of range")
    ntau = tid.size if tid.size else Nt
    out = np.zeros((ntau, N ** 4), np.complex128)
    om, mi, pf, of = _system_arrays(system)
    ms = c_double(0.0)
    gc = _colmajor(g).ravel()
    check(lib().taco_l4_batch(ctx.handle, dptr(gc), Nt, nf, dptr(om), iptr(mi), iptr(pf), N,
                              dptr(of) if of.size else None, len(system.A_eig), dt, int(method),
                              sptr(tid) if tid.size else None, tid.size, dptr(out), ctypes.byref(ms)))
    if fcr_ms is not None:
        fcr_ms.append(ms.value)
    return out.reshape(ntau, N * N, N * N).transpose(0, 2, 1)


def build_TCL4_generator_hip_fused(system: System, gamma_series, dt: float, time_index: int,
                                   method: FCRMethod = FCRMethod.Convolution,
                                   exec: Optional[Exec] = None) -> np.ndarray:
    g = _gamma_checked(system, gamma_series)
    if time_index < 0 or time_index >= g.shape[0]:
        from ._lib import TacoOutOfRange
        raise TacoOutOfRange(2, "build_TCL4_generator_hip_fused: time_index out of range")
    return build_TCL4_generator_hip_fused_batch(system, g, dt, [time_index], method, exec)[0]


def build_correction_series_hip_fused(system: System, gamma_series, dt: float,
                                      method: FCRMethod = FCRMethod.Convolution,
                                      exec: Optional[Exec] = None) -> np.ndarray:
    return build_TCL4_generator_hip_fused_batch(system, gamma_series, dt, (), method, exec)


# CUDA names of the reference, served by the same engine.
build_TCL4_generator_cuda_fused = build_TCL4_generator_hip_fused
build_TCL4_generator_cuda_fused_batch = build_TCL4_generator_hip_fused_batch
build_correction_series_cuda_fused = build_correction_series_hip_fused


def build_TCL4_generator(system: System, gamma_series, dt: float, time_index,
                         method: FCRMethod = FCRMethod.Convolution, exec: Optional[Exec] = None) -> np.ndarray:
    """tcl4.cpp:293-320 (Exec dispatch), every backend on the GPU.  Unlike the
    fused CUDA names it accepts FCRMethod.Direct, as the reference's CPU
    backends do: the F/C/R values then come from the O(T^2) sums of
    tcl4_kernels.cpp:188-281.  time_index: one index (returns N^2 x N^2) or a
    sequence of indices (returns [ntau, N^2, N^2])."""
    g = _gamma_checked(system, gamma_series)
    if np.ndim(time_index) == 0:
        t = int(time_index)
        if t < 0 or t >= g.shape[0]:
            from ._lib import TacoOutOfRange
            raise TacoOutOfRange(2, "build_TCL4_generator: time_index out of range")
        return _l4_batch(system, g, dt, [t], method, exec)[0]
    return _l4_batch(system, g, dt, list(time_index), method, exec)


def build_correction_series(system: System, gamma_series, dt: float, method: FCRMethod = FCRMethod.Convolution,
                            exec: Optional[Exec] = None) -> np.ndarray:
    return _l4_batch(system, gamma_series, dt, (), method, exec)


def build_tcl2_generator(system: System, gamma_ab, time_indices: Sequence[int] = (),
                         exec: Optional[Exec] = None) -> np.ndarray:
    """L_unitary + L_dissipator of build_tcl2_components (generator.cpp:81-177)
    from channel matrices gamma_ab [Nt, nf, num_ops, num_ops]
    (gamma_ab[t, b, alpha, beta] = Gamma_b(alpha, beta)(t)); returns
    [ntau, N^2, N^2] at time_indices (empty: every t)."""
    g = np.asarray(gamma_ab, np.complex128)
    Nt, nf, no = g.shape[0], g.shape[1], g.shape[2]
    if nf != system.nf or no != len(system.A_eig) or g.shape[3] != no:
        raise TacoInvalidArgument(1, "gamma_ab must be [Nt, nf, num_ops, num_ops]")
    N = system.N
    cols = np.ascontiguousarray(g.transpose(1, 3, 2, 0)).reshape(-1)  # [(b*no + beta)*no + alpha][t]
    ops = _ops_flat(system.A_eig)
    p2f = np.ascontiguousarray(np.asarray(system.pair_to_freq, np.int32).T.reshape(-1))
    eps = np.ascontiguousarray(system.eps, np.float64)
    tid = np.ascontiguousarray(list(time_indices), np.uint64)
    nt = len(tid) if len(tid) else Nt
    L = np.zeros((nt, N ** 4), np.complex128)
    ctx = _ctx_for(exec)
    check(lib().taco_tcl2_generator(ctx.handle, dptr(cols.view(np.float64)), Nt, nf, no, dptr(eps),
                                    dptr(ops.view(np.float64)), iptr(p2f), N, sptr(tid) if len(tid) else None,
                                    len(tid), dptr(L.view(np.float64))))
    return L.reshape(nt, N * N, N * N).transpose(0, 2, 1)


def assemble_l4_from_slices(Fs, Cs, Rs, pair_to_freq, ops, exec: Optional[Exec] = None):
    """MIKX -> GW -> L4 at one time from nf^3 slices (idx3 = (i nf + j) nf + k).
    Returns (L4, GW) as N^2 x N^2 matrices."""
    ctx = _ctx_for(exec)
    Fs = np.ascontiguousarray(np.asarray(Fs, np.complex128).ravel())
    Cs = np.ascontiguousarray(np.asarray(Cs, np.complex128).ravel())
    Rs = np.ascontiguousarray(np.asarray(Rs, np.complex128).ravel())
    nf = int(round(Fs.size ** (1.0 / 3.0)))
    if nf ** 3 != Fs.size or Cs.size != Fs.size or Rs.size != Fs.size:
        raise TacoInvalidArgument(1, "slices must hold nf^3 values each")
    p2f = np.asarray(pair_to_freq, np.int32)
    N = p2f.shape[0]
    pf = np.ascontiguousarray(p2f.T.ravel())
    of = _ops_flat(ops)
    L = np.zeros(N ** 4, np.complex128)
    GW = np.zeros(N ** 4, np.complex128)
    check(lib().taco_l4_from_slices(ctx.handle, dptr(Fs), dptr(Cs), dptr(Rs), iptr(pf), N, nf,
                                    dptr(of) if of.size else None, len(ops), dptr(L), dptr(GW)))
    return L.reshape(N * N, N * N).T, GW.reshape(N * N, N * N).T


def propagate_rk4_dense(L_series, r, t0: float, dt: float,
                        on_sample: Optional[Callable[[float, np.ndarray], None]] = None, sample_every: int = 1,
                        L_half_series=None, exec: Optional[Exec] = None) -> np.ndarray:
    """rk4_dense.cpp:246-320 on the GPU.  Returns the final r; ``on_sample``
    is called with (t, r) at steps 0, every, ... and at the end."""
    ctx = _ctx_for(exec)
    Ls = np.asarray(L_series, np.complex128)
    if Ls.ndim != 3 or Ls.shape[0] < 2:
        raise TacoInvalidArgument(1, "propagate_rk4_dense: L_series must have at least 2 matrices (endpoints)")
    steps = Ls.shape[0] - 1
    D = Ls.shape[1]
    if Ls.shape[2] != D:
        raise TacoInvalidArgument(1, "propagate_rk4_dense: L_series must be square")
    rv = np.ascontiguousarray(np.asarray(r, np.complex128)).copy()
    if rv.size != D:
        raise TacoInvalidArgument(1, "propagate_rk4_dense: L_series and r dimension mismatch")
    hp = None
    if L_half_series is not None:
        Lh = np.asarray(L_half_series, np.complex128)
        if Lh.shape != (steps, D, D):
            raise TacoInvalidArgument(1, "propagate_rk4_dense: L_half_series must match L_series size - 1 (midpoints)")
        Lhf = np.ascontiguousarray(Lh.transpose(0, 2, 1)).ravel()
        hp = dptr(Lhf)
    flat = np.ascontiguousarray(Ls.transpose(0, 2, 1)).ravel()
    samples = None
    if on_sample is not None:
        samples = np.zeros((steps // sample_every + 2, D), np.complex128)
    ns = c_size_t(0)
    check(lib().taco_rk4_propagate(ctx.handle, dptr(flat), hp, steps, D, dptr(rv), dt, sample_every,
                                   dptr(samples) if samples is not None else None, ctypes.byref(ns)))
    if on_sample is not None:
        for q in range(ns.value):
            step = q * sample_every if q < ns.value - 1 else steps
            on_sample(t0 + step * dt, samples[q])
    return rv


propagate_rk4_dense_serial = propagate_rk4_dense
propagate_rk4_dense_omp = propagate_rk4_dense


def tcl4_propagate(system: System, gamma_series, dt: float, steps: int, r0, order: int = 4,
                   exec: Optional[Exec] = None) -> np.ndarray:
    """Driver composition (TCL4_spin_boson_example.cpp:284-384): states
    r(t_0..t_steps) in the eigenbasis, shape [steps+1, N^2]."""
    ctx = _ctx_for(exec)
    g = _gamma_checked(system, gamma_series)
    Nt, nf = g.shape
    N = system.N
    om, mi, pf, of = _system_arrays(system)
    eps = np.ascontiguousarray(system.eps, np.float64)
    r = np.ascontiguousarray(np.asarray(r0, np.complex128)).copy()
    if r.size != N * N:
        raise TacoInvalidArgument(1, "r0 must have N^2 entries")
    samples = np.zeros((steps + 1, N * N), np.complex128)
    gc = _colmajor(g).ravel()
    check(lib().taco_tcl4_propagate(ctx.handle, dptr(gc), Nt, nf, dptr(om), iptr(mi), iptr(pf), N, dptr(eps),
                                    dptr(of) if of.size else None, len(system.A_eig), dt, steps, order, dptr(r),
                                    dptr(samples)))
    return samples


# --- batched runs: many independent systems in one device pass ------------------

@dataclass
class BatchItem:
    """One system of a batch: its System, Gamma table, dt, L4 output times
    and RK4 run (``steps`` steps of the ``order`` generator from ``r0``)."""
    system: System
    gamma: np.ndarray
    dt: float
    tidx: Sequence[int] = ()
    steps: int = 0
    r0: Optional[np.ndarray] = None
    order: int = 4


class Batch:
    """A prepared batch (taco_batch_* of the C ABI).  ``run()`` computes every
    system; ``run(graph=True)`` replays a captured hipGraph after the first
    (eager) run.  Per system the values equal the plain builders'
    (build_TCL4_generator_hip_fused_batch, tcl4_propagate) bit for bit."""

    def __init__(self, items: Sequence[BatchItem], exec: Optional[Exec] = None):
        from ._lib import BatchSystem

        self.ctx = _ctx_for(exec)
        self.items = list(items)
        self._keep = []
        descs = (BatchSystem * max(len(self.items), 1))()
        for q, it in enumerate(self.items):
            S = it.system
            g = _gamma_checked(S, it.gamma)
            om, mi, pf, of = _system_arrays(S)
            gc = _colmajor(g).ravel()
            tid = np.ascontiguousarray(np.asarray(list(it.tidx), dtype=np.int64), np.uint64)
            eps = np.ascontiguousarray(S.eps, np.float64)
            r0 = (np.ascontiguousarray(np.asarray(it.r0, np.complex128).ravel()) if it.r0 is not None
                  else np.zeros(S.N * S.N, np.complex128))
            if it.steps and r0.size != S.N * S.N:
                raise TacoInvalidArgument(1, "r0 must have N^2 entries")
            self._keep += [gc, om, mi, pf, of, tid, eps, r0]
            d = descs[q]
            d.gamma, d.Nt, d.nf = dptr(gc), g.shape[0], g.shape[1]
            d.omegas, d.mirror, d.pair_to_freq, d.N = dptr(om), iptr(mi), iptr(pf), S.N
            d.ops, d.num_ops, d.dt = (dptr(of) if of.size else None), len(S.A_eig), float(it.dt)
            d.tidx, d.ntau = (sptr(tid) if tid.size else None), tid.size
            d.steps, d.order, d.eps, d.r0 = int(it.steps), int(it.order), dptr(eps), dptr(r0)
        h = c_void_p()
        check(lib().taco_batch_create(self.ctx.handle, descs, len(self.items), ctypes.byref(h)))
        self.handle = h
        self._keep = []  # the engine copied every input

    def run(self, graph: bool = False) -> "Batch":
        check(lib().taco_batch_run(self.handle, 1 if graph else 0))
        return self

    def l4(self, s: int) -> np.ndarray:
        """System s's L4 at its tidx: [ntau, N^2, N^2]."""
        it = self.items[s]
        n, N2 = len(it.tidx), it.system.N ** 2
        out = np.zeros(n * N2 * N2, np.complex128)
        check(lib().taco_batch_l4(self.handle, s, dptr(out)))
        return out.reshape(n, N2, N2).transpose(0, 2, 1)

    def rho(self, s: int) -> np.ndarray:
        """System s's states at steps 0..steps: [steps+1, N^2] (eigenbasis)."""
        it = self.items[s]
        out = np.zeros((it.steps + 1, it.system.N ** 2), np.complex128)
        check(lib().taco_batch_rho(self.handle, s, dptr(out), None))
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib().taco_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batched_run(items: Sequence[BatchItem], exec: Optional[Exec] = None):
    """One-shot batch: returns ([L4 per system], [rho per system or None])."""
    b = Batch(items, exec)
    try:
        b.run()
        return ([b.l4(s) for s in range(len(items))],
                [b.rho(s) if it.steps else None for s, it in enumerate(items)])
    finally:
        b.close()


# --- device-resident inputs (torch tensors as HBM buffers) -----------------------

class DeviceSystem:
    """System arrays and a Gamma table resident on one GPU.

    torch only provides the device buffers; every kernel is the engine's."""

    def __init__(self, system: System, gamma_series, device: int = 0):
        import torch

        g = _gamma_checked(system, gamma_series)
        self.device = device
        self.N = system.N
        self.nf = system.nf
        self.Nt = g.shape[0]
        self.num_ops = len(system.A_eig)
        dev = torch.device("cuda", device)
        om, mi, pf, of = _system_arrays(system)
        self.gamma = torch.from_numpy(_colmajor(g).ravel().copy()).to(dev)  # [nf][Nt]
        self.omegas = torch.from_numpy(om.copy()).to(dev)
        self.mirror = torch.from_numpy(mi.copy()).to(dev)
        self.pair_to_freq = torch.from_numpy(pf.copy()).to(dev)
        self.ops = torch.from_numpy(of.copy()).to(dev)
        torch.cuda.synchronize(dev)


def build_l4_dev(dsys: DeviceSystem, dt: float, time_indices: Sequence[int], out=None,
                 method: FCRMethod = FCRMethod.Convolution):
    """L4(t) for t in time_indices into a device tensor [ntau, N^4] (column-major blocks)."""
    import torch

    ctx = context(dsys.device)
    tid = np.ascontiguousarray(np.asarray(list(time_indices), dtype=np.int64), np.uint64)
    ntau = tid.size if tid.size else dsys.Nt
    N4 = dsys.N ** 4
    if out is None:
        out = torch.empty((ntau, N4), dtype=torch.complex128, device=torch.device("cuda", dsys.device))
    assert out.is_contiguous() and out.numel() == ntau * N4 and out.dtype == torch.complex128
    check(lib().taco_l4_batch_dev(ctx.handle, c_void_p(dsys.gamma.data_ptr()), dsys.Nt, dsys.nf,
                                  c_void_p(dsys.omegas.data_ptr()), c_void_p(dsys.mirror.data_ptr()),
                                  c_void_p(dsys.pair_to_freq.data_ptr()), dsys.N, c_void_p(dsys.ops.data_ptr()),
                                  dsys.num_ops, dt, int(method), sptr(tid) if tid.size else None, tid.size,
                                  c_void_p(out.data_ptr()), None))
    return out


KERNEL_NAMES = ["k_phase_a", "k_phase_b", "k_phase_c", "k_combine", "k_l4", "spectra", "k_tcl2", "k_rk4"]


class Profile:
    """Per-kernel HIP-event timing of one context (taco_profile_* of the C ABI)."""

    def __init__(self, device: int = 0):
        self.ctx = context(device)

    def __enter__(self):
        check(lib().taco_profile_reset(self.ctx.handle))
        check(lib().taco_profile_enable(self.ctx.handle, 1))
        return self

    def __exit__(self, *exc):
        check(lib().taco_profile_enable(self.ctx.handle, 0))

    def read(self) -> dict:
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            ms, n, b = c_double(0), ctypes.c_long(0), c_double(0)
            check(lib().taco_profile_get(self.ctx.handle, k, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
            if n.value:
                out[name] = {"ms": ms.value, "launches": n.value, "algo_bytes": b.value}
        return out
