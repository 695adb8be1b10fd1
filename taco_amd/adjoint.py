"""Gradients of the TCL4 pipeline (no reference counterpart: jcarrots/taco has
no derivatives).

* ``Tape`` wraps the engine's forward tape and hand-written reverse pass
  (taco_tape_* of the C ABI): RK4 -> L(t) = L_TCL2(t) + L4(t) -> fused L4 ->
  F/C/R, with RK4 checkpoints every K steps.
* ``TCL4Propagate`` is the torch.autograd.Function over it: inputs (Gamma
  table, bucket frequencies, A_eig, eigenvalues, r0), output the states
  r_0..r_steps of ``tcl4_propagate`` (the drivers' composition,
  TCL4_spin_boson_example.cpp:284-384).
* ``BcfFun``, ``GammaFun``, ``EighFun`` and ``BucketOmegas`` carry the
  parameter chain in front of it: bath spectral density -> C(t)
  (correlation_fft.hpp:154-214), trapz Gamma (gamma.hpp:166-213), the
  eigensystem of H (system.hpp:37-157).  Their forwards are the engine's
  routines (bit-identical to the oracle's); their backwards are written out
  below.
* ``BathTape`` / ``BathTCL4Propagate`` fuse BcfFun -> GammaFun ->
  TCL4Propagate into one device tape (taco_tape_forward_bath): C(t) and
  Gamma are made, kept and reversed in device memory, and the reverse
  returns the bath-parameter cotangents directly.
* ``SpinPairProblem`` is the differentiation workload: H = H1 (x) I + I (x) H1,
  H1 = eps/2 sz + Delta/2 sx, A = sz (x) I + I (x) sz, Ohmic bath
  (alpha, omega_c, beta), loss sum_t ||rho_lab(t) - rho_target(t)||_F^2.

Complex cotangents follow torch's convention for real losses:
z.grad = dLoss/dRe z + i dLoss/dIm z.
"""
from __future__ import annotations

import ctypes
from ctypes import c_void_p
from dataclasses import dataclass
from typing import Optional, Sequence

import os

import numpy as np
import scipy.fft
import torch

from ._lib import TacoInvalidArgument, check, dptr, iptr, lib
from .tcl4 import (Exec, Ohmic, DrudeLorentz, System, _ctx_for, bath_gamma_hip, bcf_fft_fun,
                   compute_trapz_prefix_multi_matrix, compute_trapz_prefix_multi_matrix_hip)

__all__ = ["Tape", "BathTape", "MultiTape", "TCL4Propagate", "BathTCL4Propagate", "TCL4PropagateMany",
           "BathTCL4PropagateMany", "BcfFun",
           "GammaFun", "EighFun", "BucketOmegas", "SpinPairProblem", "TfimChainProblem"]


def _cm(a) -> np.ndarray:
    """Row-major numpy matrix -> column-major flat (Eigen's layout)."""
    return np.ascontiguousarray(np.asarray(a, np.complex128).T).ravel()


class Tape:
    """One forward run kept for the reverse pass (taco_tape_forward)."""

    def __init__(self, gamma, omegas, mirror, pair_to_freq, ops, eps, dt: float, steps: int, r0,
                 order: int = 4, checkpoint_every: int = 64, exec: Optional[Exec] = None):
        self.ctx = _ctx_for(exec)
        g = np.asarray(gamma, np.complex128)
        self.Nt, self.nf = g.shape
        p2f = np.asarray(pair_to_freq, np.int32)
        self.N = p2f.shape[0]
        self.num_ops = len(ops)
        self.steps = int(steps)
        gc = np.ascontiguousarray(g.T).ravel()
        om = np.ascontiguousarray(omegas, np.float64)
        mi = np.ascontiguousarray(mirror, np.int32)
        pf = np.ascontiguousarray(p2f.T).ravel()
        of = np.concatenate([_cm(o) for o in ops])
        ev = np.ascontiguousarray(eps, np.float64)
        r = np.ascontiguousarray(np.asarray(r0, np.complex128).ravel())
        if r.size != self.N ** 2:
            raise TacoInvalidArgument(1, "r0 must have N^2 entries")
        self.samples = np.zeros((self.steps + 1, self.N ** 2), np.complex128)
        h = c_void_p()
        check(lib().taco_tape_forward(self.ctx.handle, dptr(gc), self.Nt, self.nf, dptr(om), iptr(mi), iptr(pf),
                                      self.N, dptr(ev), dptr(of), self.num_ops, float(dt), self.steps, int(order),
                                      dptr(r), int(checkpoint_every), dptr(self.samples), ctypes.byref(h)))
        self.handle = h

    def adjoint(self, samples_bar) -> dict:
        sb = np.ascontiguousarray(np.asarray(samples_bar, np.complex128).reshape(self.steps + 1, self.N ** 2))
        gb = np.zeros(self.Nt * self.nf, np.complex128)
        ob = np.zeros(self.nf)
        ab = np.zeros(self.num_ops * self.N ** 2, np.complex128)
        eb = np.zeros(self.N)
        rb = np.zeros(self.N ** 2, np.complex128)
        check(lib().taco_tape_adjoint(self.handle, dptr(sb), dptr(gb), dptr(ob), dptr(ab), dptr(eb), dptr(rb)))
        N = self.N
        return {
            "gamma": gb.reshape(self.nf, self.Nt).T.copy(),
            "omegas": ob,
            "ops": [ab[k * N * N:(k + 1) * N * N].reshape(N, N).T.copy() for k in range(self.num_ops)],
            "eps": eb,
            "r0": rb,
        }

    def close(self):
        if getattr(self, "handle", None):
            check(lib().taco_tape_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BathTape(Tape):
    """A Tape whose Gamma table the device makes from a bath
    (taco_tape_forward_bath): C(t_n), n < Nt, of spectral density `kind`
    (0 Ohmic, 1 Drude-Lorentz) with parameters (p0, p1) at inverse temperature
    beta, and its trapz Gamma, both kept in device memory.  `adjoint` returns
    the cotangents of (p0, p1, beta) under "bath" instead of a Gamma
    cotangent, and omega_bar includes the trapz Gamma's part."""

    def __init__(self, kind: int, p0: float, p1: float, beta: float, Nt: int, omegas, mirror, pair_to_freq, ops,
                 eps, dt: float, steps: int, r0, order: int = 4, checkpoint_every: int = 64,
                 exec: Optional[Exec] = None):
        self.ctx = _ctx_for(exec)
        om = np.ascontiguousarray(omegas, np.float64)
        self.Nt, self.nf = int(Nt), om.size
        p2f = np.asarray(pair_to_freq, np.int32)
        self.N = p2f.shape[0]
        self.num_ops = len(ops)
        self.steps = int(steps)
        mi = np.ascontiguousarray(mirror, np.int32)
        pf = np.ascontiguousarray(p2f.T).ravel()
        of = np.concatenate([_cm(o) for o in ops])
        ev = np.ascontiguousarray(eps, np.float64)
        r = np.ascontiguousarray(np.asarray(r0, np.complex128).ravel())
        if r.size != self.N ** 2:
            raise TacoInvalidArgument(1, "r0 must have N^2 entries")
        self.samples = np.zeros((self.steps + 1, self.N ** 2), np.complex128)
        h = c_void_p()
        check(lib().taco_tape_forward_bath(self.ctx.handle, int(kind), float(p0), float(p1), float(beta), self.Nt,
                                           self.nf, dptr(om), iptr(mi), iptr(pf), self.N, dptr(ev), dptr(of),
                                           self.num_ops, float(dt), self.steps, int(order), dptr(r),
                                           int(checkpoint_every), dptr(self.samples), ctypes.byref(h)))
        self.handle = h

    def adjoint(self, samples_bar) -> dict:
        sb = np.ascontiguousarray(np.asarray(samples_bar, np.complex128).reshape(self.steps + 1, self.N ** 2))
        bb = np.zeros(3)
        ob = np.zeros(self.nf)
        ab = np.zeros(self.num_ops * self.N ** 2, np.complex128)
        eb = np.zeros(self.N)
        rb = np.zeros(self.N ** 2, np.complex128)
        check(lib().taco_tape_adjoint_bath(self.handle, dptr(sb), dptr(bb), dptr(ob), dptr(ab), dptr(eb), dptr(rb)))
        N = self.N
        return {
            "bath": bb,
            "omegas": ob,
            "ops": [ab[k * N * N:(k + 1) * N * N].reshape(N, N).T.copy() for k in range(self.num_ops)],
            "eps": eb,
            "r0": rb,
        }


class MultiTape:
    """Several parameter sets of one model size (the same N and steps) run
    forward and reverse together (taco_mtape_*): each set's samples and
    cotangents equal a Tape of that set alone, bit for bit, while the RK4
    recursions of all sets share one launch each way.  `sets` holds dicts
    with the Tape arguments (gamma, omegas, mirror, pair_to_freq, ops, eps,
    dt, steps, r0, optional order).  With `baths` (one (kind, p0, p1, beta)
    per set) the sets' Gamma tables come from their baths on the device, as
    in BathTape: each dict gives the grid length "Nt" instead of "gamma", and
    `adjoint` returns each set's "bath" cotangents instead of "gamma"."""

    last_slice_modes: list = []  # slice_modes of the most recently made multi-set tape

    def __init__(self, sets, checkpoint_every: int = 64, exec: Optional[Exec] = None, baths=None):
        from ._lib import BathParams, TapeSet
        self.ctx = _ctx_for(exec)
        self.meta, self._keep = [], []
        self.bath = baths is not None
        arr = (TapeSet * len(sets))()
        for k, st in enumerate(sets):
            if self.bath:
                g = None
                Nt, nf = int(st["Nt"]), len(st["omegas"])
            else:
                g = np.asarray(st["gamma"], np.complex128)
                Nt, nf = g.shape
            p2f = np.asarray(st["pair_to_freq"], np.int32)
            N = p2f.shape[0]
            ops = st["ops"]
            steps = int(st["steps"])
            buf = dict(
                gc=None if g is None else np.ascontiguousarray(g.T).ravel(),
                om=np.ascontiguousarray(st["omegas"], np.float64),
                mi=np.ascontiguousarray(st["mirror"], np.int32), pf=np.ascontiguousarray(p2f.T).ravel(),
                of=np.concatenate([_cm(o) for o in ops]), ev=np.ascontiguousarray(st["eps"], np.float64),
                r=np.ascontiguousarray(np.asarray(st["r0"], np.complex128).ravel()))
            if buf["r"].size != N * N:
                raise TacoInvalidArgument(1, "r0 must have N^2 entries")
            self._keep.append(buf)
            self.meta.append(dict(Nt=Nt, nf=nf, N=N, num_ops=len(ops), steps=steps))
            arr[k] = TapeSet(None if g is None else dptr(buf["gc"]), Nt, nf, dptr(buf["om"]), iptr(buf["mi"]),
                             iptr(buf["pf"]), N, dptr(buf["ev"]), dptr(buf["of"]), len(ops), float(st["dt"]), steps,
                             int(st.get("order", 4)), dptr(buf["r"]))
        self.samples = [np.zeros((m["steps"] + 1, m["N"] ** 2), np.complex128) for m in self.meta]
        h = c_void_p()
        if self.bath:
            bp = (BathParams * len(sets))()
            for k, (kind, p0, p1, beta) in enumerate(baths):
                bp[k] = BathParams(int(kind), float(p0), float(p1), float(beta))
            check(lib().taco_mtape_forward_bath(self.ctx.handle, arr, bp, len(sets), int(checkpoint_every),
                                                _ptrs(self.samples), ctypes.byref(h)))
        else:
            check(lib().taco_mtape_forward(self.ctx.handle, arr, len(sets), int(checkpoint_every),
                                           _ptrs(self.samples), ctypes.byref(h)))
        self.handle = h
        self._keep = None  # inputs are only read during the forward call
        MultiTape.last_slice_modes = self.slice_modes

    @property
    def keeps_slices(self) -> bool:
        return bool(lib().taco_mtape_keeps_slices(self.handle))

    @property
    def slice_modes(self) -> list:
        """Per set: ("recompute" | "whole" | "packed", device bytes of its slices)."""
        names = {0: "recompute", 1: "whole", 2: "packed"}
        out = []
        for s in range(len(self.meta)):
            b = ctypes.c_size_t(0)
            m = lib().taco_mtape_set_slices(self.handle, s, ctypes.byref(b))
            if m < 0:
                raise ValueError("multi-set tape is closed")
            out.append((names[m], int(b.value)))
        return out

    def adjoint(self, samples_bar) -> list:
        sb = [np.ascontiguousarray(np.asarray(b, np.complex128).reshape(m["steps"] + 1, m["N"] ** 2))
              for b, m in zip(samples_bar, self.meta)]
        ob = [np.zeros(m["nf"]) for m in self.meta]
        ab = [np.zeros(m["num_ops"] * m["N"] ** 2, np.complex128) for m in self.meta]
        eb = [np.zeros(m["N"]) for m in self.meta]
        rb = [np.zeros(m["N"] ** 2, np.complex128) for m in self.meta]
        if self.bath:
            bb = [np.zeros(3) for _ in self.meta]
            check(lib().taco_mtape_adjoint_bath(self.handle, _ptrs(sb), _ptrs(bb), _ptrs(ob), _ptrs(ab), _ptrs(eb),
                                                _ptrs(rb)))
        else:
            gb = [np.zeros(m["Nt"] * m["nf"], np.complex128) for m in self.meta]
            check(lib().taco_mtape_adjoint(self.handle, _ptrs(sb), _ptrs(gb), _ptrs(ob), _ptrs(ab), _ptrs(eb),
                                           _ptrs(rb)))
        out = []
        for k, m in enumerate(self.meta):
            N = m["N"]
            first = {"bath": bb[k]} if self.bath else {"gamma": gb[k].reshape(m["nf"], m["Nt"]).T.copy()}
            out.append({
                **first,
                "omegas": ob[k],
                "ops": [ab[k][q * N * N:(q + 1) * N * N].reshape(N, N).T.copy() for q in range(m["num_ops"])],
                "eps": eb[k],
                "r0": rb[k],
            })
        return out

    def close(self):
        if getattr(self, "handle", None):
            check(lib().taco_mtape_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptrs(arrays):
    """double** over numpy arrays (complex arrays as interleaved doubles)."""
    from ._lib import _dp
    return (_dp * len(arrays))(*[dptr(a) for a in arrays])


class TCL4Propagate(torch.autograd.Function):
    """r_0..r_steps of the TCL4 run as a differentiable function of
    (gamma [Nt, nf], omegas [nf], ops [num_ops, N, N], eps [N], r0 [N^2]).
    ``meta`` = (mirror, pair_to_freq, dt, steps, order, checkpoint_every,
    exec); the discrete bucket structure carries no gradient."""

    @staticmethod
    def forward(ctx, gamma, omegas, ops, eps, r0, meta):
        mirror, p2f, dt, steps, order, K, ex = meta
        tape = Tape(gamma.detach().numpy(), omegas.detach().numpy(), mirror, p2f,
                    list(ops.detach().numpy()), eps.detach().numpy(), dt, steps, r0.detach().numpy(), order, K, ex)
        ctx.tape = tape
        return torch.from_numpy(tape.samples.copy())

    @staticmethod
    def backward(ctx, gs):
        res = ctx.tape.adjoint(gs.detach().numpy())
        return (torch.from_numpy(res["gamma"]), torch.from_numpy(res["omegas"]),
                torch.from_numpy(np.stack(res["ops"])), torch.from_numpy(res["eps"]), torch.from_numpy(res["r0"]),
                None)


class BathTCL4Propagate(torch.autograd.Function):
    """BcfFun -> GammaFun -> TCL4Propagate in one step on the device
    (BathTape): r_0..r_steps as a differentiable function of the bath
    parameters (p0, p1, beta) and (omegas [nf], ops [num_ops, N, N], eps [N],
    r0 [N^2]), with C(t) and Gamma made and reversed in device memory.
    ``meta`` = (kind, Nt, mirror, pair_to_freq, dt, steps, order,
    checkpoint_every, exec)."""

    @staticmethod
    def forward(ctx, p0, p1, beta, omegas, ops, eps, r0, meta):
        kind, Nt, mirror, p2f, dt, steps, order, K, ex = meta
        tape = BathTape(kind, float(p0), float(p1), float(beta), Nt, omegas.detach().numpy(), mirror, p2f,
                        list(ops.detach().numpy()), eps.detach().numpy(), dt, steps, r0.detach().numpy(), order, K,
                        ex)
        ctx.tape = tape
        return torch.from_numpy(tape.samples.copy())

    @staticmethod
    def backward(ctx, gs):
        res = ctx.tape.adjoint(gs.detach().numpy())
        t = lambda v: torch.tensor(float(v), dtype=torch.float64)
        pb = res["bath"]
        return (t(pb[0]), t(pb[1]), t(pb[2]), torch.from_numpy(res["omegas"]),
                torch.from_numpy(np.stack(res["ops"])), torch.from_numpy(res["eps"]), torch.from_numpy(res["r0"]),
                None)


class TCL4PropagateMany(torch.autograd.Function):
    """TCL4Propagate of several parameter sets of one model size at once
    (MultiTape: one RK4 launch for all sets each way).  Inputs after ``meta``
    are the sets' (gamma, omegas, ops, eps, r0) in turn; ``meta`` = (per-set
    (mirror, pair_to_freq, dt, order) tuples, steps, checkpoint_every, exec).
    Returns each set's r_0..r_steps; every set's values and cotangents equal
    TCL4Propagate on that set alone."""

    @staticmethod
    def forward(ctx, meta, *tensors):
        per, steps, K, ex = meta
        sets = []
        for q, (mirror, p2f, dt, order) in enumerate(per):
            g, om, ops, e, r0 = tensors[5 * q:5 * q + 5]
            sets.append(dict(gamma=g.detach().numpy(), omegas=om.detach().numpy(), mirror=mirror,
                             pair_to_freq=p2f, ops=list(ops.detach().numpy()), eps=e.detach().numpy(), dt=dt,
                             steps=steps, r0=r0.detach().numpy(), order=order))
        mt = MultiTape(sets, K, ex)
        ctx.mt = mt
        return tuple(torch.from_numpy(x.copy()) for x in mt.samples)

    @staticmethod
    def backward(ctx, *gs):
        bars = [np.zeros(m.shape, np.complex128) if g is None else g.detach().numpy()
                for g, m in zip(gs, ctx.mt.samples)]
        out = [None]
        for res in ctx.mt.adjoint(bars):
            out += [torch.from_numpy(res["gamma"]), torch.from_numpy(res["omegas"]),
                    torch.from_numpy(np.stack(res["ops"])), torch.from_numpy(res["eps"]),
                    torch.from_numpy(res["r0"])]
        return tuple(out)


class BathTCL4PropagateMany(torch.autograd.Function):
    """BathTCL4Propagate of several parameter sets of one model size at once
    (MultiTape with baths).  Inputs after ``meta`` are the sets' (p0, p1,
    beta, omegas, ops, eps, r0) in turn; ``meta`` = (per-set (kind, Nt,
    mirror, pair_to_freq, dt, order) tuples, steps, checkpoint_every, exec).
    Every set's values and cotangents equal BathTCL4Propagate on that set
    alone."""

    @staticmethod
    def forward(ctx, meta, *tensors):
        per, steps, K, ex = meta
        sets, baths = [], []
        for q, (kind, Nt, mirror, p2f, dt, order) in enumerate(per):
            p0, p1, beta, om, ops, e, r0 = tensors[7 * q:7 * q + 7]
            sets.append(dict(Nt=Nt, omegas=om.detach().numpy(), mirror=mirror, pair_to_freq=p2f,
                             ops=list(ops.detach().numpy()), eps=e.detach().numpy(), dt=dt, steps=steps,
                             r0=r0.detach().numpy(), order=order))
            baths.append((kind, float(p0), float(p1), float(beta)))
        mt = MultiTape(sets, K, ex, baths=baths)
        ctx.mt = mt
        return tuple(torch.from_numpy(x.copy()) for x in mt.samples)

    @staticmethod
    def backward(ctx, *gs):
        bars = [np.zeros(m.shape, np.complex128) if g is None else g.detach().numpy()
                for g, m in zip(gs, ctx.mt.samples)]
        t = lambda v: torch.tensor(float(v), dtype=torch.float64)
        out = [None]
        for res in ctx.mt.adjoint(bars):
            pb = res["bath"]
            out += [t(pb[0]), t(pb[1]), t(pb[2]), torch.from_numpy(res["omegas"]),
                    torch.from_numpy(np.stack(res["ops"])), torch.from_numpy(res["eps"]),
                    torch.from_numpy(res["r0"])]
        return tuple(out)


def _spectral(kind: int, p0: float, p1: float, w: np.ndarray):
    """J(w) and dJ/dp0, dJ/dp1 for w > 0 (spectral_J of the bath models)."""
    w = np.asarray(w, float)
    pos = w > 0
    J = np.zeros_like(w)
    d0 = np.zeros_like(w)
    d1 = np.zeros_like(w)
    if kind == 0:  # Ohmic p0 w exp(-w/p1)
        e = np.exp(-w[pos] / p1)
        J[pos] = p0 * w[pos] * e
        d0[pos] = w[pos] * e
        d1[pos] = J[pos] * w[pos] / (p1 * p1)
    else:  # Drude-Lorentz 2 p0 p1 w / (w^2 + p1^2)
        den = w[pos] ** 2 + p1 * p1
        J[pos] = 2.0 * p0 * p1 * w[pos] / den
        d0[pos] = 2.0 * p1 * w[pos] / den
        d1[pos] = 2.0 * p0 * w[pos] * (w[pos] ** 2 - p1 * p1) / (den * den)
    return J, d0, d1


class BcfFun(torch.autograd.Function):
    """C(t_n), n < Nt = N + 1, of bcf_fft_fun (correlation_fft.hpp:154-214) as
    a function of (p0, p1, beta) for a finite beta.  C = (dw/pi) FFT(S) with
    S the sampled KMS spectrum, linear in S: the backward is
    Sbar_k = (dw/pi) Re sum_n Cbar_n e^{+2 pi i k n / Nfft} and
    p_bar = sum_k Sbar_k dS_k/dp."""

    @staticmethod
    def forward(ctx, p0, p1, beta, meta):
        # meta = (kind, N, dt[, producer]): producer an Exec runs the device
        # producer on its GPU, "host" (the default) the host bcf_fft_fun; the
        # two are bit-identical
        kind, N, dt = meta[:3]
        producer = meta[3] if len(meta) > 3 else "host"
        bath = Ohmic(float(p0), float(p1)) if kind == 0 else DrudeLorentz(float(p0), float(p1))
        if isinstance(producer, Exec):
            C, _ = bath_gamma_hip(N, dt, bath, float(beta), exec=producer)
        else:
            _, C = bcf_fft_fun(N, dt, bath, float(beta))
        ctx.save = (kind, N, dt, float(p0), float(p1), float(beta))
        return torch.from_numpy(C)

    @staticmethod
    def backward(ctx, Cbar):
        kind, N, dt, p0, p1, beta = ctx.save
        if not np.isfinite(beta):
            raise TacoInvalidArgument(1, "BcfFun backward needs a finite beta")
        Nt = N + 1
        Nfft = 2 * Nt
        if Nfft & (Nfft - 1):
            Nfft = 1 << int(Nfft - 1).bit_length()
        dw = 2.0 * np.pi / (dt * Nfft)
        wny = np.pi / dt
        cb = np.zeros(Nfft, np.complex128)
        cb[:Nt] = Cbar.detach().numpy()
        # unnormalised inverse FFT (multithreaded on the host cores)
        Sbar = (dw / np.pi) * scipy.fft.ifft(cb, norm="forward", workers=len(os.sched_getaffinity(0))).real
        k = np.arange(1, Nfft // 2)
        w = dw * k
        J, dJ0, dJ1 = _spectral(kind, p0, p1, w)
        kms = np.exp(-beta * w)
        f = (np.pi / 2.0) / (1.0 - kms)
        Sp = f * J
        # S_k = Sp, S_{Nfft-k} = Sp * kms
        dSp_dJ = f
        dSp_db = (np.pi / 2.0) * J * (-w * kms) / (1.0 - kms) ** 2
        g0 = Sbar[k] * dSp_dJ * dJ0 + Sbar[Nfft - k] * dSp_dJ * kms * dJ0
        g1 = Sbar[k] * dSp_dJ * dJ1 + Sbar[Nfft - k] * dSp_dJ * kms * dJ1
        gb = Sbar[k] * dSp_db + Sbar[Nfft - k] * (dSp_db * kms + Sp * (-w * kms))
        # S_0 = 0.5 (pi/2) J(dw) (1 + r) / (1 - r), r = exp(-beta dw)
        r = np.exp(-beta * dw)
        J0, a0, a1 = (x[0] for x in _spectral(kind, p0, p1, np.array([dw])))
        q0 = 0.5 * (np.pi / 2.0) * (1.0 + r) / (1.0 - r)
        s0b = Sbar[0]
        g0 = g0.sum() + s0b * q0 * a0
        g1 = g1.sum() + s0b * q0 * a1
        gb = gb.sum() + s0b * 0.5 * (np.pi / 2.0) * J0 * (-2.0 * dw * r / (1.0 - r) ** 2)
        # S_Ny = 0.25 pi J(wny) (1 + 2 / (e^{beta wny} - 1))
        Jn, n0, n1 = (x[0] for x in _spectral(kind, p0, p1, np.array([wny])))
        xN = np.exp(-beta * wny)  # 2 / (e^{beta wny} - 1) = 2 xN / (1 - xN)
        qn = 0.25 * np.pi * (1.0 + 2.0 * xN / (1.0 - xN))
        sNb = Sbar[Nfft // 2]
        g0 += sNb * qn * n0
        g1 += sNb * qn * n1
        gb += sNb * 0.25 * np.pi * Jn * (-2.0 * wny * xN / (1.0 - xN) ** 2)
        t = lambda v: torch.tensor(float(v), dtype=torch.float64)
        return t(g0), t(g1), t(gb), None


class GammaFun(torch.autograd.Function):
    """Gamma_b(t_k) = (dt/2) sum_{m=1..k} (phi^{m-1} C_{m-1} + phi^m C_m),
    phi = e^{i w_b dt} (compute_trapz_prefix_multi_matrix, gamma.hpp:166-213),
    as a function of C [Nt] and w [nf]."""

    @staticmethod
    def forward(ctx, C, omegas, dt, producer="host"):
        # producer: an Exec (device trapz on its GPU) or "host"; bit-identical
        Cn = C.detach().numpy()
        om = omegas.detach().numpy()
        if isinstance(producer, Exec):
            G = compute_trapz_prefix_multi_matrix_hip(Cn, dt, om, exec=producer)
        else:
            G = compute_trapz_prefix_multi_matrix(Cn, dt, om)
        ctx.save = (Cn, om, dt)
        return torch.from_numpy(G)

    @staticmethod
    def backward(ctx, Gbar):
        Cn, om, dt = ctx.save
        Gb_full = Gbar.detach().numpy()  # [Nt, nf]
        Nt = Cn.size
        # Gamma(k) depends on C_m for m <= k only: rows past the last nonzero
        # cotangent contribute nothing
        nz = np.flatnonzero(np.any(Gb_full != 0, axis=1))
        T = int(nz[-1]) + 1 if nz.size else 0
        Cbar = np.zeros(Nt, np.complex128)
        wbar = np.zeros(om.size)
        if T > 0:
            Gb = Gb_full[:T]
            m = np.arange(T)
            # Q_b(m): weight-summed cotangents; Q(0) = sum_{k>=1} Gb, Q(m) = Gb(m) + 2 sum_{k>m} Gb
            suf = np.vstack([np.cumsum(Gb[::-1], axis=0)[::-1], np.zeros((1, Gb.shape[1]))])  # sum_{k>=m}
            Q = np.empty_like(Gb)
            Q[0] = suf[1]
            Q[1:] = Gb[1:] + 2.0 * suf[2:]
            phi = np.exp(1j * np.outer(m * dt, om))  # [T, nf]
            Cbar[:T] = (dt / 2.0) * np.sum(np.conj(phi) * Q, axis=1)
            dphi = (1j * (m * dt))[:, None] * phi * Cn[:T, None]
            wbar = (dt / 2.0) * np.real(np.sum(np.conj(Q) * dphi, axis=0))
        return torch.from_numpy(Cbar), torch.from_numpy(wbar), None, None


class EighFun(torch.autograd.Function):
    """(eps, U) of a Hermitian H with the engine's eigensolver (the System
    build, system.hpp:37-51), ascending eigenvalues.  Backward: with
    G = U^H U_bar, K_ij = G_ij / (e_j - e_i) for e_i != e_j, K_ii = eps_bar_i,
    0 inside degenerate blocks (the loss must not depend on the basis chosen
    inside a degenerate eigenspace, which holds for lab-frame losses), and
    H_bar = herm(U K U^H)."""

    @staticmethod
    def forward(ctx, H, A_lab, tol):
        S = System.build(H.detach().numpy(), [np.asarray(a) for a in A_lab], tol)
        ctx.save = (S.eps.copy(), S.U.copy(), tol)
        ctx.system = S
        return torch.from_numpy(S.eps.copy()), torch.from_numpy(S.U.copy())

    @staticmethod
    def backward(ctx, eb, Ub):
        e, U, tol = ctx.save
        N = e.size
        G = U.conj().T @ (Ub.detach().numpy() if Ub is not None else np.zeros((N, N), complex))
        E = e[None, :] - e[:, None]
        K = np.zeros((N, N), complex)
        off = np.abs(E) > max(tol, 1e-12) * max(1.0, np.abs(e).max())
        K[off] = G[off] / E[off]
        K[np.diag_indices(N)] = eb.detach().numpy() if eb is not None else 0.0
        Hb = U @ K @ U.conj().T
        Hb = 0.5 * (Hb + Hb.conj().T)
        return torch.from_numpy(Hb), None, None


class BucketOmegas(torch.autograd.Function):
    """Bucket frequencies w_b (the System's values) as a function of the
    eigenvalues: w_b = e_m - e_n for a representative pair (m, n) with
    pair_to_freq[m, n] = b (every pair of a bucket moves with it here)."""

    @staticmethod
    def forward(ctx, e, omegas, p2f):
        N = p2f.shape[0]
        rep = {}
        for n in range(N):
            for m in range(N):
                rep.setdefault(int(p2f[m, n]), (m, n))
        ctx.rep = [rep[b] for b in range(len(omegas))]
        ctx.N = N
        return torch.from_numpy(np.asarray(omegas, np.float64).copy())

    @staticmethod
    def backward(ctx, wb):
        eb = np.zeros(ctx.N)
        for b, (m, n) in enumerate(ctx.rep):
            eb[m] += float(wb[b])
            eb[n] -= float(wb[b])
        return torch.from_numpy(eb), None, None


SZ = np.array([[1, 0], [0, -1]], complex)
SX = np.array([[0, 1], [1, 0]], complex)
I2 = np.eye(2, dtype=complex)


@dataclass
class SpinPairProblem:
    """Gradient workload: d = 4 spin pair, Ohmic bath, RK4 of TCL2 + L4 from
    rho0 = |0><0| in the eigenbasis, loss sum_t ||rho_lab(t) - target(t)||_F^2
    over t = 0..steps (lab basis, which does not depend on the eigenbasis
    chosen)."""
    Nt: int = 2 ** 16
    dt: float = 0.005
    steps: int = 4000
    checkpoint_every: int = 64
    order: int = 4
    freq_tol: float = 1e-9
    # "device": bath chain and TCL4 run in one device tape (BathTCL4Propagate);
    # "split": BcfFun -> GammaFun -> TCL4Propagate (host reverses of the chain)
    bath_chain: str = "device"

    names = ("alpha", "omega_c", "beta", "eps", "delta")

    def hamiltonian(self, eps, delta):
        sz = torch.from_numpy(SZ)
        sx = torch.from_numpy(SX)
        i2 = torch.from_numpy(I2)
        H1 = 0.5 * eps * sz + 0.5 * delta * sx
        return torch.kron(H1, i2) + torch.kron(i2, H1)

    @staticmethod
    def coupling():
        return np.kron(SZ, I2) + np.kron(I2, SZ)

    def rho_lab(self, params: torch.Tensor, exec: Optional[Exec] = None) -> torch.Tensor:
        """rho_lab(t) [steps+1, N, N] as a differentiable function of
        params = (alpha, omega_c, beta, eps, delta)."""
        alpha, wc, beta, eps, delta = params[0], params[1], params[2], params[3], params[4]
        H = self.hamiltonian(eps.to(torch.complex128), delta.to(torch.complex128))
        return _rho_lab(self, H, self.coupling(), 0, alpha, wc, beta, exec)

    def loss(self, params: torch.Tensor, target: torch.Tensor, exec: Optional[Exec] = None) -> torch.Tensor:
        d = self.rho_lab(params, exec) - target
        return (d.real ** 2 + d.imag ** 2).sum()


def _rho_lab(prob, H: torch.Tensor, A: np.ndarray, kind: int, p0, p1, beta, exec: Optional[Exec]) -> torch.Tensor:
    """Lab-basis rho(t), t = 0..steps, of the TCL4 run from rho0 = |0><0| (in
    the eigenbasis) for a differentiable Hamiltonian H and bath parameters
    (p0, p1, beta) of spectral density `kind`; buckets and mirror are those of
    H's current spectrum (held fixed under differentiation)."""
    e, U = EighFun.apply(H, [A], prob.freq_tol)
    S = System.build(H.detach().numpy(), [A], prob.freq_tol)
    om = BucketOmegas.apply(e, S.omegas, S.pair_to_freq)
    At = torch.from_numpy(A)
    Aeig = (U.conj().T @ At @ U).unsqueeze(0)
    N = S.N
    r0 = torch.zeros(N * N, dtype=torch.complex128)
    r0[0] = 1.0
    if prob.bath_chain == "device":
        # C(t) and Gamma made, kept and reversed in device memory (BathTape)
        meta = (kind, prob.Nt, S.mirror, S.pair_to_freq, prob.dt, prob.steps, prob.order, prob.checkpoint_every,
                exec)
        r = BathTCL4Propagate.apply(p0, p1, beta, om, Aeig, e, r0, meta)
    else:
        # the same values through the separate steps: C and Gamma from the
        # device producers, their reverses on the host
        producer = exec or Exec()
        C = BcfFun.apply(p0, p1, beta, (kind, prob.Nt - 1, prob.dt, producer))
        G = GammaFun.apply(C, om, prob.dt, producer)
        meta = (S.mirror, S.pair_to_freq, prob.dt, prob.steps, prob.order, prob.checkpoint_every, exec)
        r = TCL4Propagate.apply(G, om, Aeig, e, r0, meta)
    rho = r.reshape(-1, N, N).transpose(1, 2)  # vec column-major: rho[n, m] = r[n + m N]
    return U @ rho @ U.conj().T


def _rho_lab_many(prob, Hs, A: np.ndarray, kind: int, baths, exec: Optional[Exec]) -> list:
    """_rho_lab for several parameter sets (Hamiltonians Hs, bath parameters
    baths[q] = (p0, p1, beta)) of one model size, with the TCL4 runs of all
    sets in one multi-set tape."""
    producer = exec or Exec()
    device = prob.bath_chain == "device"
    per, flat, Us, Ns = [], [], [], []
    for H, (p0, p1, beta) in zip(Hs, baths):
        e, U = EighFun.apply(H, [A], prob.freq_tol)
        S = System.build(H.detach().numpy(), [A], prob.freq_tol)
        om = BucketOmegas.apply(e, S.omegas, S.pair_to_freq)
        Aeig = (U.conj().T @ torch.from_numpy(A) @ U).unsqueeze(0)
        r0 = torch.zeros(S.N * S.N, dtype=torch.complex128)
        r0[0] = 1.0
        if device:
            per.append((kind, prob.Nt, S.mirror, S.pair_to_freq, prob.dt, prob.order))
            flat += [p0, p1, beta, om, Aeig, e, r0]
        else:
            C = BcfFun.apply(p0, p1, beta, (kind, prob.Nt - 1, prob.dt, producer))
            G = GammaFun.apply(C, om, prob.dt, producer)
            per.append((S.mirror, S.pair_to_freq, prob.dt, prob.order))
            flat += [G, om, Aeig, e, r0]
        Us.append(U)
        Ns.append(S.N)
    meta = (per, prob.steps, prob.checkpoint_every, exec)
    rs = BathTCL4PropagateMany.apply(meta, *flat) if device else TCL4PropagateMany.apply(meta, *flat)
    out = []
    for r, U, N in zip(rs, Us, Ns):
        rho = r.reshape(-1, N, N).transpose(1, 2)
        out.append(U @ rho @ U.conj().T)
    return out


def _kron_chain(ops) -> torch.Tensor:
    out = ops[0]
    for o in ops[1:]:
        out = torch.kron(out, o)
    return out


@dataclass
class TfimChainProblem:
    """Gradient workload: open n-spin transverse-field Ising chain
    H = -J sum sz_i sz_{i+1} - h sum sx_i (d = 2^n; 16 for n = 4), coupling
    A = sum sz_i, Drude-Lorentz bath J(w) = 2 lam gam w / (w^2 + gam^2); RK4
    of TCL2 + L4 from rho0 = |0><0| in the eigenbasis, loss
    sum_t ||rho_lab(t) - target(t)||_F^2 over t = 0..steps."""
    n_spins: int = 4
    Nt: int = 2 ** 12
    dt: float = 0.005
    steps: int = 200
    checkpoint_every: int = 64
    order: int = 4
    freq_tol: float = 1e-9
    bath_chain: str = "device"  # as SpinPairProblem.bath_chain

    names = ("lambda", "gamma", "beta", "J", "h")

    def hamiltonian(self, J, h):
        n = self.n_spins
        sz, sx, i2 = (torch.from_numpy(m) for m in (SZ, SX, I2))
        d = 2 ** n
        H = torch.zeros((d, d), dtype=torch.complex128)
        for i in range(n):
            ops = [i2] * n
            ops[i] = sx
            H = H - h * _kron_chain(ops)
            if i + 1 < n:
                ops = [i2] * n
                ops[i] = sz
                ops[i + 1] = sz
                H = H - J * _kron_chain(ops)
        return H

    def coupling(self):
        n = self.n_spins
        A = np.zeros((2 ** n, 2 ** n), complex)
        for i in range(n):
            ops = [I2] * n
            ops[i] = SZ
            m = ops[0]
            for o in ops[1:]:
                m = np.kron(m, o)
            A += m
        return A

    def rho_lab(self, params: torch.Tensor, exec: Optional[Exec] = None) -> torch.Tensor:
        """rho_lab(t) [steps+1, d, d] as a differentiable function of
        params = (lambda, gamma, beta, J, h)."""
        lam, gam, beta, J, h = params[0], params[1], params[2], params[3], params[4]
        H = self.hamiltonian(J.to(torch.complex128), h.to(torch.complex128))
        return _rho_lab(self, H, self.coupling(), 1, lam, gam, beta, exec)

    def loss(self, params: torch.Tensor, target: torch.Tensor, exec: Optional[Exec] = None) -> torch.Tensor:
        d = self.rho_lab(params, exec) - target
        return (d.real ** 2 + d.imag ** 2).sum()

    def losses_many(self, params: Sequence[torch.Tensor], targets: Sequence[torch.Tensor],
                    exec: Optional[Exec] = None) -> list:
        """loss() of several parameter sets, their TCL4 runs (forward and
        reverse) in one multi-set tape; equal to loss() set by set."""
        Hs, baths = [], []
        for p in params:
            Hs.append(self.hamiltonian(p[3].to(torch.complex128), p[4].to(torch.complex128)))
            baths.append((p[0], p[1], p[2]))
        rhos = _rho_lab_many(self, Hs, self.coupling(), 1, baths, exec)
        out = []
        for rho, tgt in zip(rhos, targets):
            d = rho - tgt
            out.append((d.real ** 2 + d.imag ** 2).sum())
        return out
