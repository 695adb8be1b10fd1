"""Synthetic inputs of the TCL4 benchmark configurations.

Random numbers come from a repo-local splitmix64 stream with Box–Muller
normals, so every platform builds the same Hamiltonians for a given seed.

Configurations (system, bath, grid, time indices):

* ``spin_boson``   d=2, H = (eps/2) sz + (Delta/2) sx, A = sz, Ohmic
                   J = alpha w exp(-w/wc) (alpha 0.05, wc 5, beta 1),
                   Nt = 2^14, dt = 0.005, tidx 0:10:2000, 2000 RK4 steps.
* ``gue_pair``     d=4 GUE H (seed 2), A = sz(x)I + I(x)sz, Ohmic
                   (alpha 0.1, wc 10, beta 0.5), bath N = 200000
                   (Nt = 200001), dt = 0.002, tidx 0:1:10000.
* ``published``    the reference's published compare run: d=2, H = sx/2,
                   A = sz/2, J = w exp(-w/10), beta 0.5, dt = 0.000625,
                   N = 200000, tidx 0:1:10000 (tests/tcl4_e2e_cuda_compare.cpp:97-137).
* ``spin_pair``    d=4, H = H1(x)I + I(x)H1 with H1 of spin_boson, A as
                   gue_pair, spin_boson bath, Nt = 2^16, 4000 RK4 steps.
* ``heterogeneous_batch``  systems of d in {2,3,4,6,8}, GUE H, random
                   Hermitian A with ||A||_2 = 1, Ohmic or Drude–Lorentz.
* ``tfim``         open transverse-field Ising chain, d = 2^n.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .tcl4 import DrudeLorentz, Ohmic, System, bcf_fft_fun, compute_trapz_prefix_multi_matrix

SX = np.array([[0, 1], [1, 0]], np.complex128)
SZ = np.array([[1, 0], [0, -1]], np.complex128)
I2 = np.eye(2, dtype=np.complex128)


class SplitMix64:
    """splitmix64 stream with Box–Muller normals and uniform doubles."""
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK
        self._spare: Optional[float] = None

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def normal(self) -> float:
        if self._spare is not None:
            v, self._spare = self._spare, None
            return v
        u1 = 0.0
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        self._spare = r * math.sin(2.0 * math.pi * u2)
        return r * math.cos(2.0 * math.pi * u2)


def gue(d: int, rng: SplitMix64) -> np.ndarray:
    """H_ij = (a + ib)/sqrt(2d) (i < j), H_ii = c/sqrt(d), a, b, c ~ N(0, 1)."""
    H = np.zeros((d, d), np.complex128)
    for i in range(d):
        H[i, i] = rng.normal() / math.sqrt(d)
        for j in range(i + 1, d):
            v = complex(rng.normal(), rng.normal()) / math.sqrt(2 * d)
            H[i, j] = v
            H[j, i] = v.conjugate()
    return H


def random_hermitian_unit(d: int, rng: SplitMix64) -> np.ndarray:
    A = gue(d, rng)
    return A / np.max(np.abs(np.linalg.eigvalsh(A)))


def kron(*ms) -> np.ndarray:
    out = np.array([[1.0 + 0j]])
    for m in ms:
        out = np.kron(out, m)
    return out


def tfim(n: int, J: float, h: float) -> tuple[np.ndarray, np.ndarray]:
    """Open chain H = -J sum sz_i sz_{i+1} - h sum sx_i; A = sum sz_i."""
    d = 2 ** n
    H = np.zeros((d, d), np.complex128)
    A = np.zeros((d, d), np.complex128)
    for i in range(n):
        ops = [I2] * n
        ops[i] = SX
        H -= h * kron(*ops)
        ops = [I2] * n
        ops[i] = SZ
        A += kron(*ops)
        if i + 1 < n:
            ops = [I2] * n
            ops[i] = SZ
            ops[i + 1] = SZ
            H -= J * kron(*ops)
    return H, A


@dataclass
class Workload:
    name: str
    H: np.ndarray
    ops: list
    bath: object
    beta: float
    dt: float
    bcf_N: int                 # bath grid intervals; Nt = bcf_N + 1 samples
    tidx: list
    steps: int = 0
    meta: dict = field(default_factory=dict)
    _system: Optional[System] = None
    _gamma: Optional[np.ndarray] = None

    @property
    def system(self) -> System:
        if self._system is None:
            self._system = System.build(self.H, self.ops, 1e-9)
        return self._system

    @property
    def gamma(self) -> np.ndarray:
        """Gamma table [Nt, nf] (bcf_fft_fun + trapz, as the reference drivers)."""
        if self._gamma is None:
            _, C = bcf_fft_fun(self.bcf_N, self.dt, self.bath, self.beta)
            self._gamma = compute_trapz_prefix_multi_matrix(C, self.dt, self.system.omegas)
        return self._gamma


def spin_boson(eps: float = 1.0, delta: float = 1.0, alpha: float = 0.05, omega_c: float = 5.0, beta: float = 1.0,
               Nt: int = 2 ** 14, dt: float = 0.005, tmax: int = 2000, tstep: int = 10) -> Workload:
    H = 0.5 * eps * SZ + 0.5 * delta * SX
    return Workload("spin_boson", H, [SZ.copy()], Ohmic(alpha, omega_c), beta, dt, Nt - 1,
                    list(range(0, tmax + 1, tstep)), steps=tmax)


def gue_pair(seed: int = 2, N_bath: int = 200000, dt: float = 0.002, tmax: int = 10000, tstep: int = 1,
             alpha: float = 0.1, omega_c: float = 10.0, beta: float = 0.5) -> Workload:
    rng = SplitMix64(seed)
    H = gue(4, rng)
    A = kron(SZ, I2) + kron(I2, SZ)
    return Workload("gue_pair", H, [A], Ohmic(alpha, omega_c), beta, dt, N_bath, list(range(0, tmax + 1, tstep)),
                    meta={"seed": seed})


def published(N_bath: int = 200000, tmax: int = 10000) -> Workload:
    return Workload("published", 0.5 * SX, [0.5 * SZ], Ohmic(1.0, 10.0), 0.5, 0.000625, N_bath,
                    list(range(0, tmax + 1)))


def spin_pair(eps: float = 1.0, delta: float = 1.0, alpha: float = 0.05, omega_c: float = 5.0, beta: float = 1.0,
              Nt: int = 2 ** 16, dt: float = 0.005, steps: int = 4000) -> Workload:
    H1 = 0.5 * eps * SZ + 0.5 * delta * SX
    H = kron(H1, I2) + kron(I2, H1)
    A = kron(SZ, I2) + kron(I2, SZ)
    return Workload("spin_pair", H, [A], Ohmic(alpha, omega_c), beta, dt, Nt - 1, list(range(steps + 1)),
                    steps=steps, meta={"eps": eps, "delta": delta})


def heterogeneous_system(b: int, dims=(2, 3, 4, 6, 8), n_time: int = 1000, dt: float = 0.005,
                         log2_range=(14, 18)) -> Workload:
    """System b of the heterogeneous batch (seed 3 + b)."""
    rng = SplitMix64(3 + b)
    d = dims[int(rng.uniform() * len(dims)) % len(dims)]
    H = gue(d, rng)
    A = random_hermitian_unit(d, rng)
    if rng.uniform() < 0.5:
        bath = Ohmic(rng.uniform(0.01, 0.2), rng.uniform(1.0, 20.0))
    else:
        bath = DrudeLorentz(rng.uniform(0.01, 0.2), rng.uniform(1.0, 20.0))
    beta = rng.uniform(0.2, 5.0)
    lo, hi = log2_range
    lg = lo + int(rng.uniform() * (hi - lo + 1)) % (hi - lo + 1)
    stride = [1, 2, 4, 8, 16][int(rng.uniform() * 5) % 5]
    tidx = [q * stride for q in range(n_time)]
    Nt = 2 ** lg
    return Workload(f"hetero{b}", H, [A], bath, beta, dt, Nt - 1, tidx, steps=n_time,
                    meta={"d": d, "log2_Nt": lg, "stride": stride})


def tfim_workload(h: float, n: int = 4, J: float = 1.0, lam: float = 0.05, gamma: float = 5.0, beta: float = 1.0,
                  dt: float = 0.005, Nt: int = 2 ** 20, tmax: int = 20000) -> Workload:
    H, A = tfim(n, J, h)
    return Workload("tfim", H, [A], DrudeLorentz(lam, gamma), beta, dt, Nt - 1, list(range(tmax + 1)),
                    meta={"h": h, "J": J})
