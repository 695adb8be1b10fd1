"""Configuration-file driver: the reference's generic TCL driver
(examples/tcl_driver.cpp:149-484, schema of configs/tcl_driver.yaml) on the
GPU engine.

    python -m taco_amd.driver [--config=PATH]

Reads the YAML configuration (system.H / system.A as {re, im} matrices,
system.freq_tol; bath.beta, bath.J_expr, bath.params; numerics.dt, N, bcf_N,
gamma_rule rect|trapz, tcl4.method / nmax / fcr_fft_pad, threads;
output.tidx, gw_csv, l4_csv, profile; simulation.propagate, order, rho0,
print_series, sample_every, rho_out), builds the system and C(t) from the J
expression on the host (as the reference does), Gamma by the chosen rule,
L4 and GW at tidx on the GPU, and optionally propagates TCL2 (+ L4) with RK4
on the GPU up to tidx.  Output lines follow the reference driver's; errors
print "Error: <message>" and exit with status 1.
"""
from __future__ import annotations

import math
import sys
import time
from contextlib import contextmanager
from typing import Optional

import numpy as np

from . import io as tio
from . import tcl4

__all__ = ["load_config", "run", "main", "l4_to_gw", "parse_rho0"]


class ConfigError(RuntimeError):
    pass


def _complex_matrix(node, name: str) -> np.ndarray:
    if not isinstance(node, dict):
        raise ConfigError(f"{name} must be a map with keys {{re, im?}}")
    if "re" not in node:
        raise ConfigError(f"Missing key: {name}.re")
    re = _real_matrix(node["re"], f"{name}.re")
    if "im" in node and node["im"] is not None:
        im = _real_matrix(node["im"], f"{name}.im")
        if im.shape != re.shape:
            raise ConfigError(f"{name}: re/im dims mismatch")
    else:
        im = np.zeros_like(re)
    return re + 1j * im


def _real_matrix(node, name: str) -> np.ndarray:
    if not isinstance(node, list):
        raise ConfigError(f"{name} must be a sequence of rows")
    if not node:
        return np.zeros((0, 0))
    if not isinstance(node[0], list):
        raise ConfigError(f"{name} must be a 2D sequence")
    cols = len(node[0])
    for row in node:
        if not isinstance(row, list):
            raise ConfigError(f"{name} row is not a sequence")
        if len(row) != cols:
            raise ConfigError(f"{name} has inconsistent row sizes")
    return np.array(node, dtype=np.float64).reshape(len(node), cols)


def _index(node, Nt: int) -> int:
    """output.tidx: an integer or 'last' (default last)."""
    if Nt == 0:
        raise ConfigError("Nt is zero")
    if node is None:
        return Nt - 1
    if isinstance(node, (list, dict)):
        raise ConfigError("tidx must be a scalar (integer or 'last')")
    if isinstance(node, str) and node.strip().lower() == "last":
        return Nt - 1
    v = int(str(node))
    if v < 0:
        raise ConfigError("tidx must be >= 0")
    if v >= Nt:
        raise ConfigError("tidx out of range")
    return v


def parse_rho0(node, dim: int) -> np.ndarray:
    """simulation.rho0: a {re, im} matrix, a qubit label (0|g|ground, 1|e|excited,
    +x, -x, +y, -y) or a basis index; default |0><0|."""
    def proj(i):
        if i >= dim:
            raise ConfigError("simulation.rho0 index out of range")
        P = np.zeros((dim, dim), complex)
        P[i, i] = 1.0
        return P

    def pure(k):
        k = np.asarray(k, complex)
        return np.outer(k, k.conj())

    if node is None:
        return proj(0)
    if isinstance(node, dict) and "re" in node:
        rho = _complex_matrix(node, "simulation.rho0")
        if rho.shape != (dim, dim):
            raise ConfigError("simulation.rho0: dimension mismatch")
        return rho
    if isinstance(node, (list, dict)):
        raise ConfigError("simulation.rho0 must be a matrix map or scalar")
    s = str(node).strip()
    if dim == 2:
        v = s.lower()
        r2 = 1.0 / math.sqrt(2.0)
        if v in ("", "0", "g", "ground"):
            return pure([1, 0])
        if v in ("1", "e", "excited"):
            return pure([0, 1])
        if v == "+x":
            return pure([r2, r2])
        if v == "-x":
            return pure([r2, -r2])
        if v == "+y":
            return pure([r2, 1j * r2])
        if v == "-y":
            return pure([r2, -1j * r2])
    i = int(s)
    if i < 0:
        raise ConfigError("simulation.rho0 index must be >= 0")
    return proj(i)


def l4_to_gw(L4: np.ndarray, N: int) -> np.ndarray:
    """GW(n,i;m,j) from L4(n,m;i,j): the inverse of the exact reshuffle
    gw_to_liouvillian (tcl4_assemble.cpp:180-205); indices flat2(a, b) = a + N b."""
    Lr = np.asarray(L4).reshape(N, N, N, N)  # [m, n, j, i] = L4[n + N m, i + N j]
    return np.ascontiguousarray(Lr.transpose(3, 1, 2, 0)).reshape(N * N, N * N)


def _hermitize_and_normalize(rho: np.ndarray) -> np.ndarray:
    h = 0.5 * (rho + rho.conj().T)
    return h / np.trace(h).real


def load_config(path: str) -> dict:
    import yaml
    with open(path) as f:
        cfg = yaml.safe_load(f)
    if not isinstance(cfg, dict):
        raise ConfigError("configuration must be a map")
    return cfg


class _Profile:
    def __init__(self, on: bool, out):
        self.on, self.out = on, out

    @contextmanager
    def section(self, name: str):
        t0 = time.perf_counter()
        yield
        if self.on:
            self.out.write(f"[profile] {name}: {1e3 * (time.perf_counter() - t0):.3f} ms\n")


def run(cfg: dict, config_path: str = "<dict>", out=None, exec: Optional[tcl4.Exec] = None) -> dict:
    """Runs a parsed configuration; returns the computed arrays (C, Gamma, L4,
    GW, rho samples and the final rho when propagated)."""
    out = out or sys.stdout
    w = out.write
    sysn = cfg.get("system")
    if not sysn:
        raise ConfigError("Missing key: system")
    freq_tol = float(sysn.get("freq_tol", 1e-9))
    H = _complex_matrix(sysn.get("H"), "system.H")
    if H.shape[0] != H.shape[1]:
        raise ConfigError("system.H must be square")
    A_node = sysn.get("A")
    if not isinstance(A_node, list) or not A_node:
        raise ConfigError("system.A must be a non-empty sequence")
    A_lab = [_complex_matrix(a, f"system.A[{i}]") for i, a in enumerate(A_node)]
    for A in A_lab:
        if A.shape != H.shape:
            raise ConfigError("system.A: dimension mismatch vs H")
    bath = cfg.get("bath")
    if not bath:
        raise ConfigError("Missing key: bath")
    if "beta" not in bath or "J_expr" not in bath:
        raise ConfigError("bath needs beta and J_expr")
    beta = float(bath["beta"])
    params = bath.get("params") or {}
    if not isinstance(params, dict):
        raise ConfigError("bath.params must be a map")
    J = tcl4.SpectralExpr(str(bath["J_expr"]), **{str(k): float(v) for k, v in params.items()})
    num = cfg.get("numerics")
    if not num:
        raise ConfigError("Missing key: numerics")
    dt = float(num["dt"])
    N = int(num["N"])
    bcf_N = int(num.get("bcf_N", N))
    if bcf_N < N:
        raise ConfigError("numerics.bcf_N must be >= numerics.N")
    gamma_rule = str(num.get("gamma_rule", "rect")).lower()
    t4 = num.get("tcl4") or {}
    method = "direct" if str(t4.get("method", "convolution")).lower() == "direct" else "convolution"
    nmax = int(t4.get("nmax", 2))
    pad = int(t4.get("fcr_fft_pad", 0))
    outn = cfg.get("output") or {}
    sim = cfg.get("simulation") or {}
    prof = _Profile(bool(outn.get("profile", False)), out)

    w(f"tcl_driver config: {config_path}\n")
    w("engine: MI355X HIP (gfx950)\n")
    with prof.section("Build system"):
        S = tcl4.System.build(H, A_lab, freq_tol)
    dim, nf = S.N, S.nf
    w(f"dim={dim}, channels={len(A_lab)}, nf={nf}, dt={dt:.9f}, N={N}, bcf_N={bcf_N}, beta={beta:.9f}, "
      f"gamma_rule={gamma_rule}, tcl4_method={method}, nmax={nmax}, fcr_fft_pad={pad}\n")
    if method == "direct":
        fcr_method = tcl4.FCRMethod.Direct
    else:
        fcr_method = tcl4.FCRMethod.Convolution
    prev_pad = tcl4.get_fcr_fft_pad_factor()
    if pad > 0:
        tcl4.set_fcr_fft_pad_factor(pad)
    try:
        with prof.section("BCF FFT"):
            _, C = tcl4.bcf_fft_fun(bcf_N, dt, J, beta)
        Nt = N + 1
        if C.size < Nt:
            raise ConfigError("BCF output shorter than requested Nt=N+1")
        C = C[:Nt].copy()
        with prof.section("Gamma series"):
            if gamma_rule == "trapz":
                G = tcl4.compute_trapz_prefix_multi_matrix(C, dt, S.omegas)
            elif gamma_rule == "rect":
                G = tcl4.compute_rect_prefix_multi_matrix(C, dt, S.omegas)
            else:
                raise ConfigError("Unsupported numerics.gamma_rule (use rect or trapz)")
        tnode = outn.get("tidx", sim.get("tidx"))
        tidx = _index(tnode, G.shape[0])
        w(f"tidx={tidx} (t={tidx * dt:.9f})\n")
        with prof.section("Build GW/L4"):
            L4 = tcl4.build_TCL4_generator(S, G, dt, tidx, fcr_method, exec)
        GW = l4_to_gw(L4, dim)
        w(f"GW_raw: {GW.shape[0]}x{GW.shape[1]}, ||GW||_F={np.linalg.norm(GW):.9f}\n")
        w(f"L4: {L4.shape[0]}x{L4.shape[1]}, ||L4||_F={np.linalg.norm(L4):.9f}\n")
        if outn.get("gw_csv"):
            tio.write_csv_matrix(outn["gw_csv"], GW)
            w(f"Wrote GW_raw to {outn['gw_csv']}\n")
        if outn.get("l4_csv"):
            tio.write_csv_matrix(outn["l4_csv"], L4)
            w(f"Wrote L4 to {outn['l4_csv']}\n")
        res = {"system": S, "C": C, "gamma": G, "tidx": tidx, "L4": L4, "GW": GW}
        if bool(sim.get("propagate", False)):
            order = int(sim.get("order", 4))
            if order not in (0, 2, 4):
                raise ConfigError("simulation.order must be 0, 2, or 4")
            rho0 = parse_rho0(sim.get("rho0"), dim)
            r0 = rho0.T.reshape(-1)  # vec, column-major
            steps = tidx
            w(f"RK4: steps={steps}, t0={0.0:.9f}, tf={steps * dt:.9f}, dt={dt:.9f} (order={order})\n")
            every = int(sim.get("sample_every", 1))
            print_series = bool(sim.get("print_series", False))
            if print_series and every == 0:
                raise ConfigError("simulation.sample_every must be >= 1")
            with prof.section("RK4"):
                if steps == 0:
                    R = r0.reshape(1, -1).copy()
                elif fcr_method == tcl4.FCRMethod.Convolution:
                    R = tcl4.tcl4_propagate(S, G, dt, steps, r0, order=order, exec=exec)
                else:
                    R = _propagate_direct(S, G, dt, steps, r0, order, exec)
            if print_series:
                sz = np.diag([1.0, -1.0]).astype(complex)
                w("t,sz,tr,purity\n" if dim == 2 else "t,tr,purity\n")
                t = 0.0
                for step in range(steps + 1):
                    if step > 0:
                        t += dt
                    if step % every != 0 and step != steps:
                        continue
                    rho = _hermitize_and_normalize(R[step].reshape(dim, dim).T)
                    tr = np.trace(rho).real
                    pur = np.trace(rho @ rho).real
                    if dim == 2:
                        lab = S.U @ rho @ S.U.conj().T
                        w(f"{t:.9f},{np.trace(lab @ sz).real:.9f},{tr:.9f},{pur:.9f}\n")
                    else:
                        w(f"{t:.9f},{tr:.9f},{pur:.9f}\n")
            rho_f = _hermitize_and_normalize(R[-1].reshape(dim, dim).T)
            w("rho(tf) (eigen basis):\n" + _fmt_matrix(rho_f))
            if dim == 2:
                w("rho(tf) (lab basis):\n" + _fmt_matrix(S.U @ rho_f @ S.U.conj().T))
            if sim.get("rho_out"):
                tio.write_csv_matrix(sim["rho_out"], rho_f)
                w(f"Wrote rho(tf) to {sim['rho_out']}\n")
            res.update(samples=R, rho_final=rho_f)
        return res
    finally:
        tcl4.set_fcr_fft_pad_factor(prev_pad)


def _propagate_direct(S, G, dt: float, steps: int, r0, order: int, exec) -> np.ndarray:
    """The driver's RK4 loop (tcl_driver.cpp:369-455) with L4 from the Direct
    method: L(t) = TCL2(t) with Gamma_b(alpha, alpha) = Gamma_b(t) on every
    channel (+ L4(t)), midpoints averaged, all on the GPU."""
    no = len(S.A_eig)
    T = steps + 1
    gab = np.zeros((T, S.nf, no, no), np.complex128)
    if order != 0:
        for a in range(no):
            gab[:, :, a, a] = G[:T]
    L = tcl4.build_tcl2_generator(S, gab, exec=exec)
    if order == 4:
        L = L + tcl4.build_TCL4_generator(S, G, dt, list(range(T)), tcl4.FCRMethod.Direct, exec)
    out = []
    tcl4.propagate_rk4_dense(L, r0, 0.0, dt, on_sample=lambda t, r: out.append(r.copy()), sample_every=1,
                             exec=exec)
    return np.array(out[:T])


def _fmt_matrix(M: np.ndarray) -> str:
    return "".join(" ".join(f"({v.real:.9f},{v.imag:.9f})" for v in row) + "\n" for row in M)


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    path = "configs/tcl_driver.yaml"
    for a in argv:
        if a in ("--help", "-h"):
            print("Usage: python -m taco_amd.driver [--config=PATH]\nDefaults: config=configs/tcl_driver.yaml")
            return 0
        if a.startswith("--config="):
            path = a[len("--config="):]
    try:
        run(load_config(path), path)
        return 0
    except Exception as e:  # the reference driver reports every failure the same way
        sys.stderr.write(f"Error: {e}\n")
        return 1


if __name__ == "__main__":
    sys.exit(main())
