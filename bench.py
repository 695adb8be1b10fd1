#!/usr/bin/env python3
"""Benchmark of the TCL4 hot path on one MI355X.

Workload (default): the plain GPU drop-in at scale — d = 4 GUE Hamiltonian
(seed 2), A = sz(x)I + I(x)sz, Ohmic bath (alpha 0.1, omega_c 10, beta 0.5),
bath grid N = 200000 (Nt = 200001 Gamma rows, Nfft = 2^19), dt = 0.002,
L4(t) at t = 0..10000.  One step is one fused batch build
(F/C/R -> MIKX -> GW -> L4 for all 10001 times) from inputs already resident
in HBM.  The value is L4 time-points per second.

The CPU baseline leg times the oracle (the C++ restatement of the reference
serial/OpenMP path, with the platform libm phases the reference calls) on the
host cores, and uses the oracle as the checker of the GPU result (per-t
relative Frobenius error of L4 against the oracle in the same libm phase
convention; the error against the correctly rounded-phase oracle is reported
beside it).

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "L4 time-points/s batched vs plain at ≤1e-10 rel err; adjoint grads/s"
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="gue_pair",
                   choices=["gue_pair", "published", "spin_boson", "batched", "adjoint", "tfim_adjoint"])
    p.add_argument("--systems", type=int, default=None,
                   help="batched workload: number of systems (256); tfim_adjoint: parameter sets (64)")
    p.add_argument("--no-plain", action="store_true", help="batched workload: skip the plain per-system leg")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / checker leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--reduced", action="store_true", help="small grid (for a quick functional run)")
    p.add_argument("--plain-only", action="store_true",
                   help="default workload: skip the batched, adjoint, host-pointer and serial-CPU sub-records")
    p.add_argument("--copy-bytes", type=int, default=0,
                   help="after the timed steps, run one plain device copy of this many bytes (the reference "
                        "kernel for scaling hardware traffic counters)")
    return p.parse_args()


def make_workload(name: str, reduced: bool):
    from taco_amd import workloads as W

    if name == "gue_pair":
        return W.gue_pair(N_bath=20000, tmax=1000) if reduced else W.gue_pair()
    if name == "published":
        return W.published(N_bath=20000, tmax=2000) if reduced else W.published()
    return W.spin_boson()


def cpu_baseline_leg(w, L4_gpu: np.ndarray, threads: int) -> tuple[dict, dict]:
    """Time the oracle on the host cores and check the GPU result with it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from metrics import per_t_rel_err, ref_style_err
    from oracle import oracle

    S, G = w.system, w.gamma
    args = (G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, w.tidx)
    with oracle.phase_mode(oracle.PHASE_LIBM):
        t0 = time.perf_counter()
        ref_libm = oracle.l4_series(*args, nthreads=threads)
        cpu_s = time.perf_counter() - t0
    with oracle.phase_mode(oracle.PHASE_CR):
        ref_cr = oracle.l4_series(*args, nthreads=threads)
    e_cr, _ = per_t_rel_err(L4_gpu, ref_cr)
    e_libm, nfl = per_t_rel_err(L4_gpu, ref_libm)
    e_ref_libm_vs_cr, _ = per_t_rel_err(ref_libm, ref_cr)
    ntau = len(w.tidx)
    baseline = {
        "value": ntau / cpu_s,
        "unit": "L4 time-points/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full workload: all {S.nf ** 3} F/C/R triples (Nfft {oracle.fcr_nfft(G.shape[0])}) and "
                  f"{ntau} L4 times, oracle with libm phases, OpenMP {threads} threads, {cpu_s:.1f} s",
        "seconds": cpu_s,
    }
    parity = {
        "max_per_t_rel_err_vs_oracle_libm_phases": float(e_libm.max()),
        "max_per_t_rel_err_vs_oracle_libm_phases_t_lt_20": float(e_libm[:20].max()),
        "floored_t": nfl,
        "ref_style_max_abs_over_max1": ref_style_err(L4_gpu, ref_libm),
        "max_per_t_rel_err_vs_oracle_cr_phases": float(e_cr.max()),
        "oracle_libm_vs_cr_phases_max_per_t": float(e_ref_libm_vs_cr.max()),
        "pass_1e-10": bool(e_libm.max() <= 1e-10),
    }
    return baseline, parity, ref_libm


def kernel_table(kern: dict, steps: int, dom: str | None = None) -> tuple[dict, dict]:
    """Per-step kernel figures and the roofline of `dom` (default: the kernel
    with the most time)."""
    kernels = {}
    for name, k in kern.items():
        kernels[name] = {
            "ms_per_step": k["ms"] / steps,
            "launches_per_step": k["launches"] / steps,
            "avg_launch_us": 1000.0 * k["ms"] / k["launches"],
            "achieved_GBps": k["algo_bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else None,
        }
    if dom is None or dom not in kern:
        dom = max(kern, key=lambda n: kern[n]["ms"])
    kd = kern[dom]
    achieved = kd["algo_bytes"] / (kd["ms"] * 1e-3) / 1e9
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": achieved,
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": achieved / PEAK_HBM_GBS,
        "traffic": None,
        "algo_bytes_per_launch": kd["algo_bytes"] / kd["launches"],
        "avg_launch_us": 1000.0 * kd["ms"] / kd["launches"],
    }
    return kernels, roofline


def batched_main(a):
    print(json.dumps(batched_record(a)), flush=True)


def batched_record(a, steps: int | None = None, warmup: int | None = None):
    """Many independent systems (d in {2,3,4,6,8}, Ohmic or Drude-Lorentz,
    Nt = 2^14..2^18, 1000 L4 times at stride 1..16, 1000 RK4 steps of
    TCL2 + L4 each): one batch (inputs resident in HBM, replayed hipGraph)
    against the plain per-system path (the host-API builders, one system at a
    time: L4 at tidx, then the RK4 run)."""
    import torch

    import taco_amd
    from taco_amd import tcl4, workloads as W

    assert torch.cuda.is_available(), "bench needs an MI355X"
    # correctly rounded phases: the libm tables of the d = 6 and 8 systems
    # (tens of thousands of Omega classes x up to 2^18 times each) exceed any
    # sensible host/device table budget
    taco_amd.set_phase_mode(taco_amd.PHASE_CR)
    nsys = a.systems or 256
    t0 = time.perf_counter()
    ws = [W.heterogeneous_system(b) for b in range(nsys)]
    items = []
    for q, w in enumerate(ws):
        if q % 32 == 0:
            print(f"[bench] host prep (Gamma tables) system {q}", file=sys.stderr, flush=True)
        r0 = np.zeros(w.system.N ** 2, complex)
        r0[0] = 1.0
        items.append(taco_amd.BatchItem(w.system, w.gamma, w.dt, w.tidx, steps=w.steps, r0=r0, order=4))
    dims = {}
    for w in ws:
        dims[w.meta["d"]] = dims.get(w.meta["d"], 0) + 1
    print(f"[bench] batched: {nsys} systems, d counts {dims}, host prep {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = taco_amd.Batch(items)
    print(f"[bench] batch created in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    with tcl4.Profile(0) as prof:
        t0 = time.perf_counter()
        b.run()  # eager: tunes the FFT tiles; profiled per kernel
        eager_s = time.perf_counter() - t0
        kern = prof.read()
    print(f"[bench] eager batch run {eager_s:.2f} s", file=sys.stderr, flush=True)
    steps = a.steps if steps is None else steps
    warmup = a.warmup if warmup is None else warmup
    for _ in range(warmup):
        b.run(graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(graph=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_l4 = sum(len(it.tidx) for it in items)
    value = n_l4 * steps / elapsed
    print(f"[bench] graph replays: {steps} in {elapsed:.2f} s -> {value:.0f} L4/s", file=sys.stderr, flush=True)
    kernels, roofline = kernel_table(kern, 1)
    plain, check = None, None
    if not a.no_plain:
        t0 = time.perf_counter()
        mism = 0
        for s, it in enumerate(items):
            if s % 32 == 0:
                print(f"[bench] plain leg system {s}", file=sys.stderr, flush=True)
            L4 = taco_amd.build_TCL4_generator_hip_fused_batch(it.system, it.gamma, it.dt, list(it.tidx))
            rho = taco_amd.tcl4_propagate(it.system, it.gamma, it.dt, it.steps, it.r0, order=4)
            mism += int(not (np.array_equal(L4, b.l4(s)) and np.array_equal(rho, b.rho(s))))
        plain_s = time.perf_counter() - t0
        plain = {"value": n_l4 / plain_s, "seconds": plain_s,
                 "batched_over_plain": (n_l4 / plain_s) and value / (n_l4 / plain_s)}
        check = {"systems_bitwise_equal_to_plain": nsys - mism, "systems": nsys}
        print(f"[bench] plain per-system path {plain_s:.2f} s; {nsys - mism}/{nsys} systems bitwise equal",
              file=sys.stderr, flush=True)
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "L4 time-points/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1000.0 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {nsys} systems (seed 3+b), d counts {dims}, Nt 2^14..2^18, dt 0.005, "
                        f"1000 L4 times each (stride 1..16) + 1000 RK4 steps of TCL2+L4, inputs resident in HBM",
            "path": "batched (hipGraph replay)",
            "phases": "correctly rounded (TACO_PHASE_CR)",
        },
        "plain": plain,
        "check": check,
        "eager_run_s": eager_s,
        "roofline": roofline,
        "cpu_baseline": None,
        "kernels": kernels,
    }
    b.close()
    taco_amd.set_phase_mode(taco_amd.PHASE_LIBM)
    return rec


def adjoint_main(a):
    print(json.dumps(adjoint_record(a)), flush=True)


def adjoint_record(a, steps: int | None = None, warmup: int | None = None):
    """Gradient of the spin-pair loss (d = 4, H1 = eps/2 sz + Delta/2 sx,
    A = sz(x)I + I(x)sz, Ohmic bath alpha 0.05, omega_c 5, beta 1, bath grid
    2^16, dt 0.005, 4000 RK4 steps of TCL2 + L4, checkpoints every 64 steps,
    loss sum_t ||rho_lab(t) - rho_target(t)||_F^2 with the target from
    parameters moved by 5% (seed 4)) with respect to (alpha, omega_c, beta,
    eps, Delta).  One step = one forward + reverse pass; the value is
    gradients per second.  The CPU leg times central finite differences of
    the oracle pipeline (10 forwards) and checks the gradient against them."""
    import torch

    from taco_amd import adjoint as AD

    assert torch.cuda.is_available(), "bench needs an MI355X"
    prob = AD.SpinPairProblem()
    p0 = np.array([0.05, 5.0, 1.0, 1.0, 1.0])
    u = np.where(np.random.default_rng(4).uniform(size=5) < 0.5, -1.0, 1.0)
    with torch.no_grad():
        target = prob.rho_lab(torch.tensor(p0 * (1 + 0.05 * u), dtype=torch.float64)).detach()
    print("[bench] adjoint: target built", file=sys.stderr, flush=True)

    def fwd():
        with torch.no_grad():
            return prob.loss(torch.tensor(p0, dtype=torch.float64), target)

    def grad():
        params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
        loss = prob.loss(params, target)
        (g,) = torch.autograd.grad(loss, params)
        return float(loss), g.numpy()

    steps = a.steps if steps is None else steps
    warmup = a.warmup if warmup is None else warmup
    for _ in range(max(warmup, 1)):
        fwd()
        grad()
    t0 = time.perf_counter()
    for _ in range(steps):
        fwd()
    t_fwd = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, g = grad()
    t_grad = (time.perf_counter() - t0) / steps
    print(f"[bench] forward {t_fwd * 1e3:.1f} ms, forward+reverse {t_grad * 1e3:.1f} ms", file=sys.stderr, flush=True)
    # reverse-pass kernels of one gradient
    from taco_amd import tcl4
    with tcl4.Profile(0) as prof:
        grad()
        kern = prof.read()
    kernels, roofline = kernel_table(kern, 1)
    cpu, check = None, None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle import oracle
        from test_gpu_adjoint import oracle_spin_pair_rho_lab

        threads = a.cpu_threads or len(os.sched_getaffinity(0))
        oracle.set_phase_mode(oracle.PHASE_LIBM)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        tn = target.numpy()

        def oloss(p):
            d = oracle_spin_pair_rho_lab(oracle, prob, p) - tn
            return float(np.sum(d.real ** 2 + d.imag ** 2))

        t0 = time.perf_counter()
        fdg = np.zeros(5)
        for q in range(5):
            h = 1e-6 * abs(p0[q])
            pp, pm = p0.copy(), p0.copy()
            pp[q] += h
            pm[q] -= h
            fdg[q] = (oloss(pp) - oloss(pm)) / (2 * h)
            print(f"[bench] oracle FD {prob.names[q]} done", file=sys.stderr, flush=True)
        cpu_s = time.perf_counter() - t0
        rel = np.abs(g - fdg) / np.abs(fdg)
        cpu = {"value": 1.0 / cpu_s, "unit": "gradients/s", "cores": threads, "kind": "port",
               "sample": f"central finite differences of the oracle pipeline (10 forwards, h = 1e-6 relative), "
                         f"{cpu_s:.1f} s"}
        check = {"params": list(prob.names), "adjoint": g.tolist(), "oracle_fd": fdg.tolist(),
                 "max_rel_diff": float(rel.max()), "loss": loss}
    rec = {
        "metric": METRIC,
        "value": 1.0 / t_grad,
        "unit": "gradients/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(warmup, 1),
        "ms_per_step": 1e3 * t_grad,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": "adjoint: d=4 spin pair, Ohmic (0.05, 5, 1), Nt=2^16, dt=0.005, 4000 RK4 steps of TCL2+L4, "
                        "checkpoints every 64, 5 parameters (alpha, omega_c, beta, eps, Delta)",
            "path": "tape forward + hand-written reverse",
        },
        "forward_ms": 1e3 * t_fwd,
        "gradient_over_forward": t_grad / t_fwd,
        "check": check,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    return rec


def window_record(w, ref_libm, threads: int, steps: int = 5) -> dict:
    """The same workload in the window FFT mode (rows t < T only, Nfft =
    next_pow2(2T - 1), exact twiddles): L4 time-points/s from inputs resident
    in HBM, its per-t error against the parity-mode oracle (the reference's
    recurrence-twiddle FFT) and against the reference's Direct O(T^2) method
    on t <= 500.  A separate labelled line, never the parity headline."""
    import torch

    import taco_amd
    from taco_amd import tcl4

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from metrics import per_t_rel_err
    from oracle import oracle

    S, G = w.system, w.gamma
    ntau = len(w.tidx)
    with taco_amd.fft_mode(taco_amd.FFT_WINDOW):
        dsys = tcl4.DeviceSystem(S, G, 0)
        out = torch.empty((ntau, S.N ** 4), dtype=torch.complex128, device="cuda:0")
        tcl4.build_l4_dev(dsys, w.dt, w.tidx, out)
        tcl4.build_l4_dev(dsys, w.dt, w.tidx, out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            tcl4.build_l4_dev(dsys, w.dt, w.tidx, out)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        L4 = out.cpu().numpy().reshape(ntau, S.N ** 2, S.N ** 2).transpose(0, 2, 1)
        short = [t for t in w.tidx if t <= 500]
        L4s = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, short)
    rec = {"workload": f"{w.name} in the window FFT mode, Nfft {1 << int(np.ceil(np.log2(2 * (max(w.tidx) + 1) - 1)))}",
           "value": ntau / el, "unit": "L4 time-points/s", "ms_per_step": 1e3 * el, "steps": steps}
    if ref_libm is not None:
        e, _ = per_t_rel_err(L4, ref_libm)
        rec["max_per_t_rel_err_vs_parity_oracle"] = float(e.max())
        rec["max_per_t_rel_err_vs_parity_oracle_t_ge_20"] = float(e[20:].max()) if ntau > 20 else None
        t0 = time.perf_counter()
        with oracle.phase_mode(oracle.PHASE_LIBM):
            direct = oracle.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, short, method="direct",
                                      nthreads=threads)
        ed, _ = per_t_rel_err(L4s, direct)
        ep, _ = per_t_rel_err(ref_libm[: len(short)], direct)
        rec["max_per_t_rel_err_vs_direct_t_le_500"] = float(ed.max())
        rec["parity_oracle_vs_direct_t_le_500"] = float(ep.max())
        rec["direct_seconds"] = time.perf_counter() - t0
    return rec


def host_pointer_record(w, steps: int = 3) -> dict:
    """taco_l4_batch from host buffers: the Gamma table copied in, the L4
    series copied out and repacked into per-time matrices, as the reference's
    published timing includes (tests/tcl4_e2e_cuda_compare.cpp:208-214)."""
    import taco_amd

    S, G = w.system, w.gamma
    taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)  # tables, tiles, workspace
    t0 = time.perf_counter()
    for _ in range(steps):
        L = taco_amd.build_TCL4_generator_hip_fused_batch(S, G, w.dt, w.tidx)
        mats = [np.ascontiguousarray(L[q]) for q in range(L.shape[0])]  # per-time matrices
    el = (time.perf_counter() - t0) / steps
    del mats
    return {"workload": w.name, "d": S.N, "nf": S.nf, "ntau": len(w.tidx), "Nt": int(G.shape[0]),
            "seconds_per_call": el, "L4_per_s_with_host_copies": len(w.tidx) / el, "calls": steps}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def serial_cpu_record(w) -> dict:
    """The oracle on one thread (the reference's serial path) over the whole
    workload w."""
    from oracle import oracle

    S, G = w.system, w.gamma
    with oracle.phase_mode(oracle.PHASE_LIBM):
        t0 = time.perf_counter()
        oracle.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, w.dt, w.tidx, nthreads=1)
        el = time.perf_counter() - t0
    return {"value": len(w.tidx) / el, "unit": "L4 time-points/s", "cores": 1, "kind": "port",
            "sample": f"{w.name}: whole workload (d={S.N}, nf={S.nf}, Nt={G.shape[0]}, {len(w.tidx)} L4 times), "
                      f"oracle serial, {el:.1f} s", "seconds": el}


def tfim_adjoint_main(a):
    """Gradients of the d = 16 loss (open 4-spin transverse-field Ising chain,
    J = 1, h drawn from U[0.5, 1.5] (seed 5), A = sum sz_i, Drude-Lorentz
    bath lambda 0.05, gamma 5, beta 1) with respect to (lambda, gamma, beta,
    J, h), for --systems parameter sets (default 64) in one multi-set tape
    (their RK4 recursions one launch each way), and set by set for the first
    four (the single-tape path, timed for comparison).  The window is reduced from the 2^20-point grid and 20001 times of the full
    configuration (its forward F/C/R pass alone moves ~10 TB per parameter
    set) to a 2^12-point grid and 200 RK4 steps of TCL2 + L4, checkpoints
    every 64.  One step = one gradient per parameter set (forward + reverse);
    the value is gradients per second.  The gradient of the first set is
    checked against central differences (h = 1e-6 relative) of the same GPU
    pipeline, whose d = 16 forward the tests pin to the oracle; the oracle's
    own d = 16 forward at this size (all 531441 F/C/R triples at Nfft 2^13
    plus 201 O(N^6) L4 evaluations) is too slow to time as a baseline."""
    import torch

    from taco_amd import adjoint as AD

    assert torch.cuda.is_available(), "bench needs an MI355X"
    import taco_amd
    # correctly rounded phases: each parameter set has its own ~10^5 Omega
    # classes, whose libm table the host would rebuild for every gradient
    taco_amd.set_phase_mode(taco_amd.PHASE_CR)
    # F/C/R slices (5.1 GB per set at nf = 81 whole, 1.5 GB packed to the
    # entries L4 reads) kept within 160 GiB of the 288 GB HBM; the multi-set
    # reverse recomputes those of any set past it
    slice_budget = 160 << 30
    taco_amd.set_slice_budget(slice_budget)
    nsets = a.systems or 64
    rng = np.random.default_rng(5)
    hs = rng.uniform(0.5, 1.5, size=nsets)
    prob = AD.TfimChainProblem(Nt=2 ** 12, dt=0.005, steps=200, checkpoint_every=64)
    P0 = [np.array([0.05, 5.0, 1.0, 1.0, h]) for h in hs]
    u = np.where(np.random.default_rng(4).uniform(size=5) < 0.5, -1.0, 1.0)
    with torch.no_grad():
        targets = [prob.rho_lab(torch.tensor(p * (1 + 0.05 * u), dtype=torch.float64)).detach() for p in P0]
    print(f"[bench] tfim_adjoint: {nsets} targets built", file=sys.stderr, flush=True)

    def fwd(q):
        with torch.no_grad():
            return float(prob.loss(torch.tensor(P0[q], dtype=torch.float64), targets[q]))

    def grad(q):
        params = torch.tensor(P0[q], dtype=torch.float64, requires_grad=True)
        loss = prob.loss(params, targets[q])
        (g,) = torch.autograd.grad(loss, params)
        return float(loss), g.numpy()

    def grad_many():
        params = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in P0]
        losses = prob.losses_many(params, targets)
        gs = torch.autograd.grad(sum(losses), params)
        return [(float(l), g.numpy()) for l, g in zip(losses, gs)]

    for _ in range(a.warmup):
        fwd(0)
        grad(0)
    # set by set (a few sets: each is a whole forward + reverse)
    nseq = min(nsets, 4)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        for q in range(nseq):
            fwd(q)
    t_fwd = (time.perf_counter() - t0) / (a.steps * nseq)
    t0 = time.perf_counter()
    seq = []
    for _ in range(a.steps):
        seq = [grad(q) for q in range(nseq)]
    t_seq = (time.perf_counter() - t0) / (a.steps * nseq)
    print(f"[bench] set by set: forward {t_fwd * 1e3:.1f} ms, forward+reverse {t_seq * 1e3:.1f} ms per set",
          file=sys.stderr, flush=True)
    # every set in one multi-set tape
    grad_many()
    t0 = time.perf_counter()
    grads = []
    for _ in range(a.steps):
        grads = grad_many()
    t_grad = (time.perf_counter() - t0) / (a.steps * nsets)
    same = sum(1 for q in range(nseq) if grads[q][0] == seq[q][0] and np.array_equal(grads[q][1], seq[q][1]))
    modes = AD.MultiTape.last_slice_modes
    slice_modes = {m: sum(1 for x in modes if x[0] == m) for m in ("whole", "packed", "recompute")}
    slice_gib = sum(x[1] for x in modes) / 2 ** 30
    print(f"[bench] multi-set tape: {t_grad * 1e3:.1f} ms per set over {nsets} sets; {same}/{nseq} sets bit-equal "
          f"to the set-by-set gradients; slices {slice_modes}, {slice_gib:.1f} GiB", file=sys.stderr, flush=True)
    from taco_amd import tcl4
    # kernel figures of one multi-set gradient, per set; the roofline is phase
    # B's (the HBM-bound kernel with the most time: the L4 and F/C/R reverses
    # are FP64-latency bound and are timed under the forward kernels' names
    # with no bytes of their own)
    with tcl4.Profile(0) as prof:
        grad_many()
        kern = prof.read()
    kernels, roofline = kernel_table(kern, nsets, dom="k_phase_b")
    loss0, g0 = grads[0]
    fdg = np.zeros(5)
    for q in range(5):
        h = 1e-6 * abs(P0[0][q])
        pp, pm = P0[0].copy(), P0[0].copy()
        pp[q] += h
        pm[q] -= h
        with torch.no_grad():
            lp = float(prob.loss(torch.tensor(pp, dtype=torch.float64), targets[0]))
            lm = float(prob.loss(torch.tensor(pm, dtype=torch.float64), targets[0]))
        fdg[q] = (lp - lm) / (2 * h)
    rel = np.abs(g0 - fdg) / np.abs(fdg)
    rec = {
        "metric": METRIC,
        "value": 1.0 / t_grad,
        "unit": "gradients/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_grad * nsets,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"tfim_adjoint (reduced window): d=16 open TFIM chain (nf=81 at generic h), J=1, "
                        f"h~U[0.5,1.5] seed 5, "
                        f"{nsets} parameter sets, Drude-Lorentz (0.05, 5, 1), Nt=2^12, dt=0.005, 200 RK4 steps of "
                        f"TCL2+L4, checkpoints every 64, 5 parameters (lambda, gamma, beta, J, h)",
            "path": "multi-set tape: every set's F/C/R, L4, TCL2 and their reverses in turn, the forward RK4 "
                    "and the reverse RK4 (recompute, walk) of all sets one launch each; F/C/R slices kept "
                    f"within a {slice_budget >> 30} GiB slice budget (whole, or packed to the entries L4 reads), "
                    "recomputed in the reverse pass past it; bath correlation and Gamma made and reversed on the "
                    "device; correctly rounded phases",
            "slices": {"sets": slice_modes, "GiB": round(slice_gib, 2)},
        },
        "forward_ms": 1e3 * t_fwd,
        "gradient_over_forward": t_grad / t_fwd,
        "set_by_set": {"sets": nseq, "ms_per_gradient": 1e3 * t_seq, "forward_ms": 1e3 * t_fwd,
                       "multi_set_speedup": t_seq / t_grad, "bit_equal_sets": same},
        "check": {"params": list(prob.names), "h": float(hs[0]), "adjoint": g0.tolist(), "gpu_fd": fdg.tolist(),
                  "max_rel_diff": float(rel.max()), "loss": loss0},
        "roofline": roofline,
        "cpu_baseline": None,
        "kernels": kernels,
        "kernels_note": "per parameter set over one multi-set gradient; k_l4 and k_combine include the L4 and "
                        "F/C/R reverses' time but only the forward kernels' bytes",
    }
    print(json.dumps(rec), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {a.gpus} is not supported (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "batched":
        return batched_main(a)
    if a.workload == "adjoint":
        return adjoint_main(a)
    if a.workload == "tfim_adjoint":
        return tfim_adjoint_main(a)
    import torch

    import taco_amd
    from taco_amd import tcl4

    assert torch.cuda.is_available(), "bench needs an MI355X"
    w = make_workload(a.workload, a.reduced)
    S, G = w.system, w.gamma
    ntau = len(w.tidx)
    dsys = tcl4.DeviceSystem(S, G, 0)
    out = torch.empty((ntau, S.N ** 4), dtype=torch.complex128, device="cuda:0")
    print(f"[bench] workload {w.name}: d={S.N} nf={S.nf} Nt={G.shape[0]} ntau={ntau}; warmup {a.warmup}",
          file=sys.stderr, flush=True)
    # first call: builds the libm phase table (host std::exp per Omega class,
    # cached in the context) and autotunes the FFT tiles; later calls reuse both
    first = []
    for _ in range(max(a.warmup, 2)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tcl4.build_l4_dev(dsys, w.dt, w.tidx, out)
        torch.cuda.synchronize()
        first.append(time.perf_counter() - t0)
    print(f"[bench] first call {first[0]:.2f} s, second {first[1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    with tcl4.Profile(0) as prof:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            tcl4.build_l4_dev(dsys, w.dt, w.tidx, out)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        kern = prof.read()
    print(f"[bench] timed {a.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    if a.copy_bytes:
        n = a.copy_bytes // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ms_per_step = 1000.0 * elapsed / a.steps
    value = ntau * a.steps / elapsed
    kernels, roofline = kernel_table(kern, a.steps)
    # HBM traffic of the dominant kernel measured by rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes of this workload (committed summary), when present
    tf = os.path.join(ROOT, "profiles", "fcr_l4_gue_pair_d4_traffic.json")
    if a.workload == "gue_pair" and not a.reduced and os.path.exists(tf):
        t = json.load(open(tf))
        # only when it was measured on launches of the same size as these
        same = abs(t.get("algo_bytes_per_launch", 0.0) - roofline["algo_bytes_per_launch"]) <= (
            0.01 * roofline["algo_bytes_per_launch"])
        if t.get("kernel") == roofline["kernel"] and same:
            roofline["traffic"] = t["traffic_bytes_per_launch"]
            roofline["traffic_source"] = t["source"]
    from taco_amd import _lib
    choice = np.zeros(3, np.int32)
    nfft = 1 << int(np.ceil(np.log2(2 * G.shape[0] - 1)))
    _lib.check(_lib.lib().taco_fft_tile_choice(tcl4.context(0).handle, int(np.log2(nfft)), _lib.iptr(choice)))
    tiles = ["8192/512", "4096/256", "4096/512", "2048/256", "4096/1024", "2048/512", "8192/1024"]
    fft_tiles = {ph: tiles[c] if c >= 0 else None for ph, c in zip("ABC", choice)}
    L4 = out.cpu().numpy().reshape(ntau, S.N ** 2, S.N ** 2).transpose(0, 2, 1)
    cpu, parity, ref_libm = None, None, None
    if not a.no_cpu:
        threads = a.cpu_threads or len(os.sched_getaffinity(0))
        print(f"[bench] cpu baseline / checker leg on {threads} threads", file=sys.stderr, flush=True)
        cpu, parity, ref_libm = cpu_baseline_leg(w, L4, threads)
        print(f"[bench] cpu leg done in {cpu['seconds']:.1f} s", file=sys.stderr, flush=True)
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "L4 time-points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": max(a.warmup, 2),
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: plain per-system fused L4 build, d={S.N}, nf={S.nf}, Nt={G.shape[0]}, "
                        f"Nfft={nfft}, dt={w.dt}, tidx 0..{w.tidx[-1]} ({ntau} L4 points), inputs resident in HBM",
            "path": "plain",
            "phases": "libm (host std::exp table per Omega class, cached per context)",
            "fft_tiles": fft_tiles,
        },
        "first_call_s": first[0],
        "second_call_ms": 1e3 * first[1],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "kernels": kernels,
    }
    if cpu is not None:
        cpu["cpu_model"] = cpu_model()
    if a.workload == "gue_pair" and not a.reduced and not a.plain_only:
        # the metric's other terms, each with its own timed count and wall time
        from taco_amd import workloads as W

        del dsys, out
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        print("[bench] sub-record: host-pointer calls (d=4 bench workload, published d=2 setting)", file=sys.stderr,
              flush=True)
        rec["host_pointer"] = [host_pointer_record(w), host_pointer_record(W.published())]
        rec["host_pointer_wall_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        print("[bench] sub-record: window FFT mode (d=4 bench workload)", file=sys.stderr, flush=True)
        rec["window"] = window_record(w, None if a.no_cpu else ref_libm,
                                      a.cpu_threads or len(os.sched_getaffinity(0)))
        rec["window"]["wall_s"] = time.perf_counter() - t0
        if not a.no_cpu:
            t0 = time.perf_counter()
            print("[bench] sub-record: serial CPU (published d=2 setting)", file=sys.stderr, flush=True)
            rec["cpu_serial"] = serial_cpu_record(W.published())
            rec["cpu_serial"]["cpu_model"] = cpu_model()
            rec["cpu_serial_wall_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        print("[bench] sub-record: batched, 256 systems", file=sys.stderr, flush=True)
        rec["batched"] = batched_record(a, steps=1, warmup=1)
        rec["batched"]["wall_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        print("[bench] sub-record: adjoint, d=4 spin pair", file=sys.stderr, flush=True)
        rec["adjoint"] = adjoint_record(a, steps=5, warmup=1)
        rec["adjoint"]["wall_s"] = time.perf_counter() - t0
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
