"""The C++ drop-in header (include/taco/tcl4_hip.hpp) bound to Eigen-like and
System-like types, as the reference's callers would use it."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_test")


def build():
    subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp")], check=True, stdout=subprocess.DEVNULL)
    return EXE


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


def test_dropin_compiles_and_maps_errors_without_gpu():
    exe = build()
    if has_gpu():
        pytest.skip("a GPU is present; the gpu variant covers the drop-in")
    out = subprocess.run([exe, "nogpu"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "runtime_error" in out.stdout


@pytest.mark.gpu
def test_dropin_matches_python_path():
    import taco_amd
    from taco_amd import workloads as W
    exe = build()
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        if line.startswith("L4 "):
            _, q, e, re_, im = line.split()
            vals[(int(q), int(e))] = complex(float(re_), float(im))
    w = W.spin_boson(Nt=2048)
    L = taco_amd.build_TCL4_generator_hip_fused_batch(w.system, w.gamma, w.dt, [0, 7, 100])
    for q in range(3):
        flat = L[q].T.ravel()  # column-major
        for e in range(16):
            assert vals[(q, e)] == flat[e]
    assert "exceptions_ok 1" in out.stdout
    rk = [l for l in out.stdout.splitlines() if l.startswith("rk4 ")][0].split()
    assert abs(float(rk[1]) - np.exp(-0.4)) < 1e-6 and int(rk[2]) == 3
    lines = {l.split()[0] + ("" if not l.startswith("fcr ") else l.split()[1]): l.split()
             for l in out.stdout.splitlines()}
    # compute_triple_kernels with no template arguments, and its _cuda name
    assert int(lines["series_len"][1]) == 2048 and lines["series_cuda_equal"][1] == "1"
    got = taco_amd.compute_triple_kernels(w.system, w.gamma, w.dt)
    for t in (1, 5, 300):
        v = [float(x) for x in lines[f"fcr{t}"][2:]]
        assert complex(v[0], v[1]) == got.F[1, 2, 0, t]
        assert complex(v[2], v[3]) == got.C[0, 1, 2, t]
        assert complex(v[4], v[5]) == got.R[2, 2, 1, t]
    # the four propagate overloads on a constant generator agree; samples at steps 0, 5, 10
    ov = lines["rk4_overloads"]
    assert ov[1] == ov[2] == ov[3] == ov[4] and int(ov[5]) == 3
    # two rk4_update steps: the fourth-order Taylor factor of each eigenvalue, squared
    p4 = lambda z: 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24  # noqa: E731
    ru = lines["rk4_update"]
    assert abs(float(ru[1]) - p4(-0.01) ** 2) < 1e-15 and abs(float(ru[2]) - p4(-0.04) ** 2) < 1e-15
    # device helpers: GW from explicit MIKX equals the fused kernel's GW, raw + raw^H == GW
    assert float(lines["device_gw"][2]) < 1e-13 and float(lines["device_gw"][4]) < 1e-15
    from oracle import oracle
    Fs, Cs, Rs = (np.ascontiguousarray(x[..., 300]).reshape(-1) for x in (got.F, got.C, got.R))
    M, _, _, X = oracle.mikx(Fs, Cs, Rs, w.system.pair_to_freq, 2)
    mk = [float(x) for x in lines["mikx"][2:4] + lines["mikx"][5:7]]
    assert complex(mk[0], mk[1]) == M[0, 0] and complex(mk[2], mk[3]) == X[5]
    # compute_fcr_convolution_batched equals the full series bit for bit
    assert lines["fcr_batched_equal"][1] == "1"


def test_python_taco_package_mirrors_reference_name():
    """python/taco (the reference's package location) imports on the engine
    and exposes version() like the reference's bindings, plus the engine's
    host interface."""
    code = ("import sys; sys.path.insert(0, 'python'); import taco; "
            "print(taco.version()); print(hasattr(taco, 'System'), hasattr(taco.adjoint, 'BathTape'))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    assert lines[0].startswith("taco-mi355x")
    assert lines[1] == "True True"
