"""The C++ drop-in header (include/taco/tcl4_hip.hpp) bound to Eigen-like and
System-like types, as the reference's callers would use it."""
import os
import subprocess

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
