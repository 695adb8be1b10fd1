"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, its host-side input preparation equals the oracle's bit for
bit, and argument errors come back as status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import taco_amd
from taco_amd import _lib, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(taco_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_version_and_pad_factor():
    assert "gfx950" in taco_amd.version()
    taco_amd.set_fcr_fft_pad_factor(4)
    assert taco_amd.get_fcr_fft_pad_factor() == 4
    taco_amd.set_fcr_fft_pad_factor(0)


@pytest.mark.parametrize("kind,p0,p1,beta", [(0, 0.05, 5.0, 1.0), (0, 0.1, 10.0, 0.5), (1, 0.1, 7.0, 2.0),
                                              (0, 0.3, 4.0, np.inf)])
def test_bcf_fft_bitwise_equals_oracle(orc_any, kind, p0, p1, beta):
    orc = orc_any
    bath = taco_amd.Ohmic(p0, p1) if kind == 0 else taco_amd.DrudeLorentz(p0, p1)
    _, C = taco_amd.bcf_fft_fun(4095, 0.005, bath, beta)
    assert np.array_equal(C, orc.bcf_fft(4095, 0.005, beta, kind, p0, p1))


def test_gamma_trapz_bitwise_equals_oracle(orc_any):
    orc = orc_any
    _, C = taco_amd.bcf_fft_fun(3000, 0.002, taco_amd.Ohmic(0.1, 10.0), 0.5)
    om = np.array([-2.3, -0.4, 0.0, 0.4, 2.3, 11.7])
    assert np.array_equal(taco_amd.compute_trapz_prefix_multi_matrix(C, 0.002, om), orc.gamma_trapz(C, 0.002, om))


def test_phase_mode_switches_the_host_producers(orc):
    """The two conventions differ in last bits somewhere on the bench bath
    (libm's exp is not correctly rounded everywhere), and the mode is
    process-global and validated."""
    bath = taco_amd.Ohmic(0.1, 10.0)
    _, C_libm = taco_amd.bcf_fft_fun(200000, 0.002, bath, 0.5)
    with taco_amd.phase_mode(taco_amd.PHASE_CR):
        assert taco_amd.get_phase_mode() == taco_amd.PHASE_CR
        _, C_cr = taco_amd.bcf_fft_fun(200000, 0.002, bath, 0.5)
    assert taco_amd.get_phase_mode() == taco_amd.PHASE_LIBM
    assert not np.array_equal(C_libm, C_cr)
    assert np.abs(C_libm - C_cr).max() < 1e-12 * np.abs(C_cr).max()
    with pytest.raises(ValueError):
        taco_amd.set_phase_mode(7)


@pytest.mark.parametrize("wl", ["spin_boson", "gue_pair", "spin_pair", "published"])
def test_system_build_matches_oracle(orc, wl):
    w = getattr(W, wl)()
    S = w.system
    O = orc.system_build(w.H, w.ops)
    assert S.nf == O["nf"]
    assert np.array_equal(S.pair_to_freq, O["pair_to_freq"])
    assert np.array_equal(S.mirror, O["mirror"])
    assert S.zero_index == O["zero_index"]
    assert np.abs(S.omegas - O["omegas"]).max() < 1e-12
    assert np.abs(S.eps - O["eps"]).max() < 1e-12
    # eigenvectors agree up to the (shared) phase convention
    assert np.abs(S.U - O["U"]).max() < 1e-10 or wl == "spin_pair"  # degenerate pair: gauge unpinned
    assert np.abs(S.H @ S.U - S.U @ np.diag(S.eps)).max() < 1e-12


def test_workload_bucket_counts():
    assert W.spin_boson().system.nf == 3
    assert W.gue_pair().system.nf == 13
    assert W.spin_pair().system.nf == 5
    assert W.published().system.nf == 3


def test_errors_without_gpu_are_status_codes():
    L = _lib.lib()
    z = np.zeros(8)
    rc = L.taco_l4_batch(None, _lib.dptr(z), 4, 1, _lib.dptr(z), None, None, 1, _lib.dptr(z), 1, 0.1, 0, None,
                         0, _lib.dptr(z), None)
    assert rc == _lib.TACO_ERR_INVALID
    assert "ctx" in L.taco_last_error().decode()
    with pytest.raises(_lib.TacoInvalidArgument):
        taco_amd.bcf_fft_fun(10, -1.0, taco_amd.Ohmic(1, 1), 1.0)
    with pytest.raises(_lib.TacoInvalidArgument):
        taco_amd.System.build(np.array([[0, 1], [0, 0]], complex), [np.eye(2)])
    with pytest.raises(_lib.TacoInvalidArgument):
        taco_amd.Exec(backend=taco_amd.Backend.MpiOmp) and taco_amd.tcl4._ctx_for(
            taco_amd.Exec(backend=taco_amd.Backend.MpiOmp))


def test_ctx_create_fails_cleanly_without_device():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    rc = _lib.lib().taco_ctx_create(0, ctypes.byref(h))
    assert rc != _lib.TACO_OK

