"""The configuration-file driver (examples/tcl_driver.cpp on the GPU engine)
against the oracle: C(t) from the J expression, the rect/trapz Gamma rules,
L4 and GW at tidx, RK4 propagation (TCL2 + L4) with either F/C/R method, and
the CSV outputs."""
import io

import numpy as np
import pytest

import taco_amd
from taco_amd import driver
from taco_amd import io as tio

from .metrics import max_rel_err

pytestmark = pytest.mark.gpu


def config(tmp_path, method="convolution", rule="rect", propagate=True, rho0="+x", H=None, A=None):
    H = H if H is not None else {"re": [[0.0, 0.5], [0.5, 0.0]], "im": [[0.0, 0.0], [0.0, 0.0]]}
    A = A if A is not None else [{"re": [[0.5, 0.0], [0.0, -0.5]]}]
    return {
        "system": {"freq_tol": 1e-9, "H": H, "A": A},
        "bath": {"beta": 0.5, "J_expr": "w * exp(-w/omega_c) * step(w)", "params": {"omega_c": 10.0}},
        "numerics": {"dt": 0.01, "N": 300, "bcf_N": 1024, "gamma_rule": rule,
                     "tcl4": {"method": method, "nmax": 2, "fcr_fft_pad": 0}},
        "output": {"tidx": "last", "profile": False, "gw_csv": str(tmp_path / "gw.csv"),
                   "l4_csv": str(tmp_path / "l4.csv")},
        "simulation": {"propagate": propagate, "order": 4, "rho0": rho0, "print_series": True,
                       "sample_every": 50, "rho_out": str(tmp_path / "rho.csv")},
    }


@pytest.mark.parametrize("method,rule", [("convolution", "rect"), ("convolution", "trapz"), ("direct", "rect")])
def test_driver_matches_oracle(orc, tmp_path, method, rule):
    cfg = config(tmp_path, method, rule)
    buf = io.StringIO()
    res = driver.run(cfg, "test.yaml", out=buf)
    text = buf.getvalue()
    assert "dim=2, channels=1, nf=3" in text and "tidx=300" in text
    S = res["system"]
    # C(t): the expression J equals the Ohmic form alpha = 1 (bit for bit), sliced from bcf_N
    _, C = taco_amd.bcf_fft_fun(1024, 0.01, taco_amd.Ohmic(1.0, 10.0), 0.5)
    assert np.array_equal(res["C"], C[:301])
    Gref = (orc.gamma_rect if rule == "rect" else orc.gamma_trapz)(C[:301], 0.01, S.omegas)
    assert np.array_equal(res["gamma"], Gref)
    m = "direct" if method == "direct" else "conv"
    L4ref = orc.l4_series(Gref, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, 0.01, [300], method=m)[0]
    assert max_rel_err(res["L4"][None], L4ref[None])[0] <= 1e-10
    # GW through the exact reshuffle, and the CSV files
    assert np.array_equal(orc.gw_to_l4(res["GW"], 2), res["L4"])
    assert np.array_equal(tio.read_csv_matrix(tmp_path / "l4.csv"),
                          np.vectorize(lambda x: float(f"{x:.17f}"))(res["L4"].real) +
                          1j * np.vectorize(lambda x: float(f"{x:.17f}"))(res["L4"].imag))
    # propagation to tidx against the oracle's driver composition
    r0 = driver.parse_rho0("+x", 2).T.reshape(-1)
    Rref = orc.tcl4_propagate(Gref, S.omegas, S.mirror, S.pair_to_freq, S.eps, S.A_eig, 0.01, 300, r0,
                              method=m)
    assert max_rel_err(res["samples"], Rref)[0] <= 1e-10
    rho_csv = tio.read_csv_matrix(tmp_path / "rho.csv")
    assert np.abs(rho_csv - res["rho_final"]).max() <= 1e-16
    assert abs(np.trace(res["rho_final"]) - 1) < 1e-15
    series = text.split("t,sz,tr,purity\n")[1].split("rho(tf)")[0].strip().splitlines()
    assert len(series) == 7 and series[-1].startswith("3.000000000,")  # t = 0, 0.5, ..., 3.0


def test_driver_errors_and_cli(tmp_path, capsys):
    cfg = config(tmp_path, propagate=False)
    cfg["numerics"]["gamma_rule"] = "simpson"
    with pytest.raises(RuntimeError, match="gamma_rule"):
        driver.run(cfg, out=io.StringIO())
    cfg = config(tmp_path, propagate=False)
    cfg["bath"]["J_expr"] = "w * q"
    with pytest.raises(taco_amd._lib.TacoInvalidArgument, match="Unknown identifier"):
        driver.run(cfg, out=io.StringIO())
    cfg = config(tmp_path, propagate=False)
    cfg["output"]["tidx"] = 301
    with pytest.raises(RuntimeError, match="tidx out of range"):
        driver.run(cfg, out=io.StringIO())
    import yaml
    p = tmp_path / "c.yaml"
    cfg = config(tmp_path, propagate=False)
    p.write_text(yaml.safe_dump(cfg))
    assert driver.main([f"--config={p}"]) == 0
    assert "GW_raw: 4x4" in capsys.readouterr().out
    cfg["system"]["A"] = []
    p.write_text(yaml.safe_dump(cfg))
    assert driver.main([f"--config={p}"]) == 1
    assert capsys.readouterr().err.startswith("Error: system.A must be a non-empty sequence")
