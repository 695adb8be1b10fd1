"""GPU producers of the hot path's input: the bath correlation function
(bcf_fft_fun, correlation_fft.hpp:154-214) and the trapz Gamma table
(compute_trapz_prefix_multi_matrix, gamma.hpp:166-213) on the device must
equal the host producers bit for bit (which the CPU tests pin to the oracle);
so must the device exp against the host exp."""
import numpy as np
import pytest

import taco_amd
from taco_amd import _lib

pytestmark = pytest.mark.gpu


def test_device_cr_exp_bitwise_equals_host():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-745.2, 709.8, 200000), rng.uniform(-745.2, -707.0, 20000),
                        rng.uniform(-3, 3, 50000), [0.0, 710.0, -746.0, 709.782712893384, -745.1332191019411]])
    h = np.zeros_like(x)
    d = np.zeros_like(x)
    _lib.check(_lib.lib().taco_cr_exp(None, _lib.dptr(x), x.size, _lib.dptr(h)))
    _lib.check(_lib.lib().taco_cr_exp(taco_amd.context(0).handle, _lib.dptr(x), x.size, _lib.dptr(d)))
    assert np.array_equal(h, d)


CASES = [
    ("ohmic", taco_amd.Ohmic(0.05, 5.0), 1.0, 2 ** 14 - 1, 0.005),
    ("ohmic_bench", taco_amd.Ohmic(0.1, 10.0), 0.5, 200000, 0.002),
    ("ohmic_pair_grid", taco_amd.Ohmic(0.05, 5.0), 1.0, 2 ** 16 - 1, 0.005),
    ("drude", taco_amd.DrudeLorentz(0.05, 5.0), 1.0, 3000, 0.005),
    ("ohmic_zero_T", taco_amd.Ohmic(0.1, 3.0), float("inf"), 777, 0.01),
    ("drude_hot", taco_amd.DrudeLorentz(0.2, 20.0), 0.2, 2 ** 12, 0.005),
    ("tiny", taco_amd.Ohmic(0.3, 1.0), 2.0, 3, 0.1),
    ("two_samples", taco_amd.Ohmic(0.3, 1.0), 2.0, 1, 0.1),
    ("one_sample", taco_amd.Ohmic(0.3, 1.0), 2.0, 0, 0.1),
    ("one_sample_drude", taco_amd.DrudeLorentz(0.1, 3.0), 0.7, 0, 0.05),
]


@pytest.mark.parametrize("mode", ["PHASE_LIBM", "PHASE_CR"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_bath_gamma_bitwise_equals_host(case, mode):
    """In each phase convention: TACO_PHASE_LIBM feeds the device the host's
    std::exp values, TACO_PHASE_CR evaluates them on the device."""
    name, bath, beta, N, dt = case
    om = np.array([0.0, 1.25, -1.25, 2.5, -0.3333, 7.0])
    with taco_amd.phase_mode(getattr(taco_amd, mode)):
        check_bath_case(name, bath, beta, N, dt, om)


def check_bath_case(name, bath, beta, N, dt, om):
    _, C_h = taco_amd.bcf_fft_fun(N, dt, bath, beta)
    G_h = taco_amd.compute_trapz_prefix_multi_matrix(C_h, dt, om)
    C_d, G_d = taco_amd.bath_gamma_hip(N, dt, bath, beta, om)
    assert np.array_equal(C_d, C_h), name
    assert np.array_equal(G_d, G_h), name
    C_only, none = taco_amd.bath_gamma_hip(N, dt, bath, beta)
    assert none is None and np.array_equal(C_only, C_h)
    assert np.array_equal(taco_amd.compute_trapz_prefix_multi_matrix_hip(C_h, dt, om), G_h)

