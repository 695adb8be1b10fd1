import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def _set_modes(mode_name: str):
    from oracle import oracle

    import taco_amd

    oracle.lib()
    oracle.set_phase_mode(getattr(oracle, mode_name))
    taco_amd.set_phase_mode(getattr(taco_amd, mode_name))
    return oracle


@pytest.fixture(scope="session")
def orc():
    """The oracle and the engine in the reference's phase convention: the
    F/C/R phases, the trapz phase step and the bath's exponentials are the
    platform libm's std::exp values (tcl4_kernels.cpp:116-120, gamma.hpp:185,
    correlation_fft.hpp:154-214)."""
    yield _set_modes("PHASE_LIBM")


@pytest.fixture
def orc_cr():
    """The oracle and the engine with correctly rounded phases and
    exponentials (the engine's TACO_PHASE_CR option); restores the libm
    convention afterwards."""
    yield _set_modes("PHASE_CR")
    _set_modes("PHASE_LIBM")


@pytest.fixture(params=["PHASE_LIBM", "PHASE_CR"])
def orc_any(request):
    """The oracle and the engine in each phase convention in turn."""
    yield _set_modes(request.param)
    _set_modes("PHASE_LIBM")
