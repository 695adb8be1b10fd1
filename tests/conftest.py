import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def orc():
    """The oracle with correctly rounded phases (the engine's convention).

    The reference's phases come from the platform libm's sincos, whose last
    bits are not portable; tests that need the libm convention switch with
    ``oracle.phase_mode(oracle.PHASE_LIBM)``."""
    from oracle import oracle

    oracle.lib()
    oracle.set_phase_mode(oracle.PHASE_CR)
    return oracle
