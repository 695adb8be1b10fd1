"""Device system preparation (taco_system_build_hip): eigensystem, gauge,
buckets, mirror, zero bucket and A_eig for many Hamiltonians in one launch,
equal to the host builder (taco_system_build) bit for bit."""
import numpy as np
import pytest

import taco_amd
from taco_amd import workloads as W
from taco_amd._lib import TacoInvalidArgument

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.nf == b.nf
    assert np.array_equal(a.eps, b.eps)
    assert np.array_equal(a.U, b.U)
    for x, y in zip(a.A_eig, b.A_eig):
        assert np.array_equal(x, y)
    assert np.array_equal(a.omegas, b.omegas)
    assert np.array_equal(a.mirror, b.mirror)
    assert np.array_equal(a.pair_to_freq, b.pair_to_freq)
    assert a.zero_index == b.zero_index


@pytest.mark.parametrize("d", [1, 2, 3, 4, 6, 8, 16, 64])
def test_device_build_equals_host_gue(d):
    rng = W.SplitMix64(70 + d)
    Hs = [W.gue(d, rng) for _ in range(5)]
    ops = [W.random_hermitian_unit(d, rng), W.random_hermitian_unit(d, rng)] if d > 1 else [np.eye(1, dtype=complex)]
    dev = taco_amd.System.build_many(Hs, ops)
    for H, S in zip(Hs, dev):
        _same(S, taco_amd.System.build(H, ops))
        assert np.abs(H @ S.U - S.U @ np.diag(S.eps)).max() < 1e-12 * max(1.0, np.abs(S.eps).max())


def test_device_build_equals_host_structured():
    """Degenerate spectra (spin pair: eigenvalue 0 twice; TFIM at h = 1 with
    57 buckets), a diagonal H (no rotations) and per-system operator lists."""
    Hs = [W.spin_pair().H, W.tfim(4, 1.0, 1.0)[0], W.tfim(4, 1.0, 0.77)[0]]
    dev = taco_amd.System.build_many(Hs[1:], [W.tfim(4, 1.0, 1.0)[1]])
    for H, S in zip(Hs[1:], dev):
        _same(S, taco_amd.System.build(H, [W.tfim(4, 1.0, 1.0)[1]]))
    assert [S.nf for S in dev] == [57, 81]
    sp = W.spin_pair()
    (S,) = taco_amd.System.build_many([sp.H], [sp.ops])
    _same(S, taco_amd.System.build(sp.H, sp.ops))
    assert S.nf == 5
    D = np.diag([0.3, -1.0, 0.3, 2.0]).astype(complex)
    A = W.random_hermitian_unit(4, W.SplitMix64(3))
    (S,) = taco_amd.System.build_many([D], [A])
    _same(S, taco_amd.System.build(D, [A]))
    assert np.array_equal(S.eps, [-1.0, 0.3, 0.3, 2.0])


def test_device_build_errors():
    with pytest.raises(TacoInvalidArgument, match="Hermitian"):
        taco_amd.System.build_many([np.array([[0, 1], [0, 0]], complex)], [np.eye(2)])
    with pytest.raises(TacoInvalidArgument, match="N <= 64"):
        taco_amd.System.build_many([np.eye(65, dtype=complex)], [np.eye(65)])
