"""GPU parity of the batched run (taco_batch_* of the C ABI): many systems of
uneven N, bath, grid length and time set in one device pass.

Per system the batch must equal the plain builders bit for bit (same kernels,
same arithmetic: build_TCL4_generator_hip_fused_batch for L4 at tidx,
tcl4_propagate for rho), and through them the oracle within the per-t bar
(relative Frobenius error <= 1e-10, floored as in tests/metrics.py)."""
import numpy as np
import pytest

import taco_amd
from taco_amd import workloads as W
from taco_amd._lib import TacoInvalidArgument, TacoOutOfRange

from .metrics import max_rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def r0_for(N):
    r = np.zeros(N * N, complex)
    r[0] = 1.0
    return r


def small_batch():
    """Five systems: d = 2, 3, 4 (two), 8; Nt from 2^10 to 2^12; Ohmic and
    Drude-Lorentz; different strides, RK4 lengths and generator orders."""
    items, names = [], []
    sb = W.spin_boson(Nt=2048, tmax=300, tstep=3)
    items.append(taco_amd.BatchItem(sb.system, sb.gamma, sb.dt, sb.tidx, steps=300, r0=r0_for(2)))
    names.append("spin_boson")
    for b in (1, 2, 6):
        w = W.heterogeneous_system(b, dims=(3, 4), n_time=120, log2_range=(10, 12))
        items.append(taco_amd.BatchItem(w.system, w.gamma, w.dt, w.tidx, steps=150, r0=r0_for(w.system.N),
                                        order=4 if b != 2 else 2))
        names.append(w.name)
    rng = W.SplitMix64(77)
    d8 = W.Workload("d8", W.gue(8, rng), [W.random_hermitian_unit(8, rng)], taco_amd.DrudeLorentz(0.1, 5.0),
                    1.0, 0.01, 299, [0, 1, 7, 50, 299])
    items.append(taco_amd.BatchItem(d8.system, d8.gamma, d8.dt, d8.tidx, steps=0))
    names.append("d8")
    return items, names


@pytest.fixture(scope="module")
def batch_and_items():
    items, names = small_batch()
    b = taco_amd.Batch(items).run()
    yield b, items, names
    b.close()


def test_batch_equals_plain_builders_bitwise(batch_and_items):
    b, items, names = batch_and_items
    for s, it in enumerate(items):
        plain = taco_amd.build_TCL4_generator_hip_fused_batch(it.system, it.gamma, it.dt, list(it.tidx))
        assert np.array_equal(b.l4(s), plain), names[s]
        if it.steps:
            rho = taco_amd.tcl4_propagate(it.system, it.gamma, it.dt, it.steps, it.r0, order=it.order)
            assert np.array_equal(b.rho(s), rho), names[s]


def test_batch_matches_oracle(orc, batch_and_items):
    b, items, names = batch_and_items
    for s, it in enumerate(items):
        S, G = it.system, it.gamma
        ref = orc.l4_series(G, S.omegas, S.mirror, S.pair_to_freq, S.A_eig, it.dt, list(it.tidx))
        e, nfloor = max_rel_err(b.l4(s), ref)
        print(f"{names[s]}: d={S.N} nf={S.nf} Nt={G.shape[0]} L4 max per-t rel err {e:.3e} (floored {nfloor})")
        assert e <= TOL, names[s]
        if it.steps:
            rr = orc.tcl4_propagate(G, S.omegas, S.mirror, S.pair_to_freq, S.eps, S.A_eig, it.dt, it.steps, it.r0,
                                    order=it.order)
            er, _ = max_rel_err(b.rho(s), rr)
            print(f"{names[s]}: rho max per-t rel err {er:.3e}")
            assert er <= TOL, names[s]


def test_graph_replays_equal_eager():
    """The captured hipGraph, replayed several times with eager engine work
    and synchronisations in between, reproduces the eager results exactly."""
    items, _ = small_batch()
    items = items[:4]
    b = taco_amd.Batch(items).run()
    eager_l4 = [b.l4(s) for s in range(len(items))]
    eager_rho = [b.rho(s) for s in range(len(items))]
    w = W.spin_boson(Nt=512, tmax=40, tstep=4)
    for rep in range(3):
        b.run(graph=True)
        for s in range(len(items)):
            assert np.array_equal(b.l4(s), eager_l4[s]), (rep, s)
            assert np.array_equal(b.rho(s), eager_rho[s]), (rep, s)
        # unrelated eager work on the same context between replays
        taco_amd.build_TCL4_generator_hip_fused_batch(w.system, w.gamma, w.dt, w.tidx)
    b.close()


def test_graph_replay_after_workspace_growth():
    """Calls that grow the context's device workspace between replays (a
    larger plain build, a second and larger batch, a gradient tape) move the
    buffers the captured graph addressed; the batch captures again and its
    replays stay bit-identical to the eager run."""
    items, _ = small_batch()
    b = taco_amd.Batch(items[:2]).run()
    eager_l4 = [b.l4(s) for s in range(2)]
    eager_rho = [b.rho(s) for s in range(2)]
    b.run(graph=True)
    big = W.gue_pair(N_bath=30000, tmax=2500)
    taco_amd.build_TCL4_generator_hip_fused_batch(big.system, big.gamma, big.dt, big.tidx)
    b2 = taco_amd.Batch(items).run(graph=True)
    for rep in range(2):
        b.run(graph=True)
        for s in range(2):
            assert np.array_equal(b.l4(s), eager_l4[s]), (rep, s)
            assert np.array_equal(b.rho(s), eager_rho[s]), (rep, s)
        b2.run(graph=True)
        w = W.gue_pair(N_bath=40000 + 20000 * rep, tmax=3000)
        taco_amd.build_TCL4_generator_hip_fused_batch(w.system, w.gamma, w.dt, w.tidx)
    for s, it in enumerate(items):
        plain = taco_amd.build_TCL4_generator_hip_fused_batch(it.system, it.gamma, it.dt, list(it.tidx))
        assert np.array_equal(b2.l4(s), plain), s
    b.close()
    b2.close()


def test_batch_error_paths():
    sb = W.spin_boson(Nt=256, tmax=20, tstep=2)
    S, G = sb.system, sb.gamma
    with pytest.raises(TacoInvalidArgument):
        taco_amd.Batch([taco_amd.BatchItem(S, G, sb.dt, sb.tidx, steps=10, r0=r0_for(2), order=3)])
    with pytest.raises(TacoOutOfRange):
        taco_amd.Batch([taco_amd.BatchItem(S, G, sb.dt, [0, 256])])
    with pytest.raises(TacoOutOfRange):
        taco_amd.Batch([taco_amd.BatchItem(S, G, sb.dt, [0], steps=300, r0=r0_for(2))])
    with pytest.raises(TacoInvalidArgument):
        taco_amd.Batch([taco_amd.BatchItem(S, G, -1.0, [0])])
    # an empty batch is valid and does nothing
    taco_amd.Batch([]).run().close()
