"""The allocation-free entry point taco_l4_from_slices_noalloc (device
pointers, a stream, scratch in a file-scope device array) against
taco_l4_from_slices, the context entry point that shares its kernel: equal
bit for bit, on two calls in a row, the second with a smaller input.  The
file is named to be collected last."""
import numpy as np
import pytest
import torch

import taco_amd
from taco_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu


def slices_for(w, npos, seed):
    """Random F/C/R slices (complex, [npos][nf^3]) for the system of w."""
    rng = np.random.default_rng(seed)
    nf3 = w.system.nf ** 3
    return [rng.normal(size=(npos, nf3)) + 1j * rng.normal(size=(npos, nf3)) for _ in range(3)]


def reference_l4(S, sl, npos):
    """taco_l4_from_slices (host pointers, one time per call)."""
    N = S.N
    ops = np.ascontiguousarray(np.stack([np.asarray(a).T for a in S.A_eig]).reshape(-1))
    p2f = np.ascontiguousarray(np.asarray(S.pair_to_freq, np.int32).T.reshape(-1))
    L4 = np.zeros((npos, N ** 4), np.complex128)
    GW = np.zeros((npos, N ** 4), np.complex128)
    ctx = taco_amd.context(0).handle
    for p in range(npos):
        F, C, R = (np.ascontiguousarray(x[p]) for x in sl)
        l4 = np.zeros(N ** 4, np.complex128)
        gw = np.zeros(N ** 4, np.complex128)
        _lib.check(_lib.lib().taco_l4_from_slices(ctx, _lib.dptr(F.view(np.float64)), _lib.dptr(C.view(np.float64)),
                                                  _lib.dptr(R.view(np.float64)), _lib.iptr(p2f), N, S.nf,
                                                  _lib.dptr(ops.view(np.float64)), len(S.A_eig),
                                                  _lib.dptr(l4.view(np.float64)), _lib.dptr(gw.view(np.float64))))
        L4[p], GW[p] = l4, gw
    return L4, GW, ops, p2f


def noalloc_l4(S, sl, npos, ops, p2f):
    dev = torch.device("cuda:0")
    d_sl = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in sl]
    d_ops = torch.from_numpy(ops).to(dev)
    d_p2f = torch.from_numpy(p2f).to(dev)
    N = S.N
    d_l4 = torch.full((npos, N ** 4), complex("nan"), dtype=torch.complex128, device=dev)
    d_gw = torch.full((npos, N ** 4), complex("nan"), dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()
    _lib.check(_lib.lib().taco_l4_from_slices_noalloc(
        d_sl[0].data_ptr(), d_sl[1].data_ptr(), d_sl[2].data_ptr(), d_p2f.data_ptr(), N, S.nf, d_ops.data_ptr(),
        len(S.A_eig), npos, d_l4.data_ptr(), d_gw.data_ptr(), stream.cuda_stream))
    torch.cuda.synchronize()
    return d_l4.cpu().numpy(), d_gw.cpu().numpy()


def test_noalloc_equals_context_entry_point_twice():
    """d = 16 (one time's intermediates ~1 MB, so 200 times run in four
    launches of the 64 MiB scratch), then d = 4 with fewer times: the second
    call reuses the scratch the first left behind."""
    cases = [(W.tfim_workload(0.9, Nt=64, tmax=10), 200, 11), (W.gue_pair(N_bath=100, tmax=10), 37, 12)]
    for w, npos, seed in cases:
        S = w.system
        sl = slices_for(w, npos, seed)
        L4_ref, GW_ref, ops, p2f = reference_l4(S, sl, npos)
        L4, GW = noalloc_l4(S, sl, npos, ops, p2f)
        print(f"{w.name}: N={S.N} nf={S.nf} npos={npos}: L4 equal {np.array_equal(L4, L4_ref)}")
        assert np.array_equal(L4, L4_ref), w.name
        assert np.array_equal(GW, GW_ref), w.name


def test_noalloc_refuses_oversized_and_flags_bad_maps():
    w = W.gue_pair(N_bath=100, tmax=10)
    S = w.system
    sl = slices_for(w, 3, 5)
    _, _, ops, p2f = reference_l4(S, sl, 3)
    bad = p2f.copy()
    bad[5] = S.nf  # outside [0, nf): that time's L4 comes back NaN
    L4, GW = noalloc_l4(S, sl, 3, ops, bad)
    assert np.isnan(L4).all() and np.isnan(GW).all()
    with pytest.raises(_lib.TacoInvalidArgument):
        # N = 40, nf = 1561: 120 MB of intermediates per time (refused before any launch)
        _lib.check(_lib.lib().taco_l4_from_slices_noalloc(1, 1, 1, 1, 40, 1561, 1, 1, 1, 1, None, None))
