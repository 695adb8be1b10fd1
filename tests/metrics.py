"""Error metrics of the parity tests.

Per time t: e(t) = ||X_gpu(t) - X_ref(t)||_F / max(||X_ref(t)||_F, floor), with
floor = 1e-6 * max_s ||X_ref(s)||_F.  The floor only matters where X_ref(t) is
at rounding-noise level (L4 at t = 0, where Gamma(., 0) = 0); the count of
floored t is returned alongside.
"""
import numpy as np

FLOOR = 1e-6


def per_t_rel_err(got, ref, floor_frac=FLOOR):
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    n = ref.shape[0]
    d = np.linalg.norm((got - ref).reshape(n, -1), axis=1)
    r = np.linalg.norm(ref.reshape(n, -1), axis=1)
    floor = floor_frac * (r.max() if r.size else 0.0)
    den = np.maximum(r, floor)
    den[den == 0] = 1.0
    return d / den, int((r < floor).sum())


def max_rel_err(got, ref, floor_frac=FLOOR):
    e, nfloor = per_t_rel_err(got, ref, floor_frac)
    return float(e.max()) if e.size else 0.0, nfloor


def ref_style_err(got, ref):
    """max_abs / max(1, max|ref|) of tcl4_e2e_cuda_compare.cpp:221-223."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    return float(np.abs(got - ref).max() / max(1.0, np.abs(ref).max()))
