"""Shared comparison helpers for the parity tests."""
import numpy as np


def same_bits(a, b):
    """Bitwise equality of float64 arrays, every NaN (a lost particle's x)
    counted as one value."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    ca = np.where(np.isnan(a), np.nan, a).view(np.uint64)
    cb = np.where(np.isnan(b), np.nan, b).view(np.uint64)
    return bool(np.array_equal(ca, cb))


def assert_coords_close(gpu, ref, rtol=1e-12, floor=None, what=""):
    """Per-coordinate |a-b| <= rtol*max(|a|,|b|,sigma_c); NaN patterns equal.
    sigma_c is the RMS of coordinate c over the live reference particles."""
    gpu = np.asarray(gpu).reshape(-1, 6)
    ref = np.asarray(ref).reshape(-1, 6)
    assert gpu.shape == ref.shape
    ng, nr = np.isnan(gpu), np.isnan(ref)
    assert np.array_equal(ng, nr), f"{what}: NaN (loss) pattern differs at {np.argwhere(ng != nr)[:5]}"
    live = ~np.isnan(ref[:, 0])
    if floor is None:
        floor = np.zeros(6)
        if live.any():
            floor = np.sqrt(np.mean(np.nan_to_num(ref[live]) ** 2, axis=0))
    scale = np.maximum(np.maximum(np.abs(gpu), np.abs(ref)), np.asarray(floor)[None, :])
    ok = ~nr
    err = np.where(ok, np.abs(gpu - ref), 0.0)
    bad = err > rtol * np.where(ok, scale, 1.0)
    if bad.any():
        i, j = np.argwhere(bad)[0]
        raise AssertionError(f"{what}: {bad.sum()} coords exceed rtol={rtol}; first p{i} c{j}: "
                             f"gpu={gpu[i, j]!r} ref={ref[i, j]!r} scale={scale[i, j]!r}")
    return float(np.max(np.where(ok, err / np.where(scale > 0, scale, 1.0), 0.0)))
