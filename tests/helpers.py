"""Shared helpers of the GPU parity tests (comparison and reporting)."""
import numpy as np


def assert_same_indices(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, (name, got.shape, want.shape)
    bad = np.argwhere(got != want)
    if bad.size:
        i = tuple(bad[0])
        raise AssertionError(f"{name}: {len(bad)} of {got.size} indices differ; "
                             f"first at {i}: got {got[i]}, want {want[i]}")


def normwise_err(got, ref):
    """max |got - ref| / max |ref| (0 when both are zero)."""
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max()
    d = np.abs(np.asarray(got, np.float64) - ref).max()
    return 0.0 if scale == 0 and d == 0 else d / max(scale, 1e-300)


def rel_err(got, ref, floor=1e-3):
    """Largest elementwise relative error over values above floor * max|ref|."""
    ref = np.asarray(ref, np.float64)
    big = np.abs(ref) > floor * np.abs(ref).max()
    if not big.any():
        return 0.0
    got = np.asarray(got, np.float64)
    return float((np.abs(got[big] - ref[big]) / np.abs(ref[big])).max())


def ulp_diff(got, want):
    """Largest distance in float32 ulps (same-sign finite values)."""
    g = np.asarray(got, np.float32).view(np.int32).astype(np.int64)
    w = np.asarray(want, np.float32).view(np.int32).astype(np.int64)
    return int(np.abs(g - w).max()) if g.size else 0


def padded(oracle, img_cm, rows, cols):
    n, _, _ = oracle.pad_geometry(rows, cols)
    return oracle.pad(img_cm.ravel(), rows, cols), n
