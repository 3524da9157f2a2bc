"""Comparison helpers shared by the tests."""
import numpy as np

STATE = ("ssh", "sshp", "u", "up", "v", "vp")


def bitwise_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(
        np.ascontiguousarray(a, dtype=np.float64).view(np.uint64),
        np.ascontiguousarray(b, dtype=np.float64).view(np.uint64))


def assert_bitwise(got: dict, want: dict, keys, label=""):
    bad = []
    for k in keys:
        if not bitwise_equal(got[k], want[k]):
            d = np.abs(got[k] - want[k])
            i = np.unravel_index(np.nanargmax(d), d.shape) if np.isfinite(d).any() else None
            bad.append(f"{k}: max|diff|={np.nanmax(d):.3e} at {i}")
    assert not bad, f"{label} not bit-exact: " + "; ".join(bad)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))
