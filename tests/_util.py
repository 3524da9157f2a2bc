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


def _tile(a, ax_x, ax_y):
    """The interior of a (2 halo cells per side) repeated ax_x times along x
    and ax_y times along y, with the halo filled by the periodic extension."""
    inner = np.tile(a[2:-2, 2:-2], (ax_y, ax_x))
    return np.pad(inner, 2, mode="wrap")


PERIODIC_FIELDS = ("hhq_rest", "ssh", "sshp", "u", "up", "v", "vp", "wflux", "mu", "mu4", "taux", "tauy",
                   "rlh_s", "dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")


def periodic_pair(periodicity, nxi=40, nyi=24, seed=12):
    """A basin periodic in x (a zonal channel with land rows north and south)
    or in x and y (wet everywhere), with bathymetry, state and water flux
    varying along x; and the same basin tiled twice along x (and along y
    when periodic in y).  Both hold the periodic extension in their halos,
    as the caller's wrap of the initial fields leaves them.  Returns (b,
    b2, (tx, ty))."""
    import math

    from inmsom_bt.cases import make_basin, randomize_state

    px, py = periodicity
    b = randomize_state(make_basin(nxi, nyi, seed=seed, sigma=3.0), 5, wflux_amp=1e-5)
    wet = np.zeros((b.ny, b.nx), dtype=bool)
    wet[2:-2, 2:-2] = True
    if not py:
        wet[2, :] = wet[-3, :] = False
    b.lbasins = np.where(wet, 0, 1).astype(np.int32)
    x = np.arange(b.nx)[None, :]
    b.hhq_rest = np.where(wet, 60.0 + 15.0 * np.sin(2 * math.pi * (x - 2) / nxi)
                          + 5.0 * np.cos(np.arange(b.ny)[:, None] / 3.0), 0.0)
    for k in ("ssh", "sshp", "u", "up", "v", "vp"):
        setattr(b, k, getattr(b, k) * wet)
    b.mu = 10.0 * wet
    for k in PERIODIC_FIELDS:
        setattr(b, k, _tile(getattr(b, k), 1, 1))
    b.lbasins = _tile(b.lbasins, 1, 1).astype(np.int32)
    if not py:  # no wrap along y: the halo rows of the channel stay land
        b.lbasins[:2, :] = b.lbasins[-2:, :] = 1
    tx, ty = 2, (2 if py else 1)
    b2 = b.copy()
    b2.nx, b2.ny = tx * (b.nx - 4) + 4, ty * (b.ny - 4) + 4
    for k in PERIODIC_FIELDS:
        setattr(b2, k, _tile(getattr(b, k), tx, ty))
    b2.lbasins = _tile(b.lbasins, tx, ty).astype(np.int32)
    if not py:
        b2.lbasins[:2, :] = b2.lbasins[-2:, :] = 1
    return b, b2, (tx, ty)
