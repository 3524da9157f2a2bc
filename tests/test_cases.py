"""The synthetic-basin generator: determinism, closed land ring, depth
range, wind and Coriolis profiles, and the absence of wet cells without a
velocity point on either axis."""
import math

import numpy as np

from inmsom_bt.cases import coarsen_basin, make_basin
from inmsom_bt.workloads import single_basin


def test_deterministic_for_a_seed():
    a, b = make_basin(70, 50, seed=7), make_basin(70, 50, seed=7)
    c = make_basin(70, 50, seed=8)
    assert np.array_equal(a.hhq_rest, b.hhq_rest) and np.array_equal(a.lbasins, b.lbasins)
    assert not np.array_equal(a.hhq_rest, c.hhq_rest)


def test_land_ring_depths_and_profiles():
    b = make_basin(90, 60, seed=3)
    wet = b.lbasins == 0
    assert not wet[:2].any() and not wet[-2:].any() and not wet[:, :2].any() and not wet[:, -2:].any()
    h = b.hhq_rest[wet]
    assert h.min() >= 10.0 and h.max() <= 200.0
    assert np.all(b.hhq_rest[~wet] == 0.0)
    # wind 0.1 cos(pi y / Ly) on T rows, Coriolis 1e-4 + 2e-11 y on H rows
    y_t = (np.arange(b.ny) - 2 + 0.5) * 1000.0
    assert np.allclose(b.taux[:, 5], 0.1 * np.cos(math.pi * y_t / (60 * 1000.0)))
    y_h = (np.arange(b.ny) - 1.0) * 1000.0
    assert np.allclose(b.rlh_s[:, 5], 1e-4 + 2e-11 * y_h)
    assert np.all(b.mu[wet] == 10.0) and np.all(b.mu[~wet] == 0.0)
    # bump of 0.5 m at the centre
    assert abs(b.ssh.max() - 0.5) < 0.01


def test_no_wet_cell_without_a_velocity_point():
    b = make_basin(120, 80, seed=11, sigma=3.0, hland=70.0)
    wet = b.lbasins == 0
    p = np.pad(wet, 1)
    has_x = p[1:-1, :-2] | p[1:-1, 2:]
    has_y = p[:-2, 1:-1] | p[2:, 1:-1]
    assert np.all(~wet | (has_x & has_y))


def test_single_basin_workloads():
    b = single_basin("cfg0")
    assert (b.nx, b.ny) == (260, 260) and b.dt == 10.0 and b.g == 9.81


def test_coarsen_basin_shape():
    b = make_basin(64, 48, seed=2)
    c = coarsen_basin(b, 4)
    assert (c.nx - 4, c.ny - 4) == (16, 12)
    assert float(c.dx[3, 3]) == 4000.0 and c.dt == 40.0
