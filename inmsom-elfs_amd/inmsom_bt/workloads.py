"""The shallow-water workloads the barotropic step is measured on.

single_basin(...) is the plain leap-frog configuration: a seeded
Gaussian-smoothed bathymetry (sigma 8 cells) mapped to 0..200 m with land
below 10 m, dx = dy = 1 km, f = 1e-4 + 2e-11 y, g = 9.81,
tau_x = 0.1 cos(pi y / L_y), nu = 10 m^2/s, quadratic drag 2.5e-3,
Robert-Asselin 0.05, a 0.5 m Gaussian bump of radius 20 km at the centre.

Outer step dt = 10 s.  The barotropic mode is sub-cycled with
nstep_barotrop = 2 (leap-frog substep 5 s): a 10 s leap-frog step violates
the C-grid leap-frog limit c dt / dx <= 1/(2 sqrt 2) for depths above
~128 m at 1 km, and the reference scheme (run through the bit-exact
oracle) diverges within ~40 steps at nstep_barotrop = 1.
"""
from __future__ import annotations

import numpy as np

from .cases import Basin, make_basin

DT = 10.0
NSTEP_BAROTROP = 2

SINGLE = {
    "cfg0": dict(nxi=256, nyi=256, steps=1000),
    "cfg1": dict(nxi=2048, nyi=1536, steps=10000),
}


def ensemble(nmembers: int = 64, seed: int = 7, size_lo=(256, 256), size_hi=(1024, 768),
             sigma: float = 8.0, d: float = 1000.0, dt: float = DT) -> list:
    """Independent basins of uneven size: interior sizes uniform in
    [size_lo, size_hi] (both axes), each with its own seeded bathymetry,
    wind amplitude 0.1 N/m^2 x U(0.8, 1.2) and 1-3 surface bumps of
    0.2-0.5 m (radius 20 km) at random interior positions; grid step d,
    outer step dt."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nmembers):
        nxi = int(rng.integers(size_lo[0], size_hi[0] + 1))
        nyi = int(rng.integers(size_lo[1], size_hi[1] + 1))
        tau0 = 0.1 * float(rng.uniform(0.8, 1.2))
        nb = int(rng.integers(1, 4))
        bumps = [(float(rng.uniform(0.2, 0.5)), float(rng.uniform(0.2, 0.8)) * nxi,
                  float(rng.uniform(0.2, 0.8)) * nyi, 20.0e3) for _ in range(nb)]
        out.append(make_basin(nxi, nyi, seed=int(rng.integers(1 << 30)), d=d, sigma=sigma, hmax=200.0,
                              hland=10.0, f0=1.0e-4, beta=2.0e-11, tau0=tau0, nu=10.0, g=9.81,
                              dt=dt, bumps=bumps))
    return out


# 4 km members for the learned closure's ensemble: the 1 km generator's
# 8 km bathymetry smoothing is 2 cells here, outer step 40 s.
COARSE_FACTOR = 4
COARSE_DT = DT * COARSE_FACTOR


def coarse_ensemble(nmembers: int = 128, seed: int = 11, size_lo=(128, 128), size_hi=(512, 384)) -> list:
    return ensemble(nmembers, seed=seed, size_lo=size_lo, size_hi=size_hi, sigma=8.0 / COARSE_FACTOR,
                    d=1000.0 * COARSE_FACTOR, dt=COARSE_DT)


def single_basin(name: str = "cfg1", seed: int = 2024) -> Basin:
    spec = SINGLE[name]
    return make_basin(spec["nxi"], spec["nyi"], seed=seed, sigma=8.0, hmax=200.0, hland=10.0,
                      f0=1.0e-4, beta=2.0e-11, tau0=0.1, nu=10.0, g=9.81, dt=DT)
