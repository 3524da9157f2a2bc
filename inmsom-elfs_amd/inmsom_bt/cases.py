"""Synthetic basins for the barotropic step.

Every field is a float64 numpy array of shape (ny, nx) in C order, which
is the memory layout of the reference's column-major (m, n) arrays
(``Modules/mod_basin_grid.f90``): element ``a[n-1, m-1]`` is the
reference's ``a(m, n)``.  ``nx``/``ny`` include the two halo cells per
side (``Control/init_arrays.f90:53-69`` with one process), so the
computed interior is m = 3..nx-2, n = 3..ny-2 (1-based).

The generator follows the configuration text of the shallow-water
workloads: seeded white noise smoothed with a Gaussian of sigma = 8
cells, mapped affinely to 0..200 m, land where the depth is below 10 m,
a closed land ring in the two halo cells, a beta-plane Coriolis
parameter on H points, a zonal wind stress tau_x = 0.1 cos(pi y / L_y)
on T points, constant lateral viscosity on wet cells and a Gaussian
surface bump at rest.
"""
from __future__ import annotations

import dataclasses
import math
import struct
from typing import Optional, Sequence

import numpy as np

FIELDS_STATIC = ("hhq_rest", "dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb",
                 "dyb", "rlh_s", "taux", "tauy", "wflux", "mu", "mu4")
FIELDS_STATE = ("ssh", "sshp", "u", "up", "v", "vp")


@dataclasses.dataclass
class Basin:
    """Inputs of one barotropic run on one basin."""

    nx: int
    ny: int
    lbasins: np.ndarray          # int32, 0 = ocean, 1 = land
    hhq_rest: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    dxt: np.ndarray
    dyt: np.ndarray
    dxh: np.ndarray
    dyh: np.ndarray
    dxb: np.ndarray
    dyb: np.ndarray
    rlh_s: np.ndarray
    taux: np.ndarray
    tauy: np.ndarray
    wflux: np.ndarray
    mu: np.ndarray
    mu4: np.ndarray
    ssh: np.ndarray
    sshp: np.ndarray
    u: np.ndarray
    up: np.ndarray
    v: np.ndarray
    vp: np.ndarray
    g: float = 9.81
    rho0: float = 1025.0
    time_smooth: float = 0.05
    cb_nl: float = 2.5e-3
    ebottom: float = 25.0e-4
    r_fric: float = 1.0e-4
    dt: float = 10.0
    seed: int = 0

    @property
    def wet(self) -> np.ndarray:
        return self.lbasins == 0

    @property
    def n_wet(self) -> int:
        return int(np.count_nonzero(self.lbasins == 0))

    def state(self) -> dict:
        return {k: getattr(self, k) for k in FIELDS_STATE}

    def copy(self) -> "Basin":
        kw = {}
        for f in dataclasses.fields(self):
            val = getattr(self, f.name)
            kw[f.name] = val.copy() if isinstance(val, np.ndarray) else val
        return Basin(**kw)


def _gaussian_smooth(a: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian smoothing, reflect boundary, truncated at 4 sigma."""
    r = int(4.0 * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    out = a
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="symmetric")
        acc = np.zeros_like(out)
        n = out.shape[axis]
        for k in range(2 * r + 1):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(k, k + n)
            acc += w[k] * p[tuple(sl)]
        out = acc
    return out


def _clean_mask(wet: np.ndarray) -> np.ndarray:
    """Drop wet cells without a wet neighbour along x or along y.

    Such cells have no velocity point on one axis; the reference's
    Rayleigh coefficient divides by the count of those points
    (Control/init_pars.f90:464-467), so they are removed from synthetic
    basins.  Iterates until stable.
    """
    wet = wet.copy()
    while True:
        p = np.pad(wet, 1, constant_values=False)
        has_x = p[1:-1, :-2] | p[1:-1, 2:]
        has_y = p[:-2, 1:-1] | p[2:, 1:-1]
        keep = wet & has_x & has_y
        if np.array_equal(keep, wet):
            return keep
        wet = keep


def bathymetry(nxi: int, nyi: int, seed: int, sigma: float = 8.0,
               hmax: float = 200.0, hland: float = 10.0):
    """Interior depth and wet mask (shape (nyi, nxi))."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nyi, nxi))
    s = _gaussian_smooth(noise, sigma)
    lo, hi = float(s.min()), float(s.max())
    depth = (s - lo) / (hi - lo) * hmax
    wet = _clean_mask(depth >= hland)
    return np.where(wet, depth, 0.0), wet


def uniform_metrics(nx: int, ny: int, d: float = 1000.0) -> dict:
    one = np.full((ny, nx), d, dtype=np.float64)
    return {k: one.copy() for k in ("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")}


def stretched_metrics(nx: int, ny: int, d: float = 1000.0) -> dict:
    """Eight distinct smooth metric fields, so that a metric read at the
    wrong staggered point or transposed shows up in parity tests."""
    n, m = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64),
                       indexing="ij")
    out = {}
    for k, name in enumerate(("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")):
        ph = 0.7 * k
        out[name] = d * (1.0 + 0.15 * np.sin(2 * math.pi * m / nx + ph)
                         + 0.1 * np.cos(2 * math.pi * n / ny + 1.3 * ph)
                         + 0.02 * k)
    return out


def make_basin(nxi: int, nyi: int, seed: int = 0, *, d: float = 1000.0,
               sigma: float = 8.0, hmax: float = 200.0, hland: float = 10.0,
               f0: float = 1.0e-4, beta: float = 2.0e-11, tau0: float = 0.1,
               nu: float = 10.0, bumps: Optional[Sequence] = None,
               metrics: str = "uniform", g: float = 9.81, dt: float = 10.0,
               lvisc_4: float = 0.0) -> Basin:
    """Synthetic closed basin with an nxi x nyi interior.

    ``bumps`` is a sequence of (amplitude [m], x-centre [cells], y-centre
    [cells], radius [m]); the default is one 0.5 m bump of radius 20 km
    at the centre of the interior.
    """
    nx, ny = nxi + 4, nyi + 4
    h_int, wet_int = bathymetry(nxi, nyi, seed, sigma, hmax, hland)
    wet = np.zeros((ny, nx), dtype=bool)
    wet[2:-2, 2:-2] = wet_int
    lb = np.where(wet, 0, 1).astype(np.int32)
    hhq_rest = np.zeros((ny, nx))
    hhq_rest[2:-2, 2:-2] = h_int
    met = uniform_metrics(nx, ny, d) if metrics == "uniform" else stretched_metrics(nx, ny, d)
    lu = wet.astype(np.float64)

    j = np.arange(ny, dtype=np.float64)
    y_t = (j - 2 + 0.5) * d            # T-row centres from the interior's south edge
    y_h = (j - 2 + 1.0) * d            # H/V rows sit on the north faces
    ly = nyi * d
    rlh_s = np.repeat((f0 + beta * y_h)[:, None], nx, axis=1)
    taux = np.repeat((tau0 * np.cos(math.pi * y_t / ly))[:, None], nx, axis=1)
    tauy = np.zeros((ny, nx))
    wflux = np.zeros((ny, nx))
    mu = nu * lu
    # Inc/latdiff.fi:12 with a depth-mean over lu (init_pars.f90:167)
    mu4 = (math.sqrt(lvisc_4 / dt) * met["dx"] * met["dy"] * lu) if lvisc_4 > 0 else np.zeros((ny, nx))

    if bumps is None:
        bumps = [(0.5, 0.5 * nxi, 0.5 * nyi, 20.0e3)]
    mm, nn = np.meshgrid(np.arange(nx, dtype=np.float64) - 2 + 0.5,
                         np.arange(ny, dtype=np.float64) - 2 + 0.5)
    ssh = np.zeros((ny, nx))
    for amp, xc, yc, rad in bumps:
        r2 = ((mm - xc) * d) ** 2 + ((nn - yc) * d) ** 2
        ssh += amp * np.exp(-r2 / (rad * rad))
    ssh *= lu
    zero = np.zeros((ny, nx))
    return Basin(nx=nx, ny=ny, lbasins=lb, hhq_rest=hhq_rest, rlh_s=rlh_s, taux=taux,
                 tauy=tauy, wflux=wflux, mu=mu, mu4=mu4, ssh=ssh, sshp=ssh.copy(),
                 u=zero.copy(), up=zero.copy(), v=zero.copy(), vp=zero.copy(), g=g,
                 dt=dt, seed=seed, **met)


def randomize_state(b: Basin, seed: int, amp_ssh: float = 0.3, amp_uv: float = 0.2,
                    wflux_amp: float = 0.0) -> Basin:
    """Fill the state with smooth random fields (masked like the model's
    state: eta on T, u on both-wet U faces, v on both-wet V faces)."""
    rng = np.random.default_rng(seed)
    out = b.copy()
    wet = b.wet
    lcu = np.zeros_like(wet)
    lcv = np.zeros_like(wet)
    lcu[1:-2, 1:-2] = wet[1:-2, 1:-2] & wet[1:-2, 2:-1]
    lcv[1:-2, 1:-2] = wet[1:-2, 1:-2] & wet[2:-1, 1:-2]

    def field(amp):
        return amp * _gaussian_smooth(rng.standard_normal((b.ny, b.nx)), 3.0) * 3.0

    out.ssh = field(amp_ssh) * wet
    out.sshp = out.ssh + field(0.1 * amp_ssh) * wet
    out.u = field(amp_uv) * lcu
    out.up = out.u + field(0.1 * amp_uv) * lcu
    out.v = field(amp_uv) * lcv
    out.vp = out.v + field(0.1 * amp_uv) * lcv
    if wflux_amp:
        out.wflux = field(wflux_amp) * wet
    return out


def coarsen_basin(b: Basin, factor: int = 4) -> Basin:
    """Block-average a basin's interior by ``factor`` (a coarse model of
    the same ocean): depth and state are averaged over wet fine cells; a
    coarse cell is wet when more than half of its fine cells are."""
    nxi, nyi = b.nx - 4, b.ny - 4
    assert nxi % factor == 0 and nyi % factor == 0
    cx, cy = nxi // factor, nyi // factor
    wet_f = b.wet[2:-2, 2:-2].astype(np.float64)

    def blk(a):
        return a.reshape(cy, factor, cx, factor).sum(axis=(1, 3))

    frac = blk(wet_f) / (factor * factor)
    wet_c = _clean_mask(frac > 0.5)
    nwet = np.maximum(blk(wet_f), 1.0)
    h_c = np.where(wet_c, blk(b.hhq_rest[2:-2, 2:-2] * wet_f) / nwet, 0.0)
    h_c = np.where(wet_c, np.maximum(h_c, 1.0), 0.0)
    nx, ny = cx + 4, cy + 4
    wet = np.zeros((ny, nx), dtype=bool)
    wet[2:-2, 2:-2] = wet_c
    d = float(b.dx[2, 2]) * factor
    met = uniform_metrics(nx, ny, d)
    hq = np.zeros((ny, nx))
    hq[2:-2, 2:-2] = h_c

    def samp_rows(a):
        # rows of T (or H) points: average the fine rows of each block
        inner = a[2:-2, 2].reshape(cy, factor).mean(axis=1)
        out = np.zeros(ny)
        out[2:-2] = inner
        out[:2] = inner[0]
        out[-2:] = inner[-1]
        return np.repeat(out[:, None], nx, axis=1)

    rlh = np.zeros((ny, nx))
    # H points of the coarse grid sit on every factor-th fine H row
    rows_h = b.rlh_s[2:-2, 2][factor - 1::factor]
    rlh[2:-2] = rows_h[:, None]
    rlh[:2] = rows_h[0]
    rlh[-2:] = rows_h[-1]
    lu = wet.astype(np.float64)
    zero = np.zeros((ny, nx))
    c = Basin(nx=nx, ny=ny, lbasins=np.where(wet, 0, 1).astype(np.int32), hhq_rest=hq,
              rlh_s=rlh, taux=samp_rows(b.taux), tauy=zero.copy(), wflux=zero.copy(),
              mu=lu * float(np.max(b.mu)), mu4=zero.copy(), ssh=zero.copy(), sshp=zero.copy(),
              u=zero.copy(), up=zero.copy(), v=zero.copy(), vp=zero.copy(), g=b.g,
              rho0=b.rho0, time_smooth=b.time_smooth, cb_nl=b.cb_nl, ebottom=b.ebottom,
              r_fric=b.r_fric, dt=b.dt * factor, seed=b.seed, **met)
    return c


# ---------------------------------------------------------------- case files
CASE_MAGIC = 0x31435442  # 'BTC1'


def write_case(path: str, b: Basin, rdis: np.ndarray, nsteps: int, nstep: int = 1,
               ksw4: int = 0, mode: int = 0, out_steps: Sequence[int] = ()) -> None:
    """Binary case file read by the reference harness (layout documented
    at the top of oracle/ref_harness.f90)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<9i", CASE_MAGIC, 1, b.nx, b.ny, nsteps, nstep, ksw4, mode,
                            len(out_steps)))
        if out_steps:
            f.write(np.asarray(out_steps, dtype="<i4").tobytes())
        f.write(struct.pack("<3d", b.dt, b.cb_nl, b.ebottom))
        f.write(np.ascontiguousarray(b.lbasins, dtype="<i4").tobytes())
        for name in FIELDS_STATIC:
            f.write(np.ascontiguousarray(getattr(b, name), dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(rdis, dtype="<f8").tobytes())
        for name in FIELDS_STATE:
            f.write(np.ascontiguousarray(getattr(b, name), dtype="<f8").tobytes())


# The reference's shipped configuration (Inc/basinpar.fi:4-19): 289 x 163
# points, first computed T point at 27.525 E, 40.94 N, steps 0.05 x 0.04
# degrees, Earth radius 6371 km.
REF_NX, REF_NY = 289, 163
REF_LON0, REF_LAT0, REF_DLON, REF_DLAT = 27.525, 40.94, 0.05, 0.04
EARTH_RADIUS = 6371.0e3
EARTH_OMEGA = 7.292115974546454e-5


def read_mask(path: str, nx: int, ny: int) -> np.ndarray:
    """lbasins from a DIOGIN text mask as gridcon reads it
    (Service/grid_construction.f90:16-28): a comment line, then ny rows of
    nx digits from the northern row (n = ny) down; 0 = ocean.  Returns
    (ny, nx) int32 with row 0 = n = 1."""
    with open(path, "r", newline="") as f:
        lines = f.read().splitlines()
    rows = [ln.rstrip("\r") for ln in lines[1:1 + ny]]
    if len(rows) != ny or any(len(r) < nx for r in rows):
        raise ValueError(f"{path}: expected {ny} rows of {nx} digits")
    lb = np.array([[int(ch) for ch in r[:nx]] for r in rows], dtype=np.int32)
    return np.ascontiguousarray(lb[::-1])


def reference_grid(nx: int = None, ny: int = None, rlon: float = None, rlat: float = None, dxst: float = None,
                   dyst: float = None, curve_grid: int = 1, radius: float = None, omega: float = None) -> dict:
    """The reference's grid set-up basinpar (Service/basin_parameters.f90:
    191-300) for a regular grid (xgr_type = ygr_type = 0) with the Cartesian
    (grid_parameters_carthesian, :764-833) or the unrotated spherical
    (grid_parameters_spherical, :836-949) metric: the eight metrics and the
    Coriolis parameter rlh_s, in its operation order.  Defaults: the shipped
    Inc/basinpar.fi (289 x 163 points, 27.525 E, 40.94 N, 0.05 x 0.04
    degrees, spherical).

    The metrics are set on points 2..nx-1 x 2..ny-1 (zero outside), rlh_s
    is 2 Omega everywhere and, on H points of that range, 2 Omega times the
    sine of the latitude (spherical) or 2 Omega / sqrt(2) (Cartesian).  The
    degree-to-radian factor is the reference's 3.1415926535897/180.0d0, whose
    literal is a default (single precision) real.  cos and sin are the C
    library's (math.cos / math.sin), as the compiled reference calls them."""
    nx = REF_NX if nx is None else nx
    ny = REF_NY if ny is None else ny
    rlon = REF_LON0 if rlon is None else rlon
    rlat = REF_LAT0 if rlat is None else rlat
    dxst = REF_DLON if dxst is None else dxst
    dyst = REF_DLAT if dyst is None else dyst
    radius = EARTH_RADIUS if radius is None else radius
    omega = EARTH_OMEGA if omega is None else omega
    dpip180 = float(np.float32(3.1415926535897)) / 180.0
    mmm = nnn = 3
    xt = [rlon + float(m - mmm) * dxst for m in range(1, nx + 1)]   # xt(m) at [m - 1]
    yt = [rlat + float(n - nnn) * dyst for n in range(1, ny + 1)]
    xu = [(xt[i] + xt[i + 1]) / 2.0 for i in range(nx - 1)]          # xu(m), m = 1..nx-1
    yv = [(yt[i] + yt[i + 1]) / 2.0 for i in range(ny - 1)]
    cx = dxst * dpip180 * radius
    cy = dyst * dpip180 * radius
    out = {k: np.zeros((ny, nx)) for k in ("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")}
    rlh_s = np.full((ny, nx), 2.0 * omega)
    lat_extr = 89.99999
    sinlat_extr = math.sin(dpip180 * lat_extr)
    for n in range(2, ny):          # 1-based n = 2..ny-1 (ny_start-1..ny_end+1)
        j = n - 1
        for k in ("dx", "dxt", "dxh", "dxb"):
            out[k][j, 1:nx - 1] = cx
        for k in ("dy", "dyt", "dyh", "dyb"):
            out[k][j, 1:nx - 1] = cy
        if curve_grid == 1:
            # T and U points on the latitude yt(n), V and H points on yv(n)
            for k, y in (("dx", yt[j]), ("dxt", yt[j]), ("dxh", yv[j]), ("dxb", yv[j])):
                lat_mod = max(min(y, lat_extr), -lat_extr)
                out[k][j, 1:nx - 1] = cx * math.cos(dpip180 * lat_mod)
            for m in range(2, nx):
                xm = xu[m - 1]
                sin_lat = (math.sin(dpip180 * yv[j]) * math.cos(dpip180 * 0.0) +
                           math.cos(dpip180 * xm) * math.cos(dpip180 * yv[j]) * math.sin(dpip180 * 0.0))
                sin_lat = min(max(sin_lat, -sinlat_extr), sinlat_extr)
                rlh_s[j, m - 1] = (2.0 * omega) * sin_lat
        elif curve_grid == 0:
            rlh_s[j, 1:nx - 1] = (2.0 * omega) / math.sqrt(2.0)
        else:
            raise ValueError("curve_grid must be 0 (Cartesian) or 1 (spherical)")
    out["rlh_s"] = rlh_s
    return out


def reference_basin(mask_path: str, topo_dir: str, topo_name: str = "topo.dat", *, metrics: str = "spherical",
                    d: float = 4000.0, nu: float = 100.0, tau0: float = 0.05, g: float = 9.8,
                    dt: float = 16.0) -> Basin:
    """The reference's shipped basin: its land mask (mask.txt) and its
    bathymetry on the model grid, read as init_pars does
    (Control/init_pars.f90:151-152: prdstd of the box mmm..mm x nnn..nn
    with lu, land 0), on spherical metrics or uniform ones of step d."""
    from .fileio import rdstd

    nx, ny = REF_NX, REF_NY
    lb = read_mask(mask_path, nx, ny)
    lu = (lb == 0).astype(np.float32)
    topo = np.zeros((ny, nx), dtype=np.float32)
    rdstd(topo_dir, topo_name, 1, topo, lu, (3, nx - 2, 3, ny - 2, 1, 1))
    hhq_rest = topo.astype(np.float64)
    if metrics == "spherical":
        # the reference's own construction (basinpar), metrics and Coriolis
        met = reference_grid(nx, ny)
        rlh_s = met.pop("rlh_s")
    else:
        met = uniform_metrics(nx, ny, d)
        lat_v = REF_LAT0 + (np.arange(ny, dtype=np.float64) - 2.0 + 0.5) * REF_DLAT
        rlh_s = np.repeat((2.0 * EARTH_OMEGA * np.sin(lat_v * math.pi / 180.0))[:, None], nx, axis=1)
    lat_t = REF_LAT0 + (np.arange(ny, dtype=np.float64) - 2.0) * REF_DLAT
    y = (lat_t - lat_t[2]) / (lat_t[-3] - lat_t[2])
    taux = np.repeat((tau0 * np.cos(math.pi * y))[:, None], nx, axis=1)
    zero = np.zeros((ny, nx))
    mu = nu * lu.astype(np.float64)
    return Basin(nx=nx, ny=ny, lbasins=lb, hhq_rest=hhq_rest, rlh_s=rlh_s, taux=taux, tauy=zero.copy(),
                 wflux=zero.copy(), mu=mu, mu4=zero.copy(), ssh=zero.copy(), sshp=zero.copy(), u=zero.copy(),
                 up=zero.copy(), v=zero.copy(), vp=zero.copy(), g=g, dt=dt, seed=0, **met)
