"""Argument sets of the GrADS descriptor writer covering its branches:
linear and level axes (1, 7, 8, 15 and 20 values), time steps of every
unit, E and F ranges of the G editing, zero, negative and large values,
both calendars.  Shared by the live comparison with the reference's
ctl_file_write and the script that stores its outputs as fixtures."""
import numpy as np

CASES = []


def _case(name, undef, nx, ny, nz, nt, xtype, ytype, ztype, yr_type, date, ht, hx, hy, hz, x, y, z, title, var):
    CASES.append(dict(name=name, undef=undef, nx=nx, ny=ny, nz=nz, nt=nt, xtype=xtype, ytype=ytype, ztype=ztype,
                      yr_type=yr_type, date=date, ht=ht, hx=hx, hy=hy, hz=hz, x=list(x), y=list(y), z=list(z),
                      title=title, var=var))


_case("ssh", -1.0e32, 285, 159, 1, 5, 0, 0, 0, 1, (1979, 1, 1, 0, 0), 86400.0, 0.05, 0.04, 1.0,
      [27.525], [40.94], [0.0], "SSH, m", "ssh")
_case("uu_levels", -1.0e32, 286, 159, 20, 12, 0, 0, 1, 0, (2001, 12, 31, 23, 59), 3600.0, 0.05, 0.04, 1.0,
      [27.5], [40.94], [1000.0 * v for v in np.linspace(0.0025, 0.9975, 20)], "zonal velocity, m/s", "u")
_case("levels7", 0.03, 7, 8, 15, 1, 1, 1, 1, 1, (12, 6, 15, 6, 30), 600.0, 0.0, 0.0, 0.0,
      [0.5, -0.25, 1.0e-5, 12345678.9, 0.0, 99999999.7, -3.0e12],
      [0.1, 0.099999999, 1.0, 9.99999999, 10.0, 123.456, -0.0001, 2.5e-7],
      [float(k) * 7.25 for k in range(15)], "levels of every kind", "lev")
_case("month", 1.0e20, 100, 50, 1, 3, 0, 0, 0, 0, (1980, 2, 29, 12, 0), 30.0 * 86400.0, -180.0, -90.0, 1.0,
      [-180.0], [-90.0], [0.0], "monthly means", "tmean")
_case("year", -999.0, 10, 10, 1, 2, 0, 0, 0, 1, (1900, 7, 4, 1, 2), 360.0 * 86400.0, 1.0e-3, 123456.0, 1.0,
      [0.0], [0.0], [0.0], "", "y")
_case("hours12", -1.0e32, 2048, 1536, 1, 100, 0, 0, 0, 1, (2026, 12, 2, 0, 0), 43200.0, 0.009, 0.009, 1.0,
      [26.643750], [40.145500], [0.0], "barotropic state, 2048x1536", "ubrtr")
_case("minutes", -1.0e32, 4, 3, 1, 1, 0, 0, 0, 1, (5, 1, 1, 0, 0), 119.0, 1.5, 2.5, 1.0, [1.5], [2.5], [0.0],
      "two-minute step", "x")


def write_case(path, c, fname):
    """The harness's case file (oracle/ctl_harness.f90)."""
    with open(path, "w") as f:
        f.write(fname + "\n" + (c["title"] or " ") + "\n" + c["var"] + "\n")
        y, mo, d, h, mi = c["date"]
        f.write(" ".join(repr(float(v)) if isinstance(v, float) else str(v) for v in
                         (c["undef"], c["nx"], c["ny"], c["nz"], c["nt"], c["xtype"], c["ytype"], c["ztype"],
                          c["yr_type"], y, mo, d, h, mi, c["ht"], c["hx"], c["hy"], c["hz"])) + "\n")
        for k in ("x", "y", "z"):
            f.write(" ".join(repr(float(v)) for v in c[k]) + "\n")
