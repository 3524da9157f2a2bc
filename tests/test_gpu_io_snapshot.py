"""The device snapshot (inmsom_bt_write_snapshot) against the host writers
of the reference's formats: the records of ssh/ubrtr/vbrtr and their
GrADS descriptors are byte for byte what wdstd and ctl_file_write write
from the same state (Control/output.f90:115-260)."""
import os

import numpy as np
import pytest

from inmsom_bt import BarotropicModel, fileio, make_basin, randomize_state
from inmsom_bt.model import gridcon

pytestmark = pytest.mark.gpu


def _host_snapshot(b, st, path, nrec, axes, date, tstep):
    """What the reference's local_output would write for this state."""
    masks = gridcon(b.lbasins)
    nx, ny = b.nx, b.ny
    xt0, dxst, yt0, dyst = axes
    specs = (("ssh.dat", st["ssh"], masks["lu"], (3, nx - 2, 3, ny - 2), xt0, yt0, "SSH, m", "ssh"),
             ("ubrtr.dat", st["u"], masks["llu"], (2, nx - 2, 3, ny - 2), xt0 - 0.5 * dxst, yt0,
              "zonal barotropic transport, m2/s", "ubrtr"),
             ("vbrtr.dat", st["v"], masks["llv"], (3, nx - 2, 2, ny - 2), xt0, yt0 - 0.5 * dyst,
              "meridional barotropic transport, m2/s", "vbrtr"))
    for name, f, lu, box, x0, y0, title, var in specs:
        a = np.ascontiguousarray(f, dtype=np.float32)      # sngl()
        fileio.wdstd(path, name, nrec, a, lu, box + (1, 1))
        fileio.ctl_file_write(os.path.join(path, name), fileio.UNDEF, box[1] - box[0] + 1, box[3] - box[2] + 1, 1,
                              nrec, 0, [x0], dxst, 0, [y0], dyst, 0, [0.0], 1.0, 1, *date, tstep, title, var)


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_snapshot_matches_host_writers(mode, tmp_path):
    print(f"snapshot {mode}", flush=True)
    b = randomize_state(make_basin(90, 64, seed=12, sigma=5.0), 6)
    axes, date, tstep = (27.525, 0.05, 40.94, 0.04), (1979, 1, 1, 0, 0), 86400.0
    dev, host = tmp_path / "dev", tmp_path / "host"
    dev.mkdir()
    host.mkdir()
    with BarotropicModel(b) as m:
        for nrec in (1, 3):                                  # record 2 stays a hole
            m.run(5, mode, nstep=2)
            ms = m.write_snapshot(str(dev), nrec, *axes, date=date, tstep=tstep)
            assert ms > 0.0
            _host_snapshot(b, m.get_state(), str(host), nrec, axes, date, tstep)
    for name in ("ssh", "ubrtr", "vbrtr"):
        for ext in (".dat", ".ctl"):
            got = (dev / (name + ext)).read_bytes()
            want = (host / (name + ext)).read_bytes()
            assert got == want, f"{name}{ext} differs"
    rec = np.fromfile(dev / "ssh.dat", dtype="<f4").reshape(3, b.ny - 4, b.nx - 4)
    assert np.all(rec[1] == 0.0)                             # the hole
    assert np.any(rec[2] == fileio.UNDEF) and np.any(rec[2] != fileio.UNDEF)


def test_snapshot_errors(tmp_path):
    from inmsom_bt._capi import BtError
    b = make_basin(20, 16, seed=1)
    with BarotropicModel(b) as m:
        with pytest.raises(BtError):
            m.write_snapshot(str(tmp_path / "missing_dir"), 1)
        with pytest.raises(BtError):
            m.write_snapshot(str(tmp_path), 0)
