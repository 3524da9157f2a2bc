"""The reference's on-disk formats (Service/input_output_data.f90,
Service/rw_ctl_file.f90) through the C ABI: the GrADS descriptor writer
against the reference's own ctl_file_write (live where the harness is
built, committed fixtures otherwise), DIOGIN records against the format's
definition and the reference's shipped bathymetry file."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from ctl_cases import CASES, write_case  # noqa: E402
from inmsom_bt import fileio  # noqa: E402
from inmsom_bt._capi import BtError  # noqa: E402

ROOT = os.path.dirname(HERE)
CTL_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ctl_harness")
GOLD = os.path.join(HERE, "golden", "ctl")


def _ours(c, tmp):
    y, mo, d, h, mi = c["date"]
    fname = os.path.join(tmp, c["name"] + ".dat")
    fileio.ctl_file_write(fname, c["undef"], c["nx"], c["ny"], c["nz"], c["nt"], c["xtype"], c["x"], c["hx"],
                          c["ytype"], c["y"], c["hy"], c["ztype"], c["z"], c["hz"], c["yr_type"], y, mo, d, h, mi,
                          c["ht"], c["title"], c["var"])
    with open(os.path.join(tmp, c["name"] + ".ctl"), "rb") as f:
        return f.read()


@pytest.mark.skipif(not os.path.exists(CTL_HARNESS), reason="reference harness not built (make -C oracle ref)")
@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_ctl_write_matches_reference_live(c, tmp_path):
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir()
    our_dir.mkdir()
    case = str(tmp_path / "case.txt")
    write_case(case, c, str(ref_dir / (c["name"] + ".dat")))
    subprocess.run([CTL_HARNESS, case], check=True, cwd=str(ref_dir), capture_output=True, timeout=60)
    want = (ref_dir / (c["name"] + ".ctl")).read_bytes()
    got = _ours(c, str(our_dir))
    assert got == want, f"{c['name']}:\nours\n{got.decode()}\nreference\n{want.decode()}"


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_ctl_write_matches_reference_fixture(c, tmp_path):
    """The reference writer's outputs stored by tests/golden/make_ctl_golden.py."""
    path = os.path.join(GOLD, c["name"] + ".ctl")
    with open(path, "rb") as f:
        want = f.read()
    assert _ours(c, str(tmp_path)) == want


def test_ctl_write_rejects_short_time_step(tmp_path):
    with pytest.raises(BtError):
        fileio.ctl_file_write(str(tmp_path / "a.dat"), -1e32, 2, 2, 1, 1, 0, [0.0], 1.0, 0, [0.0], 1.0, 0, [0.0],
                              1.0, 1, 2000, 1, 1, 0, 0, 30.0, "t", "a")


def test_diogin_records_layout(tmp_path):
    """Record r of field nfield, level kr = (nfield-1)*(nze-nzb+1)+kr, first
    index fastest, land as undef on disk and 0 in the array afterwards
    (input_output_data.f90:206-327), read back with land 0."""
    rng = np.random.default_rng(4)
    nz, ny, nx = 3, 9, 11
    lu = (rng.random((ny, nx)) > 0.3).astype(np.float32)
    box = (2, 10, 3, 8, 1, 3)                     # nxb, nxe, nyb, nye, nzb, nze
    fields = [rng.standard_normal((nz, ny, nx)).astype(np.float32) for _ in range(2)]
    for nf, f in enumerate(fields, start=1):
        w = f.copy()
        fileio.wdstd(str(tmp_path), "a.dat", nf, w, lu, box)
        land = np.broadcast_to(lu < 0.5, w.shape)
        sub = np.zeros_like(land)
        sub[:, 2:8, 1:10] = True
        assert np.all(w[land & sub] == 0.0) and np.array_equal(w[~(land & sub)], f[~(land & sub)])
    raw = np.fromfile(tmp_path / "a.dat", dtype="<f4")
    rx, ry = 9, 6
    assert raw.size == 2 * nz * rx * ry
    for nf, f in enumerate(fields, start=1):
        for kr in range(1, nz + 1):
            r = (nf - 1) * nz + kr
            rec = raw[(r - 1) * rx * ry: r * rx * ry].reshape(ry, rx)
            want = f[kr - 1, 2:8, 1:10].copy()
            want[lu[2:8, 1:10] < 0.5] = fileio.UNDEF
            assert np.array_equal(rec, want)
    back = np.full((nz, ny, nx), 7.0, dtype=np.float32)
    fileio.rdstd(str(tmp_path) + "/", "/a.dat", 2, back, lu, box)
    want = fields[1].copy()
    want[np.broadcast_to(lu < 0.5, want.shape)] = 0.0
    assert np.array_equal(back[:, 2:8, 1:10], want[:, 2:8, 1:10])
    assert np.all(back[:, :2, :] == 7.0) and np.all(back[:, :, :1] == 7.0)


def test_diogin8_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((2, 5, 6))
    lu = np.ones((5, 6), dtype=np.float32)
    fileio.wdstd8(str(tmp_path), "b.dat", 3, a.copy(), lu)
    raw = np.fromfile(tmp_path / "b.dat", dtype="<f8")
    assert raw.size == 3 * 2 * 30 and np.array_equal(raw[4 * 30:], a.reshape(-1))
    assert np.all(raw[:4 * 30] == 0.0)            # records 1..4 never written: a hole of zeros
    b = np.zeros_like(a)
    fileio.rdstd8(str(tmp_path), "b.dat", 3, b, lu)
    assert np.array_equal(a, b)
    with pytest.raises(BtError):
        fileio.rdstd8(str(tmp_path), "b.dat", 4, b, lu)        # past the end of the file
    with pytest.raises(BtError):
        fileio.rdstd8(str(tmp_path), "missing.dat", 1, b, lu)
    with pytest.raises(BtError):
        fileio.wdstd8(str(tmp_path), "b.dat", 1, a.copy(), lu, (0, 6, 1, 5, 1, 2))   # nxb < 1


def test_reference_basin_from_shipped_files():
    """The shipped Black Sea files: topo.ctl describes the 285 x 159 box
    mmm..mm x nnn..nn of the 289 x 163 grid; the DIOGIN reader with the
    mask of mask.txt gives the file's depths on ocean points and 0 on land
    (Control/init_pars.f90:151-152, input_output_data.f90:164-180)."""
    from inmsom_bt.cases import REF_NX, REF_NY, reference_basin
    data = os.path.join(HERE, "golden", "blacksea")
    with open(os.path.join(data, "topo.ctl"), "rb") as f:
        ctl = f.read().decode().split()
    assert int(ctl[ctl.index("XDEF") + 1]) == REF_NX - 4 and int(ctl[ctl.index("YDEF") + 1]) == REF_NY - 4
    b = reference_basin(os.path.join(data, "mask.txt"), data)
    raw = np.fromfile(os.path.join(data, "topo.dat"), dtype="<f4").reshape(REF_NY - 4, REF_NX - 4)
    with open(os.path.join(data, "mask.txt")) as f:
        rows = f.read().splitlines()[1:1 + REF_NY]
    ocean = np.array([[ch == "0" for ch in r.rstrip()[:REF_NX]] for r in rows])[::-1]
    assert np.array_equal(b.wet, ocean)
    inner = b.hhq_rest[2:-2, 2:-2]
    assert np.array_equal(inner[ocean[2:-2, 2:-2]], raw[ocean[2:-2, 2:-2]].astype(np.float64))
    assert np.all(b.hhq_rest[~ocean] == 0.0) and raw.min() >= 3.0
    assert np.all(b.hhq_rest[ocean] >= 3.0)             # every ocean point lies inside the box
