"""The driver's output files (include/ehd_io.h) against golden text made from
the reference's own fprintf format strings (tests/golden/gen_writer_golden.py
parses them out of LBM.cu and main.cu), byte for byte; the restart reader
inverts data_end.dat; the host current follows LBM.cu:1646-1671.  CPU only."""
import json
import os

import numpy as np
import pytest

import ehd_amd as E

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "writers_golden.json")


@pytest.fixture(scope="module")
def golden():
    g = json.load(open(GOLDEN))
    nx, ny, nz = g["grid"]
    g["shape"] = (nz, ny, nx)
    g["f9"] = np.array([[float.fromhex(v) for v in row] for row in g["fields9_hex"]])
    g["t_val"] = float.fromhex(g["t"])
    return g


def _read(p):
    with open(p) as f:
        return f.read()


def test_tecplot_zones(golden, tmp_path):
    p = str(tmp_path / "data.dat")
    f9 = golden["f9"].copy()
    E.write_tecplot(p, golden["t_val"], f9, golden["shape"], golden["spacing"], first=True)
    assert _read(p) == golden["tecplot_first"]
    E.write_tecplot(p, golden["t_val"], f9, golden["shape"], golden["spacing"], append=True)
    assert _read(p) == golden["tecplot_first"] + golden["tecplot_next"]
    assert np.array_equal(f9, golden["f9"])     # the caller's arrays are untouched


def test_data_end_dmd_umax(golden, tmp_path):
    t, f9, shape = golden["t_val"], golden["f9"], golden["shape"]
    E.write_data_end(str(tmp_path / "end.dat"), t, f9, shape)
    assert _read(tmp_path / "end.dat") == golden["data_end"]
    E.write_dmd(str(tmp_path / "dmd.dat"), t, f9[5], shape)
    assert _read(tmp_path / "dmd.dat") == golden["dmd"]
    E.write_umax(str(tmp_path / "umax.dat"), t, f9[5], shape)
    E.write_umax(str(tmp_path / "umax.dat"), t, f9[5], shape, append=True)
    assert _read(tmp_path / "umax.dat") == golden["umax"] * 2
    assert "%g.dat" % t == golden["dmd_name"]


def test_restart_reader_inverts_data_end(golden, tmp_path):
    p = str(tmp_path / "data_end.dat")
    t, f9, shape = golden["t_val"], golden["f9"], golden["shape"]
    E.write_data_end(p, t, f9, shape)
    back, t2 = E.read_data_end(p, shape)
    assert t2 == float("%10.6f" % t)
    # interior values round-trip to the 6 decimals of %10.6f
    ref = np.round(f9.reshape(back.shape), 6)
    assert np.allclose(back[:, 1:-1], ref[:, 1:-1], rtol=0, atol=6e-7)
    # a truncated file is an error, not a silent partial read
    with open(p) as f:
        lines = f.readlines()
    with open(p, "w") as f:
        f.writelines(lines[:-1])
    with pytest.raises(E.EhdError):
        E.read_data_end(p, shape)


def test_host_current_closed_form():
    rng = np.random.default_rng(3)
    nz, ny, nx = 6, 4, 5
    q = rng.uniform(1, 2, (nz, ny, nx))
    ez = rng.uniform(-1, 1, (nz, ny, nx))
    K, dy = 2.5e-5, 0.2
    qt = 2 * q[nz - 2] - q[nz - 3]
    ref = 0.0
    for y in range(ny):
        for x in range(nx):
            ref += qt[y, x] * ez[nz - 1, y, x]
    assert E.host_current(q, ez, K, dy) == ref * K * dy * dy
