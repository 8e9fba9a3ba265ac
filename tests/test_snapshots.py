"""Snapshot keys and file format (snapshot(), semsyn/src/sem2d_psvsh.c:
149-150,202-204,483-510,1699-1715): plot_snap=0, startsnap=100 and
incresnap=50 by default; a file holds "ix iz u[0] u[1]" (%d %d %e %e) for
every second node column (outer) and row (inner)."""
import os

import numpy as np

import semsks
from semsks import inputs


def test_snapshot_keys_default_and_override(tmp_path, monkeypatch):
    inputs.small_case().write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    assert semsks.Model(["par=run.par"]).snapshots() == (0, 100, 50)
    assert semsks.Model(["par=run.par", "plot_snap=1", "startsnap=7", "incresnap=3"]).snapshots() == (1, 7, 3)


def test_snapshot_file_format(tmp_path):
    nx, nz = 9, 6
    rng = np.random.default_rng(5)
    u = (rng.standard_normal((3, nz, nx)) * 10.0 ** rng.integers(-8, 3, (3, nz, nx))).astype(np.float32)
    semsks._check(semsks.lib().semsks_write_snapshot(os.fspath(tmp_path).encode(), b"out3", nx, nz,
                                                     u.ctypes.data))
    lines = open(os.path.join(os.fspath(tmp_path), "out3")).read().splitlines()
    want = [(ix, iz) for ix in range(0, nx, 2) for iz in range(0, nz, 2)]
    assert len(lines) == len(want)
    for ln, (ix, iz) in zip(lines, want):
        a, b, x, y = ln.split()
        assert (int(a), int(b)) == (ix, iz)
        assert "e" in x and "e" in y and len(x.split("e")[0].lstrip("-")) == 8     # %e: d.dddddd
        for got, ref in ((float(x), u[0, iz, ix]), (float(y), u[1, iz, ix])):
            assert abs(got - ref) <= 5.1e-7 * abs(ref) + 1e-45
