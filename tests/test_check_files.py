"""check=1 setup diagnostics (grid.xz, legendre.xz, vel.xy), written by
semsks_write_check before the solve as the reference's main() and discret()
write them (semsyn/src/sem2d_psvsh.c:47,179,247-270,804-825; its plotting
scripts read vel.xy, applications/test/plot_sks.pl:239,273).

Pinned without restating the writer: the GLL nodes against their textbook
values, the Lagrange basis against numpy's polynomial evaluation in double,
partition of unity and the Kronecker property at x = -1 (the reference's
float loop variable stops at 0.95, 40 samples), and densities and
velocities against the layer each node's depth falls in.
"""
import os

import numpy as np
import pytest

import semsks
from semsks import inputs

F32 = np.float32
DEG = 6371 * 3.14159265 / 180.0


def model_in(tmp_path, monkeypatch, hpn=4, **extra):
    c = inputs.small_case(hpn=hpn, extra=dict(check=1, xmin="250.0", **extra))
    d = os.fspath(tmp_path)
    D = c.write(d)
    monkeypatch.chdir(d)
    return semsks.Model(["par=run.par"]), D, c


def layer_tops(c):
    """depths (km) of the layer interfaces below the surface, top to bottom"""
    return np.cumsum([0.0] + list(c.layers_thk[:-1]))


def layer_at(c, depth):
    """index of the layer holding this depth, or -1 within 1e-3 km of an interface"""
    tops = layer_tops(c)
    if np.any(np.abs(tops[1:] - depth) < 1e-3):
        return -1
    return int(np.searchsorted(tops, depth, side="right") - 1)


def test_check_zero_writes_nothing_and_the_default_is_one(tmp_path, monkeypatch):
    inputs.small_case().write(os.fspath(tmp_path))        # the test cases write check=0
    monkeypatch.chdir(tmp_path)
    semsks.Model(["par=run.par"]).write_check()
    for f in ("grid.xz", "legendre.xz", "vel.xy"):
        assert not os.path.exists(f)
    par = [ln for ln in open("run.par").read().splitlines() if not ln.startswith("check=")]
    open("run.par", "w").write("\n".join(par) + "\n")
    semsks.Model(["par=run.par"]).write_check()
    for f in ("grid.xz", "legendre.xz", "vel.xy"):
        assert os.path.getsize(f) > 0


def gll_nodes(n):
    """the n+1 Gauss-Lobatto-Legendre nodes: the ends and the roots of P_n'"""
    inner = np.polynomial.legendre.Legendre.basis(n).deriv().roots()
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_legendre_basis_and_gll_nodes(tmp_path, monkeypatch, n):
    m, D, _ = model_in(tmp_path, monkeypatch, hpn=n)
    m.write_check()
    gll = gll_nodes(n)
    if n == 4:
        assert np.allclose(gll, [-1.0, -np.sqrt(3.0 / 7.0), 0.0, np.sqrt(3.0 / 7.0), 1.0])
    blocks = []
    for line in open("legendre.xz"):
        if line.startswith(">"):
            blocks.append((float(line[1:]), []))
        else:
            x, v = line.split()
            blocks[-1][1].append((float(x), float(v)))
    assert len(blocks) == n + 1
    # the loop variable is a float advanced by 0.05 while <= 1.0: 40
    # samples, the last at 0.95 (float(0.95...) + 0.05 rounds above 1)
    xs, lx = [], F32(-1.0)
    while float(lx) <= 1.0:
        xs.append(float(lx))
        lx = F32(float(lx) + 0.05)
    for i, (node, pts) in enumerate(blocks):
        assert abs(node - gll[i]) <= 5.1e-6, (i, node)
        assert len(pts) == len(xs)
        others = np.delete(gll, i)
        for (xp, v), x in zip(pts, xs):
            assert abs(xp - x) <= 0.0051
            exact = np.prod((x - others) / (gll[i] - others))
            assert abs(v - exact) <= 5.1e-4 + 1e-5, (i, x, v, exact)
    # partition of unity, and the Kronecker property at x = -1
    for k in range(len(xs)):
        assert abs(sum(b[1][k][1] for b in blocks) - 1.0) <= (n + 1) * 5.1e-4
    assert np.allclose([b[1][0][1] for b in blocks], np.eye(n + 1)[0], atol=5.1e-4)


def test_grid_densities_by_layer(tmp_path, monkeypatch):
    m, D, c = model_in(tmp_path, monkeypatch)
    m.write_check()
    nx, nz, h = D["nx"], D["nz"], D["eh"] / c.hpn
    rows = [tuple(float(t) for t in ln.split()) for ln in open("grid.xz")]
    ncol, nrow = len(range(0, nx, 2)), len(range(0, nz, 2))
    assert len(rows) == ncol * (nrow + 1)
    for a, i in enumerate(range(0, nx, 2)):
        blk = rows[a * (nrow + 1):(a + 1) * (nrow + 1)]
        for b, j in enumerate(range(0, nz, 2)):
            gx, gz, rho = blk[b]
            assert abs(gx - i * h) <= 0.0051 and abs(gz - (nz - 1 - j) * h) <= 0.0051
            k = layer_at(c, (nz - 1 - j) * h)
            if k >= 0:
                assert abs(rho - D["rho"][k]) <= 5.1e-5, (i, j, rho, D["rho"][k])
        # the column's last line: the surface node's density, printed with
        # the coordinates of the last row written
        assert blk[-1][:2] == blk[-2][:2]
        assert abs(blk[-1][2] - D["rho"][0]) <= 5.1e-5


@pytest.mark.parametrize("output_vd", [0, 1])
def test_vel_by_layer(tmp_path, monkeypatch, output_vd):
    m, D, c = model_in(tmp_path, monkeypatch, output_vd=output_vd)
    m.write_check()
    nx, nz, h = D["nx"], D["nz"], D["eh"] / c.hpn
    rows = np.loadtxt("vel.xy")
    assert rows.shape == (nx * nz, 3)
    ii, jj = np.divmod(np.arange(nx * nz), nz)           # x outer, depth inner
    assert np.all(np.abs(rows[:, 0] - (250.0 + ii * h) / DEG) <= 5.1e-4)
    assert np.all(np.abs(rows[:, 1] - (nz - 1 - jj) * h) <= 5.1e-3)
    for r, j in zip(rows, jj):
        k = layer_at(c, (nz - 1 - j) * h)
        if k < 0:
            continue
        want = D["vp"][k] if output_vd == 0 else 0.0     # isotropic layers: v44 = v55
        assert abs(r[2] - want) <= 5.1e-3, (r, want)
