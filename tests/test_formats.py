"""The product's host-side input layer (semsks_load_par) against the oracle's
restatement and the reference's documented rules (no GPU needed).

* material per node, mesh sizes and receiver indices: bit-exact with the
  oracle's discret() restatement, for flat, dipping and tilted-axis models
  and for the ratio (multiplicative) semmodel layers;
* getpar semantics (getpar.c): last definition wins, par= files nest,
  '#' comments, $(name) substitution, quotes;
* error statuses that replace the reference's exit() calls.
"""
import os

import numpy as np
import pytest

import pyoracle
import semsks
from semsks import aniso, inputs


def load_both(d, extra=()):
    old = os.getcwd()
    os.chdir(d)
    try:
        m = semsks.Model(["par=run.par", *extra])
        o = pyoracle.Oracle(["par=run.par", *extra], "f32")
    finally:
        os.chdir(old)
    return m, o


def assert_same_model(m, o):
    assert (m.nx, m.nz) == (o.nx, o.nz)
    assert m.problem.nex == o.nex and m.problem.nez == o.nez
    assert np.float32(m.problem.eh) == np.float32(o.eh)
    assert np.array_equal(m.material().view(np.uint32), o.material().view(np.uint32))
    assert np.array_equal(np.stack(m.receivers()), np.stack(o.receivers()))
    assert m.nt_sem == o.nt_sem


@pytest.mark.parametrize("kind", ["flat", "dipping", "tilted", "ratio"])
def test_material_and_mesh_bit_exact(tmp_path, kind):
    kw = {}
    if kind == "dipping":
        kw = dict(interfaces=[([0.0, 20.0, 50.0], [13.0, 17.5, 15.0]), ([0.0, 50.0], [19.0, 21.0]), None])
    if kind == "tilted":
        kw = dict(aniso=[aniso.hexagonal_tilted(7.8, 4.4, 3.3, 0.06, 40.0, 10.0), None,
                         aniso.hexagonal_tilted(5.9, 3.4, 2.6, 0.03, 0.0, 45.0)])
    c = inputs.small_case(nst=5, **kw)
    d = os.fspath(tmp_path)
    c.write(d)
    if kind == "ratio":
        # second layer as multiplicative factors on the layer below it
        lines = open(os.path.join(d, "semmodel")).read().splitlines()
        parts = lines[4].split()
        parts[0] = "-" + parts[0]
        lines[4] = " ".join(parts[:1] + ["1.05"] * 14)
        open(os.path.join(d, "semmodel"), "w").write("\n".join(lines) + "\n")
    m, o = load_both(d)
    assert_same_model(m, o)
    if kind == "tilted":
        mat = m.material()
        assert np.abs(mat[:, [3, 4, 7, 8, 12, 14]]).max() > 0.01   # the six tilted-axis constants


def test_getpar_semantics(tmp_path):
    c = inputs.small_case(nst=7)
    d = os.fspath(tmp_path)
    c.write(d)
    with open(os.path.join(d, "inner.par"), "w") as f:
        f.write("# nested file\nnst=11   hpn=4\n")
    with open(os.path.join(d, "outer.par"), "w") as f:
        f.write("par=run.par\nnst=9\npar=inner.par\nmyrec='rec_file'\nrec_file=$(myrec)\n")
    m, o = load_both(d, [])
    assert m.problem.nst == o.nst == 7
    old = os.getcwd()
    os.chdir(d)
    try:
        m2 = semsks.Model(["par=outer.par"])
        o2 = pyoracle.Oracle(["par=outer.par"], "f32")
        assert m2.problem.nst == o2.nst == 11          # the nested file is the last definition
        m3 = semsks.Model(["par=outer.par", "nst=13"])
        assert m3.problem.nst == 13
    finally:
        os.chdir(old)


@pytest.mark.parametrize("what,code,ocode", [
    ("missing_par", 100, 100), ("stop", 101, 101), ("missing_file", 2, -1),
    ("np_too_big", 6, -2), ("out_of_range", 5, -3)])
def test_error_statuses(tmp_path, what, code, ocode):
    c = inputs.small_case(nst=5)
    d = os.fspath(tmp_path)
    c.write(d)
    args = ["par=run.par"]
    if what == "missing_par":
        lines = [l for l in open(os.path.join(d, "run.par")) if not l.startswith("sdt=")]
        open(os.path.join(d, "run.par"), "w").writelines(lines)
    elif what == "stop":
        args.append("STOP=1")
    elif what == "missing_file":
        args.append("semmodel=nosuchfile")
    elif what == "np_too_big":
        lines = open(os.path.join(d, "semmodel")).read().splitlines()
        p = lines[1].split()
        p[0] = "1001"
        lines[1] = " ".join(p)
        open(os.path.join(d, "semmodel"), "w").write("\n".join(lines) + "\n")
    elif what == "out_of_range":
        lines = open(os.path.join(d, "semmodel")).read().splitlines()
        lines[2] = "10.0 20.0"
        open(os.path.join(d, "semmodel"), "w").write("\n".join(lines) + "\n")
    old = os.getcwd()
    os.chdir(d)
    try:
        with pytest.raises(semsks.SemsksError) as ei:
            semsks.Model(args)
        assert ei.value.code == code
        with pytest.raises(pyoracle.OracleError) as eo:
            pyoracle.Oracle(args, "f32")
        assert eo.value.code == ocode
    finally:
        os.chdir(old)


def test_cli_exit_status_on_missing_par(tmp_path):
    import subprocess
    r = subprocess.run([semsks.CLI_PATH, "nst=3"], capture_output=True, text=True, cwd=os.fspath(tmp_path))
    assert r.returncode == 100 and "must specify" in r.stderr


def test_greenfile_record_layout(tmp_path):
    """The plane-wave producer writes [u_bot | w_bot | u_left | w_left] with w
    positive down, and [v_bot | v_left] (demult.c:63-77)."""
    c = inputs.small_case(nst=40, psi_deg=30.0, sh_zero=False)
    d = os.fspath(tmp_path)
    D = c.write(d)
    nx, nz = D["nx"], D["nz"]
    g = np.fromfile(os.path.join(d, "greenfile"), np.float32).reshape(D["nt"], 2 * (nx + nz))
    s = np.fromfile(os.path.join(d, "shgreenfile"), np.float32).reshape(D["nt"], nx + nz)
    pw = inputs.PlaneWave(p=c.p, beta=D["vs"][-1], f0=c.f0, psi_deg=c.psi_deg)
    xs = inputs.node_coords(c.nex, D["eh"], c.hpn)
    zs = inputs.node_coords(D["nez"], D["eh"], c.hpn)
    t = 30 * float(np.float32(D["sdt"]))
    u, v, w = pw.fields(xs, np.full(nx, zs[c.neh * c.hpn]), t)
    assert np.array_equal(g[30, :nx], u.astype(np.float32))
    assert np.array_equal(g[30, nx:2 * nx], (-w).astype(np.float32))
    ul, vl, wl = pw.fields(np.full(nz, xs[c.neh * c.hpn]), zs, t)
    assert np.array_equal(g[30, 2 * nx + nz:], (-wl).astype(np.float32))
    assert np.array_equal(s[30, nx:], vl.astype(np.float32))
    # SV polarisation is perpendicular to the ray, SH share sin(psi)
    s_ = c.p * D["vs"][-1]
    assert np.allclose(u * s_ + w * np.sqrt(1 - s_ * s_), 0, atol=1e-12)
