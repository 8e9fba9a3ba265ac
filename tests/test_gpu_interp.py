"""The boundary interpolation and guard on the device, off rdt = 1.

The reference reads GRT records at their own spacing dp and interpolates
them to the time step (interpl/interp_const, interpl.c:16-91, rdt = sdt/dp,
sem2d_psvsh.c:388), the coupled solver with one slot ring shared by its
P-SV and SH streams (sem2d_psvsh.c:430-431), then repairs non-finite and
|g| >= 10 values in place, left to right (sem2d_psvsh.c:433-443).  At the
two ends of a vector the reference reads g[-1] and g[mtotal], outside the
array; the device path reads 0 there, as the oracle does.

Bars (bit-exact): the device series of every level against the vectors of
the reference's own compiled interpl.c (tests/golden/ref_interpl_mesh.npz,
written by tests/golden/make_ref_vectors.py at the solver's record widths
of a 2x2-element mesh) with the oracle's guard applied, for rdt in
{1, 0.999, 0.7, 0.37}, one stream (solver 1) and the shared ring (solver 0),
records with and without spikes.  And one seismogram run with records at
dp = sdt / 0.7 against the FP64 oracle (rel-L2 and peak misfit <= 1e-4).
"""
import os

import numpy as np
import pytest

import pyoracle
import semsks
from semsks import inputs

from cases import interp_mesh_case

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("tag", ["plain", "spiky"])
@pytest.mark.parametrize("rdt", ["1", "0p999", "0p7", "0p37"])
@pytest.mark.parametrize("solver", [0, 1])
def test_device_series_matches_reference_interpl(tmp_path, monkeypatch, tag, rdt, solver):
    print("case: interp", tag, rdt, "solver", solver, flush=True)
    z = np.load(os.path.join(GOLD, "ref_interpl_mesh.npz"))
    key = "%s_r%s" % (tag, rdt)
    nt_sem = int(z[key + "_ntsem"])
    d = interp_mesh_case(tmp_path, tag, key, solver)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    assert m.nt_sem == nt_sem                        # the reader's nt_sem is the reference's
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        g, gv = e.boundary_series(s, nt_sem)
    # the guard runs on each level's vector
    want_g = np.stack([pyoracle.guard(r, "f32") for r in z[key + ("_shared_g" if solver == 0 else "_single_g")]])
    if solver == 0:
        want_gv = np.stack([pyoracle.guard(r, "f32") for r in z[key + "_shared_gv"]])
    else:
        want_gv = np.zeros_like(gv)          # sem2d.c reads no SH records
    assert np.isfinite(want_g).all() and np.isfinite(want_gv).all()
    assert np.array_equal(g.view(np.uint32), want_g.view(np.uint32)), np.argwhere(g != want_g)[:5]
    assert np.array_equal(gv.view(np.uint32), want_gv.view(np.uint32)), np.argwhere(gv != want_gv)[:5]
    if tag == "spiky":
        raw = z[key + ("_shared_g" if solver == 0 else "_single_g")]
        assert (~(np.abs(raw) < 10)).any()              # the guard had something to repair


def test_seismograms_at_coarse_record_spacing(tmp_path, monkeypatch):
    """Records of the plane wave spaced dp = sdt / 0.7 (the GRT's own sample
    interval, coarser than the time step), coupled solver: the device run
    against the FP64 oracle on the same files."""
    print("case: seismograms at dp = sdt/0.7", flush=True)
    c = inputs.small_case(solver=0, psi_deg=30.0, sh_zero=False, nst=600)
    d = os.fspath(tmp_path)
    D = c.write(d, write_records=False)
    sdt = D["sdt"]
    dp_txt = "%.6f" % (sdt / 0.7)
    dp = float(np.float32(float(dp_txt)))
    nt = int(600 * 0.7) + 8
    pw = inputs.PlaneWave(p=c.p, beta=D["vs"][-1], f0=c.f0, psi_deg=c.psi_deg)
    psv, sh = inputs.greenfile_records(pw, c.nex, D["nez"], D["eh"], c.hpn, c.neh, nt, dp)
    psv.tofile(os.path.join(d, "greenfile"))
    sh.tofile(os.path.join(d, "shgreenfile"))
    par = open(os.path.join(d, "run.par")).read().splitlines()
    par = [("dp=" + dp_txt) if ln.startswith("dp=") else ("nt=%d" % nt) if ln.startswith("nt=") else ln
           for ln in par]
    open(os.path.join(d, "run.par"), "w").write("\n".join(par) + "\n")
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        tr = e.traces(s)
    ref = pyoracle.Oracle(["par=run.par"], "f64").run()
    assert tr.shape == ref.shape
    from test_gpu_parity import misfits
    for k, (l2, pk) in enumerate(misfits(tr, ref)):
        assert np.abs(ref[:, k]).max() > 0
        assert l2 <= 1e-4 and pk <= 1e-4, (k, l2, pk)
