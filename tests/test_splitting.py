"""Splitting measurement (the cal_split.pl step) on synthetic waveforms with
a known answer: a radially polarised pulse split with a given fast azimuth
and delay must come back as that azimuth and delay (Silver & Chan's
minimum transverse energy is exact for noise-free data), and the station
loop over uvw2sac's SAC files writes split_par.txt in the reference's
format."""
import os
import subprocess

import numpy as np
import pytest

import semsks
from semsks import inputs, splitting


def pulse(n=2000, dt=0.05, t0=40.0, f0=0.15):
    t = np.arange(n) * dt
    return inputs.ricker(t - t0, f0)


@pytest.mark.parametrize("phi,delay", [(30.0, 1.2), (-50.0, 0.8), (75.0, 2.0)])
def test_recovers_known_splitting(phi, delay):
    dt, baz = 0.05, 270.0
    r = pulse(dt=dt)
    # radial along the propagation azimuth baz+180 = 90 (east): n = 0, e = r
    n, e = splitting.split(np.zeros_like(r), r, dt, phi, delay)
    _, t = splitting.rotate_rt(n, e, baz)
    assert np.abs(t).max() > 0.1                       # the split wave has transverse energy
    res = splitting.silver_chan(n, e, dt, baz, (500, 1300))
    assert abs(res["phi"] - phi) <= 1.0 and abs(res["delay"] - delay) <= 0.05 + 1e-9, res["phi"]
    assert res["e_min"] < 1e-6 * np.sum(r ** 2)
    assert res["d_phi"] >= 0 and res["d_delay"] >= 0


def test_unsplit_wave_has_no_transverse_energy():
    dt = 0.05
    r = pulse(dt=dt)
    n, e = splitting.split(np.zeros_like(r), r, dt, 90.0, 1.5)   # fast axis along the radial: a null
    _, t = splitting.rotate_rt(n, e, 270.0)
    assert np.abs(t).max() < 1e-6


def test_station_loop_over_sac_files(tmp_path):
    d = os.fspath(tmp_path)
    c = inputs.small_case(nst=3)
    c.write(d)
    nst, ntr, sdt = 2001, 3, float(np.float32(c.derived()["sdt"]))
    dts = 2 * sdt
    r = pulse(n=nst, dt=sdt, t0=20.0, f0=0.3)
    u = np.zeros((nst, ntr), np.float32); v = np.zeros_like(u); w = np.zeros_like(u)
    truth = [(20.0, 0.6), (-35.0, 1.0), (60.0, 0.4)]
    for k, (phi, delay) in enumerate(truth):
        n, e = splitting.split(np.zeros_like(r), r, sdt, phi, delay)
        u[:, k], v[:, k] = e, n
    for name, a in (("u", u), ("v", v), ("w", w)):
        a.tofile(os.path.join(d, "seis." + name))
    par = open(os.path.join(d, "run.par")).read()
    par += "\nget_rec=0\nnst=%d\nxtrace=%d\nzbeg=0\nidzplot=1\nzs=100.0\nConv_Source=0\n" % (nst, ntr)
    open(os.path.join(d, "sac.par"), "w").write(par)
    p = subprocess.run([semsks.SAC_CLI_PATH, "par=sac.par"], cwd=d, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    rows = splitting.measure_dir(d, before=6.0, after=10.0)
    assert len(rows) == ntr
    for (name, gcarc, phi, dphi, delay, ddelay), (phi0, delay0) in zip(rows, truth):
        assert abs(phi - phi0) <= 1.0 and abs(delay - delay0) <= 0.05 + 1e-9
    lines = open(os.path.join(d, "split_par.txt")).read().splitlines()
    assert len(lines) == ntr and lines[0].split()[0] == "0"
