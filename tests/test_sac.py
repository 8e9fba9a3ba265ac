"""The uvw2sac step (semsyn/plot/uvw2sac.c) as semsks_uvw2sac: no GPU needed.

The SAC files are read back with a reader written from the published SAC
version-6 layout (632-byte header: 70 float words, 40 integer words, 24
eight-byte strings with kevnm taking two; undefined = -12345), and checked
against the rules of uvw2sac.c: every second sample kept (istep =
2*sdt/sdt), the last one dropped when nt is odd, amplitudes x enlarge,
delta/b/e/o/gcarc/stlo/evdp/kstnm, component orientations.  Conv_Source=1
is checked through the closed form of a spike convolved with the unit-area
trapezoid of source.f.
"""
import os
import struct
import subprocess

import numpy as np
import semsks
from semsks import inputs


def read_sac(path):
    raw = open(path, "rb").read()
    f = struct.unpack("<70f", raw[:280])
    i = struct.unpack("<40i", raw[280:440])
    k = raw[440:632]
    n = i[9]
    y = np.frombuffer(raw[632:632 + 4 * n], "<f4")
    assert len(raw) == 632 + 4 * n
    return f, i, k, y


def setup_run(d, nst=11, ntrace=4, extra=""):
    c = inputs.small_case(nst=3)
    c.write(d)
    rng = np.random.default_rng(5)
    seis = {k: rng.standard_normal((nst, ntrace)).astype(np.float32) for k in "uvw"}
    for k, a in seis.items():
        a.tofile(os.path.join(d, "seis." + k))
    par = open(os.path.join(d, "run.par")).read()
    par += "\nget_rec=0\nnst=%d\nxtrace=%d\nzbeg=0\nidzplot=1\nzs=100.0\nConv_Source=0\ntstart=12.5\nxmin=1000\n%s\n" % (
        nst, ntrace, extra)
    open(os.path.join(d, "sac.par"), "w").write(par)
    return seis


def run_tool(d, *args):
    return subprocess.run([semsks.SAC_CLI_PATH, "par=sac.par", *args], cwd=d, capture_output=True, text=True)


def test_sac_files_follow_uvw2sac(tmp_path):
    d = os.fspath(tmp_path)
    seis = setup_run(d)
    r = run_tool(d)
    assert r.returncode == 0, r.stderr
    pars = dict(l.split("=", 1) for l in open(os.path.join(d, "sac.par")).read().split() if "=" in l)
    sdt, xbeg, idxplot = np.float32(pars["sdt"]), np.float32(pars["xbeg"]), int(pars["idxplot"])
    hpn, eh = int(pars.get("hpn", 4)), None
    # mesh spacing as the reader computes it (nez, eh from the raymodel thickness)
    c = inputs.small_case(nst=3)
    eh = np.float32(c.derived()["eh"])
    h = np.float32(eh / np.float32(hpn))
    for t in range(4):
        gc = (np.float32(1000.0) + xbeg + np.float32(idxplot * t) * h) / np.float32(111.2)
        for ext, key, az, inc in ((".ce", "u", 90.0, 90.0), (".cn", "v", 0.0, 90.0), (".cz", "w", 0.0, 0.0)):
            f, i, k, y = read_sac(os.path.join(d, "%d%s" % (t, ext)))
            assert f[0] == np.float32(2 * sdt)                       # delta = istep*dt
            assert i[9] == 5                                         # 11 records: 0,2,..,10 kept, last dropped
            assert np.array_equal(y, seis[key][0:10:2, t] * np.float32(4.0))
            assert f[5] == 0.0 and f[6] == np.float32(5 * f[0]) and f[7] == np.float32(-12.5)
            assert abs(f[53] - gc) <= 1e-6 * gc and f[32] == f[53]  # gcarc = stlo
            assert f[38] == np.float32(100.0) and f[35] == 0.0 and f[36] == 0.0 and f[31] == 0.0
            assert (f[57], f[58]) == (az, inc)
            assert i[0] == 2006 and i[6] == 6 and i[15] == 1 and i[35] == 1
            assert k[:8] == (str(t) + " " * 8)[:8].encode()
            assert f[1] == -12345.0 and i[7] == -12345 and k[8:24] == b"-12345          "
    pm = open(os.path.join(d, "pm.xy")).read().splitlines()
    assert pm[0].startswith(">> ") and len(pm) == 4 * 6
    log = open(os.path.join(d, "log.txt")).read().split()
    assert float(log[0]) == 12.5


def test_sac_source_convolution_closed_form(tmp_path):
    """A spike convolved with the unit-area trapezoid (rise Dt1, flat Dt2,
    fall Dt3) gives the trapezoid, scaled by dt, 1.4142/sqrt(r) and enlarge."""
    d = os.fspath(tmp_path)
    nst = 401
    setup_run(d, nst=nst, ntrace=2,
              extra="Conv_Source=1\nMoment=1\nDt1=0.2\nDt2=0.3\nDt3=0.25\nbeta=3.5\npoint_source=0")
    spike = np.zeros((nst, 2), np.float32)
    spike[100, :] = 1.0
    spike[-3, :] = 1e-20      # uvw2sac trims trailing zeros before convolving; keep the trace whole
    for k in "uvw":
        spike.tofile(os.path.join(d, "seis." + k))
    r = run_tool(d)
    assert r.returncode == 0, r.stderr
    f, i, k, y = read_sac(os.path.join(d, "0.ce"))
    dt = f[0]
    n1, n2, n3 = (int(np.float32(x) / dt + np.float32(0.1)) for x in (0.2, 0.3, 0.25))
    n = n1 + n2 + n3 + 1
    tri = np.array([0.0] + [min(1.0, (j - 1) / n1) if j < n1 + 1 else
                            (1.0 - (j - (n1 + n2 + 1)) / n3 if j > n1 + n2 + 1 else 1.0) for j in range(2, n + 1)])
    tri /= np.sum(0.5 * (tri[1:] + tri[:-1])) * dt                   # unit area under the trapezoid rule
    gc = f[53]
    r_km = np.hypot(gc * 111.2, 100.0)
    scale = 4.0 * dt * 1.4142 / np.sqrt(np.float32(r_km))
    k0 = 50                                                          # record 100 -> sample 50
    assert np.allclose(y[k0:k0 + n], scale * tri, rtol=2e-5, atol=1e-7 * scale * tri.max())
    assert np.abs(y[:k0]).max() == 0.0
    assert abs(np.sum(y) * dt - 4.0 * dt * 1.4142 / np.sqrt(r_km)) < 1e-3 * abs(np.sum(y) * dt)   # unit area kept


def test_sac_missing_parameter_status(tmp_path):
    d = os.fspath(tmp_path)
    setup_run(d)
    lines = [l for l in open(os.path.join(d, "sac.par")) if not l.startswith("Conv_Source")]
    open(os.path.join(d, "sac.par"), "w").writelines(lines)
    r = run_tool(d)
    assert r.returncode == 100 and "Conv_Source" in r.stderr
