"""The workflow's two steps end to end on the GPU output: the solver CLI
(`semsks par=run.par`, the drop-in for sem2d_psvsh) writes seis.u/.w/.v,
then `semsks_uvw2sac` (the drop-in for uvw2sac, semsyn/plot/uvw2sac.c:207,
300-303, 377-386, 462-467) turns them into SAC files, with and without the
source convolution of source.f.  The same two steps on the FP64 oracle's
seis.* give the reference SAC files: every SAC trace of the GPU run within
1e-4 relative L2 and peak misfit of the oracle's, with identical headers."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import semsks
from semsks import inputs, sac

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("conv", [0, 1])
def test_cli_then_uvw2sac_vs_fp64(tmp_path, monkeypatch, conv):
    print("case: CLI + uvw2sac, Conv_Source=%d" % conv, flush=True)
    d = os.fspath(tmp_path)
    c = inputs.small_case(nst=600, solver=0, psi_deg=30.0, sh_zero=False)
    c.write(d)
    monkeypatch.chdir(d)
    os.makedirs("gpu", exist_ok=True)
    os.makedirs("cpu", exist_ok=True)
    r = subprocess.run([semsks.CLI_PATH, "par=run.par", "outdir=gpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert pyoracle.main(["par=run.par"], "cpu", kind="f64") == 0
    extra = "zbeg=0\nidzplot=1\nzs=100.0\ntstart=0\nConv_Source=%d\n" % conv
    if conv:
        extra += "Moment=1\nDt1=0.4\nDt2=0.3\nDt3=0.4\nbeta=3.5\npoint_source=0\n"
    open("sac.par", "w").write(open("run.par").read() + "\n" + extra)
    for side in ("gpu", "cpu"):
        os.makedirs(side + "_sac", exist_ok=True)
        r = subprocess.run([semsks.SAC_CLI_PATH, "par=sac.par", "indir=" + side, "outdir=" + side + "_sac"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
    names = sorted(f for f in os.listdir("gpu_sac") if f.endswith((".ce", ".cn", ".cz")))
    assert len(names) == 3 * 8 and names == sorted(f for f in os.listdir("cpu_sac") if f.endswith((".ce", ".cn", ".cz")))
    for nm in names:
        hg, yg = sac.read(os.path.join("gpu_sac", nm))
        hc, yc = sac.read(os.path.join("cpu_sac", nm))
        for key in ("delta", "b", "e", "o", "gcarc", "stlo", "evdp", "cmpaz", "cmpinc", "npts", "kstnm"):
            assert hg[key] == hc[key], (nm, key)
        a, b = yg.astype(np.float64), yc.astype(np.float64)
        assert np.abs(b).max() > 0, nm
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b), nm
        assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max(), nm
