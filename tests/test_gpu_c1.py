"""The reference's own CPU-runnable case -- 200x100 elements, three flat
isotropic layers, plane SV at p = 0.06, 4000 steps, 50 surface receivers,
P-SV solver -- on the GPU against the seismograms of the FP64 oracle, which
tests/golden/make_c1_fp64.py wrote to tests/golden/c1_fp64_traces.npz:
every one of the 100 seismograms (u and -w at 50 receivers) within 1e-4
relative L2 and 1e-4 peak misfit."""
import os

import numpy as np
import pytest

import semsks
from semsks import configs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_c1_every_seismogram_vs_fp64(tmp_path, monkeypatch):
    print("case: 200x100 elements, 4000 steps, 50 receivers vs FP64 fixture", flush=True)
    ref = np.load(os.path.join(HERE, "golden", "c1_fp64_traces.npz"))
    configs.c1_case().write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    m = semsks.Model(["par=run.par"])
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        tr = e.traces(s)
    assert tr.shape == (4000, 3, 50)
    assert np.abs(tr[:, 2]).max() == 0.0                      # P-SV: no v
    for k, key in ((0, "u"), (1, "mw")):
        a, b = tr[:, k].astype(np.float64), ref[key].astype(np.float64)
        assert np.abs(b).max(axis=0).min() > 0.1               # the wave reached every receiver
        l2 = np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)
        pk = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
        assert l2.max() <= 1e-4 and pk.max() <= 1e-4, (key, l2.max(), pk.max())


def _worst_errors(tmp_path, monkeypatch, flags):
    ref = np.load(os.path.join(HERE, "golden", "c1_fp64_traces.npz"))
    os.makedirs(tmp_path, exist_ok=True)
    configs.c1_case().write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    m = semsks.Model(["par=run.par"])
    m.problem.flags = flags
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        tr = e.traces(s)
    out = []
    for k, key in ((0, "u"), (1, "mw")):
        a, b = tr[:, k].astype(np.float64), ref[key].astype(np.float64)
        out.append(float((np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)).max()))
    return out


def test_c1_compensation_reduces_the_error(tmp_path, monkeypatch):
    """The packed compensation (5-bit fields of levels t and t-dt in one
    word) is what holds the FP32 loop near the FP64 oracle: with it the worst
    seismogram misfit of u and of w is at least 3x smaller than with the
    plain FP32 update (SEMSKS_F_NO_KAHAN) on the same 4000 steps.  The plain
    update reads and writes no compensation words; it still stays within the
    FP32 drift bound of 1e-4 that the reference's own arithmetic meets."""
    print("case: 200x100 elements, 4000 steps, compensated vs plain FP32 update", flush=True)
    comp = _worst_errors(tmp_path / "k", monkeypatch, 0)
    plain = _worst_errors(tmp_path / "p", monkeypatch, semsks.F_NO_KAHAN)
    print("worst rel-L2 u, w: compensated %.2e %.2e, plain %.2e %.2e" % (comp[0], comp[1], plain[0], plain[1]))
    for c, p in zip(comp, plain):
        assert c * 3.0 <= p, (comp, plain)
        assert p <= 1e-4, plain
