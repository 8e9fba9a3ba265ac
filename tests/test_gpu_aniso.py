"""The C2 model -- crust, 4% hexagonal upper mantle with a tilted axis (15
constants, the general stress form), mantle -- on the 200x100-element mesh,
coupled P-SV+SH solver, plane S wave polarised 30 degrees off SV, 5000
steps, 50 surface receivers, on the GPU against the seismograms of the FP64
oracle, which tests/golden/make_aniso_fp64.py wrote to
tests/golden/aniso_fp64_traces.npz: every one of the 150 seismograms (u, -w
and v at 50 receivers) within 1e-4 relative L2 and 1e-4 peak misfit."""
import os

import numpy as np
import pytest

import semsks
from semsks import configs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_aniso_every_seismogram_vs_fp64(tmp_path, monkeypatch):
    print("case: C2 model on the 200x100 mesh, 5000 steps, 50 receivers, u/-w/v vs FP64 fixture", flush=True)
    ref = np.load(os.path.join(HERE, "golden", "aniso_fp64_traces.npz"))
    configs.c2_model_c1_mesh(nst=5000).write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    m = semsks.Model(["par=run.par"])
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        tr = e.traces(s)
    assert tr.shape == (5000, 3, 50)
    worst = {}
    for k, key in ((0, "u"), (1, "mw"), (2, "v")):
        a, b = tr[:, k].astype(np.float64), ref[key].astype(np.float64)
        assert np.abs(b).max(axis=0).min() > 0.1               # the wave reached every receiver
        l2 = np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)
        pk = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
        worst[key] = (l2.max(), pk.max())
        assert l2.max() <= 1e-4 and pk.max() <= 1e-4, (key, l2.max(), pk.max())
    print("worst rel-L2, peak misfit:", worst)


def test_aniso_20000_steps_vs_fp64(tmp_path, monkeypatch):
    """The same model at the production length of the coupled solver's loop:
    20000 steps (sem2d_psvsh.c:418-512 with nst = 20000), against
    tests/golden/aniso20k_fp64_traces.npz (every 5th level, written by
    tests/golden/make_aniso20k_fp64.py): the drift of the compensated FP32
    update over the whole run, every one of the 150 seismograms within 1e-4
    relative L2 and peak misfit.  The plain FP32 update (SEMSKS_F_NO_KAHAN)
    on the same run must be at least 3x worse in its worst seismogram."""
    print("case: C2 model on the 200x100 mesh, 20000 steps, 50 receivers, u/-w/v vs FP64 fixture", flush=True)
    ref = np.load(os.path.join(HERE, "golden", "aniso20k_fp64_traces.npz"))
    nst, dec = int(ref["nst"]), int(ref["decim"])
    configs.c2_model_c1_mesh(nst=nst).write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    m = semsks.Model(["par=run.par"])
    worst = {}
    for flags in (0, semsks.F_NO_KAHAN):
        p = semsks.Problem.from_buffer_copy(m.problem)
        p.flags = flags
        with semsks.Engine() as e:
            s = e.admit(p)
            e.advance(p.nst - 1)
            tr = e.traces(s)[::dec]
        assert tr.shape == ref["u"].shape[:1] + (3, 50)
        w = 0.0
        for k, key in ((0, "u"), (1, "mw"), (2, "v")):
            a, b = tr[:, k].astype(np.float64), ref[key].astype(np.float64)
            assert np.abs(b).max(axis=0).min() > 0.1
            l2 = np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)
            pk = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
            worst[(flags, key)] = (float(l2.max()), float(pk.max()))
            w = max(w, l2.max(), pk.max())
            if flags == 0:
                assert l2.max() <= 1e-4 and pk.max() <= 1e-4, (key, l2.max(), pk.max())
        worst[flags] = w
    print("worst rel-L2, peak misfit:", worst)
    assert worst[semsks.F_NO_KAHAN] >= 3 * worst[0]
