"""Stepping mechanics of the engine, bit for bit: captured hipGraph runs
against eager launches, instances of different lengths in one launch
against their solo runs, and a receiver on the interface column, which is
recorded before the next step's injection (kirrecord_zhl runs before the
next GRT_Interface4), against the FP64 oracle."""
import os

import numpy as np
import pytest

import semsks
from semsks import inputs

pytestmark = pytest.mark.gpu


def make_case(d, **kw):
    c = inputs.small_case(**kw)
    c.write(os.fspath(d))
    return os.fspath(d)


def run(problem, chunks=None, **eng):
    """traces, final field and steps per launch of one instance"""
    with semsks.Engine(**eng) as e:
        s = e.admit(problem)
        for n in (chunks or [problem.nst - 1]):
            e.advance(n)
        return e.traces(s), e.field(s), e.sweep_steps()


def same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_uneven_lengths_in_one_launch(tmp_path, monkeypatch):
    """Instances of different lengths and meshes in one launch: each stops at
    its own last level and equals its solo run (traces and final field)."""
    print("case: batch of 3 (nst 300/301/157, nex 24/30/18) vs solo", flush=True)
    probs = []
    for k, (nex, nst, p) in enumerate([(24, 300, 0.06), (30, 301, 0.05), (18, 157, 0.07)]):
        d = make_case(tmp_path / ("c%d" % k), nex=nex, nst=nst, p=p, psi_deg=25.0, sh_zero=False)
        monkeypatch.chdir(d)
        probs.append(semsks.Model(["par=run.par"]))
    solo = [run(m.problem) for m in probs]
    with semsks.Engine(max_slots=3) as e:
        slots = [e.admit(m.problem) for m in probs]
        e.advance(max(m.problem.nst for m in probs) - 1)
        for s, (t, f, _) in zip(slots, solo):
            assert same(e.traces(s), t)
            assert same(e.field(s), f)


def test_receiver_on_injection_column(tmp_path, monkeypatch):
    """A receiver on the interface column is recorded before the next step's
    injection, as kirrecord_zhl does before the next GRT_Interface4: its
    trace, and the others, match the FP64 oracle."""
    print("case: small, receiver at ix = nbl", flush=True)
    d = make_case(tmp_path, nst=500)
    monkeypatch.chdir(d)
    args = ["par=run.par", "get_rec=0", "xbeg=2.75", "idxplot=3", "xtrace=5"]
    m = semsks.Model(args)
    ix, iz = m.receivers()
    assert ix[0] == 4 and iz[0] > 4                   # nbl = neh * hpn: the left interface column
    t1, f1, s1 = run(m.problem)
    import pyoracle
    ref = pyoracle.Oracle(args, "f64", cwd=d).run()
    for k in (0, 1):
        a, b = t1[:, k, 0].astype(np.float64), ref[:, k, 0]
        assert np.abs(b).max() > 1e-3
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b), k


def test_captured_runs_equal_eager_steps(tmp_path, monkeypatch):
    """Advances of odd lengths replay the captured runs of 16, 8, 4, 2 and 1
    steps many times, with eager host work and synchronisation between the
    calls; traces and field equal those of eager launches bit for bit."""
    print("case: small, chunks 5/16/3/1/27/rest, captured vs eager", flush=True)
    d = make_case(tmp_path, nst=240, psi_deg=20.0, sh_zero=False)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    chunks = [5, 16, 3, 1, 27, 31, 2, 154]
    assert sum(chunks) == m.problem.nst - 1
    t0, f0, _ = run(m.problem, chunks=chunks, use_graph=0)
    assert np.abs(t0).max() > 1e-3
    t1, f1, _ = run(m.problem, chunks=chunks, use_graph=1)
    assert same(t1, t0) and same(f1, f0)
