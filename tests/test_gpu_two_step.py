"""The two-step sweep (step2_kernel: two time levels per pass over the
mesh, engine option steps_per_sweep = 2) against one step per launch:
traces and the final field must be identical bit for bit, for every rule
the loop applies between the two steps -- injection (GRT_Interface4
P-SV+SH and P-SV, GRT_Interface3), absorbing edges, per-node material
tiles, odd step counts, uneven batches -- and the engine must fall back to
one step where the sweep does not apply."""
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


@pytest.mark.parametrize("solver,nst", [(0, 301), (1, 300), (2, 257)])
def test_two_step_equals_one_step(tmp_path, monkeypatch, solver, nst):
    print("case: small 24x12, solver %d, nst %d, one vs two steps per sweep" % (solver, nst), flush=True)
    from semsks import aniso
    C = aniso.hexagonal_tilted(vp=7.8, vs=4.4, rho=3.3, strength=0.04, dip_deg=30.0, azimuth_deg=60.0)
    interfaces = [([0.0, 50.0], [14.0, 18.0]), None, None]       # dipping: per-node material tiles
    d = make_case(tmp_path, solver=solver, nst=nst, interfaces=interfaces, aniso=[C, None, None],
                  psi_deg=30.0 if solver == 0 else 0.0, sh_zero=solver != 0)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    t1, f1, s1 = run(m.problem, steps_per_sweep=1)
    assert s1 == 1
    assert np.abs(t1).max() > 1e-3                  # the incident wave reached the receivers
    for tz in (3, 0):
        t2, f2, s2 = run(m.problem, tile_z=tz, steps_per_sweep=2)
        assert s2 == 2
        assert same(t2, t1), tz
        assert same(f2, f1), tz
    # eager launches and odd chunks (one-step launches in between)
    t3, f3, _ = run(m.problem, chunks=[7, 1, 40, nst], use_graph=0, steps_per_sweep=2)
    assert same(t3, t1) and same(f3, f1)


def test_two_step_uneven_batch(tmp_path, monkeypatch):
    """Instances with different lengths in one sweep: the one with a single
    step left takes only the first step; each equals its one-step solo run."""
    print("case: batch of 3 (nst 300/301/157, nex 24/30/18), two-step vs one-step solo", flush=True)
    probs = []
    for k, (nex, nst, p) in enumerate([(24, 300, 0.06), (30, 301, 0.05), (18, 157, 0.07)]):
        d = make_case(tmp_path / ("c%d" % k), nex=nex, nst=nst, p=p, psi_deg=25.0, sh_zero=False)
        monkeypatch.chdir(d)
        probs.append(semsks.Model(["par=run.par"]))
    solo = [run(m.problem, steps_per_sweep=1) for m in probs]
    with semsks.Engine(max_slots=3, steps_per_sweep=2) as e:
        slots = [e.admit(m.problem) for m in probs]
        e.advance(max(m.problem.nst for m in probs) - 1)
        assert e.sweep_steps() == 2
        for s, (t, f, _) in zip(slots, solo):
            assert same(e.traces(s), t)
            assert same(e.field(s), f)


def test_receiver_on_injection_column_falls_back(tmp_path, monkeypatch):
    """A receiver on the interface column would record the first level of a
    sweep after the next step's injection, so such an instance runs one step
    per launch.  One step per launch records it before the injection, as
    kirrecord_zhl does before the next GRT_Interface4: its trace, and the
    others, match the FP64 oracle."""
    print("case: small, receiver at ix = nbl, fallback to one step", flush=True)
    d = make_case(tmp_path, nst=500)
    monkeypatch.chdir(d)
    args = ["par=run.par", "get_rec=0", "xbeg=2.75", "idxplot=3", "xtrace=5"]
    m = semsks.Model(args)
    ix, iz = m.receivers()
    assert ix[0] == 4 and iz[0] > 4                   # nbl = neh * hpn: the left interface column
    t1, f1, s1 = run(m.problem, steps_per_sweep=1)
    t0, f0, s0 = run(m.problem, steps_per_sweep=2)
    assert s0 == 1 and same(t0, t1) and same(f0, f1)
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
