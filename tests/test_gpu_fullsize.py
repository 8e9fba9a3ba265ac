"""The full-size meshes of the benchmark (2000x500 elements, 16,010,001 GLL
points per instance) through properties that do not depend on size, since
the oracle cannot step them in seconds:

  * tiling invariance: fixed tile heights and the autotuned one give
    bit-identical fields;
  * batching: an instance stepped with another one on the same mesh equals
    its solo run, bit for bit;
  * linearity: boundary records scaled by 2 (exact in FP32, and the
    interpolation weights are exact at dp = sdt) give twice the field;

and, through a fixture the oracle wrote offline in about 45 minutes, the
C2 production mesh's field after 2000 steps against FP64.

The records start just before the incident pulse reaches the interface, so
the fields carry the pulse after a few tens of steps.  One case is the
tilted-axis anisotropic model with layer cuts inside tiles (general stress,
per-node material tiles), the other the isotropic batch model (short
orthotropic stress form)."""
import ctypes as C
import os

import numpy as np
import pytest

import semsks
from semsks import configs, inputs

pytestmark = pytest.mark.gpu

NST = 25


def build(name):
    case = configs.c2_case(nst=NST) if name == "c2" else configs.c3_case(nst=NST)
    b = configs.Built(case)
    pr = b.model.problem
    k0 = max(0, int(1.2 / case.f0 / pr.sdt) - 8)       # pulse peak at t = 1.2/f0 at the first bottom node
    with semsks.Engine() as e:
        psv, sh = e.planewave(*b.planewave_args(k0 + NST + 3, p=0.05, psi=30.0), to_host=True)
    return b, np.ascontiguousarray(psv[k0:]), np.ascontiguousarray(sh[k0:])


def problem(b, psv, sh):
    p = semsks.Problem.from_buffer_copy(b.model.problem)
    p.nst, p.nt, p.dp = NST, psv.shape[0], p.sdt
    p.psv = psv.ctypes.data_as(C.POINTER(C.c_float))
    p.sh = sh.ctypes.data_as(C.POINTER(C.c_float))
    return p


def fields(problems, share=True, **eng):
    with semsks.Engine(max_slots=len(problems), **eng) as e:
        slots = []
        for p in problems:
            slots.append(e.admit(p, share=slots[0] if (share and slots) else -1))
        e.advance(NST - 1)
        return [e.field(s) for s in slots], e.sweep_steps()


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_full_mesh_invariants(name):
    print("case: %s full mesh 2000x500, %d steps: tiling, batch, linearity" % (name, NST), flush=True)
    b, psv, sh = build(name)
    psv2, sh2 = 2.0 * psv, 2.0 * sh
    pa, pb = problem(b, psv, sh), problem(b, psv2, sh2)
    (f1,), _ = fields([pa], tile_z=8)
    peak = np.abs(f1).max()
    assert np.isfinite(f1).all() and peak > 1e-3, peak      # the pulse is in the box
    (f2,), _ = fields([pa], tile_z=32)
    assert np.array_equal(f2.view(np.uint32), f1.view(np.uint32))
    (fa, fb), _ = fields([pa, pb])
    assert np.array_equal(fa.view(np.uint32), f1.view(np.uint32))
    d = fb.astype(np.float64) - 2.0 * f1.astype(np.float64)
    assert np.sqrt((d * d).sum() / (4.0 * (f1.astype(np.float64) ** 2).sum())) <= 1e-7


def build_case(case):
    b = configs.Built(case)
    pr = b.model.problem
    k0 = max(0, int(1.2 / case.f0 / pr.sdt) - 8)
    with semsks.Engine() as e:
        psv, sh = e.planewave(*b.planewave_args(k0 + NST + 3, p=0.05, psi=30.0), to_host=True)
    return b, np.ascontiguousarray(psv[k0:]), np.ascontiguousarray(sh[k0:])


def test_mixed_degrees_in_one_engine_match_solo():
    """Instances of polynomial degree 3, 4 and 5, on meshes with more tiles
    per degree than the persistent grid has waves (so every degree takes its
    tiles from its own per-XCD queues), advance in one engine -- one launch
    per degree and step -- and give the same fields as alone, bit for bit."""
    print("case: hpn 3/4/5 meshes of 900x300/1000x300/600x300 elements, %d steps, mixed vs solo" % NST, flush=True)
    cases = [inputs.Case(nex=900, nez=300, hpn=3, nst=NST, nrec=20, sh_zero=False),
             inputs.Case(nex=1000, nez=300, hpn=4, nst=NST, nrec=20, sh_zero=False),
             inputs.Case(nex=600, nez=300, hpn=5, nst=NST, nrec=20, sh_zero=False)]
    built = [build_case(c) for c in cases]
    probs = [problem(b, psv, sh) for b, psv, sh in built]
    solo = [fields([p], share=False, tile_z=4)[0][0] for p in probs]
    for f in solo:
        assert np.isfinite(f).all() and np.abs(f).max() > 1e-3
    mixed, _ = fields(probs, share=False, tile_z=4)
    for a, b in zip(mixed, solo):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_c2_full_length_stays_finite():
    """The production configuration end to end: one instance of the
    2000x500 mesh with the tilted-axis anisotropic upper mantle, plane SV at
    p = 0.05, all 20000 steps.  Every trace stays finite and bounded, every
    receiver sees the wave, and the transverse component -- zero in the
    incident wave -- is the splitting the anisotropy produces."""
    nst = 20000
    print("case: c2 full mesh 2000x500, %d steps" % nst, flush=True)
    b = configs.Built(configs.c2_case(nst=nst))
    with semsks.Engine(max_slots=1) as e:
        pr = b.problem(nst + 3)
        pr.nst = nst
        e.planewave(*b.planewave_args(nst + 3, p=0.05, psi=0.0), to_host=False)
        s = e.admit(pr)
        e.advance(nst - 1)
        tr = e.traces(s)
    assert np.isfinite(tr).all()
    peak = np.abs(tr).max(axis=(0, 2))
    assert 0.5 < peak[0] < 10.0 and 0.1 < peak[1] < 10.0, peak
    assert (np.abs(tr[:, 0]).max(axis=0) > 1e-3).all()
    assert 1e-4 * peak[0] < peak[2] < 0.1 * peak[0], peak


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_full_mesh_field_vs_fp64(tmp_path, monkeypatch, name):
    """The production mesh against the FP64 oracle: configs.c2_case (2000x500
    elements, tilted-axis upper mantle: the general stress form) and
    configs.c3_case (the batch's isotropic layers: the short orthotropic
    form), coupled solver, a plane S wave 30 degrees off SV, written in the
    reference's file formats, stepped to level 2000; the field as the
    reference's loop leaves it (semsks_get_snapshot) equals the oracle's at
    every 20th node column and row within 1e-4 of each component's peak
    (fixtures tests/golden/<name>_fullsize_fp64_field.npz, by
    make_fullsize_fp64.py)."""
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "%s_fullsize_fp64_field.npz" % name))
    nst, stride, ref = int(z["nst"]), int(z["stride"]), z["field"]
    print("case: %s full mesh 2000x500, %d steps vs FP64 field" % (name, nst - 1), flush=True)
    case = configs.c2_case(nst=nst, psi_deg=30.0, sh_zero=False) if name == "c2" else \
        configs.c3_case(nst=nst, p=0.06, psi_deg=30.0)
    case.write(os.fspath(tmp_path))
    monkeypatch.chdir(tmp_path)
    m = semsks.Model(["par=run.par"])
    with semsks.Engine(max_slots=1) as e:
        s = e.admit(m.problem)
        e.advance(nst - 1)
        got = e.snapshot(s)[:, ::stride, ::stride]
    assert got.shape == ref.shape
    for k in range(3):
        peak = np.abs(ref[k]).max()
        assert peak > 1e-2, (k, peak)
        err = np.abs(got[k] - ref[k]).max()
        assert err <= 1e-4 * peak, (k, err, peak)

