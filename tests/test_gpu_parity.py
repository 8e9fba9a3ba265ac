"""HIP path vs the CPU oracle on the same inputs (the parity tests proper).

Bars:
  * setup products that are integer or FP64 work on both sides (material per
    node, receiver indices, FP64 lumped mass, Lysmer coefficients, the FP32
    interpolated boundary series): bit-exact;
  * FP32 element force K u: max |f - f64| <= 2e-6 * max |f64|;
  * seismograms (FP32 time loop) vs the FP64 oracle: relative L2 <= 1e-4 and
    peak misfit <= 1e-4 of the peak, per component with signal.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import semsks
from semsks import configs, inputs

pytestmark = pytest.mark.gpu

REL_L2_TOL = 1e-4
PEAK_TOL = 1e-4


def rel_l2(a, b, floor=0.0):
    """||a-b|| / max(||b||, floor); floor = 1e-3 of the largest component's
    norm keeps a component with no signal yet from dividing by ~0."""
    nb = max(np.linalg.norm(b), floor)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def peak_misfit(a, b, floor=0.0):
    pb = max(np.abs(b).max(), floor)
    return float(np.abs(a - b).max() / pb) if pb > 0 else float(np.abs(a).max())


def misfits(tr, ref):
    """per component (rel-L2, peak) against the FP64 traces"""
    n_ref = max(np.linalg.norm(ref[:, k]) for k in range(3))
    p_ref = max(np.abs(ref[:, k]).max() for k in range(3))
    return [(rel_l2(tr[:, k], ref[:, k], 1e-3 * n_ref), peak_misfit(tr[:, k], ref[:, k], 1e-3 * p_ref))
            for k in range(3)]


def make_case(tmp_path, **kw):
    c = inputs.small_case(**kw)
    d = os.fspath(tmp_path)
    c.write(d)
    return d


@pytest.fixture
def small(tmp_path, monkeypatch):
    d = make_case(tmp_path)
    monkeypatch.chdir(d)
    return d


def test_setup_bit_exact(small):
    print("case: small 24x12, setup parity", flush=True)
    m = semsks.Model(["par=run.par"])
    o64 = pyoracle.Oracle(["par=run.par"], "f64")
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        mass, mdt2, ab = e.setup(s)
    om, omd, oab = o64.mass()
    assert np.array_equal(mass, om)
    assert np.array_equal(mdt2, omd.astype(np.float32))
    assert np.array_equal(ab, oab)


def test_boundary_series_bit_exact(small):
    print("case: small, boundary series", flush=True)
    m = semsks.Model(["par=run.par"])
    o32 = pyoracle.Oracle(["par=run.par"], "f32")
    n = 120
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        g, gv = e.boundary_series(s, n)
    og, ogv = o32.boundary_series(n)
    assert np.array_equal(g.view(np.uint32), og.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), ogv.view(np.uint32))
    assert np.abs(g).max() > 0.1


@pytest.mark.parametrize("seed", [1, 2])
def test_element_force(small, seed):
    print("case: small, K u seed", seed, flush=True)
    m = semsks.Model(["par=run.par"])
    o64 = pyoracle.Oracle(["par=run.par"], "f64")
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, m.nz, m.nx)).astype(np.float32)
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        f = e.stiff(s, u)
    f64 = o64.stiff(u.astype(np.float64))
    err = np.abs(f - f64).max() / np.abs(f64).max()
    assert err < 2e-6, err


def run_gpu(problem, **eng):
    with semsks.Engine(**eng) as e:
        s = e.admit(problem)
        e.advance(problem.nst - 1)
        return e.traces(s)


@pytest.mark.parametrize("solver", [0, 1, 2])
def test_seismograms_vs_fp64(tmp_path, monkeypatch, solver):
    print("case: small, seismograms, solver", solver, flush=True)
    d = make_case(tmp_path, solver=solver, sh_zero=(solver != 0), psi_deg=30.0 if solver == 0 else 0.0)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    tr = run_gpu(m.problem)
    ref = pyoracle.Oracle(["par=run.par"], "f64").run()
    assert tr.shape == ref.shape
    for k, (l2, pk) in enumerate(misfits(tr, ref)):
        if np.abs(ref[:, k]).max() == 0:
            assert np.abs(tr[:, k]).max() == 0
            continue
        assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, l2, pk)


def test_cli_dropin(small):
    print("case: small, CLI", flush=True)
    os.makedirs("gpu", exist_ok=True)
    os.makedirs("cpu", exist_ok=True)
    r = subprocess.run([semsks.CLI_PATH, "par=run.par", "outdir=gpu", "check=1"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    # check=1: the setup diagnostics go to outdir before the solve
    m = semsks.Model(["par=run.par"])
    assert sum(1 for _ in open("gpu/vel.xy")) == m.nx * m.nz
    assert sum(1 for _ in open("gpu/legendre.xz")) == 5 * 41
    assert sum(1 for _ in open("gpu/grid.xz")) == len(range(0, m.nx, 2)) * (len(range(0, m.nz, 2)) + 1)
    assert pyoracle.main(["par=run.par"], "cpu", kind="f64") == 0
    a = np.stack([np.fromfile("gpu/seis." + c, np.float32) for c in ("u", "w", "v")], axis=1)
    b = np.stack([np.fromfile("cpu/seis." + c, np.float32) for c in ("u", "w", "v")], axis=1)
    assert a.shape == b.shape == (600 * 8, 3)
    for k, (l2, pk) in enumerate(misfits(a, b)):
        assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, l2, pk)


@pytest.mark.parametrize("solver", [0, 2])
def test_snapshots_vs_fp64(tmp_path, monkeypatch, solver):
    """semsks_get_snapshot at a level is the FP64 oracle's field at the end
    of that iteration (before the next step's interface injection, which
    the field of the engine already holds; GRT_Interface4 for solver 0,
    GRT_Interface3 for solver 2); the CLI with plot_snap=1 writes it as
    out<k> (sem2d_psvsh.c:483-510,1699-1715)."""
    print("case: small, snapshots, solver", solver, flush=True)
    d = make_case(tmp_path, solver=solver, sh_zero=(solver != 0), psi_deg=30.0 if solver == 0 else 0.0)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    p = m.problem
    o64 = pyoracle.Oracle(["par=run.par"], "f64")
    nbl = p.neh * p.hpn
    iz, ix = np.meshgrid(np.arange(m.nz), np.arange(m.nx), indexing="ij")
    iface = ((iz == nbl) & (ix >= nbl)) | ((ix == nbl) & (iz > nbl))
    pending = 0.0
    snaps = {}
    with semsks.Engine() as e:
        s = e.admit(p)
        lev = 0
        for L in (0, 150, 300, p.nst - 1):
            e.advance(L - lev)
            lev = L
            snap, field = e.snapshot(s), e.field(s)
            snaps[L] = snap
            _, ref = o64.run(L + 1, want_field=True)
            assert np.array_equal(snap[:, ~iface], field[:, ~iface])
            scale = np.abs(ref).max()
            if scale == 0:                    # level 0 before the pulse arrives
                assert np.abs(snap).max() == 0
                continue
            err = np.abs(snap - ref).max()
            assert err <= 1e-4 * scale, (L, err, scale)
            pending = max(pending, np.abs(field - ref)[:, iface].max() / scale)
    # at some level the injection the field holds is far above the bar:
    # restoring the interface nodes is what makes the snapshot match
    assert pending > 1e-2, pending
    if solver != 0:
        return
    os.makedirs("gpu", exist_ok=True)
    r = subprocess.run([semsks.CLI_PATH, "par=run.par", "outdir=gpu", "plot_snap=1", "startsnap=0", "incresnap=150"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert [os.path.exists("gpu/out%d" % k) for k in range(1, 6)] == [True, True, True, True, False]
    rows = np.loadtxt("gpu/out2")
    sel = (slice(None, None, 2), slice(None, None, 2))
    want = np.stack([snaps[150][0][sel].T.ravel(), snaps[150][1][sel].T.ravel()], axis=1)
    assert np.allclose(rows[:, 2:], want, rtol=5.1e-7, atol=1e-30)


def test_device_planewave_matches_the_greenfile_writer():
    """semsks_planewave, the device generator of the batch and serving paths,
    against inputs.greenfile_records, which writes the greenfile of every
    file-based case and FP64 fixture: the same closed-form plane wave in
    FP64 at the interface nodes, rounded to float, in the greenfile layout
    (u, w down along the bottom row, then the left column; v in the SH
    file).  Within a few ulps of the peak (the device's and numpy's exp and
    cos may differ in the last bit)."""
    print("case: small, device plane-wave records vs the greenfile writer", flush=True)
    c = inputs.small_case(sh_zero=False, psi_deg=30.0)
    b = configs.Built(c)
    D, nt = b.D, 400
    with semsks.Engine() as e:
        psv, sh = e.planewave(*b.planewave_args(nt), to_host=True)
    pw = inputs.PlaneWave(p=c.p, beta=D["vs"][-1], f0=c.f0, psi_deg=c.psi_deg)
    ref_psv, ref_sh = inputs.greenfile_records(pw, c.nex, D["nez"], D["eh"], c.hpn, c.neh, nt,
                                               float(np.float32(D["sdt"])))
    assert psv.shape == ref_psv.shape and sh.shape == ref_sh.shape
    scale = np.abs(ref_psv).max()
    assert scale > 0.5 and np.abs(ref_sh).max() > 0.2
    assert np.abs(psv - ref_psv).max() <= 4e-7 * scale
    assert np.abs(sh - ref_sh).max() <= 4e-7 * scale


def test_batch_members_match_solo(tmp_path, monkeypatch):
    """Instances stepped together in one launch give the same traces as alone."""
    print("case: small, batch of 3 vs solo", flush=True)
    probs = []
    for k, (p, psi) in enumerate([(0.04, 0.0), (0.06, 30.0), (0.07, 75.0)]):
        d = make_case(tmp_path / ("c%d" % k), p=p, psi_deg=psi, sh_zero=False)
        monkeypatch.chdir(d)
        probs.append(semsks.Model(["par=run.par"]))
    solo = [run_gpu(m.problem) for m in probs]
    with semsks.Engine(max_slots=3) as e:
        slots = [e.admit(probs[0].problem)]
        slots += [e.admit(m.problem, share=slots[0]) for m in probs[1:]]
        e.advance(probs[0].problem.nst - 1)
        batch = [e.traces(s) for s in slots]
    for a, b in zip(batch, solo):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_dipping_interface_and_tilted_axis(tmp_path, monkeypatch):
    """Non-flat interfaces (per-node material path) and a tilted-axis layer
    (15 constants) against the FP64 oracle."""
    print("case: small, dipping interface + tilted axis", flush=True)
    from semsks import aniso
    C = aniso.hexagonal_tilted(vp=7.8, vs=4.4, rho=3.3, strength=0.04, dip_deg=30.0, azimuth_deg=60.0)
    xmax = 24 * 2.0 + 2.0
    interfaces = [([0.0, xmax], [14.0, 18.0]), None, None]
    d = make_case(tmp_path, interfaces=interfaces, aniso=[C, None, None], psi_deg=20.0, sh_zero=False)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    tr = run_gpu(m.problem)
    o = pyoracle.Oracle(["par=run.par"], "f64")
    ref = o.run()
    for k, (l2, pk) in enumerate(misfits(tr, ref)):
        assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, l2, pk)


def test_tile_height_and_autotuner_leave_results_unchanged(tmp_path, monkeypatch):
    """The element sweep's arithmetic does not depend on how the mesh is cut
    into tiles, so fixed tile heights and the autotuned one (which runs its
    trial steps as ordinary steps) give bit-identical traces."""
    print("case: small, tile heights 3 / 16 / autotuned", flush=True)
    d = make_case(tmp_path, psi_deg=30.0, sh_zero=False, interfaces=[([0.0, 50.0], [14.0, 18.0]), None, None])
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    runs = [run_gpu(m.problem, tile_z=tz) for tz in (3, 16, 0)]
    for r in runs[1:]:
        assert np.array_equal(r.view(np.uint32), runs[0].view(np.uint32))
    # the engine reports the height it sweeps with: the fixed one, or one of
    # the autotuner's candidates (one per band of 1 .. 64 rows, none taller
    # than the mesh) once a call long enough to tune has run
    with semsks.Engine(tile_z=3) as e:
        e.admit(m.problem)
        assert e.tile_rows() == 3
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        assert 1 <= e.tile_rows() <= min(64, m.problem.nez)
        assert np.array_equal(e.traces(s).view(np.uint32), runs[0].view(np.uint32))


def test_serving_results_equal_alone_and_in_the_mix(tmp_path, monkeypatch):
    """Continuous batching (semsks_serve): requests of different mesh size,
    model and length arrive while others run; each one's traces are the
    same, bit for bit, as when it runs alone."""
    print("case: serving, 6 requests vs solo", flush=True)
    specs = [dict(nex=24, nst=300, p=0.06), dict(nex=36, nst=200, p=0.05, psi_deg=40.0, sh_zero=False),
             dict(nex=20, nst=450, p=0.07, solver=1), dict(nex=30, nst=120, p=0.045),
             dict(nex=24, nst=260, p=0.055, psi_deg=70.0, sh_zero=False),
             dict(nex=28, nst=180, p=0.065, interfaces=[([0.0, 60.0], [14.0, 18.0]), None, None])]
    models = []
    for k, kw in enumerate(specs):
        d = make_case(tmp_path / ("r%d" % k), **kw)
        monkeypatch.chdir(d)
        models.append(semsks.Model(["par=run.par"]))
    solo = [run_gpu(m.problem) for m in models]
    arrivals = [0.0, 0.0, 0.02, 0.05, 0.05, 0.3]
    with semsks.Engine(max_slots=3) as e:          # fewer slots than requests: some must wait
        res, stats = e.serve([m.problem for m in models], arrivals, steps_per_round=16)
    assert stats["completed"] == len(models)
    for r, s in zip(res, solo):
        assert r["status"] == 0 and r["done"] >= r["admit"] >= r["arrival"]
        assert np.array_equal(r["traces"].view(np.uint32), s.view(np.uint32))
    # a sample of the served requests against the FP64 oracle (ones long
    # enough for the incident wave, peak amplitude 1, to reach the receivers)
    for k in (0, 2):
        monkeypatch.chdir(os.fspath(tmp_path / ("r%d" % k)))
        ref = pyoracle.Oracle(["par=run.par"], "f64").run()
        assert np.abs(ref).max() > 0.1
        for c, (l2, pk) in enumerate(misfits(res[k]["traces"], ref)):
            if np.abs(ref[:, c]).max() == 0:
                continue
            assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, c, l2, pk)


def test_serving_shortest_first_under_a_point_cap(tmp_path, monkeypatch):
    """semsks_serve_ex with least-work-first admission and a cap on the
    active GLL points: requests that wait are admitted smallest work first,
    the active instances never hold more than the cap (an idle engine admits
    whatever is next), and every result equals its run alone, bit for bit."""
    print("case: serving, least work first, point cap", flush=True)
    specs = [dict(nex=36, nst=300, p=0.06), dict(nex=24, nst=120, p=0.05, psi_deg=40.0, sh_zero=False),
             dict(nex=30, nst=200, p=0.07, solver=1), dict(nex=20, nst=90, p=0.045),
             dict(nex=28, nst=260, p=0.055, psi_deg=70.0, sh_zero=False)]
    models = []
    for k, kw in enumerate(specs):
        d = make_case(tmp_path / ("r%d" % k), **kw)
        monkeypatch.chdir(d)
        models.append(semsks.Model(["par=run.par"]))
    probs = [m.problem for m in models]
    solo = [run_gpu(p) for p in probs]
    pts = [(p.nex * p.hpn + 1) * (p.nez * p.hpn + 1) for p in probs]
    work = [n * (p.nst - 1) for n, p in zip(pts, probs)]
    cap = max(pts) + min(pts)          # room for the largest and one more at most
    with semsks.Engine(max_slots=8) as e:
        res, stats = e.serve(probs, [0.0] * len(probs), steps_per_round=16, order="shortest", max_active_points=cap)
    assert stats["completed"] == len(probs)
    for r, s in zip(res, solo):
        assert r["status"] == 0 and r["done"] >= r["admit"] >= r["arrival"]
        assert np.array_equal(r["traces"].view(np.uint32), s.view(np.uint32))
    # admitted in order of work (all arrived at once)
    by_admit = sorted(range(len(probs)), key=lambda i: res[i]["admit"])
    assert by_admit == sorted(range(len(probs)), key=lambda i: (work[i], i))
    # at every admission the active points stay within the cap, unless the
    # request was admitted into an idle engine
    for i in range(len(probs)):
        t = res[i]["admit"]
        active = [j for j in range(len(probs)) if res[j]["admit"] <= t < res[j]["done"]]
        assert sum(pts[j] for j in active) <= cap or active == [i], (i, active)


def test_serving_waits_for_memory_and_rejects_what_never_fits():
    """Admission by HBM footprint: with a budget for two of the small
    requests, the third waits until one retires; a request larger than the
    whole budget comes back as SEMSKS_ENOHBM once the engine is idle,
    without holding up the request behind it.  Every request generates its
    boundary records on the device, and each served result equals its run
    alone, bit for bit."""
    from semsks import configs
    print("case: serving under a 2-request HBM budget, plane-wave records on the device", flush=True)
    cases = [inputs.small_case(nex=40, nst=400),
             inputs.small_case(nex=44, nst=320, p=0.05, psi_deg=35.0, sh_zero=False),
             inputs.small_case(nex=42, nst=450, p=0.07, solver=1),
             inputs.small_case(nex=600, nst=50),                               # larger than the budget
             inputs.small_case(nex=40, nst=360, p=0.045, psi_deg=60.0, sh_zero=False)]
    built = [configs.Built(c) for c in cases]
    probs = [b.problem(b.case.nst + 3) for b in built]
    pws = [b.planewave_args(p.nt)[7:11] for b, p in zip(built, probs)]
    small_fp = [semsks.footprint(p) for k, p in enumerate(probs) if k != 3]
    scratch = max(semsks.admission_bytes(p, True) for k, p in enumerate(probs) if k != 3)
    budget = 2 * max(small_fp) + min(small_fp) // 2 + scratch     # the engine's admission scratch on top
    assert semsks.footprint(probs[3]) > budget
    arrivals = [0.0, 0.0, 0.0, 0.001, 0.002]
    with semsks.Engine(max_slots=8, hbm_budget=budget) as e:
        res, stats = e.serve(probs, arrivals, pws, steps_per_round=16)
    assert res[3]["status"] == 3 and res[3]["traces"] is not None
    assert stats["completed"] == 4 and stats["max_active"] == 2
    assert res[2]["admit"] > res[2]["arrival"]                 # waited for memory
    assert res[4]["admit"] >= res[3]["done"]                   # admitted once the large one was turned away
    for k in (0, 1, 2, 4):
        with semsks.Engine(max_slots=1) as e:
            (solo,), _ = e.serve([probs[k]], [0.0], [pws[k]], steps_per_round=16)
        assert res[k]["status"] == 0 and solo["status"] == 0
        assert np.abs(solo["traces"]).max() > 0
        assert np.array_equal(res[k]["traces"].view(np.uint32), solo["traces"].view(np.uint32)), k


@pytest.mark.parametrize("atten", [1, 4])
def test_sponge_strips_vs_fp64(small, atten):
    """atten_wave2's exponential taper on the edge strips (bottom only, and
    all four edges) against the FP64 oracle."""
    print("case: small, atten", atten, flush=True)
    args = ["par=run.par", "atten=%d" % atten, "ass=0.9"]
    m = semsks.Model(args)
    assert m.problem.atten == atten
    tr = run_gpu(m.problem)
    ref = pyoracle.Oracle(args, "f64").run()
    base = pyoracle.Oracle(["par=run.par"], "f64").run()
    assert np.abs(ref - base).max() > 1e-3 * np.abs(base).max()       # the sponge changes the traces
    for k, (l2, pk) in enumerate(misfits(tr, ref)):
        if np.abs(ref[:, k]).max() == 0:
            continue
        assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, l2, pk)


@pytest.mark.parametrize("hpn", [3, 5, 6])
def test_other_polynomial_degrees_vs_fp64(tmp_path, monkeypatch, hpn):
    """hpn = 3, 5, 6 (NGLL 4, 6, 7; lagrange.c:327-330 allows 3..6): K u and
    the seismograms against the FP64 oracle, over a dipping interface."""
    print("case: small, hpn", hpn, flush=True)
    kw = dict(psi_deg=25.0, sh_zero=False, interfaces=[([0.0, 50.0], [14.0, 17.0]), None, None])
    T = 500 * inputs.small_case(**kw).derived()["sdt"]          # the same span of time for every degree
    nst = int(T / inputs.small_case(hpn=hpn, **kw).derived()["sdt"])
    d = make_case(tmp_path, hpn=hpn, nst=nst, **kw)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    assert m.problem.hpn == hpn
    o64 = pyoracle.Oracle(["par=run.par"], "f64")
    u = np.random.default_rng(hpn).standard_normal((3, m.nz, m.nx)).astype(np.float32)
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        f = e.stiff(s, u)
    f64 = o64.stiff(u.astype(np.float64))
    assert np.abs(f - f64).max() / np.abs(f64).max() < 2e-6
    tr = run_gpu(m.problem)
    ref = o64.run()
    for k, (l2, pk) in enumerate(misfits(tr, ref)):
        if np.abs(ref[:, k]).max() == 0:
            continue
        assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, l2, pk)


def test_splitting_of_fp32_traces_matches_fp64(tmp_path, monkeypatch):
    """The FP32 traces keep the anisotropic physics: the Silver & Chan
    transverse-energy map over (fast azimuth, delay) of the GPU seismogram
    matches the FP64 oracle's, and so does its minimum."""
    from semsks import aniso, splitting
    print("case: splitting, tilted hexagonal 8% bottom layer", flush=True)
    C = aniso.hexagonal_tilted(vp=7.8, vs=4.4, rho=3.3, strength=0.08, dip_deg=20.0, azimuth_deg=40.0)
    d = make_case(tmp_path, nex=30, nst=900, aniso=[None, None, C], psi_deg=0.0, sh_zero=False, p=0.04, f0=0.6)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    gpu = run_gpu(m.problem)
    o = pyoracle.Oracle(["par=run.par"], "f64")
    ref = o.run()
    r = m.problem.ntrace // 2
    dt = o.sdt
    res = []
    for tr in (gpu, ref):
        u, v = tr[:, 0, r].astype(np.float64), tr[:, 2, r].astype(np.float64)
        ipk = int(np.argmax(np.hypot(u, v)))
        win = (max(0, ipk - int(3 / dt)), min(len(u), ipk + int(5 / dt)))
        res.append(splitting.silver_chan(v, u, dt, 270.0, win, delays=np.arange(0, 1.0001, 0.02)))
    a, b = res
    assert b["delay"] > 0                                       # the layer splits the wave
    assert np.abs(a["E"] - b["E"]).max() <= 1e-3 * b["E"].max()
    assert abs(a["phi"] - b["phi"]) <= 1.0 and abs(a["delay"] - b["delay"]) <= 0.02 + 1e-9


def test_layered_boundary_records_vs_fp64(tmp_path, monkeypatch):
    """Boundary records from the layered-medium producer (free-surface and
    interface reverberations in the greenfile) through the GPU solver,
    against the FP64 oracle on the same files."""
    print("case: small, layered boundary records", flush=True)
    d = make_case(tmp_path, boundary="layered", psi_deg=35.0, sh_zero=False, p=0.05)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    tr = run_gpu(m.problem)
    ref = pyoracle.Oracle(["par=run.par"], "f64").run()
    assert np.abs(ref).max() > 0.1
    for k, (l2, pk) in enumerate(misfits(tr, ref)):
        assert l2 <= REL_L2_TOL and pk <= PEAK_TOL, (k, l2, pk)


def test_unstable_run_is_reported(tmp_path, monkeypatch):
    """A time step far above the stability limit blows the field up; the
    traces come back with SEMSKS_ENONFINITE instead of silently."""
    print("case: small, sdt x 4 (unstable)", flush=True)
    c = inputs.small_case(nst=400)
    c.sdt = 4.0 * c.derived()["sdt"]
    d = os.fspath(tmp_path)
    c.write(d)
    monkeypatch.chdir(d)
    m = semsks.Model(["par=run.par"])
    with semsks.Engine() as e:
        s = e.admit(m.problem)
        e.advance(m.problem.nst - 1)
        with pytest.raises(semsks.SemsksError) as ei:
            e.traces(s)
        assert ei.value.code == 4
