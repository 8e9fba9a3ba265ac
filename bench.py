#!/usr/bin/env python3
"""GLL point-steps/s of the SEM time loop on one MI355X.

Default workload (the configuration the metric is quoted on): the uniform
batch -- 64 incident plane S waves (16 back-azimuths x 4 ray parameters) on
the 2000x500-element mesh (16,010,001 GLL points each), stepped together by
one persistent kernel per time step captured in a hipGraph.  The same run
also times the plain single-instance path on that mesh (the production
anisotropic model) and the faithful CPU port on the host.

value       = sum_instances nglob * steps / wall seconds of K timed steps
              (inputs resident in HBM, barrier + device sync both sides)
roofline    = algorithmic bytes per launch of the fused element-force/update
              kernel / its average launch time (HIP events on the engine's
              stream), against MI355X's 8 TB/s HBM
cpu_baseline= the faithful FP32 port (oracle/sem_oracle.c, -O3, OpenMP node
              loops as the reference) on a bounded sample of the same mesh
serving     = continuous batching (semsks_serve) of Poisson-arriving requests:
              p50/p99 end-to-end latency and throughput (default run: 20/s
              for 5 s of scaled C4-style requests; --config c5: 20/s for
              300 s of full-size ones)
traffic     = measured HBM bytes per launch, from the rocprofv3 summary in
              profiles/ for this workload
precision   = the 200x100-element case's 100 seismograms (P-SV) and the
              150 of the C2 model on that mesh (tilted-axis anisotropy,
              coupled P-SV+SH), FP32 on the GPU against the FP64 oracle's
              (committed fixtures): rel-L2 and peak misfit, worst and
              median, per component

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3|c2|c1b|c5]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "s-sks_amd", "python"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "GLL point-steps/s batched FP32 vs plain; p50/p99 latency per synthetic"


def alg_bytes(n_c: int, share: int) -> float:
    """Compulsory bytes per instance-GLL-point-step of the reference algorithm:
    read u(t), u(t-dt), write u(t+dt) (3 comps, FP32) + n_C constants and
    the inverse mass per node, shared by `share` instances."""
    return 36.0 + 4.0 * (n_c + 1) / share


def profiled_traffic(name: str):
    """HBM bytes per launch of the step kernel for this workload, from the
    rocprofv3 PMC summary kept under profiles/ (request-size counters of the
    L2-to-fabric interface; see that file for the commands)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_engine(waves, steps, warmup, share_material, tile_z=0, extra_steps=0, flags=0):
    """Admits one instance per (built, p, psi) entry (boundary records made by
    the device plane-wave generator), runs warmup + timed steps in hipGraph
    mode.  Returns throughput (pt-steps/s), ms/step and the live engine.
    flags: semsks_problem flags of every instance (F_NO_KAHAN: the plain
    FP32 update)."""
    import semsks
    nt = warmup + steps + extra_steps + 8
    eng = semsks.Engine(max_slots=len(waves), tile_z=tile_z, use_graph=1)
    slots = []
    t0 = time.time()
    for k, (b, p, psi) in enumerate(waves):
        prob = b.problem(nt)
        prob.nst = nt + 1
        prob.flags = flags
        eng.planewave(*b.planewave_args(nt, p=p, psi=psi), to_host=False)
        share = slots[0] if (share_material and slots) else -1
        slots.append(eng.admit(prob, share=share))
    log("  admitted %d instances in %.1fs" % (len(waves), time.time() - t0))
    npts = sum(b.nglob() for b, _, _ in waves)
    eng.advance(warmup, sync=True)
    eng.prepare(steps)              # the graphs the timed call replays, captured outside the timed region
    eng.sync()
    t0 = time.perf_counter()
    eng.advance(steps, sync=True)
    eng.sync()
    wall = time.perf_counter() - t0
    return dict(throughput=npts * steps / wall, ms_per_step=1e3 * wall / steps, npts=npts, engine=eng, slots=slots,
                tile_rows=eng.tile_rows())


def run_serving(rate, duration, nex_range, nez_range, nst_range, budget_gb=200.0, max_slots=1024, seed=None,
                check_sample=True, tile_z=0, order="fcfs", max_active_points=0.0):
    """Continuous batching of Poisson-arriving C4-style requests
    (semsks_serve): HBM-footprint admission against budget_gb, all admitted
    instances stepped together, each retired at its own last level.  Reports
    latency (arrival to traces on the host), admission wait, the offered
    load against what the GPU served, the backlog left when arrivals stop,
    and, for the seeded 5% sample (configs.c5_sample) that has FP64
    fixtures (tests/golden/c5_sample_fp64_traces.npz), each sampled
    request's FP32-vs-FP64 misfit."""
    import semsks
    from semsks import configs
    kw = {} if seed is None else {"seed": seed}
    reqs = configs.c5_requests(rate=rate, duration=duration, nex_range=nex_range, nez_range=nez_range,
                               nst_range=nst_range, **kw)
    t0 = time.time()
    built = [configs.Built(c) for _, c in reqs]
    probs, pws = [], []
    for b in built:
        pr = b.problem(b.case.nst + 3)
        probs.append(pr)
        a = b.planewave_args(pr.nt)
        pws.append((a[7], a[8], a[9], a[10]))
    log("  built %d requests in %.1fs" % (len(reqs), time.time() - t0))
    budget = int(budget_gb * 1e9)
    with semsks.Engine(max_slots=max_slots, hbm_budget=budget, tile_z=tile_z) as e:
        res, st = e.serve(probs, [a for a, _ in reqs], pws, steps_per_round=16, want_traces=True, order=order,
                          max_active_points=max_active_points)
    ok = [r for r in res if r["status"] == 0]
    lat = np.array([r["latency"] for r in ok])
    wait = np.array([r["admit"] - r["arrival"] for r in ok])
    work = np.array([b.nglob() * (b.case.nst - 1) for b in built], np.float64)
    arrivals = np.array([a for a, _ in reqs])
    done = np.array([r["done"] for r in res])
    in_system = (arrivals <= duration) & (done > duration)
    out = {"requests": len(reqs), "completed": st["completed"], "rate_per_s": rate, "arrival_window_s": duration,
           "wall_s": st["wall_s"], "throughput": st["point_steps"] / st["wall_s"],
           "order": order, "max_active_points": max_active_points,
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "mean_latency_s": float(lat.mean()),
           "p50_admission_wait_s": float(np.percentile(wait, 50)),
           "p99_admission_wait_s": float(np.percentile(wait, 99)), "max_active": st["max_active"],
           "offered_point_steps_per_s": float(work.sum() / duration),
           "load_over_served": float(work.sum() / duration / (st["point_steps"] / st["wall_s"])),
           "backlog_at_window_end_requests": int(in_system.sum()),
           "backlog_at_window_end_point_steps": float(work[in_system].sum()),
           "drain_after_window_s": float(done.max() - duration),
           "hbm_budget_gb": budget_gb,
           "mesh_elements": "%dx%d..%dx%d" % (nex_range[0], nez_range[0], nex_range[1], nez_range[1]),
           "steps": "%d..%d" % nst_range}
    if check_sample and seed is None:
        out["fp64_sample"] = serving_sample_check(reqs, res)
    return out


def serving_sample_check(reqs, res):
    """FP32 traces the serving loop returned for the seeded 5% sample against
    the FP64 oracle's (tests/golden/c5_sample_fp64_traces.npz, every
    decim-th level): worst rel-L2 and peak misfit per request and component."""
    from semsks import configs
    path = os.path.join(ROOT, "tests", "golden", "c5_sample_fp64_traces.npz")
    sample = configs.c5_sample(len(reqs))
    if not os.path.exists(path):
        return {"sampled": len(sample), "checked": 0, "note": "no fixture"}
    ref = np.load(path)
    dec = int(ref["decim"])
    rows, worst, quiet = [], 0.0, []
    for i in sample:
        if "r%d_u" % i not in ref.files or res[i]["status"] != 0 or int(ref["r%d_nst" % i]) != reqs[i][1].nst:
            continue
        tr = res[i]["traces"][::dec]
        row = {"request": i}
        for name, k, key in (("u", 0, "u"), ("w", 1, "mw"), ("v", 2, "v")):
            a, b = tr[:, k].astype(np.float64), ref["r%d_%s" % (i, key)].astype(np.float64)
            if np.abs(b).max() < 1e-6:
                # nothing of this component reaches the receivers within the
                # request's steps: the FP64 record is zero; the FP32 peak
                quiet.append({"request": i, "component": name, "fp32_peak": float(np.abs(a).max())})
                continue
            # receivers the wave has not reached by the last step: floor of
            # 1e-3 of the largest receiver's norm / peak
            nb, pb = np.linalg.norm(b, axis=0), np.abs(b).max(axis=0)
            l2 = np.linalg.norm(a - b, axis=0) / np.maximum(nb, 1e-3 * nb.max())
            pk = np.abs(a - b).max(axis=0) / np.maximum(pb, 1e-3 * pb.max())
            row[name] = [float(l2.max()), float(pk.max())]
            worst = max(worst, float(l2.max()), float(pk.max())) if np.isfinite(worst) else worst
            if not (np.isfinite(l2).all() and np.isfinite(pk).all()):
                worst = float("nan")             # reported, not hidden by max()
        rows.append(row)
    return {"sampled": len(sample), "checked": len(rows), "worst": worst, "per_request": rows,
            "no_arrival": quiet,
            "reference": "FP64 oracle (tests/golden/c5_sample_fp64_traces.npz, every %dth level)" % dec}


def precision_vs_fixture(case, fixture, label):
    """FP32 GPU seismograms of a case against the FP64 oracle's, kept in a
    fixture under tests/golden/ (written by the make_*_fp64.py script beside
    it): rel-L2 and peak misfit of every seismogram, worst and median over
    the receivers, per component."""
    import tempfile
    import semsks
    ref = np.load(os.path.join(ROOT, "tests", "golden", fixture))
    old = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        case.write(d)
        os.chdir(d)
        try:
            m = semsks.Model(["par=run.par"])
            with semsks.Engine() as e:
                s = e.admit(m.problem)
                e.advance(m.problem.nst - 1)
                tr = e.traces(s)
        finally:
            os.chdir(old)
    out = {"case": label, "reference": "FP64 oracle (tests/golden/%s)" % fixture, "seismograms": 0}
    for name, k, key in (("u", 0, "u"), ("w", 1, "mw"), ("v", 2, "v")):
        if key not in ref.files:
            out[name] = "zero in both (P-SV solver)" if np.abs(tr[:, k]).max() == 0 else "nonzero"
            continue
        a = tr[:, k].astype(np.float64)
        b = ref[key].astype(np.float64)
        l2 = np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)
        pk = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
        out[name] = {"rel_l2_max": float(l2.max()), "rel_l2_median": float(np.median(l2)),
                     "peak_misfit_max": float(pk.max()), "peak_misfit_median": float(np.median(pk))}
        out["seismograms"] += int(tr.shape[2])
    return out


def precision_c1():
    """The 200x100-element case (4000 steps, 50 receivers, P-SV) against
    tests/golden/c1_fp64_traces.npz (tests/golden/make_c1_fp64.py)."""
    from semsks import configs
    return precision_vs_fixture(configs.c1_case(), "c1_fp64_traces.npz",
                                "200x100 elements, 3 flat isotropic layers, plane SV p=0.06, 4000 steps, "
                                "50 receivers, P-SV")


def precision_aniso():
    """The C2 model on the 200x100-element mesh (tilted-axis hexagonal upper
    mantle, coupled P-SV+SH, 5000 steps, 50 receivers) against
    tests/golden/aniso_fp64_traces.npz (tests/golden/make_aniso_fp64.py)."""
    from semsks import configs
    return precision_vs_fixture(configs.c2_model_c1_mesh(nst=5000), "aniso_fp64_traces.npz",
                                "C2 model (crust, 4% hexagonal upper mantle, tilted axis, mantle) on 200x100 "
                                "elements, coupled P-SV+SH, plane S 30 deg off SV, p=0.05, 5000 steps, 50 receivers")


def precision_aniso20k():
    """The same model at the coupled solver's production length, 20000
    steps, against tests/golden/aniso20k_fp64_traces.npz (every 5th level;
    tests/golden/make_aniso20k_fp64.py)."""
    import tempfile
    import semsks
    from semsks import configs
    ref = np.load(os.path.join(ROOT, "tests", "golden", "aniso20k_fp64_traces.npz"))
    nst, dec = int(ref["nst"]), int(ref["decim"])
    old = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        configs.c2_model_c1_mesh(nst=nst).write(d)
        os.chdir(d)
        try:
            m = semsks.Model(["par=run.par"])
            with semsks.Engine() as e:
                s = e.admit(m.problem)
                e.advance(m.problem.nst - 1)
                tr = e.traces(s)[::dec]
        finally:
            os.chdir(old)
    out = {"case": "C2 model on 200x100 elements, coupled P-SV+SH, %d steps (every %dth level compared), "
                   "50 receivers" % (nst, dec),
           "reference": "FP64 oracle (tests/golden/aniso20k_fp64_traces.npz)", "seismograms": 0}
    for name, k, key in (("u", 0, "u"), ("w", 1, "mw"), ("v", 2, "v")):
        a, b = tr[:, k].astype(np.float64), ref[key].astype(np.float64)
        l2 = np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)
        pk = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
        out[name] = {"rel_l2_max": float(l2.max()), "rel_l2_median": float(np.median(l2)),
                     "peak_misfit_max": float(pk.max()), "peak_misfit_median": float(np.median(pk))}
        out["seismograms"] += int(tr.shape[2])
    return out


def run_c4(steps, warmup, tile_z=0, kernel_steps=10, n=96):
    """The uneven batch: 96 seeded C4 instances (200x100..1600x400 elements,
    1-4 layers, iso or tilted hexagonal up to 6%), one persistent launch per
    step.  Bytes: each instance's own 36 + 4 (n_C + 1) per point-step (no
    material shared), n_C = 15 with a tilted layer, else 9."""
    from semsks import configs
    cases = configs.c4_cases(n)
    t0 = time.time()
    built = [configs.Built(c) for c in cases]
    log("  c4: built %d meshes in %.1fs (%.3g GLL points)" % (len(built), time.time() - t0,
                                                              sum(b.nglob() for b in built)))
    waves = [(b, b.case.p, b.case.psi_deg) for b in built]
    r4 = run_engine(waves, steps, warmup, False, tile_z=tile_z, extra_steps=kernel_steps)
    k4, _ = kernel_time(r4["engine"], kernel_steps)
    r4["engine"].close()
    npts = r4["npts"]
    n_c = [15 if any(a is not None and a[1] is not None for a in (b.case.aniso or [])) else 9 for b in built]
    B = sum((36.0 + 4.0 * (nc + 1)) * b.nglob() for nc, b in zip(n_c, built)) / npts
    return dict(throughput=r4["throughput"], ms_per_step=r4["ms_per_step"], kernel_ms=k4, npts=int(npts),
                tile_rows=r4["tile_rows"], bytes_per_point_step_alg=B,
                achieved_GBs=B * npts / (k4 * 1e-3) / 1e9)


def kernel_time(eng, nsteps):
    """Average device time of the fused step kernel over nsteps eager steps."""
    import semsks
    semsks._check(semsks.lib().semsks_advance(eng._h, nsteps, 2))
    total, kern, n = eng.timing()
    return kern / max(n, 1), total / nsteps


def cpu_threads():
    """Threads of the CPU baseline: OMP_NUM_THREADS when set (the CPU share
    a GPU box gives one GPU), else every core this process may run on."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env or len(os.sched_getaffinity(0))


def cpu_baseline(case_builder, steps: int, warmup: int, kind: str = "f32"):
    """Faithful FP32 port on the host, one instance of the mesh: kind "f32"
    is the -O3 build, "f32_O0" the build with the reference makefile's own
    flags (cc -fopenmp, no -O).  The rate of a window of `steps` steps after
    `warmup` steps, from the oracle's per-step clock."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    c = case_builder()
    n = warmup + steps + 1                       # level 0 is set up, not stepped
    with tempfile.TemporaryDirectory() as d:
        c.nst = n
        c.write(d, nt=n + 7)
        t0 = time.time()
        o = pyoracle.Oracle(["par=run.par"], kind, cwd=d)
        setup = time.time() - t0
        _, clock = o.run_timed(n)
        nglob = o.nx * o.nz
        o.close()
    wall = clock[n - 1] - clock[warmup]
    return dict(value=nglob * steps / wall, steps=steps, warmup=warmup, wall=wall, setup=setup, nglob=nglob)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48,
                    help="timed steps (a multiple of 16 replays only the 16-step captured run)")
    ap.add_argument("--warmup", type=int, default=48,
                    help="untimed steps first (the tile autotuner and the hipGraph capture run here)")
    ap.add_argument("--config", default="c3")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=16, help="timed steps of the CPU baseline (-O3 build)")
    ap.add_argument("--cpu-warmup", type=int, default=10, help="untimed steps before them")
    ap.add_argument("--cpu-steps-o0", type=int, default=2,
                    help="timed steps of the port built with the reference makefile's flags (0 = skip)")
    ap.add_argument("--cpu-warmup-o0", type=int, default=1)
    ap.add_argument("--serve-window", type=float, default=3.0,
                    help="seconds of Poisson arrivals of C5-size requests in the serving leg of the default "
                         "run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0,
                    help="Poisson arrival rate of the serving leg, requests/s (the stated C5 load: 20)")
    ap.add_argument("--serve-order", choices=["fcfs", "shortest"], default="shortest",
                    help="admission order of the serving leg among waiting requests (least work first by "
                         "default; the default run also reports first come first served with no cap)")
    ap.add_argument("--serve-cap", type=float, default=128e6,
                    help="admit only while the active instances hold at most this many GLL points (0: no cap; "
                         "128M keeps the GPU near its batch rate)")
    ap.add_argument("--no-c4", dest="c4", action="store_false", help="skip the uneven-batch (C4) record")
    ap.add_argument("--only", choices=["batched"], default=None,
                    help="c3: run only the batched leg (for profiler runs)")
    ap.add_argument("--no-kahan", action="store_true",
                    help="the batched leg with the plain FP32 update (for profiler runs with --only batched)")
    ap.add_argument("--kernel-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run one device copy of this many MiB (PMC byte-counter calibration)")
    ap.add_argument("--placement-spread", type=int, default=0,
                    help="only time the batched step kernel in this many fresh engines one after another "
                         "(same tiles; the spread comes from where each engine's allocations land)")
    ap.add_argument("--stream-ceiling", type=int, default=0,
                    help="only time a device copy and a triad a = b + 2c over arrays of this many MiB "
                         "(the practical HBM rate for 1:1 and 2:1 read:write traffic), and the step "
                         "kernel's five-stream mix (s-sks_amd/build/semsks_stream_probe)")
    args = ap.parse_args()
    # the CPU baseline's OpenMP threads stay on neighbouring cores; set before
    # anything loads the OpenMP runtime
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
        return
    if args.stream_ceiling:
        import torch
        n = args.stream_ceiling * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b, c = torch.full_like(a, 2.0), torch.full_like(a, 3.0)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        res = {}
        for name, fn, nbytes in (("copy", lambda: b.copy_(a), 8 * n),
                                 ("triad", lambda: torch.add(b, c, alpha=2.0, out=a), 12 * n)):
            best = float("inf")
            for r in range(8):
                ev[0].record()
                fn()
                ev[1].record()
                torch.cuda.synchronize()
                if r >= 2:
                    best = min(best, ev[0].elapsed_time(ev[1]))
            res[name] = {"bytes": nbytes, "best_ms": best, "GB_per_s": nbytes / best * 1e-6}
        # the step kernel's own mix (28 B read, 16 B written per node, five
        # streams) in hand-written grid-stride kernels, over 2^26 nodes
        probe = os.path.join(ROOT, "s-sks_amd", "build", "semsks_stream_probe")
        if not os.path.exists(probe):
            sys.exit("%s is not built (make -C s-sks_amd)" % probe)
        out = subprocess.run([probe], capture_output=True, text=True, check=True, timeout=120).stdout
        res["hip_probe"] = json.loads(out)
        print(json.dumps(res))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only (there is no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.placement_spread:
        import semsks
        from semsks import configs
        b3 = configs.Built(configs.c3_case(nst=10))
        waves = [(b3, p, baz) for p, baz in configs.c3_waves()][: args.batch]
        tz, res = args.tile_z, []
        for i in range(args.placement_spread):
            # the first engine autotunes when no height is given; the rest reuse its pick
            r = run_engine(waves, 4, 4 if tz else 24, True, tile_z=tz, extra_steps=3 * args.kernel_steps,
                           flags=semsks.F_NO_KAHAN if args.no_kahan else 0)
            tz = tz or r["tile_rows"]
            ks = [kernel_time(r["engine"], args.kernel_steps)[0] for _ in range(3)]
            r["engine"].close()
            res.append({"kernel_ms": ks, "tile_rows": tz})
            log("  engine %d: step kernel %s ms" % (i, " ".join("%.3f" % k for k in ks)))
        print(json.dumps({"workload": "batched step kernel, %d plane waves on the 2000x500 mesh, one fresh engine "
                                      "after another" % len(waves), "engines": res}))
        return

    import semsks
    from semsks import configs

    out = {}
    if args.config == "c3":
        # batched: 64 waves on the C2 mesh with the isotropic model, one material
        log("batched: C3, %d instances" % args.batch)
        b3 = configs.Built(configs.c3_case(nst=10))
        waves = [(b3, p, baz) for p, baz in configs.c3_waves()][: args.batch]
        r3 = run_engine(waves, args.steps, args.warmup, True, tile_z=args.tile_z, extra_steps=args.kernel_steps,
                        flags=semsks.F_NO_KAHAN if args.no_kahan else 0)
        k3, _ = kernel_time(r3["engine"], args.kernel_steps)
        r3["engine"].close()
        log("  batched %.3e pt-steps/s, %.3f ms/step, kernel %.3f ms, tiles of %d rows" % (
            r3["throughput"], r3["ms_per_step"], k3, r3["tile_rows"]))
        if args.only == "batched":
            return
        # the same batch with the plain FP32 update (the reference's own
        # arithmetic): no compensation words read or written
        log("batched, plain FP32 update (F_NO_KAHAN)")
        r3f = run_engine(waves, args.steps, args.warmup, True, tile_z=args.tile_z, extra_steps=args.kernel_steps,
                         flags=semsks.F_NO_KAHAN)
        k3f, _ = kernel_time(r3f["engine"], args.kernel_steps)
        r3f["engine"].close()
        log("  batched, plain FP32 %.3e pt-steps/s, %.3f ms/step, kernel %.3f ms, tiles of %d rows" % (
            r3f["throughput"], r3f["ms_per_step"], k3f, r3f["tile_rows"]))
        # plain, like for like: one instance of the same workload (same mesh,
        # model and kernels, batch 1)
        log("plain: 1 instance of the C3 workload")
        rp = run_engine(waves[:1], args.steps, args.warmup, False, tile_z=args.tile_z, extra_steps=args.kernel_steps)
        kp, _ = kernel_time(rp["engine"], args.kernel_steps)
        rp["engine"].close()
        log("  plain %.3e pt-steps/s, %.3f ms/step, kernel %.3f ms, tiles of %d rows" % (
            rp["throughput"], rp["ms_per_step"], kp, rp["tile_rows"]))
        # the production anisotropic model (C2) alone
        log("C2: 1 instance, tilted-axis anisotropic model")
        b2 = configs.Built(configs.c2_case(nst=10))
        r2 = run_engine([(b2, 0.05, 0.0)], args.steps, args.warmup, False, tile_z=args.tile_z,
                        extra_steps=args.kernel_steps)
        k2, _ = kernel_time(r2["engine"], args.kernel_steps)
        r2["engine"].close()
        log("  C2 %.3e pt-steps/s, %.3f ms/step, kernel %.3f ms, tiles of %d rows" % (
            r2["throughput"], r2["ms_per_step"], k2, r2["tile_rows"]))
        # the same two legs on the 200x100 mesh (321,201 points per instance)
        log("plain and batched on the C1 mesh")
        b1 = configs.Built(configs.c1_case(nst=10, solver=0, sh_zero=False))
        r1p = run_engine([(b1, 0.06, 0.0)], args.steps, args.warmup, False, tile_z=args.tile_z)
        r1p["engine"].close()
        w1 = [(b1, p, baz) for p, baz in configs.c3_waves()][: args.batch]
        r1b = run_engine(w1, args.steps, args.warmup, True, tile_z=args.tile_z)
        r1b["engine"].close()
        log("  C1 plain %.3e (tiles of %d rows), batched x%d %.3e pt-steps/s (%d rows)" % (
            r1p["throughput"], r1p["tile_rows"], len(w1), r1b["throughput"], r1b["tile_rows"]))
        # a batch that fills the GPU on this mesh: the 64 waves four times over
        # (ray parameters shifted by 1e-4 s/km per copy)
        w256 = [(b1, p + 1e-4 * (k // len(w1)), baz) for k, (_, p, baz) in enumerate(w1 * 4)]
        r1c = run_engine(w256, args.steps, args.warmup, True, tile_z=args.tile_z)
        r1c["engine"].close()
        log("  C1 batched x%d %.3e pt-steps/s (%d rows)" % (len(w256), r1c["throughput"], r1c["tile_rows"]))
        # and sixteen times over: the batch size at which this mesh's batch
        # rate levels off (64 waves fill about a fifth of the GPU's waves)
        w1024 = [(b1, p + 1e-4 * (k // len(w1)), baz) for k, (_, p, baz) in enumerate(w1 * 16)]
        r1d = run_engine(w1024, args.steps, args.warmup, True, tile_z=args.tile_z)
        r1d["engine"].close()
        log("  C1 batched x%d %.3e pt-steps/s (%d rows)" % (len(w1024), r1d["throughput"], r1d["tile_rows"]))
        B = alg_bytes(9, len(waves))
        achieved = B * r3["npts"] / (k3 * 1e-3) / 1e9
        value, ms, npts, steps = r3["throughput"], r3["ms_per_step"], r3["npts"], args.steps
        hbm_bound_plain = HBM_PEAK_GBS * 1e9 / alg_bytes(9, 1)
        cfg = {"workload": "uniform batch: %d plane S waves (16 baz x 4 p) on a 2000x500-element NGLL=5 mesh, "
                           "16,010,001 GLL points each, shared isotropic 3-layer model" % len(waves),
               "instances": len(waves), "nglob_per_instance": int(b3.nglob()), "mesh": "2000x500 elements",
               "plain_workload": "1 instance of the same workload (same mesh, model and kernels, batch 1)",
               "plain_value": rp["throughput"], "plain_ms_per_step": rp["ms_per_step"],
               "batched_over_plain": r3["throughput"] / rp["throughput"],
               "batched_over_plain_note": "on the 16M-point mesh one instance already fills the GPU: 5x the plain "
                                          "rate would need %.1f TB/s at the batch's 36.6 B/pt, over the 8 TB/s HBM "
                                          "peak; on the 200x100 mesh see c1_mesh_batched_over_plain (64 waves), "
                                          "c1_mesh_batched256_over_plain and c1_mesh_batched1024_over_plain (the batch "
                                          "size at which that mesh's rate levels off)"
                                          % (5 * rp["throughput"] * B / 1e12),
               "plain_hbm_bound": hbm_bound_plain,
               "kernel_ms_batched": k3, "kernel_ms_plain": kp,
               "fp32_plain_batched_value": r3f["throughput"], "fp32_plain_batched_ms_per_step": r3f["ms_per_step"],
               "kernel_ms_fp32_plain_batched": k3f,
               "fp32_plain_batched_frac": B * r3["npts"] / (k3f * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "fp32_plain_note": "the same batch with the plain FP32 update (F_NO_KAHAN, the reference's own "
                                  "arithmetic): no compensation words read or written, 36 instead of 44 B of state "
                                  "per point-step; its 20,000-step worst misfit against FP64 is about 4x the "
                                  "default's (value is the default, compensated update)",
               "tile_rows_batched": r3["tile_rows"], "tile_rows_plain": rp["tile_rows"],
               "c2_workload": "1 instance, 2000x500 mesh, tilted-axis hexagonal 4% upper mantle (15 constants)",
               "c2_value": r2["throughput"], "c2_ms_per_step": r2["ms_per_step"], "kernel_ms_c2": k2,
               "c2_traffic_GBs": (profiled_traffic("step_kernel_plain_2000x500.md") or float("nan")) / (k2 * 1e-3) / 1e9,
               "c1_mesh_plain_value": r1p["throughput"], "c1_mesh_batched_value": r1b["throughput"],
               "c1_mesh_tile_rows_plain": r1p["tile_rows"], "c1_mesh_tile_rows_batched": r1b["tile_rows"],
               "c1_mesh_batched_over_plain": r1b["throughput"] / r1p["throughput"],
               "c1_mesh_workload": "64 plane S waves on the 200x100-element mesh vs one",
               "c1_mesh_plain_note": "one instance's state (three levels, about 16 MB) fits in the 256 MB "
                                     "Infinity Cache: the plain run on this mesh is cache-resident and "
                                     "launch-bound, the batches of 64-1024 waves (1-16 GB) are not",
               "c1_mesh_batched256_value": r1c["throughput"], "c1_mesh_tile_rows_batched256": r1c["tile_rows"],
               "c1_mesh_batched256_over_plain": r1c["throughput"] / r1p["throughput"],
               "c1_mesh_batched1024_value": r1d["throughput"], "c1_mesh_tile_rows_batched1024": r1d["tile_rows"],
               "c1_mesh_batched1024_over_plain": r1d["throughput"] / r1p["throughput"],
               "bytes_per_point_step_alg": B, "kahan": True}
        if args.c4:
            log("c4: uneven batch of 96 instances")
            c4 = run_c4(args.steps, args.warmup, tile_z=args.tile_z, kernel_steps=args.kernel_steps)
            c4["traffic_bytes_per_launch"] = profiled_traffic("step_kernel_uneven96_c4.md")
            if c4["traffic_bytes_per_launch"]:
                # the reference's bytes price per-node constants the step kernel does
                # not read (frac above 1): the measured traffic is the honest roofline
                c4["traffic_GBs"] = c4["traffic_bytes_per_launch"] / (c4["kernel_ms"] * 1e-3) / 1e9
                c4["frac_measured_traffic"] = c4["traffic_GBs"] / HBM_PEAK_GBS
            c4["frac_alg_bytes"] = c4["achieved_GBs"] / HBM_PEAK_GBS
            c4["workload"] = ("uneven batch: 96 instances, 200x100..1600x400 elements, 1-4 layers, iso or tilted "
                              "hexagonal up to 6%, one persistent kernel per step")
            log("  c4 %.3e pt-steps/s, %.3f ms/step, kernel %.3f ms" % (c4["throughput"], c4["ms_per_step"],
                                                                      c4["kernel_ms"]))
            cfg["c4"] = c4
        if args.serve_window > 0:
            log("serving: Poisson %g/s for %.0f s, C5 requests" % (args.serve_rate, args.serve_window))
            sv = run_serving(args.serve_rate, args.serve_window, (200, 1600), (100, 400), (2000, 12000), tile_z=args.tile_z,
                             order=args.serve_order, max_active_points=args.serve_cap)
            log("  %d of %d requests done, p50 %.3f s, p99 %.3f s, %.3e pt-steps/s" % (
                sv["completed"], sv["requests"], sv["p50_latency_s"], sv["p99_latency_s"], sv["throughput"]))
            cfg["serving"] = sv
            if (args.serve_order, args.serve_cap) != ("fcfs", 0.0):
                # the same stream through the plain loop: every arrival admitted
                # as soon as it fits, all admitted instances stepped together
                sf = run_serving(args.serve_rate, args.serve_window, (200, 1600), (100, 400), (2000, 12000),
                                 tile_z=args.tile_z, check_sample=False)
                log("  first come first served, no cap: p50 %.3f s, p99 %.3f s, %.3e pt-steps/s" % (
                    sf["p50_latency_s"], sf["p99_latency_s"], sf["throughput"]))
                cfg["serving_fcfs_no_cap"] = {k: sf[k] for k in (
                    "throughput", "p50_latency_s", "p99_latency_s", "mean_latency_s", "wall_s", "max_active",
                    "completed")}
        log("precision: FP32 GPU vs FP64 oracle fixtures")
        cfg["precision"] = precision_c1()
        cfg["precision_aniso"] = precision_aniso()
        cfg["precision_aniso_20000"] = precision_aniso20k()
        for key in ("precision", "precision_aniso", "precision_aniso_20000"):
            pa = cfg[key]
            log("  %s: u %.2e, w %.2e, v %s worst rel-L2" % (key, pa["u"]["rel_l2_max"], pa["w"]["rel_l2_max"],
                                                           pa["v"] if isinstance(pa["v"], str) else
                                                           "%.2e" % pa["v"]["rel_l2_max"]))
        cpu_case = lambda: configs.c3_case(nst=10)
        traffic = profiled_traffic("step_kernel_batched64_2000x500.md") if len(waves) == 64 else None
    elif args.config == "c2":
        b2 = configs.Built(configs.c2_case(nst=10))
        r2 = run_engine([(b2, 0.05, 0.0)], args.steps, args.warmup, False, tile_z=args.tile_z, extra_steps=args.kernel_steps)
        k2, _ = kernel_time(r2["engine"], args.kernel_steps)
        r2["engine"].close()
        B = alg_bytes(15, 1)
        achieved = B * r2["npts"] / (k2 * 1e-3) / 1e9
        value, ms, npts, steps = r2["throughput"], r2["ms_per_step"], r2["npts"], args.steps
        cfg = {"workload": "1 instance, 2000x500 mesh, tilted-axis hexagonal 4% upper mantle", "instances": 1,
               "kernel_ms": k2, "bytes_per_point_step_alg": B}
        cpu_case = lambda: configs.c2_case(nst=10)
        traffic = profiled_traffic("step_kernel_plain_2000x500.md")
    elif args.config == "c4":
        # uneven batch: 96 instances, 200x100 .. 1600x400 elements, 1-4 layers
        # iso/hexagonal, all stepped by one persistent launch per step
        c4 = run_c4(args.steps, args.warmup, tile_z=args.tile_z, kernel_steps=args.kernel_steps)
        value, ms, npts, steps = c4["throughput"], c4["ms_per_step"], c4["npts"], args.steps
        achieved = c4["achieved_GBs"]
        traffic = profiled_traffic("step_kernel_uneven96_c4.md")
        cfg = {"workload": "uneven batch: 96 instances, 200x100..1600x400 elements, 1-4 layers, iso or "
                           "hexagonal up to 6%, one persistent kernel per step", "instances": 96,
               "nglob_total": c4["npts"], "kernel_ms": c4["kernel_ms"],
               "bytes_per_point_step_alg": c4["bytes_per_point_step_alg"]}
        cpu_case = lambda: configs.c2_case(nst=10)
    elif args.config == "c5":
        # the serving configuration over a bounded arrival window (--serve-window;
        # the stated 300 s of 20 requests/s offers several times what one GPU
        # serves, see load_over_served): C4-size requests, 200 GB budget
        sv = run_serving(args.serve_rate, args.serve_window, (200, 1600), (100, 400), (2000, 12000), tile_z=args.tile_z,
                             order=args.serve_order, max_active_points=args.serve_cap)
        value, ms, steps = sv["throughput"], float("nan"), 0
        achieved, traffic = float("nan"), None
        cfg = {"workload": "serving: Poisson %g requests/s for %.0f s of arrivals, C4-style instances, 200 GB HBM "
                           "budget" % (args.serve_rate, args.serve_window), "serving": sv}
        cpu_case = lambda: configs.c2_case(nst=10)
    elif args.config == "c1b":
        # the literal reading: 64 waves on the C1 mesh (321,201 points each)
        b1 = configs.Built(configs.c1_case(nst=10, solver=0, sh_zero=False))
        waves = [(b1, p, baz) for p, baz in configs.c3_waves()][: args.batch]
        r1 = run_engine([(b1, 0.06, 0.0)], args.steps, args.warmup, False, tile_z=args.tile_z, extra_steps=args.kernel_steps)
        k1, _ = kernel_time(r1["engine"], args.kernel_steps)
        r1["engine"].close()
        rb = run_engine(waves, args.steps, args.warmup, True, tile_z=args.tile_z, extra_steps=args.kernel_steps)
        kb, _ = kernel_time(rb["engine"], args.kernel_steps)
        rb["engine"].close()
        B = alg_bytes(9, len(waves))
        achieved = B * rb["npts"] / (kb * 1e-3) / 1e9
        value, ms, npts, steps = rb["throughput"], rb["ms_per_step"], rb["npts"], args.steps
        cfg = {"workload": "%d plane waves on the 200x100 mesh (321,201 points each)" % len(waves),
               "plain_value": r1["throughput"], "batched_over_plain": rb["throughput"] / r1["throughput"],
               "kernel_ms_batched": kb, "kernel_ms_plain": k1, "bytes_per_point_step_alg": B}
        cpu_case = lambda: configs.c1_case(nst=10)
        traffic = None
    else:
        raise SystemExit("unknown config " + args.config)

    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        os.environ["OMP_NUM_THREADS"] = str(threads)
        log("cpu baseline: %d threads (OMP_PROC_BIND=%s), %d steps after %d warm-up steps" % (
            threads, os.environ.get("OMP_PROC_BIND"), args.cpu_steps, args.cpu_warmup))
        cb = cpu_baseline(cpu_case, args.cpu_steps, args.cpu_warmup)
        env = "OMP_NUM_THREADS=%d (nproc %d, affinity %d), OMP_PROC_BIND=%s" % (
            threads, os.cpu_count() or 0, len(os.sched_getaffinity(0)), os.environ.get("OMP_PROC_BIND"))
        cpu = {"value": cb["value"], "unit": "GLL point-steps/s", "cores": threads, "kind": "port",
               "sample": "faithful FP32 port of sem2d_psvsh (oracle/sem_oracle.c, -O3 -march=x86-64-v3, OpenMP "
                         "on the node loops, element loop serial as in every reference variant), 1 instance of "
                         "the %d-point mesh of the batch's model, %d steps timed after %d warm-up steps (%.1f s); %s"
                         % (cb["nglob"], cb["steps"], cb["warmup"], cb["wall"], env)}
        log("  cpu %.3e pt-steps/s" % cb["value"])
        if args.cpu_steps_o0 > 0:
            c0 = cpu_baseline(cpu_case, args.cpu_steps_o0, args.cpu_warmup_o0, kind="f32_O0")
            cpu["makefile_flags"] = {"value": c0["value"], "cores": threads,
                                     "sample": "the same port built with the reference makefile's flags "
                                               "(cc -fopenmp, no -O), %d steps after %d warm-up steps (%.1f s)"
                                               % (c0["steps"], c0["warmup"], c0["wall"])}
            log("  cpu, makefile flags: %.3e pt-steps/s" % c0["value"])

    line = {
        "metric": METRIC, "value": value, "unit": "GLL point-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (closed-form plane S waves, seeded models)",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["config"]["vs_cpu"] = value / cpu["value"]
    if "serving" in cfg:
        line["latency_p50_s"] = cfg["serving"]["p50_latency_s"]
        line["latency_p99_s"] = cfg["serving"]["p99_latency_s"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
