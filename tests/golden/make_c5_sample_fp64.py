"""Writes c5_sample_fp64_traces.npz: the FP64 oracle's seismograms of the
seeded 5% sample (configs.c5_sample) of the C5 requests that arrive in the
first WINDOW seconds of the seeded Poisson stream (configs.c5_requests, 20
requests/s): full length, every DECIM-th level, u, -w and v at the 50
surface receivers, in float32.

The oracle runs its colour-parallel element force (orc_run_fast: the
arithmetic of the reference-order sweep, summed in another order at shared
nodes, which tests/test_oracle_pins.py holds within 1e-12 of it).  A 0.2 Hz
Ricker pulse has no energy near the Nyquist frequency of the decimated
series (1/(2*DECIM*sdt) >= 1.4 Hz for these meshes), so the decimated traces
measure the same misfit as the full ones.

The serving leg of bench.py and tests/test_gpu_serving.py compare the FP32
traces each sampled request gets from the continuous-batching loop against
these.

  python tests/golden/make_c5_sample_fp64.py     (about two hours on 8 cores)
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "s-sks_amd", "python"))
import pyoracle  # noqa: E402
from semsks import configs  # noqa: E402

WINDOW = 3.0
DECIM = 20


def main():
    reqs = configs.c5_requests(duration=WINDOW)
    sample = configs.c5_sample(len(reqs))
    out = {"window_s": np.float64(WINDOW), "decim": np.int32(DECIM), "indices": np.array(sample, np.int32)}
    path = os.path.join(HERE, "c5_sample_fp64_traces.npz")
    if os.path.exists(path):            # resume: keep the requests an earlier run finished
        with np.load(path) as old:
            if float(old["window_s"]) == WINDOW and int(old["decim"]) == DECIM and \
                    list(old["indices"]) == list(sample):
                out.update({k: old[k] for k in old.files if k.startswith("r")})
    for i in sample:
        if "r%d_nst" % i in out:
            continue
        t0 = time.time()
        case = reqs[i][1]
        with tempfile.TemporaryDirectory() as d:
            case.write(d)
            o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
            tr = o.run(fast=True)[::DECIM]
            o.close()
        out["r%d_u" % i] = tr[:, 0].astype(np.float32)
        out["r%d_mw" % i] = tr[:, 1].astype(np.float32)
        out["r%d_v" % i] = tr[:, 2].astype(np.float32)
        out["r%d_nst" % i] = np.int32(case.nst)
        print("request %d: %dx%d elements, %d steps, %.0f s" % (i, case.nex, case.nez, case.nst, time.time() - t0),
              flush=True)
        np.savez_compressed(path, **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
