"""Writes c1_fp64_traces.npz: the seismograms of the FP64 oracle
(oracle/sem_oracle.c built with REAL=double) for the 200x100-element case
of configs.c1_case() -- three flat isotropic layers, plane SV at p = 0.06,
Ricker 0.2 Hz, 4000 steps, 50 surface receivers, P-SV solver -- in the
order the solver writes them (u, -w; v is zero for this solver) and in
float32, as seis.u/seis.w hold them.

bench.py reports the FP32 GPU traces' rel-L2 and peak misfit against these
for every seismogram, and tests/test_gpu_c1.py holds them to 1e-4.  The
oracle that writes them is pinned by tests/test_oracle_pins.py; running
this script again rewrites the file bit for bit.

  python tests/golden/make_c1_fp64.py          (a few minutes on 8 cores)
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "s-sks_amd", "python"))
import pyoracle  # noqa: E402
from semsks import configs  # noqa: E402


def c1_oracle_traces(nst=None):
    case = configs.c1_case() if nst is None else configs.c1_case(nst=nst)
    with tempfile.TemporaryDirectory() as d:
        case.write(d)
        tr = pyoracle.Oracle(["par=run.par"], "f64", cwd=d).run()
    return tr                                  # [nst][3][ntrace]: u, -w, v


if __name__ == "__main__":
    tr = c1_oracle_traces()
    assert np.abs(tr[:, 2]).max() == 0.0
    np.savez_compressed(os.path.join(HERE, "c1_fp64_traces.npz"), u=tr[:, 0].astype(np.float32),
                        mw=tr[:, 1].astype(np.float32))
    print("wrote", tr.shape)
