"""Writes aniso20k_fp64_traces.npz: the seismograms of the FP64 oracle
(oracle/sem_oracle.c built with REAL=double) for configs.c2_model_c1_mesh()
at the production length of the coupled solver's loop -- 20000 steps
(sem2d_psvsh.c:418-512 run for nst = 20000) -- on the 200x100-element
mesh: crust, 4% hexagonal upper mantle with the tilted C2 axis, mantle,
coupled P-SV+SH, plane S wave polarised 30 degrees off SV, 50 surface
receivers.

Stored in float32, in the order the solver writes them (u, -w, v), every
DECIM-th time level (levels 0, DECIM, 2*DECIM, ...): a 0.2 Hz Ricker pulse
holds no energy near the Nyquist frequency of the decimated series
(1/(2*DECIM*sdt) = 6.5 Hz), so the rel-L2 of the decimated traces measures
the same misfit as the full ones, at a fifth of the fixture's size.

tests/test_gpu_aniso.py holds all 150 FP32 GPU seismograms of this run to
these within 1e-4 (the drift the compensated FP32 update has to control
over 20000 steps).

  python tests/golden/make_aniso20k_fp64.py      (about 20 minutes on 8 cores)
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

NST = 20000
DECIM = 5


def aniso20k_oracle_traces(nst=NST):
    case = configs.c2_model_c1_mesh(nst=nst)
    with tempfile.TemporaryDirectory() as d:
        case.write(d)
        tr = pyoracle.Oracle(["par=run.par"], "f64", cwd=d).run()
    return tr                                  # [nst][3][ntrace]: u, -w, v


if __name__ == "__main__":
    tr = aniso20k_oracle_traces()[::DECIM]
    np.savez_compressed(os.path.join(HERE, "aniso20k_fp64_traces.npz"), u=tr[:, 0].astype(np.float32),
                        mw=tr[:, 1].astype(np.float32), v=tr[:, 2].astype(np.float32), nst=np.int32(NST),
                        decim=np.int32(DECIM))
    print("wrote", tr.shape, np.abs(tr).max(axis=(0, 2)))
