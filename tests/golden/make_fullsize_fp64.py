"""Writes c2_fullsize_fp64_field.npz and c3_fullsize_fp64_field.npz: the
field of the FP64 oracle (oracle/sem_oracle.c built with REAL=double, its
colour-parallel path) on the full production mesh -- 2000x500 elements,
16,010,001 GLL points, coupled P-SV+SH solver, a plane S wave polarised 30
degrees off SV -- at level NST-1 = 2000, as the reference's loop leaves it
at the end of that iteration (sem2d_psvsh.c:418-512):

  c2: configs.c2_case, crust / tilted-axis 4% hexagonal upper mantle /
      mantle, p = 0.05: the general 15-constant stress form;
  c3: configs.c3_case, the batch's three flat isotropic layers, p = 0.06
      (the batch's waves differ in p and in this polarisation angle): the
      short orthotropic stress form the 64-wave batch runs.

In 2000 steps (31 s) the pulse, which enters at the bottom and left edges
6 s after the records start, has crossed the bottom 100-150 km: the
fixtures check the production-size path (every tile column, the layer cuts,
the stress forms, the absorbing edges of a 16M-point mesh) against FP64,
not the anisotropic layer's splitting, which the 200x100-mesh fixtures
cover.

Stored in float32, every STRIDE-th node column and row ([3][101][401],
components x, y, z).  tests/test_gpu_fullsize.py holds the GPU field there
to 1e-4 of each component's peak.

  python tests/golden/make_fullsize_fp64.py [c2|c3]   (about 20 minutes each on 8 cores, 13 GB)
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

NST = 2001
STRIDE = 20


def case(name):
    if name == "c2":
        return configs.c2_case(nst=NST, psi_deg=30.0, sh_zero=False)
    return configs.c3_case(nst=NST, p=0.06, psi_deg=30.0)


if __name__ == "__main__":
    for name in sys.argv[1:] or ["c2", "c3"]:
        with tempfile.TemporaryDirectory() as d:
            case(name).write(d)
            o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
            _, f = o.run(NST, want_field=True, fast=True)
        s = f[:, ::STRIDE, ::STRIDE].astype(np.float32)
        np.savez_compressed(os.path.join(HERE, "%s_fullsize_fp64_field.npz" % name), field=s, nst=np.int32(NST),
                            stride=np.int32(STRIDE))
        print("wrote", name, s.shape, np.abs(s).max(axis=(1, 2)), flush=True)
