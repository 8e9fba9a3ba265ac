"""Shared synthetic cases for the tests (written in the reference's formats)."""
import os

import numpy as np

from semsks import inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def write_case(tmpdir, **kw):
    c = inputs.small_case(**kw)
    d = os.fspath(tmpdir)
    D = c.write(d)
    return c, D, d


def interp_mesh_case(tmpdir, tag, key, solver):
    """A 2x2-element case (nx = nz = 9) whose greenfile/shgreenfile hold the
    records of tests/golden/ref_interpl_mesh.npz (set `tag`), spaced at that
    file's dp for `key`, with nst = its nt_sem + 1."""
    z = np.load(os.path.join(GOLD, "ref_interpl_mesh.npz"))
    psv, sh = z[tag + "_psv"], z[tag + "_sh"]
    nst = int(z[key + "_ntsem"]) + 1
    c = inputs.Case(nex=2, nez=2, eh=2.0, layers_thk=(4.0,), vp=(6.0,), vs=(3.5,), rho=(2.7,), jitter=0.0,
                    nst=nst, nrec=2, sdt=float(str(z["sdt"])), solver=solver, sh_zero=False)
    d = os.fspath(tmpdir)
    D = c.write(d, nt=psv.shape[0], write_records=False)
    assert D["nx"] == 9 and D["nz"] == 9 and "%.6f" % D["sdt"] == str(z["sdt"])
    psv.astype(np.float32).tofile(os.path.join(d, "greenfile"))
    sh.astype(np.float32).tofile(os.path.join(d, "shgreenfile"))
    par = open(os.path.join(d, "run.par")).read().splitlines()
    par = [("dp=" + str(z[key + "_dp"])) if ln.startswith("dp=") else ln for ln in par]
    open(os.path.join(d, "run.par"), "w").write("\n".join(par) + "\n")
    return d
