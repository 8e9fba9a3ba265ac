"""Writes golden vectors from the reference's own compiled code.

Runs in the build container only (needs oracle/_ref/libref.so, built by
oracle/Makefile from semsyn/src/interpl.c, readmodel.c and nrutil.c where
they lie).  The vectors are data -- inputs and the reference's outputs --
committed as ref_interpl.npz and ref_readmodel.npz.

  python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

libc = C.CDLL("libc.so.6")
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]


def ref_interp_series(records_list, nsteps, rdt, shared):
    """Drives the reference interpl() exactly as the solvers do: one stream
    (sem2d.c:410-413) or two streams sharing the slot ring ns
    (sem2d_psvsh.c:430-431).  Returns the interpolated vectors per step."""
    L = pyoracle.ref_lib()
    outs = []
    files, sys_, widths = [], [], []
    tmp = tempfile.mkdtemp()
    for k, recs in enumerate(records_list):
        path = os.path.join(tmp, "g%d" % k)
        recs.astype(np.float32).tofile(path)
        files.append(libc.fopen(path.encode(), b"rb"))
        n = recs.shape[1]
        widths.append(n)
        sys_.append(L.matrix(0, n - 1, 0, 3))
        for i in range(n):
            for j in range(4):
                sys_[-1][i][j] = 0.0
        outs.append(np.zeros((nsteps, n), np.float32))
    ns_shared = (C.c_int * 4)(0, 1, 2, 3)
    ns_own = [(C.c_int * 4)(0, 1, 2, 3) for _ in records_list]
    for it in range(nsteps):
        for k in range(len(records_list)):
            ns = ns_shared if shared else ns_own[k]
            buf = (C.c_float * widths[k])()
            L.interpl(files[k], 0, it, float(np.float32(rdt)), widths[k], ns, sys_[k], buf, 0, 0)
            outs[k][it] = np.frombuffer(buf, np.float32)
    for f in files:
        libc.fclose(f)
    return outs


def ref_read_model(text, flat):
    L = pyoracle.ref_lib()
    fd, path = tempfile.mkstemp()
    os.write(fd, text.encode())
    os.close(fd)
    jo, nb, lf = C.c_int(), C.c_int(), C.c_int()
    fp = [C.POINTER(C.c_float)() for _ in range(4)]
    ip = [C.POINTER(C.c_int)() for _ in range(4)]
    L.read_model(flat, path.encode(), C.byref(jo), C.byref(nb), C.byref(lf), *[C.byref(x) for x in fp],
                 *[C.byref(x) for x in ip])
    n = jo.value
    arr = lambda p: np.array([p[i] for i in range(n)], np.float32)
    nen0 = ip[0][0]
    na = np.array([ip[2][i] for i in range(sum(ip[0][k] for k in range(lf.value)))], np.int32)
    return dict(jo=n, nb=nb.value, lfinal=lf.value, cc=arr(fp[0]), ss=arr(fp[1]), dd=arr(fp[2]),
                tth=arr(fp[3]), nen0=nen0, na=na, nrec=ip[2][nen0 - 2])


def main():
    rng = np.random.default_rng(20261203)
    out = {}
    cases = [("r1_single", 1.0, False), ("r1_shared", 1.0, True), ("r07_single", 0.7, False),
             ("r07_shared", 0.7, True), ("r037_shared", 0.37, True), ("r1p_single", 0.999, False)]
    nsteps = 40
    psv = rng.standard_normal((60, 24)).astype(np.float32)
    sh = rng.standard_normal((60, 12)).astype(np.float32)
    out["psv"] = psv
    out["sh"] = sh
    for name, rdt, shared in cases:
        recs = [psv, sh] if shared else [psv]
        res = ref_interp_series(recs, nsteps, rdt, shared)
        out[name + "_rdt"] = np.float32(rdt)
        out[name + "_g"] = res[0]
        if shared:
            out[name + "_gv"] = res[1]
    np.savez_compressed(os.path.join(HERE, "ref_interpl.npz"), **out)

    models = {
        "flat3": ("4 4\n5.8000 3.4000 2.6000 20\n6.5000 3.7500 2.9000 15\n8.0000 4.5000 3.3000 165.001\n"
                  "8.0000 4.5000 3.3000 0\n1\n3 3 4 4\n1 1 1 1\n", 1),
        "sph3": ("4 4\n5.8000 3.4000 2.6000 20\n6.5000 3.7500 2.9000 15\n8.0000 4.5000 3.3000 165.001\n"
                 "8.0000 4.5000 3.3000 0\n1\n3 3 4 4\n1 1 1 1\n", 0),
        "rays2": ("5 3\n5.8 3.4 2.6 10.5\n6.1 3.5 2.7 12.25\n6.6 3.8 2.9 30\n8.1 4.5 3.4 150.003\n"
                  "8.2 4.6 3.4 0\n2\n4 2 5 5 5\n1 1 1 1 1\n3 1 2 3\n2 2 2 2\n", 1),
    }
    md = {}
    for name, (text, flat) in models.items():
        r = ref_read_model(text, flat)
        md[name + "_text"] = np.array(text)
        md[name + "_flat"] = np.int32(flat)
        for k, v in r.items():
            md[name + "_" + k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "ref_readmodel.npz"), **md)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
