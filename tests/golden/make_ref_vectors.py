"""Writes golden vectors from the reference's own compiled code.

Runs in the build container only (needs oracle/_ref/libref.so, built by
oracle/Makefile from semsyn/src/interpl.c, readmodel.c and nrutil.c where
they lie, and oracle/_ref/elastic_rotate and elastic_dip, built from
semsyn/model/src).  The vectors are data -- inputs and the reference's outputs --
committed as ref_interpl.npz and ref_readmodel.npz.

  python tests/golden/make_ref_vectors.py          (all files)
  python tests/golden/make_ref_vectors.py mesh     (ref_interpl_mesh.npz only)
  python tests/golden/make_ref_vectors.py elastic  (ref_elastic.npz only)
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


# the mesh of the device test: 2 x 2 elements of degree 4 (nx = nz = 9), so
# the record widths are the solver's own: mtotal = 2(nx+nz) = 36, shtotal = 18
MESH_NX = MESH_NZ = 9
MESH_SDT = "0.020000"
MESH_RDT = (1.0, 0.999, 0.7, 0.37)


def mesh_records(spiky):
    """Boundary records for the device test: normal deviates (seeded), and in
    the spiky set values the solver's guard must repair (12, -12, 15, NaN),
    two of them at the first and last index of a record."""
    rng = np.random.default_rng(20261204)
    nrec = 40
    psv = (2.0 * rng.standard_normal((nrec, 2 * (MESH_NX + MESH_NZ)))).astype(np.float32)
    sh = (2.0 * rng.standard_normal((nrec, MESH_NX + MESH_NZ))).astype(np.float32)
    np.clip(psv, -6, 6, out=psv)
    np.clip(sh, -6, 6, out=sh)
    if spiky:
        psv[5, 0] = 12.0
        psv[9, psv.shape[1] - 1] = -12.0
        psv[14, 17] = np.nan
        psv[21, 20] = 25.0
        sh[7, 0] = np.nan
        sh[11, sh.shape[1] - 1] = 15.0
    return psv, sh


def mesh_vectors():
    """ref_interpl_mesh.npz: the reference interpl's vectors (before the
    solver's guard) at the mesh widths, for dp = sdt / rdt written as the
    par file holds it ("%.6f"), one stream (sem2d.c) and two streams sharing
    the slot ring (sem2d_psvsh.c), plain and spiky records."""
    out = {}
    sdt = np.float32(float(MESH_SDT))
    for spiky in (False, True):
        psv, sh = mesh_records(spiky)
        tag = "spiky" if spiky else "plain"
        out[tag + "_psv"] = psv
        out[tag + "_sh"] = sh
        for rdt_want in MESH_RDT:
            dp_txt = "%.6f" % (float(MESH_SDT) / rdt_want)
            dp = np.float32(float(dp_txt))
            rdt = np.float32(sdt / dp)                       # rdt = sdt/dp in float (sem2d_psvsh.c:388)
            nt = psv.shape[0]
            nt_sem = int(np.float32(np.float32((nt - 3) * dp) / sdt))
            key = "%s_r%s" % (tag, ("%g" % rdt_want).replace(".", "p"))
            out[key + "_dp"] = np.array(dp_txt)
            out[key + "_ntsem"] = np.int32(nt_sem)
            with np.errstate(invalid="ignore"):
                shared = ref_interp_series([psv, sh], nt_sem, float(rdt), True)
                single = ref_interp_series([psv], nt_sem, float(rdt), False)
            out[key + "_shared_g"], out[key + "_shared_gv"] = shared
            out[key + "_single_g"] = single[0]
    out["sdt"] = np.array(MESH_SDT)
    np.savez_compressed(os.path.join(HERE, "ref_interpl_mesh.npz"), **out)
    print("wrote ref_interpl_mesh.npz", len(out), "arrays")


def ref_elastic(prog, C, angle):
    """Runs the reference's elastic_rotate / elastic_dip (oracle/_ref) on a
    6x6 matrix: returns the rotated matrix it writes (%-8.2f)."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", prog)
    with tempfile.TemporaryDirectory() as d:
        src, dst = os.path.join(d, "in.ela"), os.path.join(d, "out.ela")
        with open(src, "w") as f:
            for row in C:
                f.write(" ".join("%.4f" % v for v in row) + "\n")
        subprocess.run([exe, src, dst, "%g" % angle], check=True, capture_output=True, cwd=d)
        return np.loadtxt(dst).reshape(6, 6)


def elastic_vectors():
    """ref_elastic.npz: the model builder's Bond rotations about the vertical
    axis (elastic_rotate) and within x1-x3 (elastic_dip) applied by the
    reference's own programs to a transversely isotropic matrix with a
    vertical axis and to the horizontal-axis matrix any_anisotropy.pl builds
    (step 0, sem2d's 'stronger' factor 1.3, its asthenosphere vp/vs/rho)."""
    vp, vs, rho = 8.1864, 4.5312, 3.4224
    c11, c22, c44, c55 = vp * vp * rho * 1.3, vp * vp * rho, vs * vs * rho * 1.3, vs * vs * rho
    hti = np.zeros((6, 6))
    hti[0, 0], hti[1, 1], hti[2, 2] = c11, c22, c22
    hti[0, 1] = hti[1, 0] = hti[0, 2] = hti[2, 0] = c11 - 2 * c44
    hti[1, 2] = hti[2, 1] = c22 - 2 * c55
    hti[3, 3], hti[4, 4], hti[5, 5] = c44, c55, c55
    vti = np.array([[250, 80, 60, 0, 0, 0], [80, 250, 60, 0, 0, 0], [60, 60, 210, 0, 0, 0],
                    [0, 0, 0, 60, 0, 0], [0, 0, 0, 0, 60, 0], [0, 0, 0, 0, 0, 85]], float)
    out = {"vti": vti, "hti": hti, "angles": np.array([0.0, 30.0, 70.0, 135.0, -20.0])}
    for name, C in (("vti", vti), ("hti", hti)):
        for prog in ("elastic_rotate", "elastic_dip"):
            out["%s_%s" % (name, prog)] = np.stack([ref_elastic(prog, C, a) for a in out["angles"]])
    np.savez_compressed(os.path.join(HERE, "ref_elastic.npz"), **out)
    print("wrote ref_elastic.npz")


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
    if sys.argv[1:] == ["mesh"]:
        mesh_vectors()
    elif sys.argv[1:] == ["elastic"]:
        elastic_vectors()
    else:
        main()
        mesh_vectors()
        elastic_vectors()
