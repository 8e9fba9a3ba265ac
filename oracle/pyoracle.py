"""ctypes binding of the CPU oracle (oracle/sem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result and by bench.py's cpu_baseline leg.  The product never loads
it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "f32": os.path.join(HERE, "build", "liboracle_f32.so"),
    "f32_O0": os.path.join(HERE, "build", "liboracle_f32_O0.so"),
    "f64": os.path.join(HERE, "build", "liboracle_f64.so"),
}
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

_cache: dict = {}


def load(kind: str = "f64") -> C.CDLL:
    if kind in _cache:
        return _cache[kind]
    L = C.CDLL(LIBS[kind])
    vp = C.c_void_p
    L.orc_last_error.restype = C.c_char_p
    L.orc_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.orc_destroy.argtypes = [vp]
    L.orc_dims.argtypes = [vp, C.POINTER(C.c_int)]
    L.orc_scalars.argtypes = [vp, C.POINTER(C.c_double)]
    L.orc_tables.argtypes = [vp, vp, vp, vp]
    L.orc_material.argtypes = [vp, vp]
    L.orc_mass.argtypes = [vp, vp, vp, vp]
    L.orc_receivers.argtypes = [vp, vp, vp]
    L.orc_stiff.argtypes = [vp, vp, vp]
    L.orc_update.argtypes = [vp, vp, vp, vp, vp]
    L.orc_boundary_series.argtypes = [vp, C.c_int, vp, vp]
    L.orc_run.argtypes = [vp, C.c_int, vp, vp]
    L.orc_run_fast.argtypes = [vp, C.c_int, vp, vp]
    L.orc_main.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_char_p]
    L.orc_gll.argtypes = [C.c_int, vp, vp]
    L.orc_dlag.argtypes = [C.c_int, vp, vp]
    L.orc_set_node_tensor.argtypes = [vp, C.c_long, vp, C.c_float]
    L.orc_interp_series.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_float, vp, vp]
    L.orc_guard.argtypes = [vp, C.c_int]
    L.orc_sponge.argtypes = [vp, vp]
    L.orc_set_step_clock.argtypes = [vp, C.c_int]
    L.orc_taper.argtypes = [vp, vp]
    L.orc_read_raymodel.argtypes = [C.c_char_p, C.c_int, C.c_int, vp, vp, vp, vp, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]
    _cache[kind] = L
    return L


def dtype_of(kind: str):
    return np.float64 if kind == "f64" else np.float32


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("oracle status %d: %s" % (code, msg))
        self.code = code


class Oracle:
    """One solver set up from reference-style arguments (e.g. ['par=run.par'])."""

    def __init__(self, args, kind: str = "f64", cwd: str | None = None):
        self.kind = kind
        self.L = load(kind)
        self.dt = dtype_of(kind)
        old = os.getcwd()
        if cwd:
            os.chdir(cwd)
        try:
            argv = ["sem2d_psvsh"] + list(args)
            arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
            h = C.c_void_p()
            st = self.L.orc_create(len(argv), arr, C.byref(h))
        finally:
            os.chdir(old)
        if st != 0:
            raise OracleError(st, self.L.orc_last_error().decode())
        self.h = h
        self.cwd = cwd
        d = (C.c_int * 10)()
        self.L.orc_dims(h, d)
        (self.nex, self.nez, self.nx, self.nz, self.hpn, self.ntrace, self.nt_sem, self.nst,
         self.neh, self.variant) = list(d)
        s = (C.c_double * 8)()
        self.L.orc_scalars(h, s)
        self.eh, self.h_, self.je, self.jei, self.jb, self.sdt, self.rdt, self.dp = list(s)

    def tables(self):
        n = self.hpn + 1
        gp = np.zeros(n, self.dt); gw = np.zeros(n, self.dt); dl = np.zeros((n, n), self.dt)
        self.L.orc_tables(self.h, gp.ctypes.data, gw.ctypes.data, dl.ctypes.data)
        return gp, gw, dl

    def material(self) -> np.ndarray:
        out = np.zeros((self.nx * self.nz, 16), self.dt)
        self.L.orc_material(self.h, out.ctypes.data)
        return out

    def mass(self):
        N = self.nx * self.nz
        m = np.zeros(N, self.dt); md = np.zeros(N, self.dt); ab = np.zeros((N, 3), self.dt)
        self.L.orc_mass(self.h, m.ctypes.data, md.ctypes.data, ab.ctypes.data)
        return m.reshape(self.nz, self.nx), md.reshape(self.nz, self.nx), ab.reshape(self.nz, self.nx, 3)

    def receivers(self):
        ix = np.zeros(self.ntrace, np.int32); iz = np.zeros(self.ntrace, np.int32)
        self.L.orc_receivers(self.h, ix.ctypes.data, iz.ctypes.data)
        return ix, iz

    def stiff(self, u_soa: np.ndarray) -> np.ndarray:
        """K u for a field given as SoA planes [3][nz][nx]; returns the same layout."""
        u = np.ascontiguousarray(np.moveaxis(u_soa, 0, -1), dtype=self.dt)   # AoS [nz][nx][3]
        f = np.zeros_like(u)
        self.L.orc_stiff(self.h, u.ctypes.data, f.ctypes.data)
        return np.moveaxis(f, -1, 0).copy()

    def sponge(self, u_soa: np.ndarray) -> np.ndarray:
        """atten_wave2 applied once to a field given as SoA planes [3][nz][nx]."""
        u = np.ascontiguousarray(np.moveaxis(u_soa, 0, -1), dtype=self.dt)
        self.L.orc_sponge(self.h, u.ctypes.data)
        return np.moveaxis(u, -1, 0).copy()

    def taper(self) -> np.ndarray:
        out = np.zeros(self.neh * self.hpn, np.float32)
        n = self.L.orc_taper(self.h, out.ctypes.data)
        return out[:n]

    def update(self, u, u0, G):
        def aos(a):
            return np.ascontiguousarray(np.moveaxis(a, 0, -1), dtype=self.dt)
        a, a0, g = aos(u), aos(u0), aos(G)      # keep the buffers alive across the call
        u1 = np.zeros_like(a)
        self.L.orc_update(self.h, a.ctypes.data, a0.ctypes.data, g.ctypes.data, u1.ctypes.data)
        return np.moveaxis(u1, -1, 0).copy()

    def boundary_series(self, n: int):
        old = os.getcwd()
        if self.cwd:
            os.chdir(self.cwd)
        try:
            g = np.zeros((n, 2 * (self.nx + self.nz)), self.dt)
            gv = np.zeros((n, self.nx + self.nz), self.dt)
            st = self.L.orc_boundary_series(self.h, n, g.ctypes.data, gv.ctypes.data)
        finally:
            os.chdir(old)
        if st:
            raise OracleError(st, self.L.orc_last_error().decode())
        return g, gv

    def run_timed(self, nsteps: int):
        """run(nsteps) that also returns the wall clock (s) at the end of
        every step."""
        t = np.zeros(nsteps, np.float64)
        self.L.orc_set_step_clock(t.ctypes.data, nsteps)
        try:
            tr = self.run(nsteps)
        finally:
            self.L.orc_set_step_clock(None, 0)
        return tr, t

    def run(self, nsteps: int | None = None, want_field: bool = False, fast: bool = False):
        """Traces [nsteps][3][ntrace] (and the final field).  fast: the
        element force swept in parallel colours (orc_run_fast), the same
        arithmetic summed in another order at shared nodes."""
        n = self.nst if nsteps is None else nsteps
        tr = np.zeros((n, 3, max(self.ntrace, 1)), np.float32)
        fu = np.zeros((self.nz, self.nx, 3), self.dt) if want_field else None
        old = os.getcwd()
        if self.cwd:
            os.chdir(self.cwd)
        try:
            run = self.L.orc_run_fast if fast else self.L.orc_run
            st = run(self.h, n, tr.ctypes.data, None if fu is None else fu.ctypes.data)
        finally:
            os.chdir(old)
        if st:
            raise OracleError(st, self.L.orc_last_error().decode())
        tr = tr[:, :, :self.ntrace]
        if want_field:
            return tr, np.moveaxis(fu, -1, 0).copy()
        return tr

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main(args, outdir: str, kind: str = "f32", cwd: str | None = None) -> int:
    L = load(kind)
    old = os.getcwd()
    if cwd:
        os.chdir(cwd)
    try:
        argv = ["sem2d_psvsh"] + list(args)
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        return L.orc_main(len(argv), arr, outdir.encode())
    finally:
        os.chdir(old)


def gll(hpn: int, kind: str = "f64"):
    L = load(kind)
    dt = dtype_of(kind)
    x = np.zeros(hpn + 1, dt); w = np.zeros(hpn + 1, dt)
    assert L.orc_gll(hpn, x.ctypes.data, w.ctypes.data) == 0
    d = np.zeros((hpn + 1, hpn + 1), dt)
    L.orc_dlag(hpn, x.ctypes.data, d.ctypes.data)
    return x, w, d


# ---------------------------------------------------------------------------
# the reference's own interpl.c / readmodel.c / nrutil.c (oracle/_ref)
# ---------------------------------------------------------------------------
def ref_lib() -> C.CDLL | None:
    if not os.path.exists(REF_LIB):
        return None
    if "ref" in _cache:
        return _cache["ref"]
    L = C.CDLL(REF_LIB)
    vp = C.c_void_p
    L.matrix.restype = C.POINTER(C.POINTER(C.c_float))
    L.matrix.argtypes = [C.c_long, C.c_long, C.c_long, C.c_long]
    # K&R definitions: float parameters arrive as promoted doubles
    L.interpl.argtypes = [vp, C.c_long, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
                          C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_float), C.c_int, C.c_int]
    L.interp_const.restype = C.c_int
    L.interp_const.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]
    L.read_model.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_float)),
                             C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_float)),
                             C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                             C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int))]
    _cache["ref"] = L
    return L


def interp_series(rec0, rec1, nsteps, rdt, kind="f32"):
    """Raw 4-tap interpolation of in-memory records (one stream, or two
    sharing the slot ring when rec1 is given)."""
    L = load(kind)
    dt = dtype_of(kind)
    r0 = np.ascontiguousarray(rec0, np.float32)
    r1 = None if rec1 is None else np.ascontiguousarray(rec1, np.float32)
    o0 = np.zeros((nsteps, r0.shape[1]), dt)
    o1 = None if r1 is None else np.zeros((nsteps, r1.shape[1]), dt)
    st = L.orc_interp_series(r0.ctypes.data, r0.shape[1], None if r1 is None else r1.ctypes.data,
                             0 if r1 is None else r1.shape[1], r0.shape[0], nsteps, rdt, o0.ctypes.data,
                             None if o1 is None else o1.ctypes.data)
    assert st == 0
    return o0, o1


def guard(g, kind="f32"):
    L = load(kind)
    a = np.ascontiguousarray(g, dtype_of(kind)).copy()
    L.orc_guard(a.ctypes.data, a.size)
    return a


def read_raymodel(path, flat, kind="f32"):
    L = load(kind)
    cap = 64
    arrs = [np.zeros(cap, np.float32) for _ in range(4)]
    jo, nrec = C.c_int(), C.c_int()
    st = L.orc_read_raymodel(path.encode(), flat, cap, *[a.ctypes.data for a in arrs], C.byref(jo), C.byref(nrec))
    assert st == 0
    n = jo.value
    return dict(jo=n, nrec=nrec.value, cc=arrs[0][:n], ss=arrs[1][:n], dd=arrs[2][:n], tth=arrs[3][:n])
