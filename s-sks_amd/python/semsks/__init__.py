"""ctypes binding of the C ABI in include/semsks.h (libsemsks.so).

The library is the product; this module only marshals numpy arrays across
the C ABI for tests, the benchmark and scripting.  It raises if the native
library is missing -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.normpath(os.path.join(_HERE, "..", ".."))          # s-sks_amd/
LIB_PATH = os.path.join(PKG_ROOT, "build", "libsemsks.so")
CLI_PATH = os.path.join(PKG_ROOT, "build", "semsks")
SAC_CLI_PATH = os.path.join(PKG_ROOT, "build", "semsks_uvw2sac")

MAT_STRIDE = 16
SOLVER_PSVSH, SOLVER_PSV, SOLVER_PSV_MP = 0, 1, 2
F_NO_KAHAN = 1

STATUS = {0: "OK", 1: "EINVAL", 2: "EIO", 3: "ENOHBM", 4: "ENONFINITE", 5: "ERANGE", 6: "ENP",
          7: "EDEVICE", 8: "ENOSLOT", 100: "EPAR", 101: "ESTOP"}

EXPORTS = [
    "semsks_load_par", "semsks_model_problem", "semsks_model_dims", "semsks_model_outdir",
    "semsks_model_free", "semsks_write_check", "semsks_model_snapshots", "semsks_write_snapshot",
    "semsks_write_seis", "semsks_last_error", "semsks_engine_create",
    "semsks_engine_destroy", "semsks_admit", "semsks_footprint", "semsks_admission_bytes", "semsks_advance", "semsks_prepare",
    "semsks_slot_status", "semsks_get_traces", "semsks_get_field", "semsks_get_snapshot", "semsks_retire",
    "semsks_sync",
    "semsks_last_timing", "semsks_last_timed_steps", "semsks_sweep_steps", "semsks_tile_rows", "semsks_run_problem", "semsks_stiff", "semsks_stiff_dev",
    "semsks_get_setup", "semsks_get_boundary_series", "semsks_planewave", "semsks_serve", "semsks_serve_ex",
    "semsks_uvw2sac",
]


class SemsksError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


class Problem(C.Structure):
    _fields_ = [
        ("nex", C.c_int), ("nez", C.c_int), ("hpn", C.c_int), ("neh", C.c_int),
        ("eh", C.c_float), ("sdt", C.c_float), ("nst", C.c_int), ("solver", C.c_int),
        ("absorb", C.c_int), ("flags", C.c_int),
        ("n_mat", C.c_int), ("mat_table", C.POINTER(C.c_float)), ("mat_id", C.POINTER(C.c_uint16)),
        ("ntrace", C.c_int), ("rec_ix", C.POINTER(C.c_int)), ("rec_iz", C.POINTER(C.c_int)),
        ("nt", C.c_int), ("dp", C.c_float),
        ("psv", C.POINTER(C.c_float)), ("sh", C.POINTER(C.c_float)),
        ("atten", C.c_int), ("ass", C.c_float),
    ]


class Request(C.Structure):
    _fields_ = [("problem", C.POINTER(Problem)), ("arrival_s", C.c_double), ("use_planewave", C.c_int),
                ("planewave", C.c_double * 4), ("traces", C.POINTER(C.c_float)),
                ("admit_s", C.c_double), ("done_s", C.c_double), ("slot", C.c_int), ("status", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("point_steps", C.c_double), ("completed", C.c_int),
                ("max_active", C.c_int), ("rounds", C.c_int)]


class ServeOpts(C.Structure):
    _fields_ = [("steps_per_round", C.c_int), ("order", C.c_int), ("max_active_points", C.c_double)]


SERVE_ORDER = {"fcfs": 0, "shortest": 1}


class EngineOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_slots", C.c_int), ("tile_z", C.c_int),
                ("waves_per_cu", C.c_int), ("use_graph", C.c_int), ("steps_per_sweep", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsemsks.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.semsks_last_error.restype = C.c_char_p
    L.semsks_load_par.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.semsks_model_problem.restype = C.POINTER(Problem)
    L.semsks_model_problem.argtypes = [vp]
    L.semsks_model_dims.argtypes = [vp, C.POINTER(C.c_int)]
    L.semsks_model_outdir.restype = C.c_char_p
    L.semsks_model_outdir.argtypes = [vp]
    L.semsks_model_free.argtypes = [vp]
    L.semsks_write_check.argtypes = [vp]
    L.semsks_model_snapshots.argtypes = [vp, C.POINTER(C.c_int)]
    L.semsks_write_snapshot.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, vp]
    L.semsks_write_seis.argtypes = [C.c_char_p, C.c_int, C.c_int, vp]
    L.semsks_engine_create.argtypes = [C.POINTER(EngineOpts), C.POINTER(vp)]
    L.semsks_engine_destroy.argtypes = [vp]
    L.semsks_admit.argtypes = [vp, C.POINTER(Problem), C.c_int, C.POINTER(C.c_int)]
    L.semsks_footprint.restype = C.c_size_t
    L.semsks_footprint.argtypes = [C.POINTER(Problem), C.c_int]
    L.semsks_admission_bytes.restype = C.c_size_t
    L.semsks_admission_bytes.argtypes = [C.POINTER(Problem), C.c_int]
    L.semsks_advance.argtypes = [vp, C.c_int, C.c_int]
    L.semsks_prepare.argtypes = [vp, C.c_int]
    L.semsks_slot_status.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.semsks_get_traces.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
    L.semsks_get_field.argtypes = [vp, C.c_int, vp]
    L.semsks_get_snapshot.argtypes = [vp, C.c_int, vp]
    L.semsks_retire.argtypes = [vp, C.c_int]
    L.semsks_sync.argtypes = [vp]
    L.semsks_last_timing.argtypes = [vp, C.POINTER(C.c_double)]
    L.semsks_last_timed_steps.restype = C.c_int
    L.semsks_last_timed_steps.argtypes = [vp]
    L.semsks_sweep_steps.restype = C.c_int
    L.semsks_sweep_steps.argtypes = [vp]
    L.semsks_tile_rows.restype = C.c_int
    L.semsks_tile_rows.argtypes = [vp]
    L.semsks_run_problem.argtypes = [C.POINTER(Problem), vp]
    L.semsks_stiff.argtypes = [vp, C.c_int, vp, vp]
    L.semsks_get_setup.argtypes = [vp, C.c_int, vp, vp, vp]
    L.semsks_get_boundary_series.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    L.semsks_planewave.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float,
                                   C.c_double, C.c_double, C.c_double, C.c_double, vp, vp]
    L.semsks_serve.argtypes = [vp, C.POINTER(Request), C.c_int, C.c_int, C.POINTER(ServeStats)]
    L.semsks_serve_ex.argtypes = [vp, C.POINTER(Request), C.c_int, C.POINTER(ServeOpts), C.POINTER(ServeStats)]
    _lib = L
    return L


def _check(st: int) -> None:
    if st != 0:
        raise SemsksError(st, lib().semsks_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


class Model:
    """Parsed inputs (`semsks_load_par`): the reference's par keys and files."""

    def __init__(self, args: list[str]):
        argv = ["semsks"] + list(args)
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        h = C.c_void_p()
        _check(lib().semsks_load_par(len(argv), arr, C.byref(h)))
        self._h = h
        self.problem = lib().semsks_model_problem(h).contents
        d = (C.c_int * 8)()
        lib().semsks_model_dims(h, d)
        self.nx, self.nz, self.mtotal, self.shtotal, self.nt_sem, self.plot_trace, self.ntrace, self.n_mat = list(d)
        self.outdir = lib().semsks_model_outdir(h).decode()

    def material(self) -> np.ndarray:
        """Per-node records [nz*nx][16] (rho + 15 constants)."""
        p = self.problem
        tab = np.ctypeslib.as_array(p.mat_table, shape=(p.n_mat * MAT_STRIDE,)).reshape(p.n_mat, MAT_STRIDE)
        ids = np.ctypeslib.as_array(p.mat_id, shape=(self.nx * self.nz,))
        return tab[ids]

    def write_check(self):
        """grid.xz, legendre.xz and vel.xy into the model's outdir when the
        par file leaves check at 1 (semsks_write_check)."""
        _check(lib().semsks_write_check(self._h))

    def snapshots(self):
        """(plot_snap, startsnap, incresnap) from the par keys."""
        k = (C.c_int * 3)()
        lib().semsks_model_snapshots(self._h, k)
        return tuple(k)

    def receivers(self):
        p = self.problem
        n = p.ntrace
        return (np.ctypeslib.as_array(p.rec_ix, shape=(n,)).copy(),
                np.ctypeslib.as_array(p.rec_iz, shape=(n,)).copy())

    def close(self):
        if getattr(self, "_h", None):
            lib().semsks_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ProblemArrays:
    """Owns numpy arrays behind a Problem struct built from Python."""

    def __init__(self, *, nex, nez, hpn, neh, eh, sdt, nst, mat_table, mat_id, rec_ix, rec_iz,
                 nt=0, dp=1.0, psv=None, sh=None, solver=SOLVER_PSVSH, absorb=1, flags=0):
        self.mat_table = np.ascontiguousarray(mat_table, dtype=np.float32)
        self.mat_id = np.ascontiguousarray(mat_id, dtype=np.uint16)
        self.rec_ix = np.ascontiguousarray(rec_ix, dtype=np.int32)
        self.rec_iz = np.ascontiguousarray(rec_iz, dtype=np.int32)
        self.psv = None if psv is None else np.ascontiguousarray(psv, dtype=np.float32)
        self.sh = None if sh is None else np.ascontiguousarray(sh, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        self.p = Problem(nex=nex, nez=nez, hpn=hpn, neh=neh, eh=eh, sdt=sdt, nst=nst, solver=solver,
                         absorb=absorb, flags=flags, n_mat=len(self.mat_table),
                         mat_table=self.mat_table.ctypes.data_as(fp),
                         mat_id=self.mat_id.ctypes.data_as(C.POINTER(C.c_uint16)),
                         ntrace=len(self.rec_ix), rec_ix=self.rec_ix.ctypes.data_as(C.POINTER(C.c_int)),
                         rec_iz=self.rec_iz.ctypes.data_as(C.POINTER(C.c_int)), nt=nt, dp=dp,
                         psv=None if self.psv is None else self.psv.ctypes.data_as(fp),
                         sh=None if self.sh is None else self.sh.ctypes.data_as(fp))


def footprint(p: Problem, shared: bool = False) -> int:
    return int(lib().semsks_footprint(C.byref(p), int(shared)))


def admission_bytes(p: Problem, device_records: bool = False) -> int:
    """Bytes the engine's admission scratch needs for p (semsks_admission_bytes)."""
    return int(lib().semsks_admission_bytes(C.byref(p), int(device_records)))


class Engine:
    def __init__(self, max_slots=1, hbm_budget=0, tile_z=0, waves_per_cu=0, use_graph=1, steps_per_sweep=0):
        o = EngineOpts(hbm_budget=hbm_budget, max_slots=max_slots, tile_z=tile_z,
                       waves_per_cu=waves_per_cu, use_graph=use_graph, steps_per_sweep=steps_per_sweep)
        h = C.c_void_p()
        _check(lib().semsks_engine_create(C.byref(o), C.byref(h)))
        self._h = h
        self.slots: dict[int, Problem] = {}

    def admit(self, p: Problem, share: int = -1) -> int:
        s = C.c_int(-1)
        _check(lib().semsks_admit(self._h, C.byref(p), share, C.byref(s)))
        self.slots[s.value] = Problem.from_buffer_copy(p)
        return s.value

    def advance(self, n: int, sync: bool = True):
        _check(lib().semsks_advance(self._h, n, int(sync)))

    def prepare(self, n: int):
        """Captures the graphs advance(n) will replay (semsks_prepare)."""
        _check(lib().semsks_prepare(self._h, n))

    def status(self, slot: int):
        lv, dn = C.c_int(), C.c_int()
        _check(lib().semsks_slot_status(self._h, slot, C.byref(lv), C.byref(dn)))
        return lv.value, bool(dn.value)

    def traces(self, slot: int, it0: int = 0, n: int | None = None) -> np.ndarray:
        p = self.slots[slot]
        n = p.nst - it0 if n is None else n
        out = np.zeros((n, 3, max(p.ntrace, 1)), dtype=np.float32)
        _check(lib().semsks_get_traces(self._h, slot, it0, n, out.ctypes.data))
        return out[:, :, :p.ntrace]

    def field(self, slot: int) -> np.ndarray:
        p = self.slots[slot]
        nx, nz = p.nex * p.hpn + 1, p.nez * p.hpn + 1
        out = np.zeros((3, nz, nx), dtype=np.float32)
        _check(lib().semsks_get_field(self._h, slot, out.ctypes.data))
        return out

    def snapshot(self, slot: int) -> np.ndarray:
        """The current level before the next step's interface injection
        (semsks_get_snapshot), [3][nz][nx]."""
        p = self.slots[slot]
        nx, nz = p.nex * p.hpn + 1, p.nez * p.hpn + 1
        out = np.zeros((3, nz, nx), dtype=np.float32)
        _check(lib().semsks_get_snapshot(self._h, slot, out.ctypes.data))
        return out

    def stiff(self, slot: int, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float32)
        f = np.zeros_like(u)
        _check(lib().semsks_stiff(self._h, slot, u.ctypes.data, f.ctypes.data))
        return f

    def setup(self, slot: int):
        p = self.slots[slot]
        nx, nz = p.nex * p.hpn + 1, p.nez * p.hpn + 1
        mass = np.zeros((nz, nx)); mdt2 = np.zeros((nz, nx), np.float32); ab = np.zeros((nz, nx, 3))
        _check(lib().semsks_get_setup(self._h, slot, mass.ctypes.data, mdt2.ctypes.data, ab.ctypes.data))
        return mass, mdt2, ab

    def boundary_series(self, slot: int, n: int):
        p = self.slots[slot]
        nx, nz = p.nex * p.hpn + 1, p.nez * p.hpn + 1
        g = np.zeros((n, 2 * (nx + nz)), np.float32)
        gv = np.zeros((n, nx + nz), np.float32)
        _check(lib().semsks_get_boundary_series(self._h, slot, n, g.ctypes.data, gv.ctypes.data))
        return g, gv

    def planewave(self, nex, nez, hpn, neh, eh, nt, dp, p, beta, f0, psi, to_host=True):
        nx, nz = nex * hpn + 1, nez * hpn + 1
        if to_host:
            psv = np.zeros((nt, 2 * (nx + nz)), np.float32)
            sh = np.zeros((nt, nx + nz), np.float32)
            _check(lib().semsks_planewave(self._h, nex, nez, hpn, neh, eh, nt, dp, p, beta, f0, psi,
                                          psv.ctypes.data, sh.ctypes.data))
            return psv, sh
        _check(lib().semsks_planewave(self._h, nex, nez, hpn, neh, eh, nt, dp, p, beta, f0, psi, None, None))
        return None

    def serve(self, problems, arrivals, planewaves=None, steps_per_round: int = 16, want_traces: bool = True,
              order: str = "fcfs", max_active_points: float = 0.0):
        """Continuous batching (semsks_serve_ex): problems[i] arrives
        arrivals[i] seconds after the start; planewaves[i] = (p, beta, f0,
        psi) generates its boundary records on the device; order "fcfs" or
        "shortest" (least work first) among the waiting requests, admitted
        while the active instances hold at most max_active_points GLL points
        (0: no cap).  Returns (per-request dicts, stats)."""
        n = len(problems)
        by_arrival = sorted(range(n), key=lambda i: arrivals[i])
        reqs = (Request * n)()
        keep = []
        for j, i in enumerate(by_arrival):
            p = problems[i]
            r = reqs[j]
            r.problem = C.pointer(p)
            r.arrival_s = float(arrivals[i])
            if planewaves is not None and planewaves[i] is not None:
                r.use_planewave = 1
                for k in range(4):
                    r.planewave[k] = float(planewaves[i][k])
            if want_traces and p.ntrace > 0:
                tr = np.zeros((p.nst, 3, p.ntrace), np.float32)
                keep.append(tr)
                r.traces = tr.ctypes.data_as(C.POINTER(C.c_float))
            else:
                keep.append(None)
        st = ServeStats()
        opts = ServeOpts(steps_per_round=steps_per_round, order=SERVE_ORDER[order],
                         max_active_points=float(max_active_points))
        _check(lib().semsks_serve_ex(self._h, reqs, n, C.byref(opts), C.byref(st)))
        out = [None] * n
        for j, i in enumerate(by_arrival):
            r = reqs[j]
            out[i] = dict(arrival=r.arrival_s, admit=r.admit_s, done=r.done_s, status=r.status, slot=r.slot,
                          latency=r.done_s - r.arrival_s, traces=keep[j])
        return out, dict(wall_s=st.wall_s, point_steps=st.point_steps, completed=st.completed,
                         max_active=st.max_active, rounds=st.rounds)

    def timing(self):
        t = (C.c_double * 2)()
        lib().semsks_last_timing(self._h, t)
        return t[0], t[1], lib().semsks_last_timed_steps(self._h)

    def sweep_steps(self) -> int:
        """Time steps per step-kernel launch (semsks_sweep_steps): 1."""
        return int(lib().semsks_sweep_steps(self._h))

    def tile_rows(self) -> int:
        """Element rows per tile of the current tiling."""
        return int(lib().semsks_tile_rows(self._h))

    def retire(self, slot: int):
        _check(lib().semsks_retire(self._h, slot))
        self.slots.pop(slot, None)

    def sync(self):
        _check(lib().semsks_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().semsks_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_problem(p: Problem) -> np.ndarray:
    out = np.zeros((p.nst, 3, max(p.ntrace, 1)), np.float32)
    _check(lib().semsks_run_problem(C.byref(p), out.ctypes.data))
    return out[:, :, :p.ntrace]
