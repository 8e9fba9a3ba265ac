"""Synthetic inputs in the reference's file formats.

Writes the files the SEM solver reads (`sem2d_psvsh par=run.par`):

* ``raymodel``  -- ``jo nb`` / ``jo`` lines ``vp vs rho thk`` / ``lfinal`` /
  per ray ``nen na...`` and ``ncoun nray...`` (semsyn/src/readmodel.c:61-107).
  Only ``nrec = na[nen[0]-2]`` and the thicknesses are used by the solver
  (semsyn/src/sem2d_psvsh.c:585-604).
* ``semmodel``  -- ``nll`` then per layer ``np c11 c12 c13 c16 c22 c23 c26``
  / ``c33 c36 c44 c45 c55 c66 rho`` / np x-points / np heights
  (semsyn/src/sem2d_psvsh.c:700-731; writer applications/test/iso2sem.pl:35-49).
* ``greenfile`` -- nt time-major records of ``[u_bot(nx) w_bot(nx) u_left(nz)
  w_left(nz)]`` float32, w positive down (semsyn/src/demult.c:63-77,
  semsyn/src/aserpsv_sem.c:319-488).
* ``shgreenfile`` -- nt records of ``[v_bot(nx) v_left(nz)]``.
* ``rec_file``  -- ``n`` then n receiver x positions in km
  (semsyn/src/sem2d_psvsh.c:347-358).
* ``run.par``   -- getpar ``name=value`` lines.

The boundary wavefield is evaluated at the true GLL coordinates of the
interface nodes (row ``nbl`` and column ``nbl``, ``nbl = neh*hpn``), either as
the closed-form incident plane wave of the bottom layer (``boundary="plane"``)
or as the total field of the flat 1-D model with its free surface and
interface reverberations (``boundary="layered"``, semsks.layered).  Both stand
in for the GRT producer, which needs Fortran and is not built.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

SEED = 20261203
F32 = np.float32


# --------------------------------------------------------------------------
# mesh geometry, computed with the reference's float/double promotions
# --------------------------------------------------------------------------
def gll_nodes64(hpn: int) -> np.ndarray:
    """Exact GLL nodes (float64) for degree hpn (3..6)."""
    if hpn == 3:
        r = math.sqrt(1 / 5)
        return np.array([-1, -r, r, 1.0])
    if hpn == 4:
        r = math.sqrt(3 / 7)
        return np.array([-1, -r, 0.0, r, 1.0])
    if hpn == 5:
        a, c, e = 315.0, -210.0, 15.0
    elif hpn == 6:
        a, c, e = 1386.0, -1260.0, 210.0
    else:
        raise ValueError("hpn must be in 3..6")
    d = math.sqrt(c * c - 4 * a * e)
    r1, r2 = math.sqrt((-c + d) / 2 / a), math.sqrt((-c - d) / 2 / a)
    mid = [0.0] if hpn == 6 else []
    return np.array([-1, -r1, -r2] + mid + [r2, r1, 1.0])


def mesh_dims(thickness: float, eh_in: float, nex: int, hpn: int):
    """nez, eh, nx, nz, h exactly as discret() forms them
    (semsyn/src/sem2d_psvsh.c:600-611): thickness is a float sum,
    nez=(int)((thk-1.e-3)/eh), eh=(float)((thk-1.e-3)/nez)."""
    thk = float(F32(thickness))
    eh32 = float(F32(eh_in))
    nez = int((thk - 1.0e-3) / eh32)
    eh = float(F32((thk - 1.0e-3) / nez))
    nx = nex * hpn + 1
    nz = nez * hpn + 1
    h = float(F32(F32(eh) / F32(hpn)))
    return nez, eh, nx, nz, h


def node_coords(nel: int, eh: float, hpn: int) -> np.ndarray:
    """True coordinates (km, float64) of the nel*hpn+1 GLL nodes along one axis."""
    xi = gll_nodes64(hpn)
    n = nel * hpn + 1
    i = np.arange(n)
    e = np.minimum(i // hpn, nel - 1)
    a = i - e * hpn
    return e * eh + (xi[a] + 1.0) / 2.0 * eh


def fit_bottom(top_layers, nez: int, eh: float) -> float:
    """Bottom-layer thickness for which discret()'s float arithmetic yields
    exactly nez elements of size eh (the total must sit just above nez*eh +
    1e-3 after float rounding)."""
    top = float_sum(top_layers)
    want = nez * eh + 1.0e-3
    x0 = F32(want - top)
    best = None
    x = x0
    for _ in range(4096):
        tot = float_sum(list(top_layers) + [float(x)])
        nz, e2, *_ = mesh_dims(tot, eh, 1, 4)
        if nz == nez:
            if e2 == float(F32(eh)):
                return float(x)
            if best is None:
                best = float(x)       # eh one float step away from the request
        x = np.nextafter(x, F32(np.inf))
    if best is not None:
        return best
    raise ValueError("no float thickness gives nez=%d eh=%g" % (nez, eh))


def float_sum(values) -> float:
    s = F32(0.0)
    for v in values:
        s = F32(s + F32(v))
    return float(s)


# --------------------------------------------------------------------------
# models
# --------------------------------------------------------------------------
@dataclass
class Layer:
    """One semmodel layer: constants (GPa-like units of vp^2*rho), rho, and
    the piecewise-linear top interface (x km, height km above the bottom)."""
    c: dict
    rho: float
    xs: list
    heights: list
    ratio: bool = False
    dip6: tuple | None = None   # (c14, c15, c34, c35, c46, c56) for a tilted axis


def iso_constants(vp: float, vs: float, rho: float) -> dict:
    """Isotropic constants as iso2sem.pl writes them
    (applications/test/iso2sem.pl:21-33)."""
    c11 = vp * vp * rho
    c44 = vs * vs * rho
    c13 = c11 - 2 * c44
    return dict(c11=c11, c12=c13, c13=c13, c16=0.0, c22=c11, c23=c13, c26=0.0,
                c33=c11, c36=0.0, c44=c44, c45=0.0, c55=c44, c66=c44)


def write_semmodel(path: str, layers: list[Layer], fmt: str = "%-9.4f") -> None:
    with open(path, "w") as f:
        f.write("%-4d\n" % len(layers))
        for L in layers:
            c = L.c
            np_ = len(L.xs) * (-1 if L.ratio else 1)
            head = [c["c11"], c["c12"], c["c13"], c["c16"], c["c22"], c["c23"], c["c26"]]
            tail = [c["c33"], c["c36"], c["c44"], c["c45"], c["c55"], c["c66"], L.rho]
            f.write("%-4d " % np_ + " ".join(fmt % v for v in head) + " ")
            f.write(" ".join(fmt % v for v in tail) + "\n")
            f.write(" ".join("%-11.4f" % v for v in L.xs) + "\n")
            f.write(" ".join("%-11.4f" % v for v in L.heights) + "\n")


def write_semdip(path: str, layers: list[Layer]) -> None:
    """Extension file: one line per semmodel layer with the six constants a
    tilted symmetry axis adds (c14 c15 c34 c35 c46 c56)."""
    with open(path, "w") as f:
        for L in layers:
            d = L.dip6 or (0.0,) * 6
            f.write(" ".join("%.6f" % v for v in d) + "\n")


def write_raymodel(path: str, thicknesses: list[float], vps, vss, rhos) -> None:
    """1-D layered model plus a one-ray list with nrec = number of layers.
    The SEM box spans layers 0..nrec-2 (sem2d_psvsh.c:591-599 with nfil=1)."""
    jo = len(thicknesses)
    with open(path, "w") as f:
        f.write("%d %d\n" % (jo, jo))
        for vp, vs, rho, th in zip(vps, vss, rhos, thicknesses):
            f.write("%.4f %.4f %.4f %.9g\n" % (vp, vs, rho, th))
        f.write("1\n")
        f.write("3 %d %d %d\n" % (jo - 1, jo, jo))   # nen=3, na[1] = nrec = jo
        f.write("1 1 1 1\n")


def write_recfile(path: str, xs) -> None:
    with open(path, "w") as f:
        f.write("%d\n" % len(xs))
        for x in xs:
            f.write("%.4f\n" % x)


def write_par(path: str, pars: dict) -> None:
    with open(path, "w") as f:
        for k, v in pars.items():
            f.write("%s=%s\n" % (k, v))


# --------------------------------------------------------------------------
# closed-form incident plane wave (stands in for the GRT producer)
# --------------------------------------------------------------------------
def ricker(tau: np.ndarray, f0: float) -> np.ndarray:
    a = (math.pi * f0) ** 2 * tau * tau
    return (1.0 - 2.0 * a) * np.exp(-a)


@dataclass
class PlaneWave:
    """Plane S wave (SV in the x-z plane, SH along y) travelling up and
    towards +x in the bottom layer: ray parameter p (s/km), S speed beta."""
    p: float
    beta: float
    f0: float = 0.2
    psi_deg: float = 0.0       # polarisation: SV share cos(psi), SH share sin(psi)
    amp: float = 1.0
    t_lead: float | None = None  # time from first motion to peak (default 1.2/f0)

    def fields(self, x: np.ndarray, z: np.ndarray, t: float):
        """(u, v, w_up) float64 at points (x, z) and time t."""
        s = self.p * self.beta
        if not (0.0 <= s < 1.0):
            raise ValueError("p*beta must be in [0,1)")
        ci = math.sqrt(1.0 - s * s)
        eta = ci / self.beta
        lead = self.t_lead if self.t_lead is not None else 1.2 / self.f0
        tau = t - lead - (self.p * x + eta * z)
        r = self.amp * ricker(tau, self.f0)
        psi = math.radians(self.psi_deg)
        sv = r * math.cos(psi)
        # SV polarisation perpendicular to (sin i, cos i): (cos i, -sin i)
        return sv * ci, r * math.sin(psi), -sv * s


def greenfile_records(pw: PlaneWave, nex: int, nez: int, eh: float, hpn: int,
                      neh: int, nt: int, dp: float):
    """Time-major P-SV and SH records of the plane wave at the interface nodes."""
    xs = node_coords(nex, eh, hpn)
    zs = node_coords(nez, eh, hpn)
    nx, nz = len(xs), len(zs)
    nbl = neh * hpn
    xb, zb = xs, np.full(nx, zs[nbl])          # bottom interface row iz=nbl
    xl, zl = np.full(nz, xs[nbl]), zs          # left interface column ix=nbl
    psv = np.empty((nt, 2 * (nx + nz)), dtype=F32)
    sh = np.empty((nt, nx + nz), dtype=F32)
    for it in range(nt):
        t = it * dp
        ub, vb, wb = pw.fields(xb, zb, t)
        ul, vl, wl = pw.fields(xl, zl, t)
        # greenfile w is positive down (the solver flips it)
        psv[it] = np.concatenate([ub, -wb, ul, -wl]).astype(F32)
        sh[it] = np.concatenate([vb, vl]).astype(F32)
    return psv, sh


# --------------------------------------------------------------------------
# a complete case directory
# --------------------------------------------------------------------------
@dataclass
class Case:
    nex: int = 200
    eh: float = 2.0
    hpn: int = 4
    neh: int = 1
    nst: int = 4000
    layers_thk: tuple = (20.0, 15.0, 165.001)   # top to bottom, km
    vp: tuple = (5.8, 6.5, 8.0)
    vs: tuple = (3.4, 3.75, 4.5)
    rho: tuple = (2.6, 2.9, 3.3)
    jitter: float = 0.05
    seed: int = SEED
    p: float = 0.06
    f0: float = 0.2
    psi_deg: float = 0.0
    cfl: float = 0.4
    sdt: float | None = None
    nrec: int = 50
    rec_x0: float | None = None
    rec_x1: float | None = None
    solver: int = 0            # 0 sem2d_psvsh, 1 sem2d (P-SV), 2 sem2dMP (Interface3)
    sh_zero: bool = True
    absorb: int = 1
    interfaces: list | None = None   # optional per-layer (xs, heights) overrides
    aniso: list | None = None        # optional per-layer (constants dict, dip6)
    extra: dict = field(default_factory=dict)
    nez: int | None = None           # if set, the bottom layer is fitted to give exactly nez elements
    boundary: str = "plane"          # "plane": the incident wave only; "layered": total field of the
                                     # flat 1-D model with its free surface (semsks.layered)

    def __post_init__(self):
        if self.nez is not None:
            top = list(self.layers_thk[:-1])
            self.layers_thk = tuple(top) + (fit_bottom(top, self.nez, self.eh),)

    def derived(self):
        rng = np.random.default_rng(self.seed)
        j = 1.0 + rng.uniform(-self.jitter, self.jitter, size=(len(self.vp), 3))
        vp = [float(F32(v * a)) for v, a in zip(self.vp, j[:, 0])]
        vs = [float(F32(v * a)) for v, a in zip(self.vs, j[:, 1])]
        rho = [float(F32(v * a)) for v, a in zip(self.rho, j[:, 2])]
        thk = float_sum(self.layers_thk)
        nez, eh, nx, nz, h = mesh_dims(thk, self.eh, self.nex, self.hpn)
        dxmin = (eh / 2.0) * (1.0 - math.sqrt(3.0 / 7.0)) if self.hpn == 4 else \
            (eh / 2.0) * (1.0 - abs(gll_nodes64(self.hpn)[1]))
        sdt = self.sdt if self.sdt is not None else float("%.6f" % (self.cfl * dxmin / max(vp)))
        return dict(vp=vp, vs=vs, rho=rho, nez=nez, eh=eh, nx=nx, nz=nz, h=h, sdt=sdt,
                    thickness=thk)

    def write(self, d: str, nt: int | None = None, write_records: bool = True) -> dict:
        os.makedirs(d, exist_ok=True)
        D = self.derived()
        vp, vs, rho = D["vp"], D["vs"], D["rho"]
        sdt = D["sdt"]
        nlay = len(self.layers_thk)
        # raymodel: layers top to bottom + a half-space; SEM spans all but the last
        write_raymodel(os.path.join(d, "raymodel"), list(self.layers_thk) + [0.0],
                       list(vp) + [vp[-1]], list(vs) + [vs[-1]], list(rho) + [rho[-1]])
        # semmodel: bottom up; the top of each layer is its interface height
        xmax = self.nex * D["eh"]
        tops = np.cumsum(list(reversed(self.layers_thk)))
        layers = []
        for k in range(nlay):
            src = nlay - 1 - k               # bottom layer first
            if self.interfaces and self.interfaces[k] is not None:
                xs, hs = self.interfaces[k]
            else:
                xs, hs = [0.0, xmax + self.eh], [float(tops[k])] * 2
            if self.aniso and self.aniso[k] is not None:
                c, dip6 = self.aniso[k]
            else:
                c, dip6 = iso_constants(vp[src], vs[src], rho[src]), None
            layers.append(Layer(c=c, rho=rho[src], xs=list(xs), heights=list(hs), dip6=dip6))
        write_semmodel(os.path.join(d, "semmodel"), layers)
        has_dip = any(L.dip6 is not None for L in layers)
        if has_dip:
            write_semdip(os.path.join(d, "semdip"), layers)
        # receivers on the free surface
        x0 = self.rec_x0 if self.rec_x0 is not None else 0.125 * xmax
        x1 = self.rec_x1 if self.rec_x1 is not None else 0.875 * xmax
        recs = np.linspace(x0, x1, self.nrec)
        write_recfile(os.path.join(d, "rec_file"), recs)
        # boundary records: dp = sdt, nt = nst + 3 (nt_sem = nst)
        nt = nt if nt is not None else self.nst + 3
        pw = PlaneWave(p=self.p, beta=vs[-1], f0=self.f0, psi_deg=self.psi_deg)
        if write_records:
            # record times use the float32 dp the solver reads
            if self.boundary == "layered":
                from .layered import layered_records
                psv, sh = layered_records(list(self.layers_thk), list(vp) + [vp[-1]], list(vs) + [vs[-1]],
                                          list(rho) + [rho[-1]], self.p, self.f0, self.psi_deg, self.nex,
                                          D["nez"], D["eh"], self.hpn, self.neh, nt, float(F32(sdt)))
            else:
                psv, sh = greenfile_records(pw, self.nex, D["nez"], D["eh"], self.hpn, self.neh, nt,
                                            float(F32(sdt)))
            if self.sh_zero:
                sh[:] = 0
        else:   # placeholders: the caller supplies the records another way
            psv = np.zeros((nt, 2 * (D["nx"] + D["nz"])), F32)
            sh = np.zeros((nt, D["nx"] + D["nz"]), F32)
        psv.tofile(os.path.join(d, "greenfile"))
        sh.tofile(os.path.join(d, "shgreenfile"))
        pars = dict(greenfile="greenfile", shgreenfile="shgreenfile", raymodel="raymodel",
                    semmodel="semmodel", nt=nt, dp="%.6f" % sdt, nex=self.nex,
                    eh="%.4f" % self.eh, xmin="0.0", nst=self.nst, sdt="%.6f" % sdt,
                    hpn=self.hpn, neh=self.neh, absorb=self.absorb, check=0, xbeg="0.0",
                    idxplot=1, msource=0, get_rec=1, rec_file="rec_file", plot_trace=1,
                    solver=self.solver)
        if has_dip:
            pars["semdip"] = "semdip"
        pars.update(self.extra)
        write_par(os.path.join(d, "run.par"), pars)
        D.update(nt=nt, layers=layers, recs=recs)
        return D


def config1(**kw) -> Case:
    """200x100 elements, 3 flat isotropic layers, plane SV at p=0.06,
    Ricker 0.2 Hz, 4000 steps, 50 surface receivers (P-SV solver)."""
    return Case(**kw)


def small_case(**kw) -> Case:
    """A few-second case for parity tests: 24x12 elements."""
    base = dict(nex=24, eh=2.0, layers_thk=(4.0, 4.0, 16.002), nst=600, nrec=8,
                p=0.06, f0=0.8)
    base.update(kw)
    return Case(**base)
