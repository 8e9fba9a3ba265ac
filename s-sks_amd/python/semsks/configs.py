"""The benchmark configurations as problem builders.

C1  200x100 elements, 3 flat isotropic layers, plane SV p=0.06, 4000 steps
C2  2000x500 elements (16,010,001 GLL points), upper-mantle hexagonal layer,
    4% (Thomsen eps = gamma), tilted axis (seeded dip 0-45 deg, azimuth from
    the 60 deg back-azimuth), SKS p=0.05, 20000 steps, 200 receivers
C3  64 instances on the C2 mesh with the C1-style isotropic model: 16 back-
    azimuths (polarisation split) x 4 ray parameters 0.04..0.07, 10000 steps
C4  96 instances, 200x100 .. 1600x400 elements, 1-4 layers, iso/hexagonal
    up to 6%, 2000..12000 steps
The model files go through the same reader as the CLI (semsks_load_par), so
the material per node is the reference's discret() assignment; the boundary
records come from the device plane-wave generator.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import tempfile

import numpy as np

from . import Model, Problem, inputs, aniso

SEED = inputs.SEED


def _model_dir(case: inputs.Case, d: str, nt_files: int = 4) -> Model:
    """Writes the case with a short greenfile (the real records are generated
    on the device) and loads it through semsks_load_par."""
    case.write(d, nt=nt_files, write_records=False)
    old = os.getcwd()
    os.chdir(d)
    try:
        return Model(["par=run.par"])
    finally:
        os.chdir(old)


def c1_case(**kw) -> inputs.Case:
    base = dict(nex=200, eh=2.0, layers_thk=(20.0, 15.0, 165.001), nst=4000, nrec=50, p=0.06, f0=0.2,
                solver=1)
    base.update(kw)
    return inputs.Case(**base)


def c2_case(seed: int = SEED, nst: int = 20000, **kw) -> inputs.Case:
    """4000 km x 1000 km box: crust 30 km, upper mantle 170 km with hexagonal
    anisotropy (the layer the SKS splitting samples), mantle below."""
    rng = np.random.default_rng(seed + 2)
    dip = float(rng.uniform(0.0, 45.0))
    baz = 60.0
    az = baz - 90.0        # rot_ela.pl:38, angle = angle_100 - 90
    vp, vs, rho = 8.1, 4.5, 3.35
    up = aniso.hexagonal_tilted(vp, vs, rho, 0.04, dip, az)
    base = dict(nex=2000, eh=2.0, nez=500, layers_thk=(30.0, 170.0, 800.0), vp=(6.2, 8.1, 9.0), vs=(3.6, 4.5, 5.0),
                rho=(2.8, 3.35, 3.6), jitter=0.0, nst=nst, nrec=200, p=0.05, f0=0.2, solver=0,
                aniso=[None, up, None], psi_deg=0.0)
    base.update(kw)
    c = inputs.Case(**base)
    c.meta = dict(dip=dip, azimuth=az, baz=baz)
    return c


def c2_model_c1_mesh(nst: int = 4000, seed: int = SEED, **kw) -> inputs.Case:
    """The C2 model -- crust 30 km, an upper-mantle layer with 4% hexagonal
    anisotropy whose axis has the C2 dip and azimuth, mantle below -- on the
    200x100-element mesh of C1 (400 km x 200 km): coupled P-SV+SH solver,
    plane S wave polarised 30 degrees off SV, 50 surface receivers.  Small
    enough for the FP64 oracle to step in minutes."""
    rng = np.random.default_rng(seed + 2)
    dip = float(rng.uniform(0.0, 45.0))
    az = 60.0 - 90.0
    up = aniso.hexagonal_tilted(8.1, 4.5, 3.35, 0.04, dip, az)
    base = dict(nex=200, eh=2.0, nez=100, layers_thk=(30.0, 120.0, 50.0), vp=(6.2, 8.1, 9.0), vs=(3.6, 4.5, 5.0),
                rho=(2.8, 3.35, 3.6), jitter=0.0, nst=nst, nrec=50, p=0.05, f0=0.2, solver=0,
                aniso=[None, up, None], psi_deg=30.0, sh_zero=False)
    base.update(kw)
    return inputs.Case(**base)


def c3_case(nst: int = 10000, **kw) -> inputs.Case:
    """The C2 mesh with the C1-style three-layer isotropic model."""
    base = dict(nex=2000, eh=2.0, nez=500, layers_thk=(20.0, 15.0, 965.0), nst=nst, nrec=200, p=0.05, f0=0.2,
                solver=0, sh_zero=False)
    base.update(kw)
    return inputs.Case(**base)


def c3_waves():
    """16 back-azimuths x 4 ray parameters."""
    return [(p, baz) for p in (0.04, 0.05, 0.06, 0.07) for baz in np.arange(16) * 22.5]


def c4_cases(n: int = 96, seed: int = SEED, nst_range=(2000, 12000), nex_range=(200, 1600), nez_range=(100, 400)):
    rng = np.random.default_rng(seed + 4)
    out = []
    for k in range(n):
        nex = int(rng.integers(nex_range[0], nex_range[1] + 1))
        nez = int(rng.integers(nez_range[0], nez_range[1] + 1))
        nl = int(rng.integers(1, 5))
        total = nez * 2.0
        cuts = np.sort(rng.uniform(0.05, 0.95, size=nl - 1)) * total
        edges = np.concatenate([[0.0], cuts, [total]])
        thk = tuple(float(round(x, 3)) for x in np.diff(edges)[::-1])      # top to bottom
        vp = tuple(float(v) for v in np.linspace(5.8, 8.5, nl))
        vs = tuple(float(v) for v in np.linspace(3.4, 4.7, nl))
        rho = tuple(float(v) for v in np.linspace(2.6, 3.4, nl))
        an = []
        for j in range(nl):
            if rng.uniform() < 0.5:
                src = nl - 1 - j
                an.append(aniso.hexagonal_tilted(vp[src], vs[src], rho[src], float(rng.uniform(0, 0.06)),
                                                 float(rng.uniform(0, 45)), float(rng.uniform(0, 360))))
            else:
                an.append(None)
        nst = int(rng.integers(nst_range[0], nst_range[1] + 1))
        p = float(rng.uniform(0.04, 0.07))
        psi = float(rng.uniform(0, 360))
        out.append(inputs.Case(nex=nex, eh=2.0, nez=nez, layers_thk=thk, vp=vp, vs=vs, rho=rho, jitter=0.0, nst=nst,
                               nrec=50, p=p, f0=0.2, solver=0, aniso=an, psi_deg=psi, sh_zero=False,
                               seed=seed + 100 + k))
    return out


def c5_requests(rate: float = 20.0, duration: float = 300.0, seed: int = SEED, nex_range=(200, 1600),
                nez_range=(100, 400), nst_range=(2000, 12000)):
    """Poisson arrivals at `rate` per second over `duration` seconds, each a
    C4-style instance (mesh size, 1-4 layers, iso/hexagonal, p, polarisation
    and length drawn from the seed).  Returns [(arrival_s, case)]."""
    rng = np.random.default_rng(seed + 5)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        out.append(t)
    cases = c4_cases(len(out), seed=seed + 50, nst_range=nst_range, nex_range=nex_range, nez_range=nez_range)
    return list(zip(out, cases))


def c5_sample(n: int, seed: int = SEED, frac: float = 0.05):
    """The seeded sample of C5 requests whose FP32 traces are checked against
    the FP64 oracle: request i is in it when the (seed + 55) stream's i-th
    uniform draw is below frac.  The draw of a request does not depend on
    how many requests the window holds."""
    u = np.random.default_rng(seed + 55).random(max(n, 1))
    return [i for i in range(n) if u[i] < frac]


def device_problem(model: Model, nt: int) -> Problem:
    """Copy of a loaded problem whose boundary records come from the next
    semsks_planewave(..., to_host=False) call on the engine."""
    p = Problem.from_buffer_copy(model.problem)
    p.psv = C.POINTER(C.c_float)()
    p.sh = C.POINTER(C.c_float)()
    p.nt = nt
    p.dp = p.sdt
    return p


class Built:
    """A case loaded through the CLI reader, kept alive with its files."""

    def __init__(self, case: inputs.Case):
        self.case = case
        self.tmp = tempfile.TemporaryDirectory()
        self.model = _model_dir(case, self.tmp.name)
        self.D = case.derived()

    def problem(self, nt: int | None = None) -> Problem:
        nt = nt if nt is not None else self.case.nst + 3
        return device_problem(self.model, nt)

    def planewave_args(self, nt: int, p=None, psi=None):
        c, pr = self.case, self.model.problem
        beta = self.D["vs"][-1]
        return (pr.nex, pr.nez, pr.hpn, pr.neh, pr.eh, nt, pr.sdt, c.p if p is None else p, beta, c.f0,
                c.psi_deg if psi is None else psi)

    def nglob(self) -> int:
        return self.model.nx * self.model.nz
