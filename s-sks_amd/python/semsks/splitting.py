"""Shear-wave splitting of the synthetic SKS traces (the cal_split.pl step,
applications/test/cal_split.pl:1-60, which hands each station's SAC files
to an external Silver & Chan (1991) program, an_split, that the reference
does not ship).

This module measures the splitting parameters itself, by the published
minimum-transverse-energy method of Silver & Chan (1991, JGR 96, 16429):

  for each trial fast azimuth phi and delay dt, rotate the horizontal
  motion into fast and slow directions, advance the slow component by dt,
  rotate back to radial/transverse and sum the transverse energy over the
  window; the pair with the least energy is the measurement.  The 95%
  confidence region is E <= E_min * (1 + k/(nu-k) * F_{k,nu-k}(0.95)),
  k = 2, with nu the degrees of freedom of the transverse trace (the
  second-moment estimate of its spectrum); the reported errors are half
  the extent of that region in phi and in dt, divided by 2 (one sigma).

Conventions: east = u (seis.u, the .ce SAC file), north = v (.cn); a
back-azimuth of 270 degrees (the uvw2sac default) makes the radial
direction east, i.e. the wave travelling in +x of the 2-D section.
Azimuths are degrees clockwise from north.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import sac


def rotate_rt(n: np.ndarray, e: np.ndarray, baz_deg: float):
    """(radial, transverse) of a wave arriving from back-azimuth baz: radial
    points along the propagation direction (away from the source), the
    transverse 90 degrees clockwise from it."""
    az = math.radians(baz_deg + 180.0)                 # propagation azimuth
    r = n * math.cos(az) + e * math.sin(az)
    t = -n * math.sin(az) + e * math.cos(az)
    return r, t


def split(n: np.ndarray, e: np.ndarray, dt: float, phi_deg: float, delay: float):
    """Applies splitting: the component along phi (fast) arrives first, the
    one across it (slow) `delay` seconds later."""
    p = math.radians(phi_deg)
    f = n * math.cos(p) + e * math.sin(p)
    s = -n * math.sin(p) + e * math.cos(p)
    s = _shift(s, delay / dt)
    return f * math.cos(p) - s * math.sin(p), f * math.sin(p) + s * math.cos(p)


def _shift(x: np.ndarray, lag: float) -> np.ndarray:
    """x(t - lag*dt) by a linear-phase shift on a zero-padded transform."""
    n = len(x)
    m = 1 << int(math.ceil(math.log2(2 * n)))
    X = np.fft.rfft(x, m)
    w = np.exp(-2j * math.pi * np.fft.rfftfreq(m) * lag)
    return np.fft.irfft(X * w, m)[:n]


def degrees_of_freedom(t: np.ndarray) -> float:
    """nu of a trace from its amplitude spectrum (Silver & Chan 1991, eq. A12
    style second-moment estimate): nu = 2 * (sum F^2)^2 / sum F^4 - 1,
    F the amplitude spectrum."""
    F = np.abs(np.fft.rfft(t))
    f2 = np.sum(F ** 2)
    f4 = np.sum(F ** 4)
    return float(2.0 * f2 * f2 / max(f4, 1e-300) - 1.0)


def silver_chan(n: np.ndarray, e: np.ndarray, dt: float, baz_deg: float, window=None,
                phis=None, delays=None):
    """Grid search; window = (i0, i1) sample indices.  Returns dict with phi,
    d_phi, delay, d_delay, E (energy map [phi, delay]), nu, e_min, e_95."""
    from scipy.stats import f as fdist
    n = np.asarray(n, np.float64)
    e = np.asarray(e, np.float64)
    i0, i1 = window if window is not None else (0, len(n))
    phis = np.arange(-90.0, 90.0, 1.0) if phis is None else np.asarray(phis, float)
    delays = np.arange(0.0, 4.0 + 1e-9, 0.05) if delays is None else np.asarray(delays, float)
    E = np.zeros((len(phis), len(delays)))
    for a, phi in enumerate(phis):
        p = math.radians(phi)
        f = n * math.cos(p) + e * math.sin(p)
        s0 = -n * math.sin(p) + e * math.cos(p)
        for b, dl in enumerate(delays):
            s = _shift(s0, -dl / dt)                       # undo the delay of the slow wave
            nn = f * math.cos(p) - s * math.sin(p)
            ee = f * math.sin(p) + s * math.cos(p)
            _, t = rotate_rt(nn, ee, baz_deg)
            E[a, b] = np.sum(t[i0:i1] ** 2)
    a, b = np.unravel_index(np.argmin(E), E.shape)
    phi, dl = float(phis[a]), float(delays[b])
    # transverse trace at the minimum for the degrees of freedom
    p = math.radians(phi)
    f = n * math.cos(p) + e * math.sin(p)
    s = _shift(-n * math.sin(p) + e * math.cos(p), -dl / dt)
    _, t = rotate_rt(f * math.cos(p) - s * math.sin(p), f * math.sin(p) + s * math.cos(p), baz_deg)
    nu = max(degrees_of_freedom(t[i0:i1]), 3.0)
    k = 2
    e95 = E[a, b] * (1.0 + k / (nu - k) * fdist.ppf(0.95, k, nu - k))
    inside = E <= e95
    pa = phis[inside.any(axis=1)]
    pd = delays[inside.any(axis=0)]
    return dict(phi=phi, delay=dl, d_phi=float((pa.max() - pa.min()) / 4.0), d_delay=float((pd.max() - pd.min()) / 4.0),
                E=E, nu=nu, e_min=float(E[a, b]), e_95=float(e95), phis=phis, delays=delays)


def measure_dir(d: str, before: float = 10.0, after: float = 30.0, baz: float = 270.0, out: str = "split_par.txt",
                names=None, arrivals=None):
    """cal_split.pl over the SAC files of a uvw2sac run in directory d: for
    each station k, the window [t_SKS - before, t_SKS + after] of k.ce/k.cn
    (t_SKS from `arrivals` if given -- cal_split.pl reads it from a travel-
    time table -- else the time of the largest horizontal amplitude), and
    one line per station in split_par.txt:
    name gcarc phi d_phi dt d_dt  ("%-5s %-7.2f %-6.1f %-6.1f %-7.2f %-7.2f")."""
    k = 0
    rows = []
    while os.path.exists(os.path.join(d, "%d.ce" % k)):
        he, e = sac.read(os.path.join(d, "%d.ce" % k))
        hn, n = sac.read(os.path.join(d, "%d.cn" % k))
        dt = he["delta"]
        if arrivals is not None:
            tk = arrivals[k]
        else:
            tk = float(np.argmax(np.hypot(e, n))) * dt
        i0 = max(0, int(round((tk - before) / dt)))
        i1 = min(len(e), int(round((tk + after) / dt)))
        r = silver_chan(n, e, dt, baz, (i0, i1))
        name = names[k] if names else str(k)
        rows.append((name, he["gcarc"], r["phi"], r["d_phi"], r["delay"], r["d_delay"]))
        k += 1
    with open(os.path.join(d, out), "w") as f:
        for row in rows:
            f.write("%-5s %-7.2f %-6.1f %-6.1f %-7.2f %-7.2f\n" % row)
    return rows
