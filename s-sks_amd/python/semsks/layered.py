"""Boundary wavefield of a plane S wave in a flat layered medium with a free
surface, in the reference's greenfile/shgreenfile layout.

The reference fills these files with generalised-ray synthetics of the 1-D
model (aserpsv_sem.c:415-470, TOTAL rays: every reverberation between the
free surface and the layer interfaces).  For an incident plane wave the
same total field follows from the propagator-matrix (Haskell 1953,
Thomson 1950) solution of the layered half-space, frequency by frequency:

  * in each layer the motion-stress vector f = (u_x, u_z, t_zx, t_zz)
    (z down, tractions divided by -i*omega) is a sum of up- and down-going
    P and SV waves with the horizontal slowness p of the incident wave;
    f(z + h) = F E(h) F^-1 f(z) across a layer of thickness h, E the
    diagonal of vertical phase factors;
  * the free surface has zero traction, so f(0) = (u0, w0, 0, 0);
    propagated to the top of the half-space and split into its four
    waves, the up-going SV amplitude must be the incident one and the
    up-going P must vanish: two equations for (u0, w0);
  * SH is the 2x2 analogue with (u_y, t_zy).

Time convention e^{+i omega t} (numpy's), so a wave with vertical slowness
q carries e^{-i omega q z}.  The incident wave matches inputs.PlaneWave: a
Ricker pulse (peak `amp`, centre frequency f0) whose peak passes height z
above the box bottom and abscissa x at t = lead + p*x + eta*z in the
bottom layer, SV polarised (cos i, -sin i) in (x, z up), SH share sin(psi).
"""
from __future__ import annotations

import math

import numpy as np

from .inputs import F32, node_coords


def _psv_matrix(p, a, b, rho):
    """Columns: P_down, P_up, S_down, S_up unit waves at the layer top;
    returns (F, q) with q the signed vertical slownesses."""
    xi = np.sqrt(complex(1.0 / (a * a) - p * p))
    eta = np.sqrt(complex(1.0 / (b * b) - p * p))
    mu = rho * b * b
    lam = rho * a * a - 2.0 * mu
    cols, qs = [], []
    for d, q in (((a * p, a * xi), xi), ((a * p, -a * xi), -xi), ((b * eta, -b * p), eta), ((b * eta, b * p), -eta)):
        dx, dz = d
        cols.append([dx, dz, mu * (q * dx + p * dz), lam * (p * dx + q * dz) + 2.0 * mu * q * dz])
        qs.append(q)
    return np.array(cols, dtype=complex).T, np.array(qs)


def _sh_matrix(p, b, rho):
    eta = np.sqrt(complex(1.0 / (b * b) - p * p))
    mu = rho * b * b
    F = np.array([[1.0, 1.0], [mu * eta, -mu * eta]], dtype=complex)   # down, up
    return F, np.array([eta, -eta])


class LayeredResponse:
    """Transfer functions of the layered half-space for a unit up-going SV
    (or SH) wave.  thk: layer thicknesses top to bottom (km); vp, vs, rho:
    one entry per layer plus the half-space last."""

    def __init__(self, thk, vp, vs, rho, p):
        assert len(vp) == len(vs) == len(rho) == len(thk) + 1
        self.thk = [float(h) for h in thk]
        self.vp, self.vs, self.rho = map(lambda v: [float(x) for x in v], (vp, vs, rho))
        self.p = float(p)
        self.tops = np.concatenate([[0.0], np.cumsum(self.thk)])        # layer tops; last = half-space top
        self.psv = [_psv_matrix(self.p, a, b, r) for a, b, r in zip(self.vp, self.vs, self.rho)]
        self.sh = [_sh_matrix(self.p, b, r) for b, r in zip(self.vs, self.rho)]

    def _layer_of(self, depth):
        j = int(np.searchsorted(self.tops[1:], depth, side="right"))
        return min(j, len(self.thk))

    @staticmethod
    def _prop(F, q, w, h):
        E = np.exp(-1j * w * q * h)
        return F @ np.diag(E) @ np.linalg.inv(F)

    def _surface_to(self, mats, w, depth, n):
        """Propagator from the surface to `depth` at angular frequency w."""
        P = np.eye(n, dtype=complex)
        z = 0.0
        for j, h in enumerate(self.thk):
            if depth <= z + h:
                return self._prop(*mats[j], w, depth - z) @ P
            P = self._prop(*mats[j], w, h) @ P
            z += h
        return self._prop(*mats[-1], w, depth - z) @ P

    def psv_transfer(self, w, depths):
        """(u_x, u_z down) at each depth per unit up-going SV amplitude at the
        half-space top, at angular frequency w (scalar, > 0)."""
        H = self.tops[-1]
        PH = self._surface_to(self.psv, w, H, 4)
        Fh, _ = self.psv[-1]
        A = np.linalg.inv(Fh) @ PH[:, :2]                # half-space amplitudes per unit (u0, w0)
        # rows: P_up (1) = 0, S_up (3) = 1
        M = np.array([A[1], A[3]])
        u0w0 = np.linalg.solve(M, np.array([0.0, 1.0], dtype=complex))
        f0 = np.array([u0w0[0], u0w0[1], 0.0, 0.0], dtype=complex)
        out = np.empty((len(depths), 2), dtype=complex)
        for k, d in enumerate(depths):
            out[k] = (self._surface_to(self.psv, w, d, 4) @ f0)[:2]
        return out

    def sh_transfer(self, w, depths):
        H = self.tops[-1]
        PH = self._surface_to(self.sh, w, H, 2)
        Fh, _ = self.sh[-1]
        A = np.linalg.inv(Fh) @ PH[:, 0]                 # per unit v0
        v0 = 1.0 / A[1]
        out = np.empty(len(depths), dtype=complex)
        for k, d in enumerate(depths):
            out[k] = (self._surface_to(self.sh, w, d, 2) @ np.array([v0, 0.0]))[0]
        return out


def ricker_spectrum(nfft, dp, f0):
    """dp-scaled discrete transform of the zero-centred Ricker pulse (peak
    1) sampled every dp on an nfft-point circular grid: irfft(S)/dp gives
    the samples back, and a delay tau multiplies S by e^{-i w tau}."""
    n = np.arange(nfft)
    t = np.where(n < nfft // 2, n, n - nfft) * dp
    a = (math.pi * f0 * t) ** 2
    return np.fft.rfft((1.0 - 2.0 * a) * np.exp(-a)) * dp


def layered_records(thk, vp, vs, rho, p, f0, psi_deg, nex, nez, eh, hpn, neh, nt, dp,
                    amp=1.0, t_lead=None, nfft=None):
    """Time-major P-SV and SH greenfile records (the layout of
    inputs.greenfile_records: u_bot | w_bot | u_left | w_left, w positive
    down; v_bot | v_left) of the total field of the layered medium at the
    interface nodes.  thk/vp/vs/rho as LayeredResponse; the SEM box spans
    depths 0 .. nez*eh from the surface."""
    xs = node_coords(nex, eh, hpn)
    zs = node_coords(nez, eh, hpn)               # height above the box bottom
    nx, nz = len(xs), len(zs)
    nbl = neh * hpn
    Hbox = float(zs[-1])
    depths_l = Hbox - zs                          # left column, every node
    depth_b = Hbox - float(zs[nbl])               # bottom interface row
    lead = t_lead if t_lead is not None else 1.2 / f0
    b_h = float(vs[-1])
    eta_h = math.sqrt(1.0 / (b_h * b_h) - p * p)
    resp = LayeredResponse(thk, vp, vs, rho, p)
    H = resp.tops[-1]
    if nfft is None:
        span = nt * dp + lead + p * float(xs[-1]) + 4.0 * H / min(vs) + 4.0 / f0
        nfft = 1 << int(math.ceil(math.log2(max(2 * nt, span / dp))))
    w = 2.0 * math.pi * np.fft.rfftfreq(nfft, dp)
    psi = math.radians(psi_deg)
    # incident up-going S at the half-space top: peak at t = lead + eta_h * (Hbox - H)
    src = amp * ricker_spectrum(nfft, dp, f0) * np.exp(-1j * w * (lead + eta_h * (Hbox - H)))
    depths = np.concatenate([[depth_b], depths_l])
    U = np.zeros((len(w), len(depths)), complex)
    W = np.zeros_like(U)
    V = np.zeros_like(U)
    fmax = 4.0 * f0
    for k, wk in enumerate(w):
        if wk <= 0.0 or wk > 2.0 * math.pi * fmax:
            continue
        if math.cos(psi) != 0.0:
            t = resp.psv_transfer(wk, depths)
            U[k], W[k] = t[:, 0], t[:, 1]
        if math.sin(psi) != 0.0:
            V[k] = resp.sh_transfer(wk, depths)
    base_u = src[:, None] * U * math.cos(psi)
    base_w = src[:, None] * W * math.cos(psi)
    base_v = src[:, None] * V * math.sin(psi)

    def to_time(spec, x):
        """time series of the spectrum at abscissa x (phase p*x), nt samples"""
        ph = np.exp(-1j * np.outer(w, np.atleast_1d(x) * p))
        return np.fft.irfft(spec * ph, nfft, axis=0)[:nt] / dp

    psv = np.empty((nt, 2 * (nx + nz)), dtype=F32)
    sh = np.empty((nt, nx + nz), dtype=F32)
    # bottom row: one depth, every x
    ub = to_time(np.repeat(base_u[:, :1], nx, axis=1), xs)
    wb = to_time(np.repeat(base_w[:, :1], nx, axis=1), xs)
    vb = to_time(np.repeat(base_v[:, :1], nx, axis=1), xs)
    # left column: every depth at x = xs[nbl]
    xl = float(xs[nbl])
    ul = to_time(base_u[:, 1:], np.full(nz, xl))
    wl = to_time(base_w[:, 1:], np.full(nz, xl))
    vl = to_time(base_v[:, 1:], np.full(nz, xl))
    # z down already: the greenfile w is positive down
    psv[:] = np.concatenate([ub, wb, ul, wl], axis=1).astype(F32)
    sh[:] = np.concatenate([vb, vl], axis=1).astype(F32)
    return psv, sh
