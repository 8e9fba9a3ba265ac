"""The layered-medium boundary producer (semsks.layered): propagator-matrix
total field of a plane S wave in a flat layered half-space with a free
surface, written in the greenfile layout.  Pinned by what the mathematics
fixes:
  * a homogeneous half-space: incident + free-surface reflection (SH, and SV
    at vertical incidence), sample by sample;
  * one layer over a half-space, SH at vertical incidence: the textbook
    surface amplification 2 / sqrt(cos^2(w h/b1) + (Z1/Z2)^2 sin^2(w h/b1));
  * P-SV at oblique incidence in a layered stack: the free surface reflects
    all the incident energy (up-going SV flux = down-going P + SV flux in
    the half-space), and the surface traction vanishes.
"""
import math

import numpy as np
import pytest

from semsks import inputs, layered


def test_homogeneous_half_space_closed_form():
    b, a, r = 3.5, 6.0, 2.7
    thk, vp, vs, rho = [10.0, 14.0], [a] * 3, [b] * 3, [r] * 3
    f0, nex, nez, eh, hpn, neh, nt, dp = 0.4, 20, 12, 2.0, 4, 1, 800, 0.02
    xs, zs = inputs.node_coords(nex, eh, hpn), inputs.node_coords(nez, eh, hpn)
    nx, nz, H = len(xs), len(zs), zs[-1]
    t = np.arange(nt) * dp
    lead = 1.2 / f0
    for p, psi in ((0.05, 90.0), (0.0, 0.0)):
        psv, sh = layered.layered_records(thk, vp, vs, rho, p, f0, psi, nex, nez, eh, hpn, neh, nt, dp)
        eta = math.sqrt(1 / b ** 2 - p * p)
        rec = sh[:, nx:] if psi == 90.0 else psv[:, 2 * nx:2 * nx + nz]
        for iz in (0, 10, 30, nz - 1):
            z, x = zs[iz], xs[hpn * neh]
            ana = inputs.ricker(t - lead - p * x - eta * z, f0) + inputs.ricker(t - lead - p * x - eta * (2 * H - z), f0)
            assert np.abs(rec[:, iz] - ana).max() < 1e-5
        if psi == 0.0:
            assert np.abs(psv[:, 2 * nx + nz:]).max() < 1e-6        # no vertical motion at p = 0
            z = zs[hpn * neh]
            ana = inputs.ricker(t[:, None] - lead - eta * z, f0) + inputs.ricker(t[:, None] - lead - eta * (2 * H - z), f0)
            assert np.abs(psv[:, :nx] - ana).max() < 1e-5              # bottom row


def test_single_layer_sh_amplification():
    h, b1, r1, b2, r2 = 5.0, 2.0, 2.2, 4.0, 3.0
    resp = layered.LayeredResponse([h], [3.5, 7.0], [b1, b2], [r1, r2], 0.0)
    for f in (0.05, 0.1, 0.17, 0.3, 0.55):
        w = 2 * math.pi * f
        v = resp.sh_transfer(w, [0.0])[0]
        # the unit up-going wave is referenced at the half-space top (depth h)
        k = w * h / b1
        amp = 2.0 / math.sqrt(math.cos(k) ** 2 + (r1 * b1 / (r2 * b2)) ** 2 * math.sin(k) ** 2)
        assert abs(abs(v) - amp) < 1e-10 * amp


@pytest.mark.parametrize("p", [0.03, 0.06, 0.09])
def test_psv_energy_returns_and_surface_is_traction_free(p):
    thk = [3.0, 7.5, 12.0]
    vp, vs, rho = [5.2, 6.3, 7.1, 8.1], [3.0, 3.6, 4.0, 4.5], [2.5, 2.8, 3.1, 3.4]
    resp = layered.LayeredResponse(thk, vp, vs, rho, p)
    a, b, r = vp[-1], vs[-1], rho[-1]
    xi, eta = math.sqrt(1 / a ** 2 - p * p), math.sqrt(1 / b ** 2 - p * p)
    for f in (0.07, 0.2, 0.45):
        w = 2 * math.pi * f
        H = resp.tops[-1]
        PH = resp._surface_to(resp.psv, w, H, 4)
        Fh, _ = resp.psv[-1]
        uw = resp.psv_transfer(w, [0.0])[0]
        f0 = np.array([uw[0], uw[1], 0, 0], complex)
        A = np.linalg.solve(Fh, PH @ f0)            # P_down, P_up, S_down, S_up at the half-space top
        assert abs(A[1]) < 1e-10 and abs(A[3] - 1) < 1e-10
        inc = r * b * b * eta * abs(A[3]) ** 2
        out = r * a * a * xi * abs(A[0]) ** 2 + r * b * b * eta * abs(A[2]) ** 2
        assert abs(out - inc) < 1e-9 * inc
        # traction at the surface from the field just below it
        fz = resp._surface_to(resp.psv, w, 1e-9, 4) @ f0
        assert abs(fz[2]) < 1e-6 * abs(fz[0]) and abs(fz[3]) < 1e-6 * abs(fz[0])


def test_records_layout_and_reverberations():
    """A low-velocity surface layer rings: the left-column surface record
    carries more than the single doubled pulse, with the greenfile shapes."""
    thk, vp, vs, rho = [4.0, 20.0], [4.5, 8.0, 8.0], [2.5, 4.5, 4.5], [2.3, 3.3, 3.3]
    nex, nez, eh, hpn, neh, nt, dp = 12, 12, 2.0, 4, 1, 1500, 0.02
    psv, sh = layered.layered_records(thk, vp, vs, rho, 0.0, 0.5, 45.0, nex, nez, eh, hpn, neh, nt, dp)
    nx, nz = nex * hpn + 1, nez * hpn + 1
    assert psv.shape == (nt, 2 * (nx + nz)) and sh.shape == (nt, nx + nz)
    surf = sh[:, nx + nz - 1]
    assert np.abs(surf).max() > 1.5 * math.sin(math.radians(45.0))   # amplified by the slow layer
    peaks = np.sum((np.abs(surf[1:-1]) > 0.2 * np.abs(surf).max()) &
                   (np.abs(surf[1:-1]) >= np.abs(surf[:-2])) & (np.abs(surf[1:-1]) >= np.abs(surf[2:])))
    assert peaks >= 4                                                  # reverberations
