"""Pins the oracle's material fill and edge sponge against things that do
not restate them.

* Elastic constants per node (discret's semmodel fill, Cijkl_3D and
  inv_index, sem2d_psvsh.c:733-797, :1532-1572, :1584-1597, plus the
  tilted-axis constants of the semdip file).  In a homogeneous medium the
  SEM operator applied to a long plane wave u = a sin(k n.x) gives, at nodes
  away from the domain edges, (K u)_i / M_i = k^2 (Gamma(n) a / rho) sin(k n.x_i),
  with Gamma_ik(n) = C_ijkl n_j n_l the Christoffel matrix.  Gamma is
  computed here from the transversely isotropic tensor written in closed
  form from its symmetry axis and five moduli (Thomsen parameters), with no
  Voigt table and no rotation matrix, for propagation along x, along z and
  at 45 and 30 degrees: together these fix all 15 constants that enter 2-D
  stress, and rho through the mass.
* Layer cuts (discret + lip, sem2d_psvsh.c:733-797, :831-846): every node
  more than 2e-3 km from every interface carries the record of the layer
  that its height lies in, read from the interface polylines by
  interpolation.
* The sponge (atten_taper, atten_wave2, sem2d_psvsh.c:1725-1738,
  :1780-1823): the factors against the closed form exp(n ln(ass) /
  (nh ln 2.73)), and one application to a field of ones against the factor
  of the strip each node lies in, from the node's distance to the edge.
"""
import math
import os

import numpy as np
import pytest

import pyoracle
from semsks import aniso, inputs


# ------------------------------------------------------------ Christoffel
def ti_moduli(vp, vs, rho, eps, gamma, delta):
    """Moduli of a transversely isotropic medium from the Thomsen (1986)
    definitions: vp, vs along the axis, eps, gamma, delta."""
    c33 = rho * vp * vp
    c44 = rho * vs * vs
    c11 = c33 * (1 + 2 * eps)
    c66 = c44 * (1 + 2 * gamma)
    # delta = ((c13 + c44)^2 - (c33 - c44)^2) / (2 c33 (c33 - c44))
    c13 = math.sqrt(2 * delta * c33 * (c33 - c44) + (c33 - c44) ** 2) - c44
    return c11, c33, c13, c44, c66


def ti_tensor(c11, c33, c13, c44, c66, s):
    """C_ijkl of a transversely isotropic medium with unit axis s:
    lam d_ij d_kl + mu (d_ik d_jl + d_il d_jk) + a (d_ij s_k s_l + s_i s_j d_kl)
    + b (d_ik s_j s_l + d_il s_j s_k + d_jk s_i s_l + d_jl s_i s_k) + c s_i s_j s_k s_l."""
    d = np.eye(3)
    s = np.asarray(s, float)
    lam, mu = c11 - 2 * c66, c66
    a = c13 - lam
    b = c44 - c66
    c = c11 + c33 - 2 * c13 - 4 * c44
    e = np.einsum
    return (lam * e("ij,kl->ijkl", d, d) + mu * (e("ik,jl->ijkl", d, d) + e("il,jk->ijkl", d, d))
            + a * (e("ij,k,l->ijkl", d, s, s) + e("i,j,kl->ijkl", s, s, d))
            + b * (e("ik,j,l->ijkl", d, s, s) + e("il,j,k->ijkl", d, s, s) + e("jk,i,l->ijkl", d, s, s)
                   + e("jl,i,k->ijkl", d, s, s))
            + c * e("i,j,k,l->ijkl", s, s, s, s))


def test_ti_tensor_reduces_to_its_moduli():
    """The closed form itself: axis along z gives the textbook VTI matrix."""
    c11, c33, c13, c44, c66 = 250.0, 210.0, 80.0, 60.0, 70.0
    T = ti_tensor(c11, c33, c13, c44, c66, (0, 0, 1))
    assert T[0, 0, 0, 0] == pytest.approx(c11) and T[1, 1, 1, 1] == pytest.approx(c11)
    assert T[2, 2, 2, 2] == pytest.approx(c33)
    assert T[0, 0, 2, 2] == pytest.approx(c13) and T[1, 1, 2, 2] == pytest.approx(c13)
    assert T[0, 0, 1, 1] == pytest.approx(c11 - 2 * c66)
    assert T[1, 2, 1, 2] == pytest.approx(c44) and T[0, 2, 0, 2] == pytest.approx(c44)
    assert T[0, 1, 0, 1] == pytest.approx(c66)
    assert abs(T[0, 0, 0, 2]) < 1e-12 and abs(T[0, 1, 2, 2]) < 1e-12
    # isotropic limit
    lam, mu = 40.0, 30.0
    Ti = ti_tensor(lam + 2 * mu, lam + 2 * mu, lam, mu, mu, (0.6, 0.0, 0.8))
    d = np.eye(3)
    iso = lam * np.einsum("ij,kl->ijkl", d, d) + mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d))
    assert np.allclose(Ti, iso)


def christoffel_from_oracle(o, theta_deg, wavelength):
    """Gamma(n)/rho [i][k] measured from the oracle's K and lumped mass with
    the plane waves u = e_k sin(k n.x), n = (cos theta, 0, sin theta)."""
    xs = inputs.node_coords(o.nex, o.eh, o.hpn)
    zs = inputs.node_coords(o.nez, o.eh, o.hpn)
    X, Z = np.meshgrid(xs, zs)
    th = math.radians(theta_deg)
    kw = 2 * math.pi / wavelength
    phase = np.sin(kw * (math.cos(th) * X + math.sin(th) * Z))
    m, _, _ = o.mass()
    margin = 2 * o.hpn
    sl = (slice(margin, o.nz - margin), slice(margin, o.nx - margin))
    s = phase[sl].ravel()
    G = np.zeros((3, 3))
    worst = 0.0
    for k in range(3):
        u = np.zeros((3, o.nz, o.nx))
        u[k] = phase
        f = o.stiff(u)
        for i in range(3):
            h = (f[i] / m)[sl].ravel() / kw ** 2
            G[i, k] = h @ s / (s @ s)
            worst = max(worst, np.abs(h - G[i, k] * s).max())
    return G, worst


@pytest.mark.parametrize("dip,az,delta", [(0.0, 0.0, 0.0), (30.0, 40.0, 0.0), (52.0, 205.0, 0.03)])
def test_material_fill_gives_christoffel_matrix(tmp_path, dip, az, delta):
    vp, vs, rho, eps, gam = 8.1, 4.5, 3.35, 0.04, 0.06
    layer = aniso.hexagonal_tilted(vp, vs, rho, eps, dip, az, delta=delta)
    if gam != eps:          # the builder ties gamma to eps; build this medium explicitly
        C0 = aniso.hexagonal_vertical(vp, vs, rho, eps, gam, delta)
        layer = aniso.layer_entry(aniso.rotate(C0, aniso.rotation(dip, az)))
    c = inputs.Case(nex=40, nez=20, eh=2.0, layers_thk=(40.0,), vp=(vp,), vs=(vs,), rho=(rho,), jitter=0.0,
                    nst=2, nrec=2, aniso=[layer], solver=0)
    d = os.fspath(tmp_path)
    c.write(d)
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    rho32 = float(np.float32(rho))
    mod = ti_moduli(vp, vs, rho, eps, gam, delta)
    dr, ar = math.radians(dip), math.radians(az)
    axis = (math.sin(dr) * math.cos(ar), math.sin(dr) * math.sin(ar), math.cos(dr))
    T = ti_tensor(*mod, axis)
    for theta in (0.0, 90.0, 45.0, 30.0):
        th = math.radians(theta)
        n = np.array([math.cos(th), 0.0, math.sin(th)])
        want = np.einsum("ijkl,j,l->ik", T, n, n) / rho32
        got, resid = christoffel_from_oracle(o, theta, wavelength=40.0)
        scale = np.abs(want).max()
        # the semmodel file holds the constants to 1e-4 (iso2sem.pl's %9.4f)
        assert np.abs(got - want).max() < 2e-5 * scale, (theta, got, want)
        assert resid < 1e-4 * scale
        # phase speeds: eigenvalues of Gamma / rho
        assert np.allclose(np.sort(np.linalg.eigvalsh(0.5 * (got + got.T))), np.sort(np.linalg.eigvalsh(want)),
                           rtol=2e-5)


# ------------------------------------------------------------ layer cuts
def interface_height(xs, hs, x):
    return np.interp(x, xs, hs)


@pytest.mark.parametrize("dipping", [False, True])
def test_layer_cut_follows_interface_geometry(tmp_path, dipping):
    nex, nez, eh = 30, 15, 2.0
    xmax = nex * eh
    ifc = None
    if dipping:
        ifc = [([0.0, 17.3, 41.0, xmax + eh], [4.1, 9.7, 6.3, 11.2]),       # bottom layer's top: a kinked polyline
               ([0.0, xmax + eh], [14.05, 21.9]),                            # a dipping plane
               None]
    c = inputs.Case(nex=nex, nez=nez, eh=eh, layers_thk=(7.3, 13.2, 9.5), vp=(5.9, 6.8, 8.0),
                    vs=(3.4, 3.9, 4.6), rho=(2.6, 2.9, 3.3), jitter=0.0, nst=2, nrec=2, interfaces=ifc,
                    aniso=[aniso.hexagonal_tilted(8.0, 4.6, 3.3, 0.05, 25.0, 10.0), None, None])
    d = os.fspath(tmp_path)
    D = c.write(d)
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    mat = o.material().reshape(o.nz, o.nx, 16)
    h = float(np.float32(np.float32(o.eh) / np.float32(o.hpn)))
    tops = [(L.xs, L.heights) for L in D["layers"]]            # bottom-up semmodel layers
    X = np.arange(o.nx) * h
    Zn = np.arange(o.nz) * h
    # the record of each layer, taken where the layer is thickest
    recs = []
    for j in range(len(tops)):
        lo = interface_height(*tops[j - 1], X) if j else np.full(o.nx, -1e9)
        hi = np.minimum(interface_height(*tops[j], X), Zn[-1])
        ix = int(np.argmax(hi - lo))
        iz = int(np.argmin(np.abs(Zn - 0.5 * (max(lo[ix], 0.0) + hi[ix]))))
        recs.append(mat[iz, ix])
    for j in range(len(recs)):
        for k in range(j):
            assert np.abs(recs[j] - recs[k]).max() > 1e-3     # the layers are distinguishable
    margin = 2e-3
    checked = 0
    for ix in range(o.nx):
        cuts = [interface_height(*t, X[ix]) for t in tops]
        for iz in range(o.nz):
            z = Zn[iz]
            if any(abs(z - cu) <= margin for cu in cuts):
                continue
            j = next((k for k, cu in enumerate(cuts) if z < cu), len(cuts) - 1)
            assert np.array_equal(mat[iz, ix], recs[j]), (ix, iz, z, cuts)
            checked += 1
    assert checked > 0.95 * o.nx * o.nz


# ------------------------------------------------------------ sponge
def sponge_oracle(tmp_path, atten, ass, neh=2):
    c = inputs.small_case(nst=3, neh=neh, extra=dict(atten=atten, ass=ass))
    d = os.fspath(tmp_path)
    c.write(d)
    return pyoracle.Oracle(["par=run.par"], "f64", cwd=d)


@pytest.mark.parametrize("ass", [0.93, 0.6])
def test_taper_closed_form(tmp_path, ass):
    o = sponge_oracle(tmp_path, 4, ass)
    nh = o.neh * o.hpn
    t = o.taper()
    assert len(t) == nh
    a32 = float(np.float32(ass))
    want = np.array([a32 ** (n / (nh * math.log(2.73))) for n in range(nh)])
    assert np.all(np.abs(t - want) <= np.spacing(np.float32(want)))
    assert t[0] == 1.0 and np.all(np.diff(t) < 0)       # innermost 1, damping grows toward the edge


@pytest.mark.parametrize("atten", [1, 2, 3, 4])
def test_sponge_on_ones_is_the_strip_factor(tmp_path, atten):
    ass = 0.7
    o = sponge_oracle(tmp_path, atten, ass)
    nh, nx, nz = o.neh * o.hpn, o.nx, o.nz
    a32 = float(np.float32(ass))
    fac = lambda dist: a32 ** ((nh - 1 - dist) / (nh * math.log(2.73)))   # dist = nodes from the outer edge
    got = o.sponge(np.ones((3, nz, nx)))
    want = np.ones((nz, nx))
    for iz in range(nz):
        for ix in range(nx):
            if iz < nh:                                   # bottom strip, corners included
                want[iz, ix] = fac(iz)
            elif atten > 1 and ix < nh:                   # left strip above it
                want[iz, ix] = fac(ix)
            elif atten > 2 and ix >= nx - nh:             # right strip above it
                want[iz, ix] = fac(nx - 1 - ix)
            elif atten > 3 and iz >= nz - nh:             # top strip between the side strips
                want[iz, ix] = fac(nz - 1 - iz)
    for k in range(3):
        assert np.allclose(got[k], want, rtol=2e-7, atol=0), (k, np.argwhere(~np.isclose(got[k], want)))
