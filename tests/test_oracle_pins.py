"""Pins the CPU oracle before the GPU path is compared against it.

Each oracle function is checked against something that does not restate it:
  * golden vectors written by the reference's own interpl.c / readmodel.c
    (tests/golden/make_ref_vectors.py, via oracle/_ref);
  * closed forms (GLL nodes and weights, Lagrange derivative identities);
  * an independent FP64 assembly of K from explicit strain-displacement
    matrices on tiny meshes;
  * invariants: symmetry and null space of K, the patch test, mass = integral
    of rho, discrete energy conservation of the explicit update;
  * hand-traceable rules of the reference (interpolation delays at rdt=1,
    the sequential guard, receiver indexing, the nez/eh float arithmetic);
  * physics with closed forms: a plane pulse through a Lysmer or a free
    edge, the amplitude the interface injection radiates, and a vertically
    incident pulse recorded at the free surface.
"""
import math
import os

import numpy as np
import pytest

import pyoracle
from semsks import inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- GLL tables
CLOSED = {
    3: ([-1, -math.sqrt(1 / 5), math.sqrt(1 / 5), 1], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    4: ([-1, -math.sqrt(3 / 7), 0, math.sqrt(3 / 7), 1], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
    5: ([-1, -math.sqrt(1 / 3 + 2 * math.sqrt(7) / 21), -math.sqrt(1 / 3 - 2 * math.sqrt(7) / 21),
         math.sqrt(1 / 3 - 2 * math.sqrt(7) / 21), math.sqrt(1 / 3 + 2 * math.sqrt(7) / 21), 1],
        [1 / 15, (14 - math.sqrt(7)) / 30, (14 + math.sqrt(7)) / 30, (14 + math.sqrt(7)) / 30,
         (14 - math.sqrt(7)) / 30, 1 / 15]),
}


@pytest.mark.parametrize("hpn", [3, 4, 5])
def test_gll_closed_forms(hpn):
    x, w, _ = pyoracle.gll(hpn, "f64")
    xc, wc = map(np.array, CLOSED[hpn])
    assert np.allclose(x, xc, rtol=0, atol=2e-15)
    assert np.allclose(w, wc, rtol=0, atol=2e-15)
    x32, w32, _ = pyoracle.gll(hpn, "f32")
    assert np.all(np.abs(x32 - xc.astype(np.float32)) <= np.spacing(np.float32(1)))
    assert np.all(np.abs(w32 - wc) <= 2 * np.spacing(np.float32(1)))


@pytest.mark.parametrize("hpn", [3, 4, 5, 6])
def test_gll_quadrature_exactness(hpn):
    x, w, _ = pyoracle.gll(hpn, "f64")
    assert abs(w.sum() - 2) < 1e-14
    for k in range(2 * hpn):            # exact for degree <= 2N-1
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-13


@pytest.mark.parametrize("hpn", [3, 4, 5, 6])
def test_derivative_matrix_identities(hpn):
    x, w, d = pyoracle.gll(hpn, "f64")      # d[j][i] = l'_i(x_j)
    # exact derivative of every polynomial of degree <= N
    for k in range(hpn + 1):
        assert np.allclose(d @ x ** k, k * x ** max(k - 1, 0) if k else 0 * x, atol=1e-12)
    assert np.allclose(d.sum(axis=1), 0, atol=1e-13)
    n = hpn
    assert abs(d[0, 0] + n * (n + 1) / 4) < 1e-12 and abs(d[n, n] - n * (n + 1) / 4) < 1e-12
    # independent: derivative of the Lagrange basis via numpy polynomials
    for i in range(n + 1):
        others = [x[m] for m in range(n + 1) if m != i]
        P = np.poly1d(np.poly(others)) / np.prod([x[i] - o for o in others])
        assert np.allclose(P.deriv()(x), d[:, i], atol=1e-12)


# ---------------------------------------------------------------- golden: interpl
def test_interpolation_matches_reference_interpl():
    g = np.load(os.path.join(GOLD, "ref_interpl.npz"))
    psv, sh = g["psv"], g["sh"]
    for name in ("r1_single", "r07_single", "r1p_single", "r1_shared", "r07_shared", "r037_shared"):
        rdt = float(g[name + "_rdt"])
        shared = name.endswith("shared")
        o0, o1 = pyoracle.interp_series(psv, sh if shared else None, 40, rdt, "f32")
        assert np.array_equal(o0.view(np.uint32), g[name + "_g"].view(np.uint32)), name
        if shared:
            assert np.array_equal(o1.view(np.uint32), g[name + "_gv"].view(np.uint32)), name


@pytest.mark.parametrize("solver", [0, 1])
def test_boundary_series_at_mesh_widths_matches_reference(tmp_path, solver):
    """The oracle's boundary series, read from the files through its own
    par/greenfile reader, against the reference interpl's vectors at the
    record widths of a 2x2-element mesh, dp = sdt/rdt, with the guard
    applied per level (tests/golden/ref_interpl_mesh.npz)."""
    from cases import interp_mesh_case
    z = np.load(os.path.join(GOLD, "ref_interpl_mesh.npz"))
    for tag in ("plain", "spiky"):
        for rdt in ("1", "0p999", "0p7", "0p37"):
            key = "%s_r%s" % (tag, rdt)
            d = interp_mesh_case(tmp_path / key, tag, key, solver)
            o = pyoracle.Oracle(["par=run.par"], "f32", cwd=d)
            n = int(z[key + "_ntsem"])
            assert o.nt_sem == n
            g, gv = o.boundary_series(n)
            want = z[key + ("_shared_g" if solver == 0 else "_single_g")]
            want = np.stack([pyoracle.guard(r, "f32") for r in want])
            assert np.array_equal(g.view(np.uint32), want.view(np.uint32)), key
            if solver == 0:
                wv = np.stack([pyoracle.guard(r, "f32") for r in z[key + "_shared_gv"]])
                assert np.array_equal(gv.view(np.uint32), wv.view(np.uint32)), key
            else:
                assert not gv.any()


def test_interpolation_delays_at_rdt1():
    """With dp = sdt, the single-stream solver injects record it-2 and the
    shared-ring solver record it-1 (interpl.c:59-91 with sem2d.c:410-413 /
    sem2d_psvsh.c:430-431)."""
    rng = np.random.default_rng(3)
    r = rng.standard_normal((30, 7)).astype(np.float32)
    s = rng.standard_normal((30, 5)).astype(np.float32)
    one, _ = pyoracle.interp_series(r, None, 20, 1.0, "f32")
    assert np.array_equal(one[2:], r[:18]) and not one[:2].any()
    a, b = pyoracle.interp_series(r, s, 20, 1.0, "f32")
    assert np.array_equal(a[1:], r[:19]) and np.array_equal(b[1:], s[:19]) and not a[0].any()


def test_guard_sequential_rule():
    """Non-finite or |g|>=10 values become the mean of the already-repaired
    left neighbour and the original right neighbour (sem2d_psvsh.c:433-443)."""
    g = pyoracle.guard(np.array([1, 20, 30, 5, np.nan, 7, 9.99, -10, 2], np.float32))
    assert g[1] == np.float32((1 + 30) / 2)          # 15.5, not re-checked
    assert g[2] == np.float32((15.5 + 5) / 2)        # uses the repaired g[1]
    assert g[4] == np.float32((5 + 7) / 2)
    assert g[6] == np.float32(9.99)
    assert g[7] == np.float32((9.99 + 2) / 2)


# ---------------------------------------------------------------- golden: read_model
def test_raymodel_reader_matches_reference(tmp_path):
    g = np.load(os.path.join(GOLD, "ref_readmodel.npz"))
    for name in ("flat3", "sph3", "rays2"):
        p = tmp_path / name
        p.write_text(str(g[name + "_text"]))
        r = pyoracle.read_raymodel(os.fspath(p), int(g[name + "_flat"]))
        assert r["jo"] == int(g[name + "_jo"]) and r["nrec"] == int(g[name + "_nrec"]), name
        for k in ("cc", "ss", "dd", "tth"):
            assert np.array_equal(r[k].view(np.uint32), g[name + "_" + k].view(np.uint32)), (name, k)


# ---------------------------------------------------------------- mesh + receivers
def test_mesh_dims_float_arithmetic(tmp_path):
    """nez=(int)((thk-1e-3)/eh) with a float thickness: 200.001 gives 100
    elements of 2.0 km, 24.001 gives 11 (its float value is below 24.001)."""
    assert inputs.mesh_dims(200.001, 2.0, 1, 4)[:2] == (100, 2.0)
    assert inputs.mesh_dims(24.001, 2.0, 1, 4)[0] == 11
    for thk, nez in ((24.001, 11), (24.002, 12)):
        c = inputs.small_case(layers_thk=(4.0, 4.0, thk - 8.0), nst=3)
        d = tmp_path / ("m%d" % nez)
        c.write(os.fspath(d))
        o = pyoracle.Oracle(["par=run.par"], "f32", cwd=os.fspath(d))
        assert o.nez == inputs.mesh_dims(inputs.float_sum([4.0, 4.0, thk - 8.0]), 2.0, 1, 4)[0]


def test_receiver_indexing(tmp_path):
    c = inputs.small_case(nst=3)
    d = os.fspath(tmp_path)
    D = c.write(d)
    with open(os.path.join(d, "rec_file"), "w") as f:
        f.write("4\n12.5000\n12.4999\n0.0000\n47.9999\n")
    o = pyoracle.Oracle(["par=run.par"], "f32", cwd=d)
    ix, iz = o.receivers()
    h = np.float32(D["h"])
    expect = [int(np.float32(x) / h) for x in (12.5, 12.4999, 0.0, 47.9999)]
    assert list(ix) == expect and all(iz == o.nz - 1)
    o2 = pyoracle.Oracle(["par=run.par", "get_rec=0", "xbeg=10", "idxplot=3", "xtrace=5"], "f32", cwd=d)
    ix2, _ = o2.receivers()
    assert list(ix2) == [int(np.float32(10) / h - 1) + 3 * k for k in range(5)]


# ---------------------------------------------------------------- element force
def brute_force_K(o):
    """Independent global K in FP64: explicit strain-displacement matrices
    at the collocated GLL quadrature points, K_e = sum_q w_q Je B^T C B."""
    x, w, _ = pyoracle.gll(o.hpn, "f64")
    n = o.hpn
    D = np.zeros((n + 1, n + 1))
    for i in range(n + 1):
        others = [x[m] for m in range(n + 1) if m != i]
        P = np.poly1d(np.poly(others)) / np.prod([x[i] - q for q in others])
        D[:, i] = P.deriv()(x)
    eh = o.eh
    je, jinv = eh * eh / 4.0, 2.0 / eh
    mat = o.material()
    N = o.nx * o.nz
    K = np.zeros((3 * N, 3 * N))
    for ex in range(o.nex):
        for ez in range(o.nez):
            nodes = [(ez * n + b) * o.nx + ex * n + a for b in range(n + 1) for a in range(n + 1)]
            idx = np.array([[3 * q + c for c in range(3)] for q in nodes]).ravel()
            Ke = np.zeros((len(idx), len(idx)))
            for b in range(n + 1):
                for a in range(n + 1):
                    B = np.zeros((5, len(idx)))     # rows: e_xx, e_zz, e_yz, e_xz(eng), e_xy(eng)
                    for t in range(n + 1):
                        kx = (b * (n + 1) + t) * 3          # node (t, b): d/dx
                        kz = (t * (n + 1) + a) * 3          # node (a, t): d/dz
                        B[0, kx + 0] += D[a, t] * jinv
                        B[4, kx + 1] += D[a, t] * jinv
                        B[3, kx + 2] += D[a, t] * jinv
                        B[1, kz + 2] += D[b, t] * jinv
                        B[2, kz + 1] += D[b, t] * jinv
                        B[3, kz + 0] += D[b, t] * jinv
                    r = mat[nodes[b * (n + 1) + a]]
                    C = np.zeros((5, 5))
                    vals = r[1:]
                    k = 0
                    for i in range(5):
                        for j in range(i, 5):
                            C[i, j] = C[j, i] = vals[k]
                            k += 1
                    Ke += w[a] * w[b] * je * B.T @ C @ B
            K[np.ix_(idx, idx)] += Ke
    return K


def tiny_oracle(tmp_path, aniso_layer=None, nex=3, nez=3, homogeneous=False):
    from semsks import aniso
    an = None
    if aniso_layer:
        an = [aniso.hexagonal_tilted(7.8, 4.4, 3.3, 0.05, 35.0, 70.0), None, None]
    kw = dict(interfaces=[([0.0, 8.0], [1.3, 3.9]), None, None], aniso=an)
    if homogeneous:
        kw = dict(vp=(7.0,) * 3, vs=(4.0,) * 3, rho=(3.0,) * 3, jitter=0.0)
    c = inputs.Case(nex=nex, nez=nez, eh=2.0, layers_thk=(1.5, 1.7, 4.0), nst=3, nrec=2, **kw)
    d = os.fspath(tmp_path)
    c.write(d)
    return pyoracle.Oracle(["par=run.par"], "f64", cwd=d)


@pytest.mark.parametrize("tilted", [False, True])
def test_stiffness_matches_independent_assembly(tmp_path, tilted):
    o = tiny_oracle(tmp_path, tilted)
    K = brute_force_K(o)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((3, o.nz, o.nx))
    f = o.stiff(u)
    fa = K @ np.moveaxis(u, 0, -1).ravel()
    assert np.allclose(np.moveaxis(f, 0, -1).ravel(), fa, rtol=0, atol=1e-11 * np.abs(fa).max())
    # K symmetric and positive semi-definite
    assert np.allclose(K, K.T, atol=1e-12 * np.abs(K).max())
    assert np.linalg.eigvalsh(K).min() > -1e-9 * np.abs(K).max()


def test_stiffness_null_space_and_patch(tmp_path):
    o = tiny_oracle(tmp_path, True, nex=4, nez=4)
    X = np.arange(o.nx)[None, :] * 1.0
    Z = np.arange(o.nz)[:, None] * 1.0
    from semsks.inputs import node_coords
    xs = node_coords(o.nex, o.eh, o.hpn)[None, :] + 0 * Z
    zs = node_coords(o.nez, o.eh, o.hpn)[:, None] + 0 * X
    one = np.ones((o.nz, o.nx))
    for u in (np.stack([one, 0 * one, 0 * one]), np.stack([0 * one, one, 0 * one]),
              np.stack([0 * one, 0 * one, one]), np.stack([zs, 0 * one, -xs])):   # translations, rotation about y
        f = o.stiff(u)
        assert np.abs(f).max() < 1e-10
    # patch test: a linear field has constant strain, so in a homogeneous
    # model every node off the domain edges carries no force, and the edge
    # forces are the tractions of the constant stress
    o2 = tiny_oracle(tmp_path / "h", False, nex=4, nez=4, homogeneous=True)
    u = np.stack([0.3 * xs + 0.1 * zs, -0.2 * xs + 0.05 * zs, 0.1 * xs - 0.4 * zs])
    f = o2.stiff(u)
    assert np.abs(f[:, 1:-1, 1:-1]).max() < 1e-10 * np.abs(f).max()
    assert np.abs(f[:, 0, :]).max() > 1e-3 * np.abs(f).max()
    # total edge force balances (zero net force of a self-equilibrated state)
    assert np.abs(f.sum(axis=(1, 2))).max() < 1e-10 * np.abs(f).max()


# ---------------------------------------------------------------- mass, update
def test_mass_integrates_density(tmp_path):
    c = inputs.small_case(vp=(6.0, 6.0, 6.0), vs=(3.5, 3.5, 3.5), rho=(2.7, 2.7, 2.7), jitter=0.0, nst=3)
    d = os.fspath(tmp_path)
    c.write(d)
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    m, md, ab = o.mass()
    area = (o.nex * o.eh) * (o.nez * o.eh)
    assert abs(m.sum() - float(np.float32(2.7)) * area) < 1e-12 * m.sum()
    assert np.allclose(md, np.float32(o.sdt) ** 2 / m.astype(np.float64), rtol=1e-15)
    # Lysmer coefficients live on the left, bottom and right edges only
    inner = ab.copy()
    inner[:, 0] = inner[0, :] = inner[:, -1] = 0
    assert not inner.any() and ab[0, :].all() and ab[1:, 0].all() and ab[1:, -1].all()
    # bottom normal (z) uses vp, tangential vs: ratio vp/vs at an edge node
    k = ab[0, 5]
    assert abs(k[2] / k[0] - 6.0 / 3.5) < 1e-6


def test_energy_conserved_without_absorption(tmp_path):
    """Central differences conserve E = 1/2 v'Mv + 1/2 u(n+1)'K u(n)."""
    c = inputs.small_case(nst=3, absorb=0)
    d = os.fspath(tmp_path)
    c.write(d)
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    m, _, _ = o.mass()
    zs, xs = np.meshgrid(np.linspace(0, 1, o.nz), np.linspace(0, 1, o.nx), indexing="ij")
    bump = np.exp(-((xs - 0.5) ** 2 + (zs - 0.5) ** 2) / 0.01)
    u0 = np.stack([bump, 0.3 * bump, -0.5 * bump])
    um = u0.copy()
    E = []
    for _ in range(150):
        G = o.stiff(u0)
        u1 = o.update(u0, um, G)
        v = (u1 - u0) / o.sdt
        E.append(0.5 * np.sum(m * v * v) + 0.5 * np.sum(u1 * o.stiff(u0)))
        um, u0 = u0, u1
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 1e-10 * abs(E[0])


def test_colour_parallel_fp64_path_equals_reference_order(tmp_path):
    """orc_run_fast (the FP64 oracle of the large fixtures: the element force
    swept in four colours in parallel) against orc_run (elements in the
    reference's ex-outer, ez-inner order): the same arithmetic, so the final
    fields agree to rounding of the shared-node sums (1e-12 of the peak) on
    a case with a dipping interface, a tilted axis, the coupled solver and
    absorbing edges."""
    from semsks import aniso
    c = inputs.small_case(nst=250, solver=0, psi_deg=30.0, sh_zero=False,
                          interfaces=[([0.0, 50.0], [5.3, 9.1]), None, None],
                          aniso=[aniso.hexagonal_tilted(6.5, 3.75, 2.9, 0.05, 30.0, 20.0), None, None])
    d = os.fspath(tmp_path)
    c.write(d)
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    _, fa = o.run(250, want_field=True)
    _, fb = o.run(250, want_field=True, fast=True)
    assert np.abs(fa).max() > 0.1
    assert np.abs(fa - fb).max() <= 1e-12 * np.abs(fa).max()


def test_faithful_fp32_tracks_fp64(tmp_path):
    c = inputs.small_case(nst=300, psi_deg=30.0, sh_zero=False)
    d = os.fspath(tmp_path)
    c.write(d)
    t32 = pyoracle.Oracle(["par=run.par"], "f32", cwd=d).run()
    t64 = pyoracle.Oracle(["par=run.par"], "f64", cwd=d).run()
    ref = max(np.linalg.norm(t64[:, k]) for k in range(3))
    for k in range(3):
        scale = max(np.linalg.norm(t64[:, k]), 1e-3 * ref)
        assert np.linalg.norm(t32[:, k] - t64[:, k]) <= 1e-5 * scale


# ------------------------------------------------------- absorbing edges
@pytest.mark.parametrize("pol", ["P", "S"])
def test_lysmer_edge_absorbs_normal_incidence(tmp_path, pol):
    """A plane pulse running down onto the bottom edge: with the Lysmer
    dashpots (Absorb_Matrix, Absorbs) it leaves the box, without them
    (absorb=0, free edge) it comes back whole.  A dashpot off by a factor
    2, or the P and S impedances swapped, reflects 1/3 or ~27%."""
    vp, vs, rho = 6.0, 3.5, 2.7
    found = {}
    for absorb in (1, 0):
        # 160 km wide, so that what the side edges do to a plane pulse does
        # not reach the middle columns in the time measured
        c = inputs.Case(nex=80, eh=2.0, layers_thk=(10.0, 10.0, 20.002), vp=(vp,) * 3, vs=(vs,) * 3,
                        rho=(rho,) * 3, jitter=0.0, nst=3, nrec=2, absorb=absorb)
        d = os.fspath(tmp_path / ("a%d" % absorb))
        c.write(d)
        o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
        zs = inputs.node_coords(o.nez, o.eh, o.hpn)[:, None] + np.zeros((1, o.nx))
        v = vp if pol == "P" else vs
        zc, width = 20.0, 4.0
        shape = lambda z: np.exp(-((z - zc) / width) ** 2)
        k = 2 if pol == "P" else 0                 # normal (w) or tangential (u) motion
        u0 = np.zeros((3, o.nz, o.nx)); um = np.zeros_like(u0)
        u0[k] = shape(zs)                          # u(z, t) = shape(z + v t): running down
        um[k] = shape(zs - v * o.sdt)
        nsteps = int(round((zc + 12.0) / v / o.sdt))   # pulse centre to the edge and 12 km back
        for _ in range(nsteps):
            u1 = o.update(u0, um, o.stiff(u0))
            um, u0 = u0, u1
        mid = slice(2 * o.nx // 5, 3 * o.nx // 5)
        found[absorb] = np.abs(u0[k][:, mid]).max()
    assert found[0] > 0.8          # free edge: full reflection
    assert found[1] < 0.05         # Lysmer edge: the pulse has left


# ----------------------------------------------------- boundary injection
@pytest.mark.parametrize("solver,psi", [(0, 0.0), (0, 90.0), (1, 0.0), (2, 0.0)])
def test_injected_vertical_plane_wave_amplitude(tmp_path, solver, psi):
    """A vertically incident S pulse g(t) (peak 1) injected on the
    interface row, looked at in the interior before it reaches the surface.

    GRT_Interface3 (solver 2) replaces the interface value by the scattered
    increment plus g, so the interior carries g itself (scale 1).

    GRT_Interface4 (solvers 0, 1) adds g(t) - g(t-dt) to the interface nodes
    every step.  In the central-difference update a displacement kick Dg on a
    node adds Dg/dt to its velocity, i.e. the impulse (mass per area) Dg/dt;
    the interface row's lumped mass per unit area is rho*w_0*eh (GLL end
    weight w_0 from the elements above and below, Jacobian eh/2 each), and a
    plane impulse in a medium of impedance rho*beta radiates half of it each
    way, so the up-going wave is g * w_0*eh / (2*beta*dt)."""
    beta = 3.5
    c = inputs.Case(nex=80, eh=2.0, layers_thk=(20.0, 20.0, 20.002), vp=(6.0,) * 3, vs=(beta,) * 3,
                    rho=(2.7,) * 3, jitter=0.0, p=0.0, f0=0.2, psi_deg=psi, sh_zero=(psi == 0.0),
                    solver=solver, nrec=3, nst=3)
    D = c.derived()
    lead = 1.2 / c.f0
    c.nst = int(round((lead + 30.0 / beta) / D["sdt"])) + 1      # pulse centre at z = 30 km
    d = os.fspath(tmp_path)
    c.write(d)
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    _, f = o.run(want_field=True)
    t = (c.nst - 1) * o.sdt
    zs = inputs.node_coords(o.nez, o.eh, o.hpn)
    comp = 0 if psi == 0.0 else 1
    col = f[comp][:, o.nx // 2]
    inc = inputs.ricker(t - lead - zs / beta, c.f0)
    sel = (zs > 8.0) & (zs < 52.0)
    scale = float(np.dot(col[sel], inc[sel]) / np.dot(inc[sel], inc[sel]))
    x, w, _ = pyoracle.gll(o.hpn, "f64")
    expect = 1.0 if solver == 2 else w[0] * o.eh / (2.0 * beta * o.sdt)
    assert abs(scale / expect - 1.0) < 0.01, (scale, expect)
    assert np.abs(col[sel] - scale * inc[sel]).max() < 0.03 * abs(scale)
    assert np.abs(f[1 - comp if comp < 2 else 0][:, o.nx // 2]).max() < 1e-9 * abs(scale)


# ------------------------------------------------------- receiver records
@pytest.mark.parametrize("kind,solver", [("SV", 0), ("SH", 0), ("P", 0), ("SV", 2)])
def test_surface_record_of_a_vertical_pulse(tmp_path, kind, solver):
    """kirrecord_zhl (sem2d_psvsh.c:2116-2165) writes u, -w and v; the
    injection (GRT_Interface4, :2076-2105) adds minus the greenfile's w to
    the solver's w.  So a vertically incident pulse put in one greenfile
    channel (u for SV, w for P, the shgreenfile for SH) must come out in the
    same seismogram channel with the greenfile's own sign, and at a free
    surface with twice the amplitude the interface radiates (the impulse
    w_0*eh/(2*v*dt) of GRT_Interface4, 1 for GRT_Interface3), arriving after
    the travel time to the surface; the other two channels stay silent.  A
    record with the sign of w not flipped, or two channels swapped, fails."""
    vp, vs, rho, f0 = 6.0, 3.5, 2.7, 0.25
    v = vp if kind == "P" else vs
    c = inputs.Case(nex=80, eh=2.0, layers_thk=(6.0, 6.0, 6.002), vp=(vp,) * 3, vs=(vs,) * 3, rho=(rho,) * 3,
                    jitter=0.0, p=0.0, f0=f0, psi_deg=90.0 if kind == "SH" else 0.0, sh_zero=(kind != "SH"),
                    solver=solver, nrec=5, rec_x0=64.0, rec_x1=96.0, nst=3)
    D = c.derived()
    lead = 1.2 / f0
    height = D["nez"] * D["eh"]
    c.nst = int(round((lead + height / v + 1.2 / f0 + 1.0) / D["sdt"])) + 1
    d = os.fspath(tmp_path)
    c.write(d)
    if kind == "P":     # u = 0, greenfile w = r(t - lead - z/vp) on the bottom row and the left column
        nx, nz, nbl = D["nx"], D["nz"], c.neh * c.hpn
        zs = inputs.node_coords(D["nez"], D["eh"], c.hpn)
        sdt = float(np.float32(D["sdt"]))
        psv = np.zeros((c.nst + 3, 2 * (nx + nz)), np.float32)
        for it in range(c.nst + 3):
            psv[it, nx:2 * nx] = inputs.ricker(it * sdt - lead - zs[nbl] / vp, f0)
            psv[it, 2 * nx + nz:] = inputs.ricker(it * sdt - lead - zs / vp, f0)
        psv.tofile(os.path.join(d, "greenfile"))
    o = pyoracle.Oracle(["par=run.par"], "f64", cwd=d)
    tr = o.run()
    _, w, _ = pyoracle.gll(o.hpn, "f64")
    scale = 1.0 if solver == 2 else w[0] * o.eh / (2.0 * v * o.sdt)
    t = np.arange(tr.shape[0]) * o.sdt
    top = inputs.node_coords(o.nez, o.eh, o.hpn)[-1]
    expect = 2.0 * scale * inputs.ricker(t - lead - top / v, f0)
    k = {"SV": 0, "P": 1, "SH": 2}[kind]
    peak = np.abs(expect).max()
    for r in range(tr.shape[2]):
        fit = float(tr[:, k, r] @ expect / (expect @ expect))
        assert abs(fit - 1.0) < 0.005, (r, fit)
        assert np.abs(tr[:, k, r] - expect).max() < 0.04 * peak
        for j in range(3):
            if j != k:
                assert np.abs(tr[:, j, r]).max() < 0.01 * peak, (r, j)
