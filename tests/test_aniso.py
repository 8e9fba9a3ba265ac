"""The anisotropic model builder (semsks.aniso) that writes the C2/C4
semmodel layers: Thomsen-style hexagonal constants, the Bond rotations of
the reference's builder, and the semmodel line.

* The rotations against the reference's own elastic_rotate and elastic_dip
  (semsyn/model/src/elastic_rotate.c:135-188, elastic_dip.c:158-213), run
  on a vertical-axis and on any_anisotropy.pl's horizontal-axis matrix at
  five angles (tests/golden/ref_elastic.npz, written to 2 decimals by those
  programs, built in oracle/_ref).
* Invariants a rotation keeps: the eigenvalues of the 6x6 matrix in Kelvin
  form (shear rows and columns scaled by sqrt 2).
* dip = 0 and azimuth = 0 leave hexagonal_vertical unchanged; eps, gamma,
  delta come back from the constants by Thomsen's definitions.
* The semmodel line holds the 13 constants in the order the reference's
  builder writes them (new_sem, any_anisotropy.pl:114-143), and the case
  writer puts them in the file in that order.
"""
import math
import os

import numpy as np
import pytest

from semsks import aniso, inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kelvin(C):
    w = np.array([1, 1, 1, math.sqrt(2), math.sqrt(2), math.sqrt(2)])
    return np.sort(np.linalg.eigvalsh(C * np.outer(w, w)))


@pytest.mark.parametrize("name", ["vti", "hti"])
def test_rotations_match_the_reference_programs(name):
    z = np.load(os.path.join(GOLD, "ref_elastic.npz"))
    C = z[name]
    for k, a in enumerate(z["angles"]):
        rot = aniso.bond_rotate_z(C, float(a))
        dip = aniso.bond_dip(C, float(a))
        assert np.abs(rot - z[name + "_elastic_rotate"][k]).max() < 0.006, (a, rot, z[name + "_elastic_rotate"][k])
        assert np.abs(dip - z[name + "_elastic_dip"][k]).max() < 0.006, (a, dip, z[name + "_elastic_dip"][k])


def test_rotation_keeps_kelvin_eigenvalues_and_identity_at_zero():
    C0 = aniso.hexagonal_vertical(8.1, 4.5, 3.35, 0.06, 0.04, 0.02)
    for dip, az in ((0.0, 0.0), (30.0, 60.0), (45.0, 200.0), (12.5, -33.0)):
        C = aniso.rotate(C0, aniso.rotation(dip, az))
        assert np.allclose(kelvin(C), kelvin(C0), rtol=1e-12)
        assert np.allclose(C, C.T)
    assert np.allclose(aniso.rotate(C0, aniso.rotation(0.0, 0.0)), C0, rtol=0, atol=1e-12)
    c, dip6 = aniso.hexagonal_tilted(8.1, 4.5, 3.35, 0.04, 0.0, 0.0)
    V = aniso.hexagonal_vertical(8.1, 4.5, 3.35, 0.04, 0.04)
    assert dip6 is None
    for key in aniso.SEMMODEL_ORDER:
        assert c[key] == pytest.approx(V[int(key[1]) - 1, int(key[2]) - 1], abs=1e-9)
    # the axis turns where rotation() says: R z = (sin d cos a, sin d sin a, cos d)
    R = aniso.rotation(30.0, 60.0)
    d, a = math.radians(30.0), math.radians(60.0)
    assert np.allclose(R @ [0, 0, 1], [math.sin(d) * math.cos(a), math.sin(d) * math.sin(a), math.cos(d)])


@pytest.mark.parametrize("eps,gamma,delta", [(0.04, 0.04, 0.0), (0.06, 0.02, 0.03), (0.0, 0.05, -0.02)])
def test_thomsen_parameters_come_back(eps, gamma, delta):
    vp, vs, rho = 8.1, 4.5, 3.35
    C = aniso.hexagonal_vertical(vp, vs, rho, eps, gamma, delta)
    e, g, dl = aniso.thomsen(C)
    assert e == pytest.approx(eps, abs=1e-12) and g == pytest.approx(gamma, abs=1e-12)
    assert dl == pytest.approx(delta, abs=1e-12)
    assert C[2, 2] == pytest.approx(rho * vp * vp) and C[3, 3] == pytest.approx(rho * vs * vs)
    assert C[0, 0] == pytest.approx(C[1, 1]) and C[0, 1] == pytest.approx(C[0, 0] - 2 * C[5, 5])


def test_semmodel_line_order_and_case_writer(tmp_path):
    C = np.array([[10 * (i + 1) + (j + 1) for j in range(6)] for i in range(6)], float)
    C = 0.5 * (C + C.T)
    line = aniso.semmodel_line(C, 3.3)
    # new_sem: ss[1..13] = ckl[0,0],[0,1],[0,2],[0,5],[1,1],[1,2],[1,5],[2,2],[2,5],[3,3],[3,4],[4,4],[5,5]
    idx = [(0, 0), (0, 1), (0, 2), (0, 5), (1, 1), (1, 2), (1, 5), (2, 2), (2, 5), (3, 3), (3, 4), (4, 4), (5, 5)]
    assert line == [C[i, j] for i, j in idx] + [3.3]
    layer = aniso.hexagonal_tilted(8.1, 4.5, 3.35, 0.04, 20.0, 40.0)
    c = inputs.Case(nex=4, nez=3, eh=2.0, layers_thk=(6.0,), vp=(8.1,), vs=(4.5,), rho=(3.35,), jitter=0.0,
                    nst=2, nrec=2, aniso=[layer])
    c.write(os.fspath(tmp_path))
    tokens = open(tmp_path / "semmodel").read().split()
    written = [float(t) for t in tokens[2:16]]          # after nll and np
    want = [layer[0][k] for k in aniso.SEMMODEL_ORDER] + [float(np.float32(3.35))]
    assert np.allclose(written, want, atol=6e-5)
    dip = [float(t) for t in open(tmp_path / "semdip").read().split()]
    assert np.allclose(dip, layer[1], atol=6e-7)
