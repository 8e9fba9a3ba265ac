"""Anisotropic layer constants for the semmodel format.

Hexagonal (transversely isotropic) elasticity from Thomsen-style parameters,
rotated so that the symmetry axis has a given dip from vertical and azimuth
from the profile (x) direction.  The reference builds such models with a Bond
rotation of the 6x6 matrix about x3 (semsyn/model/src/elastic_rotate.c:135-188)
and within x1-x3 (semsyn/model/src/elastic_dip.c:158-213); here the full
fourth-order tensor is rotated directly, which is the same transformation.

Voigt order: 0=xx 1=yy 2=zz 3=yz 4=xz 5=xy (x east along the profile,
y north, z up; sem2d_psvsh.c:12-16).
"""
from __future__ import annotations

import math

import numpy as np

VOIGT = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (1, 2): 3, (2, 1): 3, (0, 2): 4, (2, 0): 4, (0, 1): 5, (1, 0): 5}


def hexagonal_vertical(vp: float, vs: float, rho: float, eps: float, gamma: float, delta: float = 0.0) -> np.ndarray:
    """6x6 Voigt matrix of a medium with a vertical symmetry axis (x3)."""
    c33 = rho * vp * vp
    c44 = rho * vs * vs
    c11 = c33 * (1.0 + 2.0 * eps)
    c66 = c44 * (1.0 + 2.0 * gamma)
    c13 = math.sqrt((c33 - c44) * (2.0 * delta * c33 + c33 - c44)) - c44
    c12 = c11 - 2.0 * c66
    C = np.zeros((6, 6))
    C[0, 0] = C[1, 1] = c11
    C[2, 2] = c33
    C[0, 1] = C[1, 0] = c12
    C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = c13
    C[3, 3] = C[4, 4] = c44
    C[5, 5] = c66
    return C


def to_tensor(C: np.ndarray) -> np.ndarray:
    T = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    T[i, j, k, l] = C[VOIGT[(i, j)], VOIGT[(k, l)]]
    return T


def to_voigt(T: np.ndarray) -> np.ndarray:
    C = np.zeros((6, 6))
    for (i, j), m in VOIGT.items():
        for (k, l), n in VOIGT.items():
            C[m, n] = T[i, j, k, l]
    return C


def rotation(dip_deg: float, azimuth_deg: float) -> np.ndarray:
    """R with R @ z_hat = (sin d cos a, sin d sin a, cos d)."""
    d, a = math.radians(dip_deg), math.radians(azimuth_deg)
    Ry = np.array([[math.cos(d), 0, math.sin(d)], [0, 1, 0], [-math.sin(d), 0, math.cos(d)]])
    Rz = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    return Rz @ Ry


def rotate(C: np.ndarray, R: np.ndarray) -> np.ndarray:
    T = to_tensor(C)
    Tr = np.einsum("ip,jq,kr,ls,pqrs->ijkl", R, R, R, R, T)
    return to_voigt(Tr)


def layer_entry(C: np.ndarray):
    """(13 monoclinic constants dict, 6 tilted-axis constants or None) for
    inputs.Case(aniso=[...]).  1-based Voigt names as the semmodel uses."""
    c = lambda i, j: float(C[i - 1, j - 1])
    d = dict(c11=c(1, 1), c12=c(1, 2), c13=c(1, 3), c16=c(1, 6), c22=c(2, 2), c23=c(2, 3), c26=c(2, 6),
             c33=c(3, 3), c36=c(3, 6), c44=c(4, 4), c45=c(4, 5), c55=c(5, 5), c66=c(6, 6))
    dip6 = (c(1, 4), c(1, 5), c(3, 4), c(3, 5), c(4, 6), c(5, 6))
    if max(abs(v) for v in dip6) < 1e-12:
        dip6 = None
    return d, dip6


def bond_rotate_z(C: np.ndarray, angle_deg: float) -> np.ndarray:
    """The matrix in a frame whose x1 axis is turned angle_deg counter-
    clockwise about x3 from the old one: elastic_rotate's transformation
    (semsyn/model/src/elastic_rotate.c:135-188, new-from-old direction
    cosines cos(angle), cos(angle - 90), ...).  Turning the frame by +a
    turns the medium by -a in it."""
    return rotate(C, rotation(0.0, -angle_deg))


def bond_dip(C: np.ndarray, angle_deg: float) -> np.ndarray:
    """elastic_dip's transformation within the x1-x3 plane
    (semsyn/model/src/elastic_dip.c:158-213): a vertical symmetry axis
    tilts by angle_deg toward +x1."""
    return rotate(C, rotation(angle_deg, 0.0))


def thomsen(C: np.ndarray):
    """(eps, gamma, delta) of a matrix with a vertical symmetry axis
    (Thomsen 1986): eps = (c11 - c33) / (2 c33), gamma = (c66 - c44) / (2 c44),
    delta = ((c13 + c44)^2 - (c33 - c44)^2) / (2 c33 (c33 - c44))."""
    c11, c33, c13, c44, c66 = C[0, 0], C[2, 2], C[0, 2], C[3, 3], C[5, 5]
    return ((c11 - c33) / (2 * c33), (c66 - c44) / (2 * c44),
            ((c13 + c44) ** 2 - (c33 - c44) ** 2) / (2 * c33 * (c33 - c44)))


# the semmodel layer line: 13 constants then rho, in the order the
# reference's model builder writes them (new_sem, any_anisotropy.pl:114-143)
# and discret reads them (sem2d_psvsh.c:700-731)
SEMMODEL_ORDER = ("c11", "c12", "c13", "c16", "c22", "c23", "c26", "c33", "c36", "c44", "c45", "c55", "c66")


def semmodel_line(C: np.ndarray, rho: float) -> list:
    """[c11 c12 c13 c16 c22 c23 c26 c33 c36 c44 c45 c55 c66 rho] of a 6x6
    matrix (1-based Voigt names; 4 = yz, 5 = xz, 6 = xy)."""
    return [float(C[int(k[1]) - 1, int(k[2]) - 1]) for k in SEMMODEL_ORDER] + [float(rho)]


def hexagonal_tilted(vp: float, vs: float, rho: float, strength: float, dip_deg: float, azimuth_deg: float,
                     delta: float = 0.0):
    """Thomsen eps = gamma = strength (the 'x% anisotropy' of the configs)."""
    C0 = hexagonal_vertical(vp, vs, rho, strength, strength, delta)
    return layer_entry(rotate(C0, rotation(dip_deg, azimuth_deg)))
