"""Magnetosphere-like block: stretched spherical-polar grid, dipole background,
uniform solar wind.

Geometry: r from r_in to r_out with geometric stretching (N_r cells), the polar
angle theta measured from the +x axis (the Sun-Earth line, like the
reference's LFM grids, whose singular axis is x) over [theta0, pi - theta0],
and the azimuth phi around x over a full turn (periodic).  Corners, ghost
corners included, are x = r cos(theta), y = r sin(theta) cos(phi),
z = r sin(theta) sin(phi); the metric is built from them on the GPU like any
other grid.

Code units: length R_E, speed 100 km/s, number density 1 cm^-3 (mass density
m_p cm^-3), so time R_E / (100 km/s) = 63.71 s, pressure m_p n v^2 =
1.6726e-11 Pa and field sqrt(mu0 m_p n) v = 4.5846 nT.

The initial wind is brought to rest inside r ~ 12 R_E (a tanh taper), so the
run starts with the dipole unperturbed near the inner boundary.

Boundaries: SphIn / SphOut on the radial faces (bcs.F90:934-982); on the
theta faces either the LFM axis BCs with ring averaging (the grid reaches the
axis up to a hole of half-angle theta0 = 1e-3) or, for a wedge, zero gradient;
periodic in phi.  The background field B0 is the dipole (gam_background_dipole); the evolved field
starts as the uniform IMF, whose face fluxes are the discrete curl of the
exact edge integrals of A = (B x r)/2 (div.B = 0 to rounding).
"""
from __future__ import annotations

import math

import numpy as np

MP = 1.67262192e-27          # proton mass, kg
MU0 = 4e-7 * math.pi
KB = 1.380649e-23
RE = 6.3781e6                # m
V0 = 1.0e5                   # m/s
N0 = 1.0e6                   # m^-3
P_UNIT = MP * N0 * V0 ** 2                   # Pa
B_UNIT = math.sqrt(MU0 * MP * N0) * V0 * 1e9  # nT
T_UNIT = RE / V0                             # s


def wind_state(n_cc=5.0, v_kms=(-400.0, 0.0, 0.0), T_K=1.0e5, b_nT=(0.0, 0.0, -5.0), gamma=5.0 / 3.0):
    """(rho, v, p, B) of the solar wind in code units."""
    rho = n_cc
    v = tuple(x / 100.0 for x in v_kms)
    p = n_cc * N0 * KB * T_K / P_UNIT
    b = tuple(x / B_UNIT for x in b_nT)
    return rho, v, p, b


KBLTZ_CGS = 1.380649e-16     # erg/K
ERGCC2NPA = 1.0e8


def wind_series_from_file(path: str, gT0: float = T_UNIT, gv0: float = V0, gP0_nPa: float = P_UNIT * 1e9,
                          gB0_nT: float = B_UNIT):
    """The bcwind file at path as a configs.WindSeries in code units, converted
    as readWind does (apps/wind.F90:667-721): t / gT0, density as is, V / gv0,
    pressure D kB Temp (erg/cc -> nPa) / gP0, B / gB0, Bx0 / gB0, ByC and BzC
    unscaled, xyzBS as is (planet radii).  The defaults are this module's code
    units."""
    from kaiju_amd import configs, gamio
    w = gamio.read_wind(path)
    D = w["D"]
    return configs.WindSeries(t=(1 / gT0) * w["T"], rho=(1 / 1.0) * D, p=(1 / gP0_nPa) * D * KBLTZ_CGS * w["Temp"] * ERGCC2NPA,
                              v=(1 / gv0) * w["V"], b=(1 / gB0_nT) * w["B"], xyzBS=w["xyzBS"],
                              ByC=w["ByC"], BzC=w["BzC"], Bx0=w["Bx0"] * (1 / gB0_nT))


def dipole_moment(m_nT_RE3=-31000.0) -> float:
    """Dipole moment along z in code units (B = M (3 z r_hat - z_hat)/r^3 style)."""
    return m_nT_RE3 / B_UNIT


AXIS_THETA0 = 1.0e-3   # polar hole of the axis grid (see corners)


def corners(N, nG=4, r_in=2.0, r_out=30.0, theta0=0.15, axis=False):
    """Corner array (F,3) of the stretched spherical-polar grid, ghosts included.

    axis = False: a wedge, theta linear in j over [theta0, pi - theta0] and on
    into the ghosts.  axis = True: the LFM-grid layout, the ghost rows past
    each pole mirrored through the axis (theta(-j) = -theta(j), theta(Nt + j) =
    2 pi - theta(Nt - j)), so a ghost cell lies on its conjugate cell half a
    turn away in phi.  theta0 > 0 keeps a hole of that half-angle around the
    axis: on faces lying exactly on it the reference's metric divides zero by
    zero (face centre = integral / area)."""
    Nr, Nt, Np = N
    ir = np.arange(-nG, Nr + nG + 1)
    r = r_in * (r_out / r_in) ** (ir / Nr)
    it = np.arange(-nG, Nt + nG + 1)
    th = theta0 + (math.pi - 2 * theta0) * it / Nt
    if axis:
        lo, hi = it < 0, it > Nt
        th[lo] = -(theta0 + (math.pi - 2 * theta0) * (-it[lo]) / Nt)
        th[hi] = 2 * math.pi - (theta0 + (math.pi - 2 * theta0) * (2 * Nt - it[hi]) / Nt)
    ip = np.arange(-nG, Np + nG + 1)
    ph = 2 * math.pi * ip / Np
    R = r[:, None, None]
    T = th[None, :, None]
    P = ph[None, None, :]
    xyz = np.empty((len(r), len(th), len(ph), 3), order="F")
    xyz[..., 0] = R * np.cos(T)
    xyz[..., 1] = R * np.sin(T) * np.cos(P)
    xyz[..., 2] = R * np.sin(T) * np.sin(P)
    return xyz


def uniform_state(xyz: np.ndarray, rho, v, p, b, gamma=5.0 / 3.0, r_taper=12.0, w_taper=2.0):
    """Gas (G,5): uniform rho and p, the wind velocity v tapered to rest inside
    the inner magnetosphere by 0.5 (1 + tanh((r - r_taper)/w_taper)) at the mean
    of each cell's corners (r_taper = 0: no taper); magFlux (F,3) of the uniform
    field b as the discrete curl of the edge integrals of A = (b x r)/2 along
    the straight corner-to-corner edges (exact for a linear A)."""
    F = xyz.shape[:3]
    G = tuple(n - 1 for n in F)
    c = 0.125 * sum(xyz[a:a + G[0], b_:b_ + G[1], e:e + G[2]] for a in (0, 1) for b_ in (0, 1) for e in (0, 1))
    r = np.sqrt(np.sum(c * c, axis=-1))
    f = 0.5 * (1.0 + np.tanh((r - r_taper) / w_taper)) if r_taper > 0 else np.ones(G)
    gas = np.empty(G + (5,), order="F")
    gas[..., 0] = rho
    for d in range(3):
        gas[..., 1 + d] = rho * v[d] * f
    gas[..., 4] = p / (gamma - 1.0) + 0.5 * rho * sum(x * x for x in v) * f * f
    bb = np.asarray(b, dtype=np.float64)

    def A(x):   # (b x r)/2, x (..., 3)
        return 0.5 * np.cross(bb, x)

    L = []   # edge integrals along +d from each corner (shape F, last layer along d unused)
    for d in range(3):
        x0 = xyz
        x1 = np.roll(xyz, -1, axis=d)
        L.append(np.sum(A(0.5 * (x0 + x1)) * (x1 - x0), axis=-1))
    s = lambda a, di, dj, dk: np.roll(a, (-di, -dj, -dk), axis=(0, 1, 2))
    mag = np.zeros(F + (3,), order="F")
    mag[..., 0] = L[1] + s(L[2], 0, 1, 0) - s(L[1], 0, 0, 1) - L[2]
    mag[..., 1] = s(L[0], 0, 0, 1) - L[0] - s(L[2], 1, 0, 0) + L[2]
    mag[..., 2] = s(L[1], 1, 0, 0) - L[1] - s(L[0], 0, 1, 0) + L[0]
    # the rolled last layers wrap around and are not faces of any cell of the box:
    # zero them (the BCs and the step never read them)
    mag[-1, :, :, 1:] = 0.0
    mag[:, -1, :, 0::2] = 0.0
    mag[:, :, -1, :2] = 0.0
    return gas, mag
