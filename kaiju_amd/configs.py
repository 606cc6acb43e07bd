"""Problem specifications for the GAMERA step (grid, initial condition, model knobs).

These are plain data shared by the HIP host mirror (``kaiju_amd.gamera``), the
tests and ``bench.py``.  Each field maps onto a key the reference reads from its
XML deck (reference ``src/gamera/init.F90:312-492`` ``initModel`` and
``init.F90:572-798`` ``genGridXML``); values that the reference derives from
single-precision literals are written out explicitly so every consumer sees the
same double.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple

import numpy as np

# init.F90:32 MaxCFL = 0.3 is a default-real literal, widened to double.
CFL_DEFAULT = 0.30000001192092896


# the library's boundary types, by code (gam_set_bcs)
BC_TYPES = ("periodic", "zerograd", "zeroexten", "reflect", "sphere", "lfm", "wind", "cylpole", "lazy", "ion")


def ring_chunks(Np: int) -> Tuple[int, ...]:
    """Default chunks per ring of an LFM grid with Np cells around the axis
    (reference ringutils.F90:132-195, the non-aggressive set)."""
    table = {8: (4,), 16: (8,), 32: (8, 16), 64: (8, 16, 32, 32), 128: (8, 16, 32, 32, 64, 64, 64, 64),
             256: (8, 16, 32, 32, 64, 64, 64, 64, 128, 128, 128, 128),
             512: (8, 16, 32, 32, 64, 64, 64, 64, 128, 128, 128, 128, 256, 256, 256, 256)}
    if Np not in table:
        raise ValueError(f"no default ring chunking for Np = {Np}")
    return table[Np]


@dataclasses.dataclass
class ModelSpec:
    """Model_T knobs (reference base/types/gamtypes.F90:71-146)."""
    gamma: float = 5.0 / 3.0
    CFL: float = 0.3
    pdmb: float = 0.75
    Vd0: float = 0.5
    cHogH: float = 0.25
    cHogM: float = 0.25
    Ca: float = 1.0e12          # reference: Ca = HUGE when Boris is off
    dFloor: float = 1.0e-12     # gdefs.F90:13 TINY
    pFloor: float = 1.0e-12
    tFin: float = 1.0e30
    limDT0: float = 1.0e-3
    limCPR: float = 1.0e-1
    rmeth: str = "7UP"          # "7UP" or "8CENT" (init.F90:450-462)
    doLFMLim: bool = True
    doBoris: bool = False
    doMHD: bool = True
    doCPR: bool = False
    precision: int = 0          # 0 plain FP64; 1 FP32 bulk with FP64 CT, update and dt
    fixedTimestep: bool = False  # dt kept as set (gam_set_clock); CalcDT's checks skipped


@dataclasses.dataclass
class CaseSpec:
    name: str
    N: Tuple[int, int, int]
    bounds: Tuple[float, float, float, float, float, float]
    grid: str = "cart"          # "cart" | "warp3" | "refwarp" | "cyl" | "sph"
    warp_amp: float = 0.0
    ic: str = "ot"              # "ot" | "bw"
    ic_params: dict = dataclasses.field(default_factory=dict)
    model: ModelSpec = dataclasses.field(default_factory=ModelSpec)
    nG: int = 4
    dt0_after: int = 1          # console output (dt0 capture) after this many steps
    steps: int = 100
    # background field B0 = bu + dipole(x - xc) with moment (0,0,M0): a dict
    # with keys M0, xc, bu (the reference's Model%doBackground + Model%B0)
    background: Optional[dict] = None
    # boundary type per face INI, OUTI, INJ, OUTJ, INK, OUTK (one of BC_TYPES)
    bcs: Tuple[str, str, str, str, str, str] = ("periodic",) * 6
    # ring average about the LFM axis (Model%doRing): chunks per ring, or None
    ring: Optional[Tuple[int, ...]] = None
    # constant solar wind (rho, p, v[3], b[3]) of a "wind" outer i boundary
    wind: Optional[tuple] = None
    # or a time series of it (takes precedence over wind)
    wind_series: Optional["WindSeries"] = None

    @property
    def zones(self) -> int:
        return self.N[0] * self.N[1] * self.N[2]

    @property
    def do25D(self) -> bool:
        return self.N[2] == 1


@dataclasses.dataclass
class WindSeries:
    """A solar-wind time series in code units, as readWind
    (apps/wind.F90:617-760) holds it after its unit conversion: sample times t
    (increasing, uniformly spaced), density rho, pressure p, velocity v and
    field b (nt x 3; b's x component is replaced by Bx0 + ByC By + BzC Bz),
    the marker position xyzBS (nt x 3, the bow shock; None: the origin) and
    the front coefficients ByC, BzC, Bx0."""
    t: np.ndarray
    rho: np.ndarray
    p: np.ndarray
    v: np.ndarray
    b: np.ndarray
    xyzBS: Optional[np.ndarray] = None
    ByC: float = 0.0
    BzC: float = 0.0
    Bx0: float = 0.0

    def arrays(self):
        """FP64 arrays (t, rho, p, v, b, xyzBS or None), the (nt, 3) ones in
        Fortran order (each component's series contiguous)."""
        f = lambda a, shp: np.asfortranarray(np.asarray(a, dtype=np.float64).reshape(shp))
        nt = len(self.t)
        return (f(self.t, (nt,)), f(self.rho, (nt,)), f(self.p, (nt,)), f(self.v, (nt, 3)), f(self.b, (nt, 3)),
                None if self.xyzBS is None else f(self.xyzBS, (nt, 3)))


def orszag_tang(n=256, nk: int = 4, steps: int = 200, **mkw) -> CaseSpec:
    """2D Orszag-Tang vortex on [0,2pi]^2 x [0, nk*2pi/ni], gamma = 5/3.

    rho = gamma^2, p = gamma, v = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0);
    the reference ``initOT2D`` (prob.F90:74-119) with a unit wavenumber.
    ``n`` is the cells per axis in x and y, or an (ni, nj, nk) tuple (then
    ``nk`` is ignored).
    """
    g = 5.0 / 3.0
    L = 2.0 * math.pi
    m = ModelSpec(gamma=g, CFL=0.3, **mkw)
    ni, nj, nk = (n, n, nk) if isinstance(n, int) else tuple(n)
    return CaseSpec(name=f"ot{ni}x{nj}x{nk}", N=(ni, nj, nk),
                    bounds=(0.0, L, 0.0, L, 0.0, nk * L / ni), grid="cart", ic="ot",
                    ic_params=dict(rho0=g * g, p0=g, b0=1.0, kscale=1.0),
                    model=m, steps=steps)


def blast_wave(n: int = 256, steps: int = 100, warp_amp: float = 0.05, B0: float = 10.0,
               pin: float = 10.0, pout: float = 0.1, **mkw) -> CaseSpec:
    """3D MHD blast wave on a warped non-orthogonal grid over the unit cube.

    x = xi + a sin(2 pi xi) sin(2 pi eta) sin(2 pi zeta) (same for y, z);
    rho = 1, p = pout outside / pin inside r < 0.1, v = 0, B = B0 (1,1,0)/sqrt 2;
    the reference ``initBW`` (prob.F90:190-298) with x0 = y0 = z0 = 0.5.
    """
    m = ModelSpec(gamma=5.0 / 3.0, CFL=0.3, **mkw)
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    return CaseSpec(name=f"bw{nx}x{ny}x{nz}", N=(nx, ny, nz), bounds=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0),
                    grid="warp3" if warp_amp else "cart", warp_amp=warp_amp, ic="bw",
                    ic_params=dict(x0=0.5, y0=0.5, z0=0.5, D0=1.0, P0=pout,
                                   pRat=pin / pout, dRat=1.0, rC=0.1, B0=B0),
                    model=m, steps=steps)


def cylinder_blast(N=(16, 16, 8), steps: int = 8, r_in: float = 0.5, r_out: float = 1.5, **mkw) -> CaseSpec:
    """Blast wave in a cylindrical shell (R in [r_in, r_out], phi once around,
    z in [0, 1]; the reference's doCyl grid, init.F90:712-716) with the
    through-pole BC on the inner radial face (cylpole, bcs.F90:987-1017; the
    reference's ringBW pairs it with zero gradient outside, prob.F90:145-162)
    and periodic phi and z.  The shell keeps every cell off the pole, so the
    metric stays finite; the inner ghosts take the cells half a turn away."""
    m = ModelSpec(gamma=5.0 / 3.0, CFL=0.3, **mkw)
    nr, nph, nz = N
    return CaseSpec(name=f"cylbw{nr}x{nph}x{nz}", N=(nr, nph, nz), bounds=(r_in, r_out, 0.0, 1.0, 0.0, 1.0),
                    grid="cyl", ic="bw",
                    ic_params=dict(x0=1.0, y0=0.0, z0=0.5, D0=1.0, P0=0.1, pRat=50.0, dRat=1.0, rC=0.2, B0=2.0),
                    model=m, steps=steps, bcs=("cylpole", "zerograd", "periodic", "periodic", "periodic", "periodic"))


def magnetosphere(N=(192, 192, 256), steps: int = 500, precision: int = 1, r_in: float = 2.0, r_out: float = 30.0,
                  theta0: Optional[float] = None, Ca_kms: float = 1000.0, axis: bool = True,
                  wind: Optional[bool] = None, chunks: Optional[Tuple[int, ...]] = None, inner: str = "ion",
                  **mkw) -> CaseSpec:
    """Magnetosphere-like block (kaiju_amd/msph.py): stretched spherical-polar
    grid r in [2, 30] R_E around the x axis, Earth dipole -31000 nT R_E^3 as the
    background field, solar wind rho = 5 cm^-3, v = (-400, 0, 0) km/s (at rest
    inside ~12 R_E at the start), T = 1e5 K, IMF Bz = -5 nT, Boris speed cap
    1000 km/s, FP32 bulk.

    axis = True (the reference's LFM-grid layout): theta spans [theta0, pi -
    theta0] with theta0 = 1e-3 by default, the j faces take the LFM axis BCs and
    the rings next to the axis are averaged with the default chunking for Np
    (or chunks: the chunks of each ring, innermost first).
    axis = False: a wedge over [theta0, pi - theta0] (default theta0 = 0.15)
    with zero-gradient theta walls.  wind (default: axis): the outer radial face
    drives the wind in (WindBC with this constant wind) instead of SphOut.
    inner: the inner radial face's BC: "ion" (the Earth setup's IonInner,
    voltron/ICs/earthcmi.F90:378-490, with no ionosphere attached: the gas of
    the mirrored cell at rest, zero perturbation flux, so the field there is
    the dipole), "lazy" (the ghosts keep the initial state) or "sphere"
    (SphIn)."""
    from kaiju_amd import msph
    rho, v, p, b = msph.wind_state()
    m = ModelSpec(gamma=5.0 / 3.0, CFL=0.3, doBoris=True, Ca=Ca_kms / 100.0, precision=precision, **mkw)
    nr, nt, nphi = N
    if theta0 is None:
        theta0 = msph.AXIS_THETA0 if axis else 0.15
    wall = "lfm" if axis else "zerograd"
    if wind is None:
        wind = axis
    return CaseSpec(name=f"msph{nr}x{nt}x{nphi}" + ("" if axis else "w"), N=(nr, nt, nphi),
                    bounds=(r_in, r_out, theta0, math.pi - theta0, 0.0, 2 * math.pi),
                    grid="sph", ic="msph",
                    ic_params=dict(r_in=r_in, r_out=r_out, theta0=theta0, rho=rho, v=v, p=p, b=b, axis=axis),
                    model=m, steps=steps, background=dict(M0=msph.dipole_moment(), xc=(0.0, 0.0, 0.0),
                                                          bu=(0.0, 0.0, 0.0)),
                    bcs=(inner, "wind" if wind else "sphere", wall, wall, "periodic", "periodic"),
                    ring=(tuple(chunks) if chunks else ring_chunks(nphi)) if axis else None,
                    wind=(rho, p, v, b) if wind else None)


# The two workloads named by the benchmark configuration list.
def config_c1() -> CaseSpec:
    return orszag_tang(256, 4, 200)


def config_c2() -> CaseSpec:
    return blast_wave(256, 100)
