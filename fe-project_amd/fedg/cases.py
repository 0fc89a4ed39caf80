"""Base states and initial states of the benchmark cases (host, numpy).

Base states restate FDS/scale_atm_dyn_dgm_hydrostatic.F90 (constT :86-115,
constPT :129-170, constBVFreq :185-230) and the basic-state potential
temperature field of FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:584-619.  The
warm bubble follows the rising-thermal-bubble initial condition
(model/atm_nonhydro3d/test/case/rising_thermal_bubble/dry/mod_user.F90:216-240)
with the cosine bell of model/atm_nonhydro3d/src/preprocess/mod_mkinit_util.F90:54-165
evaluated at the LGL nodes (its L2 projection onto the same order is the
identity), and pressure left at its hydrostatic value.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import constants
from .mesh import BoxMesh


@dataclass
class BaseState:
    DENS_hyd: np.ndarray
    PRES_hyd: np.ndarray
    THERM_hyd: np.ndarray
    PRES_hyd_ref: np.ndarray
    Rtot: np.ndarray
    CVtot: np.ndarray
    CPtot: np.ndarray
    CORIOLIS: np.ndarray
    PT_hyd: np.ndarray  # potential temperature of the base state, interior nodes


def calc_rhot_hyd(PRES_hyd, c=constants.SCALE):
    """RHOT_hyd = P00/R (P_hyd/P00)^(cv/cp) (..._nonhydro3d_common.F90:584-619)."""
    return c.PRES00 / c.Rdry * (PRES_hyd / c.PRES00) ** (c.CVdry / c.CPdry)


def base_state(mesh: BoxMesh, kind="constT", T0=300.0, PT0=300.0, BVF=0.01, Psfc=1.0e5,
               c=constants.SCALE) -> BaseState:
    n = mesh.Np * mesh.Ne
    z = mesh.height
    D = mesh.field()
    P = mesh.field()
    if kind == "constT":
        H0 = c.Rdry * T0 / c.GRAV
        P[:n] = Psfc * np.exp(-z / H0)
        D[:n] = P[:n] / (c.Rdry * T0)
    elif kind == "constPT":
        RovCP, CPovR = c.Rdry / c.CPdry, c.CPdry / c.Rdry
        exner_sfc = (Psfc / c.PRES00) ** RovCP
        exner = exner_sfc - c.GRAV / (c.CPdry * PT0) * z
        P[:n] = c.PRES00 * exner ** CPovR
        D[:n] = P[:n] / (c.Rdry * exner * PT0)
    elif kind == "constBVFreq":
        RovCP, CPovR = c.Rdry / c.CPdry, c.CPdry / c.Rdry
        exner_sfc = (Psfc / c.PRES00) ** RovCP
        PT = PT0 * np.exp(BVF ** 2 / c.GRAV * z)
        exner = exner_sfc + c.GRAV ** 2 / (c.CPdry * BVF ** 2) * (1.0 / PT - 1.0 / PT0)
        P[:n] = c.PRES00 * exner ** CPovR
        D[:n] = P[:n] / (c.Rdry * exner * PT)
    else:
        raise ValueError(f"unknown base state {kind}")
    mesh.exchange(D)
    mesh.exchange(P)
    T = mesh.field()
    T[:n] = calc_rhot_hyd(P[:n], c)
    mesh.exchange(T)
    PT_hyd = T0 * (c.PRES00 / P[:n]) ** (c.Rdry / c.CPdry) if kind == "constT" else (
        PT0 * np.ones(n) if kind == "constPT" else PT0 * np.exp(BVF ** 2 / c.GRAV * z))
    return BaseState(DENS_hyd=D, PRES_hyd=P, THERM_hyd=T, PRES_hyd_ref=mesh.field(0.0),
                     Rtot=mesh.field(c.Rdry), CVtot=mesh.field(c.CVdry), CPtot=mesh.field(c.CPdry),
                     CORIOLIS=np.zeros(mesh.Nfp_v * mesh.Ne2DA), PT_hyd=PT_hyd)


def cosine_bell(mesh: BoxMesh, amp, r, center):
    """qmax * (0.5 (1 + cos(pi r)))  for r <= 1 (mkinitutil_calc_cosinebell, exponent 1)."""
    rx = ry = rz = r
    xc, yc, zc = center
    rr = np.sqrt(((mesh.x - xc) / rx) ** 2 + ((mesh.y - yc) / ry) ** 2 + ((mesh.height - zc) / rz) ** 2)
    return np.where(rr <= 1.0, amp * (0.5 * (1.0 + np.cos(np.pi * rr))), 0.0)


def warm_bubble_state(mesh: BoxMesh, bs: BaseState, bubbles, c=constants.SCALE):
    """Prognostic perturbations [DDENS, DRHOT, MOMZ, MOMX, MOMY], each (Np,NeA).

    bubbles: list of (amplitude [K], radius [m], (xc, yc, zc) [m])."""
    n = mesh.Np * mesh.Ne
    thp = np.zeros(n)
    for amp, r, ctr in bubbles:
        thp += cosine_bell(mesh, amp, r, ctr)
    RovCp = c.Rdry / c.CPdry
    Ph = bs.PRES_hyd[:n]
    PT = bs.PT_hyd + thp
    DENS = Ph / (c.Rdry * PT * (Ph / c.PRES00) ** RovCp)
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = DENS - bs.DENS_hyd[:n]
    q[1][:n] = DENS * PT - bs.DENS_hyd[:n] * bs.PT_hyd
    return q


# ---------------------------------------------------------------------------
# Benchmark configurations.

@dataclass
class Case:
    name: str
    porder: int
    NeX: int
    NeY: int
    NeZ: int
    Lx: float
    Ly: float
    Lz: float
    dt: float
    nsteps: int
    bubbles: list
    base: str = "constT"
    scheme: str = "ERK_4s4o"
    # modal filter after every step (MF_ETAC, MF_ALPHA, MF_ORDER, the same
    # horizontally and vertically), or None: MODALFILTER_FLAG = .false.
    modalfilter: tuple | None = None


def config1() -> Case:
    """Plain FP64 HEVE warm bubble, 20x20x10 km, 8x8x4 elements, p=3, RK4 dt=0.2 s, 100 steps."""
    return Case("warm bubble p3 8x8x4", 3, 8, 8, 4, 20.0e3, 20.0e3, 10.0e3, 0.2, 100,
                [(2.0, 2.0e3, (10.0e3, 10.0e3, 2.0e3))])


def config2_bubbles(seed=20261202):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        xc, yc = rng.uniform(20.0e3, 140.0e3, size=2)
        zc = rng.uniform(1.5e3, 3.0e3)
        out.append((2.0, 2.0e3, (float(xc), float(yc), float(zc))))
    return out


def config2() -> Case:
    """Plain FP64 HEVE at scale, 160x160x10 km, 64x64x16 elements, p=7, RK4 dt=0.04 s, 200 steps.

    dt = 0.04 s: with 625 m deep p=7 elements the acoustic limit of classic
    RK4 lies between 0.04 s (stable over 750 steps) and 0.05 s (blows up
    near step 180)."""
    return Case("warm bubbles p7 64x64x16", 7, 64, 64, 16, 160.0e3, 160.0e3, 10.0e3, 0.04, 200, config2_bubbles())


# Modal filter of the reference's dry rising-thermal-bubble case
# (model/atm_nonhydro3d/test/case/rising_thermal_bubble/dry/run.conf:53,71-76).
RISING_BUBBLE_FILTER = (0.0, 0.25, 24)


def config3() -> Case:
    """Precision form on the config-1 shape for one simulated hour (18000 steps of 0.2 s).

    Without a filter the unfiltered FP64 warm bubble itself blows up after
    about 5000 steps (aliasing; the oracle's run is finite at step 4000 and
    NaN by 5000), so the hour runs with the modal filter the reference's
    rising-thermal-bubble case uses, in the FP64 and the FP32 run alike."""
    c = config1()
    c.name = "warm bubble p3 8x8x4, one hour, modal filter"
    c.nsteps = 18000
    c.modalfilter = RISING_BUBBLE_FILTER
    return c


def setup_modalfilter(case: Case, dm):
    """Upload the case's filter matrices to a DeviceMesh (no-op without a filter); returns whether set."""
    if case.modalfilter is None:
        return False
    from .dyncore import modalfilter_matrix
    F = modalfilter_matrix(case.porder, *case.modalfilter)
    dm.setup_modalfilter(F, F)
    return True


def config4(dt=None) -> Case:
    """HEVI at scale: the inertia-gravity wave of Skamarock & Klemp (1994) in a
    300 x 20 x 10 km periodic channel, 240x16x8 elements, p=7, constant
    Brunt-Vaisala frequency 0.01 1/s, theta0 = 300 K, mean wind 20 m/s,
    theta' = 0.01 K sin(pi z/H) / (1 + ((x - 100 km)/5 km)^2), IMEX_ARK324."""
    return Case("inertia-gravity wave p7 240x16x8", 7, 240, 16, 8, 300.0e3, 20.0e3, 10.0e3,
                CONFIG4_DT if dt is None else dt, 900, [], base="constBVFreq", scheme="IMEX_ARK324")


# Time step of config4: the horizontal acoustics stay explicit in HEVI, and
# with 1.25 km p=7 elements IMEX_ARK324 is unstable at 0.1 s and stable at
# 0.05 s (CPU oracle runs of 60 simulated seconds on an 8x2x8 piece of the
# channel); 2 s, the step a vertical-only acoustic limit would allow, is far
# beyond the horizontal limit.
CONFIG4_DT = 0.05


def igw_state(mesh: BoxMesh, bs: BaseState, U=20.0, dtheta=0.01, xc=100.0e3, a=5.0e3, theta0=300.0, BVF=0.01,
              c=constants.SCALE):
    """Inertia-gravity-wave initial state (model/atm_nonhydro3d/test/case/inertia_gravity_wave/mod_user.F90:
    231-274 with the bell in x only and a uniform mean wind): pressure left hydrostatic (DRHOT = 0), density
    from the perturbed potential temperature, MOMX = DENS * U."""
    n = mesh.Np * mesh.Ne
    x, z = mesh.x, mesh.z
    H = float(np.max(z) - np.min(z))
    PT = theta0 * np.exp(BVF ** 2 / c.GRAV * z) + dtheta * np.sin(np.pi * z / H) / (1.0 + ((x - xc) / a) ** 2)
    P = bs.PRES_hyd[:n]
    DENS = P / (c.Rdry * PT * (P / c.PRES00) ** (c.Rdry / c.CPdry))
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = DENS - bs.DENS_hyd[:n]
    q[3][:n] = DENS * U
    for f in q:
        mesh.exchange(f)
    return q


def scaled_case(base: Case, NeX, NeY, NeZ, porder=None, nsteps=None) -> Case:
    """Same physics on a smaller element count with the element size of `base`."""
    p = base.porder if porder is None else porder
    sx, sy, sz = NeX / base.NeX, NeY / base.NeY, NeZ / base.NeZ
    Lx, Ly, Lz = base.Lx * sx, base.Ly * sy, base.Lz * sz
    bubbles = [(a, r, (min(x, 0.5 * Lx), min(y, 0.5 * Ly), min(z, 0.4 * Lz))) for a, r, (x, y, z) in base.bubbles[:1]]
    return Case(f"{base.name} scaled to {NeX}x{NeY}x{NeZ} p{p}", p, NeX, NeY, NeZ, Lx, Ly, Lz, base.dt,
                base.nsteps if nsteps is None else nsteps, bubbles, base.base, base.scheme, base.modalfilter)


def build(case: Case, c=constants.SCALE):
    """Mesh, base state and initial prognostic state of a case (host)."""
    mesh = BoxMesh(case.porder, case.NeX, case.NeY, case.NeZ, (0.0, case.Lx), (0.0, case.Ly), (0.0, case.Lz),
                   periodic=(True, True, False))
    bs = base_state(mesh, case.base, c=c)
    if case.scheme.startswith("IMEX"):
        q = igw_state(mesh, bs, xc=min(100.0e3, 0.5 * case.Lx), c=c)
    else:
        q = warm_bubble_state(mesh, bs, case.bubbles, c=c)
    return mesh, bs, q
