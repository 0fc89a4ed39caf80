"""Oracle pins of the atmosphere's diagnostic variables (vars_calc_diagnoseVar_lc,
model/atm_nonhydro3d/src/atmos/mod_atmos_vars_container.F90:1100-1366, with Umet /
Vmet of :431-444 and QDRY of Calc_SpecificHeat).  The reference's tests hold no
vectors for them; what fixes each formula here is the physics it encodes, on
inputs built from a known temperature and density:
  - DENS = DDENS + DENS_hyd exactly; U, V, W times DENS give the momenta back;
    Umet / Vmet equal U / V bit for bit (the regional mesh's Calc_UVmet is the
    identity and the two sums are the same IEEE sum);
  - T of PRES = DENS Rtot T0 is T0; T_diff of the base state vanishes;
  - PT = T (PRES00 / PRES)^(Rtot / CPtot), PT = T where PRES = PRES00, and PT_diff
    of the base state (dry air) vanishes, which ties the hydrostatic term's
    exponent CVdry / CPdry to the first term's Rdry / CPdry;
  - ENGK = |m|^2 / (2 DENS), ENGP = DENS g zlev, ENGI = DENS T0 (qdry CVdry +
    sum q_i CV_i) + DENS sum q_i ENGI0_i, QDRY = 1 - sum q_i MASS_i, and ENGT =
    ENGK + ENGI + DENS g pos_z (a plausible mistake: ENGT with zlev, or a
    dropped ENGI0 term, fails one of them);
  - an unknown name raises, as the reference's PRC_abort.
"""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases, constants

C = constants.SCALE
# QV, QC, QI: TRACER_MASS, TRACER_CV, TRACER_ENGI0 (SCALE's values: vapour carries the
# latent heat of vaporisation, ice minus that of fusion)
MASS = np.array([1.0, 1.0, 1.0])
CVQ = np.array([1384.5, 4218.0, 2106.0])
ENGI0 = np.array([2.5008e6, 0.0, -3.34e5])


def setup(seed=0, p=3, topo=False):
    mesh = fedg.BoxMesh(p, 3, 2, 3, (0.0, 6e3), (0.0, 4e3), (0.0, 4.5e3), periodic=(True, True, False))
    if topo:
        mesh.set_topography(lambda x, y: 250.0 * np.exp(-((x - 3e3) / 1.2e3) ** 2))
    bs = cases.base_state(mesh, "constBVFreq")
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    f = dict(
        DD=1e-3 * rng.standard_normal(n),
        MX=5.0 * rng.standard_normal(n),
        MY=5.0 * rng.standard_normal(n),
        MZ=rng.standard_normal(n),
        T0=rng.uniform(250.0, 300.0, n),
        Q=[0.01 * rng.uniform(0.5, 1.0, n), 1e-4 * rng.uniform(0.0, 1.0, n), 1e-5 * rng.uniform(0.0, 1.0, n)],
    )
    f["Dh"] = np.ascontiguousarray(bs.DENS_hyd[:n])
    f["Ph"] = np.ascontiguousarray(bs.PRES_hyd[:n])
    return mesh, bs, om, f


def diag(om, f, name, P, R, CP, Q=(), zlev=None, pos_z=None):
    return om.atmos_calc_diagvar(name, f["DD"], f["MX"], f["MY"], f["MZ"], P, f["Dh"], f["Ph"], R, CP, Q=Q,
                                 mass=MASS[:len(Q)] if Q else None, cv=CVQ[:len(Q)] if Q else None,
                                 engi0=ENGI0[:len(Q)] if Q else None, zlev=zlev, pos_z=pos_z)


def dry(mesh):
    n = mesh.Np * mesh.Ne
    return np.full(n, C.Rdry), np.full(n, C.CPdry)


def test_density_and_velocities():
    mesh, bs, om, f = setup()
    R, CP = dry(mesh)
    P = f["Ph"].copy()
    dens = f["DD"] + f["Dh"]
    assert np.array_equal(diag(om, f, "DENS", P, R, CP), dens)
    for name, mom in (("U", "MX"), ("V", "MY"), ("W", "MZ")):
        v = diag(om, f, name, P, R, CP)
        assert np.allclose(v * dens, f[mom], rtol=4e-16 * 4, atol=0.0)
    assert np.array_equal(diag(om, f, "Umet", P, R, CP), diag(om, f, "U", P, R, CP))
    assert np.array_equal(diag(om, f, "Vmet", P, R, CP), diag(om, f, "V", P, R, CP))


def test_temperature_and_potential_temperature():
    mesh, bs, om, f = setup(1)
    n = mesh.Np * mesh.Ne
    R, CP = dry(mesh)
    # moist mixture constants, as the specific heats of the tracers would give
    R = R * (1.0 + 0.6 * f["Q"][0])
    CP = CP * (1.0 + 0.9 * f["Q"][0])
    dens = f["DD"] + f["Dh"]
    P = dens * R * f["T0"]
    T = diag(om, f, "T", P, R, CP)
    assert np.allclose(T, f["T0"], rtol=1e-15, atol=0.0)
    with O.pow_mode(O.POW_LIBM):
        PT = diag(om, f, "PT", P, R, CP)
    assert np.allclose(PT, f["T0"] * (C.PRES00 / P) ** (R / CP), rtol=1e-14, atol=0.0)
    # where the pressure is PRES00 the potential temperature is the temperature
    P00 = np.full(n, C.PRES00)
    T00 = diag(om, f, "T", P00, R, CP)
    assert np.array_equal(diag(om, f, "PT", P00, R, CP), T00)
    assert np.array_equal(diag(om, f, "PRES_diff", P, R, CP), P - f["Ph"])


def test_deviations_of_the_base_state_vanish():
    mesh, bs, om, f = setup(2)
    R, CP = dry(mesh)
    f["DD"][:] = 0.0
    P = f["Ph"].copy()
    assert np.all(diag(om, f, "PRES_diff", P, R, CP) == 0.0)
    Td = diag(om, f, "T_diff", P, R, CP)
    assert np.max(np.abs(Td)) <= 1e-13 * 300.0
    for mode in (O.POW_LIBM, O.POW_PORTABLE):
        with O.pow_mode(mode):
            PTd = diag(om, f, "PT_diff", P, R, CP)
            PT = diag(om, f, "PT", P, R, CP)
        assert np.max(np.abs(PTd)) <= 1e-12 * np.max(PT)
        # against the base state's own potential temperature
        assert np.allclose(PT, bs.PT_hyd[: PT.size], rtol=1e-13, atol=0.0)


def test_energies_closed_forms():
    mesh, bs, om, f = setup(3, topo=True)
    n = mesh.Np * mesh.Ne
    Q = f["Q"]
    qdry = 1.0 - sum(q * m for q, m in zip(Q, MASS))
    R = C.Rdry * qdry + 461.46 * Q[0]
    CP = C.CPdry * qdry + 1846.0 * Q[0] + 4218.0 * Q[1] + 2106.0 * Q[2]
    dens = f["DD"] + f["Dh"]
    P = dens * R * f["T0"]
    zlev = np.ascontiguousarray(mesh.height[:n])
    pos_z = np.ascontiguousarray(mesh.z[:n])
    assert not np.array_equal(zlev, pos_z)  # terrain-following: the two heights differ
    m2 = f["MX"] ** 2 + f["MY"] ** 2 + f["MZ"] ** 2
    engk = diag(om, f, "ENGK", P, R, CP)
    assert np.allclose(engk, 0.5 * m2 / dens, rtol=1e-15, atol=0.0)
    engp = diag(om, f, "ENGP", P, R, CP, zlev=zlev)
    assert np.allclose(engp, dens * C.GRAV * zlev, rtol=1e-15, atol=0.0)
    qd = diag(om, f, "QDRY", P, R, CP, Q=Q)
    assert np.allclose(qd, qdry, rtol=1e-15, atol=0.0)
    engi = diag(om, f, "ENGI", P, R, CP, Q=Q)
    want = dens * f["T0"] * (qdry * C.CVdry + sum(q * c for q, c in zip(Q, CVQ))) + dens * sum(
        q * e for q, e in zip(Q, ENGI0))
    assert np.allclose(engi, want, rtol=1e-13, atol=0.0)
    # dry air: the internal energy is DENS CVdry T
    R0, CP0 = dry(mesh)
    P0 = dens * R0 * f["T0"]
    assert np.allclose(diag(om, f, "ENGI", P0, R0, CP0), dens * C.CVdry * f["T0"], rtol=1e-15, atol=0.0)
    engt = diag(om, f, "ENGT", P, R, CP, Q=Q, pos_z=pos_z)
    assert np.allclose(engt, engk + engi + dens * C.GRAV * pos_z, rtol=1e-14, atol=0.0)
    assert not np.allclose(engt, engk + engi + engp, rtol=1e-12, atol=0.0)


def test_unknown_names_are_rejected():
    mesh, bs, om, f = setup()
    R, CP = dry(mesh)
    for name in ("RH", "XYZ", "u"):
        with pytest.raises(ValueError):
            diag(om, f, name, f["Ph"], R, CP)
