"""Oracle pins of add_phy_tend (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:
1098-1178, ENTOT_CONSERVE_SCHEME_FLAG off) and of its place in the stage loop.

  - the DENS and momentum lines add the physics tendency, nothing else;
  - the THERM line: with RHOH_p = CPtot Pi X, Pi = (Rtot rhotheta / PRES00)^(Rtot /
    CVtot) (the Exner function of the state written without PRES, a closed form
    of (PRES / PRES00)^(Rtot / CPtot) for PRES = PRES00 (Rtot rhotheta /
    PRES00)^(CPtot / CVtot)), the line adds RHOT_tp + X: a wrong exponent or a
    missing CPtot moves it by a percent;
  - in the driver loop a constant DENS_tp c adds exactly c t of mass per unit
    volume over a run (the DG divergence conserves mass, and the RK weights of
    every stage sum to one), and zero tendencies change nothing bit for bit.
"""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases, constants


def mesh_state(p=3, seed=0):
    case = cases.config1()
    case = cases.scaled_case(case, 4, 3, 4, porder=p)
    mesh, bs, q = cases.build(case)
    return case, mesh, bs, q


def test_add_phy_tend_lines():
    case, mesh, bs, q = mesh_state()
    om = O.OracleMesh(mesh)
    c = constants.SCALE
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(3)
    P, _ = om.drhot2pres(q[1], bs)
    T = [mesh.field() for _ in range(5)]
    for a in T:
        a[:n] = rng.standard_normal(n)
    T0 = [a.copy() for a in T]
    tp = [mesh.field() for _ in range(6)]
    for a in tp[:5]:
        a[:n] = rng.standard_normal(n)
    # RHOH_p = CPtot Pi X with the closed-form Exner function of the state
    rhot = bs.THERM_hyd[:n] + q[1][:n]
    R, Cv, Cp = bs.Rtot[:n], bs.CVtot[:n], bs.CPtot[:n]
    Pi = np.power(R * rhot / c.PRES00, R / Cv)
    X = 1e-3 * rng.standard_normal(n)
    tp[5][:n] = Cp * Pi * X
    om.add_phy_tend(T, tp, P, bs.Rtot, bs.CPtot)
    for v in (0, 2, 3, 4):
        assert np.array_equal(T[v][:n], T0[v][:n] + tp[v][:n])
    expect = T0[1][:n] + tp[1][:n] + X
    assert np.max(np.abs(T[1][:n] - expect)) <= 1e-13 * np.max(np.abs(expect))
    # a zero heating rate and zero tendencies leave T unchanged
    T1 = [a.copy() for a in T0]
    om.add_phy_tend(T1, [mesh.field() for _ in range(6)], P, bs.Rtot, bs.CPtot)
    for a, b in zip(T1, T0):
        assert np.array_equal(a[:n], b[:n])


def test_constant_mass_source_in_the_driver_loop():
    case, mesh, bs, q = mesh_state()
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    n = mesh.Np * mesh.Ne
    w = np.tile(np.kron(np.kron(mesh.elem.w1d, mesh.elem.w1d), mesh.elem.w1d), mesh.Ne) * mesh.J[:n]
    cD = 1e-6
    tp = [mesh.field() for _ in range(6)]
    tp[0][:n] = cD
    steps = 5
    q_src = [a.copy() for a in q]
    om.set_phys_tend(tp)
    try:
        om.heve_update(q_src, bs, Dx, Dy, case.dt, steps)
    finally:
        om.set_phys_tend(None)
    q_ref = [a.copy() for a in q]
    om.heve_update(q_ref, bs, Dx, Dy, case.dt, steps)
    vol = w.sum()
    gained = (w * (q_src[0][:n] - q_ref[0][:n])).sum() / vol
    assert abs(gained - cD * case.dt * steps) <= 1e-9 * cD * case.dt * steps
    # zero tendencies through the hook: bit for bit the run without
    q_zero = [a.copy() for a in q]
    zeros = [mesh.field() for _ in range(6)]  # alive while the hook reads them
    om.set_phys_tend(zeros)
    try:
        om.heve_update(q_zero, bs, Dx, Dy, case.dt, steps)
    finally:
        om.set_phys_tend(None)
    for a, b in zip(q_zero, q_ref):
        assert np.array_equal(a[:n], b[:n])
