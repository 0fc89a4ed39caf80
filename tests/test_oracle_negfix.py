"""Oracle pins of the coupled tracer step's thermodynamics: the specific heats
of the mixture (the mixing rule of SCALE's ATMOS_THERMODYN_specific_heat, used
by update_pressure_specific_heat, FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:
653-716) and the negative fixer (FElib/src/microphysics/
scale_atm_phy_mp_dgm_common.F90:429-556).  Parity with SCALE's own
specific_heat routine is unpinned (that library is not in the reference tree);
these checks fix what the mathematics does:
  - dry air (every q = 0) gives Rdry, CVdry, CPdry exactly, one species the
    two-component closed form;
  - with no negative value the fixer changes nothing but the rounding of
    (rho q) / rho;
  - with negatives: every species keeps its element mass max(M0, 0) and
    becomes non-negative; the element's total mass changes only by the
    clipped negative species mass; the internal energy plus the potential
    energy of the moved mass keeps its element integral; DRHOT is the
    rhotheta of the new pressure and specific heats.
"""
import numpy as np

import fedg
import fedg_oracle as O
from fedg import cases, constants

C = constants.SCALE
# vapor, cloud water, cloud ice: TRACER_MASS, _R, _CV, _CP (SCALE's values for QV, QC, QI)
THERMO = (np.array([1.0, 1.0, 1.0]), np.array([461.46, 0.0, 0.0]), np.array([1384.5, 4218.0, 2106.0]),
          np.array([1846.0, 4218.0, 2106.0]))


def setup(p=3, seed=0):
    case = cases.scaled_case(cases.config1(), 3, 2, 3, porder=p)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    Q = [mesh.field() for _ in range(3)]
    Q[0][:n] = 0.01 * rng.uniform(0.5, 1.0, n)
    Q[1][:n] = 1e-4 * rng.uniform(0.0, 1.0, n)
    Q[2][:n] = 1e-5 * rng.uniform(0.0, 1.0, n)
    return mesh, bs, q, om, Q, rng


def weights(mesh):
    e = mesh.elem
    w = np.kron(np.kron(e.w1d, e.w1d), e.w1d)  # node p = i + n1 (j + n1 k): w_i w_j w_k
    return np.tile(w, mesh.Ne) * mesh.Gsqrt[:mesh.Np * mesh.Ne] * mesh.J


def elem_sum(mesh, x):
    return x.reshape(mesh.Ne, mesh.Np).sum(axis=1)


def test_specific_heat_closed_forms():
    mesh, bs, q, om, Q, rng = setup()
    n = mesh.Np * mesh.Ne
    R, CV, CP = mesh.field(), mesh.field(), mesh.field()
    zero = [mesh.field() for _ in range(3)]
    om.tracer_specific_heat(zero, THERMO, R, CV, CP)
    assert np.all(R[:n] == C.Rdry) and np.all(CV[:n] == C.CVdry) and np.all(CP[:n] == C.CPdry)
    qv = Q[0]
    om.tracer_specific_heat([qv, zero[1], zero[2]], THERMO, R, CV, CP)
    np.testing.assert_allclose(R[:n], (1 - qv[:n]) * C.Rdry + qv[:n] * 461.46, rtol=1e-15)
    np.testing.assert_allclose(CV[:n], (1 - qv[:n]) * C.CVdry + qv[:n] * 1384.5, rtol=1e-15)
    np.testing.assert_allclose(CP[:n], (1 - qv[:n]) * C.CPdry + qv[:n] * 1846.0, rtol=1e-15)
    om.tracer_specific_heat(Q, THERMO, R, CV, CP)
    qs = Q[0][:n] + Q[1][:n] + Q[2][:n]
    np.testing.assert_allclose(CV[:n], (1 - qs) * C.CVdry + Q[0][:n] * 1384.5 + Q[1][:n] * 4218.0
                               + Q[2][:n] * 2106.0, rtol=1e-15)


def thermo_state(mesh, bs, q, om, Q):
    n = mesh.Np * mesh.Ne
    R, CV, CP = mesh.field(), mesh.field(), mesh.field()
    om.tracer_specific_heat(Q, THERMO, R, CV, CP)
    P, _ = om.drhot2pres(q[1], O_bs(bs, R, CV, CP, n))
    return R, CV, CP, P


def O_bs(bs, R, CV, CP, n):
    import copy
    b = copy.copy(bs)
    b.Rtot, b.CVtot, b.CPtot = R, CV, CP
    return b


def test_fixer_without_negatives_changes_only_rounding():
    mesh, bs, q, om, Q, rng = setup()
    n = mesh.Np * mesh.Ne
    R, CV, CP, P = thermo_state(mesh, bs, q, om, Q)
    DD, Q1 = q[0].copy(), [a.copy() for a in Q]
    P1, R1, CV1, CP1 = P.copy(), R.copy(), CV.copy(), CP.copy()
    om.trc_negative_fixer(Q1, DD, P1, CV1, CP1, R1, bs.DENS_hyd, bs.PRES_hyd, mesh.J, mesh.height, THERMO)
    assert np.array_equal(DD[:n], q[0][:n])
    for a, b in zip(Q1, Q):
        np.testing.assert_allclose(a[:n], b[:n], rtol=4e-16, atol=0)
    np.testing.assert_allclose(P1[:n], P[:n], rtol=1e-15)
    np.testing.assert_allclose(R1[:n], R[:n], rtol=1e-15)


def test_fixer_removes_negatives_with_mass_and_energy_kept():
    mesh, bs, q, om, Q, rng = setup(seed=5)
    n = mesh.Np * mesh.Ne
    Q[1][:n] -= 6e-5  # some cloud water negative
    Q[2][:n] = 1e-5 * rng.standard_normal(n)  # ice of either sign, whole elements negative
    R, CV, CP, P = thermo_state(mesh, bs, q, om, Q)
    w = weights(mesh)
    dens0 = bs.DENS_hyd[:n] + q[0][:n]
    M0 = [elem_sum(mesh, w * dens0 * a[:n]) for a in Q]
    zl = mesh.height
    E0 = elem_sum(mesh, w * CV[:n] * P[:n] / R[:n])
    DD = q[0].copy()
    Q1 = [a.copy() for a in Q]
    P1, R1, CV1, CP1, DR = P.copy(), R.copy(), CV.copy(), CP.copy(), q[1].copy()
    om.trc_negative_fixer(Q1, DD, P1, CV1, CP1, R1, bs.DENS_hyd, bs.PRES_hyd, mesh.J, zl, THERMO, DRHOT=DR)
    dens1 = bs.DENS_hyd[:n] + DD[:n]
    for iq in range(3):
        assert np.all(Q1[iq][:n] >= 0.0)
        M1 = elem_sum(mesh, w * dens1 * Q1[iq][:n])
        np.testing.assert_allclose(M1, np.maximum(M0[iq], 0.0), rtol=1e-12, atol=1e-14 * np.abs(M0[iq]).max())
    clipped = sum(np.maximum(m, 0.0) - m for m in M0)
    np.testing.assert_allclose(elem_sum(mesh, w * (DD[:n] - q[0][:n])), clipped, rtol=1e-10,
                               atol=1e-12 * np.abs(M0[0]).max())
    E1 = elem_sum(mesh, w * (CV1[:n] * P1[:n] / R1[:n] + (DD[:n] - q[0][:n]) * C.GRAV * zl[:n]))
    np.testing.assert_allclose(E1, E0, rtol=1e-12)
    rhot = C.PRES00 / R1[:n] * (P1[:n] / C.PRES00) ** (CV1[:n] / CP1[:n])
    np.testing.assert_allclose(DR[:n] + bs.THERM_hyd[:n], rhot, rtol=1e-13)
