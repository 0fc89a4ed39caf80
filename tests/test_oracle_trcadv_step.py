"""The tracer time step of the oracle -- AtmDynDGMDriver_trcadv3d_update for
the tracer-only advection test (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:312-560)
with the low-storage tracer RK update (CMN/scale_timeint_rk.F90:1390-1498) --
pinned on the CPU:

* the SSP tables (Shu-Osher sigma/gamma of
  CMN/scale_timeint_rk_butcher_tab.F90:126-324 turned into Butcher a, b, c as
  ShuOsher2Butcher does) satisfy the order conditions of their stated order,
  and the third-order ones are not fourth order;
* a uniform tracer in a uniform flow stays uniform (the density-weighted
  update is consistent);
* the total tracer mass sum(w J Gsqrt dens q) is conserved with and without
  the limiter;
* with the limiter a step profile stays non-negative, without it it does not;
* a smooth profile carried once around the periodic box returns with an error
  that falls with the order.
"""
import numpy as np
import pytest

import fedg_oracle as O
from fedg import BoxMesh
from fedg.tracer import rk_tables

ORDER = {"ERK_SSP_2s2o": 2, "ERK_SSP_3s3o": 3, "ERK_SSP_4s3o": 3, "ERK_SSP_5s3o_2N2*": 3, "ERK_SSP_10s4o_2N": 4}


@pytest.mark.parametrize("name", sorted(ORDER))
def test_ssp_tables_order_conditions(name):
    t = rk_tables(name)
    A, b, c = t.a, t.b, t.c
    tol = 1e-13
    conds = [(b.sum(), 1.0), (b @ c, 0.5)]
    if ORDER[name] >= 3:
        conds += [(b @ (c * c), 1.0 / 3.0), (b @ (A @ c), 1.0 / 6.0)]
    if ORDER[name] >= 4:
        conds += [(b @ c ** 3, 0.25), (b @ (c * (A @ c)), 1.0 / 8.0), (b @ (A @ c ** 2), 1.0 / 12.0),
                  (b @ (A @ (A @ c)), 1.0 / 24.0)]
    for got, want in conds:
        assert abs(got - want) <= tol, (name, got, want)
    # row sums of A are the stage times
    np.testing.assert_allclose(A.sum(axis=1), c, rtol=0, atol=1e-15)
    if ORDER[name] == 3:
        assert abs(b @ (A @ (A @ c)) - 1.0 / 24.0) > 1e-3 or abs(b @ c ** 3 - 0.25) > 1e-3
    if ORDER[name] == 2:
        assert abs(b @ (c * c) - 1.0 / 3.0) > 1e-3


def flow(mesh, U=(12.0, 5.0, 0.0), rho=1.1, rng=None):
    n = mesh.Np * mesh.Ne
    Dh, DD = mesh.field(rho), mesh.field()
    if rng is not None:
        DD[:n] = 0.05 * rng.standard_normal(n)
    dens = Dh + DD
    M = [dens * u for u in U]
    return M, DD, Dh


def mass(mesh, dens, q):
    w = mesh.elem.w1d
    n1 = len(w)
    w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()
    n = mesh.Np * mesh.Ne
    return float((np.tile(w3, mesh.Ne) * mesh.J * dens[:n] * q[:n]).sum()) if n1 else 0.0


@pytest.mark.parametrize("scheme", ["ERK_SSP_3s3o", "ERK_SSP_10s4o_2N"])
def test_uniform_tracer_stays_uniform(scheme):
    mesh = BoxMesh(3, 3, 3, 2, (0.0, 3.0e3), (0.0, 3.0e3), (0.0, 2.0e3), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    M, DD, Dh = flow(mesh)
    Q = mesh.field(2.5e-3)
    om.tracer_advection(Q, M[0], M[1], M[2], DD, Dh, 5.0, 4, scheme=scheme)
    n = mesh.Np * mesh.Ne
    assert np.abs(Q[:n] / 2.5e-3 - 1.0).max() <= 1e-13


@pytest.mark.parametrize("limiter", [True, False], ids=["limiter", "no-limiter"])
def test_tracer_mass_conserved_and_positivity(limiter):
    mesh = BoxMesh(2, 4, 3, 2, (0.0, 4.0e3), (0.0, 3.0e3), (0.0, 2.0e3), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(8)
    M, DD, Dh = flow(mesh, U=(15.0, -8.0, 2.0), rng=rng)
    dens = Dh + DD
    n = mesh.Np * mesh.Ne
    Q = mesh.field()
    x, y = mesh.x, mesh.y
    Q[:n] = np.where((x > 1.0e3) & (x < 2.2e3) & (y > 0.5e3) & (y < 2.0e3), 1.0e-3, 0.0)
    m0 = mass(mesh, dens, Q)
    om.tracer_advection(Q, M[0], M[1], M[2], DD, Dh, 4.0, 6, disable_limiter=not limiter)
    m1 = mass(mesh, dens, Q)
    assert abs(m1 - m0) <= 1e-12 * m0
    if limiter:
        assert Q[:n].min() >= 0.0
    else:
        assert Q[:n].min() < -1e-5


def test_smooth_profile_once_around_the_box_converges_with_order():
    errs = []
    L, U = 4.0e3, 20.0
    for p in (2, 3, 4):
        mesh = BoxMesh(p, 4, 1, 1, (0.0, L), (0.0, 1.0e3), (0.0, 1.0e3), periodic=(True, True, True))
        om = O.OracleMesh(mesh)
        M, DD, Dh = flow(mesh, U=(U, 0.0, 0.0))
        n = mesh.Np * mesh.Ne
        Q = mesh.field()
        Q0 = 1.0e-3 * (1.2 + np.sin(2.0 * np.pi * mesh.x / L))
        Q[:n] = Q0
        nsteps = 100
        om.tracer_advection(Q, M[0], M[1], M[2], DD, Dh, L / U / nsteps, nsteps)
        errs.append(np.sqrt(np.mean((Q[:n] - Q0) ** 2)) / np.sqrt(np.mean(Q0 ** 2)))
    # measured: 1.7e-2, 1.7e-3, 1.6e-4 (p = 2, 3, 4 on four elements per wavelength)
    assert errs[0] < 3e-2 and errs[1] < errs[0] / 4 and errs[2] < errs[1] / 4, errs


@pytest.mark.parametrize("name", sorted(ORDER))
def test_density_weighted_update_integrates_a_constant_mass_tendency(name):
    """With a constant tendency K of the tracer mass and a density that moves
    linearly from DDENS0 to DDENS over the step, every consistent stage holds
    the exact mass m0 + K dt c_s at stage time c_s: the tracer after stage s is
    (m0 + K dt c_{s+1}) / dens(c_{s+1}), after the last (m0 + K dt) / (DENS_hyd
    + DDENS).  A wrong stage density, a var0 without the density or a dropped
    term of the accumulator breaks it."""
    mesh = BoxMesh(2, 2, 2, 2, (0.0, 2.0e3), (0.0, 2.0e3), (0.0, 2.0e3), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    t = rk_tables(name)
    rng = np.random.default_rng(4)
    n = mesh.Np * mesh.Ne
    Dh, DD0, DD = mesh.field(1.1), mesh.field(), mesh.field()
    DD0[:n], DD[:n] = 0.05 * rng.standard_normal(n), 0.05 * rng.standard_normal(n)
    q, K = mesh.field(), mesh.field()
    q[:n] = rng.uniform(1e-3, 2e-3, n)
    K[:n] = 1e-6 * rng.standard_normal(n)
    dt = 7.0
    m0 = q[:n] * (Dh + DD0)[:n]
    var0, varTmp = mesh.field(), mesh.field()
    for s in range(1, t.nstage + 1):
        om.trcvar_rk_advance(s, t, dt, q, var0, varTmp, K, DD, DD0, Dh)
        if s < t.nstage:
            cs = t.c[s]
            want = (m0 + K[:n] * dt * cs) / (Dh + DD0 + cs * (DD - DD0))[:n]
        else:
            want = (m0 + K[:n] * dt) / (Dh + DD)[:n]
        np.testing.assert_allclose(q[:n], want, rtol=1e-13, atol=0, err_msg=f"{name} stage {s}")


@pytest.mark.parametrize("porder", [3, 4])
def test_tracer_modalfilter_keeps_low_modes_and_element_mass(porder):
    """atm_dyn_dgm_tracer_modalfilter_apply filters Gsqrt dens QTRC: when that
    product is a polynomial of degree <= etac p in each direction the filter
    leaves it alone, and it never changes an element's tracer mass (the mean
    mode is undamped)."""
    from fedg import modalfilter_matrix
    L = 3.0e3
    mesh = BoxMesh(porder, 3, 2, 2, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, True))
    rng = np.random.default_rng(12)
    n = mesh.Np * mesh.Ne
    Gsqrt = mesh.field()
    Gsqrt[:n] = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, n)
    om = O.OracleMesh(mesh, Gsqrt=Gsqrt)
    Dh, DD = mesh.field(1.1), mesh.field()
    DD[:n] = 0.05 * rng.standard_normal(n)
    weight = Gsqrt[:n] * (Dh + DD)[:n]
    MF = modalfilter_matrix(porder)  # etac 2/3: modes up to 2p/3 pass unchanged
    MFv = modalfilter_matrix(porder, etac=0.4, alpha=10.0, order=8)  # only the linear mode passes in z
    x, y, z = mesh.x / L, mesh.y / L, mesh.z / L
    deg = (2 * porder) // 3
    P = (1.0 + x ** deg) * (1.0 + 0.5 * y ** deg) * (1.0 + 0.25 * z)
    Q = mesh.field()
    Q[:n] = 1e-3 * P / weight
    Q0 = Q.copy()
    om.tracer_modalfilter_apply(Q, Dh, DD, MF, MFv)
    np.testing.assert_allclose(Q[:n], Q0[:n], rtol=1e-12, atol=0)
    # random tracer: changed, element masses kept
    Q[:n] = 1e-3 * rng.uniform(0.5, 1.5, n)
    Q0 = Q.copy()
    om.tracer_modalfilter_apply(Q, Dh, DD, MF, MFv)
    assert np.abs(Q[:n] - Q0[:n]).max() > 1e-6
    w = mesh.elem.w1d
    w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()
    W = np.tile(w3, mesh.Ne) * mesh.J * weight
    m0 = (W * Q0[:n]).reshape(mesh.Ne, -1).sum(axis=1)
    m1 = (W * Q[:n]).reshape(mesh.Ne, -1).sum(axis=1)
    np.testing.assert_allclose(m1, m0, rtol=1e-13, atol=0)


def test_dynamics_step_saves_the_time_averaged_mass_flux():
    """With the tracer coupling on, the HEVE update saves sum_s b_s MOM^(s)
    and the b-weighted damping coefficients.  A uniform horizontal wind over
    the horizontally homogeneous hydrostatic base state is a steady solution,
    so after a step MFLX = MOM (sum b_s = 1) and on vertical faces alph_dens =
    rho c + 0 with c the sound speed, to rounding."""
    from fedg import cases, constants
    case = cases.scaled_case(cases.config1(), 3, 2, 3, porder=2)
    mesh, bs, _ = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    U, V = 12.0, -5.0
    dens = bs.DENS_hyd
    q = [mesh.field(), mesh.field(), mesh.field(), dens * U, dens * V]
    nA, nf, n = mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne, mesh.Np * mesh.Ne
    MF = [np.zeros(nA) for _ in range(3)]
    al = [np.zeros(nf), np.zeros(nf)]
    om.heve_set_massflux(MF, al)
    try:
        om.heve_update(q, bs, Dx, Dy, case.dt, 1)
    finally:
        om.heve_set_massflux(None)
    np.testing.assert_allclose(MF[0][:n], dens[:n] * U, rtol=1e-10)
    np.testing.assert_allclose(MF[1][:n], dens[:n] * V, rtol=1e-10)
    assert np.abs(MF[2][:n]).max() <= 1e-8 * abs(U) * dens[:n].max()
    nz = mesh.normal_fn[2 * nf:3 * nf]
    iM = mesh.vmapM - 1
    gam = constants.SCALE.CPdry / constants.SCALE.CVdry
    vert = np.abs(nz) == 1.0
    c = np.sqrt(gam * bs.PRES_hyd[iM] / dens[iM])
    np.testing.assert_allclose(al[0][vert], (c * dens[iM])[vert], rtol=1e-10)


def test_hevi_step_saves_the_time_averaged_mass_flux():
    """The same steady uniform wind through the HEVI update (IMEX_ARK324): the
    b_ex weights (lateral) and the b_im weights (vertical) each sum to one, so
    MFLX = MOM; the damping coefficient takes the HEVI form, Gnn without |nz|:
    on lateral faces alph_dens = (c + |u.n|) rho, on vertical faces the sound
    speed drops out and alph_dens = |w| rho = 0 to rounding (the HEVE form
    would give rho c there)."""
    from fedg import cases, constants
    case = cases.scaled_case(cases.config4(), 3, 2, 4, porder=3)
    mesh, bs, _ = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    U, V = 12.0, -5.0
    dens = bs.DENS_hyd
    q = [mesh.field(), mesh.field(), mesh.field(), dens * U, dens * V]
    nA, nf, n = mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne, mesh.Np * mesh.Ne
    MF = [np.zeros(nA) for _ in range(3)]
    al = [np.zeros(nf), np.zeros(nf)]
    om.heve_set_massflux(MF, al)
    try:
        om.hevi_update(q, bs, Dx, Dy, case.dt, 1, scheme="IMEX_ARK324", max_it=1, rtol=0.0)
    finally:
        om.heve_set_massflux(None)
    np.testing.assert_allclose(MF[0][:n], dens[:n] * U, rtol=1e-10)
    np.testing.assert_allclose(MF[1][:n], dens[:n] * V, rtol=1e-10)
    assert np.abs(MF[2][:n]).max() <= 1e-8 * abs(U) * dens[:n].max()
    nx, ny, nz = (mesh.normal_fn[k * nf:(k + 1) * nf] for k in range(3))
    iM = mesh.vmapM - 1
    gam = constants.SCALE.CPdry / constants.SCALE.CVdry
    c = np.sqrt(gam * bs.PRES_hyd[iM] / dens[iM])
    xf, yf, vert = np.abs(nx) == 1.0, np.abs(ny) == 1.0, np.abs(nz) == 1.0
    np.testing.assert_allclose(al[0][xf], ((c + abs(U)) * dens[iM])[xf], rtol=1e-10)
    np.testing.assert_allclose(al[0][yf], ((c + abs(V)) * dens[iM])[yf], rtol=1e-10)
    assert np.abs(al[0][vert]).max() <= 1e-8 * (c * dens[iM])[vert].max()
