"""GPU parity of the drop-in entry points and the general-metric steppers on
the inputs a terrain-following FElib host passes, against the CPU oracle,
through the C ABI.

Every input the flat box holds constant is varied here:
  - metrics: random Gsqrt (halo filled), G13 / G23, GsqrtH (the kernels' loads
    of Gsqrt, GI3, GsqrtH through EMap3Dto2D / IndexH2Dto3D);
  - a random Coriolis parameter per 2-D node;
  - moist-air Rtot / CVtot / CPtot that vary node by node (pressure and the
    fused steppers' DPRES);
  - a horizontally varying PRES_hyd with a non-zero PRES_hyd_ref
    (calc_phyd_hgrad, DPhydDx / DPhydDy);
  - walls: NOSLIP on x- / x+ and the bottom, SLIP on the top, periodic in y.

Bar: bit for bit, as for the flat box (the same expression order, the same
portable pow).  The oracle's general-metric terms are pinned by
tests/test_oracle_general_metric.py.
"""
import numpy as np
import pytest

import fedg
from fedg import cases
from fedg.mesh import BND_NONE, BND_NOSLIP, BND_SLIP
import fedg_oracle as O

pytestmark = pytest.mark.gpu

# walls in the boundary order y-, x+, y+, x-, z-, z+
WALLS = (BND_NONE, BND_NOSLIP, BND_NONE, BND_NOSLIP, BND_NOSLIP, BND_SLIP)


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def general_mesh(p, NeX, NeY, NeZ, seed):
    """Box (x non-periodic, y periodic, z walls) with random metrics."""
    dx, dz = (1.5e3, 1.0e3) if p == 3 else (1.25e3, 1.0e3)
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, dx * NeX), (0.0, dx * NeY), (0.0, dz * NeZ),
                        periodic=(False, True, False))
    rng = np.random.default_rng(seed)
    n, nA = mesh.Np * mesh.Ne, mesh.Np * mesh.NeA
    mesh.Gsqrt[:n] = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, n)
    mesh.GI3[:n] = 0.05 * rng.standard_normal(n)
    mesh.GI3[nA:nA + n] = 0.05 * rng.standard_normal(n)
    mesh.exchange(mesh.Gsqrt)
    mesh.exchange(mesh.GI3[:nA])
    mesh.exchange(mesh.GI3[nA:])
    mesh.GsqrtH[:] = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, mesh.GsqrtH.size)
    return mesh


def general_base(mesh, seed, kind="constBVFreq"):
    """Base state with a horizontally varying PRES_hyd (THERM_hyd / DENS_hyd consistent with it),
    a smooth non-zero PRES_hyd_ref, moist Rtot / CVtot / CPtot and a random Coriolis parameter."""
    from fedg import constants
    c = constants.SCALE
    bs = cases.base_state(mesh, kind)
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    x, y = mesh.x / (mesh.x.max() + 1.0), mesh.y / (mesh.y.max() + 1.0)
    fac = 1.0 + 2e-3 * np.sin(2.0 * np.pi * x) * np.cos(2.0 * np.pi * y)
    bs.PRES_hyd[:n] *= fac
    bs.DENS_hyd[:n] *= fac
    bs.THERM_hyd[:n] = cases.calc_rhot_hyd(bs.PRES_hyd[:n], c)
    bs.PRES_hyd_ref[:n] = 0.5 * bs.PRES_hyd[:n] + 30.0 * np.cos(np.pi * x)
    qv = 0.01 * rng.uniform(0.0, 1.0, n)
    bs.Rtot[:n] = c.Rdry * (1.0 - qv) + 461.46 * qv
    bs.CVtot[:n] = c.CVdry * (1.0 - qv) + 1384.5 * qv
    bs.CPtot[:n] = bs.CVtot[:n] + bs.Rtot[:n]
    for f in (bs.PRES_hyd, bs.DENS_hyd, bs.THERM_hyd, bs.PRES_hyd_ref, bs.Rtot, bs.CVtot, bs.CPtot):
        mesh.exchange(f)
    bs.CORIOLIS[:] = 1.0e-4 * (1.0 + 0.3 * rng.standard_normal(bs.CORIOLIS.size))
    return bs


def noisy_state(mesh, seed):
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = 1e-3 * rng.standard_normal(n)
    q[1][:n] = 0.3 * rng.standard_normal(n)
    q[2][:n] = 0.5 * rng.standard_normal(n)
    q[3][:n] = 10.0 + rng.standard_normal(n)
    q[4][:n] = rng.standard_normal(n)
    for f in q:
        mesh.exchange(f)
    return q


def dev(a):
    return fedg.DevField(a.size, a)


SHAPES = {3: (4, 3, 4), 7: (3, 2, 3)}


@pytest.fixture(scope="module", params=[3, 7], ids=["p3", "p7"])
def setup(request):
    p = request.param
    print(f"general inputs p={p} {SHAPES[p]}", flush=True)
    mesh = general_mesh(p, *SHAPES[p], seed=10 + p)
    bs = general_base(mesh, seed=20 + p)
    velbc = mesh.velbc(WALLS)
    om = O.OracleMesh(mesh, velbc=velbc)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh, velbc=velbc)
    assert not dm.flat
    return dict(p=p, mesh=mesh, bs=bs, om=om, Dx=Dx, Dy=Dy, dm=dm, velbc=velbc)


def state_with_pressure(s, seed):
    mesh, bs, om = s["mesh"], s["bs"], s["om"]
    q = noisy_state(mesh, seed)
    _, DP = om.drhot2pres(q[1], bs)
    mesh.exchange(DP)
    om.apply_bc(q)
    return q, DP


def tend_args(s, q, DP):
    bs = s["bs"]
    return [dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]), dev(DP), dev(bs.DENS_hyd), dev(bs.PRES_hyd),
            dev(bs.PRES_hyd_ref), dev(bs.THERM_hyd), dev(bs.CORIOLIS), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot),
            dev(s["Dx"]), dev(s["Dy"])]


NAMES = ("DENS_dt", "MOMX_dt", "MOMY_dt", "MOMZ_dt", "RHOT_dt")


def check_out(mesh, out, ref, what):
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(NAMES, out, ref):
        g = a.download()[:n] if isinstance(a, fedg.DevField) else a[:n]
        assert np.all(np.isfinite(g)), f"{what} {name}: not finite"
        assert np.array_equal(g, b[:n]), f"{what} {name}: max abs diff {np.max(np.abs(g - b[:n]))}"


def test_phyd_hgrad_bitwise(setup):
    s = setup
    mesh = s["mesh"]
    n = mesh.Np * mesh.Ne
    gx, gy = fedg.calc_phyd_hgrad(s["dm"], s["bs"].PRES_hyd, s["bs"].PRES_hyd_ref)
    assert np.max(np.abs(s["Dx"][:n])) > 1e-6  # the base state varies horizontally
    assert np.array_equal(gx[:n], s["Dx"][:n])
    assert np.array_equal(gy[:n], s["Dy"][:n])


def test_pressure_moist_bitwise(setup):
    s = setup
    mesh, bs, om = s["mesh"], s["bs"], s["om"]
    q = noisy_state(mesh, 1)
    P_o, D_o = om.drhot2pres(q[1], bs)
    P, D = dev(mesh.field()), dev(mesh.field())
    fedg.calc_pressure(s["dm"], P, D, dev(q[1]), dev(bs.PRES_hyd), dev(bs.THERM_hyd), dev(bs.Rtot),
                       dev(bs.CVtot), dev(bs.CPtot))
    n = mesh.Np * mesh.Ne
    assert np.array_equal(P.download()[:n], P_o[:n]) and np.array_equal(D.download()[:n], D_o[:n])


def test_halo_and_walls_bitwise(setup):
    """Halo exchange, then SLIP (with metrics) / NOSLIP walls written into the halo."""
    s = setup
    mesh, bs, om = s["mesh"], s["bs"], s["om"]
    q = noisy_state(mesh, 2)
    n = mesh.Np * mesh.Ne
    for f in q:
        f[n:] = 0.0
    dq = [dev(f) for f in q]
    fedg.halo_exchange(s["dm"], dq)
    fedg.apply_bc(s["dm"], dq[0], dq[3], dq[4], dq[2], dq[1], dev(bs.DENS_hyd), dev(bs.PRES_hyd))
    for f in q:
        om.exchange(f)
    om.apply_bc(q)
    for a, b in zip(dq, q):
        assert np.array_equal(a.download(), b)


def test_dropin_cal_tend_bitwise(setup):
    s = setup
    mesh, om, bs = s["mesh"], s["om"], s["bs"]
    q, DP = state_with_pressure(s, 3)
    ref = om.cal_tend(q, DP, bs, s["Dx"], s["Dy"])
    out = [dev(mesh.field()) for _ in range(5)]
    fedg.heve_cal_tend(s["dm"], out, *tend_args(s, q, DP))
    check_out(mesh, out, ref, "cal_tend")


def test_dropin_cco_and_bnd_contrib_bitwise(setup):
    s = setup
    mesh, om, bs = s["mesh"], s["om"], s["bs"]
    q, DP = state_with_pressure(s, 4)
    n = mesh.Np * mesh.Ne
    ref = om.cal_tend_cco(q, DP, bs, s["Dx"], s["Dy"])
    out = [dev(mesh.field()) for _ in range(5)]
    fedg.heve_cal_tend_cco(s["dm"], out, *tend_args(s, q, DP))
    check_out(mesh, out, ref, "cco")
    Dh = DP.copy()
    Dh[n:] = 12345.0
    Dg = dev(Dh)
    om.add_bnd_contrib(ref, q, Dh, bs, moist=True)
    fedg.heve_add_bnd_contrib(s["dm"], out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]), Dg,
                              dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(bs.THERM_hyd), dev(bs.Rtot), dev(bs.CVtot),
                              dev(bs.CPtot))
    check_out(mesh, out, ref, "bnd_contrib")
    assert np.array_equal(Dg.download()[:n + mesh.nB], Dh[:n + mesh.nB])


def test_hevi_explicit_tendency_bitwise(setup):
    s = setup
    mesh, om, bs = s["mesh"], s["om"], s["bs"]
    q, DP = state_with_pressure(s, 5)
    ref = om.hevi_cal_tend(q, DP, bs, s["Dx"], s["Dy"])
    out = [dev(mesh.field()) for _ in range(5)]
    fedg.hevi_cal_tend(s["dm"], out, *tend_args(s, q, DP))
    check_out(mesh, out, ref, "hevi_cal_tend")


@pytest.mark.parametrize("max_it,rtol", [(None, 0.0), (6, 1e-10)], ids=["one-iteration", "newton"])
def test_hevi_cal_vi_bitwise(setup, max_it, rtol):
    s = setup
    mesh, om, bs = s["mesh"], s["om"], s["bs"]
    q = noisy_state(mesh, 6)
    q0 = noisy_state(mesh, 7)
    impl_fac, dt = 0.9, 0.5
    ref, it_o, _ = om.hevi_cal_vi(q, q0, bs, impl_fac, dt, max_it=1 if max_it is None else max_it, rtol=rtol)
    out = [dev(mesh.field()) for _ in range(5)]
    it_g, _ = fedg.hevi_cal_vi(s["dm"], out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]),
                               dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(q0[0]), dev(q0[3]), dev(q0[4]), dev(q0[2]),
                               dev(q0[1]), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot), impl_fac, dt, max_it=max_it,
                               rtol=rtol)
    if max_it is not None:
        assert it_g == it_o
    g = [a.download() for a in out]
    # out is (DENS, MOMX, MOMY, MOMZ, RHOT); the oracle returns VID order
    refo = [ref[0], ref[3], ref[4], ref[2], ref[1]]
    check_out(mesh, g, refo, "cal_vi")


def test_fused_general_steps_bitwise(setup):
    """The general-metric fused HEVE stepper (moist arrays, Coriolis, DPhyd, walls) against
    the oracle's driver loop; with the per-stage mass-flux save for the tracers."""
    s = setup
    mesh, om, bs = s["mesh"], s["om"], s["bs"]
    q, _ = state_with_pressure(s, 8)
    dt = 0.02
    st = fedg.HeveStepper(s["dm"], bs, s["Dx"], s["Dy"], dt, dry=False)
    assert not st.flat
    nA, nf, n = mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne, mesh.Np * mesh.Ne
    MF = [np.zeros(nA) for _ in range(3)]
    al = [np.zeros(nf), np.zeros(nf)]
    gMF = [fedg.DevField(nA) for _ in range(3)]
    gal = [fedg.DevField(nf), fedg.DevField(nf)]
    st.set_massflux(*gMF, *gal)
    st.set_state(q)
    om.heve_set_massflux(MF, al)
    try:
        om.heve_update(q, bs, s["Dx"], s["Dy"], dt, 3)
    finally:
        om.heve_set_massflux(None)
    st.run(3)
    g = st.get_state()
    for name, a, b in zip(fedg.VARS, g, q):
        assert np.all(np.isfinite(a)), name
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    for k, (a, b) in enumerate(zip(gMF, MF)):
        assert np.array_equal(a.download()[:n], b[:n]), f"MFLX {'xyz'[k]}"
    for k, (a, b) in enumerate(zip(gal, al)):
        assert np.array_equal(a.download(), b), f"alph_dens {'MP'[k]}"
    st.set_massflux()
    st.close()


@pytest.mark.parametrize("max_it,rtol", [(1, 0.0), (6, 1e-10)], ids=["one-iteration", "newton"])
def test_hevi_general_steps_bitwise(setup, max_it, rtol):
    s = setup
    mesh, om, bs = s["mesh"], s["om"], s["bs"]
    q, _ = state_with_pressure(s, 9)
    dt = 0.05
    st = fedg.HeviStepper(s["dm"], bs, s["Dx"], s["Dy"], dt, scheme="IMEX_ARK324", max_it=max_it, rtol=rtol)
    st.set_state(q)
    st.run(2)
    g = st.get_state()
    om.hevi_update(q, bs, s["Dx"], s["Dy"], dt, 2, scheme="IMEX_ARK324", max_it=max_it, rtol=rtol)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, q):
        assert np.all(np.isfinite(a)), name
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    st.close()


@pytest.mark.parametrize("max_it,rtol", [(1, 0.0), (6, 1e-10)], ids=["one-iteration", "newton"])
def test_hevi_steps_nez8_p7_bitwise(max_it, rtol):
    """The line solve at the vertical of the 240x16x8 channel (NeZ = 8, 64 nodes per line,
    p = 7): IMEX_ARK324 steps of the inertia-gravity wave on 4x2x8 elements, with the mass-flux
    save for the tracers."""
    print(f"HEVI NeZ=8 p=7 max_it={max_it}", flush=True)
    case = cases.scaled_case(cases.config4(), 4, 2, 8, porder=7)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=max_it, rtol=rtol)
    nA, nf, n = mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne, mesh.Np * mesh.Ne
    MF = [np.zeros(nA) for _ in range(3)]
    al = [np.zeros(nf), np.zeros(nf)]
    gMF = [fedg.DevField(nA) for _ in range(3)]
    gal = [fedg.DevField(nf), fedg.DevField(nf)]
    st.set_massflux(*gMF, *gal)
    st.set_state(q)
    st.run(2)
    g = st.get_state()
    om.heve_set_massflux(MF, al)
    try:
        om.hevi_update(q, bs, Dx, Dy, case.dt, 2, scheme=case.scheme, max_it=max_it, rtol=rtol)
    finally:
        om.heve_set_massflux(None)
    for name, a, b in zip(fedg.VARS, g, q):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    for k, (a, b) in enumerate(zip(gMF, MF)):
        assert np.array_equal(a.download()[:n], b[:n]), f"MFLX {'xyz'[k]}"
    it, calls = st.newton_stats()
    assert calls == 2 * 3
    st.set_massflux()
    st.close()
    dm.close()
