"""GPU parity tests of the HEVI path against the CPU oracle, through the C ABI.

The explicit HEVI tendency, the implicit operator (impl_fac = 0), the
reference's one-iteration implicit solve and the Newton iteration to a
tolerance must all agree with the oracle (portable pow) bit for bit: every
kernel evaluates the reference's expressions in the reference's order, and the
line solve's pivoted LU performs the reference's operations on every entry.
"""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def hevi_case(p):
    if p == 3:
        return cases.Case("hevi p3", 3, 4, 3, 4, 12.0e3, 9.0e3, 10.0e3, 2.0, 1, [], base="constBVFreq",
                          scheme="IMEX_ARK324")
    return cases.Case("hevi p7", 7, 3, 2, 3, 3.75e3, 2.5e3, 3.0e3, 0.5, 1, [], base="constBVFreq",
                      scheme="IMEX_ARK324")


def state(mesh, seed, w=0.5):
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = 1e-3 * rng.standard_normal(n)
    q[1][:n] = 0.3 * rng.standard_normal(n)
    q[2][:n] = w * rng.standard_normal(n)
    q[3][:n] = 10.0 + rng.standard_normal(n)
    q[4][:n] = rng.standard_normal(n)
    for f in q:
        mesh.exchange(f)
    return q


@pytest.fixture(scope="module", params=[3, 7], ids=["p3", "p7"])
def setup(request):
    case = hevi_case(request.param)
    mesh = fedg.BoxMesh(case.porder, case.NeX, case.NeY, case.NeZ, (0.0, case.Lx), (0.0, case.Ly),
                        (0.0, case.Lz))
    bs = cases.base_state(mesh, case.base)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    return dict(case=case, mesh=mesh, bs=bs, om=om, Dx=Dx, Dy=Dy, dm=dm)


def dev(a):
    return fedg.DevField(a.size, a)


def test_hevi_explicit_tendency_bitwise(setup):
    s = setup
    mesh, bs, om = s["mesh"], s["bs"], s["om"]
    q = state(mesh, 1)
    _, DP = om.drhot2pres(q[1], bs)
    mesh.exchange(DP)
    om.apply_bc(q)
    ref = om.hevi_cal_tend(q, DP, bs, s["Dx"], s["Dy"])
    out = [dev(mesh.field()) for _ in range(5)]
    fedg.hevi_cal_tend(s["dm"], out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]), dev(DP),
                       dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(bs.PRES_hyd_ref), dev(bs.THERM_hyd),
                       dev(bs.CORIOLIS), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot), dev(s["Dx"]), dev(s["Dy"]))
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(("DENS_dt", "MOMX_dt", "MOMY_dt", "MOMZ_dt", "RHOT_dt"), out, ref):
        g = a.download()[:n]
        assert np.array_equal(g, b[:n]), f"{name}: max abs diff {np.max(np.abs(g - b[:n]))}"


def run_cal_vi(s, q, q0, impl_fac, max_it=None, rtol=0.0):
    mesh, bs = s["mesh"], s["bs"]
    out = [dev(mesh.field()) for _ in range(5)]  # DENS, MOMX, MOMY, MOMZ, RHOT
    it, resid = fedg.hevi_cal_vi(s["dm"], out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]),
                                 dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(q0[0]), dev(q0[3]), dev(q0[4]),
                                 dev(q0[2]), dev(q0[1]), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot), impl_fac,
                                 s["case"].dt, max_it=max_it, rtol=rtol)
    g = [a.download() for a in out]
    return [g[0], g[4], g[3], g[1], g[2]], it, resid  # VID order


@pytest.mark.parametrize("impl_fac", [0.0, 0.8, 2.5])
def test_hevi_cal_vi_one_iteration_bitwise(setup, impl_fac):
    s = setup
    mesh, bs, om = s["mesh"], s["bs"], s["om"]
    q = state(mesh, 2)
    q0 = state(mesh, 3) if impl_fac > 0 else q
    ref, _, _ = om.hevi_cal_vi(q, q0, bs, impl_fac, s["case"].dt)
    got, _, _ = run_cal_vi(s, q, q0, impl_fac)
    n = mesh.Np * mesh.Ne
    for v, name in enumerate(fedg.VARS):
        assert np.array_equal(got[v][:n], ref[v][:n]), \
            f"{name}: max abs diff {np.max(np.abs(got[v][:n] - ref[v][:n]))}"


def test_hevi_newton_to_tolerance(setup):
    s = setup
    mesh, bs, om = s["mesh"], s["bs"], s["om"]
    q = state(mesh, 4, w=2.0)
    q0 = state(mesh, 5, w=2.0)
    ref, it_o, res_o = om.hevi_cal_vi(q, q0, bs, 1.5, s["case"].dt, max_it=10, rtol=1e-10)
    got, it_g, res_g = run_cal_vi(s, q, q0, 1.5, max_it=10, rtol=1e-10)
    assert it_g == it_o, (res_g, res_o)
    assert np.allclose(res_g, res_o[:len(res_g)], rtol=1e-12, atol=0)
    n = mesh.Np * mesh.Ne
    for v in range(5):
        assert np.array_equal(got[v][:n], ref[v][:n])


@pytest.mark.parametrize("max_it,rtol", [(1, 0.0), (6, 1e-10)], ids=["one-iteration", "newton"])
@pytest.mark.parametrize("p", [3, 7])
def test_hevi_steps_match_oracle(p, max_it, rtol):
    """Whole IMEX_ARK324 HEVI steps of the inertia-gravity wave on the device
    against the oracle's driver restatement."""
    base = cases.config4()
    case = cases.scaled_case(base, 4, 2, 4, porder=p)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=max_it, rtol=rtol)
    st.set_state(q)
    st.run(3)
    g = st.get_state()
    qq = [a.copy() for a in q]
    om.hevi_update(qq, bs, Dx, Dy, case.dt, 3, scheme=case.scheme, max_it=max_it, rtol=rtol)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    it, calls = st.newton_stats()
    assert calls == 3 * 3  # three implicit stages per ARK324 step
    assert (it == calls) if max_it == 1 else (calls <= it <= max_it * calls)
    st.close()


@pytest.mark.parametrize("p,NeZ,max_it,rtol", [(3, 4, 1, 0.0), (7, 8, 6, 1e-10)],
                         ids=["p3-nez4-one-iteration", "p7-nez8-newton"])
def test_hevi_steps_save_massflux_bitwise(p, NeZ, max_it, rtol):
    """The tracer coupling in the HEVI stepper: per-stage save_massflux with
    b_ex / b_im weights and the HEVI damping coefficients, against the
    oracle's HEVI loop bit for bit (state, mass fluxes, coefficients); also at
    the channel's vertical (NeZ = 8, p = 7) with Newton to a tolerance."""
    print(f"save_massflux in the HEVI stepper p={p} NeZ={NeZ} max_it={max_it}", flush=True)
    case = cases.scaled_case(cases.config4(), 4, 2, NeZ, porder=p)
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
    qq = [a.copy() for a in q]
    om.heve_set_massflux(MF, al)  # the oracle's coupling hook serves both update loops
    try:
        om.hevi_update(qq, bs, Dx, Dy, case.dt, 2, scheme=case.scheme, max_it=max_it, rtol=rtol)
    finally:
        om.heve_set_massflux(None)
    g = st.get_state()
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), name
    for k, (a, b) in enumerate(zip(gMF, MF)):
        assert np.array_equal(a.download()[:n], b[:n]), f"MFLX {'xyz'[k]}"
    for k, (a, b) in enumerate(zip(gal, al)):
        assert np.array_equal(a.download(), b), f"alph_dens {'MP'[k]}"
    assert np.abs(al[0]).max() > 0.0
    st.set_massflux()
    st.close()
    dm.close()
