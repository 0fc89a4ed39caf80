"""GPU parity of the modal filter (MODALFILTER_FLAG) against the oracle.

The kernel applies the reference's x, y and z passes with each sum in the
reference's order and no fused multiply-add, so the filter alone and whole
filtered steps (HEVE and HEVI) agree with the oracle bit for bit (portable pow).
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


@pytest.mark.parametrize("weight", [True, False], ids=["weighted", "plain"])
@pytest.mark.parametrize("p", [3, 7])
def test_filter_apply_bitwise(p, weight):
    base = cases.config1() if p == 3 else cases.config2()
    case = cases.scaled_case(base, 3, 2, 3, porder=p)
    mesh, bs, q = cases.build(case)
    Fh = fedg.modalfilter_matrix(p, 0.0, 0.25, 24)
    Fv = fedg.modalfilter_matrix(p, 2.0 / 3.0, 36.0, 16)
    om = O.OracleMesh(mesh)
    om.set_modalfilter(Fh, Fv)
    dm = fedg.DeviceMesh(mesh)
    dm.setup_modalfilter(Fh, Fv)
    rng = np.random.default_rng(11)
    f = [rng.standard_normal(mesh.Np * mesh.NeA) for _ in range(5)]
    d = [fedg.DevField(a.size, a) for a in f]
    fedg.modalfilter_apply(dm, d[0], d[3], d[4], d[2], d[1], do_weight_Gsqrt=weight)
    om.modalfilter_apply(f, do_weight_Gsqrt=weight)
    for v in range(5):
        assert np.array_equal(d[v].download(), f[v]), fedg.VARS[v]


def test_filter_without_setup_is_an_error():
    case = cases.scaled_case(cases.config1(), 2, 2, 2)
    mesh, bs, q = cases.build(case)
    dm = fedg.DeviceMesh(mesh)
    d = [fedg.DevField(mesh.Np * mesh.NeA) for _ in range(5)]
    with pytest.raises(fedg.FedgError):
        fedg.modalfilter_apply(dm, *d)


@pytest.mark.parametrize("bits", [64])
@pytest.mark.parametrize("p", [3, 7])
def test_filtered_heve_steps_bitwise(p, bits):
    base = cases.config3() if p == 3 else cases.config2()
    case = cases.scaled_case(base, 4, 3, 4, porder=p)
    case.modalfilter = cases.RISING_BUBBLE_FILTER
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    assert cases.setup_modalfilter(case, dm)
    F = fedg.modalfilter_matrix(p, *case.modalfilter)
    om.set_modalfilter(F, F)
    st = fedg.HeveStepper(dm, bs, Dx, Dy, case.dt)
    st.set_modalfilter(True)
    st.set_state(q)
    nsteps = 10
    qq = [a.copy() for a in q]
    om.heve_update(qq, bs, Dx, Dy, case.dt, nsteps)
    st.run(nsteps)
    g = st.get_state()
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    P, D = st.get_pressure()
    Po, Do = om.drhot2pres(qq[1], bs)
    assert np.array_equal(D, Do[:n])
    st.close()


def test_filtered_hevi_steps_bitwise():
    base = cases.config4()
    case = cases.scaled_case(base, 4, 2, 4, porder=3)
    case.modalfilter = cases.RISING_BUBBLE_FILTER
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    cases.setup_modalfilter(case, dm)
    F = fedg.modalfilter_matrix(3, *case.modalfilter)
    om.set_modalfilter(F, F)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=4, rtol=1e-10)
    st.set_modalfilter(True)
    st.set_state(q)
    st.run(3)
    g = st.get_state()
    qq = [a.copy() for a in q]
    om.hevi_update(qq, bs, Dx, Dy, case.dt, 3, scheme=case.scheme, max_it=4, rtol=1e-10)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    st.close()
