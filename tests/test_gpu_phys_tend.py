"""Physics tendencies (add_phy_tend, FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:
846-856, 1098-1178) on the GPU against the oracle, bit for bit: the drop-in
entry point, the fused HEVE stepper (flat box at p = 3 and p = 7, and a
terrain-following mesh with moist Rtot / CVtot / CPtot) and the HEVI stepper.
The oracle's add_phy_tend is pinned by tests/test_oracle_phys_tend.py."""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases
from test_gpu_general_inputs import WALLS, general_base, general_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def dev(a):
    return fedg.DevField(a.size, a)


def phys_fields(mesh, seed, scale=1.0):
    """DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOH_p (oracle order) with the halo filled."""
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    x, z = mesh.x / (mesh.x.max() + 1.0), mesh.z / (mesh.z.max() + 1.0)
    amp = (1e-7, 1e-4, 1e-5, 1e-5, 1e-5, 1e-1)
    tp = []
    for k, a in enumerate(amp):
        f = mesh.field()
        f[:n] = scale * a * (np.sin(2.0 * np.pi * (x + 0.1 * k)) * np.cos(np.pi * z) + 0.1 * rng.standard_normal(n))
        mesh.exchange(f)
        tp.append(f)
    return tp


def device_order(tp):
    """(DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p): the argument order of the C ABI."""
    return [tp[0], tp[3], tp[4], tp[2], tp[1], tp[5]]


def test_add_phy_tend_bitwise():
    case = cases.scaled_case(cases.config1(), 4, 3, 4, porder=3)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    dm = fedg.DeviceMesh(mesh)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(7)
    P, _ = om.drhot2pres(q[1], bs)
    T = [mesh.field() for _ in range(5)]
    for a in T:
        a[:n] = rng.standard_normal(n)
    tp = phys_fields(mesh, 8, scale=10.0)
    Rt, Cp = mesh.field(), mesh.field()
    Rt[:n] = bs.Rtot[:n] * (1.0 + 0.01 * rng.uniform(size=n))
    Cp[:n] = bs.CPtot[:n] * (1.0 + 0.01 * rng.uniform(size=n))
    gT = [dev(a) for a in T]
    # C ABI order: DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt
    fedg.add_phy_tend(dm, [gT[0], gT[3], gT[4], gT[2], gT[1]], *[dev(a) for a in device_order(tp)], dev(P),
                      dev(Rt), dev(Cp))
    om.add_phy_tend(T, tp, P, Rt, Cp)
    for name, a, b in zip(fedg.VARS, gT, T):
        g = a.download()[:n]
        assert np.array_equal(g, b[:n]), f"{name}: max abs diff {np.max(np.abs(g - b[:n]))}"


def run_heve(mesh, bs, q, om, dm, dt, tp, steps, dry=True):
    Dx, Dy = om.calc_phyd_hgrad(bs)
    st = fedg.HeveStepper(dm, bs, Dx, Dy, dt, dry=dry)
    gtp = [dev(a) for a in device_order(tp)]
    st.set_phys_tend(*gtp)
    st.set_state(q)
    st.run(steps)
    g = st.get_state()
    qq = [a.copy() for a in q]
    om.set_phys_tend(tp)
    try:
        om.heve_update(qq, bs, Dx, Dy, dt, steps)
    finally:
        om.set_phys_tend(None)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    # the physics changed the run
    st2 = fedg.HeveStepper(dm, bs, Dx, Dy, dt, dry=dry)
    st2.set_state(q)
    st2.run(steps)
    assert not np.array_equal(st2.get_state()[1], g[1])
    st.set_phys_tend()
    st.close()
    st2.close()


@pytest.mark.parametrize("p,shape", [(3, (4, 3, 4)), (7, (3, 2, 3))], ids=["p3", "p7"])
def test_heve_steps_with_physics_bitwise(p, shape):
    print(f"HEVE steps with physics tendencies p={p} {shape}", flush=True)
    case = cases.scaled_case(cases.config1(), *shape, porder=p)
    if p == 7:
        case = cases.Case("bubble p7", 7, *shape, 3.75e3, 2.5e3, 3.0e3, 0.04, 1, case.bubbles)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    dm = fedg.DeviceMesh(mesh)
    assert dm.flat
    run_heve(mesh, bs, q, om, dm, case.dt, phys_fields(mesh, 11 + p), 3)


def test_heve_steps_with_physics_general_mesh_bitwise():
    print("HEVE steps with physics tendencies, terrain-following mesh p=3", flush=True)
    mesh = general_mesh(3, 4, 3, 4, seed=31)
    bs = general_base(mesh, seed=32)
    velbc = mesh.velbc(WALLS)
    om = O.OracleMesh(mesh, velbc=velbc)
    dm = fedg.DeviceMesh(mesh, velbc=velbc)
    assert not dm.flat
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(33)
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = 1e-4 * rng.standard_normal(n)
    q[1][:n] = 0.03 * rng.standard_normal(n)
    q[3][:n] = 5.0
    for f in q:
        mesh.exchange(f)
    run_heve(mesh, bs, q, om, dm, 0.1, phys_fields(mesh, 34), 3, dry=False)


def test_hevi_steps_with_physics_bitwise():
    print("HEVI steps with physics tendencies p=3", flush=True)
    case = cases.scaled_case(cases.config4(), 4, 2, 4, porder=3)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    tp = phys_fields(mesh, 41)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=1, rtol=0.0)
    gtp = [dev(a) for a in device_order(tp)]
    st.set_phys_tend(*gtp)
    st.set_state(q)
    st.run(2)
    g = st.get_state()
    qq = [a.copy() for a in q]
    om.set_phys_tend(tp)
    try:
        om.hevi_update(qq, bs, Dx, Dy, case.dt, 2, scheme=case.scheme, max_it=1, rtol=0.0)
    finally:
        om.set_phys_tend(None)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    st.set_phys_tend()
    st.close()
