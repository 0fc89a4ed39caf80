"""The HEVE and HEVI steppers at the polynomial orders the other GPU tests do
not run (p = 1, 2, 4, 5, 6), against the oracle bit for bit.

HEVE: 10 classic RK4 steps of the warm bubble on the flat-box kernel
(MODE_FUSED_FLAT: neighbours by arithmetic, several elements per workgroup
below p = 5, padding slots in the last workgroup) and on the general kernel
(vmapM / vmapP loaded).  The p = 1 box is 5 x 3 x 4 (60 elements, 32 per
workgroup, four padding slots), p = 2 is 7 x 2 x 5 (70 elements, 9 per
workgroup, eight workgroups permuted across the XCDs): tests/
test_flat_index_maps.py checks the indices these launches form on the CPU.
HEVI: 3 IMEX_ARK324 steps of the inertia-gravity wave with 3, 5 and 6 nodes
per element line (p = 2, 4, 5; the line solve's segments of 16 and 32 lanes)."""
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


HEVE_CASES = [(1, (5, 3, 4)), (2, (7, 2, 5)), (4, (4, 3, 4)), (5, (3, 3, 3)), (6, (3, 2, 3))]


@pytest.mark.parametrize("box_shape", [True, False], ids=["flat-kernel", "general-kernel"])
@pytest.mark.parametrize("p,shape", HEVE_CASES, ids=[f"p{p}" for p, _ in HEVE_CASES])
def test_heve_steps_match_oracle(p, shape, box_shape):
    print(f"HEVE p={p} {shape} box_shape={box_shape}", flush=True)
    base = cases.config1() if p <= 3 else cases.config2()
    case = cases.scaled_case(base, *shape, porder=p)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh, box_shape=box_shape)
    st = fedg.HeveStepper(dm, bs, Dx, Dy, case.dt)
    assert st.flat == box_shape
    nsteps = 10
    st.set_state(q)
    st.run(nsteps)
    g = st.get_state()
    st.close()
    qq = [a.copy() for a in q]
    om.heve_update(qq, bs, Dx, Dy, case.dt, nsteps)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.all(np.isfinite(a)), name
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    # the run moved the state: the bubble rises
    assert np.max(np.abs(g[2])) > 0.0


@pytest.mark.parametrize("p", [2, 4, 5])
def test_hevi_steps_match_oracle(p):
    print(f"HEVI p={p} 4x2x4", flush=True)
    case = cases.scaled_case(cases.config4(), 4, 2, 4, porder=p)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme)
    st.set_state(q)
    st.run(3)
    g = st.get_state()
    st.close()
    qq = [a.copy() for a in q]
    om.hevi_update(qq, bs, Dx, Dy, case.dt, 3, scheme=case.scheme)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
