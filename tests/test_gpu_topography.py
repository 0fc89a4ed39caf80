"""HEVE and HEVI steps on a terrain-following mesh from the builder
(fedg_box_topography_metrics over an Agnesi ridge) on the GPU against the
oracle, bit for bit: a warm bubble over the ridge (HEVE, classic RK4) and the
hydrostatic rest state over it (HEVI, IMEX_ARK324, Newton to a tolerance).
The metrics are pinned by tests/test_topography.py."""
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


def ridge_mesh(p, NeX=6, NeY=2, NeZ=4, Lx=12e3, Ly=3e3, Lz=8e3):
    m = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, Lx), (0.0, Ly), (0.0, Lz), periodic=(True, True, False))
    m.set_topography(lambda x, y: 400.0 / (1.0 + ((x - 0.5 * Lx) / 2e3) ** 2), zTop=Lz)
    return m


@pytest.mark.parametrize("p", [3, 7])
def test_heve_steps_over_a_ridge_bitwise(p):
    print(f"HEVE over a ridge p={p}", flush=True)
    m = ridge_mesh(p, NeX=6 if p == 3 else 3, NeY=2, NeZ=4 if p == 3 else 3)
    bs = cases.base_state(m, "constBVFreq")
    q = cases.warm_bubble_state(m, bs, [(2.0, 2.0e3, (4.0e3, 1.5e3, 2.5e3))])
    om = O.OracleMesh(m)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(m)
    assert not dm.flat
    dt = 0.05 if p == 3 else 0.01
    st = fedg.HeveStepper(dm, bs, Dx, Dy, dt)
    st.set_state(q)
    st.run(10)
    g = st.get_state()
    qq = [a.copy() for a in q]
    om.heve_update(qq, bs, Dx, Dy, dt, 10)
    n = m.Np * m.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    st.close()


def test_hevi_rest_over_a_ridge_bitwise():
    print("HEVI rest state over a ridge p=3", flush=True)
    m = ridge_mesh(3)
    bs = cases.base_state(m, "constBVFreq")
    q = [m.field() for _ in range(5)]
    om = O.OracleMesh(m)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(m)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, 0.2, scheme="IMEX_ARK324", max_it=6, rtol=1e-10)
    st.set_state(q)
    st.run(5)
    g = st.get_state()
    qq = [a.copy() for a in q]
    om.hevi_update(qq, bs, Dx, Dy, 0.2, 5, scheme="IMEX_ARK324", max_it=6, rtol=1e-10)
    n = m.Np * m.Ne
    for name, a, b in zip(fedg.VARS, g, qq):
        assert np.array_equal(a, b[:n]), f"{name}: max abs diff {np.max(np.abs(a - b[:n]))}"
    assert np.max(np.abs(g[2])) < 1e-2 and np.max(np.abs(g[3])) < 1e-2
    st.close()
