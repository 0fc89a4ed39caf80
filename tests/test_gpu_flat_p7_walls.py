"""The p = 7 flat-box stage kernel (heve7_lvl_kernel: base state from the
per-level tables) on the inputs the periodic test boxes leave at zero, against
the CPU oracle, bit for bit, through the C ABI: closed lateral faces (SLIP on
x, NOSLIP at the bottom, SLIP at the top, periodic in y; the kernel's
neighbour logic and momentum reflection on every face kind), a random
Coriolis parameter (the kernel's per-column load), a number of elements that
is not a multiple of 8 (the XCD remap's fallback), and both explicit schemes
(the RK4 stage modes first / middle / last and the single-stage scheme).
"""
import dataclasses

import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases, constants
from fedg.mesh import BND_NOSLIP, BND_NONE, BND_SLIP

pytestmark = pytest.mark.gpu

WALLS = (BND_NONE, BND_SLIP, BND_NONE, BND_SLIP, BND_NOSLIP, BND_SLIP)  # y-, x+, y+, x-, z-, z+


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


@pytest.fixture(scope="module")
def box():
    NeX, NeY, NeZ = 5, 3, 4  # 60 elements: not a multiple of 8
    mesh = fedg.BoxMesh(7, NeX, NeY, NeZ, (0.0, 2.5e3 * NeX), (0.0, 2.5e3 * NeY), (0.0, 6.25e2 * NeZ),
                        periodic=(False, True, False))
    bs = cases.base_state(mesh, "constBVFreq")
    rng = np.random.default_rng(5)
    n = mesh.Np * mesh.Ne
    bs.CORIOLIS[:] = 1.0e-4 * (1.0 + 0.3 * rng.standard_normal(bs.CORIOLIS.size))
    velbc = mesh.velbc(WALLS)
    om = O.OracleMesh(mesh, velbc=velbc)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    dm = fedg.DeviceMesh(mesh, velbc=velbc)
    q = cases.warm_bubble_state(mesh, bs, [(2.0, 2.0e3, (6.0e3, 3.0e3, 1.2e3))])
    for v, amp in ((2, 0.05), (3, 0.3), (4, 0.2)):
        q[v][:n] += amp * rng.standard_normal(n)
    for f in q:
        om.exchange(f)
    om.apply_bc(q)
    return dict(mesh=mesh, bs=bs, om=om, Dx=Dx, Dy=Dy, dm=dm, q=q)


@pytest.mark.parametrize("scheme,nsteps", [("ERK_4s4o", 3), ("ERK_1s1o", 2)], ids=["rk4", "euler"])
def test_flat_p7_walls_coriolis_match_oracle(box, scheme, nsteps):
    print(f"flat p=7 stage with walls and Coriolis: {scheme}", flush=True)
    b = box
    mesh, bs, om = b["mesh"], b["bs"], b["om"]
    dt = 0.02
    st = fedg.HeveStepper(b["dm"], bs, b["Dx"], b["Dy"], dt, scheme=scheme)
    assert st.flat
    q = [a.copy() for a in b["q"]]
    st.set_state(q)
    st.run(nsteps)
    g = st.get_state()
    _, Dg = st.get_pressure()
    c = constants.SCALE  # the flat stepper is dry
    nA = mesh.Np * mesh.NeA
    bso = dataclasses.replace(bs, Rtot=np.full(nA, c.Rdry), CVtot=np.full(nA, c.CVdry), CPtot=np.full(nA, c.CPdry))
    om.heve_update(q, bso, b["Dx"], b["Dy"], dt, nsteps, scheme=scheme)
    n = mesh.Np * mesh.Ne
    for name, a, r in zip(fedg.VARS, g, q):
        assert np.array_equal(a, r[:n]), f"{name}: max abs diff {np.max(np.abs(a - r[:n]))}"
    _, Do = om.drhot2pres(q[1], bso)
    assert np.array_equal(Dg, Do[:n])
    st.close()
