"""The coupled tracer step's thermodynamics on the GPU against the oracle, bit
for bit: fedg_tracer_specific_heat (update_pressure_specific_heat) and
fedg_trc_negative_fixer (fix_negative_val / atm_phy_mp_dgm_common_negative_fixer)
with three water species, negative cloud water and ice, at p = 3 and p = 7 on a
terrain-following mesh.  The oracle is pinned by tests/test_oracle_negfix.py."""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases
from test_oracle_negfix import THERMO

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def dev(a):
    return fedg.DevField(a.size, a)


@pytest.mark.parametrize("p,shape", [(3, (4, 3, 3)), (7, (2, 2, 2))], ids=["p3", "p7"])
def test_specific_heat_and_negative_fixer_bitwise(p, shape):
    print(f"negative fixer p={p} {shape}", flush=True)
    NeX, NeY, NeZ = shape
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, 2e3 * NeX), (0.0, 2e3 * NeY), (0.0, 1.5e3 * NeZ),
                        periodic=(True, True, False))
    mesh.set_topography(lambda x, y: 300.0 * np.exp(-((x - 2e3) / 1.5e3) ** 2))
    bs = cases.base_state(mesh, "constBVFreq")
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(100 + p)
    n = mesh.Np * mesh.Ne
    Q = [mesh.field() for _ in range(3)]
    Q[0][:n] = 0.01 * rng.uniform(0.5, 1.0, n)
    Q[1][:n] = 1e-4 * rng.uniform(-0.5, 1.0, n)
    Q[2][:n] = 1e-5 * rng.standard_normal(n)
    DD = mesh.field()
    DD[:n] = 1e-3 * rng.standard_normal(n)
    DR = mesh.field()
    DR[:n] = 0.1 * rng.standard_normal(n)
    # specific heats of the mixture, then the pressure
    R, CV, CP = mesh.field(), mesh.field(), mesh.field()
    om.tracer_specific_heat(Q, THERMO, R, CV, CP)
    dm = fedg.DeviceMesh(mesh)
    gQ = [dev(a) for a in Q]
    gR, gCV, gCP = dev(mesh.field()), dev(mesh.field()), dev(mesh.field())
    fedg.tracer_specific_heat(dm, gQ, THERMO, gR, gCV, gCP)
    for name, a, b in (("Rtot", gR, R), ("CVtot", gCV, CV), ("CPtot", gCP, CP)):
        assert np.array_equal(a.download()[:n], b[:n]), name
    import copy
    b2 = copy.copy(bs)
    b2.Rtot, b2.CVtot, b2.CPtot = R, CV, CP
    P, _ = om.drhot2pres(DR, b2)
    # the fixer on both sides
    gDD, gP, gDR = dev(DD), dev(P), dev(DR)
    gJ, gz = dev(mesh.J.copy()), dev(mesh.height.copy())
    gDh, gPh = dev(bs.DENS_hyd), dev(bs.PRES_hyd)
    fedg.trc_negative_fixer(dm, gQ, gDD, gP, gCV, gCP, gR, gDh, gPh, gJ, gz, THERMO, QLA=1, QIA=1, DRHOT=gDR)
    om.trc_negative_fixer(Q, DD, P, CV, CP, R, bs.DENS_hyd, bs.PRES_hyd, mesh.J, mesh.height, THERMO, DRHOT=DR)
    for name, a, b in (("QV", gQ[0], Q[0]), ("QC", gQ[1], Q[1]), ("QI", gQ[2], Q[2]), ("DDENS", gDD, DD),
                       ("PRES", gP, P), ("Rtot", gR, R), ("CVtot", gCV, CV), ("CPtot", gCP, CP), ("DRHOT", gDR, DR)):
        g = a.download()[:n]
        assert np.array_equal(g, b[:n]), f"{name}: max abs diff {np.max(np.abs(g - b[:n]))}"
    assert np.all(Q[1][:n] >= 0.0) and np.all(Q[2][:n] >= 0.0)
