"""GPU parity of the tracer advection tendency
(atm_dyn_dgm_trcadvect3d_heve_cal_tend, FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231)
against the CPU oracle, through the C ABI.

Bar: bit for bit.  The kernel evaluates the reference's expressions in the
reference's order without fused multiply-add and with IEEE division; the
FaceIntMat product sums each face's nodes from zero in storage order, as the
reference's CSR product does.  The metric terms are random (Gsqrt, G13, G23,
GsqrtH), so the terrain-following paths of the mass flux are exercised; the
orders cover one element per workgroup (p = 7) and many (p = 1, with a
partial last workgroup).
"""
import numpy as np
import pytest

import fedg
from fedg import BoxMesh
import fedg_oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # porder, NeX, NeY, NeZ, periodic
    (1, 5, 3, 4, (True, True, False)),
    (3, 4, 3, 4, (True, True, False)),
    (4, 3, 2, 3, (False, True, True)),
    (7, 2, 2, 2, (True, True, True)),
]


def random_metrics(mesh, rng):
    n, nA = mesh.Np * mesh.Ne, mesh.Np * mesh.NeA
    mesh.Gsqrt[:n] = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, n)
    mesh.GI3[:n] = 0.05 * rng.standard_normal(n)
    mesh.GI3[nA: nA + n] = 0.05 * rng.standard_normal(n)
    mesh.exchange(mesh.Gsqrt)
    mesh.exchange(mesh.GI3[:nA])
    mesh.exchange(mesh.GI3[nA:])
    mesh.GsqrtH[:] = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, mesh.GsqrtH.size)


@pytest.mark.parametrize("case", CASES, ids=[f"p{c[0]}" for c in CASES])
def test_trcadvect_cal_tend_bitwise(case):
    porder, NeX, NeY, NeZ, periodic = case
    print(f"trcadvect p={porder} {NeX}x{NeY}x{NeZ}", flush=True)
    mesh = BoxMesh(porder, NeX, NeY, NeZ, (0.0, 2.0e3 * NeX), (0.0, 1.5e3 * NeY), (0.0, 1.0e3 * NeZ),
                   periodic=periodic)
    rng = np.random.default_rng(100 + porder)
    random_metrics(mesh, rng)
    om = O.OracleMesh(mesh)
    dm = fedg.DeviceMesh(mesh)
    n = mesh.Np * mesh.Ne
    Q, MX, MY, MZ, R, tp = (mesh.field() for _ in range(6))
    Q[:n] = 1.0e-3 * (1.0 + 0.3 * rng.standard_normal(n))
    MX[:n], MY[:n] = 20.0 * rng.standard_normal(n), 15.0 * rng.standard_normal(n)
    MZ[:n] = 3.0 * rng.standard_normal(n)
    R[:n] = rng.uniform(0.0, 1.0, n)
    tp[:n] = 1.0e-6 * rng.standard_normal(n)
    for f in (Q, MX, MY, MZ, R, tp):
        om.exchange(f)
    nf = mesh.NfpTot * mesh.Ne
    aM, aP = rng.uniform(0.0, 40.0, nf), rng.uniform(0.0, 40.0, nf)
    ref = om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)
    d = lambda a: fedg.DevField(a.size, a)
    out = d(np.full(mesh.Np * mesh.NeA, np.nan))
    fedg.trcadvect_cal_tend(dm, out, d(Q), d(MX), d(MY), d(MZ), d(aM), d(aP), d(R), d(tp), d(mesh.J))
    g = out.download()[:n]
    assert np.all(np.isfinite(g))
    assert np.array_equal(g, ref[:n]), f"max abs diff {np.max(np.abs(g - ref[:n]))}"
    dm.close()
