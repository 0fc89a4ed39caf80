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


def _mesh_and_ctx(case, metrics, rng):
    porder, NeX, NeY, NeZ, periodic = case
    mesh = BoxMesh(porder, NeX, NeY, NeZ, (0.0, 2.0e3 * NeX), (0.0, 1.5e3 * NeY), (0.0, 1.0e3 * NeZ),
                   periodic=periodic)
    if metrics == "random":
        random_metrics(mesh, rng)
    dm = fedg.DeviceMesh(mesh)
    assert dm.flat == (metrics == "flat")
    return mesh, O.OracleMesh(mesh), dm


# "flat": the box's own metrics, which the context recognises and the
# kernels then take as constants; "random": every metric array read
@pytest.mark.parametrize("metrics", ["random", "flat"])
@pytest.mark.parametrize("case", CASES, ids=[f"p{c[0]}" for c in CASES])
def test_trcadvect_cal_tend_bitwise(case, metrics):
    porder, NeX, NeY, NeZ, periodic = case
    print(f"trcadvect p={porder} {NeX}x{NeY}x{NeZ} {metrics}", flush=True)
    rng = np.random.default_rng(100 + porder)
    mesh, om, dm = _mesh_and_ctx(case, metrics, rng)
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


def _tracer_state(mesh, rng):
    n = mesh.Np * mesh.Ne
    f = {k: mesh.field() for k in ("Q", "MX", "MY", "MZ", "tp", "Dh", "DD", "DD0", "DP", "Ph")}
    f["Q"][:n] = 1.0e-3 * (1.0 + 0.6 * rng.standard_normal(n))  # some negative values for TMAR
    f["MX"][:n], f["MY"][:n] = 20.0 * rng.standard_normal(n), 15.0 * rng.standard_normal(n)
    f["MZ"][:n] = 3.0 * rng.standard_normal(n)
    f["tp"][:n] = 1.0e-6 * rng.standard_normal(n)
    f["Dh"][:n] = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, n)
    f["DD"][:n], f["DD0"][:n] = 0.02 * rng.standard_normal(n), 0.02 * rng.standard_normal(n)
    f["Ph"][:n] = 9.0e4 + 1.0e3 * rng.standard_normal(n)
    f["DP"][:n] = 50.0 * rng.standard_normal(n)
    for a in f.values():
        mesh.exchange(a)
    return f


@pytest.mark.parametrize("metrics", ["random", "flat"])
@pytest.mark.parametrize("case", CASES, ids=[f"p{c[0]}" for c in CASES])
def test_trcadvect_fct_tmar_massflux_bitwise(case, metrics):
    porder, NeX, NeY, NeZ, periodic = case
    print(f"trcadvect fct/tmar/massflux p={porder} {NeX}x{NeY}x{NeZ} {metrics}", flush=True)
    rng = np.random.default_rng(200 + porder)
    mesh, om, dm = _mesh_and_ctx(case, metrics, rng)
    n, nA, nf = mesh.Np * mesh.Ne, mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne
    f = _tracer_state(mesh, rng)
    d = lambda a: fedg.DevField(a.size, a)
    g = {k: d(v) for k, v in f.items()}
    J = d(mesh.J)
    # save_massflux: stage 1 resets, stage 2 accumulates; HEVE and HEVI forms
    for is_hevi in (False, True):
        MF = [mesh.field(5.0) for _ in range(3)]
        al = [np.full(nf, 2.0), np.full(nf, 2.0)]
        gMF, gal = [d(a) for a in MF], [d(a) for a in al]
        for stage, tw in ((1, (0.25, 0.5)), (2, (0.375, 0.125))):
            om.trcadvect_save_massflux(MF, al, f["DD"], f["MX"], f["MY"], f["MZ"], f["DP"], f["Dh"], f["Ph"],
                                       stage, tw[0], tw[1], is_hevi)
            fedg.trcadvect_save_massflux(dm, *gMF, *gal, g["DD"], g["MX"], g["MY"], g["MZ"], g["DP"], g["Dh"],
                                         g["Ph"], None, None, None, stage, tw[0], tw[1], is_hevi)
            for a, b in zip(gMF, MF):
                assert np.array_equal(a.download()[:n], b[:n]), f"MFLX stage {stage} hevi={is_hevi}"
            for a, b in zip(gal, al):
                assert np.array_equal(a.download(), b), f"alph_dens stage {stage} hevi={is_hevi}"
    aM, aP = al
    gaM, gaP = d(aM), d(aP)
    # FCT coefficient, limiter on and off
    for c, dt in ((1.0, 30.0), (0.4, 5.0)):
        ref = om.trcadvect_calc_fct_coef(f["Q"], f["MX"], f["MY"], f["MZ"], f["tp"], aM, aP, f["Dh"], f["DD"],
                                         f["DD0"], c, dt)
        out = d(np.full(nA, np.nan))
        fedg.trcadvect_calc_fct_coef(dm, out, g["Q"], g["MX"], g["MY"], g["MZ"], g["tp"], gaM, gaP, g["Dh"],
                                     g["DD"], g["DD0"], c, dt, J)
        r = out.download()[:n]
        assert np.array_equal(r, ref[:n]), f"fct c={c}: max abs diff {np.max(np.abs(r - ref[:n]))}"
        assert 0.0 <= r.min() and r.max() <= 1.0
    out = d(np.full(nA, np.nan))
    fedg.trcadvect_calc_fct_coef(dm, out, None, None, None, None, None, None, None, None, None, None, 1.0, 1.0, None,
                                 True)
    assert np.all(out.download()[:n] == 1.0)
    # TMAR in place
    Q = f["Q"].copy()
    om.trcadvect_tmar(Q, f["Dh"], f["DD"])
    fedg.trcadvect_tmar(dm, g["Q"], g["Dh"], g["DD"], J)
    assert np.array_equal(g["Q"].download()[:n], Q[:n])
    dm.close()


@pytest.mark.parametrize("case", CASES, ids=[f"p{c[0]}" for c in CASES])
def test_tracer_modalfilter_bitwise(case):
    print(f"tracer modal filter p={case[0]}", flush=True)
    rng = np.random.default_rng(400 + case[0])
    mesh, om, dm = _mesh_and_ctx(case, "random", rng)
    n = mesh.Np * mesh.Ne
    Dh, DD, Q = mesh.field(1.1), mesh.field(), mesh.field()
    DD[:n] = 0.05 * rng.standard_normal(n)
    Q[:n] = 1e-3 * rng.uniform(0.0, 1.0, n)
    MFh = fedg.modalfilter_matrix(case[0])
    MFv = fedg.modalfilter_matrix(case[0], etac=0.4, alpha=10.0, order=8)
    ref = Q.copy()
    om.tracer_modalfilter_apply(ref, Dh, DD, MFh, MFv)
    gQ = fedg.DevField(Q.size, Q)
    fedg.tracer.tracer_modalfilter_apply(dm, gQ, fedg.DevField(Dh.size, Dh), fedg.DevField(DD.size, DD), MFh, MFv)
    g = gQ.download()[:n]
    dm.close()
    assert np.array_equal(g, ref[:n]), f"max abs diff {np.max(np.abs(g - ref[:n]))}"


@pytest.mark.parametrize("filt", [False, True], ids=["", "modalfilter"])
@pytest.mark.parametrize("limiter", [True, False], ids=["limiter", "no-limiter"])
@pytest.mark.parametrize("scheme", ["ERK_SSP_3s3o", "ERK_SSP_10s4o_2N"])
@pytest.mark.parametrize("metrics", ["random", "flat"])
def test_tracer_advection_steps_bitwise(metrics, scheme, limiter, filt):
    """Three steps of the tracer-only advection driver (FCT coefficient,
    tendency, density-weighted RK update and TMAR per stage, on the device)
    against the oracle's loop, bit for bit; with a step profile the limiter and
    TMAR are active."""
    case = (3, 4, 3, 2, (True, True, True))
    print(f"tracer advection {metrics} {scheme} limiter={limiter} filter={filt}", flush=True)
    rng = np.random.default_rng(300)
    mesh, om, dm = _mesh_and_ctx(case, metrics, rng)
    from fedg.tracer import TracerAdvection
    n = mesh.Np * mesh.Ne
    Dh, DD = mesh.field(1.1), mesh.field()
    DD[:n] = 0.03 * rng.standard_normal(n)
    M = [(Dh + DD) * u for u in (14.0, -6.0, 1.5)]
    for a in M:
        a[:n] += 2.0 * rng.standard_normal(n)
    Q = mesh.field()
    Q[:n] = np.where(mesh.x < 4.0e3, 1.0e-3, 0.0) * rng.uniform(0.8, 1.0, n)
    dt = 6.0
    mf = (fedg.modalfilter_matrix(3), fedg.modalfilter_matrix(3)) if filt else None
    ref = om.tracer_advection(Q.copy(), M[0], M[1], M[2], DD, Dh, dt, 3, scheme=scheme, disable_limiter=not limiter,
                              modal_filter=mf)
    st = TracerAdvection(dm, M[0], M[1], M[2], DD, Dh, mesh.J, dt, scheme=scheme, disable_limiter=not limiter,
                         modal_filter=mf)
    gQ = fedg.DevField(Q.size, Q)
    st.step(gQ, 3)
    g = gQ.download()[:n]
    st.close()
    dm.close()
    assert np.array_equal(g, ref[:n]), f"max abs diff {np.max(np.abs(g - ref[:n]))}"
    if limiter:
        assert g.min() >= 0.0
