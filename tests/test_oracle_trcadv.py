"""The oracle of the tracer advection tendency,
atm_dyn_dgm_trcadvect3d_heve_cal_tend (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231)
with its face flux get_delflux_generalhvc (:554-674), pinned on the CPU by what
the mathematics fixes:

* advection of a global polynomial of degree <= p by a uniform mass flux: on
  every element away from the domain boundaries the tendency is the exact
  -(U dQ/dx + V dQ/dy + W dQ/dz) + RHOQ_tp (a wrong Escale, a transposed
  derivative or a dropped RHOQ_tp fails it);
* the FCT coefficient of the upwind element: with Q = 1 and a mass flux that
  varies across the face so that the LGL-weighted face integral and the plain
  sum of the face values have opposite signs, the tendency integrated over an
  element is -S (R_e - R_upstream) in closed form, S the weighted face flux;
  unweighted face sums (a wrong FaceIntMat) or the downwind coefficient give
  the other neighbour's R;
* conservation: in a triply periodic box with random fields, random FCT
  coefficients, damping coefficients alphDENS_M/_P that are one field seen from
  both sides of each face and random (non-physical) metric terms, the sum of
  w J Gsqrt (QTRC_dt - RHOQ_tp) over the domain vanishes to rounding.
"""
import numpy as np
import pytest

import fedg_oracle as O
from fedg import BoxMesh


def w3(mesh):
    w = mesh.elem.w1d
    n1 = len(w)
    return (w[:, None, None] * w[None, :, None] * w[None, None, :]).transpose(2, 1, 0).ravel()[: n1 ** 3]


def node_ijk(mesh):
    n1 = mesh.elem.n1
    p = np.arange(mesh.Np)
    return p % n1, (p // n1) % n1, p // (n1 * n1)


def elem_integrals(mesh, f):
    """sum_p w_p J_p f_p per element (f over the Np*Ne interior part)."""
    W = np.tile(w3(mesh), mesh.Ne) * mesh.J
    return (W * f[: mesh.Np * mesh.Ne]).reshape(mesh.Ne, mesh.Np).sum(axis=1)


def alph_from_field(mesh, s, c):
    """alphDENS_M/_P = c * s(iM) / c * s(iP): one node field seen from both sides."""
    return c * s[mesh.vmapM - 1], c * s[mesh.vmapP - 1]


@pytest.mark.parametrize("porder", [2, 3])
def test_polynomial_advection_closed_form(porder):
    L = 4.0e3
    mesh = BoxMesh(porder, 4, 4, 4, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, False))
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    x, y, z = mesh.x / L, mesh.y / L, mesh.z / L
    p = porder
    Q = mesh.field()
    Q[:n] = 1.0 + 0.3 * x ** p + 0.2 * y ** (p - 1) * z + 0.1 * x * y * z ** (p - 2)
    dQ = (0.3 * p * x ** (p - 1) + 0.1 * y * z ** (p - 2),
          0.2 * (p - 1) * y ** (p - 2) * z + 0.1 * x * z ** (p - 2),
          0.2 * y ** (p - 1) + 0.1 * (p - 2) * x * y * z ** max(p - 3, 0))
    U, V, W = 12.0, -7.0, 3.0
    MX, MY, MZ = mesh.field(U), mesh.field(V), mesh.field(W)
    rng = np.random.default_rng(11)
    tp = mesh.field()
    tp[:n] = rng.standard_normal(n)
    R = mesh.field(1.0)
    aM = np.full(mesh.NfpTot * mesh.Ne, 0.5)
    aP = aM.copy()
    om.exchange(Q)
    out = om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)
    exact = -(U * dQ[0] + V * dQ[1] + W * dQ[2]) / L + tp[:n]
    # elements that touch no domain boundary (the periodic seams break the polynomial)
    e = np.arange(mesh.Ne)
    ex, ey, ez = e % 4, (e // 4) % 4, e // 16
    inner = (ex > 0) & (ex < 3) & (ey > 0) & (ey < 3) & (ez > 0) & (ez < 3)
    sel = np.repeat(inner, mesh.Np)
    err = np.abs(out[:n] - exact)[sel].max()
    scale = np.abs(exact[sel]).max()
    assert err <= 1e-11 * scale, (err, scale)
    # the tracer's own tendency term is carried through
    out0 = om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, mesh.field())
    assert np.abs(out[:n] - out0[:n] - tp[:n]).max() <= 1e-12 * scale


@pytest.mark.parametrize("porder", [2, 3])
@pytest.mark.parametrize("axis", [0, 1, 2], ids=["x", "y", "z"])
@pytest.mark.parametrize("direction", [1.0, -1.0], ids=["plus", "minus"])
def test_fct_coefficient_of_the_upwind_element(porder, axis, direction):
    L = 3.0e3
    NE = 3
    mesh = BoxMesh(porder, NE, NE, NE, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    n1 = mesh.elem.n1
    w = mesh.elem.w1d
    # h: -1 at the face's edge nodes, 0.6 inside: the LGL-weighted sum is
    # positive while the plain sum is negative
    h = np.full(n1, 0.6)
    h[0] = h[-1] = -1.0
    assert (w * h).sum() > 0 > h.sum()
    i, j, k = node_ijk(mesh)
    tang = (j, i, i)[axis]  # a node index that runs along the faces normal to `axis`
    g = direction * 5.0 * np.tile(h[tang], mesh.Ne)
    M = [mesh.field(), mesh.field(), mesh.field()]
    n = mesh.Np * mesh.Ne
    M[axis][:n] = g
    rng = np.random.default_rng(5 + axis)
    r_e = rng.uniform(0.0, 1.0, mesh.Ne)
    R = mesh.field()
    R[:n] = np.repeat(r_e, mesh.Np)
    Q = mesh.field(1.0)
    for f in (*M, R):
        om.exchange(f)
    a0 = np.zeros(mesh.NfpTot * mesh.Ne)
    out = om.trcadvect_cal_tend(Q, M[0], M[1], M[2], a0, a0, R, mesh.field())
    I = elem_integrals(mesh, out)
    # S: LGL-weighted flux through the element's face on the + side of `axis`
    h_sum = (w * h).sum() * w.sum()
    area = (L / NE) ** 2 / 4.0  # face Jacobian sJ of the reference face [-1,1]^2
    S = direction * 5.0 * h_sum * area
    e = np.arange(mesh.Ne)
    c = np.stack([e % NE, (e // NE) % NE, e // (NE * NE)])
    up, down = c.copy(), c.copy()
    up[axis] = (c[axis] - 1) % NE
    down[axis] = (c[axis] + 1) % NE
    lin = lambda cc: cc[0] + NE * (cc[1] + NE * cc[2])
    if S > 0:
        expect = -S * (r_e - r_e[lin(up)])
    else:
        expect = -S * (r_e[lin(down)] - r_e)
    assert np.abs(I - expect).max() <= 1e-11 * np.abs(S)


@pytest.mark.parametrize("porder", [2, 3])
def test_conservation_random_fields_and_metrics(porder):
    L = 3.0e3
    mesh = BoxMesh(porder, 3, 3, 3, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, True))
    rng = np.random.default_rng(17)
    n = mesh.Np * mesh.Ne
    Gsqrt = mesh.field()
    Gsqrt[:n] = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, n)
    GI3 = np.zeros(mesh.Np * mesh.NeA * 2)
    GI3[:n] = 0.05 * rng.standard_normal(n)
    GI3[mesh.Np * mesh.NeA: mesh.Np * mesh.NeA + n] = 0.05 * rng.standard_normal(n)
    mesh.exchange(Gsqrt)
    mesh.exchange(GI3[: mesh.Np * mesh.NeA])
    mesh.exchange(GI3[mesh.Np * mesh.NeA:])
    mesh.GsqrtH[:] = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, mesh.GsqrtH.size)
    om = O.OracleMesh(mesh, Gsqrt=Gsqrt, GI3=GI3)
    Q, MX, MY, MZ, R, s = (mesh.field() for _ in range(6))
    Q[:n] = 1.0 + 0.2 * rng.standard_normal(n)
    MX[:n], MY[:n], MZ[:n] = (10.0 * rng.standard_normal(n) for _ in range(3))
    R[:n] = rng.uniform(0.0, 1.0, n)
    s[:n] = rng.uniform(0.5, 1.5, n)
    for f in (Q, MX, MY, MZ, R, s):
        om.exchange(f)
    aM, aP = alph_from_field(mesh, s, 20.0)
    tp = mesh.field()
    out = om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)
    I = elem_integrals(mesh, Gsqrt[:n] * out[:n])
    assert np.abs(I).sum() > 0.0
    assert abs(I.sum()) <= 1e-12 * np.abs(I).sum(), (I.sum(), np.abs(I).sum())
    # the element sums themselves are far from zero: only the faces cancel them
    assert np.abs(I).max() > 1e-6 * np.abs(I).sum()
