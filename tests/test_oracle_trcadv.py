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


# --- FCT coefficient, TMAR, mass-flux save (..._trcadvect3d_heve.F90:232-536) ---

def upwind_alph(mesh, MX, MY, MZ):
    """alphDENS_M/_P = |normal mass flux| of each side (flat metrics): the face flux becomes pure upwind."""
    nf = mesh.NfpTot * mesh.Ne
    nx, ny, nz = (mesh.normal_fn[d * nf:(d + 1) * nf] for d in range(3))
    iM, iP = mesh.vmapM - 1, mesh.vmapP - 1
    fM = MX[iM] * nx + MY[iM] * ny + MZ[iM] * nz
    fP = MX[iP] * nx + MY[iP] * ny + MZ[iP] * nz
    return np.abs(fM), np.abs(fP)


@pytest.mark.parametrize("porder", [2, 3])
@pytest.mark.parametrize("c", [1.0, 0.4])
def test_fct_coefficient_drains_exactly_the_available_mass(porder, c):
    """Uniform flow in x over alternating empty / filled elements with upwind
    damping: a filled element has no inflow, so with the limiter its mass
    after one step of length dt is exactly the available mass minus R times
    its outflow, which the FCT coefficient sets to zero when the outflow
    exceeds it.  A wrong density of stage s-1, dt or RHOQ_tp in the available
    mass leaves a nonzero remainder; without the limiter the mass goes
    negative."""
    L = 4.0e3
    NE = 4
    mesh = BoxMesh(porder, NE, 2, 2, (0.0, L), (0.0, L / 2), (0.0, L / 2), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(23 + porder)
    Dh, DD, DD0 = mesh.field(1.1), mesh.field(), mesh.field()
    DD[:n] = 0.02 * rng.standard_normal(n)
    DD0[:n] = 0.02 * rng.standard_normal(n)
    dens_ssm1 = Dh + (1.0 - c) * DD0 + c * DD
    dens = Dh + DD
    U = 15.0
    MX, MY, MZ = dens * U, mesh.field(), mesh.field()
    e = np.arange(mesh.Ne)
    filled = (e % NE) % 2 == 0
    Q = mesh.field()
    Q[:n] = np.repeat(filled, mesh.Np) * rng.uniform(1.0e-3, 2.0e-3, n)
    tp = mesh.field()
    tp[:n] = np.repeat(filled, mesh.Np) * rng.uniform(0.0, 1.0e-6, n)
    for f in (Dh, DD, DD0, MX, Q, tp):
        om.exchange(f)
    aM, aP = upwind_alph(mesh, MX, MY, MZ)
    dt = 200.0  # long enough that the outflow over dt exceeds every filled element's mass
    R = om.trcadvect_calc_fct_coef(Q, MX, MY, MZ, tp, aM, aP, Dh, DD, DD0, c, dt)
    assert np.all((R[:n] >= 0.0) & (R[:n] <= 1.0))
    om.exchange(R)
    Rel = R[:n].reshape(mesh.Ne, mesh.Np)
    assert np.all(Rel == Rel[:, :1]), "one coefficient per element"
    # empty elements have no mass and no outflow: 0 / (0 + 1e-10)
    assert np.all(Rel[filled, 0] < 1.0) and np.all(Rel[~filled, 0] == 0.0)
    mass = elem_integrals(mesh, dens_ssm1[:n] * Q[:n])
    out = om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)
    new = mass + dt * elem_integrals(mesh, out[:n])
    scale = mass[filled].max()
    assert np.abs(new[filled]).max() <= 1e-9 * scale, new[filled]
    # the empty elements receive what the filled ones lose
    assert abs(new.sum() - (mass.sum() + dt * elem_integrals(mesh, tp[:n]).sum())) <= 1e-12 * scale * mesh.Ne
    # without the limiter the filled elements overdraw
    R1 = om.trcadvect_calc_fct_coef(Q, MX, MY, MZ, tp, aM, aP, Dh, DD, DD0, c, dt, disable_limiter=True)
    assert np.all(R1[:n] == 1.0)
    om.exchange(R1)
    out1 = om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R1, tp)
    new1 = mass + dt * elem_integrals(mesh, out1[:n])
    assert new1[filled].max() < -0.1 * scale


@pytest.mark.parametrize("porder", [2, 3])
def test_fct_full_drain_moves_each_element_mass_one_element_downstream(porder):
    """Every element filled, uniform flow in x, upwind damping and a dt long
    enough that every limiter is active: each element sends out exactly its
    available mass and receives exactly its upstream neighbour's, so after one
    step it holds the upstream element's available mass.  The net outward flux
    counts only the outflow faces (a net that also counted the inflow face
    would hold mass back)."""
    L = 4.0e3
    NE = 4
    mesh = BoxMesh(porder, NE, 2, 2, (0.0, L), (0.0, L / 2), (0.0, L / 2), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(29 + porder)
    Dh, DD = mesh.field(1.0), mesh.field()
    DD[:n] = 0.02 * rng.standard_normal(n)
    dens = Dh + DD
    MX, MY, MZ = dens * 10.0, mesh.field(), mesh.field()
    Q, tp = mesh.field(), mesh.field()
    Q[:n] = np.repeat(rng.uniform(0.5, 2.0, mesh.Ne), mesh.Np) * rng.uniform(0.9, 1.0, n) * 1e-3
    tp[:n] = rng.uniform(0.0, 1e-7, n)
    for f in (Dh, DD, MX, Q, tp):
        om.exchange(f)
    aM, aP = upwind_alph(mesh, MX, MY, MZ)
    dt = 400.0
    R = om.trcadvect_calc_fct_coef(Q, MX, MY, MZ, tp, aM, aP, Dh, DD, DD, 1.0, dt)
    assert np.all(R[:n] < 1.0)
    om.exchange(R)
    avail = elem_integrals(mesh, dens[:n] * Q[:n]) + dt * elem_integrals(mesh, tp[:n])
    new = elem_integrals(mesh, dens[:n] * Q[:n]) + dt * elem_integrals(
        mesh, om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)[:n])
    e = np.arange(mesh.Ne)
    up = (e % NE - 1) % NE + NE * (e // NE)
    np.testing.assert_allclose(new, avail[up], rtol=1e-9, atol=0)


@pytest.mark.parametrize("porder", [2, 3])
def test_fct_keeps_every_element_mass_nonnegative(porder):
    """Random positive tracer, random flow, any damping: the FCT coefficients
    bound each element's outflow by its available mass, so one step of dt
    leaves no element mass negative (inflow only adds)."""
    L = 3.0e3
    mesh = BoxMesh(porder, 3, 3, 3, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(31)
    Dh, DD = mesh.field(1.0), mesh.field()
    DD[:n] = 0.05 * rng.standard_normal(n)
    MX, MY, MZ, Q = (mesh.field() for _ in range(4))
    for M in (MX, MY, MZ):
        M[:n] = 20.0 * rng.standard_normal(n)
    Q[:n] = np.repeat(rng.uniform(0.0, 1.0, mesh.Ne) ** 3, mesh.Np) * rng.uniform(0.5, 1.0, n) * 1e-3
    tp = mesh.field()
    for f in (Dh, DD, MX, MY, MZ, Q):
        om.exchange(f)
    aM, aP = alph_from_field(mesh, Dh + DD, 30.0)
    dt = 30.0
    R = om.trcadvect_calc_fct_coef(Q, MX, MY, MZ, tp, aM, aP, Dh, DD, DD, 1.0, dt)
    om.exchange(R)
    dens = Dh + DD
    mass = elem_integrals(mesh, dens[:n] * Q[:n])
    new = mass + dt * elem_integrals(mesh, om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R, tp)[:n])
    assert (R[:n] < 1.0).any()
    assert new.min() >= -1e-12 * mass.max(), new.min()
    R1 = mesh.field(1.0)
    new1 = mass + dt * elem_integrals(mesh, om.trcadvect_cal_tend(Q, MX, MY, MZ, aM, aP, R1, tp)[:n])
    assert new1.min() < -1e-3 * mass.max()


@pytest.mark.parametrize("porder", [2, 3])
def test_tmar_cuts_negatives_and_keeps_element_mass(porder):
    L = 3.0e3
    mesh = BoxMesh(porder, 3, 2, 2, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, False))
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(41)
    Dh, DD = mesh.field(1.2), mesh.field()
    DD[:n] = 0.05 * rng.standard_normal(n)
    Q = mesh.field()
    Q[:n] = 1e-3 + 4e-4 * rng.standard_normal(n)
    Qel = Q[:n].reshape(mesh.Ne, mesh.Np)
    Qel[0] = np.abs(Qel[0])  # one element with no negative value
    Q0 = Q.copy()
    om.trcadvect_tmar(Q, Dh, DD)
    assert (Q0[:n] < 0).any() and np.all(Q[:n] >= 0.0)
    dens = (Dh + DD)[:n]
    m0, m1 = elem_integrals(mesh, dens * Q0[:n]), elem_integrals(mesh, dens * Q[:n])
    assert np.all(m0 > 0)
    assert np.abs(m1 - m0).max() <= 1e-14 * m0.max()
    assert np.array_equal(Q[: mesh.Np], Q0[: mesh.Np]), "an element without negatives is left as it is"
    # the shape of the cut field is kept: one factor per element
    pos = Q0[:n] > 0
    ratio = (Q[:n][pos] / Q0[:n][pos])
    el = np.repeat(np.arange(mesh.Ne), mesh.Np)[pos]
    for k in range(mesh.Ne):
        r = ratio[el == k]
        assert np.ptp(r) <= 1e-14 * r.max()


def mirror_face_node(mesh):
    """For each face node, the index of the same node seen from the neighbouring element."""
    key = {(int(a), int(b)): g for g, (a, b) in enumerate(zip(mesh.vmapM, mesh.vmapP))}
    n = mesh.Np * mesh.Ne
    src = np.concatenate([np.arange(n), mesh.halo_src - 1])  # halo slots resolved to interior nodes
    inv = {}
    for g, (a, b) in enumerate(zip(mesh.vmapM - 1, mesh.vmapP - 1)):
        inv[(int(a), int(src[b]))] = g
    out = np.empty(mesh.NfpTot * mesh.Ne, dtype=np.int64)
    for g, (a, b) in enumerate(zip(mesh.vmapM - 1, mesh.vmapP - 1)):
        out[g] = inv[(int(src[b]), int(a))]
    del key
    return out


@pytest.mark.parametrize("is_hevi", [False, True], ids=["heve", "hevi"])
def test_save_massflux_sound_speed_and_face_symmetry(is_hevi):
    L = 3.0e3
    mesh = BoxMesh(2, 3, 3, 3, (0.0, L), (0.0, L), (0.0, L), periodic=(True, True, True))
    om = O.OracleMesh(mesh)
    n, nf = mesh.Np * mesh.Ne, mesh.NfpTot * mesh.Ne
    from fedg import constants
    gam = constants.SCALE.CPdry / constants.SCALE.CVdry
    # uniform state, uniform flow U in x: alpha = c + |U n_x|, c = sqrt(gamma p / rho)
    rho, p, U = 1.1, 9.0e4, 12.0
    Dh, Ph, DD, DP = mesh.field(rho), mesh.field(p), mesh.field(), mesh.field()
    MX, MY, MZ = mesh.field(rho * U), mesh.field(), mesh.field()
    MF = [mesh.field(7.0) for _ in range(3)]
    al = [np.full(nf, 3.0), np.full(nf, 3.0)]
    twh, twv = 0.25, 0.5
    om.trcadvect_save_massflux(MF, al, DD, MX, MY, MZ, DP, Dh, Ph, 1, twh, twv, is_hevi)
    assert np.array_equal(MF[0][:n], twh * MX[:n]) and np.all(MF[1][:n] == 0.0)
    nx, ny, nz = (mesh.normal_fn[d * nf:(d + 1) * nf] for d in range(3))
    c = np.sqrt(gam * p / rho)
    horiz = np.abs(nz) == 0.0
    alpha = np.where(horiz, c + U * np.abs(nx), 0.0 if is_hevi else c)
    expect = np.where(horiz, twh, twv) * alpha * rho
    np.testing.assert_allclose(al[0], expect, rtol=1e-14, atol=0)
    np.testing.assert_allclose(al[1], expect, rtol=1e-14, atol=0)
    # a second stage accumulates
    om.trcadvect_save_massflux(MF, al, DD, MX, MY, MZ, DP, Dh, Ph, 2, twh, twv, is_hevi)
    np.testing.assert_allclose(MF[0][:n], 2 * twh * MX[:n], rtol=1e-15)
    np.testing.assert_allclose(al[0], 2 * expect, rtol=1e-14, atol=0)
    # random state: both sides of a face hold the same coefficients, bit for bit
    rng = np.random.default_rng(3)
    for f, s in ((DD, 0.05), (DP, 500.0), (MX, 10.0), (MY, 10.0), (MZ, 3.0)):
        f[:n] = s * rng.standard_normal(n)
        om.exchange(f)
    om.trcadvect_save_massflux(MF, al, DD, MX, MY, MZ, DP, Dh, Ph, 1, twh, twv, is_hevi)
    mir = mirror_face_node(mesh)
    assert np.array_equal(al[0], al[1][mir])
