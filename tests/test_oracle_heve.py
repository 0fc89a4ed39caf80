"""Pins of the CPU oracle's HEVE tendency, numerical flux, pressure, halo and BC.

Every check compares the oracle with something the mathematics fixes, not with
a retyped copy of its own formula:
  - Div / pressure gradient / buoyancy / Coriolis: closed-form tendencies of
    polynomial fields of degree <= p (the DG derivative is exact and continuous
    fields have no face jumps), the semantics of FElib/test/FE/
    element_operation_hexahedral/test_element_operation_hexahedral.f90:115-144
    (Div with Escale) applied through the full tendency;
  - numerical flux: the textbook Rusanov flux of the Euler equations written
    from primitive variables (u = MOM/rho, c^2 = gamma P/rho), consistency
    (identical states on both sides give exactly zero) and the two-sided
    identities of the jump;
  - conservation: the exact-mass-matrix integral of DENS_dt over a periodic box
    vanishes;
  - free stream and hydrostatic rest;
  - pressure: P = P00 (R rhotheta / P00)^(cp/cv);
  - halo exchange: exterior values equal the field at the physically
    coincident node (periodic wrap by the domain length); SLIP: exterior normal
    momentum is minus the interior one.
"""
import dataclasses

import numpy as np
import pytest

import fedg_oracle as O
from fedg import cases, constants
from fedg.mesh import BoxMesh

C0 = constants.SCALE
LX, LY, LZ = 4000.0, 3000.0, 2000.0


def mesh_of(p=3, n=(4, 4, 4), periodic=(True, True, False)):
    return BoxMesh(p, n[0], n[1], n[2], (0.0, LX), (0.0, LY), (0.0, LZ), periodic=periodic)


def uniform_base(mesh, rho=1.2, theta=300.0, c=C0):
    """Horizontally and vertically uniform base state (consistent with GRAV = 0)."""
    rhot = rho * theta
    P = c.PRES00 * (c.Rdry * rhot / c.PRES00) ** (c.CPdry / c.CVdry)
    bs = cases.BaseState(DENS_hyd=mesh.field(rho), PRES_hyd=mesh.field(P), THERM_hyd=mesh.field(rhot),
                         PRES_hyd_ref=mesh.field(0.0), Rtot=mesh.field(c.Rdry), CVtot=mesh.field(c.CVdry),
                         CPtot=mesh.field(c.CPdry), CORIOLIS=np.zeros(mesh.Nfp_v * mesh.Ne2DA),
                         PT_hyd=np.full(mesh.Np * mesh.Ne, theta))
    return bs


def interior_elements(mesh):
    ke = np.arange(mesh.Ne)
    i, j, k = ke % mesh.NeX, (ke // mesh.NeX) % mesh.NeY, ke // (mesh.NeX * mesh.NeY)
    inner = (i > 0) & (i < mesh.NeX - 1) & (j > 0) & (j < mesh.NeY - 1) & (k > 0) & (k < mesh.NeZ - 1)
    return ke[inner]


def nodes_of(mesh, kes):
    return (kes[:, None] * mesh.Np + np.arange(mesh.Np)[None, :]).ravel()


def zeros_state(mesh):
    return [mesh.field() for _ in range(5)]  # DDENS, DRHOT, MOMZ, MOMX, MOMY


def tend(om, mesh, q, DPRES, bs):
    Z = mesh.field()
    for f in q + [DPRES]:
        mesh.exchange(f)
    return om.cal_tend(q, DPRES, bs, Z, Z)  # DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("p", [2, 3])
def test_mass_tendency_is_minus_divergence(p):
    """DENS_dt = -(d/dx MOMX + d/dy MOMY + d/dz MOMZ) for polynomial momentum."""
    mesh = mesh_of(p, periodic=(False, False, False))
    c = dataclasses.replace(C0, GRAV=0.0)
    bs = uniform_base(mesh, c=c)
    om = O.OracleMesh(mesh, consts=c)
    n = mesh.Np * mesh.Ne
    x, y, z = mesh.x / LX, mesh.y / LY, mesh.z / LZ
    q = zeros_state(mesh)
    q[3][:n] = 0.3 * x ** p + 0.2 * x * y
    q[4][:n] = 0.5 * y ** p - 0.1 * z
    q[2][:n] = 0.4 * z ** p + 0.25 * x * z
    div = (0.3 * p * x ** (p - 1) + 0.2 * y) / LX + (0.5 * p * y ** (p - 1)) / LY + \
          (0.4 * p * z ** (p - 1) + 0.25 * x) / LZ
    T = tend(om, mesh, q, mesh.field(), bs)
    sel = nodes_of(mesh, interior_elements(mesh))
    scale = np.max(np.abs(div[sel]))
    assert np.max(np.abs(T[0][sel] + div[sel])) <= 1e-12 * scale * 10
    # RHOT_dt = -theta div(MOM) when rho and theta are uniform
    assert np.max(np.abs(T[4][sel] + 300.0 * div[sel])) <= 1e-12 * 300 * scale * 10


@pytest.mark.parametrize("p", [2, 3])
def test_pressure_gradient_and_buoyancy(p):
    """At rest with polynomial DPRES and DDENS (degree <= p-1 in z):
    MOMX_dt = -dP'/dx, MOMY_dt = -dP'/dy, MOMZ_dt = -dP'/dz - g DDENS."""
    mesh = mesh_of(p, periodic=(False, False, False))
    bs = uniform_base(mesh)
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    x, y, z = mesh.x / LX, mesh.y / LY, mesh.z / LZ
    DP = mesh.field()
    DP[:n] = 50.0 * x ** p + 20.0 * y * z - 30.0 * z ** p
    q = zeros_state(mesh)
    q[0][:n] = 1e-3 * (x ** p + y * z ** (p - 1))
    T = tend(om, mesh, q, DP, bs)
    sel = nodes_of(mesh, interior_elements(mesh))
    gx = 50.0 * p * x ** (p - 1) / LX
    gy = 20.0 * z / LY
    gz = (20.0 * y - 30.0 * p * z ** (p - 1)) / LZ
    s = np.max(np.abs(np.concatenate([gx[sel], gy[sel], gz[sel]])))
    assert np.max(np.abs(T[1][sel] + gx[sel])) <= 1e-11 * s
    assert np.max(np.abs(T[2][sel] + gy[sel])) <= 1e-11 * s
    assert np.max(np.abs(T[3][sel] + gz[sel] + C0.GRAV * q[0][sel])) <= 1e-11 * s
    assert np.max(np.abs(T[0][sel])) == 0.0 and np.max(np.abs(T[4][sel])) == 0.0


def test_buoyancy_filter_removes_top_vertical_mode():
    """-g VFilterPM1(DDENS): a DDENS equal to the degree-p Legendre mode in z
    inside each element gives no buoyancy."""
    import numpy.polynomial.legendre as L
    p = 3
    mesh = mesh_of(p, periodic=(True, True, True))
    bs = uniform_base(mesh)
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    zeta = np.tile(np.repeat(mesh.elem.x1d, mesh.elem.n1 ** 2), mesh.Ne)  # element-local z coordinate
    cP = np.zeros(p + 1)
    cP[p] = 1.0
    q = zeros_state(mesh)
    q[0][:n] = 1e-3 * L.legval(zeta, cP)
    T = tend(om, mesh, q, mesh.field(), bs)
    # only the face terms of the (discontinuous) density field remain in MOMZ_dt: they vanish at rest
    # because the momentum and pressure jumps are zero; the buoyancy of the filtered mode is zero
    assert np.max(np.abs(T[3][:n])) <= 1e-14


def test_coriolis():
    mesh = mesh_of(3, periodic=(True, True, True))
    bs = uniform_base(mesh)
    f0 = 1.0e-4
    bs.CORIOLIS[:] = f0
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    q = zeros_state(mesh)
    q[3][:n] = 2.0   # MOMX
    q[4][:n] = -3.0  # MOMY
    T = tend(om, mesh, q, mesh.field(), bs)
    assert np.max(np.abs(T[1][:n] - f0 * -3.0)) <= 1e-15
    assert np.max(np.abs(T[2][:n] + f0 * 2.0)) <= 1e-15


def test_free_stream_is_preserved():
    c = dataclasses.replace(C0, GRAV=0.0)
    mesh = mesh_of(3, periodic=(True, True, True))
    bs = uniform_base(mesh, c=c)
    om = O.OracleMesh(mesh, consts=c)
    n = mesh.Np * mesh.Ne
    q = zeros_state(mesh)
    q[3][:n], q[4][:n], q[2][:n] = 12.0, -7.0, 3.0
    T = tend(om, mesh, q, mesh.field(), bs)
    for t in T:
        assert np.max(np.abs(t[:n])) <= 1e-12


def test_hydrostatic_rest_state():
    mesh = mesh_of(3, n=(3, 3, 6))
    bs = cases.base_state(mesh, "constBVFreq")
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    q = zeros_state(mesh)
    _, DP = om.drhot2pres(q[1], bs)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    for f in q + [DP]:
        mesh.exchange(f)
    om.apply_bc(q)
    T = om.cal_tend(q, DP, bs, Dx, Dy)
    assert np.max(np.abs(T[0][:n])) == 0.0 and np.max(np.abs(T[4][:n])) == 0.0
    # horizontal gradient of a horizontally uniform PRES_hyd: roundoff of D's row sums
    assert np.max(np.abs(Dx[:n])) <= 1e-12 and np.max(np.abs(Dy[:n])) <= 1e-12
    # DPRES at rest is the roundoff of P(rhotheta_hyd) - PRES_hyd
    assert np.max(np.abs(DP[:n])) <= 1e-10 * np.max(bs.PRES_hyd)
    assert np.max(np.abs(T[3][:n])) <= 1e-9


def test_mass_conservation_periodic():
    """Integral of DENS_dt (exact mass matrix x Jacobian) over a periodic box is zero."""
    p = 3
    mesh = mesh_of(p, n=(3, 3, 3), periodic=(True, True, True))
    c = dataclasses.replace(C0, GRAV=0.0)
    bs = uniform_base(mesh, c=c)
    om = O.OracleMesh(mesh, consts=c)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(7)
    q = zeros_state(mesh)
    q[0][:n] = 0.01 * rng.standard_normal(n)
    q[1][:n] = 1.0 * rng.standard_normal(n)
    for v in (2, 3, 4):
        q[v][:n] = 5.0 * rng.standard_normal(n)
    DP = mesh.field()
    DP[:n] = 10.0 * rng.standard_normal(n)
    T = tend(om, mesh, q, DP, bs)
    V = np.zeros((p + 1, p + 1))
    import numpy.polynomial.legendre as L
    for k in range(p + 1):
        ck = np.zeros(p + 1)
        ck[k] = 1.0
        V[:, k] = L.legval(mesh.elem.x1d, ck) * np.sqrt((2 * k + 1) / 2)
    M1 = np.linalg.inv(V @ V.T)
    M3 = np.kron(M1, np.kron(M1, M1))
    J = mesh.J[:n]
    for v in (0, 4):  # DENS_dt, RHOT_dt: fluxes are pure divergences
        t = T[v][:n].reshape(mesh.Ne, mesh.Np)
        tot = sum(J[ke * mesh.Np] * np.sum(M3 @ t[ke]) for ke in range(mesh.Ne))
        mag = sum(J[ke * mesh.Np] * np.sum(np.abs(M3) @ np.abs(t[ke])) for ke in range(mesh.Ne))
        assert abs(tot) <= 1e-12 * mag


# ---------------------------------------------------------------------------
def random_state(mesh, seed):
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    q = zeros_state(mesh)
    q[0][:n] = 0.02 * rng.standard_normal(n)
    q[1][:n] = 3.0 * rng.standard_normal(n)
    for v in (2, 3, 4):
        q[v][:n] = 10.0 * rng.standard_normal(n)
    DP = mesh.field()
    DP[:n] = 100.0 * rng.standard_normal(n)
    return q, DP


def test_numflux_textbook_rusanov():
    mesh = mesh_of(3, n=(3, 3, 3))
    bs = cases.base_state(mesh, "constT")
    om = O.OracleMesh(mesh)
    q, DP = random_state(mesh, 3)
    for f in q + [DP]:
        mesh.exchange(f)
    om.apply_bc(q)
    df = om.numflux(q, DP, bs).reshape((mesh.NfpTot, 5, mesh.Ne), order="F")
    iM = mesh.vmapM - 1
    iP = mesh.vmapP - 1
    nf = mesh.NfpTot * mesh.Ne
    nx, ny, nz = (mesh.normal_fn[d * nf:(d + 1) * nf] for d in range(3))
    gam = C0.CPdry / C0.CVdry

    def side(ix):
        rho = bs.DENS_hyd[ix] + q[0][ix]
        u, v, w = q[3][ix] / rho, q[4][ix] / rho, q[2][ix] / rho
        un = u * nx + v * ny + w * nz
        P = bs.PRES_hyd[ix] + DP[ix]
        cs = np.sqrt(gam * P / rho)
        flux = {0: rho * un, 1: (bs.THERM_hyd[ix] + q[1][ix]) * un, 2: q[2][ix] * un + DP[ix] * nz,
                3: q[3][ix] * un + DP[ix] * nx, 4: q[4][ix] * un + DP[ix] * ny}
        return flux, np.abs(un) + cs

    FM, aM = side(iM)
    FP, aP = side(iP)
    alpha = np.maximum(aM, aP)
    for v in range(5):
        want = 0.5 * mesh.Fscale * ((FP[v] - FM[v]) - alpha * (q[v][iP] - q[v][iM]))
        got = df[:, v, :].reshape(-1, order="F")
        scale = 0.5 * mesh.Fscale * (np.abs(FP[v]) + np.abs(FM[v]) + alpha * (np.abs(q[v][iP]) + np.abs(q[v][iM])))
        assert np.max(np.abs(got - want) / scale) <= 1e-13, v


def test_numflux_consistency_and_two_sided_identities():
    """Equal states on both sides give exactly zero; across an interior face,
    del_A/Fs_A + del_B/Fs_B = 2 (F_B - F_A).n_A = -(d_A - d_B)... : the central
    parts add and the dissipation parts cancel in the difference."""
    mesh = mesh_of(3, n=(3, 3, 3), periodic=(True, True, True))
    bs = uniform_base(mesh)
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    # continuous field: identical values at coincident nodes (union of face-node pairs)
    nf = mesh.NfpTot * mesh.Ne
    parent = np.arange(mesh.Np * mesh.NeA)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    base = mesh.Np * mesh.Ne
    for t in range(nf):
        a, b = mesh.vmapM[t] - 1, mesh.vmapP[t] - 1
        if b >= base:
            b = mesh.halo_src[b - base] - 1
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[ra] = rb
    roots = np.array([find(a) for a in range(n)])
    rng = np.random.default_rng(11)
    vals = rng.standard_normal((6, mesh.Np * mesh.NeA))
    q = zeros_state(mesh)
    scales = (0.02, 3.0, 10.0, 10.0, 10.0)
    for v in range(5):
        q[v][:n] = scales[v] * vals[v][roots]
    DP = mesh.field()
    DP[:n] = 100.0 * vals[5][roots]
    for f in q + [DP]:
        mesh.exchange(f)
    df = om.numflux(q, DP, bs)
    assert np.max(np.abs(df)) == 0.0

    # two-sided identities on a discontinuous state
    q, DP = random_state(mesh, 5)
    for f in q + [DP]:
        mesh.exchange(f)
    df = om.numflux(q, DP, bs).reshape((mesh.NfpTot, 5, mesh.Ne), order="F")
    iM, iP = mesh.vmapM - 1, mesh.vmapP - 1
    iPi = np.where(iP >= base, mesh.halo_src[np.clip(iP - base, 0, mesh.nB - 1)] - 1, iP)
    # partner face node t' of t: the face node whose M node is t's exterior node and P node is t's interior node
    key = {(int(iM[t]), int(iPi[t])): t for t in range(nf)}
    partner = np.array([key[(int(iPi[t]), int(iM[t]))] for t in range(nf)])
    Fs = mesh.Fscale
    for v in range(5):
        d = df[:, v, :].reshape(-1, order="F") / Fs
        dp = d[partner]
        jump = q[v][iPi] - q[v][iM]
        # d_t - d_t' = -alpha * jump_t * (1/2 + 1/2): dissipation doubles, central parts cancel
        rho_M = bs.DENS_hyd[iM] + q[0][iM]
        rho_P = bs.DENS_hyd[iPi] + q[0][iPi]
        nf_ = mesh.NfpTot * mesh.Ne
        nvec = [mesh.normal_fn[k * nf_:(k + 1) * nf_] for k in range(3)]
        unM = (q[3][iM] * nvec[0] + q[4][iM] * nvec[1] + q[2][iM] * nvec[2]) / rho_M
        unP = (q[3][iPi] * nvec[0] + q[4][iPi] * nvec[1] + q[2][iPi] * nvec[2]) / rho_P
        gam = C0.CPdry / C0.CVdry
        cM = np.sqrt(gam * (bs.PRES_hyd[iM] + DP[iM]) / rho_M)
        cP = np.sqrt(gam * (bs.PRES_hyd[iPi] + DP[iPi]) / rho_P)
        alpha = np.maximum(np.abs(unM) + cM, np.abs(unP) + cP)
        s = np.abs(d) + np.abs(dp) + alpha * np.abs(jump)
        assert np.max(np.abs((d - dp) + alpha * jump) / s) <= 1e-13, v


# ---------------------------------------------------------------------------
def test_drhot2pres_closed_form():
    mesh = mesh_of(3, n=(2, 2, 3))
    bs = cases.base_state(mesh, "constT")
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(1)
    D = mesh.field()
    D[:n] = 5.0 * rng.standard_normal(n)
    for mode in (O.POW_LIBM, O.POW_PORTABLE):
        with O.pow_mode(mode):
            P, DP = om.drhot2pres(D, bs)
        rhot = bs.THERM_hyd[:n] + D[:n]
        ref = C0.PRES00 * (C0.Rdry * rhot / C0.PRES00) ** (C0.CPdry / C0.CVdry)
        # same operation order; numpy's pow may differ from either pow by an ulp
        assert np.max(np.abs(P[:n] - ref) / ref) <= 1e-15
        assert np.array_equal(DP[:n], P[:n] - bs.PRES_hyd[:n])


@pytest.mark.parametrize("periodic", [(True, True, False), (True, False, True), (False, False, False)])
def test_halo_exchange_matches_geometry(periodic):
    mesh = mesh_of(3, n=(3, 2, 2), periodic=periodic)
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    f = mesh.field()
    x, y, z = mesh.x, mesh.y, mesh.z
    f[:n] = np.sin(2 * np.pi * x / LX) + np.cos(2 * np.pi * y / LY) + np.sin(2 * np.pi * z / LZ) * 0.5
    om.exchange(f)
    iM, iP = mesh.vmapM - 1, mesh.vmapP - 1
    ext = iP >= n
    # exterior value equals the field at the coincident node, periodic wrap included
    assert np.max(np.abs(f[iP] - f[iM])) <= 1e-12
    # closed boundaries: the halo holds the interior node's own value exactly
    nf = mesh.NfpTot * mesh.Ne
    nvec = [mesh.normal_fn[k * nf:(k + 1) * nf] for k in range(3)]
    for ax in range(3):
        if periodic[ax]:
            continue
        coord = (x, y, z)[ax]
        L = (LX, LY, LZ)[ax]
        on = ext & (np.abs(nvec[ax]) == 1.0) & ((np.abs(coord[iM]) < 1e-9) | (np.abs(coord[iM] - L) < 1e-9 * L))
        assert on.any()
        assert np.array_equal(f[iP[on]], f[iM[on]])


def test_slip_bc_reflects_normal_momentum():
    mesh = mesh_of(3, n=(2, 2, 3))
    om = O.OracleMesh(mesh)
    q, _ = random_state(mesh, 9)
    for f in q:
        mesh.exchange(f)
    om.apply_bc(q)
    n = mesh.Np * mesh.Ne
    iM, iP = mesh.vmapM - 1, mesh.vmapP - 1
    nf = mesh.NfpTot * mesh.Ne
    nz = mesh.normal_fn[2 * nf:]
    top_bottom = (iP >= n) & (np.abs(nz) == 1.0)
    assert top_bottom.any()
    assert np.array_equal(q[2][iP[top_bottom]], -q[2][iM[top_bottom]])
    assert np.array_equal(q[3][iP[top_bottom]], q[3][iM[top_bottom]])
    assert np.array_equal(q[4][iP[top_bottom]], q[4][iM[top_bottom]])
