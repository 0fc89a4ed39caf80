"""Pins of the oracle's general-metric terms: Gsqrt, GsqrtH, G13/G23 (GI3),
the Escale diagonal, CORIOLIS, DPhydDx/Dy in the HEVE tendency and its
numerical flux, and calc_phyd_hgrad.

What fixes each answer (not a retyped copy of the oracle's formula):
  - closed forms: the reference's Div unit test
    (FElib/test/FE/element_operation_hexahedral/test_element_operation_hexahedral.f90:115-144:
    Escale = (1, 2, 0.2) on the diagonal, division by a non-constant Gsqrt)
    carried through the whole tendency, with every flux a polynomial of
    degree <= p per coordinate (the DG derivative is then exact and the
    continuous fields have no face jumps); the derivatives come from sympy;
  - terrain-following invariants on a Gal-Chen mesh (z = zeta (H - h)/H + h):
    a uniform wind is preserved (the metric identities
    d/dx Gsqrt + d/dzeta (Gsqrt G13) = 0), and the horizontal gradient of a
    pressure that depends on the height only vanishes;
  - conservation: with random, discontinuous metrics the exact-mass-matrix
    integral of Gsqrt x (every flux-form tendency) over a periodic box is
    zero, and so is the integral of GsqrtV DPhydDx / DPhydDy (the lift of the
    face jump is the central flux);
  - the Rusanov flux in contravariant vector form: normal velocity U.n with
    U = (u, v, w/Gsqrt + G13 u + G23 v), wave speed |U.n| + c |grad xi_n|
    where grad zeta = (G13, G23, 1/Gsqrt), pressure force Gsqrt P' (grad of
    the physical coordinate).
"""
import dataclasses

import numpy as np
import pytest
import sympy as sp

import fedg_oracle as O
from fedg import cases, constants
from fedg.mesh import BoxMesh

C0 = constants.SCALE
LX, LY, LZ = 4000.0, 3000.0, 2000.0
X, Y, Z = sp.symbols("X Y Z")  # x/LX, y/LY, z/LZ
ESC = (1.0, 2.0, 0.2)           # Escale factors of the reference's Div test


def mesh_of(p, n=(3, 3, 3), periodic=(False, False, False)):
    return BoxMesh(p, n[0], n[1], n[2], (0.0, LX), (0.0, LY), (0.0, LZ), periodic=periodic)


def ev(expr, mesh):
    """A sympy expression of (X, Y, Z) at the interior nodes."""
    f = sp.lambdify((X, Y, Z), expr, "numpy")
    return np.broadcast_to(np.asarray(f(mesh.x / LX, mesh.y / LY, mesh.z / LZ), dtype=np.float64),
                           mesh.x.shape).copy()


def set_field(mesh, expr):
    a = mesh.field()
    a[: mesh.Np * mesh.Ne] = ev(expr, mesh)
    mesh.exchange(a)
    return a


def set_2d(mesh, arr, expr):
    """A 2-D (x, y) field indexed (IndexH2Dto3D(p), EMap3Dto2D(ke)) like GsqrtH and CORIOLIS."""
    vals = ev(expr, mesh).reshape(mesh.Ne, mesh.Np)
    idx = (mesh.elem.IndexH2Dto3D[None, :] - 1) + mesh.Nfp_v * (mesh.EMap3Dto2D[:, None] - 1)
    arr[idx.ravel()] = vals.ravel()


def scale_escale(mesh, fac=ESC):
    n = mesh.Np * mesh.Ne
    for d, s in enumerate(fac):
        mesh.Escale[(4 * d) * n:(4 * d + 1) * n] *= s


def set_metrics(mesh, Gs, GsH, G13, G23):
    n, nA = mesh.Np * mesh.Ne, mesh.Np * mesh.NeA
    mesh.Gsqrt[:n] = ev(Gs, mesh)
    mesh.GI3[:n] = ev(G13, mesh)
    mesh.GI3[nA:nA + n] = ev(G23, mesh)
    mesh.exchange(mesh.Gsqrt)
    mesh.exchange(mesh.GI3[:nA])
    mesh.exchange(mesh.GI3[nA:])
    set_2d(mesh, mesh.GsqrtH, GsH)


def div(F, fac=ESC):
    return fac[0] / LX * sp.diff(F[0], X) + fac[1] / LY * sp.diff(F[1], Y) + fac[2] / LZ * sp.diff(F[2], Z)


def uniform_base(mesh, rho, theta, c=C0):
    rhot = rho * theta
    P = c.PRES00 * (c.Rdry * rhot / c.PRES00) ** (c.CPdry / c.CVdry)
    return cases.BaseState(DENS_hyd=mesh.field(rho), PRES_hyd=mesh.field(P), THERM_hyd=mesh.field(rhot),
                           PRES_hyd_ref=mesh.field(0.0), Rtot=mesh.field(c.Rdry), CVtot=mesh.field(c.CVdry),
                           CPtot=mesh.field(c.CPdry), CORIOLIS=np.zeros(mesh.Nfp_v * mesh.Ne2DA),
                           PT_hyd=np.full(mesh.Np * mesh.Ne, theta))


def exact_mass_matrix(mesh):
    import numpy.polynomial.legendre as L
    p = mesh.porder
    V = np.zeros((p + 1, p + 1))
    for k in range(p + 1):
        ck = np.zeros(p + 1)
        ck[k] = 1.0
        V[:, k] = L.legval(mesh.elem.x1d, ck) * np.sqrt((2 * k + 1) / 2)
    M1 = np.linalg.inv(V @ V.T)
    return np.kron(M1, np.kron(M1, M1))


def integral(mesh, f, M3):
    """(exact-mass-matrix integral, integral of |f|) over the elements."""
    t = f[: mesh.Np * mesh.Ne].reshape(mesh.Ne, mesh.Np)
    J = mesh.J[: mesh.Np * mesh.Ne].reshape(mesh.Ne, mesh.Np)[:, 0]
    tot = float(np.sum(J * (t @ M3.T).sum(axis=1)))
    mag = float(np.sum(J * (np.abs(t) @ np.abs(M3).T).sum(axis=1)))
    return tot, mag


# ---------------------------------------------------------------------------
def test_general_metric_tendency_closed_form():
    """All five HEVE tendencies of polynomial fields with Gsqrt = GsqrtH(x, y) GsqrtV(z),
    G13/G23(z), Escale diagonal scaled by (1, 2, 0.2), a Coriolis parameter f(x, y),
    given DPhydDx/Dy and buoyancy of a DDENS of degree p-1 in z:
      DENS_dt = -div(Gsqrt rho U)/Gsqrt,  U = (u, v, w/GsqrtV + G13 u + G23 v),
      RHOT_dt = -div(Gsqrt rho theta U)/Gsqrt,
      MOMZ_dt = -div(Gsqrt rho w U + (0, 0, GsqrtH P'))/Gsqrt - g DDENS,
      MOMX_dt = -DPhydDx + f MOMY - div(Gsqrt rho u U + Gsqrt P' (1, 0, G13))/Gsqrt,
      MOMY_dt = -DPhydDy - f MOMX - div(Gsqrt rho v U + Gsqrt P' (0, 1, G23))/Gsqrt,
    with div(F) = 1 dF1/dx + 2 dF2/dy + 0.2 dF3/dz (the scaled Escale)."""
    p = 4
    mesh = mesh_of(p)
    GsH = 1 + sp.Rational(1, 10) * X + sp.Rational(1, 20) * Y
    GsV = 1 - sp.Rational(1, 10) * Z
    Gs = GsH * GsV
    G13 = sp.Rational(1, 20) + sp.Rational(1, 10) * Z
    G23 = -sp.Rational(3, 100) + sp.Rational(8, 100) * Z
    set_metrics(mesh, Gs, GsH, G13, G23)
    scale_escale(mesh)
    rho0, th0 = 1.1, 300.0
    bs = uniform_base(mesh, rho0, th0)
    DD = sp.Rational(1, 100) * (X + Y + Z)
    dens = rho0 + DD
    u = 5 + 3 * X - 2 * Y
    v = -4 + 2 * Y
    w = 1 + X + sp.Rational(1, 2) * Y
    th = th0 + 2 * X + 3 * Z
    DP = 30 * X ** 2 - 20 * Y * Z + 10 * Z
    cor = sp.Rational(1, 10000) * (1 + X - Y)
    q = [set_field(mesh, DD), set_field(mesh, dens * th - rho0 * th0), set_field(mesh, dens * GsV * w),
         set_field(mesh, dens * u), set_field(mesh, dens * v)]
    DPf = set_field(mesh, DP)
    set_2d(mesh, bs.CORIOLIS, cor)
    rng = np.random.default_rng(4)
    n = mesh.Np * mesh.Ne
    Dx, Dy = mesh.field(), mesh.field()
    Dx[:n], Dy[:n] = 0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n)
    om = O.OracleMesh(mesh)
    T = om.cal_tend(q, DPf, bs, Dx, Dy)  # DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt

    U3 = w + G13 * u + G23 * v
    Fd = [Gs * dens * u, Gs * dens * v, Gs * dens * U3]
    want = {
        "DENS": -div(Fd) / Gs,
        "RHOT": -div([f * th for f in Fd]) / Gs,
        "MOMZ": -div([Fd[0] * GsV * w, Fd[1] * GsV * w, Fd[2] * GsV * w + GsH * DP]) / Gs - C0.GRAV * DD,
        "MOMX": cor * dens * v - div([Fd[0] * u + Gs * DP, Fd[1] * u, Fd[2] * u + Gs * DP * G13]) / Gs,
        "MOMY": -cor * dens * u - div([Fd[0] * v, Fd[1] * v + Gs * DP, Fd[2] * v + Gs * DP * G23]) / Gs,
    }
    got = {"DENS": T[0], "MOMX": T[1] + Dx, "MOMY": T[2] + Dy, "MOMZ": T[3], "RHOT": T[4]}
    for name, expr in want.items():
        ref = ev(expr, mesh)
        g = got[name][:n]
        scale = np.max(np.abs(ref))
        assert scale > 0.0
        err = np.max(np.abs(g - ref)) / scale
        assert err <= 1e-11, f"{name}: {err:.2e}"


def galchen(mesh, hx=0.04, hy=-0.02, h0=150.0):
    """Terrain-following metrics of z = zeta (H - h)/H + h with h = h0 + hx x + hy y on the
    (x, y, zeta) box: Gsqrt = GsqrtV = (H - h)/H, GsqrtH = 1,
    G13 = hx (zeta - H)/(H - h), G23 = hy (zeta - H)/(H - h).  Returns the height z(X, Y, Z)."""
    H = LZ
    h = h0 + hx * LX * X + hy * LY * Y
    zeta = LZ * Z
    Gs = (H - h) / H
    set_metrics(mesh, Gs, sp.Integer(1), hx * (zeta - H) / (H - h), hy * (zeta - H) / (H - h))
    return zeta * (H - h) / H + h


def test_terrain_following_free_stream_is_preserved():
    """A uniform wind over terrain (non-constant Gsqrt, G13, G23 that satisfy the metric
    identities) has zero tendency; each flux divergence term alone is not zero."""
    p = 3
    mesh = mesh_of(p)
    galchen(mesh)
    c = dataclasses.replace(C0, GRAV=0.0)
    rho0 = 1.2
    bs = uniform_base(mesh, rho0, 300.0, c=c)
    n = mesh.Np * mesh.Ne
    q = [mesh.field() for _ in range(5)]
    u0, v0, w0 = 12.0, -7.0, 3.0
    q[3][:n], q[4][:n], q[2][:n] = rho0 * u0, rho0 * v0, rho0 * w0
    for f in q:
        mesh.exchange(f)
    Z0 = mesh.field()
    om = O.OracleMesh(mesh, consts=c)
    T = om.cal_tend(q, mesh.field(), bs, Z0, Z0)
    # size of the single terms: d/dx (Gsqrt rho u) = -rho u hx / H
    term = rho0 * abs(u0) * 0.04 / LZ
    for name, t in zip(("DENS", "MOMX", "MOMY", "MOMZ", "RHOT"), T):
        s = term * (1.0 if name == "DENS" else 300.0 if name == "RHOT" else 12.0)
        assert np.max(np.abs(t[:n])) <= 1e-10 * s, f"{name}: {np.max(np.abs(t[:n])):.2e}"


def test_phyd_hgrad_closed_form_flat_and_general():
    """calc_phyd_hgrad of a horizontally varying PRES_hyd and non-zero PRES_hyd_ref:
    flat metrics: (d/dx, d/dy) (PRES_hyd - PRES_hyd_ref);
    general metrics (GsqrtV = Gsqrt/GsqrtH, G13, G23, Escale scaled by (1, 2, 0.2)):
      DPhydDx = (1 d/dx (GsqrtV dP) + 0.2 d/dz (G13 GsqrtV dP)) / GsqrtV,
      DPhydDy = (2 d/dy (GsqrtV dP) + 0.2 d/dz (G23 GsqrtV dP)) / GsqrtV."""
    p = 4
    Ph = 1.0e5 * (1 - sp.Rational(1, 10) * Z) + 500 * X ** 2 * Y - 300 * Y * Z + 80 * X * Z
    Pr = 200 * X - 100 * Y ** 2 + 40 * X * Y
    dP = Ph - Pr
    for general in (False, True):
        mesh = mesh_of(p)
        if general:
            GsH = 1 + sp.Rational(1, 10) * X + sp.Rational(1, 20) * Y
            GsV = 1 - sp.Rational(1, 10) * Z
            G13 = sp.Rational(1, 20) + sp.Rational(1, 10) * Z
            G23 = -sp.Rational(3, 100) + sp.Rational(8, 100) * Z
            set_metrics(mesh, GsH * GsV, GsH, G13, G23)
            scale_escale(mesh)
            fac = ESC
        else:
            GsV, G13, G23, fac = sp.Integer(1), sp.Integer(0), sp.Integer(0), (1.0, 1.0, 1.0)
        bs = cases.base_state(mesh, "constT")
        bs.PRES_hyd = set_field(mesh, Ph)
        bs.PRES_hyd_ref = set_field(mesh, Pr)
        om = O.OracleMesh(mesh)
        Dx, Dy = om.calc_phyd_hgrad(bs)
        n = mesh.Np * mesh.Ne
        wx = (fac[0] / LX * sp.diff(GsV * dP, X) + fac[2] / LZ * sp.diff(G13 * GsV * dP, Z)) / GsV
        wy = (fac[1] / LY * sp.diff(GsV * dP, Y) + fac[2] / LZ * sp.diff(G23 * GsV * dP, Z)) / GsV
        for name, g, expr in (("DPhydDx", Dx, wx), ("DPhydDy", Dy, wy)):
            ref = ev(expr, mesh)
            scale = np.max(np.abs(ref))
            assert scale > 1e-3
            err = np.max(np.abs(g[:n] - ref)) / scale
            assert err <= 1e-11, f"{name} general={general}: {err:.2e}"


def test_phyd_hgrad_of_height_dependent_pressure_vanishes_over_terrain():
    """Over terrain the reference pressure P(z) varies along the zeta surfaces, but its
    horizontal gradient at constant height is zero: the two terms of calc_phyd_hgrad
    (d/dx of GsqrtV P and d/dzeta of G13 GsqrtV P) cancel."""
    p = 4
    mesh = mesh_of(p)
    zh = galchen(mesh)
    bs = cases.base_state(mesh, "constT")
    bs.PRES_hyd = set_field(mesh, 1.0e5 - 11.0 * zh + 1.0e-3 * zh ** 2)
    bs.PRES_hyd_ref = mesh.field()
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    n = mesh.Np * mesh.Ne
    # one term alone: d/dx (GsqrtV P) ~ P hx / H
    term = 1.0e5 * 0.04 / LZ
    assert np.max(np.abs(Dx[:n])) <= 1e-11 * term, np.max(np.abs(Dx[:n]))
    assert np.max(np.abs(Dy[:n])) <= 1e-11 * term, np.max(np.abs(Dy[:n]))


def random_metrics(mesh, rng, GsH_one=True):
    n, nA = mesh.Np * mesh.Ne, mesh.Np * mesh.NeA
    mesh.Gsqrt[:n] = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, n)
    mesh.GI3[:n] = 0.05 * rng.standard_normal(n)
    mesh.GI3[nA:nA + n] = 0.05 * rng.standard_normal(n)
    mesh.exchange(mesh.Gsqrt)
    mesh.exchange(mesh.GI3[:nA])
    mesh.exchange(mesh.GI3[nA:])
    if not GsH_one:
        mesh.GsqrtH[:] = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, mesh.GsqrtH.size)


def test_phyd_hgrad_periodic_integral_vanishes_random_metrics():
    """With a discontinuous PRES_hyd - PRES_hyd_ref and random metrics, the face jump's lift
    turns the element derivatives into a central flux: the integral of GsqrtV DPhydDx
    (and DPhydDy) over a periodic box is zero."""
    p = 3
    mesh = mesh_of(p, n=(3, 3, 3), periodic=(True, True, True))
    rng = np.random.default_rng(12)
    random_metrics(mesh, rng)
    n = mesh.Np * mesh.Ne
    bs = cases.base_state(mesh, "constT")
    bs.PRES_hyd[:n] = 1.0e5 + 50.0 * rng.standard_normal(n)
    bs.PRES_hyd_ref[:n] = 20.0 * rng.standard_normal(n)
    mesh.exchange(bs.PRES_hyd)
    mesh.exchange(bs.PRES_hyd_ref)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    M3 = exact_mass_matrix(mesh)
    GsV = mesh.Gsqrt[:n]  # GsqrtH = 1
    for name, g in (("x", Dx), ("y", Dy)):
        tot, mag = integral(mesh, GsV * g[:n], M3)
        assert mag > 0.0
        assert abs(tot) <= 1e-12 * mag, f"{name}: {tot:.3e} of {mag:.3e}"


def test_conservation_with_random_metrics():
    """Every flux-form tendency conserves: with random (discontinuous) Gsqrt, G13, G23, no
    gravity, Coriolis or DPhyd, the exact-mass-matrix integral of Gsqrt x tendency over a
    periodic box is zero for DENS, RHOT and the three momenta."""
    p = 3
    mesh = mesh_of(p, n=(3, 3, 3), periodic=(True, True, True))
    rng = np.random.default_rng(21)
    random_metrics(mesh, rng)
    c = dataclasses.replace(C0, GRAV=0.0)
    bs = uniform_base(mesh, 1.2, 300.0, c=c)
    n = mesh.Np * mesh.Ne
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = 0.01 * rng.standard_normal(n)
    q[1][:n] = 1.0 * rng.standard_normal(n)
    for v in (2, 3, 4):
        q[v][:n] = 5.0 * rng.standard_normal(n)
    DP = mesh.field()
    DP[:n] = 10.0 * rng.standard_normal(n)
    for f in q + [DP]:
        mesh.exchange(f)
    Z0 = mesh.field()
    om = O.OracleMesh(mesh, consts=c)
    T = om.cal_tend(q, DP, bs, Z0, Z0)
    M3 = exact_mass_matrix(mesh)
    Gs = mesh.Gsqrt[:n]
    for name, t in zip(("DENS", "MOMX", "MOMY", "MOMZ", "RHOT"), T):
        tot, mag = integral(mesh, Gs * t[:n], M3)
        assert abs(tot) <= 1e-12 * mag, f"{name}: {tot:.3e} of {mag:.3e}"


def test_numflux_contravariant_rusanov_random_metrics():
    """The face flux with random metrics (GsqrtH = 1) against the Rusanov flux of the Euler
    equations in contravariant form: normal mass flux Gsqrt rho U.n with
    U = (u, v, w/Gsqrt + G13 u + G23 v); wave speed |U.n| + c |grad xi_n| with grad x = e_x,
    grad y = e_y and grad zeta = (G13, G23, 1/Gsqrt); pressure forces Gsqrt P' (1, 0, G13),
    (0, 1, G23), (0, 0, 1/Gsqrt) for MOMX, MOMY, MOMZ; dissipation on Gsqrt q."""
    mesh = mesh_of(3, n=(3, 3, 3), periodic=(True, True, False))
    rng = np.random.default_rng(33)
    random_metrics(mesh, rng)
    bs = cases.base_state(mesh, "constT")
    n = mesh.Np * mesh.Ne
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = 0.02 * rng.standard_normal(n)
    q[1][:n] = 3.0 * rng.standard_normal(n)
    for v in (2, 3, 4):
        q[v][:n] = 10.0 * rng.standard_normal(n)
    DP = mesh.field()
    DP[:n] = 100.0 * rng.standard_normal(n)
    for f in q + [DP]:
        mesh.exchange(f)
    om = O.OracleMesh(mesh)
    om.apply_bc(q)
    df = om.numflux(q, DP, bs).reshape((mesh.NfpTot, 5, mesh.Ne), order="F")
    iM, iP = mesh.vmapM - 1, mesh.vmapP - 1
    nf = mesh.NfpTot * mesh.Ne
    nvec = np.stack([mesh.normal_fn[d * nf:(d + 1) * nf] for d in range(3)], axis=1)
    gam = C0.CPdry / C0.CVdry
    nA = mesh.Np * mesh.NeA
    G13a, G23a = mesh.GI3[:nA], mesh.GI3[nA:]

    def side(ix):
        Gs, g13, g23 = mesh.Gsqrt[ix], G13a[ix], G23a[ix]
        rho = bs.DENS_hyd[ix] + q[0][ix]
        u, v, w = q[3][ix] / rho, q[4][ix] / rho, q[2][ix] / rho
        U = np.stack([u, v, w / Gs + g13 * u + g23 * v], axis=1)
        Un = np.sum(U * nvec, axis=1)
        gz = np.stack([g13, g23, 1.0 / Gs], axis=1)
        vert = np.abs(nvec[:, 2]) > 0.5
        gnorm = np.where(vert, np.linalg.norm(gz, axis=1), 1.0)
        P = bs.PRES_hyd[ix] + DP[ix]
        cs = np.sqrt(gam * P / rho)
        fx = np.stack([np.ones_like(Gs), np.zeros_like(Gs), g13], axis=1)
        fy = np.stack([np.zeros_like(Gs), np.ones_like(Gs), g23], axis=1)
        fz = np.stack([np.zeros_like(Gs), np.zeros_like(Gs), 1.0 / Gs], axis=1)
        pn = lambda f: Gs * DP[ix] * np.sum(f * nvec, axis=1)
        flux = {0: Gs * rho * Un, 1: Gs * (bs.THERM_hyd[ix] + q[1][ix]) * Un, 2: Gs * q[2][ix] * Un + pn(fz),
                3: Gs * q[3][ix] * Un + pn(fx), 4: Gs * q[4][ix] * Un + pn(fy)}
        return flux, np.abs(Un) + cs * gnorm, Gs

    FM, aM, GM = side(iM)
    FP, aP, GP = side(iP)
    alpha = np.maximum(aM, aP)
    for v in range(5):
        jump = GP * q[v][iP] - GM * q[v][iM]
        want = 0.5 * mesh.Fscale * ((FP[v] - FM[v]) - alpha * jump)
        got = df[:, v, :].reshape(-1, order="F")
        scale = 0.5 * mesh.Fscale * (np.abs(FP[v]) + np.abs(FM[v]) + alpha * (np.abs(GP * q[v][iP]) +
                                                                             np.abs(GM * q[v][iM])))
        assert np.max(np.abs(got - want) / scale) <= 1e-13, v
