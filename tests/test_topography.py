"""Terrain-following vertical coordinate of the box builder
(fedg_box_topography_metrics: MSH/scale_meshutil_vcoord.F90:81-209,
MSH/scale_mesh_topography.F90:159-258, MSH/scale_mesh_cubedom3d.F90:292-334).

Pinned by closed forms and invariants:
  - flat topography gives the flat box's metrics bit for bit;
  - a linear topography a x + b y + c (its DG gradient exact, no face jumps):
    GsqrtV = 1 - topo/zTop, G13 = -(1 - zeta/zTop)/GsqrtV a, G23 the same
    with b, zlev = zeta + (1 - zeta/zTop) topo, to rounding;
  - a piecewise-constant topography (only face jumps): the exact-mass integral
    of the element's gradient equals the face integral of the central jump,
    0.5 dy (t_east - t_west) for d/dx (a wrong sign, factor or face column of
    the lift fails it);
  - a smooth periodic mountain: G13 converges spectrally to the analytic
    metric as p grows;
  - halo slots hold the single-tile exchange's values;
  - over a mountain the hydrostatic rest state stays near rest, the spurious
    momentum after 50 steps dropping tenfold from p = 3 to p = 5.
"""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases


def box(p, NeX, NeY, NeZ, Lx, Ly, Lz, periodic):
    return fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, Lx), (0.0, Ly), (0.0, Lz), periodic=periodic)


def test_flat_topography_is_the_flat_box():
    m = box(3, 3, 2, 2, 3e3, 2e3, 2e3, (True, False, False))
    G0, GI0 = m.Gsqrt.copy(), m.GI3.copy()
    m.set_topography(np.zeros(m.Nfp_v * m.Ne2D))
    assert np.array_equal(m.Gsqrt, G0)
    assert np.array_equal(np.abs(m.GI3), np.abs(GI0))
    assert np.array_equal(m.height, m.z)


@pytest.mark.parametrize("p", [2, 4, 7])
def test_linear_topography_closed_form(p):
    m = box(p, 4, 3, 3, 4e3, 3e3, 6e3, (False, False, False))
    a, b, c0, zTop = 0.05, -0.02, 100.0, 6e3
    m.set_topography(lambda x, y: a * x + b * y + c0, zTop=zTop)
    n, nA = m.Np * m.Ne, m.Np * m.NeA
    zeta = m.z
    topo = a * m.x + b * m.y + c0
    G = 1.0 - topo / zTop
    coef = -(1.0 - zeta / zTop) / G
    np.testing.assert_allclose(m.Gsqrt[:n], G, rtol=1e-14)
    np.testing.assert_allclose(m.GI3[:n], coef * a, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(m.GI3[nA:nA + n], coef * b, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(m.height, zeta + (1.0 - zeta / zTop) * topo, rtol=1e-14)
    # halo slots: the exchange's copies
    src = m.halo_src - 1
    assert np.array_equal(m.Gsqrt[n:n + m.nB], m.Gsqrt[src])
    assert np.array_equal(m.GI3[n:n + m.nB], m.GI3[src])
    assert np.array_equal(m.GI3[nA + n:nA + n + m.nB], m.GI3[nA + src])
    assert np.array_equal(m.zlev[n:n + m.nB], m.zlev[src])


def test_piecewise_constant_topography_lifts_the_face_jumps():
    p, NeX, NeY = 3, 4, 3
    Lx, Ly, zTop = 4e3, 3e3, 5e3
    m = box(p, NeX, NeY, 1, Lx, Ly, zTop, (True, True, False))
    rng = np.random.default_rng(1)
    t_el = rng.uniform(0.0, 300.0, size=(NeY, NeX))
    Nfp = m.Nfp_v
    topo = np.repeat(t_el.ravel(), Nfp)
    m.set_topography(topo, zTop=zTop)
    n1 = p + 1
    e = m.elem
    M1 = np.linalg.inv(e.invM1D.reshape(n1, n1, order="F"))
    dx, dy = Lx / NeX, Ly / NeY
    J2 = 0.25 * dx * dy
    for jj in range(NeY):
        for ii in range(NeX):
            ke = ii + NeX * jj  # bottom level: zeta = 0 at k = 0
            te, tw = t_el[jj, (ii + 1) % NeX], t_el[jj, (ii - 1) % NeX]
            tn, ts = t_el[(jj + 1) % NeY, ii], t_el[(jj - 1) % NeY, ii]
            G = 1.0 - t_el[jj, ii] / zTop
            sl = slice(m.Np * ke, m.Np * ke + Nfp)  # nodes k = 0 of the element: coef = -1/G
            gx = -G * m.GI3[sl].reshape(n1, n1, order="F")
            gy = -G * m.GI3[m.Np * m.NeA + m.Np * ke:m.Np * m.NeA + m.Np * ke + Nfp].reshape(n1, n1, order="F")
            ix = np.einsum("ia,jb,ij->", M1, M1, gx) * J2
            iy = np.einsum("ia,jb,ij->", M1, M1, gy) * J2
            np.testing.assert_allclose(ix, 0.5 * dy * (te - tw), rtol=1e-11, atol=1e-9)
            np.testing.assert_allclose(iy, 0.5 * dx * (tn - ts), rtol=1e-11, atol=1e-9)


def test_smooth_mountain_converges_spectrally():
    Lx, Ly, zTop, h = 8e3, 2e3, 4e3, 300.0
    k = 2 * np.pi / Lx
    errs = []
    for p in (3, 5, 7):
        m = box(p, 4, 1, 2, Lx, Ly, zTop, (True, True, False))
        m.set_topography(lambda x, y: h * (1.0 + np.cos(k * x)), zTop=zTop)
        n = m.Np * m.Ne
        topo = h * (1.0 + np.cos(k * m.x))
        dtdx = -h * k * np.sin(k * m.x)
        exact = -(1.0 - m.z / zTop) / (1.0 - topo / zTop) * dtdx
        errs.append(np.max(np.abs(m.GI3[:n] - exact)) / np.max(np.abs(exact)))
    assert errs[0] < 5e-2  # 2.8e-2, 2.6e-4, 1.1e-6 measured for p = 3, 5, 7
    assert errs[1] < errs[0] / 30.0
    assert errs[2] < errs[1] / 30.0


def test_rest_over_a_mountain_stays_near_rest():
    """Hydrostatic rest state (N = 0.01 1/s) over a 400 m Agnesi ridge, 50 RK4
    steps on the oracle: the spurious momentum is small and drops with p."""
    peak = []
    for p, dt in ((3, 0.05), (5, 0.02)):
        Lx, Lz = 16e3, 8e3
        m = box(p, 8, 1, 4, Lx, 2e3, Lz, (True, True, False))
        m.set_topography(lambda x, y: 400.0 / (1.0 + ((x - 0.5 * Lx) / 2e3) ** 2), zTop=Lz)
        bs = cases.base_state(m, "constBVFreq")
        om = O.OracleMesh(m)
        Dx, Dy = om.calc_phyd_hgrad(bs)
        q = [m.field() for _ in range(5)]
        om.heve_update(q, bs, Dx, Dy, dt, 50)
        n = m.Np * m.Ne
        peak.append(max(np.max(np.abs(q[2][:n])), np.max(np.abs(q[3][:n]))))
    assert peak[0] < 2e-2
    assert peak[1] < peak[0] / 10.0
