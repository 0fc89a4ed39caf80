"""The oracle of the reference's communication-hiding pair -- cal_tend_cco
(FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:493-678) and
add_bnd_contrib_generalvc (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:242-496)
-- pinned on the CPU:

* interior part + face contribution = the whole tendency of cal_tend (already
  pinned by closed forms) to the rounding of the regrouped sums, on a state
  with every prognostic nonzero: a dropped lift, a wrong sign of the face term
  or a flux with the wrong side would fail it by many orders of magnitude;
* the halo DPRES that add_bnd_contrib recomputes equals the exchanged one bit
  for bit (same formula, same inputs);
* the hydrostatic rest state gives exact zeros from both parts;
* the moist-array path with the dry constants in the arrays is the dry path
  bit for bit.
"""
import numpy as np
import pytest

import fedg_oracle as O
from fedg import cases


def perturbed(mesh, bs, case, om, Dx, Dy, seed=3):
    q = cases.warm_bubble_state(mesh, bs, case.bubbles)
    om.heve_update(q, bs, Dx, Dy, case.dt, 2)
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    for v, amp in ((2, 0.05), (3, 0.3), (4, 0.2)):
        q[v][:n] += amp * rng.standard_normal(n)
    for f in q:
        om.exchange(f)
    _, D = om.drhot2pres(q[1], bs)
    om.exchange(D)
    om.apply_bc(q)
    return q, D


@pytest.fixture(scope="module", params=[2, 3], ids=["p2", "p3"])
def setup(request):
    case = cases.scaled_case(cases.config1(), 4, 3, 4, porder=request.param)
    mesh, bs, _ = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    return case, mesh, bs, om, Dx, Dy


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def test_cco_plus_bnd_equals_cal_tend(setup):
    case, mesh, bs, om, Dx, Dy = setup
    q, D = perturbed(mesh, bs, case, om, Dx, Dy)
    whole = om.cal_tend(q, D, bs, Dx, Dy)
    dt = om.cal_tend_cco(q, D, bs, Dx, Dy)
    face = [a.copy() for a in dt]
    om.add_bnd_contrib(dt, q, D.copy(), bs)
    n = mesh.Np * mesh.Ne
    for name, a, b, c in zip(("DENS_dt", "MOMX_dt", "MOMY_dt", "MOMZ_dt", "RHOT_dt"), dt, whole, face):
        scale = np.max(np.abs(b[:n]))
        err = np.max(np.abs(a[:n] - b[:n]))
        # the face contribution is a large part of the tendency: the check has teeth
        assert np.max(np.abs(a[:n] - c[:n])) > 1e-3 * scale, name
        assert err <= 1e-12 * scale, f"{name}: {err / scale:.2e}"


def test_bnd_halo_dpres_is_the_exchanged_one(setup):
    case, mesh, bs, om, Dx, Dy = setup
    q, D = perturbed(mesh, bs, case, om, Dx, Dy)
    D2 = D.copy()
    n = mesh.Np * mesh.Ne
    D2[n:] = -1.0
    dt = [mesh.field() for _ in range(5)]
    om.add_bnd_contrib(dt, q, D2, bs)
    assert np.array_equal(D2[n:n + mesh.nB], D[n:n + mesh.nB])
    assert np.array_equal(D2[:n], D[:n])


def test_rest_state_gives_zero(setup):
    case, mesh, bs, om, Dx, Dy = setup
    q = [mesh.field() for _ in range(5)]
    D = mesh.field()
    dt = om.cal_tend_cco(q, D, bs, Dx, Dy)
    om.add_bnd_contrib(dt, q, D, bs)
    n = mesh.Np * mesh.Ne
    for a in dt:
        assert np.max(np.abs(a[:n])) <= 1e-12


def test_moist_arrays_with_dry_constants_match_dry_path(setup):
    case, mesh, bs, om, Dx, Dy = setup
    q, D = perturbed(mesh, bs, case, om, Dx, Dy, seed=5)
    a = om.cal_tend_cco(q, D, bs, Dx, Dy)
    b = [x.copy() for x in a]
    om.add_bnd_contrib(a, q, D.copy(), bs, moist=False)
    om.add_bnd_contrib(b, q, D.copy(), bs, moist=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
