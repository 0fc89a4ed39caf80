"""Modal filter (MODALFILTER_FLAG): the builder's filter matrix, the oracle's
element filter and its place in the oracle's step, on the CPU.

The matrix is pinned by what the exponential filter means: F = V diag(sigma)
V^-1 has every Legendre polynomial as an eigenvector (sampled at the LGL
nodes), the constant passes unchanged (sigma_0 = 1), modes with
eta = k/p <= etac pass unchanged, and with etac = 0 the top mode is damped by
exactly exp(-alpha).  The oracle's element filter is checked against an
independent tensor contraction (numpy einsum) of the same matrices, with a
Gsqrt that is not 1 so that the weighting is exercised.
"""
import numpy as np
import numpy.polynomial.legendre as npleg
import pytest

import fedg
import fedg_oracle as O
from fedg import cases


def lgl(p):
    e = fedg.RefElem(p)
    return np.asarray(e.x1d)


@pytest.mark.parametrize("p", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("etac,alpha,order", [(2.0 / 3.0, 36.0, 16), (0.0, 0.25, 24), (0.5, 2.0, 4)])
def test_filter_matrix_damps_legendre_modes(p, etac, alpha, order):
    F = fedg.modalfilter_matrix(p, etac, alpha, order)
    x = lgl(p)
    for k in range(p + 1):
        Pk = npleg.legval(x, np.eye(p + 1)[k])
        y = F @ Pk
        sig = float(y @ Pk / (Pk @ Pk))
        assert np.allclose(y, sig * Pk, atol=1e-12), f"mode {k} is not an eigenvector"
        eta = k / p
        if k == 0 or eta <= etac:
            assert abs(sig - 1.0) < 1e-12, (k, sig)
        else:
            assert -1e-12 < sig < 1.0
        if k == p and etac == 0.0:
            assert abs(sig - np.exp(-alpha)) < 1e-12
    # the damping grows with the mode number
    sig = [float((F @ npleg.legval(x, np.eye(p + 1)[k])) @ npleg.legval(x, np.eye(p + 1)[k])
                 / (npleg.legval(x, np.eye(p + 1)[k]) @ npleg.legval(x, np.eye(p + 1)[k]))) for k in range(p + 1)]
    assert all(a >= b - 1e-14 for a, b in zip(sig, sig[1:]))


def test_filter_matrix_rejects_bad_arguments():
    with pytest.raises(fedg.FedgError):
        fedg.modalfilter_matrix(3, 1.0, 1.0, 4)
    with pytest.raises(fedg.FedgError):
        fedg.modalfilter_matrix(0, 0.0, 1.0, 4)


def _filter_einsum(mesh, Fh, Fv, f, G):
    """Independent restatement: per element out(i,j,k) = sum Fh(i,a) Fh(j,b) Fv(k,c) (G f)(a,b,c) / G."""
    n1, Np, Ne = mesh.elem.n1, mesh.Np, mesh.Ne
    x = (G[:Np * Ne] * f[:Np * Ne]).reshape(Ne, n1, n1, n1)  # (ke, k, j, i)
    y = np.einsum("ia,jb,kc,ecba->ekji", Fh, Fh, Fv, x)
    return y.reshape(-1) / G[:Np * Ne]


@pytest.mark.parametrize("p", [2, 3])
def test_oracle_filter_matches_tensor_contraction(p):
    case = cases.scaled_case(cases.config1(), 3, 2, 2, porder=p)
    mesh, bs, q = cases.build(case)
    rng = np.random.default_rng(3)
    G = 1.0 + 0.3 * rng.random(mesh.Np * mesh.NeA)
    om = O.OracleMesh(mesh, Gsqrt=G)
    Fh = fedg.modalfilter_matrix(p, 0.0, 0.25, 24)
    Fv = fedg.modalfilter_matrix(p, 2.0 / 3.0, 36.0, 16)
    om.set_modalfilter(Fh, Fv)
    f = [rng.standard_normal(mesh.Np * mesh.NeA) for _ in range(5)]
    ref = [_filter_einsum(mesh, Fh, Fv, a, G) for a in f]
    halo = [a[mesh.Np * mesh.Ne:].copy() for a in f]
    om.modalfilter_apply(f, do_weight_Gsqrt=True)
    n = mesh.Np * mesh.Ne
    for v in range(5):
        assert np.allclose(f[v][:n], ref[v], rtol=0, atol=1e-12 * np.max(np.abs(ref[v])))
        assert np.array_equal(f[v][n:], halo[v])  # the halo slots are not touched
    # without weighting a Gsqrt != 1 makes a difference, with Gsqrt = 1 none
    g = [rng.standard_normal(mesh.Np * mesh.NeA) for _ in range(5)]
    a = [x.copy() for x in g]
    om.modalfilter_apply(a, do_weight_Gsqrt=False)
    ref1 = [_filter_einsum(mesh, Fh, Fv, x, np.ones_like(G)) for x in g]
    for v in range(5):
        assert np.allclose(a[v][:n], ref1[v], rtol=0, atol=1e-12 * np.max(np.abs(ref1[v])))


def test_oracle_step_applies_filter_after_the_stages():
    """A filtered step equals an unfiltered step followed by the filter, bit for bit."""
    case = cases.scaled_case(cases.config1(), 3, 3, 2)
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    F = fedg.modalfilter_matrix(3, *cases.RISING_BUBBLE_FILTER)
    a = [x.copy() for x in q]
    om.heve_update(a, bs, Dx, Dy, case.dt, 1)
    om.set_modalfilter(F, F)
    om.modalfilter_apply(a)
    b = [x.copy() for x in q]
    om.heve_update(b, bs, Dx, Dy, case.dt, 1)
    n = mesh.Np * mesh.Ne
    for v in range(5):
        assert np.array_equal(a[v][:n], b[v][:n])
    om.set_modalfilter(None)


def test_filter_keeps_the_bubble_finite_for_an_hour_prefix():
    """The filtered config-3 run damps what the unfiltered one amplifies:
    after 1500 steps of the p=3 bubble the filtered state's top-mode energy is
    below the unfiltered one's."""
    case = cases.config3()
    mesh, bs, q = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    a = [x.copy() for x in q]
    om.heve_update(a, bs, Dx, Dy, case.dt, 1500)
    F = fedg.modalfilter_matrix(3, *case.modalfilter)
    om.set_modalfilter(F, F)
    b = [x.copy() for x in q]
    om.heve_update(b, bs, Dx, Dy, case.dt, 1500)
    n1, Ne = mesh.elem.n1, mesh.Ne
    x = lgl(3)
    V = np.stack([npleg.legval(x, np.eye(n1)[k]) for k in range(n1)], axis=1)
    Vi = np.linalg.inv(V)

    def top_energy(f):
        c = np.einsum("ia,jb,kc,ecba->ekji", Vi, Vi, Vi, f[:mesh.Np * Ne].reshape(Ne, n1, n1, n1))
        return float(np.sum(c[:, -1, :, :] ** 2) + np.sum(c[:, :, -1, :] ** 2) + np.sum(c[:, :, :, -1] ** 2))

    assert all(np.isfinite(f).all() for f in b)
    assert top_energy(b[2]) < top_energy(a[2])
