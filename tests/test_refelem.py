"""Reference-element operators built by the C++ builder (host only, no GPU).

Pinned the way FElib's own unit tests pin them:
  - LGL nodes / weights: the closed forms of FElib/test/FE/polynomial/
    test_polynomial.f90:8-45 (N <= 4) at its tolerance 5e-15, and for
    N = 5..7 the roots of P'_N found independently by numpy;
  - D1D: exact derivative of x^m, m <= p (test_element_operation_hexahedral.f90
    gen_dat / gen_grad: d/dx of 4x^p + 3y^p + 2z^p), zero row sums;
  - Lift_mat: equal to the dense M^-1 E (3D mass matrix inverse times the
    face mass matrices) assembled in numpy from an independent Vandermonde
    matrix (test_element_operation_hexahedral.f90:100-103, :134-135);
  - IntrpMat_VPOrdM1: identity on polynomials of degree <= p-1 in z, removes
    the degree-p Legendre mode (ELM/scale_element_operation_tensorprod3D.F90:
    5560-5563).
"""
import numpy as np
import numpy.polynomial.legendre as L
import pytest

from fedg.mesh import RefElem

CLOSED = {
    1: ([-1.0, 1.0], [1.0, 1.0]),
    2: ([-1.0, 0.0, 1.0], [1 / 3, 4 / 3, 1 / 3]),
    3: ([-1.0, -np.sqrt(1 / 5), np.sqrt(1 / 5), 1.0], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    4: ([-1.0, -np.sqrt(21.0) / 7, 0.0, np.sqrt(21.0) / 7, 1.0], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
}


def F(a, n1):
    return np.asarray(a).reshape((n1, n1), order="F")


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_lgl_closed_forms(N):
    e = RefElem(N)
    x, w = CLOSED[N]
    assert np.max(np.abs(e.x1d - x)) <= 5e-15
    assert np.max(np.abs(e.w1d - w)) <= 5e-15


@pytest.mark.parametrize("N", [5, 6, 7])
def test_lgl_vs_numpy_roots(N):
    e = RefElem(N)
    c = np.zeros(N + 1)
    c[N] = 1.0
    inner = np.sort(np.real(L.legroots(L.legder(c))))
    x = np.concatenate([[-1.0], inner, [1.0]])
    w = 2.0 / (N * (N + 1) * L.legval(x, c) ** 2)
    assert np.max(np.abs(e.x1d - x)) <= 5e-15
    assert np.max(np.abs(e.w1d - w)) <= 5e-15
    assert abs(e.w1d.sum() - 2.0) <= 1e-14


@pytest.mark.parametrize("N", [1, 3, 5, 7])
def test_D1D_exact_on_polynomials(N):
    e = RefElem(N)
    n1 = N + 1
    D = F(e.D1D, n1)
    x = e.x1d
    for m in range(N + 1):
        du = D @ x ** m
        ex = m * x ** (m - 1) if m > 0 else np.zeros(n1)
        assert np.max(np.abs(du - ex)) <= 2e-12 * max(1, m * m), m
    # negative-sum diagonal (CMN/scale_polynomial.F90:157): rows sum to zero exactly
    # up to the rounding of that one sum
    assert np.max(np.abs(D.sum(axis=1))) <= 1e-12


@pytest.mark.parametrize("N", [3, 7])
def test_3d_gradient_of_reference_test_field(N):
    """gen_dat / gen_grad of test_element_operation_hexahedral.f90:203-234."""
    e = RefElem(N)
    n1 = N + 1
    D = F(e.D1D, n1)
    x = e.x1d
    X1, X2, X3 = np.meshgrid(x, x, x, indexing="ij")  # node (i,j,k), i fastest in Fortran order
    dat = 4 * X1 ** N + 3 * X2 ** N + 2 * X3 ** N
    dx = np.einsum("il,ljk->ijk", D, dat)
    dy = np.einsum("jl,ilk->ijk", D, dat)
    dz = np.einsum("kl,ijl->ijk", D, dat)
    for got, ans in ((dx, 4 * N * X1 ** (N - 1)), (dy, 3 * N * X2 ** (N - 1)), (dz, 2 * N * X3 ** (N - 1))):
        assert np.max(np.abs(got - ans)) <= 1e-11


def vandermonde(x, N):
    """Orthonormal Legendre Vandermonde V(i,n) = sqrt((2n+1)/2) P_n(x_i)."""
    V = np.zeros((len(x), N + 1))
    for n in range(N + 1):
        c = np.zeros(N + 1)
        c[n] = 1.0
        V[:, n] = L.legval(x, c) * np.sqrt((2 * n + 1) / 2.0)
    return V


@pytest.mark.parametrize("N", [2, 3, 7])
def test_lift_equals_dense_invM_E(N):
    e = RefElem(N)
    n1, Np, Nfp = N + 1, (N + 1) ** 3, (N + 1) ** 2
    V = vandermonde(e.x1d, N)
    M1 = np.linalg.inv(V @ V.T)
    M3 = np.kron(M1, np.kron(M1, M1))  # index i + n1*(j + n1*k): kron(Mk, Mj, Mi)
    idx = np.arange(Np).reshape((n1, n1, n1), order="F")  # idx[i,j,k]
    # faces in the reference order: y=-1, x=+1, y=+1, x=-1, z=-1, z=+1; the
    # face-local node order follows the tangential indices (x or y first, then z)
    faces = [idx[:, 0, :], idx[N, :, :], idx[:, N, :], idx[0, :, :], idx[:, :, 0], idx[:, :, N]]
    M2 = np.kron(M1, M1)
    E = np.zeros((Np, 6 * Nfp))
    for f, fn in enumerate(faces):
        nodes = fn.reshape(-1, order="F")
        for a, pa in enumerate(nodes):
            for b, pb in enumerate(nodes):
                E[pa, f * Nfp + b] = M2[a, b]
    Lift_dense = np.linalg.solve(M3, E)
    # Lift_mat(p, f) is the diagonal coefficient: Lift_dense(p, face node of p's line)
    Lm = e.Lift_mat.reshape((Np, 6), order="F")
    for f, fn in enumerate(faces):
        nodes = fn.reshape(-1, order="F")
        for p in range(Np):
            i, j, k = p % n1, (p // n1) % n1, p // (n1 * n1)
            tang = {0: (i, k), 1: (j, k), 2: (i, k), 3: (j, k), 4: (i, j), 5: (i, j)}[f]
            a = tang[0] + n1 * tang[1]
            row = Lift_dense[p, f * Nfp:(f + 1) * Nfp]
            assert abs(row[a] - Lm[p, f]) <= 1e-11 * max(1.0, abs(row[a]))
            off = np.delete(row, a)
            assert np.max(np.abs(off)) <= 1e-11
    # closed form at a face's own node: (V V^T)(0,0) = sum_n (2n+1)/2 P_n(-1)^2 = (N+1)^2/2
    assert abs(Lm[idx[0, 0, 0], 3] - (N + 1) ** 2 / 2) <= 1e-12 * (N + 1) ** 2


@pytest.mark.parametrize("N", [3, 7])
def test_vfilter_pm1(N):
    e = RefElem(N)
    n1 = N + 1
    # IntrpMat_VPOrdM1_tr(l, k): drho(k) = sum_l q(l) * Itr(l, k)
    Itr = F(e.IntrpMat_VPOrdM1_tr, n1)
    x = e.x1d
    for m in range(N):
        q = x ** m
        assert np.max(np.abs(q @ Itr - q)) <= 1e-12
    c = np.zeros(N + 1)
    c[N] = 1.0
    PN = L.legval(x, c)
    assert np.max(np.abs(PN @ Itr)) <= 1e-12
