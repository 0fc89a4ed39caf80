"""GPU tests of the per-column MLP (BF16 matrix-core products, FP32
accumulation, Adam) and of the learned initial iterate of the HEVI Newton
solve, against the CPU oracle (oracle/fedg_oracle_mlp.py).

* On dyadic data (every operand and every partial sum exactly representable)
  the forward pass and every gradient agree bit for bit: a transposed operand,
  a wrong leading dimension or a lost padding column would show.
* On random data (the network of the HEVI guess, 641 -> 256 x3 -> 320) the
  outputs and gradients agree to FP32 accumulation-order rounding, and one
  Adam step matches the oracle's.
* Training on the GPU reduces the loss.
* The Newton solve started from a constant prediction is the oracle's bit for
  bit; from a random network's prediction it meets the same tolerance and
  agrees to it.
"""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
import fedg_oracle_mlp as M
from fedg import cases

pytestmark = pytest.mark.gpu


def gpu_net(dims, params=None, seed=1):
    net = fedg.Mlp(dims[0], dims[1], dims[2], dims[3], seed=seed)
    if params is not None:
        net.set_params(params)
    return net


def fwd(net, X):
    B = X.shape[0]
    dX = fedg.DevField32(X.size, X)
    dY = fedg.DevField32(B * net.n_out)
    net.forward(dX, dY, B)
    return dY.download().reshape(B, net.n_out)


def loss_grad(net, X, T):
    B = X.shape[0]
    dX, dT = fedg.DevField32(X.size, X), fedg.DevField32(T.size, T)
    loss = net.loss_grad(dX, dT, B)
    return loss, net.grads()


def dyadic(rng, shape, vals):
    return rng.choice(np.asarray(vals, dtype=np.float32), size=shape)


def test_mlp_exact_on_dyadic_data():
    dims = (20, 24, 1, 4)  # widths that are not multiples of 16: padded products
    lay = M.Layout(*dims)
    rng = np.random.default_rng(4)
    p = dyadic(rng, lay.nparam, [-1.0, -0.5, 0.0, 0.5, 1.0])
    for l in range(lay.nl):
        p[lay.boff[l]:lay.boff[l] + lay.dims[l + 1]] = dyadic(rng, lay.dims[l + 1], [-0.25, 0.0, 0.25])
    B = 128  # B * n_out = 512: the loss scale 2/n is a power of two
    X = dyadic(rng, (B, dims[0]), [-1.0, -0.75, -0.5, -0.25, 0.0, 0.25, 0.5, 0.75, 1.0])
    net = gpu_net(dims, p)
    Y = fwd(net, X)
    Yo = M.forward(lay, p, X, rnd=True)[-1]
    assert np.array_equal(Y, Yo)
    T = (np.round(Yo) + rng.integers(-1, 2, Yo.shape)).astype(np.float32)
    loss, g = loss_grad(net, X, T)
    lo, go = M.loss_grad(lay, p, X, T, rnd=True)
    assert loss == pytest.approx(lo, rel=1e-6)
    assert np.array_equal(g, go.astype(np.float32)), np.max(np.abs(g - go))
    net.close()


def test_mlp_guess_network_random_data():
    dims = (641, 256, 3, 320)
    lay = M.Layout(*dims)
    net = gpu_net(dims, seed=7)
    p = net.params()
    rng = np.random.default_rng(8)
    B = 200
    X = rng.standard_normal((B, dims[0])).astype(np.float32)
    T = (0.1 * rng.standard_normal((B, dims[-1]))).astype(np.float32)
    Y = fwd(net, X)
    Yo = M.forward(lay, p, X, rnd=True)[-1]
    assert np.linalg.norm(Y - Yo) <= 1e-4 * np.linalg.norm(Yo)
    loss, g = loss_grad(net, X, T)
    lo, go = M.loss_grad(lay, p, X, T, rnd=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    for l in range(lay.nl):
        w = slice(lay.woff[l], lay.woff[l] + lay.dims[l] * lay.dims[l + 1])
        b = slice(lay.boff[l], lay.boff[l] + lay.dims[l + 1])
        for s in (w, b):
            assert np.linalg.norm(g[s] - go[s]) <= 2e-3 * np.linalg.norm(go[s]), f"layer {l}"
    # one Adam step against the oracle's with the GPU's gradient
    net.adam_step(1e-3)
    p1 = net.params()
    q = p.copy()
    M.adam(q, g.astype(np.float32), np.zeros_like(q), np.zeros_like(q), 1, lr=1e-3)
    assert np.allclose(p1, q, rtol=0, atol=1e-6)
    net.close()


def test_mlp_training_reduces_loss():
    dims = (96, 128, 2, 32)
    rng = np.random.default_rng(9)
    A = (rng.standard_normal((dims[0], dims[-1])) / np.sqrt(dims[0])).astype(np.float32)
    B = 2048
    X = rng.standard_normal((B, dims[0])).astype(np.float32)
    T = (X @ A).astype(np.float32)
    net = gpu_net(dims, seed=3)
    dX, dT = fedg.DevField32(X.size, X), fedg.DevField32(T.size, T)
    l0 = net.loss_grad(dX, dT, B)
    for _ in range(300):
        net.loss_grad(dX, dT, B)
        net.adam_step(2e-3)
    l1 = net.loss_grad(dX, dT, B)
    print(f"training loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 0.1 * l0
    net.close()


# ---------------------------------------------------------------------------
# learned initial iterate of the HEVI Newton solve

def hevi_setup(p=3):
    case = cases.Case("hevi p3", p, 4, 3, 4, 12.0e3, 9.0e3, 10.0e3, 2.0, 1, [], base="constBVFreq",
                      scheme="IMEX_ARK324")
    mesh = fedg.BoxMesh(case.porder, case.NeX, case.NeY, case.NeZ, (0.0, case.Lx), (0.0, case.Ly), (0.0, case.Lz))
    bs = cases.base_state(mesh, case.base)
    om = O.OracleMesh(mesh)
    dm = fedg.DeviceMesh(mesh)
    return case, mesh, bs, om, dm


def state(mesh, seed, w=0.5):
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    q = [mesh.field() for _ in range(5)]
    q[0][:n] = 1e-3 * rng.standard_normal(n)
    q[1][:n] = 0.3 * rng.standard_normal(n)
    q[2][:n] = w * rng.standard_normal(n)
    q[3][:n] = 10.0 + rng.standard_normal(n)
    q[4][:n] = rng.standard_normal(n)
    for f in q:
        mesh.exchange(f)
    return q


def dev(a):
    return fedg.DevField(a.size, a)


SCALE = [1e-3, 0.3, 1.0, 10.0, 1.0, 1e-5, 1e-3, 0.05, 0.05, 0.05, 1e-4, 1e-2, 0.1, 0.1, 0.1]


def run_guess(dm, mesh, bs, q, q0, impl_fac, dt, net, Tex=None, max_it=10, rtol=1e-10):
    out = [dev(mesh.field()) for _ in range(5)]
    tex = None if Tex is None else [dev(a) for a in Tex]
    it, resid, fb = fedg.hevi_cal_vi_guess(dm, out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]),
                                           dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(q0[0]), dev(q0[3]), dev(q0[4]),
                                           dev(q0[2]), dev(q0[1]), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot),
                                           impl_fac, dt, max_it, rtol, net, SCALE, tex)
    g = [a.download() for a in out]
    return [g[0], g[4], g[3], g[1], g[2]], it, resid, fb


@pytest.fixture
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def test_guess_constant_prediction_bitwise(portable_pow):
    """All weights zero and dyadic output biases: the prediction is exact on both sides."""
    case, mesh, bs, om, dm = hevi_setup()
    L = (mesh.Ne // mesh.Ne2D) * mesh.elem.n1
    dims = (10 * L + 1, 32, 1, 5 * L)
    lay = M.Layout(*dims)
    p = np.zeros(lay.nparam, dtype=np.float32)
    rng = np.random.default_rng(12)
    last = lay.nl - 1
    p[lay.boff[last]:lay.boff[last] + dims[-1]] = dyadic(rng, dims[-1], [-0.5, -0.25, 0.0, 0.25, 0.5])
    net = gpu_net(dims, p)
    q = state(mesh, 4, w=2.0)
    q0 = state(mesh, 5, w=2.0)
    Tex = [0.01 * a for a in state(mesh, 6)]
    got, it_g, res_g, fb_g = run_guess(dm, mesh, bs, q, q0, 1.5, case.dt, net, Tex)
    ref, it_o, res_o, fb_o = M.cal_vi_guess(om, lay, p, q, q0, bs, 1.5, case.dt, Tex, SCALE, 10, 1e-10)
    print(f"iterations {it_g}, fallback lines {fb_g} of {mesh.Ne2D * mesh.elem.n1 ** 2}")
    assert fb_g == fb_o and it_g == it_o
    n = mesh.Np * mesh.Ne
    for v in range(5):
        assert np.array_equal(got[v][:n], ref[v][:n]), fedg.VARS[v]
    net.close()


def test_guess_random_network_meets_tolerance():
    case, mesh, bs, om, dm = hevi_setup()
    L = (mesh.Ne // mesh.Ne2D) * mesh.elem.n1
    dims = (10 * L + 1, 64, 2, 5 * L)
    net = gpu_net(dims, seed=21)
    q = state(mesh, 7, w=2.0)
    q0 = state(mesh, 8, w=2.0)
    got, it_g, res_g, fb_g = run_guess(dm, mesh, bs, q, q0, 1.5, case.dt, net)
    ref, it_o, res_o = om.hevi_cal_vi(q, q0, bs, 1.5, case.dt, max_it=10, rtol=1e-10)
    print(f"iterations guess {it_g} plain {it_o}, fallback lines {fb_g}")
    # the guess path's tolerance is relative to ||G(q^n)||, the plain one's to
    # the first iteration's residual (G with the uv update): close, not equal
    assert it_g <= 10 and res_g[-1] <= 2e-10 * res_o[0]
    n = mesh.Np * mesh.Ne
    for v in range(5):
        scale = np.max(np.abs(ref[v][:n])) + 1e-300
        assert np.max(np.abs(got[v][:n] - ref[v][:n])) <= 1e-6 * scale, fedg.VARS[v]
    net.close()
