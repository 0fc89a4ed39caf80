"""The oracle of the per-column MLP and of the learned HEVI guess's line
layout, pinned on the CPU.

* bf16 rounding against torch's float32 -> bfloat16 conversion (round to
  nearest even), including ties, overflow to infinity and NaN;
* back-propagation (exact arithmetic mode) against central finite
  differences of the loss in float64;
* the BF16 mode against the exact mode (the rounding is small);
* Adam's first step against its closed form (a step of lr against the sign of
  the gradient);
* the line node map against the mesh coordinates: a line's nodes share (x, y)
  and climb in z.
"""
import numpy as np
import pytest
import torch

import fedg
import fedg_oracle_mlp as M
from fedg import cases


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(5000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 5000),
        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 3.4e38, -3.4e38, 0.0, -0.0, 1e-40,
                  np.inf, -np.inf], dtype=np.float32),
    ]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = M.bf16(x)
    assert np.array_equal(got, ref)
    assert np.isnan(M.bf16(np.array([np.nan], dtype=np.float32)))[0]


def _net(seed=1, dims=(6, 9, 2, 4)):
    lay = M.Layout(*dims)
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(lay.nparam) * 0.5
    return lay, p, rng


def test_backprop_matches_finite_differences():
    lay, p, rng = _net()
    X = rng.standard_normal((5, lay.dims[0]))
    T = rng.standard_normal((5, lay.dims[-1]))
    loss, g = M.loss_grad(lay, p, X, T, rnd=False)
    h = 1e-6
    num = np.zeros_like(g)
    for i in range(lay.nparam):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (M.loss_grad(lay, p + e, X, T, rnd=False)[0] - M.loss_grad(lay, p - e, X, T, rnd=False)[0]) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    # a plausible slip in the backward pass (no ReLU mask, transposed dW) would fail here
    assert np.max(np.abs(g)) > 1e-3


def test_bf16_mode_is_close_to_exact():
    lay, p, rng = _net(2, (40, 64, 2, 12))
    X = rng.standard_normal((32, lay.dims[0]))
    T = rng.standard_normal((32, lay.dims[-1]))
    l0, g0 = M.loss_grad(lay, p, X, T, rnd=False)
    l1, g1 = M.loss_grad(lay, p.astype(np.float32), X.astype(np.float32), T, rnd=True)
    assert abs(l1 - l0) <= 3e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 3e-2 * np.linalg.norm(g0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(100).astype(np.float32)
    g = (rng.standard_normal(100) * 10.0 ** rng.integers(-3, 3, 100)).astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    p0 = p.copy()
    M.adam(p, g, m, v, 1, lr=1e-3)
    # m/c1 = g and v/c2 = g^2 at the first step: the step is lr g / (|g| + eps)
    # (float32 parameters near 1 keep the 1e-3 step to about 1e-4 relative)
    assert np.allclose(p0 - p, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-3, atol=0)


def test_line_nodes_follow_the_vertical():
    case = cases.scaled_case(cases.config4(), 3, 2, 3, porder=3)
    mesh, bs, q = cases.build(case)
    idx = M.all_line_nodes(mesh)
    NeZ = mesh.Ne // mesh.Ne2D
    assert idx.shape == (mesh.Ne2D * 16, NeZ * 4)
    assert len(np.unique(idx)) == mesh.Np * mesh.Ne  # every node on exactly one line
    x, y, z = mesh.x[idx], mesh.y[idx], mesh.z[idx]
    assert np.all(x == x[:, :1]) and np.all(y == y[:, :1])
    dz = np.diff(z, axis=1)
    assert np.all(dz >= 0)  # element faces repeat a level (dz = 0), otherwise up
    assert np.all(np.sum(dz == 0, axis=1) == NeZ - 1)


def test_features_layout():
    case = cases.scaled_case(cases.config4(), 2, 2, 2, porder=2)
    mesh, bs, q = cases.build(case)
    rng = np.random.default_rng(5)
    n = mesh.Np * mesh.NeA
    qs = [rng.standard_normal(n) for _ in range(5)]
    Tex = [rng.standard_normal(n) for _ in range(5)]
    pv = [a + 0.1 * rng.standard_normal(n) for a in qs]
    scale = [1.0, 2.0, 3.0, 4.0, 5.0, 0.5, 0.25, 0.125, 2.0, 4.0, 8.0, 0.1, 0.2, 0.3, 0.4]
    lines = [0, 3, 7]
    X, Y = M.features(mesh, qs, Tex, scale, 0.7, 0.35, lines, pv)
    L = 2 * 3
    assert X.shape == (3, 10 * L + 1) and Y.shape == (3, 5 * L)
    for r, l in enumerate(lines):
        nodes = M.line_nodes(mesh, l)
        for v in range(5):
            assert np.allclose(X[r, v * L:(v + 1) * L], qs[v][nodes] / scale[v], rtol=1e-6)
            assert np.allclose(X[r, (5 + v) * L:(6 + v) * L], 0.7 * Tex[v][nodes] / scale[5 + v], rtol=1e-6)
            assert np.allclose(Y[r, v * L:(v + 1) * L], (pv[v][nodes] - qs[v][nodes]) / scale[10 + v], rtol=1e-5,
                               atol=1e-7)
        assert X[r, -1] == np.float32(2.0)
