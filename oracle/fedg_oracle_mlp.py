"""CPU restatement of the per-column MLP and of the learned HEVI initial guess.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

No counterpart in the reference: its HEVI Newton loop starts from q^n
(FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:858-917).  What is restated
here is the network the GPU library trains and applies:

* dense layers n_in -> H (x n_hidden) -> n_out, ReLU on the hidden layers;
  parameters stored as every layer's W (N x K, row-major) then every bias;
* every product takes BF16 operands (round to nearest even) and accumulates
  in FP32 (float64 here: the GPU's FP32 accumulation order is not restated,
  which is what the tests' tolerances allow for);
* backward: dZ = dH * (H > 0), dW = bf16(dZ)^T bf16(X), db = sum(dZ) in
  FP32, dX = bf16(dZ) bf16(W);
* mean squared error over all outputs, gradient 2 (Y - T) / n;
* Adam with bias correction.

With ``bf16=False`` every rounding is skipped, which is the exact network the
finite-difference checks pin.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & np.uint64(0x7F800000)) == np.uint64(0x7F800000)  # inf and NaN
    r = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) & np.uint64(0xFFFF)
    quiet = np.where((u & np.uint64(0xFFFF)) != 0, np.uint64(0x40), np.uint64(0))
    r = np.where(special, (u >> np.uint64(16)) | quiet, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


class Layout:
    """Parameter offsets of a network n_in -> n_hidden (x n_hidden_layers) -> n_out."""

    def __init__(self, n_in, n_hidden, n_hidden_layers, n_out):
        self.dims = [n_in] + [n_hidden] * n_hidden_layers + [n_out]
        self.nl = len(self.dims) - 1
        off = 0
        self.woff = []
        for l in range(self.nl):
            self.woff.append(off)
            off += self.dims[l + 1] * self.dims[l]
        self.boff = []
        for l in range(self.nl):
            self.boff.append(off)
            off += self.dims[l + 1]
        self.nparam = off

    def W(self, p, l):
        return p[self.woff[l]:self.woff[l] + self.dims[l + 1] * self.dims[l]].reshape(self.dims[l + 1], self.dims[l])

    def b(self, p, l):
        return p[self.boff[l]:self.boff[l] + self.dims[l + 1]]


def _mm(A, B, rnd):
    """A @ B with BF16 operands (rnd) and float64 accumulation."""
    if rnd:
        A, B = bf16(A), bf16(B)
    return A.astype(np.float64) @ B.astype(np.float64)


def forward(lay: Layout, p, X, rnd=True):
    """Activations [X, H1, ..., Y] (float64 arrays; FP32 values when rnd)."""
    acts = [np.asarray(X, dtype=np.float32 if rnd else np.float64)]
    for l in range(lay.nl):
        Z = _mm(acts[-1], lay.W(p, l).T, rnd) + lay.b(p, l).astype(np.float64)
        if l < lay.nl - 1:
            Z = np.maximum(Z, 0.0)
        acts.append(Z.astype(np.float32) if rnd else Z)
    return acts


def loss_grad(lay: Layout, p, X, T, rnd=True):
    """(mean squared error, gradient of every parameter in the layout's order)."""
    acts = forward(lay, p, X, rnd)
    Y = acts[-1].astype(np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    g = np.zeros(lay.nparam)
    dZ = 2.0 * (Y - T) / n
    if rnd:
        dZ = dZ.astype(np.float32)
    for l in range(lay.nl - 1, -1, -1):
        if l < lay.nl - 1:
            dZ = np.where(acts[l + 1] > 0, dZ, 0.0)
            if rnd:
                dZ = dZ.astype(np.float32)
        g[lay.boff[l]:lay.boff[l] + lay.dims[l + 1]] = dZ.astype(np.float64).sum(axis=0)
        g[lay.woff[l]:lay.woff[l] + lay.dims[l + 1] * lay.dims[l]] = _mm(dZ.T, acts[l], rnd).ravel()
        if l > 0:
            dX = _mm(dZ, lay.W(p, l), rnd)
            dZ = dX.astype(np.float32) if rnd else dX
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of p (float32 arrays updated in place); step counts from 1."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    p -= lr * (m / c1) / (np.sqrt(v / c2) + eps)


# ---------------------------------------------------------------------------
# Learned initial iterate of the HEVI Newton solve on a BoxMesh.
# A vertical node line is (ke2D, ij): L = NeZ (p+1) nodes, node t = kz (p+1) + k
# at element ke2D + kz Ne2D, element node ij + Nfp k.

def line_nodes(mesh, line):
    """Flat node indices (into an (Np,NeA) field) of one line, in line order."""
    N = mesh.elem.n1
    Nfp = N * N
    NeZ = mesh.Ne // mesh.Ne2D
    ke2D, ij = divmod(int(line), Nfp)
    kz = np.repeat(np.arange(NeZ), N)
    k = np.tile(np.arange(N), NeZ)
    return mesh.Np * (ke2D + kz * mesh.Ne2D) + ij + Nfp * k


def all_line_nodes(mesh):
    """(nlines, L) node indices of every line, line = ke2D Nfp + ij."""
    nlines = mesh.Ne2D * mesh.elem.n1 ** 2
    return np.stack([line_nodes(mesh, l) for l in range(nlines)])


def features(mesh, qs, Tex, scale, impl_fac, dt, lines, pv=None):
    """Network inputs (and targets when pv is given) of the given lines.

    qs, Tex, pv: 5 fields (VID order); Tex may be None (first stage).
    scale: 5 state, 5 tendency-feature and 5 increment scales."""
    s, tf, d = np.asarray(scale[:5]), np.asarray(scale[5:10]), np.asarray(scale[10:15])
    idx = np.stack([line_nodes(mesh, l) for l in lines])
    L = idx.shape[1]
    X = np.zeros((len(lines), 10 * L + 1))
    Y = np.zeros((len(lines), 5 * L)) if pv is not None else None
    for v in range(5):
        X[:, v * L:(v + 1) * L] = qs[v][idx] / s[v]
        if Tex is not None:
            X[:, (5 + v) * L:(6 + v) * L] = impl_fac * Tex[v][idx] / tf[v]
        if pv is not None:
            Y[:, v * L:(v + 1) * L] = (pv[v][idx] - qs[v][idx]) / d[v]
    X[:, 10 * L] = impl_fac / dt
    return X.astype(np.float32), (None if Y is None else Y.astype(np.float32))


def line_residual2(mesh, G):
    """Per-line sum of squares of the residual G (5 fields, VID order)."""
    idx = all_line_nodes(mesh)
    return sum((g[idx] ** 2).sum(axis=1) for g in G)


def cal_vi_guess(om, lay, params, q, var0, bs, impl_fac, dt, Tex, scale, max_it, rtol):
    """Implicit tendencies with the learned initial iterate.

    The network's prediction pv = q* + d Y starts the Newton iteration on every
    line whose residual at pv is not larger than at q^n; the other lines start
    from q^n; the tolerance is relative to ||G(q^n)||.  Returns (tendencies in
    VID order, iterations, residual history, number of fallback lines)."""
    m = om.mesh
    n = m.Np * m.Ne
    v0 = [np.array(a[:n]) for a in var0]
    qv = [np.array(a[:n]) for a in q]
    G0 = om.hevi_residual([a.copy() for a in v0], qv, v0, bs, impl_fac, dt)
    r_plain = line_residual2(m, G0)
    r0 = float(np.sqrt(r_plain.sum()))
    nlines = m.Ne2D * m.elem.n1 ** 2
    X, _ = features(m, q, Tex, scale, impl_fac, dt, range(nlines))
    Y = forward(lay, params, X, rnd=True)[-1].astype(np.float64)
    idx = all_line_nodes(m)
    L = idx.shape[1]
    pv = [a.copy() for a in v0]
    for v in range(5):
        pv[v][idx] = qv[v][idx] + scale[10 + v] * Y[:, v * L:(v + 1) * L]
    Gg = om.hevi_residual([a.copy() for a in pv], qv, v0, bs, impl_fac, dt)
    r_guess = line_residual2(m, Gg)
    fb = ~(r_guess <= r_plain)
    for v in range(5):
        pv[v][idx[fb]] = v0[v][idx[fb]]
    it, resid = om.hevi_newton(pv, qv, v0, bs, impl_fac, dt, max_it, rtol, r0_ref=r0)
    tend = [(pv[v] - qv[v]) / impl_fac for v in range(5)]
    return tend, it, resid, int(fb.sum())
