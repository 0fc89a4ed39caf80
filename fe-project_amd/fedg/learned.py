"""Learned initial iterate of the HEVI Newton solve: harvest, training, use.

No counterpart in the reference, whose Newton loop starts from q^n and stops
after one iteration (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:858-917).
A per-line network (``Mlp``: BF16 matrix-core products, FP32 accumulation,
Adam, all on the GPU) maps a vertical node line's stage state, the previous
stage's explicit tendency and impl_fac/dt to the converged Newton increment.
The HEVI stepper

* harvests (features, converged increment) pairs from its own solves
  (``HeviStepper.harvest``),
* starts every line from the prediction unless the prediction's residual is
  larger than q^n's (the fallback), and
* keeps the Newton tolerance relative to ||G(q^n)||, so that results agree
  with the method without the network to that tolerance.

Scales: the harvest runs with unit scales; the per-variable state scale s,
tendency-feature scale t and increment scale d are the RMS of the harvested
features and targets, and the stored rows are rescaled with them before
training.
"""
from __future__ import annotations

import numpy as np

from .dyncore import DevField32, HeviStepper, Mlp

UNIT = [1.0] * 15


def line_length(mesh) -> int:
    return (mesh.Ne // mesh.Ne2D) * mesh.elem.n1


def harvest(dm, bs, Dx, Dy, dt, scheme, states, nsteps, per_stage, rtol=1e-10, max_it=8, seed=1, warm=0):
    """Run the plain stepper from every initial state for `nsteps` steps
    (after `warm` steps not harvested) and collect `per_stage` lines per
    implicit stage.  Returns host arrays X (rows, 10 L + 1), Y (rows, 5 L) at
    unit scales."""
    mesh = dm.mesh
    L = line_length(mesh)
    cap_per_run = per_stage * nsteps * 4
    Xs, Ys = [], []
    for r, q in enumerate(states):
        st = HeviStepper(dm, bs, Dx, Dy, dt, scheme=scheme, max_it=max_it, rtol=rtol)
        st.set_state(q)
        if warm:
            st.run(warm)
        X = DevField32(cap_per_run * (10 * L + 1))
        Y = DevField32(cap_per_run * 5 * L)
        st.set_guess(None, UNIT)
        st.harvest(X, Y, cap_per_run, per_stage, seed=seed + 7919 * r)
        st.run(nsteps)
        rows = st.harvest_count()
        st.harvest(None, None, 0, 1)
        Xs.append(X.download(rows * (10 * L + 1)).reshape(rows, 10 * L + 1))
        Ys.append(Y.download(rows * 5 * L).reshape(rows, 5 * L))
        X.free()
        Y.free()
        st.close()
    return np.concatenate(Xs), np.concatenate(Ys)


def scales_of(X, Y, L, floor=1e-30):
    """Per-variable RMS of the state features (s), the tendency features (t) and the increments (d)."""
    rms = lambda a: max(float(np.sqrt(np.mean(a.astype(np.float64) ** 2))), floor)  # noqa: E731
    s = [rms(X[:, v * L:(v + 1) * L]) for v in range(5)]
    t = [rms(X[:, (5 + v) * L:(6 + v) * L]) for v in range(5)]
    d = [rms(Y[:, v * L:(v + 1) * L]) for v in range(5)]
    return s + t + d


def normalise(X, Y, scale, L):
    """Rescale unit-scale rows to the network's inputs and targets."""
    Xn = X.astype(np.float64).copy()
    Yn = Y.astype(np.float64).copy()
    for v in range(5):
        Xn[:, v * L:(v + 1) * L] /= scale[v]
        Xn[:, (5 + v) * L:(6 + v) * L] /= scale[5 + v]
        Yn[:, v * L:(v + 1) * L] /= scale[10 + v]
    return Xn.astype(np.float32), Yn.astype(np.float32)


def train(net: Mlp, X, Y, steps, batch=512, lr=1e-3, seed=1, log_every=0):
    """Adam on contiguous minibatches of the (shuffled) rows; returns the losses."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(X.shape[0])
    X, Y = X[perm], Y[perm]
    rows = X.shape[0]
    batch = min(batch, rows)
    dX = DevField32(X.size, X)
    dY = DevField32(Y.size, Y)
    n_in, n_out = X.shape[1], Y.shape[1]
    import ctypes as C
    losses = []
    for k in range(steps):
        r0 = int(rng.integers(0, rows - batch + 1))
        px = C.c_void_p(dX.ptr.value + 4 * r0 * n_in)
        py = C.c_void_p(dY.ptr.value + 4 * r0 * n_out)
        losses.append(net.loss_grad(px, py, batch))
        net.adam_step(lr)
        if log_every and (k % log_every == 0 or k == steps - 1):
            print(f"  step {k}: loss {losses[-1]:.4e}", flush=True)
    dX.free()
    dY.free()
    return losses
