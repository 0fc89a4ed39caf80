"""The learned initial iterate of the HEVI Newton solve, end to end on the GPU:
harvest (features, converged increment) pairs from the stepper's own solves
on seeded inertia-gravity-wave runs (theta' amplitude 0.005-0.05 K, centre and
mean wind 10-30 m/s drawn from a seed), train the per-line network with the
library's BF16 forward and backward kernels and Adam, then run a held-out case
with and without the network at the same Newton tolerance.

Bar: the two runs agree to the tolerance; the trained network starts the lines
closer to the solution than q^n (the measured gain of the network: the
residual norm at the starts it provides against the one at q^n); it never
needs more Newton iterations.  The iteration counts, the fallback rate and the
residual ratio are printed.
"""
import time

import numpy as np
import pytest

import fedg
from fedg import cases, learned

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def igw_states(mesh, bs, seeds, Lx):
    out = []
    for sd in seeds:
        rng = np.random.default_rng(sd)
        out.append(cases.igw_state(mesh, bs, U=float(rng.uniform(10.0, 30.0)),
                                   dtheta=float(rng.uniform(0.005, 0.05)),
                                   xc=float(rng.uniform(0.2, 0.8) * Lx)))
    return out


@pytest.mark.parametrize("dx,dt", [(1.25e3, 0.05), (5.0e3, 0.2)], ids=["isotropic", "wide-elements"])
def test_learned_guess_end_to_end(dx, dt):
    """isotropic: the 1.25 km elements of the 240x16x8 channel at its step;
    wide-elements: 5 km x 1.25 km elements, whose horizontal acoustic limit
    allows a 4x longer step and so a stiffer vertical implicit problem."""
    base = cases.config4()
    case = cases.scaled_case(base, 8, 2, 8)  # p = 7, lines of 64 nodes
    case.Lx, case.Ly, case.dt = 8 * dx, 2 * dx, dt
    mesh, bs, _ = cases.build(case)
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    L = learned.line_length(mesh)
    assert L == 64
    t0 = time.time()
    X, Y = learned.harvest(dm, bs, Dx, Dy, case.dt, case.scheme, igw_states(mesh, bs, [11, 12, 13, 14], case.Lx),
                           nsteps=6, per_stage=256, rtol=RTOL, warm=2)
    scale = learned.scales_of(X, Y, L)
    Xn, Yn = learned.normalise(X, Y, scale, L)
    t1 = time.time()
    net = fedg.Mlp(10 * L + 1, 256, 3, 5 * L, seed=5)
    losses = learned.train(net, Xn, Yn, steps=1500, batch=512, lr=1e-3)
    t2 = time.time()
    print(f"\nharvested {X.shape[0]} rows in {t1 - t0:.1f} s; trained 1500 steps in {t2 - t1:.1f} s, "
          f"loss {np.mean(losses[:20]):.3e} -> {np.mean(losses[-20:]):.3e}")
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:20])

    (q,) = igw_states(mesh, bs, [99], case.Lx)
    res = {}
    for label, use in (("plain", False), ("learned", True)):
        st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=8, rtol=RTOL)
        st.set_state(q)
        st.run(2)
        if use:
            st.set_guess(net, scale)
        it0, calls0 = st.newton_stats()
        ta = time.time()
        st.run(6)
        tb = time.time()
        it, calls = st.newton_stats()
        res[label] = dict(state=st.get_state(), it=(it - it0) / (calls - calls0), ms=(tb - ta) / 6 * 1e3,
                          stats=st.guess_stats(), resid=st.guess_residuals())
        st.close()
    lines, fb = res["learned"]["stats"]
    rp, rs = res["learned"]["resid"]
    print(f"Newton iterations per solve: plain {res['plain']['it']:.2f}, learned {res['learned']['it']:.2f}; "
          f"fallback {fb}/{lines} lines ({fb / max(lines, 1):.1%}); residual at the starts / at q^n "
          f"{rs / rp:.3e}; ms/step plain {res['plain']['ms']:.1f}, learned {res['learned']['ms']:.1f}")
    for v, name in enumerate(fedg.VARS):
        a, b = res["learned"]["state"][v], res["plain"]["state"][v]
        err = np.linalg.norm(a - b)
        print(f"  {name}: |learned - plain| / |plain| = {err / (np.linalg.norm(b) + 1e-300):.2e}")
        assert err <= 1e-8 * np.linalg.norm(b) + 1e-12 * np.sqrt(b.size)
    assert rs < rp
    assert res["learned"]["it"] <= res["plain"]["it"]
    net.close()
