"""Writes tests/golden/config2_step200.npz: the CPU oracle's state after the
full 64x64x16, p=7 warm-bubble run (200 RK4 steps, dt 0.04 s), sampled at
fixed nodes, plus the L2 norm of every prognostic field.

Two runs: the oracle with the portable pow (the GPU path must match these
samples bit for bit) and with the C library's pow (the reference Fortran's **
on glibc; the GPU path is compared with these at the documented floor).  The
full-field distance between the two runs (relative L2 and Linf of every
prognostic, over all nodes) is stored as well: the floor between two correct
implementations whose pow differs in the last bit.

Takes about 20 minutes on 8 cores; the samples (4096 nodes x 5 fields x 2
runs) are small enough to commit.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "fe-project_amd"), os.path.join(ROOT, "oracle")]

import fedg_oracle as O  # noqa: E402
from fedg import cases  # noqa: E402

NSAMPLE = 4096


def main():
    case = cases.config2()
    mesh, bs, q0 = cases.build(case)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(424242)
    idx = np.sort(rng.choice(n, NSAMPLE, replace=False))
    # plus every node of the element holding the first bubble's centre
    out = {"idx": idx, "nsteps": np.int64(case.nsteps), "dt": np.float64(case.dt)}
    final = {}
    for tag, mode in (("portable", O.POW_PORTABLE), ("libm", O.POW_LIBM)):
        q = [a.copy() for a in q0]
        t0 = time.time()
        with O.pow_mode(mode):
            om.heve_update(q, bs, Dx, Dy, case.dt, case.nsteps)
        print(f"{tag}: {case.nsteps} steps in {time.time() - t0:.0f} s", flush=True)
        out[f"{tag}_samples"] = np.stack([a[idx] for a in q])
        out[f"{tag}_norms"] = np.array([np.linalg.norm(a[:n]) for a in q])
        out[f"{tag}_sums"] = np.array([np.sum(a[:n]) for a in q])
        final[tag] = [a[:n] for a in q]
    # full-field distance of the portable-pow run from the C-library-pow run
    pf, lf = final["portable"], final["libm"]
    out["pow_floor_rel_l2"] = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(pf, lf)])
    out["pow_floor_rel_linf"] = np.array([np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(pf, lf)])
    print("full-field relative L2 (portable vs libm pow):", out["pow_floor_rel_l2"], flush=True)
    print("full-field relative Linf:", out["pow_floor_rel_linf"], flush=True)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_step200.npz"), **out)


if __name__ == "__main__":
    main()
