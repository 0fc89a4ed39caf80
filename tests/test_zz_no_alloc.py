"""fedg_hevi_cal_vi_noalloc (the implicit HEVI solve on the library's fixed
device workspace, no allocation in the call) against fedg_hevi_cal_vi_newton
(its hipMalloc'd workspace), through the C ABI, bit for bit.

Two calls in a row share the workspace: the first on a p = 7 column stack
(NeZ = 8), the second on a smaller p = 3 mesh, which finds the first call's
contents in the workspace.  A mesh whose workspace exceeds the fixed size is
rejected with FEDG_ERR_UNSUPPORTED before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

import fedg
from fedg import cases
from fedg._lib import LIB

pytestmark = pytest.mark.gpu

FEDG_ERR_UNSUPPORTED = 4


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


def solve(dm, mesh, bs, q, q0, impl_fac, dt, max_it, rtol, noalloc):
    out = [dev(mesh.field()) for _ in range(5)]
    it, resid = fedg.hevi_cal_vi(dm, out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]),
                                 dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(q0[0]), dev(q0[3]), dev(q0[4]),
                                 dev(q0[2]), dev(q0[1]), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot), impl_fac, dt,
                                 max_it=max_it, rtol=rtol, noalloc=noalloc)
    return [a.download() for a in out], it, resid


def compare(p, NeX, NeY, NeZ, seed, max_it, rtol):
    print(f"no-alloc cal_vi p={p} {NeX}x{NeY}x{NeZ} max_it={max_it}", flush=True)
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, 1.25e3 * NeX), (0.0, 1.25e3 * NeY), (0.0, 1.25e3 * NeZ))
    bs = cases.base_state(mesh, "constBVFreq")
    dm = fedg.DeviceMesh(mesh)
    q = state(mesh, seed, w=2.0)
    q0 = state(mesh, seed + 1, w=2.0)
    ref, it_r, res_r = solve(dm, mesh, bs, q, q0, 1.5, 0.5, max_it, rtol, noalloc=False)
    got, it_g, res_g = solve(dm, mesh, bs, q, q0, 1.5, 0.5, max_it, rtol, noalloc=True)
    assert it_g == it_r
    assert np.array_equal(res_g, res_r)
    n = mesh.Np * mesh.Ne
    for name, a, b in zip(("DENS_dt", "MOMX_dt", "MOMY_dt", "MOMZ_dt", "RHOT_dt"), got, ref):
        assert np.array_equal(a[:n], b[:n]), f"{name}: max abs diff {np.max(np.abs(a[:n] - b[:n]))}"
    dm.close()


@pytest.mark.parametrize("max_it,rtol", [(1, 0.0), (6, 1e-10)], ids=["one-iteration", "newton"])
def test_noalloc_matches_allocating_solve_twice(max_it, rtol):
    compare(7, 4, 2, 8, 11, max_it, rtol)   # first call: 64 elements at p = 7
    compare(3, 3, 2, 4, 21, max_it, rtol)   # second call, smaller: 24 elements at p = 3


def test_noalloc_rejects_a_mesh_beyond_the_workspace():
    # 16 x 10 x 8 elements at p = 7 need 13*512*1280 + 2*64*1280 doubles (> 64 MiB)
    print("no-alloc cal_vi: oversize mesh is rejected", flush=True)
    mesh = fedg.BoxMesh(7, 16, 10, 8, (0.0, 2.0e4), (0.0, 1.25e4), (0.0, 1.0e4))
    bs = cases.base_state(mesh, "constBVFreq")
    dm = fedg.DeviceMesh(mesh)
    fields = [dev(mesh.field()) for _ in range(20)]
    ptrs = [f.ptr for f in fields]
    resid = (C.c_double * 4)()
    it = C.c_int(-1)
    rc = LIB.fedg_hevi_cal_vi_noalloc(dm.handle, *ptrs, 1.5, 0.5, 4, 1e-10, resid, C.byref(it))
    assert rc == FEDG_ERR_UNSUPPORTED
    assert b"64 MiB" in LIB.fedg_last_error()
    assert it.value == -1
    dm.close()
