"""Host-side box builder and the C ABI library (no GPU needed).

Builder pins (MSH/scale_mesh_cubedom3d.F90, MSH/scale_meshutil_3d.F90 semantics,
test_mesh_cubedom3d_hexahedral.f90:207-218):
  - vmapP by brute force: the exterior node of every face node is the node of
    the neighbouring element at the same physical point (periodic wrap by the
    domain length), found by a nearest-point search;
  - metric closed forms of an axis-aligned box: J = dx dy dz / 8,
    Escale = diag(2/dx, 2/dy, 2/dz), Fscale = 2/h_normal, unit axis normals;
  - halo slots: ordered by boundary (y-, x+, y+, x-, z-, z+), the periodic
    source is on the opposite boundary and a closed one is the slot's own node.
ABI: libfedg.so loads without a GPU and exports every entry point declared in
include/fedg.h.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from fedg import _lib
from fedg.mesh import BoxMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("periodic", [(True, True, False), (False, True, True), (True, True, True)])
@pytest.mark.parametrize("p", [1, 3])
def test_vmapP_by_nearest_point(p, periodic):
    L = (3000.0, 2000.0, 1000.0)
    m = BoxMesh(p, 3, 2, 2, (0.0, L[0]), (0.0, L[1]), (0.0, L[2]), periodic=periodic)
    n = m.Np * m.Ne
    X = np.stack([m.x, m.y, m.z], axis=1)
    iM, iP = m.vmapM - 1, m.vmapP - 1
    assert np.all(iM < n)
    assert np.all((iP >= 0) & (iP < n + m.nB))
    # vmapM: the element's own face nodes
    elem = np.repeat(np.arange(m.Ne), m.NfpTot)
    assert np.array_equal(iM // m.Np, elem)
    src = np.where(iP < n, iP, m.halo_src[np.clip(iP - n, 0, m.nB - 1)] - 1)
    nf = m.NfpTot * m.Ne
    nrm = np.stack([m.normal_fn[d * nf:(d + 1) * nf] for d in range(3)], axis=1)
    for t in range(nf):
        xm = X[iM[t]]
        ax = int(np.argmax(np.abs(nrm[t])))
        on_bnd = (xm[ax] <= 1e-9 and nrm[t, ax] < 0) or (xm[ax] >= L[ax] - 1e-9 and nrm[t, ax] > 0)
        if on_bnd and not periodic[ax]:
            assert src[t] == iM[t] and iP[t] >= n
            continue
        target = xm.copy()
        if on_bnd:
            target[ax] = L[ax] - target[ax] if target[ax] <= 1e-9 else target[ax] - L[ax]
        # nearest node of the neighbouring element across the face
        d = np.linalg.norm(X - target, axis=1)
        cand = np.flatnonzero(d <= 1e-6)
        cand = cand[cand // m.Np != iM[t] // m.Np] if periodic[ax] or not on_bnd else cand
        if m.NeX == 1 and ax == 0 or m.NeY == 1 and ax == 1 or m.NeZ == 1 and ax == 2:
            continue
        assert src[t] in cand, (t, src[t], cand)


def test_box_metrics_closed_form():
    L = (4000.0, 3000.0, 1500.0)
    N = (4, 3, 5)
    m = BoxMesh(3, *N, (0.0, L[0]), (0.0, L[1]), (0.0, L[2]))
    n = m.Np * m.Ne
    h = [L[a] / N[a] for a in range(3)]
    assert np.allclose(m.J[:n], h[0] * h[1] * h[2] / 8, rtol=1e-14, atol=0)
    E = m.Escale.reshape((n, 3, 3), order="F")
    for a in range(3):
        for b in range(3):
            want = 2.0 / h[a] if a == b else 0.0
            assert np.allclose(E[:, a, b], want, rtol=1e-14, atol=1e-18)
    nf = m.NfpTot * m.Ne
    nrm = np.stack([m.normal_fn[d * nf:(d + 1) * nf] for d in range(3)], axis=1).reshape(m.Ne, 6, m.Nfp_v, 3)
    ref = np.array([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, -1], [0, 0, 1]], dtype=float)
    assert np.array_equal(nrm, np.broadcast_to(ref[None, :, None, :], nrm.shape))
    Fs = m.Fscale.reshape(m.Ne, 6, m.Nfp_v)
    hn = [h[1], h[0], h[1], h[0], h[2], h[2]]
    for f in range(6):
        assert np.allclose(Fs[:, f, :], 2.0 / hn[f], rtol=1e-14)
    assert np.all(m.Gsqrt == 1.0) and np.all(m.GI3 == 0.0) and np.all(m.GsqrtH == 1.0)


def test_halo_slot_order_and_sources():
    m = BoxMesh(2, 3, 2, 2, (0.0, 3.0), (0.0, 2.0), (0.0, 2.0), periodic=(True, False, False))
    n = m.Np * m.Ne
    # slots grouped by boundary id 1..6 in ascending order
    assert np.all(np.diff(m.halo_face) >= 0)
    assert set(np.unique(m.halo_face)) == {1, 2, 3, 4, 5, 6}
    x, y, z = m.x, m.y, m.z
    src = m.halo_src - 1
    own = m.vmapB - 1
    for b, (coord, val_own) in enumerate([(y, 0.0), (x, 3.0), (y, 2.0), (x, 0.0), (z, 0.0), (z, 2.0)], start=1):
        sel = m.halo_face == b
        assert np.allclose(coord[own[sel]], val_own)
        if b in (2, 4):  # periodic in x: the source lies on the opposite boundary
            assert np.allclose(coord[src[sel]], 3.0 - val_own)
        else:
            assert np.array_equal(src[sel], own[sel])
    assert np.all(src < n)


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "fedg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*(fedg_[a-z0-9_]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_entry_point():
    names = declared_entry_points()
    assert len(names) >= 30
    lib = _lib.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(_lib.LIB, n)
    assert set(names) <= set(_lib.EXPORTED)


def test_error_reporting_without_gpu():
    rc = _lib.LIB.fedg_refelem_tables(0, None, None, None, None, None, None, None, None, None, None)
    assert rc != 0  # porder must be >= 1
    assert len(_lib.LIB.fedg_last_error()) > 0
    assert len(_lib.LIB.fedg_version()) > 0
