"""The Fortran host side of the boundary on the GPU: fe-project_amd/fortran/
fedg_dropin_harness (amdflang, the ISO_C_BINDING module fedg_c_binding and
fedg_dropin_cal_tend_ex, which takes the argument list of the reference's
atm_dyn_nonhydro3d_cal_tend_ex, FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:152-199)
builds its own mesh and context through the C ABI and computes the HEVE
tendency of fields the test writes; the result equals the ctypes path and the
CPU oracle bit for bit.  Inputs: moist Rtot / CVtot / CPtot varying node by
node, a random Coriolis parameter, a horizontally varying PRES_hyd with a
non-zero PRES_hyd_ref, a noisy state, x walls (SLIP) and a periodic y.
"""
import os
import subprocess

import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import constants
from test_gpu_general_inputs import general_base, noisy_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "fe-project_amd", "fortran", "fedg_dropin_harness")
NAMES = ("DENS_dt", "MOMX_dt", "MOMY_dt", "MOMZ_dt", "RHOT_dt")


def write_input(path, mesh, ext, periodic, fields, cor):
    c = constants.SCALE
    with open(path, "wb") as f:
        np.array([mesh.porder, mesh.NeX, mesh.NeY, mesh.NeZ, *[int(b) for b in periodic]], dtype=np.int32).tofile(f)
        np.array(ext, dtype=np.float64).tofile(f)
        np.array([c.GRAV, c.Rdry, c.CPdry, c.CVdry, c.PRES00], dtype=np.float64).tofile(f)
        for a in fields:
            assert a.size == mesh.Np * mesh.NeA
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        assert cor.size == mesh.Nfp_v * mesh.Ne2DA
        np.ascontiguousarray(cor, dtype=np.float64).tofile(f)


@pytest.mark.parametrize("p,shape", [(3, (4, 3, 4)), (7, (3, 2, 3))], ids=["p3", "p7"])
def test_fortran_cal_tend_ex_equals_ctypes_and_oracle(p, shape, tmp_path):
    assert os.access(HARNESS, os.X_OK), "the Fortran harness is built by fe-project_amd/Makefile (amdflang)"
    NeX, NeY, NeZ = shape
    periodic = (False, True, False)
    ext = (0.0, 1.25e3 * NeX, 0.0, 1.25e3 * NeY, 0.0, 1.0e3 * NeZ)
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, ext[0:2], ext[2:4], ext[4:6], periodic=periodic)
    bs = general_base(mesh, seed=40 + p)
    q = noisy_state(mesh, seed=50 + p)
    with O.pow_mode(O.POW_PORTABLE):
        om = O.OracleMesh(mesh)
        _, DP = om.drhot2pres(q[1], bs)
        mesh.exchange(DP)
        om.apply_bc(q)
        Dx, Dy = om.calc_phyd_hgrad(bs)
        ref = om.cal_tend(q, DP, bs, Dx, Dy)
    # DDENS_, MOMX_, MOMY_, MOMZ_, THERM_, DPRES_, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd,
    # Rtot, CVtot, CPtot, DPhydDx, DPhydDy
    fields = [q[0], q[3], q[4], q[2], q[1], DP, bs.DENS_hyd, bs.PRES_hyd, bs.PRES_hyd_ref, bs.THERM_hyd,
              bs.Rtot, bs.CVtot, bs.CPtot, Dx, Dy]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, mesh, ext, periodic, fields, bs.CORIOLIS)
    print(f"fortran harness p={p} {shape}", flush=True)
    r = subprocess.run([HARNESS, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = mesh.Np * mesh.Ne
    got = np.fromfile(fout, dtype=np.float64)
    assert got.size == 5 * n
    got = got.reshape(5, n)
    # the same call through ctypes
    dm = fedg.DeviceMesh(mesh)
    dev = lambda a: fedg.DevField(a.size, a)
    out = [dev(mesh.field()) for _ in range(5)]
    fedg.heve_cal_tend(dm, out, dev(q[0]), dev(q[3]), dev(q[4]), dev(q[2]), dev(q[1]), dev(DP),
                       dev(bs.DENS_hyd), dev(bs.PRES_hyd), dev(bs.PRES_hyd_ref), dev(bs.THERM_hyd),
                       dev(bs.CORIOLIS), dev(bs.Rtot), dev(bs.CVtot), dev(bs.CPtot), dev(Dx), dev(Dy))
    for k, name in enumerate(NAMES):
        c = out[k].download()[:n]
        assert np.all(np.isfinite(got[k])), name
        assert np.array_equal(got[k], c), f"{name}: Fortran vs ctypes max diff {np.max(np.abs(got[k] - c))}"
        assert np.array_equal(got[k], ref[k][:n]), f"{name}: Fortran vs oracle"
        assert np.max(np.abs(c)) > 0.0, name
