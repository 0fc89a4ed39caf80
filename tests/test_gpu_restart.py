"""Restart through a file on the device path: n steps, the device state
written to a restart file and read back, the base state and the hydrostatic
pressure gradient rebuilt from the file as the reference's reader does
(model/atm_nonhydro3d/src/atmos/mod_atmos_vars.F90:553-636), a fresh stepper
run n more steps, equal bit for bit to 2n steps without the restart.  Covers
the fused HEVE stepper on the flat p = 7 path and on the general p = 3 path,
and the HEVI stepper.
"""
import numpy as np
import pytest

import fedg
from fedg import cases, restart

pytestmark = pytest.mark.gpu


def _continue_from_file(tmp_path, mesh, bs, q, make, n):
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    ref = make(dm, bs, Dx, Dy)
    ref.set_state(q)
    ref.run(2 * n)
    want = ref.get_state()
    ref.close()

    st = make(dm, bs, Dx, Dy)
    st.set_state(q)
    st.run(n)
    base = str(tmp_path / "restart")
    restart.write_atmos_restart(base, mesh, st.get_state(), bs.PRES_hyd, bs.DENS_hyd)
    st.close()

    rs = restart.read_atmos_restart(base, mesh)
    bs2 = restart.restore_base_state(mesh, rs, coriolis=bs.CORIOLIS)
    Dx2, Dy2 = fedg.calc_phyd_hgrad(dm, bs2.PRES_hyd, bs2.PRES_hyd_ref)
    assert np.array_equal(Dx2, Dx) and np.array_equal(Dy2, Dy)
    st2 = make(dm, bs2, Dx2, Dy2)
    st2.set_state(rs.state)
    st2.run(n)
    got = st2.get_state()
    st2.close()
    for name, a, b in zip(fedg.VARS, got, want):
        assert np.array_equal(a, b), f"{name}: max abs diff {np.max(np.abs(a - b))}"


@pytest.mark.parametrize("p,NeX,NeY,NeZ", [(7, 4, 2, 4), (3, 6, 3, 4)], ids=["flat_p7", "p3"])
def test_heve_restart_continues_bitwise(tmp_path, p, NeX, NeY, NeZ):
    print(f"HEVE restart p={p} {NeX}x{NeY}x{NeZ}", flush=True)
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, 2.5e3 * NeX), (0.0, 2.5e3 * NeY), (0.0, 6.25e2 * NeZ))
    bs = cases.base_state(mesh, "constBVFreq")
    q = cases.warm_bubble_state(mesh, bs, [(2.0, 2.0e3, (4.0e3, 2.5e3, 1.2e3))])
    make = lambda dm, b, Dx, Dy: fedg.HeveStepper(dm, b, Dx, Dy, 0.05)
    _continue_from_file(tmp_path, mesh, bs, q, make, 3)


def test_hevi_restart_continues_bitwise(tmp_path):
    print("HEVI restart p=3 8x2x4", flush=True)
    mesh = fedg.BoxMesh(3, 8, 2, 4, (0.0, 8.0e3), (0.0, 2.0e3), (0.0, 4.0e3))
    bs = cases.base_state(mesh, "constBVFreq")
    q = cases.igw_state(mesh, bs, xc=4.0e3, a=1.0e3)
    make = lambda dm, b, Dx, Dy: fedg.HeviStepper(dm, b, Dx, Dy, 1.0)
    _continue_from_file(tmp_path, mesh, bs, q, make, 2)
