"""Restart files (fedg.restart) on the CPU: the Cartesian node buffer's index
map, the axes, the variable set and metadata, and the write -> read round trip.

The index map is pinned by fields whose values are the node coordinates: in the
buffer, the field x must equal the x axis along the last dimension at every
(z, y), y the y axis along the middle one and z the z axis along the first, so a
transposed or mis-strided map fails.  The axes are pinned to the closed form
x0 + (i1 + (1 + xi_i2) / 2) dx of element i1's LGL node i2.  The time label is
pinned by the reference's initial-state file name
(model/atm_nonhydro3d/test/case/inertia_gravity_wave/run.conf:3,
IN_BASENAME = "init_00000101-000000.000").
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

import fedg
from fedg import cases, restart


@pytest.fixture(scope="module", params=[(3, 4, 3, 2), (7, 2, 3, 2), (1, 5, 1, 3)], ids=["p3", "p7", "p1"])
def mesh(request):
    p, NeX, NeY, NeZ = request.param
    return fedg.BoxMesh(p, NeX, NeY, NeZ, (-500.0, 1.5e3 * NeX - 500.0), (0.0, 2.0e3 * NeY), (0.0, 7.5e2 * NeZ),
                        periodic=(True, False, False))


def test_cartesbuf_index_map_from_coordinates(mesh):
    n1 = mesh.porder + 1
    ax, ay, az = restart.cartes_axes(mesh)
    assert (ax.size, ay.size, az.size) == (mesh.NeX * n1, mesh.NeY * n1, mesh.NeZ * n1)
    bx, by, bz = (restart.put_cartesbuf(mesh, f) for f in (mesh.x, mesh.y, mesh.z))
    assert bx.shape == (az.size, ay.size, ax.size)
    assert np.array_equal(bx, np.broadcast_to(ax[None, None, :], bx.shape))
    assert np.array_equal(by, np.broadcast_to(ay[None, :, None], by.shape))
    assert np.array_equal(bz, np.broadcast_to(az[:, None, None], bz.shape))


def test_axes_closed_form(mesh):
    xi = np.asarray(mesh.elem.x1d)  # LGL nodes on [-1, 1]
    ax, ay, az = restart.cartes_axes(mesh)
    for a, (lo, hi), ne in ((ax, (-500.0, 1.5e3 * mesh.NeX - 500.0), mesh.NeX), (ay, (0.0, 2.0e3 * mesh.NeY), mesh.NeY),
                            (az, (0.0, 7.5e2 * mesh.NeZ), mesh.NeZ)):
        d = (hi - lo) / ne
        want = np.concatenate([lo + (i1 + 0.5 * (1.0 + xi)) * d for i1 in range(ne)])
        assert np.allclose(a, want, rtol=0.0, atol=1e-9 * (hi - lo))
        assert a[0] == lo and abs(a[-1] - hi) <= 1e-9 * (hi - lo)


def test_cartesbuf_round_trip_and_shape_check(mesh):
    rng = np.random.default_rng(3)
    n = mesh.Np * mesh.Ne
    f = rng.standard_normal(n + 17)  # a halo tail is ignored
    b = restart.put_cartesbuf(mesh, f)
    assert np.array_equal(restart.set_cartesbuf(mesh, b), f[:n])
    # every value lands once: the buffer is a permutation of the nodal values
    assert np.array_equal(np.sort(b.ravel()), np.sort(f[:n]))
    with pytest.raises(ValueError):
        restart.set_cartesbuf(mesh, b[:, :, 1:])
    with pytest.raises(ValueError):
        restart.put_cartesbuf(mesh, f[: n - 1])


def test_timelabel_and_file_name():
    assert restart.timelabel() == "00000101-000000.000"
    assert restart.timelabel((2000, 12, 31, 23, 59, 7), 0.25) == "20001231-235907.250"
    assert restart.file_name("out/restart") == "out/restart.pe000000.nc"
    assert restart.file_name("init", 3, postfix_timelabel=True) == "init_00000101-000000.000.pe000003.nc"


def _state(mesh, seed):
    rng = np.random.default_rng(seed)
    n = mesh.Np * mesh.Ne
    return [rng.standard_normal(n) * s for s in (1e-3, 0.5, 2.0, 10.0, 3.0)]


def test_write_read_round_trip(mesh, tmp_path):
    bs = cases.base_state(mesh, "constBVFreq")
    q = _state(mesh, 11)
    n = mesh.Np * mesh.Ne
    qv = np.abs(np.random.default_rng(2).standard_normal(n)) * 1e-3
    base = str(tmp_path / "restart")
    path = restart.write_atmos_restart(base, mesh, q, bs.PRES_hyd, bs.DENS_hyd,
                                       qtrc={"QV": (qv, "Water Vapor mixing ratio", "kg/kg")},
                                       title="ATMOS restart", date=(2000, 1, 1, 6, 0, 0), subsec=0.5)
    assert path == base + ".pe000000.nc"
    with open(path, "rb") as fh:
        assert fh.read(4) == b"CDF\x02"  # classic format, 64-bit offsets
    rs = restart.read_atmos_restart(base, mesh, tracers=("QV",))
    for a, b in zip(rs.state, q):
        assert np.array_equal(a, b)
    assert np.array_equal(rs.PRES_hyd, bs.PRES_hyd[:n])
    assert np.array_equal(rs.DENS_hyd, bs.DENS_hyd[:n])
    assert np.array_equal(rs.qtrc["QV"], qv)
    assert rs.attributes["Conventions"] == "CF-1.6"
    assert rs.attributes["time_units"] == "seconds since 2000-01-01 06:00:00"
    assert rs.attributes["title"] == "ATMOS restart"
    assert float(np.asarray(rs.attributes["time_start"]).ravel()[0]) == 0.5
    # layout as netCDF readers see it
    with netcdf_file(path, "r", mmap=False) as f:
        names = [k for k in f.variables if k not in ("x", "y", "z")]
        assert names == ["DDENS", "THERM", "MOMZ", "MOMX", "MOMY", "PRES_hyd", "DENS_hyd", "QV"]
        for name, desc, unit in restart.PRGVARS + restart.AUXVARS:
            v = f.variables[name]
            assert v.dimensions == ("z", "y", "x")
            assert v.long_name.decode() == desc and v.units.decode() == unit
            assert v.data.dtype == np.dtype(">f8")
        for name, desc, unit in restart.AXES:
            assert f.variables[name].long_name.decode() == desc and f.variables[name].units.decode() == unit
        ax = restart.cartes_axes(mesh)[0]
        assert np.array_equal(f.variables["x"][:], ax)
        assert np.array_equal(f.variables["DDENS"][:], restart.put_cartesbuf(mesh, q[0]))


def test_real4_rounds_to_float32(mesh, tmp_path):
    q = _state(mesh, 5)
    bs = cases.base_state(mesh, "constT")
    base = str(tmp_path / "r4")
    restart.write_atmos_restart(base, mesh, q, bs.PRES_hyd, bs.DENS_hyd, dtype="REAL4")
    rs = restart.read_atmos_restart(base, mesh)
    for a, b in zip(rs.state, q):
        assert np.array_equal(a, b.astype(np.float32).astype(np.float64))


def test_missing_variable_and_wrong_mesh(mesh, tmp_path):
    bs = cases.base_state(mesh, "constT")
    base = str(tmp_path / "r")
    restart.write_atmos_restart(base, mesh, _state(mesh, 1), bs.PRES_hyd, bs.DENS_hyd)
    with pytest.raises(KeyError):
        restart.read_atmos_restart(base, mesh, tracers=("QV",))
    other = fedg.BoxMesh(mesh.porder, mesh.NeX + 1, mesh.NeY, mesh.NeZ, (0.0, 1.0), (0.0, 1.0), (0.0, 1.0))
    with pytest.raises(ValueError):
        restart.read_atmos_restart(base, other)
    with pytest.raises(ValueError):
        restart.write_atmos_restart(base, mesh, _state(mesh, 1)[:4], bs.PRES_hyd, bs.DENS_hyd)


def test_restore_base_state_matches_setup(mesh, tmp_path):
    """The reader's base state equals the one the state was set up with, halos included."""
    for kind in ("constT", "constPT", "constBVFreq"):
        bs = cases.base_state(mesh, kind)
        base = str(tmp_path / kind)
        restart.write_atmos_restart(base, mesh, _state(mesh, 2), bs.PRES_hyd, bs.DENS_hyd)
        got = restart.restore_base_state(mesh, restart.read_atmos_restart(base, mesh))
        for name in ("DENS_hyd", "PRES_hyd", "THERM_hyd", "PRES_hyd_ref", "Rtot", "CVtot", "CPtot", "CORIOLIS"):
            assert np.array_equal(getattr(got, name), getattr(bs, name)), (kind, name)
        assert np.allclose(got.PT_hyd, bs.PT_hyd, rtol=1e-12, atol=0.0), kind
