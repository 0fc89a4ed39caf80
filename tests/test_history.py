"""History files (fedg.history.HistoryFile) on the CPU: the output schedule of
PARAM_FILE_HISTORY (a record every TINTERVAL, the first at step 0 when
OUTPUT_STEP0), the time axis and its bounds, instantaneous and averaged items,
the Cartesian node buffer (shared with the restart files and pinned there), the
item metadata of the reference's variable tables, and the errors.

The schedule and averaging restate SCALE's FILE_HISTORY, which is not in the
reference tree: they are pinned by their definitions (which steps write, the
dt-weighted mean of a linear-in-time field is its value at the window's
midpoint), not by files the reference holds."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import fedg
from fedg import history, restart


@pytest.fixture(scope="module")
def mesh():
    return fedg.BoxMesh(2, 3, 2, 2, (0.0, 3e3), (0.0, 2e3), (0.0, 1e3), periodic=(True, False, False))


def field(mesh, k, seed=0):
    rng = np.random.default_rng(seed)
    base = rng.standard_normal(mesh.Np * mesh.Ne)
    return base + 0.25 * k


def test_instantaneous_schedule_and_round_trip(mesh, tmp_path):
    dt = 0.5
    h = history.HistoryFile(mesh, dt, basename=str(tmp_path / "history"), tinterval=1.5, tunit="SEC",
                            datatype="REAL8", title="t")
    a = h.reg("U", *history.DIAGVARS["U"])
    b = h.reg("DDENS", *history.PRGVARS["DDENS"])
    assert (a, b) == (1, 2) and h.reg("U", "x", "y") == 1
    wrote = []
    for step in range(10):
        h.set_nowstep(step)
        for hid, seed in ((a, 1), (b, 2)):
            if h.query(hid):
                h.put(hid, field(mesh, step, seed))
        if h.write():
            wrote.append(step)
    h.close()
    assert wrote == [0, 3, 6, 9]
    assert h.path == str(tmp_path / "history.pe000000.nc")
    time, bnds, data = history.read_history(h.path, mesh)
    assert np.array_equal(time, [0.0, 1.5, 3.0, 4.5])
    assert np.array_equal(bnds, [[0.0, 0.0], [0.0, 1.5], [1.5, 3.0], [3.0, 4.5]])
    assert set(data) == {"U", "DDENS"}
    for k, step in enumerate(wrote):
        assert np.array_equal(data["U"][k], field(mesh, step, 1))
        assert np.array_equal(data["DDENS"][k], field(mesh, step, 2))
    with netcdf_file(h.path, "r", mmap=False) as f:
        v = f.variables["U"]
        assert v.dimensions == ("time", "z", "y", "x")
        assert v.long_name == b"velocity u" and v.units == b"m/s"
        assert f.variables["DDENS"].long_name == b"deviation of density"
        assert f.dimensions["time"] is None  # unlimited
        for name, ax in zip(("x", "y", "z"), restart.cartes_axes(mesh)):
            assert np.array_equal(f.variables[name][:], ax)
        assert f.variables["time"].units == b"seconds"


def test_no_step0_and_minutes(mesh, tmp_path):
    h = history.HistoryFile(mesh, 30.0, basename=str(tmp_path / "h2"), tinterval=1.0, tunit="MIN",
                            output_step0=False, date=(2000, 1, 1, 0, 0, 0))
    hid = h.reg("W", *history.DIAGVARS["W"])
    wrote = []
    for step in range(7):
        h.set_nowstep(step)
        if h.query(hid):
            h.put(hid, field(mesh, step))
        if h.write():
            wrote.append(step)
    h.close()
    assert wrote == [2, 4, 6]
    time, bnds, data = history.read_history(h.path, mesh)
    assert np.array_equal(time, [60.0, 120.0, 180.0])
    # REAL4 by default: the values are float32 roundings of the fields
    for k, step in enumerate(wrote):
        assert np.array_equal(data["W"][k], field(mesh, step).astype(np.float32).astype(np.float64))
    with netcdf_file(h.path, "r", mmap=False) as f:
        assert f.variables["time"].units == b"seconds since 2000-01-01 00:00:00"
        assert f.variables["W"].typecode() == "f"


def test_time_average(mesh, tmp_path):
    dt, tint = 2.0, 4
    h = history.HistoryFile(mesh, dt, basename=str(tmp_path / "havg"), tinterval=tint * dt, taverage=True,
                            datatype="REAL8")
    hid = h.reg("T", *history.DIAGVARS["T"])
    puts = 0
    for step in range(9):
        h.set_nowstep(step)
        if h.query(hid):
            h.put(hid, field(mesh, step))
            puts += 1
        h.write()
    h.close()
    assert puts == 9  # step 0 (instantaneous) and every step after it
    time, bnds, data = history.read_history(h.path, mesh)
    assert np.array_equal(time, [0.0, 8.0, 16.0])
    assert np.array_equal(bnds, [[0.0, 0.0], [0.0, 8.0], [8.0, 16.0]])
    assert np.array_equal(data["T"][0], field(mesh, 0))
    # steps 1..4 and 5..8: the mean of a field linear in the step is its value at the mean step
    for k, mid in ((1, 2.5), (2, 6.5)):
        assert np.allclose(data["T"][k], field(mesh, 0) + 0.25 * mid, rtol=0.0, atol=1e-14)
    with netcdf_file(h.path, "r", mmap=False) as f:
        assert f.variables["T"].cell_methods == b"time: mean"


def test_errors(mesh, tmp_path):
    with pytest.raises(ValueError):
        history.HistoryFile(mesh, 0.3, basename=str(tmp_path / "e1"), tinterval=1.0)  # not a whole number of steps
    with pytest.raises(ValueError):
        history.HistoryFile(mesh, 1.0, basename=str(tmp_path / "e2"), tunit="WEEK")
    h = history.HistoryFile(mesh, 1.0, basename=str(tmp_path / "e3"), tinterval=2.0)
    a = h.reg("U", *history.DIAGVARS["U"])
    h.reg("V", *history.DIAGVARS["V"])
    h.set_nowstep(0)
    h.put(a, field(mesh, 0))
    with pytest.raises(RuntimeError):  # V was not put for the step-0 record
        h.write()
    h.close()
    h = history.HistoryFile(mesh, 1.0, basename=str(tmp_path / "e4"))
    a = h.reg("U", *history.DIAGVARS["U"])
    h.set_nowstep(0)
    h.put(a, field(mesh, 0))
    assert h.write()
    with pytest.raises(RuntimeError):  # items are registered before the first record
        h.reg("V", *history.DIAGVARS["V"])
    h.close()


def test_variable_tables_follow_the_reference():
    # ATMOS_DIAGVARS3D_VINFO (model/atm_nonhydro3d/src/atmos/mod_atmos_vars.F90:171-182) less RH
    ref = {"DENS": ("density", "kg/m3"), "U": ("velocity u", "m/s"), "V": ("velocity v", "m/s"),
           "W": ("velocity w", "m/s"), "T": ("temperature", "K"), "Umet": ("eastward velocity", "m/s"),
           "Vmet": ("northward velocity", "m/s"), "QDRY": ("dry air", "kg/kg"), "ENGT": ("total energy", "J/m3"),
           "ENGP": ("potential energy", "J/m3"), "ENGK": ("kinetic energy", "J/m3"),
           "ENGI": ("internal energy", "J/m3")}
    for k, v in ref.items():
        assert history.DIAGVARS[k] == v
    assert list(history.PRGVARS) == ["DDENS", "THERM", "MOMZ", "MOMX", "MOMY"]
    assert history.AUXVARS["PRES"] == ("pressure", "Pa")
