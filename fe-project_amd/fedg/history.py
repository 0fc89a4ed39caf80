"""History output of the dynamical core's state (NetCDF classic data model).

Host side of the atmosphere component's history output:

  - the history items of AtmosVars_history (model/atm_nonhydro3d/src/atmos/mod_atmos_vars.F90:478-548):
    the prognostic variables, the tracers, the auxiliary variables (PRES_hyd, DENS_hyd,
    THERM_hyd, PRES, PT, RTOT, CVTOT, CPTOT, QDRY, PRES_hyd_REF) and the diagnostic
    variables of ATMOS_DIAGVARS3D_VINFO (mod_atmos_vars.F90:154-182), each diagnostic
    computed on the GPU by fedg_atmos_calc_diagvar from the device-resident state after
    Calc_diagnostics (PRES from the state, then PT; mod_atmos_vars_container.F90:318-374);
  - FILE_HISTORY_meshfield_put3D (FElib/src/file/scale_file_history_meshfield.F90:446-489):
    a 3D field goes to the file in the Cartesian node buffer of the restart files
    (File_common_meshfield_put_field3D_cartesbuf), with the axes of set_dim_axis3D
    (:664-711) built from the node coordinates;
  - the output schedule of the run configuration's PARAM_FILE_HISTORY
    (FILE_HISTORY_DEFAULT_TINTERVAL, _TUNIT, _TAVERAGE, _DATATYPE, FILE_HISTORY_OUTPUT_STEP0;
    e.g. model/atm_nonhydro3d/test/case/rising_thermal_bubble/dry/run.conf:85-99).

The output schedule, the averaging and the time axis restate the published behaviour of
SCALE's FILE_HISTORY (scale_file_history, part of the SCALE library and not of the
reference's sources): a record every TINTERVAL, the first at step 0 when OUTPUT_STEP0;
an instantaneous item holds the value put at the output step, an averaged item the
dt-weighted mean of the values put since the previous record; "time" is the end of the
record's window in seconds since the start and "time_bnds" its start and end.  The file
is written by scipy.io.netcdf_file (CDF-2) with an unlimited "time" dimension; the
reference writes NetCDF-4 through SCALE, whose container and extra attributes are not
restated (parity of the container unpinned, as for the restart files).
"""
from __future__ import annotations

import os

import numpy as np
from scipy.io import netcdf_file

from . import restart
from .dyncore import VARS, DevField, atmos_calc_diagvar, calc_pressure, tracer_specific_heat
from .mesh import BoxMesh, DeviceMesh

# (name, description, units): ATMOS_DIAGVARS3D_VINFO (mod_atmos_vars.F90:169-182) without RH,
# the auxiliary PRES and PT (scale_atm_dyn_dgm_nonhydro3d_common.F90:166-169) and the
# deviations the test cases' mod_user register (e.g. sound_wave/mod_user.F90:131)
DIAGVARS = {
    "DENS": ("density", "kg/m3"),
    "U": ("velocity u", "m/s"),
    "V": ("velocity v", "m/s"),
    "W": ("velocity w", "m/s"),
    "T": ("temperature", "K"),
    "Umet": ("eastward velocity", "m/s"),
    "Vmet": ("northward velocity", "m/s"),
    "QDRY": ("dry air", "kg/kg"),
    "ENGT": ("total energy", "J/m3"),
    "ENGP": ("potential energy", "J/m3"),
    "ENGK": ("kinetic energy", "J/m3"),
    "ENGI": ("internal energy", "J/m3"),
    "PT": ("potential temperature", "K"),
    "PRES_diff": ("pressure deviation", "Pa"),
    "T_diff": ("temperature deviation", "K"),
    "PT_diff": ("potential temperature deviation", "K"),
}
# the remaining auxiliary variables (AUXVAR_VARINFO, scale_atm_dyn_dgm_nonhydro3d_common.F90:158-179)
AUXVARS = {
    "PRES_hyd": ("hydrostatic part of pressure", "Pa"),
    "DENS_hyd": ("hydrostatic part of density", "kg/m3"),
    "THERM_hyd": ("hydrostatic part of THERM", "kg/m3"),
    "PRES": ("pressure", "Pa"),
    "RTOT": ("Total gas constant", "J/kg/K"),
    "CVTOT": ("Total heat capacity", "J/kg/K"),
    "CPTOT": ("Total heat capacity", "J/kg/K"),
    "PRES_hyd_REF": ("hydrostatic reference pressure", "Pa"),
}
PRGVARS = {name: (desc, unit) for name, desc, unit in restart.PRGVARS}

TUNITS = {"SEC": 1.0, "MIN": 60.0, "HOUR": 3600.0, "DAY": 86400.0}


class HistoryFile:
    """One rank's history file: items registered with reg(), put() at each step for which
    query() is true, write() at every step (FILE_HISTORY_meshfield_setup / _put / _write /
    _finalize).  tinterval in tunit; dt in seconds, the model's time step."""

    def __init__(self, mesh: BoxMesh, dt: float, basename: str = "history", rank: int = 0,
                 tinterval: float | None = None, tunit: str = "SEC", taverage: bool = False,
                 datatype: str = "REAL4", output_step0: bool = True, title: str = "", source: str = "fedg",
                 institution: str = "", date=(0, 1, 1, 0, 0, 0), calendar: str = ""):
        if tunit not in TUNITS:
            raise ValueError(f"tunit {tunit}: one of {', '.join(TUNITS)}")
        if datatype not in restart.DTYPES:
            raise ValueError(f"datatype {datatype}: DEFAULT, REAL4 or REAL8")
        if dt <= 0.0:
            raise ValueError("dt must be positive")
        self.mesh = mesh
        self.dt = float(dt)
        sec = (self.dt if tinterval is None else float(tinterval)) * TUNITS[tunit]
        nstep = sec / self.dt
        if nstep < 1.0 - 1e-9 or abs(nstep - round(nstep)) > 1e-6 * max(1.0, nstep):
            raise ValueError(f"the output interval {sec} s is not a whole number of time steps of {self.dt} s")
        self.tint = int(round(nstep))
        self.taverage = bool(taverage)
        self.output_step0 = bool(output_step0)
        self.dtype = restart.DTYPES[datatype]
        self.path = restart.file_name(basename, rank)
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        self.f = netcdf_file(self.path, "w", version=2)
        f = self.f
        f.title = title
        f.source = source
        f.institution = institution
        f.Conventions = "CF-1.6"
        if int(date[0]) > 0:
            y, mo, dd, h, mi, s = (int(v) for v in date)
            units = f"seconds since {y:04d}-{mo:02d}-{dd:02d} {h:02d}:{mi:02d}:{s:02d}"
        else:
            units = "seconds"
        f.createDimension("time", None)
        f.createDimension("nv", 2)
        for (name, desc, unit), a in zip(restart.AXES, restart.cartes_axes(mesh)):
            f.createDimension(name, a.size)
            v = f.createVariable(name, self.dtype, (name,))
            v.long_name = desc
            v.units = unit
            v[:] = a
        t = f.createVariable("time", "f8", ("time",))
        t.long_name = "time"
        t.units = units
        t.bounds = "time_bnds"
        if calendar:
            t.calendar = calendar
        tb = f.createVariable("time_bnds", "f8", ("time", "nv"))
        tb.long_name = "bounds of time"
        tb.units = units
        self._items = []  # [name, variable, sum, weight]
        self._ids = {}
        self.step = 0
        self.nrec = 0
        self._win0 = 0  # first step of the current window

    # -- registration
    def reg(self, name: str, desc: str, units: str, standard_name: str = "") -> int:
        """FILE_HISTORY_reg: register an item (before the first write); returns its id (from 1)."""
        if self.nrec:
            raise RuntimeError("items are registered before the first record")
        if name in self._ids:
            return self._ids[name]
        v = self.f.createVariable(name, self.dtype, ("time", "z", "y", "x"))
        v.long_name = desc
        v.units = units
        if standard_name:
            v.standard_name = standard_name
        v.cell_methods = "time: mean" if self.taverage else "time: point"
        self._items.append([name, v, None, 0.0])
        self._ids[name] = len(self._items)
        return len(self._items)

    def item_id(self, name: str) -> int:
        return self._ids.get(name, 0)

    # -- stepping
    def set_nowstep(self, step: int):
        """FILE_HISTORY_Set_NowDate: the step the following puts belong to."""
        self.step = int(step)

    def _output_step(self) -> bool:
        if self.step == 0:
            return self.output_step0
        return self.step % self.tint == 0

    def query(self, hid: int) -> bool:
        """FILE_HISTORY_query: whether a put of item hid is needed at this step."""
        if hid < 1 or hid > len(self._items):
            return False
        if self.taverage and self.step > 0:
            return True
        return self._output_step()

    def put(self, hid: int, val):
        """FILE_HISTORY_meshfield_put3D: a nodal field of Np*Ne values (numpy or DevField)."""
        it = self._items[hid - 1]
        v = val.download() if isinstance(val, DevField) else np.asarray(val, dtype=np.float64)
        buf = restart.put_cartesbuf(self.mesh, v)
        if self.taverage and self.step > 0:
            it[2] = buf * self.dt if it[2] is None else it[2] + buf * self.dt
            it[3] += self.dt
        else:
            it[2] = buf
            it[3] = 1.0

    def write(self):
        """FILE_HISTORY_write: at an output step, one record of every item put for it."""
        if not self._output_step():
            return False
        k = self.nrec
        t0 = self._win0 * self.dt
        t1 = self.step * self.dt
        self.f.variables["time"][k] = t1
        self.f.variables["time_bnds"][k] = (t0 if self.step > 0 else t1, t1)
        for it in self._items:
            name, var, s, w = it
            if s is None:
                raise RuntimeError(f"history item {name}: nothing was put for the record at step {self.step}")
            var[k] = (s / w if (self.taverage and self.step > 0) else s).astype(self.dtype)
            it[2], it[3] = None, 0.0
        self.nrec += 1
        self._win0 = self.step
        return True

    def close(self):
        """FILE_HISTORY_meshfield_finalize."""
        if self.f is not None:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_history(path: str, mesh: BoxMesh, names=None):
    """(time, time_bnds, {name: (nrec, Np*Ne) nodal values}) of a history file."""
    with netcdf_file(path, "r", mmap=False) as f:
        time = np.array(f.variables["time"][:], dtype=np.float64)
        bnds = np.array(f.variables["time_bnds"][:], dtype=np.float64)
        if names is None:
            names = [k for k, v in f.variables.items() if v.dimensions == ("time", "z", "y", "x")]
        out = {}
        for name in names:
            data = np.array(f.variables[name][:], dtype=np.float64)
            out[name] = np.stack([restart.set_cartesbuf(mesh, data[k]) for k in range(data.shape[0])])
    return time, bnds, out


class AtmosHistory:
    """AtmosVars_history over a device-resident state: the items named in `items` are
    registered in `hist`; history(step, state) puts the ones due at that step and writes
    the record.  state: the five prognostic fields in the steppers' order (DDENS, DRHOT,
    MOMZ, MOMX, MOMY, e.g. HeveStepper.get_state()).  tracers: {name: (device field,
    TRACER_MASS, TRACER_R, TRACER_CV, TRACER_CP, TRACER_ENGI0, description, units)}, whose
    Rtot / CVtot / CPtot then come from fedg_tracer_specific_heat; without tracers the air
    is dry.  zlev: the node heights (ENGP), default the mesh's; pos_z: the computational
    z (ENGT), default the mesh's z."""

    def __init__(self, dm: DeviceMesh, bs, hist: HistoryFile, items, tracers=None, zlev=None, pos_z=None):
        self.dm, self.bs, self.hist = dm, bs, hist
        m = dm.mesh
        self.n = m.Np * m.Ne
        self.tracers = dict(tracers or {})
        self.ids = {}
        for name in items:
            if name in PRGVARS:
                desc, unit = PRGVARS[name]
            elif name in AUXVARS:
                desc, unit = AUXVARS[name]
            elif name in DIAGVARS:
                desc, unit = DIAGVARS[name]
            elif name in self.tracers:
                desc, unit = self.tracers[name][6:8]
            else:
                raise ValueError(f"history item {name} is neither a prognostic, auxiliary, diagnostic nor tracer "
                                 "variable")
            self.ids[name] = hist.reg(name, desc, unit)
        n = self.n
        up = lambda a: DevField(n, np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:n]))
        self._Dh, self._Ph, self._Th = up(bs.DENS_hyd), up(bs.PRES_hyd), up(bs.THERM_hyd)
        self._zlev = up(m.height if zlev is None else zlev)
        self._zpos = up(m.z if pos_z is None else pos_z)
        self._q = [DevField(n) for _ in range(5)]
        self._P, self._DP, self._out = DevField(n), DevField(n), DevField(n)
        self._R, self._CV, self._CP = (DevField(n) for _ in range(3))

    def _thermo(self):
        if not self.tracers:
            for f, a in ((self._R, self.bs.Rtot), (self._CV, self.bs.CVtot), (self._CP, self.bs.CPtot)):
                f.upload(np.ascontiguousarray(np.asarray(a, dtype=np.float64)[: self.n]))
            return [], None, None, None
        tr = list(self.tracers.values())
        Q = [t[0] for t in tr]
        th = [np.array([t[k] for t in tr]) for k in (1, 2, 3, 4)]
        tracer_specific_heat(self.dm, Q, th, self._R, self._CV, self._CP)
        return Q, th[0], th[2], np.array([t[5] for t in tr])

    def history(self, step: int, state):
        """Puts the items due at `step` and writes the record if it is an output step."""
        h = self.hist
        h.set_nowstep(step)
        due = [name for name, hid in self.ids.items() if h.query(hid)]
        if due:
            self._put(due, state)
        return h.write()

    def _put(self, due, state):
        h, dm = self.hist, self.dm
        host = dict(zip(VARS, state))
        for f, name in zip(self._q, VARS):
            f.upload(np.ascontiguousarray(host[name][: self.n]))
        DD, DRHOT, MZ, MX, MY = self._q
        # Calc_diagnostics: specific heat, then PRES of the state, then PT
        Q, mass, cv, e0 = self._thermo()
        calc_pressure(dm, self._P, self._DP, DRHOT, self._Ph, self._Th, self._R, self._CV, self._CP)
        for name in due:
            hid = self.ids[name]
            if name in PRGVARS:
                h.put(hid, host["DRHOT" if name == "THERM" else name])
            elif name in self.tracers:
                h.put(hid, self.tracers[name][0])
            elif name in AUXVARS:
                src = {"PRES_hyd": self.bs.PRES_hyd, "DENS_hyd": self.bs.DENS_hyd, "THERM_hyd": self.bs.THERM_hyd,
                       "PRES_hyd_REF": self.bs.PRES_hyd_ref, "PRES": self._P, "RTOT": self._R,
                       "CVTOT": self._CV, "CPTOT": self._CP}[name]
                h.put(hid, src if isinstance(src, DevField) else np.asarray(src)[: self.n])
            else:
                atmos_calc_diagvar(dm, name, self._out, DD, MX, MY, MZ, self._P, self._Dh, self._Ph, self._R,
                                   self._CV, self._CP, Q=Q, mass=mass, cv=cv, engi0=e0, zlev=self._zlev,
                                   pos_z=self._zpos)
                h.put(hid, self._out)

    def close(self):
        for f in [*self._q, self._P, self._DP, self._out, self._R, self._CV, self._CP, self._Dh, self._Ph, self._Th,
                  self._zlev, self._zpos]:
            f.free()
