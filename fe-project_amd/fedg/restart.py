"""Restart files of the dynamical core's state (NetCDF classic data model).

Host side of the atmosphere component's restart I/O:

  - the variable set, order and metadata of AtmosVars_Write_restart_file_prep /
    AtmosVars_Write_restart_file (model/atm_nonhydro3d/src/atmos/mod_atmos_vars.F90:641-715):
    the five prognostic variables DDENS, THERM, MOMZ, MOMX, MOMY (PRGVAR ids 1..5),
    the auxiliary PRES_hyd and DENS_hyd (AUXVAR ids 1..2), then the QA tracers,
    with the names, descriptions and units of
    FElib/src/fluid_dyn_solver/scale_atm_dyn_dgm_nonhydro3d_common.F90:146-166;
  - the reader of AtmosVars_Read_restart_file (mod_atmos_vars.F90:553-636): the same
    variables by name, then THERM_hyd from PRES_hyd, PRES_hyd_REF = 0 and the
    horizontal gradient of PRES_hyd (restore_base_state);
  - FILE_restart_meshfield_component create / def_var / write / open / read / close
    (FElib/src/file/scale_file_restart_meshfield.F90:340-560) with the file name
    <basename>[_<timelabel>].pe<rank:06d>.nc and the global attributes of
    FILE_base_meshfield_put_global_attribute_time (scale_file_base_meshfield.F90:1102-1140);
  - the Cartesian node buffer every 3D field is stored in
    (File_common_meshfield_put_field3D_cartesbuf / set_cartesbuf_field3D,
    FElib/src/file/scale_file_common_meshfield.F90:1191-1330, 1395-) and the axes
    built from the node coordinates (File_common_meshfield_get_axis3D, :1036-1112).

A field of Np * Ne nodal values (node i2 + N1 j2 + N1^2 k2 of element
ei + NeX ej + NeX NeY ek, the mesh's own order) is stored as the Fortran array
buf(NeX N1, NeY N1, NeZ N1) with buf(i2 + i1 N1, j2 + j1 N1, k2 + k1 N1) (0-based
here) = value of node (i2, j2, k2) of element (i1, j1, k1).  NetCDF lists a Fortran
array's dimensions in reverse, so the variable has dimensions ("z", "y", "x") and
numpy sees it as shape (NeZ N1, NeY N1, NeX N1).  Nodes on element faces appear
once per element, as in the reference: the buffer holds every nodal value and the
round trip is exact.

The files are written by scipy.io.netcdf_file in the NetCDF classic format with
64-bit offsets (CDF-2), which the netCDF library and xarray read as they read the
reference's files.  The reference writes NetCDF-4/HDF5 through the SCALE library
(not part of the reference's sources); that container format is not written
here, and attributes SCALE's FILE_Create adds beyond title, source and
institution are not restated.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np
from scipy.io import netcdf_file

from . import constants
from .cases import BaseState, calc_rhot_hyd
from .mesh import BoxMesh

# (name, description, units): PRGVAR_VARINFO and AUXVAR_VARINFO(1:2),
# scale_atm_dyn_dgm_nonhydro3d_common.F90:146-166; the restart file's long_name
# is the description (AtmosVars_Write_restart_file_prep passes %DESC).
PRGVARS = (("DDENS", "deviation of density", "kg/m3"),
           ("THERM", "THERM", "-"),
           ("MOMZ", "momentum z", "kg/m2/s"),
           ("MOMX", "momentum x", "kg/m2/s"),
           ("MOMY", "momentum y", "kg/m2/s"))
AUXVARS = (("PRES_hyd", "hydrostatic part of pressure", "Pa"),
           ("DENS_hyd", "hydrostatic part of density", "kg/m3"))

# mesh dimension names and metadata (FElib/src/mesh/scale_mesh_base3d.F90:127-134)
AXES = (("x", "X-coordinate", "m"), ("y", "Y-coordinate", "m"), ("z", "Z-coordinate", "m"))

DTYPES = {"DEFAULT": "f8", "REAL8": "f8", "REAL4": "f4"}


def timelabel(date=(0, 1, 1, 0, 0, 0), subsec=0.0) -> str:
    """YYYYMMDD-hhmmss.sss, the label of TIME_gettimelabel (e.g. 00000101-000000.000)."""
    y, mo, d, h, mi, s = (int(v) for v in date)
    ms = int(round(float(subsec) * 1000.0))
    return f"{y:04d}{mo:02d}{d:02d}-{h:02d}{mi:02d}{s:02d}.{ms:03d}"


def file_name(basename: str, rank: int = 0, postfix_timelabel: bool = False, date=(0, 1, 1, 0, 0, 0),
              subsec: float = 0.0) -> str:
    """<basename>[_<timelabel>].pe<rank:06d>.nc (FILE_restart_meshfield_component_create :394-399,
    and the per-rank file names the reference's scripts open, e.g. model/atm_nonhydro3d/test/case/
    common/cmd_mkgraph.py:43)."""
    base = f"{basename}_{timelabel(date, subsec)}" if postfix_timelabel else basename
    return f"{base}.pe{int(rank):06d}.nc"


def _shape(mesh: BoxMesh):
    n1 = mesh.porder + 1
    return n1, (mesh.NeZ * n1, mesh.NeY * n1, mesh.NeX * n1)


def put_cartesbuf(mesh: BoxMesh, val: np.ndarray) -> np.ndarray:
    """Nodal field (Np*Ne values, halo tail ignored) -> Cartesian buffer, numpy shape (z, y, x)
    (File_common_meshfield_put_field3D_cartesbuf, scale_file_common_meshfield.F90:1191-1330)."""
    n1, shp = _shape(mesh)
    n = mesh.Np * mesh.Ne
    v = np.asarray(val)
    if v.shape[0] < n:
        raise ValueError(f"field has {v.shape[0]} values, the mesh {n}")
    # [ek, ej, ei, k2, j2, i2] -> [ek, k2, ej, j2, ei, i2]
    v = v[:n].reshape(mesh.NeZ, mesh.NeY, mesh.NeX, n1, n1, n1)
    return np.ascontiguousarray(v.transpose(0, 3, 1, 4, 2, 5)).reshape(shp)


def set_cartesbuf(mesh: BoxMesh, buf: np.ndarray) -> np.ndarray:
    """Cartesian buffer (z, y, x) -> nodal field of Np*Ne values
    (File_common_meshfield_set_cartesbuf_field3D, scale_file_common_meshfield.F90:1395-)."""
    n1, shp = _shape(mesh)
    b = np.asarray(buf)
    if b.shape != shp:
        raise ValueError(f"buffer shape {b.shape}, the mesh needs {shp}")
    b = b.reshape(mesh.NeZ, n1, mesh.NeY, n1, mesh.NeX, n1)
    return np.ascontiguousarray(b.transpose(0, 2, 4, 1, 3, 5)).reshape(-1).astype(np.float64)


def cartes_axes(mesh: BoxMesh):
    """x, y, z axes of the Cartesian buffer from the node coordinates
    (File_common_meshfield_get_axis3D, scale_file_common_meshfield.F90:1036-1112): x from the
    nodes along the y- face edge of the elements of the first row (j = k = 1), y from the
    x- face of the first column, z from the vertical node column of the first elements."""
    n1 = mesh.porder + 1
    Np, NeX, NeY = mesh.Np, mesh.NeX, mesh.NeY
    x = mesh.x.reshape(-1, Np)
    y = mesh.y.reshape(-1, Np)
    z = mesh.z.reshape(-1, Np)
    ax = np.concatenate([x[i, :n1] for i in range(NeX)])
    ay = np.concatenate([y[j * NeX, 0: n1 * n1: n1] for j in range(NeY)])
    az = np.concatenate([z[k * NeX * NeY, 0: Np: n1 * n1] for k in range(mesh.NeZ)])
    return ax, ay, az


class RestartFile:
    """One rank's restart file (FILE_restart_meshfield_component).

    Write: create(), def_var() per variable, end_def(), write_var() per variable, close().
    Read: open(), read_var() per variable, close()."""

    def __init__(self, mesh: BoxMesh, rank: int = 0):
        self.mesh = mesh
        self.rank = int(rank)
        self.f = None
        self.path = None
        self._vars = {}
        self._defs = []

    # -- write
    def create(self, basename: str, title: str = "", dtype: str = "DEFAULT", date=(0, 1, 1, 0, 0, 0),
               subsec: float = 0.0, postfix_timelabel: bool = False, calendar: str = "",
               source: str = "fedg", institution: str = ""):
        """FILE_restart_meshfield_component_create + Put_GlobalAttribute_time."""
        if dtype not in DTYPES:
            raise ValueError(f"dtype {dtype}: DEFAULT, REAL4 or REAL8")
        self.dtype = DTYPES[dtype]
        self.path = file_name(basename, self.rank, postfix_timelabel, date, subsec)
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        self.f = netcdf_file(self.path, "w", version=2)
        self.f.title = title
        self.f.source = source
        self.f.institution = institution
        self.f.Conventions = "CF-1.6"
        self.f.grid_name = "hoge"
        if int(date[0]) > 0:
            y, mo, dd, h, mi, s = (int(v) for v in date)
            self.f.time_units = f"seconds since {y:04d}-{mo:02d}-{dd:02d} {h:02d}:{mi:02d}:{s:02d}"
            if calendar:
                self.f.calendar = calendar
        else:
            self.f.time_units = "seconds"
        self.f.time_start = np.array([float(subsec)])
        # axes (def_axes, scale_file_base_meshfield.F90:1144-1183)
        axes = cartes_axes(self.mesh)
        for (name, desc, unit), a in zip(AXES, axes):
            self.f.createDimension(name, a.size)
            v = self.f.createVariable(name, self.dtype, (name,))
            v.long_name = desc
            v.units = unit
            v[:] = a
        return self.path

    def def_var(self, name: str, desc: str, units: str):
        if self.f is None or self.f.mode != "w":
            raise RuntimeError("def_var before create")
        v = self.f.createVariable(name, self.dtype, ("z", "y", "x"))
        v.long_name = desc
        v.units = units
        self._vars[name] = v
        self._defs.append(name)

    def end_def(self):
        """Definitions are complete (scipy writes the header at close)."""

    def write_var(self, name: str, val: np.ndarray):
        if name not in self._vars:
            raise KeyError(f"{name} is not defined in {self.path}")
        self._vars[name][:] = put_cartesbuf(self.mesh, val).astype(self.dtype)

    # -- read
    def open(self, basename: str, postfix_timelabel: bool = False, date=(0, 1, 1, 0, 0, 0), subsec: float = 0.0):
        """FILE_restart_meshfield_component_open: the rank's file, checked against the mesh's axes."""
        self.path = file_name(basename, self.rank, postfix_timelabel, date, subsec)
        self.f = netcdf_file(self.path, "r", mmap=False)
        _, shp = _shape(self.mesh)
        dims = tuple(self.f.dimensions[a] for a in ("z", "y", "x"))
        if dims != shp:
            raise ValueError(f"{self.path}: dimensions (z, y, x) = {dims}, the mesh needs {shp}")
        return self.path

    def read_var(self, name: str, allow_missing: bool = False):
        if name not in self.f.variables:
            if allow_missing:
                return None
            raise KeyError(f"variable {name} is not in {self.path}")
        return set_cartesbuf(self.mesh, np.array(self.f.variables[name][:], dtype=np.float64))

    def attributes(self) -> dict:
        return {k: (v.decode() if isinstance(v, bytes) else v) for k, v in self.f._attributes.items()}

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None


@dataclass
class AtmosRestart:
    """What AtmosVars_Read_restart_file reads: the prognostic state (VARS order
    DDENS, THERM, MOMZ, MOMX, MOMY), PRES_hyd, DENS_hyd and the tracers, as Np*Ne arrays."""
    state: list
    PRES_hyd: np.ndarray
    DENS_hyd: np.ndarray
    qtrc: dict
    attributes: dict


def write_atmos_restart(basename: str, mesh: BoxMesh, state, PRES_hyd, DENS_hyd, qtrc=None, rank: int = 0,
                        title: str = "", dtype: str = "DEFAULT", date=(0, 1, 1, 0, 0, 0), subsec: float = 0.0,
                        postfix_timelabel: bool = False) -> str:
    """AtmosVars_Write_restart_file_prep + _file + _post (mod_atmos_vars.F90:641-730).

    state: the five prognostic fields in PRGVAR order (DDENS, THERM, MOMZ, MOMX, MOMY), e.g.
    HeveStepper.get_state(); qtrc: {tracer name: (field, description, units)}, written after
    the auxiliary fields in the order given."""
    if len(state) != len(PRGVARS):
        raise ValueError("state needs the five prognostic fields DDENS, THERM, MOMZ, MOMX, MOMY")
    rf = RestartFile(mesh, rank)
    path = rf.create(basename, title, dtype, date, subsec, postfix_timelabel)
    qtrc = qtrc or {}
    try:
        for name, desc, unit in PRGVARS + AUXVARS:
            rf.def_var(name, desc, unit)
        for name, (_, desc, unit) in qtrc.items():
            rf.def_var(name, desc, unit)
        rf.end_def()
        for (name, _, _), val in zip(PRGVARS, state):
            rf.write_var(name, val)
        rf.write_var("PRES_hyd", PRES_hyd)
        rf.write_var("DENS_hyd", DENS_hyd)
        for name, (val, _, _) in qtrc.items():
            rf.write_var(name, val)
    finally:
        rf.close()
    return path


def read_atmos_restart(basename: str, mesh: BoxMesh, tracers=(), rank: int = 0, postfix_timelabel: bool = False,
                       date=(0, 1, 1, 0, 0, 0), subsec: float = 0.0) -> AtmosRestart:
    """The reads of AtmosVars_Read_restart_file (mod_atmos_vars.F90:576-596); a missing
    variable is an error, as in the reference."""
    rf = RestartFile(mesh, rank)
    rf.open(basename, postfix_timelabel, date, subsec)
    try:
        state = [rf.read_var(name) for name, _, _ in PRGVARS]
        P, D = rf.read_var("PRES_hyd"), rf.read_var("DENS_hyd")
        q = {name: rf.read_var(name) for name in tracers}
        attrs = rf.attributes()
    finally:
        rf.close()
    return AtmosRestart(state, P, D, q, attrs)


def restore_base_state(mesh: BoxMesh, rs: AtmosRestart, c=constants.SCALE, coriolis=None) -> BaseState:
    """Base state after a restart read (mod_atmos_vars.F90:598-633) for dry air: halos filled,
    THERM_hyd = P00/Rdry (PRES_hyd/P00)^(CVdry/CPdry) (update_therm_hyd), PRES_hyd_REF = 0,
    Rtot / CVtot / CPtot of dry air.  The hydrostatic pressure gradient follows from
    calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref) (update_phyd_hgrad); the Coriolis
    parameter is a setup field, not a restart variable."""
    n = mesh.Np * mesh.Ne
    D, P = mesh.field(), mesh.field()
    D[:n], P[:n] = rs.DENS_hyd, rs.PRES_hyd
    mesh.exchange(D)
    mesh.exchange(P)
    T = mesh.field()
    T[:n] = calc_rhot_hyd(P[:n], c)
    mesh.exchange(T)
    cor = np.zeros(mesh.Nfp_v * mesh.Ne2DA) if coriolis is None else np.ascontiguousarray(coriolis, dtype=np.float64)
    PT_hyd = c.PRES00 / c.Rdry * (P[:n] / c.PRES00) ** (c.CVdry / c.CPdry) / D[:n]
    return BaseState(DENS_hyd=D, PRES_hyd=P, THERM_hyd=T, PRES_hyd_ref=mesh.field(0.0),
                     Rtot=mesh.field(c.Rdry), CVtot=mesh.field(c.CVdry), CPtot=mesh.field(c.CPdry),
                     CORIOLIS=cor, PT_hyd=PT_hyd)
