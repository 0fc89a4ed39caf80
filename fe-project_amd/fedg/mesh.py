"""Reference element, box mesh and device context (host side).

Mirrors FElib's setup objects: ``HexahedralElement`` + ``ElementOperationTensorProd3D_P<p>``
(ELM/scale_element_hexahedral.F90, ELM/scale_element_operation_tensorprod3D.F90)
and ``MeshCubeDom3D`` with one local mesh (MSH/scale_mesh_cubedom3d.F90).  All
arrays are numpy arrays in the reference's Fortran (column-major) layout, so
``a.reshape(shape, order="F")`` gives the Fortran-shaped view.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import constants
from ._lib import LIB, BoxArrays, BoxDesc, BoxSizes, MeshDesc, check

BND_NONE, BND_SLIP, BND_NOSLIP = 0, 1, 2


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


@dataclass
class RefElem:
    """Operators of the order-p LGL hexahedron (1D tables, Lift_mat, VFilterPM1)."""
    porder: int
    x1d: np.ndarray = field(init=False)
    w1d: np.ndarray = field(init=False)
    D1D: np.ndarray = field(init=False)
    V1D: np.ndarray = field(init=False)
    invV1D: np.ndarray = field(init=False)
    invM1D: np.ndarray = field(init=False)
    Lift_mat: np.ndarray = field(init=False)
    IntrpMat_VPOrdM1_tr: np.ndarray = field(init=False)
    IndexH2Dto3D: np.ndarray = field(init=False)
    IndexH2Dto3D_bnd: np.ndarray = field(init=False)

    def __post_init__(self):
        n1 = self.porder + 1
        self.n1, self.Np, self.Nfp = n1, n1 ** 3, n1 ** 2
        self.NfpTot = 6 * n1 ** 2
        f = lambda n: np.zeros(n, dtype=np.float64)
        self.x1d, self.w1d = f(n1), f(n1)
        self.D1D, self.V1D, self.invV1D, self.invM1D = f(n1 * n1), f(n1 * n1), f(n1 * n1), f(n1 * n1)
        self.Lift_mat = f(6 * self.Np)
        self.IntrpMat_VPOrdM1_tr = f(n1 * n1)
        self.IndexH2Dto3D = np.zeros(self.Np, dtype=np.int32)
        self.IndexH2Dto3D_bnd = np.zeros(self.NfpTot, dtype=np.int32)
        check(LIB.fedg_refelem_tables(self.porder, _dp(self.x1d), _dp(self.w1d), _dp(self.D1D), _dp(self.V1D),
                                      _dp(self.invV1D), _dp(self.invM1D), _dp(self.Lift_mat),
                                      _dp(self.IntrpMat_VPOrdM1_tr), _ip(self.IndexH2Dto3D),
                                      _ip(self.IndexH2Dto3D_bnd)), "fedg_refelem_tables")


class BoxMesh:
    """One-tile structured box (MeshCubeDom3D with NprcX = NprcY = 1)."""

    def __init__(self, porder, NeX, NeY, NeZ, xlim, ylim, zlim, periodic=(True, True, False)):
        self.desc = BoxDesc(porder, NeX, NeY, NeZ, float(xlim[0]), float(xlim[1]), float(ylim[0]), float(ylim[1]),
                            float(zlim[0]), float(zlim[1]), int(periodic[0]), int(periodic[1]), int(periodic[2]))
        s = BoxSizes()
        check(LIB.fedg_box_get_sizes(C.byref(self.desc), C.byref(s)), "fedg_box_get_sizes")
        self.sizes = s
        for n, _ in BoxSizes._fields_:
            setattr(self, n, int(getattr(s, n)))
        self.porder, self.NeX, self.NeY, self.NeZ = porder, NeX, NeY, NeZ
        self.periodic = tuple(bool(x) for x in periodic)
        Np, NfpTot, Ne, NeA, nB = self.Np, self.NfpTot, self.Ne, self.NeA, self.nB
        f = lambda n: np.zeros(n, dtype=np.float64)
        i = lambda n: np.zeros(n, dtype=np.int32)
        self.pos_en, self.J, self.Escale = f(Np * Ne * 3), f(Np * Ne), f(Np * Ne * 9)
        self.sJ, self.Fscale, self.normal_fn = f(NfpTot * Ne), f(NfpTot * Ne), f(NfpTot * Ne * 3)
        self.vmapM, self.vmapP = i(NfpTot * Ne), i(NfpTot * Ne)
        self.vmapB, self.halo_src, self.halo_face = i(nB), i(nB), i(nB)
        self.EMap3Dto2D = i(Ne)
        self.Gsqrt, self.GI3, self.GsqrtH = f(Np * NeA), f(Np * NeA * 2), f(self.Nfp_v * self.Ne2D)
        arr = BoxArrays(_dp(self.pos_en), _dp(self.J), _dp(self.Escale), _dp(self.sJ), _dp(self.Fscale),
                        _dp(self.normal_fn), _ip(self.vmapM), _ip(self.vmapP), _ip(self.vmapB), _ip(self.halo_src),
                        _ip(self.halo_face), _ip(self.EMap3Dto2D), _dp(self.Gsqrt), _dp(self.GI3), _dp(self.GsqrtH))
        check(LIB.fedg_box_build(C.byref(self.desc), C.byref(arr)), "fedg_box_build")
        self.elem = RefElem(porder)

    def set_topography(self, topo, zTop=None):
        """Terrain-following vertical coordinate over a topography (MeshTopography_set_vcoordinate,
        MSH/scale_mesh_topography.F90:159-258): topo is a callable f(x, y) evaluated at the horizontal
        nodes or an (Nfp_v*Ne2D,) array; zTop defaults to the box top.  Replaces Gsqrt and GI3 by the
        coordinate's metrics and sets zlev, the height of every node (``height``)."""
        zTop = float(self.desc.zmax if zTop is None else zTop)
        n2 = self.Nfp_v * self.Ne2D
        if callable(topo):
            xs = self.x[: self.Np * self.Ne2D].reshape(self.Ne2D, self.Np)[:, : self.Nfp_v].ravel()
            ys = self.y[: self.Np * self.Ne2D].reshape(self.Ne2D, self.Np)[:, : self.Nfp_v].ravel()
            t = np.ascontiguousarray(topo(xs, ys), dtype=np.float64)
        else:
            t = np.ascontiguousarray(topo, dtype=np.float64)
        if t.size != n2:
            raise ValueError(f"topography needs {n2} values (Nfp_v*Ne2D)")
        zlev = np.zeros(self.Np * self.NeA)
        check(LIB.fedg_box_topography_metrics(C.byref(self.desc), _dp(t), zTop, _dp(self.Gsqrt), _dp(self.GI3),
                                              _dp(zlev)), "fedg_box_topography_metrics")
        self.topo, self.zTop, self.zlev = t, zTop, zlev
        return self

    # coordinates of the nodes, (Np*Ne,) each
    @property
    def height(self):
        """Height of the nodes: zlev over a topography, else the box's z."""
        z = getattr(self, "zlev", None)
        return self.z if z is None else z[: self.Np * self.Ne]

    @property
    def x(self):
        return self.pos_en[: self.Np * self.Ne]

    @property
    def y(self):
        return self.pos_en[self.Np * self.Ne: 2 * self.Np * self.Ne]

    @property
    def z(self):
        return self.pos_en[2 * self.Np * self.Ne:]

    def velbc(self, bnd=None):
        """Velocity BC per halo slot from a per-boundary list (y-, x+, y+, x-, z-, z+).

        Default: periodic lateral boundaries, SLIP bottom and top, as in the
        rising-thermal-bubble cases (PARAM_ATMOS_DYN_BND btm/top_vel_bc='SLIP')."""
        if bnd is None:
            # boundary order y-, x+, y+, x-, z-, z+ -> axis y, x, y, x, z, z
            axis = (1, 0, 1, 0, 2, 2)
            bnd = [BND_NONE if self.periodic[a] else BND_SLIP for a in axis]
        return np.asarray(bnd, dtype=np.int32)[self.halo_face - 1].astype(np.int32)

    def field(self, fill=0.0):
        """A zero (Np,NeA) field."""
        return np.full(self.Np * self.NeA, fill, dtype=np.float64)

    def exchange(self, f):
        """Single-tile halo exchange on a host field (in place)."""
        base = self.Np * self.Ne
        f[base: base + self.nB] = f[self.halo_src - 1]
        return f


class DeviceMesh:
    """fedg_ctx: the mesh, element and operators uploaded once to the GPU."""

    def __init__(self, mesh: BoxMesh, velbc=None, consts=constants.SCALE, box_shape=True):
        self.mesh = mesh
        self.consts = consts
        e = mesh.elem
        self._velbc = mesh.velbc() if velbc is None else np.asarray(velbc, dtype=np.int32)
        d = MeshDesc()
        d.porder = mesh.porder
        d.Ne, d.NeA, d.Ne2D, d.Ne2DA, d.nB = mesh.Ne, mesh.NeA, mesh.Ne2D, mesh.Ne2DA, mesh.nB
        d.Escale, d.Gsqrt, d.GI3, d.GsqrtH = _dp(mesh.Escale), _dp(mesh.Gsqrt), _dp(mesh.GI3), _dp(mesh.GsqrtH)
        d.G11 = d.G12 = d.G22 = None
        d.normal_fn, d.Fscale = _dp(mesh.normal_fn), _dp(mesh.Fscale)
        d.vmapM, d.vmapP, d.EMap3Dto2D = _ip(mesh.vmapM), _ip(mesh.vmapP), _ip(mesh.EMap3Dto2D)
        d.halo_src, d.velbc = _ip(mesh.halo_src), _ip(self._velbc)
        d.D1D, d.Lift_mat, d.IntrpMat_VPOrdM1_tr = _dp(e.D1D), _dp(e.Lift_mat), _dp(e.IntrpMat_VPOrdM1_tr)
        if box_shape:
            d.NeX, d.NeY, d.NeZ = mesh.NeX, mesh.NeY, mesh.NeZ
        d.GRAV, d.Rdry, d.CPdry, d.CVdry, d.PRES00 = (consts.GRAV, consts.Rdry, consts.CPdry, consts.CVdry,
                                                        consts.PRES00)
        h = C.c_void_p()
        check(LIB.fedg_ctx_create(C.byref(d), C.byref(h)), "fedg_ctx_create")
        self.handle = h

    def setup_modalfilter(self, MF_h, MF_v):
        """Setup_ModalFilter: (n1,n1) filter matrices (e.g. from ``modalfilter_matrix``)."""
        h = np.asfortranarray(MF_h, dtype=np.float64).ravel(order="F")
        v = np.asfortranarray(MF_v, dtype=np.float64).ravel(order="F")
        check(LIB.fedg_ctx_setup_modalfilter(self.handle, _dp(h), _dp(v)), "fedg_ctx_setup_modalfilter")

    @property
    def flat(self) -> bool:
        """Whether the context recognised the flat-terrain uniform box (constant-metric kernels)."""
        return bool(LIB.fedg_ctx_is_flat(self.handle))

    def close(self):
        if self.handle:
            LIB.fedg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
