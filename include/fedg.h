/*
 * fedg: MI355X (gfx950) implementation of FElib's regional nonhydrostatic 3D
 * DG dynamical-core step (ywkawai/FE-Project, FElib/src/fluid_dyn_solver).
 *
 * C ABI only: plain pointers, sizes and status codes; no C++ or torch types.
 * Arrays keep the reference's Fortran column-major layout (X(Np,NeA) is
 * X[p + Np*ke]) and its 1-based integer maps, so a Fortran host passes
 * lcmesh%... arrays through ISO_C_BINDING unchanged.
 *
 * Pointer kinds:
 *   host   - ordinary CPU memory
 *   device - GPU memory (fedg_dev_alloc, hipMalloc, or an OpenMP/OpenACC
 *            device copy); the entry points never copy it to the host.
 *
 * Every entry point returns FEDG_OK (0) or an error code; fedg_last_error()
 * gives the message.  The reference aborts (PRC_abort) on such errors; the
 * Fortran binding (fe-project_amd/fortran/fedg_c_binding.F90) lets the host do the
 * same with a nonzero status.  All device work is stream-ordered on the
 * context's stream and complete when the function returns, as the reference's
 * OpenACC path waits on its async queue before returning
 * (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_gpu.F90:471,484).
 */
#ifndef FEDG_H
#define FEDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FEDG_OK = 0,
  FEDG_ERR_ARG = 1,         /* invalid argument or unsupported configuration */
  FEDG_ERR_HIP = 2,         /* a HIP runtime call failed */
  FEDG_ERR_NOMEM = 3,       /* device allocation failed */
  FEDG_ERR_UNSUPPORTED = 4  /* feature outside the built scope */
};

/* Prognostic variable ids (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:68-77), 1-based. */
enum { FEDG_DDENS_ID = 1, FEDG_DRHOT_ID = 2, FEDG_MOMZ_ID = 3, FEDG_MOMX_ID = 4, FEDG_MOMY_ID = 5 };

/* Velocity boundary kinds per halo slot (FDS/scale_atm_dyn_dgm_bnd.F90:323-360). */
enum { FEDG_BND_NONE = 0, FEDG_BND_SLIP = 1, FEDG_BND_NOSLIP = 2 };

const char *fedg_last_error(void);
const char *fedg_version(void);

/* ------------------------------------------------------------------ */
/* Device memory helpers (for hosts without their own device allocator) */
int fedg_dev_alloc(void **ptr, size_t bytes);
int fedg_dev_free(void *ptr);
int fedg_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes);
int fedg_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes);
int fedg_memset_dev(void *dst_dev, int value, size_t bytes);
int fedg_device_sync(void);
int fedg_device_name(char *buf, int buflen);
/* Access-pattern probe (no counterpart in the reference; for measurement).
 * Times, on the current device, a kernel with the loads and stores of one
 * p = 7 fused RK4 stage on a periodic NeX x NeY x NeZ box and no arithmetic:
 * mode 0 / 1 / 2 = first / middle / last stage (reads 6 / 16 / 11 fields of
 * every node, writes 11 / 11 / 6, gathers six fields at the 384 face nodes of
 * every element from its neighbours).  Allocates and frees 27 fields of
 * NeX NeY NeZ 512 doubles.  *ms = mean kernel time over reps launches.  This
 * is the practical HBM ceiling of a stage with that traffic. */
int fedg_stage_probe(int mode, int NeX, int NeY, int NeZ, int reps, double *ms);

/* ------------------------------------------------------------------ */
/* Reference element operators (host).  Replaces the setup of
 * ElementOperationTensorProd3D_P<p> (ELM/scale_element_operation_tensorprod3D.F90:5512-5569)
 * and HexahedralElement (ELM/scale_element_hexahedral.F90:102-404).
 * n1 = porder+1.  Any output pointer may be NULL.
 *   x1d, w1d (n1): LGL nodes/weights; D1D (n1,n1); V1D, invV1D (n1,n1);
 *   invM1D (n1,n1) = V1D V1D^T; Lift_mat (n1,n1,n1,6); IntrpMat_VPOrdM1_tr (n1,n1);
 *   IndexH2Dto3D (n1^3), IndexH2Dto3D_bnd (6 n1^2): 1-based. */
int fedg_refelem_tables(int porder, double *x1d, double *w1d, double *D1D, double *V1D,
                        double *invV1D, double *invM1D, double *Lift_mat, double *IntrpMat_VPOrdM1_tr,
                        int32_t *IndexH2Dto3D, int32_t *IndexH2Dto3D_bnd);

/* Modal filter matrix of a 1D LGL element of order porder: ModalFilter%FilterMat
 * of ModalFilter_Init_line (ELM/scale_element_modalfilter.F90:60-90) with the
 * exponential filter of get_exp_filter (:196-236), F = V diag(sigma) V^-1,
 * sigma_p = exp(-alpha ((eta_p - etac)/(1 - etac))^ord) for eta_p = p/porder
 * > etac, p > 0, else 1.  F: (n1,n1) column-major.  The driver's defaults are
 * etac 2/3, alpha 36, ord 16 (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:1249-1294). */
int fedg_modalfilter_matrix(int porder, double etac, double alpha, int ord, double *F);

/* ------------------------------------------------------------------ */
/* Structured box mesh (host).  Replaces MeshCubeDom3D_setupLocalDom for one
 * tile (MSH/scale_mesh_cubedom3d.F90:339-469) with the metrics of
 * MSH/scale_mesh_base3d.F90:200-318, the face maps of
 * MSH/scale_meshutil_3d.F90:264-641 and the single-tile halo exchange of
 * DAT/scale_meshfieldcomm_cubedom3d.F90:245-344. */
typedef struct {
  int32_t porder;
  int32_t NeX, NeY, NeZ;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int32_t periodic_x, periodic_y, periodic_z;
} fedg_box_desc;

typedef struct {
  int32_t n1, Np, Nfp_h, Nfp_v, NfpTot;
  int32_t Ne, NeA, Ne2D, Ne2DA, nB;
} fedg_box_sizes;

/* Host output arrays (caller-allocated, any may be NULL). */
typedef struct {
  double *pos_en;       /* (Np,Ne,3)     */
  double *J;            /* (Np,Ne)       */
  double *Escale;       /* (Np,Ne,3,3)   */
  double *sJ;           /* (NfpTot,Ne)   */
  double *Fscale;       /* (NfpTot,Ne)   */
  double *normal_fn;    /* (NfpTot,Ne,3) */
  int32_t *vmapM;       /* (NfpTot,Ne) 1-based into Np*NeA */
  int32_t *vmapP;       /* (NfpTot,Ne) 1-based into Np*NeA */
  int32_t *vmapB;       /* (nB) 1-based into Np*Ne */
  int32_t *halo_src;    /* (nB) 1-based interior node copied into halo slot */
  int32_t *halo_face;   /* (nB) domain boundary 1..6 (y-,x+,y+,x-,z-,z+) of halo slot */
  int32_t *EMap3Dto2D;  /* (Ne) 1-based */
  double *Gsqrt;        /* (Np,NeA)      */
  double *GI3;          /* (Np,NeA,2)    */
  double *GsqrtH;       /* (Nfp_v,Ne2D)  */
} fedg_box_arrays;

int fedg_box_get_sizes(const fedg_box_desc *desc, fedg_box_sizes *sizes);
int fedg_box_build(const fedg_box_desc *desc, fedg_box_arrays *out);
/* Terrain-following vertical coordinate of the box (host): MeshTopography_set_vcoordinate
 * (MSH/scale_mesh_topography.F90:159-258) with MeshUtil_VCoord_GetMetric
 * (MSH/scale_meshutil_vcoord.F90:81-209, TERRAIN_FOLLOWING) and
 * Set_geometric_with_vcoord (MSH/scale_mesh_cubedom3d.F90:292-334).  The box
 * of desc is the computational one (zeta from zmin = 0 to zmax <= zTop);
 * topo (Nfp_v,Ne2D) holds the topography at the 2-D nodes of the horizontal
 * mesh (IndexH2Dto3D order).  Outputs (any may be NULL): Gsqrt (Np,NeA) =
 * GsqrtV = 1 - topo/zTop; GI3 (Np,NeA,2) = G13, G23 = -(1 - zeta/zTop) /
 * GsqrtV d topo/dx, dy with the DG gradient of topo (central face jumps
 * lifted); zlev (Np,NeA) = zeta + (1 - zeta/zTop) topo.  Halo slots hold the
 * single-tile exchange's values.  The other metrics of fedg_box_build
 * (Escale, normals, Fscale, GsqrtH) are unchanged by the vertical coordinate. */
int fedg_box_topography_metrics(const fedg_box_desc *desc, const double *topo, double zTop, double *Gsqrt,
                                double *GI3, double *zlev);

/* ------------------------------------------------------------------ */
/* Context: flattened LocalMesh3D + ElementBase3D + ElementOperationTensorProd3D,
 * uploaded once (the reference keeps them on the device with !$acc enter data,
 * DAT/scale_localmeshfield_base.F90:119-128).  All pointers are host pointers
 * to the reference's arrays; the context keeps device copies. */
typedef struct {
  int32_t porder;
  int32_t Ne, NeA, Ne2D, Ne2DA, nB;
  const double *Escale;      /* (Np,Ne,3,3)   lcmesh%Escale   */
  const double *Gsqrt;       /* (Np,NeA)      lcmesh%Gsqrt    */
  const double *GI3;         /* (Np,NeA,2)    lcmesh%GI3      */
  const double *GsqrtH;      /* (Nfp_v,Ne2D)  lcmesh%GsqrtH   */
  const double *G11, *G12, *G22; /* (Nfp_v,Ne2D) lcmesh%GIJ(:,:,1,1),(1,2),(2,2); NULL = 1,0,1 */
  const double *normal_fn;   /* (NfpTot,Ne,3) lcmesh%normal_fn */
  const double *Fscale;      /* (NfpTot,Ne)   lcmesh%Fscale   */
  const int32_t *vmapM;      /* (NfpTot,Ne)   lcmesh%vmapM (1-based) */
  const int32_t *vmapP;      /* (NfpTot,Ne)   lcmesh%vmapP (1-based) */
  const int32_t *EMap3Dto2D; /* (Ne)          lcmesh%EMap3Dto2D      */
  const int32_t *halo_src;   /* (nB) interior node that fills each halo slot (1-based) */
  const int32_t *velbc;      /* (nB) FEDG_BND_* of each halo slot (AtmDynBnd VelBC_list) */
  const double *D1D;         /* (n1,n1)       element3D_operation%D1D */
  const double *Lift_mat;    /* (n1,n1,n1,6)  element3D_operation%Lift_mat */
  const double *IntrpMat_VPOrdM1_tr; /* (n1,n1) element3D_operation%IntrpMat_VPOrdM1_tr */
  /* box shape when the mesh came from fedg_box_build (0 otherwise); lets the
   * fused stepper resolve neighbours arithmetically */
  int32_t NeX, NeY, NeZ;
  /* physical constants (SCALE scale_const: CONST_GRAV, CONST_Rdry, CONST_CPdry,
   * CONST_CVdry, CONST_PRE00); zero selects the SCALE defaults below */
  double GRAV, Rdry, CPdry, CVdry, PRES00;
} fedg_mesh_desc;

#define FEDG_DEFAULT_GRAV 9.80665
#define FEDG_DEFAULT_RDRY 287.04
#define FEDG_DEFAULT_CPDRY 1004.64
#define FEDG_DEFAULT_CVDRY (1004.64 - 287.04)
#define FEDG_DEFAULT_PRES00 100000.0

typedef struct fedg_ctx fedg_ctx;

int fedg_ctx_create(const fedg_mesh_desc *desc, fedg_ctx **ctx);
int fedg_ctx_destroy(fedg_ctx *ctx);

/* Modal filter.  Setup_ModalFilter (ELM/scale_element_operation_tensorprod3D.F90:5464-5509):
 * the host's MFilter_h1D and MFilter_v1D ((n1,n1) column-major host arrays,
 * e.g. from fedg_modalfilter_matrix).  Apply: atm_dyn_dgm_modalfilter_apply
 * (FDS/scale_atm_dyn_dgm_modalfilter.F90:48-130) on device fields (Np,NeA),
 * elements 1..Ne, in place; returns after the work is complete. */
int fedg_ctx_setup_modalfilter(fedg_ctx *ctx, const double *MFilter_h1D, const double *MFilter_v1D);
int fedg_modalfilter_apply(fedg_ctx *ctx, double *DDENS, double *MOMX, double *MOMY, double *MOMZ, double *DRHOT,
                           int do_weight_Gsqrt);
/* sizes: Np, NfpTot, Ne, NeA, nB ... of the context */
int fedg_ctx_sizes(const fedg_ctx *ctx, fedg_box_sizes *sizes);
/* 1 when the context recognised the flat-terrain uniform box (unit Gsqrt and
 * GsqrtH, zero GI3, axis-aligned unit normals, Escale and Fscale constant per
 * direction and face, the box's own face map); its kernels then take those
 * values as constants.  No counterpart in the reference (a query of this
 * library's specialisation). */
int fedg_ctx_is_flat(const fedg_ctx *ctx);

/* ------------------------------------------------------------------ */
/* Drop-in entry points.  Device pointers; fields are (Np,NeA) unless noted.
 * Each computes what the cited reference routine computes, for elements
 * NeS..NeE = 1..Ne, reading halo slots through vmapP as the reference does. */

/* atm_dyn_dgm_nonhydro3d_common_DRHOT2PRES
 * (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:428-479): PRES, DPRES from DRHOT. */
int fedg_drhot2pres(fedg_ctx *ctx, double *PRES, double *DPRES, const double *DRHOT,
                    const double *PRES_hyd, const double *THERM_hyd, const double *Rtot,
                    const double *CVtot, const double *CPtot);

/* MeshFieldComm_Exchange on one tile (DAT/scale_meshfieldcomm_base.F90:571-805,
 * 846-912): fills the halo tail of each of nfields device fields. */
int fedg_halo_exchange(fedg_ctx *ctx, int nfields, double *const *fields);

/* ATMOS_dyn_bnd_applyBC_prgvars_lc (FDS/scale_atm_dyn_dgm_bnd.F90:270-367). */
int fedg_apply_bc_prgvars(fedg_ctx *ctx, double *DDENS, double *MOMX, double *MOMY, double *MOMZ,
                          double *THERM, const double *DENS_hyd, const double *PRES_hyd);

/* atm_dyn_dgm_nonhydro3d_rhot_heve_cal_tend
 * (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:292-489 with its numerical
 * flux FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:946-1138); the
 * argument list is that of the abstract interface atm_dyn_nonhydro3d_cal_tend_ex
 * (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:152-199) minus the derived-type
 * arguments, which the context carries.  CORIOLIS is (Nfp_v,Ne2DA). */
int fedg_heve_cal_tend(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                       double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                       const double *MOMY_, const double *MOMZ_, const double *DRHOT_,
                       const double *DPRES_, const double *DENS_hyd, const double *PRES_hyd,
                       const double *PRES_hyd_ref, const double *THERM_hyd, const double *CORIOLIS,
                       const double *Rtot, const double *CVtot, const double *CPtot,
                       const double *DPhydDx, const double *DPhydDy);

/* The same tendency split so that the halo exchange can overlap the interior
 * work (the reference's communication-hiding pair):
 *  - fedg_heve_cal_tend_cco = atm_dyn_dgm_nonhydro3d_rhot_heve_cal_tend_cco
 *    (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:493-678): the volume
 *    terms only (pointwise fluxes, Div_var5_2, VFilterPM1, Coriolis and
 *    DPhydDx/Dy), no face flux; needs no halo values.  Same arguments as
 *    fedg_heve_cal_tend.
 *  - fedg_heve_add_bnd_contrib = atm_dyn_dgm_nonhydro3d_rhot_heve_add_bnd_contrib_generalvc
 *    (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:242-496): after
 *    the exchange, DPRES of the halo tail from the halo's DRHOT and base state
 *    (written to DPRES), then the Rusanov flux at every face node (vmapM /
 *    vmapP), its lift (Lift_var5) and dt -= lift / Gsqrt for the five
 *    tendencies in place (PRGVAR_dt(:,:,1..5) of the reference =
 *    DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt).  Rtot, CVtot, CPtot are all
 *    given (halo included) or all NULL (dry-air constants).
 * cal_tend_cco followed by add_bnd_contrib equals fedg_heve_cal_tend up to the
 * rounding of the regrouped sums (the reference's own two paths differ the
 * same way). */
int fedg_heve_cal_tend_cco(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                           double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                           const double *MOMY_, const double *MOMZ_, const double *DRHOT_,
                           const double *DPRES_, const double *DENS_hyd, const double *PRES_hyd,
                           const double *PRES_hyd_ref, const double *THERM_hyd, const double *CORIOLIS,
                           const double *Rtot, const double *CVtot, const double *CPtot,
                           const double *DPhydDx, const double *DPhydDy);
int fedg_heve_add_bnd_contrib(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                              double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                              const double *MOMY_, const double *MOMZ_, const double *DRHOT_, double *DPRES,
                              const double *DENS_hyd, const double *PRES_hyd, const double *THERM_hyd,
                              const double *Rtot, const double *CVtot, const double *CPtot);

/* atm_dyn_dgm_trcadvect3d_heve_cal_tend
 * (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231 with its face flux
 * atm_dyn_dgm_trcadvect3d_heve_get_delflux_generalhvc, :554-674): the
 * advection tendency of one tracer, QTRC_dt = -Div(Gsqrt (MOMX, MOMY,
 * MOMZ/GsqrtV + G13 MOMX + G23 MOMY) QTRC)/Gsqrt + RHOQ_tp, with the central
 * face flux damped by alphDENS_M/_P (NfpTot,Ne) and weighted by fct_coef on
 * the upwind side of the face's net outward flux.  The FaceIntMat argument of
 * the reference (its _Init, :75-137) is built from the context's order; J is
 * lcmesh%J (Np,Ne).  QTRC_, the momenta, fct_coef and RHOQ_tp are
 * (Np,NeA) with the halo filled. */
int fedg_trcadvect_cal_tend(fedg_ctx *ctx, double *QTRC_dt, const double *QTRC_,
                            const double *MOMX_, const double *MOMY_, const double *MOMZ_,
                            const double *alphDENS_M, const double *alphDENS_P,
                            const double *fct_coef, const double *RHOQ_tp, const double *J);

/* atm_dyn_dgm_trcadvect3d_heve_calc_fct_coef
 * (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:232-305, with
 * get_netOutwardFlux_generalhvc, :676-797): per element
 * fct_coef = max(0, min(1, Q / (netOutwardFlux + 1e-10))), Q the tracer mass
 * available over dt, sum(Gsqrt J IntWeight_lgl (dens_ssm1 QTRC_ / dt + RHOQ_tp_))
 * with dens_ssm1 = DENS_hyd + (1 - rk_c_ssm1) DDENS0_ + rk_c_ssm1 DDENS_, and
 * netOutwardFlux the sum over the six faces of the positive LGL-weighted face
 * integrals of the face flux of fedg_trcadvect_cal_tend.  disable_limiter != 0:
 * fct_coef = 1 (every other pointer may then be NULL).  Interior elements are
 * written; the halo is the caller's exchange, as in the reference driver. */
int fedg_trcadvect_calc_fct_coef(fedg_ctx *ctx, double *fct_coef, const double *QTRC_,
                                 const double *MOMX_, const double *MOMY_, const double *MOMZ_,
                                 const double *RHOQ_tp_, const double *alphDENS_M,
                                 const double *alphDENS_P, const double *DENS_hyd,
                                 const double *DDENS_, const double *DDENS0_, double rk_c_ssm1,
                                 double dt, const double *J, int disable_limiter);

/* atm_dyn_dgm_trcadvect3d_TMAR (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:308-337):
 * in place, per element, QTRC_ = Q0 / (Q1 + 1e-32) * max(0, QTRC_) with Q0, Q1 the
 * element masses sum(Gsqrt J IntWeight_lgl dens QTRC_) before and after the cut,
 * dens = DENS_hyd + DDENS_. */
int fedg_trcadvect_tmar(fedg_ctx *ctx, double *QTRC_, const double *DENS_hyd, const double *DDENS_,
                        const double *J);

/* atm_dyn_dgm_trcadvect3d_save_massflux (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:340-404)
 * with atm_dyn_dgm_trcadvect3d_heve_cal_alphdens_dyn (:447-536): at RK stage
 * rkstage (from 1; 1 resets), MFLX_x/y_tavg += tavg_weight_h MOMX/MOMY,
 * MFLX_z_tavg += tavg_weight_v MOMZ on the interior elements, and
 * alph_dens_M/_P (NfpTot,Ne) += tavg_weight alpha dens Gsqrt on both sides of
 * every face node, alpha the larger of the two sides' sqrt(Gnn gamma P / dens)
 * + |normal velocity| (Gnn without |nz| when is_hevi).  As in the reference,
 * gamma is the dry-air CPdry / CVdry; Rtot, CVtot, CPtot are accepted and
 * unused. */
int fedg_trcadvect_save_massflux(fedg_ctx *ctx, double *MFLX_x_tavg, double *MFLX_y_tavg,
                                 double *MFLX_z_tavg, double *alph_dens_M, double *alph_dens_P,
                                 const double *DDENS, const double *MOMX, const double *MOMY,
                                 const double *MOMZ, const double *DPRES, const double *DENS_hyd,
                                 const double *PRES_hyd, const double *Rtot, const double *CVtot,
                                 const double *CPtot, int rkstage, double tavg_weight_h,
                                 double tavg_weight_v, int is_hevi);

/* Coupled dynamics + tracer step, after the tracer update
 * (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:540-553):
 *  - fedg_tracer_specific_heat = the specific-heat part of
 *    update_pressure_specific_heat (:653-716): ATMOS_THERMODYN_specific_heat of
 *    the SCALE library (not in the reference tree; its mixing rule qdry = 1 -
 *    sum q_i TRACER_MASS_i, Rtot = sum q_i TRACER_R_i + qdry Rdry, likewise CVtot
 *    and CPtot) on every node of elements 1..Ne; the pressure follows with
 *    fedg_drhot2pres, as the driver's calc_pressure;
 *  - fedg_trc_negative_fixer = fix_negative_val (:596-651) with
 *    atm_phy_mp_dgm_common_negative_fixer
 *    (FElib/src/microphysics/scale_atm_phy_mp_dgm_common.F90:429-556): per
 *    element, each water species' negative mass removed with the element's
 *    mass kept (max(M0, 0) / M1 scaling), DDENS, the internal energy (kept as
 *    an element integral, then the potential-energy change Grav zlev
 *    removed), PRES, Rtot, CVtot, CPtot and, when DRHOT is not NULL, DRHOT
 *    updated in place.
 * QTRC: nq device (Np,NeA) fields (1 <= nq <= 8); TRACER_MASS/_R/_CV/_CP:
 * nq host values each; J (Np,Ne) lcmesh%J; zlev (Np,Ne) the node heights; the
 * fixer's tracers are the water species (nq = 1 + QLA + QIA); dt is accepted
 * and unused, as in the reference. */
int fedg_tracer_specific_heat(fedg_ctx *ctx, int nq, double *const *QTRC, const double *TRACER_MASS,
                              const double *TRACER_R, const double *TRACER_CV, const double *TRACER_CP,
                              double *Rtot, double *CVtot, double *CPtot);
int fedg_trc_negative_fixer(fedg_ctx *ctx, int nq, double *const *QTRC, double *DDENS, double *PRES,
                            double *CVtot, double *CPtot, double *Rtot, const double *DENS_hyd,
                            const double *PRES_hyd, const double *J, const double *zlev, double dt, int QLA,
                            int QIA, const double *TRACER_MASS, const double *TRACER_R, const double *TRACER_CV,
                            const double *TRACER_CP, double *DRHOT);

/* A diagnostic variable of the atmosphere's history output: AtmosVars_CalcDiagvar
 * -> AtmosVarsContainer_CalcDiagvar -> vars_calc_diagnoseVar_lc
 * (model/atm_nonhydro3d/src/atmos/mod_atmos_vars_container.F90:378-452, 1100-1366),
 * by its name (field_name) on every node of elements 1..Ne of var_out (Np,Ne):
 * DENS, U, V, W, PRES_diff, T, T_diff, PT, PT_diff, ENGK, ENGP (with zlev, the
 * node heights), ENGI, ENGT (with pos_z, lcmesh%pos_en(:,:,3): the
 * computational coordinate, as the reference's ENGT), Umet and Vmet (the
 * regional mesh's Calc_UVmet, mod_atmos_mesh_rm.F90:299-323, is the identity),
 * and QDRY (the AUXVAR of Calc_SpecificHeat: 1 - sum QTRC_i TRACER_MASS_i, in
 * tracer order).  QDRY_ of ENGI / ENGT is that same value.  "PRES" leaves
 * var_out as it is (an empty case in the reference); RH / RHL are rejected
 * (they need the SCALE library's saturation vapour pressure); any other name is
 * an error, as the reference's PRC_abort.  The tracers (nq of them, 0 for dry
 * air) are read only by QDRY, ENGI and ENGT; CVtot is accepted and unused, as
 * in the reference; fields a formula does not read may be NULL. */
int fedg_atmos_calc_diagvar(fedg_ctx *ctx, const char *field_name, double *var_out, const double *DDENS_,
                            const double *MOMX_, const double *MOMY_, const double *MOMZ_,
                            const double *PRES_, int nq, double *const *QTRC, const double *TRACER_MASS,
                            const double *TRACER_CV, const double *TRACER_ENGI0, const double *DENS_hyd,
                            const double *PRES_hyd, const double *Rtot, const double *CVtot,
                            const double *CPtot, const double *zlev, const double *pos_z);

/* atm_dyn_dgm_trcadvect3d_heve_cal_alphdens_advtest
 * (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:407-445): the damping
 * coefficients of the tracer-only advection test, alpha = max(|u.n| of both
 * sides), alph_dens_M/_P (NfpTot,Ne) = alpha dens Gsqrt of each side. */
int fedg_trcadvect_cal_alphdens_advtest(fedg_ctx *ctx, double *alph_dens_M, double *alph_dens_P,
                                        const double *DDENS_, const double *MOMX_, const double *MOMY_,
                                        const double *MOMZ_, const double *DENS_hyd);

/* atm_dyn_dgm_tracer_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:232-270)
 * with ModalFilter_tracer (ELM/scale_element_operation_tensorprod3D.F90, the
 * x, y, z passes of element_operation_kernel_matvec_ModalFilter_P*): per
 * element, weight = Gsqrt (DENS_hyd + DDENS_), QTRC_ = filter(weight QTRC_) /
 * weight, in place.  MFilter_tracer_h1D, MFilter_tracer_v1D: host (n1,n1)
 * column-major (fedg_modalfilter_matrix with the tracer filter's parameters,
 * PARAM_ATMOS_DYN_TRACER_MODALFILTER). */
int fedg_tracer_modalfilter_apply(fedg_ctx *ctx, double *QTRC_, const double *DENS_hyd, const double *DDENS_,
                                  const double *MFilter_tracer_h1D, const double *MFilter_tracer_v1D);

/* timeint_rk%Advance_trcvar for the low-storage schemes
 * (CMN/scale_timeint_rk.F90:1390-1498, rk_advance_trcvar_low_storage): the
 * density-weighted RK update of one tracer at stage nowstage (1-based) of
 * nstage.  coef_sig_ex, coef_gam_ex: (nstage+1, nstage) column-major host
 * arrays and coef_c_ex (nstage), the timeint_rk object's tables; dt its step.
 * q, var0, varTmp, tend (the stage's tendency), DDENS, DDENS0, DENS_hyd:
 * device (Np,NeA), interior elements updated. */
int fedg_trcvar_rk_advance(fedg_ctx *ctx, int nowstage, int nstage, const double *coef_sig_ex,
                           const double *coef_gam_ex, const double *coef_c_ex, double dt, double *q,
                           double *var0, double *varTmp, const double *tend, const double *DDENS,
                           const double *DDENS0, const double *DENS_hyd);

/* atm_dyn_dgm_nonhydro3d_rhot_hevi_cal_tend
 * (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:289-482 with its numerical
 * flux FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi_numflux.F90:232-416): the
 * explicit tendency of the HEVI scheme; same arguments as fedg_heve_cal_tend
 * (the cal_tend_ex interface, FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:152-199). */
int fedg_hevi_cal_tend(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                       double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                       const double *MOMY_, const double *MOMZ_, const double *DRHOT_,
                       const double *DPRES_, const double *DENS_hyd, const double *PRES_hyd,
                       const double *PRES_hyd_ref, const double *THERM_hyd, const double *CORIOLIS,
                       const double *Rtot, const double *CVtot, const double *CPtot,
                       const double *DPhydDx, const double *DPhydDy);

/* atm_dyn_dgm_nonhydro3d_rhot_hevi_cal_vi
 * (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965, with
 * FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi_common_2.F90:111-1328 and the
 * block LU of FDS/scale_atm_dyn_dgm_hevi_common_linalgebra.F90:2142-2445): the
 * implicit tendencies of the vertical terms, one Newton iteration from var0
 * (DDENS0_ .. DRHOT0_ = q^n) towards q = q* + impl_fac T_v(q), q* the stage
 * state; impl_fac = 0 returns T_v(var0).  Arguments in the order of the
 * abstract interface atm_dyn_nonhydro3d_cal_vi
 * (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:201-250) minus the derived
 * types; fields (Np,NeA) on the device; the mesh must be one box column
 * stack (elements numbered horizontally first). */
int fedg_hevi_cal_vi(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt,
                     double *RHOT_dt, const double *DDENS_, const double *MOMX_, const double *MOMY_,
                     const double *MOMZ_, const double *DRHOT_, const double *DENS_hyd,
                     const double *PRES_hyd, const double *DDENS0_, const double *MOMX0_,
                     const double *MOMY0_, const double *MOMZ0_, const double *DRHOT0_,
                     const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac,
                     double dt);

/* No counterpart in the reference (which stops after one iteration,
 * ..._rhot_hevi.F90:885): fedg_hevi_cal_vi iterated until the residual norm
 * ||G_k|| <= rtol ||G_0||, or at the rounding floor of G = impl_fac T - q +
 * q* (||G_k|| <= 4 eps ||q*||, q* the stage state, or an update that has
 * not halved ||G||), or after max_it iterations; resid (host, max_it) gets the
 * norms, *iterations the number of Newton updates applied. */
int fedg_hevi_cal_vi_newton(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                            double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                            const double *MOMY_, const double *MOMZ_, const double *DRHOT_,
                            const double *DENS_hyd, const double *PRES_hyd, const double *DDENS0_,
                            const double *MOMX0_, const double *MOMY0_, const double *MOMZ0_,
                            const double *DRHOT0_, const double *Rtot, const double *CVtot,
                            const double *CPtot, double impl_fac, double dt, int max_it, double rtol,
                            double *resid, int *iterations);

/* No counterpart in the reference: fedg_hevi_cal_vi_newton (same arguments,
 * same results bit for bit) for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Its implicit-solve workspace is a fixed 64 MiB device array of
 * the library instead of a hipMalloc'd one; the call allocates and frees
 * nothing.  Meshes whose workspace (13 Np Ne + 2 Nfp Ne doubles, each part
 * rounded to 256 B, plus 2 KiB) exceeds 64 MiB are rejected with
 * FEDG_ERR_UNSUPPORTED (at p = 7 about 1200 elements).  Every call on one
 * device shares the workspace: calls that overlap in time on the same device
 * (several host threads, several streams) race and must be serialised by the
 * caller. */
int fedg_hevi_cal_vi_noalloc(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                             double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                             const double *MOMY_, const double *MOMZ_, const double *DRHOT_,
                             const double *DENS_hyd, const double *PRES_hyd, const double *DDENS0_,
                             const double *MOMX0_, const double *MOMY0_, const double *MOMZ0_,
                             const double *DRHOT0_, const double *Rtot, const double *CVtot,
                             const double *CPtot, double impl_fac, double dt, int max_it, double rtol,
                             double *resid, int *iterations);

/* atm_dyn_dgm_nonhydro3d_common_calc_phyd_hgrad_lc
 * (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:624-777), halos of PRES_hyd and
 * PRES_hyd_ref filled. */
int fedg_calc_phyd_hgrad(fedg_ctx *ctx, double *DPhydDx, double *DPhydDy, const double *PRES_hyd,
                         const double *PRES_hyd_ref);

/* add_phy_tend (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:1098-1178, the
 * ENTOT_CONSERVE_SCHEME_FLAG = .false. branch, which the rhotheta form of the
 * dynamics takes): the physics tendencies added to the dynamics tendencies
 * of elements 1..Ne in place, DENS_dt += DENS_tp, MOMX/Y/Z_dt += MOMX/Y/Z_tp,
 * RHOT_dt = RHOT_dt + RHOT_tp + RHOH_p / (CPtot EXNER) with EXNER =
 * (PRES / PRES00)^(Rtot / CPtot).  Device (Np,NeA) fields. */
int fedg_add_phy_tend(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt,
                      double *RHOT_dt, const double *DENS_tp, const double *MOMX_tp, const double *MOMY_tp,
                      const double *MOMZ_tp, const double *RHOT_tp, const double *RHOH_p, const double *PRES,
                      const double *Rtot, const double *CPtot);

/* ------------------------------------------------------------------ */
/* Time integration (timeint_rk, CMN/scale_timeint_rk.F90:35-104, 135-229).
 * The integrator owns var0/varTmp/tend_buf2D_ex(/_im) on the device, sized
 * (Np,NeA,5[,nbuf]) like the reference. */
typedef struct fedg_tint fedg_tint;
/* scheme: "ERK_4s4o"/"ERK_RK4", "ERK_1s1o"/"ERK_Euler", "IMEX_ARK232", "IMEX_ARK324"
 * (CMN/scale_timeint_rk_butcher_tab.F90:60-324) */
int fedg_tint_create(fedg_ctx *ctx, const char *scheme, double dt, fedg_tint **tint);
int fedg_tint_destroy(fedg_tint *tint);
int fedg_tint_info(const fedg_tint *tint, int *nstage, int *tend_buf_size, int *imex);
/* device pointer to tend_buf2D_ex(:,:,varID,tend_buf_indmap(stage)) (1-based ids) */
double *fedg_tint_tend_ex(fedg_tint *tint, int stage, int varID);
double *fedg_tint_tend_im(fedg_tint *tint, int stage, int varID);
/* timeint_rk_store_var0_2D (CMN/scale_timeint_rk.F90:624-647) */
int fedg_tint_store_var0(fedg_tint *tint, const double *q, int varID);
/* Advance_varlist: rk_advance_general2D (CMN/scale_timeint_rk.F90:536-573, 2201-2355)
 * for nvar device fields q[i] with ids varIDs[i] at 1-based stage */
int fedg_tint_advance(fedg_tint *tint, int stage, int nvar, double *const *q, const int *varIDs);
/* StoreImplicit: rk_storeimpl_general2D (CMN/scale_timeint_rk.F90:2510-2560) */
int fedg_tint_store_implicit(fedg_tint *tint, int stage, double *q, int varID);

/* ------------------------------------------------------------------ */
/* Fused HEVE stepper: the whole AtmDynDGMDriver_nonhydro3d_update HEVE step
 * (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:614-963) with device-resident
 * state, one kernel per RK stage (halo, pressure, boundary condition,
 * tendency and RK update fused; each stage reads the state once).
 * Explicit schemes with one tendency buffer (ERK_4s4o, ERK_1s1o).
 * Base-state fields are host pointers (Np,NeA) with filled halos; CORIOLIS is
 * (Nfp_v,Ne2DA); Rtot/CVtot/CPtot (Np,NeA) or NULL for the dry-air constants. */
typedef struct fedg_heve fedg_heve;
int fedg_heve_create(fedg_ctx *ctx, const char *scheme, double dt, const double *DENS_hyd,
                     const double *PRES_hyd, const double *THERM_hyd, const double *CORIOLIS,
                     const double *DPhydDx, const double *DPhydDy, const double *Rtot,
                     const double *CVtot, const double *CPtot, fedg_heve **st);
int fedg_heve_destroy(fedg_heve *st);
/* 1 if the flat-terrain uniform-box kernel is used, 0 for the general-metric kernel */
int fedg_heve_is_flat(const fedg_heve *st);
/* Tracer coupling (QA > 0): the driver's per-stage call of
 * atm_dyn_dgm_trcadvect3d_save_massflux (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:899-915)
 * on every stage state (halo exchanged, velocity BC applied) with weights
 * b_ex(s), accumulating MFLX_x/y/z_tavg (Np,NeA) and alph_dens_M/_P
 * (NfpTot,Ne) on the device for the tracer step that follows (the first stage
 * of each step resets them).  All five device arrays, or all NULL to stop.
 * FP64 stages only.  The stepper keeps the pointers and writes through them at
 * every stage: the arrays must stay allocated until the next call or
 * fedg_heve_destroy; MFLX_* hold Np*NeA values, alph_dens_* NfpTot*Ne. */
int fedg_heve_set_massflux(fedg_heve *st, double *MFLX_x_tavg, double *MFLX_y_tavg, double *MFLX_z_tavg,
                           double *alph_dens_M, double *alph_dens_P);
/* Physics tendencies of the driver's add_phy_tend
 * (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:846-856, 1098-1178): added
 * to every stage's dynamics tendency before the RK update, with the stage's
 * PRES in EXNER.  Six device (Np,NeA) arrays (DENS_tp, MOMX_tp, MOMY_tp,
 * MOMZ_tp, RHOT_tp, RHOH_p), constant over the step as in the reference, or
 * all NULL to stop.  The stepper keeps the pointers: the arrays stay
 * allocated until the next call or fedg_heve_destroy.  FP64 stages only; at
 * p = 7 the stages then run on the general-layout kernel. */
int fedg_heve_set_phys_tend(fedg_heve *st, const double *DENS_tp, const double *MOMX_tp, const double *MOMY_tp,
                            const double *MOMZ_tp, const double *RHOT_tp, const double *RHOH_p);
/* No counterpart in the reference (one global precision switch, Mkinclude:34-35):
 * bits = 32 evaluates each stage's tendency -- fluxes, face flux, derivative
 * contractions, lift, buoyancy -- in FP32 on the perturbation variables about
 * the FP64 base state, with the pressure perturbation formed without
 * cancellation, P_hyd expm1(cp/cv log1p(DRHOT/THERM_hyd)); the state, the
 * base state and the RK update stay FP64.  Flat-terrain box, dry air.
 * bits = 64 (default): the reference's FP64 arithmetic. */
int fedg_heve_set_precision(fedg_heve *st, int bits);
/* MODALFILTER_FLAG: filter the state after every step, with Gsqrt weighting
 * (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:940-950); needs
 * fedg_ctx_setup_modalfilter first */
int fedg_heve_set_modalfilter(fedg_heve *st, int on);
/* host state in/out: q[v] (Np,Ne or more) for v = DDENS, DRHOT, MOMZ, MOMX, MOMY */
int fedg_heve_set_state(fedg_heve *st, const double *const *q_host);
int fedg_heve_get_state(fedg_heve *st, double *const *q_host);
/* device pointers of the current state (valid until the next step) */
int fedg_heve_state_dev(fedg_heve *st, double **q_dev);
/* nsteps full RK steps on the device; synchronous */
int fedg_heve_run(fedg_heve *st, int nsteps);
/* nsteps steps bracketed by HIP events on the stepper's stream: total
 * milliseconds and the mean duration of each stage kernel launch */
int fedg_heve_run_timed(fedg_heve *st, int nsteps, double *ms_total, double *ms_per_stage_kernel);
/* PRES and DPRES of the current state (host (Np,Ne)), the reference's
 * post-step calc_pressure (:954-959) */
int fedg_heve_get_pressure(fedg_heve *st, double *PRES_host, double *DPRES_host);

/* ------------------------------------------------------------------ */
/* HEVI step driver: AtmDynDGMDriver_nonhydro3d_update with hevi_flag
 * (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:694-870): StoreVar0, then per
 * IMEX stage cal_vi (fedg_hevi_cal_vi; Newton iterations max_it, tolerance
 * rtol, the reference: 1 and 0), StoreImplicit, halo exchange, DRHOT2PRES,
 * boundary conditions, the explicit HEVI tendency and Advance, on the
 * context's stream with state resident on the device.  Base-state arguments
 * are host (Np,NeA) arrays with filled halos (CORIOLIS (Nfp_v,Ne2DA)). */
typedef struct fedg_hevi fedg_hevi;
int fedg_hevi_create(fedg_ctx *ctx, const char *scheme, double dt, const double *DENS_hyd,
                     const double *PRES_hyd, const double *THERM_hyd, const double *CORIOLIS,
                     const double *DPhydDx, const double *DPhydDy, const double *Rtot, const double *CVtot,
                     const double *CPtot, int max_it, double rtol, fedg_hevi **out);
int fedg_hevi_destroy(fedg_hevi *st);
/* q: 5 host arrays (Np,Ne) in variable-id order */
int fedg_hevi_set_state(fedg_hevi *st, const double *const *q);
int fedg_hevi_get_state(fedg_hevi *st, double *const *q);
/* The tracer coupling of fedg_heve_set_massflux for the HEVI stepper: the
 * same per-stage save_massflux with weights b_ex(s) horizontally and b_im(s)
 * vertically and the HEVI form of the damping coefficients (no |nz| in Gnn). */
int fedg_hevi_set_massflux(fedg_hevi *st, double *MFLX_x_tavg, double *MFLX_y_tavg, double *MFLX_z_tavg,
                           double *alph_dens_M, double *alph_dens_P);
/* The physics tendencies of fedg_heve_set_phys_tend for the HEVI stepper:
 * added to every stage's explicit tendency (the reference's add_phy_tend
 * after cal_tend_ex). */
int fedg_hevi_set_phys_tend(fedg_hevi *st, const double *DENS_tp, const double *MOMX_tp, const double *MOMY_tp,
                            const double *MOMZ_tp, const double *RHOT_tp, const double *RHOH_p);
int fedg_hevi_run(fedg_hevi *st, int nsteps);
/* nsteps steps bracketed by HIP events: total milliseconds and the mean
 * duration per stage of the implicit solve and of the explicit tendency */
int fedg_hevi_run_timed(fedg_hevi *st, int nsteps, double *ms_total, double *ms_implicit_per_stage,
                        double *ms_explicit_per_stage);
/* Newton iterations summed over the implicit stages run so far, and their number */
int fedg_hevi_newton_stats(const fedg_hevi *st, long *iterations, long *calls);
/* as fedg_heve_set_modalfilter */
int fedg_hevi_set_modalfilter(fedg_hevi *st, int on);
/* No counterpart in the reference: the explicit tendency of the HEVI step on
 * the flat-terrain box at p = 7 with dry air, evaluated in FP32 perturbation
 * form about the FP64 base state (bits = 32; the state, the IMEX combination
 * and the implicit Newton solve stay FP64) or in FP64 (bits = 64, the
 * default, bit for bit the reference's).  bits = 32 elsewhere:
 * FEDG_ERR_UNSUPPORTED. */
int fedg_hevi_set_precision(fedg_hevi *st, int bits);


/* ------------------------------------------------------------------ */
/* Learned initial iterate of the HEVI Newton solve.  No counterpart in the
 * reference (its Newton loop starts from q^n and stops after one iteration,
 * FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:868-917).  A dense ReLU
 * network n_in -> n_hidden (x n_hidden_layers) -> n_out, BF16 products on
 * the matrix cores with FP32 accumulation, FP32 master weights, Adam.  For a
 * vertical node line of L = NeZ (p+1) nodes: n_in = 10 L + 1 (the line's
 * stage state q* / s per variable, then impl_fac T_ex / t per variable, then
 * impl_fac/dt; T_ex is the previous stage's explicit tendency, zero at the
 * first stage), n_out = 5 L ((q - q*) / d of the converged iterate q).
 * Variables in VID order, line node t = kz (p+1) + k (k the node level in
 * element kz); s, t, d: per-variable scales of the state, the tendency
 * feature and the increment.
 * Device arrays are fp32, row-major (batch x features). */
typedef struct fedg_mlp fedg_mlp;
int fedg_mlp_create(int n_in, int n_hidden, int n_hidden_layers, int n_out, unsigned seed, fedg_mlp **out);
int fedg_mlp_destroy(fedg_mlp *m);
int fedg_mlp_param_count(const fedg_mlp *m, long *n);
/* host copies: every layer's W (n_out x n_in row-major), then every bias */
int fedg_mlp_get_params(fedg_mlp *m, float *params_host, float *grads_host);
int fedg_mlp_set_params(fedg_mlp *m, const float *params_host);
int fedg_mlp_forward(fedg_mlp *m, const float *X_dev, float *Y_dev, int batch);
/* mean squared error against T_dev and the gradient of every parameter */
int fedg_mlp_loss_grad(fedg_mlp *m, const float *X_dev, const float *T_dev, int batch, float *loss);
int fedg_mlp_adam_step(fedg_mlp *m, float lr, float beta1, float beta2, float eps);
/* Use mlp (or none: NULL) for the initial iterate of every implicit stage of
 * the HEVI stepper; lines whose predicted iterate has a larger residual than
 * q^n start from q^n.  scale: 15 values, s, t, then d (VID order each).  The
 * tolerance stays relative to the residual norm at q^n. */
int fedg_hevi_set_guess(fedg_hevi *st, fedg_mlp *mlp, const double *scale);
/* fedg_hevi_cal_vi_newton started from the network's prediction for a host
 * that drives its own stage loop: Tex (5 device fields in VID order, or NULL)
 * is the previous stage's explicit tendency; *fallbacks receives the number
 * of lines that started from q^n; the tolerance stays relative to ||G(q^n)||. */
int fedg_hevi_cal_vi_guess(fedg_ctx *ctx, double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt,
                           double *RHOT_dt, const double *DDENS_, const double *MOMX_, const double *MOMY_,
                           const double *MOMZ_, const double *DRHOT_, const double *DENS_hyd,
                           const double *PRES_hyd, const double *DDENS0_, const double *MOMX0_,
                           const double *MOMY0_, const double *MOMZ0_, const double *DRHOT0_,
                           const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac,
                           double dt, int max_it, double rtol, fedg_mlp *mlp, const double *scale,
                           const double *const *Tex, double *resid, int *iterations, long *fallbacks);
int fedg_hevi_guess_stats(const fedg_hevi *st, long *lines, long *fallbacks);
/* ||G(q^n)||^2 and ||G||^2 at the iterates the lines started from (the
 * prediction or, where it was worse, q^n), each summed over the solves */
int fedg_hevi_guess_residuals(const fedg_hevi *st, double *r2_plain, double *r2_start);
/* Collect training pairs from the following steps: per implicit stage,
 * per_stage lines' (features, converged targets) appended to X_dev / Y_dev
 * until capacity rows are stored. */
int fedg_hevi_harvest(fedg_hevi *st, float *X_dev, float *Y_dev, int capacity, int per_stage, unsigned seed);
int fedg_hevi_harvest_count(const fedg_hevi *st, int *count);

#ifdef __cplusplus
}
#endif
#endif /* FEDG_H */
