/*
 * CPU oracle, HEVI part: literal C restatement of FElib's horizontally
 * explicit / vertically implicit step.  TEST INFRASTRUCTURE ONLY (see
 * fedg_oracle.h): the product never calls it.
 *
 *   explicit tendency   FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:289-482
 *   explicit face flux  FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi_numflux.F90:232-416
 *   implicit tendency   FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965
 *                       FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi_common_2.F90:111-1328
 *   block LU kernels    FDS/scale_atm_dyn_dgm_hevi_common_linalgebra.F90:2142-2445
 *   IMEX driver         FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:694-870
 *
 * Same loop order and operation order as the Fortran (-ffp-contract=off).
 * Layout of the implicit solve: one column per horizontal node line
 * (ij = 0..Nfp-1 of horizontal element ke2D), NeZ elements stacked, N = n1
 * nodes per element; the unknowns of one element are ordered (var, pv) with
 * the variable fastest, variables DENS, MOMZ, RHOT (the _LC ids of
 * construct_matbnd); matrices are column-major, A[r + 3N*c].
 *
 * Operator entries of the implicit matrix that the reference takes from 3-D
 * matrices are taken from the 1-D tables: elem%Dx3 on a vertical line holds
 * the 1-D derivative entries exactly (ELM/scale_element_hexahedral.F90:275);
 * elem%Lift on a vertical face is Lift_mat (the tensor-product Lift_mat is
 * copied from elem%Lift, ELM/scale_element_operation_tensorprod3D.F90:5540-5553);
 * IntrpMat_VPOrdM1 (a 3-D modal truncation, nonhydro3d_common.F90:112-124)
 * equals the 1-D truncation on a vertical line in exact arithmetic only, so
 * its last bits are not reproduced (parity unpinned at the ulp level for that
 * one coefficient; the tests pin the assembled matrix against the Jacobian).
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "fedg_oracle.h"
#include "oracle_ops.h"

#define opow oracle_opow

#define DENS_VID 0
#define RHOT_VID 1
#define MOMZ_VID 2
#define MOMX_VID 3
#define MOMY_VID 4

static inline double fmax_f(double a, double b) { return a > b ? a : b; }

/* ---------------------------------------------------------------------- */
/* Explicit part: ..._rhot_hevi_numflux.F90:232-416 (get_generalvc) */
static void hevi_numflux(const oracle_mesh *m, double *del_flux, const double *DDENS_, const double *MOMX_,
                         const double *MOMY_, const double *MOMZ_, const double *DRHOT_, const double *DPRES,
                         const double *DENS_hyd, const double *PRES_hyd, const double *THERM_hyd) {
  const int NfpTot = m->NfpTot, Ne = m->Ne;
  const long NpNeA = (long)m->Np * m->NeA;
  const long nfall = (long)NfpTot * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double gamm = m->CPdry / m->CVdry;
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < Ne; ++ke) {
    for (int fp = 0; fp < NfpTot; ++fp) {
      long i = fp + (long)NfpTot * ke;
      long iM = m->vmapM[i] - 1, iP = m->vmapP[i] - 1;
      double nx = m->normal_fn[i], ny = m->normal_fn[i + nfall], nz = m->normal_fn[i + 2 * nfall];
      double Gs_M = m->Gsqrt[iM], Gs_P = m->Gsqrt[iP];
      double RGsqrtV_M = 1.0 / Gs_M, RGsqrtV_P = 1.0 / Gs_P; /* GsqrtV_ = Gsqrt_ here */
      double G13_M = G13[iM], G13_P = G13[iP], G23_M = G23[iM], G23_P = G23[iP];
      double GDD_M = Gs_M * DDENS_[iM], GDD_P = Gs_P * DDENS_[iP];
      double GMX_M = Gs_M * MOMX_[iM], GMX_P = Gs_P * MOMX_[iP];
      double GMY_M = Gs_M * MOMY_[iM], GMY_P = Gs_P * MOMY_[iP];
      double GMZ_M = Gs_M * MOMZ_[iM], GMZ_P = Gs_P * MOMZ_[iP];
      double GDT_M = Gs_M * DRHOT_[iM], GDT_P = Gs_P * DRHOT_[iP];
      double Phyd_M = PRES_hyd[iM], Phyd_P = PRES_hyd[iP];
      double dp_M = DPRES[iM], dp_P = DPRES[iP];
      double GDens_M = GDD_M + Gs_M * DENS_hyd[iM];
      double GDens_P = GDD_P + Gs_P * DENS_hyd[iP];
      double GRhot_M = Gs_M * THERM_hyd[iM] + GDT_M;
      double GRhot_P = Gs_P * THERM_hyd[iP] + GDT_P;
      double Velh_M = (GMX_M * nx + GMY_M * ny) / GDens_M;
      double Velh_P = (GMX_P * nx + GMY_P * ny) / GDens_P;
      double Vel_M = Velh_M + (((GMZ_M * RGsqrtV_M + G13_M * GMX_M + G23_M * GMY_M) * nz)) / GDens_M;
      double Vel_P = Velh_P + (((GMZ_P * RGsqrtV_P + G13_P * GMX_P + G23_P * GMY_P) * nz)) / GDens_P;
      double swV = 1.0 - nz * nz;
      double alpha = swV * fmax_f(sqrt(gamm * (Phyd_M + dp_M) * Gs_M / GDens_M) + fabs(Vel_M),
                                  sqrt(gamm * (Phyd_P + dp_P) * Gs_P / GDens_P) + fabs(Vel_P));
      double *df = del_flux + (long)NfpTot * 5 * ke;
      double tmp1 = m->Fscale[i] * 0.5;
      double tmp2 = -alpha * (GDD_P - GDD_M);
      df[fp + NfpTot * DENS_VID] = tmp1 * (GDens_P * Velh_P - GDens_M * Velh_M + tmp2);
      tmp2 = -alpha * (GDT_P - GDT_M);
      df[fp + NfpTot * RHOT_VID] = tmp1 * (GRhot_P * Velh_P - GRhot_M * Velh_M + tmp2);
      tmp2 = -alpha * (GMZ_P - GMZ_M);
      df[fp + NfpTot * MOMZ_VID] = tmp1 * (GMZ_P * Vel_P - GMZ_M * Vel_M + tmp2);
      double tmp3 = Gs_P * dp_P;
      double tmp4 = Gs_M * dp_M;
      double mx = (nx + G13_P * nz) * tmp3 - (nx + G13_M * nz) * tmp4;
      double my = (ny + G23_P * nz) * tmp3 - (ny + G23_M * nz) * tmp4;
      tmp2 = -alpha * (GMX_P - GMX_M);
      df[fp + NfpTot * MOMX_VID] = tmp1 * (GMX_P * Vel_P - GMX_M * Vel_M + mx + tmp2);
      tmp2 = -alpha * (GMY_P - GMY_M);
      df[fp + NfpTot * MOMY_VID] = tmp1 * (GMY_P * Vel_P - GMY_M * Vel_M + my + tmp2);
    }
  }
}

void oracle_hevi_numflux(const oracle_mesh *m, double *del_flux, const double *DDENS_, const double *MOMX_,
                         const double *MOMY_, const double *MOMZ_, const double *DRHOT_, const double *DPRES,
                         const double *DENS_hyd, const double *PRES_hyd, const double *THERM_hyd) {
  hevi_numflux(m, del_flux, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES, DENS_hyd, PRES_hyd, THERM_hyd);
}

/* ..._rhot_hevi.F90:289-482 (cal_tend) */
void oracle_hevi_cal_tend(const oracle_mesh *m, double *DENS_dt, double *MOMX_dt, double *MOMY_dt,
                          double *MOMZ_dt, double *RHOT_dt, const double *DDENS_, const double *MOMX_,
                          const double *MOMY_, const double *MOMZ_, const double *DRHOT_, const double *DPRES_,
                          const double *DENS_hyd, const double *PRES_hyd, const double *PRES_hyd_ref,
                          const double *THERM_hyd, const double *CORIOLIS, const double *Rtot,
                          const double *CVtot, const double *CPtot, const double *DPhydDx, const double *DPhydDy) {
  (void)PRES_hyd_ref; (void)Rtot; (void)CVtot; (void)CPtot;
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Nfp_v = m->Nfp_v, Ne = m->Ne;
  const long NpNeA = (long)Np * m->NeA, NpNe = (long)Np * Ne;
  const double *GI3_1 = m->GI3, *GI3_2 = m->GI3 + NpNeA;
  double *del_flux = (double *)malloc(sizeof(double) * (size_t)NfpTot * 5 * Ne);
  hevi_numflux(m, del_flux, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES_, DENS_hyd, PRES_hyd, THERM_hyd);
#pragma omp parallel
  {
    double *Flux = (double *)calloc((size_t)Np * 15, sizeof(double));
    double *DFlux = (double *)malloc(sizeof(double) * Np * 20);
    double *RDENS = (double *)malloc(sizeof(double) * Np);
    double *RGsqrtV = (double *)malloc(sizeof(double) * Np);
    double *RGsqrt = (double *)malloc(sizeof(double) * Np);
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      const int ke2d = m->EMap3Dto2D[ke] - 1;
      const long o = (long)Np * ke;
      for (int p = 0; p < Np; ++p) {
        double GsqrtV = m->Gsqrt[o + p] / m->GsqrtH[(m->IndexH2Dto3D[p] - 1) + (long)Nfp_v * ke2d];
        RGsqrtV[p] = 1.0 / GsqrtV;
        RGsqrt[p] = 1.0 / m->Gsqrt[o + p];
        RDENS[p] = 1.0 / (DDENS_[o + p] + DENS_hyd[o + p]);
      }
#define FL(p, d, v) Flux[(p) + Np * ((d) + 3 * (v))]
#define DF(p, d, v) DFlux[(p) + Np * ((d) + 4 * (v))]
      for (int p = 0; p < Np; ++p) {
        double Gs = m->Gsqrt[o + p];
        FL(p, 0, DENS_VID) = Gs * MOMX_[o + p];
        FL(p, 1, DENS_VID) = Gs * MOMY_[o + p];
        FL(p, 2, DENS_VID) = Gs * (MOMZ_[o + p] * RGsqrtV[p] + GI3_1[o + p] * MOMX_[o + p] + GI3_2[o + p] * MOMY_[o + p]);
      }
      for (int p = 0; p < Np; ++p) {
        double pt_ = (THERM_hyd[o + p] + DRHOT_[o + p]) * RDENS[p];
        FL(p, 0, RHOT_VID) = FL(p, 0, DENS_VID) * pt_;
        FL(p, 1, RHOT_VID) = FL(p, 1, DENS_VID) * pt_;
        /* Flux(:,3,RHOT_VID) is left unset in the reference and its Dz is unused */
        double w_ = MOMZ_[o + p] * RDENS[p];
        FL(p, 0, MOMZ_VID) = FL(p, 0, DENS_VID) * w_;
        FL(p, 1, MOMZ_VID) = FL(p, 1, DENS_VID) * w_;
        FL(p, 2, MOMZ_VID) = FL(p, 2, DENS_VID) * w_;
      }
      for (int p = 0; p < Np; ++p) {
        double GsqrtDPRES_ = m->Gsqrt[o + p] * DPRES_[o + p];
        double u_ = MOMX_[o + p] * RDENS[p];
        double v_ = MOMY_[o + p] * RDENS[p];
        FL(p, 0, MOMX_VID) = FL(p, 0, DENS_VID) * u_ + GsqrtDPRES_;
        FL(p, 1, MOMX_VID) = FL(p, 1, DENS_VID) * u_;
        FL(p, 2, MOMX_VID) = FL(p, 2, DENS_VID) * u_ + GsqrtDPRES_ * GI3_1[o + p];
        FL(p, 0, MOMY_VID) = FL(p, 0, DENS_VID) * v_;
        FL(p, 1, MOMY_VID) = FL(p, 1, DENS_VID) * v_ + GsqrtDPRES_;
        FL(p, 2, MOMY_VID) = FL(p, 2, DENS_VID) * v_ + GsqrtDPRES_ * GI3_2[o + p];
      }
      const double *dfe = del_flux + (long)NfpTot * 5 * ke;
      for (int v = 0; v < 5; ++v) op_lift(n1, m->Lift_mat, dfe + (long)NfpTot * v, &DF(0, 3, v));
      for (int v = 0; v < 5; ++v) {
        op_dirX(n1, m->D1D, &FL(0, 0, v), &DF(0, 0, v));
        op_dirY(n1, m->D1D, &FL(0, 1, v), &DF(0, 1, v));
        op_dirZ(n1, m->D1D, &FL(0, 2, v), &DF(0, 2, v));
      }
      const double *E11 = m->Escale + o, *E22 = m->Escale + o + NpNe * 4, *E33 = m->Escale + o + NpNe * 8;
      for (int p = 0; p < Np; ++p) {
        DENS_dt[o + p] = -(E11[p] * DF(p, 0, DENS_VID) + E22[p] * DF(p, 1, DENS_VID) + DF(p, 3, DENS_VID)) * RGsqrt[p];
        RHOT_dt[o + p] = -(E11[p] * DF(p, 0, RHOT_VID) + E22[p] * DF(p, 1, RHOT_VID) + DF(p, 3, RHOT_VID)) * RGsqrt[p];
      }
      for (int p = 0; p < Np; ++p)
        MOMZ_dt[o + p] = -(E11[p] * DF(p, 0, MOMZ_VID) + E22[p] * DF(p, 1, MOMZ_VID) + E33[p] * DF(p, 2, MOMZ_VID) +
                           DF(p, 3, MOMZ_VID)) * RGsqrt[p];
      for (int p = 0; p < Np; ++p) {
        double cor = CORIOLIS[(m->IndexH2Dto3D[p] - 1) + (long)Nfp_v * ke2d];
        MOMX_dt[o + p] = -DPhydDx[o + p] + cor * MOMY_[o + p];
        MOMY_dt[o + p] = -DPhydDy[o + p] - cor * MOMX_[o + p];
      }
      for (int p = 0; p < Np; ++p) {
        MOMX_dt[o + p] = MOMX_dt[o + p] - (E11[p] * DF(p, 0, MOMX_VID) + E22[p] * DF(p, 1, MOMX_VID) +
                                           E33[p] * DF(p, 2, MOMX_VID) + DF(p, 3, MOMX_VID)) * RGsqrt[p];
        MOMY_dt[o + p] = MOMY_dt[o + p] - (E11[p] * DF(p, 0, MOMY_VID) + E22[p] * DF(p, 1, MOMY_VID) +
                                           E33[p] * DF(p, 2, MOMY_VID) + DF(p, 3, MOMY_VID)) * RGsqrt[p];
      }
#undef FL
#undef DF
    }
    free(Flux); free(DFlux); free(RDENS); free(RGsqrtV); free(RGsqrt);
  }
  free(del_flux);
}

/* ---------------------------------------------------------------------- */
/* Implicit part */

typedef struct {
  const oracle_mesh *m;
  int N, Nfp, Ne2D, NeZ, n3; /* n3 = 3N */
  double impl_fac, dt;
  const double *const *qs;    /* q* (stage state), VID order, (Np,NeA) */
  const double *const *q0;    /* var0, VID order */
  double *const *pv;          /* PROG_VARS (iterate), VID order, (Np,Ne) */
  const double *DENS_hyd, *PRES_hyd, *Rtot, *CVtot, *CPtot;
  double *GsqrtV;             /* (Np,Ne) */
  double *alph;               /* (NfpTot,Ne) */
  /* per node coefficients of eval_Ax */
  double *DENSc, *WT, *POT, *DPDRHOT; /* (Np,Ne) */
  double *b;                  /* var3 rhs (Np,Ne,3): node-major (3*p + v) within an element */
  double *buv;                /* uv rhs (Np,Ne,2) */
} hevi_ctx;

/* vertical face maps of LocalMesh3D_getVmapZ3D (MSH/scale_localmesh_3d.F90:190-241),
 * 0-based node indices into (Np,Ne); horizontal faces map to themselves */
static inline long vz_M(const hevi_ctx *h, int ke, int fp) {
  const int Nfp = h->Nfp, N = h->N;
  const int f = fp / Nfp, a = fp % Nfp;
  long Np = (long)Nfp * N;
  int pm;
  switch (f) {
    case 0: pm = (a % N) + Nfp * (a / N); break;                    /* y-: (i,0,k) */
    case 1: pm = (N - 1) + N * (a % N) + Nfp * (a / N); break;      /* x+: (N-1,j,k) */
    case 2: pm = (a % N) + N * (N - 1) + Nfp * (a / N); break;      /* y+: (i,N-1,k) */
    case 3: pm = N * (a % N) + Nfp * (a / N); break;                /* x-: (0,j,k) */
    case 4: pm = a; break;                                         /* z-: (i,j,0) */
    default: pm = a + Nfp * (N - 1); break;                        /* z+: (i,j,N-1) */
  }
  return Np * ke + pm;
}
static inline long vz_P(const hevi_ctx *h, int ke, int fp) {
  const int Nfp = h->Nfp, N = h->N, Ne2D = h->Ne2D;
  const int f = fp / Nfp, a = fp % Nfp;
  const int kz = ke / Ne2D;
  long Np = (long)Nfp * N;
  if (f == 4 && kz > 0) return Np * (ke - Ne2D) + a + Nfp * (N - 1);
  if (f == 5 && kz < h->NeZ - 1) return Np * (ke + Ne2D) + a;
  return vz_M(h, ke, fp);
}

/* vi_cal_del_flux_dyn_uv (..._common_2.F90:1015-1158): alph from var0 and
 * the horizontal-momentum face jumps of the iterate */
static void hevi_del_flux_uv(hevi_ctx *h, double *dfuv) {
  const oracle_mesh *m = h->m;
  const int NfpTot = m->NfpTot, Ne = m->Ne, Nfp_v = m->Nfp_v;
  const long NpNeA = (long)m->Np * m->NeA, nfall = (long)NfpTot * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double gamm = m->CPdry / m->CVdry, rgamm = m->CVdry / m->CPdry;
  const double PRES0ovRdry = m->PRES00 / m->Rdry, rP0 = 1.0 / m->PRES00;
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < Ne; ++ke) {
    const int ke2D = m->EMap3Dto2D[ke] - 1;
    for (int fp = 0; fp < NfpTot; ++fp) {
      const long i = fp + (long)NfpTot * ke;
      const long iM = vz_M(h, ke, fp), iP = vz_P(h, ke, fp);
      const double nz = m->normal_fn[i + 2 * nfall];
      double wt0[2], pres0[2], rdens0[2], Gnn[2];
      const long ix[2] = {iM, iP};
      const double G11_ = m->G11[(m->IndexH2Dto3D_bnd[fp] - 1) + (long)Nfp_v * ke2D];
      const double G12_ = m->G12[(m->IndexH2Dto3D_bnd[fp] - 1) + (long)Nfp_v * ke2D];
      const double G22_ = m->G22[(m->IndexH2Dto3D_bnd[fp] - 1) + (long)Nfp_v * ke2D];
      for (int s = 0; s < 2; ++s) {
        const long n = ix[s];
        const double gam = 1.0; /* lcmesh%gam of the regional mesh */
        const double rgam2 = 1.0 / (gam * gam);
        const double RGsqrtV = 1.0 / h->GsqrtV[n];
        const double MOMX0 = h->q0[MOMX_VID][n], MOMY0 = h->q0[MOMY_VID][n], DRHOT0 = h->q0[RHOT_VID][n];
        const double Rt = h->Rtot[n], CPovCV = h->CPtot[n] / h->CVtot[n];
        rdens0[s] = 1.0 / (h->DENS_hyd[n] + h->q0[DENS_VID][n]);
        wt0[s] = (h->q0[MOMZ_VID][n] * RGsqrtV + G13[n] * MOMX0 + G23[n] * MOMY0) * rdens0[s];
        const double rhot_hyd = PRES0ovRdry * opow(h->PRES_hyd[n] / m->PRES00, rgamm);
        pres0[s] = m->PRES00 * opow(Rt * rP0 * (rhot_hyd + DRHOT0), CPovCV);
        const double Gxz = rgam2 * (G11_ * G13[n] + G12_ * G23[n]);
        const double Gyz = rgam2 * (G12_ * G13[n] + G22_ * G23[n]);
        Gnn[s] = (1.0 * (RGsqrtV * RGsqrtV) + G13[n] * Gxz + G23[n] * Gyz) * fabs(nz);
      }
      const double a = (nz * nz) * fmax_f(fabs(wt0[0]) + sqrt(Gnn[0] * gamm * pres0[0] * rdens0[0]),
                                          fabs(wt0[1]) + sqrt(Gnn[1] * gamm * pres0[1] * rdens0[1]));
      h->alph[i] = a;
      const double tmp1 = -0.5 * m->Fscale[i] * a;
      dfuv[fp + NfpTot * (0 + 2 * (long)ke)] = tmp1 * (h->pv[MOMX_VID][iP] - h->pv[MOMX_VID][iM]);
      dfuv[fp + NfpTot * (1 + 2 * (long)ke)] = tmp1 * (h->pv[MOMY_VID][iP] - h->pv[MOMY_VID][iM]);
    }
  }
}

/* eval_Ax_uv (..._common_2.F90:482-586): MOMX_t, MOMY_t and the uv rhs */
static void hevi_eval_Ax_uv(hevi_ctx *h, double *MOMX_t, double *MOMY_t, int want_b) {
  const oracle_mesh *m = h->m;
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Ne = m->Ne;
  double *dfuv = (double *)malloc(sizeof(double) * (size_t)NfpTot * 2 * Ne);
  hevi_del_flux_uv(h, dfuv);
#pragma omp parallel
  {
    double *LD = (double *)malloc(sizeof(double) * Np * 2);
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      const long o = (long)Np * ke;
      op_lift(n1, m->Lift_mat, dfuv + (long)NfpTot * (0 + 2 * ke), LD);
      op_lift(n1, m->Lift_mat, dfuv + (long)NfpTot * (1 + 2 * ke), LD + Np);
      for (int p = 0; p < Np; ++p) {
        const double RGsqrtV = 1.0 / h->GsqrtV[o + p];
        MOMX_t[o + p] = -LD[p] * RGsqrtV;
        MOMY_t[o + p] = -LD[p + Np] * RGsqrtV;
      }
      if (want_b)
        for (int p = 0; p < Np; ++p) {
          h->buv[2 * (o + p) + 0] = h->impl_fac * MOMX_t[o + p] - h->pv[MOMX_VID][o + p] + h->qs[MOMX_VID][o + p];
          h->buv[2 * (o + p) + 1] = h->impl_fac * MOMY_t[o + p] - h->pv[MOMY_VID][o + p] + h->qs[MOMY_VID][o + p];
        }
    }
    free(LD);
  }
  free(dfuv);
}

/* vi_cal_del_flux_dyn (..._common_2.F90:1161-1328) */
static void hevi_del_flux(hevi_ctx *h, double *df) {
  const oracle_mesh *m = h->m;
  const int NfpTot = m->NfpTot, Ne = m->Ne;
  const long NpNeA = (long)m->Np * m->NeA, nfall = (long)NfpTot * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double rgamm = m->CVdry / m->CPdry;
  const double PRES0ovRdry = m->PRES00 / m->Rdry, rP0 = 1.0 / m->PRES00;
  double *const *PV = h->pv;
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < Ne; ++ke) {
    const int kz = ke / h->Ne2D;
    for (int fp = 0; fp < NfpTot; ++fp) {
      const long i = fp + (long)NfpTot * ke;
      const long iM = vz_M(h, ke, fp), iP = vz_P(h, ke, fp);
      const long ix[2] = {iM, iP};
      double DD[2], MZ[2], MW[2], DT[2], pott[2], dpres[2], GsqrtV_[2];
      for (int s = 0; s < 2; ++s) {
        const long n = ix[s];
        GsqrtV_[s] = h->GsqrtV[n];
        DD[s] = PV[DENS_VID][n];
        MZ[s] = PV[MOMZ_VID][n];
        DT[s] = PV[RHOT_VID][n];
        const double dens = h->DENS_hyd[n] + DD[s];
        const double rhot_hyd = PRES0ovRdry * opow(h->PRES_hyd[n] / m->PRES00, rgamm);
        pott[s] = (rhot_hyd + DT[s]) / dens;
        dpres[s] = m->PRES00 * opow(h->Rtot[n] * rP0 * dens * pott[s], h->CPtot[n] / h->CVtot[n]);
        dpres[s] = dpres[s] - h->PRES_hyd[n];
        MW[s] = MZ[s] + GsqrtV_[s] * G13[n] * PV[MOMX_VID][n] + GsqrtV_[s] * G23[n] * PV[MOMY_VID][n];
      }
      if ((kz == 0 || kz == h->NeZ - 1) && iM == iP) {
        MZ[1] = -MZ[0] - 2.0 * GsqrtV_[0] * (G13[iM] * PV[MOMX_VID][iM] + G23[iM] * PV[MOMY_VID][iM]);
        MW[1] = -MW[0];
      }
      const double nz = m->normal_fn[i + 2 * nfall];
      const double a = h->alph[i];
      const double tmp1 = 0.5 * m->Fscale[i];
      double *d = df + (long)NfpTot * 3 * ke;
      d[fp + NfpTot * 0] = tmp1 * (+(MW[1] - MW[0]) * nz - a * (DD[1] - DD[0]));
      d[fp + NfpTot * 1] = tmp1 * (+(dpres[1] - dpres[0]) * nz - a * (MZ[1] - MZ[0]));
      d[fp + NfpTot * 2] = tmp1 * (+(pott[1] * MW[1] - pott[0] * MW[0]) * nz - a * (DT[1] - DT[0]));
    }
  }
}

/* eval_Ax (..._common_2.F90:111-326): DENS_t, MOMZ_t, RHOT_t, the var3 rhs
 * and the linearisation coefficients DENS, WT, POT, DPDRHOT */
static void hevi_eval_Ax(hevi_ctx *h, double *DENS_t, double *MOMZ_t, double *RHOT_t, int want_b) {
  const oracle_mesh *m = h->m;
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Ne = m->Ne;
  const long NpNeA = (long)Np * m->NeA, NpNe = (long)Np * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double rgamm = m->CVdry / m->CPdry, rP0 = 1.0 / m->PRES00, P0ovR = m->PRES00 / m->Rdry;
  double *df = (double *)malloc(sizeof(double) * (size_t)NfpTot * 3 * Ne); /* DENS, MOMZ, RHOT */
  hevi_del_flux(h, df);
  double *const *PV = h->pv;
#pragma omp parallel
  {
    double *F = (double *)malloc(sizeof(double) * Np * 3);
    double *DF = (double *)malloc(sizeof(double) * Np * 6);
    double *RHOT = (double *)malloc(sizeof(double) * Np);
    double *drho = (double *)malloc(sizeof(double) * Np);
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      const long o = (long)Np * ke;
      const double *GV = h->GsqrtV + o;
      for (int p = 0; p < Np; ++p)
        F[p] = PV[MOMZ_VID][o + p] + GV[p] * G13[o + p] * PV[MOMX_VID][o + p] + GV[p] * G23[o + p] * PV[MOMY_VID][o + p];
      for (int p = 0; p < Np; ++p)
        RHOT[p] = P0ovR * opow(h->PRES_hyd[o + p] / m->PRES00, rgamm) + PV[RHOT_VID][o + p];
      for (int p = 0; p < Np; ++p) {
        const double pt_ = RHOT[p] / (PV[DENS_VID][o + p] + h->DENS_hyd[o + p]);
        F[p + 2 * Np] = pt_ * F[p];
      }
      for (int p = 0; p < Np; ++p)
        F[p + Np] = m->PRES00 * opow(h->Rtot[o + p] * rP0 * RHOT[p], h->CPtot[o + p] / h->CVtot[o + p]) -
                    h->PRES_hyd[o + p];
      /* F: 0 DENS, 1 MOMZ, 2 RHOT; DF(:,1..2,v) */
      const double *de = df + (long)NfpTot * 3 * ke;
      op_dirZ(n1, m->D1D, F, DF);
      op_lift(n1, m->Lift_mat, de + NfpTot * 0, DF + Np);
      op_dirZ(n1, m->D1D, F + 2 * Np, DF + 4 * Np);
      op_lift(n1, m->Lift_mat, de + NfpTot * 2, DF + 5 * Np);
      op_dirZ(n1, m->D1D, F + Np, DF + 2 * Np);
      op_lift(n1, m->Lift_mat, de + NfpTot * 1, DF + 3 * Np);
      const double *E33 = m->Escale + o + NpNe * 8;
      for (int p = 0; p < Np; ++p) {
        const double RGsqrtV = 1.0 / GV[p];
        DENS_t[o + p] = -(E33[p] * DF[p] + DF[p + Np]) * RGsqrtV;
        RHOT_t[o + p] = -(E33[p] * DF[p + 4 * Np] + DF[p + 5 * Np]) * RGsqrtV;
      }
      op_dirZ_tr(n1, m->IntrpMat_tr, PV[DENS_VID] + o, drho);
      for (int p = 0; p < Np; ++p) {
        const double RGsqrtV = 1.0 / GV[p];
        MOMZ_t[o + p] = -(E33[p] * DF[p + 2 * Np] + DF[p + 3 * Np]) * RGsqrtV - m->GRAV * drho[p];
      }
      if (want_b) {
        for (int p = 0; p < Np; ++p) {
          const double dens = h->DENS_hyd[o + p] + PV[DENS_VID][o + p];
          h->DENSc[o + p] = dens;
          h->POT[o + p] = RHOT[p] / dens;
          h->WT[o + p] = F[p] / dens;
        }
        for (int p = 0; p < Np; ++p) {
          const double CPovCV = h->CPtot[o + p] / h->CVtot[o + p];
          h->DPDRHOT[o + p] = CPovCV * m->PRES00 * opow(h->Rtot[o + p] / m->PRES00 * RHOT[p], CPovCV) / RHOT[p];
        }
        for (int p = 0; p < Np; ++p) {
          h->b[3 * (o + p) + 0] = h->impl_fac * DENS_t[o + p] - PV[DENS_VID][o + p] + h->qs[DENS_VID][o + p];
          h->b[3 * (o + p) + 1] = h->impl_fac * MOMZ_t[o + p] - PV[MOMZ_VID][o + p] + h->qs[MOMZ_VID][o + p];
          h->b[3 * (o + p) + 2] = h->impl_fac * RHOT_t[o + p] - PV[RHOT_VID][o + p] + h->qs[RHOT_VID][o + p];
        }
      }
    }
    free(F); free(DF); free(RHOT); free(drho);
  }
  free(df);
}

/* ---------------------------------------------------------------------- */
/* construct_matbnd (..._common_2.F90:693-901) for column ij of horizontal
 * element ke2D at level kz: D (3N x 3N), L and U (3N x 3) */
static void hevi_matbnd(const hevi_ctx *h, int ke2D, int ij, int kz, double *D, double *Lm, double *Um) {
  const oracle_mesh *m = h->m;
  const int N = h->N, Nfp = h->Nfp, n3 = h->n3, NfpTot = m->NfpTot;
  const long Np = m->Np, NpNe = Np * m->Ne, nfall = (long)NfpTot * m->Ne;
  const int ke = ke2D + kz * h->Ne2D;
  const double impl_fac = h->impl_fac;
#define MD(v, pv, v2, pv2) D[((v) + 3 * (pv)) + (long)n3 * ((v2) + 3 * (pv2))]
#define ML(M, v, pv, v2) M[((v) + 3 * (pv)) + (long)n3 * (v2)]
  const double *E33 = m->Escale + NpNe * 8;
  const long o = Np * ke;
  double fac_dz[64 * 64];
  for (int pv2 = 0; pv2 < N; ++pv2)
    for (int pv = 0; pv < N; ++pv) {
      const long p = o + ij + (long)Nfp * pv;
      const double fDx3 = impl_fac * m->D1D[pv + N * pv2];
      fac_dz[pv + N * pv2] = E33[p] / h->GsqrtV[p] * fDx3;
    }
  for (int pv2 = 0; pv2 < N; ++pv2)
    for (int pv = 0; pv < N; ++pv) {
      const long q2 = o + ij + (long)Nfp * pv2;
      const double Id = (pv == pv2) ? 1.0 : 0.0;
      const double fdz = fac_dz[pv + N * pv2];
      /* IntrpMat_VPOrdM1(pv,pv2) = IntrpMat_VPOrdM1_tr(pv2,pv) */
      const double fI = impl_fac * m->GRAV * m->IntrpMat_tr[pv2 + N * pv];
      MD(0, pv, 0, pv2) = Id;
      MD(0, pv, 1, pv2) = fdz;
      MD(0, pv, 2, pv2) = 0.0;
      MD(1, pv, 1, pv2) = Id;
      MD(1, pv, 0, pv2) = fI;
      MD(1, pv, 2, pv2) = fdz * h->DPDRHOT[q2];
      MD(2, pv, 0, pv2) = -fdz * h->POT[q2] * h->WT[q2];
      MD(2, pv, 1, pv2) = fdz * h->POT[q2];
      MD(2, pv, 2, pv2) = Id + fdz * h->WT[q2];
    }
  for (int f1 = 1; f1 <= 2; ++f1) {
    int kz2, pv1, pv2, f2;
    if (f1 == 1) {
      kz2 = kz - 1 > 0 ? kz - 1 : 0;
      pv1 = 0; pv2 = N - 1; f2 = 2;
    } else {
      kz2 = kz + 1 < h->NeZ - 1 ? kz + 1 : h->NeZ - 1;
      pv1 = N - 1; pv2 = 0; f2 = 1;
    }
    const int ke_2 = ke2D + kz2 * h->Ne2D;
    const int bc = (kz == 0 && f1 == 1) || (kz == h->NeZ - 1 && f1 == 2);
    if (bc) { pv2 = pv1; f2 = f1; }
    double tmp1[64], tmp2[64];
    for (int pv = 0; pv < N; ++pv) {
      const int fp = 4 * Nfp + (f1 - 1) * Nfp + ij;
      const int fp2 = 4 * Nfp + (f2 - 1) * Nfp + ij;
      const long p = ij + (long)Nfp * pv;
      const long gfp = fp + (long)NfpTot * ke, gfp2 = fp2 + (long)NfpTot * ke_2;
      const double Lift = m->Lift_mat[p + Np * (3 + f1)];
      const double fac = 0.5 * impl_fac / h->GsqrtV[o + p] * Lift * m->Fscale[gfp];
      tmp1[pv] = fac * fmax_f(h->alph[gfp], h->alph[gfp2]);
      tmp2[pv] = fac * m->normal_fn[gfp + 2 * nfall];
    }
    const long q1 = o + ij + (long)Nfp * pv1;
    if (bc) {
      for (int pv = 0; pv < N; ++pv) {
        MD(2, pv, 0, pv1) = MD(2, pv, 0, pv1) + 2.0 * tmp2[pv] * h->POT[q1] * h->WT[q1];
        MD(0, pv, 1, pv1) = MD(0, pv, 1, pv1) - 2.0 * tmp2[pv];
        MD(1, pv, 1, pv1) = MD(1, pv, 1, pv1) + 2.0 * tmp1[pv];
        MD(2, pv, 1, pv1) = MD(2, pv, 1, pv1) - 2.0 * tmp2[pv] * h->POT[q1];
        MD(2, pv, 2, pv1) = MD(2, pv, 2, pv1) - 2.0 * tmp2[pv] * h->WT[q1];
      }
    } else {
      for (int pv = 0; pv < N; ++pv) {
        MD(0, pv, 0, pv1) = MD(0, pv, 0, pv1) + tmp1[pv];
        MD(2, pv, 0, pv1) = MD(2, pv, 0, pv1) + tmp2[pv] * h->POT[q1] * h->WT[q1];
        MD(0, pv, 1, pv1) = MD(0, pv, 1, pv1) - tmp2[pv];
        MD(1, pv, 1, pv1) = MD(1, pv, 1, pv1) + tmp1[pv];
        MD(2, pv, 1, pv1) = MD(2, pv, 1, pv1) - tmp2[pv] * h->POT[q1];
        MD(1, pv, 2, pv1) = MD(1, pv, 2, pv1) - tmp2[pv] * h->DPDRHOT[q1];
        MD(2, pv, 2, pv1) = MD(2, pv, 2, pv1) + tmp1[pv] - tmp2[pv] * h->WT[q1];
      }
      double *M = (f1 == 1) ? Lm : Um;
      const long q2 = Np * ke_2 + ij + (long)Nfp * pv2;
      for (int pv = 0; pv < N; ++pv) {
        ML(M, 0, pv, 0) = -tmp1[pv];
        ML(M, 1, pv, 0) = 0.0;
        ML(M, 2, pv, 0) = -tmp2[pv] * h->POT[q2] * h->WT[q2];
        ML(M, 0, pv, 1) = +tmp2[pv];
        ML(M, 1, pv, 1) = -tmp1[pv];
        ML(M, 2, pv, 1) = +tmp2[pv] * h->POT[q2];
        ML(M, 0, pv, 2) = 0.0;
        ML(M, 1, pv, 2) = +tmp2[pv] * h->DPDRHOT[q2];
        ML(M, 2, pv, 2) = -tmp1[pv] + tmp2[pv] * h->WT[q2];
      }
    }
  }
#undef MD
#undef ML
}

/* construct_matbnd_uv (..._common_2.F90:905-1011) */
static void hevi_matbnd_uv(const hevi_ctx *h, int ke2D, int ij, int kz, double *D, double *Lv, double *Uv) {
  const oracle_mesh *m = h->m;
  const int N = h->N, Nfp = h->Nfp, NfpTot = m->NfpTot;
  const long Np = m->Np;
  const int ke = ke2D + kz * h->Ne2D;
  const long o = Np * ke;
  for (int pv2 = 0; pv2 < N; ++pv2)
    for (int pv = 0; pv < N; ++pv) D[pv + N * pv2] = (pv == pv2) ? 1.0 : 0.0;
  for (int f1 = 1; f1 <= 2; ++f1) {
    int kz2, pv1, f2;
    if (f1 == 1) {
      kz2 = kz - 1 > 0 ? kz - 1 : 0;
      pv1 = 0; f2 = 2;
    } else {
      kz2 = kz + 1 < h->NeZ - 1 ? kz + 1 : h->NeZ - 1;
      pv1 = N - 1; f2 = 1;
    }
    const int ke_2 = ke2D + kz2 * h->Ne2D;
    const int bc = (kz == 0 && f1 == 1) || (kz == h->NeZ - 1 && f1 == 2);
    if (bc) continue;
    for (int pv = 0; pv < N; ++pv) {
      const int fp = 4 * Nfp + (f1 - 1) * Nfp + ij;
      const int fp2 = 4 * Nfp + (f2 - 1) * Nfp + ij;
      const long p = ij + (long)Nfp * pv;
      const long gfp = fp + (long)NfpTot * ke, gfp2 = fp2 + (long)NfpTot * ke_2;
      const double fac = 0.5 * h->impl_fac / h->GsqrtV[o + p];
      const double t = fac * m->Lift_mat[p + Np * (3 + f1)] * m->Fscale[gfp] * fmax_f(h->alph[gfp], h->alph[gfp2]);
      D[pv + N * pv1] = D[pv + N * pv1] + t;
      if (f1 == 1) Lv[pv] = -t;
      else Uv[pv] = -t;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* solve_Nnode8_var3 / solve_Nnode8_uv (..._linalgebra.F90:2142-2445), one
 * column: LU with partial pivoting of the n x n block A (column-major,
 * overwritten), then nr right-hand sides X (n x nr, column-major).  When
 * is_top is false, the rhs are b followed by the columns of G. */
static void hevi_lu_solve(int n, double *A, int nr, double *X) {
  int ipiv[64];
  for (int k = 0; k < n; ++k) {
    double tmpv = fabs(A[k + n * k]);
    ipiv[k] = k;
    for (int i = k + 1; i < n; ++i) {
      const double tmp = fabs(A[i + n * k]);
      if (tmp > tmpv) {
        ipiv[k] = i;
        tmpv = tmp;
      }
    }
    const int ip = ipiv[k];
    if (ip != k)
      for (int j = 0; j < n; ++j) {
        const double t = A[k + n * j];
        A[k + n * j] = A[ip + n * j];
        A[ip + n * j] = t;
      }
    tmpv = 1.0 / A[k + n * k];
    A[k + n * k] = tmpv;
    for (int i = k + 1; i < n; ++i) A[i + n * k] = A[i + n * k] * tmpv;
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i) A[i + n * j] = A[i + n * j] - A[i + n * k] * A[k + n * j];
  }
  for (int i = 0; i < n - 1; ++i) {
    const int ip = ipiv[i];
    if (ip != i)
      for (int r = 0; r < nr; ++r) {
        const double t = X[i + n * r];
        X[i + n * r] = X[ip + n * r];
        X[ip + n * r] = t;
      }
  }
  for (int i = 0; i < n; ++i)
    for (int r = 0; r < nr; ++r) {
      double t = X[i + n * r];
      for (int j = 0; j < i; ++j) t = t - X[j + n * r] * A[i + n * j];
      X[i + n * r] = t;
    }
  for (int i = n - 1; i >= 0; --i)
    for (int r = 0; r < nr; ++r) {
      double t = X[i + n * r];
      for (int j = i + 1; j < n; ++j) t = t - A[i + n * j] * X[j + n * r];
      X[i + n * r] = t * A[i + n * i];
    }
}

/* solve_uv (..._common_2.F90:589-689): block Thomas over the column, then
 * MOMX, MOMY += increment */
static void hevi_solve_uv(hevi_ctx *h) {
  const int N = h->N, Nfp = h->Nfp, NeZ = h->NeZ, Ne2D = h->Ne2D;
  const long Np = h->m->Np;
#pragma omp parallel
  {
    double *D = (double *)malloc(sizeof(double) * N * N);
    double *Lv = (double *)malloc(sizeof(double) * N);
    double *X = (double *)malloc(sizeof(double) * N * 3 * NeZ); /* per level: b(:,1), b(:,2), G */
#pragma omp for schedule(static) collapse(2)
    for (int ke2D = 0; ke2D < Ne2D; ++ke2D)
      for (int ij = 0; ij < Nfp; ++ij) {
        for (int kz = 0; kz < NeZ; ++kz) {
          double *Xk = X + (long)N * 3 * kz;
          const long o = Np * (ke2D + (long)kz * Ne2D);
          for (int pv = 0; pv < N; ++pv) {
            Xk[pv] = h->buv[2 * (o + ij + (long)Nfp * pv) + 0];
            Xk[pv + N] = h->buv[2 * (o + ij + (long)Nfp * pv) + 1];
          }
          hevi_matbnd_uv(h, ke2D, ij, kz, D, Lv, Xk + 2 * N); /* U lands in the G slot */
          if (kz > 0) {
            const double *Xp = X + (long)N * 3 * (kz - 1);
            for (int pv = 0; pv < N; ++pv) {
              const double tmp = Lv[pv];
              D[pv + N * 0] = D[pv + N * 0] - tmp * Xp[(N - 1) + 2 * N];
              Xk[pv] = Xk[pv] - tmp * Xp[N - 1];
              Xk[pv + N] = Xk[pv + N] - tmp * Xp[(N - 1) + N];
            }
          }
          hevi_lu_solve(N, D, kz == NeZ - 1 ? 2 : 3, Xk);
        }
        for (int kz = NeZ - 2; kz >= 0; --kz) {
          double *Xk = X + (long)N * 3 * kz;
          const double *Xn = X + (long)N * 3 * (kz + 1);
          for (int pv = 0; pv < N; ++pv) {
            const double tmp = Xk[pv + 2 * N];
            Xk[pv] = Xk[pv] - tmp * Xn[0];
            Xk[pv + N] = Xk[pv + N] - tmp * Xn[N];
          }
        }
        for (int kz = 0; kz < NeZ; ++kz) {
          const double *Xk = X + (long)N * 3 * kz;
          const long o = Np * (ke2D + (long)kz * Ne2D);
          for (int pv = 0; pv < N; ++pv) {
            const long q = o + ij + (long)Nfp * pv;
            h->pv[MOMX_VID][q] = h->pv[MOMX_VID][q] + Xk[pv];
            h->pv[MOMY_VID][q] = h->pv[MOMY_VID][q] + Xk[pv + N];
          }
        }
      }
    free(D); free(Lv); free(X);
  }
}

/* solve (..._common_2.F90:329-479): block Thomas over the column with the
 * 3N x 3N blocks, then DDENS, MOMZ, DRHOT += increment */
static void hevi_solve_var3(hevi_ctx *h) {
  const int N = h->N, Nfp = h->Nfp, NeZ = h->NeZ, Ne2D = h->Ne2D, n3 = h->n3;
  const long Np = h->m->Np;
#pragma omp parallel
  {
    double *D = (double *)malloc(sizeof(double) * n3 * n3);
    double *Lm = (double *)malloc(sizeof(double) * n3 * 3);
    double *X = (double *)malloc(sizeof(double) * n3 * 4 * NeZ); /* per level: b, G(:,1..3) */
#pragma omp for schedule(static) collapse(2)
    for (int ke2D = 0; ke2D < Ne2D; ++ke2D)
      for (int ij = 0; ij < Nfp; ++ij) {
        for (int kz = 0; kz < NeZ; ++kz) {
          double *Xk = X + (long)n3 * 4 * kz;
          const long o = Np * (ke2D + (long)kz * Ne2D);
          for (int pv = 0; pv < N; ++pv)
            for (int v = 0; v < 3; ++v) Xk[v + 3 * pv] = h->b[3 * (o + ij + (long)Nfp * pv) + v];
          hevi_matbnd(h, ke2D, ij, kz, D, Lm, Xk + n3); /* U lands in the G slots */
          if (kz > 0) {
            const double *Xp = X + (long)n3 * 4 * (kz - 1);
            const int p1 = 3 * (N - 1), p2 = p1 + 1, p3 = p1 + 2;
            for (int r = 0; r < n3; ++r) {
              const double t1 = Lm[r], t2 = Lm[r + n3], t3 = Lm[r + 2 * n3];
              for (int c = 0; c < 3; ++c)
                D[r + n3 * c] = D[r + n3 * c] - t1 * Xp[p1 + n3 * (1 + c)] - t2 * Xp[p2 + n3 * (1 + c)] -
                                t3 * Xp[p3 + n3 * (1 + c)];
              Xk[r] = Xk[r] - t1 * Xp[p1] - t2 * Xp[p2] - t3 * Xp[p3];
            }
          }
          hevi_lu_solve(n3, D, kz == NeZ - 1 ? 1 : 4, Xk);
        }
        for (int kz = NeZ - 2; kz >= 0; --kz) {
          double *Xk = X + (long)n3 * 4 * kz;
          const double *Xn = X + (long)n3 * 4 * (kz + 1);
          for (int r = 0; r < n3; ++r)
            Xk[r] = Xk[r] - Xk[r + n3] * Xn[0] - Xk[r + 2 * n3] * Xn[1] - Xk[r + 3 * n3] * Xn[2];
        }
        for (int kz = 0; kz < NeZ; ++kz) {
          const double *Xk = X + (long)n3 * 4 * kz;
          const long o = Np * (ke2D + (long)kz * Ne2D);
          for (int pv = 0; pv < N; ++pv) {
            const long q = o + ij + (long)Nfp * pv;
            h->pv[DENS_VID][q] = h->pv[DENS_VID][q] + Xk[0 + 3 * pv];
            h->pv[MOMZ_VID][q] = h->pv[MOMZ_VID][q] + Xk[1 + 3 * pv];
            h->pv[RHOT_VID][q] = h->pv[RHOT_VID][q] + Xk[2 + 3 * pv];
          }
        }
      }
    free(D); free(Lm); free(X);
  }
}

static double sumsq(const double *a, long n) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += a[i] * a[i];
  return s;
}

static void hevi_ctx_init(hevi_ctx *h, const oracle_mesh *m, double impl_fac, double dt) {
  memset(h, 0, sizeof(*h));
  h->m = m;
  h->N = m->n1;
  h->Nfp = m->Nfp_v;
  h->Ne2D = m->Ne2D;
  h->NeZ = m->Ne / m->Ne2D;
  h->n3 = 3 * m->n1;
  h->impl_fac = impl_fac;
  h->dt = dt;
  const long n = (long)m->Np * m->Ne;
  h->GsqrtV = (double *)malloc(sizeof(double) * n);
  h->alph = (double *)malloc(sizeof(double) * (size_t)m->NfpTot * m->Ne);
  h->DENSc = (double *)malloc(sizeof(double) * n);
  h->WT = (double *)malloc(sizeof(double) * n);
  h->POT = (double *)malloc(sizeof(double) * n);
  h->DPDRHOT = (double *)malloc(sizeof(double) * n);
  h->b = (double *)malloc(sizeof(double) * n * 3);
  h->buv = (double *)malloc(sizeof(double) * n * 2);
  for (int ke = 0; ke < m->Ne; ++ke) {
    const int ke2D = m->EMap3Dto2D[ke] - 1;
    for (int p = 0; p < m->Np; ++p) {
      const long q = (long)m->Np * ke + p;
      h->GsqrtV[q] = m->Gsqrt[q] / m->GsqrtH[(m->IndexH2Dto3D[p] - 1) + (long)m->Nfp_v * ke2D];
    }
  }
}

static void hevi_ctx_free(hevi_ctx *h) {
  free(h->GsqrtV); free(h->alph); free(h->DENSc); free(h->WT); free(h->POT); free(h->DPDRHOT);
  free(h->b); free(h->buv);
}

/* Newton iterations of cal_vi (..._rhot_hevi.F90:858-917) on the iterate
 * pv (VID order, (Np,Ne), initialised by the caller).  The reference runs
 * exactly one (max_it = 1, rtol <= 0).  With rtol > 0 the loop stops once
 * ||G||_2 / ||G_0||_2 <= rtol, ||G||_2 <= 4 eps ||q*||_2 or an update has
 * not halved ||G|| (the rounding floor), G = (b_uv, b) being the residual that eval_Ax_uv
 * and eval_Ax form at the start of an iteration; resid[k] receives ||G_k||.
 * Returns the number of iterations done. */
int oracle_hevi_newton(const oracle_mesh *m, double *const pv[5], const double *const qs[5],
                       const double *const q0[5], const double *DENS_hyd, const double *PRES_hyd,
                       const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac, double dt,
                       int max_it, double rtol, double *resid) {
  return oracle_hevi_newton_r0(m, pv, qs, q0, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, impl_fac, dt, max_it, rtol,
                               0.0, resid);
}

/* As oracle_hevi_newton; r0_ext > 0 is the norm the tolerance is relative to
 * (the residual at q^n when the iterate starts from a prediction, which may
 * already meet the tolerance at the first check). */
int oracle_hevi_newton_r0(const oracle_mesh *m, double *const pv[5], const double *const qs[5],
                          const double *const q0[5], const double *DENS_hyd, const double *PRES_hyd,
                          const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac, double dt,
                          int max_it, double rtol, double r0_ext, double *resid) {
  hevi_ctx h;
  hevi_ctx_init(&h, m, impl_fac, dt);
  h.qs = qs; h.q0 = q0; h.pv = pv;
  h.DENS_hyd = DENS_hyd; h.PRES_hyd = PRES_hyd; h.Rtot = Rtot; h.CVtot = CVtot; h.CPtot = CPtot;
  const long n = (long)m->Np * m->Ne;
  double *T = (double *)malloc(sizeof(double) * n * 5);
  double r0 = r0_ext;
  /* rounding floor: stop also at ||G|| <= 4 eps ||q*||, and once an update
   * has not halved the residual */
  double atol = 0.0;
  if (rtol > 0.0) {
    double q2 = 0.0;
    for (int v = 0; v < 5; ++v)
      for (long i = 0; i < n; ++i) q2 += qs[v][i] * qs[v][i];
    atol = 4.0 * DBL_EPSILON * sqrt(q2);
  }
  double r_prev = 0.0;
  int it = 0;
  for (; it < max_it; ++it) {
    hevi_eval_Ax_uv(&h, T + 3 * n, T + 4 * n, 1);
    /* the var3 residual is formed after the uv update in the reference; the
     * convergence test uses the residual of the iterate before the update */
    double r2 = sumsq(h.buv, 2 * n);
    hevi_solve_uv(&h);
    hevi_eval_Ax(&h, T, T + 2 * n, T + n, 1);
    r2 += sumsq(h.b, 3 * n);
    const double r = sqrt(r2);
    if (resid) resid[it] = r;
    if (it == 0 && r0_ext <= 0.0) r0 = r;
    if (rtol > 0.0 && (it > 0 || r0_ext > 0.0) && (r <= rtol * r0 || r <= atol)) break;
    if (rtol > 0.0 && it > 0 && r > 0.5 * r_prev) break;
    r_prev = r;
    hevi_solve_var3(&h);
  }
  free(T);
  hevi_ctx_free(&h);
  return it;
}

/* cal_vi (..._rhot_hevi.F90:772-965): implicit tendencies of the stage
 * state q (VID order, (Np,NeA)); var0 = q^n */
int oracle_hevi_cal_vi(const oracle_mesh *m, double *const tend[5], const double *const q[5],
                       const double *const var0[5], const double *DENS_hyd, const double *PRES_hyd,
                       const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac, double dt,
                       int max_it, double rtol, double *resid) {
  const long n = (long)m->Np * m->Ne;
  double *buf = (double *)malloc(sizeof(double) * n * 5);
  double *pv[5];
  for (int v = 0; v < 5; ++v) {
    pv[v] = buf + n * v;
    memcpy(pv[v], var0[v], sizeof(double) * n);
  }
  int it = 0;
  if (fabs(impl_fac) > 0.0) {
    it = oracle_hevi_newton(m, pv, q, var0, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, impl_fac, dt, max_it, rtol,
                            resid);
    for (int v = 0; v < 5; ++v)
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) tend[v][i] = (pv[v][i] - q[v][i]) / impl_fac;
  } else {
    hevi_ctx h;
    hevi_ctx_init(&h, m, impl_fac, dt);
    h.qs = q; h.q0 = var0; h.pv = pv;
    h.DENS_hyd = DENS_hyd; h.PRES_hyd = PRES_hyd; h.Rtot = Rtot; h.CVtot = CVtot; h.CPtot = CPtot;
    hevi_eval_Ax_uv(&h, tend[MOMX_VID], tend[MOMY_VID], 0);
    hevi_eval_Ax(&h, tend[DENS_VID], tend[MOMZ_VID], tend[RHOT_VID], 0);
    hevi_ctx_free(&h);
  }
  free(buf);
  return it;
}

/* The implicit linear system of one Newton iteration at iterate pv, for
 * tests: every column's blocks in the layout of the solver above.
 *   D: (3N*3N, NeZ, Nfp, Ne2D)    L, U: (3N*3, ...)    b: (3N, ...)
 *   Duv: (N*N, ...)    Luv, Uuv: (N, ...)    buv: (N*2, ...)
 * The uv system is the one at pv; the var3 system is formed after the uv
 * increment has been added to pv (as the reference does), and pv returns
 * with that uv update. */
void oracle_hevi_linear_system(const oracle_mesh *m, double *const pv[5], const double *const qs[5],
                               const double *const q0[5], const double *DENS_hyd, const double *PRES_hyd,
                               const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac,
                               double dt, double *D, double *L, double *U, double *b, double *Duv, double *Luv,
                               double *Uuv, double *buv) {
  hevi_ctx h;
  hevi_ctx_init(&h, m, impl_fac, dt);
  h.qs = qs; h.q0 = q0; h.pv = pv;
  h.DENS_hyd = DENS_hyd; h.PRES_hyd = PRES_hyd; h.Rtot = Rtot; h.CVtot = CVtot; h.CPtot = CPtot;
  const long n = (long)m->Np * m->Ne;
  const int N = h.N, n3 = h.n3, Nfp = h.Nfp, NeZ = h.NeZ;
  double *T = (double *)malloc(sizeof(double) * n * 5);
  hevi_eval_Ax_uv(&h, T + 3 * n, T + 4 * n, 1);
  for (int ke2D = 0; ke2D < h.Ne2D; ++ke2D)
    for (int ij = 0; ij < Nfp; ++ij)
      for (int kz = 0; kz < NeZ; ++kz) {
        const long c = kz + (long)NeZ * (ij + (long)Nfp * ke2D);
        double Lv[64], Uv[64];
        memset(Lv, 0, sizeof(Lv));
        memset(Uv, 0, sizeof(Uv));
        hevi_matbnd_uv(&h, ke2D, ij, kz, Duv + c * N * N, Lv, Uv);
        memcpy(Luv + c * N, Lv, sizeof(double) * N);
        memcpy(Uuv + c * N, Uv, sizeof(double) * N);
        const long o = (long)m->Np * (ke2D + (long)kz * h.Ne2D);
        for (int pv_ = 0; pv_ < N; ++pv_)
          for (int r = 0; r < 2; ++r) buv[c * 2 * N + pv_ + N * r] = h.buv[2 * (o + ij + (long)Nfp * pv_) + r];
      }
  hevi_solve_uv(&h);
  hevi_eval_Ax(&h, T, T + 2 * n, T + n, 1);
  for (int ke2D = 0; ke2D < h.Ne2D; ++ke2D)
    for (int ij = 0; ij < Nfp; ++ij)
      for (int kz = 0; kz < NeZ; ++kz) {
        const long c = kz + (long)NeZ * (ij + (long)Nfp * ke2D);
        double Lm[64 * 3], Um[64 * 3];
        memset(Lm, 0, sizeof(Lm));
        memset(Um, 0, sizeof(Um));
        hevi_matbnd(&h, ke2D, ij, kz, D + c * n3 * n3, Lm, Um);
        memcpy(L + c * n3 * 3, Lm, sizeof(double) * n3 * 3);
        memcpy(U + c * n3 * 3, Um, sizeof(double) * n3 * 3);
        const long o = (long)m->Np * (ke2D + (long)kz * h.Ne2D);
        for (int pv_ = 0; pv_ < N; ++pv_)
          for (int v = 0; v < 3; ++v) b[c * n3 + v + 3 * pv_] = h.b[3 * (o + ij + (long)Nfp * pv_) + v];
      }
  free(T);
  hevi_ctx_free(&h);
}

/* Residual G(pv) = (b_uv, b) of eval_Ax_uv / eval_Ax without any update,
 * node layout: G[v][node] for VID v (DDENS, DRHOT, MOMZ, MOMX, MOMY), for
 * finite-difference checks of the linear system. */
void oracle_hevi_residual(const oracle_mesh *m, double *const pv[5], const double *const qs[5],
                          const double *const q0[5], const double *DENS_hyd, const double *PRES_hyd,
                          const double *Rtot, const double *CVtot, const double *CPtot, double impl_fac, double dt,
                          double *const G[5]) {
  hevi_ctx h;
  hevi_ctx_init(&h, m, impl_fac, dt);
  h.qs = qs; h.q0 = q0; h.pv = pv;
  h.DENS_hyd = DENS_hyd; h.PRES_hyd = PRES_hyd; h.Rtot = Rtot; h.CVtot = CVtot; h.CPtot = CPtot;
  const long n = (long)m->Np * m->Ne;
  double *T = (double *)malloc(sizeof(double) * n * 5);
  hevi_eval_Ax_uv(&h, T + 3 * n, T + 4 * n, 1);
  hevi_eval_Ax(&h, T, T + 2 * n, T + n, 1);
  for (long i = 0; i < n; ++i) {
    G[DENS_VID][i] = h.b[3 * i + 0];
    G[MOMZ_VID][i] = h.b[3 * i + 1];
    G[RHOT_VID][i] = h.b[3 * i + 2];
    G[MOMX_VID][i] = h.buv[2 * i + 0];
    G[MOMY_VID][i] = h.buv[2 * i + 1];
  }
  free(T);
  hevi_ctx_free(&h);
}

/* ---------------------------------------------------------------------- */
/* FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:694-870, HEVI branch (IMEX) */
void oracle_hevi_update(const oracle_mesh *m, const oracle_rk_tab *t, int nsteps, double *const q[5],
                        double *PRES, double *DPRES, const double *DENS_hyd, const double *PRES_hyd,
                        const double *PRES_hyd_ref, const double *THERM_hyd, const double *CORIOLIS,
                        const double *Rtot, const double *CVtot, const double *CPtot, const double *DPhydDx,
                        const double *DPhydDy, int max_it, double rtol, double *work) {
  const long NpNeA = (long)m->Np * m->NeA;
  const long n = (long)m->Np * m->Ne;
  double *var0 = work;
  double *varTmp = work + 5 * NpNeA;
  double *tend_ex = work + 10 * NpNeA;                               /* (Np,NeA,5,nbuf) */
  double *tend_im = tend_ex + (long)5 * NpNeA * t->tend_buf_size; /* (Np,NeA,5,nbuf) */
  for (int step = 0; step < nsteps; ++step) {
    for (int v = 0; v < 5; ++v) memcpy(var0 + v * NpNeA, q[v], sizeof(double) * n);
    for (int s = 1; s <= t->nstage; ++s) {
      const int ind = t->tend_buf_indmap[s - 1];
      const double impl_fac = t->a_im[(s - 1) * 4 + (s - 1)] * t->dt;
      double *Ti = tend_im + (long)ind * 5 * NpNeA;
      double *Te = tend_ex + (long)ind * 5 * NpNeA;
      {
        double *tv[5];
        const double *qv[5], *v0[5];
        for (int v = 0; v < 5; ++v) {
          tv[v] = Ti + v * NpNeA;
          qv[v] = q[v];
          v0[v] = var0 + v * NpNeA;
        }
        oracle_hevi_cal_vi(m, tv, qv, v0, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, impl_fac, t->dt, max_it, rtol,
                           NULL);
      }
      for (int v = 0; v < 5; ++v)
        oracle_rk_store_impl(t, s, n, q[v], var0 + v * NpNeA, varTmp + v * NpNeA, tend_im + v * NpNeA, 5 * NpNeA);
      for (int v = 0; v < 5; ++v) oracle_halo_exchange(m, q[v]);
      oracle_drhot2pres(m, PRES, DPRES, q[RHOT_VID], PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot);
      oracle_halo_exchange(m, DPRES);
      oracle_apply_bc(m, q[DENS_VID], q[MOMX_VID], q[MOMY_VID], q[MOMZ_VID], q[RHOT_VID]);
      oracle_hevi_cal_tend(m, Te + DENS_VID * NpNeA, Te + MOMX_VID * NpNeA, Te + MOMY_VID * NpNeA,
                           Te + MOMZ_VID * NpNeA, Te + RHOT_VID * NpNeA, q[DENS_VID], q[MOMX_VID], q[MOMY_VID],
                           q[MOMZ_VID], q[RHOT_VID], DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd, CORIOLIS,
                           Rtot, CVtot, CPtot, DPhydDx, DPhydDy);
      {
        double *Tv[5];
        for (int v = 0; v < 5; ++v) Tv[v] = Te + v * NpNeA;
        oracle_phys_tend_hook(m, Tv, PRES, Rtot, CPtot);
      }
      /* FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:899-915, tracer coupling */
      oracle_save_massflux_hook(m, q, DPRES, DENS_hyd, PRES_hyd, s, t->b_ex[s - 1],
                                t->imex ? t->b_im[s - 1] : t->b_ex[s - 1], 1);
      for (int v = 0; v < 5; ++v)
        oracle_rk_advance(t, s, n, q[v], var0 + v * NpNeA, varTmp + v * NpNeA, tend_ex + v * NpNeA,
                          tend_im + v * NpNeA, 5 * NpNeA);
    }
    if (m->MF_h)
      oracle_modalfilter_apply(m, q[DENS_VID], q[MOMX_VID], q[MOMY_VID], q[MOMZ_VID], q[RHOT_VID], 1);
    oracle_drhot2pres(m, PRES, DPRES, q[RHOT_VID], PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot);
  }
}
