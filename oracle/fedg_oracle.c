/*
 * CPU oracle: literal C restatement of the FElib HEVE step.  TEST
 * INFRASTRUCTURE ONLY (see fedg_oracle.h): the product never calls it.
 *
 * Build: gcc -O3 -fopenmp -ffp-contract=off (oracle/Makefile).  Sums are
 * accumulated left to right in the order the reference writes them, each
 * product rounded before it is added, as gfortran evaluates the Fortran source
 * without -ffp-contract=fast.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *  - tensor-product operators (dirX/dirY/dirZ, Lift): the reference's own
 *    unit test semantics (FElib/test/FE/element_operation_hexahedral:
 *    derivative of 4x^p+3y^p+2z^p, Lift vs a dense M^-1 E), plus a numpy
 *    evaluation of the same contractions (tests/test_refelem.py);
 *  - Div with metrics (tests/test_oracle_general_metric.py): the reference's
 *    Div test (Escale = (1,2,0.2), division by a non-constant Gsqrt,
 *    test_element_operation_hexahedral.f90:115-144) carried through all five
 *    tendencies with Gsqrt = GsqrtH(x,y) GsqrtV(z), G13/G23, Coriolis f(x,y)
 *    and DPhydDx/Dy against sympy closed forms; a uniform wind preserved over
 *    Gal-Chen terrain; conservation of all five tendencies with random
 *    discontinuous metrics; the face flux against the contravariant Rusanov
 *    flux with random metrics;
 *  - calc_phyd_hgrad: closed forms of horizontally varying PRES_hyd and
 *    PRES_hyd_ref with flat and general metrics (scaled Escale), zero for a
 *    height-only pressure over terrain, and a vanishing periodic integral of
 *    GsqrtV x gradient with random metrics (the lift of the face jump);
 *  - VFilterPM1: identity on polynomials of degree <= p-1 in z, removes the
 *    degree-p Legendre mode (closed form);
 *  - numflux: consistency (equal states give only the pressure and advective
 *    jump terms, which vanish), per-face antisymmetry of the jump between the
 *    two elements sharing a face, and a brute-force numpy evaluation of the
 *    Rusanov formula at single face nodes;
 *  - cal_tend: hydrostatic rest state gives tendencies of exactly zero except
 *    the roundoff of the discrete hydrostatic balance, which the test measures
 *    against the -DPhydDx/-g*drho terms it cancels; free-stream preservation;
 *    mass conservation (sum of w*J*DENS_dt = 0 in a periodic box);
 *  - drhot2pres: closed form P = P00 (R rhotheta / P00)^(cp/cv);
 *  - RK tables / advance: the reference's test FElib/test/common/timeint_rk
 *    (order of convergence on the harmonic oscillator) and the Butcher
 *    coefficients in closed form;
 *  - tracer advection (cal_tend, calc_fct_coef, TMAR, save_massflux,
 *    tests/test_oracle_trcadv.py): advection of a global polynomial exact
 *    away from the boundaries; the FCT coefficient of the upwind element
 *    against LGL-weighted face sums whose sign differs from the plain sum;
 *    conservation with random metrics; the limiter drains exactly the
 *    available mass (and moves each element's mass one element downstream
 *    when every limiter is active); positivity; TMAR keeps element masses and
 *    cuts negatives; damping coefficients = sound speed + |normal velocity|
 *    and equal on both sides of every face; the low-storage tracer RK update
 *    integrates a constant mass tendency exactly under a moving density, and
 *    the tracer-only driver loop keeps a uniform tracer, conserves mass,
 *    stays non-negative with the limiter and converges with the order
 *    (tests/test_oracle_trcadv_step.py); the tracer filter keeps low modes
 *    and element masses; the dynamics loop's mass-flux save gives MFLX = MOM
 *    for a steady wind;
 *  - diagnostic variables (oracle_atmos_calc_diagvar,
 *    tests/test_oracle_diagvar.py): the physics each formula encodes on
 *    states of known temperature (ideal gas T, PT = T at PRES00, vanishing
 *    deviations of the base state, kinetic / potential / internal energy in
 *    closed form, ENGT with the computational z over terrain);
 *  - halo exchange / BC: the reference's mesh/field test semantics
 *    (FElib/test/FE/field_cubedom3d_hexahedral) on integer-valued fields, and
 *    the SLIP condition's closed form (exterior normal momentum = -interior).
 * The physical constants are SCALE's CONST_* values (external, not vendored
 * in the reference): parity unpinned for them; both the oracle and the
 * product take them as inputs.
 */
#include "fedg_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* The portable double-double pow that the GPU kernels use (host build of the
 * same source).  It is selected for the pressure's pow in parity mode; it is
 * pinned against 60-digit references in tests/test_pow.py. */
#include "../fe-project_amd/csrc/fedg_pow.h"

#define DENS_VID 0
#define RHOT_VID 1
#define MOMZ_VID 2
#define MOMX_VID 3
#define MOMY_VID 4

int oracle_num_threads(void) { return omp_get_max_threads(); }

/* 0: the C library's pow, as the reference Fortran's ** compiles with
 * gfortran/glibc; 1: the portable correctly rounded pow (parity mode). */
static int g_pow_mode = 0;
void oracle_set_pow_mode(int mode) { g_pow_mode = mode; }
int oracle_get_pow_mode(void) { return g_pow_mode; }
double oracle_opow(double x, double y) { return g_pow_mode ? fedg_pow(x, y) : pow(x, y); }
#define opow oracle_opow
void oracle_pow_array(const double *x, const double *y, double *out, long n, int mode) {
  for (long i = 0; i < n; ++i) out[i] = mode ? fedg_pow(x[i], y[i]) : pow(x[i], y[i]);
}

/* ---------------------------------------------------------------------- */
/* FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:428-479 (non-OpenACC branch) */
void oracle_drhot2pres(const oracle_mesh *m, double *PRES, double *DPRES,
                       const double *DRHOT, const double *PRES_hyd, const double *THERM_hyd,
                       const double *Rtot, const double *CVtot, const double *CPtot) {
  const double rP0 = 1.0 / m->PRES00;
  const long n = (long)m->Np * m->Ne;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double RHOT = THERM_hyd[i] + DRHOT[i];
    PRES[i] = m->PRES00 * opow(Rtot[i] * rP0 * RHOT, CPtot[i] / CVtot[i]);
    DPRES[i] = PRES[i] - PRES_hyd[i];
  }
}

/* DAT/scale_meshfieldcomm_base.F90:571-609 (pack by VMapB), :886-894 (copy
 * on the same rank), :768-789 (unpack into the halo tail), resolved through
 * the single-tile source map. */
void oracle_halo_exchange(const oracle_mesh *m, double *val) {
  const long base = (long)m->Np * m->Ne;
  double *buf = (double *)malloc(sizeof(double) * (size_t)m->nB);
  for (int b = 0; b < m->nB; ++b) buf[b] = val[m->halo_src[b] - 1];
  for (int b = 0; b < m->nB; ++b) val[base + b] = buf[b];
  free(buf);
}

/* FDS/scale_atm_dyn_dgm_bnd.F90:270-367 */
void oracle_apply_bc(const oracle_mesh *m, double *DDENS, double *MOMX, double *MOMY,
                     double *MOMZ, double *THERM) {
  (void)DDENS;
  (void)THERM;
  const int Np = m->Np, NfpTot = m->NfpTot, Ne = m->Ne, Nfp_v = m->Nfp_v;
  const long NpNeA = (long)Np * m->NeA;
  const long nfall = (long)NfpTot * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double *nx = m->normal_fn, *ny = m->normal_fn + nfall, *nz = m->normal_fn + 2 * nfall;
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < Ne; ++ke) {
    for (int p = 0; p < NfpTot; ++p) {
      long i = p + (long)ke * NfpTot;
      long iP = m->vmapP[i] - 1;
      long i_ = (long)m->vmapP[i] - (long)Np * Ne; /* 1-based halo index */
      if (i_ > 0) {
        long iM = m->vmapM[i] - 1;
        int bc = m->velbc[i_ - 1];
        if (bc == 1) { /* SLIP */
          int ke2D = m->EMap3Dto2D[ke] - 1;
          long h = (m->IndexH2Dto3D_bnd[p] - 1) + (long)Nfp_v * ke2D;
          double GsqrtV = m->Gsqrt[iM] / m->GsqrtH[h];
          double G11_ = m->G11[h], G12_ = m->G12[h], G22_ = m->G22[h];
          double MOMW = MOMZ[iM] / GsqrtV + G13[iM] * MOMX[iM] + G23[iM] * MOMY[iM];
          double t1 = GsqrtV * G13[iM];
          double t2 = GsqrtV * G23[iM];
          double fac = nz[i] * (GsqrtV * GsqrtV) /
                       (1.0 + G11_ * (t1 * t1) + 2.0 * G12_ * ((GsqrtV * GsqrtV) * G13[iM] * G23[iM]) +
                        G22_ * (t2 * t2));
          double mom_normal = MOMX[iM] * nx[i] + MOMY[iM] * ny[i] + MOMW * nz[i];
          MOMX[iP] = MOMX[iM] - 2.0 * mom_normal * (nx[i] + fac * (G11_ * G13[iM] + G12_ * G23[iM]));
          MOMY[iP] = MOMY[iM] - 2.0 * mom_normal * (ny[i] + fac * (G12_ * G13[iM] + G22_ * G23[iM]));
          MOMZ[iP] = MOMZ[iM] - 2.0 * mom_normal * fac / GsqrtV;
        } else if (bc == 2) { /* NOSLIP */
          MOMX[iP] = -MOMX[iM];
          MOMY[iP] = -MOMY[iM];
          MOMZ[iP] = -MOMZ[iM];
        }
      }
    }
  }
}

#include "oracle_ops.h"

/* ---------------------------------------------------------------------- */
/* FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:946-1138.  cs2 != 0:
 * the flux of add_bnd_contrib_generalvc (:242-496), identical except that the
 * sound speed is grouped as sqrt(Gnn * Cs2), Cs2 = gamm (P_hyd + DPRES) Gsqrt
 * / GsqrtDens (:351-352, :408-409, :425-431). */
static void heve_numflux_range(const oracle_mesh *m, int ke_s, int ke_e, double *del_flux,
                               const double *DDENS_, const double *MOMX_, const double *MOMY_,
                               const double *MOMZ_, const double *DRHOT_, const double *DPRES,
                               const double *DENS_hyd, const double *PRES_hyd, const double *THERM_hyd, int cs2) {
  const int NfpTot = m->NfpTot, Ne = m->Ne;
  const long NpNeA = (long)m->Np * m->NeA;
  const long nfall = (long)NfpTot * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double gamm = m->CPdry / m->CVdry;
#pragma omp parallel for schedule(static)
  for (int ke = ke_s; ke < ke_e; ++ke) {
    for (int fp = 0; fp < NfpTot; ++fp) {
      long i = fp + (long)NfpTot * ke;
      long iM = m->vmapM[i] - 1, iP = m->vmapP[i] - 1;
      double nx = m->normal_fn[i], ny = m->normal_fn[i + nfall], nz = m->normal_fn[i + 2 * nfall];

      double Gsqrt_M = m->Gsqrt[iM], Gsqrt_P = m->Gsqrt[iP];
      double RGsqrtV_M = 1.0 / Gsqrt_M, RGsqrtV_P = 1.0 / Gsqrt_P;
      double G13_M = G13[iM], G13_P = G13[iP], G23_M = G23[iM], G23_P = G23[iP];

      double GsqrtDDENS_M = Gsqrt_M * DDENS_[iM], GsqrtDDENS_P = Gsqrt_P * DDENS_[iP];
      double GsqrtMOMX_M = Gsqrt_M * MOMX_[iM], GsqrtMOMX_P = Gsqrt_P * MOMX_[iP];
      double GsqrtMOMY_M = Gsqrt_M * MOMY_[iM], GsqrtMOMY_P = Gsqrt_P * MOMY_[iP];
      double GsqrtMOMZ_M = Gsqrt_M * MOMZ_[iM], GsqrtMOMZ_P = Gsqrt_P * MOMZ_[iP];
      double GsqrtDRHOT_M = Gsqrt_M * DRHOT_[iM], GsqrtDRHOT_P = Gsqrt_P * DRHOT_[iP];

      double Phyd_M = PRES_hyd[iM], Phyd_P = PRES_hyd[iP];
      double dpres_M = DPRES[iM], dpres_P = DPRES[iP];

      double GsqrtDens_M = GsqrtDDENS_M + Gsqrt_M * DENS_hyd[iM];
      double GsqrtDens_P = GsqrtDDENS_P + Gsqrt_P * DENS_hyd[iP];
      double GsqrtRhot_M = Gsqrt_M * THERM_hyd[iM] + GsqrtDRHOT_M;
      double GsqrtRhot_P = Gsqrt_P * THERM_hyd[iP] + GsqrtDRHOT_P;

      double Vel_M = (GsqrtMOMX_M * nx + GsqrtMOMY_M * ny +
                      ((GsqrtMOMZ_M * RGsqrtV_M + G13_M * GsqrtMOMX_M + G23_M * GsqrtMOMY_M) * nz)) / GsqrtDens_M;
      double Vel_P = (GsqrtMOMX_P * nx + GsqrtMOMY_P * ny +
                      ((GsqrtMOMZ_P * RGsqrtV_P + G13_P * GsqrtMOMX_P + G23_P * GsqrtMOMY_P) * nz)) / GsqrtDens_P;

      double tmp1 = fabs(nx) + fabs(ny);
      double Gnn_M = tmp1 + (1.0 * (RGsqrtV_M * RGsqrtV_M) + G13_M * G13_M + G23_M * G23_M) * fabs(nz);
      double Gnn_P = tmp1 + (1.0 * (RGsqrtV_P * RGsqrtV_P) + G13_P * G13_P + G23_P * G23_P) * fabs(nz);
      double aM, aP;
      if (cs2) {
        double Cs2_M = gamm * (Phyd_M + dpres_M) * Gsqrt_M / GsqrtDens_M;
        double Cs2_P = gamm * (Phyd_P + dpres_P) * Gsqrt_P / GsqrtDens_P;
        aM = sqrt(Gnn_M * Cs2_M) + fabs(Vel_M);
        aP = sqrt(Gnn_P * Cs2_P) + fabs(Vel_P);
      } else {
        aM = sqrt(Gnn_M * gamm * (Phyd_M + dpres_M) * Gsqrt_M / GsqrtDens_M) + fabs(Vel_M);
        aP = sqrt(Gnn_P * gamm * (Phyd_P + dpres_P) * Gsqrt_P / GsqrtDens_P) + fabs(Vel_P);
      }
      double alpha = aM > aP ? aM : aP;

      double *df = del_flux + (long)NfpTot * 5 * ke;
      tmp1 = m->Fscale[i] * 0.5;
      double tmp2 = -alpha * (GsqrtDDENS_P - GsqrtDDENS_M);
      df[fp + NfpTot * DENS_VID] = tmp1 * (GsqrtDens_P * Vel_P - GsqrtDens_M * Vel_M + tmp2);
      tmp2 = -alpha * (GsqrtDRHOT_P - GsqrtDRHOT_M);
      df[fp + NfpTot * RHOT_VID] = tmp1 * (GsqrtRhot_P * Vel_P - GsqrtRhot_M * Vel_M + tmp2);

      double tmp3 = Gsqrt_P * dpres_P;
      double tmp4 = Gsqrt_M * dpres_M;
      double mom1 = (tmp3 * RGsqrtV_P - tmp4 * RGsqrtV_M) * nz;
      double mom2 = (nx + G13_P * nz) * tmp3 - (nx + G13_M * nz) * tmp4;
      double mom3 = (ny + G23_P * nz) * tmp3 - (ny + G23_M * nz) * tmp4;

      tmp2 = -alpha * (GsqrtMOMZ_P - GsqrtMOMZ_M);
      df[fp + NfpTot * MOMZ_VID] = tmp1 * (GsqrtMOMZ_P * Vel_P - GsqrtMOMZ_M * Vel_M + mom1 + tmp2);
      tmp2 = -alpha * (GsqrtMOMX_P - GsqrtMOMX_M);
      df[fp + NfpTot * MOMX_VID] = tmp1 * (GsqrtMOMX_P * Vel_P - GsqrtMOMX_M * Vel_M + mom2 + tmp2);
      tmp2 = -alpha * (GsqrtMOMY_P - GsqrtMOMY_M);
      df[fp + NfpTot * MOMY_VID] = tmp1 * (GsqrtMOMY_P * Vel_P - GsqrtMOMY_M * Vel_M + mom3 + tmp2);
    }
  }
}

void oracle_heve_numflux(const oracle_mesh *m, double *del_flux,
                         const double *DDENS_, const double *MOMX_, const double *MOMY_,
                         const double *MOMZ_, const double *DRHOT_, const double *DPRES,
                         const double *DENS_hyd, const double *PRES_hyd, const double *THERM_hyd) {
  heve_numflux_range(m, 0, m->Ne, del_flux, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES, DENS_hyd, PRES_hyd,
                     THERM_hyd, 0);
}

/* ---------------------------------------------------------------------- */
/* FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:372-484 (element loop),
 * with Div_var5 = ELM/scale_element_operation_tensorprod3D.F90:2760-2783 and
 * VFilterPM1 = :2828-2839. del_flux holds the numflux of elements [ke_s, ke_e).
 * del_flux NULL: the interior form cal_tend_cco (..._rhot_heve.F90:493-678),
 * Div_var5_2 (ELM/..._tensorprod3D.F90:2788-2805) without the lift term. */
static void heve_elements(const oracle_mesh *m, int ke_s, int ke_e, const double *del_flux,
                          double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt, double *RHOT_dt,
                          const double *DDENS_, const double *MOMX_, const double *MOMY_,
                          const double *MOMZ_, const double *DRHOT_, const double *DPRES_,
                          const double *DENS_hyd, const double *THERM_hyd, const double *CORIOLIS,
                          const double *DPhydDx, const double *DPhydDy) {
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Nfp_v = m->Nfp_v, Ne = m->Ne;
  const long NpNeA = (long)Np * m->NeA;
  const long NpNe = (long)Np * Ne;
  const double *GI3_1 = m->GI3, *GI3_2 = m->GI3 + NpNeA;
  const double Grav = m->GRAV;
#pragma omp parallel
  {
    double *Flux = (double *)malloc(sizeof(double) * Np * 15);  /* Flux(Np,3,5)  */
    double *DFlux = (double *)malloc(sizeof(double) * Np * 20); /* DFlux(Np,4,5) */
    double *drho = (double *)malloc(sizeof(double) * Np);
    double *RDENS = (double *)malloc(sizeof(double) * Np);
    double *RGsqrtV = (double *)malloc(sizeof(double) * Np);
    double *RGsqrt = (double *)malloc(sizeof(double) * Np);
#pragma omp for schedule(static)
    for (int ke = ke_s; ke < ke_e; ++ke) {
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
        double Gsqrt_ = m->Gsqrt[o + p];
        FL(p, 0, DENS_VID) = Gsqrt_ * MOMX_[o + p];
        FL(p, 1, DENS_VID) = Gsqrt_ * MOMY_[o + p];
        FL(p, 2, DENS_VID) = Gsqrt_ * (MOMZ_[o + p] * RGsqrtV[p] + GI3_1[o + p] * MOMX_[o + p] + GI3_2[o + p] * MOMY_[o + p]);
      }
      for (int p = 0; p < Np; ++p) {
        double pt_ = (THERM_hyd[o + p] + DRHOT_[o + p]) * RDENS[p];
        FL(p, 0, RHOT_VID) = FL(p, 0, DENS_VID) * pt_;
        FL(p, 1, RHOT_VID) = FL(p, 1, DENS_VID) * pt_;
        FL(p, 2, RHOT_VID) = FL(p, 2, DENS_VID) * pt_;
        double w_ = MOMZ_[o + p] * RDENS[p];
        FL(p, 0, MOMZ_VID) = FL(p, 0, DENS_VID) * w_;
        FL(p, 1, MOMZ_VID) = FL(p, 1, DENS_VID) * w_;
        FL(p, 2, MOMZ_VID) = FL(p, 2, DENS_VID) * w_ + m->Gsqrt[o + p] * DPRES_[o + p] * RGsqrtV[p];
      }
      for (int p = 0; p < Np; ++p) {
        double GsqrtDPRES_ = m->Gsqrt[o + p] * DPRES_[o + p];
        double u_ = MOMX_[o + p] * RDENS[p];
        FL(p, 0, MOMX_VID) = FL(p, 0, DENS_VID) * u_ + GsqrtDPRES_;
        FL(p, 1, MOMX_VID) = FL(p, 1, DENS_VID) * u_;
        FL(p, 2, MOMX_VID) = FL(p, 2, DENS_VID) * u_ + GsqrtDPRES_ * GI3_1[o + p];
        double v_ = MOMY_[o + p] * RDENS[p];
        FL(p, 0, MOMY_VID) = FL(p, 0, DENS_VID) * v_;
        FL(p, 1, MOMY_VID) = FL(p, 1, DENS_VID) * v_ + GsqrtDPRES_;
        FL(p, 2, MOMY_VID) = FL(p, 2, DENS_VID) * v_ + GsqrtDPRES_ * GI3_2[o + p];
      }
      /* Div_var5: Lift first, then the three line contractions, per variable */
      if (del_flux) {
        const double *dfe = del_flux + (long)NfpTot * 5 * ke;
        for (int v = 0; v < 5; ++v) op_lift(n1, m->Lift_mat, dfe + (long)NfpTot * v, &DF(0, 3, v));
      }
      for (int v = 0; v < 5; ++v) {
        op_dirX(n1, m->D1D, &FL(0, 0, v), &DF(0, 0, v));
        op_dirY(n1, m->D1D, &FL(0, 1, v), &DF(0, 1, v));
        op_dirZ(n1, m->D1D, &FL(0, 2, v), &DF(0, 2, v));
      }
      const double *E11 = m->Escale + o, *E22 = m->Escale + o + NpNe * 4, *E33 = m->Escale + o + NpNe * 8;
      if (del_flux) {
        for (int p = 0; p < Np; ++p) {
          DENS_dt[o + p] = -(E11[p] * DF(p, 0, DENS_VID) + E22[p] * DF(p, 1, DENS_VID) +
                             E33[p] * DF(p, 2, DENS_VID) + DF(p, 3, DENS_VID)) * RGsqrt[p];
          RHOT_dt[o + p] = -(E11[p] * DF(p, 0, RHOT_VID) + E22[p] * DF(p, 1, RHOT_VID) +
                             E33[p] * DF(p, 2, RHOT_VID) + DF(p, 3, RHOT_VID)) * RGsqrt[p];
        }
      } else { /* ..._rhot_heve.F90:614-626 */
        for (int p = 0; p < Np; ++p) {
          DENS_dt[o + p] = -(E11[p] * DF(p, 0, DENS_VID) + E22[p] * DF(p, 1, DENS_VID) +
                             E33[p] * DF(p, 2, DENS_VID)) * RGsqrt[p];
          RHOT_dt[o + p] = -(E11[p] * DF(p, 0, RHOT_VID) + E22[p] * DF(p, 1, RHOT_VID) +
                             E33[p] * DF(p, 2, RHOT_VID)) * RGsqrt[p];
        }
      }
      op_dirZ_tr(n1, m->IntrpMat_tr, DDENS_ + o, drho);
      for (int p = 0; p < Np; ++p) {
        double X = E11[p] * DF(p, 0, MOMZ_VID) + E22[p] * DF(p, 1, MOMZ_VID) + E33[p] * DF(p, 2, MOMZ_VID);
        if (del_flux) X = X + DF(p, 3, MOMZ_VID);
        MOMZ_dt[o + p] = -X * RGsqrt[p] - Grav * drho[p];
      }
      for (int p = 0; p < Np; ++p) {
        double cor = CORIOLIS[(m->IndexH2Dto3D[p] - 1) + (long)Nfp_v * ke2d];
        MOMX_dt[o + p] = -DPhydDx[o + p] + cor * MOMY_[o + p];
        MOMY_dt[o + p] = -DPhydDy[o + p] - cor * MOMX_[o + p];
      }
      for (int p = 0; p < Np; ++p) {
        double XU = E11[p] * DF(p, 0, MOMX_VID) + E22[p] * DF(p, 1, MOMX_VID) + E33[p] * DF(p, 2, MOMX_VID);
        double XV = E11[p] * DF(p, 0, MOMY_VID) + E22[p] * DF(p, 1, MOMY_VID) + E33[p] * DF(p, 2, MOMY_VID);
        if (del_flux) {
          XU = XU + DF(p, 3, MOMX_VID);
          XV = XV + DF(p, 3, MOMY_VID);
        }
        MOMX_dt[o + p] = MOMX_dt[o + p] - XU * RGsqrt[p];
        MOMY_dt[o + p] = MOMY_dt[o + p] - XV * RGsqrt[p];
      }
#undef FL
#undef DF
    }
    free(Flux);
    free(DFlux);
    free(drho);
    free(RDENS);
    free(RGsqrtV);
    free(RGsqrt);
  }
}

/* The reference computes the face flux of every element first
 * (get_ebnd_flux, :350-357) and then runs the element loop. */
void oracle_heve_cal_tend_range(const oracle_mesh *m, int ke_s, int ke_e,
                          double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt, double *RHOT_dt,
                          const double *DDENS_, const double *MOMX_, const double *MOMY_,
                          const double *MOMZ_, const double *DRHOT_, const double *DPRES_,
                          const double *DENS_hyd, const double *PRES_hyd, const double *PRES_hyd_ref,
                          const double *THERM_hyd, const double *CORIOLIS,
                          const double *Rtot, const double *CVtot, const double *CPtot,
                          const double *DPhydDx, const double *DPhydDy) {
  (void)PRES_hyd_ref;
  (void)Rtot;
  (void)CVtot;
  (void)CPtot; /* part of the interface; the dry-air constants are used instead */
  double *del_flux = (double *)malloc(sizeof(double) * (size_t)m->NfpTot * 5 * m->Ne);
  heve_numflux_range(m, ke_s, ke_e, del_flux, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES_, DENS_hyd, PRES_hyd,
                     THERM_hyd, 0);
  heve_elements(m, ke_s, ke_e, del_flux, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, DDENS_, MOMX_, MOMY_,
                MOMZ_, DRHOT_, DPRES_, DENS_hyd, THERM_hyd, CORIOLIS, DPhydDx, DPhydDy);
  free(del_flux);
}

void oracle_heve_cal_tend(const oracle_mesh *m,
                          double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt, double *RHOT_dt,
                          const double *DDENS_, const double *MOMX_, const double *MOMY_,
                          const double *MOMZ_, const double *DRHOT_, const double *DPRES_,
                          const double *DENS_hyd, const double *PRES_hyd, const double *PRES_hyd_ref,
                          const double *THERM_hyd, const double *CORIOLIS,
                          const double *Rtot, const double *CVtot, const double *CPtot,
                          const double *DPhydDx, const double *DPhydDy) {
  oracle_heve_cal_tend_range(m, 0, m->Ne, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, DDENS_, MOMX_, MOMY_,
                             MOMZ_, DRHOT_, DPRES_, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd, CORIOLIS, Rtot,
                             CVtot, CPtot, DPhydDx, DPhydDy);
}

/* FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:493-678 (cal_tend_cco):
 * the interior part of the tendency, no face flux */
void oracle_heve_cal_tend_cco(const oracle_mesh *m,
                              double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt, double *RHOT_dt,
                              const double *DDENS_, const double *MOMX_, const double *MOMY_,
                              const double *MOMZ_, const double *DRHOT_, const double *DPRES_,
                              const double *DENS_hyd, const double *THERM_hyd, const double *CORIOLIS,
                              const double *DPhydDx, const double *DPhydDy) {
  heve_elements(m, 0, m->Ne, NULL, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_,
                DPRES_, DENS_hyd, THERM_hyd, CORIOLIS, DPhydDx, DPhydDy);
}

/* FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:242-496
 * (add_bnd_contrib_generalvc): DPRES of the halo tail (:331-335), the face
 * flux with the Cs2 grouping, Lift_var5 (ELM/..._tensorprod3D.F90:2810-2823)
 * and dt -= RGsqrt * Lift (:488-492).  Rtot NULL: the dry-air constants. */
void oracle_heve_add_bnd_contrib(const oracle_mesh *m,
                                 double *DENS_dt, double *MOMX_dt, double *MOMY_dt, double *MOMZ_dt, double *RHOT_dt,
                                 const double *DDENS_, const double *MOMX_, const double *MOMY_,
                                 const double *MOMZ_, const double *DRHOT_, double *DPRES,
                                 const double *DENS_hyd, const double *PRES_hyd, const double *THERM_hyd,
                                 const double *Rtot, const double *CVtot, const double *CPtot) {
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Ne = m->Ne;
  const double rP0 = 1.0 / m->PRES00;
  const long h0 = (long)Np * Ne;
  for (long i = h0; i < h0 + m->nB; ++i) {
    double Rt = Rtot ? Rtot[i] : m->Rdry, Cv = CVtot ? CVtot[i] : m->CVdry, Cp = CPtot ? CPtot[i] : m->CPdry;
    double PRES = m->PRES00 * opow(Rt * rP0 * (THERM_hyd[i] + DRHOT_[i]), Cp / Cv);
    DPRES[i] = PRES - PRES_hyd[i];
  }
  double *del_flux = (double *)malloc(sizeof(double) * (size_t)NfpTot * 5 * Ne);
  heve_numflux_range(m, 0, Ne, del_flux, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_, DPRES, DENS_hyd, PRES_hyd, THERM_hyd,
                     1);
  double *dt[5];
  dt[DENS_VID] = DENS_dt;
  dt[RHOT_VID] = RHOT_dt;
  dt[MOMZ_VID] = MOMZ_dt;
  dt[MOMX_VID] = MOMX_dt;
  dt[MOMY_VID] = MOMY_dt;
#pragma omp parallel
  {
    double *Lift = (double *)malloc(sizeof(double) * Np);
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      const long o = (long)Np * ke;
      for (int v = 0; v < 5; ++v) {
        op_lift(n1, m->Lift_mat, del_flux + (long)NfpTot * (v + 5 * ke), Lift);
        for (int p = 0; p < Np; ++p) {
          double RGsqrt = 1.0 / m->Gsqrt[o + p];
          dt[v][o + p] = dt[v][o + p] - RGsqrt * Lift[p];
        }
      }
    }
    free(Lift);
  }
  free(del_flux);
}

/* FaceIntMat of atm_dyn_dgm_trcadvect3d_heve_Init (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:75-137):
 * the nonzero of each column, w(a0)*w(a1) of the 1D LGL weights w1d */
static double *trc_face_weights(const oracle_mesh *m, const double *w1d) {
  const int n1 = m->n1, Nfp_h = m->Nfp_h, Nfp_v = m->Nfp_v;
  double *IntW = (double *)malloc(sizeof(double) * m->NfpTot);
  for (int f = 0; f < 4; ++f)
    for (int k = 0; k < n1; ++k)
      for (int i = 0; i < n1; ++i) IntW[f * Nfp_h + i + n1 * k] = w1d[i] * w1d[k];
  for (int f = 0; f < 2; ++f)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) IntW[4 * Nfp_h + f * Nfp_v + i + n1 * j] = w1d[i] * w1d[j];
  return IntW;
}

/* The face flux of element ke shared by get_delflux_generalhvc (:554-674)
 * and get_netOutwardFlux_generalhvc (:676-797): MomFlxM, numflux and the six
 * LGL-weighted face integrals outward(f) of J(iM) Fscale numflux, applied as the
 * reference's CSR product (CMN/scale_sparsemat.F90:439-474): row f sums its
 * face's nodes in storage order from s = 0. */
static void trc_face_flux(const oracle_mesh *m, const double *IntW, int ke, const double *QTRC_, const double *MOMX_,
                          const double *MOMY_, const double *MOMZ_, const double *alphDENS_M,
                          const double *alphDENS_P, const double *J, double *MomFlxM, double *numflux,
                          double *outward) {
  const int NfpTot = m->NfpTot, Nfp_h = m->Nfp_h, Nfp_v = m->Nfp_v, Ne = m->Ne;
  const long NpNeA = (long)m->Np * m->NeA;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  const double *nx = m->normal_fn, *ny = m->normal_fn + (long)NfpTot * Ne, *nz = m->normal_fn + 2L * NfpTot * Ne;
  const int ke2d = m->EMap3Dto2D[ke] - 1;
  const long of = (long)NfpTot * ke;
  double b[6 * 64]; /* six faces of at most 8 x 8 nodes (p <= 7) */
  for (int fp = 0; fp < NfpTot; ++fp) {
    const long iM = m->vmapM[of + fp] - 1, iP = m->vmapP[of + fp] - 1;
    const double GsH = m->GsqrtH[(m->IndexH2Dto3D_bnd[fp] - 1) + (long)Nfp_v * ke2d];
    const double Gsqrt_M = m->Gsqrt[iM], Gsqrt_P = m->Gsqrt[iP];
    const double GsqrtV_M = Gsqrt_M / GsH, GsqrtV_P = Gsqrt_P / GsH;
    const double GMX_M = Gsqrt_M * MOMX_[iM], GMX_P = Gsqrt_P * MOMX_[iP];
    const double GMY_M = Gsqrt_M * MOMY_[iM], GMY_P = Gsqrt_P * MOMY_[iP];
    const double GMZ_M = Gsqrt_M * MOMZ_[iM], GMZ_P = Gsqrt_P * MOMZ_[iP];
    MomFlxM[fp] = GMX_M * nx[of + fp] + GMY_M * ny[of + fp] +
                  ((GMZ_M / GsqrtV_M + G13[iM] * GMX_M + G23[iM] * GMY_M) * nz[of + fp]);
    const double MomFlxP = GMX_P * nx[of + fp] + GMY_P * ny[of + fp] +
                           ((GMZ_P / GsqrtV_P + G13[iP] * GMX_P + G23[iP] * GMY_P) * nz[of + fp]);
    numflux[fp] = 0.5 * ((QTRC_[iP] * MomFlxP + QTRC_[iM] * MomFlxM[fp]) - alphDENS_P[of + fp] * QTRC_[iP] +
                         alphDENS_M[of + fp] * QTRC_[iM]);
    b[fp] = J[iM] * m->Fscale[of + fp] * numflux[fp];
  }
  for (int f = 0; f < 6; ++f) {
    const int is = f < 4 ? f * Nfp_h : 4 * Nfp_h + (f - 4) * Nfp_v, n = f < 4 ? Nfp_h : Nfp_v;
    double s = 0.0;
    for (int c = is; c < is + n; ++c) s = s + IntW[c] * b[c];
    outward[f] = s;
  }
}

/* FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231 (cal_tend) with
 * get_delflux_generalhvc (:554-674): del_flux with the FCT coefficient of the
 * upwind side, sign(1, outward(f)) of the face's net flux; Div
 * (ELM/..._tensorprod3D.F90) is the three line contractions plus the lift of
 * del_flux. */
void oracle_trcadvect_cal_tend(const oracle_mesh *m, const double *w1d, double *QTRC_dt, const double *QTRC_,
                               const double *MOMX_, const double *MOMY_, const double *MOMZ_,
                               const double *alphDENS_M, const double *alphDENS_P, const double *fct_coef,
                               const double *RHOQ_tp, const double *J) {
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Nfp_h = m->Nfp_h, Nfp_v = m->Nfp_v, Ne = m->Ne;
  const long NpNeA = (long)Np * m->NeA;
  const long NpNe = (long)Np * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  double *IntW = trc_face_weights(m, w1d);
#pragma omp parallel
  {
    double *MomFlxM = (double *)malloc(sizeof(double) * NfpTot), *numflux = (double *)malloc(sizeof(double) * NfpTot);
    double *del_flux = (double *)malloc(sizeof(double) * NfpTot);
    double *Flux = (double *)malloc(sizeof(double) * Np * 3), *DFlux = (double *)malloc(sizeof(double) * Np * 4);
    double outward[6];
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      const int ke2d = m->EMap3Dto2D[ke] - 1;
      const long of = (long)NfpTot * ke, o = (long)Np * ke;
      trc_face_flux(m, IntW, ke, QTRC_, MOMX_, MOMY_, MOMZ_, alphDENS_M, alphDENS_P, J, MomFlxM, numflux, outward);
      for (int fp = 0; fp < NfpTot; ++fp) {
        const long iM = m->vmapM[of + fp] - 1, iP = m->vmapP[of + fp] - 1;
        const double RM = fct_coef[iM], RP = fct_coef[iP];
        const int f = fp < 4 * Nfp_h ? fp / Nfp_h : 4 + (fp - 4 * Nfp_h) / Nfp_v;
        del_flux[fp] = m->Fscale[of + fp] *
                       (numflux[fp] * 0.5 * (RP + RM - (RP - RM) * copysign(1.0, outward[f])) -
                        QTRC_[iM] * MomFlxM[fp]);
      }
      for (int p = 0; p < Np; ++p) {
        const double Gsqrt_ = m->Gsqrt[o + p];
        const double RGsqrt = 1.0 / Gsqrt_;
        const double RGsqrtV = m->GsqrtH[(m->IndexH2Dto3D[p] - 1) + (long)Nfp_v * ke2d] * RGsqrt;
        Flux[p] = Gsqrt_ * MOMX_[o + p] * QTRC_[o + p];
        Flux[p + Np] = Gsqrt_ * MOMY_[o + p] * QTRC_[o + p];
        Flux[p + 2 * Np] =
            Gsqrt_ * (MOMZ_[o + p] * RGsqrtV + G13[o + p] * MOMX_[o + p] + G23[o + p] * MOMY_[o + p]) * QTRC_[o + p];
      }
      op_dirX(n1, m->D1D, Flux, DFlux);
      op_dirY(n1, m->D1D, Flux + Np, DFlux + Np);
      op_dirZ(n1, m->D1D, Flux + 2 * Np, DFlux + 2 * Np);
      op_lift(n1, m->Lift_mat, del_flux, DFlux + 3 * Np);
      const double *E11 = m->Escale + o, *E22 = m->Escale + o + NpNe * 4, *E33 = m->Escale + o + NpNe * 8;
      for (int p = 0; p < Np; ++p)
        QTRC_dt[o + p] = -(E11[p] * DFlux[p] + E22[p] * DFlux[p + Np] + E33[p] * DFlux[p + 2 * Np] +
                           DFlux[p + 3 * Np]) / m->Gsqrt[o + p] +
                         RHOQ_tp[o + p];
    }
    free(MomFlxM);
    free(numflux);
    free(del_flux);
    free(Flux);
    free(DFlux);
  }
  free(IntW);
}

/* IntWeight_lgl of the hexahedron (ELM/scale_element_hexahedral.F90:286-298) */
static double trc_vol_weight(const double *w1d, int n1, int p) {
  const int i = p % n1, j = (p / n1) % n1, k = p / (n1 * n1);
  return w1d[i] * w1d[j] * w1d[k];
}

/* FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:232-305 (calc_fct_coef) with
 * get_netOutwardFlux_generalhvc (:676-797): net = sum_f max(0, outward(f)),
 * Q = sum_p Gsqrt J IntWeight_lgl (dens_ssm1 QTRC / dt + RHOQ_tp), both summed
 * in order from zero as Fortran's sum does. */
void oracle_trcadvect_calc_fct_coef(const oracle_mesh *m, const double *w1d, double *fct_coef, const double *QTRC_,
                                    const double *MOMX_, const double *MOMY_, const double *MOMZ_,
                                    const double *RHOQ_tp_, const double *alphDENS_M, const double *alphDENS_P,
                                    const double *DENS_hyd, const double *DDENS_, const double *DDENS0_,
                                    double rk_c_ssm1, double dt, const double *J, int disable_limiter) {
  const int Np = m->Np, NfpTot = m->NfpTot, Ne = m->Ne;
  if (disable_limiter) {
    for (long t = 0; t < (long)Np * Ne; ++t) fct_coef[t] = 1.0;
    return;
  }
  double *IntW = trc_face_weights(m, w1d);
#pragma omp parallel
  {
    double *MomFlxM = (double *)malloc(sizeof(double) * NfpTot), *numflux = (double *)malloc(sizeof(double) * NfpTot);
    double outward[6];
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      const long o = (long)Np * ke;
      trc_face_flux(m, IntW, ke, QTRC_, MOMX_, MOMY_, MOMZ_, alphDENS_M, alphDENS_P, J, MomFlxM, numflux, outward);
      double net = 0.0;
      for (int f = 0; f < 6; ++f) net = net + (outward[f] > 0.0 ? outward[f] : 0.0);
      double Q = 0.0;
      for (int p = 0; p < Np; ++p) {
        const double dens_ssm1 = DENS_hyd[o + p] + (1.0 - rk_c_ssm1) * DDENS0_[o + p] + rk_c_ssm1 * DDENS_[o + p];
        Q = Q + m->Gsqrt[o + p] * J[o + p] * trc_vol_weight(w1d, m->n1, p) *
                    (dens_ssm1 * QTRC_[o + p] / dt + RHOQ_tp_[o + p]);
      }
      double r = Q / (net + 1.0e-10);
      r = r < 1.0 ? r : 1.0;
      r = r > 0.0 ? r : 0.0;
      for (int p = 0; p < Np; ++p) fct_coef[o + p] = r;
    }
    free(MomFlxM);
    free(numflux);
  }
  free(IntW);
}

/* FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:308-337 (TMAR) */
void oracle_trcadvect_tmar(const oracle_mesh *m, const double *w1d, double *QTRC_, const double *DENS_hyd,
                           const double *DDENS_, const double *J) {
  const int Np = m->Np;
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < m->Ne; ++ke) {
    const long o = (long)Np * ke;
    double Q0 = 0.0, Q1 = 0.0;
    for (int p = 0; p < Np; ++p) {
      const double dens = DENS_hyd[o + p] + DDENS_[o + p];
      const double q = QTRC_[o + p] > 0.0 ? QTRC_[o + p] : 0.0;
      const double g = m->Gsqrt[o + p] * J[o + p] * trc_vol_weight(w1d, m->n1, p) * dens;
      Q0 = Q0 + g * QTRC_[o + p];
      Q1 = Q1 + g * q;
    }
    const double r = Q0 / (Q1 + 1.0e-32);
    for (int p = 0; p < Np; ++p) QTRC_[o + p] = r * (QTRC_[o + p] > 0.0 ? QTRC_[o + p] : 0.0);
  }
}

/* SCALE's ATMOS_THERMODYN_specific_heat (scale_atmos_thermodyn, not in the
 * reference tree): the published mixing rule, the tracer sums in order */
static void spec_heat(int nq, const double *q, const double *Mq, const double *Rq, const double *CVq,
                      const double *CPq, const oracle_mesh *m, double *R, double *CV, double *CP) {
  double qdry = 1.0, r = 0.0, cv = 0.0, cp = 0.0;
  for (int iq = 0; iq < nq; ++iq) {
    qdry = qdry - q[iq] * Mq[iq];
    r = r + q[iq] * Rq[iq];
    cv = cv + q[iq] * CVq[iq];
    cp = cp + q[iq] * CPq[iq];
  }
  *R = r + qdry * m->Rdry;
  *CV = cv + qdry * m->CVdry;
  *CP = cp + qdry * m->CPdry;
}

/* FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:653-716 (update_pressure_specific_heat,
 * its specific-heat part) over the nodes of elements 1..Ne */
void oracle_tracer_specific_heat(const oracle_mesh *m, int nq, double *const *Q, const double *Mq, const double *Rq,
                                 const double *CVq, const double *CPq, double *Rtot, double *CVtot, double *CPtot) {
  const long n = (long)m->Np * m->Ne;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < n; ++t) {
    double q[8];
    for (int iq = 0; iq < nq; ++iq) q[iq] = Q[iq][t];
    spec_heat(nq, q, Mq, Rq, CVq, CPq, m, &Rtot[t], &CVtot[t], &CPtot[t]);
  }
}

/* model/atm_nonhydro3d/src/atmos/mod_atmos_vars_container.F90:1100-1366
 * (vars_calc_diagnoseVar_lc), :431-444 (Umet, Vmet of the regional mesh:
 * MOMX, MOMY over DENS_hyd + DDENS) and QDRY of Calc_SpecificHeat; the box
 * mesh's G_ij is the identity.  Returns 0, or -1 for a name the reference
 * does not diagnose here. */
int oracle_atmos_calc_diagvar(const oracle_mesh *m, const char *name, double *out, const double *DD,
                              const double *MX, const double *MY, const double *MZ, const double *P, int nq,
                              double *const *Q, const double *Mq, const double *CVq, const double *E0q,
                              const double *Dh, const double *Ph, const double *R, const double *CP,
                              const double *zlev, const double *zpos) {
  const long n = (long)m->Np * m->Ne;
  const double G_11 = 1.0, G_12 = 0.0, G_22 = 1.0;
  int ok = 1;
#define NODES _Pragma("omp parallel for schedule(static)") for (long i = 0; i < n; ++i)
  if (!strcmp(name, "DENS")) {
    NODES out[i] = DD[i] + Dh[i];
  } else if (!strcmp(name, "U")) {
    NODES { double DENS = DD[i] + Dh[i]; out[i] = MX[i] / DENS; }
  } else if (!strcmp(name, "V")) {
    NODES { double DENS = DD[i] + Dh[i]; out[i] = MY[i] / DENS; }
  } else if (!strcmp(name, "W")) {
    NODES { double DENS = DD[i] + Dh[i]; out[i] = MZ[i] / DENS; }
  } else if (!strcmp(name, "Umet")) {
    NODES out[i] = MX[i] / (Dh[i] + DD[i]);
  } else if (!strcmp(name, "Vmet")) {
    NODES out[i] = MY[i] / (Dh[i] + DD[i]);
  } else if (!strcmp(name, "PRES")) {
    /* empty case */
  } else if (!strcmp(name, "PRES_diff")) {
    NODES out[i] = P[i] - Ph[i];
  } else if (!strcmp(name, "T")) {
    NODES out[i] = P[i] / (R[i] * (DD[i] + Dh[i]));
  } else if (!strcmp(name, "T_diff")) {
    NODES out[i] = P[i] / (R[i] * (DD[i] + Dh[i])) - Ph[i] / (m->Rdry * Dh[i]);
  } else if (!strcmp(name, "PT")) {
    NODES {
      double DENS = DD[i] + Dh[i];
      out[i] = P[i] / (R[i] * DENS) * oracle_opow(m->PRES00 / P[i], R[i] / CP[i]);
    }
  } else if (!strcmp(name, "PT_diff")) {
    NODES {
      double DENS = DD[i] + Dh[i];
      out[i] = P[i] / (R[i] * DENS) * oracle_opow(m->PRES00 / P[i], R[i] / CP[i]) -
               m->PRES00 / m->Rdry * oracle_opow(Ph[i] / m->PRES00, m->CVdry / m->CPdry) / Dh[i];
    }
  } else if (!strcmp(name, "QDRY")) {
    NODES {
      double qdry = 1.0;
      for (int iq = 0; iq < nq; ++iq) qdry = qdry - Q[iq][i] * Mq[iq];
      out[i] = qdry;
    }
  } else if (!strcmp(name, "ENGK")) {
    NODES {
      double DENS = DD[i] + Dh[i];
      double mom_u1 = G_11 * MX[i] + G_12 * MY[i];
      double mom_u2 = G_12 * MX[i] + G_22 * MY[i];
      out[i] = 0.5 * (MX[i] * mom_u1 + MY[i] * mom_u2 + MZ[i] * MZ[i]) / DENS;
    }
  } else if (!strcmp(name, "ENGP")) {
    NODES { double DENS = DD[i] + Dh[i]; out[i] = DENS * m->GRAV * zlev[i]; }
  } else if (!strcmp(name, "ENGI") || !strcmp(name, "ENGT")) {
    const int engt = !strcmp(name, "ENGT");
    NODES {
      double DENS = DD[i] + Dh[i];
      double QDRY = 1.0;
      for (int iq = 0; iq < nq; ++iq) QDRY = QDRY - Q[iq][i] * Mq[iq];
      double v = QDRY * P[i] / R[i] * m->CVdry;
      for (int iq = 0; iq < nq; ++iq) v = v + Q[iq][i] * (P[i] / R[i] * CVq[iq] + DENS * E0q[iq]);
      if (engt) {
        double mom_u1 = G_11 * MX[i] + G_12 * MY[i];
        double mom_u2 = G_12 * MX[i] + G_22 * MY[i];
        v = 0.5 * (MX[i] * mom_u1 + MY[i] * mom_u2 + MZ[i] * MZ[i]) / DENS + v + DENS * m->GRAV * zpos[i];
      }
      out[i] = v;
    }
  } else {
    ok = 0;
  }
#undef NODES
  return ok ? 0 : -1;
}

/* FElib/src/microphysics/scale_atm_phy_mp_dgm_common.F90:429-556
 * (atm_phy_mp_dgm_common_negative_fixer; its water species are the nq =
 * 1 + QLA + QIA tracers), called by fix_negative_val
 * (FDS/scale_atm_dyn_dgm_driver_trcadv3d.F90:596-651) */
void oracle_trc_negative_fixer(const oracle_mesh *m, const double *w1d, int nq, double *const *Q, double *DDENS,
                               double *PRES, double *CVtot, double *CPtot, double *Rtot, const double *DENS_hyd,
                               const double *PRES_hyd, const double *J, const double *zlev, const double *Mq,
                               const double *Rq, const double *CVq, const double *CPq, double *DRHOT) {
  const int Np = m->Np;
  const double TRC_EPS = 1e-128;
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < m->Ne; ++ke) {
    const long o = (long)Np * ke;
    double DENS[512], DDENS0[512], IE0[512], IE[512], TEMP[512], intw[512], TM0[512];
    double TM1[8][512], Qt[8][512];
    for (int p = 0; p < Np; ++p) {
      DENS[p] = DENS_hyd[o + p] + DDENS[o + p];
      DDENS0[p] = DDENS[o + p];
      IE0[p] = CVtot[o + p] * PRES[o + p] / Rtot[o + p];
      TEMP[p] = IE0[p] / (DENS[p] * CVtot[o + p]);
      IE[p] = IE0[p];
      intw[p] = m->Gsqrt[o + p] * J[o + p] * trc_vol_weight(w1d, m->n1, p);
    }
    for (int iq = 0; iq < nq; ++iq) {
      for (int p = 0; p < Np; ++p) {
        TM0[p] = DENS[p] * Q[iq][o + p];
        TM1[iq][p] = TM0[p] > TRC_EPS ? TM0[p] : TRC_EPS;
      }
      double M0 = 0.0, M1 = 0.0;
      for (int p = 0; p < Np; ++p) M0 = M0 + intw[p] * TM0[p];
      for (int p = 0; p < Np; ++p) M1 = M1 + intw[p] * TM1[iq][p];
      for (int p = 0; p < Np; ++p) {
        TM1[iq][p] = (M0 > 0.0 ? M0 : 0.0) / M1 * TM1[iq][p];
        DDENS[o + p] = DDENS[o + p] + (TM1[iq][p] - TM0[p]);
        IE[p] = IE[p] + (TM1[iq][p] - TM0[p]) * CVq[iq] * TEMP[p];
      }
    }
    for (int p = 0; p < Np; ++p) DENS[p] = DENS_hyd[o + p] + DDENS[o + p];
    for (int iq = 0; iq < nq; ++iq)
      for (int p = 0; p < Np; ++p) {
        Qt[iq][p] = TM1[iq][p] / DENS[p];
        Q[iq][o + p] = Qt[iq][p];
      }
    double E0 = 0.0, E1 = 0.0;
    for (int p = 0; p < Np; ++p) E0 = E0 + intw[p] * IE0[p];
    for (int p = 0; p < Np; ++p) E1 = E1 + intw[p] * IE[p];
    for (int p = 0; p < Np; ++p) {
      IE[p] = E0 / E1 * IE[p];
      double q[8], R, CV, CP;
      for (int iq = 0; iq < nq; ++iq) q[iq] = Qt[iq][p];
      spec_heat(nq, q, Mq, Rq, CVq, CPq, m, &R, &CV, &CP);
      Rtot[o + p] = R;
      CVtot[o + p] = CV;
      CPtot[o + p] = CP;
      IE[p] = IE[p] - (DDENS[o + p] - DDENS0[p]) * m->GRAV * zlev[o + p];
      PRES[o + p] = IE[p] * R / CV;
      if (DRHOT)
        DRHOT[o + p] = m->PRES00 / R * opow(PRES[o + p] / m->PRES00, CV / CP) -
                       m->PRES00 / m->Rdry * opow(PRES_hyd[o + p] / m->PRES00, m->CVdry / m->CPdry);
    }
  }
}

/* FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:407-445 (cal_alphdens_advtest) */
void oracle_trcadvect_cal_alphdens_advtest(const oracle_mesh *m, double *alph_M, double *alph_P, const double *DDENS_,
                                           const double *MOMX_, const double *MOMY_, const double *MOMZ_,
                                           const double *DENS_hyd) {
  const long nf = (long)m->NfpTot * m->Ne;
  const double *nx = m->normal_fn, *ny = m->normal_fn + nf, *nz = m->normal_fn + 2 * nf;
  for (long i = 0; i < nf; ++i) {
    const long iM = m->vmapM[i] - 1, iP = m->vmapP[i] - 1;
    const double densM = DDENS_[iM] + DENS_hyd[iM], densP = DDENS_[iP] + DENS_hyd[iP];
    const double VelM = (MOMX_[iM] * nx[i] + MOMY_[iM] * ny[i] + MOMZ_[iM] * nz[i]) / densM;
    const double VelP = (MOMX_[iP] * nx[i] + MOMY_[iP] * ny[i] + MOMZ_[iP] * nz[i]) / densP;
    const double alpha = fabs(VelM) > fabs(VelP) ? fabs(VelM) : fabs(VelP);
    alph_M[i] = alpha * densM * m->Gsqrt[iM];
    alph_P[i] = alpha * densP * m->Gsqrt[iP];
  }
}

/* CMN/scale_timeint_rk.F90:1390-1498 (rk_advance_trcvar_low_storage2D) over
 * nodes [0, n); sig, gam: (nstage+1, nstage) column-major; c: (nstage) */
void oracle_trcvar_rk_advance(long n, int nowstage, int nstage, const double *sig, const double *gam,
                              const double *c, double dt, double *q, double *var0, double *varTmp,
                              const double *tend, const double *DDENS, const double *DDENS0, const double *DENS_hyd) {
  const int ld = nstage + 1;
#define SIG(i, j) sig[((i)-1) + ld * ((j)-1)]
#define GAM(i, j) gam[((i)-1) + ld * ((j)-1)]
  const double EPS = 1.0e-16; /* CONST_EPS */
  const double sig_ss = SIG(nowstage + 1, nowstage), sig_Ns = SIG(nstage + 1, nowstage);
  const double one_Minus_sig_ss = 1.0 - sig_ss;
  const double gam_ss = dt * GAM(nowstage + 1, nowstage), gam_Ns = dt * GAM(nstage + 1, nowstage);
  const double c_ssm1 = c[nowstage - 1];
#undef SIG
#undef GAM
  if (nowstage == nstage) {
    for (long i = 0; i < n; ++i) {
      const double dens_ssm1 = DENS_hyd[i] + DDENS0[i] + c_ssm1 * (DDENS[i] - DDENS0[i]);
      q[i] = (varTmp[i] + sig_ss * q[i] * dens_ssm1 + gam_ss * tend[i]) / (DENS_hyd[i] + DDENS[i]);
    }
    return;
  }
  const double c_ss = c[nowstage];
  if (nowstage == 1)
    for (long i = 0; i < n; ++i) {
      var0[i] = q[i] * (DENS_hyd[i] + DDENS0[i]);
      varTmp[i] = 0.0;
    }
  if (fabs(sig_Ns) > EPS || fabs(gam_Ns) > EPS)
    for (long i = 0; i < n; ++i) {
      const double dens_ssm1 = DENS_hyd[i] + DDENS0[i] + c_ssm1 * (DDENS[i] - DDENS0[i]);
      varTmp[i] = varTmp[i] + sig_Ns * q[i] * dens_ssm1 + gam_Ns * tend[i];
    }
  for (long i = 0; i < n; ++i) {
    const double dens_ssm1 = DENS_hyd[i] + DDENS0[i] + c_ssm1 * (DDENS[i] - DDENS0[i]);
    const double dens_ss = DENS_hyd[i] + DDENS0[i] + c_ss * (DDENS[i] - DDENS0[i]);
    q[i] = (one_Minus_sig_ss * var0[i] + sig_ss * q[i] * dens_ssm1 + gam_ss * tend[i]) / dens_ss;
  }
}

/* FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:340-404 (save_massflux) with
 * cal_alphdens_dyn (:447-536); gamma = CPdry / CVdry as there */
void oracle_trcadvect_save_massflux(const oracle_mesh *m, double *MFLX_x, double *MFLX_y, double *MFLX_z,
                                    double *alph_M, double *alph_P, const double *DDENS, const double *MOMX,
                                    const double *MOMY, const double *MOMZ, const double *DPRES,
                                    const double *DENS_hyd, const double *PRES_hyd, int rkstage, double twh,
                                    double twv, int is_hevi) {
  const int Np = m->Np, NfpTot = m->NfpTot, Nfp_v = m->Nfp_v, Ne = m->Ne;
  const double *nx = m->normal_fn, *ny = m->normal_fn + (long)NfpTot * Ne, *nz = m->normal_fn + 2L * NfpTot * Ne;
  const double gamm = m->CPdry / m->CVdry;
  for (long t = 0; t < (long)Np * Ne; ++t) {
    if (rkstage == 1) {
      MFLX_x[t] = twh * MOMX[t];
      MFLX_y[t] = twh * MOMY[t];
      MFLX_z[t] = twv * MOMZ[t];
    } else {
      MFLX_x[t] = MFLX_x[t] + twh * MOMX[t];
      MFLX_y[t] = MFLX_y[t] + twh * MOMY[t];
      MFLX_z[t] = MFLX_z[t] + twv * MOMZ[t];
    }
  }
  if (rkstage == 1)
    for (long t = 0; t < (long)NfpTot * Ne; ++t) alph_M[t] = alph_P[t] = 0.0;
  for (int ke = 0; ke < Ne; ++ke) {
    const int ke2d = m->EMap3Dto2D[ke] - 1;
    for (int fp = 0; fp < NfpTot; ++fp) {
      const long g = (long)NfpTot * ke + fp;
      const long iM = m->vmapM[g] - 1, iP = m->vmapP[g] - 1;
      const long h = (m->IndexH2Dto3D_bnd[fp] - 1) + (long)Nfp_v * ke2d;
      const double densM = DDENS[iM] + DENS_hyd[iM], densP = DDENS[iP] + DENS_hyd[iP];
      double Gnn = m->G11[h] * fabs(nx[g]) + m->G22[h] * fabs(ny[g]);
      if (!is_hevi) Gnn = Gnn + fabs(nz[g]);
      const double VelM = (MOMX[iM] * nx[g] + MOMY[iM] * ny[g] + MOMZ[iM] * nz[g]) / densM;
      const double VelP = (MOMX[iP] * nx[g] + MOMY[iP] * ny[g] + MOMZ[iP] * nz[g]) / densP;
      const double cM = sqrt(Gnn * gamm * (PRES_hyd[iM] + DPRES[iM]) / densM) + fabs(VelM);
      const double cP = sqrt(Gnn * gamm * (PRES_hyd[iP] + DPRES[iP]) / densP) + fabs(VelP);
      const double alpha = cM > cP ? cM : cP;
      const double tw = twh * (fabs(nx[g]) + fabs(ny[g])) + twv * fabs(nz[g]);
      alph_M[g] = alph_M[g] + tw * alpha * densM * m->Gsqrt[iM];
      alph_P[g] = alph_P[g] + tw * alpha * densP * m->Gsqrt[iP];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:624-777 with gam = 1 (the
 * regional mesh sets lcmesh%gam = 1, MSH/scale_mesh_base3d.F90:316) */
void oracle_calc_phyd_hgrad(const oracle_mesh *m, double *DPhydDx, double *DPhydDy,
                            const double *PRES_hyd, const double *PRES_hyd_ref) {
  const int n1 = m->n1, Np = m->Np, NfpTot = m->NfpTot, Nfp_v = m->Nfp_v, Ne = m->Ne;
  const long NpNeA = (long)Np * m->NeA, NpNe = (long)Np * Ne, nfall = (long)NfpTot * Ne;
  const double *G13 = m->GI3, *G23 = m->GI3 + NpNeA;
  double *dfh = (double *)malloc(sizeof(double) * (size_t)NfpTot * 2 * Ne);
#pragma omp parallel for schedule(static)
  for (int ke = 0; ke < Ne; ++ke) {
    int ke2D = m->EMap3Dto2D[ke] - 1;
    for (int fp = 0; fp < NfpTot; ++fp) {
      long i = fp + (long)NfpTot * ke;
      long iM = m->vmapM[i] - 1, iP = m->vmapP[i] - 1;
      double GH = m->GsqrtH[(m->IndexH2Dto3D_bnd[fp] - 1) + (long)Nfp_v * ke2D];
      double GsqrtV_M = m->Gsqrt[iM] / 1.0 / GH;
      double GsqrtV_P = m->Gsqrt[iP] / 1.0 / GH;
      double DP_M = PRES_hyd[iM] - PRES_hyd_ref[iM];
      double DP_P = PRES_hyd[iP] - PRES_hyd_ref[iP];
      double nx = m->normal_fn[i], ny = m->normal_fn[i + nfall], nz = m->normal_fn[i + 2 * nfall];
      double tmp1 = m->Fscale[i] * 0.5 * GsqrtV_P * DP_P;
      double tmp2 = m->Fscale[i] * 0.5 * GsqrtV_M * DP_M;
      dfh[fp + NfpTot * (0 + 2 * ke)] = (nx + G13[iP] * nz) * tmp1 - (nx + G13[iM] * nz) * tmp2;
      dfh[fp + NfpTot * (1 + 2 * ke)] = (ny + G23[iP] * nz) * tmp1 - (ny + G23[iM] * nz) * tmp2;
    }
  }
#pragma omp parallel
  {
    double *F = (double *)malloc(sizeof(double) * Np * 4);
    double *D = (double *)malloc(sizeof(double) * Np * 8);
    double *RG = (double *)malloc(sizeof(double) * Np);
#pragma omp for schedule(static)
    for (int ke = 0; ke < Ne; ++ke) {
      int ke2d = m->EMap3Dto2D[ke] - 1;
      long o = (long)Np * ke;
      for (int p = 0; p < Np; ++p) {
        double GsqrtV = m->Gsqrt[o + p] / (1.0 * m->GsqrtH[(m->IndexH2Dto3D[p] - 1) + (long)Nfp_v * ke2d]);
        RG[p] = 1.0 / GsqrtV;
        F[p] = GsqrtV * (PRES_hyd[o + p] - PRES_hyd_ref[o + p]);
      }
      for (int p = 0; p < Np; ++p) {
        F[p + Np] = F[p];
        F[p + 2 * Np] = G13[o + p] * F[p];
        F[p + 3 * Np] = G23[o + p] * F[p];
      }
      /* Div(Flux, del_flux_hyd(:,1,ke)) -> D(:,1..4,1); Dz(Fz) -> D(:,3,2); Lift -> D(:,4,2) */
      op_lift(n1, m->Lift_mat, dfh + (long)NfpTot * (0 + 2 * ke), D + 3 * Np);
      op_dirX(n1, m->D1D, F, D);
      op_dirY(n1, m->D1D, F + Np, D + Np);
      op_dirZ(n1, m->D1D, F + 2 * Np, D + 2 * Np);
      op_dirZ(n1, m->D1D, F + 3 * Np, D + 6 * Np);
      op_lift(n1, m->Lift_mat, dfh + (long)NfpTot * (1 + 2 * ke), D + 7 * Np);
      const double *E11 = m->Escale + o, *E22 = m->Escale + o + NpNe * 4, *E33 = m->Escale + o + NpNe * 8;
      for (int p = 0; p < Np; ++p) {
        double gx = E11[p] * D[p] + E33[p] * D[p + 2 * Np] + D[p + 3 * Np];
        double gy = E22[p] * D[p + Np] + E33[p] * D[p + 6 * Np] + D[p + 7 * Np];
        DPhydDx[o + p] = gx * RG[p];
        DPhydDy[o + p] = gy * RG[p];
      }
    }
    free(F);
    free(D);
    free(RG);
  }
  free(dfh);
}

/* ---------------------------------------------------------------------- */
/* CMN/scale_timeint_rk_butcher_tab.F90:60-324 */
int oracle_rk_tab_get(const char *name, double dt, oracle_rk_tab *t) {
  memset(t, 0, sizeof(*t));
  t->dt = dt;
#define A_EX(s, u) t->a_ex[((s)-1) * 4 + ((u)-1)]
#define A_IM(s, u) t->a_im[((s)-1) * 4 + ((u)-1)]
  if (!strcmp(name, "ERK_4s4o") || !strcmp(name, "ERK_RK4")) {
    t->nstage = 4;
    t->tend_buf_size = 1;
    t->imex = 0;
    A_EX(2, 1) = 0.5;
    A_EX(3, 2) = 0.5;
    A_EX(4, 3) = 1.0;
    t->b_ex[0] = 1.0 / 6.0;
    t->b_ex[1] = 2.0 / 6.0;
    t->b_ex[2] = 2.0 / 6.0;
    t->b_ex[3] = 1.0 / 6.0;
    for (int s = 0; s < 4; ++s) t->tend_buf_indmap[s] = 0;
  } else if (!strcmp(name, "ERK_1s1o") || !strcmp(name, "ERK_Euler")) {
    t->nstage = 1;
    t->tend_buf_size = 1;
    t->b_ex[0] = 1.0;
    t->tend_buf_indmap[0] = 0;
  } else if (!strcmp(name, "IMEX_ARK232")) {
    double alp = (3.0 + 2.0 * sqrt(2.0)) / 6.0;
    double gam = 1.0 - 1.0 / sqrt(2.0);
    double del = 1.0 / (2.0 * sqrt(2.0));
    t->nstage = 3;
    t->tend_buf_size = 3;
    t->imex = 1;
    A_EX(2, 1) = 2.0 * gam;
    A_EX(3, 1) = 1.0 - alp;
    A_EX(3, 2) = alp;
    t->b_ex[0] = del;
    t->b_ex[1] = del;
    t->b_ex[2] = gam;
    A_IM(2, 1) = gam;
    A_IM(2, 2) = gam;
    A_IM(3, 1) = del;
    A_IM(3, 2) = del;
    A_IM(3, 3) = gam;
    for (int s = 0; s < 3; ++s) t->b_im[s] = t->b_ex[s];
    for (int s = 0; s < 3; ++s) t->tend_buf_indmap[s] = s;
  } else if (!strcmp(name, "IMEX_ARK324")) {
    t->nstage = 4;
    t->tend_buf_size = 4;
    t->imex = 1;
    A_EX(2, 1) = 1767732205903.0 / 2027836641118.0;
    A_EX(3, 1) = 5535828885825.0 / 10492691773637.0;
    A_EX(3, 2) = 788022342437.0 / 10882634858940.0;
    A_EX(4, 1) = 6485989280629.0 / 16251701735622.0;
    A_EX(4, 2) = -4246266847089.0 / 9704473918619.0;
    A_EX(4, 3) = 10755448449292.0 / 10357097424841.0;
    t->b_ex[0] = 1471266399579.0 / 7840856788654.0;
    t->b_ex[1] = -4482444167858.0 / 7529755066697.0;
    t->b_ex[2] = 11266239266428.0 / 11593286722821.0;
    t->b_ex[3] = 1767732205903.0 / 4055673282236.0;
    A_IM(2, 1) = 1767732205903.0 / 4055673282236.0;
    A_IM(2, 2) = 1767732205903.0 / 4055673282236.0;
    A_IM(3, 1) = 2746238789719.0 / 10658868560708.0;
    A_IM(3, 2) = -640167445237.0 / 6845629431997.0;
    A_IM(3, 3) = 1767732205903.0 / 4055673282236.0;
    for (int s = 0; s < 4; ++s) A_IM(4, s + 1) = t->b_ex[s];
    for (int s = 0; s < 4; ++s) t->b_im[s] = t->b_ex[s];
    for (int s = 0; s < 4; ++s) t->tend_buf_indmap[s] = s;
  } else {
    return 1;
  }
#undef A_EX
#undef A_IM
  return 0;
}

/* CMN/scale_timeint_rk.F90:2201-2355 (rk_advance_general2D), 1-based stage */
void oracle_rk_advance(const oracle_rk_tab *t, int stage, long n, double *q, double *var0,
                       double *varTmp, const double *tend_ex, const double *tend_im, long bs) {
  const int ind = t->tend_buf_indmap[stage - 1];
  if (t->nstage == 1)
    for (long i = 0; i < n; ++i) varTmp[i] = q[i];
  if (stage == t->nstage) {
    double cbe = t->b_ex[stage - 1] * t->dt;
    if (t->imex) {
      double cbi = t->b_im[stage - 1] * t->dt;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) q[i] = varTmp[i] + cbe * tend_ex[ind * bs + i] + cbi * tend_im[ind * bs + i];
    } else {
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) q[i] = varTmp[i] + cbe * tend_ex[ind * bs + i];
    }
    return;
  }
  if (stage == 1 && !t->imex)
    for (long i = 0; i < n; ++i) {
      var0[i] = q[i];
      varTmp[i] = q[i];
    }
  double cbe = t->b_ex[stage - 1] * t->dt;
  if (t->imex) {
    double cbi = t->b_im[stage - 1] * t->dt;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      q[i] = var0[i];
      varTmp[i] = varTmp[i] + cbe * tend_ex[ind * bs + i] + cbi * tend_im[ind * bs + i];
    }
  } else {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      q[i] = var0[i];
      varTmp[i] = varTmp[i] + cbe * tend_ex[ind * bs + i];
    }
  }
  if (t->tend_buf_size == 1 && !t->imex) {
    double ca = t->dt * t->a_ex[stage * 4 + (stage - 1)];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) q[i] = var0[i] + ca * tend_ex[i];
  } else if (!t->imex) {
    for (int s = 1; s <= stage; ++s) {
      double ca = t->dt * t->a_ex[stage * 4 + (s - 1)];
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) q[i] = q[i] + ca * tend_ex[(s - 1) * bs + i];
    }
  } else {
    for (int s = 1; s <= stage; ++s) {
      double ca = t->dt * t->a_ex[stage * 4 + (s - 1)];
      double ci = t->dt * t->a_im[stage * 4 + (s - 1)];
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) q[i] = q[i] + ca * tend_ex[(s - 1) * bs + i] + ci * tend_im[(s - 1) * bs + i];
    }
  }
}

/* CMN/scale_timeint_rk.F90:2510-2560 (rk_storeimpl_general2D) */
void oracle_rk_store_impl(const oracle_rk_tab *t, int stage, long n, double *q, double *var0,
                          double *varTmp, const double *tend_im, long bs) {
  if (!t->imex) return;
  const int ind = t->tend_buf_indmap[stage - 1];
  const double ca = t->dt * t->a_im[(stage - 1) * 4 + (stage - 1)];
  if (stage == 1)
    for (long i = 0; i < n; ++i) {
      var0[i] = q[i];
      varTmp[i] = q[i];
    }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) q[i] = q[i] + ca * tend_im[ind * bs + i];
}

/* ---------------------------------------------------------------------- */
/* FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:614-963, HEVE branch */
/* Tracer coupling of the driver (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:899-915):
 * when set, oracle_heve_update calls save_massflux on every stage state. */
static double *g_mflx[3], *g_alph[2];
void oracle_save_massflux_hook(const oracle_mesh *m, double *const q[5], const double *DPRES, const double *DENS_hyd,
                               const double *PRES_hyd, int stage, double twh, double twv, int is_hevi) {
  if (g_mflx[0])
    oracle_trcadvect_save_massflux(m, g_mflx[0], g_mflx[1], g_mflx[2], g_alph[0], g_alph[1], q[DENS_VID],
                                   q[MOMX_VID], q[MOMY_VID], q[MOMZ_VID], DPRES, DENS_hyd, PRES_hyd, stage, twh, twv,
                                   is_hevi);
}
void oracle_heve_set_massflux(double *MFLX_x, double *MFLX_y, double *MFLX_z, double *alph_M, double *alph_P) {
  g_mflx[0] = MFLX_x;
  g_mflx[1] = MFLX_y;
  g_mflx[2] = MFLX_z;
  g_alph[0] = alph_M;
  g_alph[1] = alph_P;
}

/* FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:1098-1178 (add_phy_tend), the
 * ENTOT_CONSERVE_SCHEME_FLAG = .false. branch; the OpenACC (:1157-1163) and
 * CPU (:1166-1172) forms compute the same expression per node */
void oracle_add_phy_tend(const oracle_mesh *m, double *const T[5], const double *const tp[6], const double *PRES,
                         const double *Rtot, const double *CPtot) {
  const double rP0 = 1.0 / m->PRES00;
  const long n = (long)m->Np * m->Ne;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    T[DENS_VID][i] = T[DENS_VID][i] + tp[DENS_VID][i];
    T[MOMZ_VID][i] = T[MOMZ_VID][i] + tp[MOMZ_VID][i];
    T[MOMX_VID][i] = T[MOMX_VID][i] + tp[MOMX_VID][i];
    T[MOMY_VID][i] = T[MOMY_VID][i] + tp[MOMY_VID][i];
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    const double EXNER = opow(PRES[i] * rP0, Rtot[i] / CPtot[i]);
    T[RHOT_VID][i] = T[RHOT_VID][i] + tp[RHOT_VID][i] + tp[5][i] / (CPtot[i] * EXNER);
  }
}
static const double *g_tp[6];
void oracle_set_phys_tend(const double *tp0, const double *tp1, const double *tp2, const double *tp3,
                          const double *tp4, const double *tp5) {
  g_tp[0] = tp0;
  g_tp[1] = tp1;
  g_tp[2] = tp2;
  g_tp[3] = tp3;
  g_tp[4] = tp4;
  g_tp[5] = tp5;
}
void oracle_phys_tend_hook(const oracle_mesh *m, double *const T[5], const double *PRES, const double *Rtot,
                           const double *CPtot) {
  if (g_tp[0]) oracle_add_phy_tend(m, T, g_tp, PRES, Rtot, CPtot);
}

void oracle_heve_update(const oracle_mesh *m, const oracle_rk_tab *t, int nsteps,
                        double *const q[5], double *PRES, double *DPRES,
                        const double *DENS_hyd, const double *PRES_hyd, const double *PRES_hyd_ref,
                        const double *THERM_hyd, const double *CORIOLIS,
                        const double *Rtot, const double *CVtot, const double *CPtot,
                        const double *DPhydDx, const double *DPhydDy, double *work) {
  const long NpNeA = (long)m->Np * m->NeA;
  const long n = (long)m->Np * m->Ne; /* nodes NeS..NeE */
  double *var0 = work;
  double *varTmp = work + 5 * NpNeA;
  double *tend = work + 10 * NpNeA; /* tend_buf2D_ex(Np,NeA,5,tend_buf_size) */
  for (int step = 0; step < nsteps; ++step) {
    for (int s = 1; s <= t->nstage; ++s) {
      for (int v = 0; v < 5; ++v) oracle_halo_exchange(m, q[v]);
      oracle_drhot2pres(m, PRES, DPRES, q[RHOT_VID], PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot);
      oracle_halo_exchange(m, DPRES);
      oracle_apply_bc(m, q[DENS_VID], q[MOMX_VID], q[MOMY_VID], q[MOMZ_VID], q[RHOT_VID]);
      const int ind = t->tend_buf_indmap[s - 1];
      double *T = tend + (long)ind * 5 * NpNeA;
      oracle_heve_cal_tend(m, T + DENS_VID * NpNeA, T + MOMX_VID * NpNeA, T + MOMY_VID * NpNeA,
                           T + MOMZ_VID * NpNeA, T + RHOT_VID * NpNeA, q[DENS_VID], q[MOMX_VID], q[MOMY_VID],
                           q[MOMZ_VID], q[RHOT_VID], DPRES, DENS_hyd, PRES_hyd, PRES_hyd_ref, THERM_hyd,
                           CORIOLIS, Rtot, CVtot, CPtot, DPhydDx, DPhydDy);
      {
        double *Tv[5];
        for (int v = 0; v < 5; ++v) Tv[v] = T + v * NpNeA;
        oracle_phys_tend_hook(m, Tv, PRES, Rtot, CPtot);
      }
      oracle_save_massflux_hook(m, q, DPRES, DENS_hyd, PRES_hyd, s, t->b_ex[s - 1], t->b_ex[s - 1], 0);
      for (int v = 0; v < 5; ++v)
        oracle_rk_advance(t, s, n, q[v], var0 + v * NpNeA, varTmp + v * NpNeA, tend + v * NpNeA, NULL,
                          5 * NpNeA);
    }
    if (m->MF_h)
      oracle_modalfilter_apply(m, q[DENS_VID], q[MOMX_VID], q[MOMY_VID], q[MOMZ_VID], q[RHOT_VID], 1);
    oracle_drhot2pres(m, PRES, DPRES, q[RHOT_VID], PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot);
  }
}

/* element_operation_kernel_matvec_ModalFilter (x: Mat_h1D, y: Mat_h1D_tr, z:
 * Mat_v1D_tr), each sum taken over l = 1..n1 in order */
static void modalfilter_elem(int N, const double *Fh, const double *Fv, const double *in, double *work,
                             double *out) {
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double s = Fh[i] * in[N * (j + N * k)];
        for (int l = 1; l < N; ++l) s = s + Fh[i + N * l] * in[l + N * (j + N * k)];
        out[i + N * (j + N * k)] = s;
      }
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double s = out[i + N * N * k] * Fh[j];
        for (int l = 1; l < N; ++l) s = s + out[i + N * (l + N * k)] * Fh[j + N * l];
        work[i + N * (j + N * k)] = s;
      }
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double s = work[i + N * j] * Fv[k];
        for (int l = 1; l < N; ++l) s = s + work[i + N * (j + N * l)] * Fv[k + N * l];
        out[i + N * (j + N * k)] = s;
      }
}

/* FDS/scale_atm_dyn_dgm_modalfilter.F90:232-270 (atm_dyn_dgm_tracer_modalfilter_apply):
 * weight = Gsqrt (DENS_hyd + DDENS), QTRC = filter(weight QTRC) / weight with
 * the tracer's matrices MF_h, MF_v ((n1,n1) column-major) */
void oracle_tracer_modalfilter_apply(const oracle_mesh *m, double *QTRC, const double *DENS_hyd, const double *DDENS,
                                     const double *MF_h, const double *MF_v) {
  const int N = m->n1, Np = m->Np;
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * Np * 4);
    double *work = tmp + Np, *out = tmp + 2 * Np, *weight = tmp + 3 * Np;
#pragma omp for schedule(static)
    for (int ke = 0; ke < m->Ne; ++ke) {
      const long o = (long)Np * ke;
      for (int kk = 0; kk < Np; ++kk) {
        weight[kk] = m->Gsqrt[o + kk] * (DENS_hyd[o + kk] + DDENS[o + kk]);
        tmp[kk] = weight[kk] * QTRC[o + kk];
      }
      modalfilter_elem(N, MF_h, MF_v, tmp, work, out);
      for (int kk = 0; kk < Np; ++kk) QTRC[o + kk] = out[kk] / weight[kk];
    }
    free(tmp);
  }
}

void oracle_modalfilter_apply(const oracle_mesh *m, double *DDENS, double *MOMX, double *MOMY, double *MOMZ,
                              double *DRHOT, int do_weight_Gsqrt) {
  const int N = m->n1, Np = m->Np;
  double *f[5] = {DDENS, MOMX, MOMY, MOMZ, DRHOT};
#pragma omp parallel
  {
    double *tmp = (double *)malloc(sizeof(double) * Np * 3);
    double *work = tmp + Np, *out = tmp + 2 * Np;
#pragma omp for schedule(static)
    for (int ke = 0; ke < m->Ne; ++ke) {
      const double *G = m->Gsqrt + (long)Np * ke;
      for (int v = 0; v < 5; ++v) {
        double *x = f[v] + (long)Np * ke;
        for (int kk = 0; kk < Np; ++kk) tmp[kk] = do_weight_Gsqrt ? G[kk] * x[kk] : x[kk];
        modalfilter_elem(N, m->MF_h, m->MF_v, tmp, work, out);
        for (int kk = 0; kk < Np; ++kk) {
          if (do_weight_Gsqrt) {
            const double RGsqrt = 1.0 / G[kk];
            x[kk] = out[kk] * RGsqrt;
          } else {
            x[kk] = out[kk];
          }
        }
      }
    }
    free(tmp);
  }
}
