// Device helpers shared by the HEVE and HEVI kernels: variable ids, the
// face-flux formulas, boundary reflection, pressure, face-node maps and the
// 32-bit-offset global loads and stores.
#pragma once
#include "fedg_internal.h"
#include "fedg_pow.h"

namespace fedg {
namespace {

enum { VD = 0, VT = 1, VW = 2, VU = 3, VV = 4 };  // DDENS, DRHOT, MOMZ, MOMX, MOMY

// Values of one side of a face node (reference: ..._heve_numflux.F90:1028-1063).
struct Side {
  double Gsqrt, G13, G23;
  double DD, MX, MY, MZ, DT, DP;
  double Dhyd, Phyd, Thyd;
};

// Rusanov flux jump for the five variables at one face node, pre-scaled by
// 0.5*Fscale (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:1024-1135).
// CS2: the boundary-contribution form of the same flux
// (..._heve_numflux.F90:242-496), which groups the sound speed as
// sqrt(Gnn * Cs2) with Cs2 = gamm (P_hyd + DPRES) Gsqrt / GsqrtDens (:351-352,
// :408-409, :425-431) instead of sqrt(Gnn gamm (P_hyd + DPRES) Gsqrt / GsqrtDens)
// (:1071-1080); every other operation is the same.
template <bool CS2 = false>
__device__ __forceinline__ void rusanov(const Side& M, const Side& P, double nx, double ny, double nz,
                                        double Fscale, double gamm, double df[5]) {
  const double RGsqrtV_M = 1.0 / M.Gsqrt, RGsqrtV_P = 1.0 / P.Gsqrt;
  const double GsqrtDDENS_M = M.Gsqrt * M.DD, GsqrtDDENS_P = P.Gsqrt * P.DD;
  const double GsqrtMOMX_M = M.Gsqrt * M.MX, GsqrtMOMX_P = P.Gsqrt * P.MX;
  const double GsqrtMOMY_M = M.Gsqrt * M.MY, GsqrtMOMY_P = P.Gsqrt * P.MY;
  const double GsqrtMOMZ_M = M.Gsqrt * M.MZ, GsqrtMOMZ_P = P.Gsqrt * P.MZ;
  const double GsqrtDRHOT_M = M.Gsqrt * M.DT, GsqrtDRHOT_P = P.Gsqrt * P.DT;
  const double GsqrtDens_M = GsqrtDDENS_M + M.Gsqrt * M.Dhyd;
  const double GsqrtDens_P = GsqrtDDENS_P + P.Gsqrt * P.Dhyd;
  const double GsqrtRhot_M = M.Gsqrt * M.Thyd + GsqrtDRHOT_M;
  const double GsqrtRhot_P = P.Gsqrt * P.Thyd + GsqrtDRHOT_P;
  const double Vel_M = (GsqrtMOMX_M * nx + GsqrtMOMY_M * ny +
                        ((GsqrtMOMZ_M * RGsqrtV_M + M.G13 * GsqrtMOMX_M + M.G23 * GsqrtMOMY_M) * nz)) / GsqrtDens_M;
  const double Vel_P = (GsqrtMOMX_P * nx + GsqrtMOMY_P * ny +
                        ((GsqrtMOMZ_P * RGsqrtV_P + P.G13 * GsqrtMOMX_P + P.G23 * GsqrtMOMY_P) * nz)) / GsqrtDens_P;
  double tmp1 = fabs(nx) + fabs(ny);
  const double Gnn_M = tmp1 + (1.0 * (RGsqrtV_M * RGsqrtV_M) + M.G13 * M.G13 + M.G23 * M.G23) * fabs(nz);
  const double Gnn_P = tmp1 + (1.0 * (RGsqrtV_P * RGsqrtV_P) + P.G13 * P.G13 + P.G23 * P.G23) * fabs(nz);
  double aM, aP;
  if constexpr (CS2) {
    const double Cs2_M = gamm * (M.Phyd + M.DP) * M.Gsqrt / GsqrtDens_M;
    const double Cs2_P = gamm * (P.Phyd + P.DP) * P.Gsqrt / GsqrtDens_P;
    aM = sqrt(Gnn_M * Cs2_M) + fabs(Vel_M);
    aP = sqrt(Gnn_P * Cs2_P) + fabs(Vel_P);
  } else {
    aM = sqrt(Gnn_M * gamm * (M.Phyd + M.DP) * M.Gsqrt / GsqrtDens_M) + fabs(Vel_M);
    aP = sqrt(Gnn_P * gamm * (P.Phyd + P.DP) * P.Gsqrt / GsqrtDens_P) + fabs(Vel_P);
  }
  const double alpha = aM > aP ? aM : aP;

  tmp1 = Fscale * 0.5;
  double tmp2 = -alpha * (GsqrtDDENS_P - GsqrtDDENS_M);
  df[VD] = tmp1 * (GsqrtDens_P * Vel_P - GsqrtDens_M * Vel_M + tmp2);
  tmp2 = -alpha * (GsqrtDRHOT_P - GsqrtDRHOT_M);
  df[VT] = tmp1 * (GsqrtRhot_P * Vel_P - GsqrtRhot_M * Vel_M + tmp2);
  const double tmp3 = P.Gsqrt * P.DP;
  const double tmp4 = M.Gsqrt * M.DP;
  const double mom1 = (tmp3 * RGsqrtV_P - tmp4 * RGsqrtV_M) * nz;
  const double mom2 = (nx + P.G13 * nz) * tmp3 - (nx + M.G13 * nz) * tmp4;
  const double mom3 = (ny + P.G23 * nz) * tmp3 - (ny + M.G23 * nz) * tmp4;
  tmp2 = -alpha * (GsqrtMOMZ_P - GsqrtMOMZ_M);
  df[VW] = tmp1 * (GsqrtMOMZ_P * Vel_P - GsqrtMOMZ_M * Vel_M + mom1 + tmp2);
  tmp2 = -alpha * (GsqrtMOMX_P - GsqrtMOMX_M);
  df[VU] = tmp1 * (GsqrtMOMX_P * Vel_P - GsqrtMOMX_M * Vel_M + mom2 + tmp2);
  tmp2 = -alpha * (GsqrtMOMY_P - GsqrtMOMY_M);
  df[VV] = tmp1 * (GsqrtMOMY_P * Vel_P - GsqrtMOMY_M * Vel_M + mom3 + tmp2);
}

// HEVI explicit face flux (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi_numflux.F90:232-416):
// dissipation only across lateral faces (swV = 1 - nz^2), mass and rhotheta
// advected by the horizontal velocity only, no vertical pressure jump.
__device__ __forceinline__ void rusanov_hevi(const Side& M, const Side& P, double nx, double ny, double nz,
                                             double Fscale, double gamm, double df[5]) {
  const double RGsqrtV_M = 1.0 / M.Gsqrt, RGsqrtV_P = 1.0 / P.Gsqrt;
  const double GDD_M = M.Gsqrt * M.DD, GDD_P = P.Gsqrt * P.DD;
  const double GMX_M = M.Gsqrt * M.MX, GMX_P = P.Gsqrt * P.MX;
  const double GMY_M = M.Gsqrt * M.MY, GMY_P = P.Gsqrt * P.MY;
  const double GMZ_M = M.Gsqrt * M.MZ, GMZ_P = P.Gsqrt * P.MZ;
  const double GDT_M = M.Gsqrt * M.DT, GDT_P = P.Gsqrt * P.DT;
  const double GDens_M = GDD_M + M.Gsqrt * M.Dhyd;
  const double GDens_P = GDD_P + P.Gsqrt * P.Dhyd;
  const double GRhot_M = M.Gsqrt * M.Thyd + GDT_M;
  const double GRhot_P = P.Gsqrt * P.Thyd + GDT_P;
  const double Velh_M = (GMX_M * nx + GMY_M * ny) / GDens_M;
  const double Velh_P = (GMX_P * nx + GMY_P * ny) / GDens_P;
  const double Vel_M = Velh_M + (((GMZ_M * RGsqrtV_M + M.G13 * GMX_M + M.G23 * GMY_M) * nz)) / GDens_M;
  const double Vel_P = Velh_P + (((GMZ_P * RGsqrtV_P + P.G13 * GMX_P + P.G23 * GMY_P) * nz)) / GDens_P;
  const double swV = 1.0 - nz * nz;
  const double aM = sqrt(gamm * (M.Phyd + M.DP) * M.Gsqrt / GDens_M) + fabs(Vel_M);
  const double aP = sqrt(gamm * (P.Phyd + P.DP) * P.Gsqrt / GDens_P) + fabs(Vel_P);
  const double alpha = swV * (aM > aP ? aM : aP);
  const double tmp1 = Fscale * 0.5;
  double tmp2 = -alpha * (GDD_P - GDD_M);
  df[VD] = tmp1 * (GDens_P * Velh_P - GDens_M * Velh_M + tmp2);
  tmp2 = -alpha * (GDT_P - GDT_M);
  df[VT] = tmp1 * (GRhot_P * Velh_P - GRhot_M * Velh_M + tmp2);
  tmp2 = -alpha * (GMZ_P - GMZ_M);
  df[VW] = tmp1 * (GMZ_P * Vel_P - GMZ_M * Vel_M + tmp2);
  const double tmp3 = P.Gsqrt * P.DP;
  const double tmp4 = M.Gsqrt * M.DP;
  const double mx = (nx + P.G13 * nz) * tmp3 - (nx + M.G13 * nz) * tmp4;
  const double my = (ny + P.G23 * nz) * tmp3 - (ny + M.G23 * nz) * tmp4;
  tmp2 = -alpha * (GMX_P - GMX_M);
  df[VU] = tmp1 * (GMX_P * Vel_P - GMX_M * Vel_M + mx + tmp2);
  tmp2 = -alpha * (GMY_P - GMY_M);
  df[VV] = tmp1 * (GMY_P * Vel_P - GMY_M * Vel_M + my + tmp2);
}

// Exterior momentum of a closed boundary (FDS/scale_atm_dyn_dgm_bnd.F90:335-359),
// from the interior values of the same node.
__device__ __forceinline__ void reflect_momentum(int bc, double GsqrtM, double GsqrtH, double G11_, double G12_,
                                                 double G22_, double G13, double G23, double nx, double ny,
                                                 double nz, double MX, double MY, double MZ, double& MXo,
                                                 double& MYo, double& MZo) {
  if (bc == FEDG_BND_SLIP) {
    const double GsqrtV = GsqrtM / GsqrtH;
    const double MOMW = MZ / GsqrtV + G13 * MX + G23 * MY;
    const double t1 = GsqrtV * G13, t2 = GsqrtV * G23;
    const double fac = nz * (GsqrtV * GsqrtV) /
                       (1.0 + G11_ * (t1 * t1) + 2.0 * G12_ * ((GsqrtV * GsqrtV) * G13 * G23) + G22_ * (t2 * t2));
    const double mn = MX * nx + MY * ny + MOMW * nz;
    MXo = MX - 2.0 * mn * (nx + fac * (G11_ * G13 + G12_ * G23));
    MYo = MY - 2.0 * mn * (ny + fac * (G12_ * G13 + G22_ * G23));
    MZo = MZ - 2.0 * mn * fac / GsqrtV;
  } else if (bc == FEDG_BND_NOSLIP) {
    MXo = -MX;
    MYo = -MY;
    MZo = -MZ;
  }
}

// Pressure perturbation of one node (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:470-474).
__device__ __forceinline__ double dpres_of(double DRHOT, double THERM_hyd, double PRES_hyd, double Rtot,
                                           double CVtot, double CPtot, const Consts& c) {
  const double rP0 = 1.0 / c.PRES00;
  const double RHOT = THERM_hyd + DRHOT;
  const double PRES = c.PRES00 * fedg_pow(Rtot * rP0 * RHOT, CPtot / CVtot);
  return PRES - PRES_hyd;
}
// PRES of a node as DRHOT2PRES forms it (the PRES of dpres_of)
__device__ __forceinline__ double pres_of(double DRHOT, double THERM_hyd, double Rtot, double CVtot, double CPtot,
                                          const Consts& c) {
  const double rP0 = 1.0 / c.PRES00;
  const double RHOT = THERM_hyd + DRHOT;
  return c.PRES00 * fedg_pow(Rtot * rP0 * RHOT, CPtot / CVtot);
}
// The THERM line of add_phy_tend (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:
// 1143-1175, ENTOT_CONSERVE_SCHEME_FLAG off): T + RHOT_tp + RHOH_p / (CPtot
// EXNER), EXNER = (PRES / PRES00)^(Rtot / CPtot)
__device__ __forceinline__ double phys_add_rhot(double T, double RHOT_tp, double RHOH_p, double PRES, double Rtot,
                                                double CPtot, const Consts& c) {
  const double rP0 = 1.0 / c.PRES00;
  const double EXNER = fedg_pow(PRES * rP0, Rtot / CPtot);
  return T + RHOT_tp + RHOH_p / (CPtot * EXNER);
}
// The same with the pow tables at `tab` (a copy in LDS): the same bits.
__device__ __forceinline__ double dpres_of_tab(double DRHOT, double THERM_hyd, double PRES_hyd, double Rtot,
                                               double CVtot, double CPtot, const Consts& c, const double* tab) {
  const double rP0 = 1.0 / c.PRES00;
  const double RHOT = THERM_hyd + DRHOT;
  const double PRES = c.PRES00 * fedg_pow_tab(Rtot * rP0 * RHOT, CPtot / CVtot, tab);
  return PRES - PRES_hyd;
}
// Copy the 5 x 128 pow tables to `tab` (LDS) with the threads of a workgroup
// of `nthr` threads; the caller synchronises before the first read.
__device__ __forceinline__ void pow_tables_to(double* tab, int t, int nthr) {
  for (int q = t; q < 640; q += nthr) {
    const int e = q & 127;
    double v;
    switch (q >> 7) {
      case 0: v = fedg_pow_invc[e]; break;
      case 1: v = fedg_pow_logc_hi[e]; break;
      case 2: v = fedg_pow_logc_lo[e]; break;
      case 3: v = fedg_pow_exp_hi[e]; break;
      default: v = fedg_pow_exp_lo[e]; break;
    }
    tab[q] = v;
  }
}

// Element-local node of face node a on face f (faces in the reference order
// y-, x+, y+, x-, z-, z+; ELM/scale_element_hexahedral.F90:173-180).
template <int N1>
__device__ __forceinline__ int face_node(int f, int a) {
  constexpr int N = N1 - 1, n2 = N1 * N1;
  const int a0 = a % N1, a1 = a / N1;
  switch (f) {
    case 0: return a0 + n2 * a1;
    case 1: return N + N1 * a0 + n2 * a1;
    case 2: return a0 + N1 * N + n2 * a1;
    case 3: return N1 * a0 + n2 * a1;
    case 4: return a;
    default: return a + n2 * N;
  }
}

// Spread the workgroups of one launch so that XCD x (workgroups x, x+8, ...
// under round-robin dispatch) works on one contiguous 1/8 of the elements:
// neighbouring elements then share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Loads and stores with a 32-bit byte offset from a base pointer held in
// SGPRs (fields are below 4 GB; fedg_ctx_create checks it), so that the
// compiler uses the saddr + 32-bit vgpr offset form and keeps no 64-bit
// addresses in VGPRs.
__device__ __forceinline__ double ldg(const double* base, uint32_t idx) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + (idx << 3));
}
__device__ __forceinline__ void stg(double* base, uint32_t idx, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + (idx << 3)) = v;
}
// The same with the non-temporal hint, for data a kernel streams once (read
// once, or written for the next launch): it does not displace the lines that
// other workgroups are about to read again (the face values of neighbours).
__device__ __forceinline__ double ldg_nt(const double* base, uint32_t idx) {
  return __builtin_nontemporal_load(reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + (idx << 3)));
}
__device__ __forceinline__ void stg_nt(double* base, uint32_t idx, double v) {
  __builtin_nontemporal_store(v, reinterpret_cast<double*>(reinterpret_cast<char*>(base) + (idx << 3)));
}

}  // namespace
}  // namespace fedg
