// CDNA4 (gfx950) kernels for the vertically implicit part of FElib's HEVI
// scheme (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965 and
// FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi_common_2.F90:111-1328).
//
// One Newton iteration is four launches:
//   1. hevi_rhs_uv   (element, one thread per node): the damping coefficient
//      alph at every vertical face node from q^n, the horizontal-momentum
//      face jumps of the iterate, their lift, MOMX_t / MOMY_t and the uv rhs;
//   2. hevi_solve_uv (64/8 lines per wavefront, a lane per block row):
//      assembles the N x N blocks of construct_matbnd_uv level by level,
//      block-Thomas elimination with the pivoted LU of solve_NnodeN_uv, back
//      substitution, MOMX, MOMY += increment;
//   3. hevi_rhs3     (element): vertical fluxes of DDENS, MOMZ, DRHOT, their
//      face jumps, Dz and lift, VFilterPM1 buoyancy, the var3 rhs and the
//      linearisation coefficients POT, WT, DPDRHOT;
//   4. hevi_solve3   (two lines per wavefront at p = 7, a lane per block
//      row): the 3N x 3N blocks of construct_matbnd, block-Thomas with the LU
//      of solve_NnodeN_var3, DDENS, MOMZ, DRHOT += increment.
// A line is the column of one horizontal node (ij) of one horizontal element
// through all NeZ elements; everything the implicit operator couples lies on
// it.  A block's rows stay in registers of their lanes; pivoting exchanges
// row positions, the pivot row and the solved entries travel by cross-lane
// shuffles, and the levels' solutions wait in LDS for the block back
// substitution.  Every entry sees the same operations in the same order as
// the reference's serial loops, so results are bit for bit those of the CPU
// restatement.
#include <cfloat>
#include <cmath>
#include <mutex>

#include "fedg_device.h"
#include "fedg_hevi.h"
#include "fedg_mlp.h"


namespace fedg {
namespace {

struct HeviArgs {
  DevMesh m;
  Consts c;
  const double* qs[5];     // stage state q* (VID order), (Np,NeA)
  const double* q0[5];     // var0 = q^n
  double* pv[5];           // Newton iterate, (Np,Ne), written by the solves
  const double* pin[5];    // the iterate as read: pv, or q^n for a variable no solve has updated yet
  const double* DENS_hyd;
  const double* PRES_hyd;
  const double* RHOT_hyd;  // PRES00/Rdry (PRES_hyd/PRES00)^(CVdry/CPdry) per node, or null: formed in the kernels
  int alph_ready;          // alph holds the damping coefficients of q0 (a residual kernel formed them)
  const double* Rtot;
  const double* CVtot;
  const double* CPtot;
  double* alph;            // (2*Nfp, Ne): z- then z+ face nodes
  double* b3;              // (3, Np*Ne): DENS, MOMZ, RHOT rhs
  double* buv;             // (2, Np*Ne)
  double* coef;            // (3, Np*Ne): POT, WT, DPDRHOT
  double* T[5];            // implicit-operator tendencies (VID order), written when want_T
  double impl_fac;
  int want_b;              // form the rhs (Newton) or only the operator (impl_fac = 0)
  int NeZ;
};

__device__ __forceinline__ double gsqrtv(const HeviArgs& a, int ke, int p, int N1) {
  const int ke2d = a.m.EMap3Dto2D[ke];
  const int ij = p % (N1 * N1);
  return ldg(a.m.Gsqrt, (uint32_t)a.m.Np * ke + p) / ldg(a.m.GsqrtH, (uint32_t)ij + (uint32_t)a.m.Nfp * ke2d);
}

// vertical neighbour node of the face node (bottom: f = 0, top: f = 1); self at the column ends
__device__ __forceinline__ uint32_t vz_nbr(const HeviArgs& a, int ke, int f, int ij, int N1) {
  const int Ne2D = a.m.Ne2D;
  const int kz = ke / Ne2D;
  const uint32_t Np = a.m.Np, Nfp = a.m.Nfp;
  if (f == 0) {
    if (kz > 0) return Np * (ke - Ne2D) + ij + Nfp * (N1 - 1);
    return Np * ke + ij;
  }
  if (kz < a.NeZ - 1) return Np * (ke + Ne2D) + ij;
  return Np * ke + ij + Nfp * (N1 - 1);
}

// ---------------------------------------------------------------------------
// 1. vi_cal_del_flux_dyn_uv (:1015-1158) + eval_Ax_uv (:482-586)
// FLAT (the flat-terrain uniform box the context detected): the metric
// arrays hold Gsqrt = GsqrtH = 1, G13 = G23 = 0, G11 = G22 = 1, G12 = 0,
// nz = -1 / +1 on the bottom / top faces, E33 and Fscale constant and a
// tensor-product Lift, so those values are taken from the context's
// constants instead of loaded; every operation is kept, so the results are
// the same bits.  The outputs (and, in hevi_rhs3_body, the coefficients) and
// the stage state q* are streamed with the non-temporal hint, so that the
// iterate lines the vertical neighbours gather stay cached.
template <int N1, bool FLAT>
__device__ __forceinline__ void hevi_rhs_uv_body(const HeviArgs& a) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1;
  __shared__ double sdf[2][2][Nfp];  // face (bottom, top), var (MOMX, MOMY), node
  __shared__ double sPow[640];        // the pow tables (the same values as the global ones)
  const int p = threadIdx.x;
  const int ke = blockIdx.x;
  const int ij = p % Nfp, k = p / Nfp;
  const DevMesh& m = a.m;
  const Consts& c = a.c;
  const uint32_t o = (uint32_t)Np * ke + p;
  const double gamm = c.CPdry / c.CVdry, rgamm = c.CVdry / c.CPdry;
  const double PRES0ovRdry = c.PRES00 / c.Rdry, rP0 = 1.0 / c.PRES00;
  if (!a.alph_ready) {
    pow_tables_to(sPow, p, Np);
    __syncthreads();
  }
  if (k == 0 || k == N1 - 1) {
    for (int f = 0; f < 2; ++f) {
      if ((f == 0) != (k == 0)) continue;
      const int fp = (4 + f) * Nfp + ij;  // reference face-node index
      const uint32_t gfp = (uint32_t)fp + (uint32_t)(6 * Nfp) * ke;
      const uint32_t iM = o, iP = vz_nbr(a, ke, f, ij, N1);
      const uint32_t ia = (uint32_t)(f * Nfp + ij) + (uint32_t)(2 * Nfp) * ke;
      double al;
      if (a.alph_ready) {
        al = ldg(a.alph, ia);
      } else {
        const double nz = FLAT ? (f == 0 ? -1.0 : 1.0) : ldg(m.nz, gfp);
        const int ke2d = FLAT ? 0 : m.EMap3Dto2D[ke];
        const double G11_ = FLAT ? 1.0 : ldg(m.G11, (uint32_t)ij + (uint32_t)Nfp * ke2d);
        const double G12_ = FLAT ? 0.0 : ldg(m.G12, (uint32_t)ij + (uint32_t)Nfp * ke2d);
        const double G22_ = FLAT ? 1.0 : ldg(m.G22, (uint32_t)ij + (uint32_t)Nfp * ke2d);
        double wt0[2], pres0[2], rdens0[2], Gnn[2];
        const uint32_t ix[2] = {iM, iP};
        for (int s = 0; s < 2; ++s) {
          const uint32_t n = ix[s];
          const int ken = n / Np;
          const double gam = 1.0;
          const double rgam2 = 1.0 / (gam * gam);
          const double RGsqrtV = 1.0 / (FLAT ? 1.0 : gsqrtv(a, ken, n - (uint32_t)Np * ken, N1));
          const double MOMX0 = ldg(a.q0[VU], n), MOMY0 = ldg(a.q0[VV], n), DRHOT0 = ldg(a.q0[VT], n);
          const double Rt = ldg(a.Rtot, n), CPovCV = ldg(a.CPtot, n) / ldg(a.CVtot, n);
          const double g13 = FLAT ? 0.0 : ldg(m.G13, n), g23 = FLAT ? 0.0 : ldg(m.G23, n);
          rdens0[s] = 1.0 / (ldg(a.DENS_hyd, n) + ldg(a.q0[VD], n));
          wt0[s] = (ldg(a.q0[VW], n) * RGsqrtV + g13 * MOMX0 + g23 * MOMY0) * rdens0[s];
          const double rhot_hyd = a.RHOT_hyd ? ldg(a.RHOT_hyd, n)
                                             : PRES0ovRdry * fedg_pow_tab(ldg(a.PRES_hyd, n) / c.PRES00, rgamm, sPow);
          pres0[s] = c.PRES00 * fedg_pow_tab(Rt * rP0 * (rhot_hyd + DRHOT0), CPovCV, sPow);
          const double Gxz = rgam2 * (G11_ * g13 + G12_ * g23);
          const double Gyz = rgam2 * (G12_ * g13 + G22_ * g23);
          Gnn[s] = (1.0 * (RGsqrtV * RGsqrtV) + g13 * Gxz + g23 * Gyz) * fabs(nz);
        }
        const double sM = fabs(wt0[0]) + sqrt(Gnn[0] * gamm * pres0[0] * rdens0[0]);
        const double sP = fabs(wt0[1]) + sqrt(Gnn[1] * gamm * pres0[1] * rdens0[1]);
        al = (nz * nz) * (sM > sP ? sM : sP);
        stg(a.alph, ia, al);
      }
      const double tmp1 = -0.5 * (FLAT ? m.fF[4 + f] : ldg(m.Fscale, gfp)) * al;
      sdf[f][0][ij] = tmp1 * (ldg(a.pin[VU], iP) - ldg(a.pin[VU], iM));
      sdf[f][1][ij] = tmp1 * (ldg(a.pin[VV], iP) - ldg(a.pin[VV], iM));
    }
  }
  __syncthreads();
  // Lift of the vertical-face jumps (the lateral faces carry none)
  const double L4 = FLAT ? m.liftT[k + N1 * 4] : ldg(m.Lift, (uint32_t)(p + 4 * Np));
  const double L5 = FLAT ? m.liftT[k + N1 * 5] : ldg(m.Lift, (uint32_t)(p + 5 * Np));
  const double LDx = L4 * sdf[0][0][ij] + L5 * sdf[1][0][ij];
  const double LDy = L4 * sdf[0][1][ij] + L5 * sdf[1][1][ij];
  const double RGsqrtV = 1.0 / (FLAT ? 1.0 : gsqrtv(a, ke, p, N1));
  const double Tx = -LDx * RGsqrtV, Ty = -LDy * RGsqrtV;
  if (a.want_b) {
    const uint32_t n = (uint32_t)Np * m.Ne;
    stg_nt(a.buv, o, a.impl_fac * Tx - ldg(a.pin[VU], o) + ldg_nt(a.qs[VU], o));
    stg_nt(a.buv, o + n, a.impl_fac * Ty - ldg(a.pin[VV], o) + ldg_nt(a.qs[VV], o));
  } else {
    stg_nt(a.T[VU], o, Tx);
    stg_nt(a.T[VV], o, Ty);
  }
}

// ---------------------------------------------------------------------------
// 3. vi_cal_del_flux_dyn (:1161-1328) + eval_Ax (:111-326)
template <int N1, bool FLAT>
__device__ __forceinline__ void hevi_rhs3_body(const HeviArgs& a) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1;
  __shared__ double sF[3][Np];      // DENS, MOMZ, RHOT vertical fluxes
  __shared__ double sRho[Np];
  __shared__ double sdf[2][3][Nfp];  // face (bottom, top), var (DENS, MOMZ, RHOT), node
  __shared__ double sD[N1 * N1];
  __shared__ double sI[N1 * N1];
  __shared__ double sPow[640];  // the pow tables (the same values as the global ones)
  const int p = threadIdx.x;
  const int ke = blockIdx.x;
  const int ij = p % Nfp, k = p / Nfp;
  const DevMesh& m = a.m;
  const Consts& c = a.c;
  const uint32_t o = (uint32_t)Np * ke + p;
  const int kz = ke / m.Ne2D;
  const double rgamm = c.CVdry / c.CPdry;
  const double PRES0ovRdry = c.PRES00 / c.Rdry, rP0 = 1.0 / c.PRES00, P0ovR = c.PRES00 / c.Rdry;
  if (p < N1 * N1) {
    sD[p] = m.D1D[p];
    sI[p] = m.Itr[p];
  }
  pow_tables_to(sPow, p, Np);
  __syncthreads();
  // vertical face jumps of the iterate
  if (k == 0 || k == N1 - 1) {
    for (int f = 0; f < 2; ++f) {
      if ((f == 0) != (k == 0)) continue;
      const int fp = (4 + f) * Nfp + ij;
      const uint32_t gfp = (uint32_t)fp + (uint32_t)(6 * Nfp) * ke;
      const uint32_t iM = o, iP = vz_nbr(a, ke, f, ij, N1);
      const uint32_t ix[2] = {iM, iP};
      double DD[2], MZ[2], MW[2], DT[2], pott[2], dpres[2], GV[2];
      for (int s = 0; s < 2; ++s) {
        const uint32_t n = ix[s];
        const int ken = n / Np;
        GV[s] = FLAT ? 1.0 : gsqrtv(a, ken, n - (uint32_t)Np * ken, N1);
        DD[s] = ldg(a.pin[VD], n);
        MZ[s] = ldg(a.pin[VW], n);
        DT[s] = ldg(a.pin[VT], n);
        const double dens = ldg(a.DENS_hyd, n) + DD[s];
        const double PH = ldg(a.PRES_hyd, n);
        const double rhot_hyd = a.RHOT_hyd ? ldg(a.RHOT_hyd, n) : PRES0ovRdry * fedg_pow_tab(PH / c.PRES00, rgamm, sPow);
        pott[s] = (rhot_hyd + DT[s]) / dens;
        dpres[s] = c.PRES00 * fedg_pow_tab(ldg(a.Rtot, n) * rP0 * dens * pott[s], ldg(a.CPtot, n) / ldg(a.CVtot, n), sPow);
        dpres[s] = dpres[s] - PH;
        const double g13 = FLAT ? 0.0 : ldg(m.G13, n), g23 = FLAT ? 0.0 : ldg(m.G23, n);
        MW[s] = MZ[s] + GV[s] * g13 * ldg(a.pin[VU], n) + GV[s] * g23 * ldg(a.pin[VV], n);
      }
      if ((kz == 0 || kz == a.NeZ - 1) && iM == iP) {
        const double g13 = FLAT ? 0.0 : ldg(m.G13, iM), g23 = FLAT ? 0.0 : ldg(m.G23, iM);
        MZ[1] = -MZ[0] - 2.0 * GV[0] * (g13 * ldg(a.pin[VU], iM) + g23 * ldg(a.pin[VV], iM));
        MW[1] = -MW[0];
      }
      const double nz = FLAT ? (f == 0 ? -1.0 : 1.0) : ldg(m.nz, gfp);
      const double al = ldg(a.alph, (uint32_t)(f * Nfp + ij) + (uint32_t)(2 * Nfp) * ke);
      const double tmp1 = 0.5 * (FLAT ? m.fF[4 + f] : ldg(m.Fscale, gfp));
      sdf[f][0][ij] = tmp1 * (+(MW[1] - MW[0]) * nz - al * (DD[1] - DD[0]));
      sdf[f][1][ij] = tmp1 * (+(dpres[1] - dpres[0]) * nz - al * (MZ[1] - MZ[0]));
      sdf[f][2][ij] = tmp1 * (+(pott[1] * MW[1] - pott[0] * MW[0]) * nz - al * (DT[1] - DT[0]));
    }
  }
  // vertical fluxes of the node
  const double GV = FLAT ? 1.0 : gsqrtv(a, ke, p, N1);
  const double PVD = ldg(a.pin[VD], o), PVW = ldg(a.pin[VW], o), PVT = ldg(a.pin[VT], o);
  const double G13o = FLAT ? 0.0 : ldg(m.G13, o), G23o = FLAT ? 0.0 : ldg(m.G23, o);
  const double FD = PVW + GV * G13o * ldg(a.pin[VU], o) + GV * G23o * ldg(a.pin[VV], o);
  const double PH = ldg(a.PRES_hyd, o), DH = ldg(a.DENS_hyd, o);
  const double RHOT = (a.RHOT_hyd ? ldg(a.RHOT_hyd, o) : P0ovR * fedg_pow_tab(PH / c.PRES00, rgamm, sPow)) + PVT;
  const double pt_ = RHOT / (PVD + DH);
  const double Rt = ldg(a.Rtot, o), CPt = ldg(a.CPtot, o), CVt = ldg(a.CVtot, o);
  const double FW = c.PRES00 * fedg_pow_tab(Rt * rP0 * RHOT, CPt / CVt, sPow) - PH;
  sF[0][p] = FD;
  sF[1][p] = FW;
  sF[2][p] = pt_ * FD;
  sRho[p] = PVD;
  __syncthreads();
  // Dz (matvec_dirZ), VFilterPM1 and the lift of the vertical-face jumps
  double dz[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    double s = sF[v][ij] * sD[k];
#pragma unroll
    for (int l = 1; l < N1; ++l) s = s + sF[v][ij + Nfp * l] * sD[k + N1 * l];
    dz[v] = s;
  }
  double drho = sRho[ij] * sI[N1 * k];
#pragma unroll
  for (int l = 1; l < N1; ++l) drho = drho + sRho[ij + Nfp * l] * sI[l + N1 * k];
  const double L4 = FLAT ? m.liftT[k + N1 * 4] : ldg(m.Lift, (uint32_t)(p + 4 * Np));
  const double L5 = FLAT ? m.liftT[k + N1 * 5] : ldg(m.Lift, (uint32_t)(p + 5 * Np));
  const double E33 = FLAT ? m.fE[2] : ldg(m.E33, o);
  const double RGsqrtV = 1.0 / GV;
  const double TD = -(E33 * dz[0] + (L4 * sdf[0][0][ij] + L5 * sdf[1][0][ij])) * RGsqrtV;
  const double TT = -(E33 * dz[2] + (L4 * sdf[0][2][ij] + L5 * sdf[1][2][ij])) * RGsqrtV;
  const double TW = -(E33 * dz[1] + (L4 * sdf[0][1][ij] + L5 * sdf[1][1][ij])) * RGsqrtV - c.GRAV * drho;
  if (a.want_b) {
    const uint32_t n = (uint32_t)Np * m.Ne;
    const double dens = DH + PVD;
    stg_nt(a.coef, o, RHOT / dens);
    stg_nt(a.coef, o + n, FD / dens);
    const double CPovCV = CPt / CVt;
    stg_nt(a.coef, o + 2 * n, CPovCV * c.PRES00 * fedg_pow_tab(Rt / c.PRES00 * RHOT, CPovCV, sPow) / RHOT);
    stg_nt(a.b3, o, a.impl_fac * TD - PVD + ldg_nt(a.qs[VD], o));
    stg_nt(a.b3, o + n, a.impl_fac * TW - PVW + ldg_nt(a.qs[VW], o));
    stg_nt(a.b3, o + 2 * n, a.impl_fac * TT - PVT + ldg_nt(a.qs[VT], o));
  } else {
    stg_nt(a.T[VD], o, TD);
    stg_nt(a.T[VW], o, TW);
    stg_nt(a.T[VT], o, TT);
  }
}

// ---------------------------------------------------------------------------
// Wavefront-cooperative pivoted LU + solve of solve_NnodeN_var3 / _uv
// (FDS/scale_atm_dyn_dgm_hevi_common_linalgebra.F90:2142-2445) for an n x n
// block (rows in the lanes' registers) and NR right-hand sides (solution to
// LDS).
// ---------------------------------------------------------------------------
// Line solves.  A wavefront holds 64/SEG lines, one per segment of SEG lanes;
// lane r of a segment owns row r of the level's n x n block in registers and
// the rows' right-hand-side entries.  Partial pivoting moves rows by
// exchanging their positions (pos), not their data: the operations every
// entry sees -- pivot choice, reciprocal of the pivot, scaling of the column,
// rank-1 update, forward substitution by columns, back substitution in
// ascending column order -- are those of the reference's serial loops
// (FDS/scale_atm_dyn_dgm_hevi_common_linalgebra.F90:2142-2445), so results
// are bit for bit the CPU restatement's (rows outside the update subtract an
// exact zero, which can only turn a -0 entry into +0).

// Cross-lane exchanges inside a segment: DPP permutations within 16-lane
// rows, ds_swizzle between the two rows of a 32-lane segment (every DPP
// control used here reads a valid lane, so no lane keeps an old value).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}

// maximum over the segment of a double that is never NaN
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  return __hiloint2double(dpp_i<CTRL>(__double2hiint(v)), dpp_i<CTRL>(__double2loint(v)));
}
// v_max_f64: the values reduced are |A(i,k)| (never -0), the sentinels -1
// and -2 and +inf, never NaN, so the maximum is the one of b > a ? b : a
// bit for bit, in one instruction instead of a compare and two selects
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
template <int SEG>
__device__ __forceinline__ double seg_max(double v) {
  v = dmax(v, dpp_d<0xB1>(v));
  v = dmax(v, dpp_d<0x4E>(v));
  v = dmax(v, dpp_d<0x141>(v));
  if constexpr (SEG >= 16) v = dmax(v, dpp_d<0x140>(v));
  if constexpr (SEG >= 32) {
    constexpr int X16 = (0x10 << 10) | 0x1F;
    v = dmax(v, __hiloint2double(__builtin_amdgcn_ds_swizzle(__double2hiint(v), X16),
                                 __builtin_amdgcn_ds_swizzle(__double2loint(v), X16)));
  }
  if constexpr (SEG >= 64) v = dmax(v, __shfl_xor(v, 32, 64));
  return v;
}
template <int SEG>
__device__ __forceinline__ int seg_min_i(int v) {
  v = min(v, dpp_i<0xB1>(v));
  v = min(v, dpp_i<0x4E>(v));
  v = min(v, dpp_i<0x141>(v));
  if constexpr (SEG >= 16) v = min(v, dpp_i<0x140>(v));
  if constexpr (SEG >= 32) v = min(v, __builtin_amdgcn_ds_swizzle(v, (0x10 << 10) | 0x1F));
  if constexpr (SEG >= 64) v = min(v, __shfl_xor(v, 32, 64));
  return v;
}
// lowest lane of the segment whose predicate is set (the predicate holds for
// at least one lane of every segment); *several: more than one lane holds it
template <int SEG>
__device__ __forceinline__ int seg_first(bool pred, bool* several) {
  const unsigned long long bal = __ballot(pred);
  const int sb = threadIdx.x & ~(SEG - 1);
  if constexpr (SEG == 64) {
    *several = __popcll(bal) > 1;
    return __builtin_ctzll(bal);
  } else {
    const unsigned s = (unsigned)(bal >> sb) & (unsigned)((1ull << SEG) - 1);
    *several = __popc(s) > 1;
    return sb + __builtin_ctz(s);
  }
}

// LU with partial pivoting of the segment's n x n block (row r in A of lane
// r), applied to NR right-hand sides as it goes (x[q] of lane r: entry r of
// rhs q; the row operations on them are the forward substitution), then back
// substitution.  Rows are never moved: a row's position (pos) changes when
// it is swapped with the pivot row, and the pivot row and the entries solved
// travel by cross-lane permutes or LDS.  Every lane of the single-wavefront
// workgroup calls it.
//  - the pivot search reduces |A(i,k)| alone and finds the row by a ballot;
//    only when several rows hold the maximum does a second reduction pick
//    the first position among them;
//  - every lane forms 1 / A(i,k) before the search, and the pivot lane's
//    reciprocal is the one broadcast;
//  - all NR right-hand sides are carried (unused ones hold zeros);
//  - after the elimination each row lane stores its row of U (the reciprocal
//    of the pivot on the diagonal) in the segment's triangular LDS block U
//    (row i at i n - i (i - 1) / 2), and the back substitution runs on the NR
//    right-hand-side lanes alone, reading U from LDS.
// X (LDS, column-major n x NR) receives the solution.  U holds
// n (n + 1) / 2 + SEG doubles (the last SEG are write sinks).
template <int n, int SEG, int NR>
__device__ __forceinline__ void lu_solve_tri(double (&A)[n], double (&x)[NR], bool act, double* U, double* X) {
  const int lane = threadIdx.x;
  int pos = act ? (lane % SEG) : (1 << 20);
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const double inv = 1.0 / A[k];
    const bool elig = act && pos >= k;
    double av = -2.0;
    if (elig) {
      const double t = fabs(A[k]);
      av = (t == t) ? t : (pos == k ? __builtin_huge_val() : -1.0);
    }
    const double mx = seg_max<SEG>(av);
    bool several;
    int lp = seg_first<SEG>(av == mx, &several);
    if (__any(several)) {
      const int first = seg_min_i<SEG>(av == mx ? pos : (1 << 30));
      lp = seg_first<SEG>(av == mx && pos == first, &several);
    }
    const bool isp = lane == lp;
    const bool below = elig && !isp;
    const double dinv = __shfl(inv, lp, 64);
    const double l = A[k] * dinv;
    const double l0 = below ? l : 0.0;
#pragma unroll
    for (int j = k + 1; j < n; ++j) {
      const double pj = __shfl(A[j], lp, 64);
      A[j] = A[j] - l0 * pj;
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double pq = __shfl(x[q], lp, 64);
      x[q] = x[q] - pq * l0;
    }
    const int ppos = __shfl(pos, lp, 64);
    if (act) {
      if (pos == k) pos = ppos;
      else if (isp) pos = k;
    }
    if (isp) A[k] = dinv;
    else if (below) A[k] = l;
  }
  // rows of U and the forward-reduced right-hand sides to LDS
  constexpr int nt = n * (n + 1) / 2;
  const int sink = nt + lane % SEG;
  const int base = act ? pos * n - pos * (pos - 1) / 2 - pos : 0;
#pragma unroll
  for (int j = 0; j < n; ++j) U[(act && j >= pos) ? base + j : sink] = A[j];
  if (act) {
#pragma unroll
    for (int q = 0; q < NR; ++q) X[pos + n * q] = x[q];
  }
  __syncthreads();
  // back substitution on the right-hand-side lanes, each row's sum in
  // ascending column order
  const int rq = lane % SEG;
  if (rq < NR) {
    double* Xq = X + n * rq;
#pragma unroll
    for (int i = n - 1; i >= 0; --i) {
      const double* Ui = U + (i * n - i * (i - 1) / 2 - i);
      double t = Xq[i];
#pragma unroll
      for (int j = i + 1; j < n; ++j) t = t - Ui[j] * Xq[j];
      Xq[i] = t * Ui[i];
    }
  }
  __syncthreads();
}

template <int N1>
struct LineSeg {
  static constexpr int n3 = 3 * N1;
  static constexpr int SEG3 = n3 <= 8 ? 8 : (n3 <= 16 ? 16 : (n3 <= 32 ? 32 : 64));
  static constexpr int SEGU = N1 <= 8 ? 8 : (N1 <= 16 ? 16 : 32);
};

// ---------------------------------------------------------------------------
// 2. construct_matbnd_uv (:905-1011) + solve_uv (:589-689)
template <int N1>
__global__ void __launch_bounds__(64) hevi_solve_uv_kernel(const HeviArgs a) {
  constexpr int N = N1, Nfp = N1 * N1, Np = N1 * N1 * N1;
  constexpr int SEG = LineSeg<N1>::SEGU, LPW = 64 / SEG;
  extern __shared__ double lds[];
  const int lane = threadIdx.x, sg = lane / SEG, r = lane % SEG;
  const bool act = r < N;
  const DevMesh& m = a.m;
  const int nline = m.Ne2D * Nfp;
  const int NeZ = a.NeZ;
  const uint32_t nn = (uint32_t)Np * m.Ne;
  constexpr int nU = N * (N + 1) / 2 + SEG;
  double* U = lds + (size_t)sg * (nU + (size_t)NeZ * 3 * N);  // the block's U rows (triangular)
  double* X = U + nU;  // per level: b(:,1), b(:,2), G
  const int pv = act ? r : 0;
  for (int g = blockIdx.x; g * LPW < nline; g += gridDim.x) {
    const int line = g * LPW + sg;
    const bool lv = line < nline;
    const int ke2D = lv ? line / Nfp : 0, ij = lv ? line % Nfp : 0;
    for (int kz = 0; kz < NeZ; ++kz) {
      const int ke = ke2D + kz * m.Ne2D;
      double A[N];
#pragma unroll
      for (int c = 0; c < N; ++c) A[c] = (c == pv) ? 1.0 : 0.0;
      double x[3] = {0.0, 0.0, 0.0};
      double lvv = 0.0;
      if (lv && act) {
        const uint32_t q = (uint32_t)Np * ke + ij + Nfp * pv;
        x[0] = ldg(a.buv, q);
        x[1] = ldg(a.buv, q + nn);
        // face damping of the row: f1 = 1 couples to the level below, 2 above
#pragma unroll
        for (int f1 = 1; f1 <= 2; ++f1) {
          const int kz2 = (f1 == 1) ? (kz - 1 > 0 ? kz - 1 : 0) : (kz + 1 < NeZ - 1 ? kz + 1 : NeZ - 1);
          const int f2 = (f1 == 1) ? 2 : 1;
          const bool bc = (kz == 0 && f1 == 1) || (kz == NeZ - 1 && f1 == 2);
          if (bc) continue;
          const int ke_2 = ke2D + kz2 * m.Ne2D;
          const int p = ij + Nfp * pv;
          const uint32_t gfp = (uint32_t)((3 + f1) * Nfp + ij) + (uint32_t)(6 * Nfp) * ke;
          const double fac = 0.5 * a.impl_fac / gsqrtv(a, ke, p, N1);
          const double am = ldg(a.alph, (uint32_t)((f1 - 1) * Nfp + ij) + (uint32_t)(2 * Nfp) * ke);
          const double ap = ldg(a.alph, (uint32_t)((f2 - 1) * Nfp + ij) + (uint32_t)(2 * Nfp) * ke_2);
          const double t = fac * ldg(m.Lift, (uint32_t)(p + (3 + f1) * Np)) * ldg(m.Fscale, gfp) * (am > ap ? am : ap);
          if (f1 == 1) {
            A[0] = A[0] + t;
            lvv = -t;
          } else {
            A[N - 1] = A[N - 1] + t;
            x[2] = -t;  // U lands in the G slot
          }
        }
      }
      if (kz > 0) {
        const double* Xp = X + (size_t)(kz - 1) * 3 * N;
        A[0] = A[0] - lvv * Xp[(N - 1) + 2 * N];
        x[0] = x[0] - lvv * Xp[N - 1];
        x[1] = x[1] - lvv * Xp[(N - 1) + N];
      }
      // at the top level the G column is zero and its solution unused
      lu_solve_tri<N, SEG, 3>(A, x, act, U, X + (size_t)kz * 3 * N);
    }
    for (int kz = NeZ - 2; kz >= 0; --kz) {
      double* Xk = X + (size_t)kz * 3 * N;
      const double* Xn = X + (size_t)(kz + 1) * 3 * N;
      if (act) {
        const double tmp = Xk[r + 2 * N];
        Xk[r] = Xk[r] - tmp * Xn[0];
        Xk[r + N] = Xk[r + N] - tmp * Xn[N];
      }
      __syncthreads();
    }
    if (lv && act) {
      for (int kz = 0; kz < NeZ; ++kz) {
        const uint32_t q = (uint32_t)Np * (ke2D + kz * m.Ne2D) + ij + Nfp * pv;
        const double* Xk = X + (size_t)kz * 3 * N;
        stg(a.pv[VU], q, ldg(a.pin[VU], q) + Xk[pv]);
        stg(a.pv[VV], q, ldg(a.pin[VV], q) + Xk[pv + N]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 4. construct_matbnd (:693-901) + solve (:329-479).
// Unknowns of an element: (var, pv), var fastest: DENS, MOMZ, RHOT.
template <int N1>
__global__ void __launch_bounds__(64) hevi_solve3_kernel(const HeviArgs a) {
  constexpr int N = N1, Nfp = N1 * N1, Np = N1 * N1 * N1, n3 = 3 * N1;
  constexpr int SEG = LineSeg<N1>::SEG3, LPW = 64 / SEG;
  extern __shared__ double lds[];
  const int lane = threadIdx.x, sg = lane / SEG, r = lane % SEG;
  const bool act = r < n3;
  const DevMesh& m = a.m;
  const int nline = m.Ne2D * Nfp;
  const int NeZ = a.NeZ;
  const uint32_t nn = (uint32_t)Np * m.Ne;
  const double impl_fac = a.impl_fac;
  constexpr int nU = n3 * (n3 + 1) / 2 + SEG;
  double* sD = lds;      // D1D
  double* sI = sD + N * N;  // IntrpMat_VPOrdM1_tr
  double* cf = sI + N * N + (size_t)sg * (13 * N + nU + (size_t)NeZ * 4 * n3);  // [3 levels][POT, WT, DPDRHOT][N]
  double* t12 = cf + 9 * N;  // [face][tmp1, tmp2][N]
  double* U = t12 + 4 * N;   // the block's U rows (triangular)
  double* X = U + nU;        // NeZ * 4 * n3: b, G(:,1..3) per level
  for (int e = lane; e < N * N; e += 64) {
    sD[e] = m.D1D[e];
    sI[e] = m.Itr[e];
  }
  const int pv = act ? r / 3 : 0, v = r % 3;
  for (int g = blockIdx.x; g * LPW < nline; g += gridDim.x) {
    const int line = g * LPW + sg;
    const bool lv = line < nline;
    const int ke2D = lv ? line / Nfp : 0, ij = lv ? line % Nfp : 0;
    for (int kz = 0; kz < NeZ; ++kz) {
      const int ke = ke2D + kz * m.Ne2D;
      __syncthreads();
      if (lv) {
        // POT, WT, DPDRHOT of the line in this level (slot 1) and the levels below (0) / above (2)
        for (int e = r; e < 9 * N; e += SEG) {
          const int lvl = e / (3 * N), rr = e % (3 * N), c = rr / N, pq = rr % N;
          int kzz = kz + lvl - 1;
          kzz = kzz < 0 ? 0 : (kzz > NeZ - 1 ? NeZ - 1 : kzz);
          const uint32_t q = (uint32_t)Np * (ke2D + kzz * m.Ne2D) + ij + Nfp * pq;
          cf[e] = ldg(a.coef, q + (uint32_t)c * nn);
        }
        // tmp1, tmp2 of both vertical faces
        if (r < 2 * N) {
          const int f1 = 1 + r / N, pq = r % N;
          const int kz2 = (f1 == 1) ? (kz - 1 > 0 ? kz - 1 : 0) : (kz + 1 < NeZ - 1 ? kz + 1 : NeZ - 1);
          const bool bc = (kz == 0 && f1 == 1) || (kz == NeZ - 1 && f1 == 2);
          const int f2 = bc ? f1 : ((f1 == 1) ? 2 : 1);
          const int ke_2 = ke2D + kz2 * m.Ne2D;
          const int p = ij + Nfp * pq;
          const uint32_t gfp = (uint32_t)((3 + f1) * Nfp + ij) + (uint32_t)(6 * Nfp) * ke;
          const double Lift = ldg(m.Lift, (uint32_t)(p + (3 + f1) * Np));
          const double fac = 0.5 * impl_fac / gsqrtv(a, ke, p, N1) * Lift * ldg(m.Fscale, gfp);
          const double am = ldg(a.alph, (uint32_t)((f1 - 1) * Nfp + ij) + (uint32_t)(2 * Nfp) * ke);
          const double ap = ldg(a.alph, (uint32_t)((f2 - 1) * Nfp + ij) + (uint32_t)(2 * Nfp) * ke_2);
          t12[(f1 - 1) * 2 * N + pq] = fac * (am > ap ? am : ap);
          t12[(f1 - 1) * 2 * N + N + pq] = fac * ldg(m.nz, gfp);
        }
      }
      __syncthreads();
      double A[n3];
      double x[4] = {0.0, 0.0, 0.0, 0.0};
      double lm[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int cc = 0; cc < n3; ++cc) A[cc] = 0.0;
      if (lv && act) {
        const double* POT0 = cf + 3 * N;  // this level
        const double* WT0 = POT0 + N;
        const double* DPD0 = WT0 + N;
        const uint32_t prow = (uint32_t)Np * ke + ij + Nfp * pv;
        const double Eg = ldg(m.E33, prow) / gsqrtv(a, ke, ij + Nfp * pv, N1);
        x[0] = ldg(a.b3, prow + (uint32_t)v * nn);
        const double tm1[2] = {t12[pv], t12[2 * N + pv]};      // tmp1 of faces 1, 2
        const double tm2[2] = {t12[N + pv], t12[3 * N + pv]};  // tmp2
        // D block with the face terms of its first / last node column
#pragma unroll
        for (int cc = 0; cc < n3; ++cc) {
          const int v2 = cc % 3, pv2 = cc / 3;
          const double Id = (pv == pv2) ? 1.0 : 0.0;
          const double fdz = Eg * (impl_fac * sD[pv + N * pv2]);
          const double POT = POT0[pv2], WT = WT0[pv2], DPD = DPD0[pv2];
          double d;
          if (v == 0) d = (v2 == 0) ? Id : (v2 == 1 ? fdz : 0.0);
          else if (v == 1) d = (v2 == 1) ? Id : (v2 == 0 ? impl_fac * a.c.GRAV * sI[pv2 + N * pv] : fdz * DPD);
          else d = (v2 == 0) ? -fdz * POT * WT : (v2 == 1 ? fdz * POT : Id + fdz * WT);
#pragma unroll
          for (int f1 = 1; f1 <= 2; ++f1) {
            const int pv1 = (f1 == 1) ? 0 : N - 1;
            if (pv2 != pv1) continue;
            const bool bc = (kz == 0 && f1 == 1) || (kz == NeZ - 1 && f1 == 2);
            const double tmp1 = tm1[f1 - 1], tmp2 = tm2[f1 - 1];
            if (bc) {
              if (v == 2 && v2 == 0) d = d + 2.0 * tmp2 * POT * WT;
              else if (v == 0 && v2 == 1) d = d - 2.0 * tmp2;
              else if (v == 1 && v2 == 1) d = d + 2.0 * tmp1;
              else if (v == 2 && v2 == 1) d = d - 2.0 * tmp2 * POT;
              else if (v == 2 && v2 == 2) d = d - 2.0 * tmp2 * WT;
            } else {
              if (v == 0 && v2 == 0) d = d + tmp1;
              else if (v == 2 && v2 == 0) d = d + tmp2 * POT * WT;
              else if (v == 0 && v2 == 1) d = d - tmp2;
              else if (v == 1 && v2 == 1) d = d + tmp1;
              else if (v == 2 && v2 == 1) d = d - tmp2 * POT;
              else if (v == 1 && v2 == 2) d = d - tmp2 * DPD;
              else if (v == 2 && v2 == 2) d = d + tmp1 - tmp2 * WT;
            }
          }
          A[cc] = d;
        }
        // L (coupling to the top node of the level below) and U (bottom node
        // of the level above, the G right-hand sides of this level)
#pragma unroll
        for (int f1 = 1; f1 <= 2; ++f1) {
          const bool bc = (kz == 0 && f1 == 1) || (kz == NeZ - 1 && f1 == 2);
          if (bc) continue;
          const int pv2 = (f1 == 1) ? N - 1 : 0;
          const double* cfo = cf + (f1 == 1 ? 0 : 6 * N);
          const double POT = cfo[pv2], WT = cfo[N + pv2], DPD = cfo[2 * N + pv2];
          const double tmp1 = tm1[f1 - 1], tmp2 = tm2[f1 - 1];
#pragma unroll
          for (int v2 = 0; v2 < 3; ++v2) {
            double l;
            if (v2 == 0) l = (v == 0) ? -tmp1 : (v == 1 ? 0.0 : -tmp2 * POT * WT);
            else if (v2 == 1) l = (v == 0) ? +tmp2 : (v == 1 ? -tmp1 : +tmp2 * POT);
            else l = (v == 0) ? 0.0 : (v == 1 ? +tmp2 * DPD : -tmp1 + tmp2 * WT);
            if (f1 == 1) lm[v2] = l;
            else x[1 + v2] = l;
          }
        }
      }
      if (kz > 0) {
        const double* Xp = X + (size_t)(kz - 1) * 4 * n3;
        constexpr int p1 = 3 * (N - 1), p2 = p1 + 1, p3 = p1 + 2;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          A[c] = A[c] - lm[0] * Xp[p1 + n3 * (1 + c)] - lm[1] * Xp[p2 + n3 * (1 + c)] - lm[2] * Xp[p3 + n3 * (1 + c)];
        x[0] = x[0] - lm[0] * Xp[p1] - lm[1] * Xp[p2] - lm[2] * Xp[p3];
      }
      // at the top level the G columns are zero and their solution unused
      lu_solve_tri<n3, SEG, 4>(A, x, act, U, X + (size_t)kz * 4 * n3);
    }
    for (int kz = NeZ - 2; kz >= 0; --kz) {
      double* Xk = X + (size_t)kz * 4 * n3;
      const double* Xn = X + (size_t)(kz + 1) * 4 * n3;
      if (act) Xk[r] = Xk[r] - Xk[r + n3] * Xn[0] - Xk[r + 2 * n3] * Xn[1] - Xk[r + 3 * n3] * Xn[2];
      __syncthreads();
    }
    if (lv && act) {
      double* dst = v == 0 ? a.pv[VD] : (v == 1 ? a.pv[VW] : a.pv[VT]);
      const double* src = v == 0 ? a.pin[VD] : (v == 1 ? a.pin[VW] : a.pin[VT]);
      for (int kz = 0; kz < NeZ; ++kz) {
        const uint32_t q = (uint32_t)Np * (ke2D + kz * m.Ne2D) + ij + Nfp * pv;
        stg(dst, q, ldg(src, q) + X[(size_t)kz * 4 * n3 + r]);
      }
    }
  }
}

// tendency = (iterate - q*) / impl_fac  (..._rhot_hevi.F90:937-947)
__global__ void hevi_tend_kernel(double* T0, double* T1, double* T2, double* T3, double* T4, const double* p0,
                                 const double* p1, const double* p2, const double* p3, const double* p4,
                                 const double* q0, const double* q1, const double* q2, const double* q3,
                                 const double* q4, double impl_fac, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  T0[t] = (p0[t] - q0[t]) / impl_fac;
  T1[t] = (p1[t] - q1[t]) / impl_fac;
  T2[t] = (p2[t] - q2[t]) / impl_fac;
  T3[t] = (p3[t] - q3[t]) / impl_fac;
  T4[t] = (p4[t] - q4[t]) / impl_fac;
}

// the same tendency, then the stepper's StoreImplicit on the node: var0 and
// varTmp (first stage) take q, and q = q + impl_fac T (streamed: non-temporal
// loads and stores)
struct TendStore {
  double* T[5];
  const double* p[5];
  double* q[5];
  double* var0[5];
  double* varTmp[5];
  double impl_fac;
  int copy;
};
__global__ void hevi_tend_store_kernel(const TendStore a, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double q = __builtin_nontemporal_load(a.q[v] + t);
    const double T = (__builtin_nontemporal_load(a.p[v] + t) - q) / a.impl_fac;
    __builtin_nontemporal_store(T, a.T[v] + t);
    if (a.copy) {
      __builtin_nontemporal_store(q, a.var0[v] + t);
      __builtin_nontemporal_store(q, a.varTmp[v] + t);
    }
    __builtin_nontemporal_store(q + a.impl_fac * T, a.q[v] + t);
  }
}

// sum of squares of a vector, per-block partial sums (residual norms)
__global__ void sumsq_kernel(const double* x, long n, double* partial) {
  __shared__ double s[256];
  double acc = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    acc += x[t] * x[t];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// Learned initial iterate: features, targets, application, per-line residual
// and fallback.  One workgroup per line, one thread per node of the line
// (L = NeZ*N1 threads); node t of a line is (kz = t / N1, pv = t % N1).
struct LineArgs {
  const double* qs[5];
  const double* q0[5];
  const double* Tex[5];
  double* pv[5];
  double inv_s[5];      // 1 / state scale
  double inv_t[5];      // 1 / tendency-feature scale
  double inv_d[5], d[5];  // increment scale
  double impl_fac, dt;
  int Np, Nfp, Ne2D, N1, L, nlines;
  long line0, stride;
};

__device__ __forceinline__ uint32_t line_node(const LineArgs& a, long line, int t) {
  const int ke2D = (int)(line / a.Nfp), ij = (int)(line % a.Nfp);
  const int kz = t / a.N1, pv = t % a.N1;
  return (uint32_t)a.Np * (ke2D + kz * a.Ne2D) + ij + a.Nfp * pv;
}

__global__ void line_features_kernel(const LineArgs a, float* X, int n_in, float* Y, int n_out) {
  const int row = blockIdx.x, t = threadIdx.x;
  const long line = (a.line0 + (long)row * a.stride) % a.nlines;
  const uint32_t n = line_node(a, line, t);
  float* x = X + (size_t)row * n_in;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double qs = ldg(a.qs[v], n);
    x[v * a.L + t] = (float)(qs * a.inv_s[v]);
    const double te = a.Tex[v] ? ldg(a.Tex[v], n) : 0.0;
    x[(5 + v) * a.L + t] = (float)(a.impl_fac * te * a.inv_t[v]);
    if (Y) Y[(size_t)row * n_out + v * a.L + t] = (float)((ldg(a.pv[v], n) - qs) * a.inv_d[v]);
  }
  if (t == 0) x[10 * a.L] = (float)(a.impl_fac / a.dt);
}

__global__ void line_apply_kernel(const LineArgs a, const float* Y, int n_out) {
  const int row = blockIdx.x, t = threadIdx.x;
  const long line = (a.line0 + (long)row * a.stride) % a.nlines;
  const uint32_t n = line_node(a, line, t);
#pragma unroll
  for (int v = 0; v < 5; ++v)
    stg(a.pv[v], n, ldg(a.qs[v], n) + a.d[v] * (double)Y[(size_t)row * n_out + v * a.L + t]);
}

// residual norm^2 of every line from the uv and var3 right-hand sides
__global__ void line_resid_kernel(const LineArgs a, const double* buv, const double* b3, long nn, double* r) {
  __shared__ double sh[512];
  const long line = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t n = line_node(a, line, t);
  double acc = 0.0;
  for (int c = 0; c < 2; ++c) {
    const double x = ldg(buv, n + (uint32_t)(c * nn));
    acc += x * x;
  }
  for (int c = 0; c < 3; ++c) {
    const double x = ldg(b3, n + (uint32_t)(c * nn));
    acc += x * x;
  }
  sh[t] = acc;
  __syncthreads();
  for (int w = 256; w > 0; w >>= 1) {
    if (t < w && t + w < a.L) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) r[line] = sh[0];
}

// lines whose predicted iterate has the larger residual restart from q^n;
// r_guess becomes the residual of the start the line keeps
__global__ void line_select_kernel(const LineArgs a, double* r_guess, const double* r_plain, double* flag) {
  const long line = blockIdx.x;
  const int t = threadIdx.x;
  const double rg = r_guess[line], rp = r_plain[line];
  const bool fb = !(rg <= rp);
  if (fb) {
    const uint32_t n = line_node(a, line, t);
#pragma unroll
    for (int v = 0; v < 5; ++v) stg(a.pv[v], n, ldg(a.q0[v], n));
  }
  __syncthreads();  // every thread has read r_guess[line]
  if (t == 0) {
    flag[line] = fb ? 1.0 : 0.0;
    r_guess[line] = fb ? rp : rg;
  }
}

__global__ void sum_kernel(const double* x, long n, double* partial) {
  __shared__ double s[256];
  double acc = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) acc += x[t];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

int host_sum(fedg_ctx* ctx, const double* x, long n, double* partial, double* partial_h, double* out) {
  hipLaunchKernelGGL(sum_kernel, dim3(256), dim3(256), 0, ctx->stream, x, n, partial);
  FEDG_HIP_CHECK(hipGetLastError());
  FEDG_HIP_CHECK(hipMemcpyAsync(partial_h, partial, 256 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  double s = 0.0;
  for (int b = 0; b < 256; ++b) s += partial_h[b];
  *out = s;
  return FEDG_OK;
}

LineArgs line_args(const DevMesh& m, const HeviArgs& a, const double* const Tex[5], const double scale[15], double dt) {
  LineArgs l{};
  for (int v = 0; v < 5; ++v) {
    l.qs[v] = a.qs[v];
    l.q0[v] = a.q0[v];
    l.Tex[v] = Tex ? Tex[v] : nullptr;
    l.pv[v] = a.pv[v];
    l.inv_s[v] = 1.0 / scale[v];
    l.inv_t[v] = 1.0 / scale[5 + v];
    l.d[v] = scale[10 + v];
    l.inv_d[v] = 1.0 / scale[10 + v];
  }
  l.impl_fac = a.impl_fac;
  l.dt = dt;
  l.Np = m.Np;
  l.Nfp = m.Nfp;
  l.Ne2D = m.Ne2D;
  l.N1 = m.n1;
  l.L = a.NeZ * m.n1;
  l.nlines = m.Ne2D * m.Nfp;
  l.line0 = 0;
  l.stride = 1;
  return l;
}

template <int N1, bool FLAT>
__global__ void __launch_bounds__(N1* N1* N1) hevi_rhs_uv_kernel(const HeviArgs a) {
  hevi_rhs_uv_body<N1, FLAT>(a);
}
template <int N1, bool FLAT>
__global__ void __launch_bounds__(N1* N1* N1) hevi_rhs3_kernel(const HeviArgs a) {
  hevi_rhs3_body<N1, FLAT>(a);
}
// p = 7 on the flat box at no more than 128 VGPRs, so that two 512-thread
// workgroups share a CU (the kernels held 135-143 VGPRs when every node formed
// the hydrostatic RHOT with two pows; with it loaded, 70-72)
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
hevi_rhs_uv8_flat_kernel(const HeviArgs a) {
  hevi_rhs_uv_body<8, true>(a);
}
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
hevi_rhs38_flat_kernel(const HeviArgs a) {
  hevi_rhs3_body<8, true>(a);
}

// the residual kernels, on the flat box without the metric loads
template <int N1>
void launch_rhs_uv(fedg_ctx* ctx, const HeviArgs& a) {
  const DevMesh& m = ctx->m;
  if (N1 == 8 && m.flat && m.lift_tensor)
    hipLaunchKernelGGL(hevi_rhs_uv8_flat_kernel, dim3(m.Ne), dim3(512), 0, ctx->stream, a);
  else if (m.flat && m.lift_tensor)
    hipLaunchKernelGGL((hevi_rhs_uv_kernel<N1, true>), dim3(m.Ne), dim3(N1 * N1 * N1), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((hevi_rhs_uv_kernel<N1, false>), dim3(m.Ne), dim3(N1 * N1 * N1), 0, ctx->stream, a);
}
template <int N1>
void launch_rhs3(fedg_ctx* ctx, const HeviArgs& a) {
  const DevMesh& m = ctx->m;
  if (N1 == 8 && m.flat && m.lift_tensor)
    hipLaunchKernelGGL(hevi_rhs38_flat_kernel, dim3(m.Ne), dim3(512), 0, ctx->stream, a);
  else if (m.flat && m.lift_tensor)
    hipLaunchKernelGGL((hevi_rhs3_kernel<N1, true>), dim3(m.Ne), dim3(N1 * N1 * N1), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((hevi_rhs3_kernel<N1, false>), dim3(m.Ne), dim3(N1 * N1 * N1), 0, ctx->stream, a);
}

template <int N1>
size_t lds_uv(int NeZ) {
  constexpr int nU = N1 * (N1 + 1) / 2 + LineSeg<N1>::SEGU;
  return sizeof(double) * (size_t)(64 / LineSeg<N1>::SEGU) * (nU + (size_t)NeZ * 3 * N1);
}
template <int N1>
size_t lds_3(int NeZ) {
  constexpr int n3 = 3 * N1;
  constexpr int nU = n3 * (n3 + 1) / 2 + LineSeg<N1>::SEG3;
  return sizeof(double) *
         (2 * N1 * N1 + (size_t)(64 / LineSeg<N1>::SEG3) * (13 * N1 + nU + (size_t)NeZ * 4 * n3));
}

template <int N1>
int run_newton(fedg_ctx* ctx, HeviArgs& a, int max_it, double rtol, double* resid, double* partial,
               double* partial_h, int* it_done, double r0_ext = 0.0) {
  const DevMesh& m = ctx->m;
  const int nline = m.Ne2D * N1 * N1;
  const int guv = (nline + 64 / LineSeg<N1>::SEGU - 1) / (64 / LineSeg<N1>::SEGU);
  const int g3 = (nline + 64 / LineSeg<N1>::SEG3 - 1) / (64 / LineSeg<N1>::SEG3);
  const int nbuv = guv < 65536 ? guv : 65536, nb3 = g3 < 65536 ? g3 : 65536;
  const size_t luv = lds_uv<N1>(a.NeZ), l3 = lds_3<N1>(a.NeZ);
  if (l3 > 64 * 1024 || luv > 64 * 1024) {
    set_error("HEVI: too many elements in the vertical for the line solve's LDS (NeZ)");
    return FEDG_ERR_UNSUPPORTED;
  }
  const long n = (long)m.Np * m.Ne;
  // r0_ext > 0: the residual norm at q^n, the reference of the tolerance when
  // the iteration starts from a predicted iterate (which may already meet it)
  double r0 = r0_ext;
  // rounding floor of the residual: G = impl_fac T - q + q* cancels terms of
  // the size of the state, so ||G|| cannot drop much below eps ||q*||; the
  // iteration also stops at 4 eps ||q*||, and once an update has not halved
  // the residual (it has reached its floor)
  double atol = 0.0;
  if (rtol > 0.0) {
    double q2 = 0.0;
    for (int v = 0; v < 5; ++v) {
      hipLaunchKernelGGL(sumsq_kernel, dim3(256), dim3(256), 0, ctx->stream, a.qs[v], n, partial);
      FEDG_HIP_CHECK(hipGetLastError());
      FEDG_HIP_CHECK(hipMemcpyAsync(partial_h, partial, 256 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int b = 0; b < 256; ++b) q2 += partial_h[b];
    }
    atol = 4.0 * DBL_EPSILON * sqrt(q2);
  }
  double r_prev = 0.0;
  int it = 0;
  for (; it < max_it; ++it) {
    launch_rhs_uv<N1>(ctx, a);
    FEDG_HIP_CHECK(hipGetLastError());
    a.alph_ready = 1;  // the damping coefficients depend on q0 alone
    double r2 = 0.0;
    const bool track = rtol > 0.0 || resid != nullptr;
    if (track) {
      hipLaunchKernelGGL(sumsq_kernel, dim3(256), dim3(256), 0, ctx->stream, a.buv, 2 * n, partial);
      FEDG_HIP_CHECK(hipGetLastError());
      FEDG_HIP_CHECK(hipMemcpyAsync(partial_h, partial, 256 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int b = 0; b < 256; ++b) r2 += partial_h[b];
    }
    hipLaunchKernelGGL(hevi_solve_uv_kernel<N1>, dim3(nbuv), dim3(64), luv, ctx->stream, a);
    FEDG_HIP_CHECK(hipGetLastError());
    a.pin[VU] = a.pv[VU];
    a.pin[VV] = a.pv[VV];
    launch_rhs3<N1>(ctx, a);
    FEDG_HIP_CHECK(hipGetLastError());
    if (track) {
      hipLaunchKernelGGL(sumsq_kernel, dim3(256), dim3(256), 0, ctx->stream, a.b3, 3 * n, partial);
      FEDG_HIP_CHECK(hipGetLastError());
      FEDG_HIP_CHECK(hipMemcpyAsync(partial_h, partial, 256 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int b = 0; b < 256; ++b) r2 += partial_h[b];
      const double r = sqrt(r2);
      if (resid) resid[it] = r;
      if (it == 0 && r0_ext <= 0.0) r0 = r;
      if (rtol > 0.0 && (it > 0 || r0_ext > 0.0) && (r <= rtol * r0 || r <= atol)) break;
      if (rtol > 0.0 && it > 0 && r > 0.5 * r_prev) break;
      r_prev = r;
    }
    hipLaunchKernelGGL(hevi_solve3_kernel<N1>, dim3(nb3), dim3(64), l3, ctx->stream, a);
    FEDG_HIP_CHECK(hipGetLastError());
    a.pin[VD] = a.pv[VD];
    a.pin[VW] = a.pv[VW];
    a.pin[VT] = a.pv[VT];
  }
  *it_done = it;
  return FEDG_OK;
}

// Predicted Newton iterate from the network, with the per-line fallback to
// q^n; returns ||G(q^n)|| (the tolerance's reference) in *r0.
int guess_buffers(fedg_ctx* ctx, HeviWork* w, int n_in, int n_out) {
  const DevMesh& m = ctx->m;
  const long nlines = (long)m.Ne2D * m.Nfp;
  if (w->X && w->n_in == n_in && w->n_out == n_out) return FEDG_OK;
  if (w->X) (void)hipFree(w->X);
  if (w->Y) (void)hipFree(w->Y);
  if (w->r_plain) (void)hipFree(w->r_plain);
  if (w->r_guess) (void)hipFree(w->r_guess);
  if (w->flag) (void)hipFree(w->flag);
  w->X = w->Y = nullptr;
  w->r_plain = w->r_guess = w->flag = nullptr;
  w->chunk = (int)(nlines < 16384 ? nlines : 16384);
  w->n_in = n_in;
  w->n_out = n_out;
  FEDG_HIP_CHECK(hipMalloc((void**)&w->X, sizeof(float) * (size_t)w->chunk * n_in));
  FEDG_HIP_CHECK(hipMalloc((void**)&w->Y, sizeof(float) * (size_t)w->chunk * n_out));
  FEDG_HIP_CHECK(hipMalloc((void**)&w->r_plain, sizeof(double) * nlines));
  FEDG_HIP_CHECK(hipMalloc((void**)&w->r_guess, sizeof(double) * nlines));
  FEDG_HIP_CHECK(hipMalloc((void**)&w->flag, sizeof(double) * nlines));
  return FEDG_OK;
}

template <int N1>
int predict_iterate(fedg_ctx* ctx, HeviWork* w, HeviArgs& a, HeviGuess* g, double* r0) {
  const DevMesh& m = ctx->m;
  const long nlines = (long)m.Ne2D * m.Nfp;
  const int L = a.NeZ * N1;
  int n_in = 0, n_out = 0;
  mlp_dims(g->mlp, &n_in, &n_out);
  if (n_in != 10 * L + 1 || n_out != 5 * L || L > 512) {
    set_error("HEVI guess: the network's sizes do not match 10 NeZ (p+1) + 1 inputs and 5 NeZ (p+1) outputs");
    return FEDG_ERR_ARG;
  }
  const int rcb = guess_buffers(ctx, w, n_in, n_out);
  if (rcb != FEDG_OK) return rcb;
  LineArgs la = line_args(m, a, g->Tex, g->scale, g->dt);
  const long nn = (long)m.Np * m.Ne;
  // residual of every line at q^n (pv holds q^n)
  launch_rhs_uv<N1>(ctx, a);
  FEDG_HIP_CHECK(hipGetLastError());
  a.alph_ready = 1;
  launch_rhs3<N1>(ctx, a);
  FEDG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(line_resid_kernel, dim3(nlines), dim3(L), 0, ctx->stream, la, a.buv, a.b3, nn, w->r_plain);
  FEDG_HIP_CHECK(hipGetLastError());
  double r2 = 0.0;
  const int rcs = host_sum(ctx, w->r_plain, nlines, w->partial, w->partial_h, &r2);
  if (rcs != FEDG_OK) return rcs;
  *r0 = sqrt(r2);
  // network prediction, chunk by chunk
  for (long l0 = 0; l0 < nlines; l0 += w->chunk) {
    const int rows = (int)((nlines - l0) < w->chunk ? (nlines - l0) : w->chunk);
    la.line0 = l0;
    la.stride = 1;
    hipLaunchKernelGGL(line_features_kernel, dim3(rows), dim3(L), 0, ctx->stream, la, w->X, n_in, (float*)nullptr,
                       n_out);
    FEDG_HIP_CHECK(hipGetLastError());
    FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const int rcf = mlp_forward(g->mlp, w->X, w->Y, rows);
    if (rcf != FEDG_OK) return rcf;
    hipLaunchKernelGGL(line_apply_kernel, dim3(rows), dim3(L), 0, ctx->stream, la, w->Y, n_out);
    FEDG_HIP_CHECK(hipGetLastError());
  }
  // residual at the prediction; lines where it is not smaller restart from q^n
  la.line0 = 0;
  launch_rhs_uv<N1>(ctx, a);
  FEDG_HIP_CHECK(hipGetLastError());
  launch_rhs3<N1>(ctx, a);
  FEDG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(line_resid_kernel, dim3(nlines), dim3(L), 0, ctx->stream, la, a.buv, a.b3, nn, w->r_guess);
  FEDG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(line_select_kernel, dim3(nlines), dim3(L), 0, ctx->stream, la, w->r_guess, w->r_plain, w->flag);
  FEDG_HIP_CHECK(hipGetLastError());
  double nfb = 0.0, r2s = 0.0;
  const int rcs2 = host_sum(ctx, w->flag, nlines, w->partial, w->partial_h, &nfb);
  if (rcs2 != FEDG_OK) return rcs2;
  const int rcs3 = host_sum(ctx, w->r_guess, nlines, w->partial, w->partial_h, &r2s);
  if (rcs3 != FEDG_OK) return rcs3;
  g->lines += nlines;
  g->fallbacks += (long)(nfb + 0.5);
  g->r2_plain += r2;
  g->r2_start += r2s;
  return FEDG_OK;
}

template <int N1>
int run_operator(fedg_ctx* ctx, HeviArgs& a) {
  launch_rhs_uv<N1>(ctx, a);
  FEDG_HIP_CHECK(hipGetLastError());
  launch_rhs3<N1>(ctx, a);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

}  // namespace

namespace {
// the hydrostatic RHOT the residual kernels form from PRES_hyd (the same
// expression and pow, so the same bits), once per base state
__global__ void rhot_hyd_kernel(double* out, const double* PRES_hyd, const Consts c, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double rgamm = c.CVdry / c.CPdry;
  const double PRES0ovRdry = c.PRES00 / c.Rdry;
  out[t] = PRES0ovRdry * fedg_pow(PRES_hyd[t] / c.PRES00, rgamm);
}
}  // namespace

int launch_rhot_hyd(fedg_ctx* ctx, double* RHOT_hyd, const double* PRES_hyd, long n) {
  if (n <= 0) return FEDG_OK;
  hipLaunchKernelGGL(rhot_hyd_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, RHOT_hyd, PRES_hyd, ctx->c, n);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int hevi_cal_vi(fedg_ctx* ctx, HeviWork* w, double* const tend[5], const double* const q[5],
                const double* const var0[5], const double* DENS_hyd, const double* PRES_hyd, const double* Rtot,
                const double* CVtot, const double* CPtot, double impl_fac, int max_it, double rtol, double* resid,
                int* it_done, HeviGuess* guess, const HeviStore* store, const double* RHOT_hyd,
                int alph_ready) {
  const DevMesh& m = ctx->m;
  if (store && !(fabs(impl_fac) > 0.0)) {
    set_error("hevi_cal_vi: the fused StoreImplicit needs impl_fac != 0");
    return FEDG_ERR_ARG;
  }
  HeviArgs a{};
  a.m = m;
  a.c = ctx->c;
  for (int v = 0; v < 5; ++v) {
    a.qs[v] = q[v];
    a.q0[v] = var0[v];
    a.pv[v] = w->pv[v];
    a.pin[v] = var0[v];
    a.T[v] = tend[v];
  }
  a.DENS_hyd = DENS_hyd;
  a.PRES_hyd = PRES_hyd;
  a.RHOT_hyd = RHOT_hyd;
  a.alph_ready = alph_ready;
  a.Rtot = Rtot;
  a.CVtot = CVtot;
  a.CPtot = CPtot;
  a.alph = w->alph;
  a.b3 = w->b3;
  a.buv = w->buv;
  a.coef = w->coef;
  a.impl_fac = impl_fac;
  a.NeZ = m.Ne / m.Ne2D;
  const long n = (long)m.Np * m.Ne;
  *it_done = 0;
  // The first Newton iterate is q^n: the kernels read it from var0 until a
  // solve has written the variable's update to pv (run_newton moves a.pin).
  // The predicted iterate and a solve of no iteration start from a copy.
  const bool implicit = fabs(impl_fac) > 0.0;
  if (implicit && ((guess && guess->mlp) || max_it < 1)) {
    for (int v = 0; v < 5; ++v) {
      FEDG_HIP_CHECK(hipMemcpyAsync(w->pv[v], var0[v], sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
      a.pin[v] = w->pv[v];
    }
  }
  auto dispatch = [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    if (implicit) {
      a.want_b = 1;
      double r0 = 0.0;
      if (guess && guess->mlp) {
        const int rc = predict_iterate<N1>(ctx, w, a, guess, &r0);
        if (rc != FEDG_OK) return rc;
      }
      const int rc = run_newton<N1>(ctx, a, max_it, rtol, resid, w->partial, w->partial_h, it_done, r0);
      if (rc != FEDG_OK) return rc;
      const int bs = 256;
      if (store) {
        TendStore ts{};
        for (int v = 0; v < 5; ++v) {
          ts.T[v] = tend[v];
          ts.p[v] = w->pv[v];
          ts.q[v] = store->q[v];
          ts.var0[v] = store->var0[v];
          ts.varTmp[v] = store->varTmp[v];
          if (store->q[v] != q[v]) {
            set_error("hevi_cal_vi: the StoreImplicit state is not the solve's stage state");
            return FEDG_ERR_ARG;
          }
        }
        ts.copy = store->var0[0] != nullptr;
        ts.impl_fac = impl_fac;
        hipLaunchKernelGGL(hevi_tend_store_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, ctx->stream, ts, n);
        FEDG_HIP_CHECK(hipGetLastError());
        return FEDG_OK;
      }
      hipLaunchKernelGGL(hevi_tend_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, ctx->stream, tend[VD], tend[VT],
                         tend[VW], tend[VU], tend[VV], w->pv[VD], w->pv[VT], w->pv[VW], w->pv[VU], w->pv[VV], q[VD],
                         q[VT], q[VW], q[VU], q[VV], impl_fac, n);
      FEDG_HIP_CHECK(hipGetLastError());
      return FEDG_OK;
    }
    a.want_b = 0;
    return run_operator<N1>(ctx, a);
  };
  switch (m.n1) {
    case 2: return dispatch(std::integral_constant<int, 2>{});
    case 3: return dispatch(std::integral_constant<int, 3>{});
    case 4: return dispatch(std::integral_constant<int, 4>{});
    case 5: return dispatch(std::integral_constant<int, 5>{});
    case 6: return dispatch(std::integral_constant<int, 6>{});
    case 7: return dispatch(std::integral_constant<int, 7>{});
    case 8: return dispatch(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

int hevi_harvest(fedg_ctx* ctx, HeviWork* w, const double* const q[5], const double* const var0[5],
                 const double* const Tex[5], const double scale[15], double impl_fac, double dt, long line0,
                 long stride, int nrows, float* X, float* Y) {
  const DevMesh& m = ctx->m;
  HeviArgs a{};
  a.m = m;
  for (int v = 0; v < 5; ++v) {
    a.qs[v] = q[v];
    a.q0[v] = var0[v];
    a.pv[v] = w->pv[v];
  }
  a.impl_fac = impl_fac;
  a.NeZ = m.Ne / m.Ne2D;
  LineArgs la = line_args(m, a, Tex, scale, dt);
  if (la.L > 512 || nrows < 1) return FEDG_ERR_ARG;
  la.line0 = line0;
  la.stride = stride;
  hipLaunchKernelGGL(line_features_kernel, dim3(nrows), dim3(la.L), 0, ctx->stream, la, X, 10 * la.L + 1, Y,
                     5 * la.L);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

// ---------------------------------------------------------------------------
// Workspace of fedg_hevi_cal_vi_noalloc: a file-scope device array in place
// of hipMalloc, for callers that may not allocate (HIP graph capture, serving
// loops).  Every device that runs the library holds its own copy until the
// process ends; every call on a device shares it and finds the previous
// call's contents (the solve writes each part before reading it: the iterate
// is read from var0 until a solve has written it, the rhs, damping and
// linearisation coefficients and the partial sums are written by their
// kernels, so nothing needs clearing).
constexpr size_t kScratchDoubles = (size_t)8 << 20;  // 64 MiB

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kScratchDevices];  // address of `scratch` on each device, set once

// a part of the workspace rounded up to 32 doubles (256 B alignment)
size_t part(size_t n) { return (n + 31) & ~(size_t)31; }

// doubles of the workspace a solve on mesh m needs (the HeviWork layout)
size_t noalloc_need(const DevMesh& m) {
  const size_t n = (size_t)m.Np * m.Ne;
  return 5 * part(n) + part((size_t)2 * m.Nfp * m.Ne) + part(3 * n) + part(2 * n) + part(3 * n) + part(256);
}
}  // namespace

// Looks up the device address of `scratch` on `device`, which must be the
// current device; once per device (later calls return at once).
int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices) {
    set_error("device index beyond the workspace table of fedg_hevi_cal_vi_noalloc (64 devices)");
    return FEDG_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return FEDG_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    set_error(std::string("hipGetSymbolAddress of the no-allocation workspace: ") + hipGetErrorString(e));
    return FEDG_ERR_HIP;
  }
  scratch_dev[device] = static_cast<double*>(ptr);
  return FEDG_OK;
}

}  // namespace fedg

extern "C" int fedg_hevi_cal_vi_noalloc(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt,
                                        double* MOMZ_dt, double* RHOT_dt, const double* DDENS_,
                                        const double* MOMX_, const double* MOMY_, const double* MOMZ_,
                                        const double* DRHOT_, const double* DENS_hyd, const double* PRES_hyd,
                                        const double* DDENS0_, const double* MOMX0_, const double* MOMY0_,
                                        const double* MOMZ0_, const double* DRHOT0_, const double* Rtot,
                                        const double* CVtot, const double* CPtot, double impl_fac, double dt,
                                        int max_it, double rtol, double* resid, int* iterations) {
  using namespace fedg;
  (void)dt;
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ ||
      !DRHOT_ || !DENS_hyd || !PRES_hyd || !DDENS0_ || !MOMX0_ || !MOMY0_ || !MOMZ0_ || !DRHOT0_ || !Rtot ||
      !CVtot || !CPtot || max_it < 1 || !std::isfinite(impl_fac))
    return FEDG_ERR_ARG;
  const DevMesh& m = ctx->m;
  if (m.Ne2D <= 0 || m.Ne % m.Ne2D != 0) {
    set_error("HEVI: the mesh is not a stack of Ne2D-element horizontal layers");
    return FEDG_ERR_UNSUPPORTED;
  }
  const size_t need = noalloc_need(m);
  if (need > kScratchDoubles) {
    set_error("fedg_hevi_cal_vi_noalloc: the mesh needs " + std::to_string(need * sizeof(double)) +
              " bytes of workspace, the fixed workspace holds " + std::to_string(kScratchDoubles * sizeof(double)) +
              " (64 MiB); use fedg_hevi_cal_vi_newton");
    return FEDG_ERR_UNSUPPORTED;
  }
  double* p = (ctx->device >= 0 && ctx->device < kScratchDevices) ? scratch_dev[ctx->device] : nullptr;
  if (p == nullptr) {
    set_error("fedg_hevi_cal_vi_noalloc: the workspace of this context's device was not set up");
    return FEDG_ERR_HIP;
  }
  const size_t n = (size_t)m.Np * m.Ne;
  HeviWork w{};
  for (int v = 0; v < 5; ++v) {
    w.pv[v] = p;
    p += part(n);
  }
  w.alph = p;
  p += part((size_t)2 * m.Nfp * m.Ne);
  w.b3 = p;
  p += part(3 * n);
  w.buv = p;
  p += part(2 * n);
  w.coef = p;
  p += part(3 * n);
  w.partial = p;
  double* tend[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  const double* q[5] = {DDENS_, DRHOT_, MOMZ_, MOMX_, MOMY_};
  const double* q0[5] = {DDENS0_, DRHOT0_, MOMZ0_, MOMX0_, MOMY0_};
  int it = 0;
  const int rc = hevi_cal_vi(ctx, &w, tend, q, q0, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, impl_fac, max_it, rtol,
                             resid, &it, nullptr);
  if (rc != FEDG_OK) return rc;
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (iterations) *iterations = it;
  return FEDG_OK;
}
