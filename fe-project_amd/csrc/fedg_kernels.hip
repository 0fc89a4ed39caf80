// CDNA4 (gfx950) kernels for the HEVE step of FElib's regional nonhydrostatic
// DG dynamical core.
//
// Element kernel: one workgroup handles EPB elements with one thread per node
// (Np = N1^3 threads per element, 512 at p=7 = 8 wavefronts, wavefront w =
// horizontal node plane k=w).  Per element it
//   1. loads the stage state and base state of its nodes (coalesced, 4 KB per
//      field per element at p=7),
//   2. evaluates the Rusanov flux at its NfpTot face nodes (one thread per
//      face node, exterior values gathered from the neighbour element),
//   3. forms the 15 pointwise fluxes and applies the 1D derivative matrix along
//      x, y and z through LDS (one variable at a time, double-buffered), adds
//      the lift of the face flux and the buoyancy / Coriolis / hydrostatic
//      terms,
//   4. either writes the five tendencies (drop-in cal_tend) or applies the RK
//      stage update in place of the reference's separate streaming pass
//      (fused stepper).
// The arithmetic of every node follows the reference's expression order
// (compiled with -ffp-contract=off; division and sqrt are IEEE correctly
// rounded), and the pressure's pow is the portable double-double routine of
// fedg_pow.h, so results agree bit for bit with a CPU evaluation of the same
// expressions that uses the same pow.
#include <type_traits>
#include <vector>

#include "fedg_device.h"

namespace fedg {
namespace {

// MODE_DROPIN_HEVI: the explicit tendency of the HEVI scheme
// (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:289-482): horizontal-only
// dissipation in the face flux, no vertical mass / rhotheta flux, no vertical
// pressure gradient and no buoyancy (those are the implicit part).
// MODE_DROPIN_CCO: the interior part of the drop-in tendency
// (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve.F90:493-678, cal_tend_cco): no
// face flux and no lift; the face contribution is added afterwards by
// heve_bnd_kernel (add_bnd_contrib_generalvc), so that a halo exchange can
// overlap the interior work.
enum { MODE_DROPIN = 0, MODE_FUSED_GENERAL = 1, MODE_FUSED_FLAT = 2, MODE_DROPIN_HEVI = 3, MODE_DROPIN_CCO = 4 };

// Own-node values staged in LDS for the face phase: DDENS, DRHOT, MOMZ, MOMX,
// MOMY (VID order), then DENS_hyd, PRES_hyd, THERM_hyd, DPRES.
enum { S_DH = 5, S_PH = 6, S_TH = 7, S_DP = 8, S_N = 9 };

template <int N1, int EPB, int MODE>
__global__ void __launch_bounds__(EPB * N1 * N1 * N1, 2)
heve_kernel(const DevMesh m, const Consts c, const TendIn in, const RkOut rk) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1, NfpTot = 6 * Nfp;
  constexpr bool HEVI = MODE == MODE_DROPIN_HEVI;
  constexpr bool CCO = MODE == MODE_DROPIN_CCO;
  constexpr bool DROPIN = MODE == MODE_DROPIN || HEVI || CCO;
  constexpr bool FLAT = MODE == MODE_FUSED_FLAT;
  constexpr bool FUSED = !DROPIN;
  // exterior values are loaded before the first barrier when every face node
  // has its own thread (p >= 5); otherwise inside the face loop
  constexpr bool PRELOAD = NfpTot <= Np;

  __shared__ double sDt[N1 * N1];
  __shared__ double sI[N1 * N1];
  __shared__ double sS[EPB][S_N][Np];
  __shared__ double sF[2][EPB][3][Np];
  __shared__ double sDF[EPB][5][NfpTot];

  const int tid = threadIdx.x;
  const int e = tid / Np;
  const int p = tid - e * Np;
  int blk = blockIdx.x;
  if constexpr (FLAT) blk = xcd_remap(blk, gridDim.x);
  const int ke = blk * EPB + e;
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / Nfp;
  const double gamm = c.CPdry / c.CVdry;

  for (int t = tid; t < N1 * N1; t += blockDim.x) {
    sDt[(t % N1) * N1 + t / N1] = m.D1D[t];
    sI[t] = m.Itr[t];
  }

  // ---- exterior side of this thread's face node: index and boundary kind ----
  const int fp0 = p;
  const int f0 = fp0 / Nfp, a0 = fp0 - f0 * Nfp;
  const bool has_face = valid && fp0 < NfpTot;
  uint32_t iP0 = 0;
  int bc0 = FEDG_BND_NONE;
  const int pm0 = face_node<N1>(f0 < 6 ? f0 : 0, a0);
  auto exterior = [&](int f, int a, int pm, uint32_t& iP, int& bc) {
    const uint32_t gf = (uint32_t)(f * Nfp + a) + (uint32_t)NfpTot * ke;
    if constexpr (DROPIN) {
      iP = (uint32_t)m.vmapP[gf];
      bc = FEDG_BND_NONE;
    } else if constexpr (FLAT) {
      const int ei = ke % m.NeX, ej = (ke / m.NeX) % m.NeY, ek = ke / (m.NeX * m.NeY);
      int ni = ei, nj = ej, nk = ek;
      bool closed = false;
      switch (f) {
        case 0: nj -= 1; if (nj < 0) { if (m.periodic[1]) nj = m.NeY - 1; else closed = true; } break;
        case 1: ni += 1; if (ni >= m.NeX) { if (m.periodic[0]) ni = 0; else closed = true; } break;
        case 2: nj += 1; if (nj >= m.NeY) { if (m.periodic[1]) nj = 0; else closed = true; } break;
        case 3: ni -= 1; if (ni < 0) { if (m.periodic[0]) ni = m.NeX - 1; else closed = true; } break;
        case 4: nk -= 1; if (nk < 0) { if (m.periodic[2]) nk = m.NeZ - 1; else closed = true; } break;
        default: nk += 1; if (nk >= m.NeZ) { if (m.periodic[2]) nk = 0; else closed = true; } break;
      }
      if (closed) {
        iP = (uint32_t)Np * ke + pm;
        bc = m.bnd_bc[f];
      } else {
        constexpr int opp[6] = {2, 3, 0, 1, 5, 4};
        iP = (uint32_t)Np * (ni + m.NeX * (nj + m.NeY * nk)) + face_node<N1>(opp[f], a);
        bc = FEDG_BND_NONE;
      }
    } else {
      iP = (uint32_t)m.nbr[gf];
      bc = m.nbr_bc[gf];
    }
  };
  struct Ext {
    double DD, DT, MZ, MX, MY, Dh, Ph, Th, DP, Gs, g13, g23;
  };
  auto load_ext = [&](uint32_t iP, Ext& x) {
    x.DD = ldg(in.q[VD], iP); x.DT = ldg(in.q[VT], iP); x.MZ = ldg(in.q[VW], iP);
    x.MX = ldg(in.q[VU], iP); x.MY = ldg(in.q[VV], iP);
    x.Dh = ldg(in.DENS_hyd, iP); x.Ph = ldg(in.PRES_hyd, iP); x.Th = ldg(in.THERM_hyd, iP);
    x.DP = ldg(in.DPRES, iP);
    if constexpr (!FLAT) {
      x.Gs = ldg(m.Gsqrt, iP); x.g13 = ldg(m.G13, iP); x.g23 = ldg(m.G23, iP);
    } else {
      x.Gs = 1.0; x.g13 = 0.0; x.g23 = 0.0;
    }
  };
  Ext X0;
  if constexpr (PRELOAD && !CCO) {
    if (has_face) {
      exterior(f0, a0, pm0, iP0, bc0);
      load_ext(iP0, X0);
    }
  }

  // ---- own node: state, base state, pressure ----
  const uint32_t o = (uint32_t)Np * (valid ? ke : 0) + p;
  double q[5] = {0, 0, 0, 0, 0};
  double Dh = 0, Ph = 0, Th = 0, DP = 0;
  double Gs = 1.0, g13 = 0.0, g23 = 0.0, GsH = 1.0;
  int ke2d = 0;
  double cor = 0.0, dpx = 0.0, dpy = 0.0;
  if (valid) {
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = ldg(in.q[v], o);
    Dh = ldg(in.DENS_hyd, o);
    Ph = ldg(in.PRES_hyd, o);
    Th = ldg(in.THERM_hyd, o);
    if constexpr (FLAT) {
      ke2d = ke % (m.NeX * m.NeY);
    } else {
      Gs = ldg(m.Gsqrt, o);
      g13 = ldg(m.G13, o);
      g23 = ldg(m.G23, o);
      ke2d = m.EMap3Dto2D[ke];
      GsH = ldg(m.GsqrtH, (uint32_t)(i + N1 * j) + (uint32_t)Nfp * ke2d);
    }
    cor = ldg(in.CORIOLIS, (uint32_t)(i + N1 * j) + (uint32_t)Nfp * ke2d);
    dpx = ldg(in.DPhydDx, o);
    dpy = ldg(in.DPhydDy, o);
    DP = ldg(in.DPRES, o);
  }
  double L[6] = {0, 0, 0, 0, 0, 0};
  if constexpr (!CCO) {
#pragma unroll
    for (int f = 0; f < 6; ++f) L[f] = ldg(m.Lift, (uint32_t)(p + f * Np));
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) sS[e][v][p] = q[v];
  sS[e][S_DH][p] = Dh;
  sS[e][S_PH][p] = Ph;
  sS[e][S_TH][p] = Th;
  sS[e][S_DP][p] = DP;
  __syncthreads();

  // ---- face phase: Rusanov flux at the element's face nodes ----
  if (valid && !CCO) {
    for (int fp = p; fp < NfpTot; fp += Np) {
      const int f = fp / Nfp, a = fp - f * Nfp;
      const int pm = PRELOAD ? pm0 : face_node<N1>(f, a);
      const uint32_t gf = (uint32_t)fp + (uint32_t)NfpTot * ke;
      uint32_t iP;
      int bc;
      Ext X;
      if constexpr (PRELOAD) {
        iP = iP0;
        bc = bc0;
        X = X0;
      } else {
        exterior(f, a, pm, iP, bc);
        load_ext(iP, X);
      }
      const uint32_t iM = DROPIN ? (uint32_t)m.vmapM[gf] : (uint32_t)Np * ke + pm;
      const int pmM = DROPIN ? (int)(iM - (uint32_t)Np * ke) : pm;
      Side M, P;
      double nx, ny, nz, Fs;
      if constexpr (FLAT) {
        nx = (f == 1) ? 1.0 : (f == 3) ? -1.0 : 0.0;
        ny = (f == 2) ? 1.0 : (f == 0) ? -1.0 : 0.0;
        nz = (f == 5) ? 1.0 : (f == 4) ? -1.0 : 0.0;
        Fs = m.fF[f];
        M.Gsqrt = 1.0; M.G13 = 0.0; M.G23 = 0.0;
      } else {
        nx = ldg(m.nx, gf);
        ny = ldg(m.ny, gf);
        nz = ldg(m.nz, gf);
        Fs = ldg(m.Fscale, gf);
        M.Gsqrt = ldg(m.Gsqrt, iM); M.G13 = ldg(m.G13, iM); M.G23 = ldg(m.G23, iM);
      }
      M.DD = sS[e][VD][pmM]; M.DT = sS[e][VT][pmM]; M.MZ = sS[e][VW][pmM];
      M.MX = sS[e][VU][pmM]; M.MY = sS[e][VV][pmM];
      M.Dhyd = sS[e][S_DH][pmM]; M.Phyd = sS[e][S_PH][pmM]; M.Thyd = sS[e][S_TH][pmM];
      M.DP = sS[e][S_DP][pmM];
      P.Gsqrt = X.Gs; P.G13 = X.g13; P.G23 = X.g23;
      P.DD = X.DD; P.DT = X.DT; P.MZ = X.MZ; P.MX = X.MX; P.MY = X.MY;
      P.Dhyd = X.Dh; P.Phyd = X.Ph; P.Thyd = X.Th;
      P.DP = X.DP;
      if (bc != FEDG_BND_NONE) {
        // exterior momentum from the interior node, as applyBC writes it into the halo
        double GH = 1.0, g11 = 1.0, g12 = 0.0, g22 = 1.0;
        if constexpr (!FLAT) {
          const uint32_t h = (uint32_t)(pm % Nfp) + (uint32_t)Nfp * m.EMap3Dto2D[ke];
          GH = ldg(m.GsqrtH, h);
          g11 = ldg(m.G11, h);
          g12 = ldg(m.G12, h);
          g22 = ldg(m.G22, h);
        }
        reflect_momentum(bc, M.Gsqrt, GH, g11, g12, g22, M.G13, M.G23, nx, ny, nz, M.MX, M.MY, M.MZ, P.MX, P.MY,
                         P.MZ);
      }
      double df[5];
      if constexpr (HEVI) rusanov_hevi(M, P, nx, ny, nz, Fs, gamm, df);
      else rusanov(M, P, nx, ny, nz, Fs, gamm, df);
#pragma unroll
      for (int v = 0; v < 5; ++v) sDF[e][v][fp] = df[v];
    }
  }

  // ---- volume phase ----
  // RK operands, loaded now so their latency overlaps the contractions
  double rkA[5] = {0, 0, 0, 0, 0}, rkB[5] = {0, 0, 0, 0, 0};
  if constexpr (FUSED) {
    if (valid) {
      if (rk.mode == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          rkA[v] = ldg(rk.var0[v], o);
          rkB[v] = ldg(rk.varTmp[v], o);
        }
      } else if (rk.mode == 2) {
#pragma unroll
        for (int v = 0; v < 5; ++v) rkB[v] = ldg(rk.varTmp[v], o);
      }
    }
  }
  const double RDENS = 1.0 / (q[VD] + Dh);
  const double GsqrtV = Gs / GsH;
  const double RGsqrtV = 1.0 / GsqrtV;
  const double RGsqrt = 1.0 / Gs;
  double E11, E22, E33;
  if constexpr (FLAT) {
    E11 = m.fE[0];
    E22 = m.fE[1];
    E33 = m.fE[2];
  } else {
    E11 = valid ? ldg(m.E11, o) : 0.0;
    E22 = valid ? ldg(m.E22, o) : 0.0;
    E33 = valid ? ldg(m.E33, o) : 0.0;
  }
  // VFilterPM1 of DDENS (ELM/..._tensorprod3D.F90:2828-2839)
  double drho = sS[e][VD][i + N1 * j] * sI[N1 * k];
#pragma unroll
  for (int l = 1; l < N1; ++l) drho = drho + sS[e][VD][i + N1 * (j + N1 * l)] * sI[l + N1 * k];

  // One variable at a time: pointwise flux (FDS/..._rhot_heve.F90:376-418),
  // Div_var5 (ELM/..._tensorprod3D.F90:2760-2783: Dx Fx, Dy Fy, Dz Fz and the
  // lift of del_flux), tendency (:424-483) and its RK update
  // (CMN/scale_timeint_rk.F90:2232-2350 with one tendency buffer).
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double m0 = Gs * q[VU];
    const double m1 = Gs * q[VV];
    const double m2 = Gs * (q[VW] * RGsqrtV + g13 * q[VU] + g23 * q[VV]);
    double f0, f1, f2;
    if (v == VD) {
      f0 = m0;
      f1 = m1;
      f2 = m2;
    } else if (v == VT) {
      const double pt_ = (Th + q[VT]) * RDENS;
      f0 = m0 * pt_;
      f1 = m1 * pt_;
      f2 = m2 * pt_;
    } else if (v == VW) {
      const double w_ = q[VW] * RDENS;
      f0 = m0 * w_;
      f1 = m1 * w_;
      f2 = HEVI ? m2 * w_ : m2 * w_ + Gs * DP * RGsqrtV;
    } else if (v == VU) {
      const double GsqrtDPRES_ = Gs * DP;
      const double u_ = q[VU] * RDENS;
      f0 = m0 * u_ + GsqrtDPRES_;
      f1 = m1 * u_;
      f2 = m2 * u_ + GsqrtDPRES_ * g13;
    } else {
      const double GsqrtDPRES_ = Gs * DP;
      const double v_ = q[VV] * RDENS;
      f0 = m0 * v_;
      f1 = m1 * v_ + GsqrtDPRES_;
      f2 = m2 * v_ + GsqrtDPRES_ * g23;
    }
    const int b = v & 1;
    sF[b][e][0][p] = f0;
    sF[b][e][1][p] = f1;
    sF[b][e][2][p] = f2;
    __syncthreads();
    const double* Fx = sF[b][e][0];
    const double* Fy = sF[b][e][1];
    const double* Fz = sF[b][e][2];
    // one direction at a time, so that only one line of operands is live
    // (sDt holds D1D transposed: row i of D1D is sDt[N1*i .. N1*i+N1-1])
    double dx = sDt[N1 * i] * Fx[N1 * (j + N1 * k)];
#pragma unroll
    for (int l = 1; l < N1; ++l) dx = dx + sDt[l + N1 * i] * Fx[l + N1 * (j + N1 * k)];
    __builtin_amdgcn_sched_barrier(0);
    double dy = Fy[i + N1 * (N1 * k)] * sDt[N1 * j];
#pragma unroll
    for (int l = 1; l < N1; ++l) dy = dy + Fy[i + N1 * (l + N1 * k)] * sDt[l + N1 * j];
    __builtin_amdgcn_sched_barrier(0);
    double dz = Fz[i + N1 * j] * sDt[N1 * k];
#pragma unroll
    for (int l = 1; l < N1; ++l) dz = dz + Fz[i + N1 * (j + N1 * l)] * sDt[l + N1 * k];
    __builtin_amdgcn_sched_barrier(0);
    const double* dfv = sDF[e][v];
    double X;
    if constexpr (CCO) {
      // Div_var5_2: the three line contractions only (..._rhot_heve.F90:607-669)
      X = E11 * dx + E22 * dy + E33 * dz;
    } else {
      const double lift = L[0] * dfv[i + N1 * k] + L[1] * dfv[Nfp + j + N1 * k] + L[2] * dfv[2 * Nfp + i + N1 * k] +
                          L[3] * dfv[3 * Nfp + j + N1 * k] + L[4] * dfv[4 * Nfp + i + N1 * j] +
                          L[5] * dfv[5 * Nfp + i + N1 * j];
      // HEVI: DENS and RHOT have no vertical flux term, MOMZ no buoyancy
      X = (HEVI && (v == VD || v == VT)) ? E11 * dx + E22 * dy + lift : E11 * dx + E22 * dy + E33 * dz + lift;
    }
    double T;
    if (v == VD || v == VT) {
      T = -X * RGsqrt;
    } else if (v == VW) {
      T = HEVI ? -X * RGsqrt : -X * RGsqrt - c.GRAV * drho;
    } else if (v == VU) {
      T = -dpx + cor * q[VV];
      T = T - X * RGsqrt;
    } else {
      T = -dpy - cor * q[VU];
      T = T - X * RGsqrt;
    }
    if constexpr (FUSED) {
      // add_phy_tend (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:1098-1178):
      // the physics tendencies join the stage's tendency before the update
      if (valid && in.tp[0] != nullptr) {
        if (v == VT) {
          double Rt = c.Rdry, Cv = c.CVdry, Cp = c.CPdry;
          if (in.Rtot != nullptr) {
            Rt = ldg(in.Rtot, o);
            Cv = ldg(in.CVtot, o);
            Cp = ldg(in.CPtot, o);
          }
          T = phys_add_rhot(T, ldg(in.tp[VT], o), ldg(in.tp[5], o), pres_of(q[VT], Th, Rt, Cv, Cp, c), Rt, Cp, c);
        } else {
          T = T + ldg(in.tp[v], o);
        }
      }
    }
    if (valid) {
      if constexpr (DROPIN) {
        stg(rk.out[v], o, T);
      } else {
        double qn;
        switch (rk.mode) {
          case 0:
            stg(rk.varTmp[v], o, q[v] + rk.cb * T);
            qn = q[v] + rk.ca * T;
            break;
          case 1:
            stg(rk.varTmp[v], o, rkB[v] + rk.cb * T);
            qn = rkA[v] + rk.ca * T;
            break;
          case 2:
            qn = rkB[v] + rk.cb * T;
            break;
          default:
            qn = q[v] + rk.cb * T;
            break;
        }
        stg(rk.out[v], o, qn);
        if (v == VT) {
          // pressure perturbation of the new state, read by the next stage
          // (own node and the neighbours' face nodes) instead of recomputed
          double Rt = c.Rdry, Cv = c.CVdry, Cp = c.CPdry;
          if (in.Rtot != nullptr) {
            Rt = ldg(in.Rtot, o);
            Cv = ldg(in.CVtot, o);
            Cp = ldg(in.CPtot, o);
          }
          stg(rk.DPout, o, dpres_of(qn, Th, Ph, Rt, Cv, Cp, c));
        }
      }
    }
    // finish this variable before the next one's LDS traffic; otherwise the
    // scheduler defers the arithmetic and every variable's operands stay live
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ---------------------------------------------------------------------------
// FP32 perturbation form of the fused HEVE stage (flat-terrain box, dry air).
// The tendency -- pointwise fluxes, face flux, derivative contractions, lift,
// buoyancy -- is evaluated in FP32 on the perturbation variables about the
// FP64 base state; the state, the base state and the RK update stay FP64.
// The pressure perturbation, which FP64 forms as the small difference
// P(rhotheta) - P_hyd, is formed without that cancellation:
//   DPRES = P_hyd ((1 + DRHOT/THERM_hyd)^(cp/cv) - 1)
//         = P_hyd expm1((cp/cv) log1p(DRHOT/THERM_hyd)),
// exact for THERM_hyd = P00/R (P_hyd/P00)^(cv/cp) (the base state's
// definition, ..._nonhydro3d_common.F90:584-619).
__device__ __forceinline__ float dpres32(float DT, float Th, float Ph, float gamm) {
  return Ph * expm1f(gamm * log1pf(DT / Th));
}

// Rusanov flux jump in FP32 for the flat box (Gsqrt = 1, G13 = G23 = 0,
// axis-aligned unit normals; same formula as rusanov()).
struct Side32 {
  float DD, DT, MZ, MX, MY, Dh, Ph, Th, DP;
};
__device__ __forceinline__ void rusanov32(const Side32& M, const Side32& P, float nx, float ny, float nz, float Fs,
                                          float gamm, float df[5]) {
  const float DensM = M.DD + M.Dh, DensP = P.DD + P.Dh;
  const float RhotM = M.Th + M.DT, RhotP = P.Th + P.DT;
  const float VelM = (M.MX * nx + M.MY * ny + M.MZ * nz) / DensM;
  const float VelP = (P.MX * nx + P.MY * ny + P.MZ * nz) / DensP;
  const float aM = sqrtf(gamm * (M.Ph + M.DP) / DensM) + fabsf(VelM);
  const float aP = sqrtf(gamm * (P.Ph + P.DP) / DensP) + fabsf(VelP);
  const float alpha = aM > aP ? aM : aP;
  const float h = 0.5f * Fs;
  df[VD] = h * (DensP * VelP - DensM * VelM - alpha * (P.DD - M.DD));
  df[VT] = h * (RhotP * VelP - RhotM * VelM - alpha * (P.DT - M.DT));
  const float dp = P.DP - M.DP;
  df[VW] = h * (P.MZ * VelP - M.MZ * VelM + dp * nz - alpha * (P.MZ - M.MZ));
  df[VU] = h * (P.MX * VelP - M.MX * VelM + dp * nx - alpha * (P.MX - M.MX));
  df[VV] = h * (P.MY * VelP - M.MY * VelM + dp * ny - alpha * (P.MY - M.MY));
}

// HEVI explicit face flux in FP32 for the flat box (rusanov_hevi with
// Gsqrt = 1, G13 = G23 = 0): dissipation across lateral faces only (swV =
// 1 - nz^2), mass and rhotheta carried by the horizontal velocity, no
// vertical pressure jump.
__device__ __forceinline__ void rusanov32_hevi(const Side32& M, const Side32& P, float nx, float ny, float nz,
                                               float Fs, float gamm, float df[5]) {
  const float DensM = M.DD + M.Dh, DensP = P.DD + P.Dh;
  const float RhotM = M.Th + M.DT, RhotP = P.Th + P.DT;
  const float VelhM = (M.MX * nx + M.MY * ny) / DensM;
  const float VelhP = (P.MX * nx + P.MY * ny) / DensP;
  const float VelM = VelhM + (M.MZ * nz) / DensM;
  const float VelP = VelhP + (P.MZ * nz) / DensP;
  const float swV = 1.f - nz * nz;
  const float aM = sqrtf(gamm * (M.Ph + M.DP) / DensM) + fabsf(VelM);
  const float aP = sqrtf(gamm * (P.Ph + P.DP) / DensP) + fabsf(VelP);
  const float alpha = swV * (aM > aP ? aM : aP);
  const float h = 0.5f * Fs;
  df[VD] = h * (DensP * VelhP - DensM * VelhM - alpha * (P.DD - M.DD));
  df[VT] = h * (RhotP * VelhP - RhotM * VelhM - alpha * (P.DT - M.DT));
  df[VW] = h * (P.MZ * VelP - M.MZ * VelM - alpha * (P.MZ - M.MZ));
  const float dp = P.DP - M.DP;
  df[VU] = h * (P.MX * VelP - M.MX * VelM + dp * nx - alpha * (P.MX - M.MX));
  df[VV] = h * (P.MY * VelP - M.MY * VelM + dp * ny - alpha * (P.MY - M.MY));
}

// ---------------------------------------------------------------------------
// p = 7 fused stage kernel, one node per thread: one element per 512-thread
// workgroup, wavefront k holding the horizontal node plane k.  The arithmetic
// of every node is that of heve7_body, operation for operation; the layout
// keeps fewer values per thread (at most 128 VGPRs), so that four wavefronts
// share each SIMD (two elements per CU) and hide each other's LDS and memory
// latency.
//  - face work: wavefront 0 evaluates the bottom face and wavefront 7 the top
//    face, whose interior nodes are their own; wavefronts 1, 2, 3, 4 the
//    lateral faces y-, x+, y+, x- (interior and exterior sides loaded), so
//    that wavefronts w and w + 4, which usually share a SIMD, hold one
//    lateral face between them;
//  - the z rows D(k,:), the VFilterPM1 column and the vertical Lift entries
//    are wavefront-uniform; D rows i and j and the lateral Lift entries are
//    lane-varying;
//  - the pointwise fluxes of two variables go to LDS between barriers (three
//    barriers per stage), the contractions, lift, tendency and RK update of
//    each variable follow, the next variable's RK operands loaded ahead;
//  - the RK operands are loaded and the stage's outputs stored with the
//    non-temporal hint (streamed once), so that the state and DPRES lines the
//    neighbouring elements gather stay in the caches.
template <typename R, bool LVL, bool HEVI = false>
__device__ __forceinline__ void heve7_body(const DevMesh& m, const Consts& c, const TendIn& in, const RkOut& rk,
                                            const int ke) {
  constexpr int N1 = 8, Np = 512, Nfp = 64, NfpTot = 384;
  constexpr bool F64 = std::is_same<R, double>::value;
  __shared__ R sDt[N1 * N1];
  __shared__ R sI[N1 * N1];
  __shared__ R sF[4][3][Np];
  __shared__ R sDF[5][NfpTot];
  __shared__ R sRho[Np];
  __shared__ double sPow[F64 ? 640 : 1];

  const int t = threadIdx.x;
  const int i = t & 7, j = (t >> 3) & 7;
  const int k = __builtin_amdgcn_readfirstlane(t >> 6);  // wavefront index = node plane
  const int p = t;                                        // node i + N1 j + Nfp k
  const uint32_t base = (uint32_t)Np * ke;
  const uint32_t o = base + p;
  const R gamm = (R)(c.CPdry / c.CVdry);
  if (t < N1 * N1) {
    sDt[(t % N1) * N1 + t / N1] = (R)m.D1D[t];
    sI[t] = (R)m.Itr[t];
  }
  if constexpr (F64) pow_tables_to(sPow, t, 512);

  // ---- own state ----
  double q[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) q[v] = ldg(in.q[v], o);
  [[maybe_unused]] R rf[5];
  if constexpr (!F64) {
#pragma unroll
    for (int v = 0; v < 5; ++v) rf[v] = (R)q[v];
  }
  R* const r = [&]() -> R* {
    if constexpr (F64) return q;
    else return rf;
  }();
  const int ek0 = ke / (m.NeX * m.NeY);
  auto hyd = [&](const double* fld, int fi, uint32_t g, int lev, int node) -> double {
    if constexpr (LVL) return ldg(in.lvl, (uint32_t)((fi * m.NeZ + lev) * Np + node));
    else return ldg(fld, g);
  };
  const R Dh = (R)hyd(in.DENS_hyd, 0, o, ek0, p);
  const R Th = (R)hyd(in.THERM_hyd, 2, o, ek0, p);
  const R DP = (R)ldg(in.DPRES, o);

  // ---- this wavefront's face: the interior (M) and exterior (P) sides ----
  const int ei = ke % m.NeX, ej = (ke / m.NeX) % m.NeY, ek = ke / (m.NeX * m.NeY);
  const int lane = t & 63;
  // wavefront -> face: 0 bottom (4), 7 top (5), 1/2/3/4 lateral y-/x+/y+/x- (0..3), 5/6 none
  const int f = (k == 0) ? 4 : (k == 7) ? 5 : (k >= 1 && k <= 4) ? k - 1 : -1;
  R Mv[9], Pv[9];
  int bc = FEDG_BND_NONE;
  int fp = 0;  // face node slot of sDF
  if (f >= 0 && f < 4) {
    const int pm = face_node<N1>(f, lane);
    uint32_t iP;
    int nodeP;
    int ni = ei, nj = ej;
    bool closed = false;
    switch (f) {
      case 0: nj -= 1; if (nj < 0) { if (m.periodic[1]) nj = m.NeY - 1; else closed = true; } break;
      case 1: ni += 1; if (ni >= m.NeX) { if (m.periodic[0]) ni = 0; else closed = true; } break;
      case 2: nj += 1; if (nj >= m.NeY) { if (m.periodic[1]) nj = 0; else closed = true; } break;
      default: ni -= 1; if (ni < 0) { if (m.periodic[0]) ni = m.NeX - 1; else closed = true; } break;
    }
    constexpr int opp[4] = {2, 3, 0, 1};
    if (closed) {
      iP = base + pm;
      nodeP = pm;
      bc = m.bnd_bc[f];
    } else {
      nodeP = face_node<N1>(opp[f], lane);
      iP = (uint32_t)Np * (ni + m.NeX * (nj + m.NeY * ek)) + nodeP;
    }
    const uint32_t iM = base + pm;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      Mv[v] = (R)ldg(in.q[v], iM);
      Pv[v] = (R)ldg(in.q[v], iP);
    }
    Mv[5] = (R)hyd(in.DENS_hyd, 0, iM, ek, pm); Mv[6] = (R)hyd(in.PRES_hyd, 1, iM, ek, pm);
    Mv[7] = (R)hyd(in.THERM_hyd, 2, iM, ek, pm);
    Mv[8] = (R)ldg(in.DPRES, iM);
    Pv[5] = (R)hyd(in.DENS_hyd, 0, iP, ek, nodeP); Pv[6] = (R)hyd(in.PRES_hyd, 1, iP, ek, nodeP);
    Pv[7] = (R)hyd(in.THERM_hyd, 2, iP, ek, nodeP);
    Pv[8] = (R)ldg(in.DPRES, iP);
    fp = f * Nfp + lane;
  } else if (f >= 4) {
    // the interior node is this thread's node (i, j, 0) or (i, j, 7)
    int nk = ek + (f == 4 ? -1 : 1);
    bool closed = false;
    if (nk < 0 || nk >= m.NeZ) {
      if (m.periodic[2]) nk = (nk + m.NeZ) % m.NeZ;
      else closed = true;
    }
    uint32_t iP;
    int nodeP = p, levP = ek;
    if (closed) {
      iP = o;
      bc = m.bnd_bc[f];
    } else {
      nodeP = i + N1 * j + Nfp * (f == 4 ? 7 : 0);
      levP = nk;
      iP = (uint32_t)Np * (ei + m.NeX * (ej + m.NeY * nk)) + (uint32_t)nodeP;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) Pv[v] = (R)ldg(in.q[v], iP);
    Pv[5] = (R)hyd(in.DENS_hyd, 0, iP, levP, nodeP); Pv[6] = (R)hyd(in.PRES_hyd, 1, iP, levP, nodeP);
    Pv[7] = (R)hyd(in.THERM_hyd, 2, iP, levP, nodeP);
    Pv[8] = (R)ldg(in.DPRES, iP);
#pragma unroll
    for (int v = 0; v < 5; ++v) Mv[v] = r[v];
    Mv[5] = Dh;
    Mv[6] = (R)hyd(in.PRES_hyd, 1, o, ek, p);
    Mv[7] = Th;
    Mv[8] = DP;
    fp = f * Nfp + i + N1 * j;
  }

  // VFilterPM1 input
  sRho[p] = r[VD];

  // ---- face flux of this wavefront's face ----
  if (f >= 0) {
    R df[5];
    if constexpr (F64) {
      Side M, P;
      const double nx = (f == 1) ? 1.0 : (f == 3) ? -1.0 : 0.0;
      const double ny = (f == 2) ? 1.0 : (f == 0) ? -1.0 : 0.0;
      const double nz = (f == 5) ? 1.0 : (f == 4) ? -1.0 : 0.0;
      M.Gsqrt = 1.0; M.G13 = 0.0; M.G23 = 0.0;
      P.Gsqrt = 1.0; P.G13 = 0.0; P.G23 = 0.0;
      M.DD = Mv[0]; M.DT = Mv[1]; M.MZ = Mv[2]; M.MX = Mv[3]; M.MY = Mv[4];
      M.Dhyd = Mv[5]; M.Phyd = Mv[6]; M.Thyd = Mv[7]; M.DP = Mv[8];
      P.DD = Pv[0]; P.DT = Pv[1]; P.MZ = Pv[2]; P.MX = Pv[3]; P.MY = Pv[4];
      P.Dhyd = Pv[5]; P.Phyd = Pv[6]; P.Thyd = Pv[7]; P.DP = Pv[8];
      if (bc != FEDG_BND_NONE)
        reflect_momentum(bc, 1.0, 1.0, 1.0, 0.0, 1.0, 0.0, 0.0, nx, ny, nz, M.MX, M.MY, M.MZ, P.MX, P.MY, P.MZ);
      if constexpr (HEVI) rusanov_hevi(M, P, nx, ny, nz, m.fF[f], gamm, df);
      else rusanov(M, P, nx, ny, nz, m.fF[f], gamm, df);
    } else {
      const float nx = (f == 1) ? 1.f : (f == 3) ? -1.f : 0.f;
      const float ny = (f == 2) ? 1.f : (f == 0) ? -1.f : 0.f;
      const float nz = (f == 5) ? 1.f : (f == 4) ? -1.f : 0.f;
      Side32 M, P;
      M.DD = Mv[0]; M.DT = Mv[1]; M.MZ = Mv[2]; M.MX = Mv[3]; M.MY = Mv[4];
      M.Dh = Mv[5]; M.Ph = Mv[6]; M.Th = Mv[7]; M.DP = Mv[8];
      P.DD = Pv[0]; P.DT = Pv[1]; P.MZ = Pv[2]; P.MX = Pv[3]; P.MY = Pv[4];
      P.Dh = Pv[5]; P.Ph = Pv[6]; P.Th = Pv[7]; P.DP = Pv[8];
      if (bc == FEDG_BND_SLIP) {
        const float mn = M.MX * nx + M.MY * ny + M.MZ * nz;
        P.MX = M.MX - 2.f * mn * nx;
        P.MY = M.MY - 2.f * mn * ny;
        P.MZ = M.MZ - 2.f * mn * nz;
      } else if (bc == FEDG_BND_NOSLIP) {
        P.MX = -M.MX; P.MY = -M.MY; P.MZ = -M.MZ;
      }
      if constexpr (HEVI) rusanov32_hevi(M, P, nx, ny, nz, (float)m.fF[f], gamm, df);
      else rusanov32(M, P, nx, ny, nz, (float)m.fF[f], gamm, df);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) sDF[v][fp] = df[v];
  }
  __builtin_amdgcn_sched_barrier(0);

  const int ke2d = ke % (m.NeX * m.NeY);
  const uint32_t o2d = (uint32_t)(i + N1 * j) + (uint32_t)Nfp * ke2d;
  double Dk[N1];  // D(k,:), wavefront-uniform
#pragma unroll
  for (int l = 0; l < N1; ++l) Dk[l] = m.D1D[k + N1 * l];
  const double* LT = m.liftT;
  const double Ly0 = LT[j + N1 * 0], Lx1 = LT[i + N1 * 1], Ly2 = LT[j + N1 * 2], Lx3 = LT[i + N1 * 3];
  const double Lz4 = LT[k + N1 * 4], Lz5 = LT[k + N1 * 5];

  const R E11 = (R)m.fE[0], E22 = (R)m.fE[1], E33 = (R)m.fE[2];
  const R RD = R(1) / (r[VD] + Dh);
  const R GsqrtV = R(1) / R(1);
  const R RGsqrtV = R(1) / GsqrtV;
  const R RGsqrt = R(1) / R(1);
  const R Gs = R(1), g13 = R(0), g23 = R(0);
  const R GRAV = (R)c.GRAV;

  auto flux = [&](int v, R& f0, R& f1, R& f2) {
    const R m0 = Gs * r[VU];
    const R m1 = Gs * r[VV];
    const R m2 = Gs * (r[VW] * RGsqrtV + g13 * r[VU] + g23 * r[VV]);
    if (v == VD) {
      f0 = m0; f1 = m1; f2 = m2;
    } else if (v == VT) {
      const R pt_ = (Th + r[VT]) * RD;
      f0 = m0 * pt_; f1 = m1 * pt_; f2 = m2 * pt_;
    } else if (v == VW) {
      const R w_ = r[VW] * RD;
      f0 = m0 * w_; f1 = m1 * w_;
      if constexpr (HEVI) f2 = m2 * w_;
      else f2 = m2 * w_ + Gs * DP * RGsqrtV;
    } else if (v == VU) {
      const R G = Gs * DP;
      const R u_ = r[VU] * RD;
      f0 = m0 * u_ + G; f1 = m1 * u_; f2 = m2 * u_ + G * g13;
    } else {
      const R G = Gs * DP;
      const R v_ = r[VV] * RD;
      f0 = m0 * v_; f1 = m1 * v_ + G; f2 = m2 * v_ + G * g23;
    }
  };

  struct Ops {
    double rA, rB, s, cor, Ph;
  };
  auto load_ops = [&](int v) {
    Ops op{0.0, 0.0, 0.0, 0.0, 0.0};
    if (rk.mode == 1) op.rA = ldg_nt(rk.var0[v], o);
    if (rk.mode == 1 || rk.mode == 2) op.rB = ldg_nt(rk.varTmp[v], o);
    if (v == VU) {
      op.s = hyd(in.DPhydDx, 3, o, ek0, p);
      op.cor = ldg(in.CORIOLIS, o2d);
    } else if (v == VV) {
      op.s = hyd(in.DPhydDy, 4, o, ek0, p);
      op.cor = ldg(in.CORIOLIS, o2d);
    } else if (v == VT) {
      op.Ph = hyd(in.PRES_hyd, 1, o, ek0, p);
    }
    return op;
  };
  Ops nxt = load_ops(0);
  R drho = R(0);
  auto wr = [&](int v) {
    R f0, f1, f2;
    flux(v, f0, f1, f2);
    const int b = v & 3;
    sF[b][0][p] = f0; sF[b][1][p] = f1; sF[b][2][p] = f2;
  };
  auto cu = [&](int v) {
    const Ops cur = nxt;
    if (v < 4) nxt = load_ops(v + 1);
    const R s = (R)cur.s, cor = (R)cur.cor;
    const int b = v & 3;
    if (v == 0) {
      R zl[N1];
#pragma unroll
      for (int l = 0; l < N1; ++l) zl[l] = sRho[i + N1 * (j + N1 * l)];
      drho = zl[0] * sI[N1 * k];
#pragma unroll
      for (int l = 1; l < N1; ++l) drho = drho + zl[l] * sI[l + N1 * k];
    }
    const R* Fx = sF[b][0];
    const R* Fy = sF[b][1];
    const R* Fz = sF[b][2];
    R dx, dy, dz;
    {
      R Di[N1];
#pragma unroll
      for (int l = 0; l < N1; ++l) Di[l] = sDt[l + N1 * i];
      dx = Di[0] * Fx[N1 * (j + N1 * k)];
#pragma unroll
      for (int l = 1; l < N1; ++l) dx = dx + Di[l] * Fx[l + N1 * (j + N1 * k)];
    }
    {
      R Dj[N1];
#pragma unroll
      for (int l = 0; l < N1; ++l) Dj[l] = sDt[l + N1 * j];
      dy = Fy[i + N1 * (N1 * k)] * Dj[0];
#pragma unroll
      for (int l = 1; l < N1; ++l) dy = dy + Fy[i + N1 * (l + N1 * k)] * Dj[l];
    }
    {
      R zl[N1];
#pragma unroll
      for (int l = 0; l < N1; ++l) zl[l] = Fz[i + N1 * (j + N1 * l)];
      dz = zl[0] * Dk[0];
#pragma unroll
      for (int l = 1; l < N1; ++l) dz = dz + zl[l] * Dk[l];
    }
    const R* dfv = sDF[v];
    const R fz4 = dfv[4 * Nfp + i + N1 * j], fz5 = dfv[5 * Nfp + i + N1 * j];
    const R lift = Ly0 * dfv[i + N1 * k] + Lx1 * dfv[Nfp + j + N1 * k] + Ly2 * dfv[2 * Nfp + i + N1 * k] +
                   Lx3 * dfv[3 * Nfp + j + N1 * k] + Lz4 * fz4 + Lz5 * fz5;
    const bool horiz = HEVI && (v == VD || v == VT);
    const R X = horiz ? E11 * dx + E22 * dy + lift : E11 * dx + E22 * dy + E33 * dz + lift;
    R T;
    if (v == VD || v == VT) {
      T = -X * RGsqrt;
    } else if (v == VW) {
      if constexpr (HEVI) T = -X * RGsqrt;
      else T = -X * RGsqrt - GRAV * drho;
    } else if (v == VU) {
      T = -s + cor * r[VV];
      T = T - X * RGsqrt;
    } else {
      T = -s - cor * r[VU];
      T = T - X * RGsqrt;
    }
    const double TD = (double)T;
    if (rk.mode == 4) {
      stg_nt(rk.out[v], o, TD);
      return;
    }
    double nq;
    switch (rk.mode) {
      case 0:
        stg_nt(rk.varTmp[v], o, q[v] + rk.cb * TD);
        nq = q[v] + rk.ca * TD;
        break;
      case 1:
        stg_nt(rk.varTmp[v], o, cur.rB + rk.cb * TD);
        nq = cur.rA + rk.ca * TD;
        break;
      case 2:
        nq = cur.rB + rk.cb * TD;
        break;
      default:
        nq = q[v] + rk.cb * TD;
        break;
    }
    stg_nt(rk.out[v], o, nq);
    if (v == VT) {
      if constexpr (F64) {
        if (in.Rtot != nullptr)
          stg_nt(rk.DPout, o, dpres_of_tab(nq, Th, cur.Ph, ldg(in.Rtot, o), ldg(in.CVtot, o), ldg(in.CPtot, o), c, sPow));
        else
          stg_nt(rk.DPout, o, dpres_of_tab(nq, Th, cur.Ph, c.Rdry, c.CVdry, c.CPdry, c, sPow));
      } else {
        stg_nt(rk.DPout, o, (double)dpres32((float)nq, Th, (float)cur.Ph, gamm));
      }
    }
  };
  // the fluxes of two variables per barrier in four LDS buffers: variable 4
  // reuses buffer 0, whose last reads (variable 0) precede the second barrier
  wr(0); wr(1);
  __syncthreads();
  cu(0);
  __builtin_amdgcn_sched_barrier(0);
  cu(1);
  wr(2); wr(3);
  __syncthreads();
  cu(2);
  __builtin_amdgcn_sched_barrier(0);
  cu(3);
  wr(4);
  __syncthreads();
  cu(4);
}

// 512 threads, at most 128 VGPRs: two workgroups (16 wavefronts) per CU
#define FEDG_HEVE7_BOUNDS __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
// at most 80 VGPRs: three workgroups per CU (the FP32 HEVI tendency)
#define FEDG_HEVE7_BOUNDS3 __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6)))
__global__ void FEDG_HEVE7_BOUNDS heve7_flat_kernel(const DevMesh m, const Consts c, const TendIn in, const RkOut rk) {
  heve7_body<double, false>(m, c, in, rk, xcd_remap(blockIdx.x, gridDim.x));
}

// base state and its horizontal gradient read from per-level tables
// (horizontally homogeneous base state): the same values, a fraction of the
// bytes (the tables stay in L2)
__global__ void FEDG_HEVE7_BOUNDS heve7_lvl_kernel(const DevMesh m, const Consts c, const TendIn in, const RkOut rk) {
  heve7_body<double, true>(m, c, in, rk, xcd_remap(blockIdx.x, gridDim.x));
}

// The same p = 7 stage with the tendency in FP32 about the FP64 base state
// (perturbation form; state, base state and RK update FP64; the pressure
// perturbation of the new state in FP32 as dpres32).
__global__ void FEDG_HEVE7_BOUNDS heve7_f32_kernel(const DevMesh m, const Consts c, const TendIn in, const RkOut rk) {
  heve7_body<float, false>(m, c, in, rk, xcd_remap(blockIdx.x, gridDim.x));
}
__global__ void FEDG_HEVE7_BOUNDS heve7_f32_lvl_kernel(const DevMesh m, const Consts c, const TendIn in,
                                                       const RkOut rk) {
  heve7_body<float, true>(m, c, in, rk, xcd_remap(blockIdx.x, gridDim.x));
}

// The explicit HEVI tendency on the same layout (rk.mode 4): FP64, or FP32 in
// perturbation form about the FP64 base state; LVL: base state from the
// per-level tables.
template <bool LVL>
__global__ void FEDG_HEVE7_BOUNDS heve7_hevi_kernel(const DevMesh m, const Consts c, const TendIn in, const RkOut rk) {
  heve7_body<double, LVL, true>(m, c, in, rk, xcd_remap(blockIdx.x, gridDim.x));
}
template <bool LVL>
__global__ void FEDG_HEVE7_BOUNDS3 heve7_hevi32_kernel(const DevMesh m, const Consts c, const TendIn in,
                                                       const RkOut rk) {
  heve7_body<float, LVL, true>(m, c, in, rk, xcd_remap(blockIdx.x, gridDim.x));
}

template <int N1, int EPB>
__global__ void __launch_bounds__(EPB* N1* N1* N1, 2) heve32_kernel(const DevMesh m, const Consts c, const TendIn in,
                                                                    const RkOut rk) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1, NfpTot = 6 * Nfp;
  __shared__ float sDt[N1 * N1];
  __shared__ float sI[N1 * N1];
  __shared__ float sS[EPB][S_N][Np];
  __shared__ float sF[2][EPB][3][Np];
  __shared__ float sDF[EPB][5][NfpTot];

  const int tid = threadIdx.x;
  const int e = tid / Np;
  const int p = tid - e * Np;
  const int ke = xcd_remap(blockIdx.x, gridDim.x) * EPB + e;
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / Nfp;
  const float gamm = (float)(c.CPdry / c.CVdry);
  for (int t = tid; t < N1 * N1; t += blockDim.x) {
    sDt[(t % N1) * N1 + t / N1] = (float)m.D1D[t];
    sI[t] = (float)m.Itr[t];
  }
  const uint32_t o = (uint32_t)Np * (valid ? ke : 0) + p;
  double q[5] = {0, 0, 0, 0, 0};
  float qf[5] = {0, 0, 0, 0, 0};
  float Dh = 1.f, Ph = 1.f, Th = 1.f, DP = 0.f, cor = 0.f, dpx = 0.f, dpy = 0.f;
  if (valid) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      q[v] = ldg(in.q[v], o);
      qf[v] = (float)q[v];
    }
    Dh = (float)ldg(in.DENS_hyd, o);
    Ph = (float)ldg(in.PRES_hyd, o);
    Th = (float)ldg(in.THERM_hyd, o);
    DP = dpres32(qf[VT], Th, Ph, gamm);
    cor = (float)ldg(in.CORIOLIS, (uint32_t)(i + N1 * j) + (uint32_t)Nfp * (ke % (m.NeX * m.NeY)));
    dpx = (float)ldg(in.DPhydDx, o);
    dpy = (float)ldg(in.DPhydDy, o);
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) sS[e][v][p] = qf[v];
  sS[e][S_DH][p] = Dh;
  sS[e][S_PH][p] = Ph;
  sS[e][S_TH][p] = Th;
  sS[e][S_DP][p] = DP;
  float L[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) L[f] = (float)ldg(m.Lift, (uint32_t)(p + f * Np));
  __syncthreads();

  // face phase (flat box: arithmetic neighbours, closed boundaries reflect)
  if (valid) {
    const int ei = ke % m.NeX, ej = (ke / m.NeX) % m.NeY, ek = ke / (m.NeX * m.NeY);
    for (int fp = p; fp < NfpTot; fp += Np) {
      const int f = fp / Nfp, a = fp - f * Nfp;
      const int pm = face_node<N1>(f, a);
      int ni = ei, nj = ej, nk = ek;
      bool closed = false;
      switch (f) {
        case 0: nj -= 1; if (nj < 0) { if (m.periodic[1]) nj = m.NeY - 1; else closed = true; } break;
        case 1: ni += 1; if (ni >= m.NeX) { if (m.periodic[0]) ni = 0; else closed = true; } break;
        case 2: nj += 1; if (nj >= m.NeY) { if (m.periodic[1]) nj = 0; else closed = true; } break;
        case 3: ni -= 1; if (ni < 0) { if (m.periodic[0]) ni = m.NeX - 1; else closed = true; } break;
        case 4: nk -= 1; if (nk < 0) { if (m.periodic[2]) nk = m.NeZ - 1; else closed = true; } break;
        default: nk += 1; if (nk >= m.NeZ) { if (m.periodic[2]) nk = 0; else closed = true; } break;
      }
      const float nx = (f == 1) ? 1.f : (f == 3) ? -1.f : 0.f;
      const float ny = (f == 2) ? 1.f : (f == 0) ? -1.f : 0.f;
      const float nz = (f == 5) ? 1.f : (f == 4) ? -1.f : 0.f;
      Side32 M, P;
      M.DD = sS[e][VD][pm]; M.DT = sS[e][VT][pm]; M.MZ = sS[e][VW][pm]; M.MX = sS[e][VU][pm];
      M.MY = sS[e][VV][pm]; M.Dh = sS[e][S_DH][pm]; M.Ph = sS[e][S_PH][pm]; M.Th = sS[e][S_TH][pm];
      M.DP = sS[e][S_DP][pm];
      if (closed) {
        P = M;
        const int bc = m.bnd_bc[f];
        if (bc == FEDG_BND_SLIP) {
          const float mn = M.MX * nx + M.MY * ny + M.MZ * nz;
          P.MX = M.MX - 2.f * mn * nx;
          P.MY = M.MY - 2.f * mn * ny;
          P.MZ = M.MZ - 2.f * mn * nz;
        } else if (bc == FEDG_BND_NOSLIP) {
          P.MX = -M.MX; P.MY = -M.MY; P.MZ = -M.MZ;
        }
      } else {
        constexpr int opp[6] = {2, 3, 0, 1, 5, 4};
        const uint32_t iP = (uint32_t)Np * (ni + m.NeX * (nj + m.NeY * nk)) + face_node<N1>(opp[f], a);
        P.DD = (float)ldg(in.q[VD], iP); P.DT = (float)ldg(in.q[VT], iP); P.MZ = (float)ldg(in.q[VW], iP);
        P.MX = (float)ldg(in.q[VU], iP); P.MY = (float)ldg(in.q[VV], iP);
        P.Dh = (float)ldg(in.DENS_hyd, iP); P.Ph = (float)ldg(in.PRES_hyd, iP); P.Th = (float)ldg(in.THERM_hyd, iP);
        P.DP = dpres32(P.DT, P.Th, P.Ph, gamm);
      }
      float df[5];
      rusanov32(M, P, nx, ny, nz, (float)m.fF[f], gamm, df);
#pragma unroll
      for (int v = 0; v < 5; ++v) sDF[e][v][fp] = df[v];
    }
  }

  // RK operands (FP64)
  double rkA[5] = {0, 0, 0, 0, 0}, rkB[5] = {0, 0, 0, 0, 0};
  if (valid) {
    if (rk.mode == 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        rkA[v] = ldg(rk.var0[v], o);
        rkB[v] = ldg(rk.varTmp[v], o);
      }
    } else if (rk.mode == 2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) rkB[v] = ldg(rk.varTmp[v], o);
    }
  }
  const float RD = 1.f / (qf[VD] + Dh);
  const float E11 = (float)m.fE[0], E22 = (float)m.fE[1], E33 = (float)m.fE[2];
  const float grav = (float)c.GRAV;
  float drho = sS[e][VD][i + N1 * j] * sI[N1 * k];
#pragma unroll
  for (int l = 1; l < N1; ++l) drho += sS[e][VD][i + N1 * (j + N1 * l)] * sI[l + N1 * k];

#pragma unroll
  for (int v = 0; v < 5; ++v) {
    float f0, f1, f2;
    if (v == VD) {
      f0 = qf[VU]; f1 = qf[VV]; f2 = qf[VW];
    } else if (v == VT) {
      const float pt = (Th + qf[VT]) * RD;
      f0 = qf[VU] * pt; f1 = qf[VV] * pt; f2 = qf[VW] * pt;
    } else {
      const float u = qf[v] * RD;
      f0 = qf[VU] * u + (v == VU ? DP : 0.f);
      f1 = qf[VV] * u + (v == VV ? DP : 0.f);
      f2 = qf[VW] * u + (v == VW ? DP : 0.f);
    }
    const int b = v & 1;
    sF[b][e][0][p] = f0;
    sF[b][e][1][p] = f1;
    sF[b][e][2][p] = f2;
    __syncthreads();
    const float* Fx = sF[b][e][0];
    const float* Fy = sF[b][e][1];
    const float* Fz = sF[b][e][2];
    float dx = 0.f, dy = 0.f, dz = 0.f;
#pragma unroll
    for (int l = 0; l < N1; ++l) dx += sDt[l + N1 * i] * Fx[l + N1 * (j + N1 * k)];
#pragma unroll
    for (int l = 0; l < N1; ++l) dy += Fy[i + N1 * (l + N1 * k)] * sDt[l + N1 * j];
#pragma unroll
    for (int l = 0; l < N1; ++l) dz += Fz[i + N1 * (j + N1 * l)] * sDt[l + N1 * k];
    const float* dfv = sDF[e][v];
    const float lift = L[0] * dfv[i + N1 * k] + L[1] * dfv[Nfp + j + N1 * k] + L[2] * dfv[2 * Nfp + i + N1 * k] +
                       L[3] * dfv[3 * Nfp + j + N1 * k] + L[4] * dfv[4 * Nfp + i + N1 * j] +
                       L[5] * dfv[5 * Nfp + i + N1 * j];
    float T = -(E11 * dx + E22 * dy + E33 * dz + lift);
    if (v == VW) T -= grav * drho;
    else if (v == VU) T += -dpx + cor * qf[VV];
    else if (v == VV) T += -dpy - cor * qf[VU];
    const double Td = (double)T;
    if (valid) {
      double qn;
      switch (rk.mode) {
        case 0:
          stg(rk.varTmp[v], o, q[v] + rk.cb * Td);
          qn = q[v] + rk.ca * Td;
          break;
        case 1:
          stg(rk.varTmp[v], o, rkB[v] + rk.cb * Td);
          qn = rkA[v] + rk.ca * Td;
          break;
        case 2:
          qn = rkB[v] + rk.cb * Td;
          break;
        default:
          qn = q[v] + rk.cb * Td;
          break;
      }
      stg(rk.out[v], o, qn);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Face contribution of the drop-in tendency
// (FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_heve_numflux.F90:242-496,
// add_bnd_contrib_generalvc): the Rusanov flux jump at every face node from
// vmapM / vmapP (boundary faces read the halo tail), its lift (Lift_var5,
// ELM/scale_element_operation_tensorprod3D.F90:2810-2823) and
// dt -= lift / Gsqrt.  One thread per node; the face nodes are shared out
// over the element's threads.
template <int N1, int EPB>
__global__ void __launch_bounds__(EPB * N1 * N1 * N1) heve_bnd_kernel(const DevMesh m, const Consts c,
                                                                     const TendIn in, const RkOut rk) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1, NfpTot = 6 * Nfp;
  __shared__ double sDF[EPB][5][NfpTot];
  const int tid = threadIdx.x;
  const int e = tid / Np;
  const int p = tid - e * Np;
  const int ke = blockIdx.x * EPB + e;
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / Nfp;
  const double gamm = c.CPdry / c.CVdry;
  auto side = [&](uint32_t n, Side& S) {
    S.Gsqrt = ldg(m.Gsqrt, n); S.G13 = ldg(m.G13, n); S.G23 = ldg(m.G23, n);
    S.DD = ldg(in.q[VD], n); S.DT = ldg(in.q[VT], n); S.MZ = ldg(in.q[VW], n);
    S.MX = ldg(in.q[VU], n); S.MY = ldg(in.q[VV], n);
    S.Dhyd = ldg(in.DENS_hyd, n); S.Phyd = ldg(in.PRES_hyd, n); S.Thyd = ldg(in.THERM_hyd, n);
    S.DP = ldg(in.DPRES, n);
  };
  if (valid) {
    for (int fp = p; fp < NfpTot; fp += Np) {
      const uint32_t gf = (uint32_t)fp + (uint32_t)NfpTot * ke;
      Side M, P;
      side((uint32_t)m.vmapM[gf], M);
      side((uint32_t)m.vmapP[gf], P);
      double df[5];
      rusanov<true>(M, P, ldg(m.nx, gf), ldg(m.ny, gf), ldg(m.nz, gf), ldg(m.Fscale, gf), gamm, df);
#pragma unroll
      for (int v = 0; v < 5; ++v) sDF[e][v][fp] = df[v];
    }
  }
  __syncthreads();
  if (!valid) return;
  const uint32_t o = (uint32_t)Np * ke + p;
  double L[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) L[f] = ldg(m.Lift, (uint32_t)(p + f * Np));
  const double RGsqrt = 1.0 / ldg(m.Gsqrt, o);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double* dfv = sDF[e][v];
    const double lift = L[0] * dfv[i + N1 * k] + L[1] * dfv[Nfp + j + N1 * k] + L[2] * dfv[2 * Nfp + i + N1 * k] +
                        L[3] * dfv[3 * Nfp + j + N1 * k] + L[4] * dfv[4 * Nfp + i + N1 * j] +
                        L[5] * dfv[5 * Nfp + i + N1 * j];
    stg(rk.out[v], o, ldg(rk.out[v], o) - RGsqrt * lift);
  }
}

// Tracer advection tendency, HEVE (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:149-231
// with its face flux get_delflux_generalhvc, :554-674): central flux of the
// mass flux with the density's dissipation coefficients alphDENS_M/P, the
// FCT coefficient of the upwind side chosen by the sign of the net outward
// flux of each face (FaceIntMat of the LGL face weights, :75-137), Div with
// the lift, then QTRC_dt = -(...)/Gsqrt + RHOQ_tp.  One thread per node.
struct TrcArgs {
  double* dt;                       // QTRC_dt (Np,NeA); fct_coef for the FCT kernel
  const double *Q, *MX, *MY, *MZ;   // QTRC_ and the mass fluxes (Np,NeA)
  const double *aM, *aP;            // alphDENS_M / _P (NfpTot,Ne)
  const double* R;                  // fct_coef (Np,NeA)
  const double* tp;                 // RHOQ_tp (Np,NeA)
  const double* J;                  // (Np,Ne)
  double w[8];                      // 1D LGL integration weights
  // FCT coefficient only: density of the stage s-1, DENS_hyd + (1-c) DDENS0 + c DDENS
  const double *Dh, *DD, *DD0;
  double c, dtt;
  int disable;                      // disable_limiter: fct_coef = 1
};
enum { TRC_TEND = 0, TRC_FCT = 1 };
// KIND = TRC_TEND: the tendency of cal_tend.  KIND = TRC_FCT: the FCT
// coefficient of calc_fct_coef (:232-305): the sum of the positive net
// outward fluxes of the six faces (get_netOutwardFlux_generalhvc, :676-797,
// the same face flux), against the tracer mass available over dt,
// sum(Gsqrt J IntWeight_lgl (dens_ssm1 QTRC / dt + RHOQ_tp)), summed node by
// node in order by one thread.
// FLAT: the flat-terrain uniform box the context detected (Gsqrt = GsqrtH = 1,
// GI3 = 0, axis-aligned unit normals, Escale and Fscale constant per
// direction and face): the same expressions with those exact constants in
// place of the loads.
template <int N1, int EPB, int KIND, bool FLAT>
__global__ void __launch_bounds__(EPB * N1 * N1 * N1) trcadv_kernel(const DevMesh m, const TrcArgs a) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1, NfpTot = 6 * Nfp;
  constexpr bool TEND = KIND == TRC_TEND;
  constexpr int EF = TEND ? EPB : 1;  // arrays only the tendency uses
  constexpr int NFT = (NfpTot + Np - 1) / Np;  // face nodes per thread
  __shared__ double sDt[N1 * N1];
  __shared__ double sF[EF][3][TEND ? Np : 1];
  __shared__ double sDF[EF][TEND ? NfpTot : 1];  // del_flux
  // numflux, QTRC_M MomFlxM and the two fct_coef of this thread's face nodes
  double rNF[NFT], rQF[NFT], rRP[NFT], rRM[NFT];
  __shared__ double sJF[EPB][NfpTot];            // J(iM) * Fscale * numflux
  __shared__ double sOut[EPB][6];                // net outward flux of each face
  __shared__ double sM[TEND ? 1 : EPB][TEND ? 1 : Np];  // FCT: mass terms of the nodes
  __shared__ double sC[EPB];                            // FCT: the element's coefficient
  const int tid = threadIdx.x;
  const int e = tid / Np;
  const int p = tid - e * Np;
  const int ke = xcd_remap(blockIdx.x, gridDim.x) * EPB + e;  // x-neighbours on one XCD's L2
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / Nfp;
  const uint32_t o = (uint32_t)Np * (valid ? ke : 0) + p;
  if constexpr (!TEND) {
    if (a.disable) {  // uniform over the grid
      if (valid) stg(a.dt, o, 1.0);
      return;
    }
  }
  if constexpr (TEND)
    for (int t = tid; t < N1 * N1; t += blockDim.x) sDt[(t % N1) * N1 + t / N1] = m.D1D[t];
  const int ke2d = valid ? m.EMap3Dto2D[ke] : 0;
  if (valid) {
#pragma unroll
    for (int u = 0; u < NFT; ++u) {
      const int fp = p + u * Np;
      if (fp >= NfpTot) break;
      const uint32_t gf = (uint32_t)fp + (uint32_t)NfpTot * ke;
      const uint32_t iM = (uint32_t)m.vmapM[gf], iP = (uint32_t)m.vmapP[gf];
      const int f = fp / Nfp, an = fp - f * Nfp;
      const double GsH = FLAT ? 1.0 : ldg(m.GsqrtH, (uint32_t)(face_node<N1>(f, an) % Nfp) + (uint32_t)Nfp * ke2d);
      const double Gsqrt_M = FLAT ? 1.0 : ldg(m.Gsqrt, iM), Gsqrt_P = FLAT ? 1.0 : ldg(m.Gsqrt, iP);
      const double GsqrtV_M = Gsqrt_M / GsH, GsqrtV_P = Gsqrt_P / GsH;
      const double G13_M = FLAT ? 0.0 : ldg(m.G13, iM), G13_P = FLAT ? 0.0 : ldg(m.G13, iP);
      const double G23_M = FLAT ? 0.0 : ldg(m.G23, iM), G23_P = FLAT ? 0.0 : ldg(m.G23, iP);
      const double Q_M = ldg(a.Q, iM), Q_P = ldg(a.Q, iP);
      const double GMX_M = Gsqrt_M * ldg(a.MX, iM), GMX_P = Gsqrt_P * ldg(a.MX, iP);
      const double GMY_M = Gsqrt_M * ldg(a.MY, iM), GMY_P = Gsqrt_P * ldg(a.MY, iP);
      const double GMZ_M = Gsqrt_M * ldg(a.MZ, iM), GMZ_P = Gsqrt_P * ldg(a.MZ, iP);
      const double nx = FLAT ? (f == 1 ? 1.0 : f == 3 ? -1.0 : 0.0) : ldg(m.nx, gf);
      const double ny = FLAT ? (f == 2 ? 1.0 : f == 0 ? -1.0 : 0.0) : ldg(m.ny, gf);
      const double nz = FLAT ? (f == 5 ? 1.0 : f == 4 ? -1.0 : 0.0) : ldg(m.nz, gf);
      const double MF_M = GMX_M * nx + GMY_M * ny + ((GMZ_M / GsqrtV_M + G13_M * GMX_M + G23_M * GMY_M) * nz);
      const double MF_P = GMX_P * nx + GMY_P * ny + ((GMZ_P / GsqrtV_P + G13_P * GMX_P + G23_P * GMY_P) * nz);
      const double nf = 0.5 * ((Q_P * MF_P + Q_M * MF_M) - ldg(a.aP, gf) * Q_P + ldg(a.aM, gf) * Q_M);
      if constexpr (TEND) {
        rNF[u] = nf;
        rQF[u] = Q_M * MF_M;
        rRP[u] = ldg(a.R, iP);
        rRM[u] = ldg(a.R, iM);
      }
      sJF[e][fp] = ldg(a.J, iM) * (FLAT ? m.fF[f] : ldg(m.Fscale, gf)) * nf;
    }
  }
  // volume flux, formed before the face phase's barrier so that its loads
  // overlap the face gathers (sF is not read before the third barrier)
  double Gs = 1.0, tpo = 0.0;
  if constexpr (TEND) {
    if (valid) {
      Gs = FLAT ? 1.0 : ldg(m.Gsqrt, o);
      tpo = ldg(a.tp, o);
      const double RGsqrt = 1.0 / Gs;
      const double RGsqrtV = (FLAT ? 1.0 : ldg(m.GsqrtH, (uint32_t)(i + N1 * j) + (uint32_t)Nfp * ke2d)) * RGsqrt;
      const double Qo = ldg(a.Q, o), mx = ldg(a.MX, o), my = ldg(a.MY, o), mz = ldg(a.MZ, o);
      const double G13 = FLAT ? 0.0 : ldg(m.G13, o), G23 = FLAT ? 0.0 : ldg(m.G23, o);
      sF[e][0][p] = Gs * mx * Qo;
      sF[e][1][p] = Gs * my * Qo;
      sF[e][2][p] = Gs * (mz * RGsqrtV + G13 * mx + G23 * my) * Qo;
    }
  }
  if constexpr (!TEND) {
    if (valid) {
      const double IW = (a.w[i] * a.w[j]) * a.w[k];
      const double dens = ldg(a.Dh, o) + (1.0 - a.c) * ldg(a.DD0, o) + a.c * ldg(a.DD, o);
      sM[e][p] = (FLAT ? 1.0 : ldg(m.Gsqrt, o)) * ldg(a.J, o) * IW * (dens * ldg(a.Q, o) / a.dtt + ldg(a.tp, o));
    }
  }
  __syncthreads();
  // FaceIntMat x (J Fscale numflux): row f sums its face's nodes in order from 0
  if (valid && p < 6) {
    double s = 0.0;
#pragma unroll 8
    for (int an = 0; an < Nfp; ++an) s = s + (a.w[an % N1] * a.w[an / N1]) * sJF[e][p * Nfp + an];
    sOut[e][p] = s;
  }
  if constexpr (!TEND) {
    __syncthreads();
    if (valid && p == 0) {
      double net = 0.0;
      for (int f = 0; f < 6; ++f) net = net + (sOut[e][f] > 0.0 ? sOut[e][f] : 0.0);
      double q = 0.0;
#pragma unroll 16
      for (int n = 0; n < Np; ++n) q = q + sM[e][n];
      double r = q / (net + 1.0e-10);
      r = r < 1.0 ? r : 1.0;
      sC[e] = r > 0.0 ? r : 0.0;
    }
    __syncthreads();
    if (valid) stg(a.dt, o, sC[e]);
    return;
  } else {
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int u = 0; u < NFT; ++u) {
        const int fp = p + u * Np;
        if (fp >= NfpTot) break;
        const uint32_t gf = (uint32_t)fp + (uint32_t)NfpTot * ke;
        const double RP = rRP[u], RM = rRM[u];
        const double sg = copysign(1.0, sOut[e][fp / Nfp]);
        sDF[e][fp] = (FLAT ? m.fF[fp / Nfp] : ldg(m.Fscale, gf)) * (rNF[u] * 0.5 * (RP + RM - (RP - RM) * sg) - rQF[u]);
      }
    }
    __syncthreads();
    if (!valid) return;
    const double* Fx = sF[e][0];
    const double* Fy = sF[e][1];
    const double* Fz = sF[e][2];
    double dx = sDt[N1 * i] * Fx[N1 * (j + N1 * k)];
#pragma unroll
    for (int l = 1; l < N1; ++l) dx = dx + sDt[l + N1 * i] * Fx[l + N1 * (j + N1 * k)];
    double dy = Fy[i + N1 * (N1 * k)] * sDt[N1 * j];
#pragma unroll
    for (int l = 1; l < N1; ++l) dy = dy + Fy[i + N1 * (l + N1 * k)] * sDt[l + N1 * j];
    double dz = Fz[i + N1 * j] * sDt[N1 * k];
#pragma unroll
    for (int l = 1; l < N1; ++l) dz = dz + Fz[i + N1 * (j + N1 * l)] * sDt[l + N1 * k];
    double L[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) L[f] = ldg(m.Lift, (uint32_t)(p + f * Np));
    const double* dfv = sDF[e];
    const double lift = L[0] * dfv[i + N1 * k] + L[1] * dfv[Nfp + j + N1 * k] + L[2] * dfv[2 * Nfp + i + N1 * k] +
                        L[3] * dfv[3 * Nfp + j + N1 * k] + L[4] * dfv[4 * Nfp + i + N1 * j] +
                        L[5] * dfv[5 * Nfp + i + N1 * j];
    const double E11 = FLAT ? m.fE[0] : ldg(m.E11, o), E22 = FLAT ? m.fE[1] : ldg(m.E22, o);
    const double E33 = FLAT ? m.fE[2] : ldg(m.E33, o);
    const double X = E11 * dx + E22 * dy + E33 * dz + lift;
    stg(a.dt, o, -X / Gs + tpo);
  }
}

// TMAR, the second step of the limiter (:308-337): negative values cut to
// zero, then the element's tracer mass restored by one factor,
// QTRC = Q0 / (Q1 + 1e-32) * max(0, QTRC), the two masses summed node by node
// in order by one thread.
template <int N1, int EPB>
__global__ void __launch_bounds__(EPB * N1 * N1 * N1)
    tmar_kernel(const DevMesh m, double* Q, const double* Dh, const double* DD, const double* J, TrcW w) {
  constexpr int Np = N1 * N1 * N1;
  __shared__ double s0[EPB][Np], s1[EPB][Np];
  __shared__ double sr[EPB];
  const int tid = threadIdx.x;
  const int e = tid / Np;
  const int p = tid - e * Np;
  const int ke = blockIdx.x * EPB + e;
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / (N1 * N1);
  const uint32_t o = (uint32_t)Np * (valid ? ke : 0) + p;
  double qp = 0.0;
  if (valid) {
    const double dens = ldg(Dh, o) + ldg(DD, o);
    const double q = ldg(Q, o);
    qp = q > 0.0 ? q : 0.0;
    const double g = ldg(m.Gsqrt, o) * ldg(J, o) * ((w.w[i] * w.w[j]) * w.w[k]) * dens;
    s0[e][p] = g * q;
    s1[e][p] = g * qp;
  }
  __syncthreads();
  if (valid && p == 0) {
    double q0 = 0.0, q1 = 0.0;
#pragma unroll 16
    for (int n = 0; n < Np; ++n) {
      q0 = q0 + s0[e][n];
      q1 = q1 + s1[e][n];
    }
    sr[e] = q0 / (q1 + 1.0e-32);
  }
  __syncthreads();
  if (valid) stg(Q, o, sr[e] * qp);
}

// ATMOS_THERMODYN_specific_heat of the SCALE library (scale_atmos_thermodyn,
// not part of the reference tree; the published mixing rule): qdry = 1 -
// sum q_i M_i, Rtot = sum q_i R_i + qdry Rdry, likewise CVtot and CPtot, the
// sums over the tracers in order.
__device__ __forceinline__ void specific_heat(const double* q, const TrcThermo& th, const Consts& c, double& R,
                                              double& CV, double& CP) {
  double qdry = 1.0, r = 0.0, cv = 0.0, cp = 0.0;
  for (int iq = 0; iq < th.nq; ++iq) {
    qdry = qdry - q[iq] * th.Mq[iq];
    r = r + q[iq] * th.Rq[iq];
    cv = cv + q[iq] * th.CVq[iq];
    cp = cp + q[iq] * th.CPq[iq];
  }
  R = r + qdry * c.Rdry;
  CV = cv + qdry * c.CVdry;
  CP = cp + qdry * c.CPdry;
}

struct SpecHeat {
  const double* Q[kMaxTrc];
  double *R, *CV, *CP;
  TrcThermo th;
};
__global__ void __launch_bounds__(256) specific_heat_kernel(const SpecHeat a, const Consts c, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double q[kMaxTrc];
  for (int iq = 0; iq < a.th.nq; ++iq) q[iq] = a.Q[iq][t];
  double R, CV, CP;
  specific_heat(q, a.th, c, R, CV, CP);
  a.R[t] = R;
  a.CV[t] = CV;
  a.CP[t] = CP;
}

// vars_calc_diagnoseVar_lc (model/atm_nonhydro3d/src/atmos/mod_atmos_vars_container.F90:1100-1366)
// and the Umet/Vmet branch of AtmosVarsContainer_CalcDiagvar (:431-444), one
// node per thread; each formula keeps the reference's order of operations.
// The horizontal metric G_ij of the box mesh is the identity, written as the
// reference's products (G_11 MOMX + G_12 MOMY) so that the bits agree.
__global__ void __launch_bounds__(256) diagvar_kernel(const DiagVar a, const Consts c, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  constexpr double G11 = 1.0, G12 = 0.0, G22 = 1.0;
  double r = 0.0;
  switch (a.kind) {
    case kDiagDENS: r = a.DD[t] + a.Dh[t]; break;
    case kDiagU: r = a.MX[t] / (a.DD[t] + a.Dh[t]); break;
    case kDiagV: r = a.MY[t] / (a.DD[t] + a.Dh[t]); break;
    case kDiagW: r = a.MZ[t] / (a.DD[t] + a.Dh[t]); break;
    case kDiagUmet: r = a.MX[t] / (a.Dh[t] + a.DD[t]); break;
    case kDiagVmet: r = a.MY[t] / (a.Dh[t] + a.DD[t]); break;
    case kDiagPRES_diff: r = a.P[t] - a.Ph[t]; break;
    case kDiagT: r = a.P[t] / (a.R[t] * (a.DD[t] + a.Dh[t])); break;
    case kDiagT_diff:
      r = a.P[t] / (a.R[t] * (a.DD[t] + a.Dh[t])) - a.Ph[t] / (c.Rdry * a.Dh[t]);
      break;
    case kDiagPT:
    case kDiagPT_diff: {
      const double P = a.P[t], R = a.R[t];
      const double dens = a.DD[t] + a.Dh[t];
      r = P / (R * dens) * fedg_pow(c.PRES00 / P, R / a.CP[t]);
      if (a.kind == kDiagPT_diff)
        r = r - c.PRES00 / c.Rdry * fedg_pow(a.Ph[t] / c.PRES00, c.CVdry / c.CPdry) / a.Dh[t];
      break;
    }
    case kDiagQDRY: {
      double qdry = 1.0;
      for (int iq = 0; iq < a.nq; ++iq) qdry = qdry - a.Q[iq][t] * a.Mq[iq];
      r = qdry;
      break;
    }
    default: {  // ENGK, ENGP, ENGI, ENGT
      const double dens = a.DD[t] + a.Dh[t];
      double engk = 0.0, engi = 0.0;
      if (a.kind == kDiagENGK || a.kind == kDiagENGT) {
        const double mx = a.MX[t], my = a.MY[t], mz = a.MZ[t];
        const double u1 = G11 * mx + G12 * my;
        const double u2 = G12 * mx + G22 * my;
        engk = 0.5 * (mx * u1 + my * u2 + mz * mz) / dens;
      }
      if (a.kind == kDiagENGI || a.kind == kDiagENGT) {
        const double P = a.P[t], R = a.R[t];
        double qdry = 1.0;
        for (int iq = 0; iq < a.nq; ++iq) qdry = qdry - a.Q[iq][t] * a.Mq[iq];
        engi = qdry * P / R * c.CVdry;
        for (int iq = 0; iq < a.nq; ++iq) engi = engi + a.Q[iq][t] * (P / R * a.CVq[iq] + dens * a.E0q[iq]);
      }
      if (a.kind == kDiagENGK) r = engk;
      else if (a.kind == kDiagENGI) r = engi;
      else if (a.kind == kDiagENGP) r = dens * c.GRAV * a.zlev[t];
      else r = engk + engi + dens * c.GRAV * a.zpos[t];
    }
  }
  a.out[t] = r;
}

// atm_phy_mp_dgm_common_negative_fixer
// (FElib/src/microphysics/scale_atm_phy_mp_dgm_common.F90:429-556), one
// element per workgroup, one node per thread; every element sum is taken
// over the nodes in order by one thread, as Fortran's sum over p = 1..Np.
// The water species are all nq tracers (the reference's loop over
// 1 + QLA + QIA; its later loop over QA reads only those).
template <int N1, int EPB>
__global__ void __launch_bounds__(EPB * N1 * N1 * N1) negfix_kernel(const DevMesh m, const Consts c, const NegFix a) {
  constexpr int Np = N1 * N1 * N1;
  constexpr double TRC_EPS = 1.0e-128;
  __shared__ double s0[EPB][Np], s1[EPB][Np];
  __shared__ double sr[EPB][2];
  const int tid = threadIdx.x;
  const int e = tid / Np;
  const int p = tid - e * Np;
  const int ke = blockIdx.x * EPB + e;
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / (N1 * N1);
  const uint32_t o = (uint32_t)Np * (valid ? ke : 0) + p;
  const int nq = a.th.nq;
  double DENS = 0.0, DD0 = 0.0, DD = 0.0, IE0 = 0.0, TEMP = 0.0, IE = 0.0, intw = 0.0;
  double Dh = 0.0;
  double tm1[kMaxTrc];
  if (valid) {
    Dh = ldg(a.Dh, o);
    DD = ldg(a.DD, o);
    DENS = Dh + DD;
    DD0 = DD;
    const double CV = ldg(a.CV, o), P = ldg(a.PRES, o), R = ldg(a.R, o);
    IE0 = CV * P / R;
    TEMP = IE0 / (DENS * CV);
    IE = IE0;
    intw = ldg(m.Gsqrt, o) * ldg(a.J, o) * ((a.w.w[i] * a.w.w[j]) * a.w.w[k]);
  }
  auto esum = [&](double v0, double v1, double& r0, double& r1) {
    s0[e][p] = v0;
    s1[e][p] = v1;
    __syncthreads();
    if (valid && p == 0) {
      double q0 = 0.0, q1 = 0.0;
#pragma unroll 16
      for (int n = 0; n < Np; ++n) {
        q0 = q0 + s0[e][n];
        q1 = q1 + s1[e][n];
      }
      sr[e][0] = q0;
      sr[e][1] = q1;
    }
    __syncthreads();
    r0 = sr[e][0];
    r1 = sr[e][1];
    __syncthreads();
  };
  for (int iq = 0; iq < nq; ++iq) {
    double tm0 = 0.0, t1 = 0.0;
    if (valid) {
      tm0 = DENS * ldg(a.Q[iq], o);
      t1 = tm0 > TRC_EPS ? tm0 : TRC_EPS;
    }
    double M0, M1;
    esum(intw * tm0, intw * t1, M0, M1);
    if (valid) {
      t1 = (M0 > 0.0 ? M0 : 0.0) / M1 * t1;
      DD = DD + (t1 - tm0);
      IE = IE + (t1 - tm0) * a.th.CVq[iq] * TEMP;
    }
    tm1[iq] = t1;
  }
  double qn[kMaxTrc];
  if (valid) {
    DENS = Dh + DD;
    for (int iq = 0; iq < nq; ++iq) {
      qn[iq] = tm1[iq] / DENS;
      stg(a.Q[iq], o, qn[iq]);
    }
  }
  double E0, E1;
  esum(intw * IE0, intw * IE, E0, E1);
  if (!valid) return;
  IE = E0 / E1 * IE;
  double R, CV, CP;
  specific_heat(qn, a.th, c, R, CV, CP);
  stg(a.R, o, R);
  stg(a.CV, o, CV);
  stg(a.CP, o, CP);
  IE = IE - (DD - DD0) * c.GRAV * ldg(a.zlev, o);
  const double P = IE * R / CV;
  stg(a.PRES, o, P);
  stg(a.DD, o, DD);
  if (a.DRHOT) {
    stg(a.DRHOT, o, c.PRES00 / R * fedg_pow(P / c.PRES00, CV / CP) -
                        c.PRES00 / c.Rdry * fedg_pow(ldg(a.Ph, o) / c.PRES00, c.CVdry / c.CPdry));
  }
}

// save_massflux (:340-404): time-weighted mass fluxes over the RK stages,
// and the damping coefficients of cal_alphdens_dyn (:447-536) accumulated
// per face node.
__global__ void __launch_bounds__(256) massflux_kernel(const MflxArgs a, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t o = (uint32_t)t;
  if (a.rkstage == 1) {
    stg(a.FX, o, a.twh * ldg(a.MX, o));
    stg(a.FY, o, a.twh * ldg(a.MY, o));
    stg(a.FZ, o, a.twv * ldg(a.MZ, o));
  } else {
    stg(a.FX, o, ldg(a.FX, o) + a.twh * ldg(a.MX, o));
    stg(a.FY, o, ldg(a.FY, o) + a.twh * ldg(a.MY, o));
    stg(a.FZ, o, ldg(a.FZ, o) + a.twv * ldg(a.MZ, o));
  }
}
// Low-storage RK update of a tracer weighted by the density of each stage
// (CMN/scale_timeint_rk.F90:1390-1498, rk_advance_trcvar_low_storage):
// stage 1 stores var0 = q (DENS_hyd + DDENS0) and clears varTmp; varTmp
// collects sig_Ns q dens_ssm1 + gam_Ns tend when either is nonzero; q becomes
// ((1 - sig_ss) var0 + sig_ss q dens_ssm1 + gam_ss tend) / dens_ss, at the
// last stage (varTmp + sig_ss q dens_ssm1 + gam_ss tend) / (DENS_hyd + DDENS).
__global__ void __launch_bounds__(256) trc_rk_kernel(const TrcRk r, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t o = (uint32_t)t;
  const double Dh = ldg(r.Dh, o), DD = ldg(r.DD, o), DD0 = ldg(r.DD0, o);
  const double q = ldg(r.q, o), tend = ldg(r.tend, o);
  const double dens_ssm1 = Dh + DD0 + r.c_ssm1 * (DD - DD0);
  if (r.last) {
    stg(r.q, o, (ldg(r.varTmp, o) + r.sig_ss * q * dens_ssm1 + r.gam_ss * tend) / (Dh + DD));
    return;
  }
  double v0;
  double vt = 0.0;
  if (r.first) {
    v0 = q * (Dh + DD0);
    stg(r.var0, o, v0);
  } else {
    v0 = ldg(r.var0, o);
    vt = ldg(r.varTmp, o);
  }
  if (r.acc) vt = vt + r.sig_Ns * q * dens_ssm1 + r.gam_Ns * tend;
  if (r.first || r.acc) stg(r.varTmp, o, vt);
  const double dens_ss = Dh + DD0 + r.c_ss * (DD - DD0);
  stg(r.q, o, (r.one_minus_sig_ss * v0 + r.sig_ss * q * dens_ssm1 + r.gam_ss * tend) / dens_ss);
}

// cal_alphdens_advtest (FDS/scale_atm_dyn_dgm_trcadvect3d_heve.F90:407-445):
// alpha = max(|u_n M|, |u_n P|), alph_dens = alpha dens Gsqrt on each side.
__global__ void __launch_bounds__(256) alphdens_advtest_kernel(const DevMesh m, const MflxArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.NfpTot * m.Ne) return;
  const uint32_t gf = (uint32_t)t;
  const uint32_t iM = (uint32_t)m.vmapM[gf], iP = (uint32_t)m.vmapP[gf];
  const double densM = ldg(a.DD, iM) + ldg(a.Dh, iM);
  const double densP = ldg(a.DD, iP) + ldg(a.Dh, iP);
  const double nx = ldg(m.nx, gf), ny = ldg(m.ny, gf), nz = ldg(m.nz, gf);
  const double VelM = (ldg(a.MX, iM) * nx + ldg(a.MY, iM) * ny + ldg(a.MZ, iM) * nz) / densM;
  const double VelP = (ldg(a.MX, iP) * nx + ldg(a.MY, iP) * ny + ldg(a.MZ, iP) * nz) / densP;
  const double aM = fabs(VelM), aP = fabs(VelP);
  const double alpha = aM > aP ? aM : aP;
  stg(a.alM, gf, alpha * densM * ldg(m.Gsqrt, iM));
  stg(a.alP, gf, alpha * densP * ldg(m.Gsqrt, iP));
}

template <int N1>
__global__ void __launch_bounds__(256) alphdens_kernel(const DevMesh m, const MflxArgs a, double gamm) {
  constexpr int Nfp = N1 * N1, NfpTot = 6 * Nfp;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NfpTot * m.Ne) return;
  const uint32_t gf = (uint32_t)t;
  const int ke = t / NfpTot, fp = t - ke * NfpTot;
  const int f = fp / Nfp, an = fp - f * Nfp;
  const uint32_t h = (uint32_t)(face_node<N1>(f, an) % Nfp) + (uint32_t)Nfp * m.EMap3Dto2D[ke];
  const uint32_t iM = (uint32_t)m.vmapM[gf], iP = (uint32_t)m.vmapP[gf];
  // a.nbr: the fields have no halo tail (the fused stepper's buffers): the
  // exterior value is read at the interior node the exchange would copy and
  // the velocity BC that applyBC would apply is applied here
  const uint32_t jP = a.nbr ? (uint32_t)m.nbr[gf] : iP;
  const double nx = ldg(m.nx, gf), ny = ldg(m.ny, gf), nz = ldg(m.nz, gf);
  const double MXM = ldg(a.MX, iM), MYM = ldg(a.MY, iM), MZM = ldg(a.MZ, iM);
  double MXP = ldg(a.MX, jP), MYP = ldg(a.MY, jP), MZP = ldg(a.MZ, jP);
  if (a.nbr && m.nbr_bc[gf] != FEDG_BND_NONE)
    reflect_momentum(m.nbr_bc[gf], ldg(m.Gsqrt, iM), ldg(m.GsqrtH, h), ldg(m.G11, h), ldg(m.G12, h), ldg(m.G22, h),
                     ldg(m.G13, iM), ldg(m.G23, iM), nx, ny, nz, MXM, MYM, MZM, MXP, MYP, MZP);
  const double densM = ldg(a.DD, iM) + ldg(a.Dh, iM);
  const double densP = ldg(a.DD, jP) + ldg(a.Dh, jP);
  double Gnn = ldg(m.G11, h) * fabs(nx) + ldg(m.G22, h) * fabs(ny);
  if (!a.hevi) Gnn = Gnn + fabs(nz);
  const double VelM = (MXM * nx + MYM * ny + MZM * nz) / densM;
  const double VelP = (MXP * nx + MYP * ny + MZP * nz) / densP;
  const double cM = sqrt(Gnn * gamm * (ldg(a.Ph, iM) + ldg(a.DP, iM)) / densM) + fabs(VelM);
  const double cP = sqrt(Gnn * gamm * (ldg(a.Ph, jP) + ldg(a.DP, jP)) / densP) + fabs(VelP);
  const double alpha = cM > cP ? cM : cP;
  const double tw = a.twh * (fabs(nx) + fabs(ny)) + a.twv * fabs(nz);
  const double aM0 = a.rkstage == 1 ? 0.0 : ldg(a.alM, gf);
  const double aP0 = a.rkstage == 1 ? 0.0 : ldg(a.alP, gf);
  stg(a.alM, gf, aM0 + tw * alpha * densM * ldg(m.Gsqrt, iM));
  stg(a.alP, gf, aP0 + tw * alpha * densP * ldg(m.Gsqrt, iP));
}

template <int N1>
constexpr int epb_for() {
  return (N1 * N1 * N1 >= 128) ? 1 : (256 / (N1 * N1 * N1));
}

template <int MODE>
int launch_heve_mode(fedg_ctx* ctx, const TendIn& in, const RkOut& rk) {
  const DevMesh& m = ctx->m;
  auto go = [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    const int nblk = (m.Ne + EPB - 1) / EPB;
    hipLaunchKernelGGL((heve_kernel<N1, EPB, MODE>), dim3(nblk), dim3(EPB * N1 * N1 * N1), 0, ctx->stream, m,
                       ctx->c, in, rk);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  };
  switch (m.n1) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

// ---------------------------------------------------------------------------
__global__ void drhot2pres_kernel(double* PRES, double* DPRES, const double* DRHOT, const double* PRES_hyd,
                                  const double* THERM_hyd, const double* Rtot, const double* CVtot,
                                  const double* CPtot, long n, const Consts c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double rP0 = 1.0 / c.PRES00;
  const double RHOT = THERM_hyd[t] + DRHOT[t];
  // null Rtot / CVtot / CPtot: dry air (the caller verified the fields hold
  // the dry constants); null PRES: only DPRES is wanted
  const double Rt = Rtot ? Rtot[t] : c.Rdry;
  const double CVt = CVtot ? CVtot[t] : c.CVdry, CPt = CPtot ? CPtot[t] : c.CPdry;
  const double P = c.PRES00 * fedg_pow(Rt * rP0 * RHOT, CPt / CVt);
  if (PRES) PRES[t] = P;
  DPRES[t] = P - PRES_hyd[t];
}

__global__ void dpres_kernel(double* DPRES, const double* DRHOT, const double* PRES_hyd, const double* THERM_hyd,
                             const double* Rtot, const double* CVtot, const double* CPtot, long n, const Consts c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double Rt = c.Rdry, Cv = c.CVdry, Cp = c.CPdry;
  if (Rtot != nullptr) {
    Rt = Rtot[t];
    Cv = CVtot[t];
    Cp = CPtot[t];
  }
  DPRES[t] = dpres_of(DRHOT[t], THERM_hyd[t], PRES_hyd[t], Rt, Cv, Cp, c);
}

__global__ void halo_kernel(double* f, const int32_t* src, long base, int nB) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nB) return;
  f[base + b] = f[src[b]];
}

__global__ void apply_bc_kernel(const DevMesh m, double* MOMX, double* MOMY, double* MOMZ) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)m.NfpTot * m.Ne;
  if (t >= n) return;
  const long base = (long)m.Np * m.Ne;
  const long iP = m.vmapP[t];
  if (iP < base) return;
  const int bc = m.velbc[iP - base];
  if (bc == FEDG_BND_NONE) return;
  const long iM = m.vmapM[t];
  const int ke = (int)(t / m.NfpTot);
  const int fp = (int)(t - (long)ke * m.NfpTot);
  const int pm = (int)(iM - (long)m.Np * ke);
  const size_t h = (size_t)(pm % m.Nfp) + (size_t)m.Nfp * m.EMap3Dto2D[ke];
  double mx, my, mz;
  reflect_momentum(bc, m.Gsqrt[iM], m.GsqrtH[h], m.G11[h], m.G12[h], m.G22[h], m.G13[iM], m.G23[iM], m.nx[t],
                   m.ny[t], m.nz[t], MOMX[iM], MOMY[iM], MOMZ[iM], mx, my, mz);
  (void)fp;
  MOMX[iP] = mx;
  MOMY[iP] = my;
  MOMZ[iP] = mz;
}

// Hydrostatic pressure gradient (FDS/scale_atm_dyn_dgm_nonhydro3d_common.F90:624-777).
template <int N1, int EPB>
__global__ void __launch_bounds__(EPB * N1 * N1 * N1)
phyd_hgrad_kernel(const DevMesh m, double* DPhydDx, double* DPhydDy, const double* PRES_hyd,
                  const double* PRES_hyd_ref) {
  constexpr int Np = N1 * N1 * N1, Nfp = N1 * N1, NfpTot = 6 * Nfp;
  __shared__ double sD[N1 * N1];
  __shared__ double sF[EPB][4][Np];
  __shared__ double sH[EPB][2][NfpTot];
  const int tid = threadIdx.x;
  const int e = tid / Np, p = tid - e * Np;
  const int ke = blockIdx.x * EPB + e;
  const bool valid = ke < m.Ne;
  const int i = p % N1, j = (p / N1) % N1, k = p / Nfp;
  for (int t = tid; t < N1 * N1; t += blockDim.x) sD[t] = m.D1D[t];
  const size_t o = (size_t)Np * (valid ? ke : 0) + p;
  double RG = 0.0;
  if (valid) {
    const int ke2d = m.EMap3Dto2D[ke];
    for (int fp = p; fp < NfpTot; fp += Np) {
      const size_t gf = fp + (size_t)NfpTot * ke;
      const size_t iM = m.vmapM[gf], iP = m.vmapP[gf];
      const int pm = (int)(iM - (size_t)Np * ke);
      const double GH = m.GsqrtH[(size_t)(pm % Nfp) + (size_t)Nfp * ke2d];
      const double GsqrtV_M = m.Gsqrt[iM] / 1.0 / GH;
      const double GsqrtV_P = m.Gsqrt[iP] / 1.0 / GH;
      const double DP_M = PRES_hyd[iM] - PRES_hyd_ref[iM];
      const double DP_P = PRES_hyd[iP] - PRES_hyd_ref[iP];
      const double tmp1 = m.Fscale[gf] * 0.5 * GsqrtV_P * DP_P;
      const double tmp2 = m.Fscale[gf] * 0.5 * GsqrtV_M * DP_M;
      sH[e][0][fp] = (m.nx[gf] + m.G13[iP] * m.nz[gf]) * tmp1 - (m.nx[gf] + m.G13[iM] * m.nz[gf]) * tmp2;
      sH[e][1][fp] = (m.ny[gf] + m.G23[iP] * m.nz[gf]) * tmp1 - (m.ny[gf] + m.G23[iM] * m.nz[gf]) * tmp2;
    }
    const double GsqrtV = m.Gsqrt[o] / (1.0 * m.GsqrtH[(i + N1 * j) + (size_t)Nfp * ke2d]);
    RG = 1.0 / GsqrtV;
    const double F1 = GsqrtV * (PRES_hyd[o] - PRES_hyd_ref[o]);
    sF[e][0][p] = F1;
    sF[e][1][p] = F1;
    sF[e][2][p] = m.G13[o] * F1;
    sF[e][3][p] = m.G23[o] * F1;
  }
  __syncthreads();
  if (!valid) return;
  const double* F1 = sF[e][0];
  const double* F2 = sF[e][1];
  const double* F3 = sF[e][2];
  const double* Fz = sF[e][3];
  double dx = sD[i] * F1[N1 * (j + N1 * k)];
  double dy = F2[i + N1 * (N1 * k)] * sD[j];
  double dz = F3[i + N1 * j] * sD[k];
  double dz2 = Fz[i + N1 * j] * sD[k];
  for (int l = 1; l < N1; ++l) {
    dx = dx + sD[i + N1 * l] * F1[l + N1 * (j + N1 * k)];
    dy = dy + F2[i + N1 * (l + N1 * k)] * sD[j + N1 * l];
    dz = dz + F3[i + N1 * (j + N1 * l)] * sD[k + N1 * l];
    dz2 = dz2 + Fz[i + N1 * (j + N1 * l)] * sD[k + N1 * l];
  }
  double L[6];
  for (int f = 0; f < 6; ++f) L[f] = m.Lift[p + (size_t)Np * f];
  double lift[2];
  for (int s = 0; s < 2; ++s) {
    const double* h = sH[e][s];
    lift[s] = L[0] * h[i + N1 * k] + L[1] * h[Nfp + j + N1 * k] + L[2] * h[2 * Nfp + i + N1 * k] +
              L[3] * h[3 * Nfp + j + N1 * k] + L[4] * h[4 * Nfp + i + N1 * j] + L[5] * h[5 * Nfp + i + N1 * j];
  }
  const double gx = m.E11[o] * dx + m.E33[o] * dz + lift[0];
  const double gy = m.E22[o] * dy + m.E33[o] * dz2 + lift[1];
  DPhydDx[o] = gx * RG;
  DPhydDy[o] = gy * RG;
}

__global__ void axpy_set_kernel(double* y, const double* a, double ca, const double* x, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a[t] + ca * x[t];
}
__global__ void axpy_acc_kernel(double* y, double ca, const double* x, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = y[t] + ca * x[t];
}
__global__ void axpy2_acc_kernel(double* y, double ca, const double* x, double cb, const double* z, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = y[t] + ca * x[t] + cb * z[t];
}
__global__ void axpy2_set_kernel(double* y, const double* a, double ca, const double* x, double cb, const double* z,
                                 long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a[t] + ca * x[t] + cb * z[t];
}
// one pass of the IMEX stage advance over the five prognostics
// (blockIdx.y: variable); the expressions of axpy2_acc / axpy2_set.  S: the
// stage (its loads unrolled so that they are all in flight at once), 0: last.
// Every operand is streamed once: loads and stores carry the non-temporal hint
// (as do StoreImplicit's).
template <int S>
__global__ void __launch_bounds__(256) imex_advance_kernel(const ImexAdvance a, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (t >= n) return;
  const double Te = __builtin_nontemporal_load(a.Te[v] + t), Ti = __builtin_nontemporal_load(a.Ti[v] + t);
  if constexpr (S == 0) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(a.vt[v] + t) + a.cbe * Te + a.cbi * Ti, a.q[v] + t);
  } else {
    double e[S], im[S];
#pragma unroll
    for (int s2 = 0; s2 < S; ++s2) {
      e[s2] = __builtin_nontemporal_load(a.Tes[s2][v] + t);
      im[s2] = __builtin_nontemporal_load(a.Tis[s2][v] + t);
    }
    __builtin_nontemporal_store(__builtin_nontemporal_load(a.vt[v] + t) + a.cbe * Te + a.cbi * Ti, a.vt[v] + t);
    double q = __builtin_nontemporal_load(a.v0[v] + t);
#pragma unroll
    for (int s2 = 0; s2 < S; ++s2) q = q + a.ae[s2] * e[s2] + a.ai[s2] * im[s2];
    __builtin_nontemporal_store(q, a.q[v] + t);
  }
}

struct PhysTend {
  double* T[5];
  const double* tp[6];
  const double *PRES, *Rtot, *CPtot;
};
// add_phy_tend (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:1098-1178,
// ENTOT_CONSERVE_SCHEME_FLAG off) on one node per thread
__global__ void __launch_bounds__(256) add_phy_tend_kernel(const PhysTend a, const Consts c, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t o = (uint32_t)t;
  a.T[VD][t] = a.T[VD][t] + ldg(a.tp[VD], o);
  a.T[VW][t] = a.T[VW][t] + ldg(a.tp[VW], o);
  a.T[VU][t] = a.T[VU][t] + ldg(a.tp[VU], o);
  a.T[VV][t] = a.T[VV][t] + ldg(a.tp[VV], o);
  a.T[VT][t] = phys_add_rhot(a.T[VT][t], ldg(a.tp[VT], o), ldg(a.tp[5], o), ldg(a.PRES, o), ldg(a.Rtot, o),
                             ldg(a.CPtot, o), c);
}

struct StoreImpl {
  double* q[5];
  const double* Ti[5];
  double* var0[5];
  double* varTmp[5];
  double impl_fac;
  int copy;
};
__global__ void __launch_bounds__(256) store_implicit_kernel(const StoreImpl a, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (t >= n) return;
  const double q = __builtin_nontemporal_load(a.q[v] + t);
  if (a.copy) {
    __builtin_nontemporal_store(q, a.var0[v] + t);
    __builtin_nontemporal_store(q, a.varTmp[v] + t);
  }
  __builtin_nontemporal_store(q + a.impl_fac * __builtin_nontemporal_load(a.Ti[v] + t), a.q[v] + t);
}

// Access-pattern probe (measurement only): the loads and stores of one p = 7
// fused stage on a periodic NeX x NeY x NeZ box without its arithmetic.  Every
// workgroup reads NR fields of its element's 512 nodes, the exterior values
// of NG fields at its 384 face nodes (from the six neighbours) and writes NW
// fields: the practical HBM ceiling for that traffic mix.
struct ProbeArgs {
  const double* r[16];
  double* w[11];
  int NeX, NeY, NeZ;
};
// The fields beyond the first NG (the RK operands) are read, and all fields
// written, with the non-temporal hint, as the stage kernel streams them.
template <int NR, int NW, int NG>
__global__ void __launch_bounds__(256) stage_probe_kernel(const ProbeArgs p) {
  const int e = blockIdx.x, t = threadIdx.x;
  const size_t base = (size_t)e * 512;
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int f = 0; f < NR; ++f) {
    if (f < NG) {
      a0 += p.r[f][base + t];
      a1 += p.r[f][base + t + 256];
    } else {
      a0 += __builtin_nontemporal_load(p.r[f] + base + t);
      a1 += __builtin_nontemporal_load(p.r[f] + base + t + 256);
    }
  }
  const int ei = e % p.NeX, ej = (e / p.NeX) % p.NeY, ek = e / (p.NeX * p.NeY);
  for (int idx = t; idx < 384; idx += 256) {
    const int f = idx / 64, a = idx % 64;
    int ni = ei, nj = ej, nk = ek;
    if (f == 0) nj = (nj + p.NeY - 1) % p.NeY;
    else if (f == 1) ni = (ni + 1) % p.NeX;
    else if (f == 2) nj = (nj + 1) % p.NeY;
    else if (f == 3) ni = (ni + p.NeX - 1) % p.NeX;
    else if (f == 4) nk = (nk + p.NeZ - 1) % p.NeZ;
    else nk = (nk + 1) % p.NeZ;
    const int opp = (f < 4) ? (f ^ 2) : (9 - f);
    const size_t g = (size_t)(ni + p.NeX * (nj + p.NeY * nk)) * 512 + face_node<8>(opp, a);
#pragma unroll
    for (int q = 0; q < NG; ++q) a0 += p.r[q][g];
  }
#pragma unroll
  for (int f = 0; f < NW; ++f) {
    __builtin_nontemporal_store(a0 + f, p.w[f] + base + t);
    __builtin_nontemporal_store(a1 + f, p.w[f] + base + t + 256);
  }
}

__global__ void copy_kernel(double* y, const double* x, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = x[t];
}

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

int launch_heve_dropin(fedg_ctx* ctx, const TendIn& in, double* const out[5]) {
  RkOut rk{};
  for (int v = 0; v < 5; ++v) rk.out[v] = out[v];
  return launch_heve_mode<MODE_DROPIN>(ctx, in, rk);
}

int launch_heve_cco(fedg_ctx* ctx, const TendIn& in, double* const out[5]) {
  RkOut rk{};
  for (int v = 0; v < 5; ++v) rk.out[v] = out[v];
  return launch_heve_mode<MODE_DROPIN_CCO>(ctx, in, rk);
}

int launch_heve_bnd(fedg_ctx* ctx, const TendIn& in, double* const dt[5], double* DPRES) {
  const DevMesh& m = ctx->m;
  // DPRES of the halo tail from its state and base state (..._heve_numflux.F90:331-335)
  if (m.nB > 0) {
    const long base = (long)m.Np * m.Ne;
    const auto off = [&](const double* f) -> const double* { return f ? f + base : nullptr; };
    hipLaunchKernelGGL(dpres_kernel, dim3(nblocks(m.nB, 256)), dim3(256), 0, ctx->stream, DPRES + base,
                       in.q[VT] + base, in.PRES_hyd + base, in.THERM_hyd + base, off(in.Rtot), off(in.CVtot),
                       off(in.CPtot), (long)m.nB, ctx->c);
    FEDG_HIP_CHECK(hipGetLastError());
  }
  RkOut rk{};
  for (int v = 0; v < 5; ++v) rk.out[v] = dt[v];
  TendIn tin = in;
  tin.DPRES = DPRES;
  auto go = [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    const int nblk = (m.Ne + EPB - 1) / EPB;
    hipLaunchKernelGGL((heve_bnd_kernel<N1, EPB>), dim3(nblk), dim3(EPB * N1 * N1 * N1), 0, ctx->stream, m, ctx->c,
                       tin, rk);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  };
  switch (m.n1) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

template <class F>
static int dispatch_n1(int n1, F&& go) {
  switch (n1) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

int launch_trcadv(fedg_ctx* ctx, const TrcLaunch& t) {
  const DevMesh& m = ctx->m;
  TrcArgs a{};
  a.dt = t.dt; a.Q = t.Q; a.MX = t.MX; a.MY = t.MY; a.MZ = t.MZ;
  a.aM = t.aM; a.aP = t.aP; a.R = t.R; a.tp = t.tp; a.J = t.J;
  a.Dh = t.Dh; a.DD = t.DD; a.DD0 = t.DD0; a.c = t.c; a.dtt = t.dtt; a.disable = t.disable;
  for (int l = 0; l < 8; ++l) a.w[l] = t.w.w[l];
  return dispatch_n1(m.n1, [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    const int nblk = (m.Ne + EPB - 1) / EPB;
    const dim3 grid(nblk), block(EPB * N1 * N1 * N1);
    if (t.fct && m.flat)
      hipLaunchKernelGGL((trcadv_kernel<N1, EPB, TRC_FCT, true>), grid, block, 0, ctx->stream, m, a);
    else if (t.fct)
      hipLaunchKernelGGL((trcadv_kernel<N1, EPB, TRC_FCT, false>), grid, block, 0, ctx->stream, m, a);
    else if (m.flat)
      hipLaunchKernelGGL((trcadv_kernel<N1, EPB, TRC_TEND, true>), grid, block, 0, ctx->stream, m, a);
    else
      hipLaunchKernelGGL((trcadv_kernel<N1, EPB, TRC_TEND, false>), grid, block, 0, ctx->stream, m, a);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  });
}

int launch_tmar(fedg_ctx* ctx, double* Q, const double* Dh, const double* DD, const double* J, const TrcW& w) {
  const DevMesh& m = ctx->m;
  return dispatch_n1(m.n1, [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    const int nblk = (m.Ne + EPB - 1) / EPB;
    hipLaunchKernelGGL((tmar_kernel<N1, EPB>), dim3(nblk), dim3(EPB * N1 * N1 * N1), 0, ctx->stream, m, Q, Dh, DD, J,
                       w);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  });
}

int launch_specific_heat(fedg_ctx* ctx, const double* const Q[kMaxTrc], const TrcThermo& th, double* Rtot,
                         double* CVtot, double* CPtot) {
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  SpecHeat a{};
  for (int iq = 0; iq < th.nq; ++iq) a.Q[iq] = Q[iq];
  a.R = Rtot;
  a.CV = CVtot;
  a.CP = CPtot;
  a.th = th;
  hipLaunchKernelGGL(specific_heat_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, a, ctx->c, n);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_diagvar(fedg_ctx* ctx, const DiagVar& a) {
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  hipLaunchKernelGGL(diagvar_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, a, ctx->c, n);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_negative_fixer(fedg_ctx* ctx, const NegFix& a) {
  const DevMesh& m = ctx->m;
  return dispatch_n1(m.n1, [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    const int nblk = (m.Ne + EPB - 1) / EPB;
    hipLaunchKernelGGL((negfix_kernel<N1, EPB>), dim3(nblk), dim3(EPB * N1 * N1 * N1), 0, ctx->stream, m, ctx->c,
                       a);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  });
}

int launch_trc_rk(fedg_ctx* ctx, const TrcRk& r) {
  const int n = ctx->m.Np * ctx->m.Ne;
  hipLaunchKernelGGL(trc_rk_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, r, n);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_alphdens_advtest(fedg_ctx* ctx, const MflxArgs& a) {
  const int nf = ctx->m.NfpTot * ctx->m.Ne;
  hipLaunchKernelGGL(alphdens_advtest_kernel, dim3((nf + 255) / 256), dim3(256), 0, ctx->stream, ctx->m, a);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_save_massflux(fedg_ctx* ctx, const MflxArgs& a) {
  const DevMesh& m = ctx->m;
  const int n = m.Np * m.Ne;
  hipLaunchKernelGGL(massflux_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, a, n);
  FEDG_HIP_CHECK(hipGetLastError());
  const double gamm = ctx->c.CPdry / ctx->c.CVdry;
  return dispatch_n1(m.n1, [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    const int nf = 6 * N1 * N1 * m.Ne;
    hipLaunchKernelGGL((alphdens_kernel<N1>), dim3((nf + 255) / 256), dim3(256), 0, ctx->stream, m, a, gamm);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  });
}

int launch_hevi_explicit(fedg_ctx* ctx, const TendIn& in, double* const out[5]) {
  RkOut rk{};
  for (int v = 0; v < 5; ++v) rk.out[v] = out[v];
  return launch_heve_mode<MODE_DROPIN_HEVI>(ctx, in, rk);
}

// mode 0 / 1 / 2: the traffic of the first / middle / last RK4 stage
// (reads 6 / 16 / 11 fields, writes 11 / 11 / 6, six fields gathered at the
// face nodes); arrays of NeX NeY NeZ 512 doubles, allocated here
int run_stage_probe(int mode, int NeX, int NeY, int NeZ, int reps, double* ms) {
  if (mode < 0 || mode > 2 || NeX < 1 || NeY < 1 || NeZ < 1 || reps < 1 || !ms) return FEDG_ERR_ARG;
  const size_t ne = (size_t)NeX * NeY * NeZ;
  if (ne > (size_t)0x7fffffff) return FEDG_ERR_ARG;
  const size_t n = ne * 512;
  std::vector<double*> buf(27, nullptr);
  int rc = FEDG_OK;
  for (auto& b : buf) {
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess || hipMemset(b, 0, n * sizeof(double)) != hipSuccess) {
      set_error("stage probe: cannot allocate its 27 fields");
      rc = FEDG_ERR_NOMEM;
      break;
    }
  }
  ProbeArgs p;
  for (int f = 0; f < 16; ++f) p.r[f] = buf[f];
  for (int f = 0; f < 11; ++f) p.w[f] = buf[16 + f];
  p.NeX = NeX; p.NeY = NeY; p.NeZ = NeZ;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == FEDG_OK) {
    auto go = [&]() {
      if (mode == 0) hipLaunchKernelGGL((stage_probe_kernel<6, 11, 6>), dim3((unsigned)ne), dim3(256), 0, 0, p);
      else if (mode == 1) hipLaunchKernelGGL((stage_probe_kernel<16, 11, 6>), dim3((unsigned)ne), dim3(256), 0, 0, p);
      else hipLaunchKernelGGL((stage_probe_kernel<11, 6, 6>), dim3((unsigned)ne), dim3(256), 0, 0, p);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = FEDG_ERR_HIP;
    if (rc == FEDG_OK) {
      go();
      go();
      (void)hipEventRecord(e0, 0);
      for (int r = 0; r < reps; ++r) go();
      (void)hipEventRecord(e1, 0);
      float t = 0.f;
      if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
        set_error("stage probe: launch failed");
        rc = FEDG_ERR_HIP;
      }
      *ms = (double)t / reps;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (auto b : buf)
    if (b) (void)hipFree(b);
  return rc;
}

int launch_imex_advance(fedg_ctx* ctx, const ImexAdvance& a) {
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  if (a.s < 1 || a.s > kImexMaxStages) {
    set_error("IMEX advance: stage out of range");
    return FEDG_ERR_ARG;
  }
  const dim3 g(nblocks(n, 256), 5);
  switch (a.last ? 0 : a.s) {
    case 0: hipLaunchKernelGGL(imex_advance_kernel<0>, g, dim3(256), 0, ctx->stream, a, n); break;
    case 1: hipLaunchKernelGGL(imex_advance_kernel<1>, g, dim3(256), 0, ctx->stream, a, n); break;
    case 2: hipLaunchKernelGGL(imex_advance_kernel<2>, g, dim3(256), 0, ctx->stream, a, n); break;
    case 3: hipLaunchKernelGGL(imex_advance_kernel<3>, g, dim3(256), 0, ctx->stream, a, n); break;
    case 4: hipLaunchKernelGGL(imex_advance_kernel<4>, g, dim3(256), 0, ctx->stream, a, n); break;
    case 5: hipLaunchKernelGGL(imex_advance_kernel<5>, g, dim3(256), 0, ctx->stream, a, n); break;
    case 6: hipLaunchKernelGGL(imex_advance_kernel<6>, g, dim3(256), 0, ctx->stream, a, n); break;
    default: hipLaunchKernelGGL(imex_advance_kernel<7>, g, dim3(256), 0, ctx->stream, a, n); break;
  }
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_add_phy_tend(fedg_ctx* ctx, double* const T[5], const double* const tp[6], const double* PRES,
                        const double* Rtot, const double* CPtot) {
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  PhysTend a{};
  for (int v = 0; v < 5; ++v) a.T[v] = T[v];
  for (int v = 0; v < 6; ++v) a.tp[v] = tp[v];
  a.PRES = PRES;
  a.Rtot = Rtot;
  a.CPtot = CPtot;
  hipLaunchKernelGGL(add_phy_tend_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, a, ctx->c, n);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_store_implicit(fedg_ctx* ctx, double* const q[5], const double* const Ti[5], double* const var0[5],
                          double* const varTmp[5], double impl_fac) {
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  StoreImpl a{};
  a.copy = var0 != nullptr && varTmp != nullptr;
  for (int v = 0; v < 5; ++v) {
    a.q[v] = q[v];
    a.Ti[v] = Ti[v];
    a.var0[v] = a.copy ? var0[v] : nullptr;
    a.varTmp[v] = a.copy ? varTmp[v] : nullptr;
  }
  a.impl_fac = impl_fac;
  hipLaunchKernelGGL(store_implicit_kernel, dim3(nblocks(n, 256), 5), dim3(256), 0, ctx->stream, a, n);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_heve_fused32(fedg_ctx* ctx, const TendIn& in, const RkOut& rk) {
  const DevMesh& m = ctx->m;
  if (!m.flat || in.Rtot != nullptr) {
    set_error("the FP32 perturbation form is built for the flat-terrain box with dry air");
    return FEDG_ERR_UNSUPPORTED;
  }
  if (m.n1 == 8 && m.lift_tensor) {
    if (in.lvl)
      hipLaunchKernelGGL(heve7_f32_lvl_kernel, dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
    else
      hipLaunchKernelGGL(heve7_f32_kernel, dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  }
  auto go = [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    const int nblk = (m.Ne + EPB - 1) / EPB;
    hipLaunchKernelGGL((heve32_kernel<N1, EPB>), dim3(nblk), dim3(EPB * N1 * N1 * N1), 0, ctx->stream, m, ctx->c,
                       in, rk);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  };
  switch (m.n1) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

int launch_heve_fused(fedg_ctx* ctx, const TendIn& in, const RkOut& rk, int use_flat) {
  const DevMesh& m = ctx->m;
  // the p = 7 layout carries no physics tendencies: with them, the general layout
  if (use_flat && m.n1 == 8 && m.lift_tensor && in.tp[0] == nullptr) {
    if (in.lvl)
      hipLaunchKernelGGL(heve7_lvl_kernel, dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
    else
      hipLaunchKernelGGL(heve7_flat_kernel, dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  }
  if (use_flat) return launch_heve_mode<MODE_FUSED_FLAT>(ctx, in, rk);
  return launch_heve_mode<MODE_FUSED_GENERAL>(ctx, in, rk);
}

int launch_hevi_explicit7(fedg_ctx* ctx, const TendIn& in, double* const out[5], int fp32) {
  const DevMesh& m = ctx->m;
  if (!(m.flat && m.n1 == 8 && m.lift_tensor) || in.Rtot != nullptr) {
    set_error("the p = 7 explicit HEVI tendency is built for the flat-terrain box with dry air");
    return FEDG_ERR_UNSUPPORTED;
  }
  RkOut rk{};
  for (int v = 0; v < 5; ++v) rk.out[v] = out[v];
  rk.mode = 4;
  if (fp32) {
    if (in.lvl)
      hipLaunchKernelGGL((heve7_hevi32_kernel<true>), dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
    else
      hipLaunchKernelGGL((heve7_hevi32_kernel<false>), dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
  } else {
    if (in.lvl)
      hipLaunchKernelGGL((heve7_hevi_kernel<true>), dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in, rk);
    else
      hipLaunchKernelGGL((heve7_hevi_kernel<false>), dim3(m.Ne), dim3(512), 0, ctx->stream, m, ctx->c, in,
                         rk);
  }
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_drhot2pres(fedg_ctx* ctx, double* PRES, double* DPRES, const double* DRHOT, const double* PRES_hyd,
                      const double* THERM_hyd, const double* Rtot, const double* CVtot, const double* CPtot,
                      long n) {
  if (n <= 0) return FEDG_OK;
  hipLaunchKernelGGL(drhot2pres_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, PRES, DPRES, DRHOT,
                     PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot, n, ctx->c);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_dpres(fedg_ctx* ctx, double* DPRES, const double* DRHOT, const double* PRES_hyd, const double* THERM_hyd,
                 const double* Rtot, const double* CVtot, const double* CPtot, long n) {
  if (n <= 0) return FEDG_OK;
  hipLaunchKernelGGL(dpres_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, DPRES, DRHOT, PRES_hyd,
                     THERM_hyd, Rtot, CVtot, CPtot, n, ctx->c);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_halo_exchange(fedg_ctx* ctx, double* field) {
  const DevMesh& m = ctx->m;
  if (m.nB <= 0) return FEDG_OK;
  hipLaunchKernelGGL(halo_kernel, dim3(nblocks(m.nB, 256)), dim3(256), 0, ctx->stream, field, m.halo_src,
                     (long)m.Np * m.Ne, m.nB);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_apply_bc(fedg_ctx* ctx, double* MOMX, double* MOMY, double* MOMZ) {
  const DevMesh& m = ctx->m;
  const long n = (long)m.NfpTot * m.Ne;
  hipLaunchKernelGGL(apply_bc_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, m, MOMX, MOMY, MOMZ);
  FEDG_HIP_CHECK(hipGetLastError());
  return FEDG_OK;
}

int launch_phyd_hgrad(fedg_ctx* ctx, double* DPhydDx, double* DPhydDy, const double* PRES_hyd,
                      const double* PRES_hyd_ref) {
  const DevMesh& m = ctx->m;
  auto go = [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    hipLaunchKernelGGL((phyd_hgrad_kernel<N1, EPB>), dim3((m.Ne + EPB - 1) / EPB), dim3(EPB * N1 * N1 * N1), 0,
                       ctx->stream, m, DPhydDx, DPhydDy, PRES_hyd, PRES_hyd_ref);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  };
  switch (m.n1) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

// Modal filter of atm_dyn_dgm_modalfilter_apply (FDS/scale_atm_dyn_dgm_modalfilter.F90:48-130):
// per element and variable, the x pass (MFilter_h1D), y pass (its transpose
// applied from the right) and z pass (MFilter_v1D) of
// element_operation_kernel_matvec_ModalFilter_P* (ELM/..._tensorprod3D_kernel.F90:1061-1112),
// each sum over l in order, on Gsqrt-weighted values when weight != 0.
struct Fields5 {
  double* f[5];
};

template <int N1, int EPB>
__global__ void __launch_bounds__(EPB* N1* N1* N1) modalfilter_kernel(Fields5 F, const double* __restrict__ Gsqrt,
                                                                     const double* __restrict__ MFh,
                                                                     const double* __restrict__ MFv, int Ne,
                                                                     int weight) {
  constexpr int Np = N1 * N1 * N1;
  __shared__ double sh[2 * N1 * N1];
  __shared__ double a[EPB * Np], b[EPB * Np];
  const int le = threadIdx.x / Np, t = threadIdx.x % Np;
  const int ke = blockIdx.x * EPB + le;
  for (int u = threadIdx.x; u < 2 * N1 * N1; u += blockDim.x) sh[u] = u < N1 * N1 ? MFh[u] : MFv[u - N1 * N1];
  const bool act = ke < Ne;
  const int i = t % N1, j = (t / N1) % N1, k = t / (N1 * N1);
  const size_t o = (size_t)ke * Np + t;
  const double G = (act && weight) ? Gsqrt[o] : 1.0;
  const double* Fh = sh;
  const double* Fv = sh + N1 * N1;
  double* A = a + le * Np;
  double* B = b + le * Np;
  for (int v = 0; v < 5; ++v) {
    double x = act ? F.f[v][o] : 0.0;
    if (weight) x = G * x;
    A[t] = x;
    __syncthreads();
    double s = Fh[i] * A[N1 * (j + N1 * k)];
#pragma unroll
    for (int l = 1; l < N1; ++l) s = s + Fh[i + N1 * l] * A[l + N1 * (j + N1 * k)];
    B[t] = s;
    __syncthreads();
    s = B[i + N1 * N1 * k] * Fh[j];
#pragma unroll
    for (int l = 1; l < N1; ++l) s = s + B[i + N1 * (l + N1 * k)] * Fh[j + N1 * l];
    A[t] = s;
    __syncthreads();
    s = A[i + N1 * j] * Fv[k];
#pragma unroll
    for (int l = 1; l < N1; ++l) s = s + A[i + N1 * (j + N1 * l)] * Fv[k + N1 * l];
    if (weight) {
      const double RG = 1.0 / G;
      s = s * RG;
    }
    if (act) F.f[v][o] = s;
    __syncthreads();
  }
}

// Tracer modal filter (FDS/scale_atm_dyn_dgm_modalfilter.F90:232-270): the
// same three passes with the tracer's matrices on weight QTRC, weight =
// Gsqrt (DENS_hyd + DDENS), then QTRC = out / weight.
struct TrcMF {
  double h[64], v[64];  // MFilter_tracer_h1D, MFilter_tracer_v1D (n1,n1)
};
template <int N1, int EPB>
__global__ void __launch_bounds__(EPB* N1* N1* N1)
    tracer_modalfilter_kernel(double* __restrict__ Q, const double* __restrict__ Gsqrt, const double* __restrict__ Dh,
                              const double* __restrict__ DD, const TrcMF mf, int Ne) {
  constexpr int Np = N1 * N1 * N1;
  __shared__ double a[EPB * Np], b[EPB * Np];
  const int le = threadIdx.x / Np, t = threadIdx.x % Np;
  const int ke = blockIdx.x * EPB + le;
  const bool act = ke < Ne;
  const int i = t % N1, j = (t / N1) % N1, k = t / (N1 * N1);
  const size_t o = (size_t)(act ? ke : 0) * Np + t;
  const double* Fh = mf.h;
  const double* Fv = mf.v;
  double* A = a + le * Np;
  double* B = b + le * Np;
  double w = 1.0;
  if (act) {
    w = Gsqrt[o] * (Dh[o] + DD[o]);
    A[t] = w * Q[o];
  } else {
    A[t] = 0.0;
  }
  __syncthreads();
  double s = Fh[i] * A[N1 * (j + N1 * k)];
#pragma unroll
  for (int l = 1; l < N1; ++l) s = s + Fh[i + N1 * l] * A[l + N1 * (j + N1 * k)];
  B[t] = s;
  __syncthreads();
  s = B[i + N1 * N1 * k] * Fh[j];
#pragma unroll
  for (int l = 1; l < N1; ++l) s = s + B[i + N1 * (l + N1 * k)] * Fh[j + N1 * l];
  __syncthreads();
  A[t] = s;
  __syncthreads();
  s = A[i + N1 * j] * Fv[k];
#pragma unroll
  for (int l = 1; l < N1; ++l) s = s + A[i + N1 * (j + N1 * l)] * Fv[k + N1 * l];
  if (act) Q[o] = s / w;
}

int launch_tracer_modalfilter(fedg_ctx* ctx, double* Q, const double* Dh, const double* DD, const double* MFh,
                              const double* MFv) {
  const DevMesh& m = ctx->m;
  TrcMF mf{};
  for (int u = 0; u < m.n1 * m.n1; ++u) {
    mf.h[u] = MFh[u];
    mf.v[u] = MFv[u];
  }
  return dispatch_n1(m.n1, [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    hipLaunchKernelGGL((tracer_modalfilter_kernel<N1, EPB>), dim3((m.Ne + EPB - 1) / EPB), dim3(EPB * N1 * N1 * N1),
                       0, ctx->stream, Q, m.Gsqrt, Dh, DD, mf, m.Ne);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  });
}

int launch_modalfilter(fedg_ctx* ctx, double* DDENS, double* MOMX, double* MOMY, double* MOMZ, double* DRHOT,
                       int do_weight_Gsqrt) {
  const DevMesh& m = ctx->m;
  if (!m.MFh || !m.MFv) {
    set_error("modal filter: matrices not set up (fedg_ctx_setup_modalfilter)");
    return FEDG_ERR_ARG;
  }
  Fields5 F{{DDENS, MOMX, MOMY, MOMZ, DRHOT}};
  auto go = [&](auto n1c) -> int {
    constexpr int N1 = decltype(n1c)::value;
    constexpr int EPB = epb_for<N1>();
    hipLaunchKernelGGL((modalfilter_kernel<N1, EPB>), dim3((m.Ne + EPB - 1) / EPB), dim3(EPB * N1 * N1 * N1), 0,
                       ctx->stream, F, m.Gsqrt, m.MFh, m.MFv, m.Ne, do_weight_Gsqrt ? 1 : 0);
    FEDG_HIP_CHECK(hipGetLastError());
    return FEDG_OK;
  };
  switch (m.n1) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default:
      set_error("polynomial order not built (supported: p = 1..7)");
      return FEDG_ERR_UNSUPPORTED;
  }
}

#define FEDG_STREAM_LAUNCH(kern, n, ...)                                                                   \
  do {                                                                                                     \
    if ((n) <= 0) return FEDG_OK;                                                                          \
    hipLaunchKernelGGL(kern, dim3(nblocks((n), 256)), dim3(256), 0, ctx->stream, __VA_ARGS__);             \
    FEDG_HIP_CHECK(hipGetLastError());                                                                     \
    return FEDG_OK;                                                                                        \
  } while (0)

int launch_axpy_set(fedg_ctx* ctx, double* y, const double* a, double ca, const double* x, long n) {
  FEDG_STREAM_LAUNCH(axpy_set_kernel, n, y, a, ca, x, n);
}
int launch_axpy_acc(fedg_ctx* ctx, double* y, double ca, const double* x, long n) {
  FEDG_STREAM_LAUNCH(axpy_acc_kernel, n, y, ca, x, n);
}
int launch_axpy2_acc(fedg_ctx* ctx, double* y, double ca, const double* x, double cb, const double* z, long n) {
  FEDG_STREAM_LAUNCH(axpy2_acc_kernel, n, y, ca, x, cb, z, n);
}
int launch_axpy2_set(fedg_ctx* ctx, double* y, const double* a, double ca, const double* x, double cb,
                     const double* z, long n) {
  FEDG_STREAM_LAUNCH(axpy2_set_kernel, n, y, a, ca, x, cb, z, n);
}
int launch_copy(fedg_ctx* ctx, double* y, const double* x, long n) { FEDG_STREAM_LAUNCH(copy_kernel, n, y, x, n); }

}  // namespace fedg
