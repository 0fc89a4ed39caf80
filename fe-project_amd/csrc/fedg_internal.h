// Internal definitions shared by the C-ABI layer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/fedg.h"

namespace fedg {

void set_error(const std::string& msg);

#define FEDG_HIP_CHECK(expr)                                                         \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::fedg::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));          \
      return FEDG_ERR_HIP;                                                           \
    }                                                                                \
  } while (0)

struct Consts {
  double GRAV, Rdry, CPdry, CVdry, PRES00;
};

// Device copies of the flattened mesh / element / operator data.
struct DevMesh {
  int n1, Np, Nfp, NfpTot;
  int Ne, NeA, Ne2D, Ne2DA, nB;
  int NeX, NeY, NeZ;           // box shape (0 if unknown)
  double* E11;                 // (Np,Ne) Escale(:,:,1,1)
  double* E22;                 // (Np,Ne) Escale(:,:,2,2)
  double* E33;                 // (Np,Ne) Escale(:,:,3,3)
  double* Gsqrt;               // (Np,NeA)
  double* G13;                 // (Np,NeA)
  double* G23;                 // (Np,NeA)
  double* GsqrtH;              // (Nfp,Ne2D)
  double* G11; double* G12; double* G22;  // (Nfp,Ne2D)
  double* nx; double* ny; double* nz;     // (NfpTot,Ne)
  double* Fscale;              // (NfpTot,Ne)
  int32_t* vmapM;              // (NfpTot,Ne) 0-based
  int32_t* vmapP;              // (NfpTot,Ne) 0-based
  int32_t* EMap3Dto2D;         // (Ne) 0-based
  int32_t* halo_src;           // (nB) 0-based
  int32_t* velbc;              // (nB)
  int32_t* nbr;                // (NfpTot,Ne) 0-based interior node of the exterior value
  int8_t* nbr_bc;              // (NfpTot,Ne) velocity BC applied to the exterior value
  double* D1D;                 // (n1,n1)
  double* Lift;                // (n1,n1,n1,6)
  double* Itr;                 // (n1,n1)
  // Lift_mat as six 1D columns, when it has the tensor-product structure of
  // the exact-mass-matrix lift: Lift(i,j,k,f) = liftT(x,f) with x = j (f = 1,3),
  // i (f = 2,4), k (f = 5,6); lift_tensor = 0 otherwise
  int lift_tensor;
  double* liftT;               // (n1,6)
  double* MFh;                 // (n1,n1) modal filter MFilter_h1D, or null
  double* MFv;                 // (n1,n1) MFilter_v1D
  // flat uniform box description (valid when flat != 0)
  int flat;
  double fE[3];                // E11,E22,E33 of every element
  double fF[6];                // Fscale of every face
  int bnd_bc[6];               // velocity BC of the six domain boundaries (closed ones)
  int periodic[3];             // x, y, z
};

}  // namespace fedg

namespace fedg {
struct HeviWork;
}

struct fedg_ctx {
  fedg::DevMesh m;
  fedg::Consts c;
  hipStream_t stream;
  int device;
  int ncu;               // compute units of the device
  fedg::HeviWork* hevi;  // implicit-solve workspace, allocated on first use
  int have_w1d;          // w1d holds the 1D LGL weights of the order (first tracer call)
  double w1d[8];
};

namespace fedg {

// ---- kernel launchers (fedg_kernels.hip) ----
struct TendIn {
  const double* q[5];        // stage state, VID order (DDENS, DRHOT, MOMZ, MOMX, MOMY)
  const double* DPRES;       // pressure perturbation of the stage state
  const double* DENS_hyd;
  const double* PRES_hyd;
  const double* THERM_hyd;
  const double* CORIOLIS;
  const double* DPhydDx;
  const double* DPhydDy;
  const double* Rtot;        // fused: per node or null (dry constants)
  const double* CVtot;
  const double* CPtot;
  // fused p = 7 flat stage: DENS_hyd, PRES_hyd, THERM_hyd, DPhydDx, DPhydDy as
  // (5, NeZ, Np) tables by element level when every element of a level holds
  // the same values (horizontally homogeneous base state), else null
  const double* lvl;
  // fused: physics tendencies added to the dynamics tendency before the RK
  // update (add_phy_tend): DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp (VID
  // order), then RHOH_p; null when there are none
  const double* tp[6];
};

struct RkOut {
  double* out[5];            // drop-in: tendencies; fused: next stage state
  const double* var0[5];     // fused: q^n
  double* varTmp[5];         // fused: RK accumulator
  double* DPout;             // fused: pressure perturbation of the next stage state
  double ca, cb;             // dt*a(s+1,s), dt*b(s)
  int mode;                  // 0: first stage, 1: middle, 2: last, 3: single-stage scheme,
                             // 4: the tendency itself into out (p = 7 flat-box kernels)
};

int launch_heve_dropin(fedg_ctx* ctx, const TendIn& in, double* const out[5]);
int launch_hevi_explicit(fedg_ctx* ctx, const TendIn& in, double* const out[5]);
// the same tendency on the p = 7 flat-box layout (dry air, in.Rtot null;
// in.lvl: base state from per-level tables), FP64 (bit for bit the same) or
// with the tendency in FP32 perturbation form (fp32)
int launch_hevi_explicit7(fedg_ctx* ctx, const TendIn& in, double* const out[5], int fp32);
// drop-in tendency split for exchange overlap: interior part, then the face
// contribution added in place (DPRES: its halo tail is computed first)
int launch_heve_cco(fedg_ctx* ctx, const TendIn& in, double* const out[5]);
int launch_heve_bnd(fedg_ctx* ctx, const TendIn& in, double* const dt[5], double* DPRES);
// tracer advection (HEVE): tendency, FCT coefficient, TMAR, mass-flux save
struct TrcW {
  double w[8];  // 1D LGL integration weights
};
struct TrcLaunch {
  double* dt;   // QTRC_dt, or fct_coef when fct
  const double *Q, *MX, *MY, *MZ, *aM, *aP, *R, *tp, *J;
  const double *Dh, *DD, *DD0;  // fct: DENS_hyd, DDENS_, DDENS0_
  double c, dtt;                // fct: rk_c_ssm1, dt
  int fct, disable;
  TrcW w;
};
struct MflxArgs {
  double *FX, *FY, *FZ;      // MFLX_x/y/z_tavg (Np,NeA)
  double *alM, *alP;         // alph_dens_M / _P (NfpTot,Ne)
  const double *DD, *MX, *MY, *MZ, *DP, *Dh, *Ph;
  int rkstage, hevi;
  double twh, twv;
  int nbr;  // fields without a halo tail: exterior values through DevMesh::nbr / nbr_bc
};
int launch_trcadv(fedg_ctx* ctx, const TrcLaunch& t);
int launch_tmar(fedg_ctx* ctx, double* Q, const double* Dh, const double* DD, const double* J, const TrcW& w);
int launch_save_massflux(fedg_ctx* ctx, const MflxArgs& a);
int launch_alphdens_advtest(fedg_ctx* ctx, const MflxArgs& a);
struct TrcRk {
  double *q, *var0, *varTmp;
  const double *tend, *DD, *DD0, *Dh;
  double sig_ss, one_minus_sig_ss, gam_ss, sig_Ns, gam_Ns, c_ssm1, c_ss;
  int first, last, acc;
};
int launch_trc_rk(fedg_ctx* ctx, const TrcRk& r);
// water species of the coupled dynamics + tracer step (at most kMaxTrc)
constexpr int kMaxTrc = 8;
struct TrcThermo {
  int nq;                                                  // tracers
  double Mq[kMaxTrc], Rq[kMaxTrc], CVq[kMaxTrc], CPq[kMaxTrc];  // TRACER_MASS, _R, _CV, _CP
};
// ATMOS_THERMODYN_specific_heat on every node of elements 1..Ne
int launch_specific_heat(fedg_ctx* ctx, const double* const Q[kMaxTrc], const TrcThermo& th, double* Rtot,
                         double* CVtot, double* CPtot);
// diagnostic variables of vars_calc_diagnoseVar_lc on elements 1..Ne
enum DiagKind {
  kDiagDENS, kDiagU, kDiagV, kDiagW, kDiagUmet, kDiagVmet, kDiagPRES_diff, kDiagT, kDiagT_diff, kDiagPT,
  kDiagPT_diff, kDiagQDRY, kDiagENGK, kDiagENGP, kDiagENGI, kDiagENGT
};
struct DiagVar {
  int kind, nq;
  double* out;
  const double *DD, *MX, *MY, *MZ, *P, *Dh, *Ph, *R, *CP, *zlev, *zpos;
  const double* Q[kMaxTrc];
  double Mq[kMaxTrc], CVq[kMaxTrc], E0q[kMaxTrc];  // TRACER_MASS, TRACER_CV, TRACER_ENGI0
};
int launch_diagvar(fedg_ctx* ctx, const DiagVar& a);
struct NegFix {
  double* Q[kMaxTrc];
  double *DD, *PRES, *CV, *CP, *R, *DRHOT;  // DRHOT may be null
  const double *Dh, *Ph, *J, *zlev;
  TrcThermo th;
  TrcW w;
};
// atm_phy_mp_dgm_common_negative_fixer on elements 1..Ne
int launch_negative_fixer(fedg_ctx* ctx, const NegFix& a);
int launch_tracer_modalfilter(fedg_ctx* ctx, double* Q, const double* Dh, const double* DD, const double* MFh,
                              const double* MFv);
int launch_heve_fused(fedg_ctx* ctx, const TendIn& in, const RkOut& rk, int use_flat);
int run_stage_probe(int mode, int NeX, int NeY, int NeZ, int reps, double* ms);
// FP32 perturbation form of the fused stage (flat box, dry air; at p = 7 the
// stage reads DPRES and writes RkOut.DPout like the FP64 stage, otherwise it
// forms DPRES itself and leaves DPout)
int launch_heve_fused32(fedg_ctx* ctx, const TendIn& in, const RkOut& rk);
// DPRES of a state (dry constants when Rtot is null), as DRHOT2PRES computes it
int launch_dpres(fedg_ctx* ctx, double* DPRES, const double* DRHOT, const double* PRES_hyd,
                 const double* THERM_hyd, const double* Rtot, const double* CVtot, const double* CPtot, long n);
int launch_drhot2pres(fedg_ctx* ctx, double* PRES, double* DPRES, const double* DRHOT,
                      const double* PRES_hyd, const double* THERM_hyd, const double* Rtot,
                      const double* CVtot, const double* CPtot, long n);
int launch_halo_exchange(fedg_ctx* ctx, double* field);
int launch_apply_bc(fedg_ctx* ctx, double* MOMX, double* MOMY, double* MOMZ);
int launch_phyd_hgrad(fedg_ctx* ctx, double* DPhydDx, double* DPhydDy, const double* PRES_hyd,
                      const double* PRES_hyd_ref);
// generic streaming RK update pieces
int launch_axpy_set(fedg_ctx* ctx, double* y, const double* a, double ca, const double* x, long n);  // y = a + ca*x
int launch_axpy_acc(fedg_ctx* ctx, double* y, double ca, const double* x, long n);                   // y = y + ca*x
int launch_axpy2_acc(fedg_ctx* ctx, double* y, double ca, const double* x, double cb, const double* z, long n);
int launch_axpy2_set(fedg_ctx* ctx, double* y, const double* a, double ca, const double* x, double cb,
                     const double* z, long n);  // y = a + ca*x + cb*z
int launch_copy(fedg_ctx* ctx, double* y, const double* x, long n);
// IMEX stage advance of the five prognostics in one pass, per node the
// operations of the per-variable passes in their order:
//   s < nstage:  vt = vt + cbe Te(s) + cbi Ti(s);  q = v0, then for s2 = 1..s
//                q = q + ae(s2) Te(s2) + ai(s2) Ti(s2)
//   s = nstage:  q = vt + cbe Te(s) + cbi Ti(s)
constexpr int kImexMaxStages = 7;
struct ImexAdvance {
  double* q[5];
  double* vt[5];
  const double* v0[5];
  const double* Te[5];                      // stage s
  const double* Ti[5];
  const double* Tes[kImexMaxStages][5];     // stages 1..s (s < nstage)
  const double* Tis[kImexMaxStages][5];
  double ae[kImexMaxStages], ai[kImexMaxStages];
  double cbe, cbi;
  int s, last;
};
int launch_imex_advance(fedg_ctx* ctx, const ImexAdvance& a);
// add_phy_tend on the five tendencies T (VID order) of elements 1..Ne:
// tp = DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOH_p; PRES, Rtot, CPtot
int launch_add_phy_tend(fedg_ctx* ctx, double* const T[5], const double* const tp[6], const double* PRES,
                        const double* Rtot, const double* CPtot);
// IMEX StoreImplicit of the five prognostics in one pass: when var0 / varTmp
// are given (first stage) they take q, then q = q + impl_fac Ti
int launch_store_implicit(fedg_ctx* ctx, double* const q[5], const double* const Ti[5], double* const var0[5],
                          double* const varTmp[5], double impl_fac);
// atm_dyn_dgm_modalfilter_apply on elements 1..Ne (needs m.MFh, m.MFv)
int launch_modalfilter(fedg_ctx* ctx, double* DDENS, double* MOMX, double* MOMY, double* MOMZ, double* DRHOT,
                       int do_weight_Gsqrt);

}  // namespace fedg
