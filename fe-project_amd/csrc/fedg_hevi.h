// Vertically implicit HEVI solve (fedg_hevi.hip): workspace and launcher.
#pragma once
#include "fedg_internal.h"

namespace fedg {

// Device workspace of the implicit solve, allocated once per context.
struct HeviWork {
  double* pv[5];      // Newton iterate (Np,Ne) per variable
  double* alph;       // (2*Nfp, Ne) damping coefficient at the vertical face nodes
  double* b3;         // (3, Np*Ne) DDENS, MOMZ, DRHOT rhs
  double* buv;        // (2, Np*Ne) MOMX, MOMY rhs
  double* coef;       // (3, Np*Ne) POT, WT, DPDRHOT
  double* partial;    // 256 partial sums (residual norms)
  double partial_h[256];
  // learned initial guess (allocated on first use)
  float* X;           // (chunk, n_in) features
  float* Y;           // (chunk, n_out) predictions
  int chunk, n_in, n_out;
  double* r_plain;    // (nlines) residual norm^2 of each line at q^n
  double* r_guess;    // (nlines) at the predicted iterate
  double* flag;       // (nlines) 1 where the line fell back to q^n
};

struct MlpImpl;

// Learned initial iterate of the Newton solve: the network maps a line's
// stage state q* / s, impl_fac T_ex / t and impl_fac/dt to (q - q*) / d of
// the converged iterate q (s, t, d: per-variable scales of the state, the
// tendency feature and the increment).  Lines whose predicted iterate has a larger residual than q^n
// start from q^n (fallback).
struct HeviGuess {
  MlpImpl* mlp;
  const double* Tex[5];  // explicit tendency of the previous stage (VID order), or null
  double scale[15];      // s (VID order), then t, then d
  double dt;
  long lines;            // out: lines solved
  long fallbacks;        // out: lines that fell back to q^n
  double r2_plain;       // out: ||G(q^n)||^2 summed over the solves
  double r2_start;       // out: ||G||^2 at the iterates the lines started from
};

// The IMEX StoreImplicit of the HEVI stepper (CMN/scale_timeint_rk.F90,
// StoreImplicit) done by the pass that forms the implicit tendency: q = q +
// impl_fac T per node after T is written, and, when var0 / varTmp are given
// (first stage), their copies of q before it.  q is the solve's stage state
// (the q argument of hevi_cal_vi).
struct HeviStore {
  double* q[5];
  double* var0[5];    // or null
  double* varTmp[5];  // or null
};

int hevi_work_get(fedg_ctx* ctx, HeviWork** w);
void hevi_work_free(fedg_ctx* ctx);

// cal_vi of FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965 on the
// context's stream.  tend, q, var0 in VID order (DDENS, DRHOT, MOMZ, MOMX,
// MOMY).  impl_fac == 0: the implicit operator itself.  max_it Newton
// iterations (the reference: 1), stopping early once the residual norm has
// dropped by rtol when rtol > 0; resid (host, max_it) receives the norms.
// store (impl_fac != 0 only): the stepper's StoreImplicit, fused into the
// pass that forms the tendency.
int hevi_cal_vi(fedg_ctx* ctx, HeviWork* w, double* const tend[5], const double* const q[5],
                const double* const var0[5], const double* DENS_hyd, const double* PRES_hyd, const double* Rtot,
                const double* CVtot, const double* CPtot, double impl_fac, int max_it, double rtol, double* resid,
                int* it_done, HeviGuess* guess = nullptr, const HeviStore* store = nullptr,
                const double* RHOT_hyd = nullptr, int alph_ready = 0);
// alph_ready: the workspace's vertical-face damping coefficients (alph) were
// formed from this same var0 by an earlier call on the context (a later
// stage of the same step), so the residual kernels load them
// RHOT_hyd = PRES00/Rdry (PRES_hyd/PRES00)^(CVdry/CPdry) over n nodes: the
// value the residual kernels form from PRES_hyd, bit for bit (their RHOT_hyd
// argument of hevi_cal_vi)
int launch_rhot_hyd(fedg_ctx* ctx, double* RHOT_hyd, const double* PRES_hyd, long n);

// Features and targets of nrows lines (line = (line0 + r*stride) mod nlines)
// of a converged solve (iterate in w->pv) into X (nrows x n_in) and Y
// (nrows x n_out), device fp32; the network's input/output layout.
int hevi_harvest(fedg_ctx* ctx, HeviWork* w, const double* const q[5], const double* const var0[5],
                 const double* const Tex[5], const double scale[15], double impl_fac, double dt, long line0,
                 long stride, int nrows, float* X, float* Y);

// Device address of the no-allocation workspace on `device` (the current
// device), looked up once per device; called by fedg_ctx_create.
int noalloc_scratch_setup(int device);

}  // namespace fedg
