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
};

int hevi_work_get(fedg_ctx* ctx, HeviWork** w);
void hevi_work_free(fedg_ctx* ctx);

// cal_vi of FDS/scale_atm_dyn_dgm_nonhydro3d_rhot_hevi.F90:772-965 on the
// context's stream.  tend, q, var0 in VID order (DDENS, DRHOT, MOMZ, MOMX,
// MOMY).  impl_fac == 0: the implicit operator itself.  max_it Newton
// iterations (the reference: 1), stopping early once the residual norm has
// dropped by rtol when rtol > 0; resid (host, max_it) receives the norms.
int hevi_cal_vi(fedg_ctx* ctx, HeviWork* w, double* const tend[5], const double* const q[5],
                const double* const var0[5], const double* DENS_hyd, const double* PRES_hyd, const double* Rtot,
                const double* CVtot, const double* CPtot, double impl_fac, int max_it, double rtol, double* resid,
                int* it_done);

}  // namespace fedg
