// C-ABI layer: device context, uploads, drop-in entry points, time
// integrator and the fused HEVE stepper.  Declarations and reference
// citations are in include/fedg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fedg_builder.h"
#include "fedg_internal.h"
#include "fedg_hevi.h"
#include "fedg_mlp.h"

namespace fedg {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

namespace {

template <typename T>
int dev_upload(T** dst, const T* src, size_t count) {
  *dst = nullptr;
  if (count == 0) return FEDG_OK;
  hipError_t e = hipMalloc((void**)dst, sizeof(T) * count);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    *dst = nullptr;
    return FEDG_ERR_NOMEM;
  }
  if (src) FEDG_HIP_CHECK(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
  return FEDG_OK;
}

#define TRY(x)                       \
  do {                               \
    int rc_ = (x);                   \
    if (rc_ != FEDG_OK) return rc_;  \
  } while (0)

void free_mesh(DevMesh& m) {
  void* ptrs[] = {m.E11, m.E22, m.E33, m.Gsqrt, m.G13, m.G23, m.GsqrtH, m.G11, m.G12, m.G22, m.nx, m.ny, m.nz,
                  m.Fscale, m.vmapM, m.vmapP, m.EMap3Dto2D, m.halo_src, m.velbc, m.nbr, m.nbr_bc, m.D1D,
                  m.Lift, m.Itr, m.liftT, m.MFh, m.MFv};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  memset(&m, 0, sizeof(m));
}

bool all_equal(const double* a, size_t n, double v) {
  for (size_t t = 0; t < n; ++t)
    if (a[t] != v) return false;
  return true;
}

}  // namespace
}  // namespace fedg

using namespace fedg;

struct fedg_tint {
  fedg_ctx* ctx;
  int nstage, tend_buf_size, imex;
  double dt;
  double a_ex[4][4], b_ex[4], a_im[4][4], b_im[4];
  int indmap[4];
  double* var0;    // (Np,NeA,5)
  double* varTmp;  // (Np,NeA,5)
  double* tend_ex; // (Np,NeA,5,nbuf)
  double* tend_im; // (Np,NeA,5,nbuf) or null
};

struct fedg_heve {
  fedg_ctx* ctx;
  fedg_tint tab;            // Butcher table (no buffers)
  int flat;
  double* block;            // one allocation for every device array below
  double* buf[3][5];        // stage states (q^n and two intermediates)
  double* DP[3];            // pressure perturbation of each stage state
  double* VT[5];            // RK accumulator
  int q0;                   // index of q^n in buf
  double *DENS_hyd, *PRES_hyd, *THERM_hyd, *DPhydDx, *DPhydDy, *CORIOLIS;
  double *Rtot, *CVtot, *CPtot;
  double* pres_scratch;     // (Np,Ne) x2 for PRES/DPRES
  int fp32;                 // stage tendency in the FP32 perturbation form
  int mfilter;              // modal filter after every step
  double* lvl;              // per-level base-state tables (horizontally homogeneous base state) or null
  // time-averaged mass fluxes for the tracers (QA > 0): MFLX_x/y/z_tavg and
  // ALPH_DENS_M/P_tavg saved at every stage, or null
  double *mflx[3], *alph[2];
  // physics tendencies (DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOH_p) or null
  const double* tp[6];
};

namespace {

// Horizontally homogeneous base state (every element of a level holds the
// same values, bit for bit): the p = 7 flat-box kernels read DENS_hyd,
// PRES_hyd, THERM_hyd, DPhydDx, DPhydDy (host arrays f) from per-level
// (5, NeZ, Np) tables, which stay in L2, instead of five full fields.
// *out stays null when the base state varies horizontally.
int level_tables(const DevMesh& m, const double* const f[5], double** out) {
  *out = nullptr;
  if (!(m.NeX > 0 && m.NeY > 0 && m.NeZ > 0 && m.Ne == m.NeX * m.NeY * m.NeZ)) return FEDG_OK;
  const size_t Np = (size_t)m.Np, per = (size_t)m.NeX * m.NeY;
  bool hom = true;
  for (int a = 0; a < 5 && hom; ++a)
    for (size_t ke = 0; ke < (size_t)m.Ne && hom; ++ke) {
      const size_t rep = (ke / per) * per;
      hom = memcmp(f[a] + ke * Np, f[a] + rep * Np, Np * sizeof(double)) == 0;
    }
  if (!hom) return FEDG_OK;
  std::vector<double> tab((size_t)5 * m.NeZ * Np);
  for (int a = 0; a < 5; ++a)
    for (int kz = 0; kz < m.NeZ; ++kz)
      memcpy(&tab[((size_t)a * m.NeZ + kz) * Np], f[a] + (size_t)kz * per * Np, Np * sizeof(double));
  return dev_upload<double>(out, tab.data(), tab.size());
}

int butcher(const char* name, double dt, fedg_tint* t) {
  memset(t->a_ex, 0, sizeof(t->a_ex));
  memset(t->a_im, 0, sizeof(t->a_im));
  memset(t->b_ex, 0, sizeof(t->b_ex));
  memset(t->b_im, 0, sizeof(t->b_im));
  t->dt = dt;
  // CMN/scale_timeint_rk_butcher_tab.F90:60-324
  if (!strcmp(name, "ERK_4s4o") || !strcmp(name, "ERK_RK4")) {
    t->nstage = 4;
    t->tend_buf_size = 1;
    t->imex = 0;
    t->a_ex[1][0] = 0.5;
    t->a_ex[2][1] = 0.5;
    t->a_ex[3][2] = 1.0;
    t->b_ex[0] = 1.0 / 6.0;
    t->b_ex[1] = 2.0 / 6.0;
    t->b_ex[2] = 2.0 / 6.0;
    t->b_ex[3] = 1.0 / 6.0;
    for (int s = 0; s < 4; ++s) t->indmap[s] = 0;
  } else if (!strcmp(name, "ERK_1s1o") || !strcmp(name, "ERK_Euler")) {
    t->nstage = 1;
    t->tend_buf_size = 1;
    t->imex = 0;
    t->b_ex[0] = 1.0;
    t->indmap[0] = 0;
  } else if (!strcmp(name, "IMEX_ARK232")) {
    const double alp = (3.0 + 2.0 * sqrt(2.0)) / 6.0;
    const double gam = 1.0 - 1.0 / sqrt(2.0);
    const double del = 1.0 / (2.0 * sqrt(2.0));
    t->nstage = 3;
    t->tend_buf_size = 3;
    t->imex = 1;
    t->a_ex[1][0] = 2.0 * gam;
    t->a_ex[2][0] = 1.0 - alp;
    t->a_ex[2][1] = alp;
    t->b_ex[0] = del;
    t->b_ex[1] = del;
    t->b_ex[2] = gam;
    t->a_im[1][0] = gam;
    t->a_im[1][1] = gam;
    t->a_im[2][0] = del;
    t->a_im[2][1] = del;
    t->a_im[2][2] = gam;
    for (int s = 0; s < 3; ++s) t->b_im[s] = t->b_ex[s];
    for (int s = 0; s < 3; ++s) t->indmap[s] = s;
  } else if (!strcmp(name, "IMEX_ARK324")) {
    t->nstage = 4;
    t->tend_buf_size = 4;
    t->imex = 1;
    t->a_ex[1][0] = 1767732205903.0 / 2027836641118.0;
    t->a_ex[2][0] = 5535828885825.0 / 10492691773637.0;
    t->a_ex[2][1] = 788022342437.0 / 10882634858940.0;
    t->a_ex[3][0] = 6485989280629.0 / 16251701735622.0;
    t->a_ex[3][1] = -4246266847089.0 / 9704473918619.0;
    t->a_ex[3][2] = 10755448449292.0 / 10357097424841.0;
    t->b_ex[0] = 1471266399579.0 / 7840856788654.0;
    t->b_ex[1] = -4482444167858.0 / 7529755066697.0;
    t->b_ex[2] = 11266239266428.0 / 11593286722821.0;
    t->b_ex[3] = 1767732205903.0 / 4055673282236.0;
    t->a_im[1][0] = 1767732205903.0 / 4055673282236.0;
    t->a_im[1][1] = 1767732205903.0 / 4055673282236.0;
    t->a_im[2][0] = 2746238789719.0 / 10658868560708.0;
    t->a_im[2][1] = -640167445237.0 / 6845629431997.0;
    t->a_im[2][2] = 1767732205903.0 / 4055673282236.0;
    for (int s = 0; s < 4; ++s) t->a_im[3][s] = t->b_ex[s];
    for (int s = 0; s < 4; ++s) t->b_im[s] = t->b_ex[s];
    for (int s = 0; s < 4; ++s) t->indmap[s] = s;
  } else {
    set_error(std::string("unsupported time integration scheme: ") + name);
    return FEDG_ERR_ARG;
  }
  return FEDG_OK;
}

}  // namespace

extern "C" {

const char* fedg_last_error(void) { return g_err.c_str(); }
const char* fedg_version(void) { return "fedg 0.1 (gfx950)"; }

int fedg_dev_alloc(void** ptr, size_t bytes) {
  if (!ptr) return FEDG_ERR_ARG;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    *ptr = nullptr;
    return FEDG_ERR_NOMEM;
  }
  return FEDG_OK;
}
int fedg_dev_free(void* ptr) {
  if (ptr) FEDG_HIP_CHECK(hipFree(ptr));
  return FEDG_OK;
}
int fedg_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  FEDG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FEDG_OK;
}
int fedg_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  FEDG_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FEDG_OK;
}
int fedg_memset_dev(void* dst, int value, size_t bytes) {
  FEDG_HIP_CHECK(hipMemset(dst, value, bytes));
  return FEDG_OK;
}
int fedg_device_sync(void) {
  FEDG_HIP_CHECK(hipDeviceSynchronize());
  return FEDG_OK;
}
int fedg_stage_probe(int mode, int NeX, int NeY, int NeZ, int reps, double* ms) {
  return fedg::run_stage_probe(mode, NeX, NeY, NeZ, reps, ms);
}
int fedg_device_name(char* buf, int buflen) {
  int dev = 0;
  FEDG_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  FEDG_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (prop.name[0])
    snprintf(buf, (size_t)buflen, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  else  // some runtimes leave the marketing name empty
    snprintf(buf, (size_t)buflen, "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
  return FEDG_OK;
}

// ---------------------------------------------------------------------------
int fedg_refelem_tables(int porder, double* x1d, double* w1d, double* D1D, double* V1D, double* invV1D,
                        double* invM1D, double* Lift_mat, double* IntrpMat_VPOrdM1_tr, int32_t* IndexH2Dto3D,
                        int32_t* IndexH2Dto3D_bnd) {
  RefElem e;
  try {
    build_refelem(porder, e);
  } catch (const std::exception& ex) {
    set_error(ex.what());
    return FEDG_ERR_ARG;
  }
  auto cp = [](auto* dst, const auto& v) {
    if (dst) memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(x1d, e.x1d);
  cp(w1d, e.w1d);
  cp(D1D, e.D1D);
  cp(V1D, e.V1D);
  cp(invV1D, e.invV1D);
  cp(invM1D, e.invM1D);
  cp(Lift_mat, e.Lift_mat);
  cp(IntrpMat_VPOrdM1_tr, e.IntrpMat_VPOrdM1_tr);
  cp(IndexH2Dto3D, e.IndexH2Dto3D);
  cp(IndexH2Dto3D_bnd, e.IndexH2Dto3D_bnd);
  return FEDG_OK;
}

// ModalFilter_Init_line with get_exp_filter (ELM/scale_element_modalfilter.F90:60-90, 196-236):
// FilterMat = V diag(sigma) V^-1, sigma_p = exp(-alpha ((eta - etac)/(1 - etac))^ord)
// for eta = (p-1)/porder > etac (p > 1), else 1.  F is (n1,n1) column-major.
int fedg_modalfilter_matrix(int porder, double etac, double alpha, int ord, double* F) {
  if (!F || ord < 0 || !(etac < 1.0)) {
    set_error("modal filter: bad arguments (F, ord >= 0, etac < 1)");
    return FEDG_ERR_ARG;
  }
  RefElem e;
  try {
    build_refelem(porder, e);
  } catch (const std::exception& ex) {
    set_error(ex.what());
    return FEDG_ERR_ARG;
  }
  const int n1 = e.n1;
  std::vector<double> sigma(n1, 1.0);
  for (int p = 0; p < n1; ++p) {
    const double eta = (double)p / (double)porder;
    if (eta > etac && p != 0) {
      const double x = (eta - etac) / (1.0 - etac);
      double r = 1.0, b = x;  // x**ord by binary powering
      for (int k = ord; k > 0; k >>= 1) {
        if (k & 1) r *= b;
        b *= b;
      }
      sigma[p] = std::exp(-alpha * r);
    }
  }
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i) {
      double acc = 0.0;
      for (int l = 0; l < n1; ++l) acc += e.V1D[i + n1 * l] * (sigma[l] * e.invV1D[l + n1 * j]);
      F[i + n1 * j] = acc;
    }
  return FEDG_OK;
}

int fedg_box_get_sizes(const fedg_box_desc* d, fedg_box_sizes* s) {
  if (!d || !s || d->porder < 1 || d->porder > 15 || d->NeX < 1 || d->NeY < 1 || d->NeZ < 1) {
    set_error("invalid box description");
    return FEDG_ERR_ARG;
  }
  const int n1 = d->porder + 1;
  s->n1 = n1;
  s->Np = n1 * n1 * n1;
  s->Nfp_h = n1 * n1;
  s->Nfp_v = n1 * n1;
  s->NfpTot = 6 * n1 * n1;
  s->Ne = d->NeX * d->NeY * d->NeZ;
  s->NeA = s->Ne + 2 * (d->NeX + d->NeY) * d->NeZ + 2 * d->NeX * d->NeY;
  s->Ne2D = d->NeX * d->NeY;
  s->Ne2DA = d->NeX * d->NeY + 2 * (d->NeX + d->NeY);
  s->nB = 2 * (d->NeX + d->NeY) * d->NeZ * s->Nfp_h + 2 * d->NeX * d->NeY * s->Nfp_v;
  if ((double)s->Np * (double)s->NeA > 2.0e9) {
    set_error("mesh too large for 32-bit node indices");
    return FEDG_ERR_ARG;
  }
  return FEDG_OK;
}

int fedg_box_build(const fedg_box_desc* d, fedg_box_arrays* out) {
  fedg_box_sizes s;
  TRY(fedg_box_get_sizes(d, &s));
  RefElem e;
  BoxMesh m;
  try {
    build_refelem(d->porder, e);
    BoxDesc bd{d->porder, d->NeX, d->NeY, d->NeZ, d->xmin, d->xmax, d->ymin, d->ymax, d->zmin, d->zmax,
               d->periodic_x, d->periodic_y, d->periodic_z};
    if (build_box(bd, e, m) != 0) {
      set_error("box mesh construction failed");
      return FEDG_ERR_ARG;
    }
  } catch (const std::exception& ex) {
    set_error(ex.what());
    return FEDG_ERR_ARG;
  }
  auto cp = [](auto* dst, const auto& v) {
    if (dst) memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(out->pos_en, m.pos_en);
  cp(out->J, m.J);
  cp(out->Escale, m.Escale);
  cp(out->sJ, m.sJ);
  cp(out->Fscale, m.Fscale);
  cp(out->normal_fn, m.normal_fn);
  cp(out->vmapM, m.vmapM);
  cp(out->vmapP, m.vmapP);
  cp(out->vmapB, m.vmapB);
  cp(out->halo_src, m.halo_src);
  cp(out->halo_face, m.halo_face);
  cp(out->EMap3Dto2D, m.EMap3Dto2D);
  cp(out->Gsqrt, m.Gsqrt);
  cp(out->GI3, m.GI3);
  cp(out->GsqrtH, m.GsqrtH);
  return FEDG_OK;
}

int fedg_box_topography_metrics(const fedg_box_desc* d, const double* topo, double zTop, double* Gsqrt, double* GI3,
                                double* zlev) {
  if (!d || !topo || !(zTop > 0.0)) return FEDG_ERR_ARG;
  fedg_box_sizes s;
  TRY(fedg_box_get_sizes(d, &s));
  if (!(d->zmin == 0.0) || !(zTop >= d->zmax)) {
    set_error("terrain-following coordinate: the box must start at zeta = 0 and zTop must be at or above its top");
    return FEDG_ERR_ARG;
  }
  RefElem e;
  BoxMesh m;
  std::vector<double> zl;
  try {
    build_refelem(d->porder, e);
    BoxDesc bd{d->porder, d->NeX, d->NeY, d->NeZ, d->xmin, d->xmax, d->ymin, d->ymax, d->zmin, d->zmax,
               d->periodic_x, d->periodic_y, d->periodic_z};
    if (build_box(bd, e, m) != 0 || set_topography(bd, e, topo, zTop, m, zl) != 0) {
      set_error("terrain-following metrics: construction failed");
      return FEDG_ERR_ARG;
    }
  } catch (const std::exception& ex) {
    set_error(ex.what());
    return FEDG_ERR_ARG;
  }
  for (size_t t = 0; t < m.Gsqrt.size(); ++t)
    if (!(m.Gsqrt[t] > 0.0)) {
      set_error("terrain-following coordinate: the topography reaches zTop (GsqrtV <= 0)");
      return FEDG_ERR_ARG;
    }
  if (Gsqrt) memcpy(Gsqrt, m.Gsqrt.data(), m.Gsqrt.size() * sizeof(double));
  if (GI3) memcpy(GI3, m.GI3.data(), m.GI3.size() * sizeof(double));
  if (zlev) memcpy(zlev, zl.data(), zl.size() * sizeof(double));
  return FEDG_OK;
}

// ---------------------------------------------------------------------------
int fedg_ctx_create(const fedg_mesh_desc* d, fedg_ctx** out) {
  if (!d || !out) return FEDG_ERR_ARG;
  *out = nullptr;
  if (d->porder < 1 || d->porder > 7) {
    set_error("polynomial order not built (supported: p = 1..7)");
    return FEDG_ERR_UNSUPPORTED;
  }
  if (!d->Escale || !d->Gsqrt || !d->GI3 || !d->GsqrtH || !d->normal_fn || !d->Fscale || !d->vmapM || !d->vmapP ||
      !d->EMap3Dto2D || !d->D1D || !d->Lift_mat || !d->IntrpMat_VPOrdM1_tr || d->Ne < 1 || d->NeA < d->Ne ||
      (d->nB > 0 && (!d->halo_src || !d->velbc))) {
    set_error("fedg_ctx_create: missing array or inconsistent sizes");
    return FEDG_ERR_ARG;
  }
  fedg_ctx* ctx = new fedg_ctx();
  memset(&ctx->m, 0, sizeof(ctx->m));
  DevMesh& m = ctx->m;
  const int n1 = d->porder + 1, Np = n1 * n1 * n1, Nfp = n1 * n1, NfpTot = 6 * Nfp;
  m.n1 = n1;
  m.Np = Np;
  m.Nfp = Nfp;
  m.NfpTot = NfpTot;
  m.Ne = d->Ne;
  m.NeA = d->NeA;
  m.Ne2D = d->Ne2D;
  m.Ne2DA = d->Ne2DA;
  m.nB = d->nB;
  m.NeX = d->NeX;
  m.NeY = d->NeY;
  m.NeZ = d->NeZ;
  ctx->c.GRAV = d->GRAV != 0.0 ? d->GRAV : FEDG_DEFAULT_GRAV;
  ctx->c.Rdry = d->Rdry != 0.0 ? d->Rdry : FEDG_DEFAULT_RDRY;
  ctx->c.CPdry = d->CPdry != 0.0 ? d->CPdry : FEDG_DEFAULT_CPDRY;
  ctx->c.CVdry = d->CVdry != 0.0 ? d->CVdry : FEDG_DEFAULT_CVDRY;
  ctx->c.PRES00 = d->PRES00 != 0.0 ? d->PRES00 : FEDG_DEFAULT_PRES00;
  if (hipGetDevice(&ctx->device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    set_error("no usable HIP device");
    delete ctx;
    return FEDG_ERR_HIP;
  }
  if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
      ctx->ncu < 1)
    ctx->ncu = 256;
  {
    // the no-allocation workspace's address on this (current) device
    const int rcs = noalloc_scratch_setup(ctx->device);
    if (rcs != FEDG_OK) {
      (void)hipStreamDestroy(ctx->stream);
      delete ctx;
      return rcs;
    }
  }
  const size_t NpNe = (size_t)Np * m.Ne, NpNeA = (size_t)Np * m.NeA, nf = (size_t)NfpTot * m.Ne;
  const size_t n2d = (size_t)Nfp * m.Ne2D;
  int rc = FEDG_OK;
  auto up = [&](auto** dst, const auto* src, size_t n) {
    if (rc == FEDG_OK) rc = dev_upload(dst, src, n);
  };
  up(&m.E11, d->Escale, NpNe);
  up(&m.E22, d->Escale + NpNe * 4, NpNe);
  up(&m.E33, d->Escale + NpNe * 8, NpNe);
  up(&m.Gsqrt, d->Gsqrt, NpNeA);
  up(&m.G13, d->GI3, NpNeA);
  up(&m.G23, d->GI3 + NpNeA, NpNeA);
  up(&m.GsqrtH, d->GsqrtH, n2d);
  std::vector<double> ones(n2d, 1.0), zeros(n2d, 0.0);
  up(&m.G11, d->G11 ? d->G11 : ones.data(), n2d);
  up(&m.G12, d->G12 ? d->G12 : zeros.data(), n2d);
  up(&m.G22, d->G22 ? d->G22 : ones.data(), n2d);
  up(&m.nx, d->normal_fn, nf);
  up(&m.ny, d->normal_fn + nf, nf);
  up(&m.nz, d->normal_fn + 2 * nf, nf);
  up(&m.Fscale, d->Fscale, nf);
  // integer maps to 0-based device copies
  std::vector<int32_t> vm(nf), vp(nf), em(m.Ne), hs(m.nB > 0 ? m.nB : 0), nbr(nf);
  std::vector<int8_t> nbc(nf, 0);
  bool ok_maps = true;
  for (size_t t = 0; t < nf; ++t) {
    vm[t] = d->vmapM[t] - 1;
    vp[t] = d->vmapP[t] - 1;
    if (vm[t] < 0 || (size_t)vm[t] >= NpNe || vp[t] < 0 || (size_t)vp[t] >= NpNe + (size_t)m.nB) ok_maps = false;
  }
  for (int t = 0; t < m.Ne; ++t) {
    em[t] = d->EMap3Dto2D[t] - 1;
    if (em[t] < 0 || em[t] >= m.Ne2D) ok_maps = false;
  }
  for (int b = 0; b < m.nB; ++b) {
    hs[b] = d->halo_src[b] - 1;
    if (hs[b] < 0 || (size_t)hs[b] >= NpNe) ok_maps = false;
  }
  if (!ok_maps) {
    set_error("fedg_ctx_create: index map out of range");
    free_mesh(m);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FEDG_ERR_ARG;
  }
  // exterior value of every face node resolved to an interior node, with the
  // velocity BC that applyBC would apply in the halo slot
  for (size_t t = 0; t < nf; ++t) {
    if ((size_t)vp[t] < NpNe) {
      nbr[t] = vp[t];
    } else {
      const int b = (int)((size_t)vp[t] - NpNe);
      nbr[t] = hs[b];
      nbc[t] = (int8_t)d->velbc[b];
    }
  }
  up(&m.vmapM, vm.data(), nf);
  up(&m.vmapP, vp.data(), nf);
  up(&m.EMap3Dto2D, em.data(), (size_t)m.Ne);
  up(&m.halo_src, hs.data(), (size_t)m.nB);
  up(&m.velbc, d->velbc, (size_t)m.nB);
  up(&m.nbr, nbr.data(), nf);
  up(&m.nbr_bc, nbc.data(), nf);
  up(&m.D1D, d->D1D, (size_t)n1 * n1);
  up(&m.Lift, d->Lift_mat, (size_t)Np * 6);
  up(&m.Itr, d->IntrpMat_VPOrdM1_tr, (size_t)n1 * n1);
  {
    // tensor-product structure of Lift_mat: face f's entry depends only on
    // the node's index along the face's normal axis (y for faces 1,3; x for
    // faces 2,4; z for faces 5,6)
    std::vector<double> lt((size_t)n1 * 6);
    const double* L = d->Lift_mat;
    auto line = [&](int f, int i, int j, int k) { return f == 0 || f == 2 ? j : f == 1 || f == 3 ? i : k; };
    for (int f = 0; f < 6; ++f)
      for (int x = 0; x < n1; ++x) {
        const int i = (f == 1 || f == 3) ? x : 0, j = (f == 0 || f == 2) ? x : 0, k = (f >= 4) ? x : 0;
        lt[x + (size_t)n1 * f] = L[i + n1 * (j + n1 * k) + (size_t)Np * f];
      }
    bool tensor = true;
    for (int f = 0; f < 6 && tensor; ++f)
      for (int k = 0; k < n1 && tensor; ++k)
        for (int j = 0; j < n1 && tensor; ++j)
          for (int i = 0; i < n1; ++i)
            if (L[i + n1 * (j + n1 * k) + (size_t)Np * f] != lt[line(f, i, j, k) + (size_t)n1 * f]) {
              tensor = false;
              break;
            }
    m.lift_tensor = tensor ? 1 : 0;
    up(&m.liftT, lt.data(), lt.size());
  }
  if (rc != FEDG_OK) {
    free_mesh(m);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }

  // Detect the flat-terrain uniform box: unit Gsqrt/GsqrtH, zero GI3, metric
  // tensor (1,0,1), axis-aligned unit normals, Escale diagonal and Fscale
  // constant over the mesh, and a face map that is the box's own.
  m.flat = 0;
  if (d->NeX > 0 && d->NeY > 0 && d->NeZ > 0 && d->NeX * d->NeY * d->NeZ == d->Ne && d->Ne2D == d->NeX * d->NeY) {
    bool flat = all_equal(d->Gsqrt, NpNeA, 1.0) && all_equal(d->GI3, 2 * NpNeA, 0.0) &&
                all_equal(d->GsqrtH, n2d, 1.0) && (!d->G11 || all_equal(d->G11, n2d, 1.0)) &&
                (!d->G12 || all_equal(d->G12, n2d, 0.0)) && (!d->G22 || all_equal(d->G22, n2d, 1.0));
    const double E0[3] = {d->Escale[0], d->Escale[NpNe * 4], d->Escale[NpNe * 8]};
    for (int a = 0; a < 3 && flat; ++a)
      for (int b = 0; b < 3 && flat; ++b) {
        const double* E = d->Escale + NpNe * (a + 3 * b);
        flat = all_equal(E, NpNe, a == b ? E0[a] : 0.0);
      }
    double F0[6];
    for (int f = 0; f < 6; ++f) F0[f] = d->Fscale[f * Nfp];
    const double nref[6][3] = {{0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 1}};
    for (size_t t = 0; t < nf && flat; ++t) {
      const int f = (int)((t % NfpTot) / Nfp);
      if (d->Fscale[t] != F0[f]) flat = false;
      for (int dd = 0; dd < 3; ++dd)
        if (d->normal_fn[t + nf * dd] != nref[f][dd]) flat = false;
    }
    // the box's own face map, node by node (the flat kernels form the own and
    // the neighbour's face node arithmetically), and one velocity BC per
    // closed boundary
    const int n1 = m.n1, n2 = n1 * n1, Nn = n1 - 1;
    auto face_node_h = [&](int f, int a) {
      const int a0 = a % n1, a1 = a / n1;
      switch (f) {
        case 0: return a0 + n2 * a1;
        case 1: return Nn + n1 * a0 + n2 * a1;
        case 2: return a0 + n1 * Nn + n2 * a1;
        case 3: return n1 * a0 + n2 * a1;
        case 4: return a;
        default: return a + n2 * Nn;
      }
    };
    const int opp[6] = {2, 3, 0, 1, 5, 4};
    int bnd_bc[6] = {-1, -1, -1, -1, -1, -1};
    int periodic[3] = {1, 1, 1};
    for (int ke = 0; ke < m.Ne && flat; ++ke) {
      const int ei = ke % m.NeX, ej = (ke / m.NeX) % m.NeY, ek = ke / (m.NeX * m.NeY);
      for (int fp = 0; fp < NfpTot && flat; ++fp) {
        const size_t t = fp + (size_t)NfpTot * ke;
        const int f = fp / Nfp;
        const int a = fp - f * Nfp;
        if (vm[t] != Np * ke + face_node_h(f, a)) flat = false;
        int ni = ei, nj = ej, nk = ek, axis;
        switch (f) {
          case 0: nj -= 1; axis = 1; break;
          case 1: ni += 1; axis = 0; break;
          case 2: nj += 1; axis = 1; break;
          case 3: ni -= 1; axis = 0; break;
          case 4: nk -= 1; axis = 2; break;
          default: nk += 1; axis = 2; break;
        }
        const bool outside = ni < 0 || ni >= m.NeX || nj < 0 || nj >= m.NeY || nk < 0 || nk >= m.NeZ;
        if (!outside) {
          if (vp[t] != Np * (ni + m.NeX * (nj + m.NeY * nk)) + face_node_h(opp[f], a)) flat = false;
          continue;
        }
        if ((size_t)vp[t] < NpNe) {
          flat = false;
          continue;
        }
        const int own_elem = ke;
        const int src_elem = nbr[t] / Np;
        if (src_elem == own_elem) {
          periodic[axis] = 0;
          if (nbr[t] != vm[t]) flat = false;
          if (bnd_bc[f] < 0) bnd_bc[f] = nbc[t];
          else if (bnd_bc[f] != nbc[t]) flat = false;
        } else {
          if (nbc[t] != FEDG_BND_NONE) flat = false;
          const int wi = (ni + m.NeX) % m.NeX, wj = (nj + m.NeY) % m.NeY, wk = (nk + m.NeZ) % m.NeZ;
          if (nbr[t] != Np * (wi + m.NeX * (wj + m.NeY * wk)) + face_node_h(opp[f], a)) flat = false;
        }
      }
    }
    if (flat) {
      m.flat = 1;
      for (int a = 0; a < 3; ++a) m.fE[a] = E0[a];
      for (int f = 0; f < 6; ++f) {
        m.fF[f] = F0[f];
        m.bnd_bc[f] = bnd_bc[f] < 0 ? FEDG_BND_NONE : bnd_bc[f];
      }
      for (int a = 0; a < 3; ++a) m.periodic[a] = periodic[a];
    }
  }
  *out = ctx;
  return FEDG_OK;
}

int fedg_ctx_setup_modalfilter(fedg_ctx* ctx, const double* MFilter_h1D, const double* MFilter_v1D) {
  if (!ctx || !MFilter_h1D || !MFilter_v1D) return FEDG_ERR_ARG;
  DevMesh& m = ctx->m;
  const size_t nn = (size_t)m.n1 * m.n1;
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (m.MFh) (void)hipFree(m.MFh);
  if (m.MFv) (void)hipFree(m.MFv);
  m.MFh = m.MFv = nullptr;
  int rc = dev_upload(&m.MFh, MFilter_h1D, nn);
  if (rc == FEDG_OK) rc = dev_upload(&m.MFv, MFilter_v1D, nn);
  return rc;
}

int fedg_modalfilter_apply(fedg_ctx* ctx, double* DDENS, double* MOMX, double* MOMY, double* MOMZ, double* DRHOT,
                           int do_weight_Gsqrt) {
  if (!ctx || !DDENS || !MOMX || !MOMY || !MOMZ || !DRHOT) return FEDG_ERR_ARG;
  TRY(launch_modalfilter(ctx, DDENS, MOMX, MOMY, MOMZ, DRHOT, do_weight_Gsqrt));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_ctx_destroy(fedg_ctx* ctx) {
  if (!ctx) return FEDG_OK;
  hevi_work_free(ctx);
  free_mesh(ctx->m);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FEDG_OK;
}

int fedg_ctx_is_flat(const fedg_ctx* ctx) { return ctx ? ctx->m.flat : 0; }

int fedg_ctx_sizes(const fedg_ctx* ctx, fedg_box_sizes* s) {
  if (!ctx || !s) return FEDG_ERR_ARG;
  const DevMesh& m = ctx->m;
  s->n1 = m.n1;
  s->Np = m.Np;
  s->Nfp_h = m.Nfp;
  s->Nfp_v = m.Nfp;
  s->NfpTot = m.NfpTot;
  s->Ne = m.Ne;
  s->NeA = m.NeA;
  s->Ne2D = m.Ne2D;
  s->Ne2DA = m.Ne2DA;
  s->nB = m.nB;
  return FEDG_OK;
}

// ---------------------------------------------------------------------------
int fedg_drhot2pres(fedg_ctx* ctx, double* PRES, double* DPRES, const double* DRHOT, const double* PRES_hyd,
                    const double* THERM_hyd, const double* Rtot, const double* CVtot, const double* CPtot) {
  if (!ctx || !PRES || !DPRES || !DRHOT || !PRES_hyd || !THERM_hyd || !Rtot || !CVtot || !CPtot)
    return FEDG_ERR_ARG;
  TRY(launch_drhot2pres(ctx, PRES, DPRES, DRHOT, PRES_hyd, THERM_hyd, Rtot, CVtot, CPtot,
                        (long)ctx->m.Np * ctx->m.Ne));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_halo_exchange(fedg_ctx* ctx, int nfields, double* const* fields) {
  if (!ctx || nfields < 0 || (nfields > 0 && !fields)) return FEDG_ERR_ARG;
  for (int v = 0; v < nfields; ++v) TRY(launch_halo_exchange(ctx, fields[v]));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_apply_bc_prgvars(fedg_ctx* ctx, double* DDENS, double* MOMX, double* MOMY, double* MOMZ, double* THERM,
                          const double* DENS_hyd, const double* PRES_hyd) {
  (void)DDENS;
  (void)THERM;
  (void)DENS_hyd;
  (void)PRES_hyd;
  if (!ctx || !MOMX || !MOMY || !MOMZ) return FEDG_ERR_ARG;
  TRY(launch_apply_bc(ctx, MOMX, MOMY, MOMZ));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_heve_cal_tend(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                       double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                       const double* MOMZ_, const double* DRHOT_, const double* DPRES_, const double* DENS_hyd,
                       const double* PRES_hyd, const double* PRES_hyd_ref, const double* THERM_hyd,
                       const double* CORIOLIS, const double* Rtot, const double* CVtot, const double* CPtot,
                       const double* DPhydDx, const double* DPhydDy) {
  (void)PRES_hyd_ref;
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ ||
      !DRHOT_ || !DPRES_ || !DENS_hyd || !PRES_hyd || !THERM_hyd || !CORIOLIS || !DPhydDx || !DPhydDy)
    return FEDG_ERR_ARG;
  TendIn in{};
  in.q[0] = DDENS_;
  in.q[1] = DRHOT_;
  in.q[2] = MOMZ_;
  in.q[3] = MOMX_;
  in.q[4] = MOMY_;
  in.DPRES = DPRES_;
  in.DENS_hyd = DENS_hyd;
  in.PRES_hyd = PRES_hyd;
  in.THERM_hyd = THERM_hyd;
  in.CORIOLIS = CORIOLIS;
  in.DPhydDx = DPhydDx;
  in.DPhydDy = DPhydDy;
  in.Rtot = Rtot;
  in.CVtot = CVtot;
  in.CPtot = CPtot;
  double* out[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  TRY(launch_heve_dropin(ctx, in, out));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_heve_cal_tend_cco(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                           double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                           const double* MOMZ_, const double* DRHOT_, const double* DPRES_, const double* DENS_hyd,
                           const double* PRES_hyd, const double* PRES_hyd_ref, const double* THERM_hyd,
                           const double* CORIOLIS, const double* Rtot, const double* CVtot, const double* CPtot,
                           const double* DPhydDx, const double* DPhydDy) {
  (void)PRES_hyd_ref;
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ ||
      !DRHOT_ || !DPRES_ || !DENS_hyd || !PRES_hyd || !THERM_hyd || !CORIOLIS || !DPhydDx || !DPhydDy)
    return FEDG_ERR_ARG;
  TendIn in{};
  in.q[0] = DDENS_;
  in.q[1] = DRHOT_;
  in.q[2] = MOMZ_;
  in.q[3] = MOMX_;
  in.q[4] = MOMY_;
  in.DPRES = DPRES_;
  in.DENS_hyd = DENS_hyd;
  in.PRES_hyd = PRES_hyd;
  in.THERM_hyd = THERM_hyd;
  in.CORIOLIS = CORIOLIS;
  in.DPhydDx = DPhydDx;
  in.DPhydDy = DPhydDy;
  in.Rtot = Rtot;
  in.CVtot = CVtot;
  in.CPtot = CPtot;
  double* out[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  TRY(launch_heve_cco(ctx, in, out));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_heve_add_bnd_contrib(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                              double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                              const double* MOMZ_, const double* DRHOT_, double* DPRES, const double* DENS_hyd,
                              const double* PRES_hyd, const double* THERM_hyd, const double* Rtot,
                              const double* CVtot, const double* CPtot) {
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ ||
      !DRHOT_ || !DPRES || !DENS_hyd || !PRES_hyd || !THERM_hyd)
    return FEDG_ERR_ARG;
  if ((Rtot == nullptr) != (CVtot == nullptr) || (Rtot == nullptr) != (CPtot == nullptr)) {
    set_error("fedg_heve_add_bnd_contrib: Rtot, CVtot, CPtot are all given or all NULL");
    return FEDG_ERR_ARG;
  }
  TendIn in{};
  in.q[0] = DDENS_;
  in.q[1] = DRHOT_;
  in.q[2] = MOMZ_;
  in.q[3] = MOMX_;
  in.q[4] = MOMY_;
  in.DENS_hyd = DENS_hyd;
  in.PRES_hyd = PRES_hyd;
  in.THERM_hyd = THERM_hyd;
  in.Rtot = Rtot;
  in.CVtot = CVtot;
  in.CPtot = CPtot;
  double* dt[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  TRY(launch_heve_bnd(ctx, in, dt, DPRES));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

// FaceIntMat and IntWeight_lgl (scale_atm_dyn_dgm_trcadvect3d_heve.F90:75-137,
// ELM/scale_element_hexahedral.F90:286-298) are products of the 1D LGL
// weights of the element's order
static int trc_weights(fedg_ctx* ctx, TrcW& w) {
  if (!ctx->have_w1d) {
    RefElem e;
    try {
      build_refelem(ctx->m.n1 - 1, e);
    } catch (const std::exception& ex) {
      set_error(ex.what());
      return FEDG_ERR_ARG;
    }
    for (int l = 0; l < 8; ++l) ctx->w1d[l] = l < ctx->m.n1 ? e.w1d[l] : 0.0;
    ctx->have_w1d = 1;
  }
  for (int l = 0; l < 8; ++l) w.w[l] = ctx->w1d[l];
  return FEDG_OK;
}

int fedg_trcadvect_cal_tend(fedg_ctx* ctx, double* QTRC_dt, const double* QTRC_, const double* MOMX_,
                            const double* MOMY_, const double* MOMZ_, const double* alphDENS_M,
                            const double* alphDENS_P, const double* fct_coef, const double* RHOQ_tp,
                            const double* J) {
  if (!ctx || !QTRC_dt || !QTRC_ || !MOMX_ || !MOMY_ || !MOMZ_ || !alphDENS_M || !alphDENS_P || !fct_coef ||
      !RHOQ_tp || !J)
    return FEDG_ERR_ARG;
  TrcLaunch t{};
  TRY(trc_weights(ctx, t.w));
  t.dt = QTRC_dt;
  t.Q = QTRC_;
  t.MX = MOMX_;
  t.MY = MOMY_;
  t.MZ = MOMZ_;
  t.aM = alphDENS_M;
  t.aP = alphDENS_P;
  t.R = fct_coef;
  t.tp = RHOQ_tp;
  t.J = J;
  TRY(launch_trcadv(ctx, t));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_trcadvect_calc_fct_coef(fedg_ctx* ctx, double* fct_coef, const double* QTRC_, const double* MOMX_,
                                 const double* MOMY_, const double* MOMZ_, const double* RHOQ_tp_,
                                 const double* alphDENS_M, const double* alphDENS_P, const double* DENS_hyd,
                                 const double* DDENS_, const double* DDENS0_, double rk_c_ssm1, double dt,
                                 const double* J, int disable_limiter) {
  if (!ctx || !fct_coef) return FEDG_ERR_ARG;
  TrcLaunch t{};
  t.dt = fct_coef;
  t.fct = 1;
  t.disable = disable_limiter ? 1 : 0;
  if (!t.disable) {
    if (!QTRC_ || !MOMX_ || !MOMY_ || !MOMZ_ || !RHOQ_tp_ || !alphDENS_M || !alphDENS_P || !DENS_hyd || !DDENS_ ||
        !DDENS0_ || !J)
      return FEDG_ERR_ARG;
    if (!(dt != 0.0)) {
      set_error("fedg_trcadvect_calc_fct_coef: dt must be nonzero");
      return FEDG_ERR_ARG;
    }
    TRY(trc_weights(ctx, t.w));
    t.Q = QTRC_;
    t.MX = MOMX_;
    t.MY = MOMY_;
    t.MZ = MOMZ_;
    t.tp = RHOQ_tp_;
    t.aM = alphDENS_M;
    t.aP = alphDENS_P;
    t.Dh = DENS_hyd;
    t.DD = DDENS_;
    t.DD0 = DDENS0_;
    t.c = rk_c_ssm1;
    t.dtt = dt;
    t.J = J;
  }
  TRY(launch_trcadv(ctx, t));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_trcadvect_tmar(fedg_ctx* ctx, double* QTRC_, const double* DENS_hyd, const double* DDENS_,
                        const double* J) {
  if (!ctx || !QTRC_ || !DENS_hyd || !DDENS_ || !J) return FEDG_ERR_ARG;
  TrcW w{};
  TRY(trc_weights(ctx, w));
  TRY(launch_tmar(ctx, QTRC_, DENS_hyd, DDENS_, J, w));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

static int trc_thermo(int nq, const double* Mq, const double* Rq, const double* CVq, const double* CPq,
                      TrcThermo& th) {
  if (nq < 1 || nq > kMaxTrc || !Mq || !Rq || !CVq || !CPq) {
    set_error("water species: 1 to " + std::to_string(kMaxTrc) + " tracers with their TRACER_MASS, _R, _CV, _CP");
    return FEDG_ERR_ARG;
  }
  th = TrcThermo{};
  th.nq = nq;
  for (int iq = 0; iq < nq; ++iq) {
    th.Mq[iq] = Mq[iq];
    th.Rq[iq] = Rq[iq];
    th.CVq[iq] = CVq[iq];
    th.CPq[iq] = CPq[iq];
  }
  return FEDG_OK;
}

int fedg_tracer_specific_heat(fedg_ctx* ctx, int nq, double* const* QTRC, const double* TRACER_MASS,
                              const double* TRACER_R, const double* TRACER_CV, const double* TRACER_CP, double* Rtot,
                              double* CVtot, double* CPtot) {
  if (!ctx || !QTRC || !Rtot || !CVtot || !CPtot) return FEDG_ERR_ARG;
  TrcThermo th;
  TRY(trc_thermo(nq, TRACER_MASS, TRACER_R, TRACER_CV, TRACER_CP, th));
  const double* Q[kMaxTrc] = {};
  for (int iq = 0; iq < nq; ++iq) {
    if (!QTRC[iq]) return FEDG_ERR_ARG;
    Q[iq] = QTRC[iq];
  }
  TRY(launch_specific_heat(ctx, Q, th, Rtot, CVtot, CPtot));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_atmos_calc_diagvar(fedg_ctx* ctx, const char* field_name, double* var_out, const double* DDENS_,
                            const double* MOMX_, const double* MOMY_, const double* MOMZ_, const double* PRES_, int nq,
                            double* const* QTRC, const double* TRACER_MASS, const double* TRACER_CV,
                            const double* TRACER_ENGI0, const double* DENS_hyd, const double* PRES_hyd,
                            const double* Rtot, const double* CVtot, const double* CPtot, const double* zlev,
                            const double* pos_z) {
  (void)CVtot;  // an argument of the reference's routine that none of its formulas reads
  if (!ctx || !field_name || !var_out) return FEDG_ERR_ARG;
  static const struct {
    const char* name;
    int kind;
  } names[] = {{"DENS", kDiagDENS},   {"U", kDiagU},         {"V", kDiagV},
               {"W", kDiagW},         {"Umet", kDiagUmet},   {"Vmet", kDiagVmet},
               {"PRES_diff", kDiagPRES_diff}, {"T", kDiagT}, {"T_diff", kDiagT_diff},
               {"PT", kDiagPT},       {"PT_diff", kDiagPT_diff}, {"QDRY", kDiagQDRY},
               {"ENGK", kDiagENGK},   {"ENGP", kDiagENGP},   {"ENGI", kDiagENGI},
               {"ENGT", kDiagENGT}};
  const std::string name(field_name);
  // the reference's case('PRES') is empty: var_out is left as it is
  if (name == "PRES") return FEDG_OK;
  int kind = -1;
  for (const auto& e : names)
    if (name == e.name) kind = e.kind;
  if (kind < 0) {
    if (name == "RH" || name == "RHL")
      set_error("diagnostic variable " + name + " needs the saturation vapour pressure over liquid water of the "
                "SCALE library (ATMOS_SATURATION_psat_liq), which this library does not restate");
    else
      set_error("the name of the diagnostic variable is not supported: " + name);
    return FEDG_ERR_ARG;
  }
  DiagVar a{};
  a.kind = kind;
  a.out = var_out;
  a.DD = DDENS_;
  a.MX = MOMX_;
  a.MY = MOMY_;
  a.MZ = MOMZ_;
  a.P = PRES_;
  a.Dh = DENS_hyd;
  a.Ph = PRES_hyd;
  a.R = Rtot;
  a.CP = CPtot;
  a.zlev = zlev;
  a.zpos = pos_z;
  const bool engi = kind == kDiagENGI || kind == kDiagENGT;
  const bool engk = kind == kDiagENGK || kind == kDiagENGT;
  // the fields each formula reads
  bool ok = true;
  if (kind != kDiagPRES_diff && kind != kDiagQDRY) ok = ok && DDENS_ && DENS_hyd;
  if (kind == kDiagU || kind == kDiagUmet || engk) ok = ok && MOMX_;
  if (kind == kDiagV || kind == kDiagVmet || engk) ok = ok && MOMY_;
  if (kind == kDiagW || engk) ok = ok && MOMZ_;
  if (kind == kDiagPRES_diff || kind == kDiagT || kind == kDiagT_diff || kind == kDiagPT || kind == kDiagPT_diff ||
      engi)
    ok = ok && PRES_;
  if (kind == kDiagPRES_diff || kind == kDiagT_diff) ok = ok && PRES_hyd;
  if (kind == kDiagT || kind == kDiagT_diff || kind == kDiagPT || kind == kDiagPT_diff || engi) ok = ok && Rtot;
  if (kind == kDiagPT || kind == kDiagPT_diff) ok = ok && CPtot;
  if (kind == kDiagPT_diff) ok = ok && PRES_hyd;
  if (kind == kDiagENGP) ok = ok && zlev;
  if (kind == kDiagENGT) ok = ok && pos_z;
  if (!ok) {
    set_error("diagnostic variable " + name + ": a field its formula reads is NULL");
    return FEDG_ERR_ARG;
  }
  if (kind == kDiagQDRY || engi) {
    if (nq < 0 || nq > kMaxTrc || (nq > 0 && (!QTRC || !TRACER_MASS || (engi && (!TRACER_CV || !TRACER_ENGI0))))) {
      set_error("diagnostic variable " + name + ": 0 to " + std::to_string(kMaxTrc) +
                " tracers with TRACER_MASS (and TRACER_CV, TRACER_ENGI0 for the energies)");
      return FEDG_ERR_ARG;
    }
    a.nq = nq;
    for (int iq = 0; iq < nq; ++iq) {
      if (!QTRC[iq]) return FEDG_ERR_ARG;
      a.Q[iq] = QTRC[iq];
      a.Mq[iq] = TRACER_MASS[iq];
      if (engi) {
        a.CVq[iq] = TRACER_CV[iq];
        a.E0q[iq] = TRACER_ENGI0[iq];
      }
    }
  }
  TRY(launch_diagvar(ctx, a));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_trc_negative_fixer(fedg_ctx* ctx, int nq, double* const* QTRC, double* DDENS, double* PRES, double* CVtot,
                            double* CPtot, double* Rtot, const double* DENS_hyd, const double* PRES_hyd,
                            const double* J, const double* zlev, double dt, int QLA, int QIA,
                            const double* TRACER_MASS, const double* TRACER_R, const double* TRACER_CV,
                            const double* TRACER_CP, double* DRHOT) {
  (void)dt;  // the reference takes dt and does not use it
  if (!ctx || !QTRC || !DDENS || !PRES || !CVtot || !CPtot || !Rtot || !DENS_hyd || !PRES_hyd || !J || !zlev)
    return FEDG_ERR_ARG;
  if (QLA < 0 || QIA < 0 || nq != 1 + QLA + QIA) {
    set_error("negative fixer: the tracers are the water species, nq = 1 + QLA + QIA (the reference leaves "
              "any further tracer undefined)");
    return FEDG_ERR_ARG;
  }
  NegFix a{};
  TRY(trc_thermo(nq, TRACER_MASS, TRACER_R, TRACER_CV, TRACER_CP, a.th));
  for (int iq = 0; iq < nq; ++iq) {
    if (!QTRC[iq]) return FEDG_ERR_ARG;
    a.Q[iq] = QTRC[iq];
  }
  a.DD = DDENS;
  a.PRES = PRES;
  a.CV = CVtot;
  a.CP = CPtot;
  a.R = Rtot;
  a.DRHOT = DRHOT;
  a.Dh = DENS_hyd;
  a.Ph = PRES_hyd;
  a.J = J;
  a.zlev = zlev;
  TRY(trc_weights(ctx, a.w));
  TRY(launch_negative_fixer(ctx, a));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_trcadvect_save_massflux(fedg_ctx* ctx, double* MFLX_x_tavg, double* MFLX_y_tavg, double* MFLX_z_tavg,
                                 double* alph_dens_M, double* alph_dens_P, const double* DDENS, const double* MOMX,
                                 const double* MOMY, const double* MOMZ, const double* DPRES, const double* DENS_hyd,
                                 const double* PRES_hyd, const double* Rtot, const double* CVtot,
                                 const double* CPtot, int rkstage, double tavg_weight_h, double tavg_weight_v,
                                 int is_hevi) {
  (void)Rtot;  // the reference's cal_alphdens_dyn takes them but uses the dry-air gamma
  (void)CVtot;
  (void)CPtot;
  if (!ctx || !MFLX_x_tavg || !MFLX_y_tavg || !MFLX_z_tavg || !alph_dens_M || !alph_dens_P || !DDENS || !MOMX ||
      !MOMY || !MOMZ || !DPRES || !DENS_hyd || !PRES_hyd)
    return FEDG_ERR_ARG;
  if (rkstage < 1) {
    set_error("fedg_trcadvect_save_massflux: rkstage counts from 1");
    return FEDG_ERR_ARG;
  }
  MflxArgs a{};
  a.FX = MFLX_x_tavg;
  a.FY = MFLX_y_tavg;
  a.FZ = MFLX_z_tavg;
  a.alM = alph_dens_M;
  a.alP = alph_dens_P;
  a.DD = DDENS;
  a.MX = MOMX;
  a.MY = MOMY;
  a.MZ = MOMZ;
  a.DP = DPRES;
  a.Dh = DENS_hyd;
  a.Ph = PRES_hyd;
  a.rkstage = rkstage;
  a.hevi = is_hevi ? 1 : 0;
  a.twh = tavg_weight_h;
  a.twv = tavg_weight_v;
  TRY(launch_save_massflux(ctx, a));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_trcadvect_cal_alphdens_advtest(fedg_ctx* ctx, double* alph_dens_M, double* alph_dens_P,
                                        const double* DDENS_, const double* MOMX_, const double* MOMY_,
                                        const double* MOMZ_, const double* DENS_hyd) {
  if (!ctx || !alph_dens_M || !alph_dens_P || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ || !DENS_hyd) return FEDG_ERR_ARG;
  MflxArgs a{};
  a.alM = alph_dens_M;
  a.alP = alph_dens_P;
  a.DD = DDENS_;
  a.MX = MOMX_;
  a.MY = MOMY_;
  a.MZ = MOMZ_;
  a.Dh = DENS_hyd;
  TRY(launch_alphdens_advtest(ctx, a));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_trcvar_rk_advance(fedg_ctx* ctx, int nowstage, int nstage, const double* coef_sig_ex,
                           const double* coef_gam_ex, const double* coef_c_ex, double dt, double* q, double* var0,
                           double* varTmp, const double* tend, const double* DDENS, const double* DDENS0,
                           const double* DENS_hyd) {
  if (!ctx || !coef_sig_ex || !coef_gam_ex || !coef_c_ex || !q || !var0 || !varTmp || !tend || !DDENS || !DDENS0 ||
      !DENS_hyd)
    return FEDG_ERR_ARG;
  if (nstage < 1 || nowstage < 1 || nowstage > nstage) {
    set_error("fedg_trcvar_rk_advance: need 1 <= nowstage <= nstage");
    return FEDG_ERR_ARG;
  }
  // coef_sig_ex, coef_gam_ex: (nstage+1, nstage) column-major, as timeint_rk holds them
  const int ld = nstage + 1;
  auto SIG = [&](int i, int j) { return coef_sig_ex[(i - 1) + ld * (j - 1)]; };
  auto GAM = [&](int i, int j) { return coef_gam_ex[(i - 1) + ld * (j - 1)]; };
  const double EPS = 1.0e-16;  // CONST_EPS of SCALE's scale_const
  TrcRk r{};
  r.q = q;
  r.var0 = var0;
  r.varTmp = varTmp;
  r.tend = tend;
  r.DD = DDENS;
  r.DD0 = DDENS0;
  r.Dh = DENS_hyd;
  r.sig_ss = SIG(nowstage + 1, nowstage);
  r.sig_Ns = SIG(nstage + 1, nowstage);
  r.one_minus_sig_ss = 1.0 - r.sig_ss;
  r.gam_ss = dt * GAM(nowstage + 1, nowstage);
  r.gam_Ns = dt * GAM(nstage + 1, nowstage);
  r.c_ssm1 = coef_c_ex[nowstage - 1];
  r.last = nowstage == nstage;
  r.c_ss = r.last ? 0.0 : coef_c_ex[nowstage];
  r.first = nowstage == 1;
  r.acc = !r.last && (fabs(r.sig_Ns) > EPS || fabs(r.gam_Ns) > EPS);
  TRY(launch_trc_rk(ctx, r));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_tracer_modalfilter_apply(fedg_ctx* ctx, double* QTRC_, const double* DENS_hyd, const double* DDENS_,
                                  const double* MFilter_tracer_h1D, const double* MFilter_tracer_v1D) {
  if (!ctx || !QTRC_ || !DENS_hyd || !DDENS_ || !MFilter_tracer_h1D || !MFilter_tracer_v1D) return FEDG_ERR_ARG;
  TRY(launch_tracer_modalfilter(ctx, QTRC_, DENS_hyd, DDENS_, MFilter_tracer_h1D, MFilter_tracer_v1D));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_hevi_cal_tend(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                       double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                       const double* MOMZ_, const double* DRHOT_, const double* DPRES_, const double* DENS_hyd,
                       const double* PRES_hyd, const double* PRES_hyd_ref, const double* THERM_hyd,
                       const double* CORIOLIS, const double* Rtot, const double* CVtot, const double* CPtot,
                       const double* DPhydDx, const double* DPhydDy) {
  (void)PRES_hyd_ref;
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ ||
      !DRHOT_ || !DPRES_ || !DENS_hyd || !PRES_hyd || !THERM_hyd || !CORIOLIS || !DPhydDx || !DPhydDy)
    return FEDG_ERR_ARG;
  TendIn in{};
  in.q[0] = DDENS_;
  in.q[1] = DRHOT_;
  in.q[2] = MOMZ_;
  in.q[3] = MOMX_;
  in.q[4] = MOMY_;
  in.DPRES = DPRES_;
  in.DENS_hyd = DENS_hyd;
  in.PRES_hyd = PRES_hyd;
  in.THERM_hyd = THERM_hyd;
  in.CORIOLIS = CORIOLIS;
  in.DPhydDx = DPhydDx;
  in.DPhydDy = DPhydDy;
  in.Rtot = Rtot;
  in.CVtot = CVtot;
  in.CPtot = CPtot;
  double* out[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  TRY(launch_hevi_explicit(ctx, in, out));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

}  // extern "C"

namespace fedg {
int hevi_work_get(fedg_ctx* ctx, HeviWork** w) {
  if (ctx->hevi) {
    *w = ctx->hevi;
    return FEDG_OK;
  }
  const DevMesh& m = ctx->m;
  if (m.Ne2D <= 0 || m.Ne % m.Ne2D != 0) {
    set_error("HEVI: the mesh is not a stack of Ne2D-element horizontal layers");
    return FEDG_ERR_UNSUPPORTED;
  }
  HeviWork* h = new HeviWork();
  const size_t n = (size_t)m.Np * m.Ne;
  int rc = FEDG_OK;
  auto al = [&](double** p, size_t cnt) {
    if (rc == FEDG_OK && hipMalloc((void**)p, sizeof(double) * cnt) != hipSuccess) {
      set_error("HEVI: device allocation of the implicit-solve workspace failed");
      rc = FEDG_ERR_NOMEM;
    }
  };
  for (int v = 0; v < 5; ++v) al(&h->pv[v], n);
  al(&h->alph, (size_t)2 * m.Nfp * m.Ne);
  al(&h->b3, 3 * n);
  al(&h->buv, 2 * n);
  al(&h->coef, 3 * n);
  al(&h->partial, 256);
  ctx->hevi = h;
  if (rc != FEDG_OK) {
    hevi_work_free(ctx);
    return rc;
  }
  *w = h;
  return FEDG_OK;
}

void hevi_work_free(fedg_ctx* ctx) {
  HeviWork* h = ctx->hevi;
  if (!h) return;
  for (int v = 0; v < 5; ++v)
    if (h->pv[v]) (void)hipFree(h->pv[v]);
  double* bufs[] = {h->alph, h->b3, h->buv, h->coef, h->partial, h->r_plain, h->r_guess, h->flag};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (h->X) (void)hipFree(h->X);
  if (h->Y) (void)hipFree(h->Y);
  delete h;
  ctx->hevi = nullptr;
}
}  // namespace fedg

extern "C" {

static int hevi_cal_vi_common(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                              double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                              const double* MOMZ_, const double* DRHOT_, const double* DENS_hyd,
                              const double* PRES_hyd, const double* DDENS0_, const double* MOMX0_,
                              const double* MOMY0_, const double* MOMZ0_, const double* DRHOT0_, const double* Rtot,
                              const double* CVtot, const double* CPtot, double impl_fac, int max_it, double rtol,
                              double* resid, int* iterations, HeviGuess* guess = nullptr) {
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DDENS_ || !MOMX_ || !MOMY_ || !MOMZ_ ||
      !DRHOT_ || !DENS_hyd || !PRES_hyd || !DDENS0_ || !MOMX0_ || !MOMY0_ || !MOMZ0_ || !DRHOT0_ || !Rtot ||
      !CVtot || !CPtot || max_it < 1 || !std::isfinite(impl_fac))
    return FEDG_ERR_ARG;
  HeviWork* w = nullptr;
  TRY(hevi_work_get(ctx, &w));
  double* tend[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  const double* q[5] = {DDENS_, DRHOT_, MOMZ_, MOMX_, MOMY_};
  const double* q0[5] = {DDENS0_, DRHOT0_, MOMZ0_, MOMX0_, MOMY0_};
  int it = 0;
  TRY(hevi_cal_vi(ctx, w, tend, q, q0, DENS_hyd, PRES_hyd, Rtot, CVtot, CPtot, impl_fac, max_it, rtol, resid, &it,
                  guess));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (iterations) *iterations = it;
  return FEDG_OK;
}

int fedg_hevi_cal_vi(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                     double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                     const double* MOMZ_, const double* DRHOT_, const double* DENS_hyd, const double* PRES_hyd,
                     const double* DDENS0_, const double* MOMX0_, const double* MOMY0_, const double* MOMZ0_,
                     const double* DRHOT0_, const double* Rtot, const double* CVtot, const double* CPtot,
                     double impl_fac, double dt) {
  (void)dt;
  return hevi_cal_vi_common(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_,
                            DENS_hyd, PRES_hyd, DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot,
                            impl_fac, 1, 0.0, nullptr, nullptr);
}

int fedg_hevi_cal_vi_newton(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                            double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                            const double* MOMZ_, const double* DRHOT_, const double* DENS_hyd,
                            const double* PRES_hyd, const double* DDENS0_, const double* MOMX0_,
                            const double* MOMY0_, const double* MOMZ0_, const double* DRHOT0_, const double* Rtot,
                            const double* CVtot, const double* CPtot, double impl_fac, double dt, int max_it,
                            double rtol, double* resid, int* iterations) {
  (void)dt;
  return hevi_cal_vi_common(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_,
                            DENS_hyd, PRES_hyd, DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot,
                            impl_fac, max_it, rtol, resid, iterations);
}

struct fedg_mlp {
  fedg::MlpImpl* impl;
};

int fedg_hevi_cal_vi_guess(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                           double* RHOT_dt, const double* DDENS_, const double* MOMX_, const double* MOMY_,
                           const double* MOMZ_, const double* DRHOT_, const double* DENS_hyd, const double* PRES_hyd,
                           const double* DDENS0_, const double* MOMX0_, const double* MOMY0_, const double* MOMZ0_,
                           const double* DRHOT0_, const double* Rtot, const double* CVtot, const double* CPtot,
                           double impl_fac, double dt, int max_it, double rtol, fedg_mlp* mlp, const double* scale,
                           const double* const* Tex, double* resid, int* iterations, long* fallbacks) {
  if (!mlp || !scale || !(dt > 0.0) || !(fabs(impl_fac) > 0.0)) return FEDG_ERR_ARG;
  HeviGuess g{};
  g.mlp = mlp->impl;
  for (int v = 0; v < 15; ++v) {
    if (!(scale[v] > 0.0)) return FEDG_ERR_ARG;
    g.scale[v] = scale[v];
  }
  for (int v = 0; v < 5; ++v) g.Tex[v] = Tex ? Tex[v] : nullptr;
  g.dt = dt;
  TRY(hevi_cal_vi_common(ctx, DENS_dt, MOMX_dt, MOMY_dt, MOMZ_dt, RHOT_dt, DDENS_, MOMX_, MOMY_, MOMZ_, DRHOT_,
                         DENS_hyd, PRES_hyd, DDENS0_, MOMX0_, MOMY0_, MOMZ0_, DRHOT0_, Rtot, CVtot, CPtot, impl_fac,
                         max_it, rtol, resid, iterations, &g));
  if (fallbacks) *fallbacks = g.fallbacks;
  return FEDG_OK;
}

int fedg_calc_phyd_hgrad(fedg_ctx* ctx, double* DPhydDx, double* DPhydDy, const double* PRES_hyd,
                         const double* PRES_hyd_ref) {
  if (!ctx || !DPhydDx || !DPhydDy || !PRES_hyd || !PRES_hyd_ref) return FEDG_ERR_ARG;
  TRY(launch_phyd_hgrad(ctx, DPhydDx, DPhydDy, PRES_hyd, PRES_hyd_ref));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

// ---------------------------------------------------------------------------
int fedg_tint_create(fedg_ctx* ctx, const char* scheme, double dt, fedg_tint** out) {
  if (!ctx || !scheme || !out) return FEDG_ERR_ARG;
  *out = nullptr;
  fedg_tint* t = new fedg_tint();
  memset(t, 0, sizeof(*t));
  t->ctx = ctx;
  int rc = butcher(scheme, dt, t);
  if (rc != FEDG_OK) {
    delete t;
    return rc;
  }
  const size_t n = (size_t)ctx->m.Np * ctx->m.NeA * 5;
  rc = dev_upload<double>(&t->var0, nullptr, n);
  if (rc == FEDG_OK) rc = dev_upload<double>(&t->varTmp, nullptr, n);
  if (rc == FEDG_OK) rc = dev_upload<double>(&t->tend_ex, nullptr, n * t->tend_buf_size);
  if (rc == FEDG_OK && t->imex) rc = dev_upload<double>(&t->tend_im, nullptr, n * t->tend_buf_size);
  if (rc != FEDG_OK) {
    fedg_tint_destroy(t);
    return rc;
  }
  *out = t;
  return FEDG_OK;
}

int fedg_tint_destroy(fedg_tint* t) {
  if (!t) return FEDG_OK;
  double* ptrs[] = {t->var0, t->varTmp, t->tend_ex, t->tend_im};
  for (double* p : ptrs)
    if (p) (void)hipFree(p);
  delete t;
  return FEDG_OK;
}

int fedg_tint_info(const fedg_tint* t, int* nstage, int* tend_buf_size, int* imex) {
  if (!t) return FEDG_ERR_ARG;
  if (nstage) *nstage = t->nstage;
  if (tend_buf_size) *tend_buf_size = t->tend_buf_size;
  if (imex) *imex = t->imex;
  return FEDG_OK;
}

static double* tend_ptr(fedg_tint* t, double* base, int buf, int varID) {
  const size_t NpNeA = (size_t)t->ctx->m.Np * t->ctx->m.NeA;
  return base + ((size_t)buf * 5 + (size_t)(varID - 1)) * NpNeA;
}

double* fedg_tint_tend_ex(fedg_tint* t, int stage, int varID) {
  if (!t || stage < 1 || stage > t->nstage || varID < 1 || varID > 5) return nullptr;
  return tend_ptr(t, t->tend_ex, t->indmap[stage - 1], varID);
}
double* fedg_tint_tend_im(fedg_tint* t, int stage, int varID) {
  if (!t || !t->tend_im || stage < 1 || stage > t->nstage || varID < 1 || varID > 5) return nullptr;
  return tend_ptr(t, t->tend_im, t->indmap[stage - 1], varID);
}

int fedg_tint_store_var0(fedg_tint* t, const double* q, int varID) {
  if (!t || !q || varID < 1 || varID > 5) return FEDG_ERR_ARG;
  fedg_ctx* ctx = t->ctx;
  const size_t NpNeA = (size_t)ctx->m.Np * ctx->m.NeA;
  TRY(launch_copy(ctx, t->var0 + (size_t)(varID - 1) * NpNeA, q, (long)ctx->m.Np * ctx->m.Ne));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

// rk_advance_general2D (CMN/scale_timeint_rk.F90:2201-2355) as streaming passes
static int tint_advance_one(fedg_tint* t, int s, double* q, int varID) {
  fedg_ctx* ctx = t->ctx;
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  const size_t NpNeA = (size_t)ctx->m.Np * ctx->m.NeA;
  double* v0 = t->var0 + (size_t)(varID - 1) * NpNeA;
  double* vt = t->varTmp + (size_t)(varID - 1) * NpNeA;
  const int ind = t->indmap[s - 1];
  double* Te = tend_ptr(t, t->tend_ex, ind, varID);
  double* Ti = t->imex ? tend_ptr(t, t->tend_im, ind, varID) : nullptr;
  if (t->nstage == 1) TRY(launch_copy(ctx, vt, q, n));
  if (s == t->nstage) {
    const double cbe = t->b_ex[s - 1] * t->dt;
    if (t->imex) return launch_axpy2_set(ctx, q, vt, cbe, Te, t->b_im[s - 1] * t->dt, Ti, n);
    return launch_axpy_set(ctx, q, vt, cbe, Te, n);
  }
  if (s == 1 && !t->imex) {
    TRY(launch_copy(ctx, v0, q, n));
    TRY(launch_copy(ctx, vt, q, n));
  }
  const double cbe = t->b_ex[s - 1] * t->dt;
  TRY(launch_copy(ctx, q, v0, n));
  if (t->imex) TRY(launch_axpy2_acc(ctx, vt, cbe, Te, t->b_im[s - 1] * t->dt, Ti, n));
  else TRY(launch_axpy_acc(ctx, vt, cbe, Te, n));
  if (t->tend_buf_size == 1 && !t->imex) {
    return launch_axpy_set(ctx, q, v0, t->dt * t->a_ex[s][s - 1], tend_ptr(t, t->tend_ex, 0, varID), n);
  } else if (!t->imex) {
    for (int s2 = 1; s2 <= s; ++s2)
      TRY(launch_axpy_acc(ctx, q, t->dt * t->a_ex[s][s2 - 1], tend_ptr(t, t->tend_ex, s2 - 1, varID), n));
  } else {
    for (int s2 = 1; s2 <= s; ++s2)
      TRY(launch_axpy2_acc(ctx, q, t->dt * t->a_ex[s][s2 - 1], tend_ptr(t, t->tend_ex, s2 - 1, varID),
                           t->dt * t->a_im[s][s2 - 1], tend_ptr(t, t->tend_im, s2 - 1, varID), n));
  }
  return FEDG_OK;
}

// tint_advance_one of an IMEX scheme for the five prognostics (VID order) in
// one pass (the same operations per node; launch_imex_advance)
static int imex_advance_all(fedg_tint* t, int s, double* const q[5]) {
  if (!t->imex || t->nstage < 2 || s > kImexMaxStages) {
    for (int v = 0; v < 5; ++v) TRY(tint_advance_one(t, s, q[v], v + 1));
    return FEDG_OK;
  }
  const size_t NpNeA = (size_t)t->ctx->m.Np * t->ctx->m.NeA;
  const int ind = t->indmap[s - 1];
  ImexAdvance a{};
  for (int v = 0; v < 5; ++v) {
    a.q[v] = q[v];
    a.vt[v] = t->varTmp + (size_t)v * NpNeA;
    a.v0[v] = t->var0 + (size_t)v * NpNeA;
    a.Te[v] = tend_ptr(t, t->tend_ex, ind, v + 1);
    a.Ti[v] = tend_ptr(t, t->tend_im, ind, v + 1);
  }
  a.cbe = t->b_ex[s - 1] * t->dt;
  a.cbi = t->b_im[s - 1] * t->dt;
  a.s = s;
  a.last = s == t->nstage;
  if (!a.last) {
    for (int s2 = 1; s2 <= s; ++s2) {
      a.ae[s2 - 1] = t->dt * t->a_ex[s][s2 - 1];
      a.ai[s2 - 1] = t->dt * t->a_im[s][s2 - 1];
      for (int v = 0; v < 5; ++v) {
        a.Tes[s2 - 1][v] = tend_ptr(t, t->tend_ex, s2 - 1, v + 1);
        a.Tis[s2 - 1][v] = tend_ptr(t, t->tend_im, s2 - 1, v + 1);
      }
    }
  }
  return launch_imex_advance(t->ctx, a);
}

int fedg_tint_advance(fedg_tint* t, int stage, int nvar, double* const* q, const int* varIDs) {
  if (!t || stage < 1 || stage > t->nstage || nvar < 0 || (nvar > 0 && (!q || !varIDs))) return FEDG_ERR_ARG;
  for (int v = 0; v < nvar; ++v) {
    if (varIDs[v] < 1 || varIDs[v] > 5 || !q[v]) return FEDG_ERR_ARG;
    TRY(tint_advance_one(t, stage, q[v], varIDs[v]));
  }
  FEDG_HIP_CHECK(hipStreamSynchronize(t->ctx->stream));
  return FEDG_OK;
}

int fedg_tint_store_implicit(fedg_tint* t, int stage, double* q, int varID) {
  if (!t || !q || stage < 1 || stage > t->nstage || varID < 1 || varID > 5) return FEDG_ERR_ARG;
  if (!t->imex) return FEDG_OK;
  fedg_ctx* ctx = t->ctx;
  const long n = (long)ctx->m.Np * ctx->m.Ne;
  const size_t NpNeA = (size_t)ctx->m.Np * ctx->m.NeA;
  if (stage == 1) {
    TRY(launch_copy(ctx, t->var0 + (size_t)(varID - 1) * NpNeA, q, n));
    TRY(launch_copy(ctx, t->varTmp + (size_t)(varID - 1) * NpNeA, q, n));
  }
  const int ind = t->indmap[stage - 1];
  TRY(launch_axpy_acc(ctx, q, t->dt * t->a_im[stage - 1][stage - 1], tend_ptr(t, t->tend_im, ind, varID), n));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

// ---------------------------------------------------------------------------
int fedg_heve_create(fedg_ctx* ctx, const char* scheme, double dt, const double* DENS_hyd, const double* PRES_hyd,
                     const double* THERM_hyd, const double* CORIOLIS, const double* DPhydDx, const double* DPhydDy,
                     const double* Rtot, const double* CVtot, const double* CPtot, fedg_heve** out) {
  if (!ctx || !scheme || !out || !DENS_hyd || !PRES_hyd || !THERM_hyd || !CORIOLIS || !DPhydDx || !DPhydDy)
    return FEDG_ERR_ARG;
  if ((Rtot || CVtot || CPtot) && !(Rtot && CVtot && CPtot)) {
    set_error("Rtot, CVtot and CPtot must be given together");
    return FEDG_ERR_ARG;
  }
  *out = nullptr;
  fedg_heve* st = new fedg_heve();
  memset(st, 0, sizeof(*st));
  st->ctx = ctx;
  int rc = butcher(scheme, dt, &st->tab);
  if (rc == FEDG_OK && (st->tab.imex || st->tab.tend_buf_size != 1)) {
    set_error("the fused HEVE stepper takes explicit schemes with one tendency buffer (ERK_4s4o, ERK_1s1o)");
    rc = FEDG_ERR_UNSUPPORTED;
  }
  const DevMesh& m = ctx->m;
  const size_t NpNe = (size_t)m.Np * m.Ne, cor = (size_t)m.Nfp * m.Ne2DA;
  // Rtot/CVtot/CPtot equal to the dry constants everywhere are dropped (the
  // kernel then uses the constants, which is the same arithmetic)
  bool dry = true;
  if (Rtot)
    dry = all_equal(Rtot, NpNe, ctx->c.Rdry) && all_equal(CVtot, NpNe, ctx->c.CVdry) &&
          all_equal(CPtot, NpNe, ctx->c.CPdry);
  st->flat = m.flat && dry;
  const size_t nfield = 15 + 3 + 5 + 5 + (dry ? 0 : 3) + 2;
  const size_t total = NpNe * nfield + cor;
  if (rc == FEDG_OK) rc = dev_upload<double>(&st->block, nullptr, total);
  if (rc != FEDG_OK) {
    if (st->block) (void)hipFree(st->block);
    delete st;
    return rc;
  }
  double* p = st->block;
  for (int b = 0; b < 3; ++b)
    for (int v = 0; v < 5; ++v, p += NpNe) st->buf[b][v] = p;
  for (int v = 0; v < 5; ++v, p += NpNe) st->VT[v] = p;
  for (int b = 0; b < 3; ++b, p += NpNe) st->DP[b] = p;
  st->DENS_hyd = p; p += NpNe;
  st->PRES_hyd = p; p += NpNe;
  st->THERM_hyd = p; p += NpNe;
  st->DPhydDx = p; p += NpNe;
  st->DPhydDy = p; p += NpNe;
  if (!dry) {
    st->Rtot = p; p += NpNe;
    st->CVtot = p; p += NpNe;
    st->CPtot = p; p += NpNe;
  }
  st->pres_scratch = p; p += 2 * NpNe;
  st->CORIOLIS = p;
  struct {
    double* d;
    const double* h;
    size_t n;
  } ups[] = {{st->DENS_hyd, DENS_hyd, NpNe}, {st->PRES_hyd, PRES_hyd, NpNe}, {st->THERM_hyd, THERM_hyd, NpNe},
             {st->DPhydDx, DPhydDx, NpNe},   {st->DPhydDy, DPhydDy, NpNe},   {st->CORIOLIS, CORIOLIS, cor},
             {st->Rtot, Rtot, NpNe},         {st->CVtot, CVtot, NpNe},       {st->CPtot, CPtot, NpNe}};
  for (auto& u : ups) {
    if (!u.d) continue;
    hipError_t e = hipMemcpy(u.d, u.h, sizeof(double) * u.n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
      fedg_heve_destroy(st);
      return FEDG_ERR_HIP;
    }
  }
  // horizontally homogeneous base state: per-level tables for the flat p = 7 stage
  if (st->flat && m.n1 == 8) {
    const double* f[5] = {DENS_hyd, PRES_hyd, THERM_hyd, DPhydDx, DPhydDy};
    rc = level_tables(m, f, &st->lvl);
    if (rc != FEDG_OK) {
      fedg_heve_destroy(st);
      return rc;
    }
  }
  st->q0 = 0;
  *out = st;
  return FEDG_OK;
}

int fedg_heve_destroy(fedg_heve* st) {
  if (!st) return FEDG_OK;
  if (st->block) (void)hipFree(st->block);
  if (st->lvl) (void)hipFree(st->lvl);
  delete st;
  return FEDG_OK;
}

int fedg_heve_is_flat(const fedg_heve* st) { return st ? st->flat : 0; }

int fedg_heve_set_massflux(fedg_heve* st, double* MFLX_x_tavg, double* MFLX_y_tavg, double* MFLX_z_tavg,
                           double* alph_dens_M, double* alph_dens_P) {
  if (!st) return FEDG_ERR_ARG;
  const bool any = MFLX_x_tavg || MFLX_y_tavg || MFLX_z_tavg || alph_dens_M || alph_dens_P;
  const bool all = MFLX_x_tavg && MFLX_y_tavg && MFLX_z_tavg && alph_dens_M && alph_dens_P;
  if (any && !all) {
    set_error("fedg_heve_set_massflux: give all five arrays or none");
    return FEDG_ERR_ARG;
  }
  if (all && st->fp32) {
    set_error("fedg_heve_set_massflux: the mass-flux save needs the FP64 stage (its DPRES)");
    return FEDG_ERR_UNSUPPORTED;
  }
  st->mflx[0] = MFLX_x_tavg;
  st->mflx[1] = MFLX_y_tavg;
  st->mflx[2] = MFLX_z_tavg;
  st->alph[0] = alph_dens_M;
  st->alph[1] = alph_dens_P;
  return FEDG_OK;
}

// the six physics-tendency pointers in the stepper's order, checked: all or none
static int phys_pointers(const double* DENS_tp, const double* MOMX_tp, const double* MOMY_tp, const double* MOMZ_tp,
                         const double* RHOT_tp, const double* RHOH_p, const double* tp[6]) {
  const double* in[6] = {DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOH_p};
  int given = 0;
  for (int v = 0; v < 6; ++v) given += in[v] != nullptr;
  if (given != 0 && given != 6) {
    set_error("physics tendencies: give all six arrays or none");
    return FEDG_ERR_ARG;
  }
  for (int v = 0; v < 6; ++v) tp[v] = in[v];
  return FEDG_OK;
}

int fedg_heve_set_phys_tend(fedg_heve* st, const double* DENS_tp, const double* MOMX_tp, const double* MOMY_tp,
                            const double* MOMZ_tp, const double* RHOT_tp, const double* RHOH_p) {
  if (!st) return FEDG_ERR_ARG;
  const double* tp[6];
  TRY(phys_pointers(DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p, tp));
  if (tp[0] && st->fp32) {
    set_error("fedg_heve_set_phys_tend: physics tendencies join the FP64 stage only");
    return FEDG_ERR_UNSUPPORTED;
  }
  for (int v = 0; v < 6; ++v) st->tp[v] = tp[v];
  return FEDG_OK;
}

int fedg_add_phy_tend(fedg_ctx* ctx, double* DENS_dt, double* MOMX_dt, double* MOMY_dt, double* MOMZ_dt,
                      double* RHOT_dt, const double* DENS_tp, const double* MOMX_tp, const double* MOMY_tp,
                      const double* MOMZ_tp, const double* RHOT_tp, const double* RHOH_p, const double* PRES,
                      const double* Rtot, const double* CPtot) {
  if (!ctx || !DENS_dt || !MOMX_dt || !MOMY_dt || !MOMZ_dt || !RHOT_dt || !DENS_tp || !MOMX_tp || !MOMY_tp ||
      !MOMZ_tp || !RHOT_tp || !RHOH_p || !PRES || !Rtot || !CPtot)
    return FEDG_ERR_ARG;
  double* T[5] = {DENS_dt, RHOT_dt, MOMZ_dt, MOMX_dt, MOMY_dt};
  const double* tp[6] = {DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOH_p};
  TRY(launch_add_phy_tend(ctx, T, tp, PRES, Rtot, CPtot));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_heve_set_precision(fedg_heve* st, int bits) {
  if (!st || (bits != 32 && bits != 64)) return FEDG_ERR_ARG;
  if (bits == 32 && !st->flat) {
    set_error("the FP32 perturbation form is built for the flat-terrain box with dry air");
    return FEDG_ERR_UNSUPPORTED;
  }
  if (bits == 32 && st->mflx[0]) {
    set_error("the mass-flux save needs the FP64 stage (fedg_heve_set_massflux with NULL first)");
    return FEDG_ERR_UNSUPPORTED;
  }
  if (bits == 32 && st->tp[0]) {
    set_error("physics tendencies join the FP64 stage only (fedg_heve_set_phys_tend with NULL first)");
    return FEDG_ERR_UNSUPPORTED;
  }
  st->fp32 = bits == 32;
  // the pressure perturbation of q^n as the FP64 stage forms it (the FP32
  // stages below p = 7 do not keep it)
  const DevMesh& m = st->ctx->m;
  TRY(launch_dpres(st->ctx, st->DP[st->q0], st->buf[st->q0][1], st->PRES_hyd, st->THERM_hyd, st->Rtot, st->CVtot,
                   st->CPtot, (long)m.Np * m.Ne));
  return FEDG_OK;
}

int fedg_heve_set_modalfilter(fedg_heve* st, int on) {
  if (!st) return FEDG_ERR_ARG;
  if (on && !st->ctx->m.MFh) {
    set_error("modal filter: matrices not set up (fedg_ctx_setup_modalfilter)");
    return FEDG_ERR_ARG;
  }
  st->mfilter = on ? 1 : 0;
  return FEDG_OK;
}

int fedg_heve_set_state(fedg_heve* st, const double* const* q) {
  if (!st || !q) return FEDG_ERR_ARG;
  fedg_ctx* ctx = st->ctx;
  const size_t NpNe = (size_t)ctx->m.Np * ctx->m.Ne;
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (int v = 0; v < 5; ++v) FEDG_HIP_CHECK(hipMemcpy(st->buf[st->q0][v], q[v], sizeof(double) * NpNe, hipMemcpyHostToDevice));
  // the stage kernels read the pressure perturbation of their input state
  TRY(launch_dpres(ctx, st->DP[st->q0], st->buf[st->q0][1], st->PRES_hyd, st->THERM_hyd, st->Rtot, st->CVtot,
                   st->CPtot, (long)NpNe));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return FEDG_OK;
}

int fedg_heve_get_state(fedg_heve* st, double* const* q) {
  if (!st || !q) return FEDG_ERR_ARG;
  const size_t NpNe = (size_t)st->ctx->m.Np * st->ctx->m.Ne;
  FEDG_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
  for (int v = 0; v < 5; ++v) FEDG_HIP_CHECK(hipMemcpy(q[v], st->buf[st->q0][v], sizeof(double) * NpNe, hipMemcpyDeviceToHost));
  return FEDG_OK;
}

int fedg_heve_state_dev(fedg_heve* st, double** q) {
  if (!st || !q) return FEDG_ERR_ARG;
  for (int v = 0; v < 5; ++v) q[v] = st->buf[st->q0][v];
  return FEDG_OK;
}

// One RK step: one fused kernel per stage.  Stage inputs and outputs are
// distinct buffers because neighbouring elements read the stage input.
static int heve_step(fedg_heve* st, std::vector<hipEvent_t>* evs, size_t* ev_i) {
  fedg_ctx* ctx = st->ctx;
  const fedg_tint& t = st->tab;
  TendIn in{};
  in.DENS_hyd = st->DENS_hyd;
  in.PRES_hyd = st->PRES_hyd;
  in.THERM_hyd = st->THERM_hyd;
  in.CORIOLIS = st->CORIOLIS;
  in.DPhydDx = st->DPhydDx;
  in.DPhydDy = st->DPhydDy;
  in.Rtot = st->Rtot;
  in.CVtot = st->CVtot;
  in.CPtot = st->CPtot;
  in.lvl = st->lvl;
  for (int v = 0; v < 6; ++v) in.tp[v] = st->tp[v];
  const int q0 = st->q0;
  int cur = q0;
  int last_out = q0;
  for (int s = 1; s <= t.nstage; ++s) {
    RkOut rk{};
    int outb;
    if (s == t.nstage) {
      outb = (t.nstage > 1) ? q0 : (q0 + 1) % 3;
      rk.mode = (t.nstage > 1) ? 2 : 3;
    } else {
      outb = (s == 1) ? (q0 + 1) % 3 : 3 - q0 - cur;  // the buffer that is neither q^n nor the input
      rk.mode = (s == 1) ? 0 : 1;
    }
    for (int v = 0; v < 5; ++v) {
      in.q[v] = st->buf[cur][v];
      rk.out[v] = st->buf[outb][v];
      rk.var0[v] = st->buf[q0][v];
      rk.varTmp[v] = st->VT[v];
    }
    in.DPRES = st->DP[cur];
    rk.DPout = st->DP[outb];
    rk.cb = t.b_ex[s - 1] * t.dt;
    rk.ca = (s < t.nstage) ? t.dt * t.a_ex[s][s - 1] : 0.0;
    if (st->mflx[0]) {
      // FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:899-915: the stage state
      // as the stage's cal_tend sees it (exterior values of the exchange and
      // the velocity BC, resolved through the neighbour map: the stepper's
      // buffers hold no halo tail), weighted by b_ex(s)
      double* const* qs = st->buf[cur];
      MflxArgs a{};
      a.nbr = 1;
      a.FX = st->mflx[0];
      a.FY = st->mflx[1];
      a.FZ = st->mflx[2];
      a.alM = st->alph[0];
      a.alP = st->alph[1];
      a.DD = qs[0];
      a.MX = qs[3];
      a.MY = qs[4];
      a.MZ = qs[2];
      a.DP = st->DP[cur];
      a.Dh = st->DENS_hyd;
      a.Ph = st->PRES_hyd;
      a.rkstage = s;
      a.hevi = 0;
      a.twh = t.b_ex[s - 1];
      a.twv = t.b_ex[s - 1];
      TRY(launch_save_massflux(ctx, a));
    }
    if (evs) FEDG_HIP_CHECK(hipEventRecord((*evs)[(*ev_i)++], ctx->stream));
    if (st->fp32) TRY(launch_heve_fused32(ctx, in, rk));
    else TRY(launch_heve_fused(ctx, in, rk, st->flat));
    if (evs) FEDG_HIP_CHECK(hipEventRecord((*evs)[(*ev_i)++], ctx->stream));
    cur = outb;
    last_out = outb;
  }
  st->q0 = last_out;
  if (st->mfilter) {  // FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:940-950
    double* const* q = st->buf[st->q0];
    TRY(launch_modalfilter(ctx, q[0], q[3], q[4], q[2], q[1], 1));
    const DevMesh& m = ctx->m;
    TRY(launch_dpres(ctx, st->DP[st->q0], q[1], st->PRES_hyd, st->THERM_hyd, st->Rtot, st->CVtot, st->CPtot,
                     (long)m.Np * m.Ne));
  }
  return FEDG_OK;
}

int fedg_heve_run(fedg_heve* st, int nsteps) {
  if (!st || nsteps < 0) return FEDG_ERR_ARG;
  for (int n = 0; n < nsteps; ++n) TRY(heve_step(st, nullptr, nullptr));
  FEDG_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
  return FEDG_OK;
}

int fedg_heve_run_timed(fedg_heve* st, int nsteps, double* ms_total, double* ms_per_stage_kernel) {
  if (!st || nsteps < 1) return FEDG_ERR_ARG;
  fedg_ctx* ctx = st->ctx;
  const size_t nev = (size_t)nsteps * st->tab.nstage * 2;
  std::vector<hipEvent_t> evs(nev);
  for (auto& e : evs) FEDG_HIP_CHECK(hipEventCreate(&e));
  hipEvent_t t0, t1;
  FEDG_HIP_CHECK(hipEventCreate(&t0));
  FEDG_HIP_CHECK(hipEventCreate(&t1));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  FEDG_HIP_CHECK(hipEventRecord(t0, ctx->stream));
  size_t ev_i = 0;
  int rc = FEDG_OK;
  for (int n = 0; n < nsteps && rc == FEDG_OK; ++n) rc = heve_step(st, &evs, &ev_i);
  FEDG_HIP_CHECK(hipEventRecord(t1, ctx->stream));
  FEDG_HIP_CHECK(hipEventSynchronize(t1));
  if (rc == FEDG_OK) {
    float tot = 0.f;
    FEDG_HIP_CHECK(hipEventElapsedTime(&tot, t0, t1));
    double ksum = 0.0;
    for (size_t e = 0; e + 1 < ev_i; e += 2) {
      float ms = 0.f;
      FEDG_HIP_CHECK(hipEventElapsedTime(&ms, evs[e], evs[e + 1]));
      ksum += ms;
    }
    if (ms_total) *ms_total = tot;
    if (ms_per_stage_kernel) *ms_per_stage_kernel = ksum / (double)(ev_i / 2);
  }
  for (auto& e : evs) (void)hipEventDestroy(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return rc;
}

int fedg_heve_get_pressure(fedg_heve* st, double* PRES_host, double* DPRES_host) {
  if (!st || !PRES_host || !DPRES_host) return FEDG_ERR_ARG;
  fedg_ctx* ctx = st->ctx;
  const size_t NpNe = (size_t)ctx->m.Np * ctx->m.Ne;
  double* PRES = st->pres_scratch;
  double* DPRES = st->pres_scratch + NpNe;
  // dry constants as fields when the stepper dropped them
  double *R = st->Rtot, *Cv = st->CVtot, *Cp = st->CPtot;
  std::vector<double*> tmp;
  if (!R) {
    for (int a = 0; a < 3; ++a) {
      double* d = nullptr;
      TRY(dev_upload<double>(&d, nullptr, NpNe));
      std::vector<double> h(NpNe, a == 0 ? ctx->c.Rdry : a == 1 ? ctx->c.CVdry : ctx->c.CPdry);
      FEDG_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(double) * NpNe, hipMemcpyHostToDevice));
      tmp.push_back(d);
    }
    R = tmp[0];
    Cv = tmp[1];
    Cp = tmp[2];
  }
  int rc = launch_drhot2pres(ctx, PRES, DPRES, st->buf[st->q0][1], st->PRES_hyd, st->THERM_hyd, R, Cv, Cp,
                             (long)NpNe);
  if (rc == FEDG_OK) {
    FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    FEDG_HIP_CHECK(hipMemcpy(PRES_host, PRES, sizeof(double) * NpNe, hipMemcpyDeviceToHost));
    FEDG_HIP_CHECK(hipMemcpy(DPRES_host, DPRES, sizeof(double) * NpNe, hipMemcpyDeviceToHost));
  }
  for (double* d : tmp) (void)hipFree(d);
  return rc;
}


// ---------------------------------------------------------------------------
// HEVI step driver (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:694-870, the
// hevi_flag branch): per stage the implicit tendency (cal_vi), StoreImplicit,
// halo exchange, pressure, boundary conditions, the explicit HEVI tendency
// and the IMEX advance, all on the context's stream.
struct fedg_hevi {
  fedg_ctx* ctx;
  fedg_tint* tint;
  double* block;
  double* q[5];  // (Np,NeA) state, VID order
  double *PRES, *DPRES;
  double *DENS_hyd, *PRES_hyd, *THERM_hyd, *DPhydDx, *DPhydDy, *Rtot, *CVtot, *CPtot, *CORIOLIS;
  double* RHOT_hyd;  // the hydrostatic RHOT of the vertical residual kernels, formed once from PRES_hyd
  int max_it;
  double rtol;
  int mfilter;  // modal filter after every step
  long newton_iters, newton_calls;
  // learned initial guess and training-pair harvest
  fedg::MlpImpl* mlp;
  double scale[15];  // state, tendency-feature and increment scales (5 each, VID order)
  long guess_lines, guess_fallbacks;
  double guess_r2_plain, guess_r2_start;
  float* hx;
  float* hy;
  int hcap, hcount, hper;
  unsigned hseed;
  // time-averaged mass fluxes for the tracers (QA > 0), or null
  double *mflx[3], *alph[2];
  // flat-terrain box at p = 7 with dry air: the explicit tendency on the p = 7
  // layout (base state from per-level tables when homogeneous), in FP64 or,
  // with fp32, in the FP32 perturbation form
  int flat7;
  double* lvl;
  int fp32;
  // physics tendencies (DENS_tp, RHOT_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOH_p) or null
  const double* tp[6];
};

int fedg_hevi_destroy(fedg_hevi* st) {
  if (!st) return FEDG_OK;
  if (st->tint) fedg_tint_destroy(st->tint);
  if (st->block) (void)hipFree(st->block);
  if (st->lvl) (void)hipFree(st->lvl);
  delete st;
  return FEDG_OK;
}

int fedg_hevi_create(fedg_ctx* ctx, const char* scheme, double dt, const double* DENS_hyd, const double* PRES_hyd,
                     const double* THERM_hyd, const double* CORIOLIS, const double* DPhydDx, const double* DPhydDy,
                     const double* Rtot, const double* CVtot, const double* CPtot, int max_it, double rtol,
                     fedg_hevi** out) {
  if (!ctx || !scheme || !out || !DENS_hyd || !PRES_hyd || !THERM_hyd || !CORIOLIS || !DPhydDx || !DPhydDy ||
      !Rtot || !CVtot || !CPtot || max_it < 1)
    return FEDG_ERR_ARG;
  *out = nullptr;
  fedg_hevi* st = new fedg_hevi();
  memset(st, 0, sizeof(*st));
  st->ctx = ctx;
  st->max_it = max_it;
  st->rtol = rtol;
  for (int v = 0; v < 15; ++v) st->scale[v] = 1.0;
  int rc = fedg_tint_create(ctx, scheme, dt, &st->tint);
  if (rc == FEDG_OK && !st->tint->imex) {
    set_error("the HEVI step takes an IMEX scheme (IMEX_ARK232, IMEX_ARK324)");
    rc = FEDG_ERR_ARG;
  }
  HeviWork* w = nullptr;
  if (rc == FEDG_OK) rc = hevi_work_get(ctx, &w);
  const DevMesh& m = ctx->m;
  const size_t nA = (size_t)m.Np * m.NeA, cor = (size_t)m.Nfp * m.Ne2DA;
  if (rc == FEDG_OK) rc = dev_upload<double>(&st->block, nullptr, nA * 16 + cor);
  if (rc != FEDG_OK) {
    fedg_hevi_destroy(st);
    return rc;
  }
  double* p = st->block;
  for (int v = 0; v < 5; ++v, p += nA) st->q[v] = p;
  st->PRES = p; p += nA;
  st->DPRES = p; p += nA;
  double** dst[] = {&st->DENS_hyd, &st->PRES_hyd, &st->THERM_hyd, &st->DPhydDx, &st->DPhydDy,
                    &st->Rtot, &st->CVtot, &st->CPtot};
  const double* src[] = {DENS_hyd, PRES_hyd, THERM_hyd, DPhydDx, DPhydDy, Rtot, CVtot, CPtot};
  for (int a = 0; a < 8; ++a, p += nA) {
    *dst[a] = p;
    hipError_t e = hipMemcpy(p, src[a], sizeof(double) * nA, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
      fedg_hevi_destroy(st);
      return FEDG_ERR_HIP;
    }
  }
  st->RHOT_hyd = p;
  p += nA;
  rc = launch_rhot_hyd(ctx, st->RHOT_hyd, st->PRES_hyd, (long)nA);
  if (rc != FEDG_OK) {
    fedg_hevi_destroy(st);
    return rc;
  }
  st->CORIOLIS = p;
  if (hipMemcpy(p, CORIOLIS, sizeof(double) * cor, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("hipMemcpy of CORIOLIS failed");
    fedg_hevi_destroy(st);
    return FEDG_ERR_HIP;
  }
  const size_t NpNe = (size_t)m.Np * m.Ne;
  const bool dry = all_equal(Rtot, NpNe, ctx->c.Rdry) && all_equal(CVtot, NpNe, ctx->c.CVdry) &&
                   all_equal(CPtot, NpNe, ctx->c.CPdry);
  st->flat7 = m.flat && m.n1 == 8 && m.lift_tensor && dry;
  if (st->flat7) {
    const double* f[5] = {DENS_hyd, PRES_hyd, THERM_hyd, DPhydDx, DPhydDy};
    rc = level_tables(m, f, &st->lvl);
    if (rc != FEDG_OK) {
      fedg_hevi_destroy(st);
      return rc;
    }
  }
  *out = st;
  return FEDG_OK;
}

int fedg_hevi_set_precision(fedg_hevi* st, int bits) {
  if (!st || (bits != 32 && bits != 64)) return FEDG_ERR_ARG;
  if (bits == 32 && !st->flat7) {
    set_error("the FP32 explicit HEVI tendency is built for the flat-terrain box at p = 7 with dry air");
    return FEDG_ERR_UNSUPPORTED;
  }
  st->fp32 = bits == 32;
  return FEDG_OK;
}

int fedg_hevi_set_phys_tend(fedg_hevi* st, const double* DENS_tp, const double* MOMX_tp, const double* MOMY_tp,
                            const double* MOMZ_tp, const double* RHOT_tp, const double* RHOH_p) {
  if (!st) return FEDG_ERR_ARG;
  return phys_pointers(DENS_tp, MOMX_tp, MOMY_tp, MOMZ_tp, RHOT_tp, RHOH_p, st->tp);
}

int fedg_hevi_set_state(fedg_hevi* st, const double* const* q) {
  if (!st || !q) return FEDG_ERR_ARG;
  const size_t n = (size_t)st->ctx->m.Np * st->ctx->m.Ne;
  FEDG_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
  for (int v = 0; v < 5; ++v) FEDG_HIP_CHECK(hipMemcpy(st->q[v], q[v], sizeof(double) * n, hipMemcpyHostToDevice));
  return FEDG_OK;
}

int fedg_hevi_get_state(fedg_hevi* st, double* const* q) {
  if (!st || !q) return FEDG_ERR_ARG;
  const size_t n = (size_t)st->ctx->m.Np * st->ctx->m.Ne;
  FEDG_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
  for (int v = 0; v < 5; ++v) FEDG_HIP_CHECK(hipMemcpy(q[v], st->q[v], sizeof(double) * n, hipMemcpyDeviceToHost));
  return FEDG_OK;
}

int fedg_hevi_set_modalfilter(fedg_hevi* st, int on) {
  if (!st) return FEDG_ERR_ARG;
  if (on && !st->ctx->m.MFh) {
    set_error("modal filter: matrices not set up (fedg_ctx_setup_modalfilter)");
    return FEDG_ERR_ARG;
  }
  st->mfilter = on ? 1 : 0;
  return FEDG_OK;
}

int fedg_hevi_newton_stats(const fedg_hevi* st, long* iterations, long* calls) {
  if (!st) return FEDG_ERR_ARG;
  if (iterations) *iterations = st->newton_iters;
  if (calls) *calls = st->newton_calls;
  return FEDG_OK;
}

static int hevi_step(fedg_hevi* st, hipEvent_t* ev) {
  fedg_ctx* ctx = st->ctx;
  fedg_tint* t = st->tint;
  const DevMesh& m = ctx->m;
  const long n = (long)m.Np * m.Ne;
  const size_t nA = (size_t)m.Np * m.NeA;
  HeviWork* w = nullptr;
  TRY(hevi_work_get(ctx, &w));
  // StoreVar0.  With an explicit first stage (the IMEX tables here) the
  // first stage evaluates the implicit operator without changing q, so it
  // reads var0 from q itself and its StoreImplicit pass writes the copy.
  const bool var0_at_stage1 = t->a_im[0][0] == 0.0;
  if (!var0_at_stage1)
    for (int v = 0; v < 5; ++v) TRY(launch_copy(ctx, t->var0 + v * nA, st->q[v], n));
  for (int s = 1; s <= t->nstage; ++s) {
    const int ind = t->indmap[s - 1];
    const double impl_fac = t->a_im[s - 1][s - 1] * t->dt;
    double* Ti[5];
    double* Te[5];
    const double* qv[5];
    const double* v0[5];
    for (int v = 0; v < 5; ++v) {
      Ti[v] = tend_ptr(t, t->tend_im, ind, v + 1);
      Te[v] = tend_ptr(t, t->tend_ex, ind, v + 1);
      qv[v] = st->q[v];
      v0[v] = (s == 1 && var0_at_stage1) ? st->q[v] : t->var0 + v * nA;
    }
    int it = 0;
    // explicit tendency of the previous stage: a feature of the learned guess
    const double* Tprev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    if (s > 1)
      for (int v = 0; v < 5; ++v) Tprev[v] = tend_ptr(t, t->tend_ex, t->indmap[s - 2], v + 1);
    HeviGuess g{};
    g.mlp = st->mlp;
    for (int v = 0; v < 5; ++v) g.Tex[v] = Tprev[v];
    for (int v = 0; v < 15; ++v) g.scale[v] = st->scale[v];
    g.dt = t->dt;
    // StoreImplicit fused into the pass that forms the implicit tendency
    // (not while harvesting: the harvest reads the stage state after the solve)
    const bool fuse_store = fabs(impl_fac) > 0.0 && !st->hx;
    HeviStore sto{};
    for (int v = 0; v < 5; ++v) {
      sto.q[v] = st->q[v];
      sto.var0[v] = s == 1 ? t->var0 + v * nA : nullptr;
      sto.varTmp[v] = s == 1 ? t->varTmp + v * nA : nullptr;
    }
    if (ev) FEDG_HIP_CHECK(hipEventRecord(ev[0], ctx->stream));
    TRY(hevi_cal_vi(ctx, w, Ti, qv, v0, st->DENS_hyd, st->PRES_hyd, st->Rtot, st->CVtot, st->CPtot, impl_fac,
                    st->max_it, st->rtol, nullptr, &it, st->mlp ? &g : nullptr, fuse_store ? &sto : nullptr,
                    st->RHOT_hyd, s > 1));  // stage 1 formed alph from var0
    if (ev) FEDG_HIP_CHECK(hipEventRecord(ev[1], ctx->stream));
    st->guess_lines += g.lines;
    st->guess_fallbacks += g.fallbacks;
    st->guess_r2_plain += g.r2_plain;
    st->guess_r2_start += g.r2_start;
    if (st->hx && fabs(impl_fac) > 0.0 && st->hcount < st->hcap) {
      const long nlines = (long)m.Ne2D * m.Nfp;
      const int L = (m.Ne / m.Ne2D) * m.n1;
      const int rows = std::min(st->hper, st->hcap - st->hcount);
      st->hseed = st->hseed * 1664525u + 1013904223u;
      const long line0 = (long)(st->hseed % (unsigned)nlines);
      long stride = std::max(1L, nlines / rows) | 1L;  // odd stride spreads the lines
      TRY(hevi_harvest(ctx, w, qv, v0, Tprev, st->scale, impl_fac, t->dt, line0, stride, rows,
                       st->hx + (size_t)st->hcount * (10 * L + 1), st->hy + (size_t)st->hcount * (5 * L)));
      st->hcount += rows;
    }
    if (fabs(impl_fac) > 0.0) {
      st->newton_iters += it;
      st->newton_calls += 1;
    }
    if (!fuse_store)  // StoreImplicit
      TRY(launch_store_implicit(ctx, st->q, Ti, s == 1 ? sto.var0 : nullptr, s == 1 ? sto.varTmp : nullptr,
                                impl_fac));
    for (int v = 0; v < 5; ++v) TRY(launch_halo_exchange(ctx, st->q[v]));
    // PRES only for the physics tendencies; dry constants on the verified dry box
    TRY(launch_drhot2pres(ctx, st->tp[0] ? st->PRES : nullptr, st->DPRES, st->q[1], st->PRES_hyd, st->THERM_hyd,
                          st->flat7 ? nullptr : st->Rtot, st->flat7 ? nullptr : st->CVtot,
                          st->flat7 ? nullptr : st->CPtot, n));
    TRY(launch_halo_exchange(ctx, st->DPRES));
    TRY(launch_apply_bc(ctx, st->q[3], st->q[4], st->q[2]));
    TendIn in{};
    for (int v = 0; v < 5; ++v) in.q[v] = st->q[v];
    in.DPRES = st->DPRES;
    in.DENS_hyd = st->DENS_hyd;
    in.PRES_hyd = st->PRES_hyd;
    in.THERM_hyd = st->THERM_hyd;
    in.CORIOLIS = st->CORIOLIS;
    in.DPhydDx = st->DPhydDx;
    in.DPhydDy = st->DPhydDy;
    in.Rtot = st->Rtot;
    in.CVtot = st->CVtot;
    in.CPtot = st->CPtot;
    if (ev) FEDG_HIP_CHECK(hipEventRecord(ev[2], ctx->stream));
    if (st->flat7) {
      in.Rtot = in.CVtot = in.CPtot = nullptr;  // dry constants
      in.lvl = st->lvl;
      TRY(launch_hevi_explicit7(ctx, in, Te, st->fp32));
    } else {
      TRY(launch_hevi_explicit(ctx, in, Te));
    }
    if (ev) FEDG_HIP_CHECK(hipEventRecord(ev[3], ctx->stream));
    // add_phy_tend on the explicit tendency (FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:846-856)
    if (st->tp[0]) TRY(launch_add_phy_tend(ctx, Te, st->tp, st->PRES, st->Rtot, st->CPtot));
    if (st->mflx[0]) {
      // FDS/scale_atm_dyn_dgm_driver_nonhydro3d.F90:899-915: weights b_ex(s)
      // horizontally and b_im(s) vertically (IMEX), is_hevi; the state and
      // DPRES carry their exchanged halo and the velocity BC here
      MflxArgs a{};
      a.FX = st->mflx[0];
      a.FY = st->mflx[1];
      a.FZ = st->mflx[2];
      a.alM = st->alph[0];
      a.alP = st->alph[1];
      a.DD = st->q[0];
      a.MX = st->q[3];
      a.MY = st->q[4];
      a.MZ = st->q[2];
      a.DP = st->DPRES;
      a.Dh = st->DENS_hyd;
      a.Ph = st->PRES_hyd;
      a.rkstage = s;
      a.hevi = 1;
      a.twh = t->b_ex[s - 1];
      a.twv = t->imex ? t->b_im[s - 1] : t->b_ex[s - 1];
      TRY(launch_save_massflux(ctx, a));
    }
    TRY(imex_advance_all(t, s, st->q));
    if (ev) ev += 4;
  }
  if (st->mfilter) TRY(launch_modalfilter(ctx, st->q[0], st->q[3], st->q[4], st->q[2], st->q[1], 1));
  return FEDG_OK;
}

int fedg_hevi_set_massflux(fedg_hevi* st, double* MFLX_x_tavg, double* MFLX_y_tavg, double* MFLX_z_tavg,
                           double* alph_dens_M, double* alph_dens_P) {
  if (!st) return FEDG_ERR_ARG;
  const bool any = MFLX_x_tavg || MFLX_y_tavg || MFLX_z_tavg || alph_dens_M || alph_dens_P;
  const bool all = MFLX_x_tavg && MFLX_y_tavg && MFLX_z_tavg && alph_dens_M && alph_dens_P;
  if (any && !all) {
    set_error("fedg_hevi_set_massflux: give all five arrays or none");
    return FEDG_ERR_ARG;
  }
  st->mflx[0] = MFLX_x_tavg;
  st->mflx[1] = MFLX_y_tavg;
  st->mflx[2] = MFLX_z_tavg;
  st->alph[0] = alph_dens_M;
  st->alph[1] = alph_dens_P;
  return FEDG_OK;
}

int fedg_hevi_run(fedg_hevi* st, int nsteps) {
  if (!st || nsteps < 0) return FEDG_ERR_ARG;
  for (int k = 0; k < nsteps; ++k) TRY(hevi_step(st, nullptr));
  FEDG_HIP_CHECK(hipStreamSynchronize(st->ctx->stream));
  return FEDG_OK;
}

int fedg_hevi_run_timed(fedg_hevi* st, int nsteps, double* ms_total, double* ms_implicit_per_stage,
                        double* ms_explicit_per_stage) {
  if (!st || nsteps < 1) return FEDG_ERR_ARG;
  fedg_ctx* ctx = st->ctx;
  const int ns = st->tint->nstage;
  std::vector<hipEvent_t> evs((size_t)nsteps * ns * 4);
  for (auto& e : evs) FEDG_HIP_CHECK(hipEventCreate(&e));
  hipEvent_t t0, t1;
  FEDG_HIP_CHECK(hipEventCreate(&t0));
  FEDG_HIP_CHECK(hipEventCreate(&t1));
  FEDG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  FEDG_HIP_CHECK(hipEventRecord(t0, ctx->stream));
  int rc = FEDG_OK;
  for (int k = 0; k < nsteps && rc == FEDG_OK; ++k) rc = hevi_step(st, evs.data() + (size_t)k * ns * 4);
  FEDG_HIP_CHECK(hipEventRecord(t1, ctx->stream));
  FEDG_HIP_CHECK(hipEventSynchronize(t1));
  if (rc == FEDG_OK) {
    float tot = 0.f;
    FEDG_HIP_CHECK(hipEventElapsedTime(&tot, t0, t1));
    double im = 0.0, ex = 0.0;
    for (size_t e = 0; e < evs.size(); e += 4) {
      float a = 0.f, b = 0.f;
      FEDG_HIP_CHECK(hipEventElapsedTime(&a, evs[e], evs[e + 1]));
      FEDG_HIP_CHECK(hipEventElapsedTime(&b, evs[e + 2], evs[e + 3]));
      im += a;
      ex += b;
    }
    const double cnt = (double)nsteps * ns;
    if (ms_total) *ms_total = tot;
    if (ms_implicit_per_stage) *ms_implicit_per_stage = im / cnt;
    if (ms_explicit_per_stage) *ms_explicit_per_stage = ex / cnt;
  }
  for (auto& e : evs) (void)hipEventDestroy(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return rc;
}


// ---------------------------------------------------------------------------
// Per-column MLP and its use as the HEVI Newton initial guess

int fedg_mlp_create(int n_in, int n_hidden, int n_hidden_layers, int n_out, unsigned seed, fedg_mlp** out) {
  if (!out) return FEDG_ERR_ARG;
  *out = nullptr;
  MlpImpl* p = nullptr;
  TRY(mlp_create(n_in, n_hidden, n_hidden_layers, n_out, seed, &p));
  *out = new fedg_mlp{p};
  return FEDG_OK;
}

int fedg_mlp_destroy(fedg_mlp* m) {
  if (!m) return FEDG_OK;
  mlp_destroy(m->impl);
  delete m;
  return FEDG_OK;
}

int fedg_mlp_param_count(const fedg_mlp* m, long* n) {
  if (!m || !n) return FEDG_ERR_ARG;
  *n = mlp_param_count(m->impl);
  return FEDG_OK;
}

int fedg_mlp_get_params(fedg_mlp* m, float* params_host, float* grads_host) {
  if (!m) return FEDG_ERR_ARG;
  return mlp_get(m->impl, params_host, grads_host);
}

int fedg_mlp_set_params(fedg_mlp* m, const float* params_host) {
  if (!m || !params_host) return FEDG_ERR_ARG;
  return mlp_set(m->impl, params_host);
}

int fedg_mlp_forward(fedg_mlp* m, const float* X, float* Y, int batch) {
  if (!m || !X || !Y || batch < 1) return FEDG_ERR_ARG;
  return mlp_forward(m->impl, X, Y, batch);
}

int fedg_mlp_loss_grad(fedg_mlp* m, const float* X, const float* T, int batch, float* loss) {
  if (!m || !X || !T || batch < 1) return FEDG_ERR_ARG;
  return mlp_loss_grad(m->impl, X, T, batch, loss);
}

int fedg_mlp_adam_step(fedg_mlp* m, float lr, float beta1, float beta2, float eps) {
  if (!m || !(lr > 0.f)) return FEDG_ERR_ARG;
  return mlp_adam(m->impl, lr, beta1, beta2, eps);
}

int fedg_hevi_set_guess(fedg_hevi* st, fedg_mlp* mlp, const double* scale) {
  if (!st || !scale) return FEDG_ERR_ARG;
  for (int v = 0; v < 15; ++v)
    if (!(scale[v] > 0.0)) return FEDG_ERR_ARG;
  for (int v = 0; v < 15; ++v) st->scale[v] = scale[v];
  st->mlp = mlp ? mlp->impl : nullptr;
  return FEDG_OK;
}

int fedg_hevi_guess_stats(const fedg_hevi* st, long* lines, long* fallbacks) {
  if (!st) return FEDG_ERR_ARG;
  if (lines) *lines = st->guess_lines;
  if (fallbacks) *fallbacks = st->guess_fallbacks;
  return FEDG_OK;
}

int fedg_hevi_guess_residuals(const fedg_hevi* st, double* r2_plain, double* r2_start) {
  if (!st) return FEDG_ERR_ARG;
  if (r2_plain) *r2_plain = st->guess_r2_plain;
  if (r2_start) *r2_start = st->guess_r2_start;
  return FEDG_OK;
}

int fedg_hevi_harvest(fedg_hevi* st, float* X_dev, float* Y_dev, int capacity, int per_stage, unsigned seed) {
  if (!st || capacity < 0 || (capacity > 0 && (!X_dev || !Y_dev || per_stage < 1))) return FEDG_ERR_ARG;
  st->hx = capacity > 0 ? X_dev : nullptr;
  st->hy = capacity > 0 ? Y_dev : nullptr;
  st->hcap = capacity;
  st->hcount = 0;
  st->hper = per_stage;
  st->hseed = seed;
  return FEDG_OK;
}

int fedg_hevi_harvest_count(const fedg_hevi* st, int* count) {
  if (!st || !count) return FEDG_ERR_ARG;
  *count = st->hcount;
  return FEDG_OK;
}

}  // extern "C"
