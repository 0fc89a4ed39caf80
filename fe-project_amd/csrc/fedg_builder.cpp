// Host-side operator and mesh construction.  See fedg_builder.h for the layout
// conventions.  Each routine cites the reference routine whose arithmetic it
// restates; the order of floating-point operations follows the reference where
// it affects the last bits of the tables (normals, Fscale, Escale, D1D).
#include "fedg_builder.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace fedg {

namespace {

// P_0..P_N at x by the three-term recurrence
// (CMN/scale_polynomial.F90:199-222, Polynomial_GenLegendrePoly_sub).
void legendre_all(int N, double x, double* P) {
  P[0] = 1.0;
  if (N == 0) return;
  P[1] = x;
  for (int n = 2; n <= N; ++n)
    P[n] = ((double)(2 * n - 1) * x * P[n - 1] - (double)(n - 1) * P[n - 2]) / (double)n;
}

// P_N and P'_N at x (derivative recurrence of CMN/scale_polynomial.F90:240-263).
void legendre_and_deriv(int N, double x, double* PN, double* dPN) {
  std::vector<double> P(N + 1), G(N + 1);
  legendre_all(N, x, P.data());
  G[0] = 0.0;
  if (N >= 1) G[1] = 1.0;
  for (int n = 2; n <= N; ++n) G[n] = 2.0 * x * G[n - 1] - G[n - 2] + P[n - 1];
  *PN = P[N];
  *dPN = G[N];
}

// Legendre-Gauss-Lobatto points: x_0=-1, x_N=1 and the N-1 roots of P'_N in
// ascending order.  The reference obtains the interior roots as eigenvalues of
// the Jacobi(1,1) matrix with LAPACK dstev (CMN/scale_polynomial.F90:334-373);
// here they come from Newton's method on P'_N, which reaches the same roots to
// the last bit or two.
std::vector<double> lgl_points(int N) {
  std::vector<double> x(N + 1);
  x[0] = -1.0;
  x[N] = 1.0;
  const double pi = 3.14159265358979323846;
  for (int j = 1; j < N; ++j) {
    double xi = -std::cos(pi * (double)j / (double)N);
    for (int it = 0; it < 100; ++it) {
      double PN, dPN;
      legendre_and_deriv(N, xi, &PN, &dPN);
      // (1-x^2) P''_N = 2x P'_N - N(N+1) P_N  (Legendre equation)
      double d2 = (2.0 * xi * dPN - (double)(N * (N + 1)) * PN) / (1.0 - xi * xi);
      double dx = dPN / d2;
      xi -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    x[j] = xi;
  }
  return x;
}

// Gauss-Jordan inverse with partial pivoting of a column-major n x n matrix
// (stands in for linalgebra_inv = LAPACK dgetrf/dgetri, CMN/scale_linalgebra.F90).
std::vector<double> inverse(const std::vector<double>& A, int n) {
  std::vector<double> a(A), inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i + n * i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r + n * c]) > std::fabs(a[piv + n * c])) piv = r;
    if (a[piv + n * c] == 0.0) throw std::runtime_error("singular matrix");
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[c + n * k], a[piv + n * k]);
        std::swap(inv[c + n * k], inv[piv + n * k]);
      }
    double d = a[c + n * c];
    for (int k = 0; k < n; ++k) {
      a[c + n * k] /= d;
      inv[c + n * k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r + n * c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r + n * k] -= f * a[c + n * k];
        inv[r + n * k] -= f * inv[c + n * k];
      }
    }
  }
  return inv;
}

}  // namespace

void build_refelem(int porder, RefElem& e) {
  if (porder < 1 || porder > 15) throw std::runtime_error("polynomial order must be in 1..15");
  const int N = porder, n1 = N + 1;
  e.p = N;
  e.n1 = n1;
  e.Np = n1 * n1 * n1;
  e.Nfp_h = n1 * n1;
  e.Nfp_v = n1 * n1;
  e.NfpTot = 4 * e.Nfp_h + 2 * e.Nfp_v;

  e.x1d = lgl_points(N);
  // P1D(n,l) = P_{l-1}(x_n)
  std::vector<double> P1D(n1 * n1);
  for (int n = 0; n < n1; ++n) {
    std::vector<double> P(n1);
    legendre_all(N, e.x1d[n], P.data());
    for (int l = 0; l < n1; ++l) P1D[n + n1 * l] = P[l];
  }
  // Gauss-Lobatto weights (CMN/scale_polynomial.F90:278-296)
  e.w1d.resize(n1);
  for (int n = 0; n < n1; ++n) {
    double PN = P1D[n + n1 * N];
    e.w1d[n] = 2.0 / ((double)(N * (N + 1)) * (PN * PN));
  }
  // Derivative of the Lagrange basis at the LGL points with the negative-sum
  // diagonal (CMN/scale_polynomial.F90:126-161); lr(k,n) = l_k'(x_n), D1D(i,l) = lr(l,i).
  std::vector<double> lr(n1 * n1, 0.0);
  for (int n = 0; n < n1; ++n) {
    double lr_nn = 0.0;
    for (int k = 0; k < n1; ++k) {
      if (k == n) continue;
      double v = P1D[n + n1 * N] / (P1D[k + n1 * N] * (e.x1d[n] - e.x1d[k]));
      lr[k + n1 * n] = v;
      lr_nn = lr_nn + v;
    }
    lr[n + n1 * n] = -lr_nn;
  }
  e.D1D.resize(n1 * n1);
  for (int i = 0; i < n1; ++i)
    for (int l = 0; l < n1; ++l) e.D1D[i + n1 * l] = lr[l + n1 * i];

  // Vandermonde of the orthonormal Legendre basis (ELM/scale_element_line.F90:150-158)
  e.V1D.resize(n1 * n1);
  for (int n = 0; n < n1; ++n)
    for (int l = 0; l < n1; ++l) e.V1D[n + n1 * l] = P1D[n + n1 * l] * std::sqrt((double)l + 0.5);
  e.invV1D = inverse(e.V1D, n1);

  // Inverse of the exact (non-lumped) 1D mass matrix, M^-1 = V V^T
  // (ELM/scale_element_base.F90:224-240).
  e.invM1D.assign(n1 * n1, 0.0);
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) {
      double s = 0.0;
      for (int l = 0; l < n1; ++l) s += e.V1D[i + n1 * l] * e.V1D[j + n1 * l];
      e.invM1D[i + n1 * j] = s;
    }

  // Lift_mat(i,j,k,f): with the exact mass matrices of the reference element
  // (LumpedMassMatFlag=.false., the default of every atmosphere mesh), the
  // lifting matrix M^-1 E (ELM/scale_element_hexahedral.F90:300-399) restricted
  // to the line through a face node equals one column of the 1D M^-1; the
  // tensor-product kernel keeps exactly those entries
  // (ELM/scale_element_operation_tensorprod3D.F90:5512-5569).
  e.Lift_mat.assign((size_t)n1 * n1 * n1 * 6, 0.0);
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) {
        size_t p = i + n1 * (j + n1 * k);
        size_t S = (size_t)n1 * n1 * n1;
        e.Lift_mat[p + S * 0] = e.invM1D[j + n1 * 0];   // face 1: y = -1
        e.Lift_mat[p + S * 1] = e.invM1D[i + n1 * N];   // face 2: x = +1
        e.Lift_mat[p + S * 2] = e.invM1D[j + n1 * N];   // face 3: y = +1
        e.Lift_mat[p + S * 3] = e.invM1D[i + n1 * 0];   // face 4: x = -1
        e.Lift_mat[p + S * 4] = e.invM1D[k + n1 * 0];   // face 5: z = -1
        e.Lift_mat[p + S * 5] = e.invM1D[k + n1 * N];   // face 6: z = +1
      }

  // Vertical modal truncation V diag(1,..,1,0) V^-1, stored transposed
  // (ELM/scale_element_operation_tensorprod3D.F90:5560-5563).
  std::vector<double> invVm(e.invV1D);
  for (int l = 0; l < n1; ++l) invVm[N + n1 * l] = 0.0;
  e.IntrpMat_VPOrdM1_tr.assign(n1 * n1, 0.0);
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) {
      double s = 0.0;
      for (int m = 0; m < n1; ++m) s += e.V1D[i + n1 * m] * invVm[m + n1 * j];
      e.IntrpMat_VPOrdM1_tr[j + n1 * i] = s;
    }

  // Face masks and index maps (ELM/scale_element_hexahedral.F90:161-230)
  auto nid = [n1](int i, int j, int k) { return 1 + i + n1 * (j + n1 * k); };
  e.Fmask_h.resize(e.Nfp_h * 4);
  e.Fmask_v.resize(e.Nfp_v * 2);
  for (int k = 0; k < n1; ++k)
    for (int a = 0; a < n1; ++a) {
      int m = a + n1 * k;
      e.Fmask_h[m + e.Nfp_h * 0] = nid(a, 0, k);
      e.Fmask_h[m + e.Nfp_h * 1] = nid(N, a, k);
      e.Fmask_h[m + e.Nfp_h * 2] = nid(a, N, k);
      e.Fmask_h[m + e.Nfp_h * 3] = nid(0, a, k);
    }
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i) {
      int m = i + n1 * j;
      e.Fmask_v[m + e.Nfp_v * 0] = nid(i, j, 0);
      e.Fmask_v[m + e.Nfp_v * 1] = nid(i, j, N);
    }
  e.IndexH2Dto3D.resize(e.Np);
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) e.IndexH2Dto3D[i + n1 * (j + n1 * k)] = 1 + i + n1 * j;
  // quadrilateral edge masks (ELM/scale_element_quadrilateral.F90:145-148)
  std::vector<int> qmask(n1 * 4);
  for (int a = 0; a < n1; ++a) {
    qmask[a + n1 * 0] = 1 + a;
    qmask[a + n1 * 1] = 1 + N + n1 * a;
    qmask[a + n1 * 2] = 1 + a + n1 * N;
    qmask[a + n1 * 3] = 1 + n1 * a;
  }
  e.IndexH2Dto3D_bnd.resize(e.NfpTot);
  for (int f = 0; f < 4; ++f)
    for (int v = 0; v < n1; ++v)
      for (int h = 0; h < n1; ++h) e.IndexH2Dto3D_bnd[h + n1 * v + e.Nfp_h * f] = qmask[h + n1 * f];
  for (int f = 0; f < 2; ++f)
    for (int h2 = 0; h2 < n1; ++h2)
      for (int h1 = 0; h1 < n1; ++h1)
        e.IndexH2Dto3D_bnd[h1 + n1 * h2 + e.Nfp_v * f + 4 * e.Nfp_h] = 1 + h1 + n1 * h2;
}

int build_box(const BoxDesc& d, const RefElem& e, BoxMesh& m) {
  const int NeX = d.NeX, NeY = d.NeY, NeZ = d.NeZ;
  if (NeX < 1 || NeY < 1 || NeZ < 1) return 1;
  if (d.porder != e.p) return 1;
  const int n1 = e.n1, Np = e.Np, NfpTot = e.NfpTot, Nfp_h = e.Nfp_h, Nfp_v = e.Nfp_v;
  const int Ne = NeX * NeY * NeZ;
  m.Ne = Ne;
  m.NeA = Ne + 2 * (NeX + NeY) * NeZ + 2 * NeX * NeY;  // MSH/scale_mesh_cubedom3d.F90:384
  m.Ne2D = NeX * NeY;
  m.Ne2DA = NeX * NeY + 2 * (NeX + NeY);
  m.nB = 2 * (NeX + NeY) * NeZ * Nfp_h + 2 * NeX * NeY * Nfp_v;
  const size_t NpNe = (size_t)Np * Ne;

  // Vertex coordinates (MSH/scale_meshutil_3d.F90:38-118) with the uniform
  // vertical levels of MSH/scale_mesh_cubedom3d.F90:156-165.
  const int NvX = NeX + 1, NvY = NeY + 1, NvZ = NeZ + 1;
  std::vector<double> FZ(NvZ);
  FZ[0] = d.zmin;
  FZ[NeZ] = d.zmax;
  double dz = (d.zmax - d.zmin) / (double)NeZ;
  for (int k = 1; k < NeZ; ++k) FZ[k] = FZ[k - 1] + dz;
  auto vposx = [&](int i) { return (d.xmax - d.xmin) * (double)i / (double)(NvX - 1) + d.xmin; };
  auto vposy = [&](int j) { return (d.ymax - d.ymin) * (double)j / (double)(NvY - 1) + d.ymin; };

  m.pos_en.assign(NpNe * 3, 0.0);
  m.J.assign(NpNe, 0.0);
  m.Escale.assign(NpNe * 9, 0.0);
  m.sJ.assign((size_t)NfpTot * Ne, 0.0);
  m.Fscale.assign((size_t)NfpTot * Ne, 0.0);
  m.normal_fn.assign((size_t)NfpTot * Ne * 3, 0.0);
  m.EMap3Dto2D.resize(Ne);

  // face-local node -> element node (1-based) in the reference's face order
  std::vector<int> fmask(NfpTot);
  for (int f = 0; f < 4; ++f)
    for (int a = 0; a < Nfp_h; ++a) fmask[a + Nfp_h * f] = e.Fmask_h[a + Nfp_h * f];
  for (int f = 0; f < 2; ++f)
    for (int a = 0; a < Nfp_v; ++a) fmask[4 * Nfp_h + a + Nfp_v * f] = e.Fmask_v[a + Nfp_v * f];

  for (int kk = 0; kk < NeZ; ++kk)
    for (int jj = 0; jj < NeY; ++jj)
      for (int ii = 0; ii < NeX; ++ii) {
        const int ke = ii + NeX * (jj + NeY * kk);
        m.EMap3Dto2D[ke] = 1 + ii + NeX * jj;
        // vertices 1,2,3,5 of EToV are enough for the affine box map
        double vx1 = vposx(ii), vx2 = vposx(ii + 1);
        double vy1 = vposy(jj), vy3 = vposy(jj + 1);
        double vz1 = FZ[kk], vz5 = FZ[kk + 1];
        // MSH/scale_mesh_cubedom3d.F90:545-574
        const double xX = 0.5 * (vx2 - vx1), xY = 0.0, xZ = 0.0;
        const double yX = 0.0, yY = 0.5 * (vy3 - vy1), yZ = 0.0;
        const double zX = 0.0, zY = 0.0, zZ = 0.5 * (vz5 - vz1);
        for (int p = 0; p < Np; ++p) {
          int i = p % n1, j = (p / n1) % n1, k = p / (n1 * n1);
          size_t o = p + (size_t)Np * ke;
          m.pos_en[o + NpNe * 0] = vx1 + 0.5 * (e.x1d[i] + 1.0) * (vx2 - vx1);
          m.pos_en[o + NpNe * 1] = vy1 + 0.5 * (e.x1d[j] + 1.0) * (vy3 - vy1);
          m.pos_en[o + NpNe * 2] = vz1 + 0.5 * (e.x1d[k] + 1.0) * (vz5 - vz1);
          // MSH/scale_mesh_base3d.F90:264-280
          double Jv = xX * (yY * zZ - zY * yZ) - yX * (xY * zZ - zY * xZ) + zX * (xY * yZ - yY * xZ);
          m.J[o] = Jv;
          double* E = m.Escale.data();
          E[o + NpNe * (0 + 3 * 0)] = (yY * zZ - zY * yZ) / Jv;
          E[o + NpNe * (0 + 3 * 1)] = -(xY * zZ - zY * xZ) / Jv;
          E[o + NpNe * (0 + 3 * 2)] = (xY * yZ - yY * xZ) / Jv;
          E[o + NpNe * (1 + 3 * 0)] = -(yX * zZ - zX * yZ) / Jv;
          E[o + NpNe * (1 + 3 * 1)] = (xX * zZ - zX * xZ) / Jv;
          E[o + NpNe * (1 + 3 * 2)] = -(xX * yZ - yX * xZ) / Jv;
          E[o + NpNe * (2 + 3 * 0)] = (yX * zY - zX * yY) / Jv;
          E[o + NpNe * (2 + 3 * 1)] = -(xX * zY - zX * xY) / Jv;
          E[o + NpNe * (2 + 3 * 2)] = (xX * yY - yX * xY) / Jv;
        }
        // normals, sJ, Fscale (MSH/scale_mesh_base3d.F90:282-301,
        // MSH/scale_mesh_cubedom3d.F90:577-602)
        for (int fp = 0; fp < NfpTot; ++fp) {
          int face = fp < 4 * Nfp_h ? fp / Nfp_h : 4 + (fp - 4 * Nfp_h) / Nfp_v;
          size_t o = (size_t)(fmask[fp] - 1) + (size_t)Np * ke;
          double n[3];
          for (int dd = 0; dd < 3; ++dd) {
            const double* E = m.Escale.data();
            double e1 = E[o + NpNe * (0 + 3 * dd)], e2 = E[o + NpNe * (1 + 3 * dd)], e3 = E[o + NpNe * (2 + 3 * dd)];
            switch (face) {
              case 0: n[dd] = -e2; break;
              case 1: n[dd] = +e1; break;
              case 2: n[dd] = +e2; break;
              case 3: n[dd] = -e1; break;
              case 4: n[dd] = -e3; break;
              default: n[dd] = +e3; break;
            }
          }
          double s = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
          size_t of = fp + (size_t)NfpTot * ke;
          size_t NfNe = (size_t)NfpTot * Ne;
          for (int dd = 0; dd < 3; ++dd) m.normal_fn[of + NfNe * dd] = n[dd] / s;
          s = s * m.J[o];
          m.sJ[of] = s;
          m.Fscale[of] = s / m.J[o];
        }
      }

  // Interior face maps (MSH/scale_meshutil_3d.F90:264-475): on the conforming
  // box the closest neighbour face node is the one with the same tangential
  // indices.  Faces on the domain boundary are then sent into halo slots in
  // the order of MeshUtil3D_genPatchBoundaryMap (:478-641): lateral boundaries
  // y=ymin, x=xmax, y=ymax, x=xmin, then z=zmin, z=zmax, each listing elements
  // in ascending order.
  m.vmapM.resize((size_t)NfpTot * Ne);
  m.vmapP.resize((size_t)NfpTot * Ne);
  auto elem_id = [&](int i, int j, int k) { return i + NeX * (j + NeY * k); };
  const int opp[6] = {2, 3, 0, 1, 5, 4};
  for (int kk = 0; kk < NeZ; ++kk)
    for (int jj = 0; jj < NeY; ++jj)
      for (int ii = 0; ii < NeX; ++ii) {
        const int ke = elem_id(ii, jj, kk);
        for (int fp = 0; fp < NfpTot; ++fp) {
          int face = fp < 4 * Nfp_h ? fp / Nfp_h : 4 + (fp - 4 * Nfp_h) / Nfp_v;
          int a = fp < 4 * Nfp_h ? fp % Nfp_h : (fp - 4 * Nfp_h) % Nfp_v;
          m.vmapM[fp + (size_t)NfpTot * ke] = fmask[fp] + Np * ke;
          int ni = ii, nj = jj, nk = kk;
          switch (face) {
            case 0: nj -= 1; break;
            case 1: ni += 1; break;
            case 2: nj += 1; break;
            case 3: ni -= 1; break;
            case 4: nk -= 1; break;
            default: nk += 1; break;
          }
          int val;
          if (ni < 0 || ni >= NeX || nj < 0 || nj >= NeY || nk < 0 || nk >= NeZ) {
            val = 0;  // set below
          } else {
            int of = opp[face];
            int fpn = of < 4 ? a + Nfp_h * of : 4 * Nfp_h + a + Nfp_v * (of - 4);
            val = fmask[fpn] + Np * elem_id(ni, nj, nk);
          }
          m.vmapP[fp + (size_t)NfpTot * ke] = val;
        }
      }
  m.vmapB.resize(m.nB);
  m.halo_face.resize(m.nB);
  std::vector<int> seg_start(6), seg_len(6);
  int counter = 0;
  for (int b = 0; b < 6; ++b) {
    seg_start[b] = counter;
    for (int ke = 0; ke < Ne; ++ke) {
      int ii = ke % NeX, jj = (ke / NeX) % NeY, kk = ke / (NeX * NeY);
      bool on = false;
      switch (b) {
        case 0: on = (jj == 0); break;
        case 1: on = (ii == NeX - 1); break;
        case 2: on = (jj == NeY - 1); break;
        case 3: on = (ii == 0); break;
        case 4: on = (kk == 0); break;
        default: on = (kk == NeZ - 1); break;
      }
      if (!on) continue;
      int nfp = b < 4 ? Nfp_h : Nfp_v;
      int fbase = b < 4 ? Nfp_h * b : 4 * Nfp_h + Nfp_v * (b - 4);
      for (int a = 0; a < nfp; ++a) {
        int fp = fbase + a;
        m.vmapP[fp + (size_t)NfpTot * ke] = (int)NpNe + counter + 1;
        m.vmapB[counter] = fmask[fp] + Np * ke;
        m.halo_face[counter] = b + 1;
        ++counter;
      }
    }
    seg_len[b] = counter - seg_start[b];
  }
  if (counter != m.nB) return 2;
  // Halo contents after MeshFieldComm_Exchange on a single tile
  // (DAT/scale_meshfieldcomm_cubedom3d.F90:245-344, DAT/scale_meshfieldcomm_base.F90:886-894,
  // tile connectivity MSH/scale_meshutil_3d.F90:750-895): a periodic boundary
  // receives the opposite boundary's face values, a closed one its own.
  const int periodic[6] = {d.periodic_y, d.periodic_x, d.periodic_y, d.periodic_x, d.periodic_z, d.periodic_z};
  m.halo_src.resize(m.nB);
  for (int b = 0; b < 6; ++b) {
    int s = periodic[b] ? opp[b] : b;
    if (seg_len[s] != seg_len[b]) return 3;
    for (int a = 0; a < seg_len[b]; ++a) m.halo_src[seg_start[b] + a] = m.vmapB[seg_start[s] + a];
  }

  // Flat terrain metrics (MSH/scale_mesh_base3d.F90:304-316; MSH/scale_meshutil_vcoord.F90:81-176 with topo=0)
  m.Gsqrt.assign((size_t)Np * m.NeA, 1.0);
  m.GI3.assign((size_t)Np * m.NeA * 2, 0.0);
  m.GsqrtH.assign((size_t)Nfp_v * m.Ne2D, 1.0);
  return 0;
}

// Terrain-following coordinate z = zeta + (1 - zeta / zTop) topo
// (MeshUtil_VCoord_GetMetric, MSH/scale_meshutil_vcoord.F90:81-176):
//   GsqrtV = dz/dzeta = 1 - topo / zTop,
//   G13, G23 = -(1 - zeta / zTop) / GsqrtV * d topo / dx, dy,
// the horizontal gradient of the topography taken as the DG gradient on the
// 2-D quadrilateral mesh: Escale D topo plus the lift of the central jump
// 0.5 (topo+ - topo-) n across the element faces (cal_del_flux, :181-209),
// with the exact-mass lift of the element (its 1-D columns invM1D(:, face
// node), ELM/scale_element_quadrilateral.F90:218-241); then the metric tensor
// GIJ (the identity in this Cartesian box) applied as in
// calc_vcoordinate_metrics_lc (MSH/scale_mesh_topography.F90:213-258), and the
// halo slots filled by the single-tile exchange before
// Set_geometric_with_vcoord stores Gsqrt = GsqrtV and GI3
// (MSH/scale_mesh_cubedom3d.F90:292-334).
int set_topography(const BoxDesc& d, const RefElem& e, const double* topo, double zTop, BoxMesh& m,
                   std::vector<double>& zlev) {
  if (!topo || !(zTop > 0.0) || d.porder != e.p) return 1;
  const int NeX = d.NeX, NeY = d.NeY, NeZ = d.NeZ, n1 = e.n1, N = n1 - 1, Np = e.Np, Nfp = e.Nfp_v;
  const int Ne2D = NeX * NeY, Ne = m.Ne;
  const size_t NpNe = (size_t)Np * Ne;
  const double dx = (d.xmax - d.xmin) / NeX, dy = (d.ymax - d.ymin) / NeY;
  const double Ex = 2.0 / dx, Ey = 2.0 / dy;  // Escale(1,1), Escale(2,2) of the 2-D element
  const double Fx = 2.0 / dx, Fy = 2.0 / dy;  // Fscale of its x and y faces (sJ / J)
  const double* D = e.D1D.data();
  const double* iM = e.invM1D.data();
  // exterior topography value across a face: the neighbour's node (periodic
  // wrap) or, at a closed boundary, the node itself (the exchange's copy)
  auto topo_at = [&](int ii, int jj, int a) { return topo[a + (size_t)Nfp * (ii + NeX * jj)]; };
  std::vector<double> gx((size_t)Nfp * Ne2D), gy((size_t)Nfp * Ne2D);
  for (int jj = 0; jj < NeY; ++jj)
    for (int ii = 0; ii < NeX; ++ii) {
      const int ke2 = ii + NeX * jj;
      const int iw = ii > 0 ? ii - 1 : (d.periodic_x ? NeX - 1 : -1);
      const int ie = ii < NeX - 1 ? ii + 1 : (d.periodic_x ? 0 : -1);
      const int js = jj > 0 ? jj - 1 : (d.periodic_y ? NeY - 1 : -1);
      const int jn = jj < NeY - 1 ? jj + 1 : (d.periodic_y ? 0 : -1);
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n1; ++i) {
          double fx = 0.0, fy = 0.0;
          for (int l = 0; l < n1; ++l) {
            fx += D[i + n1 * l] * topo_at(ii, jj, l + n1 * j);
            fy += D[j + n1 * l] * topo_at(ii, jj, i + n1 * l);
          }
          // faces x+ (i = N) and x- (i = 0) along the node's row j, y+ and y- along column i
          const double tE = ie >= 0 ? topo_at(ie, jj, 0 + n1 * j) : topo_at(ii, jj, N + n1 * j);
          const double tW = iw >= 0 ? topo_at(iw, jj, N + n1 * j) : topo_at(ii, jj, 0 + n1 * j);
          const double tN = jn >= 0 ? topo_at(ii, jn, i + n1 * 0) : topo_at(ii, jj, i + n1 * N);
          const double tS = js >= 0 ? topo_at(ii, js, i + n1 * N) : topo_at(ii, jj, i + n1 * 0);
          const double dE = 0.5 * (tE - topo_at(ii, jj, N + n1 * j)) * (+1.0);
          const double dW = 0.5 * (tW - topo_at(ii, jj, 0 + n1 * j)) * (-1.0);
          const double dN = 0.5 * (tN - topo_at(ii, jj, i + n1 * N)) * (+1.0);
          const double dS = 0.5 * (tS - topo_at(ii, jj, i + n1 * 0)) * (-1.0);
          const double liftx = iM[i + n1 * N] * (Fx * dE) + iM[i + n1 * 0] * (Fx * dW);
          const double lifty = iM[j + n1 * N] * (Fy * dN) + iM[j + n1 * 0] * (Fy * dS);
          gx[(i + n1 * j) + (size_t)Nfp * ke2] = Ex * fx + liftx;
          gy[(i + n1 * j) + (size_t)Nfp * ke2] = Ey * fy + lifty;
        }
    }
  zlev.assign((size_t)Np * m.NeA, 0.0);
  const size_t NpNeA = (size_t)Np * m.NeA;
  for (int ke = 0; ke < Ne; ++ke) {
    const int ke2 = m.EMap3Dto2D[ke] - 1;
    for (int p = 0; p < Np; ++p) {
      const int h = p % Nfp;  // IndexH2Dto3D
      const size_t o = p + (size_t)Np * ke;
      const double zeta = m.pos_en[o + NpNe * 2];
      const double tp = topo[h + (size_t)Nfp * ke2];
      double coef = 1.0 - zeta / zTop;
      zlev[o] = zeta + coef * tp;
      const double GsqrtV = 1.0 - tp / zTop;
      coef = -coef / GsqrtV;
      const double G13 = coef * gx[h + (size_t)Nfp * ke2];
      const double G23 = coef * gy[h + (size_t)Nfp * ke2];
      m.Gsqrt[o] = GsqrtV;
      // GIJ = identity: tmp_G13 = G11 G13 + G12 G23, tmp_G23 = G21 G13 + G22 G23
      m.GI3[o] = 1.0 * G13 + 0.0 * G23;
      m.GI3[o + NpNeA] = 0.0 * G13 + 1.0 * G23;
    }
  }
  for (int b = 0; b < m.nB; ++b) {
    const size_t dst = NpNe + b, src = (size_t)(m.halo_src[b] - 1);
    zlev[dst] = zlev[src];
    m.Gsqrt[dst] = m.Gsqrt[src];
    m.GI3[dst] = m.GI3[src];
    m.GI3[dst + NpNeA] = m.GI3[src + NpNeA];
  }
  (void)NeZ;
  return 0;
}

}  // namespace fedg
