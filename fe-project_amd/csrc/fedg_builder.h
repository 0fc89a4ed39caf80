// Host-side construction of the tensor-product LGL hexahedral element operators,
// the structured box mesh (node coordinates, metrics, face maps, halo layout),
// hydrostatic base states and warm-bubble initial states.
//
// Every array follows the reference's Fortran column-major layout with 0-based
// C offsets, e.g. a Fortran array X(Np,Ne) is stored as X[p + Np*ke], and integer
// maps (vmapM, vmapP, vmapB, EMap3Dto2D, IndexH2Dto3D*) keep the reference's
// 1-based values so that arrays coming from a Fortran host pass through unchanged.
#pragma once
#include <cstdint>
#include <vector>

namespace fedg {

struct RefElem {
  int p = 0;        // polynomial order (horizontal = vertical)
  int n1 = 0;       // nodes per direction = p+1
  int Np = 0;       // n1^3
  int Nfp_h = 0;    // n1*n1 (lateral faces)
  int Nfp_v = 0;    // n1*n1 (top/bottom faces)
  int NfpTot = 0;   // 4*Nfp_h + 2*Nfp_v
  std::vector<double> x1d, w1d;        // LGL nodes / weights
  std::vector<double> D1D;             // D1D(i,l) = d l_l/dx at x_i        (n1,n1)
  std::vector<double> V1D, invV1D;     // orthonormal Legendre Vandermonde  (n1,n1)
  std::vector<double> invM1D;          // V1D V1D^T                         (n1,n1)
  std::vector<double> Lift_mat;        // Lift_mat(i,j,k,f)                 (n1,n1,n1,6)
  std::vector<double> IntrpMat_VPOrdM1_tr;  // (n1,n1)
  std::vector<int32_t> Fmask_h, Fmask_v;     // 1-based (Nfp_h,4), (Nfp_v,2)
  std::vector<int32_t> IndexH2Dto3D;         // 1-based (Np)
  std::vector<int32_t> IndexH2Dto3D_bnd;     // 1-based (NfpTot)
};

void build_refelem(int porder, RefElem& e);

struct BoxDesc {
  int porder;
  int NeX, NeY, NeZ;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int periodic_x, periodic_y, periodic_z;
};

struct BoxMesh {
  int Ne = 0, NeA = 0, Ne2D = 0, Ne2DA = 0, nB = 0;
  std::vector<double> pos_en;      // (Np,Ne,3)
  std::vector<double> J;           // (Np,Ne)
  std::vector<double> Escale;      // (Np,Ne,3,3)
  std::vector<double> sJ;          // (NfpTot,Ne)
  std::vector<double> Fscale;      // (NfpTot,Ne)
  std::vector<double> normal_fn;   // (NfpTot,Ne,3)
  std::vector<int32_t> vmapM, vmapP;  // (NfpTot,Ne), 1-based into Np*NeA
  std::vector<int32_t> vmapB;         // (nB), 1-based into Np*Ne
  std::vector<int32_t> halo_src;      // (nB), 1-based interior node copied into halo slot b
  std::vector<int32_t> halo_face;     // (nB), domain boundary id 1..6 of halo slot b
  std::vector<int32_t> EMap3Dto2D;    // (Ne), 1-based
  std::vector<double> Gsqrt;          // (Np,NeA)
  std::vector<double> GI3;            // (Np,NeA,2)
  std::vector<double> GsqrtH;         // (Nfp_v,Ne2D)
};

int build_box(const BoxDesc& d, const RefElem& e, BoxMesh& m);

// Terrain-following vertical coordinate over a topography topo (Nfp_v, Ne2D)
// of the box's horizontal mesh and a model top zTop: GsqrtV, G13, G23 into
// m.Gsqrt and m.GI3 (halo slots filled as the single-tile exchange fills
// them) and zlev (Np, NeA), the height of every node.  m comes from build_box.
int set_topography(const BoxDesc& d, const RefElem& e, const double* topo, double zTop, BoxMesh& m,
                   std::vector<double>& zlev);

}  // namespace fedg
