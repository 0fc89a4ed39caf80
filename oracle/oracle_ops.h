/* Internal to the CPU oracle (test infrastructure only): tensor-product
 * element operators shared by the HEVE and HEVI restatements, and the pow
 * selected by oracle_set_pow_mode. */
#ifndef ORACLE_OPS_H
#define ORACLE_OPS_H

double oracle_opow(double x, double y);

/* ---------------------------------------------------------------------- */
/* Tensor-product operators: ELM/scale_element_operation_tensorprod3D_kernel.F90
 * (P7 instance at :2226-2339, :2415-2480; the .erb template generates the
 * same sums for every p). */

/* matvec_dirX: out(i,jk) = sum_l Mat(i,l) * in(l,jk) */
static inline void op_dirX(int n1, const double *Mat, const double *in, double *out) {
  for (int jk = 0; jk < n1 * n1; ++jk)
    for (int i = 0; i < n1; ++i) {
      double s = Mat[i] * in[n1 * jk];
      for (int l = 1; l < n1; ++l) s = s + Mat[i + n1 * l] * in[l + n1 * jk];
      out[i + n1 * jk] = s;
    }
}
/* matvec_dirY: out(i,j,k) = sum_l in(i,l,k) * Mat_tr(l,j), Mat_tr(l,j) = Mat(j,l) */
static inline void op_dirY(int n1, const double *Mat, const double *in, double *out) {
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) {
        double s = in[i + n1 * (0 + n1 * k)] * Mat[j];
        for (int l = 1; l < n1; ++l) s = s + in[i + n1 * (l + n1 * k)] * Mat[j + n1 * l];
        out[i + n1 * (j + n1 * k)] = s;
      }
}
/* matvec_dirZ: out(i,j,k) = sum_l in(i,j,l) * Mat_tr(l,k); Mat_tr passed as stored */
static inline void op_dirZ_tr(int n1, const double *Mat_tr, const double *in, double *out) {
  const int n2 = n1 * n1;
  for (int k = 0; k < n1; ++k)
    for (int ij = 0; ij < n2; ++ij) {
      double s = in[ij] * Mat_tr[0 + n1 * k];
      for (int l = 1; l < n1; ++l) s = s + in[ij + n2 * l] * Mat_tr[l + n1 * k];
      out[ij + n2 * k] = s;
    }
}
/* dirZ with the derivative matrix: Mat_tr(l,k) = Mat(k,l) */
static inline void op_dirZ(int n1, const double *Mat, const double *in, double *out) {
  const int n2 = n1 * n1;
  for (int k = 0; k < n1; ++k)
    for (int ij = 0; ij < n2; ++ij) {
      double s = in[ij] * Mat[k];
      for (int l = 1; l < n1; ++l) s = s + in[ij + n2 * l] * Mat[k + n1 * l];
      out[ij + n2 * k] = s;
    }
}
/* matvec_Lift_hexahedral: six face terms per node, faces in reference order */
static inline void op_lift(int n1, const double *Lift, const double *vin, double *out) {
  const int n2 = n1 * n1, Np = n2 * n1;
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) {
        int p = i + n1 * (j + n1 * k);
        out[p] = Lift[p] * vin[i + n1 * k]
               + Lift[p + Np] * vin[j + n1 * k + n2]
               + Lift[p + 2 * Np] * vin[i + n1 * k + 2 * n2]
               + Lift[p + 3 * Np] * vin[j + n1 * k + 3 * n2]
               + Lift[p + 4 * Np] * vin[i + n1 * j + 4 * n2]
               + Lift[p + 5 * Np] * vin[i + n1 * j + 5 * n2];
      }
}

#endif
