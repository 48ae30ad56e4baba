/*
 * CPU oracle for the matrix-free spectral/hp operator path (TEST INFRASTRUCTURE).
 *
 * This is a plain-C restatement of the reference CPU algorithms, kept only as the
 * checker for the HIP product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links or calls it.
 *
 * Layout conventions (boundary layout of the reference Collections layer):
 *   coefficients  u[e*M + (r*nm + q)*nm + p]            (p fastest)
 *   quad points   f[e*Q + (k*nq + j)*nq + i]            (i fastest)
 *   jac           deformed: jac[e*Q + q]   regular: jac[e]
 *   df            deformed: df[d*(E*Q) + e*Q + q]  regular: df[d*E + e]
 *                 with d = c*dim + a holding d(xi_a)/d(x_c)
 * These follow Collections/CoalescedGeomData.cpp:53-112 (GetJac) and :251-313
 * (GetDerivFactors).
 */
#ifndef ITHACA_MF_ORACLE_H
#define ITHACA_MF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Polylib (LibUtilities/Polylib/Polylib.cpp) ---- */
void or_jacobfd(int np, const double *z, double *poly, double *polyd, int n,
                double alpha, double beta);
void or_zwglj(double *z, double *w, int np);
void or_Dglj(double *D, const double *z, int np);

/* ---- Basis (LibUtilities/Foundations/Basis.cpp:392-421, eModified_A on GLL) ---- */
void or_basis_modified_a(int nm, int nq, double *z, double *w, double *D,
                         double *bdata, double *dbdata);

/* ---- Geometry (SpatialDomains/GeomFactors.cpp:399-478) ---- */
/* verts: E x 8 x 3 (hex) / E x 4 x 2 (quad), vertex order of HexGeom/QuadGeom.
 * deformed != 0 -> per-point factors, else the value at the first point. */
void or_geom_hex(const double *verts, int E, int nq, const double *z, int deformed,
                 double *jac, double *df);
void or_geom_quad(const double *verts, int E, int nq, const double *z, int deformed,
                  double *jac, double *df);
/* physical coordinates of the quadrature points: xyz[c*(E*Q) + e*Q + q] */
void or_coords_hex(const double *verts, int E, int nq, const double *z, double *xyz);
void or_coords_quad(const double *verts, int E, int nq, const double *z, double *xy);

/* ---- MatrixFree operators (width-1 element interleave, reference loop order) ---- */
void or_set_threads(int nthreads);

void or_bwdtrans_hex(int nm, int nq, const double *bdata, int E,
                     const double *in, double *out);
void or_iproduct_hex(int nm, int nq, const double *bdata, const double *w, int E,
                     int deformed, const double *jac, const double *in, double *out);
void or_physderiv_hex(int nq, const double *D, int E, int deformed, const double *df,
                      const double *in, double *out0, double *out1, double *out2);
void or_iprod_deriv_hex(int nm, int nq, const double *bdata, const double *dbdata,
                        const double *w, int E, int deformed, const double *jac,
                        const double *df, const double *in0, const double *in1,
                        const double *in2, double *out);
void or_helmholtz_hex(int nm, int nq, const double *bdata, const double *dbdata,
                      const double *D, const double *w, int E, int deformed,
                      const double *jac, const double *df, double lambda,
                      const double *in, double *out);

void or_bwdtrans_quad(int nm, int nq, const double *bdata, int E,
                      const double *in, double *out);
void or_iproduct_quad(int nm, int nq, const double *bdata, const double *w, int E,
                      int deformed, const double *jac, const double *in, double *out);
void or_physderiv_quad(int nq, const double *D, int E, int deformed, const double *df,
                       const double *in, double *out0, double *out1);
void or_iprod_deriv_quad(int nm, int nq, const double *bdata, const double *dbdata,
                         const double *w, int E, int deformed, const double *jac,
                         const double *df, const double *in0, const double *in1,
                         double *out);
/* The same with the AVX2 SIMD ABI: four elements interleaved per vector (mf_oracle_avx2.c). */
void or_helmholtz_hex_avx2(int nm, int nq, const double *bdata, const double *dbdata,
                           const double *D, const double *w, int E, int deformed,
                           const double *jac, const double *df, double lambda,
                           const double *in, double *out);
void or_helmholtz_quad(int nm, int nq, const double *bdata, const double *dbdata,
                       const double *D, const double *w, int E, int deformed,
                       const double *jac, const double *df, double lambda,
                       const double *in, double *out);

/* ---- Triangles (mf_oracle_tri.c): Modified_A x Modified_B on GLL x Gauss-Radau ---- */
void or_jacobz(int n, double *z, double alpha, double beta);
void or_deriv_factors(int dim, const double *deriv, double *jac, double *fac);
void or_zwgrjm(double *z, double *w, int np, double alpha, double beta);
void or_Dnodes(double *D, const double *z, int np);
void or_basis_modified_b(int nm, int nq, double *z, double *w, double *D, double *bdata,
                         double *dbdata);
/* the Modified_B values on the nq points z (nm(nm+1)/2 rows of nq) */
void or_modified_b_tables(int nm, int nq, const double *z, double *bdata);
void or_geom_tri(const double *verts, int E, int nq0, int nq1, int deformed, double *jac,
                 double *df);
void or_coords_tri(const double *verts, int E, int nq0, int nq1, const double *z0,
                   const double *z1, double *xy);
void or_bwdtrans_tri(int nm, int nq0, int nq1, const double *b0, const double *b1, int E,
                     const double *in, double *out);
void or_iproduct_tri(int nm, int nq0, int nq1, const double *b0, const double *b1,
                     const double *w0, const double *w1, int E, int deformed,
                     const double *jac, const double *in, double *out);
void or_physderiv_tri(int nq0, int nq1, const double *D0, const double *D1,
                      const double *z0, const double *z1, int E, int deformed,
                      const double *df, const double *in, double *out0, double *out1);
void or_helmholtz_tri(int nm, int nq0, int nq1, const double *b0, const double *db0,
                      const double *b1, const double *db1, const double *D0,
                      const double *D1, const double *z0, const double *z1,
                      const double *w0, const double *w1, int E, int deformed,
                      const double *jac, const double *df, double lambda,
                      const double *in, double *out);

/* ---- Prisms (mf_oracle_prism.c): Modified_A x Modified_A x Modified_B ---- */
void or_geom_prism(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                   const double *z1, const double *z2, int deformed, double *jac, double *df);
void or_coords_prism(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                     const double *z1, const double *z2, double *xyz);
void or_bwdtrans_prism(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                       const double *b2, int E, const double *in, double *out);
void or_iproduct_prism(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                       const double *b2, const double *w0, const double *w1, const double *w2,
                       int E, int deformed, const double *jac, const double *in, double *out);
void or_physderiv_prism(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                        const double *D2, const double *z0, const double *z2, int E,
                        int deformed, const double *df, const double *in, double *out0,
                        double *out1, double *out2);
void or_helmholtz_prism(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                        const double *b1, const double *db1, const double *b2,
                        const double *db2, const double *D0, const double *D1,
                        const double *D2, const double *z0, const double *z2,
                        const double *w0, const double *w1, const double *w2, int E,
                        int deformed, const double *jac, const double *df, double lambda,
                        const double *in, double *out);

/* reference-space derivatives of one element's nq0 x nq1 x nq2 point values along the three
 * directions (PhysDerivTensor3DKernel order) */
void or_deriv3_elmt(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                    const double *D2, const double *in, double *d0, double *d1, double *d2);

/* ---- Tetrahedra (mf_oracle_tet.c): Modified_A x Modified_B x Modified_C ---- */
void or_basis_modified_c(int nm, int nq, double *z, double *w, double *D, double *bdata,
                         double *dbdata);
void or_geom_tet(const double *verts, int E, int nq0, int nq1, int nq2, int deformed,
                 double *jac, double *df);
void or_coords_tet(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                   const double *z1, const double *z2, double *xyz);
void or_bwdtrans_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, int E, const double *in, double *out);
void or_iproduct_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, const double *w0, const double *w1, const double *w2,
                     int E, int deformed, const double *jac, const double *in, double *out);
void or_physderiv_tet(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                      const double *D2, const double *z0, const double *z1, const double *z2,
                      int E, int deformed, const double *df, const double *in, double *out0,
                      double *out1, double *out2);
void or_helmholtz_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                      const double *b1, const double *db1, const double *b2, const double *db2,
                      const double *D0, const double *D1, const double *D2, const double *z0,
                      const double *z1, const double *z2, const double *w0, const double *w1,
                      const double *w2, int E, int deformed, const double *jac,
                      const double *df, double lambda, const double *in, double *out);

/* ---- Pyramids (mf_oracle_pyr.c): Modified_A x Modified_A x ModifiedPyr_C ---- */
void or_basis_modified_pyr_c(int nm, int nq, double *z, double *w, double *D, double *bdata,
                             double *dbdata);
void or_geom_pyr(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                 const double *z1, const double *z2, int deformed, double *jac, double *df);
void or_coords_pyr(const double *verts, int E, int nq0, int nq1, int nq2, const double *z0,
                   const double *z1, const double *z2, double *xyz);
void or_bwdtrans_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, int E, const double *in, double *out);
void or_iproduct_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *b1,
                     const double *b2, const double *w0, const double *w1, const double *w2,
                     int E, int deformed, const double *jac, const double *in, double *out);
void or_physderiv_pyr(int nq0, int nq1, int nq2, const double *D0, const double *D1,
                      const double *D2, const double *z0, const double *z1, const double *z2,
                      int E, int deformed, const double *df, const double *in, double *out0,
                      double *out1, double *out2);
void or_helmholtz_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                      const double *b1, const double *db1, const double *b2, const double *db2,
                      const double *D0, const double *D1, const double *D2, const double *z0,
                      const double *z1, const double *z2, const double *w0, const double *w1,
                      const double *w2, int E, int deformed, const double *jac,
                      const double *df, double lambda, const double *in, double *out);

/* IProductWRTDerivBase of the collapsed shapes: out = sum_c (d phi / d x_c, in_c) */
void or_iprodderiv_tri(int nm, int nq0, int nq1, const double *b0, const double *db0,
                       const double *b1, const double *db1, const double *z0, const double *z1,
                       const double *w0, const double *w1, int E, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, double *out);
void or_iprodderiv_prism(int nm, int nq0, int nq1, int nq2, const double *b0,
                         const double *db0, const double *b1, const double *db1,
                         const double *b2, const double *db2, const double *z0,
                         const double *z2, const double *w0, const double *w1, const double *w2,
                         int E, int deformed, const double *jac, const double *df,
                         const double *in0, const double *in1, const double *in2, double *out);
void or_iprodderiv_tet(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                       const double *b1, const double *db1, const double *b2, const double *db2,
                       const double *z0, const double *z1, const double *z2, const double *w0,
                       const double *w1, const double *w2, int E, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out);
void or_iprodderiv_pyr(int nm, int nq0, int nq1, int nq2, const double *b0, const double *db0,
                       const double *b1, const double *db1, const double *b2, const double *db2,
                       const double *z0, const double *z1, const double *z2, const double *w0,
                       const double *w1, const double *w2, int E, int deformed,
                       const double *jac, const double *df, const double *in0,
                       const double *in1, const double *in2, double *out);

void or_geom_seg(const double *verts, int E, int ncoord, int nq, int deformed, double *jac,
                 double *df);
void or_bwdtrans_seg(int nm, int nq, const double *b, int E, const double *in, double *out);
void or_iproduct_seg(int nm, int nq, const double *b, const double *w, int E, int deformed,
                     const double *jac, const double *in, double *out);
void or_physderiv_seg(int nq, const double *D, int E, int deformed, int ncoord,
                      const double *df, const double *in, double *out);
void or_iprodderiv_seg(int nm, int nq, const double *db, const double *w, int E, int deformed,
                       int ncoord, const double *jac, const double *df, const double *in,
                       double *out);

#ifdef __cplusplus
}
#endif
#endif
