/*
 * ithaca_mf.h -- C ABI of the MI355X (gfx950) matrix-free spectral/hp operators.
 *
 * Drop-in boundary for the reference's MatrixFree operator layer.  Each entry point
 * names the reference interface it replaces (paths relative to the reference root).
 * Plain C types only: no C++, HIP or torch types cross this boundary; streams are
 * passed as `void *` (a hipStream_t, NULL = the default stream).
 *
 * Layouts (un-interleaved, element-major; Collections/CoalescedGeomData.cpp:53-112,
 * :251-313):
 *   coefficients  u[e*M + (r*nm + q)*nm + p]          M = nm^dim
 *   quad points   f[e*Q + (k*nq + j)*nq + i]          Q = nq^dim
 *   jac           deformed: jac[e*Q + q]             regular: jac[e]
 *   df            deformed: df[d*(nElmt*Q) + e*Q + q] regular: df[d*nElmt + e]
 *                 d = c*dim + a holds d(xi_a)/d(x_c)   (dim*dim components)
 * Status codes: 0 = OK; non-zero = error, message via imf_last_error().
 */
#ifndef ITHACA_MF_H
#define ITHACA_MF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct imf_op_s *imf_op;

enum imf_status {
    IMF_OK = 0,
    IMF_ERR_ARG = 1,         /* bad argument / size mismatch */
    IMF_ERR_UNSUPPORTED = 2, /* unknown key or (modes, points) combination */
    IMF_ERR_HIP = 3,         /* HIP runtime failure */
    IMF_ERR_STATE = 4        /* missing geometry, wrong precision, ... */
};

enum imf_precision { IMF_FP64 = 0, IMF_FP32 = 1 };

/* Replaces MatrixFree::GetOperatorFactory().CreateInstance(key, basis, nElmt)
 * (MatrixFreeOps/Operator.hpp:303-310, keys from Operator.cpp:12-56 such as
 * "Helmholtz_Hex_Deformed", "IProduct_Quad_Regular", "BwdTrans_Hex_Regular") together
 * with the Helper<DIM,DEFORMED> constructor that copies the basis tables
 * (Operator.hpp:205-275).  Per direction d < dim: bdata[d] (nm*nq, bdata[p*nq+i]),
 * dbdata[d] (nm*nq), D[d] (nq*nq, D[a*nq+b] = l_a'(z_b)), w[d] (nq).  Like the
 * reference kernels the isotropic operators support nm in 2..8 with nq in nm..2*nm
 * (Helmholtz.h:669-760).  No padding restriction on nElmt (the reference's vec_t::width
 * padding is handled internally). */
/* Also accepted, as in the reference's MatrixFree factory:
 *   "PhysDeriv_{Hex,Quad}_{Regular,Deformed}" (MatrixFreeOps/PhysDeriv.cpp:14-36;
 *     PhysDeriv.h:260-642): in = nElmt*Q point values, out = dim consecutive arrays of
 *     nElmt*Q (the reference's out[0..dim-1]), out_c = sum_a df[c*dim+a] d/dxi_a in;
 *     needs SetDF only; any nq in 2..16 (nm is not restricted).
 *   "IProductWRTDerivBase_{Hex,Quad}_{Regular,Deformed}" (IProductWRTDerivBaseHex.cpp,
 *     IProductWRTDerivBaseQuad.cpp; IProductWRTDerivBase.h:371-, :1068-1343): in = dim
 *     consecutive arrays of nElmt*Q (the reference's in[0..dim-1]), out = nElmt*M;
 *     needs SetJac and SetDF.  Quads take 2D coordinates (dim*dim = 4 factors).
 * Neither has a transpose (imf_apply_transpose_device reports IMF_ERR_UNSUPPORTED).
 * Different modes, points or tables per direction (hexahedra, FP64; BwdTrans, IProduct,
 * Helmholtz and Mass): the reference's MatrixFree kernels are isotropic and use direction
 * 0's tables for all three (Helmholtz.h:844-856); such expansions run there through the
 * StdRegions matrix path.  Here they get runtime-size kernels (2 <= nm[d] <= nq[d] <= 16),
 * reading SetJac / SetDF arrays over nq[0] nq[1] nq[2] points per element; the coefficient
 * layout is (r*nm[1] + q)*nm[0] + p and the point layout (k*nq[1] + j)*nq[0] + i.
 * Triangles, "{BwdTrans,IProduct,Helmholtz,PhysDeriv}_Tri_{Regular,Deformed}" (FP64;
 * BwdTrans.h:333-470, IProductTri.cpp, Helmholtz.h:288-620, PhysDeriv.h; kernels
 * BwdTransKernels.hpp:77-126, IProductKernels.hpp:134-236, PhysDerivKernels.hpp:150-216):
 * dim = 2, nm[0] = nm[1] modes (nm(nm+1)/2 coefficients, p-major with q fastest), nq[0]
 * GLL points along eta0 and nq[1] Gauss-Radau (alpha = 1) points along eta1 (the
 * reference requires nq[1] = nq[0] - 1; here 2 <= nm <= 8, nm <= nq[0] <= 16,
 * nm - 1 <= nq[1] <= 16).  Direction 0 takes the Modified_A tables of
 * imf_basis_modified_a, direction 1 the Modified_B tables of imf_basis_modified_b (w[1]
 * the Gauss-Radau weights; the factor 1/2 of the collapse is applied here, as the
 * reference's Helper does).  Points are ordered j*nq[0] + i (eta0 fastest); SetJac /
 * SetDF take one value (4 factors) per element for _Regular, per point for _Deformed.
 * Transposes: BwdTrans and IProduct (Euclidean), Helmholtz (symmetric).
 * Prisms, "{BwdTrans,IProduct,Helmholtz,PhysDeriv}_Prism_{Regular,Deformed}" (FP64;
 * BwdTrans.cpp:31-33, IProductPrism.cpp, Helmholtz.cpp:25-29, PhysDeriv.cpp:38-42; kernels
 * BwdTransKernels.hpp:224-300, IProductKernels.hpp:315-449, PhysDerivKernels.hpp:219-294,
 * Helmholtz.h:1291-1460): dim = 3, nm[0] = nm[1] = nm[2] modes (nm*nm(nm+1)/2 coefficients,
 * mode (p, q, r) with r < nm - p at nm*sum_{p'<p}(nm - p') + q(nm - p) + r), nq[0], nq[1]
 * GLL points along eta0, xi1 and nq[2] Gauss-Radau (alpha = 1) points along eta2 (the
 * reference requires nq[2] = nq[0] - 1; here 2 <= nm <= 8, nm <= nq[0], nq[1] <= 16,
 * nm - 1 <= nq[2] <= 16, and one element's arrays must fit the LDS of a workgroup).
 * Directions 0 and 1 take Modified_A tables, direction 2 the Modified_B tables of
 * imf_basis_modified_b (w[2] without the collapse factor 1/2, applied here).  Points are
 * ordered (k*nq[1] + j)*nq[0] + i; SetJac / SetDF take one value (9 factors) per element
 * for _Regular, per point for _Deformed; transposes as for triangles.
 * Tetrahedra, "{BwdTrans,IProduct,Helmholtz,PhysDeriv}_Tet_{Regular,Deformed}" (FP64;
 * BwdTrans.cpp, IProductTet.cpp, Helmholtz.cpp, PhysDeriv.cpp; kernels
 * BwdTransKernels.hpp:377-484, IProductKernels.hpp:604-760, PhysDerivKernels.hpp:556-694,
 * Helmholtz.h:2266-2455): dim = 3, nm[0] = nm[1] = nm[2] modes (nm(nm+1)(nm+2)/6
 * coefficients, mode (p, q, r) with p + q + r < nm, r fastest, then q, then p), nq[0] GLL
 * points along eta0, nq[1] Gauss-Radau alpha = 1 along eta1 and nq[2] Gauss-Radau
 * alpha = 2 along eta2 (the reference requires nq[1] = nq[2] = nq[0] - 1; here
 * 2 <= nm <= 8, nm <= nq[0] <= 16, nm - 1 <= nq[1], nq[2] <= 16, within one workgroup's
 * LDS).  Tables: Modified_A (imf_basis_modified_a), Modified_B (imf_basis_modified_b)
 * and Modified_C (imf_basis_modified_c), the raw Gauss-Radau weights as w[1] and w[2]
 * (the factors 1/2 and 1/4 are applied here).  Points (k*nq[1] + j)*nq[0] + i; SetJac /
 * SetDF one value (9 factors) per element for _Regular, per point for _Deformed.
 * Pyramids, "{BwdTrans,IProduct,Helmholtz,PhysDeriv}_Pyr_{Regular,Deformed}" (FP64;
 * BwdTrans.cpp, IProductPyr.cpp, Helmholtz.cpp:31-35, PhysDeriv.cpp; kernels
 * BwdTransKernels.hpp:131-222, IProductKernels.hpp:455-598, PhysDerivKernels.hpp:464-550,
 * Helmholtz.h:1771-1935): dim = 3, nm[0] = nm[1] = nm[2] modes (nm(nm+1)(2nm+1)/6
 * coefficients, mode (p, q, r) with r < nm - max(p, q), r fastest, then q, then p), nq[0],
 * nq[1] GLL points along eta0, eta1 and nq[2] Gauss-Radau alpha = 2 along eta2 (the
 * reference requires nq[2] = nq[0] - 1; here 2 <= nm <= 8, nm <= nq[0], nq[1] <= 16,
 * nm - 1 <= nq[2] <= 16, within one workgroup's LDS).  Tables: Modified_A for directions
 * 0 and 1, ModifiedPyr_C (imf_basis_modified_pyr_c) for direction 2 with its raw
 * Gauss-Radau weights as w[2] (the factor 1/4 is applied here).  Layouts as for prisms.
 * Segments, "BwdTrans_Seg_Regular", "{IProduct,PhysDeriv,IProductWRTDerivBase}_Seg_{Regular,
 * Deformed}" (FP64; BwdTrans.cpp, IProductSeg.cpp, PhysDeriv.cpp, IProductWRTDerivBaseSeg.cpp;
 * BwdTrans.h:16-, IProduct.h:17-, PhysDeriv.h:17-, IProductWRTDerivBase.h:18-): one Modified_A
 * direction of nm[0] modes on nq[0] GLL points (1 <= nm <= 16, nm <= nq <= 32; PhysDeriv any
 * nq in 1..32) with tables bdata[0], dbdata[0] (IProductWRTDerivBase), D[0], w[0]; dim is the
 * coordinate dimension (1 to 3) of the mesh the segments lie in: SetDF takes dim factors per
 * element (_Regular) or per point (_Deformed), dim-major, PhysDeriv's out and
 * IProductWRTDerivBase's in are dim arrays of nElmt*nq back to back.  No segment Helmholtz
 * (the reference registers none).  Transposes: BwdTrans and IProduct.
 * "IProductWRTDerivBase_{Tri,Prism,Tet,Pyr}_{Regular,Deformed}" (IProductWRTDerivBase
 * {Tri,Prism,Tet,Pyr}.cpp; IProductWRTDerivBase.h:671-, :1630-, :2056-, :2484-): the
 * same sizes, tables (dbdata required) and layouts as the shape's other operators; in =
 * dim consecutive arrays of nElmt*Q, out = nElmt*M, out = sum_a (D_a B)^T W J t_a with
 * t_a = sum_c df[c*dim+a] in_c pushed through the transpose of the collapsed chain rule;
 * needs SetJac and SetDF; no transpose.
 * Also accepted, with no MatrixFreeOps counterpart: "Mass_Hex_Deformed" and
 * "Mass_Hex_Regular", the elemental mass action out = B^T diag(w J) B in of
 * StdRegions/StdExpansion.cpp:647-661 (MassMatrixOp_MatFree: BwdTrans, then
 * IProductWRTBase) fused into one kernel; needs SetJac only; in and out are
 * coefficient arrays (nElmt*M). */
int imf_create(const char *key, int dim, const int *nm, const int *nq,
               const double *const *bdata, const double *const *dbdata,
               const double *const *D, const double *const *w, int nElmt,
               int precision, imf_op *op);

/* Releases device memory held by the operator. */
int imf_destroy(imf_op op);

/* Operator::Ndof (Operator.hpp:29-32; BwdTrans.h:190-193): nm^dim * nElmt.
 * Defined for every operator type, including Helmholtz. */
double imf_ndof(imf_op op);

/* Operator::NeedsJac / NeedsDF (Operator.hpp:35-45, 88-91, 163-171). */
int imf_needs_jac(imf_op op);
int imf_needs_df(imf_op op);

/* Number of input / output scalars per apply (nElmt * M or nElmt * Q). */
size_t imf_in_size(imf_op op);
size_t imf_out_size(imf_op op);

/* Helper::SetJac / SetDF (Operator.hpp:277-291), in the un-interleaved GetJac /
 * GetDerivFactors layouts above.  n is the number of doubles; the data are copied
 * to device memory (host pointers) or device-to-device (_device variants). */
int imf_set_jac(imf_op op, const double *jac, size_t n);
int imf_set_df(imf_op op, const double *df, size_t n);
int imf_set_jac_device(imf_op op, const double *d_jac, size_t n);
int imf_set_df_device(imf_op op, const double *d_df, size_t n);

/* Helmholtz::SetLambda (Operator.hpp:193-196). */
int imf_set_lambda(imf_op op, double lambda);

/* IProductWRTDerivBase of triangles and quadrilaterals in 3D space (a surface mesh), FP64:
 * ncoord = 3 fields in, and SetDF then takes 6 factors per point (regular: per element),
 * t_a = sum_c df[2c+a] in_c, as MatrixFreeOps/IProductWRTDerivBase.h:545- (quads) and :906-
 * (triangles) do for three inputs.  Call before SetDF; ncoord = 2 restores the default.
 * Other operators and shapes: IMF_ERR_UNSUPPORTED. */
int imf_set_coord_dim(imf_op op, int ncoord);

/* operator()(in, out) (Operator.hpp:70-72, 99-101, 176-178) on host arrays: copies in,
 * applies, copies out; synchronous.  out is overwritten.  FP64 operators only.
 * BwdTrans, IProduct, Helmholtz and Mass on quads and hexes run in sixteen element chunks
 * through pinned staging buffers (allocated on the first call, kept with the operator),
 * so the host copies, the transfers each way and the kernels of neighbouring chunks
 * overlap; the result is bitwise that of imf_apply_device. */
int imf_apply(imf_op op, const double *in, double *out);

/* operator()(in, out) on device arrays of the operator's precision (double for
 * IMF_FP64, float for IMF_FP32), asynchronous on `stream`.  Allocates nothing. */
int imf_apply_device(imf_op op, const void *d_in, void *d_out, void *stream);

/* Transposed (adjoint) apply under the Euclidean inner product, device arrays, no
 * reference counterpart (used by the adjoint/autograd layer):
 *   BwdTrans^T:  f (nElmt*Q) -> B^T f            (IProduct with unit weights, J = 1)
 *   IProduct^T:  g (nElmt*M) -> diag(w J) B g    (BwdTrans followed by w*J scaling)
 *   Helmholtz^T, Mass^T: identical to the operator (both are symmetric). */
int imf_apply_transpose_device(imf_op op, const void *d_in, void *d_out, void *stream);

/* operator()(in_v, out_v) for v < nvec vectors at d_in + v*ld_in and d_out + v*ld_out
 * (ld in scalars of the operator's precision, ld_in >= imf_in_size, ld_out >=
 * imf_out_size), device arrays, asynchronous, allocates nothing.  No reference
 * counterpart (the reference applies one field at a time); the Mass operator does all
 * vectors in one launch, reading the geometry once; other operators loop over vectors. */
int imf_apply_many_device(imf_op op, int nvec, const void *d_in, long ld_in, void *d_out,
                          long ld_out, void *stream);

/* Message of the last failing call on this thread ("" if none). */
const char *imf_last_error(void);

/* ---- C0 global assembly and the matrix-free Helmholtz solve (hexahedra, quads) -------
 * Replaces AssemblyMapCG's local-to-global map and signs (MultiRegions/AssemblyMap/
 * AssemblyMapCG.cpp:1332-1900; edge and face signs at :1786, :1832).  Its global order
 * (graph-based, bandwidth-reducing) differs; the map's contract is the same: continuity
 * across elements, with the signs of the modes' orientation.
 * Global numbering of the modified (C0) basis built from the elements' vertex ids
 * (vertex_ids[e*8 + v], any 64-bit ids, vertex order of the reference hex): vertex modes
 * shared by id, edge-interior modes by the vertex pair with sign (-1)^m against the global
 * edge direction, face-interior modes by the 4 vertex ids in a frame fixed by the ids
 * (sign per reversed axis, index swap for swapped axes).  Local arrays are element-major
 * (e*nm^3 + (r*nm+q)*nm + p), global arrays have imf_assembly_nglobal entries. */
typedef struct imf_asm_s *imf_asm;

/* Host only (no device): the map itself.  l2g, sign, on_boundary (1 if the local mode lies
 * on a face owned by one element) each nElmt*nm^3 entries or NULL; *nglobal out. */
int imf_assembly_map_hex(int nElmt, int nm, const long long *vertex_ids, int *l2g,
                         signed char *sign, unsigned char *on_boundary, long *nglobal);

/* The same for quadrilaterals: vertex_ids[e*4 + v] counter-clockwise from (-1,-1) (the
 * reference quad's vertex order), local arrays e*nm^2 + q*nm + p, edge-interior modes
 * signed (-1)^m against the global edge direction, on_boundary = on an edge owned by one
 * element. */
int imf_assembly_map_quad(int nElmt, int nm, const long long *vertex_ids, int *l2g,
                          signed char *sign, unsigned char *on_boundary, long *nglobal);

/* Device assembly map.  dirichlet = 1: homogeneous Dirichlet conditions on every mode of
 * the boundary faces (those modes are excluded from the solve and set to zero);
 * dirichlet = 0: natural (Neumann) conditions. */
int imf_assembly_create_hex(int nElmt, int nm, const long long *vertex_ids, int dirichlet,
                            imf_asm *assembly);
int imf_assembly_create_quad(int nElmt, int nm, const long long *vertex_ids, int dirichlet,
                             imf_asm *assembly);

/* Variable order (dim = 3: hexahedra, dim = 2: quadrilaterals): element e has nm[e] modes
 * per direction (nm[e] >= 2) and its local coefficients start at sum_{f<e} nm[f]^dim (the
 * reference's ExpList coefficient offsets).  As AssemblyMapCG.cpp:1380-1440 and :1771-1791
 * do for modal expansions of different order, a shared edge carries the fewest interior
 * modes of the elements around it and a shared face the fewest of its two elements; an
 * element's modes beyond those get l2g = 0 and sign = 0, so they take no part in the global
 * system (GlobalToLocal sets them to zero, Assemble skips them, on_boundary is 0 for them).
 * With every nm[e] equal these are the maps of the uniform calls above. */
int imf_assembly_map_varp(int dim, int nElmt, const int *nm, const long long *vertex_ids,
                          int *l2g, signed char *sign, unsigned char *on_boundary, long *nglobal);
int imf_assembly_create_varp(int dim, int nElmt, const int *nm, const long long *vertex_ids,
                             int dirichlet, imf_asm *assembly);

/* Hexahedra with different modes per direction (expansions such as NUMMODES="4,6,8",
 * Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_Heterogeneous.tst): element e has
 * nm3[3e + d] modes along its local direction d and nm3[3e] nm3[3e+1] nm3[3e+2] local
 * coefficients (p fastest).  An edge along direction d carries nm3[3e + d] - 2 interior
 * modes, a face the interior modes of its two in-plane directions, each counted in the
 * face's frame fixed by the vertex ids; where elements disagree the fewest are kept, as in
 * the variable-order maps above (sign 0 beyond them). */
int imf_assembly_map_aniso(int nElmt, const int *nm3, const long long *vertex_ids, int *l2g,
                           signed char *sign, unsigned char *on_boundary, long *nglobal);
int imf_assembly_create_aniso(int nElmt, const int *nm3, const long long *vertex_ids,
                              int dirichlet, imf_asm *assembly);
long imf_assembly_nglobal(imf_asm assembly);

/* Boundary conditions that differ by region (ContField with a mix of Dirichlet, Neumann and
 * Robin regions, e.g. Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_AllBCs.tst):
 * imf_assembly_set_dirichlet_mask replaces the Dirichlet set by mask[g] != 0 over the
 * nglobal modes (the closure of the Dirichlet faces; NULL: none); those modes are excluded
 * from the solve as the create calls' dirichlet = 1 excludes every boundary mode.
 * imf_assembly_set_robin gives the Robin trace terms alpha * (trace mass) as a sparse
 * matrix R over the global modes (COO, nnz entries, repeated entries summed; nnz = 0
 * removes it); the solve's operator becomes mask0(Assemble(H(GlobalToLocal(x))) + R x), as
 * GlobalLinSysIterativeFull::v_DoMatrixMultiply adds AddRobinTraceContribution per element
 * before the assembly (GlobalLinSysIterativeFull.cpp:232-252), and diag(R) joins the Jacobi
 * preconditioner.  Neumann and Robin right-hand-side integrals are the caller's (host-side
 * set-up, as the Dirichlet values).  Both calls are synchronous and copy their arrays. */
int imf_assembly_set_dirichlet_mask(imf_asm assembly, const unsigned char *mask);
int imf_assembly_set_robin(imf_asm assembly, long nnz, const int *rows, const int *cols,
                           const double *vals);
int imf_assembly_destroy(imf_asm assembly);

/* AssemblyMapCG::GlobalToLocal (MultiRegions/AssemblyMap/AssemblyMapCG.cpp:2853-2876):
 * local = sign * global[l2g].  Device arrays, asynchronous. */
int imf_global_to_local_device(imf_asm assembly, const double *d_global, double *d_local,
                               void *stream);

/* AssemblyMapCG::Assemble (AssemblyMapCG.cpp:2885-2910): global = sum of sign * local over
 * each global mode's local copies (overwrites global; deterministic order).  Asynchronous. */
int imf_assemble_device(imf_asm assembly, const double *d_local, double *d_global, void *stream);

/* ContField::HelmSolve's iterative path with a matrix-free operator: conjugate gradients
 * of NekLinSysIterCG::DoConjugateGradient (LibUtilities/LinearAlgebra/NekLinSysIterCG.cpp:
 * 104-260, one fused reduction per iteration) on y = Assemble(H(GlobalToLocal(x)))
 * (GlobalLinSysIterativeFull::v_DoMatrixMultiply, GlobalLinSysIterativeFull.cpp:215-232)
 * with the assembled diagonal as preconditioner; stops when (r,r) < tol^2 (b,b)
 * (Set_Rhs_Magnitude, NekLinSysIter.cpp:128-153; the reference's default tol is 1e-9).
 * helmholtz: an FP64 "Helmholtz_Hex_*" (or, for a quad assembly, "Helmholtz_Quad_*")
 * operator on the same elements and modes, with
 * SetJac/SetDF/SetLambda done.  d_rhs, d_x: global arrays.  Synchronous (reads the three
 * dot products every iteration); allocates its workspace on the first call.  Returns
 * IMF_ERR_STATE with *iterations = max_iter + 1 when the tolerance is not reached. */
int imf_helmsolve_cg_device(imf_asm assembly, imf_op helmholtz, const double *d_rhs,
                            double *d_x, double tol, int max_iter, int *iterations,
                            double *residual, void *stream);

/* The same solve with one operator per run of equal-order elements, as the reference starts
 * a new Collection where the element's coefficient or point count changes
 * (ExpList::CreateCollections, MultiRegions/ExpList.cpp:5005-5110): helmholtz[g] acts on the
 * next helmholtz[g]'s nElmt elements, in element order, and must have their modes; together
 * they cover the assembly's elements. */
int imf_helmsolve_cg_groups_device(imf_asm assembly, int nops, const imf_op *helmholtz,
                                   const double *d_rhs, double *d_x, double tol, int max_iter,
                                   int *iterations, double *residual, void *stream);

/* Message of the last failing assembly call on this thread. */
const char *imf_assembly_last_error(void);

/* ---- set-up helpers (no reference FFI counterpart; host-side restatements) ---- */

/* LibUtilities::Basis::GenBasis for eModified_A on nq Gauss-Lobatto-Legendre points
 * (Foundations/Basis.cpp:392-421, Polylib.cpp zwglj/Dglj).  Outputs z[nq], w[nq],
 * D[nq*nq], bdata[nm*nq], dbdata[nm*nq]. */
int imf_basis_modified_a(int nm, int nq, double *z, double *w, double *D,
                         double *bdata, double *dbdata);
/* LibUtilities::Basis::GenBasis for eModified_B on nq Gauss-Radau-Jacobi points with
 * z = -1 and weight (1 - z) (eGaussRadauMAlpha1Beta0; Basis.cpp:423-509, Polylib.cpp
 * zwgrjm): the collapsed direction of a triangle.  Outputs z[nq], w[nq] (the quadrature
 * weights of the points, without the factor 1/2 of the collapse), D[nq*nq],
 * bdata[nm(nm+1)/2 * nq] and dbdata (modes packed p-major, q fastest). */
int imf_basis_modified_b(int nm, int nq, double *z, double *w, double *D,
                         double *bdata, double *dbdata);
/* Modified_C on nq Gauss-Radau (alpha = 2, beta = 0) points (Basis.cpp:513-563,
 * GaussPoints.cpp:123-125): phi^c_{pqr} = phi^b_{p+q,r}, the collapsed eta2 direction of a
 * tetrahedron.  z, w (the Gauss-Radau weights, without the collapse factor 1/4), D (nq*nq),
 * bdata and dbdata (nm(nm+1)(nm+2)/6 * nq, mode (p, q, r) with r fastest).  Returns
 * IMF_ERR_ARG unless 2 <= nm and 1 <= nq <= 64. */
int imf_basis_modified_c(int nm, int nq, double *z, double *w, double *D,
                         double *bdata, double *dbdata);
/* ModifiedPyr_C on nq Gauss-Radau (alpha = 2, beta = 0) points (Basis.cpp:569-668), the
 * eta2 direction of a pyramid: mode (p, q, r) with r < nm - max(p, q), r fastest, then q,
 * then p; phi^b_{max(p,q),r} when min(p, q) <= 1, else ((1-z)/2)^{p+q-2} times
 * (1+z)/2 P_{r-1}^{(2p+2q-3,1)} for r >= 1.  bdata and dbdata hold nm(nm+1)(2nm+1)/6 * nq
 * values.  Returns IMF_ERR_ARG unless 2 <= nm and 1 <= nq <= 64. */
int imf_basis_modified_pyr_c(int nm, int nq, double *z, double *w, double *D,
                             double *bdata, double *dbdata);

/* GeomFactors::ComputeDerivFactors (SpatialDomains/GeomFactors.cpp:399-478) for
 * straight-sided (trilinear hex / bilinear quad) elements, evaluated on the GPU.
 * d_verts: nElmt x 8 x 3 (hex) or nElmt x 4 x 2 (quad) in geometric vertex order;
 * z: host GLL points (nq).  Writes d_jac / d_df in the layouts above. */
int imf_geom_hex_device(const double *d_verts, int nElmt, int nq, const double *z,
                        int deformed, double *d_jac, double *d_df, void *stream);
int imf_geom_quad_device(const double *d_verts, int nElmt, int nq, const double *z,
                         int deformed, double *d_jac, double *d_df, void *stream);
/* Straight-sided triangles (TriGeom, SpatialDomains/TriGeom.cpp; GeomFactors.cpp:420-474):
 * d_verts nElmt x 3 x 2 in the element's vertex order; the map is linear, so the factors
 * are constant: one value per element (the reference's eRegular factors), or nq0*nq1
 * copies when deformed != 0.  Layouts as above (4 factors, d = c*2 + a). */
int imf_geom_tri_device(const double *d_verts, int nElmt, int nq0, int nq1, int deformed,
                        double *d_jac, double *d_df, void *stream);
/* Straight segments (SegGeom; GeomFactors.cpp:399-478 for one expansion direction):
 * d_verts nElmt x 2 x ncoord (the two end points); J = |t| and df[c] = t_c / |t|^2 with
 * t = (x1 - x0) / 2, per element (deformed = 0) or repeated per point (deformed = 1). */
int imf_geom_seg_device(const double *d_verts, int nElmt, int ncoord, int nq, int deformed,
                        double *d_jac, double *d_df, void *stream);
/* Straight-sided pyramids (PyrGeom, SpatialDomains/PyrGeom.cpp; GeomFactors.cpp):
 * d_verts nElmt x 5 x 3 in StdPyrExp vertex order (xi = (-1,-1,-1), (1,-1,-1), (1,1,-1),
 * (-1,1,-1), apex (-1,-1,1)).  A base that is not a parallelogram makes the map
 * non-affine, so deformed != 0 gives factors at each of the nq0*nq1*nq2 points (nq0, nq1
 * GLL, nq2 Gauss-Radau alpha = 2); deformed = 0 gives one value per element, at point 0. */
int imf_geom_pyr_device(const double *d_verts, int nElmt, int nq0, int nq1, int nq2,
                        int deformed, double *d_jac, double *d_df, void *stream);
/* Straight-sided tetrahedra (TetGeom, SpatialDomains/TetGeom.cpp; GeomFactors.cpp:420-474):
 * d_verts nElmt x 4 x 3 in StdTetExp vertex order (xi = (-1,-1,-1), (1,-1,-1), (-1,1,-1),
 * (-1,-1,1)); the map is affine, so the factors are constant: one value per element, or
 * nq0*nq1*nq2 copies when deformed != 0.  Layouts as above (9 factors, d = c*3 + a). */
int imf_geom_tet_device(const double *d_verts, int nElmt, int nq0, int nq1, int nq2,
                        int deformed, double *d_jac, double *d_df, void *stream);
/* Straight-sided prisms (PrismGeom, SpatialDomains/PrismGeom.cpp; GeomFactors.cpp:420-474):
 * d_verts nElmt x 6 x 3 in StdPrismExp vertex order: 0..3 the xi2 = -1 quad at (xi0, xi1)
 * = (-1,-1), (1,-1), (1,1), (-1,1); 4 and 5 the collapsed edge at xi2 = 1, xi1 = -1 and 1.
 * The map is not affine, so deformed != 0 gives factors at each of the nq0*nq1*nq2 points
 * (nq0, nq1 GLL, nq2 Gauss-Radau, as the prism operators use); deformed = 0 gives one
 * value per element, at point 0.  Layouts as above (9 factors, d = c*3 + a). */
int imf_geom_prism_device(const double *d_verts, int nElmt, int nq0, int nq1, int nq2,
                          int deformed, double *d_jac, double *d_df, void *stream);
/* The hex factors on nq[d] points z[d] (host) along direction d; points ordered
 * (k*nq[1] + j)*nq[0] + i.  For the operators with per-direction points. */
int imf_geom_hex_aniso_device(const double *d_verts, int nElmt, const int *nq,
                              const double *const *z, int deformed, double *d_jac, double *d_df,
                              void *stream);

#ifdef __cplusplus
}
#endif
#endif
