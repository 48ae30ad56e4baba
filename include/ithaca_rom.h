/*
 * ithaca_rom.h -- C ABI of the MI355X reduced-basis kernels.
 *
 * The reference ROM (solvers/IncNavierStokesSolver/EquationSystems/CoupledLinearNS_ROM.cpp)
 * runs its offline and online phases on the CPU with Eigen 3.3.7:
 *   offline  ROM_offline_phase  :3916-4001  (BDCSVD of the snapshot matrix, truncation)
 *            gen_reference_matrices / gen_proj_adv_terms  :2730-2830, :3639-3911
 *            (Galerkin projections RB^T A RB)
 *   online   ROM_online_phase   :4003-4100  one parameter at a time:
 *            gen_affine_mat_proj :4694-4704 (A(mu) = sum_q theta_q(mu) A_q),
 *            colPivHouseholderQr().solve :4053.
 * The entry points below replace those loops with batched device kernels.  All arrays are
 * device pointers unless stated; streams are hipStream_t passed as void* (NULL = default).
 * Status codes as in ithaca_mf.h (0 = OK), message via irom_last_error().
 */
#ifndef ITHACA_ROM_H
#define ITHACA_ROM_H

#include <stddef.h>

#include "ithaca_mf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct irom_batch_s *irom_batch;

enum irom_method {
    IROM_LU_FP64 = 0,    /* assemble + LU (partial pivoting) + solve, all FP64 */
    IROM_LU_FP32_IR = 1, /* FP32 LU + FP64 residuals / iterative refinement */
    IROM_LU_FP32_IR_FALLBACK = 2  /* as IROM_LU_FP32_IR; every instance whose refinement
                                     does not converge is solved again by FP64 LU in a
                                     second persistent launch (status 3) */
};

/* ---- batched online solves: A(mu_i) x_i = b_i, i < nInst ----------------------------
 * Replaces the per-parameter loop of ROM_online_phase (CoupledLinearNS_ROM.cpp:4003-4100):
 * gen_affine_mat_proj (:4694-4704) and colPivHouseholderQr().solve (:4053).
 * Instance i has size N[i] (1..256) and Q[i] affine terms.  Packed device layouts (offsets
 * in elements, instances in the given order):
 *   A      : A_q of instance i, row-major N x N, at a_off[i] + q*N*N      (double)
 *   theta  : theta_q(mu_i) at t_off[i] + q                                 (double)
 *   b, x   : vectors at v_off[i]                                           (double)
 * The handle owns its workspace and scheduling data; solve calls allocate nothing. */
int irom_batch_create(int nInst, const int *N, const int *Q, irom_batch *out);
int irom_batch_destroy(irom_batch b);
/* offsets (host arrays of nInst entries) and total sizes of the packed arrays */
int irom_batch_layout(irom_batch b, long *a_off, long *t_off, long *v_off, size_t *a_size,
                      size_t *t_size, size_t *v_size);

/* Solve every instance (transpose != 0 solves A(mu)^T x = b, the adjoint system).
 * tol: refinement target (IROM_LU_FP32_IR*): an instance has converged when the normwise
 * backward error ||b - A x||inf / (||A||inf ||x||inf + ||b||inf) is at most tol and the last
 * correction is at most 100 tol ||x||inf or above 0.9 of the one before (the accuracy of
 * the FP64 residual is reached); max_iter caps the number of corrections.
 * One persistent launch (plus a counter reset; two for IROM_LU_FP32_IR_FALLBACK).
 * Per-instance refinement iterations and status (0 ok, 1 zero pivot, 2 refinement not
 * converged, 3 refinement not converged and solved by the FP64 fallback) go to d_iters /
 * d_status when non-NULL (int arrays of nInst). */
int irom_batch_solve(irom_batch b, int method, int transpose, const double *d_A,
                     const double *d_theta, const double *d_b, double *d_x, double tol,
                     int max_iter, int *d_iters, int *d_status, void *stream);

/* Tile variant used by irom_batch_solve for a method: set explicitly, or choose by timing
 * every variant on the given data (the autotuner; returns the chosen index).  */
int irom_batch_num_variants(void);
const char *irom_batch_variant_name(int v);
int irom_batch_set_variant(irom_batch b, int method, int v);
int irom_batch_autotune(irom_batch b, int method, const double *d_A, const double *d_theta,
                        const double *d_b, double *d_x, int reps, int *chosen, void *stream);

/* hipGraph: capture one irom_batch_solve with these arguments into a graph held by the
 * handle, then replay it.  No memset nodes are recorded (counters are reset by a kernel). */
int irom_batch_capture(irom_batch b, int method, int transpose, const double *d_A,
                       const double *d_theta, const double *d_b, double *d_x, double tol,
                       int max_iter, int *d_iters, int *d_status);
int irom_batch_replay(irom_batch b, void *stream);

/* Parameter gradient of J = l^T x(mu) for theta_q = mu_q: with A(mu)^T lambda = l,
 * dJ/dmu_q = -lambda^T A_q x.  Writes d_grad[t_off[i] + q].  No reference counterpart. */
int irom_batch_param_grad(irom_batch b, const double *d_A, const double *d_lambda,
                          const double *d_x, double *d_grad, void *stream);

/* ---- POD offline: Gram / Galerkin products on MFMA --------------------------------
 * C (row-major m x n, leading dimension ldc) (+)= X^T Y, X: K x m and Y: K x n column-major
 * (columns contiguous, leading dimensions ldx / ldy), K up to ~1e8.  Replaces the dense
 * products of the reference's offline phase (snapshot correlation / BDCSVD input,
 * CoupledLinearNS_ROM.cpp:3925; RB^T A RB, :2788-2802, :3661-3663).  Modes:
 *   IROM_GRAM_FP64   FP64 MFMA, FP64 accumulation
 *   IROM_GRAM_FP32X3 hi/lo FP32 split, 3 FP32 MFMA products, FP32 accumulation within a
 *                    16-deep panel, FP64 across panels
 *   IROM_GRAM_BF16X3 hi/lo BF16 split, 3 BF16 MFMA products, FP32 accumulation within a
 *                    32-deep panel, FP64 across panels
 * X == Y (same pointer, m == n, ldx == ldy) computes only the upper blocks and mirrors;
 * so does a mode OR-ed with IROM_GRAM_SYMMETRIC_RESULT when m == n (the caller states
 * that X^T Y is symmetric, e.g. S^T (M S) with M symmetric).
 * d_work must hold irom_gram_workspace_size(m, n, K) bytes (split-K partial blocks). */
enum irom_gram_mode { IROM_GRAM_FP64 = 0, IROM_GRAM_FP32X3 = 1, IROM_GRAM_BF16X3 = 2,
                      IROM_GRAM_SYMMETRIC_RESULT = 16 };
size_t irom_gram_workspace_size(int m, int n, long K);
int irom_gram(int mode, int m, int n, long K, const double *d_X, long ldx, const double *d_Y,
              long ldy, double *d_C, long ldc, int accumulate, void *d_work, size_t work_bytes,
              void *stream);
const char *irom_gram_last_error(void);

/* ---- Galerkin projection of one affine operator term ---------------------------------
 * C (row-major r x r, ldc) = V^T (A V) for the r basis vectors of V (column-major N x r:
 * vector j at d_V + j*ldv, N = imf_in_size(op)) and the FP64 matrix-free operator op, a
 * term that maps coefficients to coefficients ("Helmholtz_Hex_*" with its lambda, e.g.
 * lambda = 0 for the stiffness term, or "Mass_Hex_*" for the mass term).  d_W (N x r,
 * ldw) receives A V: one launch applies op to all r vectors, staging each element's
 * geometry once; then C = V^T W by irom_gram in `mode` (IROM_GRAM_SYMMETRIC_RESULT may be
 * OR-ed in: both terms are symmetric).  d_work holds irom_gram_workspace_size(r, r, N)
 * bytes.  Replaces the per-term dense products RB^T A RB of the reference's offline phase
 * (CoupledLinearNS_ROM.cpp:2788-2802); the reduced operator of a parameter is then
 * sum_q theta_q C_q (gen_affine_mat_proj, :4694-4704).  Asynchronous; allocates nothing.
 * Message of a failure via irom_project_last_error(). */
int irom_galerkin_project(imf_op op, int r, const double *d_V, long ldv, double *d_W, long ldw,
                          double *d_C, long ldc, int mode, void *d_work, size_t work_bytes,
                          void *stream);
const char *irom_project_last_error(void);

/* ---- POD offline: FP64 symmetric eigen-decomposition --------------------------------
 * Eigenpairs of the symmetric positive semi-definite n x n matrix d_G (row-major, device,
 * n <= 4096): h_evals (host, n) in descending order, d_evecs (device, column-major n x n,
 * column j the eigenvector of h_evals[j]).  Blocked one-sided Jacobi in one persistent
 * launch (block pairs diagonalised through their Gram matrices); column pairs rotate while
 * |w_p.w_q| > tol ||w_p|| ||w_q|| and |w_p.w_q| > (tol ||G||_F)^2; at most max_sweeps
 * (<= 64) sweeps, the count in *sweeps.  Replaces the reference's BDCSVD of the snapshot matrix
 * (CoupledLinearNS_ROM.cpp:3925) through the method of snapshots.  Synchronous.
 * irom_eig_sym keeps a workspace per device (allocated on the first call that needs a
 * larger one, kept until the process ends); irom_eig_sym_ws takes the caller's workspace
 * of irom_eig_workspace_size(n) bytes and allocates nothing. */
size_t irom_eig_workspace_size(int n);
int irom_eig_sym(int n, const double *d_G, double *h_evals, double *d_evecs, double tol,
                 int max_sweeps, int *sweeps, void *stream);
int irom_eig_sym_ws(int n, const double *d_G, double *h_evals, double *d_evecs, double tol,
                    int max_sweeps, int *sweeps, void *d_work, size_t work_bytes, void *stream);

/* irom_eig_sym without any allocation, for callers that cannot allocate device memory
 * (tight serving loops, code that forbids hipMalloc).  No counterpart in the reference.
 * The workspace is a static device array sized for n <= 512 (4.2 MB per GPU, held until
 * the process ends); larger n is rejected with an error.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other host threads or streams) race. */
int irom_eig_sym_noalloc(int n, const double *d_G, double *h_evals, double *d_evecs,
                         double tol, int max_sweeps, int *sweeps, void *stream);
const char *irom_eig_last_error(void);

const char *irom_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
