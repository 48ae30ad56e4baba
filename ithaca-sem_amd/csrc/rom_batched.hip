// Batched online reduced solves for gfx950: A(mu_i) x_i = b_i with A(mu) = sum_q theta_q A_q.
//
// The reference solves one parameter at a time on the CPU (CoupledLinearNS_ROM.cpp:
// gen_affine_mat_proj :4694-4704, colPivHouseholderQr().solve :4053).  Here one persistent
// launch processes the whole batch: 256-thread workgroups pull instances (sorted by size,
// largest first) from a device counter, assemble A(mu) into a per-workgroup workspace,
// factor it with a blocked right-looking LU with partial pivoting (panel in LDS,
// trailing update in register tiles), and solve with blocked triangular solves (diagonal
// blocks wave-synchronous, off-diagonal blocks as one GEMV each).
//
// Mixed precision (MIXED): the LU factors and the correction solves run in FP32; the
// assembled A(mu) is kept in FP64 and the residual r = b - A x, the convergence test and
// the update x += d are FP64 (classical iterative refinement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ithaca_rom.h"

namespace irom {

constexpr int NMAX = 256;
constexpr int NT = 256;

struct BatchDev {
    const int *order;  // processing order (largest N first)
    const int *N;
    const int *Q;
    const long *aoff;
    const long *toff;
    const long *voff;
    int nInst;
};

struct Args {
    BatchDev bd;
    int transpose;
    const double *A;
    const double *theta;
    const double *b;
    double *x;
    double tol;
    int max_iter;
    int *iters;
    int *status;
    int *counter;
    double *ws64;   // per workgroup NMAX*NMAX (assembled FP64 A / FP64 factors)
    float *ws32;    // per workgroup NMAX*NMAX (FP32 factors, MIXED only)
};

// Block-wide argmax of v (ties -> smallest index).  Ends synchronised.
__device__ __forceinline__ void block_argmax(double v, int ix, double *sv, int *si, double &bv,
                                             int &bi)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(ix, off);
        if (ov > v || (ov == v && oi < ix)) {
            v = ov;
            ix = oi;
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sv[w] = v;
        si[w] = ix;
    }
    __syncthreads();
    bv = sv[0];
    bi = si[0];
#pragma unroll
    for (int k = 1; k < NT / 64; ++k)
        if (sv[k] > bv || (sv[k] == bv && si[k] < bi)) {
            bv = sv[k];
            bi = si[k];
        }
    __syncthreads();
}

__device__ __forceinline__ double block_max(double v, double *sv)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sv[0];
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) r = fmax(r, sv[k]);
    __syncthreads();
    return r;
}

// In-place LU with partial pivoting of the n x n row-major matrix L (ld n).
// ipiv[k] = row swapped with k at step k.  Returns through *zero a zero-pivot flag.
template <typename T, int NB, int TM, int TN>
__device__ void block_lu(T *__restrict__ L, int n, T *P, int *ipiv, double *sv, int *si,
                         int *zero)
{
    constexpr int NBP = NB + 1;
    const int tid = threadIdx.x;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = min(NB, n - k0), m = n - k0;
        for (int idx = tid; idx < m * nb; idx += NT) {
            const int r = idx / nb, c = idx - r * nb;
            P[r * NBP + c] = L[(long)(k0 + r) * n + k0 + c];
        }
        __syncthreads();
        for (int c = 0; c < nb; ++c) {
            double v = -1.0;
            int ix = 0x7fffffff;
            if (tid >= c && tid < m) {
                v = fabs((double)P[tid * NBP + c]);
                ix = tid;
            }
            double bv;
            int bi;
            block_argmax(v, ix, sv, si, bv, bi);
            if (tid == 0) {
                ipiv[k0 + c] = k0 + bi;
                if (!(bv > 0.0)) *zero = 1;
            }
            if (bi != c && tid < nb) {
                const T t = P[c * NBP + tid];
                P[c * NBP + tid] = P[bi * NBP + tid];
                P[bi * NBP + tid] = t;
            }
            __syncthreads();
            const T piv = P[c * NBP + c];
            if (tid > c && tid < m) {
                const T l = P[tid * NBP + c] / piv;
                P[tid * NBP + c] = l;
                for (int cc = c + 1; cc < nb; ++cc) P[tid * NBP + cc] -= l * P[c * NBP + cc];
            }
            __syncthreads();
        }
        for (int idx = tid; idx < m * nb; idx += NT) {
            const int r = idx / nb, c = idx - r * nb;
            L[(long)(k0 + r) * n + k0 + c] = P[r * NBP + c];
        }
        // row interchanges of this panel applied to the other columns
        for (int j = tid; j < n; j += NT) {
            if (j >= k0 && j < k0 + nb) continue;
            for (int c = 0; c < nb; ++c) {
                const int p = ipiv[k0 + c];
                if (p != k0 + c) {
                    const T t = L[(long)(k0 + c) * n + j];
                    L[(long)(k0 + c) * n + j] = L[(long)p * n + j];
                    L[(long)p * n + j] = t;
                }
            }
        }
        __syncthreads();
        // U12 = L11^{-1} A12 (unit lower L11 from the panel); each thread owns a column and
        // re-reads the rows it has already produced (L1-resident, same thread).
        for (int j = k0 + nb + tid; j < n; j += NT) {
            for (int c = 1; c < nb; ++c) {
                T s = L[(long)(k0 + c) * n + j];
                for (int cp = 0; cp < c; ++cp) s -= P[c * NBP + cp] * L[(long)(k0 + cp) * n + j];
                L[(long)(k0 + c) * n + j] = s;
            }
        }
        __syncthreads();
        // A22 -= L21 U12
        const int m2 = n - k0 - nb;
        if (m2 > 0) {
            const int tr_n = (m2 + TM - 1) / TM, tc_n = (m2 + TN - 1) / TN;
            for (int t = tid; t < tr_n * tc_n; t += NT) {
                const int tr = t / tc_n, tc = t - tr * tc_n;
                const int i0 = k0 + nb + tr * TM, j0 = k0 + nb + tc * TN;
                T acc[TM][TN];
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq)
                        acc[a][bq] = (i0 + a < n && j0 + bq < n) ? L[(long)(i0 + a) * n + j0 + bq] : T(0);
                for (int c = 0; c < nb; ++c) {
                    T av[TM], bv[TN];
#pragma unroll
                    for (int a = 0; a < TM; ++a)
                        av[a] = (i0 + a < n) ? P[(nb + tr * TM + a) * NBP + c] : T(0);
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq)
                        bv[bq] = (j0 + bq < n) ? L[(long)(k0 + c) * n + j0 + bq] : T(0);
#pragma unroll
                    for (int a = 0; a < TM; ++a)
#pragma unroll
                        for (int bq = 0; bq < TN; ++bq) acc[a][bq] -= av[a] * bv[bq];
                }
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq)
                        if (i0 + a < n && j0 + bq < n) L[(long)(i0 + a) * n + j0 + bq] = acc[a][bq];
            }
        }
        __syncthreads();
    }
}

// Solve L U y = y in place (y already permuted), factors row-major in L (ld n).  The
// diagonal blocks are staged in D (NB*NB scalars of LDS) and solved by wave 0 without
// barriers; the off-diagonal blocks are one GEMV each.
template <typename T, int NB>
__device__ void block_lu_solve(const T *__restrict__ L, int n, T *y, T *D)
{
    const int tid = threadIdx.x, lane = tid & 63;
    // forward, unit lower
    for (int kb = 0; kb < n; kb += NB) {
        const int nb = min(NB, n - kb);
        for (int idx = tid; idx < nb * nb; idx += NT) {
            const int r = idx / nb, c = idx - r * nb;
            D[r * NB + c] = L[(long)(kb + r) * n + kb + c];
        }
        __syncthreads();
        if (tid < 64) {
            T yr = lane < nb ? y[kb + lane] : T(0);
            for (int c = 0; c < nb; ++c) {
                const T yc = __shfl(yr, c);
                if (lane > c && lane < nb) yr -= D[lane * NB + c] * yc;
            }
            if (lane < nb) y[kb + lane] = yr;
        }
        __syncthreads();
        for (int i = kb + nb + tid; i < n; i += NT) {
            T s = y[i];
            for (int c = 0; c < nb; ++c) s -= L[(long)i * n + kb + c] * y[kb + c];
            y[i] = s;
        }
        __syncthreads();
    }
    // backward, non-unit upper
    const int last = ((n - 1) / NB) * NB;
    for (int kb = last; kb >= 0; kb -= NB) {
        const int nb = min(NB, n - kb);
        for (int idx = tid; idx < nb * nb; idx += NT) {
            const int r = idx / nb, c = idx - r * nb;
            D[r * NB + c] = L[(long)(kb + r) * n + kb + c];
        }
        __syncthreads();
        if (tid < 64) {
            T yr = lane < nb ? y[kb + lane] : T(0);
            for (int c = nb - 1; c >= 0; --c) {
                if (lane == c) yr = yr / D[c * NB + c];
                const T yc = __shfl(yr, c);
                if (lane < c) yr -= D[lane * NB + c] * yc;
            }
            if (lane < nb) y[kb + lane] = yr;
        }
        __syncthreads();
        for (int i = tid; i < kb; i += NT) {
            T s = y[i];
            for (int c = 0; c < nb; ++c) s -= L[(long)i * n + kb + c] * y[kb + c];
            y[i] = s;
        }
        __syncthreads();
    }
}

template <typename T, bool MIXED, int NB, int TM, int TN>
__global__ __launch_bounds__(NT, 2) void batched_solve_kernel(Args a)
{
    constexpr int NBP = NB + 1;
    __shared__ T P[NMAX * NBP];
    __shared__ double b64[NMAX], x64[NMAX], r64[NMAX];
    __shared__ T yv[NMAX];
    __shared__ int ipiv[NMAX], perm[NMAX];
    __shared__ double sv[NT / 64];
    __shared__ int si[NT / 64];
    __shared__ int s_inst, s_zero;

    const int tid = threadIdx.x;
    double *A64 = a.ws64 + (long)blockIdx.x * NMAX * NMAX;
    T *LU = MIXED ? (T *)(a.ws32 + (long)blockIdx.x * NMAX * NMAX) : (T *)A64;

    for (;;) {
        if (tid == 0) {
            s_inst = atomicAdd(a.counter, 1);
            s_zero = 0;
        }
        __syncthreads();
        const int k = s_inst;
        __syncthreads();
        if (k >= a.bd.nInst) break;
        const int inst = a.bd.order[k];
        const int n = a.bd.N[inst], nq = a.bd.Q[inst];
        const double *Ai = a.A + a.bd.aoff[inst];
        const double *th = a.theta + a.bd.toff[inst];
        const long vo = a.bd.voff[inst];

        // assemble A(mu) (or its transpose)
        const long nn = (long)n * n;
        for (long idx = tid; idx < nn; idx += NT) {
            const int i = (int)(idx / n), j = (int)(idx - (long)i * n);
            const long src = a.transpose ? (long)j * n + i : idx;
            double s = 0.0;
            for (int q = 0; q < nq; ++q) s += th[q] * Ai[q * nn + src];
            A64[idx] = s;
            if (MIXED) LU[idx] = (T)s;
        }
        for (int i = tid; i < n; i += NT) b64[i] = a.b[vo + i];
        __syncthreads();

        block_lu<T, NB, TM, TN>(LU, n, P, ipiv, sv, si, &s_zero);
        if (tid == 0) {
            for (int i = 0; i < n; ++i) perm[i] = i;
            for (int c = 0; c < n; ++c) {
                const int p = ipiv[c];
                const int t = perm[c];
                perm[c] = perm[p];
                perm[p] = t;
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += NT) yv[i] = (T)b64[perm[i]];
        __syncthreads();
        block_lu_solve<T, NB>(LU, n, yv, P);
        for (int i = tid; i < n; i += NT) x64[i] = (double)yv[i];
        __syncthreads();

        int it = 0, st = s_zero ? 1 : 0;
        if (MIXED) {
            // ||A||_inf and ||b||_inf for the backward-error test
            double arow = 0.0;
            const int w = tid >> 6, lane = tid & 63;
            for (int i = w; i < n; i += NT / 64) {
                double s = 0.0;
                for (int j = lane; j < n; j += 64) s += fabs(A64[(long)i * n + j]);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                arow = fmax(arow, s);
            }
            double bmax = 0.0;
            for (int i = tid; i < n; i += NT) bmax = fmax(bmax, fabs(b64[i]));
            const double anorm = block_max(arow, sv);
            const double bnorm = block_max(bmax, sv);
            bool conv = false;
            for (; it <= a.max_iter; ++it) {
                // r = b - A x  (one wave per row, lanes over columns)
                for (int i = w; i < n; i += NT / 64) {
                    double s = 0.0;
                    for (int j = lane; j < n; j += 64) s += A64[(long)i * n + j] * x64[j];
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
                    if (lane == 0) r64[i] = b64[i] - s;
                }
                __syncthreads();
                double rm = 0.0, xm = 0.0;
                for (int i = tid; i < n; i += NT) {
                    rm = fmax(rm, fabs(r64[i]));
                    xm = fmax(xm, fabs(x64[i]));
                }
                const double rn = block_max(rm, sv);
                const double xn = block_max(xm, sv);
                if (rn <= a.tol * (anorm * xn + bnorm)) {
                    conv = true;
                    break;
                }
                if (it == a.max_iter) break;
                for (int i = tid; i < n; i += NT) yv[i] = (T)r64[perm[i]];
                __syncthreads();
                block_lu_solve<T, NB>(LU, n, yv, P);
                for (int i = tid; i < n; i += NT) x64[i] += (double)yv[i];
                __syncthreads();
            }
            if (!conv && !st) st = 2;
        }
        for (int i = tid; i < n; i += NT) a.x[vo + i] = x64[i];
        if (tid == 0) {
            if (a.iters) a.iters[inst] = it;
            if (a.status) a.status[inst] = st;
        }
        __syncthreads();
    }
}

__global__ void reset_counter(int *c) { *c = 0; }

// dJ/dmu_q = -lambda^T A_q x for every instance and term (one workgroup per instance).
__global__ __launch_bounds__(NT) void param_grad_kernel(BatchDev bd, const double *A,
                                                         const double *lam, const double *x,
                                                         double *grad)
{
    __shared__ double ls[NMAX], xs[NMAX], sv[NT / 64];
    const int inst = blockIdx.x;
    const int n = bd.N[inst], nq = bd.Q[inst];
    const long vo = bd.voff[inst];
    for (int i = threadIdx.x; i < n; i += NT) {
        ls[i] = lam[vo + i];
        xs[i] = x[vo + i];
    }
    __syncthreads();
    const long nn = (long)n * n;
    for (int q = 0; q < nq; ++q) {
        const double *Aq = A + bd.aoff[inst] + q * nn;
        double s = 0.0;
        for (int j = threadIdx.x; j < n; j += NT) {
            double t = 0.0;
            for (int i = 0; i < n; ++i) t += ls[i] * Aq[(long)i * n + j];
            s += t * xs[j];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0.0;
            for (int k = 0; k < NT / 64; ++k) tot += sv[k];
            grad[bd.toff[inst] + q] = -tot;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Variants (tile shapes) and host side
// ---------------------------------------------------------------------------
using KernelFn = void (*)(Args);

struct Variant {
    const char *name;
    KernelFn fp64, mixed;
};

#define IROM_V(NB, TM, TN)                                                               \
    Variant{"NB" #NB "_T" #TM "x" #TN, &batched_solve_kernel<double, false, NB, TM, TN>, \
            &batched_solve_kernel<float, true, NB, TM, TN>}

const Variant kVariants[] = {IROM_V(16, 4, 4), IROM_V(32, 4, 4), IROM_V(16, 2, 4),
                             IROM_V(32, 2, 4), IROM_V(16, 4, 2), IROM_V(32, 2, 2)};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace irom

struct irom_batch_s {
    int nInst = 0;
    std::vector<int> N, Q;
    std::vector<long> aoff, toff, voff;
    size_t a_size = 0, t_size = 0, v_size = 0;
    int *d_order = nullptr, *d_N = nullptr, *d_Q = nullptr, *d_counter = nullptr;
    long *d_aoff = nullptr, *d_toff = nullptr, *d_voff = nullptr;
    double *d_ws64 = nullptr;
    float *d_ws32 = nullptr;
    int nWG = 0, nCU = 0;
    int variant[2] = {0, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &m)
{
    g_err = m;
    return code;
}

int hfail(hipError_t e, const char *what)
{
    return fail(3, std::string(what) + ": " + hipGetErrorString(e));
}

#define IROM_CHECK(call, what)                      \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return hfail(e_, what); \
    } while (0)

irom::BatchDev dev_view(const irom_batch b)
{
    return irom::BatchDev{b->d_order, b->d_N, b->d_Q, b->d_aoff, b->d_toff, b->d_voff, b->nInst};
}

int launch_solve(irom_batch b, int method, int variant, int transpose, const double *d_A,
                 const double *d_theta, const double *d_b, double *d_x, double tol, int max_iter,
                 int *d_iters, int *d_status, hipStream_t s)
{
    if (method != IROM_LU_FP64 && method != IROM_LU_FP32_IR) return fail(1, "unknown method");
    if (!d_A || !d_theta || !d_b || !d_x) return fail(1, "null device array");
    irom::Args a{dev_view(b), transpose, d_A, d_theta, d_b, d_x, tol, max_iter, d_iters,
                 d_status, b->d_counter, b->d_ws64, b->d_ws32};
    irom::reset_counter<<<1, 1, 0, s>>>(b->d_counter);
    const irom::Variant &v = irom::kVariants[variant];
    irom::KernelFn fn = method == IROM_LU_FP64 ? v.fp64 : v.mixed;
    // persistent grid: as many workgroups as are co-resident, capped by the workspace
    int per_cu = 1;
    IROM_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, irom::NT, 0),
               "occupancy");
    const int grid = std::max(1, std::min({b->nWG, std::max(1, per_cu) * b->nCU, b->nInst}));
    fn<<<grid, irom::NT, 0, s>>>(a);
    IROM_CHECK(hipGetLastError(), "batched solve launch");
    return 0;
}

}  // namespace

extern "C" {

const char *irom_last_error(void) { return g_err.c_str(); }

int irom_batch_create(int nInst, const int *N, const int *Q, irom_batch *out)
{
    if (nInst <= 0 || !N || !Q || !out) return fail(1, "irom_batch_create: bad argument");
    auto *b = new irom_batch_s();
    b->nInst = nInst;
    b->N.assign(N, N + nInst);
    b->Q.assign(Q, Q + nInst);
    long ao = 0, to = 0, vo = 0;
    for (int i = 0; i < nInst; ++i) {
        if (N[i] < 1 || N[i] > irom::NMAX || Q[i] < 1) {
            delete b;
            return fail(1, "irom_batch_create: N must be in 1..256 and Q >= 1");
        }
        b->aoff.push_back(ao);
        b->toff.push_back(to);
        b->voff.push_back(vo);
        ao += (long)Q[i] * N[i] * N[i];
        to += Q[i];
        vo += N[i];
    }
    b->a_size = ao;
    b->t_size = to;
    b->v_size = vo;
    // largest (most work) first
    std::vector<int> order(nInst);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return N[x] > N[y]; });
    int dev = 0, ncu = 0;
    IROM_CHECK(hipGetDevice(&dev), "hipGetDevice");
    IROM_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "attr");
    b->nCU = ncu;
    b->nWG = std::min(nInst, 4 * ncu);
    IROM_CHECK(hipMalloc(&b->d_order, nInst * sizeof(int)), "hipMalloc");
    IROM_CHECK(hipMalloc(&b->d_N, nInst * sizeof(int)), "hipMalloc");
    IROM_CHECK(hipMalloc(&b->d_Q, nInst * sizeof(int)), "hipMalloc");
    IROM_CHECK(hipMalloc(&b->d_counter, sizeof(int)), "hipMalloc");
    IROM_CHECK(hipMalloc(&b->d_aoff, nInst * sizeof(long)), "hipMalloc");
    IROM_CHECK(hipMalloc(&b->d_toff, nInst * sizeof(long)), "hipMalloc");
    IROM_CHECK(hipMalloc(&b->d_voff, nInst * sizeof(long)), "hipMalloc");
    const size_t ws = (size_t)b->nWG * irom::NMAX * irom::NMAX;
    IROM_CHECK(hipMalloc(&b->d_ws64, ws * sizeof(double)), "hipMalloc(workspace)");
    IROM_CHECK(hipMalloc(&b->d_ws32, ws * sizeof(float)), "hipMalloc(workspace)");
    IROM_CHECK(hipMemcpy(b->d_order, order.data(), nInst * sizeof(int), hipMemcpyHostToDevice), "copy");
    IROM_CHECK(hipMemcpy(b->d_N, N, nInst * sizeof(int), hipMemcpyHostToDevice), "copy");
    IROM_CHECK(hipMemcpy(b->d_Q, Q, nInst * sizeof(int), hipMemcpyHostToDevice), "copy");
    IROM_CHECK(hipMemcpy(b->d_aoff, b->aoff.data(), nInst * sizeof(long), hipMemcpyHostToDevice), "copy");
    IROM_CHECK(hipMemcpy(b->d_toff, b->toff.data(), nInst * sizeof(long), hipMemcpyHostToDevice), "copy");
    IROM_CHECK(hipMemcpy(b->d_voff, b->voff.data(), nInst * sizeof(long), hipMemcpyHostToDevice), "copy");
    b->variant[0] = 1;  // NB32 4x4
    b->variant[1] = 1;
    *out = b;
    return 0;
}

int irom_batch_destroy(irom_batch b)
{
    if (!b) return 0;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    for (void *p : {(void *)b->d_order, (void *)b->d_N, (void *)b->d_Q, (void *)b->d_counter,
                    (void *)b->d_aoff, (void *)b->d_toff, (void *)b->d_voff, (void *)b->d_ws64,
                    (void *)b->d_ws32})
        (void)hipFree(p);
    delete b;
    return 0;
}

int irom_batch_layout(irom_batch b, long *a_off, long *t_off, long *v_off, size_t *a_size,
                      size_t *t_size, size_t *v_size)
{
    if (!b) return fail(1, "null batch");
    if (a_off) std::memcpy(a_off, b->aoff.data(), b->nInst * sizeof(long));
    if (t_off) std::memcpy(t_off, b->toff.data(), b->nInst * sizeof(long));
    if (v_off) std::memcpy(v_off, b->voff.data(), b->nInst * sizeof(long));
    if (a_size) *a_size = b->a_size;
    if (t_size) *t_size = b->t_size;
    if (v_size) *v_size = b->v_size;
    return 0;
}

int irom_batch_solve(irom_batch b, int method, int transpose, const double *d_A,
                     const double *d_theta, const double *d_b, double *d_x, double tol,
                     int max_iter, int *d_iters, int *d_status, void *stream)
{
    if (!b) return fail(1, "null batch");
    return launch_solve(b, method, b->variant[method == IROM_LU_FP32_IR], transpose, d_A,
                        d_theta, d_b, d_x, tol, max_iter, d_iters, d_status,
                        (hipStream_t)stream);
}

int irom_batch_num_variants(void) { return irom::kNumVariants; }

const char *irom_batch_variant_name(int v)
{
    return (v >= 0 && v < irom::kNumVariants) ? irom::kVariants[v].name : "";
}

int irom_batch_set_variant(irom_batch b, int method, int v)
{
    if (!b || v < 0 || v >= irom::kNumVariants || (method != 0 && method != 1))
        return fail(1, "irom_batch_set_variant: bad argument");
    b->variant[method] = v;
    return 0;
}

int irom_batch_autotune(irom_batch b, int method, const double *d_A, const double *d_theta,
                        const double *d_b, double *d_x, int reps, int *chosen, void *stream)
{
    if (!b || (method != 0 && method != 1)) return fail(1, "irom_batch_autotune: bad argument");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    IROM_CHECK(hipEventCreate(&e0), "event");
    IROM_CHECK(hipEventCreate(&e1), "event");
    float best = 1e30f;
    int bestv = b->variant[method];
    reps = std::max(1, reps);
    for (int v = 0; v < irom::kNumVariants; ++v) {
        if (int st = launch_solve(b, method, v, 0, d_A, d_theta, d_b, d_x, 1e-14, 10, nullptr,
                                  nullptr, s))
            return st;
        IROM_CHECK(hipEventRecord(e0, s), "event");
        for (int r = 0; r < reps; ++r)
            if (int st = launch_solve(b, method, v, 0, d_A, d_theta, d_b, d_x, 1e-14, 10,
                                      nullptr, nullptr, s))
                return st;
        IROM_CHECK(hipEventRecord(e1, s), "event");
        IROM_CHECK(hipEventSynchronize(e1), "event");
        float ms = 0;
        IROM_CHECK(hipEventElapsedTime(&ms, e0, e1), "event");
        if (ms < best) {
            best = ms;
            bestv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->variant[method] = bestv;
    if (chosen) *chosen = bestv;
    return 0;
}

int irom_batch_capture(irom_batch b, int method, int transpose, const double *d_A,
                       const double *d_theta, const double *d_b, double *d_x, double tol,
                       int max_iter, int *d_iters, int *d_status)
{
    if (!b) return fail(1, "null batch");
    if (!b->cap_stream) IROM_CHECK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking), "stream");
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    IROM_CHECK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal), "capture");
    int st = launch_solve(b, method, b->variant[method == IROM_LU_FP32_IR], transpose, d_A,
                          d_theta, d_b, d_x, tol, max_iter, d_iters, d_status, b->cap_stream);
    hipError_t e = hipStreamEndCapture(b->cap_stream, &b->graph);
    if (st) return st;
    if (e != hipSuccess) return hfail(e, "hipStreamEndCapture");
    IROM_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return 0;
}

int irom_batch_replay(irom_batch b, void *stream)
{
    if (!b || !b->exec) return fail(4, "irom_batch_replay: nothing captured");
    IROM_CHECK(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
    return 0;
}

int irom_batch_param_grad(irom_batch b, const double *d_A, const double *d_lambda,
                          const double *d_x, double *d_grad, void *stream)
{
    if (!b || !d_A || !d_lambda || !d_x || !d_grad) return fail(1, "irom_batch_param_grad: bad argument");
    irom::param_grad_kernel<<<b->nInst, irom::NT, 0, (hipStream_t)stream>>>(dev_view(b), d_A,
                                                                           d_lambda, d_x, d_grad);
    IROM_CHECK(hipGetLastError(), "param_grad launch");
    return 0;
}

}  // extern "C"
