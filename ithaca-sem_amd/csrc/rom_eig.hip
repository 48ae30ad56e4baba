// FP64 symmetric eigen-decomposition of the (small, dense) snapshot Gram matrix on the GPU.
//
// The reference obtains its POD basis from an Eigen BDCSVD of the snapshot matrix
// (CoupledLinearNS_ROM.cpp:3925-3954).  By the method of snapshots the same modes follow
// from the eigenpairs of the n x n Gram matrix G = S^T M S (n = number of snapshots).
// Algorithm: one-sided Jacobi (Hestenes) on G, which for a symmetric positive
// semi-definite matrix yields G V = W with orthogonal columns, lambda_i = ||w_i||_2 and
// eigenvectors v_i.  Sweeps use the round-robin ("circle") pair ordering, so every round
// rotates n/2 disjoint column pairs in parallel: one kernel per round, one workgroup per
// pair, and the n-1 rounds of a sweep are replayed from a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ithaca_rom.h"

namespace irom {

constexpr int ET = 256;

__device__ __forceinline__ double block_sum(double v, double *sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < ET / 64; ++k) r += sh[k];
    __syncthreads();
    return r;
}

// player at position pos (0..n-1) in round r of the circle schedule for n (even) players
__device__ __forceinline__ int player(int pos, int r, int n)
{
    return pos == 0 ? 0 : ((pos - 1 + r) % (n - 1)) + 1;
}

// One round: workgroup k rotates columns (p, q) = (player(k), player(n-1-k)) of W and V
// (column-major, ld = n) so that w_p . w_q = 0.  flag is set when any pair was rotated.
__global__ __launch_bounds__(ET) void jacobi_round(int n, int r, double tol, double *__restrict__ W,
                                                   double *__restrict__ V, int *flag)
{
    __shared__ double sh[3][ET / 64];
    const int k = blockIdx.x;
    const int p = player(k, r, n), q = player(n - 1 - k, r, n);
    double *wp = W + (long)p * n, *wq = W + (long)q * n;
    double a = 0.0, b = 0.0, g = 0.0;
    for (int i = threadIdx.x; i < n; i += ET) {
        const double x = wp[i], y = wq[i];
        a += x * x;
        b += y * y;
        g += x * y;
    }
    a = block_sum(a, sh[0]);
    b = block_sum(b, sh[1]);
    g = block_sum(g, sh[2]);
    if (g == 0.0 || !(fabs(g) > tol * sqrt(a * b))) return;
    const double zeta = (b - a) / (2.0 * g);
    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
    double *vp = V + (long)p * n, *vq = V + (long)q * n;
    for (int i = threadIdx.x; i < n; i += ET) {
        const double x = wp[i], y = wq[i];
        wp[i] = c * x - s * y;
        wq[i] = s * x + c * y;
        const double u = vp[i], w = vq[i];
        vp[i] = c * u - s * w;
        vq[i] = s * u + c * w;
    }
    if (threadIdx.x == 0) atomicOr(flag, 1);
}

__global__ void set_flag(int *flag) { *flag = 0; }

// W <- G (padded to np x np with zeros), V <- I (np x np); G row-major n x n.
__global__ void jacobi_init(int n, int np, const double *__restrict__ G, double *__restrict__ W,
                            double *__restrict__ V)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)np * np) return;
    const int col = (int)(g / np), row = (int)(g - (long)col * np);
    W[g] = (row < n && col < n) ? G[(long)row * n + col] : 0.0;
    V[g] = row == col ? 1.0 : 0.0;
}

__global__ __launch_bounds__(ET) void column_norms(int np, const double *__restrict__ W,
                                                   double *__restrict__ nrm)
{
    __shared__ double sh[ET / 64];
    const double *w = W + (long)blockIdx.x * np;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += ET) s += w[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) nrm[blockIdx.x] = sqrt(s);
}

// out column j (ld n) <- V column perm[j] restricted to the first n rows
__global__ void gather_columns(int n, int np, const int *__restrict__ perm,
                               const double *__restrict__ V, double *__restrict__ out)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)n * n) return;
    const int j = (int)(g / n), i = (int)(g - (long)j * n);
    out[g] = V[(long)perm[j] * np + i];
}

}  // namespace irom

namespace {
thread_local std::string g_eerr;
int efail(int c, const std::string &m)
{
    g_eerr = m;
    return c;
}
}  // namespace

#define IROM_ECHECK(call, what)                                                   \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) return efail(3, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

const char *irom_eig_last_error(void) { return g_eerr.c_str(); }

int irom_eig_sym(int n, const double *d_G, double *h_evals, double *d_evecs, double tol,
                 int max_sweeps, int *sweeps, void *stream)
{
    if (n <= 0 || n > 4096 || !d_G || !h_evals || !d_evecs) return efail(1, "irom_eig_sym: bad argument");
    hipStream_t s = (hipStream_t)stream;
    const int np = n + (n & 1);  // even number of players
    double *W = nullptr, *V = nullptr, *nrm = nullptr;
    int *flag = nullptr, *perm_d = nullptr;
    IROM_ECHECK(hipMalloc(&W, sizeof(double) * np * np), "hipMalloc");
    IROM_ECHECK(hipMalloc(&V, sizeof(double) * np * np), "hipMalloc");
    IROM_ECHECK(hipMalloc(&nrm, sizeof(double) * np), "hipMalloc");
    IROM_ECHECK(hipMalloc(&flag, sizeof(int)), "hipMalloc");
    IROM_ECHECK(hipMalloc(&perm_d, sizeof(int) * n), "hipMalloc");
    const long tot = (long)np * np;
    irom::jacobi_init<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, np, d_G, W, V);
    // one sweep = n-1 rounds, captured once and replayed
    hipStream_t cs;
    IROM_ECHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream");
    hipGraph_t graph;
    hipGraphExec_t exec;
    IROM_ECHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "capture");
    irom::set_flag<<<1, 1, 0, cs>>>(flag);
    for (int r = 0; r < np - 1; ++r)
        irom::jacobi_round<<<np / 2, irom::ET, 0, cs>>>(np, r, tol, W, V, flag);
    IROM_ECHECK(hipStreamEndCapture(cs, &graph), "end capture");
    IROM_ECHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
    IROM_ECHECK(hipStreamSynchronize(s), "sync");
    int sw = 0, h_flag = 1;
    for (; sw < max_sweeps && h_flag; ++sw) {
        IROM_ECHECK(hipGraphLaunch(exec, cs), "graph launch");
        IROM_ECHECK(hipMemcpyAsync(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, cs), "copy");
        IROM_ECHECK(hipStreamSynchronize(cs), "sync");
    }
    if (sweeps) *sweeps = sw;
    irom::column_norms<<<np, irom::ET, 0, cs>>>(np, W, nrm);
    std::vector<double> lam(np);
    IROM_ECHECK(hipMemcpyAsync(lam.data(), nrm, sizeof(double) * np, hipMemcpyDeviceToHost, cs), "copy");
    IROM_ECHECK(hipStreamSynchronize(cs), "sync");
    // the padding column (odd n) carries a zero eigenvalue: drop it
    std::vector<int> idx(np);
    std::iota(idx.begin(), idx.end(), 0);
    if (np != n) idx.pop_back();
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return lam[a] > lam[b]; });
    for (int j = 0; j < n; ++j) h_evals[j] = lam[idx[j]];
    IROM_ECHECK(hipMemcpyAsync(perm_d, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, cs), "copy");
    irom::gather_columns<<<(unsigned)(((long)n * n + 255) / 256), 256, 0, cs>>>(n, np, perm_d, V, d_evecs);
    IROM_ECHECK(hipStreamSynchronize(cs), "sync");
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    (void)hipFree(W);
    (void)hipFree(V);
    (void)hipFree(nrm);
    (void)hipFree(flag);
    (void)hipFree(perm_d);
    return sw < max_sweeps || !h_flag ? 0 : efail(2, "irom_eig_sym: not converged");
}

}  // extern "C"
