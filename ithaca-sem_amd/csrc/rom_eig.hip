// FP64 symmetric eigen-decomposition of the (small, dense) snapshot Gram matrix on the GPU.
//
// The reference obtains its POD basis from an Eigen BDCSVD of the snapshot matrix
// (CoupledLinearNS_ROM.cpp:3925-3954).  By the method of snapshots the same modes follow
// from the eigenpairs of the n x n Gram matrix G = S^T M S (n = number of snapshots).
//
// Algorithm: blocked one-sided (Hestenes) Jacobi on G.  For a symmetric positive
// semi-definite G, rotating the columns of W = G V until they are mutually orthogonal
// gives lambda_j = ||w_j|| and eigenvectors v_j.  The np columns are split into nb blocks
// of B columns; a sweep is nb-1 rounds of the round-robin ("circle") pairing of blocks.
// In a round, a workgroup takes a block pair: its 2B columns of W are staged in LDS, one
// wave per column pair runs the inner round-robin over the 2B local columns (dot products
// reduced by lane exchanges inside the wave, the rotations applied in LDS and accumulated
// in a 2B x 2B matrix R), then W is written back and V's 2B columns become V R.  All
// rounds and sweeps run in one persistent launch; workgroups meet at a grid barrier
// between rounds (an atomic arrival counter and a generation word, with a bounded wait),
// and stop together after the first sweep without a rotation.  Eigenvalues are sorted on
// the device (bitonic sort in one workgroup) and the eigenvectors gathered in that order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ithaca_rom.h"

namespace irom {

constexpr int EIG_MAX_SWEEPS = 64;   // length of the per-sweep rotation flag array
constexpr int EIG_LDS = 128 * 1024;  // LDS for the staged columns of W

__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// player at position pos (0..m-1) in round r of the circle schedule for m (even) players
__device__ __forceinline__ int player(int pos, int r, int m)
{
    return pos == 0 ? 0 : ((pos - 1 + r) % (m - 1)) + 1;
}

// Grid barrier of nblk co-resident workgroups.  bar[0] counts arrivals, bar[1] is the
// generation; `gen` is this workgroup's copy of the generation.  The wait is bounded (about
// a second): on expiry *err is set and the workgroup goes on, so every wave ends.
__device__ __forceinline__ void grid_sync(unsigned *bar, unsigned nblk, unsigned &gen, int *err)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&bar[0], 1u) == nblk - 1) {
            atomicExch(&bar[0], 0u);
            __threadfence();
            atomicAdd(&bar[1], 1u);
        } else {
            long spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1L << 23)) {
                    atomicOr(err, 1);
                    break;
                }
            }
        }
        __threadfence();
    }
    ++gen;
    __syncthreads();
}

struct EigArgs {
    int np;          // padded order (multiple of 2B)
    int nb;          // number of column blocks (np / B, even)
    double tol;      // rotate while |w_p.w_q| > tol ||w_p|| ||w_q||
    int max_sweeps;  // <= EIG_MAX_SWEEPS
    double *W;       // np x np column-major (W = G V)
    double *V;       // np x np column-major
    int *rot;        // [EIG_MAX_SWEEPS] per-sweep "some pair was rotated" flags (zeroed)
    unsigned *bar;   // [2] grid barrier (zeroed)
    int *sweeps;     // sweeps performed
    int *err;        // barrier expiry
};

// One persistent launch for all sweeps.  B columns per block, one wave per local column
// pair (64*B threads); W's 2B staged columns in dynamic LDS ([2B][np]), R in static LDS.
template <int B>
__global__ __launch_bounds__(64 * B) void block_jacobi_kernel(EigArgs a)
{
    extern __shared__ double lw[];
    __shared__ double R[2 * B][2 * B + 1];  // R[row][col]
    __shared__ int s_any;
    constexpr int C2 = 2 * B;
    const int np = a.np, nb = a.nb;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned gen = 0;
    int sw = 0;
    for (; sw < a.max_sweeps; ++sw) {
        for (int r = 0; r < nb - 1; ++r) {
            for (int k = blockIdx.x; k < nb / 2; k += gridDim.x) {
                const int bi = player(k, r, nb), bj = player(nb - 1 - k, r, nb);
                auto gcol = [&](int c) { return c < B ? bi * B + c : bj * B + (c - B); };
                for (int idx = threadIdx.x; idx < C2 * np; idx += blockDim.x) {
                    const int c = idx / np, i = idx - c * np;
                    lw[idx] = a.W[(long)gcol(c) * np + i];
                }
                for (int idx = threadIdx.x; idx < C2 * C2; idx += blockDim.x) {
                    const int rr = idx / C2, cc = idx - rr * C2;
                    R[rr][cc] = rr == cc ? 1.0 : 0.0;
                }
                if (threadIdx.x == 0) s_any = 0;
                __syncthreads();
                // inner round-robin over the 2B local columns: wave wv rotates one pair
                for (int s = 0; s < C2 - 1; ++s) {
                    const int p = player(wv, s, C2), q = player(C2 - 1 - wv, s, C2);
                    double *wp = lw + (long)p * np, *wq = lw + (long)q * np;
                    double x2 = 0.0, y2 = 0.0, xy = 0.0;
                    for (int i = lane; i < np; i += 64) {
                        const double x = wp[i], y = wq[i];
                        x2 = fma(x, x, x2);
                        y2 = fma(y, y, y2);
                        xy = fma(x, y, xy);
                    }
                    x2 = wave_sum64(x2);
                    y2 = wave_sum64(y2);
                    xy = wave_sum64(xy);
                    if (xy != 0.0 && fabs(xy) > a.tol * sqrt(x2 * y2)) {
                        const double zeta = (y2 - x2) / (2.0 * xy);
                        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                        for (int i = lane; i < np; i += 64) {
                            const double x = wp[i], y = wq[i];
                            wp[i] = c * x - sn * y;
                            wq[i] = sn * x + c * y;
                        }
                        if (lane < C2) {
                            const double u = R[lane][p], w = R[lane][q];
                            R[lane][p] = c * u - sn * w;
                            R[lane][q] = sn * u + c * w;
                        }
                        if (lane == 0) s_any = 1;
                    }
                    __syncthreads();
                }
                // W back; V's 2B columns <- V R (row i of the block: 2B loads, 2B outputs)
                for (int idx = threadIdx.x; idx < C2 * np; idx += blockDim.x) {
                    const int c = idx / np, i = idx - c * np;
                    a.W[(long)gcol(c) * np + i] = lw[idx];
                }
                if (s_any) {
                    for (int i = threadIdx.x; i < np; i += blockDim.x) {
                        double v[C2];
#pragma unroll
                        for (int d = 0; d < C2; ++d) v[d] = a.V[(long)gcol(d) * np + i];
#pragma unroll 1
                        for (int c = 0; c < C2; ++c) {
                            double s = 0.0;
#pragma unroll
                            for (int d = 0; d < C2; ++d) s = fma(v[d], R[d][c], s);
                            a.V[(long)gcol(c) * np + i] = s;
                        }
                    }
                    if (threadIdx.x == 0) atomicOr(&a.rot[sw], 1);
                }
                __syncthreads();  // LDS reused by the next block pair of this workgroup
            }
            grid_sync(a.bar, gridDim.x, gen, a.err);
        }
        // every flag of this sweep was set before the last barrier: all workgroups read
        // the same value and stop together
        if (__hip_atomic_load(&a.rot[sw], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            ++sw;
            break;
        }
        if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.sweeps = sw;
}

__global__ void eig_init(int n, int np, const double *__restrict__ G, double *__restrict__ W,
                         double *__restrict__ V, int *__restrict__ zero_ints, int nzero)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nzero) zero_ints[g] = 0;
    if (g >= (long)np * np) return;
    const int col = (int)(g / np), row = (int)(g - (long)col * np);
    W[g] = (row < n && col < n) ? G[(long)row * n + col] : 0.0;
    V[g] = row == col ? 1.0 : 0.0;
}

constexpr int ET = 256;

__global__ __launch_bounds__(ET) void column_norms(int np, const double *__restrict__ W,
                                                   double *__restrict__ nrm)
{
    __shared__ double sh[ET / 64];
    const double *w = W + (long)blockIdx.x * np;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += ET) s = fma(w[i], w[i], s);
    s = wave_sum64(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < ET / 64; ++k) t += sh[k];
        nrm[blockIdx.x] = sqrt(t);
    }
}

// Descending order of the first n norms (ties by index), by a bitonic sort of m (a power of
// two >= n) keys in one workgroup; writes the sorted values and the permutation.
__global__ __launch_bounds__(1024) void sort_desc(int n, int m, const double *__restrict__ nrm,
                                                  double *__restrict__ evals, int *__restrict__ perm)
{
    extern __shared__ unsigned char sm[];
    double *key = reinterpret_cast<double *>(sm);
    int *idx = reinterpret_cast<int *>(key + m);
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        key[i] = i < n ? nrm[i] : -1.0;  // norms are >= 0: the padding sorts last
        idx[i] = i;
    }
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    // "before": larger key, or equal key and smaller index
                    const bool desc = (i & k) == 0;
                    const bool lbefore = key[l] > key[i] || (key[l] == key[i] && idx[l] < idx[i]);
                    if (desc ? lbefore : !lbefore) {
                        const double tk = key[i];
                        key[i] = key[l];
                        key[l] = tk;
                        const int ti = idx[i];
                        idx[i] = idx[l];
                        idx[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        evals[i] = key[i];
        perm[i] = idx[i];
    }
}

// out column j (ld n) <- V column perm[j] restricted to the first n rows, divided by its
// norm vn[perm[j]]: the accumulated products of rotations leave each column's length off
// one by O(rotations * eps); the normalisation removes that drift (directions unchanged)
__global__ void gather_columns(int n, int np, const int *__restrict__ perm,
                               const double *__restrict__ V, const double *__restrict__ vn,
                               double *__restrict__ out)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)n * n) return;
    const int j = (int)(g / n), i = (int)(g - (long)j * n);
    const int c = perm[j];
    out[g] = V[(long)c * np + i] / vn[c];
}

// Block width for order n: the widest B with 2B staged columns of np rows in EIG_LDS.
struct EigPlan {
    int B, np, nb, m;
};

constexpr EigPlan eig_plan(int n)
{
    EigPlan p{1, 0, 0, 1};
    for (int B = 16; B >= 1; B /= 2) {
        const int np = (n + 2 * B - 1) / (2 * B) * (2 * B);
        if ((long)2 * B * np * 8 <= EIG_LDS || B == 1) {
            p.B = B;
            p.np = np;
            p.nb = np / B;
            break;
        }
    }
    while (p.m < n) p.m <<= 1;
    return p;
}

// Workspace layout (bytes, each part 256-byte aligned): W, V (np^2 doubles each), norms
// (2 np: of W's and of V's columns), sorted values (n), perm (n ints), rot flags, barrier words, sweeps, err.
struct EigWs {
    double *W, *V, *nrm, *ev;
    int *perm, *ints;  // ints: rot[EIG_MAX_SWEEPS], bar[2], sweeps, err
};

constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }

constexpr size_t eig_ws_bytes(int n)
{
    const EigPlan p = eig_plan(n);
    return 2 * al256(sizeof(double) * (size_t)p.np * p.np) + al256(sizeof(double) * 2 * p.np) +
           al256(sizeof(double) * n) + al256(sizeof(int) * n) +
           al256(sizeof(int) * (EIG_MAX_SWEEPS + 4));
}

inline EigWs eig_ws(int n, void *base)
{
    const EigPlan p = eig_plan(n);
    char *c = (char *)base;
    EigWs w;
    w.W = (double *)c;
    c += al256(sizeof(double) * (size_t)p.np * p.np);
    w.V = (double *)c;
    c += al256(sizeof(double) * (size_t)p.np * p.np);
    w.nrm = (double *)c;  // W column norms, then V column norms
    c += al256(sizeof(double) * 2 * p.np);
    w.ev = (double *)c;
    c += al256(sizeof(double) * n);
    w.perm = (int *)c;
    c += al256(sizeof(int) * n);
    w.ints = (int *)c;
    return w;
}

}  // namespace irom

namespace {
thread_local std::string g_eerr;
int efail(int c, const std::string &m)
{
    g_eerr = m;
    return c;
}

#define IROM_ECHECK(call, what)                                                   \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) return efail(3, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <int B>
hipError_t launch_block_jacobi(const irom::EigArgs &a, int ncu, hipStream_t s)
{
    const size_t lds = sizeof(double) * 2 * B * (size_t)a.np;
    hipError_t e = hipFuncSetAttribute((const void *)&irom::block_jacobi_kernel<B>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    // every workgroup must be resident at once (grid barrier): at most one per CU here
    const int grid = std::max(1, std::min(a.nb / 2, ncu));
    void *args[] = {(void *)&a};
    return hipLaunchCooperativeKernel((const void *)&irom::block_jacobi_kernel<B>, dim3(grid),
                                      dim3(64 * B), args, lds, s);
}

// The whole decomposition on the workspace `ws` (eig_ws_bytes(n) bytes of device memory).
int eig_run(int n, const double *d_G, double *h_evals, double *d_evecs, double tol,
            int max_sweeps, int *sweeps, void *ws, hipStream_t s)
{
    if (max_sweeps < 1 || max_sweeps > irom::EIG_MAX_SWEEPS)
        return efail(1, "irom_eig_sym: max_sweeps must be in 1..64");
    const irom::EigPlan p = irom::eig_plan(n);
    const irom::EigWs w = irom::eig_ws(n, ws);
    int dev = 0, ncu = 0;
    IROM_ECHECK(hipGetDevice(&dev), "hipGetDevice");
    IROM_ECHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "attr");
    int *rot = w.ints, *bar = w.ints + irom::EIG_MAX_SWEEPS, *sw_d = bar + 2, *err_d = sw_d + 1;
    const long tot = std::max<long>((long)p.np * p.np, irom::EIG_MAX_SWEEPS + 4);
    irom::eig_init<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, p.np, d_G, w.W, w.V, w.ints,
                                                                 irom::EIG_MAX_SWEEPS + 4);
    IROM_ECHECK(hipGetLastError(), "eig_init");
    irom::EigArgs a{p.np, p.nb, tol, max_sweeps, w.W, w.V, rot, (unsigned *)bar, sw_d, err_d};
    hipError_t e;
    switch (p.B) {
    case 16: e = launch_block_jacobi<16>(a, ncu, s); break;
    case 8: e = launch_block_jacobi<8>(a, ncu, s); break;
    case 4: e = launch_block_jacobi<4>(a, ncu, s); break;
    case 2: e = launch_block_jacobi<2>(a, ncu, s); break;
    default: e = launch_block_jacobi<1>(a, ncu, s); break;
    }
    IROM_ECHECK(e, "block_jacobi launch");
    irom::column_norms<<<p.np, irom::ET, 0, s>>>(p.np, w.W, w.nrm);
    irom::column_norms<<<p.np, irom::ET, 0, s>>>(p.np, w.V, w.nrm + p.np);
    const size_t sort_lds = (sizeof(double) + sizeof(int)) * (size_t)p.m;
    irom::sort_desc<<<1, 1024, sort_lds, s>>>(n, p.m, w.nrm, w.ev, w.perm);
    irom::gather_columns<<<(unsigned)(((long)n * n + 255) / 256), 256, 0, s>>>(
        n, p.np, w.perm, w.V, w.nrm + p.np, d_evecs);
    IROM_ECHECK(hipGetLastError(), "eig finish");
    int info[2] = {0, 0};
    IROM_ECHECK(hipMemcpyAsync(h_evals, w.ev, sizeof(double) * n, hipMemcpyDeviceToHost, s), "copy");
    IROM_ECHECK(hipMemcpyAsync(info, sw_d, sizeof(int) * 2, hipMemcpyDeviceToHost, s), "copy");
    IROM_ECHECK(hipStreamSynchronize(s), "sync");
    if (sweeps) *sweeps = info[0];
    if (info[1]) return efail(3, "irom_eig_sym: grid barrier wait expired");
    return info[0] < max_sweeps ? 0 : efail(2, "irom_eig_sym: not converged");
}

// Allocation-free variant: a file-scope device array holds the workspace of the largest
// order it accepts.  Each GPU has its own copy at its own address, taken once per device
// with hipGetSymbolAddress (the array's name in host code is not a device address).
constexpr int EIG_NOALLOC_MAX_N = 512;
constexpr size_t EIG_SCRATCH_BYTES = irom::eig_ws_bytes(EIG_NOALLOC_MAX_N);
static __device__ __align__(256) unsigned char scratch[EIG_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void *g_scratch_ptr[64];

int scratch_address(int dev, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return efail(3, std::string("irom_eig_sym_noalloc: hipGetSymbolAddress: ") +
                                hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Workspace of irom_eig_sym, per device, grown on demand and kept until the process ends.
struct EigCache {
    void *p = nullptr;
    size_t bytes = 0;
};
EigCache g_eig_cache[64];

}  // namespace

extern "C" {

const char *irom_eig_last_error(void) { return g_eerr.c_str(); }

size_t irom_eig_workspace_size(int n)
{
    return (n > 0 && n <= 4096) ? irom::eig_ws_bytes(n) : 0;
}

int irom_eig_sym(int n, const double *d_G, double *h_evals, double *d_evecs, double tol,
                 int max_sweeps, int *sweeps, void *stream)
{
    if (n <= 0 || n > 4096 || !d_G || !h_evals || !d_evecs) return efail(1, "irom_eig_sym: bad argument");
    int dev = 0;
    IROM_ECHECK(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= 64) return efail(1, "irom_eig_sym: device index out of range");
    EigCache &c = g_eig_cache[dev];
    const size_t need = irom::eig_ws_bytes(n);
    if (c.bytes < need) {
        // the previous workspace may still be read by work queued on another stream
        IROM_ECHECK(hipDeviceSynchronize(), "sync");
        if (c.p) (void)hipFree(c.p);
        c.p = nullptr;
        c.bytes = 0;
        IROM_ECHECK(hipMalloc(&c.p, need), "hipMalloc(eig workspace)");
        c.bytes = need;
    }
    return eig_run(n, d_G, h_evals, d_evecs, tol, max_sweeps, sweeps, c.p, (hipStream_t)stream);
}

int irom_eig_sym_ws(int n, const double *d_G, double *h_evals, double *d_evecs, double tol,
                    int max_sweeps, int *sweeps, void *d_work, size_t work_bytes, void *stream)
{
    if (n <= 0 || n > 4096 || !d_G || !h_evals || !d_evecs || !d_work)
        return efail(1, "irom_eig_sym_ws: bad argument");
    if (work_bytes < irom::eig_ws_bytes(n))
        return efail(1, "irom_eig_sym_ws: workspace too small (see irom_eig_workspace_size)");
    return eig_run(n, d_G, h_evals, d_evecs, tol, max_sweeps, sweeps, d_work, (hipStream_t)stream);
}

int irom_eig_sym_noalloc(int n, const double *d_G, double *h_evals, double *d_evecs,
                         double tol, int max_sweeps, int *sweeps, void *stream)
{
    if (n <= 0 || !d_G || !h_evals || !d_evecs) return efail(1, "irom_eig_sym_noalloc: bad argument");
    if (n > EIG_NOALLOC_MAX_N)
        return efail(1, "irom_eig_sym_noalloc: n = " + std::to_string(n) +
                            " exceeds the static workspace (n <= 512)");
    int dev = 0;
    IROM_ECHECK(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= 64) return efail(1, "irom_eig_sym_noalloc: device index out of range");
    void *ws = nullptr;
    if (int st = scratch_address(dev, &ws)) return st;
    return eig_run(n, d_G, h_evals, d_evecs, tol, max_sweeps, sweeps, ws, (hipStream_t)stream);
}

}  // extern "C"
