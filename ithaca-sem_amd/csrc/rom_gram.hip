// Gram / projection GEMMs of the POD offline stage on MFMA: C = X^T Y with X (K x m) and
// Y (K x n) column-major, K very large (the local coefficient count), m, n <= a few hundred.
//
// The reference forms its POD basis with a dense thin SVD of the snapshot matrix
// (CoupledLinearNS_ROM.cpp:3925-3954, Eigen BDCSVD) and projects with dense products
// RB^T A RB (:2788-2802, :3661-3663).  Here the snapshot Gram matrix S^T (M S) and the
// Galerkin products V^T (A_q V) are split-K GEMMs: each workgroup owns a 128 x 128 block of
// C and a slice of K, stages 16-deep K panels of X and Y in LDS (padded against bank
// conflicts) and accumulates with MFMA; partial blocks are reduced in FP64 by a second
// kernel (deterministic, no atomics).  Precision modes:
//   FP64   v_mfma_f64_16x16x4f64, FP64 accumulation.
//   FP32   inputs split x = hi + lo in FP32, three v_mfma_f32_16x16x4f32 products
//          (hi*hi + hi*lo + lo*hi), FP32 accumulation within a 16-deep panel, FP64 across
//          panels.
//   BF16   inputs split x = hi + lo in BF16, three v_mfma_f32_16x16x32_bf16 products,
//          FP32 accumulation within a 32-deep panel, FP64 across panels.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/ithaca_rom.h"

namespace irom {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int GB = 128;   // C block edge per workgroup
constexpr int GT = 256;   // threads (4 waves, each a 64 x 64 quadrant)

// ---------------------------------------------------------------------------
// FP64: K panels of 16, X panel Xs[col][k] with row stride 17 (bank spread).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GT, 2) void gram_f64_kernel(int m, int n, long K, long kchunk,
                                                         const double *__restrict__ X, long ldx,
                                                         const double *__restrict__ Y, long ldy,
                                                         double *__restrict__ part, int symmetric)
{
    constexpr int KB = 16, KP = KB + 1;
    __shared__ double Xs[GB * KP], Ys[GB * KP];
    const int bi = blockIdx.x, bj = blockIdx.y, split = blockIdx.z;
    if (symmetric && bj < bi) return;
    const int i0 = bi * GB, j0 = bj * GB;
    const long kbeg = (long)split * kchunk, kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;  // quadrant of this wave

    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0, 0, 0, 0};

    for (long k0 = kbeg; k0 < kend; k0 += KB) {
        // stage: 128 columns x 16 k of X and of Y; thread -> (column, 8 consecutive k)
        {
            const int col = tid >> 1, kk = (tid & 1) * 8;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const long k = k0 + kk + t;
                const bool kin = k < kend;
                const int gi = i0 + col, gj = j0 + col;
                Xs[col * KP + kk + t] = (kin && gi < m) ? X[k + (long)gi * ldx] : 0.0;
                Ys[col * KP + kk + t] = (kin && gj < n) ? Y[k + (long)gj * ldy] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KB; ks += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = Xs[(wr + a * 16 + (lane & 15)) * KP + ks + (lane >> 4)];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Ys[(wc + b * 16 + (lane & 15)) * KP + ks + (lane >> 4)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    // partial block -> part[split][i][j] (row-major m x n).  The f64 16x16x4 accumulator
    // holds D[lane/16 + 4r][lane%16] in component r.
    double *P = part + (long)split * m * n;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wr + a * 16 + (lane >> 4) + 4 * r;
                const int j = j0 + wc + b * 16 + (lane & 15);
                if (i < m && j < n) P[(long)i * n + j] = acc[a][b][r];
            }
}

// ---------------------------------------------------------------------------
// FP32 split: x = hi + lo, three f32 MFMA products per k-step, FP64 accumulation per panel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GT, 2) void gram_f32x3_kernel(int m, int n, long K, long kchunk,
                                                           const double *__restrict__ X, long ldx,
                                                           const double *__restrict__ Y, long ldy,
                                                           double *__restrict__ part, int symmetric)
{
    constexpr int KB = 16, KP = KB + 1;
    __shared__ float Xh[GB * KP], Xl[GB * KP], Yh[GB * KP], Yl[GB * KP];
    const int bi = blockIdx.x, bj = blockIdx.y, split = blockIdx.z;
    if (symmetric && bj < bi) return;
    const int i0 = bi * GB, j0 = bj * GB;
    const long kbeg = (long)split * kchunk, kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;

    d4 tot[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) tot[a][b] = d4{0, 0, 0, 0};

    for (long k0 = kbeg; k0 < kend; k0 += KB) {
        {
            const int col = tid >> 1, kk = (tid & 1) * 8;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const long k = k0 + kk + t;
                const bool kin = k < kend;
                const int gi = i0 + col, gj = j0 + col;
                const double xv = (kin && gi < m) ? X[k + (long)gi * ldx] : 0.0;
                const double yv = (kin && gj < n) ? Y[k + (long)gj * ldy] : 0.0;
                const float xh = (float)xv, yh = (float)yv;
                Xh[col * KP + kk + t] = xh;
                Xl[col * KP + kk + t] = (float)(xv - (double)xh);
                Yh[col * KP + kk + t] = yh;
                Yl[col * KP + kk + t] = (float)(yv - (double)yh);
            }
        }
        __syncthreads();
        f4 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KB; ks += 4) {
            float ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int o = (wr + a * 16 + (lane & 15)) * KP + ks + (lane >> 4);
                ah[a] = Xh[o];
                al[a] = Xl[o];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int o = (wc + b * 16 + (lane & 15)) * KP + ks + (lane >> 4);
                bh[b] = Yh[o];
                bl[b] = Yl[o];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bh[b], acc[a][b], 0, 0, 0);
                }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) tot[a][b][r] += (double)acc[a][b][r];
        __syncthreads();
    }
    double *P = part + (long)split * m * n;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wr + a * 16 + 4 * (lane >> 4) + r;
                const int j = j0 + wc + b * 16 + (lane & 15);
                if (i < m && j < n) P[(long)i * n + j] = tot[a][b][r];
            }
}

// ---------------------------------------------------------------------------
// BF16 split: x = hi + lo in BF16, three v_mfma_f32_16x16x32_bf16 per 32-deep step.
// LDS holds each column's 32 k-values as bf16 (row stride 40 elements = 80 bytes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ __bf16 to_bf16(float v) { return __float2bfloat16(v); }
__device__ __forceinline__ float from_bf16(__bf16 v) { return __bfloat162float(v); }

__global__ __launch_bounds__(GT, 2) void gram_bf16x3_kernel(int m, int n, long K, long kchunk,
                                                            const double *__restrict__ X, long ldx,
                                                            const double *__restrict__ Y, long ldy,
                                                            double *__restrict__ part, int symmetric)
{
    constexpr int KB = 32, KP = 40;
    __shared__ __attribute__((aligned(16))) __bf16 Xh[GB * KP], Xl[GB * KP], Yh[GB * KP], Yl[GB * KP];
    const int bi = blockIdx.x, bj = blockIdx.y, split = blockIdx.z;
    if (symmetric && bj < bi) return;
    const int i0 = bi * GB, j0 = bj * GB;
    const long kbeg = (long)split * kchunk, kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;

    d4 tot[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) tot[a][b] = d4{0, 0, 0, 0};

    for (long k0 = kbeg; k0 < kend; k0 += KB) {
        {
            const int col = tid >> 1, kk = (tid & 1) * 16;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const long k = k0 + kk + t;
                const bool kin = k < kend;
                const int gi = i0 + col, gj = j0 + col;
                const double xv = (kin && gi < m) ? X[k + (long)gi * ldx] : 0.0;
                const double yv = (kin && gj < n) ? Y[k + (long)gj * ldy] : 0.0;
                const __bf16 xh = to_bf16((float)xv), yh = to_bf16((float)yv);
                Xh[col * KP + kk + t] = xh;
                Xl[col * KP + kk + t] = to_bf16((float)(xv - (double)from_bf16(xh)));
                Yh[col * KP + kk + t] = yh;
                Yl[col * KP + kk + t] = to_bf16((float)(yv - (double)from_bf16(yh)));
            }
        }
        __syncthreads();
        f4 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
        bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int o = (wr + a * 16 + (lane & 15)) * KP + 8 * (lane >> 4);  // 16-byte aligned
            ah[a] = *reinterpret_cast<const bf8 *>(&Xh[o]);
            al[a] = *reinterpret_cast<const bf8 *>(&Xl[o]);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int o = (wc + b * 16 + (lane & 15)) * KP + 8 * (lane >> 4);
            bh[b] = *reinterpret_cast<const bf8 *>(&Yh[o]);
            bl[b] = *reinterpret_cast<const bf8 *>(&Yl[o]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
            }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) tot[a][b][r] += (double)acc[a][b][r];
        __syncthreads();
    }
    double *P = part + (long)split * m * n;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wr + a * 16 + 4 * (lane >> 4) + r;
                const int j = j0 + wc + b * 16 + (lane & 15);
                if (i < m && j < n) P[(long)i * n + j] = tot[a][b][r];
            }
}

// C (row-major m x n, ldc) (+)= sum over splits of the partial blocks; for symmetric
// products the lower triangle is mirrored from the computed upper blocks.
__global__ void gram_reduce_kernel(int m, int n, int nsplit, const double *__restrict__ part,
                                   double *__restrict__ C, long ldc, int accumulate, int symmetric)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)m * n) return;
    int i = (int)(g / n), j = (int)(g - (long)i * n);
    int si = i, sj = j;
    if (symmetric && (j / GB) < (i / GB)) {
        si = j;
        sj = i;
    }
    double s = 0.0;
    for (int p = 0; p < nsplit; ++p) s += part[(long)p * m * n + (long)si * n + sj];
    if (accumulate) C[(long)i * ldc + j] += s;
    else C[(long)i * ldc + j] = s;
}

}  // namespace irom

namespace {
thread_local std::string g_gerr;
}

extern "C" {

const char *irom_gram_last_error(void) { return g_gerr.c_str(); }

size_t irom_gram_workspace_size(int m, int n, long K)
{
    const long nsplit = std::max<long>(1, std::min<long>(512, (K + (1 << 16) - 1) >> 16));
    return (size_t)nsplit * m * n * sizeof(double);
}

int irom_gram(int mode, int m, int n, long K, const double *d_X, long ldx, const double *d_Y,
              long ldy, double *d_C, long ldc, int accumulate, void *d_work, size_t work_bytes,
              void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (m <= 0 || n <= 0 || K <= 0 || !d_X || !d_Y || !d_C || !d_work) {
        g_gerr = "irom_gram: bad argument";
        return 1;
    }
    long nsplit = std::max<long>(1, std::min<long>(512, (K + (1 << 16) - 1) >> 16));
    if ((size_t)nsplit * m * n * sizeof(double) > work_bytes) {
        g_gerr = "irom_gram: workspace too small (see irom_gram_workspace_size)";
        return 1;
    }
    long kchunk = (K + nsplit - 1) / nsplit;
    kchunk = (kchunk + 31) / 32 * 32;
    nsplit = (K + kchunk - 1) / kchunk;
    const int symmetric = (d_X == d_Y && m == n && ldx == ldy) ? 1 : 0;
    dim3 grid((m + irom::GB - 1) / irom::GB, (n + irom::GB - 1) / irom::GB, (unsigned)nsplit);
    double *part = (double *)d_work;
    if (mode == IROM_GRAM_FP64)
        irom::gram_f64_kernel<<<grid, irom::GT, 0, s>>>(m, n, K, kchunk, d_X, ldx, d_Y, ldy, part, symmetric);
    else if (mode == IROM_GRAM_FP32X3)
        irom::gram_f32x3_kernel<<<grid, irom::GT, 0, s>>>(m, n, K, kchunk, d_X, ldx, d_Y, ldy, part, symmetric);
    else if (mode == IROM_GRAM_BF16X3)
        irom::gram_bf16x3_kernel<<<grid, irom::GT, 0, s>>>(m, n, K, kchunk, d_X, ldx, d_Y, ldy, part, symmetric);
    else {
        g_gerr = "irom_gram: unknown mode";
        return 1;
    }
    const long tot = (long)m * n;
    irom::gram_reduce_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(m, n, (int)nsplit, part, d_C,
                                                                         ldc, accumulate, symmetric);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_gerr = std::string("irom_gram: ") + hipGetErrorString(e);
        return 3;
    }
    return 0;
}

}  // extern "C"
