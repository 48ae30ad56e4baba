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

struct GramArgs {
    int m, n;
    long K, kchunk;
    const double *X;
    long ldx;
    const double *Y;
    long ldy;
    double *part;
    int nbj;      // C blocks along n
    int ntiles;   // blocks computed per K slice
    int nsplit;   // K slices
    int sym;      // only blocks bi <= bj (the caller's product is symmetric)
    int vec;      // X, Y rows 16-byte aligned: 16-byte loads
};

// Block id -> (K slice, C block).  Workgroups are dispatched to the 8 XCDs round-robin by
// linear id, so ids L with the same L % 8 share an L2: every C block of one K slice is
// given the same L % 8 and adjacent dispatch order, and the slice's X/Y panels are read
// from HBM once per XCD instead of once per C block.
__device__ __forceinline__ bool gram_block(const GramArgs &g, int &split, int &bi, int &bj)
{
    const int L = blockIdx.x, xcd = L & 7, r = L >> 3;
    split = 8 * (r / g.ntiles) + xcd;
    int t = r % g.ntiles;
    if (split >= g.nsplit) return false;
    if (!g.sym) {
        bi = t / g.nbj;
        bj = t - bi * g.nbj;
        return true;
    }
    for (bi = 0; t >= g.nbj - bi; ++bi) t -= g.nbj - bi;  // upper blocks, row by row
    bj = bi + t;
    return true;
}

// Per-mode LDS panel formats, MFMA steps and workgroup shapes.
//   MODE 0 (FP64):   KB = 16, double panels, v_mfma_f64_16x16x4f64; 4 waves, each a
//                    64 x 64 quadrant of the C block, two workgroups per CU
//   MODE 1 (FP32x3): KB = 16, hi/lo float panels, 3 x v_mfma_f32_16x16x4f32
//   MODE 2 (BF16x3): KB = 32, hi/lo bf16 panels, 3 x v_mfma_f32_16x16x32_bf16
// The split modes convert every panel element to two parts (VALU work between the MFMA
// steps) and keep FP32 panel sums next to FP64 block sums, so they run 8 waves, each a
// 32 x 64 strip: two waves per SIMD, one converting while the other multiplies.
template <int MODE> struct GramMode;
template <> struct GramMode<0> { static constexpr int KB = 16, KP = 17, NW = 4; using E = double; };
template <> struct GramMode<1> { static constexpr int KB = 16, KP = 17, NW = 8; using E = float; };
template <> struct GramMode<2> { static constexpr int KB = 32, KP = 40, NW = 8; using E = __bf16; };

__device__ __forceinline__ __bf16 to_bf16(float v) { return __float2bfloat16(v); }
__device__ __forceinline__ float from_bf16(__bf16 v) { return __bfloat162float(v); }

// Panel row of the q-th group of KB/2 consecutive threads.  For the BF16 panels (KB = 32,
// 80-byte LDS rows, 16 threads per row) the 4 rows one wave stores are spread 4 apart,
// which puts them in 4 disjoint 16-bank ranges (consecutive rows would overlap the first
// and the fourth); loads stay whole-row and coalesced.
template <int KB>
__device__ __forceinline__ int panel_row(int q)
{
    if constexpr (KB == 32) return (q & ~15) | ((q & 3) << 2) | ((q >> 2) & 3);
    return q;
}

// One KB-deep panel of GB rows of X (column-major K x m: row i of the panel is column i0+i
// of X): NL = KB*GB/(2T) pairs of doubles per thread of a T-thread workgroup, consecutive
// threads on consecutive k of one column so that a wave reads whole 128-byte lines.
template <int KB, int T>
__device__ __forceinline__ void panel_load(const double *__restrict__ X, long ldx, int i0, int m,
                                           long k0, long kend, int vec,
                                           double2 (&v)[KB * GB / (2 * T)])
{
    constexpr int NL = KB * GB / (2 * T);
    // interior panels (the common case) take one uniform branch of unconditional loads
    if (vec && i0 + GB <= m && k0 + KB <= kend) {
#pragma unroll
        for (int s = 0; s < NL; ++s) {
            const int idx = s * T + threadIdx.x;
            const int c = panel_row<KB>(idx / (KB / 2)), ko = 2 * (idx % (KB / 2));
            v[s] = *reinterpret_cast<const double2 *>(X + (long)(i0 + c) * ldx + k0 + ko);
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < NL; ++s) {
        const int idx = s * T + threadIdx.x;
        const int c = panel_row<KB>(idx / (KB / 2)), ko = 2 * (idx % (KB / 2));
        const int gi = i0 + c;
        const long k = k0 + ko;
        const double *p = X + (long)gi * ldx + k;
        v[s].x = (gi < m && k < kend) ? p[0] : 0.0;
        v[s].y = (gi < m && k + 1 < kend) ? p[1] : 0.0;
    }
}

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

template <int MODE, int KB, int KP, int T, typename E>
__device__ __forceinline__ void panel_store(const double2 (&v)[KB * GB / (2 * T)], E *hi, E *lo)
{
    constexpr int NL = KB * GB / (2 * T);
#pragma unroll
    for (int s = 0; s < NL; ++s) {
        const int idx = s * T + threadIdx.x;
        const int c = panel_row<KB>(idx / (KB / 2)), ko = 2 * (idx % (KB / 2));
        const double x[2] = {v[s].x, v[s].y};
        if constexpr (MODE == 2) {
            // x ~ hi + lo: hi = the leading 8 bits of x rounded to FP32 (half-up on the bit
            // pattern: an add and a mask, no correctly rounded FP64 -> BF16 sequence, which
            // hi does not need), lo = bf16(xf - hi), the difference exact in FP32; each pair
            // is stored as one 32-bit word (KP and ko even)
            const float x0 = (float)x[0], x1 = (float)x[1];
            const unsigned b0 = (__float_as_uint(x0) + 0x8000u) & 0xffff0000u;
            const unsigned b1 = (__float_as_uint(x1) + 0x8000u) & 0xffff0000u;
            const int o = c * KP + ko;
            *reinterpret_cast<unsigned *>(&hi[o]) = (b0 >> 16) | b1;
            *reinterpret_cast<bf2 *>(&lo[o]) =
                bf2{to_bf16(x0 - __uint_as_float(b0)), to_bf16(x1 - __uint_as_float(b1))};
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int o = c * KP + ko + h;
                if constexpr (MODE == 0) {
                    hi[o] = x[h];
                } else {
                    const float xh = (float)x[h];
                    hi[o] = xh;
                    lo[o] = (float)(x[h] - (double)xh);
                }
            }
        }
    }
}

// C block (bi, bj) of X^T Y over one K slice, written to part[split] (row-major m x n).
// The next panel is loaded into registers while the current one is multiplied.  Wave w
// owns rows (w/2)*WR .. +WR and columns (w%2)*64 .. +64 of the block (TA x 4 MFMA tiles).
template <int MODE>
__global__ __launch_bounds__(64 * GramMode<MODE>::NW, 2) void gram_kernel(GramArgs g)
{
    using Md = GramMode<MODE>;
    using E = typename Md::E;
    constexpr int KB = Md::KB, KP = Md::KP, T = 64 * Md::NW;
    constexpr int WR = GB / (Md::NW / 2), TA = WR / 16, NL = KB * GB / (2 * T);
    constexpr int NA = MODE == 0 ? 1 : 2;  // hi (and lo) panels
    // FP32x3: two LDS panel buffers, the next panel converted and stored while the current
    // one is multiplied, one barrier per panel (measured faster there, slower for the
    // other modes, which keep one buffer and two barriers per panel)
    constexpr int NBUF = MODE == 1 ? 2 : 1;
    __shared__ __attribute__((aligned(16))) E Xs[NBUF][NA][GB * KP], Ys[NBUF][NA][GB * KP];
    int split, bi, bj;
    if (!gram_block(g, split, bi, bj)) return;
    const int i0 = bi * GB, j0 = bj * GB;
    const long kbeg = (long)split * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * WR, wc = (wave & 1) * 64;  // this wave's strip

    d4 tot[TA][4];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) tot[a][b] = d4{0, 0, 0, 0};

    auto multiply = [&](const E *xh, const E *xl, const E *yh, const E *yl) {
        if constexpr (MODE == 0) {
#pragma unroll
            for (int ks = 0; ks < KB; ks += 4) {
                double av[TA], bv[4];
#pragma unroll
                for (int a = 0; a < TA; ++a)
                    av[a] = xh[(wr + a * 16 + (lane & 15)) * KP + ks + (lane >> 4)];
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    bv[b] = yh[(wc + b * 16 + (lane & 15)) * KP + ks + (lane >> 4)];
#pragma unroll
                for (int a = 0; a < TA; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        tot[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], tot[a][b],
                                                                         0, 0, 0);
            }
        } else {
            f4 acc[TA][4];
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
            if constexpr (MODE == 1) {
#pragma unroll
                for (int ks = 0; ks < KB; ks += 4) {
                    float ah[TA], al[TA], bh[4], bl[4];
#pragma unroll
                    for (int a = 0; a < TA; ++a) {
                        const int o = (wr + a * 16 + (lane & 15)) * KP + ks + (lane >> 4);
                        ah[a] = xh[o];
                        al[a] = xl[o];
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int o = (wc + b * 16 + (lane & 15)) * KP + ks + (lane >> 4);
                        bh[b] = yh[o];
                        bl[b] = yl[o];
                    }
#pragma unroll
                    for (int a = 0; a < TA; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[a], bh[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bl[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bh[b], acc[a][b], 0, 0, 0);
                        }
                }
            } else {
                bf8 ah[TA], al[TA], bh[4], bl[4];
#pragma unroll
                for (int a = 0; a < TA; ++a) {
                    const int o = (wr + a * 16 + (lane & 15)) * KP + 8 * (lane >> 4);
                    ah[a] = *reinterpret_cast<const bf8 *>(&xh[o]);
                    al[a] = *reinterpret_cast<const bf8 *>(&xl[o]);
                }
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int o = (wc + b * 16 + (lane & 15)) * KP + 8 * (lane >> 4);
                    bh[b] = *reinterpret_cast<const bf8 *>(&yh[o]);
                    bl[b] = *reinterpret_cast<const bf8 *>(&yl[o]);
                }
#pragma unroll
                for (int a = 0; a < TA; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int a = 0; a < TA; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) tot[a][b][r] += (double)acc[a][b][r];
        }
    };

    double2 xv[NL], yv[NL];
    panel_load<KB, T>(g.X, g.ldx, i0, g.m, kbeg, kend, g.vec, xv);
    panel_load<KB, T>(g.Y, g.ldy, j0, g.n, kbeg, kend, g.vec, yv);
    if constexpr (NBUF == 2) {
        panel_store<MODE, KB, KP, T, E>(xv, Xs[0][0], Xs[0][NA - 1]);
        panel_store<MODE, KB, KP, T, E>(yv, Ys[0][0], Ys[0][NA - 1]);
        if (kbeg + KB < kend) {
            panel_load<KB, T>(g.X, g.ldx, i0, g.m, kbeg + KB, kend, g.vec, xv);
            panel_load<KB, T>(g.Y, g.ldy, j0, g.n, kbeg + KB, kend, g.vec, yv);
        }
        __syncthreads();
        int cur = 0;
        for (long k0 = kbeg; k0 < kend; k0 += KB, cur ^= 1) {
            if (k0 + KB < kend) {  // the other buffer was last read before the last barrier
                panel_store<MODE, KB, KP, T, E>(xv, Xs[cur ^ 1][0], Xs[cur ^ 1][NA - 1]);
                panel_store<MODE, KB, KP, T, E>(yv, Ys[cur ^ 1][0], Ys[cur ^ 1][NA - 1]);
                if (k0 + 2 * KB < kend) {
                    panel_load<KB, T>(g.X, g.ldx, i0, g.m, k0 + 2 * KB, kend, g.vec, xv);
                    panel_load<KB, T>(g.Y, g.ldy, j0, g.n, k0 + 2 * KB, kend, g.vec, yv);
                }
            }
            multiply(Xs[cur][0], Xs[cur][NA - 1], Ys[cur][0], Ys[cur][NA - 1]);
            __syncthreads();
        }
    } else {
        for (long k0 = kbeg; k0 < kend; k0 += KB) {
            panel_store<MODE, KB, KP, T, E>(xv, Xs[0][0], Xs[0][NA - 1]);
            panel_store<MODE, KB, KP, T, E>(yv, Ys[0][0], Ys[0][NA - 1]);
            __syncthreads();
            if (k0 + KB < kend) {
                panel_load<KB, T>(g.X, g.ldx, i0, g.m, k0 + KB, kend, g.vec, xv);
                panel_load<KB, T>(g.Y, g.ldy, j0, g.n, k0 + KB, kend, g.vec, yv);
            }
            multiply(Xs[0][0], Xs[0][NA - 1], Ys[0][0], Ys[0][NA - 1]);
            __syncthreads();
        }
    }
    // The f64 16x16x4 accumulator holds D[lane/16 + 4r][lane%16] in component r; the f32
    // 16x16x4 and bf16 16x16x32 ones hold D[4*(lane/16) + r][lane%16].
    double *P = g.part + (long)split * g.m * g.n;
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wr + a * 16 + (MODE == 0 ? (lane >> 4) + 4 * r
                                                            : 4 * (lane >> 4) + r);
                const int j = j0 + wc + b * 16 + (lane & 15);
                if (i < g.m && j < g.n) P[(long)i * g.n + j] = tot[a][b][r];
            }
}

// C (row-major m x n, ldc) (+)= sum over splits of the partial blocks; for symmetric
// products every entry below the diagonal is taken from its mirror above it (computed in
// an upper or diagonal block), so the result is exactly symmetric.
__global__ void gram_reduce_kernel(int m, int n, int nsplit, const double *__restrict__ part,
                                   double *__restrict__ C, long ldc, int accumulate, int symmetric)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)m * n) return;
    int i = (int)(g / n), j = (int)(g - (long)i * n);
    int si = i, sj = j;
    if (symmetric && j < i) {
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
    const int base = mode & 0xf;
    const int symmetric = (m == n && ((d_X == d_Y && ldx == ldy) ||
                                      (mode & IROM_GRAM_SYMMETRIC_RESULT))) ? 1 : 0;
    irom::GramArgs ga;
    ga.m = m;
    ga.n = n;
    ga.K = K;
    ga.kchunk = kchunk;
    ga.X = d_X;
    ga.ldx = ldx;
    ga.Y = d_Y;
    ga.ldy = ldy;
    ga.part = (double *)d_work;
    const int nbi = (m + irom::GB - 1) / irom::GB;
    ga.nbj = (n + irom::GB - 1) / irom::GB;
    ga.ntiles = symmetric ? nbi * (nbi + 1) / 2 : nbi * ga.nbj;
    ga.nsplit = (int)nsplit;
    ga.sym = symmetric;
    ga.vec = (ldx % 2 == 0 && ldy % 2 == 0 && (uintptr_t)d_X % 16 == 0 &&
              (uintptr_t)d_Y % 16 == 0) ? 1 : 0;
    const dim3 grid((unsigned)(8L * ga.ntiles * ((nsplit + 7) / 8)));
    double *part = (double *)d_work;
    if (base == IROM_GRAM_FP64)
        irom::gram_kernel<0><<<grid, 64 * irom::GramMode<0>::NW, 0, s>>>(ga);
    else if (base == IROM_GRAM_FP32X3)
        irom::gram_kernel<1><<<grid, 64 * irom::GramMode<1>::NW, 0, s>>>(ga);
    else if (base == IROM_GRAM_BF16X3)
        irom::gram_kernel<2><<<grid, 64 * irom::GramMode<2>::NW, 0, s>>>(ga);
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
