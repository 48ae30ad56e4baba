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
    // FP64 fallback of the mixed method.  MIXED kernel: instances whose refinement does
    // not converge are appended to fail_list (count in *fail_count) when fail_list is set.
    // FP64 kernel: with fail_count set, it solves fail_list[0 .. *fail_count) instead of
    // the batch order and reports status 3 (solved by the fallback) or 1 (zero pivot).
    int *fail_list;
    int *fail_count;
};

// Value of v in lane l of this wave (l uniform).
__device__ __forceinline__ float read_lane(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double read_lane(double v, int l)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// v of another lane of the same 16-lane row, by a DPP lane pattern (every lane active).
template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Reduction over the 64 lanes of a wave (all active), result in every lane: within each
// row of 16 lanes by DPP exchanges (lane xor 1, xor 2, half-row mirror, row mirror: each
// step combines two disjoint groups, so every lane ends with its row's total), then the
// four row totals read as scalars.  No LDS round trip, unlike a shuffle.
template <typename V, typename Op>
__device__ __forceinline__ V wave_reduce(V v, Op op)
{
    v = op(v, dpp<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, dpp<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, dpp<0x141>(v));  // row_half_mirror
    v = op(v, dpp<0x140>(v));  // row_mirror
    return op(op(read_lane(v, 0), read_lane(v, 16)), op(read_lane(v, 32), read_lane(v, 48)));
}
template <typename V>
__device__ __forceinline__ V wave_max(V v)
{
    return wave_reduce(v, [](V x, V y) { return x > y ? x : y; });
}
template <typename V>
__device__ __forceinline__ V wave_sum(V v)
{
    return wave_reduce(v, [](V x, V y) { return x + y; });
}

__device__ __forceinline__ double block_max(double v, double *sv)
{
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sv[0];
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) r = fmax(r, sv[k]);
    __syncthreads();
    return r;
}

// LDS column stride of the panel and of U12: reads of a 4-wide register tile may run up
// to 3 entries past the last valid row / column (their products are discarded).
constexpr int PLD = NMAX + 4;

// Per-workgroup LDS bookkeeping of the row interchanges.
struct PivLds {
    int posP[NMAX];    // panel-start row whose content is at row r (rows of the panel step)
    int posAll[NMAX];  // original row whose content is at row r (the final permutation)
    int list[NMAX];    // rows below the panel's top block that receive a top row
    int cnt;
    double wv[2][NT / 64];  // per-wave pivot candidates (double-buffered by column parity)
    int wi[2][NT / 64];     // their rows
    int wo[2][NT / 64];     // panel-start rows of their contents
    int corg[2];            // panel-start row of row c's content
};

// In-place LU with partial pivoting of the n x n row-major matrix L (ld n), right-looking
// with panels of NB columns.  On return L holds the unit-lower and upper factors of
// P A = L U with the rows permuted, pv.posAll[i] = row of A that became row i, and *zero
// is set when a pivot is zero or NaN.  Per panel:
//   1. the m x nb panel is copied to LDS (column-major, P) and factored by wave 0 alone:
//      lane-per-row pivot search reduced by shuffles, interchanges and rank-1 updates
//      ordered by wave barriers only; lane 0 records which rows the interchanges touch;
//   2. every other column applies the panel's interchanges as one gather (all loads of
//      a column issued before its stores); the columns right of the panel then solve
//      U12 = L11^{-1} A12 in registers and write U12 to global and to LDS (U);
//   3. A22 -= L21 U12 in TM x TN register tiles, both operands read from LDS, the next
//      tile's global values loaded while the current one is computed.
template <typename T, int NB, int TM, int TN>
__device__ __forceinline__ void block_lu(T *__restrict__ L, int n, T *P, T *U, T *R, PivLds &pv, int *zero)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < n; i += NT) pv.posAll[i] = i;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = min(NB, n - k0), m = n - k0;
        // 1. panel to LDS (loads issued 8 at a time per thread)
        for (int base = tid; base < m * nb; base += 8 * NT) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = base + u * NT;
                const int r = idx / nb, c = idx - r * nb;
                v[u] = idx < m * nb ? L[(long)(k0 + r) * n + k0 + c] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = base + u * NT;
                const int r = idx / nb, c = idx - r * nb;
                if (idx < m * nb) P[c * PLD + r] = v[u];
            }
        }
        __syncthreads();
        // Panel factorisation by the whole workgroup: thread r owns panel row r (m <= NT)
        // in registers, with the panel-start row its content came from (org).  Per column,
        // one barrier: each wave's best candidate lane publishes its row, value and index,
        // and the owner of row c publishes row c (LDS scratch R, double-buffered by column
        // parity); after the barrier every thread picks the block winner, row c and the
        // pivot row are exchanged, and every row below c is updated with the pivot row.
        {
            constexpr int NW = NT / 64;
            const int r = tid;
            const bool own = r < m;
            T v[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) v[c] = (own && c < nb) ? P[c * PLD + r] : T(0);
            int org = k0 + r;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < nb) {
                    const int par = c & 1;
                    T *wrow = R + par * (NW + 1) * NB;  // NW candidate rows, then row c
                    T a = (own && r >= c) ? fabs(v[c]) : T(-1);
                    if (a != a) a = T(-1);  // NaN: never a pivot
                    const T wmax = wave_max(a);
                    const unsigned long long hit = __ballot(a == wmax);
                    const int wl = __builtin_ffsll((long long)hit) - 1;  // lowest lane
                    if (lane == wl) {
#pragma unroll
                        for (int cc = 0; cc < NB; ++cc) wrow[wave * NB + cc] = v[cc];
                        pv.wv[par][wave] = wmax;
                        pv.wi[par][wave] = r;
                        pv.wo[par][wave] = org;
                    }
                    if (r == c) {
#pragma unroll
                        for (int cc = 0; cc < NB; ++cc) wrow[NW * NB + cc] = v[cc];
                        pv.corg[par] = org;
                    }
                    __syncthreads();
                    double best = pv.wv[par][0];
                    int p = pv.wi[par][0], pw = 0;
#pragma unroll
                    for (int w = 1; w < NW; ++w) {
                        const double bw = pv.wv[par][w];
                        const int iw = pv.wi[par][w];
                        if (bw > best || (bw == best && iw < p)) {
                            best = bw;
                            p = iw;
                            pw = w;
                        }
                    }
                    if (tid == 0 && !(best > 0.0)) *zero = 1;
                    const T *prow = wrow + pw * NB;
                    if (best < 0.0) {  // no usable pivot: keep row c
                        p = c;
                        prow = wrow + NW * NB;
                    }
                    if (p != c) {
                        if (r == c) {
#pragma unroll
                            for (int cc = 0; cc < NB; ++cc) v[cc] = prow[cc];
                            org = pv.wo[par][pw];
                        } else if (r == p) {
#pragma unroll
                            for (int cc = 0; cc < NB; ++cc) v[cc] = wrow[NW * NB + cc];
                            org = pv.corg[par];
                        }
                    }
                    if (own && r > c) {
                        const T l = v[c] / prow[c];
                        v[c] = l;
#pragma unroll
                        for (int cc = c + 1; cc < NB; ++cc) v[cc] -= l * prow[cc];
                    }
                }
            }
            // factored panel back to LDS; posP from org; the final permutation advances by
            // this panel's; rows below the top block holding a top row are listed.
            int prev = 0;
            if (own) {
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < nb) P[c * PLD + r] = v[c];
                pv.posP[k0 + r] = org;
                prev = pv.posAll[org];
            }
            if (tid == 0) pv.cnt = 0;
            __syncthreads();
            if (own) {
                pv.posAll[k0 + r] = prev;
                if (r >= nb && org != k0 + r) pv.list[atomicAdd(&pv.cnt, 1)] = k0 + r;
            }
            __syncthreads();
        }
        // panel back to global
        for (int idx = tid; idx < m * nb; idx += NT) {
            const int r = idx / nb, c = idx - r * nb;
            L[(long)(k0 + r) * n + k0 + c] = P[c * PLD + r];
        }
        // 2. the other columns: q < n2 -> column k0+nb+q (right), else column q-n2 (left)
        const int n2 = n - k0 - nb, cnt = pv.cnt;
        for (int q = tid; q < n - nb; q += NT) {
            const bool right = q < n2;
            const int j = right ? k0 + nb + q : q - n2;
            // top rows' new values: from below the top block now, from LDS after the stage
            T fin[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int src = c < nb ? pv.posP[k0 + c] : 0;
                fin[c] = (c < nb && src >= k0 + nb) ? L[(long)src * n + j] : T(0);
            }
            // top rows' old values, staged in this thread's LDS column of U
            T old[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) old[c] = c < nb ? L[(long)(k0 + c) * n + j] : T(0);
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < nb) U[c * PLD + q] = old[c];
            // rows below the top block that receive a top row (their old values are read
            // above into fin where needed, so they may be overwritten now)
            for (int e = 0; e < cnt; ++e) {
                const int r = pv.list[e];
                L[(long)r * n + j] = U[(pv.posP[r] - k0) * PLD + q];
            }
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int src = c < nb ? pv.posP[k0 + c] : 0;
                if (c < nb && src < k0 + nb) fin[c] = U[(src - k0) * PLD + q];
            }
            if (right) {
                // U12 = L11^{-1} A12 (unit lower, L11[c][cp] = P[cp*PLD + c]) in this
                // thread's LDS column
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < nb) U[c * PLD + q] = fin[c];
                for (int c = 1; c < nb; ++c) {
                    T s = U[c * PLD + q];
                    for (int cp = 0; cp < c; ++cp) s -= P[cp * PLD + c] * U[cp * PLD + q];
                    U[c * PLD + q] = s;
                }
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < nb) fin[c] = U[c * PLD + q];
            }
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < nb) L[(long)(k0 + c) * n + j] = fin[c];
        }
        __syncthreads();
        // 3. A22 -= L21 U12
        if (n2 > 0) {
            const int tr_n = (n2 + TM - 1) / TM, tc_n = (n2 + TN - 1) / TN;
            const int ntile = tr_n * tc_n;
            auto load_tile = [&](int t, T (&dst)[TM][TN]) {
                const int tr = t / tc_n, tc = t - tr * tc_n;
                const int i0 = tr * TM, j0 = tc * TN;
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq)
                        dst[a][bq] = (t < ntile && i0 + a < n2 && j0 + bq < n2)
                                         ? L[(long)(k0 + nb + i0 + a) * n + k0 + nb + j0 + bq]
                                         : T(0);
            };
            T nxt[TM][TN];
            load_tile(tid, nxt);
            for (int t = tid; t < ntile; t += NT) {
                T acc[TM][TN];
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq) acc[a][bq] = nxt[a][bq];
                load_tile(t + NT, nxt);
                const int tr = t / tc_n, tc = t - tr * tc_n;
                const int i0 = tr * TM, j0 = tc * TN;
                for (int c = 0; c < nb; ++c) {
                    T av[TM], bv[TN];
#pragma unroll
                    for (int a = 0; a < TM; ++a) av[a] = P[c * PLD + nb + i0 + a];
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq) bv[bq] = U[c * PLD + j0 + bq];
#pragma unroll
                    for (int a = 0; a < TM; ++a)
#pragma unroll
                        for (int bq = 0; bq < TN; ++bq) acc[a][bq] -= av[a] * bv[bq];
                }
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int bq = 0; bq < TN; ++bq)
                        if (i0 + a < n2 && j0 + bq < n2)
                            L[(long)(k0 + nb + i0 + a) * n + k0 + nb + j0 + bq] = acc[a][bq];
            }
        }
        __syncthreads();
    }
}

// Solve L U y = y in place (y in LDS, already permuted), factors row-major in L (ld n).
// Thread i owns row i (n <= NT): it streams its row of the factor in chunks of SB columns,
// the next chunk loaded while the current one is used.  The SB x SB diagonal block of a
// chunk lies in one wave and is solved there with lane reads; one barrier per chunk
// publishes the solved entries (ys) to the rows that still need them.
template <typename T>
__device__ __forceinline__ void lu_solve_rows(const T *__restrict__ L, int n, T *y, T *ys)
{
    constexpr int SB = 16;
    const int i = threadIdx.x, w = i >> 6;
    const bool act = i < n;
    T yi = act ? y[i] : T(0);
    T cur[SB], nxt[SB];
    // forward, unit lower: row i uses columns < i
    auto load_lo = [&](int kb, T (&dst)[SB]) {
#pragma unroll
        for (int c = 0; c < SB; ++c)
            dst[c] = (act && kb + c < i) ? L[(long)i * n + kb + c] : T(0);
    };
    load_lo(0, nxt);
    for (int kb = 0; kb < n; kb += SB) {
#pragma unroll
        for (int c = 0; c < SB; ++c) cur[c] = nxt[c];
        load_lo(kb + SB, nxt);
        if (w == kb / 64) {
#pragma unroll
            for (int c = 0; c < SB; ++c) {
                const int r = kb + c;
                if (r < n) {
                    const T yc = read_lane(yi, r & 63);
                    if (i > r && i < kb + SB) yi -= cur[c] * yc;
                }
            }
        }
        if (act && i >= kb && i < kb + SB) ys[i] = yi;
        __syncthreads();
        if (act && i >= kb + SB) {
#pragma unroll
            for (int c = 0; c < SB; ++c) yi -= cur[c] * ys[kb + c];
        }
    }
    // backward, upper with diagonal: row i uses columns >= i
    auto load_up = [&](int kb, T (&dst)[SB]) {
#pragma unroll
        for (int c = 0; c < SB; ++c)
            dst[c] = (act && kb >= 0 && kb + c >= i && kb + c < n) ? L[(long)i * n + kb + c]
                                                                   : T(0);
    };
    const int last = ((n - 1) / SB) * SB;
    load_up(last, nxt);
    for (int kb = last; kb >= 0; kb -= SB) {
#pragma unroll
        for (int c = 0; c < SB; ++c) cur[c] = nxt[c];
        load_up(kb - SB, nxt);
        if (w == kb / 64) {
#pragma unroll
            for (int c = SB - 1; c >= 0; --c) {
                const int r = kb + c;
                if (r < n) {
                    if (i == r) yi = yi / cur[c];
                    const T yc = read_lane(yi, r & 63);
                    if (i >= kb && i < r) yi -= cur[c] * yc;
                }
            }
        }
        if (act && i >= kb && i < kb + SB) ys[i] = yi;
        __syncthreads();
        if (act && i < kb) {
#pragma unroll
            for (int c = 0; c < SB; ++c)
                if (kb + c < n) yi -= cur[c] * ys[kb + c];
        }
    }
    if (act) y[i] = yi;
    __syncthreads();
}

// r = b - A x for the FP64 n x n row-major A (global) and x, b in LDS; one wave per group
// of RB rows, all loads of a group issued together.  With rowsum, also the largest
// absolute row sum of A (for the backward-error test), reduced over the wave.
template <bool ROWSUM>
__device__ __forceinline__ double residual(const double *__restrict__ A, int n, const double *x,
                           const double *b, double *r)
{
    constexpr int RB = 4;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double amax = 0.0;
    for (int i0 = w * RB; i0 < n; i0 += RB * (NT / 64)) {
        double a[RB][NMAX / 64];
#pragma unroll
        for (int k = 0; k < RB; ++k)
#pragma unroll
            for (int s = 0; s < NMAX / 64; ++s) {
                const int i = i0 + k, j = lane + 64 * s;
                a[k][s] = (i < n && j < n) ? A[(long)i * n + j] : 0.0;
            }
        double xs[NMAX / 64];
#pragma unroll
        for (int s = 0; s < NMAX / 64; ++s) xs[s] = lane + 64 * s < n ? x[lane + 64 * s] : 0.0;
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            double sum = 0.0, asum = 0.0;
#pragma unroll
            for (int s = 0; s < NMAX / 64; ++s) {
                sum += a[k][s] * xs[s];
                if (ROWSUM) asum += fabs(a[k][s]);
            }
            sum = wave_sum(sum);
            if (ROWSUM) asum = wave_sum(asum);
            if (i0 + k < n) {
                if (lane == 0) r[i0 + k] = b[i0 + k] - sum;
                if (ROWSUM) amax = fmax(amax, asum);
            }
        }
    }
    return amax;
}

// Workgroups per CU the static LDS allows (P and U plus about 5 KB of pivot scratch, of
// the 160 KB of a CU), at most 4; the launch bound lets the compiler size registers for it.
template <typename T, int NB>
constexpr int min_blocks()
{
    const int b = 160 * 1024 / (2 * (int)sizeof(T) * NB * PLD + 5 * 1024);
    return b < 1 ? 1 : (b > 4 ? 4 : b);
}

template <typename T, bool MIXED, int NB, int TM, int TN>
__global__ __launch_bounds__(NT, (min_blocks<T, NB>())) void
batched_solve_kernel(Args a)
{
    __shared__ __attribute__((aligned(16))) T P[NB * PLD];
    __shared__ __attribute__((aligned(16))) T U[NB * PLD];
    // P and U are free once the factorisation is done: the solve and refinement vectors
    // live there, which keeps the static LDS (and so the co-resident workgroups per CU)
    // at what the factorisation needs
    static_assert(sizeof(T) * NB * PLD >= 3 * NMAX * sizeof(double) && NB * PLD >= 2 * NMAX,
                  "solve vectors must fit in the panel buffers");
    double *const b64 = reinterpret_cast<double *>(P), *const x64 = b64 + NMAX,
                  *const r64 = x64 + NMAX;
    T *const yv = U, *const ys = U + NMAX;
    __shared__ T R[2 * (NT / 64 + 1) * NB];
    __shared__ PivLds pv;
    __shared__ double sv[NT / 64];
    __shared__ int s_inst, s_zero;

    constexpr int QB = 4;
    const int tid = threadIdx.x;
    double *A64 = a.ws64 + (long)blockIdx.x * NMAX * NMAX;
    T *LU = MIXED ? (T *)(a.ws32 + (long)blockIdx.x * NMAX * NMAX) : (T *)A64;

    const bool fallback = !MIXED && a.fail_count != nullptr;
    const int total = fallback ? *(volatile int *)a.fail_count : a.bd.nInst;
    for (;;) {
        if (tid == 0) {
            s_inst = atomicAdd(a.counter, 1);
            s_zero = 0;
        }
        __syncthreads();
        const int k = s_inst;
        __syncthreads();
        if (k >= total) break;
        const int inst = fallback ? a.fail_list[k] : a.bd.order[k];
        const int n = a.bd.N[inst], nq = a.bd.Q[inst];
        const double *Ai = a.A + a.bd.aoff[inst];
        const double *th = a.theta + a.bd.toff[inst];
        const long vo = a.bd.voff[inst];

        // assemble A(mu) (or its transpose): 4 entries per thread, QB terms at a time, all
        // 4*QB loads issued before the sums (one round trip for up to QB terms)
        const long nn = (long)n * n;
        for (long i0 = tid; i0 < nn; i0 += 4 * NT) {
            double s[4] = {0.0, 0.0, 0.0, 0.0};
            for (int q0 = 0; q0 < nq; q0 += QB) {
                double v[QB][4];
#pragma unroll
                for (int qq = 0; qq < QB; ++qq)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const long idx = i0 + u * NT;
                        v[qq][u] = (idx < nn && q0 + qq < nq) ? Ai[(q0 + qq) * nn + idx] : 0.0;
                    }
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const double t = q0 + qq < nq ? th[q0 + qq] : 0.0;
#pragma unroll
                    for (int u = 0; u < 4; ++u) s[u] += t * v[qq][u];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const long idx = i0 + u * NT;
                if (idx < nn) {
                    const int i = (int)(idx / n), j = (int)(idx - (long)i * n);
                    const long dst = a.transpose ? (long)j * n + i : idx;
                    A64[dst] = s[u];
                    if (MIXED) LU[dst] = (T)s[u];
                }
            }
        }
        __syncthreads();

        block_lu<T, NB, TM, TN>(LU, n, P, U, R, pv, &s_zero);  // ends with a barrier
        for (int i = tid; i < n; i += NT) {
            b64[i] = a.b[vo + i];
            yv[i] = (T)a.b[vo + pv.posAll[i]];
        }
        __syncthreads();
        lu_solve_rows<T>(LU, n, yv, ys);
        for (int i = tid; i < n; i += NT) x64[i] = (double)yv[i];
        __syncthreads();

        int it = 0, st = s_zero ? 1 : 0;
        if (MIXED) {
            double bmax = 0.0;
            for (int i = tid; i < n; i += NT) bmax = fmax(bmax, fabs(b64[i]));
            const double bnorm = block_max(bmax, sv);
            double anorm = 0.0;
            bool conv = false;
            // Stopping rule: the normwise backward error is at most tol, and the last
            // correction is either at most 100 tol relative to x or no longer shrinking
            // (above 0.9 of the one before: the iterate has reached the accuracy of the FP64
            // residual).  A backward-error test alone would stop while the forward error is
            // still about cond(A) * tol; refinement that converges slowly (cond(A) near
            // 1 / eps_FP32) reaches max_iter instead and reports status 2.
            double dn = 0.0, dprev = INFINITY;
            for (; it <= a.max_iter; ++it) {
                const double am = it == 0 ? residual<true>(A64, n, x64, b64, r64)
                                          : residual<false>(A64, n, x64, b64, r64);
                __syncthreads();
                if (it == 0) anorm = block_max(am, sv);
                double rm = 0.0, xm = 0.0;
                for (int i = tid; i < n; i += NT) {
                    rm = fmax(rm, fabs(r64[i]));
                    xm = fmax(xm, fabs(x64[i]));
                }
                const double rn = block_max(rm, sv);
                const double xn = block_max(xm, sv);
                if (it > 0 && rn <= a.tol * (anorm * xn + bnorm) &&
                    (dn <= 100.0 * a.tol * xn || dn > 0.9 * dprev)) {
                    conv = true;
                    break;
                }
                if (it == a.max_iter) break;
                for (int i = tid; i < n; i += NT) yv[i] = (T)r64[pv.posAll[i]];
                __syncthreads();
                lu_solve_rows<T>(LU, n, yv, ys);
                double dm = 0.0;
                for (int i = tid; i < n; i += NT) {
                    const double d = (double)yv[i];
                    x64[i] += d;
                    dm = fmax(dm, fabs(d));
                }
                dprev = it == 0 ? INFINITY : dn;
                dn = block_max(dm, sv);  // ends with a barrier
            }
            if (!conv && !st) {
                st = 2;
                if (a.fail_list) {
                    st = 3;
                    if (tid == 0) a.fail_list[atomicAdd(a.fail_count, 1)] = inst;
                }
            }
        }
        if (fallback && !st) st = 3;
        for (int i = tid; i < n; i += NT) a.x[vo + i] = x64[i];
        if (tid == 0) {
            if (a.iters) a.iters[inst] = it;
            if (a.status) a.status[inst] = st;
        }
        __syncthreads();
    }
}

__global__ void reset_counter(int *c) { *c = 0; }
__global__ void reset_counters(int *c, int *d)
{
    *c = 0;
    *d = 0;
}

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
                             IROM_V(32, 2, 4), IROM_V(16, 4, 2), IROM_V(32, 2, 2),
                             IROM_V(8, 4, 4),  IROM_V(8, 2, 4)};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace irom

struct irom_batch_s {
    int nInst = 0;
    std::vector<int> N, Q;
    std::vector<long> aoff, toff, voff;
    size_t a_size = 0, t_size = 0, v_size = 0;
    int *d_order = nullptr, *d_N = nullptr, *d_Q = nullptr, *d_counter = nullptr;
    int *d_fail = nullptr;  // FP64-fallback list (nInst) followed by its count
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

int launch_kernel(irom_batch b, irom::KernelFn fn, const irom::Args &a, hipStream_t s)
{
    // persistent grid: as many workgroups as are co-resident, capped by the workspace
    int per_cu = 1;
    IROM_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, irom::NT, 0),
               "occupancy");
    const int grid = std::max(1, std::min({b->nWG, std::max(1, per_cu) * b->nCU, b->nInst}));
    fn<<<grid, irom::NT, 0, s>>>(a);
    IROM_CHECK(hipGetLastError(), "batched solve launch");
    return 0;
}

int launch_solve(irom_batch b, int method, int variant, int transpose, const double *d_A,
                 const double *d_theta, const double *d_b, double *d_x, double tol, int max_iter,
                 int *d_iters, int *d_status, hipStream_t s)
{
    if (method != IROM_LU_FP64 && method != IROM_LU_FP32_IR && method != IROM_LU_FP32_IR_FALLBACK)
        return fail(1, "unknown method");
    if (!d_A || !d_theta || !d_b || !d_x) return fail(1, "null device array");
    irom::Args a{dev_view(b), transpose, d_A, d_theta, d_b, d_x, tol, max_iter, d_iters,
                 d_status, b->d_counter, b->d_ws64, b->d_ws32, nullptr, nullptr};
    const irom::Variant &v = irom::kVariants[variant];
    if (method == IROM_LU_FP64) {
        irom::reset_counter<<<1, 1, 0, s>>>(b->d_counter);
        return launch_kernel(b, v.fp64, a, s);
    }
    if (method == IROM_LU_FP32_IR) {
        irom::reset_counter<<<1, 1, 0, s>>>(b->d_counter);
        return launch_kernel(b, v.mixed, a, s);
    }
    // mixed, then FP64 LU for the instances whose refinement did not converge (the
    // second launch reads the list length from device memory; an empty list exits at once)
    a.fail_list = b->d_fail;
    a.fail_count = b->d_fail + b->nInst;
    irom::reset_counters<<<1, 1, 0, s>>>(b->d_counter, a.fail_count);
    if (int st = launch_kernel(b, v.mixed, a, s)) return st;
    irom::reset_counter<<<1, 1, 0, s>>>(b->d_counter);
    const irom::Variant &v64 = irom::kVariants[b->variant[0]];
    return launch_kernel(b, v64.fp64, a, s);
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
    IROM_CHECK(hipMalloc(&b->d_fail, (nInst + 1) * sizeof(int)), "hipMalloc");
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
    b->variant[0] = 0;  // NB16 4x4: the fastest on the 4096-instance N ~ U{16..256} batch
    b->variant[1] = 0;
    *out = b;
    return 0;
}

int irom_batch_destroy(irom_batch b)
{
    if (!b) return 0;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    for (void *p : {(void *)b->d_order, (void *)b->d_N, (void *)b->d_Q, (void *)b->d_counter, (void *)b->d_fail,
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
    return launch_solve(b, method, b->variant[method != IROM_LU_FP64], transpose, d_A,
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
    if (!b || v < 0 || v >= irom::kNumVariants || method < 0 || method > 2)
        return fail(1, "irom_batch_set_variant: bad argument");
    b->variant[method != IROM_LU_FP64] = v;
    return 0;
}

int irom_batch_autotune(irom_batch b, int method, const double *d_A, const double *d_theta,
                        const double *d_b, double *d_x, int reps, int *chosen, void *stream)
{
    if (!b || method < 0 || method > 2) return fail(1, "irom_batch_autotune: bad argument");
    const int slot = method != IROM_LU_FP64;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    IROM_CHECK(hipEventCreate(&e0), "event");
    IROM_CHECK(hipEventCreate(&e1), "event");
    float best = 1e30f;
    int bestv = b->variant[slot];
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
    b->variant[slot] = bestv;
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
    int st = launch_solve(b, method, b->variant[method != IROM_LU_FP64], transpose, d_A,
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
