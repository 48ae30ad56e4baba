// FP64 symmetric eigen-decomposition of the (small, dense) snapshot Gram matrix on the GPU.
//
// The reference obtains its POD basis from an Eigen BDCSVD of the snapshot matrix
// (CoupledLinearNS_ROM.cpp:3925-3954).  By the method of snapshots the same modes follow
// from the eigenpairs of the n x n Gram matrix G = S^T M S (n = number of snapshots).
//
// Algorithm: blocked one-sided (Hestenes) Jacobi on G.  For a symmetric positive
// semi-definite G, rotating the columns of W = G V until they are mutually orthogonal
// gives lambda_j = ||w_j|| and eigenvectors v_j.  The np columns are split into nb blocks
// of 16; a sweep is nb-1 rounds of the round-robin ("circle") pairing of blocks.  In a
// round, S workgroups take a block pair: each forms the 32 x 32 Gram matrix
// P = W_pair^T W_pair with FP64 MFMA (the waves sum disjoint row ranges), diagonalises it
// in LDS by a parallel two-sided Jacobi (16 disjoint rotations per step, one thread per
// 2 x 2 block of P, rotations accumulated in R), and the pair's columns of W and V become
// W R and V R (MFMA again) on its own share of the rows.  The S workgroups of a pair
// compute the same P and R bit for bit; a per-pair arrival count keeps any share from
// writing W before every share has read it.  The rotation of a column pair is the
// one-sided Jacobi rotation of the same two columns, so one inner sweep is exactly the
// column-by-column method; the columns themselves are touched twice per round instead of
// once per inner step.  All rounds and sweeps run in one persistent launch.  Rounds are
// not separated by a grid barrier: each column block counts the shares done with it, and
// a pair starts once its two blocks' previous round is done; workgroups meet at a grid
// barrier (an atomic arrival counter and a generation word, with a bounded wait) after
// each sweep, and stop together after the first sweep without a rotation.  Eigenvalues
// are sorted on the device (bitonic sort in one workgroup) and the eigenvectors gathered
// in that order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ithaca_rom.h"

namespace irom {

constexpr int EIG_MAX_SWEEPS = 64;  // length of the per-sweep rotation flag array

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

// Thread 0 waits (bounded like grid_sync) until both counters reach `want`, then acquires
// at device scope, so that the writes released before those counts are visible to the
// workgroup.
__device__ __forceinline__ void wait_counts(const unsigned *c0, const unsigned *c1, unsigned want,
                                            int *err)
{
    if (threadIdx.x == 0) {
        long spins = 0;
        while (__hip_atomic_load(c0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want ||
               __hip_atomic_load(c1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1L << 23)) {
                atomicOr(err, 1);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int JB = 16;       // columns per block
constexpr int JC = 2 * JB;   // columns of a block pair
static_assert(JC == 32, "the circle-schedule position arithmetic assumes 32 columns per pair");
constexpr int JP = JC + 1;   // padded LDS row
constexpr int NW = 8;        // waves per workgroup
constexpr int JT = 64 * NW;  // threads per workgroup
constexpr int GU = 2;        // 16-row blocks per wave and load group in the update

struct EigArgs {
    int n;           // order
    int np;          // padded order (multiple of 2 JB)
    int nb;          // number of column blocks (np / JB, even)
    double tol;      // rotate while |w_p.w_q| > tol ||w_p|| ||w_q|| and > (tol ||G||_F)^2
    int max_sweeps;  // <= EIG_MAX_SWEEPS
    int max_inner;   // inner sweeps over a pair's Gram matrix per round
    const double *G; // n x n (for ||G||_F)
    double *W;       // np x np column-major (W = G V)
    double *V;       // np x np column-major
    int *rot;        // [EIG_MAX_SWEEPS] per-sweep "some pair was rotated" flags (zeroed)
    unsigned *bar;   // [2] grid barrier (zeroed)
    int *sweeps;     // sweeps performed
    int *err;        // barrier expiry
    int split;       // workgroups per block pair (each updates a share of the rows)
    unsigned *cnt;   // [nb / 2] arrivals per block pair (zeroed)
    unsigned *ver;   // [nb] per column block: shares done with it, over all rounds (zeroed)
};

// 1/x and 1/sqrt(x) from the hardware approximations (about 2^-23 relative) refined by
// Newton steps; each step squares the relative error.
template <int STEPS>
__device__ __forceinline__ double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < STEPS; ++i) y = fma(y, fma(-x, y, 1.0), y);
    return y;
}
template <int STEPS>
__device__ __forceinline__ double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < STEPS; ++i) y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return y;
}

// Rotation (c, s) that makes columns p, q of W orthogonal, from app = |w_p|^2,
// aqq = |w_q|^2, apq = w_p.w_q: w_p <- c w_p - s w_q, w_q <- s w_p + c w_q.  Pairs already
// orthogonal to tol (relative; tol2 = tol^2), or whose product is below the absolute
// floor, keep c = 1, s = 0.  t = tan(theta) is the smaller root of t^2 + 2 zeta t - 1 = 0,
// zeta = (aqq - app) / (2 apq), computed to about 1e-14: nearly parallel columns (the
// columns of a rank-deficient Gram matrix) need the angle to about 1/kappa of the pair, so
// an FP32 angle would stall them; a rotation by a slightly different angle is still a
// rotation.  c = (1 + t^2)^(-1/2) is refined to full precision (two Newton steps from the
// hardware estimate), so c^2 + s^2 = 1 to rounding and V stays orthogonal.  No branches:
// the special cases (no rotation, |zeta| > 1e60 where t ~ 1/(2 zeta), a NaN zeta giving
// t = 0) are selected at the end.
__device__ __forceinline__ bool rot_params(double app, double aqq, double apq, double tol2,
                                           double floor_, double &c, double &s)
{
    const bool rot = apq != 0.0 && apq * apq > tol2 * (app * aqq) && fabs(apq) > floor_;
    const double zeta = (aqq - app) * rcp_nr<1>(rot ? 2.0 * apq : 1.0);
    const double az = fabs(zeta);
    const bool big = !(az <= 1e60);
    const double zc = big ? 1.0 : az;  // finite on the ordinary path
    const double h = fma(zc, zc, 1.0);
    const double tn = rcp_nr<1>(zc + h * rsq_nr<1>(h));
    const double tb = az == az ? 0.5 * rcp_nr<1>(zeta) : 0.0;
    const double t = big ? tb : (zeta < 0.0 ? -tn : tn);
    const double cc = rsq_nr<2>(fma(t, t, 1.0));
    c = rot ? cc : 1.0;
    s = rot ? cc * t : 0.0;
    return rot;
}

// One persistent launch for all sweeps; JT threads, one block pair at a time.
__global__ __launch_bounds__(JT) void gram_jacobi_kernel(EigArgs a)
{
    __shared__ double Pw[NW / 2][JC][JP];  // partial Gram matrices (waves w and w + NW/2)
    __shared__ double Pb[2][JC][JP];       // the pair's Gram matrix, current and next
    __shared__ double R[JC][JP];           // accumulated rotation
    __shared__ double Cs[JB][2];           // the rotations (c, s) of one inner step
    __shared__ double s_red[NW];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int i16 = lane & 15, g4 = lane >> 4;
    const int np = a.np, nb = a.nb, nrb = np / 16, S = a.split, nitems = nb / 2 * S;

    // ||G||_F, summed in the same order by every workgroup
    double fro = 0.0;
    for (long i = tid; i < (long)a.n * a.n; i += JT) fro = fma(a.G[i], a.G[i], fro);
    fro = wave_sum64(fro);
    if (lane == 0) s_red[wv] = fro;
    __syncthreads();
    double fsum = 0.0;
    for (int w = 0; w < NW; ++w) fsum += s_red[w];
    const double tg = a.tol * sqrt(fsum);
    const double floor_ = tg * tg, tol2 = a.tol * a.tol;

    unsigned gen = 0;
    int sw = 0;
    for (; sw < a.max_sweeps; ++sw) {
        for (int r = 0; r < nb - 1; ++r) {
            // work item (k, sp): block pair k of the round, row share sp.  The S workgroups of
            // a pair form the same P and R (same data, same order of operations); each
            // updates the rows [rb0, rb1) (16-row blocks) of the pair's columns.
            for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
                const int k = item / S, sp = item - k * S;
                const int rb0 = sp * nrb / S, rb1 = (sp + 1) * nrb / S;
                const int bi = player(k, r, nb), bj = player(nb - 1 - k, r, nb);
                // the two blocks' previous round is complete: every share of the pairs that
                // held them has written its rows (no grid barrier between rounds)
                wait_counts(&a.ver[bi], &a.ver[bj], (unsigned)S * (unsigned)(sw * (nb - 1) + r),
                            a.err);
                auto gcol = [&](int c) { return c < JB ? bi * JB + c : bj * JB + (c - JB); };
                // 1. P = W_pair^T W_pair.  Lane (i16, g4) holds rows 4 g4 .. 4 g4 + 3 of a
                // 16-row block for column i16 of each block; MFMA step t pairs row 4 g4 + t
                // of both operands (any one-to-one pairing of rows gives the same sum).
                // Groups of GP row blocks per wave: all loads of a group before its MFMAs.
                d4 a00 = {0, 0, 0, 0}, a01 = {0, 0, 0, 0}, a11 = {0, 0, 0, 0};
                {
                    const double *w0 = a.W + (long)gcol(i16) * np + 4 * g4;
                    const double *w1 = a.W + (long)gcol(JB + i16) * np + 4 * g4;
                    constexpr int GP = 4;
                    for (int rg = wv; rg < nrb; rg += NW * GP) {
                        double2 xv[GP][2], yv[GP][2];
#pragma unroll
                        for (int u = 0; u < GP; ++u) {
                            // rows past the end: load the last block (in bounds), use zeros
                            const int rb = min(rg + NW * u, nrb - 1);
                            xv[u][0] = *reinterpret_cast<const double2 *>(w0 + rb * 16);
                            xv[u][1] = *reinterpret_cast<const double2 *>(w0 + rb * 16 + 2);
                            yv[u][0] = *reinterpret_cast<const double2 *>(w1 + rb * 16);
                            yv[u][1] = *reinterpret_cast<const double2 *>(w1 + rb * 16 + 2);
                        }
#pragma unroll
                        for (int u = 0; u < GP; ++u) {
                            const double m = rg + NW * u < nrb ? 1.0 : 0.0;
                            const double x[4] = {m * xv[u][0].x, m * xv[u][0].y, m * xv[u][1].x, m * xv[u][1].y};
                            const double y[4] = {m * yv[u][0].x, m * yv[u][0].y, m * yv[u][1].x, m * yv[u][1].y};
#pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                a00 = __builtin_amdgcn_mfma_f64_16x16x4f64(x[t], x[t], a00, 0, 0, 0);
                                a01 = __builtin_amdgcn_mfma_f64_16x16x4f64(x[t], y[t], a01, 0, 0, 0);
                                a11 = __builtin_amdgcn_mfma_f64_16x16x4f64(y[t], y[t], a11, 0, 0, 0);
                            }
                        }
                    }
                }
                // component q of lane l: D[l/16 + 4q][l%16]; waves w < NW/2 store, then waves
                // w + NW/2 add theirs
                const int hw = wv % (NW / 2);
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    if ((wv >= NW / 2) == (half == 1)) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int ii = g4 + 4 * q;
                            double *d0 = &Pw[hw][ii][i16], *d1 = &Pw[hw][ii][JB + i16];
                            double *d2 = &Pw[hw][JB + ii][JB + i16];
                            *d0 = half ? *d0 + a00[q] : a00[q];
                            *d1 = half ? *d1 + a01[q] : a01[q];
                            *d2 = half ? *d2 + a11[q] : a11[q];
                        }
                    }
                    __syncthreads();
                }
                // sum of the partials in a fixed order; the upper triangle, mirrored
                for (int idx = tid; idx < JC * JC; idx += JT) {
                    const int ii = idx / JC, jj = idx - ii * JC;
                    const int u = min(ii, jj), v = max(ii, jj);
                    double sum = Pw[0][u][v];
#pragma unroll
                    for (int w = 1; w < NW / 2; ++w) sum += Pw[w][u][v];
                    Pb[0][ii][jj] = sum;
                    R[ii][jj] = ii == jj ? 1.0 : 0.0;
                }

                // The update's first load group (rows of W and V, lane (i16, g4): row i16
                // of a 16-row block, column 4 kk + g4), issued now so that it arrives while
                // the Gram matrix is diagonalised.
                // column 4 kk + g4 of the pair: kk < 4 in block bi, else in block bj
                const long kb0 = ((long)bi * JB + g4) * np + i16, kb1 = ((long)bj * JB + g4) * np + i16;
                auto ka = [&](int kk) { return (kk < 4 ? kb0 : kb1) + (long)(4 * (kk & 3)) * np; };
                double aw[GU][8], av[GU][8];
                auto load_group = [&](int g) {
#pragma unroll
                    for (int u = 0; u < GU; ++u) {
                        // rows past the share: the last block (in bounds), not used
                        const int rb = min(rb0 + wv + NW * (g * GU + u), nrb - 1);
#pragma unroll
                        for (int kk = 0; kk < 8; ++kk) {
                            aw[u][kk] = a.W[ka(kk) + rb * 16];
                            av[u][kk] = a.V[ka(kk) + rb * 16];
                        }
                    }
                };
                load_group(0);
                __syncthreads();
                // every row of W this workgroup reads from the other shares is consumed
                // (P is formed): count the arrival; the shares write W only after all S
                // arrivals of this round
                if (S > 1 && tid == 0) atomicAdd(&a.cnt[k], 1u);

                // 2. Pairs already orthogonal skip the rest.  Otherwise two-sided Jacobi on
                // P: in step st the 16 pairs of the circle schedule rotate at once; threads
                // 0..15 form the 16 rotations into Cs, then thread (k1, k2) of the first
                // half forms the 2 x 2 block of J^T P J on rows {p1, q1} and columns
                // {p2, q2} (P double-buffered), the same thread of the second half updates
                // that block of R <- R J.  The blocks partition the matrices.
                int need = 0;
                for (int idx = tid; idx < JC * JC; idx += JT) {
                    const int ii = idx / JC, jj = idx - ii * JC;
                    if (ii < jj) {
                        const double apq = Pb[0][ii][jj];
                        if (apq != 0.0 && apq * apq > tol2 * (Pb[0][ii][ii] * Pb[0][jj][jj]) &&
                            fabs(apq) > floor_)
                            need = 1;
                    }
                }
                bool pair_rot = false;
                if (__syncthreads_or(need)) {
                    const bool rhalf = tid >= JT / 2;
                    const int hb = rhalf ? tid - JT / 2 : tid;
                    const int k1 = hb >> 4, k2 = hb & 15, ka = tid & (JB - 1);
                    int cur = 0;
#pragma unroll 1
                    for (int it = 0; it < a.max_inner; ++it) {
                        int any = 0;
                        int pl[6] = {player(k1, 0, JC), player(JC - 1 - k1, 0, JC), player(k2, 0, JC),
                                     player(JC - 1 - k2, 0, JC), player(ka, 0, JC),
                                     player(JC - 1 - ka, 0, JC)};
#pragma unroll 1
                        for (int st = 0; st < JC - 1; ++st) {
                            const int p1 = pl[0], q1 = pl[1], p2 = pl[2], q2 = pl[3];
                            double(*P)[JP] = Pb[cur];
                            // each rotation once: thread k < 16 forms pair k's (c, s)
                            if (tid < JB) {
                                const int pa = pl[4], qa = pl[5];
                                double c, sn;
                                if (rot_params(P[pa][pa], P[qa][qa], P[pa][qa], tol2, floor_, c, sn))
                                    any = 1;
                                Cs[tid][0] = c;
                                Cs[tid][1] = sn;
                            }
                            __syncthreads();
                            const double c2 = Cs[k2][0], s2 = Cs[k2][1];
                            if (!rhalf) {
                                double(*Pn)[JP] = Pb[cur ^ 1];
                                const double c1 = Cs[k1][0], s1 = Cs[k1][1];
                                const double mpp = P[p1][p2], mpq = P[p1][q2], mqp = P[q1][p2], mqq = P[q1][q2];
                                const double xpp = c2 * mpp - s2 * mpq, xpq = s2 * mpp + c2 * mpq;
                                const double xqp = c2 * mqp - s2 * mqq, xqq = s2 * mqp + c2 * mqq;
                                Pn[p1][p2] = c1 * xpp - s1 * xqp;
                                Pn[p1][q2] = c1 * xpq - s1 * xqq;
                                Pn[q1][p2] = s1 * xpp + c1 * xqp;
                                Pn[q1][q2] = s1 * xpq + c1 * xqq;
                            } else {
                                const double rpp = R[p1][p2], rpq = R[p1][q2], rqp = R[q1][p2], rqq = R[q1][q2];
                                R[p1][p2] = c2 * rpp - s2 * rpq;
                                R[p1][q2] = s2 * rpp + c2 * rpq;
                                R[q1][p2] = c2 * rqp - s2 * rqq;
                                R[q1][q2] = s2 * rqp + c2 * rqq;
                            }
                            // the circle schedule: every position but the first advances
                            // by one, 31 wrapping to 1 (integer arithmetic: a conditional
                            // form is compiled into divergent branches)
#pragma unroll
                            for (int u = 0; u < 6; ++u) {
                                int v = pl[u];
                                v += (v + JC - 1) >> 5;  // +1 unless v = 0
                                v -= (v >> 5) * (JC - 1);  // JC -> 1
                                pl[u] = v;
                            }
                            cur ^= 1;
                            __syncthreads();
                        }
                        if (!__syncthreads_or(any)) break;
                        pair_rot = true;
                    }
                }

                // 3. The pair's columns of W and V <- (W, V) R: A = the rows loaded above,
                // B = R; each wave writes only its own rows.
                if (pair_rot) {
                    if (S > 1) {  // the other shares have read W for their P
                        if (tid == 0) {
                            const unsigned target = (unsigned)S * (unsigned)(sw * (nb - 1) + r + 1);
                            long spins = 0;
                            while (__hip_atomic_load(&a.cnt[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                                __builtin_amdgcn_s_sleep(1);
                                if (++spins > (1L << 23)) {
                                    atomicOr(a.err, 1);
                                    break;
                                }
                            }
                        }
                        __syncthreads();
                    }
                    const long kc0 = (long)gcol(i16) * np + g4, kc1 = (long)gcol(JB + i16) * np + g4;
                    for (int g = 0; rb0 + wv + NW * g * GU < rb1; ++g) {
                        if (g > 0) load_group(g);
#pragma unroll
                        for (int u = 0; u < GU; ++u) {
                            const int rb = rb0 + wv + NW * (g * GU + u);
                            const int r0 = rb * 16;
                            if (rb < rb1) {
                            d4 ow0 = {0, 0, 0, 0}, ow1 = {0, 0, 0, 0}, ov0 = {0, 0, 0, 0}, ov1 = {0, 0, 0, 0};
#pragma unroll
                            for (int kk = 0; kk < 8; ++kk) {
                                const double rb0 = R[4 * kk + g4][i16], rb1 = R[4 * kk + g4][JB + i16];
                                ow0 = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[u][kk], rb0, ow0, 0, 0, 0);
                                ow1 = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[u][kk], rb1, ow1, 0, 0, 0);
                                ov0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][kk], rb0, ov0, 0, 0, 0);
                                ov1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][kk], rb1, ov1, 0, 0, 0);
                            }
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                a.W[kc0 + r0 + 4 * q] = ow0[q];
                                a.W[kc1 + r0 + 4 * q] = ow1[q];
                                a.V[kc0 + r0 + 4 * q] = ov0[q];
                                a.V[kc1 + r0 + 4 * q] = ov1[q];
                            }
                            }
                        }
                    }
                    if (tid == 0) atomicOr(&a.rot[sw], 1);
                }
                // this share is done with blocks bi and bj: release its writes and count it
                __syncthreads();
                if (tid == 0) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                    atomicAdd(&a.ver[bi], 1u);
                    atomicAdd(&a.ver[bj], 1u);
                }
            }
        }
        grid_sync(a.bar, gridDim.x, gen, a.err);
        // every flag of this sweep was set before the barrier: all workgroups read the same
        // value and stop together
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

// Padded order (a multiple of the 32 columns of a block pair), block count, and the
// power of two m >= n of the sort.
struct EigPlan {
    int np, nb, m;
};

constexpr EigPlan eig_plan(int n)
{
    EigPlan p{(n + JC - 1) / JC * JC, 0, 1};
    p.nb = p.np / JB;
    while (p.m < n) p.m <<= 1;
    return p;
}

// Workspace layout (bytes, each part 256-byte aligned): W, V (np^2 doubles each), norms
// (2 np: of W's and of V's columns), sorted values (n), perm (n ints), then the ints: rot
// flags, barrier words, sweeps, err, the per-pair arrival counters (np / JC) and the
// per-block version counters (np / JB).
struct EigWs {
    double *W, *V, *nrm, *ev;
    int *perm, *ints;  // ints: rot[EIG_MAX_SWEEPS], bar[2], sweeps, err, cnt[np / JC], ver[np / JB]
};

constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }

constexpr size_t eig_ws_bytes(int n)
{
    const EigPlan p = eig_plan(n);
    return 2 * al256(sizeof(double) * (size_t)p.np * p.np) + al256(sizeof(double) * 2 * p.np) +
           al256(sizeof(double) * n) + al256(sizeof(int) * n) +
           al256(sizeof(int) * (EIG_MAX_SWEEPS + 4 + p.np / JC + p.np / JB));
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

// Row shares per block pair: the most (up to 4) for which every work item of a round has
// a workgroup of its own (the shares of a pair wait for each other), one per CU.
int eig_split(int nb, int ncu)
{
    int sp = 4;
    while (sp > 1 && (nb / 2) * sp > ncu) sp >>= 1;
    return sp;
}

hipError_t launch_gram_jacobi(const irom::EigArgs &a, int ncu, hipStream_t s)
{
    // every workgroup must be resident at once (grid barrier): at most one per CU here;
    // with more than one share per pair, one workgroup per work item
    const int grid = std::max(1, std::min(a.nb / 2 * a.split, ncu));
    void *args[] = {(void *)&a};
    return hipLaunchCooperativeKernel((const void *)&irom::gram_jacobi_kernel, dim3(grid),
                                      dim3(irom::JT), args, 0, s);
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
    const int nints = irom::EIG_MAX_SWEEPS + 4 + p.np / irom::JC + p.np / irom::JB;
    const long tot = std::max<long>((long)p.np * p.np, nints);
    irom::eig_init<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, p.np, d_G, w.W, w.V, w.ints,
                                                                 nints);
    IROM_ECHECK(hipGetLastError(), "eig_init");
    // one inner sweep per round: more inner sweeps saved 1-2 outer sweeps of 13 but cost
    // more than they saved
    irom::EigArgs a{n, p.np, p.nb, tol, max_sweeps, 1, d_G, w.W, w.V, rot,
                    (unsigned *)bar, sw_d, err_d, eig_split(p.nb, ncu),
                    (unsigned *)(err_d + 1), (unsigned *)(err_d + 1) + p.np / irom::JC};
    IROM_ECHECK(launch_gram_jacobi(a, ncu, s), "gram_jacobi launch");
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
