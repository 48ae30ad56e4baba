// Hexahedral sum-factorised operators for gfx950.
//
// Work decomposition ("pencil rotation").  One element is owned by NQ*NQ threads.
// Every 1D contraction is done in registers on a pencil (a line of the element along
// the contracted direction) held by one thread; the data are re-oriented between
// contractions through LDS (one write + one read per value, odd strides, conflict
// free), instead of each output reading all NQ inputs from LDS.  The last pencil
// orientation of the BwdTrans is the k-column of thread (i,j), which is also the
// orientation in which the geometry stream is read coalesced ([e][comp][k][j][i]).
//
// Reference semantics (all paths relative to the reference root):
//   BwdTrans   BwdTransKernels.hpp:302-372      u = (B x B x B) u_hat
//   IProduct   IProductKernels.hpp:236-311      out = (B x B x B)^T diag(w J) f
//   PhysDeriv  PhysDerivKernels.hpp:219-294     d_a = D_a u
//   Helmholtz  Helmholtz.h:764-996              out = lambda M u_hat + L u_hat
// The Helmholtz kernel applies dbdata = D bdata (Basis.cpp:418) in factored form:
// sum_i dB[p,i] g_i = sum_l B[p,l] (D^T g)_l, so the three IProducts of the reference
// collapse into one D^T application per direction followed by one (B x B x B)^T.
#pragma once

#include "imf_common.hpp"

namespace imf {

template <typename T>
__device__ __forceinline__ T ld_stream(const T *p)
{
    return __builtin_nontemporal_load(p);
}

template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// ---------------------------------------------------------------------------
// Building blocks.  `t` is the thread's index inside its element (0 <= t < NQ*NQ).
// Coefficient pencils live in LDS with an odd row stride NM+1 (value (t, p) of the
// element's mode set, t = r*NM + q, at t*(NM+1) + p) so that the p-pencil reads of
// neighbouring threads hit different banks.
// ---------------------------------------------------------------------------

template <int NM, int NQ>
struct HexLds {
    static constexpr int M = NM * NM * NM, Q = NQ * NQ * NQ;
    static constexpr int MP = NM * NM * (NM + 1);  // padded pencil span of a mode set
    static constexpr int QB = Q > MP ? Q : MP;     // one LDS buffer
};

// Block-cooperative, fully coalesced load of the block's contiguous span of nv elements'
// coefficients (nv*M scalars from src) into the padded pencil buffers buf(slot).
// All of a thread's loads are issued before its first LDS store, so the span costs one
// memory round trip rather than one per loop iteration.
template <int NM, int NQ, typename T, typename BufFn>
__device__ __forceinline__ void load_modes_block(const T *__restrict__ src, int nv, BufFn buf)
{
    constexpr int M = NM * NM * NM;
    constexpr int BLOCK = HexCfg<NQ>::BLOCK, PER = (HexCfg<NQ>::EPB * M + BLOCK - 1) / BLOCK;
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int f = threadIdx.x + u * BLOCK;
        v[u] = f < nv * M ? ld_stream(src + f) : T(0);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int f = threadIdx.x + u * BLOCK;
        if (f < nv * M) {
            const int sl = f / M, o = f - sl * M;
            const int tp = o / NM, p = o - tp * NM;
            buf(sl)[tp * (NM + 1) + p] = v[u];
        }
    }
}

// The same from a global vector: coefficient f of the block is gsign[f] * glob[gmap[f]]
// (gmap, gsign offset to the block's first element; the map reads are coalesced, the
// gather follows the global numbering).
template <int NM, int NQ, typename T, typename BufFn>
__device__ __forceinline__ void load_modes_block_gather(const T *__restrict__ glob,
                                                        const int *__restrict__ gmap,
                                                        const signed char *__restrict__ gsign,
                                                        int nv, BufFn buf)
{
    constexpr int M = NM * NM * NM;
    constexpr int BLOCK = HexCfg<NQ>::BLOCK, PER = (HexCfg<NQ>::EPB * M + BLOCK - 1) / BLOCK;
    int gi[PER];
    signed char sg[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int f = threadIdx.x + u * BLOCK;
        gi[u] = f < nv * M ? gmap[f] : 0;
        sg[u] = f < nv * M ? gsign[f] : 0;
    }
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) v[u] = T(sg[u]) * glob[gi[u]];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int f = threadIdx.x + u * BLOCK;
        if (f < nv * M) {
            const int sl = f / M, o = f - sl * M;
            const int tp = o / NM, p = o - tp * NM;
            buf(sl)[tp * (NM + 1) + p] = v[u];
        }
    }
}

// Block-cooperative, fully coalesced store of nv elements' coefficients from the padded
// pencil buffers buf(slot) to dst.
template <int NM, int NQ, typename T, typename BufFn>
__device__ __forceinline__ void store_modes_block(T *__restrict__ dst, int nv, BufFn buf)
{
    constexpr int M = NM * NM * NM;
    for (int f = threadIdx.x; f < nv * M; f += blockDim.x) {
        const int sl = f / M, o = f - sl * M;
        const int tp = o / NM, p = o - tp * NM;
        st_stream(dst + f, buf(sl)[tp * (NM + 1) + p]);
    }
}

// Mode pencils -> u (registers, k-column of thread (i,j)).  pencil(p) returns the p-th
// coefficient of this thread's p-pencil (t = r*NM + q); uses buf0 and buf1.
template <int NM, int NQ, typename T, typename PencilFn>
__device__ __forceinline__ void hex_bwd_to_columns(const Tab<NM, NQ, T> &tab, PencilFn pencil,
                                                   bool active, int t, T *buf0, T *buf1,
                                                   T (&u)[NQ])
{
    // p -> i on p-pencils (thread (q, r))
    if (active && t < NM * NM) {
        T uh[NM];
#pragma unroll
        for (int p = 0; p < NM; ++p) uh[p] = pencil(p);
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = uh[0] * tab.B[i];
#pragma unroll
            for (int p = 1; p < NM; ++p) s = fma(uh[p], tab.B[p * NQ + i], s);
            buf0[t * NQ + i] = s;  // s1[r][q][i], t = r*NM + q
        }
    }
    __syncthreads();
    // q -> j on q-pencils (thread (i, r))
    if (active && t < NQ * NM) {
        const int i = t % NQ, r = t / NQ;
        T s1[NM];
#pragma unroll
        for (int q = 0; q < NM; ++q) s1[q] = buf0[(r * NM + q) * NQ + i];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            T s = s1[0] * tab.B[j];
#pragma unroll
            for (int q = 1; q < NM; ++q) s = fma(s1[q], tab.B[q * NQ + j], s);
            buf1[(r * NQ + j) * NQ + i] = s;  // s2[r][j][i]
        }
    }
    __syncthreads();
    // r -> k on r-pencils (thread (i, j)): the element's k-columns
    if (active) {
        T s2[NM];
#pragma unroll
        for (int r = 0; r < NM; ++r) s2[r] = buf1[r * NQ * NQ + t];
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            T s = s2[0] * tab.B[k];
#pragma unroll
            for (int r = 1; r < NM; ++r) s = fma(s2[r], tab.B[r * NQ + k], s);
            u[k] = s;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NQ; ++k) u[k] = T(0);
    }
}

// k-columns acc (registers of thread (i,j)) -> (B x B x B)^T acc; emit(p, v) receives the
// p-th coefficient of this thread's p-pencil (t = r*NM + q).  Uses buf0 and buf1; emit may
// write to buf0.
template <int NM, int NQ, typename T, typename EmitFn>
__device__ __forceinline__ void hex_columns_to_modes(const Tab<NM, NQ, T> &tab,
                                                     const T (&acc)[NQ], bool active, int t,
                                                     T *buf0, T *buf1, EmitFn emit)
{
    // k -> r in registers
    if (active) {
#pragma unroll
        for (int r = 0; r < NM; ++r) {
            T s = acc[0] * tab.B[r * NQ];
#pragma unroll
            for (int k = 1; k < NQ; ++k) s = fma(acc[k], tab.B[r * NQ + k], s);
            buf0[r * NQ * NQ + t] = s;  // z2[r][j][i]
        }
    }
    __syncthreads();
    // j -> q on j-pencils (thread (i, r))
    if (active && t < NQ * NM) {
        const int i = t % NQ, r = t / NQ;
        T v[NQ];
#pragma unroll
        for (int j = 0; j < NQ; ++j) v[j] = buf0[(r * NQ + j) * NQ + i];
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            T s = v[0] * tab.B[q * NQ];
#pragma unroll
            for (int j = 1; j < NQ; ++j) s = fma(v[j], tab.B[q * NQ + j], s);
            buf1[(r * NM + q) * NQ + i] = s;  // z1[r][q][i]
        }
    }
    __syncthreads();
    // i -> p on i-pencils (thread (q, r))
    if (active && t < NM * NM) {
        T v[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) v[i] = buf1[t * NQ + i];
#pragma unroll
        for (int p = 0; p < NM; ++p) {
            T s = v[0] * tab.B[p * NQ];
#pragma unroll
            for (int i = 1; i < NQ; ++i) s = fma(v[i], tab.B[p * NQ + i], s);
            emit(p, s);
        }
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// BwdTrans (DIR = kForward, GEO ignored) and the Euclidean transpose of IProduct
// (DIR = kTranspose: out = diag(w J) B g, GEO = kDeformed or kRegular).  These light
// kernels read and write global memory directly from the pencil owners (staging their
// spans through LDS costs occupancy and is slower here).
template <int NM, int NQ, typename T, int DIR, int GEO>
__global__ __launch_bounds__(HexCfg<NQ>::BLOCK) void hex_bwdtrans_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = HexCfg<NQ>;
    constexpr int M = NM * NM * NM, Q = NQ * NQ * NQ;
    __shared__ T lds[C::EPB][2][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;

    // this thread's p-pencil, loaded before the contraction (as one unconditional block)
    T uh[NM];
    if (valid && t < NM * NM) {
        const T *src = in + e * M + t * NM;
#pragma unroll
        for (int p = 0; p < NM; ++p) uh[p] = ld_stream(src + p);
    } else {
#pragma unroll
        for (int p = 0; p < NM; ++p) uh[p] = T(0);
    }
    T u[NQ];
    hex_bwd_to_columns<NM, NQ, T>(tab, [&](int p) { return uh[p]; }, active, t, lds[es][0],
                                  lds[es][1], u);
    if (!valid) return;
    if (DIR == kTranspose) {
        if (GEO == kDeformed) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) u[k] *= ld_stream(geom + e * Q + k * C::TPE + t);
        } else {
            const T wij = tab.w[t % NQ] * tab.w[t / NQ] * geom[e];
#pragma unroll
            for (int k = 0; k < NQ; ++k) u[k] *= wij * tab.w[k];
        }
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) st_stream(out + e * Q + k * C::TPE + t, u[k]);
}

// IProductWRTBase: out = (B x B x B)^T diag(w J) f.  GEO = kUnit gives the Euclidean
// transpose of BwdTrans (w J == 1).
template <int NM, int NQ, typename T, int GEO>
__global__ __launch_bounds__(HexCfg<NQ>::BLOCK) void hex_iproduct_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = HexCfg<NQ>;
    constexpr int M = NM * NM * NM, Q = NQ * NQ * NQ;
    __shared__ T lds[C::EPB][2][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;

    T acc[NQ];
    if (valid) {
        if (GEO == kDeformed) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) {
                const int q = k * C::TPE + t;
                acc[k] = ld_stream(in + e * Q + q) * ld_stream(geom + e * Q + q);
            }
        } else if (GEO == kRegular) {
            const T wij = tab.w[t % NQ] * tab.w[t / NQ] * geom[e];
#pragma unroll
            for (int k = 0; k < NQ; ++k)
                acc[k] = ld_stream(in + e * Q + k * C::TPE + t) * (wij * tab.w[k]);
        } else {
#pragma unroll
            for (int k = 0; k < NQ; ++k) acc[k] = ld_stream(in + e * Q + k * C::TPE + t);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NQ; ++k) acc[k] = T(0);
    }
    T *dst = out + e * M + t * NM;
    hex_columns_to_modes<NM, NQ, T>(
        tab, acc, active, t, lds[es][0], lds[es][1], [&](int p, T v) {
            if (valid) st_stream(dst + p, v);
        });
}

// Mass: out_v = (B x B x B)^T diag(w J) (B x B x B) in_v for nvec vectors, the fused form
// of the reference's MassMatrixOp_MatFree (StdRegions/StdExpansion.cpp:647-661:
// BwdTrans, then IProductWRTBase).  The thread's geometry column (NQ values of w J) is
// read once and reused for every vector, so each further vector streams only its
// coefficients in and out.
template <int NM, int NQ, typename T, int GEO>
__global__ __launch_bounds__(HexCfg<NQ>::BLOCK) void hex_mass_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, int nvec, const T *__restrict__ in, long ldin,
    const T *__restrict__ geom, T *__restrict__ out, long ldout)
{
    using C = HexCfg<NQ>;
    constexpr int M = NM * NM * NM, Q = NQ * NQ * NQ;
    __shared__ T lds[C::EPB][2][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;

    T jw[NQ];
    if (valid) {
        if (GEO == kDeformed) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) jw[k] = geom[e * Q + k * C::TPE + t];
        } else {
            const T wij = tab.w[t % NQ] * tab.w[t / NQ] * geom[e];
#pragma unroll
            for (int k = 0; k < NQ; ++k) jw[k] = wij * tab.w[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NQ; ++k) jw[k] = T(0);
    }
    // Buffer reuse across vectors: each phase writes a buffer only after a barrier that
    // follows the previous vector's last read of it.
    for (int v = 0; v < nvec; ++v) {
        T uh[NM];
        if (valid && t < NM * NM) {
            const T *src = in + v * ldin + e * M + t * NM;
#pragma unroll
            for (int p = 0; p < NM; ++p) uh[p] = ld_stream(src + p);
        } else {
#pragma unroll
            for (int p = 0; p < NM; ++p) uh[p] = T(0);
        }
        T u[NQ];
        hex_bwd_to_columns<NM, NQ, T>(tab, [&](int p) { return uh[p]; }, active, t,
                                      lds[es][0], lds[es][1], u);
#pragma unroll
        for (int k = 0; k < NQ; ++k) u[k] *= jw[k];
        T *dst = out + v * ldout + e * M + t * NM;
        hex_columns_to_modes<NM, NQ, T>(
            tab, u, active, t, lds[es][0], lds[es][1], [&](int p, T x) {
                if (valid) st_stream(dst + p, x);
            });
    }
}

// Helmholtz: out = lambda M u_hat + sum_ab (D_a B)^T diag(w J G^{ab}) (D_b B) u_hat.
// Packed geometry: deformed [e][GS] with GS = 7Q rounded up to 128-byte lines and the
// components (Jw, JwG00, JwG01, JwG02, JwG11, JwG12, JwG22) at c*Q; regular [e][7]
// (J, J G00, ...), weights applied in the kernel.
// GATHER: `in` is a global vector read through (gmap, gsign), as GlobalToLocal would
// expand it, so the solver's operator needs no local copy of its input.
// MANY: nvec vectors per launch (vector v at in + v*ldin, out + v*ldout), the operator
// applied to each in turn by the same workgroup; deformed geometry is staged in LDS once
// (when the element's 7Q values fit beside the work buffers) and read from there for
// every vector, so each further vector streams only its coefficients in and out.
template <int NM, int NQ, typename T, int GEO>
constexpr bool helm_geo_in_lds()
{
    using Lb = HexLds<NM, NQ>;
    return GEO == kDeformed &&
           HexCfg<NQ>::EPB * (3 * Lb::QB + 7 * Lb::Q) * (int)sizeof(T) <= 64 * 1024;
}

// Number of the 7 geometry arrays a single-vector deformed launch prefetches into
// registers: as many as fit in 56 VGPRs per thread (NQ values each).  For nm = 6, nq = 7
// in FP64 that is 4 (162 VGPRs, 3 waves per SIMD), measured 7 % faster than none (103
// VGPRs, 4 waves) and than all 7 (182 VGPRs, 2 waves).
template <int NQ, typename T, int GEO, bool MANY>
constexpr int helm_geo_prefetch()
{
    if (GEO != kDeformed || MANY) return 0;
    const int per = NQ * (int)sizeof(T) / 4;
    return per > 56 ? 0 : (56 / per > 7 ? 7 : 56 / per);
}

template <int NM, int NQ, typename T, int GEO, bool GATHER = false, bool MANY = false>
__global__ __launch_bounds__(HexCfg<NQ>::BLOCK) void hex_helmholtz_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, T lambda, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out, const int *__restrict__ gmap = nullptr,
    const signed char *__restrict__ gsign = nullptr, int nvec = 1, long ldin = 0,
    long ldout = 0)
{
    using C = HexCfg<NQ>;
    using Lb = HexLds<NM, NQ>;
    constexpr int M = Lb::M, Q = Lb::Q, TPE = C::TPE;
    constexpr long GS = geo_stride(7 * Q, sizeof(T));
    constexpr bool GLDS = MANY && helm_geo_in_lds<NM, NQ, T, GEO>();
    // Single-vector deformed launches read the first NP geometry arrays of this thread's
    // points into registers before anything else, so they arrive with the coefficients, in
    // one memory round trip, while the BwdTrans runs; the rest are read where they are used.
    constexpr int NP = helm_geo_prefetch<NQ, T, GEO, MANY>();
    constexpr bool PRE = NP > 0;
    __shared__ T lds[C::EPB][3][Lb::QB];
    __shared__ T lgeo[GLDS ? C::EPB : 1][GLDS ? 7 * Q : 1];
    const int eb = threadIdx.x / TPE;
    const int t = threadIdx.x - eb * TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e0 = (long)blockIdx.x * C::EPB, e = e0 + eb;
    const int nv = (int)min((long)C::EPB, (long)nElmt - e0);
    const bool valid = active && eb < nv;
    T *buf0 = lds[es][0], *buf1 = lds[es][1], *buf2 = lds[es][2];
    const int ci = t % NQ, cj = t / NQ;  // this thread's column (i, j)
    if constexpr (GLDS) {
        // the block's elements' geometry, coalesced, once (made visible by the first
        // barrier below)
        for (int f = threadIdx.x; f < nv * 7 * Q; f += blockDim.x) {
            const int sl = f / (7 * Q), o = f - sl * (7 * Q);
            lgeo[sl][o] = ld_stream(geom + (e0 + sl) * GS + o);
        }
    }
    T gp[PRE ? NP : 1][PRE ? NQ : 1];
    if constexpr (PRE) {
        if (valid) {
            const T *g = geom + e * GS + t;
#pragma unroll
            for (int k = 0; k < NQ; ++k)
#pragma unroll
                for (int c = 0; c < NP; ++c) gp[c][k] = ld_stream(g + c * Q + k * TPE);
        }
    }
    const int nrep = MANY ? nvec : 1;
  for (int vv = 0; vv < nrep; ++vv) {
    if (MANY && vv > 0) {
        in += ldin;
        out += ldout;
        __syncthreads();  // the previous vector's store has read buf2 of every slot
    }

    // 1. BwdTrans into k-columns; keep a copy of u in LDS for the cross derivatives.
    if constexpr (GATHER)
        load_modes_block_gather<NM, NQ, T>(in, gmap + e0 * M, gsign + e0 * M, nv,
                                           [&](int sl) { return lds[sl][2]; });
    else
        load_modes_block<NM, NQ, T>(in + e0 * M, nv, [&](int sl) { return lds[sl][2]; });
    __syncthreads();
    T u[NQ];
    hex_bwd_to_columns<NM, NQ, T>(
        tab, [&](int p) { return buf2[t * (NM + 1) + p]; }, active, t, buf0, buf1, u);
    // buf2 (the mode pencils) was last read before the first barrier above
    if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) buf2[k * TPE + t] = u[k];
    }
    __syncthreads();

    // 2. Reference-space derivatives (PhysDerivTensor3DKernel).  d0 on i-pencils of
    // thread (j,k) = (t%NQ, t/NQ), d1 on j-pencils of thread (i,k) = (t%NQ, t/NQ),
    // d2 on this thread's k-column.
    if (active) {
        T v[NQ];
        {
            const int pj = t % NQ, pk = t / NQ;
#pragma unroll
            for (int l = 0; l < NQ; ++l) v[l] = buf2[(pk * NQ + pj) * NQ + l];
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                T s = v[0] * tab.D[i];
#pragma unroll
                for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[l * NQ + i], s);
                buf0[(pk * NQ + pj) * NQ + i] = s;
            }
        }
        {
            const int pi = t % NQ, pk = t / NQ;
#pragma unroll
            for (int l = 0; l < NQ; ++l) v[l] = buf2[(pk * NQ + l) * NQ + pi];
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                T s = v[0] * tab.D[j];
#pragma unroll
                for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[l * NQ + j], s);
                buf1[(pk * NQ + j) * NQ + pi] = s;
            }
        }
    }
    __syncthreads();

    // 3. Pointwise metric contraction and mass term on this thread's column.
    T g0[NQ], g1[NQ], g2[NQ], acc[NQ];
    if (valid) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            T d2 = u[0] * tab.D[k];
#pragma unroll
            for (int l = 1; l < NQ; ++l) d2 = fma(u[l], tab.D[l * NQ + k], d2);
            const T d0 = buf0[k * TPE + t];
            const T d1 = buf1[k * TPE + t];
            T jw, m00, m01, m02, m11, m12, m22;
            if constexpr (PRE) {
                const T *g = geom + e * GS + k * TPE + t;
                auto get = [&](int c) { return c < NP ? gp[c < NP ? c : 0][k] : ld_stream(g + c * Q); };
                jw = get(0);
                m00 = get(1);
                m01 = get(2);
                m02 = get(3);
                m11 = get(4);
                m12 = get(5);
                m22 = get(6);
            } else if (GLDS) {
                const T *g = &lgeo[GLDS ? es : 0][k * TPE + t];
                jw = g[0];
                m00 = g[Q];
                m01 = g[2 * Q];
                m02 = g[3 * Q];
                m11 = g[4 * Q];
                m12 = g[5 * Q];
                m22 = g[6 * Q];
            } else if (GEO == kDeformed) {
                const T *g = geom + e * GS + k * TPE + t;
                jw = ld_stream(g);
                m00 = ld_stream(g + Q);
                m01 = ld_stream(g + 2 * Q);
                m02 = ld_stream(g + 3 * Q);
                m11 = ld_stream(g + 4 * Q);
                m12 = ld_stream(g + 5 * Q);
                m22 = ld_stream(g + 6 * Q);
            } else {
                const T wq = tab.w[ci] * tab.w[cj] * tab.w[k];
                const T *g = geom + e * 7;
                jw = g[0] * wq;
                m00 = g[1] * wq;
                m01 = g[2] * wq;
                m02 = g[3] * wq;
                m11 = g[4] * wq;
                m12 = g[5] * wq;
                m22 = g[6] * wq;
            }
            g0[k] = fma(m02, d2, fma(m01, d1, m00 * d0));
            g1[k] = fma(m12, d2, fma(m11, d1, m01 * d0));
            g2[k] = fma(m22, d2, fma(m12, d1, m02 * d0));
            acc[k] = lambda * jw * u[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NQ; ++k) g0[k] = g1[k] = g2[k] = acc[k] = T(0);
    }
    // g0, g1 replace d0, d1 at this thread's own positions (no other thread reads them)
    if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            buf0[k * TPE + t] = g0[k];
            buf1[k * TPE + t] = g1[k];
        }
    }
    // D^T along k in registers.
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        T s = acc[k];
#pragma unroll
        for (int l = 0; l < NQ; ++l) s = fma(g2[l], tab.D[k * NQ + l], s);
        acc[k] = s;
    }
    __syncthreads();

    // 4. D^T along i (pencils of thread (j,k)) into buf2, then D^T along j (pencils of
    // thread (i,k)) added to it: three buffers instead of four per element.
    if (active) {
        T v[NQ];
        const int pj = t % NQ, pk = t / NQ;
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = buf0[(pk * NQ + pj) * NQ + l];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = v[0] * tab.D[i * NQ];
#pragma unroll
            for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[i * NQ + l], s);
            buf2[(pk * NQ + pj) * NQ + i] = s;
        }
    }
    __syncthreads();
    if (active) {
        T v[NQ];
        const int pi = t % NQ, pk = t / NQ;
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = buf1[(pk * NQ + l) * NQ + pi];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            T s = v[0] * tab.D[j * NQ];
#pragma unroll
            for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[j * NQ + l], s);
            buf2[(pk * NQ + j) * NQ + pi] += s;
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) acc[k] += buf2[k * TPE + t];
    }

    // 5. (B x B x B)^T of the combined quadrature-space field, then one coalesced store.
    hex_columns_to_modes<NM, NQ, T>(tab, acc, active, t, buf0, buf1,
                                    [&](int p, T v) { buf2[t * (NM + 1) + p] = v; });
    __syncthreads();
    store_modes_block<NM, NQ, T>(out + e0 * M, nv, [&](int sl) { return lds[sl][2]; });
  }
}

// Reference-space derivatives of the element's quadrature values u (held as this thread's
// k-column) along i and j, left in bi / bj in [k][j][i] order (PhysDerivTensor3DKernel,
// PhysDerivKernels.hpp:219-294); bu holds u in the same order.  Ends synchronised.
template <int NQ, typename T, typename DT>
__device__ __forceinline__ void hex_deriv_ij(const DT &D, bool active, int t, const T *bu,
                                             T *bi, T *bj)
{
    if (active) {
        T v[NQ];
        {
            const int pj = t % NQ, pk = t / NQ;
#pragma unroll
            for (int l = 0; l < NQ; ++l) v[l] = bu[(pk * NQ + pj) * NQ + l];
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                T s = v[0] * D[i];
#pragma unroll
                for (int l = 1; l < NQ; ++l) s = fma(v[l], D[l * NQ + i], s);
                bi[(pk * NQ + pj) * NQ + i] = s;
            }
        }
        {
            const int pi = t % NQ, pk = t / NQ;
#pragma unroll
            for (int l = 0; l < NQ; ++l) v[l] = bu[(pk * NQ + l) * NQ + pi];
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                T s = v[0] * D[j];
#pragma unroll
                for (int l = 1; l < NQ; ++l) s = fma(v[l], D[l * NQ + j], s);
                bj[(pk * NQ + j) * NQ + pi] = s;
            }
        }
    }
    __syncthreads();
}

// PhysDeriv (PhysDeriv.h:509-600, PhysDerivKernels.hpp:296-374): out_c = sum_a
// df[c*3+a] d_a with d_a the collocation derivatives of the quadrature values u.  in and
// out_c (at out + c*nElmt*Q) are point arrays; df is the GetDerivFactors array in the
// operator's precision: [d][nElmt*Q] (deformed) or [d][nElmt] (regular).
template <int NQ, typename T, int GEO>
__global__ __launch_bounds__(HexCfg<NQ>::BLOCK) void hex_physderiv_kernel(
    const DTab<NQ, T> tab, int nElmt, const T *__restrict__ in, const T *__restrict__ df,
    T *__restrict__ out)
{
    using C = HexCfg<NQ>;
    constexpr int Q = NQ * NQ * NQ, TPE = C::TPE;
    __shared__ T lds[C::EPB][3][Q];
    const int eb = threadIdx.x / TPE;
    const int t = threadIdx.x - eb * TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;
    T *bu = lds[es][0], *b0 = lds[es][1], *b1 = lds[es][2];

    T u[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) u[k] = valid ? ld_stream(in + e * Q + k * TPE + t) : T(0);
    if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) bu[k * TPE + t] = u[k];
    }
    __syncthreads();
    hex_deriv_ij<NQ, T>(tab.D, active, t, bu, b0, b1);
    if (!valid) return;
    const long nt = (long)nElmt * Q;
    // all loads before any store: gfx9 counts both in one in-order counter, so a load
    // issued after a store would wait for it
    T o[3][NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        T d2 = u[0] * tab.D[k];
#pragma unroll
        for (int l = 1; l < NQ; ++l) d2 = fma(u[l], tab.D[l * NQ + k], d2);
        const T d0 = b0[k * TPE + t], d1 = b1[k * TPE + t];
        T f[9];
        if (GEO == kDeformed) {
#pragma unroll
            for (int d = 0; d < 9; ++d) f[d] = ld_stream(df + d * nt + e * Q + k * TPE + t);
        } else {
#pragma unroll
            for (int d = 0; d < 9; ++d) f[d] = df[(long)d * nElmt + e];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c][k] = fma(d2, f[c * 3 + 2], fma(d1, f[c * 3 + 1], d0 * f[c * 3]));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < NQ; ++k) st_stream(out + c * nt + e * Q + k * TPE + t, o[c][k]);
}

// IProductWRTDerivBase (IProductWRTDerivBase.h:1231-1343): out = sum_a (D_a B)^T
// diag(w J) tmp_a with tmp_a = sum_c df[c*3+a] in_c (in_c at in + c*nElmt*Q), the three
// reference IProducts with dbdata collapsed into one D^T per direction and one
// (B x B x B)^T (dbdata = D bdata).  Packed geometry: deformed [e][GS], GS = 9Q rounded
// up to 128-byte lines, w J df[d] at d*Q; regular [e][9] = J df[d] (weights applied here).
template <int NM, int NQ, typename T, int GEO>
__global__ __launch_bounds__(HexCfg<NQ>::BLOCK) void hex_iprod_deriv_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = HexCfg<NQ>;
    constexpr int M = NM * NM * NM, Q = NQ * NQ * NQ, TPE = C::TPE;
    constexpr long GS = geo_stride(9 * Q, sizeof(T));
    __shared__ T lds[C::EPB][4][Q];
    const int eb = threadIdx.x / TPE;
    const int t = threadIdx.x - eb * TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;
    T *buf0 = lds[es][0], *buf1 = lds[es][1], *buf2 = lds[es][2], *buf3 = lds[es][3];
    const long nt = (long)nElmt * Q;

    T g0[NQ], g1[NQ], g2[NQ];
    if (valid) {
        const int ci = t % NQ, cj = t / NQ;
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const long o = e * Q + k * TPE + t;
            const T x0 = ld_stream(in + o), x1 = ld_stream(in + nt + o),
                    x2 = ld_stream(in + 2 * nt + o);
            T f[9];
            if (GEO == kDeformed) {
#pragma unroll
                for (int d = 0; d < 9; ++d) f[d] = ld_stream(geom + e * GS + d * Q + k * TPE + t);
            } else {
                const T wq = tab.w[ci] * tab.w[cj] * tab.w[k];
#pragma unroll
                for (int d = 0; d < 9; ++d) f[d] = geom[e * 9 + d] * wq;
            }
            g0[k] = fma(f[6], x2, fma(f[3], x1, f[0] * x0));
            g1[k] = fma(f[7], x2, fma(f[4], x1, f[1] * x0));
            g2[k] = fma(f[8], x2, fma(f[5], x1, f[2] * x0));
        }
    } else {
#pragma unroll
        for (int k = 0; k < NQ; ++k) g0[k] = g1[k] = g2[k] = T(0);
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            buf2[k * TPE + t] = g0[k];
            buf3[k * TPE + t] = g1[k];
        }
    }
    // D^T along k in registers
    T acc[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        T s = g2[0] * tab.D[k * NQ];
#pragma unroll
        for (int l = 1; l < NQ; ++l) s = fma(g2[l], tab.D[k * NQ + l], s);
        acc[k] = s;
    }
    __syncthreads();
    // D^T along i (pencils of thread (j,k)) and along j (pencils of thread (i,k))
    if (active) {
        T v[NQ];
        {
            const int pj = t % NQ, pk = t / NQ;
#pragma unroll
            for (int l = 0; l < NQ; ++l) v[l] = buf2[(pk * NQ + pj) * NQ + l];
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                T s = v[0] * tab.D[i * NQ];
#pragma unroll
                for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[i * NQ + l], s);
                buf0[(pk * NQ + pj) * NQ + i] = s;
            }
        }
        {
            const int pi = t % NQ, pk = t / NQ;
#pragma unroll
            for (int l = 0; l < NQ; ++l) v[l] = buf3[(pk * NQ + l) * NQ + pi];
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                T s = v[0] * tab.D[j * NQ];
#pragma unroll
                for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[j * NQ + l], s);
                buf1[(pk * NQ + j) * NQ + pi] = s;
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) acc[k] += buf0[k * TPE + t] + buf1[k * TPE + t];
    }
    T *dst = out + e * M + t * NM;
    hex_columns_to_modes<NM, NQ, T>(tab, acc, active, t, buf2, buf3, [&](int p, T v) {
        if (valid) st_stream(dst + p, v);
    });
}

}  // namespace imf
