/* TEST INFRASTRUCTURE: the oracle's hex Helmholtz (or_helmholtz_hex) with the reference's
 * AVX2 SIMD ABI -- four elements interleaved per 256-bit vector lane (vec_t::width = 4,
 * LibUtilities/SimdLib/avx2.hpp; load_interleave / deinterleave_store, avx2.hpp:591-624),
 * the same loop order and summation order as the width-1 restatement in mf_oracle.c
 * (BwdTransKernels.hpp:302-372, IProductKernels.hpp:236-311, PhysDerivKernels.hpp:219-294,
 * Helmholtz.h:796-990).  Used as the second CPU baseline beside the width-1 (SIMD off)
 * build, which is the reference's default (cmake/NektarSIMD.cmake:7-9). */
#include <stdlib.h>
#include <string.h>

#include "mf_oracle.h"

typedef double v4d __attribute__((vector_size(32)));
#define W 4

static void bwd_hex4(int nm, int nq, const double *B, const v4d *in, v4d *sum_irq,
                     v4d *sum_jir, v4d *out)
{
    for (int i = 0, cnt_irq = 0; i < nq; ++i)
        for (int r = 0, cnt_rqp = 0; r < nm; ++r)
            for (int q = 0; q < nm; ++q, ++cnt_irq) {
                v4d tmp = in[cnt_rqp] * B[i];
                ++cnt_rqp;
                for (int p = 1; p < nm; ++p, ++cnt_rqp) tmp += in[cnt_rqp] * B[p * nq + i];
                sum_irq[cnt_irq] = tmp;
            }
    for (int j = 0, cnt_jir = 0; j < nq; ++j)
        for (int i = 0, cnt_irq = 0; i < nq; ++i)
            for (int r = 0; r < nm; ++r, ++cnt_jir) {
                v4d tmp = sum_irq[cnt_irq] * B[j];
                ++cnt_irq;
                for (int q = 1; q < nm; ++q) tmp += sum_irq[cnt_irq++] * B[q * nq + j];
                sum_jir[cnt_jir] = tmp;
            }
    for (int k = 0, cnt_kji = 0; k < nq; ++k)
        for (int j = 0, cnt_jir = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i, ++cnt_kji) {
                v4d tmp = sum_jir[cnt_jir] * B[k];
                ++cnt_jir;
                for (int r = 1; r < nm; ++r) tmp += sum_jir[cnt_jir++] * B[r * nq + k];
                out[cnt_kji] = tmp;
            }
}

/* scale: multiply by sc; append: add to out */
static void iprod_hex4(int nm, int nq, const double *B0, const double *B1, const double *B2,
                       const double *w, int deformed, const v4d *jac, const v4d *in,
                       v4d *sums_kj, v4d *sums_k, v4d *out, int scale, int append, double sc)
{
    for (int p = 0; p < nm; ++p) {
        int cnt_kji = 0, cnt_kj = 0;
        for (int k = 0; k < nq; ++k)
            for (int j = 0; j < nq; ++j, ++cnt_kj) {
                v4d sum_kj = in[0] * 0.0;
                for (int i = 0; i < nq; ++i, ++cnt_kji) {
                    v4d jv = deformed ? jac[nq * nq * k + nq * j + i] : jac[0];
                    v4d prod = in[cnt_kji] * B0[i + nq * p] * jv;
                    sum_kj += prod * w[i];
                }
                sums_kj[cnt_kj] = sum_kj;
            }
        for (int q = 0; q < nm; ++q) {
            cnt_kj = 0;
            for (int k = 0; k < nq; ++k) {
                v4d sum_k = in[0] * 0.0;
                for (int j = 0; j < nq; ++j, ++cnt_kj) {
                    v4d prod = sums_kj[cnt_kj] * B1[q * nq + j];
                    sum_k += prod * w[j];
                }
                sums_k[k] = sum_k;
            }
            for (int r = 0; r < nm; ++r) {
                v4d sum = in[0] * 0.0;
                for (int k = 0; k < nq; ++k) {
                    v4d prod = sums_k[k] * B2[r * nq + k];
                    sum += prod * w[k];
                }
                v4d *o = &out[r * nm * nm + q * nm + p];
                if (scale && append) *o += sum * sc;
                else if (append) *o = *o + sum;
                else if (scale) *o = sum * sc;
                else *o = sum;
            }
        }
    }
}

static void deriv_hex4(int nq, const double *D, const v4d *in, v4d *d0, v4d *d1, v4d *d2)
{
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq * nq; ++j) {
            v4d s = in[0] * 0.0;
            for (int k = 0; k < nq; ++k) s += D[k * nq + i] * in[j * nq + k];
            d0[j * nq + i] = s;
        }
    for (int block = 0; block < nq; ++block) {
        int start = block * nq * nq;
        for (int i = 0; i < nq; ++i)
            for (int j = 0; j < nq; ++j) {
                v4d s = in[0] * 0.0;
                for (int k = 0; k < nq; ++k) s += in[start + k * nq + i] * D[k * nq + j];
                d1[start + j * nq + i] = s;
            }
    }
    for (int i = 0; i < nq * nq; ++i)
        for (int j = 0; j < nq; ++j) {
            v4d s = in[0] * 0.0;
            for (int k = 0; k < nq; ++k) s += in[k * nq * nq + i] * D[k * nq + j];
            d2[j * nq * nq + i] = s;
        }
}

void or_helmholtz_hex_avx2(int nm, int nq, const double *bdata, const double *dbdata,
                           const double *D, const double *w, int E, int deformed,
                           const double *jac, const double *df, double lambda,
                           const double *in, double *out)
{
    const int M = nm * nm * nm, Q = nq * nq * nq, NB = (E + W - 1) / W;
    const int nj = deformed ? Q : 1;
    const size_t np = deformed ? (size_t)E * Q : (size_t)E;
#pragma omp parallel
    {
        v4d *buf = (v4d *)aligned_alloc(32, sizeof(v4d) * (size_t)(8 * Q + 2 * M + 10 * nj));
        v4d *wsp1 = buf, *wsp2 = wsp1 + Q, *bwd = wsp2 + Q, *d0 = bwd + Q, *d1 = d0 + Q,
            *d2 = d1 + Q, *uin = d2 + Q, *o = uin + M, *jv = o + M, *dfv = jv + nj;
#pragma omp for schedule(static)
        for (int bk = 0; bk < NB; ++bk) {
            const int e0 = bk * W;
            /* load_interleave: lane l carries element e0 + l (missing elements: zeros) */
            for (int m = 0; m < M; ++m)
                for (int l = 0; l < W; ++l)
                    uin[m][l] = e0 + l < E ? in[(size_t)(e0 + l) * M + m] : 0.0;
            for (int q = 0; q < nj; ++q)
                for (int l = 0; l < W; ++l) {
                    const int e = e0 + l < E ? e0 + l : E - 1;
                    jv[q][l] = jac[(size_t)e * nj + q];
                    for (int a = 0; a < 9; ++a)
                        dfv[a * nj + q][l] = df[a * np + (size_t)e * nj + q];
                }
            bwd_hex4(nm, nq, bdata, uin, wsp1, wsp2, bwd);
            iprod_hex4(nm, nq, bdata, bdata, bdata, w, deformed, jv, bwd, wsp1, wsp2, o, 1, 0,
                       lambda);
            deriv_hex4(nq, D, bwd, d0, d1, d2);
            for (int q = 0; q < Q; ++q) {
                const int iq = deformed ? q : 0;
                v4d f[9];
                for (int a = 0; a < 9; ++a) f[a] = dfv[a * nj + iq];
                v4d m00 = f[0] * f[0]; m00 += f[3] * f[3]; m00 += f[6] * f[6];
                v4d m01 = f[0] * f[1]; m01 += f[3] * f[4]; m01 += f[6] * f[7];
                v4d m02 = f[0] * f[2]; m02 += f[3] * f[5]; m02 += f[6] * f[8];
                v4d m11 = f[1] * f[1]; m11 += f[4] * f[4]; m11 += f[7] * f[7];
                v4d m12 = f[1] * f[2]; m12 += f[4] * f[5]; m12 += f[7] * f[8];
                v4d m22 = f[2] * f[2]; m22 += f[5] * f[5]; m22 += f[8] * f[8];
                v4d a0 = d0[q], a1 = d1[q], a2 = d2[q];
                v4d t = m00 * a0; t += m01 * a1; t += m02 * a2; d0[q] = t;
                t = m01 * a0; t += m11 * a1; t += m12 * a2; d1[q] = t;
                t = m02 * a0; t += m12 * a1; t += m22 * a2; d2[q] = t;
            }
            iprod_hex4(nm, nq, dbdata, bdata, bdata, w, deformed, jv, d0, wsp1, wsp2, o, 0, 1, 1.0);
            iprod_hex4(nm, nq, bdata, dbdata, bdata, w, deformed, jv, d1, wsp1, wsp2, o, 0, 1, 1.0);
            iprod_hex4(nm, nq, bdata, bdata, dbdata, w, deformed, jv, d2, wsp1, wsp2, o, 0, 1, 1.0);
            /* deinterleave_store */
            for (int l = 0; l < W && e0 + l < E; ++l)
                for (int m = 0; m < M; ++m) out[(size_t)(e0 + l) * M + m] = o[m][l];
        }
        free(buf);
    }
}
