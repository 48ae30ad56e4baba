// C0 global assembly of hexahedral expansions and the matrix-free conjugate-gradient
// Helmholtz solve on the GPU.
//
// Reference (paths relative to the reference root):
//   AssemblyMapCG::v_GlobalToLocal / v_Assemble   MultiRegions/AssemblyMap/AssemblyMapCG.cpp:2853-2910
//   GlobalLinSysIterativeFull::v_DoMatrixMultiply MultiRegions/GlobalLinSysIterativeFull.cpp:215-232
//   NekLinSysIterCG::DoConjugateGradient          LibUtilities/LinearAlgebra/NekLinSysIterCG.cpp:104-260
//   NekLinSysIter::Set_Rhs_Magnitude              LibUtilities/LinearAlgebra/NekLinSysIter.cpp:128-153
// The global numbering is built here from the elements' vertex ids: vertex modes are
// shared through the vertex id, edge-interior modes through the (unordered) vertex pair with
// a sign (-1)^m where an element traverses the edge against the global direction, and
// face-interior modes through the face's 4 vertex ids in a frame fixed by those ids (origin
// at the smallest id, first axis towards its smaller neighbour), with a sign for each
// reversed axis and an index swap when the element's face axes are swapped.  This is the
// sign/permutation structure of the modified (C0) basis; the reference derives the same
// from its mesh graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/ithaca_mf.h"
#include "mf_op.hpp"

namespace {

thread_local std::string g_aerr;

int afail(int code, const std::string &m)
{
    g_aerr = m;
    return code;
}

// Local vertex of the reference hex with coordinate bits (b0, b1, b2): xi0 = +1 for v%4 in
// {1,2}, xi1 = +1 for v%4 in {2,3}, xi2 = +1 for v >= 4 (StdRegions/StdHexExp.cpp:1042-1082).
inline int vert_of(int b0, int b1, int b2)
{
    const int v = b1 ? (b0 ? 2 : 3) : (b0 ? 1 : 0);
    return v + 4 * b2;
}

struct H2 {
    size_t operator()(const std::array<long long, 2> &a) const
    {
        return std::hash<long long>()(a[0] * 1000003LL ^ a[1]);
    }
};
struct H4 {
    size_t operator()(const std::array<long long, 4> &a) const
    {
        size_t h = 0;
        for (long long v : a) h = h * 1000003ULL ^ std::hash<long long>()(v);
        return h;
    }
};

// l2g[loff[e] + m] = global index, sgn = +-1, bnd = 1 if the local mode lies on a face owned
// by one element only.  Element e has nm3[3e + d] modes along its local direction d and its
// coefficients start at loff[e] = sum_{f<e} nm3[3f] nm3[3f+1] nm3[3f+2] (ExpList's
// coefficient offsets), p fastest.  Variable order follows AssemblyMapCG.cpp:1380-1440 and
// :1771-1791: an edge carries the fewest interior modes of the elements that share it, a
// face the fewest of its two elements along each of its two axes (counted in the face's own
// frame, fixed by the vertex ids), and an element's modes beyond those map to global 0 with
// sign 0 (they take no part in the global system and are not marked as boundary modes).
int build_map_hex(int E, const int *nm3, const long long *vid, std::vector<int> &l2g,
                  std::vector<signed char> &sgn, std::vector<unsigned char> &bnd, long &nglobal)
{
    if (E <= 0 || !nm3 || !vid) return afail(IMF_ERR_ARG, "assembly: bad argument");
    std::vector<long> loff(E + 1, 0);
    for (long e = 0; e < E; ++e) {
        for (int d = 0; d < 3; ++d)
            if (nm3[3 * e + d] < 2) return afail(IMF_ERR_ARG, "assembly: fewer than 2 modes on an element");
        loff[e + 1] = loff[e] + (long)nm3[3 * e] * nm3[3 * e + 1] * nm3[3 * e + 2];
    }
    std::unordered_map<long long, int> vmap;
    std::unordered_map<std::array<long long, 2>, int, H2> emap;
    std::unordered_map<std::array<long long, 4>, int, H4> fmap;
    std::vector<int> ecnt, fuse;               // per id: interior modes, users
    std::vector<std::array<int, 2>> fcnt;      // per face: interior modes along its axes A, B
    // the face of element e normal to local direction df on side s in the face's own frame:
    // corner ids g[c1][c2] over the in-plane directions d1 < d2, the origin (o1, o2) at the
    // smallest id, tr = the first axis runs along d2
    struct FaceFrame {
        long long g[2][2];
        std::array<long long, 4> key;
        int d1, d2, o1, o2;
        bool tr;
    };
    auto face_frame = [&](const long long *v, int df, int side) {
        FaceFrame f;
        f.d1 = df == 0 ? 1 : 0;
        f.d2 = df == 2 ? 1 : 2;
        for (int c1 = 0; c1 < 2; ++c1)
            for (int c2 = 0; c2 < 2; ++c2) {
                int bits[3];
                bits[df] = side;
                bits[f.d1] = c1;
                bits[f.d2] = c2;
                f.g[c1][c2] = v[vert_of(bits[0], bits[1], bits[2])];
                f.key[c1 * 2 + c2] = f.g[c1][c2];
            }
        std::sort(f.key.begin(), f.key.end());
        f.o1 = 0;
        f.o2 = 0;
        for (int c1 = 0; c1 < 2; ++c1)
            for (int c2 = 0; c2 < 2; ++c2)
                if (f.g[c1][c2] < f.g[f.o1][f.o2]) {
                    f.o1 = c1;
                    f.o2 = c2;
                }
        f.tr = f.g[f.o1][1 - f.o2] < f.g[1 - f.o1][f.o2];
        return f;
    };
    for (long e = 0; e < E; ++e) {
        const long long *v = vid + e * 8;
        const int *nme = nm3 + 3 * e;
        for (int a = 0; a < 8; ++a) vmap.emplace(v[a], (int)vmap.size());
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < 4; ++c) {  // the 4 edges along direction d
                int bits[3], bb[3];
                const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
                bits[o1] = c & 1;
                bits[o2] = c >> 1;
                bits[d] = 0;
                std::copy(bits, bits + 3, bb);
                bb[d] = 1;
                long long a0 = v[vert_of(bits[0], bits[1], bits[2])];
                long long a1 = v[vert_of(bb[0], bb[1], bb[2])];
                if (a0 == a1) return afail(IMF_ERR_ARG, "assembly: degenerate edge");
                const int ni = nme[d] - 2;
                auto ins = emap.emplace(std::array<long long, 2>{std::min(a0, a1), std::max(a0, a1)},
                                        (int)emap.size());
                if (ins.second) ecnt.push_back(ni);
                else ecnt[ins.first->second] = std::min(ecnt[ins.first->second], ni);
            }
        for (int df = 0; df < 3; ++df)
            for (int s = 0; s < 2; ++s) {
                const FaceFrame f = face_frame(v, df, s);
                const int n1 = nme[f.d1] - 2, n2 = nme[f.d2] - 2;
                const std::array<int, 2> nab = f.tr ? std::array<int, 2>{n2, n1}
                                                    : std::array<int, 2>{n1, n2};
                auto ins = fmap.emplace(f.key, (int)fmap.size());
                if (ins.second) {
                    fcnt.push_back(nab);
                    fuse.push_back(1);
                } else {
                    const int id = ins.first->second;
                    fcnt[id][0] = std::min(fcnt[id][0], nab[0]);
                    fcnt[id][1] = std::min(fcnt[id][1], nab[1]);
                    if (++fuse[id] > 2)
                        return afail(IMF_ERR_ARG, "assembly: a face is shared by more than two elements");
                }
            }
    }
    const long nV = (long)vmap.size(), nE = (long)emap.size(), nF = (long)fmap.size();
    std::vector<long> eoff(nE + 1), foff(nF + 1), ioff(E + 1);
    eoff[0] = nV;
    for (long i = 0; i < nE; ++i) eoff[i + 1] = eoff[i] + ecnt[i];
    foff[0] = eoff[nE];
    for (long i = 0; i < nF; ++i) foff[i + 1] = foff[i] + (long)fcnt[i][0] * fcnt[i][1];
    ioff[0] = foff[nF];
    for (long e = 0; e < E; ++e)
        ioff[e + 1] = ioff[e] + (long)(nm3[3 * e] - 2) * (nm3[3 * e + 1] - 2) * (nm3[3 * e + 2] - 2);
    nglobal = ioff[E];
    if (nglobal > 2147483647L) return afail(IMF_ERR_ARG, "assembly: more than 2^31 global modes");
    l2g.assign((size_t)loff[E], 0);
    sgn.assign((size_t)loff[E], 1);
    bnd.assign((size_t)loff[E], 0);
    for (long e = 0; e < E; ++e) {
        const long long *v = vid + e * 8;
        const int n0 = nm3[3 * e], n1 = nm3[3 * e + 1], n2 = nm3[3 * e + 2];
        for (int r = 0; r < n2; ++r)
            for (int q = 0; q < n1; ++q)
                for (int p = 0; p < n0; ++p) {
                    const int m[3] = {p, q, r};
                    const size_t li = (size_t)loff[e] + ((long)r * n1 + q) * n0 + p;
                    int nv = 0;
                    for (int d = 0; d < 3; ++d) nv += m[d] < 2;
                    long gi = 0;
                    int sg = 1;
                    if (nv == 3) {
                        gi = vmap[v[vert_of(p, q, r)]];
                    } else if (nv == 2) {
                        int di = 0;
                        while (m[di] < 2) ++di;
                        int b0[3] = {m[0], m[1], m[2]}, b1[3] = {m[0], m[1], m[2]};
                        b0[di] = 0;
                        b1[di] = 1;
                        const long long g0 = v[vert_of(b0[0], b0[1], b0[2])];
                        const long long g1 = v[vert_of(b1[0], b1[1], b1[2])];
                        const int idx = emap[{std::min(g0, g1), std::max(g0, g1)}];
                        if (m[di] - 2 < ecnt[idx]) {
                            gi = eoff[idx] + (m[di] - 2);
                            if (g0 > g1 && (m[di] & 1)) sg = -1;
                        } else {
                            sg = 0;
                        }
                    } else if (nv == 1) {
                        int df = 0;
                        while (m[df] >= 2) ++df;
                        const FaceFrame f = face_frame(v, df, m[df]);
                        const int idx = fmap[f.key];
                        const int mA = f.tr ? m[f.d2] : m[f.d1], mB = f.tr ? m[f.d1] : m[f.d2];
                        const bool fA = f.tr ? f.o2 == 1 : f.o1 == 1, fB = f.tr ? f.o1 == 1 : f.o2 == 1;
                        if (mA - 2 < fcnt[idx][0] && mB - 2 < fcnt[idx][1]) {
                            gi = foff[idx] + (long)(mA - 2) * fcnt[idx][1] + (mB - 2);
                            if (fA && (mA & 1)) sg = -sg;
                            if (fB && (mB & 1)) sg = -sg;
                        } else {
                            sg = 0;
                        }
                    } else {
                        gi = ioff[e] + ((long)(r - 2) * (n1 - 2) + (q - 2)) * (n0 - 2) + (p - 2);
                    }
                    l2g[li] = (int)gi;
                    sgn[li] = (signed char)sg;
                }
        // modes on boundary faces (faces with one user): the vertex-type index of the
        // face's normal direction equals the face side
        for (int df = 0; df < 3; ++df)
            for (int s = 0; s < 2; ++s) {
                const FaceFrame f = face_frame(v, df, s);
                if (fuse[fmap[f.key]] != 1) continue;
                for (int r = 0; r < n2; ++r)
                    for (int q = 0; q < n1; ++q)
                        for (int p = 0; p < n0; ++p) {
                            const int m[3] = {p, q, r};
                            const size_t li = (size_t)loff[e] + ((long)r * n1 + q) * n0 + p;
                            if (m[df] == s && sgn[li] != 0) bnd[li] = 1;
                        }
            }
    }
    return IMF_OK;
}

// Local vertex of the reference quad with coordinate bits (b0, b1), counter-clockwise from
// (-1,-1) (StdRegions/StdQuadExp.cpp vertex order, the bottom face of the hex above).
inline int vert2(int b0, int b1) { return vert_of(b0, b1, 0); }

// The quad analogue of build_map_hex: vertex, edge ((-1)^m sign when the element runs the
// edge against the min -> max id direction) and interior modes; boundary = edges with one
// user; variable order as above (an edge carries the fewest interior modes of its two
// elements).
int build_map_quad(int E, const int *nmv, const long long *vid, std::vector<int> &l2g,
                   std::vector<signed char> &sgn, std::vector<unsigned char> &bnd, long &nglobal)
{
    if (E <= 0 || !nmv || !vid) return afail(IMF_ERR_ARG, "assembly: bad argument");
    std::vector<long> loff(E + 1, 0);
    for (long e = 0; e < E; ++e) {
        if (nmv[e] < 2) return afail(IMF_ERR_ARG, "assembly: fewer than 2 modes on an element");
        loff[e + 1] = loff[e] + (long)nmv[e] * nmv[e];
    }
    // end points of the edge where coordinate dn is fixed at s (running along 1 - dn)
    auto ends = [&](const long long *v, int dn, int s, long long &g0, long long &g1) {
        int b0[2], b1[2];
        b0[dn] = b1[dn] = s;
        b0[1 - dn] = 0;
        b1[1 - dn] = 1;
        g0 = v[vert2(b0[0], b0[1])];
        g1 = v[vert2(b1[0], b1[1])];
    };
    std::unordered_map<long long, int> vmap;
    std::unordered_map<std::array<long long, 2>, int, H2> emap;
    std::vector<int> ecnt, euse;
    for (long e = 0; e < E; ++e) {
        const long long *v = vid + e * 4;
        for (int a = 0; a < 4; ++a) vmap.emplace(v[a], (int)vmap.size());
        for (int dn = 0; dn < 2; ++dn)
            for (int sd = 0; sd < 2; ++sd) {
                long long g0, g1;
                ends(v, dn, sd, g0, g1);
                if (g0 == g1) return afail(IMF_ERR_ARG, "assembly: degenerate edge");
                const std::array<long long, 2> k{std::min(g0, g1), std::max(g0, g1)};
                auto ins = emap.emplace(k, (int)emap.size());
                if (ins.second) {
                    ecnt.push_back(nmv[e] - 2);
                    euse.push_back(1);
                } else {
                    const int id = ins.first->second;
                    ecnt[id] = std::min(ecnt[id], nmv[e] - 2);
                    if (++euse[id] > 2)
                        return afail(IMF_ERR_ARG, "assembly: an edge is shared by more than two elements");
                }
            }
    }
    const long nV = (long)vmap.size(), nE = (long)emap.size();
    std::vector<long> eoff(nE + 1), ioff(E + 1);
    eoff[0] = nV;
    for (long i = 0; i < nE; ++i) eoff[i + 1] = eoff[i] + ecnt[i];
    ioff[0] = eoff[nE];
    for (long e = 0; e < E; ++e) ioff[e + 1] = ioff[e] + (long)(nmv[e] - 2) * (nmv[e] - 2);
    nglobal = ioff[E];
    if (nglobal > 2147483647L) return afail(IMF_ERR_ARG, "assembly: more than 2^31 global modes");
    l2g.assign((size_t)loff[E], 0);
    sgn.assign((size_t)loff[E], 1);
    bnd.assign((size_t)loff[E], 0);
    for (long e = 0; e < E; ++e) {
        const long long *v = vid + e * 4;
        const int nm = nmv[e], ni = nm - 2;
        for (int q = 0; q < nm; ++q)
            for (int p = 0; p < nm; ++p) {
                const int m[2] = {p, q};
                const size_t li = (size_t)loff[e] + q * nm + p;
                const int nv = (p < 2) + (q < 2);
                long gi = 0;
                int sg = 1;
                if (nv == 2) {
                    gi = vmap[v[vert2(p, q)]];
                } else if (nv == 1) {
                    const int di = p >= 2 ? 0 : 1, dn = 1 - di;
                    long long g0, g1;
                    ends(v, dn, m[dn], g0, g1);
                    const int idx = emap[{std::min(g0, g1), std::max(g0, g1)}];
                    if (m[di] - 2 < ecnt[idx]) {
                        gi = eoff[idx] + (m[di] - 2);
                        if (g0 > g1 && (m[di] & 1)) sg = -1;
                    } else {
                        sg = 0;
                    }
                } else {
                    gi = ioff[e] + (long)(q - 2) * ni + (p - 2);
                }
                l2g[li] = (int)gi;
                sgn[li] = (signed char)sg;
            }
        for (int dn = 0; dn < 2; ++dn)
            for (int sd = 0; sd < 2; ++sd) {
                long long g0, g1;
                ends(v, dn, sd, g0, g1);
                if (euse[emap[{std::min(g0, g1), std::max(g0, g1)}]] != 1) continue;
                for (int q = 0; q < nm; ++q)
                    for (int p = 0; p < nm; ++p) {
                        const size_t li = (size_t)loff[e] + q * nm + p;
                        if ((dn == 0 ? p : q) == sd && sgn[li] != 0) bnd[li] = 1;
                    }
            }
    }
    return IMF_OK;
}

// nm: one order per element, or (hexes, per_dir) three per element (directions 0, 1, 2)
int build_map(int shape, int E, const int *nm, bool per_dir, const long long *vid,
              std::vector<int> &l2g, std::vector<signed char> &sgn,
              std::vector<unsigned char> &bnd, long &nglobal)
{
    if (shape != imf::kHex) {
        if (per_dir) return afail(IMF_ERR_ARG, "assembly: per-direction orders need hexahedra");
        return build_map_quad(E, nm, vid, l2g, sgn, bnd, nglobal);
    }
    if (per_dir) return build_map_hex(E, nm, vid, l2g, sgn, bnd, nglobal);
    if (E <= 0 || !nm) return afail(IMF_ERR_ARG, "assembly: bad argument");
    std::vector<int> nm3(3 * (size_t)E);
    for (long e = 0; e < E; ++e) nm3[3 * e] = nm3[3 * e + 1] = nm3[3 * e + 2] = nm[e];
    return build_map_hex(E, nm3.data(), vid, l2g, sgn, bnd, nglobal);
}

// --------------------------------------------------------------------------- kernels
// local = sign * global[l2g]   (Vmath::Gathr with signs)
__global__ void g2l_kernel(long n, const int *__restrict__ l2g, const signed char *__restrict__ sg,
                           const double *__restrict__ glob, double *__restrict__ loc)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) loc[i] = sg[i] * glob[l2g[i]];
}

// global[g] = sum over the local modes mapped to g of sign * local (Vmath::Assmb), as a
// deterministic gather over the inverse map (CSR: off[g]..off[g+1] into inv).
// With sg == NULL the contributions are added unsigned (the assembled diagonal:
// sign^2 = 1).
__device__ __forceinline__ double assemble_one(long g, const int *__restrict__ off,
                                               const int *__restrict__ inv,
                                               const signed char *__restrict__ sg,
                                               const double *__restrict__ loc,
                                               const unsigned char *__restrict__ mask)
{
    double s = 0.0;
    for (int k = off[g]; k < off[g + 1]; ++k) {
        const int i = inv[k];
        s += sg ? sg[i] * loc[i] : loc[i];
    }
    return (mask && mask[g]) ? 0.0 : s;
}

__global__ void assemble_kernel(long ng, const int *__restrict__ off, const int *__restrict__ inv,
                                const signed char *__restrict__ sg, const double *__restrict__ loc,
                                const unsigned char *__restrict__ mask, double *__restrict__ glob)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) glob[g] = assemble_one(g, off, inv, sg, loc, mask);
}

__global__ void zero_kernel(long n, double *__restrict__ x)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

struct DiagTab {
    double B[128], dB[128], w[16];
};

// Local diagonal of the Helmholtz operator lambda M + sum_ab (D_a B)^T diag(w J G^ab) (D_b B)
// from the operator's packed geometry (deformed [e][GS], GS = 7Q rounded to 128-byte lines;
// regular [e][7] with the weights applied here); one thread per (element, mode).
__global__ void helm_diag_hex(DiagTab t, int nm, int nq, long nElmt, int deformed, double lambda,
                              const double *__restrict__ geom, double *__restrict__ diag)
{
    const long M = (long)nm * nm * nm, Q = (long)nq * nq * nq;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nElmt * M) return;
    const long e = g / M;
    const int md = (int)(g - e * M), p = md % nm, q = (md / nm) % nm, r = md / (nm * nm);
    const long GS = imf::geo_stride(7 * Q, sizeof(double));
    double s = 0.0;
    for (int k = 0; k < nq; ++k)
        for (int j = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i) {
                const long pt = ((long)k * nq + j) * nq + i;
                const double bp = t.B[p * nq + i], bq = t.B[q * nq + j], br = t.B[r * nq + k];
                const double phi = bp * bq * br;
                const double d0 = t.dB[p * nq + i] * bq * br, d1 = bp * t.dB[q * nq + j] * br,
                             d2 = bp * bq * t.dB[r * nq + k];
                double c[7];
                if (deformed) {
                    for (int a = 0; a < 7; ++a) c[a] = geom[e * GS + a * Q + pt];
                } else {
                    const double wq = t.w[i] * t.w[j] * t.w[k];
                    for (int a = 0; a < 7; ++a) c[a] = geom[e * 7 + a] * wq;
                }
                s += lambda * c[0] * phi * phi + c[1] * d0 * d0 + c[4] * d1 * d1 + c[6] * d2 * d2 +
                     2.0 * (c[2] * d0 * d1 + c[3] * d0 * d2 + c[5] * d1 * d2);
            }
    diag[g] = s;
}

// The quad analogue (deformed [e][4Q]: Jw, Jw G00, Jw G01, Jw G11; regular [e][4]).
__global__ void helm_diag_quad(DiagTab t, int nm, int nq, long nElmt, int deformed,
                               double lambda, const double *__restrict__ geom,
                               double *__restrict__ diag)
{
    const long M = (long)nm * nm, Q = (long)nq * nq;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nElmt * M) return;
    const long e = g / M;
    const int md = (int)(g - e * M), p = md % nm, q = md / nm;
    double s = 0.0;
    for (int j = 0; j < nq; ++j)
        for (int i = 0; i < nq; ++i) {
            const long pt = (long)j * nq + i;
            const double bp = t.B[p * nq + i], bq = t.B[q * nq + j];
            const double phi = bp * bq, d0 = t.dB[p * nq + i] * bq, d1 = bp * t.dB[q * nq + j];
            double c[4];
            if (deformed) {
                for (int a = 0; a < 4; ++a) c[a] = geom[e * 4 * Q + a * Q + pt];
            } else {
                const double wq = t.w[i] * t.w[j];
                for (int a = 0; a < 4; ++a) c[a] = geom[e * 4 + a] * wq;
            }
            s += lambda * c[0] * phi * phi + c[1] * d0 * d0 + c[3] * d1 * d1 +
                 2.0 * c[2] * d0 * d1;
        }
    diag[g] = s;
}

__global__ void inv_masked(long n, const unsigned char *__restrict__ mask, double *__restrict__ d)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (mask && mask[i]) ? 0.0 : 1.0 / d[i];
}

// r = mask0(b), x = 0
__global__ void cg_init(long n, const unsigned char *__restrict__ mask, const double *__restrict__ b,
                        double *__restrict__ r, double *__restrict__ x)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = (mask && mask[i]) ? 0.0 : b[i];
    x[i] = 0.0;
}

// w = dinv * r
__global__ void cg_precon(long n, const double *__restrict__ dinv, const double *__restrict__ r,
                          double *__restrict__ w)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = dinv[i] * r[i];
}

// p = w + beta p; q = s + beta q; x += alpha p; r -= alpha q; w = dinv r
// (NekLinSysIterCG.cpp:186-204, with the Jacobi preconditioner fused)
__global__ void cg_update(long n, double alpha, double beta, const double *__restrict__ dinv,
                          double *__restrict__ p, double *__restrict__ q, double *__restrict__ w,
                          const double *__restrict__ s, double *__restrict__ x, double *__restrict__ r)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pi = w[i] + beta * p[i], qi = s[i] + beta * q[i];
    p[i] = pi;
    q[i] = qi;
    x[i] += alpha * pi;
    const double ri = r[i] - alpha * qi;
    r[i] = ri;
    w[i] = dinv[i] * ri;
}

constexpr int RB = 256;

// partial sums of (a.b, c.b, a.a) per block (Dirichlet entries are zero in all vectors)
__global__ void dot3_partial(long n, const double *__restrict__ a, const double *__restrict__ b,
                             const double *__restrict__ c, double *__restrict__ part)
{
    __shared__ double sh[3][RB / 64];
    double s0 = 0, s1 = 0, s2 = 0;
    for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
        const double ai = a[i], bi = b[i];
        s0 += ai * bi;
        s1 += c[i] * bi;
        s2 += ai * ai;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
        sh[2][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
        for (int w = 0; w < RB / 64; ++w) t += sh[threadIdx.x][w];
        part[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// s = Assemble(loc) as assemble_kernel, and in the same pass the block partial sums of
// (r,w), (s,w), (r,r) for the conjugate-gradient step (grid-stride over the global modes,
// part laid out as in dot3_partial, gridDim.x blocks: a fixed order, so reproducible).
__global__ void assemble_dot3_kernel(long ng, const int *__restrict__ off,
                                     const int *__restrict__ inv,
                                     const signed char *__restrict__ sg,
                                     const double *__restrict__ loc,
                                     const unsigned char *__restrict__ mask,
                                     const double *__restrict__ extra,
                                     const double *__restrict__ r, const double *__restrict__ w,
                                     double *__restrict__ s_out, double *__restrict__ part)
{
    __shared__ double sh[3][RB / 64];
    double s0 = 0, s1 = 0, s2 = 0;
    for (long g = (long)blockIdx.x * RB + threadIdx.x; g < ng; g += (long)gridDim.x * RB) {
        double sv = assemble_one(g, off, inv, sg, loc, mask);
        // the Robin trace term of the global mode, when there is one (zero on Dirichlet modes)
        if (extra && !(mask && mask[g])) sv += extra[g];
        s_out[g] = sv;
        const double rg = r[g], wg = w[g];
        s0 += rg * wg;
        s1 += sv * wg;
        s2 += rg * rg;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
        sh[2][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
        for (int k = 0; k < RB / 64; ++k) t += sh[threadIdx.x][k];
        part[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void dot3_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double sh[RB / 64];
    for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int i = threadIdx.x; i < nb; i += RB) s += part[(long)c * nb + i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int w = 0; w < RB / 64; ++w) t += sh[w];
            out[c] = t;
        }
        __syncthreads();
    }
}

inline unsigned nblk(long n) { return (unsigned)((n + RB - 1) / RB); }

// y = R x for the Robin matrix R in CSR form over the global modes
__global__ void csr_spmv(long ng, const int *__restrict__ off, const int *__restrict__ col,
                         const double *__restrict__ val, const double *__restrict__ x,
                         double *__restrict__ y)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    double s = 0.0;
    for (int k = off[g]; k < off[g + 1]; ++k) s += val[k] * x[col[k]];
    y[g] = s;
}

// d += diag(R)
__global__ void csr_diag_add(long ng, const int *__restrict__ off, const int *__restrict__ col,
                             const double *__restrict__ val, double *__restrict__ d)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    for (int k = off[g]; k < off[g + 1]; ++k)
        if (col[k] == g) d[g] += val[k];
}

// blocks of the fused assembly + dot pass (16 per CU on a 256-CU device)
constexpr long kAsmParts = 4096;

}  // namespace

struct imf_asm_s {
    int shape = 0, nElmt = 0, nm = 0, dirichlet = 0;  // nm = 0: orders vary by element
    long nloc = 0, ng = 0;
    std::vector<int> nmv;    // modes per direction of each element (isotropic elements)
    std::vector<int> nm3;    // hexahedra: modes along directions 0, 1, 2 of each element
    std::vector<long> loff;  // first local coefficient of each element (nElmt + 1)
    int *d_l2g = nullptr;
    signed char *d_sg = nullptr;
    int *d_off = nullptr, *d_inv = nullptr;  // inverse map (CSR), 32-bit
    unsigned char *d_mask = nullptr;  // global Dirichlet mask (NULL without Dirichlet)
    // Robin trace terms: a sparse matrix over the global modes (CSR), and R w of the solve
    long r_nnz = 0;
    int *d_roff = nullptr, *d_rcol = nullptr;
    double *d_rval = nullptr, *d_rw = nullptr;
    // solver workspace (allocated on the first solve)
    double *d_ws = nullptr, *d_loc = nullptr, *d_part = nullptr, *d_dots = nullptr;
    int nparts = 0;
    // operators, their geometry versions and lambdas the cached diagonal belongs to
    std::vector<std::tuple<const void *, unsigned long, double>> diag_key;
};

namespace {

#define ACHECK(call, what)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return afail(IMF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

// out = mask0(Assemble(H(GlobalToLocal(w))))  (GlobalLinSysIterativeFull::v_DoMatrixMultiply)
// with the partial sums of (r,w), (out,w), (r,r) taken in the assembly pass.  ops[g] acts on
// the next ops[g]->nElmt elements (one operator per order, as the reference keeps one
// collection per expansion type).
int apply_global_dots(imf_asm a, int nops, const imf_op *ops, const double *w, const double *r,
                      double *out, hipStream_t s)
{
    double *l0 = a->d_loc, *l1 = a->d_loc + a->nloc;
    const bool any_aniso = std::any_of(ops, ops + nops, [](const imf_op o) { return o->aniso; });
    if (a->shape == imf::kHex && !any_aniso) {  // GlobalToLocal fused into the coefficient load
        long e0 = 0;
        for (int g = 0; g < nops; ++g) {
            const long o = a->loff[e0];
            if (int st = imf_apply_gather_device(ops[g], w, a->d_l2g + o, a->d_sg + o, l1 + o, s))
                return st;
            e0 += ops[g]->nElmt;
        }
    } else {
        g2l_kernel<<<nblk(a->nloc), RB, 0, s>>>(a->nloc, a->d_l2g, a->d_sg, w, l0);
        long e0 = 0;
        for (int g = 0; g < nops; ++g) {
            const long o = a->loff[e0];
            if (int st = imf_apply_device(ops[g], l0 + o, l1 + o, s)) return st;
            e0 += ops[g]->nElmt;
        }
    }
    if (a->r_nnz)
        csr_spmv<<<nblk(a->ng), RB, 0, s>>>(a->ng, a->d_roff, a->d_rcol, a->d_rval, w, a->d_rw);
    const int nb = (int)std::min<long>(kAsmParts, nblk(a->ng));
    assemble_dot3_kernel<<<nb, RB, 0, s>>>(a->ng, a->d_off, a->d_inv, a->d_sg, l1, a->d_mask,
                                           a->r_nnz ? a->d_rw : nullptr, r, w, out, a->d_part);
    dot3_final<<<1, RB, 0, s>>>(nb, a->d_part, a->d_dots);
    ACHECK(hipGetLastError(), "assembly kernels");
    return IMF_OK;
}

}  // namespace

extern "C" {

const char *imf_assembly_last_error(void) { return g_aerr.c_str(); }

static int map_any(int shape, int nElmt, const int *nm, const long long *vertex_ids, int *l2g,
                   signed char *sign, unsigned char *on_boundary, long *nglobal,
                   bool per_dir = false)
{
    std::vector<int> m;
    std::vector<signed char> sg;
    std::vector<unsigned char> bd;
    long ng = 0;
    if (int st = build_map(shape, nElmt, nm, per_dir, vertex_ids, m, sg, bd, ng)) return st;
    if (l2g) std::memcpy(l2g, m.data(), m.size() * sizeof(int));
    if (sign) std::memcpy(sign, sg.data(), sg.size());
    if (on_boundary) std::memcpy(on_boundary, bd.data(), bd.size());
    if (nglobal) *nglobal = ng;
    return IMF_OK;
}

static int create_any(int shape, int nElmt, const int *nm, const long long *vertex_ids,
                      int dirichlet, imf_asm *out, bool per_dir = false)
{
    if (!out) return afail(IMF_ERR_ARG, "imf_assembly_create: null output");
    std::vector<int> m;
    std::vector<signed char> sg;
    std::vector<unsigned char> bd;
    long ng = 0;
    if (int st = build_map(shape, nElmt, nm, per_dir, vertex_ids, m, sg, bd, ng)) return st;
    if (m.size() > 2147483647UL)
        return afail(IMF_ERR_ARG, "imf_assembly_create: more than 2^31 local modes");
    auto *a = new imf_asm_s();
    a->shape = shape;
    a->nElmt = nElmt;
    if (per_dir) {
        a->nm3.assign(nm, nm + 3 * (size_t)nElmt);
        // an element with equal orders in its three directions also counts as isotropic
        a->nmv.resize(nElmt);
        for (int e = 0; e < nElmt; ++e)
            a->nmv[e] = (nm[3 * e] == nm[3 * e + 1] && nm[3 * e] == nm[3 * e + 2]) ? nm[3 * e] : 0;
    } else {
        a->nmv.assign(nm, nm + nElmt);
        if (shape == imf::kHex) {
            a->nm3.resize(3 * (size_t)nElmt);
            for (int e = 0; e < nElmt; ++e) a->nm3[3 * e] = a->nm3[3 * e + 1] = a->nm3[3 * e + 2] = nm[e];
        }
    }
    a->nm = std::all_of(a->nmv.begin(), a->nmv.end(), [&](int v) { return v == a->nmv[0]; })
                ? a->nmv[0] : 0;
    a->loff.assign(nElmt + 1, 0);
    for (int e = 0; e < nElmt; ++e)
        a->loff[e + 1] = a->loff[e] + (shape == imf::kHex
                                           ? (long)a->nm3[3 * e] * a->nm3[3 * e + 1] * a->nm3[3 * e + 2]
                                           : (long)nm[e] * nm[e]);
    a->dirichlet = dirichlet ? 1 : 0;
    a->nloc = (long)m.size();
    a->ng = ng;
    // inverse map (CSR) for the deterministic assembly; modes with sign 0 (beyond the order
    // of a shared edge or face) are left out
    long nused = 0;
    std::vector<int> off(ng + 1, 0);
    for (long i = 0; i < a->nloc; ++i)
        if (sg[i]) ++off[m[i] + 1], ++nused;
    std::vector<int> inv(std::max<long>(nused, 1));
    for (long g = 0; g < ng; ++g) off[g + 1] += off[g];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (long i = 0; i < a->nloc; ++i)
        if (sg[i]) inv[fill[m[i]]++] = (int)i;
    std::vector<unsigned char> gmask;
    if (a->dirichlet) {
        gmask.assign(ng, 0);
        for (long i = 0; i < a->nloc; ++i)
            if (bd[i] && sg[i]) gmask[m[i]] = 1;
    }
    auto cleanup = [&](int st) {
        imf_assembly_destroy(a);
        return st;
    };
    hipError_t e;
    if ((e = hipMalloc(&a->d_l2g, a->nloc * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&a->d_sg, a->nloc)) != hipSuccess ||
        (e = hipMalloc(&a->d_off, (ng + 1) * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&a->d_inv, inv.size() * sizeof(int))) != hipSuccess ||
        (a->dirichlet && (e = hipMalloc(&a->d_mask, ng)) != hipSuccess))
        return cleanup(afail(IMF_ERR_HIP, std::string("hipMalloc(assembly): ") + hipGetErrorString(e)));
    if ((e = hipMemcpy(a->d_l2g, m.data(), a->nloc * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(a->d_sg, sg.data(), a->nloc, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(a->d_off, off.data(), (ng + 1) * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(a->d_inv, inv.data(), inv.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess ||
        (a->dirichlet && (e = hipMemcpy(a->d_mask, gmask.data(), ng, hipMemcpyHostToDevice)) != hipSuccess))
        return cleanup(afail(IMF_ERR_HIP, std::string("hipMemcpy(assembly): ") + hipGetErrorString(e)));
    *out = a;
    return IMF_OK;
}

static std::vector<int> uniform(int nElmt, int nm)
{
    return std::vector<int>(nElmt > 0 ? nElmt : 0, nm);
}

static int shape_of_dim(int dim, int &shape)
{
    if (dim != 2 && dim != 3) return afail(IMF_ERR_ARG, "assembly: dim must be 2 (quad) or 3 (hex)");
    shape = dim == 3 ? imf::kHex : imf::kQuad;
    return IMF_OK;
}

int imf_assembly_map_hex(int nElmt, int nm, const long long *vertex_ids, int *l2g,
                         signed char *sign, unsigned char *on_boundary, long *nglobal)
{
    return map_any(imf::kHex, nElmt, uniform(nElmt, nm).data(), vertex_ids, l2g, sign,
                   on_boundary, nglobal);
}

int imf_assembly_map_quad(int nElmt, int nm, const long long *vertex_ids, int *l2g,
                          signed char *sign, unsigned char *on_boundary, long *nglobal)
{
    return map_any(imf::kQuad, nElmt, uniform(nElmt, nm).data(), vertex_ids, l2g, sign,
                   on_boundary, nglobal);
}

int imf_assembly_map_varp(int dim, int nElmt, const int *nm, const long long *vertex_ids,
                          int *l2g, signed char *sign, unsigned char *on_boundary, long *nglobal)
{
    int shape;
    if (int st = shape_of_dim(dim, shape)) return st;
    return map_any(shape, nElmt, nm, vertex_ids, l2g, sign, on_boundary, nglobal);
}

int imf_assembly_create_hex(int nElmt, int nm, const long long *vertex_ids, int dirichlet,
                            imf_asm *out)
{
    return create_any(imf::kHex, nElmt, uniform(nElmt, nm).data(), vertex_ids, dirichlet, out);
}

int imf_assembly_create_quad(int nElmt, int nm, const long long *vertex_ids, int dirichlet,
                             imf_asm *out)
{
    return create_any(imf::kQuad, nElmt, uniform(nElmt, nm).data(), vertex_ids, dirichlet, out);
}

int imf_assembly_create_varp(int dim, int nElmt, const int *nm, const long long *vertex_ids,
                             int dirichlet, imf_asm *out)
{
    int shape;
    if (int st = shape_of_dim(dim, shape)) return st;
    return create_any(shape, nElmt, nm, vertex_ids, dirichlet, out);
}

long imf_assembly_nglobal(imf_asm a) { return a ? a->ng : 0; }

int imf_assembly_map_aniso(int nElmt, const int *nm3, const long long *vertex_ids, int *l2g,
                           signed char *sign, unsigned char *on_boundary, long *nglobal)
{
    return map_any(imf::kHex, nElmt, nm3, vertex_ids, l2g, sign, on_boundary, nglobal, true);
}

int imf_assembly_create_aniso(int nElmt, const int *nm3, const long long *vertex_ids,
                              int dirichlet, imf_asm *out)
{
    return create_any(imf::kHex, nElmt, nm3, vertex_ids, dirichlet, out, true);
}

int imf_assembly_destroy(imf_asm a)
{
    if (!a) return IMF_OK;
    for (void *p : {(void *)a->d_l2g, (void *)a->d_sg, (void *)a->d_off, (void *)a->d_inv,
                    (void *)a->d_mask, (void *)a->d_ws, (void *)a->d_loc, (void *)a->d_part,
                    (void *)a->d_dots, (void *)a->d_roff, (void *)a->d_rcol, (void *)a->d_rval,
                    (void *)a->d_rw})
        if (p) (void)hipFree(p);
    delete a;
    return IMF_OK;
}

int imf_assembly_set_dirichlet_mask(imf_asm a, const unsigned char *mask)
{
    if (!a) return afail(IMF_ERR_ARG, "imf_assembly_set_dirichlet_mask: null assembly");
    a->diag_key.clear();  // the cached preconditioner excludes the Dirichlet modes
    if (!mask) {
        if (a->d_mask) (void)hipFree(a->d_mask);
        a->d_mask = nullptr;
        a->dirichlet = 0;
        return IMF_OK;
    }
    if (!a->d_mask) ACHECK(hipMalloc(&a->d_mask, a->ng), "hipMalloc(Dirichlet mask)");
    std::vector<unsigned char> m(mask, mask + a->ng);
    for (auto &v : m) v = v ? 1 : 0;
    ACHECK(hipMemcpy(a->d_mask, m.data(), a->ng, hipMemcpyHostToDevice), "hipMemcpy(Dirichlet mask)");
    a->dirichlet = 1;
    return IMF_OK;
}

int imf_assembly_set_robin(imf_asm a, long nnz, const int *rows, const int *cols,
                           const double *vals)
{
    if (!a || nnz < 0 || (nnz > 0 && (!rows || !cols || !vals)))
        return afail(IMF_ERR_ARG, "imf_assembly_set_robin: bad argument");
    for (long k = 0; k < nnz; ++k)
        if (rows[k] < 0 || rows[k] >= a->ng || cols[k] < 0 || cols[k] >= a->ng)
            return afail(IMF_ERR_ARG, "imf_assembly_set_robin: index outside the global modes");
    for (void *p : {(void *)a->d_roff, (void *)a->d_rcol, (void *)a->d_rval, (void *)a->d_rw})
        if (p) (void)hipFree(p);
    a->d_roff = a->d_rcol = nullptr;
    a->d_rval = a->d_rw = nullptr;
    a->r_nnz = 0;
    a->diag_key.clear();  // the cached preconditioner includes diag(R)
    if (nnz == 0) return IMF_OK;
    // CSR by row, entries of a row in the order given (repeated entries are summed by the
    // product)
    std::vector<int> off(a->ng + 1, 0), col(nnz);
    std::vector<double> val(nnz);
    for (long k = 0; k < nnz; ++k) ++off[rows[k] + 1];
    for (long g = 0; g < a->ng; ++g) off[g + 1] += off[g];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (long k = 0; k < nnz; ++k) {
        const int at = fill[rows[k]]++;
        col[at] = cols[k];
        val[at] = vals[k];
    }
    ACHECK(hipMalloc(&a->d_roff, (a->ng + 1) * sizeof(int)), "hipMalloc(Robin)");
    ACHECK(hipMalloc(&a->d_rcol, nnz * sizeof(int)), "hipMalloc(Robin)");
    ACHECK(hipMalloc(&a->d_rval, nnz * sizeof(double)), "hipMalloc(Robin)");
    ACHECK(hipMalloc(&a->d_rw, a->ng * sizeof(double)), "hipMalloc(Robin)");
    ACHECK(hipMemcpy(a->d_roff, off.data(), (a->ng + 1) * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy(Robin)");
    ACHECK(hipMemcpy(a->d_rcol, col.data(), nnz * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy(Robin)");
    ACHECK(hipMemcpy(a->d_rval, val.data(), nnz * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy(Robin)");
    a->r_nnz = nnz;
    return IMF_OK;
}

int imf_global_to_local_device(imf_asm a, const double *d_global, double *d_local, void *stream)
{
    if (!a || !d_global || !d_local) return afail(IMF_ERR_ARG, "imf_global_to_local_device: null argument");
    g2l_kernel<<<nblk(a->nloc), RB, 0, (hipStream_t)stream>>>(a->nloc, a->d_l2g, a->d_sg, d_global, d_local);
    ACHECK(hipGetLastError(), "g2l_kernel");
    return IMF_OK;
}

int imf_assemble_device(imf_asm a, const double *d_local, double *d_global, void *stream)
{
    if (!a || !d_global || !d_local) return afail(IMF_ERR_ARG, "imf_assemble_device: null argument");
    assemble_kernel<<<nblk(a->ng), RB, 0, (hipStream_t)stream>>>(a->ng, a->d_off, a->d_inv, a->d_sg,
                                                                d_local, nullptr, d_global);
    ACHECK(hipGetLastError(), "assemble_kernel");
    return IMF_OK;
}

int imf_helmsolve_cg_groups_device(imf_asm a, int nops, const imf_op *ops, const double *d_rhs,
                                   double *d_x, double tol, int max_iter, int *iterations,
                                   double *residual, void *stream)
{
    if (!a || !ops || nops < 1 || !d_rhs || !d_x)
        return afail(IMF_ERR_ARG, "imf_helmsolve_cg_device: null argument");
    long e0 = 0;
    for (int g = 0; g < nops; ++g) {
        const imf_op op = ops[g];
        if (!op) return afail(IMF_ERR_ARG, "imf_helmsolve_cg_device: null operator");
        bool orders_match = e0 + op->nElmt <= a->nElmt;
        for (long e = e0; orders_match && e < e0 + op->nElmt; ++e) {
            if (a->shape == imf::kHex) {
                for (int d = 0; d < 3; ++d)
                    orders_match = orders_match &&
                                   a->nm3[3 * e + d] == (op->aniso ? op->nmd[d] : op->nm);
            } else {
                orders_match = a->nmv[e] == op->nm;
            }
        }
        if (op->kind != imf::kHelmholtz || op->shape != a->shape || op->prec != IMF_FP64 ||
            !orders_match)
            return afail(IMF_ERR_ARG, "imf_helmsolve_cg_device: needs FP64 Helmholtz operators "
                                      "of the assembly's shape on its elements and modes");
        if (!op->geom_ready) return afail(IMF_ERR_STATE, "Jacobian and derivative factors not set");
        if (!op->aniso && (op->dbdata.size() != (size_t)op->nm * op->nq || op->nq > 16))
            return afail(IMF_ERR_ARG, "imf_helmsolve_cg_device: operator created without dbdata");
        e0 += op->nElmt;
    }
    if (e0 != a->nElmt)
        return afail(IMF_ERR_ARG, "imf_helmsolve_cg_device: the operators do not cover the "
                                  "assembly's elements");
    hipStream_t s = (hipStream_t)stream;
    const long ng = a->ng;
    if (!a->d_ws) {
        a->nparts = (int)std::min<long>(1024, (ng + RB - 1) / RB);
        ACHECK(hipMalloc(&a->d_ws, 6 * ng * sizeof(double)), "hipMalloc(cg)");
        ACHECK(hipMalloc(&a->d_loc, 2 * a->nloc * sizeof(double)), "hipMalloc(cg)");
        // block partial sums: dot3_partial's or the fused assembly pass's blocks
        const long np = std::max<long>(a->nparts, std::min<long>(kAsmParts, nblk(ng)));
        ACHECK(hipMalloc(&a->d_part, 3 * np * sizeof(double)), "hipMalloc(cg)");
        ACHECK(hipMalloc(&a->d_dots, 4 * sizeof(double)), "hipMalloc(cg)");
        a->diag_key.clear();
    }
    double *r = a->d_ws, *p = r + ng, *q = p + ng, *w = q + ng, *sv = w + ng, *dinv = sv + ng;
    // assembled diagonal (Jacobi preconditioner), cached per geometry pack (a process-wide
    // id, never reused) and lambda of every operator
    std::vector<std::tuple<const void *, unsigned long, double>> key;
    for (int g = 0; g < nops; ++g)
        key.emplace_back((const void *)ops[g], ops[g]->geom_version, ops[g]->lambda);
    if (a->diag_key != key) {
        e0 = 0;
        for (int g = 0; g < nops; ++g) {
            const imf_op op = ops[g];
            DiagTab t{};
            const int nm = op->nm, nq = op->nq;
            std::copy(op->bdata.begin(), op->bdata.end(), t.B);
            std::copy(op->dbdata.begin(), op->dbdata.end(), t.dB);
            std::copy(op->w.begin(), op->w.end(), t.w);
            double *dl = a->d_loc + a->loff[e0];
            const long n = a->loff[e0 + op->nElmt] - a->loff[e0];
            if (op->aniso)
                ACHECK(imf::launch_aniso_diag(aniso_args(op, nullptr, nullptr, 1, 0, 0), dl, s),
                       "anisotropic diagonal");
            else if (a->shape == imf::kHex)
                helm_diag_hex<<<nblk(n), RB, 0, s>>>(t, nm, nq, op->nElmt, op->deformed, op->lambda,
                                                    (const double *)op->d_geom, dl);
            else
                helm_diag_quad<<<nblk(n), RB, 0, s>>>(t, nm, nq, op->nElmt, op->deformed, op->lambda,
                                                     (const double *)op->d_geom, dl);
            e0 += op->nElmt;
        }
        assemble_kernel<<<nblk(ng), RB, 0, s>>>(ng, a->d_off, a->d_inv, nullptr, a->d_loc, nullptr, dinv);
        if (a->r_nnz)
            csr_diag_add<<<nblk(ng), RB, 0, s>>>(ng, a->d_roff, a->d_rcol, a->d_rval, dinv);
        inv_masked<<<nblk(ng), RB, 0, s>>>(ng, a->d_mask, dinv);
        ACHECK(hipGetLastError(), "diagonal kernels");
        a->diag_key = key;
    }
    double h[4];
    auto read_dots = [&]() -> int {
        ACHECK(hipMemcpyAsync(h, a->d_dots, 3 * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        ACHECK(hipStreamSynchronize(s), "sync");
        return IMF_OK;
    };
    auto dots = [&](const double *x0, const double *x1, const double *x2) -> int {
        dot3_partial<<<a->nparts, RB, 0, s>>>(ng, x0, x1, x2, a->d_part);
        dot3_final<<<1, RB, 0, s>>>(a->nparts, a->d_part, a->d_dots);
        ACHECK(hipGetLastError(), "dot kernels");
        return read_dots();
    };
    // s = A w with (r,w), (s,w), (r,r) in the same pass
    auto apply_dots = [&]() -> int {
        if (int st = apply_global_dots(a, nops, ops, w, r, sv, s)) return st;
        return read_dots();
    };
    // rhs magnitude over all global modes (Set_Rhs_Magnitude)
    if (int st = dots(d_rhs, d_rhs, d_rhs)) return st;
    const double rhs_mag = h[0] > 1e-6 ? h[0] : 1.0;
    cg_init<<<nblk(ng), RB, 0, s>>>(ng, a->d_mask, d_rhs, r, d_x);
    ACHECK(hipGetLastError(), "cg_init");
    if (int st = dots(r, r, r)) return st;
    double eps = h[2];
    const double thr = tol * tol * rhs_mag;
    int it = 0;
    if (eps >= thr) {
        cg_precon<<<nblk(ng), RB, 0, s>>>(ng, dinv, r, w);
        if (int st = apply_dots()) return st;  // (r,w), (s,w), (r,r)
        double rho = h[0], mu = h[1];
        double beta = 0.0, alpha = rho / mu;
        it = 1;
        zero_kernel<<<nblk(2 * ng), RB, 0, s>>>(2 * ng, p);  // p and q
        for (;;) {
            if (it > max_iter) {
                if (iterations) *iterations = it;
                if (residual) *residual = std::sqrt(eps / rhs_mag);
                return afail(IMF_ERR_STATE, "Exceeded maximum number of iterations");
            }
            cg_update<<<nblk(ng), RB, 0, s>>>(ng, alpha, beta, dinv, p, q, w, sv, d_x, r);
            if (int st = apply_dots()) return st;
            const double rho_new = h[0];
            mu = h[1];
            eps = h[2];
            ++it;
            if (eps < thr) break;
            beta = rho_new / rho;
            alpha = rho_new / (mu - rho_new * beta / alpha);
            rho = rho_new;
        }
    }
    if (iterations) *iterations = it;
    if (residual) *residual = std::sqrt(eps / rhs_mag);
    return IMF_OK;
}

int imf_helmsolve_cg_device(imf_asm a, imf_op op, const double *d_rhs, double *d_x, double tol,
                            int max_iter, int *iterations, double *residual, void *stream)
{
    return imf_helmsolve_cg_groups_device(a, 1, &op, d_rhs, d_x, tol, max_iter, iterations,
                                          residual, stream);
}

}  // extern "C"
