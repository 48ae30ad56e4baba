"""Independent dense references used to pin the oracle (pure numpy/scipy).

Nothing here shares code with the oracle or the product: GLL nodes come from numpy's
Legendre module, the Modified_A basis and its derivative from closed forms with scipy's
Jacobi polynomials, and the physical gradient from the inverse of the map Jacobian
(not from the adjugate/derivative-factor formulae).  Element operators are assembled as
explicit matrices (Kronecker products), so a dropped term, a transposed operand or a
wrong index in the sum-factorised code shows up as a mismatch.
"""
import numpy as np
from numpy.polynomial import legendre as L
from scipy.special import eval_jacobi


def gll(nq):
    n = nq - 1
    c = np.zeros(nq); c[-1] = 1.0           # P_n
    inner = np.sort(L.legroots(L.legder(c)))  # roots of P_n'
    z = np.concatenate([[-1.0], inner, [1.0]])
    w = 2.0 / (n * (n + 1) * L.legval(z, c) ** 2)
    return z, w


def modified_a(nm, z):
    """B[i, p] = phi_p(z_i) and dB[i, p] = phi_p'(z_i) in closed form."""
    B = np.zeros((len(z), nm)); dB = np.zeros_like(B)
    B[:, 0] = 0.5 * (1 - z); dB[:, 0] = -0.5
    B[:, 1] = 0.5 * (1 + z); dB[:, 1] = 0.5
    for p in range(2, nm):
        n = p - 2
        P = eval_jacobi(n, 1, 1, z)
        dP = 0.5 * (n + 3) * eval_jacobi(n - 1, 2, 2, z) if n > 0 else np.zeros_like(z)
        B[:, p] = 0.25 * (1 - z * z) * P
        dB[:, p] = -0.5 * z * P + 0.25 * (1 - z * z) * dP
    return B, dB


def hex_map(verts, xi):
    """x(xi) and A[c, a] = dx_c/dxi_a for a trilinear hex (verts 8x3)."""
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    x = np.zeros(3); A = np.zeros((3, 3))
    for v, c in enumerate(corners):
        f = np.where(c, 0.5 * (1 + xi), 0.5 * (1 - xi))
        d = np.where(c, 0.5, -0.5)
        x += np.prod(f) * verts[v]
        grad = np.array([d[0] * f[1] * f[2], f[0] * d[1] * f[2], f[0] * f[1] * d[2]])
        A += np.outer(verts[v], grad)
    return x, A


def quad_map(verts, xi):
    corners = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])
    nc = np.shape(verts)[1]                # 2, or 3 for a quad in 3D space
    x = np.zeros(nc); A = np.zeros((nc, 2))
    for v, c in enumerate(corners):
        f = np.where(c, 0.5 * (1 + xi), 0.5 * (1 - xi))
        d = np.where(c, 0.5, -0.5)
        x += np.prod(f) * verts[v]
        A += np.outer(verts[v], np.array([d[0] * f[1], f[0] * d[1]]))
    return x, A


def element_mats(shape, nm, nq, verts):
    """Dense per-element matrices: B (Q x M), grads G_a (Q x M) in reference space,
    quadrature weights W (Q), |det A| (Q), A^{-1} (Q x dim x dim), coordinates.  For hexes
    nm and nq may be tuples of the three directions' modes and points (coefficient index
    (r nm1 + q) nm0 + p, point index (k nq1 + j) nq0 + i)."""
    if shape == "hex" and isinstance(nm, tuple):
        zz = [gll(n) for n in nq]
        Bd = [modified_a(m, z_[0]) for m, z_ in zip(nm, zz)]
        (B0, dB0), (B1, dB1), (B2, dB2) = Bd
        B = np.kron(np.kron(B2, B1), B0)
        G = [np.kron(np.kron(B2, B1), dB0), np.kron(np.kron(B2, dB1), B0),
             np.kron(np.kron(dB2, B1), B0)]
        W = np.kron(np.kron(zz[2][1], zz[1][1]), zz[0][1])
        pts = [(zz[0][0][i], zz[1][0][j], zz[2][0][k]) for k in range(nq[2])
               for j in range(nq[1]) for i in range(nq[0])]
        dets, Ainv, X = [], [], []
        for p in pts:
            x, A = hex_map(verts, np.array(p))
            dets.append(abs(np.linalg.det(A))); Ainv.append(np.linalg.inv(A)); X.append(x)
        return dict(B=B, G=G, W=W, J=np.array(dets), Ainv=np.array(Ainv), X=np.array(X),
                    z=[z_[0] for z_ in zz], w=[z_[1] for z_ in zz])
    z, w = gll(nq)
    B1, dB1 = modified_a(nm, z)
    if shape == "hex":
        B = np.kron(np.kron(B1, B1), B1)
        G = [np.kron(np.kron(B1, B1), dB1), np.kron(np.kron(B1, dB1), B1),
             np.kron(np.kron(dB1, B1), B1)]
        W = np.kron(np.kron(w, w), w)
        pts = [(z[i], z[j], z[k]) for k in range(nq) for j in range(nq) for i in range(nq)]
        mp = hex_map
    else:
        B = np.kron(B1, B1)
        G = [np.kron(B1, dB1), np.kron(dB1, B1)]
        W = np.kron(w, w)
        pts = [(z[i], z[j]) for j in range(nq) for i in range(nq)]
        mp = quad_map
    dets, Ainv, X = [], [], []
    for p in pts:
        x, A = mp(verts, np.array(p))
        dets.append(abs(np.linalg.det(A))); Ainv.append(np.linalg.inv(A)); X.append(x)
    return dict(B=B, G=G, W=W, J=np.array(dets), Ainv=np.array(Ainv), X=np.array(X), z=z, w=w)


def helmholtz_matrix(m, lam):
    """lambda B^T W J B + sum_c (d_c B)^T W J (d_c B), d_c = sum_a (A^{-1})[a, c] G_a."""
    WJ = m["W"] * m["J"]
    H = lam * m["B"].T @ (WJ[:, None] * m["B"])
    dim = len(m["G"])
    for c in range(dim):
        dc = sum(m["Ainv"][:, a, c][:, None] * m["G"][a] for a in range(dim))
        H += dc.T @ (WJ[:, None] * dc)
    return H


# ---- triangles: Modified_A (eta0, GLL) x Modified_B (eta1, Gauss-Radau alpha = 1) ----

def gauss_radau_a1(nq):
    """Nodes -1 and the roots of P_{nq-1}^{(1,1)} (scipy), weights w_i = int l_i (1-z) dz
    by a 40-point Gauss-Legendre rule (exact for these polynomial degrees)."""
    from scipy.special import roots_jacobi
    z = np.concatenate([[-1.0], np.sort(roots_jacobi(nq - 1, 1.0, 1.0)[0])]) if nq > 1 else np.zeros(1)
    xg, wg = L.leggauss(40)
    w = np.zeros(nq)
    for i in range(nq):
        li = np.ones_like(xg)
        for k in range(nq):
            if k != i:
                li *= (xg - z[k]) / (z[i] - z[k])
        w[i] = np.sum(wg * li * (1 - xg))
    return z, w


def modified_b(nm, z):
    """B[i, mode] in the (p outer, q inner) packing, closed forms with scipy Jacobi."""
    cols = []
    a, b = 0.5 * (1 - z), 0.5 * (1 + z)
    for p in range(nm):
        for q in range(nm - p):
            if p == 0:
                v = a if q == 0 else b if q == 1 else a * b * eval_jacobi(q - 2, 1, 1, z)
            elif p == 1:
                v = a if q == 0 else a * b * eval_jacobi(q - 1, 1, 1, z)
            else:
                v = a ** p if q == 0 else a ** p * b * eval_jacobi(q - 1, 2 * p - 1, 1, z)
            cols.append(v)
    return np.array(cols).T


def tri_element_mats(nm, nq0, nq1, verts):
    """Dense triangle operators: B (Q x M, the collapsed vertex mode 1 includes phi_1(eta0)),
    physical gradients G_c (Q x M) from an exact monomial fit in (xi0, xi1), quadrature
    weights W (with the collapse factor 1/2), the constant |det A| and coordinates."""
    z0, w0 = gll(nq0)
    z1, w1 = gauss_radau_a1(nq1)
    Ba, _ = modified_a(nm, z0)
    Bb = modified_b(nm, z1)
    E0 = np.repeat(z0[None, :], nq1, 0).ravel()        # point (j, i): i fastest
    E1 = np.repeat(z1[:, None], nq0, 1).ravel()
    I0 = np.tile(np.arange(nq0), nq1)
    I1 = np.repeat(np.arange(nq1), nq0)
    M = nm * (nm + 1) // 2
    B = np.zeros((nq0 * nq1, M))
    mode = 0
    for p in range(nm):
        for q in range(nm - p):
            B[:, mode] = Ba[I0, p] * Bb[I1, mode]
            mode += 1
    B[:, 1] += Ba[I0, 1] * Bb[I1, 1]
    xi0 = 0.5 * (1 + E0) * (1 - E1) - 1.0
    xi1 = E1
    pw = [(a, c) for a in range(nm) for c in range(nm - a)]
    V = np.stack([xi0 ** a * xi1 ** c for a, c in pw], 1)
    coef = np.linalg.lstsq(V, B, rcond=None)[0]
    dV0 = np.stack([a * xi0 ** max(a - 1, 0) * xi1 ** c for a, c in pw], 1)
    dV1 = np.stack([c * xi0 ** a * xi1 ** max(c - 1, 0) for a, c in pw], 1)
    G = [dV0 @ coef, dV1 @ coef]
    v = np.asarray(verts, dtype=float)
    nc = v.shape[1]                                                # 2, or 3 for a surface
    A = np.stack([0.5 * (v[1] - v[0]), 0.5 * (v[2] - v[0])], 1)   # A[c, a] = dx_c / dxi_a
    Ainv = np.linalg.pinv(A)                                       # the inverse when nc = 2
    Gp = [Ainv[0, c] * G[0] + Ainv[1, c] * G[1] for c in range(nc)]
    X = np.stack([-(xi0 + xi1) / 2 * v[0][c] + (1 + xi0) / 2 * v[1][c] + (1 + xi1) / 2 * v[2][c]
                  for c in range(nc)])
    W = 0.5 * w0[I0] * w1[I1]
    J = np.sqrt(np.linalg.det(A.T @ A))                            # |det A| when nc = 2
    return dict(B=B, Gphys=Gp, W=W, J=J, X=X, A=A, Ainv=Ainv, z0=z0, z1=z1, w0=w0,
                w1_raw=w1, fit_residual=np.abs(V @ coef - B).max())


def tri_helmholtz_matrix(m, lam):
    WJ = m["W"] * m["J"]
    H = lam * m["B"].T @ (WJ[:, None] * m["B"])
    for g in m["Gphys"]:
        H += g.T @ (WJ[:, None] * g)
    return H


# ---- prisms: Modified_A (eta0) x Modified_A (xi1) x Modified_B (eta2) ----

def prism_element_mats(nm, nq, nq2, verts):
    """Dense prism operators (points (k nq + j) nq + i): B with the singular-edge modes
    (0, q, 1) also carrying phi_1(eta0), physical gradients from an exact monomial fit in
    (xi0, xi1, xi2) (degree <= nm-1 in (xi0, xi2) jointly and in xi1), weights with the
    collapse factor 1/2, and |det| of the map Jacobian per point."""
    z0, w0 = gll(nq)
    z2, w2 = gauss_radau_a1(nq2)
    Ba, _ = modified_a(nm, z0)
    Bb = modified_b(nm, z2)
    I0 = np.tile(np.arange(nq), nq * nq2)
    I1 = np.tile(np.repeat(np.arange(nq), nq), nq2)
    I2 = np.repeat(np.arange(nq2), nq * nq)
    M = nm * nm * (nm + 1) // 2
    B = np.zeros((nq * nq * nq2, M))
    mode, mpr = 0, 0
    for p in range(nm):
        for q in range(nm):
            for r in range(nm - p):
                B[:, mode] = Ba[I0, p] * Ba[I1, q] * Bb[I2, mpr + r]
                mode += 1
        mpr += nm - p
    for q in range(nm):
        B[:, q * nm + 1] += Ba[I0, 1] * Ba[I1, q] * Bb[I2, 1]
    e0, xi1, xi2 = z0[I0], z0[I1], z2[I2]
    xi0 = 0.5 * (1 + e0) * (1 - xi2) - 1.0
    pw = [(a, b, c) for a in range(nm) for c in range(nm - a) for b in range(nm)]
    V = np.stack([xi0 ** a * xi1 ** b * xi2 ** c for a, b, c in pw], 1)
    coef = np.linalg.lstsq(V, B, rcond=None)[0]
    dV = [np.stack([a * xi0 ** max(a - 1, 0) * xi1 ** b * xi2 ** c for a, b, c in pw], 1),
          np.stack([b * xi0 ** a * xi1 ** max(b - 1, 0) * xi2 ** c for a, b, c in pw], 1),
          np.stack([c * xi0 ** a * xi1 ** b * xi2 ** max(c - 1, 0) for a, b, c in pw], 1)]
    G = [d @ coef for d in dV]
    v = np.asarray(verts, dtype=float).reshape(6, 3)
    la, lb, lc = -(xi0 + xi2) / 2, (1 + xi0) / 2, (1 + xi2) / 2
    sl, sh = (1 - xi1) / 2, (1 + xi1) / 2
    N = [la * sl, lb * sl, lb * sh, la * sh, lc * sl, lc * sh]
    dN = [(-0.5 * sl, -0.5 * la, -0.5 * sl), (0.5 * sl, -0.5 * lb, 0 * sl),
          (0.5 * sh, 0.5 * lb, 0 * sl), (-0.5 * sh, 0.5 * la, -0.5 * sh),
          (0 * sl, -0.5 * lc, 0.5 * sl), (0 * sl, 0.5 * lc, 0.5 * sh)]
    X = np.stack([sum(N[n] * v[n][c] for n in range(6)) for c in range(3)])
    J, Gp = [], [np.zeros_like(B) for _ in range(3)]
    for t in range(len(xi0)):
        A = np.array([[sum(dN[n][a][t] * v[n][c] for n in range(6)) for a in range(3)]
                      for c in range(3)])          # A[c, a] = dx_c / dxi_a
        J.append(abs(np.linalg.det(A)))
        Ai = np.linalg.inv(A)
        for c in range(3):
            Gp[c][t] = sum(Ai[a, c] * G[a][t] for a in range(3))
    W = w0[I0] * w0[I1] * 0.5 * w2[I2]
    return dict(B=B, Gphys=Gp, W=W, J=np.array(J), X=X, fit_residual=np.abs(V @ coef - B).max())


def prism_helmholtz_matrix(m, lam):
    WJ = m["W"] * m["J"]
    H = lam * m["B"].T @ (WJ[:, None] * m["B"])
    for g in m["Gphys"]:
        H += g.T @ (WJ[:, None] * g)
    return H


# ---- tetrahedra: Modified_A (eta0) x Modified_B (eta1) x Modified_C (eta2) ----

def gauss_radau(nq, alpha):
    """Nodes -1 and the roots of P_{nq-1}^{(alpha,1)} (scipy), weights w_i = int l_i
    (1-z)^alpha dz by a 40-point Gauss-Legendre rule (exact for these degrees)."""
    from scipy.special import roots_jacobi
    z = (np.concatenate([[-1.0], np.sort(roots_jacobi(nq - 1, float(alpha), 1.0)[0])])
         if nq > 1 else np.zeros(1))
    xg, wg = L.leggauss(40)
    w = np.zeros(nq)
    for i in range(nq):
        li = np.ones_like(xg)
        for k in range(nq):
            if k != i:
                li *= (xg - z[k]) / (z[i] - z[k])
        w[i] = np.sum(wg * li * (1 - xg) ** alpha)
    return z, w


def tri_index(nm, p, q):
    return sum(nm - pp for pp in range(p)) + q


def modified_c(nm, z):
    """C[i, mode] for modes (p, q, r), r fastest: phi^c_pqr = phi^b_{p+q, r} (closed form)."""
    Bb = modified_b(nm, z)
    cols = [Bb[:, tri_index(nm, p + q, r)]
            for p in range(nm) for q in range(nm - p) for r in range(nm - p - q)]
    return np.array(cols).T


def tet_element_mats(nm, nq0, nq1, nq2, verts):
    """Dense tetrahedron operators (points (k nq1 + j) nq0 + i): B with the top vertex
    (mode 1), the bottom vertex of the collapsed edge (mode nm) and the singular edge
    (modes nm + r) carrying the extra lower-direction products, physical gradients from an
    exact monomial fit of total degree <= nm - 1 in (xi0, xi1, xi2), weights with the
    collapse factors 1/2 and 1/4, and |det| of the affine map."""
    z0, w0 = gll(nq0)
    z1, w1 = gauss_radau(nq1, 1)
    z2, w2 = gauss_radau(nq2, 2)
    Ba, _ = modified_a(nm, z0)
    Bb = modified_b(nm, z1)
    Bc = modified_b(nm, z2)  # phi^c_{pqr} = column tri_index(p + q, r) of this
    I0 = np.tile(np.arange(nq0), nq1 * nq2)
    I1 = np.tile(np.repeat(np.arange(nq1), nq0), nq2)
    I2 = np.repeat(np.arange(nq2), nq0 * nq1)
    M = nm * (nm + 1) * (nm + 2) // 6
    B = np.zeros((nq0 * nq1 * nq2, M))
    mode = 0
    for p in range(nm):
        for q in range(nm - p):
            for r in range(nm - p - q):
                B[:, mode] = (Ba[I0, p] * Bb[I1, tri_index(nm, p, q)]
                              * Bc[I2, tri_index(nm, p + q, r)])
                mode += 1
    a0, a1 = Ba[I0, 0], Ba[I0, 1]
    b00, b01 = Bb[I1, tri_index(nm, 0, 0)], Bb[I1, tri_index(nm, 0, 1)]
    C = lambda r: Bc[I2, tri_index(nm, 0, r)]
    B[:, 1] += C(1) * (a0 * b01 + a1 * b00 + a1 * b01)
    B[:, nm] += C(0) * a1 * b01
    for r in range(1, nm - 1):
        B[:, nm + r] += C(r + 1) * a1 * b01
    xi2 = z2[I2]
    xi1 = 0.5 * (1 + z1[I1]) * (1 - xi2) - 1
    xi0 = 0.5 * (1 + z0[I0]) * (-xi1 - xi2) - 1
    pw = [(a, b, c) for a in range(nm) for b in range(nm - a) for c in range(nm - a - b)]
    V = np.stack([xi0 ** a * xi1 ** b * xi2 ** c for a, b, c in pw], 1)
    coef = np.linalg.lstsq(V, B, rcond=None)[0]
    dV = [np.stack([a * xi0 ** max(a - 1, 0) * xi1 ** b * xi2 ** c for a, b, c in pw], 1),
          np.stack([b * xi0 ** a * xi1 ** max(b - 1, 0) * xi2 ** c for a, b, c in pw], 1),
          np.stack([c * xi0 ** a * xi1 ** b * xi2 ** max(c - 1, 0) for a, b, c in pw], 1)]
    G = [d @ coef for d in dV]
    v = np.asarray(verts, dtype=float).reshape(4, 3)
    A = np.array([[0.5 * (v[a + 1, c] - v[0, c]) for a in range(3)] for c in range(3)])
    Ai = np.linalg.inv(A)
    Gp = [sum(Ai[a, c] * G[a] for a in range(3)) for c in range(3)]
    W = w0[I0] * w1[I1] * 0.5 * w2[I2] * 0.25
    Nv = [-0.5 * (1 + xi0 + xi1 + xi2), 0.5 * (1 + xi0), 0.5 * (1 + xi1), 0.5 * (1 + xi2)]
    X = np.stack([sum(Nv[n] * v[n, c] for n in range(4)) for c in range(3)])
    return dict(B=B, Gphys=Gp, W=W, J=abs(np.linalg.det(A)) * np.ones(len(W)), X=X,
                fit_residual=np.abs(V @ coef - B).max(), z2=z2, w2=w2)


def tet_helmholtz_matrix(m, lam):
    WJ = m["W"] * m["J"]
    H = lam * m["B"].T @ (WJ[:, None] * m["B"])
    for g in m["Gphys"]:
        H += g.T @ (WJ[:, None] * g)
    return H


# ---- pyramids: Modified_A (eta0) x Modified_A (eta1) x ModifiedPyr_C (eta2) ----
# The pyramid space is rational in xi, so instead of a monomial fit the modes are evaluated
# as closed-form functions of xi (through eta(xi)) and differentiated by complex step.

def jacobi_c(n, a, b, x):
    """P_n^{(a,b)}(x) by the three-term recurrence (complex arguments allowed)."""
    x = np.asarray(x)
    if n == 0:
        return np.ones_like(x)
    p0, p1 = np.ones_like(x), 0.5 * (a - b + (a + b + 2) * x)
    for k in range(1, n):
        s = 2 * k + a + b
        c = 2 * (k + 1) * (k + a + b + 1) * s
        p0, p1 = p1, ((s + 1) * (s * (s + 2) * x + a * a - b * b) * p1
                      - 2 * (k + a) * (k + b) * (s + 2) * p0) / c
    return p1


def _mod_a(p, x):
    if p == 0:
        return 0.5 * (1 - x)
    if p == 1:
        return 0.5 * (1 + x)
    return 0.25 * (1 - x) * (1 + x) * jacobi_c(p - 2, 1, 1, x)


def _mod_b(s, r, z):
    a, b = 0.5 * (1 - z), 0.5 * (1 + z)
    if s == 0:
        return a if r == 0 else b if r == 1 else a * b * jacobi_c(r - 2, 1, 1, z)
    if s == 1:
        return a if r == 0 else a * b * jacobi_c(r - 1, 1, 1, z)
    return a ** s if r == 0 else a ** s * b * jacobi_c(r - 1, 2 * s - 1, 1, z)


def pyr_c(p, q, r, z):
    """phi^{pyr c}_{pqr}: phi^b_{max(p,q), r} if min(p, q) <= 1, else ((1-z)/2)^{p+q-2}
    times (1+z)/2 P_{r-1}^{(2p+2q-3,1)} for r >= 1."""
    if min(p, q) <= 1:
        return _mod_b(max(p, q), r, z)
    a = (0.5 * (1 - z)) ** (p + q - 2)
    return a if r == 0 else a * 0.5 * (1 + z) * jacobi_c(r - 1, 2 * p + 2 * q - 3, 1, z)


def pyr_modes(nm):
    return [(p, q, r) for p in range(nm) for q in range(nm) for r in range(nm - max(p, q))]


def pyr_mode_fn(p, q, r, xi0, xi1, xi2):
    """Mode (p, q, r) as a function of xi; the top vertex (0, 0, 1) is the whole (1+eta2)/2
    (its CORRECT terms)."""
    e0 = 2 * (1 + xi0) / (1 - xi2) - 1
    e1 = 2 * (1 + xi1) / (1 - xi2) - 1
    v = _mod_a(p, e0) * _mod_a(q, e1) * pyr_c(p, q, r, xi2)
    if (p, q, r) == (0, 0, 1):
        v = v + pyr_c(0, 0, 1, xi2) * (_mod_a(0, e0) * _mod_a(1, e1) + _mod_a(1, e0) * _mod_a(0, e1)
                                       + _mod_a(1, e0) * _mod_a(1, e1))
    return v


def pyr_map_fn(v, xi0, xi1, xi2):
    """x(xi) = sum of the five vertex modes times the vertex coordinates."""
    e0 = 2 * (1 + xi0) / (1 - xi2) - 1
    e1 = 2 * (1 + xi1) / (1 - xi2) - 1
    a0, a1, b0, b1 = 0.5 * (1 - e0), 0.5 * (1 + e0), 0.5 * (1 - e1), 0.5 * (1 + e1)
    c0, c1 = 0.5 * (1 - xi2), 0.5 * (1 + xi2)
    N = [a0 * b0 * c0, a1 * b0 * c0, a1 * b1 * c0, a0 * b1 * c0, c1]
    return [sum(N[n] * v[n, c] for n in range(5)) for c in range(3)]


def lagrange_dmat(z):
    """Dm[i, a] = l_a'(z_i) for the Lagrange basis through z, from the product rule."""
    n = len(z)
    Dm = np.zeros((n, n))
    for a in range(n):
        den = np.prod([z[a] - z[m] for m in range(n) if m != a])
        for i in range(n):
            s = 0.0
            for m in range(n):
                if m == a:
                    continue
                s += np.prod([z[i] - z[k] for k in range(n) if k not in (a, m)])
            Dm[i, a] = s / den
    return Dm


def pyr_element_mats(nm, nq0, nq1, nq2, verts, deriv="collocation"):
    """Dense pyramid operators.  deriv = "collocation": reference-space derivatives are
    those of the interpolant through the points along each eta direction (what MatrixFree
    computes; the pyramid's high modes have degree up to 2 nm - 4 in eta2, more than nq2
    points resolve), turned into xi derivatives with d eta / d xi by complex step;
    deriv = "exact": the modes differentiated in xi by complex step."""
    z0, w0 = gll(nq0)
    z1, w1 = gll(nq1)
    z2, w2 = gauss_radau(nq2, 2)
    I0 = np.tile(np.arange(nq0), nq1 * nq2)
    I1 = np.tile(np.repeat(np.arange(nq1), nq0), nq2)
    I2 = np.repeat(np.arange(nq2), nq0 * nq1)
    xi2 = z2[I2]
    xi0 = 0.5 * (1 + z0[I0]) * (1 - xi2) - 1
    xi1 = 0.5 * (1 + z1[I1]) * (1 - xi2) - 1
    modes = pyr_modes(nm)
    h = 1e-30
    xi = [xi0.astype(complex), xi1.astype(complex), xi2.astype(complex)]
    B = np.stack([pyr_mode_fn(p, q, r, xi0, xi1, xi2) for p, q, r in modes], 1)
    G = []
    if deriv == "exact":
        for a in range(3):
            s = [x.copy() for x in xi]
            s[a] = s[a] + 1j * h
            G.append(np.stack([pyr_mode_fn(p, q, r, *s).imag / h for p, q, r in modes], 1))
    else:
        U = B.reshape(nq2, nq1, nq0, -1)
        dU = [np.einsum("ia,kjam->kjim", lagrange_dmat(z0), U).reshape(len(xi0), -1),
              np.einsum("ja,kaim->kjim", lagrange_dmat(z1), U).reshape(len(xi0), -1),
              np.einsum("ka,ajim->kjim", lagrange_dmat(z2), U).reshape(len(xi0), -1)]
        eta = lambda x0, x1, x2: [2 * (1 + x0) / (1 - x2) - 1, 2 * (1 + x1) / (1 - x2) - 1, x2]
        for a in range(3):
            s = [x.copy() for x in xi]
            s[a] = s[a] + 1j * h
            de = [e.imag / h for e in eta(*s)]  # d eta_b / d xi_a
            G.append(sum(de[b][:, None] * dU[b] for b in range(3)))
    v = np.asarray(verts, dtype=float).reshape(5, 3)
    A = np.zeros((len(xi0), 3, 3))  # A[t, c, a] = dx_c / dxi_a
    for a in range(3):
        s = [x.copy() for x in xi]
        s[a] = s[a] + 1j * h
        X = pyr_map_fn(v, *s)
        for c in range(3):
            A[:, c, a] = X[c].imag / h
    J = np.abs(np.linalg.det(A))
    Ai = np.linalg.inv(A)  # Ai[t, a, c] = dxi_a / dx_c
    Gp = [sum(Ai[:, a, c][:, None] * G[a] for a in range(3)) for c in range(3)]
    W = w0[I0] * w1[I1] * w2[I2] * 0.25
    return dict(B=B, Gphys=Gp, W=W, J=J, X=np.stack(pyr_map_fn(v, xi0, xi1, xi2)))


def pyr_helmholtz_matrix(m, lam):
    WJ = m["W"] * m["J"]
    H = lam * m["B"].T @ (WJ[:, None] * m["B"])
    for g in m["Gphys"]:
        H += g.T @ (WJ[:, None] * g)
    return H
