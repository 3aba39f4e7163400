"""Dense, face-centric DG assembly used to pin the oracle (not collected by pytest).

Independent of the oracle's code: it uses a Gauss-Lobatto nodal basis (as
deal.II's FE_DGQ), a dense QGauss(k+1) mass matrix inverted with
numpy.linalg.solve, numpy's own Gauss rule, explicit basis evaluation without
sum factorisation, and visits each face once with a plus and a minus side as
WaveEquationOperation::evaluate_inner_face does
(source/wave_equation_operations.cc:480-551, 569-629).  Input and output are
values at the cell's Gauss points so they compare with the oracle's layout.
"""
import itertools

import numpy as np


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def lobatto01(n):
    if n == 1:
        return np.array([0.5])
    # roots of P'_{n-1} plus the end points
    c = np.zeros(n)
    c[-1] = 1.0
    inner = np.polynomial.legendre.Legendre(c).deriv().roots()
    x = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    return 0.5 * (x + 1.0)


def lagrange_matrix(nodes, pts):
    """V[q, i] = L_i(pts[q]), dV[q, i] = L_i'(pts[q]) via polynomial fits."""
    n = len(nodes)
    V = np.zeros((len(pts), n))
    dV = np.zeros((len(pts), n))
    for i in range(n):
        y = np.zeros(n)
        y[i] = 1.0
        p = np.polynomial.Polynomial.fit(nodes, y, n - 1, domain=[0, 1], window=[0, 1])
        V[:, i] = p(pts)
        dV[:, i] = p.deriv()(pts)
    return V, dV


def multilinear(V, xi):
    dim = V.shape[1]
    x = np.zeros(dim)
    J = np.zeros((dim, dim))
    for v in range(1 << dim):
        bits = [(v >> d) & 1 for d in range(dim)]
        s = np.prod([xi[d] if bits[d] else 1 - xi[d] for d in range(dim)])
        x += s * V[v]
        for e in range(dim):
            g = np.prod([(1.0 if bits[d] else -1.0) if d == e else (xi[d] if bits[d] else 1 - xi[d])
                         for d in range(dim)])
            J[:, e] += g * V[v]
    return x, J


def apply(vertices, neighbor, rho, speed, k, u, sign=1.0):
    n_cells, nv, dim = vertices.shape
    n = k + 1
    nf = dim + 1
    xg, wg = gauss01(n)
    nodes = lobatto01(n)
    B, dB = lagrange_matrix(nodes, xg)
    Bf, _ = lagrange_matrix(nodes, np.array([0.0, 1.0]))
    Q = list(itertools.product(range(n), repeat=dim))  # (q_{dim-1},...,q_0) order fixed below
    Q = [tuple(reversed(t)) for t in Q]  # xi_0 fastest
    nd = len(Q)
    Phi = np.array([[np.prod([B[q[d], i[d]] for d in range(dim)]) for i in Q] for q in Q])
    dPhi = [np.array([[np.prod([(dB if d == e else B)[q[d], i[d]] for d in range(dim)])
                       for i in Q] for q in Q]) for e in range(dim)]
    # coefficients of the Lobatto basis from Gauss-point values
    uc = u.reshape(n_cells, nf, nd)
    a = np.einsum("ij,cfj->cfi", np.linalg.inv(Phi), uc)
    R = np.zeros_like(a)
    M = np.zeros((n_cells, nd, nd))
    for c in range(n_cells):
        Vc = vertices[c]
        rh, cc = rho[c], speed[c]
        vq = [Phi @ a[c, d] for d in range(dim)]
        gref = [dPhi[e] @ a[c, dim] for e in range(dim)]
        for iq, q in enumerate(Q):
            xi = np.array([xg[q[d]] for d in range(dim)])
            _, J = multilinear(Vc, xi)
            Ji = np.linalg.inv(J)
            jxw = abs(np.linalg.det(J)) * np.prod([wg[q[d]] for d in range(dim)])
            M[c] += jxw * np.outer(Phi[iq], Phi[iq])
            gp = Ji.T @ np.array([gref[e][iq] for e in range(dim)])
            for d in range(dim):
                R[c, d] += Phi[iq] * (-gp[d] / rh) * jxw
            vv = np.array([vq[d][iq] for d in range(dim)])
            gtest = Ji @ vv * rh * cc * cc * jxw
            for e in range(dim):
                R[c, dim] += dPhi[e][iq] * gtest[e]
    # faces, each once
    T = list(itertools.product(range(n), repeat=dim - 1))
    for c in range(n_cells):
        for f in range(2 * dim):
            d, s = f // 2, f % 2
            nb = neighbor[c, f]
            if nb >= 0 and nb < c:
                continue
            others = [e for e in range(dim) if e != d]
            for t in T:
                xi = np.zeros(dim)
                xi[d] = s
                wt = 1.0
                for j, e in enumerate(others):
                    xi[e] = xg[t[j]]
                    wt *= wg[t[j]]
                _, J = multilinear(vertices[c], xi)
                Ji = np.linalg.inv(J)
                nref = np.zeros(dim)
                nref[d] = 1.0 if s else -1.0
                nn = Ji.T @ nref
                jxw = abs(np.linalg.det(J)) * np.linalg.norm(nn) * wt
                nn = nn / np.linalg.norm(nn)

                def trace(cell, side):
                    vals = []
                    for i in Q:
                        val = Bf[side, i[d]]
                        for j, e in enumerate(others):
                            val *= B[t[j], i[e]]
                        vals.append(val)
                    vals = np.array(vals)
                    return vals, vals @ a[cell].T

                phip, up = trace(c, s)
                rp, cp = rho[c], speed[c]
                taup = 1.0 / cp / rp
                nvp = up[:dim] @ nn
                if nb >= 0:
                    phim, um = trace(nb, 1 - s)
                    rm, cm = rho[nb], speed[nb]
                    taum = 1.0 / cm / rm
                    nvm = -um[:dim] @ nn
                    lam = (nvp + nvm + taup * up[dim] + taum * um[dim]) / (taup + taum)
                    valm = np.zeros(nf)
                    valm[:dim] = -(um[dim] - lam) / rm * nn
                    valm[dim] = cm * cm * rm * (-nvm + taum * (lam - um[dim]))
                    R[nb] += np.outer(valm * jxw, phim)
                else:
                    bid = -nb
                    if bid == 1:
                        lam = nvp / taup + up[dim]
                    elif bid == 2:
                        lam = 0.0
                    else:
                        lam = taup / (taup + 1 / cp / rp) * up[dim] + nvp / (taup + 1 / cp / rp)
                valp = np.zeros(nf)
                valp[:dim] = (up[dim] - lam) / rp * nn
                valp[dim] = cp * cp * rp * (-nvp + taup * (lam - up[dim]))
                R[c] += np.outer(valp * jxw, phip)
    out = np.zeros_like(a)
    for c in range(n_cells):
        out[c] = sign * np.linalg.solve(M[c], R[c].T).T
    return np.einsum("ij,cfj->cfi", Phi, out).reshape(-1)
