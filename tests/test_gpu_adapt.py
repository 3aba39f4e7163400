"""GPU solution transfer of one adaptive level (exw_adapt_box_transfer,
SolutionTransfer::interpolate in adapt_mesh, source/explicit_wave.cc:306-322)
against an independent evaluation here: children take the parent's polynomial
(Lagrange evaluation at the mapped nodes), a coarsened parent the L2 projection
of its children on the reference cell (computed with a QGauss(k+3) rule per
child, not the library's 1D matrices); refining and coarsening back returns
the original state."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def X():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import exwave_amd
    exwave_amd.load()
    return exwave_amd


@pytest.fixture(scope="module")
def ctx(X):
    return X.Context(0)


def _lag(nodes, x):
    return np.array([[np.prod([(xx - nodes[j]) / (nodes[i] - nodes[j])
                               for j in range(len(nodes)) if j != i])
                      for i in range(len(nodes))] for xx in np.atleast_1d(x)])


def _expected(dim, k, n, ro, rn, u_old):
    """new-mesh coefficients from the old ones, cell by cell"""
    N = k + 1
    nd = N ** dim
    xg, wg = O.gauss(N)
    xq, wq = O.gauss(k + 3)
    nch = 1 << dim
    uo = u_old.reshape(-1, dim + 1, nd)
    out = []
    co = 0

    def tens(mats, vals):  # apply mats[d] (rows: out) along xi_d, vals xi_0 fastest
        t = vals.reshape((N,) * dim)
        for d in range(dim):
            ax = dim - 1 - d
            t = np.moveaxis(np.tensordot(mats[d], t, axes=([1], [ax])), 0, ax)
        return t.ravel()
    for B in range(n ** dim):
        if ro[B] == rn[B]:
            m = nch if ro[B] else 1
            out.extend(uo[co:co + m])
        elif rn[B]:
            for ch in range(nch):
                mats = [_lag(xg, 0.5 * (xg + ((ch >> d) & 1))) for d in range(dim)]
                out.append(np.array([tens(mats, uo[co, f]) for f in range(dim + 1)]))
        else:
            # parent_i = (1/w_i) sum_b int_{child b} l_i u_b, reference measure 2^-dim
            acc = np.zeros((dim + 1, nd))
            for ch in range(nch):
                # child values at the child's quadrature points, parent basis there
                Ec = [_lag(xg, xq) for d in range(dim)]
                Ep = [_lag(xg, 0.5 * (xq + ((ch >> d) & 1))) for d in range(dim)]
                wprod = wq
                for d in range(1, dim):
                    wprod = np.multiply.outer(wq, wprod)
                wprod = wprod.ravel() / nch
                for f in range(dim + 1):
                    # evaluate the child at the quadrature points
                    t = uo[co + ch, f].reshape((N,) * dim)
                    for d in range(dim):
                        ax = dim - 1 - d
                        t = np.moveaxis(np.tensordot(Ec[d], t, axes=([1], [ax])), 0, ax)
                    g = t.ravel() * wprod
                    # integrate against every parent basis function
                    t = g.reshape((k + 3,) * dim)
                    for d in range(dim):
                        ax = dim - 1 - d
                        t = np.moveaxis(np.tensordot(Ep[d].T, t, axes=([1], [ax])), 0, ax)
                    acc[f] += t.ravel()
            wn = wg
            for d in range(1, dim):
                wn = np.multiply.outer(wg, wn)
            out.append(acc / wn.ravel()[None, :])
        co += nch if ro[B] else 1
    return np.concatenate([np.asarray(o).reshape(-1) for o in out])


@pytest.mark.parametrize("dim,k,n,seed", [(2, 3, 4, 1), (3, 2, 3, 2), (3, 4, 2, 3), (2, 8, 3, 4)])
def test_transfer_matches_independent_evaluation(X, ctx, dim, k, n, seed):
    import torch
    print("transfer", dim, k, n, seed, flush=True)
    rng = np.random.default_rng(seed)
    ro = (rng.uniform(size=n ** dim) < 0.5).astype(np.int32)
    rn = ro.copy()
    flip = rng.uniform(size=n ** dim) < 0.5
    rn[flip] = 1 - rn[flip]
    assert (ro != rn).any() and (ro == rn).any()
    Vo, _, _ = X.mesh_box_refined(dim, n, ro)
    Vn, _, _ = X.mesh_box_refined(dim, n, rn)
    nd = (k + 1) ** dim
    u_old = rng.standard_normal(len(Vo) * (dim + 1) * nd)
    old = torch.from_numpy(u_old).cuda()
    new = torch.zeros(len(Vn) * (dim + 1) * nd, dtype=torch.float64, device="cuda")
    X.adapt_box_transfer(ctx, dim, k, n, ro, rn, old, new)
    want = _expected(dim, k, n, ro, rn, u_old)
    got = new.cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_refine_then_coarsen_is_identity(X, ctx):
    import torch
    dim, k, n = 3, 3, 2
    r0 = np.zeros(n ** dim, np.int32)
    r1 = np.ones(n ** dim, np.int32)
    nd = (k + 1) ** dim
    u = np.random.default_rng(5).standard_normal(n ** dim * (dim + 1) * nd)
    a = torch.from_numpy(u).cuda()
    b = torch.zeros(n ** dim * 8 * (dim + 1) * nd, dtype=torch.float64, device="cuda")
    c = torch.zeros_like(a)
    X.adapt_box_transfer(ctx, dim, k, n, r0, r1, a, b)
    X.adapt_box_transfer(ctx, dim, k, n, r1, r0, b, c)
    assert np.max(np.abs(c.cpu().numpy() - u)) <= 1e-13 * np.max(np.abs(u))


@pytest.mark.parametrize("dim,k,n,seed", [(2, 3, 3, 1), (3, 3, 2, 2), (2, 12, 2, 3),
                                          (3, 9, 1, 4)])
def test_forest_transfer_matches_oracle(X, ctx, dim, k, n, seed):
    """exw_adapt_forest_transfer over several levels (kept, refined and
    coarsened cells of a random adaptation) against the oracle's transfer"""
    import torch
    print("forest transfer", dim, k, n, seed, flush=True)
    rng = np.random.default_rng(seed)
    tree = O.uniform_forest(dim, n, 1)
    for _ in range(2):
        lev = O.forest_levels(dim, n, tree)
        fl = rng.choice([-1, 0, 1], size=lev.size, p=[0.4, 0.3, 0.3]).astype(np.int32)
        fl[(fl == 1) & (lev >= 3)] = 0
        fl[(fl == -1) & (lev <= 0)] = 0
        new = X.adapt_forest(dim, n, tree, fl)
        nd = (dim + 1) * (k + 1) ** dim
        n_old, n_new = lev.size, O.forest_levels(dim, n, new).size
        u = rng.standard_normal(n_old * nd)
        want = O.adapt_forest_transfer(dim, k, n, tree, new, u, n_new * nd)
        a = torch.from_numpy(u).cuda()
        b = torch.zeros(n_new * nd, dtype=torch.float64, device="cuda")
        X.adapt_forest_transfer(ctx, dim, k, n, tree, new, a, b)
        got = b.cpu().numpy()
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
        tree = new
