"""The gathered-trace Cartesian kernel agrees bit for bit with the staged one.

On Cartesian cells of 3D k >= 6 every sweep runs dg_operator_cartg_kernel,
which sums the neighbours' face traces from memory instead of staging the
neighbour fields in LDS.  It evaluates the same expressions in the same order
as dg_operator_cart_kernel (selected with EXW_CART_GATHER=0, which the library
reads at every launch), so the outputs must be identical in FP64 (the cell's
fields through LDS-DMA, and through registers where the vectors are only
8-byte aligned) and with the FP32 operator: with and without an accumulator,
with a c operand, with cell scales of which some are zero (inactive cells keep
their outputs), on every boundary id, with random materials and with hanging
faces.  A check against the oracle (<= 1e-12
relative, FP64) anchors the comparison.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # intervals, degree, boundary id, material seed, refine seed (None: conforming)
    (3, 6, 1, 2, None),
    (3, 6, 3, None, 7),
    (2, 7, 2, 3, None),
    (2, 8, 1, None, 5),
    (2, 9, 3, 4, None),
    (1, 6, 2, None, None),  # one cell: every face a boundary
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


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


def _mesh(X, ctx, n, k, bid, mseed, rseed):
    dim = 3
    if rseed is None:
        om = O.Mesh(dim, n, 0, O.TRANSFORM_NONE, 0.0, bid)
        hanging = None
    else:
        ref = (np.random.default_rng(rseed).uniform(size=(n,) * dim) < 0.5).astype(np.int32)
        ref.flat[0], ref.flat[-1] = 1, 0
        om = O.Mesh.box_refined(dim, n, ref, O.TRANSFORM_NONE, 0.0, bid)
        hanging = om.hanging
    if mseed is not None:
        rng = np.random.default_rng(mseed)
        om.set_material(rng.uniform(0.5, 2.0, om.n_cells), rng.uniform(0.5, 2.0, om.n_cells))
    op = X.WaveOperator(ctx, dim, k, om.vertices, om.neighbor, om.rho, om.speed,
                        hanging=hanging)
    assert op.is_cartesian
    return om, op


def _with_kernel(gather, fn):
    old = os.environ.get("EXW_CART_GATHER")
    os.environ["EXW_CART_GATHER"] = "1" if gather else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["EXW_CART_GATHER"]
        else:
            os.environ["EXW_CART_GATHER"] = old


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_k%d_b%d_%s" % (
    c[0], c[1], c[2], "nc" if c[4] is not None else "conf"))
def test_gathered_and_staged_kernels_agree(X, ctx, case):
    import torch
    n, k, bid, mseed, rseed = case
    _say("cartg vs cart", case)
    om, op = _mesh(X, ctx, n, k, bid, mseed, rseed)
    N = op.n_dofs
    rng = np.random.default_rng(300 + k)
    u = rng.standard_normal(N)
    bh = rng.standard_normal(N)
    ch = rng.standard_normal(N)
    a1, a2 = 0.37, -0.81
    src, b, c = (torch.from_numpy(x).cuda() for x in (u, bh, ch))
    sc = np.random.default_rng(400 + k).uniform(0.5, 2.0, om.n_cells)
    sc[::3] = 0.0
    scale = torch.from_numpy(sc).cuda()

    def sweeps(precision):
        out = []
        # accumulator and c operand (never the pipelined kernel)
        for kw in (dict(b=b, c=c), dict(b=None, c=c), dict(b=b, c=c, cell_scale=scale)):
            o1 = torch.full((N,), 7.0, dtype=torch.float64, device="cuda")
            o2 = torch.full((N,), 7.0, dtype=torch.float64, device="cuda")
            op.apply_fused(src, o1=o1, a1=a1, o2=o2, a2=a2, precision=precision, **kw)
            torch.cuda.synchronize()
            out += [o1.cpu().numpy(), o2.cpu().numpy()]
        # the plain operator
        d = torch.empty(N, dtype=torch.float64, device="cuda")
        op.apply(src, d, precision=precision)
        torch.cuda.synchronize()
        out.append(d.cpu().numpy())
        # vectors at an 8-byte offset (not 16-byte aligned)
        sb = torch.empty(N + 1, dtype=torch.float64, device="cuda")
        bb = torch.empty(N + 1, dtype=torch.float64, device="cuda")
        sb[1:].copy_(src)
        bb[1:].copy_(b)
        o1 = torch.empty(N + 1, dtype=torch.float64, device="cuda")
        o2 = torch.empty(N + 1, dtype=torch.float64, device="cuda")
        op.apply_fused(sb[1:], b=bb[1:], o1=o1[1:], a1=a1, o2=o2[1:], a2=a2, o2_from_o1=True,
                       precision=precision)
        torch.cuda.synchronize()
        out += [o1[1:].cpu().numpy(), o2[1:].cpu().numpy()]
        return out

    for precision in (X.FP64, X.FP32):
        g = _with_kernel(True, lambda: sweeps(precision))
        s = _with_kernel(False, lambda: sweeps(precision))
        for i, (x, y) in enumerate(zip(g, s)):
            assert np.array_equal(x, y), "output %d differs (precision %d)" % (i, precision)
        if precision == X.FP64:
            ref = om.apply(k, u)
            tol = 1e-12 * np.max(np.abs(ref))
            assert np.max(np.abs(g[6] - ref)) <= tol, "gathered kernel vs oracle"
            assert np.max(np.abs(g[0] - (bh + a1 * ref))) <= 1e-12 * np.max(np.abs(g[0]))
            nd = N // om.n_cells
            dead = np.repeat(sc == 0.0, nd)
            assert np.all(g[4][dead] == 7.0) and np.all(g[5][dead] == 7.0), \
                "inactive cells were written"
        # o1 = b + a1 k through registers (unaligned) and through LDS-DMA (aligned)
        assert np.array_equal(g[7], g[0]), "unaligned path differs (precision %d)" % precision
