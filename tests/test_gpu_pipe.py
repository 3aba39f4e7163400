"""The two Cartesian stage kernels agree bit for bit, through the C-ABI.

An FP64 low-storage RK stage on Cartesian cells of 3D k=3..5, and any FP64
sweep on 2D k=6-8, runs the software-pipelined kernel
(dg_operator_pipe_kernel), also with cell scales.  With a c operand (o2 = c +
a2 k, c set to the pipelined o1 so that the result is the same), or with
vectors that are not 16-byte aligned (3D), the same sweep runs
dg_operator_cart_kernel instead.  Both evaluate the same expressions in the
same order (the pipelined kernel's flux coefficients are precomputed by the
Cartesian kernel's face-lane code), so the outputs must be identical, also on
meshes with hanging faces, with every boundary id, and with cell scales of
which some are zero (inactive cells: their outputs are left untouched).  A
first check against the oracle (<= 1e-12 relative) keeps the comparison
anchored.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # dim, intervals, degree, boundary id, material seed, refine seed (None: conforming)
    (3, 4, 4, 2, 1, None),
    (3, 4, 4, 1, None, None),
    (3, 3, 3, 3, 2, None),
    (3, 3, 4, 2, 3, 11),
    (3, 3, 5, 2, 6, None),
    (3, 3, 5, 1, None, 13),
    (2, 5, 6, 2, 4, None),
    (2, 4, 7, 3, None, None),
    (2, 4, 8, 1, 5, 12),
    (3, 1, 4, 3, 9, None),  # one cell: one group with one cell, every face a boundary
    (2, 1, 7, 1, None, None),
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


def _mesh(X, ctx, dim, n, k, bid, mseed, rseed):
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


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d_n%d_k%d_b%d_%s" % (
    c[0], c[1], c[2], c[3], "nc" if c[5] is not None else "conf"))
def test_pipelined_and_cartesian_kernels_agree(X, ctx, case):
    import torch
    dim, n, k, bid, mseed, rseed = case
    _say("pipe vs cart", case)
    om, op = _mesh(X, ctx, dim, n, k, bid, mseed, rseed)
    N = op.n_dofs
    rng = np.random.default_rng(100 + k)
    u = rng.standard_normal(N)
    bh = rng.standard_normal(N)
    a1, a2 = 0.37, -0.81
    src, b = torch.from_numpy(u).cuda(), torch.from_numpy(bh).cuda()
    assert src.data_ptr() % 16 == 0 and b.data_ptr() % 16 == 0

    def sweep(src, b, cell_scale=None, c=None, fill=0.0):
        o1 = torch.full((N,), fill, dtype=torch.float64, device="cuda")
        o2 = torch.full((N,), fill, dtype=torch.float64, device="cuda")
        op.apply_fused(src, b=b, o1=o1, a1=a1, c=c, o2=o2, a2=a2, o2_from_o1=c is None,
                       cell_scale=cell_scale)
        torch.cuda.synchronize()
        return o1.cpu().numpy(), o2.cpu().numpy()

    p1, p2 = sweep(src, b)  # pipelined
    # anchor: o1 = b + a1 M^-1 A u against the oracle
    ref = bh + a1 * om.apply(k, u)
    assert np.max(np.abs(p1 - ref)) <= 1e-12 * np.max(np.abs(ref)), "pipelined kernel vs oracle"
    # Cartesian kernel: a c operand equal to the pipelined o1.  (On hanging
    # faces the subface increment of o2 is then added as a2 dk to o1 + a2 k
    # instead of (a1 + a2) dk to the sweep's o1 + a2 k: equal, rounded apart.)
    def same_o2(x, y):
        if rseed is None:
            return np.array_equal(x, y)
        return np.max(np.abs(x - y)) <= 1e-14 * np.max(np.abs(y))

    pc = torch.from_numpy(p1).cuda()
    c1, c2 = sweep(src, b, c=pc)
    assert np.array_equal(p1, c1) and same_o2(c2, p2), "c-operand path differs"
    # Cartesian kernel (8-byte offset vectors: not 16-byte aligned, 3D)
    sb = torch.empty(N + 1, dtype=torch.float64, device="cuda")
    bb = torch.empty(N + 1, dtype=torch.float64, device="cuda")
    sb[1:].copy_(src)
    bb[1:].copy_(b)
    m1, m2 = sweep(sb[1:], bb[1:])
    assert np.array_equal(p1, m1) and np.array_equal(p2, m2), "unaligned path differs"
    # cell scales: all ones (same result), then some zero (inactive cells keep
    # their outputs) with the scaled coefficients, against the Cartesian kernel
    ones = torch.ones(om.n_cells, dtype=torch.float64, device="cuda")
    s1, s2 = sweep(src, b, ones)
    assert np.array_equal(p1, s1) and np.array_equal(p2, s2), "unit cell scales differ"
    sc = np.random.default_rng(200 + k).uniform(0.5, 2.0, om.n_cells)
    sc[::3] = 0.0
    scale = torch.from_numpy(sc).cuda()
    q1, q2 = sweep(src, b, scale, fill=7.0)
    nd = N // om.n_cells
    dead = np.repeat(sc == 0.0, nd)
    assert np.all(q1[dead] == 7.0) and np.all(q2[dead] == 7.0), "inactive cells were written"
    r1, r2 = sweep(src, b, scale, c=torch.from_numpy(q1).cuda(), fill=7.0)
    assert np.array_equal(q1, r1) and same_o2(r2, q2), "scaled cells differ"
    want = bh + a1 * np.repeat(sc, nd) * om.apply(k, u)
    live = ~dead
    if live.any():  # (a one-cell mesh has no active cell here)
        assert np.max(np.abs(q1[live] - want[live])) <= 1e-12 * np.max(np.abs(want[live]))
    # the last low-storage RK stage (one output, y = b + f1 k) through exw_rk_stage
    y = torch.empty(N, dtype=torch.float64, device="cuda")
    op.rk_stage(src, b, y, None, a1, 0.0, True)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), p1), "last RK stage differs from the fused stage's o1"
    # without an accumulator (o1 = a1 k; 2D: pipelined, 3D: Cartesian), against
    # the c-operand path of the Cartesian kernel
    o1 = torch.empty(N, dtype=torch.float64, device="cuda")
    op.apply_fused(src, o1=o1, a1=a1)
    torch.cuda.synchronize()
    z1 = o1.cpu().numpy()
    o1c = torch.empty(N, dtype=torch.float64, device="cuda")
    o2c = torch.empty(N, dtype=torch.float64, device="cuda")
    op.apply_fused(src, o1=o1c, a1=a1, c=o1, o2=o2c, a2=0.0)
    torch.cuda.synchronize()
    assert np.array_equal(z1, o1c.cpu().numpy()), "no-accumulator paths differ"
