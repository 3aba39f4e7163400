"""Several global ADER steps at once (exw_ader_steps) against single steps.

On box meshes without hanging faces of 3D degree 3-5 the update sweep of each
step runs fused with the next step's CK time integral: ader_sweep_ck_kernel
(staged neighbour fields; the FP32 operator) and ader_sweepg_ck_kernel
(gathered neighbour traces; FP64, two cells per group at degree 3).  Both
evaluate the separate kernels' expressions in the same order, so n steps must
give exactly the bits of n exw_ader_step calls, in FP64 and with the FP32
operator, for every step count (the two CK fields alternate).  Meshes
outside the fused kernel (hanging faces, deformed cells, 2D) take one step
after the other in place, with the same bits.  The oracle's ADER steps anchor
the chain (<= 1e-12 relative).
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # dim, intervals, degree, boundary id, material seed, refine seed, transform
    (3, 3, 4, 2, 1, None, 0),
    (3, 3, 3, 1, None, None, 0),
    (3, 2, 5, 3, 2, None, 0),
    (3, 1, 4, 2, None, None, 0),  # one cell: every face a boundary
    (3, 2, 4, 2, 3, 7, 0),        # hanging faces: one step after the other
    (3, 2, 4, 1, None, None, 1),  # deformed cells: one step after the other
    (2, 4, 5, 2, 4, None, 0),     # 2D: one step after the other
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


def _mesh(X, ctx, dim, n, k, bid, mseed, rseed, tr):
    transform = O.TRANSFORM_NONE if tr == 0 else O.TRANSFORM_REFERENCE_SINE
    amp = 0.0 if tr == 0 else 0.05
    if rseed is None:
        om = O.Mesh(dim, n, 0, transform, amp, bid)
        hanging = None
    else:
        ref = (np.random.default_rng(rseed).uniform(size=(n,) * dim) < 0.5).astype(np.int32)
        ref.flat[0], ref.flat[-1] = 1, 0
        om = O.Mesh.box_refined(dim, n, ref, transform, amp, bid)
        hanging = om.hanging
    if mseed is not None:
        rng = np.random.default_rng(mseed)
        om.set_material(rng.uniform(0.5, 2.0, om.n_cells), rng.uniform(0.5, 2.0, om.n_cells))
    op = X.WaveOperator(ctx, dim, k, om.vertices, om.neighbor, om.rho, om.speed,
                        hanging=hanging)
    assert op.is_cartesian == (tr == 0)
    return om, op


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d_n%d_k%d_b%d_%s_t%d" % (
    c[0], c[1], c[2], c[3], "nc" if c[5] is not None else "conf", c[6]))
def test_ader_steps_equal_single_steps(X, ctx, case):
    import torch
    dim, n, k, bid, mseed, rseed, tr = case
    _say("ader steps", case)
    om, op = _mesh(X, ctx, dim, n, k, bid, mseed, rseed, tr)
    N = op.n_dofs
    u0 = np.random.default_rng(500 + k).standard_normal(N)
    dt = 0.1 / k ** 1.5 * op.min_vertex_distance_over_c
    for precision in (X.FP64, X.FP32):
        # single steps
        a = torch.from_numpy(u0).cuda()
        b = torch.empty_like(a)
        work = torch.empty(2 * N, dtype=torch.float64, device="cuda")
        singles = []
        for s in range(4):
            op.time_step(X.ADER, dt, a, b, work, precision=precision)
            a, b = b, a
            singles.append(a.cpu().numpy())
        # n steps at once, n = 0 .. 4
        for ns in range(5):
            v = torch.from_numpy(u0).cuda()
            w = torch.full((2 * N,), 3.0, dtype=torch.float64, device="cuda")
            op.ader_steps(ns, dt, v, w, precision=precision)
            torch.cuda.synchronize()
            got = v.cpu().numpy()
            want = u0 if ns == 0 else singles[ns - 1]
            assert np.array_equal(got, want), "%d steps differ (precision %d)" % (ns, precision)
    # the oracle's steps anchor the chain (FP64)
    ref = u0.copy()
    for s in range(3):
        nxt = np.empty_like(ref)
        om.time_step(k, O.ADER, dt, ref, nxt)
        ref = nxt
    v = torch.from_numpy(u0).cuda()
    w = torch.empty(2 * N, dtype=torch.float64, device="cuda")
    op.ader_steps(3, dt, v, w)
    torch.cuda.synchronize()
    got3 = v.cpu().numpy()
    assert np.max(np.abs(got3 - ref)) <= 1e-12 * np.max(np.abs(ref)), "3 steps vs oracle"


def test_ader_steps_argument_errors(X, ctx):
    import torch
    om = O.Mesh(3, 1, 0, O.TRANSFORM_NONE, 0.0, 2)
    op = X.WaveOperator(ctx, 3, 4, om.vertices, om.neighbor, om.rho, om.speed)
    v = op.new_vector()
    w = op.new_vector(2)
    with pytest.raises(X.ExwError):
        op.ader_steps(-1, 0.01, v, w)
    with pytest.raises(X.ExwError):
        op.ader_steps(1, 0.01, v, w, precision=X.PRECISION_AUTO)
