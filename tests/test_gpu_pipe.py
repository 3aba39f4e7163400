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


def test_headline_size_64cube(X, ctx):
    """The headline workload at its full size: 64^3 Cartesian cells numbered in
    8^3 tiles, k=4 (131,072,000 DoFs, 131,072 groups of the pipelined kernel
    remapped over the XCDs, vector offsets up to 1 GB).
      * one fused LSRK45 stage of the pipelined kernel is bit-identical to the
        same stage on the one-cell Cartesian kernel (8-byte offset vectors);
      * five LSRK45 steps at CFL 0.2 from p0 = sin(pi x)sin(pi y)sin(pi z) +
        sin(2 pi x)sin(3 pi y)sin(pi z): the discrete energy never increases
        beyond the round-off of its sum (upwind flux, p = 0 walls) and stays
        within 1e-9 relative of E0;
      * the L2 errors against the analytical solution: at t = 0 the closed
        form of the Gauss-point interpolant, after five steps within 1.5x of it."""
    import ctypes as C

    import torch
    from exwave_amd import serving
    _say("headline size: 64^3 k=4, tiles of 8^3")
    n1d, k = 64, 4
    V, nb = X.mesh_hypercube(3, n1d, 0, X.TRANSFORM_NONE, 0.0, 2)
    V, nb = X.tile_order(V, nb, n1d, 8)
    ones = np.ones(V.shape[0])
    op = X.WaveOperator(ctx, 3, k, V, nb, ones, ones)
    assert op.is_cartesian and op.n_dofs == 131_072_000
    N = op.n_dofs
    modes = X.Modes.make(3, [[1, 1, 1], [2, 3, 1]])
    u0 = op.new_vector()
    op.project(modes, 0.0, u0)
    # one stage (LSRK45 stage 2 coefficients) on both kernels
    f1, f2 = 0.41, -0.23
    b = u0 * 0.5
    y1, b1 = op.new_vector(), op.new_vector()
    op.rk_stage(u0, b, y1, b1, f1, f2, False)
    sb = torch.empty(N + 1, dtype=torch.float64, device="cuda")
    bb = torch.empty(N + 1, dtype=torch.float64, device="cuda")
    sb[1:].copy_(u0)
    bb[1:].copy_(b)
    y2, b2 = op.new_vector(), op.new_vector()
    op.rk_stage(sb[1:], bb[1:], y2, b2, f1, f2, False)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2) and torch.equal(b1, b2), "pipelined and Cartesian stages differ"
    del sb, bb, y2, b2, b
    _say("  stage kernels bit-identical")
    # closed-form t = 0 error of the interpolant
    r = serving.make_request(dict(dim=3, degree=k, n_refinements=6, rho=1.0, c=1.0,
                                  modes=[1, 1, 1], final_time=1.0, cfl=0.2), X.FP64)
    r.modes = modes
    e0, n0, sc, pr = C.c_double(), C.c_double(), C.c_double(), C.c_int()
    X.check(X.load().exw_request_initial_error(C.byref(r), C.byref(e0), C.byref(n0), C.byref(sc),
                                               C.byref(pr)))
    err0 = op.l2_errors(modes, 0.0, u0)
    # (the error, 2.5e-10 against |p| = 0.7, is a sum of squares of differences
    # each rounded at 1e-16 |p|: agreement to ~1e-7 relative is all there is)
    assert abs(err0[0] - e0.value) <= 1e-5 * e0.value
    assert abs(err0[2] - n0.value) <= 1e-12 * n0.value
    dt = 0.2 / k ** 1.5 * op.min_vertex_distance_over_c
    del b1
    vec_n, vec_np = y1, u0
    work = torch.empty(3 * N, dtype=torch.float64, device="cuda")
    E = [op.energy(vec_np)]
    t = 0.0
    for s in range(5):
        vec_n, vec_np = vec_np, vec_n
        op.time_step(X.LSRK45R2, dt, vec_n, vec_np, work)
        t += dt
        E.append(op.energy(vec_np))
    _say("  energies", E)
    # non-increasing up to the round-off of the energy sum itself
    assert all(E[i + 1] <= E[i] * (1 + 1e-13) for i in range(5)), E
    assert E[0] - E[-1] <= 1e-9 * E[0]
    err = op.l2_errors(modes, t, vec_np)
    _say("  errors t=0", err0, "t=%g" % t, err)
    # the interpolation error plus five steps of time-stepping error
    assert err[0] <= 1.5 * err0[0] and err[1] <= 1.5 * err0[0]
