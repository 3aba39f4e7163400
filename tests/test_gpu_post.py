"""GPU parity of the post-processed pressure (exw_post_pressure, exw_post_error)
against the oracle's restatement of compute_post_pressure
(source/wave_equation_operations.cc:786-848).

The kernel runs CG on the zero-sum subspace of the local Neumann problem to
|r|/|r0| < 1e-13; it is compared with the oracle's same-subspace CG at that
tolerance (<= 1e-10 relative in max-norm: both stop on a residual, the
remaining difference is CG round-off), and with the oracle's restatement of the
reference's own solver (first unknown fixed, 1e-11) to <= 1e-8, the accuracy
that solver itself reaches.  The end-to-end 'error p post' column of the
reference fixtures is checked in test_gpu_parity.py and test_gpu_lts.py.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # dim, intervals, transform, factor, degree, material seed
    (2, 4, O.TRANSFORM_NONE, 0.0, 1, None),
    (2, 4, O.TRANSFORM_REFERENCE_SINE, 0.2, 3, 1),
    (2, 3, O.TRANSFORM_REFERENCE_SINE, 0.2, 5, None),
    (2, 3, O.TRANSFORM_AXIS_2PI, 0.04, 8, 2),
    (3, 3, O.TRANSFORM_REFERENCE_SINE, 0.1, 1, 3),
    (3, 3, O.TRANSFORM_NONE, 0.0, 2, None),
    (3, 2, O.TRANSFORM_REFERENCE_SINE, 0.1, 4, 4),
    (3, 2, O.TRANSFORM_AXIS_2PI, 0.05, 6, None),
    (3, 2, O.TRANSFORM_REFERENCE_SINE, 0.1, 8, 5),
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


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d_n%d_t%d_k%d" % (c[0], c[1], c[2], c[4]))
def test_post_pressure_matches_oracle(X, ctx, case):
    import torch
    dim, nint, tr, fac, k, seed = case
    _say("post pressure", case)
    om = O.Mesh(dim, nint, 0, tr, fac)
    if seed is not None:
        rng = np.random.default_rng(seed)
        om.set_material(rng.uniform(0.5, 2.0, om.n_cells), rng.uniform(0.5, 2.0, om.n_cells))
    op = X.WaveOperator(ctx, dim, k, om.vertices, om.neighbor, om.rho, om.speed)
    e = O.Exact.reference(dim, 2)
    u = om.project(k, e, 0.3)
    u += 0.05 * np.random.default_rng(11).standard_normal(u.shape) * np.abs(u).max()
    assert op.n_post == om.n_cells * (k + 2) ** dim
    src = torch.from_numpy(u).cuda()
    tmp = torch.zeros_like(src)
    post = torch.zeros(op.n_post, dtype=torch.float64, device="cuda")
    its = torch.zeros(om.n_cells, dtype=torch.int32, device="cuda")
    op.post_pressure(src, tmp, post, its)
    torch.cuda.synchronize()
    got, git = post.cpu().numpy(), its.cpu().numpy()
    assert _rel(tmp.cpu().numpy(), om.apply(k, u)) <= 1e-12
    want, wit = om.post_pressure(k, u, rel_tol=1e-13, mean_free=True, iterations=True)
    assert np.all(git >= 1) and np.all(git <= 6 * (k + 2) ** dim + 1)
    assert np.abs(git - wit).max() <= max(2, 0.05 * wit.max()), (git, wit)
    assert _rel(got, want) <= 1e-10, _rel(got, want)
    # the reference's own solver (first unknown fixed, |r|/|r0| < 1e-11)
    ref = om.post_pressure(k, u)
    assert _rel(got, ref) <= 1e-8, _rel(got, ref)
    # the post error
    modes = X.Modes.reference(dim, 2)
    t = 0.3
    gerr = op.post_error(modes, t, post)
    werr = om.post_error(k, e, t, want)
    assert abs(gerr - werr) <= 1e-10 * werr, (gerr, werr)


def test_post_pressure_constant_and_zero(X, ctx):
    """A constant pressure with zero velocity: apply gives v_t = 0 inside and
    the boundary flux only; p* keeps every cell average.  An all-zero state
    gives p* = 0 (zero right-hand side: one CG iteration, the division guard)."""
    import torch
    om = O.Mesh(2, 3, 0, O.TRANSFORM_REFERENCE_SINE, 0.1)
    k = 2
    op = X.WaveOperator(ctx, 2, k, om.vertices, om.neighbor, om.rho, om.speed)
    u = np.zeros(om.n_dofs(k))
    src = torch.from_numpy(u).cuda()
    tmp = torch.zeros_like(src)
    post = torch.full((op.n_post,), 7.0, dtype=torch.float64, device="cuda")
    its = torch.zeros(om.n_cells, dtype=torch.int32, device="cuda")
    op.post_pressure(src, tmp, post, its)
    torch.cuda.synchronize()
    assert np.all(post.cpu().numpy() == 0.0)
    assert np.all(its.cpu().numpy() == 1)


def test_post_pressure_rejects_aliasing(X, ctx):
    import torch
    om = O.Mesh(2, 2, 0, O.TRANSFORM_NONE, 0.0)
    op = X.WaveOperator(ctx, 2, 1, om.vertices, om.neighbor, om.rho, om.speed)
    src = torch.zeros(om.n_dofs(1), dtype=torch.float64, device="cuda")
    post = torch.zeros(op.n_post, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        op.post_pressure(src, src, post)


@pytest.mark.parametrize("case", [(2, 5, O.TRANSFORM_REFERENCE_SINE, 0.2, 3, 1),
                                  (3, 3, O.TRANSFORM_NONE, 0.0, 4, 2),
                                  (3, 2, O.TRANSFORM_AXIS_2PI, 0.05, 6, None)],
                         ids=lambda c: "d%d_k%d" % (c[0], c[4]))
def test_estimate_error_matches_oracle(X, ctx, case):
    """exw_estimate_error (estimate_error, source/wave_equation_operations.cc:
    871-898) against the oracle, <= 1e-12 relative per cell."""
    import torch
    dim, nint, tr, fac, k, seed = case
    _say("estimate error", case)
    om = O.Mesh(dim, nint, 0, tr, fac)
    if seed is not None:
        rng = np.random.default_rng(seed)
        om.set_material(rng.uniform(0.5, 2.0, om.n_cells), rng.uniform(0.5, 2.0, om.n_cells))
    op = X.WaveOperator(ctx, dim, k, om.vertices, om.neighbor, om.rho, om.speed)
    u = np.random.default_rng(3).standard_normal(om.n_dofs(k))
    src = torch.from_numpy(u).cuda()
    tmp = torch.zeros_like(src)
    err = torch.zeros(om.n_cells, dtype=torch.float64, device="cuda")
    op.estimate_error(src, tmp, err)
    torch.cuda.synchronize()
    want = om.estimate_error(k, u)
    assert _rel(err.cpu().numpy(), want) <= 1e-12


def test_estimate_error_hanging_faces(X, ctx):
    """The indicator on a mesh with hanging faces (apply includes the subface
    integrals) against the oracle."""
    import torch
    ref = np.zeros((3, 3, 3), dtype=np.int32)
    ref[:, :, :1] = 1
    om = O.Mesh.box_refined(3, 3, ref)
    k = 3
    op = X.WaveOperator(ctx, 3, k, om.vertices, om.neighbor, om.rho, om.speed,
                        hanging=om.hanging)
    u = np.random.default_rng(8).standard_normal(om.n_dofs(k))
    src = torch.from_numpy(u).cuda()
    tmp = torch.zeros_like(src)
    err = torch.zeros(om.n_cells, dtype=torch.float64, device="cuda")
    op.estimate_error(src, tmp, err)
    torch.cuda.synchronize()
    assert _rel(err.cpu().numpy(), om.estimate_error(k, u)) <= 1e-12
