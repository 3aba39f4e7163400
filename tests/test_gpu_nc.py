"""GPU parity on meshes with 2:1 hanging faces (one level of local
refinement): the operator sweep (Cartesian or general kernel) plus the
subface kernel (nc_face_kernel) against the oracle's restatement of deal.II's
subface integrals, through the C-ABI.  Tolerances as for conforming meshes:
FP64 <= 1e-12 relative in max-norm, FP32 operator <= 2e-5."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # dim, base intervals, degree, transform, factor, refine seed, material seed
    (2, 4, 1, O.TRANSFORM_NONE, 0.0, 1, None),
    (2, 4, 3, O.TRANSFORM_NONE, 0.0, 2, 5),
    (2, 3, 5, O.TRANSFORM_REFERENCE_SINE, 0.1, 3, 6),
    (2, 3, 8, O.TRANSFORM_AXIS_2PI, 0.04, 4, None),
    (3, 3, 1, O.TRANSFORM_NONE, 0.0, 5, 7),
    (3, 3, 2, O.TRANSFORM_REFERENCE_SINE, 0.1, 6, None),
    (3, 3, 4, O.TRANSFORM_NONE, 0.0, 7, 8),
    (3, 2, 6, O.TRANSFORM_REFERENCE_SINE, 0.1, 8, 9),
    (3, 2, 8, O.TRANSFORM_NONE, 0.0, 9, None),
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


def _pair(X, ctx, dim, n, k, tr, fac, rseed, mseed, bid=2):
    ref = (np.random.default_rng(rseed).uniform(size=(n,) * dim) < 0.5).astype(np.int32)
    ref.flat[0] = 1
    ref.flat[-1] = 0
    om = O.Mesh.box_refined(dim, n, ref, tr, fac, bid)
    if mseed is not None:
        rng = np.random.default_rng(mseed)
        om.set_material(rng.uniform(0.5, 2.0, om.n_cells), rng.uniform(0.5, 2.0, om.n_cells))
    H = om.hanging
    assert len(H) > 0
    op = X.WaveOperator(ctx, dim, k, om.vertices, om.neighbor, om.rho, om.speed, hanging=H)
    return om, op


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d_n%d_k%d_t%d" % c[:4])
def test_apply_parity_hanging(X, ctx, case):
    import torch
    dim, n, k, tr, fac, rseed, mseed = case
    _say("nc apply", case)
    om, op = _pair(X, ctx, dim, n, k, tr, fac, rseed, mseed)
    if tr == O.TRANSFORM_NONE:  # box cells: the Cartesian sweep plus the subface kernel
        assert op.is_cartesian
    u = np.random.default_rng(rseed + 100).standard_normal(om.n_dofs(k))
    want = om.apply(k, u)
    src = torch.from_numpy(u).cuda()
    dst = torch.zeros_like(src)
    op.apply(src, dst)
    torch.cuda.synchronize()
    assert _rel(dst.cpu().numpy(), want) <= 1e-12
    op.apply(src, dst, X.FP32)
    torch.cuda.synchronize()
    assert _rel(dst.cpu().numpy(), want) <= 2e-5


@pytest.mark.parametrize("integ", [O.LSRK45R2, O.ADER, O.LSRK45R3, O.SSPRK, O.CLRK4])
def test_time_step_parity_hanging(X, ctx, integ):
    """The fused integrators (every epilogue output form: o1, o1+o2 from o1,
    o1+o2 from c, three outputs) with the subface increments."""
    import torch
    _say("nc time step", integ)
    for dim, n, k, tr, fac in [(2, 4, 3, O.TRANSFORM_REFERENCE_SINE, 0.1),
                               (3, 3, 3, O.TRANSFORM_NONE, 0.0)]:
        om, op = _pair(X, ctx, dim, n, k, tr, fac, 21, 22)
        N = om.n_dofs(k)
        e = O.Exact.make(dim, [[1] * dim])
        u0 = om.project(k, e, 0.0) + 1e-3 * np.random.default_rng(5).standard_normal(N)
        dt = 0.1 / k ** 1.5 * om.min_vertex_distance_over_c
        vn, vnp = u0.copy(), np.zeros(N)
        om.time_step(k, integ, dt, vn, vnp, ader_recon=1 if integ == O.ADER else 0)
        dn, dnp = torch.from_numpy(u0.copy()).cuda(), torch.zeros(N, dtype=torch.float64,
                                                                  device="cuda")
        work = torch.zeros(max(op.work_vectors(integ), 1) * N, dtype=torch.float64,
                           device="cuda")
        op.time_step(integ, dt, dn, dnp, work, ader_recon=integ == O.ADER)
        torch.cuda.synchronize()
        assert _rel(dnp.cpu().numpy(), vnp) <= 1e-12, (dim, integ)


def test_hanging_records_are_checked(X, ctx):
    """A record whose fine cells are listed in the wrong subface order, or an
    EXW_NC_FACE entry without a record, is rejected before anything runs."""
    ref = np.zeros((4, 4), dtype=np.int32)
    ref[:, :2] = 1
    om = O.Mesh.box_refined(2, 4, ref)
    H = om.hanging.copy()
    H[0, 2], H[0, 3] = H[0, 3], H[0, 2]
    with pytest.raises(X.ExwError):
        X.WaveOperator(ctx, 2, 2, om.vertices, om.neighbor, om.rho, om.speed, hanging=H)
    with pytest.raises(X.ExwError):
        X.WaveOperator(ctx, 2, 2, om.vertices, om.neighbor, om.rho, om.speed,
                       hanging=om.hanging[1:])
    op = X.WaveOperator(ctx, 2, 2, om.vertices, om.neighbor, om.rho, om.speed,
                        hanging=om.hanging)
    with pytest.raises(X.ExwError):  # bad LTS arguments are rejected on such meshes too
        X.LocalTimeStepping(op, 0.1, -0.01, 2, 1, 0.0, False)
