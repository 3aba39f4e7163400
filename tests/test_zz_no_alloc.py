"""exw_l2_errors_noalloc (the entry point that allocates nothing) against
exw_l2_errors through the C-ABI: two calls in a row on the same GPU, the
second on a smaller operator, each equal to the allocating entry point's
result (same partial sums in the same order: equal to the last ulp of the
square root)."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_l2_errors_noalloc_matches_allocating_entry_point():
    import torch
    assert torch.cuda.is_available()
    import exwave_amd as X
    L = X.load()
    ctx = X.Context(0)
    cases = [(3, 6, 3, O.TRANSFORM_REFERENCE_SINE, [[1, 1, 1], [2, 3, 1]], 0.13),
             (2, 3, 2, O.TRANSFORM_NONE, [[2, 1]], 0.31)]  # second call: smaller input
    out = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    for dim, n, k, tr, mm, t in cases:
        _say("noalloc l2 errors", dim, n, k, tr, mm, t)
        om = O.Mesh(dim, n, 0, tr, 0.1, 2)
        rng = np.random.default_rng(dim * 10 + k)
        om.set_material(rng.uniform(0.5, 2, om.n_cells), rng.uniform(0.5, 2, om.n_cells))
        op = X.WaveOperator(ctx, dim, k, om.vertices, om.neighbor, om.rho, om.speed)
        modes = X.Modes.make(dim, mm)
        u = op.new_vector()
        op.project(modes, 0.0, u)
        u += 1e-3 * torch.from_numpy(rng.standard_normal(op.n_dofs)).cuda()
        want = op.l2_errors(modes, t, u)
        X.check(L.exw_l2_errors_noalloc(op.h, C.byref(modes), t, C.c_void_p(u.data_ptr()),
                                        C.c_void_p(out.data_ptr())))
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        _say("  want", want, "got", got)
        assert np.all(np.abs(got - want) <= 1e-15 * np.abs(want)), (got, want)
        # and against the oracle, the checker of both
        e = O.Exact.make(dim, mm)
        ref = om.l2_errors(k, e, t, u.cpu().numpy())
        assert np.max(np.abs(got - ref) / ref) <= 1e-11
