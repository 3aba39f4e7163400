"""GPU parity of ADER with local time stepping (exw_lts_*) against the oracle's
restatement of WaveEquationOperationADERLTS + ClusterManager, and the
reference's aderlts_2d_recon fixture end to end."""
import numpy as np
import pytest

import oracle as O
from tests.test_oracle import _problem, lts_info

pytestmark = pytest.mark.gpu

CASES = [  # dim, intervals, transform factor, degree, max clusters, max diff
    (2, 15, 0.2, 5, 5, 1),   # the fixture's mesh: 3 clusters
    (2, 9, 0.3, 2, 3, 1),    # 7 clusters
    (2, 8, 0.3, 3, 5, 1),    # 6 clusters
    (3, 4, 0.25, 2, 5, 1),   # 4 clusters in 3D
    (3, 5, 0.25, 2, 2, 3),   # cluster_diff 2: step multiples 1, 3
]


@pytest.fixture(scope="module")
def X():
    import torch
    assert torch.cuda.is_available()
    import exwave_amd
    exwave_amd.load()
    return exwave_amd


@pytest.fixture(scope="module")
def ctx(X):
    return X.Context(0)


@pytest.mark.parametrize("recon", [0, 1])
@pytest.mark.parametrize("case", CASES)
def test_lts_steps_match_oracle(X, ctx, case, recon):
    import torch
    dim, nint, f, k, maxc, maxd = case
    print("case", case, "recon", recon, flush=True)
    op = X.WaveOperator.hypercube(ctx, dim, k, nint, 0, X.TRANSFORM_REFERENCE_SINE, f, 2)
    m = O.Mesh(dim, nint, 0, O.TRANSFORM_REFERENCE_SINE, f, 2)
    cfl = 0.2 / k ** 1.5
    fastest = cfl * m.min_vertex_distance_over_c
    lts = X.LocalTimeStepping(op, cfl, fastest, maxc, maxd, 0.0, bool(recon))
    info = lts_info(m, k, 0.2, maxc, maxd, recon)
    cl, fl = lts.cells()
    assert np.array_equal(cl, info["cluster"]) and np.array_equal(fl, info["flags"])
    assert lts.n_clusters == info["n_clusters"] and lts.global_dt == info["global_dt"]
    modes = [[1] * dim, [2] + [1] * (dim - 1)]
    e = O.Exact.make(dim, modes)
    u0 = m.project(k, e, 0.0)
    L = O.lib()
    h = L.ox_lts_create(m._m, k, cfl, fastest, maxc, maxd, 0.0, recon)
    try:
        src = torch.from_numpy(u0.copy()).cuda()
        dst = torch.zeros_like(src)
        uo = u0.copy()
        t = 0.0
        for step in range(3):
            t += lts.global_dt
            lts.step(t, src, dst)
            out = np.zeros_like(uo)
            L.ox_lts_step(h, t, O._p(uo), O._p(out))
            uo = out
            got = dst.cpu().numpy()
            rel = np.max(np.abs(got - uo)) / np.max(np.abs(uo))
            assert rel <= 1e-12, (step, rel)
            src, dst = dst, src
    finally:
        L.ox_lts_free(h)


def test_lts_fixture_end_to_end(X, ctx, fixtures):
    """exw_run with the reference's aderlts_2d_recon parameters reproduces every
    'error' line (4 printed digits, error p post included) and the step count."""
    fx = fixtures["aderlts_2d_recon"]
    prm = fx["params"]
    op_ = _problem(prm)
    p = X.Problem()
    for name in ("dim", "degree", "n_intervals", "n_refinements", "transform",
                 "transform_factor", "boundary_id", "integrator", "cfl_number",
                 "max_time_steps", "final_time", "output_every_time", "ader_recon",
                 "lts_max_clusters", "lts_max_diff"):
        setattr(p, name, getattr(op_, name))
    p.precision = X.FP64
    p.modes = X.Modes.make(prm["dimension"], [[prm["membrane_modes"]] * prm["dimension"]])
    p.post_pressure = 1
    steps, rows, _ = X.run(ctx, p)
    assert steps == fx["n_steps"]
    assert len(rows) == len(fx["error_lines"])
    for (st, t, ep, ev, mp, epp), ref in zip(rows, fx["error_lines"]):
        assert "%.2f" % t == ref[0]
        got = ["%.4e" % ep, "%.4e" % epp, "%.4e" % ev, "%.4e" % mp]
        assert got == ref[1:5], (t, got, ref)
