"""exw_run with adaptive levels (adapt_mesh before the loop and every
adaptive_refinement_interval steps, source/explicit_wave.cc:704-716, 789-791)
against the oracle's run of the same problem: step count and every error line.
The marking depends on comparisons of per-cell indicators, so the problems are
chosen without exact ties between the indicators (no mirror symmetry)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # dim, intervals, degree, transform, factor, integrator, interval, modes,
    #          n_refinements, n_adaptive_refinements
    (2, 8, 3, O.TRANSFORM_AXIS_2PI, 0.03, O.LSRK45R2, 10, [[2, 1], [1, 3]], 0, 1),
    (2, 6, 4, O.TRANSFORM_AXIS_2PI, 0.03, O.ADER, 5, [[1, 2], [3, 1]], 0, 1),
    (3, 4, 2, O.TRANSFORM_AXIS_2PI, 0.03, O.LSRK45R2, 8, [[1, 2, 3], [2, 1, 1]], 0, 1),
    # ADER-LTS: clusters rebuilt after every adaptation (make_dofs -> setup)
    (2, 4, 3, O.TRANSFORM_AXIS_2PI, 0.03, O.ADER_LTS, 6, [[2, 1], [1, 3]], 0, 2),
    # several levels on the forest (levels n_refinements .. + n_adaptive)
    (2, 4, 3, O.TRANSFORM_AXIS_2PI, 0.03, O.LSRK45R2, 10, [[2, 1], [1, 3]], 1, 2),
    (2, 5, 2, O.TRANSFORM_AXIS_2PI, 0.03, O.ADER, 6, [[1, 2], [3, 1]], 0, 3),
    (3, 2, 2, O.TRANSFORM_REFERENCE_SINE, 0.05, O.LSRK45R2, 8, [[1, 2, 3], [2, 1, 1]], 1, 2),
]


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


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d_k%d_i%d_r%d_a%d" % (
    c[0], c[2], c[5], c[8], c[9]))
def test_adaptive_run_matches_oracle(X, ctx, case):
    dim, nint, k, tr, fac, integ, interval, modes, nref, nada = case
    print("adaptive run", case, flush=True)
    po = O.Problem()
    px = X.Problem()
    for p in (po, px):
        p.dim, p.degree, p.n_intervals, p.n_refinements = dim, k, nint, nref
        p.transform, p.transform_factor, p.boundary_id = tr, fac, 2
        p.integrator, p.cfl_number = integ, 0.2
        p.max_time_steps, p.final_time, p.output_every_time = 100000, 0.15, 0.05
        p.post_pressure, p.n_adaptive_refinements, p.adaptive_refinement_interval = 1, nada, interval
        p.lts_max_clusters, p.lts_max_diff, p.ader_recon = 3, 1, 0
    po.exact = O.Exact.make(dim, modes)
    px.precision = X.FP64
    px.modes = X.Modes.make(dim, modes)
    # the data give distinct indicators: the initial marking is the same under
    # 1e-11 relative perturbations (a tie would be broken by round-off)
    m0 = O.Mesh.forest(dim, nint, O.uniform_forest(dim, nint, nref), tr, fac, 2)
    err = m0.estimate_error(k, m0.project(k, po.exact, 0.0))
    lev = np.full(m0.n_cells, nref, np.int32)
    base = X.refinement_flags(err, lev, nref, nref + nada, -1.0)
    rng = np.random.default_rng(0)
    for _ in range(10):
        pert = err * (1 + 1e-11 * rng.standard_normal(err.size))
        assert np.array_equal(X.refinement_flags(pert, lev, nref, nref + nada, -1.0), base)
    so, ro, _ = O.run(po)
    sx, rx, _ = X.run(ctx, px)
    assert sx == so and len(rx) == len(ro)
    for a, b in zip(rx, ro):
        assert a[0] == b[0] and abs(a[1] - b[1]) <= 1e-13
        for j in (2, 3, 4, 5):
            assert abs(a[j] - b[j]) <= 1e-9 * abs(b[j]) + 1e-15, (a, b)
