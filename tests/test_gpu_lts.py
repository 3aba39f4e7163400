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
            if step == 1:  # the second step in place (src == dst)
                lts.step(t, src, src)
                cur = src
            else:
                lts.step(t, src, dst)
                cur = dst
            out = np.zeros_like(uo)
            L.ox_lts_step(h, t, O._p(uo), O._p(out))
            uo = out
            got = cur.cpu().numpy()
            rel = np.max(np.abs(got - uo)) / np.max(np.abs(uo))
            assert rel <= 1e-12, (step, rel)
            if step != 1:
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


def _two_material_box(dim, n):
    """x < 1/2 refined once with rho = c = 1; x >= 1/2 with rho = 2, c = 4."""
    ref = np.zeros((n,) * dim, dtype=np.int32)
    ref[..., : n // 2] = 1
    m = O.Mesh.box_refined(dim, n, ref.ravel(), O.TRANSFORM_NONE, 0.0, 2)
    xc = m.vertices[:, :, 0].mean(axis=1)
    m.set_material(np.where(xc < 0.5, 1.0, 2.0), np.where(xc < 0.5, 1.0, 4.0))
    return m


NC_CASES = [  # dim, base intervals, degree, recon, source, max clusters, max diff
    (3, 4, 3, 0, True, 2, 1),   # the two-material refined-box workload, small
    (3, 4, 3, 1, True, 2, 1),
    (3, 4, 2, 1, False, 2, 1),
    (2, 8, 3, 1, True, 2, 1),
    (2, 8, 4, 0, True, 2, 1),
    (2, 6, 2, 1, True, 1, 3),   # one cluster: global ADER on the hanging-face mesh
]


@pytest.mark.parametrize("case", NC_CASES)
def test_lts_hanging_faces_speed_aware_source(X, ctx, case):
    """ADER-LTS over 2:1 hanging faces with clusters by h/c (the coarse c=4 half
    is the fast cluster, step ratio 2) and a Ricker point source at a vertex of
    the refined half: the cluster report equals the oracle's and three steps
    match the oracle to 1e-12."""
    import torch
    dim, n, k, recon, with_src, maxc, maxd = case
    print("lts nc case", case, flush=True)
    m = _two_material_box(dim, n)
    op = X.WaveOperator(ctx, dim, k, m.vertices, m.neighbor, m.rho, m.speed, hanging=m.hanging)
    cfl = 0.2 / k ** 1.5
    fastest = cfl * m.min_vertex_distance_over_c
    assert abs(op.min_vertex_distance_over_c - m.min_vertex_distance_over_c) <= 1e-15
    src = O.Source.make([0.25, 0.5, 0.5][:dim], 1.0, 10.0) if with_src else None
    if with_src:
        op.set_source(X.Source.make([0.25, 0.5, 0.5][:dim], 1.0, 10.0))
    lts = X.LocalTimeStepping(op, cfl, fastest, maxc, maxd, 0.0, bool(recon), speed_aware=True)
    ref = O.Lts(m, k, cfl, fastest, maxc, maxd, 0.0, recon, speed_aware=True, source=src)
    cl, fl = lts.cells()
    rcl, rfl = ref.cells()
    assert np.array_equal(cl, rcl) and np.array_equal(fl, rfl)
    assert lts.n_clusters == ref.n_clusters and lts.global_dt == ref.global_dt
    if maxc == 2:
        xc = m.vertices[:, :, 0].mean(axis=1)
        assert lts.n_clusters == 2 and np.all(cl[xc >= 0.5] == 0) and np.all(cl[xc < 0.5] == 1)
    e = O.Exact.make(dim, [[1] * dim, [2] + [1] * (dim - 1)])
    u = m.project(k, e, 0.0)
    s_, d_ = torch.from_numpy(u.copy()).cuda(), torch.zeros(u.size, dtype=torch.float64,
                                                          device="cuda")
    t = 0.0
    for step in range(3):
        t += lts.global_dt
        lts.step(t, s_, d_)
        u = ref.step(t, u)
        got = d_.cpu().numpy()
        rel = np.max(np.abs(got - u)) / np.max(np.abs(u))
        print("  step", step, "rel", rel, flush=True)
        assert rel <= 1e-12, (step, rel)
        s_, d_ = d_, s_


def test_source_add_matches_oracle(X, ctx):
    """exw_source_add after global ADER steps on the refined two-material box
    against the oracle's ox_source_add (same cells, same vector)."""
    import torch
    dim, k = 3, 3
    m = _two_material_box(dim, 4)
    op = X.WaveOperator(ctx, dim, k, m.vertices, m.neighbor, m.rho, m.speed, hanging=m.hanging)
    op.set_source(X.Source.make([0.25, 0.5, 0.5], 1.0, 10.0))
    src = O.Source.make([0.25, 0.5, 0.5], 1.0, 10.0)
    dt = 0.2 / k ** 1.5 * m.min_vertex_distance_over_c
    u = np.zeros(m.n_dofs(k))
    d = torch.zeros(u.size, dtype=torch.float64, device="cuda")
    dn = torch.zeros_like(d)
    work = torch.zeros(2 * u.size, dtype=torch.float64, device="cuda")
    t = 0.09
    for step in range(3):
        vn, vnp = u.copy(), np.zeros_like(u)
        m.time_step(k, O.ADER, dt, vn, vnp)
        m.source_add(k, src, t, t + dt, vnp)
        u = vnp
        op.time_step(X.ADER, dt, d, dn, work)
        op.source_add(t, t + dt, dn)
        d, dn = dn, d
        t += dt
    got = d.cpu().numpy()
    assert np.max(np.abs(u)) > 0
    assert np.max(np.abs(got - u)) <= 1e-12 * np.max(np.abs(u))
