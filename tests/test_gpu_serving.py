"""GPU tests of the serving loop (exw_server_*) and the one-at-a-time path
(exw_solve):
  * exw_solve (FP64) against the oracle running the same request,
  * every request gives the same result alone and in the mix (bit-identical
    for the same precision), and the same as exw_solve (graph replays with
    admissions in between against eager launches),
  * admission under a budget that holds one solve at a time,
  * with the automatic precision choice every result stays within 1.1x of the
    FP64 error, and the choice of FP32 for coarse solves actually happens.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

REQS = [
    dict(id=0, arrival=0.0, dim=2, degree=3, n_refinements=3, rho=1.3, c=0.7, modes=[2, 1],
         final_time=0.25, cfl=0.2),
    dict(id=1, arrival=0.0, dim=2, degree=2, n_refinements=4, rho=0.6, c=1.9, modes=[1, 3],
         final_time=0.3, cfl=0.2),
    dict(id=2, arrival=0.0, dim=3, degree=4, n_refinements=2, rho=1.0, c=1.0, modes=[1, 1, 2],
         final_time=0.2, cfl=0.2),
    dict(id=3, arrival=0.0, dim=2, degree=3, n_refinements=3, rho=2.0, c=0.5, modes=[4, 4],
         final_time=0.15, cfl=0.2),
    dict(id=4, arrival=0.0, dim=3, degree=2, n_refinements=3, rho=0.8, c=1.4, modes=[2, 1, 1],
         final_time=0.1, cfl=0.2),
    dict(id=5, arrival=0.0, dim=2, degree=6, n_refinements=3, rho=1.1, c=0.9, modes=[3, 2],
         final_time=0.2, cfl=0.2),
    # second instances in the pools that run the pipelined kernel (3D k=3/4,
    # 2D k=6): they start at a cell > 0 of their pool and their flux
    # coefficients are filled at admission with first > 0
    dict(id=6, arrival=0.0, dim=3, degree=4, n_refinements=2, rho=1.7, c=0.6, modes=[2, 1, 1],
         final_time=0.15, cfl=0.2),
    dict(id=7, arrival=0.0, dim=2, degree=6, n_refinements=2, rho=0.9, c=1.6, modes=[1, 2],
         final_time=0.1, cfl=0.2),
    dict(id=8, arrival=0.0, dim=3, degree=3, n_refinements=2, rho=1.2, c=1.1, modes=[1, 2, 1],
         final_time=0.12, cfl=0.2),
    dict(id=9, arrival=0.0, dim=3, degree=3, n_refinements=1, rho=0.7, c=0.8, modes=[1, 1, 1],
         final_time=0.2, cfl=0.2),
    # two instances in the 3D k=6 pool (the gathered-trace kernel over a cell
    # list, the second starting at a cell > 0) and one in the 3D k=5 pool
    dict(id=10, arrival=0.0, dim=3, degree=6, n_refinements=1, rho=1.4, c=0.9, modes=[1, 2, 1],
         final_time=0.05, cfl=0.2),
    dict(id=11, arrival=0.0, dim=3, degree=6, n_refinements=1, rho=0.8, c=1.3, modes=[2, 1, 1],
         final_time=0.04, cfl=0.2),
    dict(id=12, arrival=0.0, dim=3, degree=5, n_refinements=1, rho=1.1, c=1.0, modes=[1, 1, 2],
         final_time=0.05, cfl=0.2),
]


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available()
    import exwave_amd
    from exwave_amd import serving
    exwave_amd.load()
    return exwave_amd, serving


@pytest.fixture(scope="module")
def ctx(S):
    return S[0].Context(0)


def _oracle_solve(q):
    dim, k = q["dim"], q["degree"]
    m = O.Mesh(dim, 1 << q["n_refinements"], 0, O.TRANSFORM_NONE, 0.0, 2)
    m.set_material(np.full(m.n_cells, q["rho"]), np.full(m.n_cells, q["c"]))
    e = O.Exact.make(dim, [q["modes"]], rho=q["rho"], c=q["c"])
    u = m.project(k, e, 0.0)
    h = 1.0 / (1 << q["n_refinements"])
    dt = q["cfl"] / k ** 1.5 * h / q["c"]
    t, n = 0.0, 0
    while t < q["final_time"]:
        t += dt
        n += 1
    for _ in range(n):
        vn, vnp = u, np.zeros_like(u)
        m.time_step(k, O.LSRK45R2, dt, vn, vnp)
        u = vnp
    err = m.l2_errors(k, e, t, u)
    return n, t, err


@pytest.mark.parametrize("i", [0, 2, 4])
def test_solve_matches_oracle(S, ctx, i):
    X, srv = S
    q = REQS[i]
    res, _ = srv.run_one_at_a_time(ctx, [q], X.FP64)
    n, t, err = _oracle_solve(q)
    r = res[q["id"]]
    assert r["n_steps"] == n and r["time"] == t
    got = np.array([r["err_p"], r["err_v"], r["mag_p"]])
    assert np.max(np.abs(got - err) / err) < 1e-9, (got, err)


def _key(r):
    return (r["n_steps"], r["time"], r["err_p"], r["err_v"], r["mag_p"])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_alone_equals_mix(S, ctx, prec):
    X, srv = S
    p = X.FP64 if prec == "fp64" else X.FP32
    solo, _ = srv.run_one_at_a_time(ctx, REQS, p)
    alone = {}
    for q in REQS:
        r, _ = srv.run_server(ctx, [q], lambda _q: p, realtime=False)
        alone.update(r)
    mix, _ = srv.run_server(ctx, REQS, lambda _q: p, realtime=False)
    for q in REQS:
        i = q["id"]
        assert _key(alone[i]) == _key(mix[i]), (i, alone[i], mix[i])
        assert _key(solo[i]) == _key(mix[i]), (i, solo[i], mix[i])


def test_budget_admission(S, ctx):
    """A budget that holds about one solve at a time still serves every request,
    with the same results."""
    X, srv = S
    free, _ = srv.run_server(ctx, REQS, lambda _q: X.FP64, realtime=False)
    tight, _ = srv.run_server(ctx, REQS, lambda _q: X.FP64, hbm_budget_bytes=3 << 20,
                              realtime=False)
    for q in REQS:
        assert _key(free[q["id"]]) == _key(tight[q["id"]])


def test_auto_precision_error_bar(S, ctx):
    X, srv = S
    f64, _ = srv.run_server(ctx, REQS, lambda _q: X.FP64, realtime=False)
    auto, _ = srv.run_server(ctx, REQS, lambda _q: X.PRECISION_AUTO, realtime=False)
    n32 = 0
    for q in REQS:
        a, b = f64[q["id"]], auto[q["id"]]
        assert b["precision"] == srv.initial_error(q)[3]
        n32 += b["precision"] == X.FP32
        assert b["err_p"] <= 1.1 * a["err_p"] and b["err_v"] <= 1.1 * a["err_v"], (q, a, b)
    assert n32 >= 2


def test_many_instances_with_turnover(S, ctx):
    """Forty small solves of mixed length through one pool: instances finish
    and new ones are admitted into the freed cells while others keep stepping
    (the active-cell list changes every few steps); every result equals the
    solve alone."""
    X, srv = S
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(40):
        reqs.append(dict(id=100 + i, arrival=0.0, dim=2, degree=3,
                         n_refinements=int(rng.integers(2, 5)), rho=float(rng.uniform(0.5, 2)),
                         c=float(rng.uniform(0.5, 2)), modes=[int(rng.integers(1, 4)), 1],
                         final_time=float(rng.uniform(0.05, 0.3)), cfl=0.2))
    mix, _ = srv.run_server(ctx, reqs, lambda _q: X.FP64, hbm_budget_bytes=2 << 20,
                            realtime=False)
    solo, _ = srv.run_one_at_a_time(ctx, reqs, X.FP64)
    for q in reqs:
        assert _key(mix[q["id"]]) == _key(solo[q["id"]]), q


def test_turnover_pipelined_pools(S, ctx):
    """Turnover through the pools swept by the pipelined kernel (cell lists,
    per-cell scales, flux coefficients filled at admission for solves that
    start anywhere in the pool): 2D k=6, 3D k=3 and 3D k=4 solves of mixed
    size and length under a budget that forces reuse of freed cells; every
    result is bit-identical to the solve alone."""
    X, srv = S
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(24):
        kind = i % 3
        dim, k = [(2, 6), (3, 3), (3, 4)][kind]
        nref = int(rng.integers(2, 4)) if dim == 2 else int(rng.integers(1, 3))
        modes = [int(rng.integers(1, 4)) for _ in range(dim)]
        reqs.append(dict(id=200 + i, arrival=0.0, dim=dim, degree=k, n_refinements=nref,
                         rho=float(rng.uniform(0.5, 2)), c=float(rng.uniform(0.5, 2)),
                         modes=modes, final_time=float(rng.uniform(0.03, 0.12)), cfl=0.2))
    mix, _ = srv.run_server(ctx, reqs, lambda _q: X.FP64, hbm_budget_bytes=3 << 20,
                            realtime=False)
    solo, _ = srv.run_one_at_a_time(ctx, reqs, X.FP64)
    for q in reqs:
        assert _key(mix[q["id"]]) == _key(solo[q["id"]]), q


def test_footprint_matches_pool_size(S, ctx):
    """The size check at submission uses the pool's own per-cell bytes: a
    request whose pool cells (with the pipelined kernel's flux coefficients)
    exceed the budget is rejected at submit instead of waiting forever; one
    that fits is served."""
    X, srv = S
    # 2D k=6, 16x16 cells: 3940 B per pool cell with coefficients (1,008,640 B),
    # 3652 B without (934,912 B)
    q = dict(id=0, arrival=0.0, dim=2, degree=6, n_refinements=4, rho=1.0, c=1.0, modes=[1, 1],
             final_time=0.02, cfl=0.2)
    s = srv.Server(ctx, hbm_budget_bytes=1_000_000)
    with pytest.raises(X.ExwError):
        s.submit(srv.make_request(q, X.FP64))
    s.close()
    # 2D k=1 (no coefficients: the one-cell kernel) with 64x64 cells fits 2 MiB
    q2 = dict(id=1, arrival=0.0, dim=2, degree=1, n_refinements=6, rho=1.0, c=1.0, modes=[1, 1],
              final_time=0.02, cfl=0.2)
    res, _ = srv.run_server(ctx, [q2], lambda _q: X.FP64, hbm_budget_bytes=2 << 20,
                            realtime=False, stall_timeout=60.0)
    solo, _ = srv.run_one_at_a_time(ctx, [q2], X.FP64)
    assert _key(res[1]) == _key(solo[1])
