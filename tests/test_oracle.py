"""CPU tests that pin the oracle (oracle/exwave_oracle.c) before it is used as
the checker of the HIP path:
  * the reference's own regression fixtures (tests/*.output, 'error' lines),
  * an independent dense brute-force assembly (tests/brute_force.py),
  * closed forms: Gauss rule, polynomial exactness, energy dissipation,
  * time-integrator convergence orders against the matrix exponential.
"""
import numpy as np
import pytest

import oracle as O
from tests import brute_force as BF


def _problem(params, transform=O.TRANSFORM_REFERENCE_SINE):
    p = O.Problem()
    p.dim = params["dimension"]
    p.degree = params["fe_degree"]
    p.n_intervals = params["n_initial_intervals"]
    p.n_refinements = params["n_refinements"]
    p.transform = transform
    p.transform_factor = params["grid_transform_factor"]
    p.boundary_id = params.get("boundary_id", 2)
    p.integrator = O.INTEGRATORS[params["time_integrator"]]
    p.cfl_number = params["cfl_number"]
    p.max_time_steps = params["max_time_steps"]
    p.final_time = params["final_time"]
    p.output_every_time = params["output_every_time"]
    p.ader_recon = 1 if params.get("use_ader_post", False) else 0
    p.exact = O.Exact.reference(p.dim, params["membrane_modes"])
    p.lts_max_clusters = params.get("max_n_clusters", 10)
    p.lts_max_diff = params.get("max_diff_clusters", 7)
    return p


FIXTURES = ["lsrk45r2_2d_cartesian", "clrk4_3d", "ader_2d_norecon_ref2", "ader_2d_recon_ref2",
            "ader_2d_recon_ref3", "ader_3d_norecon", "aderlts_2d_recon", "aderadconfull_2d"]


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_fixture(fixtures, name):
    """Every 'error' line (error p, error p post, error v, solution mag p) at 4
    printed digits, the number of steps and the logged dt
    (source/explicit_wave.cc:509-519)."""
    fx = fixtures[name]
    prm = fx["params"]
    problem = _problem(prm)
    problem.post_pressure = 1
    steps, rows, _ = O.run(problem)
    assert steps == fx["n_steps"]
    # the unrounded CFL step printed after make_dofs (explicit_wave.cc:695)
    k = prm["fe_degree"]
    m = O.Mesh(prm["dimension"], prm["n_initial_intervals"], prm["n_refinements"],
               O.TRANSFORM_REFERENCE_SINE, prm["grid_transform_factor"])
    assert m.n_cells == prm["n_cells"] and m.n_dofs(k) == prm["n_dofs"]
    dt = prm["cfl_number"] / k ** 1.5 * m.min_vertex_distance_over_c
    if "lts" in fx:  # LTS prints the slowest cluster's step; the CFL step is cluster 0's
        assert "%g" % dt == fx["lts"]["clusters"][0][2]
    else:
        assert "%g" % dt == fx["time_step_sizes"][-1]
    assert len(rows) == len(fx["error_lines"])
    for (st, t, ep, ev, mp, epp), ref in zip(rows, fx["error_lines"]):
        assert "%.2f" % t == ref[0]
        got = ["%.4e" % ep, "%.4e" % epp, "%.4e" % ev, "%.4e" % mp]
        want = ref[1:5]
        if name == "ader_3d_norecon" and ref[0] == "1.00":
            # this k=7, CFL 0.2 ADER run carries a weakly unstable mode that is
            # seeded by round-off after t~0.7 (two builds of the oracle that
            # differ only in FMA contraction differ by 3e-5 here, <1e-10 at
            # t<=0.7), so the last line is compared to 1e-4 relative
            for g, w in zip([ep, epp, ev, mp], want):
                assert abs(g - float(w)) <= 1e-4 * float(w) + 0.5e-4 * 10 ** np.floor(
                    np.log10(float(w)))
        else:
            assert got == want, (t, got, want)


def test_gauss_rule():
    for n in range(1, 14):
        x, w = O.gauss(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.allclose(x, 0.5 * (xr + 1), atol=1e-15, rtol=0)
        assert np.allclose(w, 0.5 * wr, atol=1e-15, rtol=0)
        # exact for degree 2n-1
        for p in range(2 * n):
            assert abs(np.dot(w, x ** p) - 1.0 / (p + 1)) < 1e-14


def _random_fields(m, k, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(m.n_dofs(k))


@pytest.mark.parametrize("dim,nint,k,bid", [(2, 3, 2, 2), (2, 2, 3, 1), (2, 2, 2, 3),
                                           (3, 2, 2, 2), (3, 2, 1, 3)])
def test_brute_force_dense(dim, nint, k, bid):
    """oracle apply == dense face-centric assembly in a Gauss-Lobatto basis on a
    deformed mesh with a different rho, c per cell (material interfaces)."""
    m = O.Mesh(dim, nint, 0, O.TRANSFORM_REFERENCE_SINE, 0.12, bid)
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.5, 2.0, m.n_cells)
    c = rng.uniform(0.5, 2.0, m.n_cells)
    m.set_material(rho, c)
    u = _random_fields(m, k, 3)
    got = m.apply(k, u)
    want = BF.apply(m.vertices, m.neighbor, rho, c, k, u)
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))
    got_m = m.apply(k, u, -1.0)
    assert np.array_equal(got_m, -got)


def test_polynomial_exactness():
    """On an affine (Cartesian) mesh a continuous polynomial field of degree <= k
    with v.n = 0 on the boundary makes every face term vanish for the soft-wall
    boundary (id 1), so apply is the strong derivative (-grad p/rho, -rho c^2 div v)."""
    k, rho, c = 3, 1.7, 0.8
    for dim in (2, 3):
        m = O.Mesh(dim, 3, 0, O.TRANSFORM_NONE, 0.0, 1)
        m.set_material(np.full(m.n_cells, rho), np.full(m.n_cells, c))
        x, _ = O.gauss(k + 1)
        nd = (k + 1) ** dim
        V = m.vertices
        u = np.zeros((m.n_cells, dim + 1, nd))
        want = np.zeros_like(u)
        for cc in range(m.n_cells):
            lo = V[cc, 0]
            hi = V[cc, -1]
            for q in range(nd):
                idx = [(q // (k + 1) ** d) % (k + 1) for d in range(dim)]
                X = [lo[d] + (hi[d] - lo[d]) * x[idx[d]] for d in range(dim)]
                # p = (1+x+y^2)(1 + z^3)..., v_d = x_d(1-x_d) * (other coords)
                others = lambda d: np.prod([1.0 + 0.5 * X[e] for e in range(dim) if e != d])
                p = 1.0 + X[0] + X[1] ** 2 + (X[2] ** 3 if dim == 3 else 0.0)
                gp = [1.0, 2 * X[1]] + ([3 * X[2] ** 2] if dim == 3 else [])
                div = 0.0
                for d in range(dim):
                    u[cc, d, q] = X[d] * (1 - X[d]) * others(d)
                    div += (1 - 2 * X[d]) * others(d)
                    want[cc, d, q] = -gp[d] / rho
                u[cc, dim, q] = p
                want[cc, dim, q] = -rho * c * c * div
        got = m.apply(k, u.reshape(-1))
        assert np.max(np.abs(got - want.reshape(-1))) < 1e-10


@pytest.mark.parametrize("bid", [1, 2, 3])
def test_energy_dissipation(bid):
    """u^T W M (M^-1 A u) <= 0 for W = (rho on v, 1/(rho c^2) on p): the upwind
    flux dissipates energy; a sign error in any flux term makes it positive."""
    dim, k = 2, 3
    m = O.Mesh(dim, 4, 0, O.TRANSFORM_REFERENCE_SINE, 0.1, bid)
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.5, 2.0, m.n_cells)
    c = rng.uniform(0.5, 2.0, m.n_cells)
    m.set_material(rho, c)
    for seed in range(5):
        u = _random_fields(m, k, seed)
        du = m.apply(k, u)
        # energy rate = d/dt E = sum_K int rho v.dv + p dp /(rho c^2) = (E(u+eps du)-E(u-eps du))/(4 eps)
        eps = 1e-6
        rate = (m.energy(k, u + eps * du) - m.energy(k, u - eps * du)) / (2 * eps)
        assert rate < 0.0


def _dense_operator(m, k):
    N = m.n_dofs(k)
    L = np.zeros((N, N))
    e = np.zeros(N)
    for i in range(N):
        e[:] = 0
        e[i] = 1
        L[:, i] = m.apply(k, e)
    return L


@pytest.mark.parametrize("integ,order", [(O.EXPL_EULER, 1), (O.CLRK4, 4), (O.LSRK45R2, 4),
                                         (O.LSRK33R2, 3), (O.LSRK59R2, 5), (O.LSRK45R3, 4),
                                         (O.SSPRK, 4), (O.ADER_ADCONFULL, 3)])
def test_integrator_order(integ, order):
    """One step of each integrator vs exp(dt L): the local error scales like
    dt^(order+1) (a wrong tableau entry breaks an order condition)."""
    from scipy.linalg import expm
    dim, k = 2, 2
    m = O.Mesh(dim, 2, 0, O.TRANSFORM_REFERENCE_SINE, 0.1, 2)
    L = _dense_operator(m, k)
    u0 = _random_fields(m, k, 5)
    errs = []
    dts = [0.02, 0.01]
    for dt in dts:
        vn = u0.copy()
        vnp = np.zeros_like(u0)
        m.time_step(k, integ, dt, vn, vnp)
        errs.append(np.linalg.norm(vnp - expm(dt * L) @ u0))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - (order + 1)) < 0.35, (errs, rate)


def test_ader_ck_taylor():
    """With exact spatial derivatives the CK time integral of a globally smooth
    polynomial-in-space field whose derivatives stay representable equals the
    Taylor time integral: check on a degree-k field where L^j u stays degree k-?"""
    # u(t) for the operator restricted to one cell with t1=t2 gives zero
    dim, k = 2, 4
    m = O.Mesh(dim, 2, 0, O.TRANSFORM_NONE, 0.0, 2)
    u = _random_fields(m, k, 9)
    z = m.ader_ck(k, u, None, 0.3, 0.3, 0.1)
    assert np.max(np.abs(z)) == 0.0
    # linearity in u and additivity in time: int_0^a + int_a^b = int_0^b
    a, b = 0.004, 0.009
    i1 = m.ader_ck(k, u, None, a, 0.0, 0.0)
    i2 = m.ader_ck(k, u, None, b, a, 0.0)
    i3 = m.ader_ck(k, u, None, b, 0.0, 0.0)
    assert np.max(np.abs(i1 + i2 - i3)) < 1e-15 * np.max(np.abs(i3)) * 100


@pytest.mark.parametrize("integ,stages,order", [(O.EXPL_EULER, 1, 1), (O.CLRK4, 4, 4),
                                                (O.LSRK45R2, 5, 4), (O.LSRK33R2, 3, 3),
                                                (O.LSRK59R2, 9, 5), (O.LSRK45R3, 5, 4),
                                                (O.SSPRK, 8, 4), (O.ADER_ADCONFULL, 3, 3)])
def test_stability_polynomial(integ, stages, order):
    """For u' = L u one step is R(dt L) u with R of degree `stages`; recover the
    coefficients gamma_j by a Vandermonde fit over several dt and require
    gamma_j = 1/j! for j <= order (the linear order conditions)."""
    from math import factorial
    dim, k = 2, 2
    m = O.Mesh(dim, 2, 0, O.TRANSFORM_REFERENCE_SINE, 0.1, 2)
    L = _dense_operator(m, k)
    h0 = 1.0 / np.max(np.abs(np.linalg.eigvals(L)))
    u0 = _random_fields(m, k, 5)
    S = stages
    tau = 0.6 + 0.4 * np.cos(np.pi * (np.arange(S + 1) + 0.5) / (S + 1))  # in [0.2, 1]
    Y = []
    for t in tau:
        vn = u0.copy()
        vnp = np.zeros_like(u0)
        m.time_step(k, integ, t * h0, vn, vnp)
        Y.append(vnp)
    V = np.vander(tau, S + 1, increasing=True)
    X = np.linalg.solve(V, np.array(Y))
    w = u0.copy()
    gam = []
    for j in range(S + 1):
        gam.append(X[j] @ w / (w @ w))
        w = h0 * (L @ w)
    for j in range(order + 1):
        assert abs(gam[j] - 1.0 / factorial(j)) < 1e-7, (j, gam)
    if integ == O.LSRK45R2:
        assert abs(gam[5] - 0.0048544) < 1e-6
    if integ == O.ADER_ADCONFULL:  # k = 2: exactly the cubic Taylor polynomial of exp
        assert abs(gam[3] - 1.0 / 6.0) < 1e-7


def test_ssprk84_shu_osher_rows():
    """Each stage of the Shu-Osher form is a combination of the earlier stages
    whose A row sums to one, so a step with dt = 0 returns u itself (a
    mistyped A entry breaks it; the B entries are pinned by the order tests)."""
    m = O.Mesh(2, 2, 0, O.TRANSFORM_NONE, 0.0, 2)
    k = 2
    u0 = _random_fields(m, k, 3)
    vn, vnp = u0.copy(), np.zeros_like(u0)
    m.time_step(k, O.SSPRK, 0.0, vn, vnp)
    assert np.max(np.abs(vnp - u0)) <= 1e-14 * np.max(np.abs(u0))


def lts_info(m, k, cfl_number, max_clusters, max_diff, recon=1):
    """Clusters, flags, schedule and steps of the oracle's ADER-LTS set-up."""
    import ctypes as C
    L = O.lib()
    cfl = cfl_number / k ** 1.5
    fastest = cfl * m.min_vertex_distance_over_c
    h = L.ox_lts_create(m._m, k, cfl, fastest, max_clusters, max_diff, 0.0, recon)
    try:
        nc, nu = L.ox_lts_n_clusters(h), L.ox_lts_n_updates(h)
        cl = np.zeros(m.n_cells, np.int32)
        fl = np.zeros(m.n_cells, np.int32)
        ip = C.POINTER(C.c_int)
        L.ox_lts_cells(h, cl.ctypes.data_as(ip), fl.ctypes.data_as(ip))
        order = np.zeros(nu, np.int32)
        times = np.zeros(2 * nu)
        L.ox_lts_schedule(h, order.ctypes.data_as(ip), O._p(times))
        return dict(n_clusters=nc, cluster_diff=L.ox_lts_cluster_diff(h),
                    iterations=L.ox_lts_iterations(h), global_dt=L.ox_lts_global_dt(h),
                    fastest=fastest, cluster=cl, flags=fl, order=order,
                    times=times.reshape(-1, 2))
    finally:
        L.ox_lts_free(h)


def test_lts_clusters_fixture(fixtures):
    """The reference's cluster report for aderlts_2d_recon: cluster count, the
    categorisation iterations, the per-cluster time steps (multiples 1, 2, 3 of
    the fastest step) and the cell counts, which the run prints in cell batches
    of VectorizedArray width 2 formed per category (cluster, faster/slower
    neighbour flags), include/cluster_manager.templates.h:252, 402-416."""
    fx = fixtures["aderlts_2d_recon"]
    prm, lts = fx["params"], fx["lts"]
    k = prm["fe_degree"]
    m = O.Mesh(2, prm["n_initial_intervals"], 0, O.TRANSFORM_REFERENCE_SINE,
               prm["grid_transform_factor"])
    info = lts_info(m, k, prm["cfl_number"], prm["max_n_clusters"], prm["max_diff_clusters"])
    assert info["n_clusters"] == lts["n_clusters"]
    assert info["cluster_diff"] == lts["cluster_diff"]
    assert info["iterations"] == lts["iterations"]
    nv = prm["n_vect"]
    for c, batches, dt in lts["clusters"]:
        cats = [int(np.sum((info["cluster"] == c) & (info["flags"] == f))) for f in range(3)]
        assert sum((x + nv - 1) // nv for x in cats) == batches, (c, cats, batches)
        assert "%g" % ((lts["cluster_diff"] * c + 1) * info["fastest"]) == dt
    assert "%g" % info["global_dt"] == fx["time_step_sizes"][-1]


def test_lts_invariants():
    """Clustering invariants (include/cluster_manager.templates.h:100-240):
    face neighbours differ by at most one cluster, no cell has both a faster
    and a slower neighbour, flags match the neighbours; every cluster reaches
    the global step exactly once its updates are summed, and each update of a
    cluster starts where its previous one ended."""
    for nint, f, k in [(15, 0.2, 5), (12, 0.25, 3), (9, 0.3, 2)]:
        m = O.Mesh(2, nint, 0, O.TRANSFORM_REFERENCE_SINE, f)
        info = lts_info(m, k, 0.2, 5, 2)
        cl, fl, nb = info["cluster"], info["flags"], m.neighbor
        for c in range(m.n_cells):
            nbs = [x for x in nb[c] if x >= 0]
            assert all(abs(cl[x] - cl[c]) <= 1 for x in nbs)
            fa = any(cl[x] < cl[c] for x in nbs)
            sl = any(cl[x] > cl[c] for x in nbs)
            assert not (fa and sl)
            assert fl[c] == int(sl) + 2 * int(fa)
        D, nc, diff = info["fastest"], info["n_clusters"], info["cluster_diff"]
        level = np.zeros(nc)
        for u, c in enumerate(info["order"]):
            assert abs(info["times"][u, 0] - level[c] * D) < 1e-12 * D * 1e3
            level[c] = info["times"][u, 1] / D
        assert np.allclose(level, diff * (nc - 1) + 1)
        assert abs(info["global_dt"] - D * (diff * (nc - 1) + 1)) < 1e-15


def test_lts_single_cluster_is_ader():
    """With one cluster (Cartesian mesh: all cells alike) ADER-LTS is the
    global ADER step with the same dt."""
    m = O.Mesh(2, 4, 0, O.TRANSFORM_NONE, 0.0)
    k = 3
    info = lts_info(m, k, 0.2, 5, 1)
    assert info["n_clusters"] == 1
    e = O.Exact.make(2, [[1, 2]])
    u0 = m.project(k, e, 0.0)
    import ctypes as C
    L = O.lib()
    cfl = 0.2 / k ** 1.5
    dt = cfl * m.min_vertex_distance_over_c
    for recon in (0, 1):
        h = L.ox_lts_create(m._m, k, cfl, dt, 5, 1, 0.0, recon)
        out = np.zeros_like(u0)
        L.ox_lts_step(h, dt, O._p(u0.copy()), O._p(out))
        L.ox_lts_free(h)
        ref = u0 - m.apply_ader(k, u0, dt, recon)
        assert np.max(np.abs(out - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_lts_accuracy_two_clusters():
    """On a deformed mesh with several clusters one LTS step (every cluster
    advanced by its own multiple of the fastest step, fluxes between clusters
    split in time through the flux memory) is at least as accurate against
    the exact solution as one global ADER step of the same total length."""
    m = O.Mesh(2, 6, 0, O.TRANSFORM_REFERENCE_SINE, 0.25, 2)
    k = 3
    info = lts_info(m, k, 0.2, 5, 1)
    assert info["n_clusters"] >= 2
    e = O.Exact.make(2, [[1, 1]])
    u0 = m.project(k, e, 0.0)
    import ctypes as C
    L = O.lib()
    cfl = 0.2 / k ** 1.5
    dt = cfl * m.min_vertex_distance_over_c
    h = L.ox_lts_create(m._m, k, cfl, dt, 5, 1, 0.0, 0)
    gdt = L.ox_lts_global_dt(h)
    out = np.zeros_like(u0)
    L.ox_lts_step(h, gdt, O._p(u0.copy()), O._p(out))
    L.ox_lts_free(h)
    # one global ADER step with the slowest step is a different scheme with
    # the same order: both approximate the exact solution at t = gdt
    ref = u0 - m.apply_ader(k, u0, gdt, 0)
    err_lts = m.l2_errors(k, e, gdt, out)
    err_glob = m.l2_errors(k, e, gdt, ref)
    assert err_lts[0] < 2 * err_glob[0] + 1e-9 and err_lts[1] < 2 * err_glob[1] + 1e-9


def test_gauss_lobatto():
    """Interior points are the roots of P'_{n-1} mapped to [0,1]."""
    for n in range(2, 12):
        x = O.gauss_lobatto(n)
        inner = np.sort(np.polynomial.legendre.Legendre.basis(n - 1).deriv().roots().real)
        want = np.concatenate([[0.0], 0.5 * (inner + 1.0), [1.0]])
        assert np.allclose(x, want, atol=1e-14, rtol=0), n


def _cell_points(m, pts1d):
    """Physical coordinates of the tensor points pts1d^dim (xi_0 fastest) of every
    cell of an axis-aligned mesh: [cell][point][d]."""
    V = m.vertices
    lo, hi = V[:, 0, :], V[:, -1, :]
    G = np.meshgrid(*([pts1d] * m.dim), indexing="ij")  # axis j is xi_{dim-1-j}
    xi = np.stack([G[m.dim - 1 - d].ravel() for d in range(m.dim)], axis=-1)
    return lo[:, None, :] + (hi - lo)[:, None, :] * xi[None, :, :]


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 3), (3, 2)])
def test_post_pressure_reproduces_polynomials(dim, k):
    """For a continuous pressure p in P_k and zero velocity the operator gives
    rho v_t = -grad p exactly on cells away from the boundary (no jumps, Gauss
    collocation exact), so the local Neumann problem returns p* = p there: the
    constraint, the sign and the rho factor of the right-hand side and the
    average adjustment all have to be right. Cell-wise random rho, speed."""
    rng = np.random.default_rng(dim * 10 + k)
    m = O.Mesh(dim, 5, 0, O.TRANSFORM_NONE, 0.0)
    m.set_material(rng.uniform(0.5, 3.0, m.n_cells), rng.uniform(0.5, 2.0, m.n_cells))
    coef = rng.standard_normal((k + 1,) * dim)

    def p(x):  # total degree <= k
        out = np.zeros(x.shape[:-1])
        for idx in np.ndindex(*coef.shape):
            if sum(idx) <= k:
                out += coef[idx] * np.prod([x[..., d] ** idx[d] for d in range(dim)], axis=0)
        return out

    xg, _ = O.gauss(k + 1)
    nd = (k + 1) ** dim
    u = np.zeros((m.n_cells, dim + 1, nd))
    u[:, dim, :] = p(_cell_points(m, xg))
    post = m.post_pressure(k, u.ravel()).reshape(m.n_cells, -1)
    want = p(_cell_points(m, O.gauss_lobatto(k + 2)))
    interior = (m.neighbor >= 0).all(axis=1)
    assert interior.sum() > 0
    err = np.abs(post[interior] - want[interior]).max() / np.abs(want).max()
    assert err < 1e-11, err


def test_post_pressure_keeps_cell_average():
    """On every cell the QGauss(k+2) average of p* equals that of p_h
    (source/wave_equation_operations.cc:823-841), on a deformed mesh."""
    k, dim = 3, 2
    m = O.Mesh(dim, 4, 0, O.TRANSFORM_REFERENCE_SINE, 0.2)
    e = O.Exact.reference(dim, 2)
    u = m.project(k, e, 0.3)
    post = m.post_pressure(k, u).reshape(m.n_cells, -1)
    # average of p_h with the same rule: Gauss(k+1) is exact for its integral
    # against the bilinear Jacobian only on affine cells, so compare the
    # average of p* from GL nodes interpolated to Gauss(k+2) with p_h's.
    xq, wq = O.gauss(k + 2)
    xg, _ = O.gauss(k + 1)
    xl = O.gauss_lobatto(k + 2)

    def lag(nodes, x):
        A = np.zeros((len(x), len(nodes)))
        for q, xx in enumerate(x):
            for i in range(len(nodes)):
                A[q, i] = np.prod([(xx - nodes[j]) / (nodes[i] - nodes[j])
                                   for j in range(len(nodes)) if j != i])
        return A
    Ig, Il = lag(xg, xq), lag(xl, xq)
    V = m.vertices
    uu = u.reshape(m.n_cells, dim + 1, -1)
    for c in range(m.n_cells):
        ph = Ig @ uu[c, dim].reshape(k + 1, k + 1) @ Ig.T      # [q1][q0]
        ps = Il @ post[c].reshape(k + 2, k + 2) @ Il.T
        # bilinear Jacobian determinant at (xi0 = xq[j], xi1 = xq[i])
        a, b, cc, d = V[c]
        det = np.zeros((k + 2, k + 2))
        for i, y in enumerate(xq):
            for j, x in enumerate(xq):
                dx0 = (b - a) * (1 - y) + (d - cc) * y
                dx1 = (cc - a) * (1 - x) + (d - b) * x
                det[i, j] = abs(dx0[0] * dx1[1] - dx0[1] * dx1[0])
        W = np.outer(wq, wq) * det
        assert abs((W * ph).sum() - (W * ps).sum()) <= 1e-13 * max(1.0, abs(W * ph).sum())


def test_estimate_error_closed_form():
    """p = a.x, v = 0: on every cell without a boundary face the operator
    gives v_t = -a / rho, so the indicator is |a| (1 + 1/rho) sqrt(|K|)
    (the reference's literal grad p - v_t), with random rho per cell on a box
    mesh."""
    dim, k = 3, 2
    m = O.Mesh(dim, 4, 0, O.TRANSFORM_NONE, 0.0, 2)
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.5, 2.0, m.n_cells)
    m.set_material(rho, rng.uniform(0.5, 2.0, m.n_cells))
    a = np.array([0.3, -1.2, 0.7])
    x, _ = O.gauss(k + 1)
    V = m.vertices
    lo, hi = V[:, 0, :], V[:, -1, :]
    G = np.meshgrid(x, x, x, indexing="ij")
    xi = np.stack([G[2].ravel(), G[1].ravel(), G[0].ravel()], axis=-1)
    pts = lo[:, None, :] + (hi - lo)[:, None, :] * xi[None]
    u = np.zeros((m.n_cells, dim + 1, (k + 1) ** dim))
    u[:, dim] = pts @ a
    err = m.estimate_error(k, u.ravel())
    vol = np.prod(hi - lo, axis=1)
    want = np.linalg.norm(a) * (1 + 1 / rho) * np.sqrt(vol)
    interior = (m.neighbor >= 0).all(axis=1)
    assert interior.sum() > 0
    assert np.allclose(err[interior], want[interior], rtol=1e-12, atol=0)


@pytest.mark.parametrize("dim,k,tr", [(2, 1, 0), (2, 3, 1), (2, 5, 1), (3, 2, 1), (3, 3, 0)])
def test_energy_of_constant_fields(dim, k, tr):
    """The discrete energy 1/2 sum_K int_K (rho |v|^2 + p^2 / (rho c^2)) of fields that are
    constant on every cell is a closed form: with one material over the unit cube (deformed
    cells tile it exactly, and Gauss quadrature integrates det J of a multilinear map
    exactly) 1/2 (rho |v|^2 + p^2 / (rho c^2)); with a material per box cell the sum of
    that over the cells times their volume 1/n^dim.  A dropped det J, a quadrature weight
    off, rho in place of 1/rho or a missing 1/2 fails it."""
    n = 3
    transform = O.TRANSFORM_NONE if tr == 0 else O.TRANSFORM_REFERENCE_SINE
    m = O.Mesh(dim, n, 0, transform, 0.0 if tr == 0 else 0.1, 2)
    nc, nd, nf = m.n_cells, (k + 1) ** dim, dim + 1
    rng = np.random.default_rng(40 + dim * 10 + k)
    v = rng.normal(size=dim)
    p = 0.8
    u = np.empty((nc, nf, nd))
    u[:, :dim, :] = v[None, :, None]
    u[:, dim, :] = p
    rho0, c0 = 1.7, 0.6
    m.set_material(np.full(nc, rho0), np.full(nc, c0))
    want = 0.5 * (rho0 * v @ v + p * p / (rho0 * c0 * c0))
    got = m.energy(k, u.reshape(-1))
    assert abs(got - want) <= 1e-13 * want
    if tr == 0:
        rho = rng.uniform(0.5, 2.0, nc)
        c = rng.uniform(0.5, 2.0, nc)
        m.set_material(rho, c)
        want = np.sum(0.5 * (rho * (v @ v) + p * p / (rho * c * c))) / n ** dim
        got = m.energy(k, u.reshape(-1))
        assert abs(got - want) <= 1e-13 * want
