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
    return p


FIXTURES = ["lsrk45r2_2d_cartesian", "clrk4_3d", "ader_2d_norecon_ref2", "ader_2d_recon_ref2",
            "ader_2d_recon_ref3", "ader_3d_norecon"]


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_fixture(fixtures, name):
    """Every 'error' line (error p, error v, solution mag p) at 4 printed digits,
    the number of steps and the logged dt (source/explicit_wave.cc:509-519)."""
    fx = fixtures[name]
    prm = fx["params"]
    steps, rows, _ = O.run(_problem(prm))
    assert steps == fx["n_steps"]
    # the unrounded CFL step printed after make_dofs (explicit_wave.cc:695)
    k = prm["fe_degree"]
    m = O.Mesh(prm["dimension"], prm["n_initial_intervals"], prm["n_refinements"],
               O.TRANSFORM_REFERENCE_SINE, prm["grid_transform_factor"])
    assert m.n_cells == prm["n_cells"] and m.n_dofs(k) == prm["n_dofs"]
    dt = prm["cfl_number"] / k ** 1.5 * m.min_vertex_distance_over_c
    assert "%g" % dt == fx["time_step_sizes"][-1]
    assert len(rows) == len(fx["error_lines"])
    for (st, t, ep, ev, mp), ref in zip(rows, fx["error_lines"]):
        assert "%.2f" % t == ref[0]
        got = ["%.4e" % ep, "%.4e" % ev, "%.4e" % mp]
        want = [ref[1], ref[3], ref[4]]
        if name == "ader_3d_norecon" and ref[0] == "1.00":
            # this k=7, CFL 0.2 ADER run carries a weakly unstable mode that is
            # seeded by round-off after t~0.7 (two builds of the oracle that
            # differ only in FMA contraction differ by 3e-5 here, <1e-10 at
            # t<=0.7), so the last line is compared to 1e-4 relative
            for g, w in zip([ep, ev, mp], want):
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
                                         (O.LSRK33R2, 3), (O.LSRK59R2, 5)])
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
                                                (O.LSRK59R2, 9, 5)])
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
