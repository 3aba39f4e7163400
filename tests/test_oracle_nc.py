"""CPU tests that pin the oracle's non-conforming (2:1 hanging) faces before it
checks the HIP path on such meshes.  The reference gets these faces from
deal.II (p4est refinement, MatrixFree subface evaluation); its own tests have
no locally refined mesh, so the pins are mathematical:
  * continuous polynomial fields of degree k are reproduced exactly across
    hanging faces (no spurious jumps: the subface mapping and weights);
  * the energy-weighted operator is negative semi-definite (both sides of a
    subface integrate the same flux with the same weights and normals);
  * a smooth solution converges at order k+1 on refined-half meshes;
  * the mesh generator's records are consistent with its geometry.
"""
import numpy as np
import pytest

import oracle as O


def _half_refined(dim, n, transform=O.TRANSFORM_NONE, factor=0.0, bid=2):
    ref = np.zeros((n,) * dim, dtype=np.int32)
    ref[..., : n // 2] = 1  # axis -1 is x (lexicographic, xi_0 fastest)
    return O.Mesh.box_refined(dim, n, ref, transform, factor, bid)


def _checker_refined(dim, n, seed):
    ref = (np.random.default_rng(seed).uniform(size=(n,) * dim) < 0.4).astype(np.int32)
    return O.Mesh.box_refined(dim, n, ref)


def _node_points(m, k):
    """Physical coordinates of every cell's Gauss nodes (affine box cells):
    [cell][node][d], node lexicographic with xi_0 fastest."""
    x, _ = O.gauss(k + 1)
    V = m.vertices
    lo, hi = V[:, 0, :], V[:, -1, :]
    G = np.meshgrid(*([x] * m.dim), indexing="ij")
    xi = np.stack([G[m.dim - 1 - d].ravel() for d in range(m.dim)], axis=-1)
    return lo[:, None, :] + (hi - lo)[:, None, :] * xi[None, :, :]


def test_mesh_records_match_geometry():
    """Every fine face of a hanging record is the right quarter (half in 2D)
    of the coarse face, and every conforming neighbour pair shares its face."""
    for dim in (2, 3):
        m = _checker_refined(dim, 4, dim)
        V = m.vertices
        nb = m.neighbor
        nv = 1 << dim

        def face_vertices(c, f):
            d, s = f >> 1, f & 1
            return V[c][[v for v in range(nv) if ((v >> d) & 1) == s]]

        for c in range(m.n_cells):
            for f in range(2 * dim):
                if nb[c, f] >= 0:
                    assert np.allclose(face_vertices(c, f), face_vertices(nb[c, f], f ^ 1))
        H = m.hanging
        assert len(H) > 0
        for rec in H:
            c, f = rec[0], rec[1]
            d = f >> 1
            fc = face_vertices(c, f)
            lo, hi = fc.min(axis=0), fc.max(axis=0)
            tang = [e for e in range(dim) if e != d]
            for b, fine in enumerate(rec[2:]):
                assert nb[fine, f ^ 1] == O.NC_FACE and nb[c, f] == O.NC_FACE
                ff = face_vertices(fine, f ^ 1)
                for j, e in enumerate(tang):
                    half = (b >> j) & 1
                    mid = 0.5 * (lo[e] + hi[e])
                    want = (mid, hi[e]) if half else (lo[e], mid)
                    assert np.isclose(ff[:, e].min(), want[0]) and np.isclose(ff[:, e].max(),
                                                                                  want[1])
                assert np.allclose(ff[:, d], lo[d])


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 3), (3, 2)])
def test_polynomials_pass_hanging_faces(dim, k):
    """p and v of total degree <= k, continuous: on every cell away from the
    boundary (including both sides of hanging faces) the operator returns the
    strong form v_t = -grad p / rho, p_t = -rho c^2 div v exactly, with
    cell-wise random rho and c."""
    rng = np.random.default_rng(10 * dim + k)
    m = _checker_refined(dim, 4, 7 + dim)
    assert len(m.hanging) > 0
    rho = rng.uniform(0.5, 3.0, m.n_cells)
    c = rng.uniform(0.5, 2.0, m.n_cells)
    m.set_material(rho, c)
    monos = [idx for idx in np.ndindex(*((k + 1,) * dim)) if sum(idx) <= k]
    cp = rng.standard_normal(len(monos))
    cv = rng.standard_normal((dim, len(monos)))

    def poly(coef, x):
        return sum(a * np.prod([x[..., d] ** e[d] for d in range(dim)], axis=0)
                   for a, e in zip(coef, monos))

    def dpoly(coef, x, d):
        out = np.zeros(x.shape[:-1])
        for a, e in zip(coef, monos):
            if e[d] == 0:
                continue
            t = a * e[d] * x[..., d] ** (e[d] - 1)
            for dd in range(dim):
                if dd != d:
                    t = t * x[..., dd] ** e[dd]
            out += t
        return out

    X = _node_points(m, k)
    nd = (k + 1) ** dim
    u = np.zeros((m.n_cells, dim + 1, nd))
    for d in range(dim):
        u[:, d] = poly(cv[d], X)
    u[:, dim] = poly(cp, X)
    du = m.apply(k, u.ravel()).reshape(m.n_cells, dim + 1, nd)
    want = np.zeros_like(du)
    for d in range(dim):
        want[:, d] = -dpoly(cp, X, d) / rho[:, None]
    div = sum(dpoly(cv[d], X, d) for d in range(dim))
    want[:, dim] = -(rho * c * c)[:, None] * div
    # cells without a boundary face (hanging faces count as interior)
    interior = ((m.neighbor >= 0) | (m.neighbor == O.NC_FACE)).all(axis=1)
    nc_cells = (m.neighbor == O.NC_FACE).any(axis=1) & interior
    assert nc_cells.sum() > 0
    scale = np.abs(want[interior]).max()
    err = np.abs(du[interior] - want[interior]).max() / scale
    assert err < 1e-11, err


@pytest.mark.parametrize("dim,k,bid", [(2, 2, 1), (2, 3, 3), (3, 1, 2)])
def test_energy_operator_is_dissipative(dim, k, bid):
    """The symmetric part of W A (W: the energy weights rho jxw on v,
    jxw/(rho c^2) on p) is negative semi-definite on a mesh with hanging faces
    and random materials: the coarse and fine sides of every subface integrate
    the same upwind flux with the same points, weights and normals."""
    rng = np.random.default_rng(3 * dim + k)
    m = _checker_refined(dim, 3 if dim == 3 else 4, 11 + dim)
    if bid != 2:
        m = O.Mesh.box_refined(dim, 3 if dim == 3 else 4,
                               (rng.uniform(size=(3 if dim == 3 else 4,) * dim) < 0.4)
                               .astype(np.int32), 0, 0.0, bid)
    assert len(m.hanging) > 0
    m.set_material(rng.uniform(0.5, 2.0, m.n_cells), rng.uniform(0.5, 2.0, m.n_cells))
    N = m.n_dofs(k)
    A = np.zeros((N, N))
    Wd = np.zeros(N)
    e = np.zeros(N)
    for i in range(N):
        e[:] = 0.0
        e[i] = 1.0
        A[:, i] = m.apply(k, e)
        Wd[i] = 2.0 * m.energy(k, e)
    S = Wd[:, None] * A
    S = 0.5 * (S + S.T)
    ev = np.linalg.eigvalsh(S)
    assert ev.max() <= 1e-11 * np.abs(ev).max(), (ev.max(), np.abs(ev).max())
    assert ev.min() < -1e-3 * np.abs(ev).max()  # the upwind flux does dissipate


def test_convergence_on_refined_half():
    """LSRK45 on the membrane mode (2,1) over a mesh whose x < 1/2 half is
    refined once: the L2 error of p falls by about 2^(k+1) per base refinement
    (k = 2), as on conforming meshes."""
    k, T = 2, 0.1
    e = O.Exact.reference(2, 2)
    errs = []
    for n in (4, 8, 16):
        m = _half_refined(2, n)
        dt0 = 0.1 / k ** 1.5 * m.min_vertex_distance_over_c
        nst = int(np.ceil(T / dt0))
        dt = T / nst
        u = m.project(k, e, 0.0)
        for _ in range(nst):
            un = np.zeros_like(u)
            m.time_step(k, O.LSRK45R2, dt, u, un)
            u = un
        errs.append(m.l2_errors(k, e, T, u)[0])
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > k + 1 - 0.4), (errs, rates)


def _unused_lts_rejects_hanging_faces():
    m = _half_refined(2, 4)
    h = O.lib().ox_lts_create(m._m, 2, 0.1, 0.01, 2, 1, 0.0, 0)
    assert not h
