"""CPU-only checks of the HIP library: it loads, exports every entry point
declared in include/exwave_amd.h, and its host-side mesh generator reproduces
the oracle's mesh (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "exwave_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(exw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import exwave_amd.capi as capi
    lib = capi.load()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes signature table covers all of them
    assert sorted(s[0] for s in capi.SIGNATURES) == names


@pytest.mark.parametrize("dim,nint,nref,tr,f", [(2, 5, 2, 1, 0.1), (3, 3, 1, 1, 0.05),
                                               (3, 4, 0, 2, 0.05), (2, 6, 0, 3, 0.05),
                                               (3, 2, 0, 0, 0.0)])
def test_mesh_generator_matches_oracle(dim, nint, nref, tr, f):
    import exwave_amd as X
    V, nb = X.mesh_hypercube(dim, nint, nref, tr, f, 2)
    m = O.Mesh(dim, nint, nref, tr, f, 2)
    assert np.array_equal(nb, m.neighbor)
    assert np.max(np.abs(V - m.vertices)) < 1e-15


@pytest.mark.parametrize("dim,n,tr,f,seed", [(2, 4, 0, 0.0, 1), (2, 5, 1, 0.1, 2), (3, 3, 0, 0.0, 3),
                                          (3, 4, 2, 0.05, 4)])
def test_refined_mesh_generator_matches_oracle(dim, n, tr, f, seed):
    """exw_mesh_box_refined (library host code) and the oracle's generator give
    the same cells, neighbours and hanging records."""
    import exwave_amd as X
    ref = (np.random.default_rng(seed).uniform(size=(n,) * dim) < 0.5).astype(np.int32)
    V, nb, H = X.mesh_box_refined(dim, n, ref, tr, f, 2)
    m = O.Mesh.box_refined(dim, n, ref, tr, f, 2)
    assert len(H) > 0
    assert np.array_equal(nb, m.neighbor)
    assert np.array_equal(H, m.hanging)
    assert np.max(np.abs(V - m.vertices)) < 1e-15


def test_mesh_generator_rejects_bad_input():
    import exwave_amd as X
    with pytest.raises(X.ExwError):
        X.mesh_hypercube(4, 2)
    with pytest.raises(X.ExwError):
        X.mesh_hypercube(2, 2, 0, 0, 0.0, 7)


@pytest.mark.parametrize("dim,nint,f,k,maxc,maxd", [(2, 15, 0.2, 5, 5, 1), (2, 12, 0.25, 3, 5, 2),
                                                    (2, 9, 0.3, 2, 3, 1), (3, 6, 0.2, 3, 5, 1),
                                                    (3, 5, 0.25, 2, 2, 3)])
def test_lts_categorize_matches_oracle(dim, nint, f, k, maxc, maxd):
    """The library's host-side cluster categorisation (exw_lts_categorize) gives
    the oracle's clusters and faster/slower flags."""
    import exwave_amd as X
    from tests.test_oracle import lts_info
    V, nb = X.mesh_hypercube(dim, nint, 0, X.TRANSFORM_REFERENCE_SINE, f, 2)
    cfl = 0.2 / k ** 1.5
    cl, fl, nc, diff, it = X.lts_categorize(dim, V, nb, cfl, maxc, maxd)
    m = O.Mesh(dim, nint, 0, O.TRANSFORM_REFERENCE_SINE, f, 2)
    info = lts_info(m, k, 0.2, maxc, maxd)
    assert nc == info["n_clusters"] and diff == info["cluster_diff"]
    assert it == info["iterations"]
    assert np.array_equal(cl, info["cluster"]) and np.array_equal(fl, info["flags"])


@pytest.mark.parametrize("dim,k,r,modes,amps", [(2, 3, 3, [[2, 1]], [1.0]), (3, 2, 2, [[1, 1, 2]], [1.0]),
                                                (2, 6, 4, [[3, 4]], [1.0]),
                                                (2, 4, 3, [[1, 1], [2, 3]], [1.0, -0.5]),
                                                (3, 5, 2, [[1, 2, 1], [2, 1, 3]], [0.7, 0.3])])
def test_request_initial_error_matches_oracle(dim, k, r, modes, amps):
    """The closed-form initial interpolation error and norm used by the
    automatic precision choice equal the oracle's L2 error of the projected
    initial field at t = 0 (QGauss(k+2))."""
    import ctypes as C
    import exwave_amd.capi as capi
    req = capi.Request()
    req.dim, req.degree, req.n_refinements = dim, k, r
    req.rho = req.c = 1.0
    req.modes = capi.Modes.make(dim, modes, amps)
    req.final_time, req.cfl_number, req.precision = 0.5, 0.2, capi.PRECISION_AUTO
    e0, n0, sc, pr = C.c_double(), C.c_double(), C.c_double(), C.c_int()
    capi.check(capi.load().exw_request_initial_error(C.byref(req), C.byref(e0), C.byref(n0),
                                                     C.byref(sc), C.byref(pr)))
    m = O.Mesh(dim, 1 << r, 0, O.TRANSFORM_NONE, 0.0, 2)
    e = O.Exact.make(dim, modes, amps)
    u = m.project(k, e, 0.0)
    err = m.l2_errors(k, e, 0.0, u)
    assert abs(e0.value - err[0]) <= 1e-9 * err[0] + 1e-15, (e0.value, err[0])
    assert abs(n0.value - err[2]) <= 1e-12 * err[2]
    assert pr.value == (capi.FP32 if sc.value >= 10.0 else capi.FP64)


@pytest.mark.parametrize("dim,k", [(2, 2), (3, 3)])
def test_write_vtu(tmp_path, dim, k):
    """exw_write_vtu: counts, point coordinates (the multilinear map at the
    equispaced points), solution and post-pressure values (the polynomials
    evaluated independently here), cell data, and sub-cells with positive
    orientation in VTK's corner order."""
    import xml.etree.ElementTree as ET
    import exwave_amd as X
    m = O.Mesh(dim, 2, 0, O.TRANSFORM_REFERENCE_SINE, 0.1)
    rng = np.random.default_rng(dim)
    u = rng.standard_normal(m.n_dofs(k))
    post = rng.standard_normal(m.n_cells * (k + 2) ** dim)
    est = rng.uniform(size=m.n_cells)
    fn = tmp_path / "out.vtu"
    X.write_vtu(fn, dim, k, m.vertices, u, error_estimate=est, post_pressure=post)
    root = ET.parse(fn).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    npt, ncl = int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))
    assert npt == m.n_cells * (k + 1) ** dim and ncl == m.n_cells * k ** dim

    def arr(path, name=None):
        for a in piece.iter("DataArray"):
            if name is None or a.get("Name") == name:
                return np.array(a.text.split(), dtype=float)
    pts = arr(None, "Points").reshape(npt, 3)
    conn = arr(None, "connectivity").astype(int).reshape(ncl, 1 << dim)
    types = arr(None, "types")
    assert np.all(types == (9 if dim == 2 else 12))
    # expected points and values
    xe = np.arange(k + 1) / k
    G = np.meshgrid(*([xe] * dim), indexing="ij")
    xi = np.stack([G[dim - 1 - d].ravel() for d in range(dim)], axis=-1)  # xi_0 fastest

    def lag(nodes, x):
        return np.array([[np.prod([(xx - nodes[j]) / (nodes[i] - nodes[j])
                                   for j in range(len(nodes)) if j != i])
                          for i in range(len(nodes))] for xx in x])
    xg, _ = O.gauss(k + 1)
    Lg = lag(xg, xe)
    Lp = lag(O.gauss_lobatto(k + 2), xe)

    def eval_tensor(L, vals):  # vals: n^dim, xi_0 fastest
        n = L.shape[1]
        t = vals.reshape((n,) * dim)  # axes (xi_{dim-1}, ..., xi_0)
        for ax in range(dim):
            t = np.tensordot(L, t, axes=([1], [ax]))
            t = np.moveaxis(t, 0, ax)
        return t.ravel()
    V = m.vertices
    nd = (k + 1) ** dim
    uu = u.reshape(m.n_cells, dim + 1, nd)
    want_p, want_x, want_post = [], [], []
    for c in range(m.n_cells):
        w = np.ones((len(xi), 1 << dim))
        for v in range(1 << dim):
            for d in range(dim):
                w[:, v] *= xi[:, d] if (v >> d) & 1 else 1 - xi[:, d]
        want_x.append(w @ V[c])
        want_p.append(eval_tensor(Lg, uu[c, dim]))
        want_post.append(eval_tensor(Lp, post[c * (k + 2) ** dim:(c + 1) * (k + 2) ** dim]))
    assert np.allclose(pts[:, :dim], np.concatenate(want_x), atol=1e-9)
    assert np.allclose(arr(None, "solution_pressure"), np.concatenate(want_p), atol=1e-8)
    assert np.allclose(arr(None, "post_pressure"), np.concatenate(want_post), atol=1e-8)
    assert np.allclose(arr(None, "Error_estimate"), np.repeat(est, k ** dim), atol=1e-9)
    # orientation of every sub-cell
    P0 = pts[conn[:, 0], :dim]
    if dim == 2:
        a, b = pts[conn[:, 1], :2] - P0, pts[conn[:, 3], :2] - P0
        assert np.all(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0] > 0)
    else:
        a, b, c3 = (pts[conn[:, i]] - pts[conn[:, 0]] for i in (1, 3, 4))
        assert np.all(np.einsum("ij,ij->i", np.cross(a, b), c3) > 0)


def test_refinement_flags_fixed_number():
    """refine_and_coarsen_fixed_number(0.1, 0.6) on distinct indicators: the
    int(0.1 n) largest are refined, the int(0.6 n) smallest coarsened (cells of
    an intermediate level, thresholds off)."""
    import exwave_amd as X
    rng = np.random.default_rng(0)
    for n in (7, 10, 64, 333):
        e = rng.permutation(n).astype(float) + 1.0
        fl = X.refinement_flags(e, np.ones(n, np.int32), 0, 2, -1.0)
        order = np.argsort(-e)
        nr, ncs = int(0.1 * n), int(0.6 * n)
        assert set(np.nonzero(fl == 1)[0]) == set(order[:nr])
        assert set(np.nonzero(fl == -1)[0]) == set(order[n - ncs:])
        assert (fl == 0).sum() == n - nr - ncs


def test_refinement_flags_rules():
    """adapt_mesh's rules: no refinement at the maximum level, no coarsening at
    the minimum level, no refinement below 0.1 max_error_init, coarsening below
    0.05 max_error_init (source/explicit_wave.cc:281-296); ties at a threshold
    are flagged (>= / <=)."""
    import exwave_amd as X
    n = 20
    e = np.arange(1.0, n + 1)
    lv = np.zeros(n, np.int32)
    lv[-1] = 1  # the largest indicator sits on the finest level
    fl = X.refinement_flags(e, lv, 0, 1, -1.0)
    assert fl[-1] == 0 and fl[-2] == 1  # int(0.1*20) = 2 largest: 20 (max level), 19
    assert np.all(fl[:12] == 0)  # coarsening flags on the minimum level are cleared
    fl = X.refinement_flags(e, np.ones(n, np.int32), 0, 2, 200.0)
    assert np.all(fl[e < 10.0] == -1)  # below 0.05 * 200
    assert fl[18] == 0 and fl[19] == 1  # 19 < 0.1 * 200 is not refined, 20 is
    e2 = np.array([1.0, 1.0, 1.0, 5.0, 5.0, 2.0, 3.0, 4.0, 6.0, 6.0])
    fl = X.refinement_flags(e2, np.ones(10, np.int32), 0, 2, -1.0)
    assert fl[8] == 1 and fl[9] == 1  # tie at the refinement threshold: both


def test_adapt_box_refine():
    import exwave_amd as X
    dim, n = 2, 3
    ro = np.zeros((n, n), np.int32)
    ro[0, 0] = ro[1, 1] = 1
    V, nb, H = X.mesh_box_refined(dim, n, ro)
    fl = np.zeros(len(V), np.int32)
    # base 0 refined: all four children coarsen -> unrefined; base 4 (refined):
    # three of four -> stays refined; base 2 unrefined flagged -> refined
    fl[0:4] = -1
    first4 = 4 + 1 + 1 + 1  # cells of bases 0 (4), 1, 2, 3 precede base 4
    fl[first4:first4 + 3] = -1
    fl[4 + 1] = 1  # base 2 is the third output cell block: bases 0 (4 cells), 1 -> index 5
    rn = X.adapt_box_refine(dim, n, ro, fl)
    assert rn[0] == 0 and rn[4] == 1 and rn[2] == 1 and rn.sum() == 2


def _check(dim, V, nb):
    import ctypes as C

    import exwave_amd.capi as capi
    L = capi.load()
    V = np.ascontiguousarray(V, dtype=np.float64)
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    return L.exw_mesh_check(dim, V.shape[0], V.ctypes.data_as(C.POINTER(C.c_double)),
                            nb.ctypes.data_as(C.POINTER(C.c_int)), 0, None)


@pytest.mark.parametrize("dim", [2, 3])
def test_mesh_check_rejects_non_standard_orientation(dim):
    """exw_mesh_check (run by exw_op_create): generated meshes pass; two cells
    side by side whose shared face runs the other way in one of them (the
    second cell mirrored in xi_1: a flipped face, faces still paired as 1/0),
    a cell whose faces pair as f/f rather than f/f^1 (a rotated or mirrored
    cell inside a mesh), and a neighbour entry that is not returned are all
    rejected before anything runs."""
    import exwave_amd as X
    V, nb = X.mesh_hypercube(dim, 3, 0, X.TRANSFORM_REFERENCE_SINE, 0.1, 2)
    assert _check(dim, V, nb) == 0
    nv = 1 << dim
    # two unit cells along x, every other face a boundary
    V2 = np.zeros((2, nv, dim))
    for c in range(2):
        for v in range(nv):
            for d in range(dim):
                V2[c, v, d] = ((v >> d) & 1) + (c if d == 0 else 0)
    nb2 = np.full((2, 2 * dim), -2, dtype=np.int32)
    nb2[0, 1], nb2[1, 0] = 1, 0
    assert _check(dim, V2, nb2) == 0
    V2f = V2.copy()
    V2f[1] = V2[1][[v ^ 2 for v in range(nv)]]  # mirrored in xi_1
    assert _check(dim, V2f, nb2) == -1  # EXW_EINVAL
    assert "orientation" in X.load().exw_last_error().decode()
    # inside a mesh, the mirrored cell's xi_1 faces pair with its neighbours'
    # faces of the same index
    c = 4
    Vf = V.copy()
    Vf[c] = V[c][[v ^ 2 for v in range(nv)]]
    nbf = nb.copy()
    nbf[c, 2], nbf[c, 3] = nb[c, 3], nb[c, 2]
    assert _check(dim, Vf, nbf) == -1
    assert "orientation" in X.load().exw_last_error().decode()
    # a neighbour entry that is not returned
    nbr = nb.copy()
    j = int(np.flatnonzero(nbr[:, 1] >= 0)[0])
    nbr[j, 1] = nbr[j, 1] + 1 if nbr[j, 1] + 1 < V.shape[0] else 0
    assert _check(dim, V, nbr) < 0


def test_host_entry_points_reject_empty_and_malformed_input():
    """Empty meshes, empty indicator arrays and malformed forests are
    rejected with EXW_EINVAL by the host-side entry points (no device use)."""
    import ctypes as C

    import exwave_amd as X
    L = X.load()
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    V = np.zeros((1, 4, 2))
    nb = -np.ones((1, 4), np.int32)
    vp, nbp = V.ctypes.data_as(dp), nb.ctypes.data_as(ip)
    assert L.exw_mesh_check(2, 0, vp, nbp, 0, None) < 0          # no cells
    assert L.exw_mesh_check(4, 1, vp, nbp, 0, None) < 0          # bad dimension
    out = np.zeros(4, np.int32)
    cl, fl = out.ctypes.data_as(ip), out.ctypes.data_as(ip)
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert L.exw_lts_categorize(2, 0, vp, nbp, 0.1, 3, 1, cl, fl, C.byref(a), C.byref(b),
                                C.byref(c)) < 0
    err = np.zeros(1)
    assert L.exw_refinement_flags(0, err.ctypes.data_as(dp), out.ctypes.data_as(ip), 0, 1,
                                  1.0, out.ctypes.data_as(ip)) < 0
    # forests: wrong length, entries other than 0/1, faces two levels apart
    for tree in ([1, 0, 0, 0], [2], [1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0]):
        with pytest.raises(X.ExwError):
            X.mesh_forest(2, 1, tree)
    with pytest.raises(X.ExwError):
        X.adapt_forest(2, 1, [1, 0, 0, 0], [0, 0, 0])
    # a one-cell forest, flagged, refines; a one-level family coarsens back
    t1 = X.adapt_forest(2, 1, [0], [1])
    assert list(t1) == [1, 0, 0, 0, 0]
    assert list(X.adapt_forest(2, 1, t1, [-1, -1, -1, -1])) == [0]
    assert list(X.adapt_forest(2, 1, t1, [-1, -1, -1, 0])) == list(t1)
