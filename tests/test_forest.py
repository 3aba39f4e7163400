"""Forests of refinement trees (several levels of adaptive refinement,
adapt_mesh in source/explicit_wave.cc:265-330 with min/max level from :272-273).

The oracle (dense finest-grid levels) and the library (hash-mapped tree nodes)
compute the new trees in different ways and must agree exactly.  Both are
pinned by properties checked here from the trees alone, in plain Python: the
mesh is 2:1 balanced across faces, edges and corners; every flagged cell is
refined; every other refinement is forced by the balance; families are
coarsened only when all children were active and flagged; and every family
that was eligible but kept would break the balance if it were coarsened too.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import exwave_amd as X
import oracle as O


def _cells(dim, n, tree):
    """(level, coords) of the active cells, depth first (a restatement of the
    encoding, independent of the C parsers)"""
    out, pos = [], 0

    def rec(l, I):
        nonlocal pos
        t = tree[pos]
        pos += 1
        if t == 0:
            out.append((l, tuple(I)))
            return
        for ch in range(1 << dim):
            rec(l + 1, [2 * I[d] + ((ch >> d) & 1) for d in range(dim)])

    for B in range(n ** dim):
        rec(0, [(B // n ** d) % n for d in range(dim)])
    assert pos == len(tree)
    return out


def _grid(dim, n, cells, L):
    G = n << L
    lev = np.full((G,) * dim, -1, dtype=np.int64)
    for l, I in cells:
        s = 1 << (L - l)
        lev[tuple(slice(i * s, (i + 1) * s) for i in I)] = l
    assert (lev >= 0).all()
    return lev


def _balanced(lev):
    dim = lev.ndim
    for o in itertools.product((-1, 0, 1), repeat=dim):
        if not any(o):
            continue
        a = lev[tuple(slice(max(0, -x), lev.shape[0] - max(0, x)) for x in o)]
        b = lev[tuple(slice(max(0, x), lev.shape[0] - max(0, -x)) for x in o)]
        if (np.abs(a - b) > 1).any():
            return False
    return True


def _random_flags(rng, lev, max_level, p=(0.35, 0.35, 0.3)):
    fl = rng.choice([-1, 0, 1], size=lev.size, p=list(p)).astype(np.int32)
    fl[(fl == 1) & (lev >= max_level)] = 0
    fl[(fl == -1) & (lev <= 0)] = 0
    return fl


def _check_properties(dim, n, old, flags, new):
    oc, nc = _cells(dim, n, old), _cells(dim, n, new)
    L = max(l for l, _ in oc + nc)
    lo, ln = _grid(dim, n, oc, L), _grid(dim, n, nc, L)
    assert _balanced(ln)
    # every cell changes by at most one level; flagged cells are refined
    assert (np.abs(ln - lo) <= 1).all()
    for (l, I), f in zip(oc, flags):
        s = 1 << (L - l)
        blk = ln[tuple(slice(i * s, (i + 1) * s) for i in I)]
        if f == 1:
            assert (blk == l + 1).all()
        if (blk < l).any():  # coarsened: a whole family, all flagged
            assert (blk == l - 1).all() and f == -1
    # minimality: a refinement that was not flagged is forced by the balance
    flagged = {c for c, f in zip(oc, flags) if f == 1}
    for c in oc:
        l, I = c
        s = 1 << (L - l)
        idx = tuple(slice(i * s, (i + 1) * s) for i in I)
        if c not in flagged and (ln[idx] == l + 1).all():
            trial = ln.copy()
            trial[idx] = l
            assert not _balanced(trial), ("unforced refinement", c)
    # maximality of the coarsening: an eligible family that was kept would
    # break the balance if it were coarsened as well
    fam = {}
    for c, f in zip(oc, flags):
        l, I = c
        if l > 0:
            fam.setdefault((l - 1, tuple(i >> 1 for i in I)), []).append(f == -1)
    for (l, P), fs in fam.items():
        if len(fs) != 1 << dim or not all(fs):
            continue
        s = 1 << (L - l)
        idx = tuple(slice(i * s, (i + 1) * s) for i in P)
        if (ln[idx] == l + 1).all():  # kept (and not refined further)
            trial = ln.copy()
            trial[idx] = l
            assert not _balanced(trial), ("coarsening not done", (l, P))


@pytest.mark.parametrize("dim,n,steps,seed", [(2, 3, 6, 0), (3, 2, 4, 1), (2, 1, 7, 2),
                                              (3, 1, 4, 3), (2, 2, 6, 4)])
def test_adapt_forest_library_equals_oracle_and_properties(dim, n, steps, seed):
    rng = np.random.default_rng(seed)
    tree = O.uniform_forest(dim, n, 1)
    for _ in range(steps):
        lev = O.forest_levels(dim, n, tree)
        fl = _random_flags(rng, lev, 4)
        t_o = O.adapt_forest(dim, n, tree, fl)
        t_h = X.adapt_forest(dim, n, tree, fl)
        assert t_o.size == t_h.size and (t_o == t_h).all()
        _check_properties(dim, n, tree, fl, t_o)
        tree = t_o


def test_deep_refinement_forces_a_cascade():
    """refining, three times in a row, the cell of base cell 0 that touches
    the centre of a 2x2 box refines the other base cells through the 2:1 rule
    (across the corner for the diagonal one)"""
    dim, n = 2, 2
    tree = O.uniform_forest(dim, n, 0)
    x = 0.5 - 1e-9
    for _ in range(3):
        cells = _cells(dim, n, tree)
        fl = np.zeros(len(cells), np.int32)
        hit = [i for i, (l, I) in enumerate(cells)
               if all(I[d] <= x * (n << l) < I[d] + 1 for d in range(dim))]
        assert len(hit) == 1
        fl[hit[0]] = 1
        new = X.adapt_forest(dim, n, tree, fl)
        assert (new == O.adapt_forest(dim, n, tree, fl)).all()
        _check_properties(dim, n, tree, fl, new)
        tree = new
    cells = _cells(dim, n, tree)
    assert max(l for l, _ in cells) == 3
    # no base cell other than 0 is left at level 0; the diagonal one (3) was
    # reached only through the corner at the centre
    assert all(l >= 1 for l, _ in cells)
    assert min(l for l, I in cells if I[0] >= (n << l) // 2 and I[1] >= (n << l) // 2) == 1


@pytest.mark.parametrize("dim,n,seed", [(2, 3, 5), (3, 2, 6)])
def test_mesh_forest_library_equals_oracle(dim, n, seed):
    rng = np.random.default_rng(seed)
    tree = O.uniform_forest(dim, n, 1)
    for _ in range(3):
        tree = O.adapt_forest(dim, n, tree, _random_flags(rng, O.forest_levels(dim, n, tree), 3))
    for tr, fac, bid in ((0, 0.0, 2), (1, 0.05, 1), (3, 0.04, 3)):
        V, nb, lev, H = X.mesh_forest(dim, n, tree, tr, fac, bid)
        om = O.Mesh.forest(dim, n, tree, tr, fac, bid)
        m = om._m.contents
        nv, ns = 1 << dim, 1 << (dim - 1)
        assert m.n_cells == V.shape[0]
        Vo = np.ctypeslib.as_array(m.vertices, shape=(m.n_cells * nv * dim,))
        nbo = np.ctypeslib.as_array(m.neighbor, shape=(m.n_cells * 2 * dim,))
        Ho = np.ctypeslib.as_array(m.hanging, shape=(max(m.n_hanging, 1) * (2 + ns),))
        assert np.array_equal(Vo, V.ravel())
        assert np.array_equal(nbo, nb.ravel())
        assert np.array_equal(lev, O.forest_levels(dim, n, tree))
        assert m.n_hanging == H.shape[0] > 0
        assert sorted(map(tuple, Ho[: m.n_hanging * (2 + ns)].reshape(-1, 2 + ns))) == \
            sorted(map(tuple, H))
        # geometry: every hanging record passes the library's own checks
        ip = C.POINTER(C.c_int)
        assert X.load().exw_mesh_check(dim, V.shape[0], V.ctypes.data_as(C.POINTER(C.c_double)),
                                       nb.ctypes.data_as(ip), H.shape[0],
                                       H.ctypes.data_as(ip)) == 0


@pytest.mark.parametrize("dim,n", [(2, 4), (3, 3)])
def test_one_level_forest_is_the_box_mesh(dim, n):
    rng = np.random.default_rng(7)
    ref = rng.integers(0, 2, n ** dim).astype(np.int32)
    tree = []
    for r in ref:
        tree += [1] + [0] * (1 << dim) if r else [0]
    V1, nb1, H1 = X.mesh_box_refined(dim, n, ref, 3, 0.04, 1)
    V2, nb2, lev, H2 = X.mesh_forest(dim, n, tree, 3, 0.04, 1)
    assert np.array_equal(V1, V2) and np.array_equal(nb1, nb2)
    assert sorted(map(tuple, H1)) == sorted(map(tuple, H2))
    assert np.array_equal(lev, np.repeat(ref, np.where(ref == 1, 1 << dim, 1)))
    # one adaptation of depth <= 1 equals the box rules
    lev = O.forest_levels(dim, n, np.array(tree, np.int32))
    fl = rng.choice([-1, 0, 1], size=lev.size).astype(np.int32)
    fl[(fl == 1) & (lev >= 1)] = 0
    fl[(fl == -1) & (lev <= 0)] = 0
    new = X.adapt_forest(dim, n, tree, fl)
    rnew = X.adapt_box_refine(dim, n, ref, fl)
    want = []
    for r in rnew:
        want += [1] + [0] * (1 << dim) if r else [0]
    assert np.array_equal(new, np.array(want, np.int32))


def _node_values(dim, n, cells, k, fn):
    """[cell][field][node] values of fn(x) (one scalar per field) at the Gauss
    nodes of the Cartesian forest cells, xi_0 fastest"""
    xg, _ = O.gauss(k + 1)
    out = []
    for l, I in cells:
        h = 1.0 / (n << l)
        pts = [np.array([(I[d] + xg[(q // (k + 1) ** d) % (k + 1)]) * h for d in range(dim)])
               for q in range((k + 1) ** dim)]
        for f in range(dim + 1):
            out += [fn(x, f) for x in pts]
    return np.array(out)


@pytest.mark.parametrize("dim,k", [(2, 3), (3, 2)])
def test_forest_transfer_oracle(dim, k):
    """the oracle's forest transfer: exact on fields that are one polynomial
    of degree k per direction (children interpolate it, parents project it
    back), and the identity after refining and coarsening back"""
    n = 2
    rng = np.random.default_rng(3 + dim)
    tree = O.uniform_forest(dim, n, 1)
    tree = O.adapt_forest(dim, n, tree, _random_flags(rng, O.forest_levels(dim, n, tree), 2))
    fl = _random_flags(rng, O.forest_levels(dim, n, tree), 3)
    new = O.adapt_forest(dim, n, tree, fl)
    oc, nc = _cells(dim, n, tree), _cells(dim, n, new)
    assert len(oc) != len(nc)

    def poly(x, f):
        return np.prod([(x[d] ** k - 0.3 * x[d] + f + 0.1 * d) for d in range(dim)])
    uo = _node_values(dim, n, oc, k, poly)
    un = O.adapt_forest_transfer(dim, k, n, tree, new, uo, len(nc) * (dim + 1) * (k + 1) ** dim)
    assert np.allclose(un, _node_values(dim, n, nc, k, poly), rtol=0, atol=1e-12)
    # arbitrary data: refine then coarsen back is the identity on those cells
    u = rng.standard_normal(uo.size)
    un = O.adapt_forest_transfer(dim, k, n, tree, new, u, un.size)
    back = O.adapt_forest_transfer(dim, k, n, new, tree, un, u.size)
    dpc = u.size // len(oc)
    newset = set(nc)
    checked = 0
    for i, c in enumerate(oc):
        l, I = c
        kids = [(l + 1, tuple(2 * I[d] + ((ch >> d) & 1) for d in range(dim)))
                for ch in range(1 << dim)]
        if c in newset or all(kid in newset for kid in kids):
            checked += 1
            assert np.allclose(back[i * dpc:(i + 1) * dpc], u[i * dpc:(i + 1) * dpc],
                               rtol=0, atol=1e-12)
    assert checked > 0


def test_malformed_and_unbalanced_forests_are_rejected():
    dim, n = 2, 1
    with pytest.raises(X.ExwError):
        X.mesh_forest(dim, n, [1, 0, 0, 0])           # too short
    with pytest.raises(X.ExwError):
        X.mesh_forest(dim, n, [2])                    # not 0/1
    # level 0 next to level 2 across a face: refine child 0, then its child 3
    bad = [1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0]
    with pytest.raises(X.ExwError):
        X.mesh_forest(dim, n, bad)
    with pytest.raises(ValueError):
        O.Mesh.forest(dim, n, bad)


def test_oracle_adaptive_lts_run_rebuilds_clusters():
    """ADER-LTS with adaptivity (the operator's setup rebuilds the clusters in
    make_dofs after every adaptation): the run keeps the global step of the
    new clusters (twice the refined cells' CFL step here) and its errors stay
    at the level of the adaptive global-ADER run of the same problem"""
    def prob(integ):
        p = O.Problem()
        p.dim, p.degree, p.n_intervals, p.n_refinements = 2, 3, 4, 0
        p.transform, p.transform_factor, p.boundary_id = O.TRANSFORM_AXIS_2PI, 0.03, 2
        p.integrator, p.cfl_number = integ, 0.2
        p.max_time_steps, p.final_time, p.output_every_time = 100000, 0.15, 0.05
        p.post_pressure, p.n_adaptive_refinements, p.adaptive_refinement_interval = 0, 2, 6
        p.lts_max_clusters, p.lts_max_diff, p.ader_recon = 3, 1, 0
        p.exact = O.Exact.make(2, [[2, 1], [1, 3]])
        return p
    sa, ra, _ = O.run(prob(O.ADER))
    sl, rl, _ = O.run(prob(O.ADER_LTS))
    assert len(ra) == len(rl) == 4
    # two clusters after the first adaptation: half as many global steps
    assert sl < sa and abs(2 * sl - sa) <= 2
    for a, b in zip(ra, rl):
        for j in (2, 3):
            assert np.isfinite(b[j]) and abs(b[j] - a[j]) <= 0.1 * a[j]
