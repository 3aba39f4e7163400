"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module; it is the checker, never the product path.
The C sources are ``exwave_oracle.{c,h}`` next to this file; ``build()``
compiles them with gcc into ``oracle/_build/``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libexwave_oracle.so")

EXPL_EULER, CLRK4, LSRK45R2, LSRK33R2, LSRK59R2, ADER, ADER_LTS, LSRK45R3, SSPRK, ADER_ADCONFULL = range(10)
INTEGRATORS = {"ExplEuler": EXPL_EULER, "clRK4": CLRK4, "LSRK45R2": LSRK45R2,
               "LSRK33R2": LSRK33R2, "LSRK59R2": LSRK59R2, "ADER": ADER, "ADERLTS": ADER_LTS,
               "LSRK45R3": LSRK45R3, "SSPRK": SSPRK, "ADERADCONFULL": ADER_ADCONFULL}
TRANSFORM_NONE, TRANSFORM_REFERENCE_SINE, TRANSFORM_AXIS_2PI, TRANSFORM_PRODUCT_2PI = range(4)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _SO


class _Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_cells", C.c_int),
                ("vertices", C.POINTER(C.c_double)), ("neighbor", C.POINTER(C.c_int)),
                ("rho", C.POINTER(C.c_double)), ("speed", C.POINTER(C.c_double)),
                ("min_vertex_distance_over_c", C.c_double),
                ("n_hanging", C.c_int), ("hanging", C.POINTER(C.c_int)),
                ("nc_rec", C.POINTER(C.c_int)), ("nc_pos", C.POINTER(C.c_int))]


NC_FACE = -1000  # neighbour entry of a non-conforming face


class Exact(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_modes", C.c_int), ("amp", C.c_double * 8),
                ("m", (C.c_int * 3) * 8), ("rho", C.c_double), ("c", C.c_double)]

    @staticmethod
    def make(dim, modes, amps=None, rho=1.0, c=1.0):
        e = Exact()
        e.dim = dim
        e.n_modes = len(modes)
        for k, mm in enumerate(modes):
            e.amp[k] = 1.0 if amps is None else amps[k]
            for d in range(dim):
                e.m[k][d] = int(mm[d])
        e.rho, e.c = rho, c
        return e

    @staticmethod
    def reference(dim, membrane_modes):
        """input_exact_solution, include/input_parameters.h:94-128"""
        return Exact.make(dim, [[membrane_modes] * dim])


class Source(C.Structure):
    """ox_source: Ricker point source of pressure at x."""
    _fields_ = [("x", C.c_double * 3), ("amplitude", C.c_double), ("f0", C.c_double),
                ("t0", C.c_double)]

    @staticmethod
    def make(x, amplitude=1.0, f0=10.0, t0=None):
        s = Source()
        for i, v in enumerate(x):
            s.x[i] = float(v)
        s.amplitude, s.f0 = float(amplitude), float(f0)
        s.t0 = 1.0 / f0 if t0 is None else float(t0)
        return s

    def value(self, t):
        return lib().ox_ricker(C.byref(self), t)

    def integral(self, t1, t2):
        return lib().ox_ricker_integral(C.byref(self), t1, t2)


LTS_SPEED_AWARE = 1


class Problem(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("n_intervals", C.c_int),
                ("n_refinements", C.c_int), ("transform", C.c_int),
                ("transform_factor", C.c_double), ("boundary_id", C.c_int),
                ("integrator", C.c_int), ("cfl_number", C.c_double),
                ("max_time_steps", C.c_int), ("final_time", C.c_double),
                ("output_every_time", C.c_double), ("ader_recon", C.c_int),
                ("exact", Exact), ("lts_max_clusters", C.c_int), ("lts_max_diff", C.c_int),
                ("post_pressure", C.c_int), ("n_adaptive_refinements", C.c_int),
                ("adaptive_refinement_interval", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(build())
        L = _lib
        dp = C.POINTER(C.c_double)
        L.ox_mesh_hypercube.restype = C.POINTER(_Mesh)
        L.ox_mesh_hypercube.argtypes = [C.c_int] * 5 + [C.c_double, C.c_int]
        L.ox_mesh_hypercube.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
        L.ox_mesh_box_refined.restype = C.POINTER(_Mesh)
        L.ox_mesh_box_refined.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                          C.POINTER(C.c_int)]
        L.ox_mesh_free.argtypes = [C.POINTER(_Mesh)]
        L.ox_mesh_set_material.argtypes = [C.POINTER(_Mesh), dp, dp]
        L.ox_apply.argtypes = [C.POINTER(_Mesh), C.c_int, dp, dp, C.c_double]
        L.ox_rk_update.argtypes = [C.c_long, C.c_double, C.c_double, C.c_int, dp, dp, dp]
        L.ox_time_step.argtypes = [C.POINTER(_Mesh), C.c_int, C.c_int, C.c_double, dp, dp, dp,
                                   dp, C.c_int]
        L.ox_ader_ck.argtypes = [C.POINTER(_Mesh), C.c_int, dp, dp, C.c_double, C.c_double,
                                 C.c_double, dp]
        L.ox_apply_ader.argtypes = [C.POINTER(_Mesh), C.c_int, dp, dp, C.c_double, C.c_int, dp]
        L.ox_project.argtypes = [C.POINTER(_Mesh), C.c_int, C.POINTER(Exact), C.c_double, dp]
        L.ox_l2_errors.argtypes = [C.POINTER(_Mesh), C.c_int, C.POINTER(Exact), C.c_double, dp,
                                   dp]
        L.ox_energy.restype = C.c_double
        L.ox_energy.argtypes = [C.POINTER(_Mesh), C.c_int, dp]
        L.ox_run.restype = C.c_int
        L.ox_run.argtypes = [C.POINTER(Problem), C.c_char_p, C.c_int, dp]
        L.ox_n_dofs.restype = C.c_long
        L.ox_n_dofs.argtypes = [C.POINTER(_Mesh), C.c_int]
        L.ox_gauss.argtypes = [C.c_int, dp, dp]
        L.ox_lagrange.argtypes = [C.c_int, dp, C.c_double, dp, dp]
        ip = C.POINTER(C.c_int)
        L.ox_lts_create.restype = C.c_void_p
        L.ox_lts_create.argtypes = [C.POINTER(_Mesh), C.c_int, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_double, C.c_int]
        L.ox_lts_free.argtypes = [C.c_void_p]
        for f in ("ox_lts_n_clusters", "ox_lts_cluster_diff", "ox_lts_iterations",
                  "ox_lts_n_updates"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.c_void_p]
        L.ox_lts_global_dt.restype = C.c_double
        L.ox_lts_global_dt.argtypes = [C.c_void_p]
        L.ox_lts_schedule.argtypes = [C.c_void_p, ip, dp]
        L.ox_lts_cells.argtypes = [C.c_void_p, ip, ip]
        L.ox_lts_step.argtypes = [C.c_void_p, C.c_double, dp, dp]
        L.ox_lts_create2.restype = C.c_void_p
        L.ox_lts_create2.argtypes = [C.POINTER(_Mesh), C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.c_double, C.c_int, C.c_int]
        L.ox_lts_set_source.argtypes = [C.c_void_p, C.POINTER(Source)]
        L.ox_ricker.restype = C.c_double
        L.ox_ricker.argtypes = [C.POINTER(Source), C.c_double]
        L.ox_ricker_integral.restype = C.c_double
        L.ox_ricker_integral.argtypes = [C.POINTER(Source), C.c_double, C.c_double]
        L.ox_source_cells.restype = C.c_int
        L.ox_source_cells.argtypes = [C.POINTER(_Mesh), C.c_int, C.POINTER(Source), ip, dp]
        L.ox_source_add.argtypes = [C.POINTER(_Mesh), C.c_int, C.POINTER(Source), C.c_double,
                                    C.c_double, dp]
        L.ox_gauss_lobatto.argtypes = [C.c_int, dp]
        L.ox_estimate_error.argtypes = [C.POINTER(_Mesh), C.c_int, dp, dp]
        L.ox_adapt_transfer.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, dp]
        L.ox_forest_levels.restype = C.c_int
        L.ox_forest_levels.argtypes = [C.c_int, C.c_int, ip, C.c_int, ip]
        L.ox_mesh_forest.restype = C.POINTER(_Mesh)
        L.ox_mesh_forest.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, ip, C.c_int]
        L.ox_adapt_forest.restype = C.c_int
        L.ox_adapt_forest.argtypes = [C.c_int, C.c_int, ip, ip, ip, C.c_int]
        L.ox_adapt_forest_transfer.restype = C.c_int
        L.ox_adapt_forest_transfer.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, dp]
        L.ox_post_pressure.argtypes = [C.POINTER(_Mesh), C.c_int, dp, dp]
        L.ox_post_set_solver.argtypes = [C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_int]
        L.ox_post_error.restype = C.c_double
        L.ox_post_error.argtypes = [C.POINTER(_Mesh), C.c_int, C.POINTER(Exact), C.c_double, dp]
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def forest_levels(dim, n_intervals, tree):
    t = np.ascontiguousarray(tree, dtype=np.int32)
    n = lib().ox_forest_levels(dim, n_intervals, _ip(t), t.size, None)
    if n < 0:
        raise ValueError("not a forest")
    lev = np.zeros(n, np.int32)
    lib().ox_forest_levels(dim, n_intervals, _ip(t), t.size, _ip(lev))
    return lev


def adapt_forest(dim, n_intervals, tree, flags):
    """ox_adapt_forest: the new depth-first trees after the flags."""
    t = np.ascontiguousarray(tree, dtype=np.int32)
    f = np.ascontiguousarray(flags, dtype=np.int32)
    n = lib().ox_adapt_forest(dim, n_intervals, _ip(t), _ip(f), None, 0)
    if n < 0:
        raise ValueError("ox_adapt_forest failed")
    out = np.zeros(n, np.int32)
    assert lib().ox_adapt_forest(dim, n_intervals, _ip(t), _ip(f), _ip(out), n) == n
    return out


def adapt_forest_transfer(dim, degree, n_intervals, tree_old, tree_new, old_u, n_new_dofs):
    to = np.ascontiguousarray(tree_old, dtype=np.int32)
    tn = np.ascontiguousarray(tree_new, dtype=np.int32)
    out = np.zeros(n_new_dofs)
    u = np.ascontiguousarray(old_u, dtype=np.float64)
    if lib().ox_adapt_forest_transfer(dim, degree, n_intervals, _ip(to), _ip(tn), _p(u), _p(out)):
        raise ValueError("trees differ by more than one level")
    return out


def uniform_forest(dim, n_intervals, depth):
    """the depth-first trees of the box refined `depth` times everywhere"""
    one = []

    def rec(l):
        one.append(1 if l < depth else 0)
        if l < depth:
            for _ in range(1 << dim):
                rec(l + 1)
    rec(0)
    return np.array(one * n_intervals ** dim, np.int32)


class Mesh:
    """A hypercube mesh exactly as the reference builds it (see exwave_oracle.c)."""

    def __init__(self, dim, n_intervals, n_refinements=0, transform=0, factor=0.0,
                 boundary_id=2):
        self.dim = dim
        self._m = lib().ox_mesh_hypercube(dim, n_intervals, n_refinements, transform, factor,
                                          boundary_id)
        self.n_cells = self._m.contents.n_cells

    @classmethod
    def forest(cls, dim, n_intervals, tree, transform=0, factor=0.0, boundary_id=2):
        """ox_mesh_forest: the box with base cells refined by the depth-first
        trees in `tree` (1 refined, 0 active)."""
        t = np.ascontiguousarray(tree, dtype=np.int32).ravel()
        self = cls.__new__(cls)
        self.dim = dim
        m = lib().ox_mesh_forest(dim, n_intervals, transform, factor, boundary_id, _ip(t), t.size)
        if not m:
            raise ValueError("ox_mesh_forest: not a 2:1 balanced forest")
        self._m = m
        self.n_cells = self._m.contents.n_cells
        return self

    @classmethod
    def box_refined(cls, dim, n_intervals, refine, transform=0, factor=0.0, boundary_id=2):
        """ox_mesh_box_refined: base cells flagged in refine (n_intervals^dim,
        lexicographic) split once; faces to unrefined cells hang."""
        r = np.ascontiguousarray(refine, dtype=np.int32).ravel()
        assert r.size == n_intervals ** dim
        self = cls.__new__(cls)
        self.dim = dim
        self._m = lib().ox_mesh_box_refined(dim, n_intervals, transform, factor, boundary_id,
                                            r.ctypes.data_as(C.POINTER(C.c_int)))
        self.n_cells = self._m.contents.n_cells
        return self

    @property
    def hanging(self):
        """(n_hanging, 2 + 2^(dim-1)) records: coarse cell, coarse face, fine cells."""
        n = self._m.contents.n_hanging
        w = 2 + (1 << (self.dim - 1))
        if n == 0:
            return np.zeros((0, w), dtype=np.int32)
        return np.ctypeslib.as_array(self._m.contents.hanging, shape=(n, w)).copy()

    def __del__(self):
        if getattr(self, "_m", None) is not None and _lib is not None:
            _lib.ox_mesh_free(self._m)
            self._m = None

    @property
    def vertices(self):
        nv = 1 << self.dim
        return np.ctypeslib.as_array(self._m.contents.vertices,
                                     shape=(self.n_cells, nv, self.dim)).copy()

    @property
    def neighbor(self):
        return np.ctypeslib.as_array(self._m.contents.neighbor,
                                     shape=(self.n_cells, 2 * self.dim)).copy()

    @property
    def rho(self):
        return np.ctypeslib.as_array(self._m.contents.rho, shape=(self.n_cells,)).copy()

    @property
    def speed(self):
        return np.ctypeslib.as_array(self._m.contents.speed, shape=(self.n_cells,)).copy()

    @property
    def min_vertex_distance_over_c(self):
        return self._m.contents.min_vertex_distance_over_c

    def set_material(self, rho, speed):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        speed = np.ascontiguousarray(speed, dtype=np.float64)
        lib().ox_mesh_set_material(self._m, _p(rho), _p(speed))

    def n_dofs(self, degree):
        return lib().ox_n_dofs(self._m, degree)

    def apply(self, degree, src, sign=1.0):
        dst = np.empty_like(src)
        lib().ox_apply(self._m, degree, _p(src), _p(dst), sign)
        return dst

    def time_step(self, degree, integrator, dt, vec_n, vec_np, ader_recon=0):
        t1 = np.zeros_like(vec_n)
        t2 = np.zeros_like(vec_n)
        lib().ox_time_step(self._m, degree, integrator, dt, _p(vec_n), _p(vec_np), _p(t1),
                           _p(t2), ader_recon)

    def ader_ck(self, degree, src, recon, t2, t1, te):
        dst = np.empty_like(src)
        rp = _p(recon) if recon is not None else None
        lib().ox_ader_ck(self._m, degree, _p(src), rp, t2, t1, te, _p(dst))
        return dst

    def apply_ader(self, degree, src, dt, recon=0):
        dst = np.empty_like(src)
        scr = np.empty(2 * src.size)
        lib().ox_apply_ader(self._m, degree, _p(src), _p(dst), dt, recon, _p(scr))
        return dst

    def project(self, degree, exact, t=0.0):
        dst = np.empty(self.n_dofs(degree))
        lib().ox_project(self._m, degree, C.byref(exact), t, _p(dst))
        return dst

    def l2_errors(self, degree, exact, t, u):
        out = np.zeros(3)
        lib().ox_l2_errors(self._m, degree, C.byref(exact), t, _p(u), _p(out))
        return out

    def energy(self, degree, u):
        return lib().ox_energy(self._m, degree, _p(u))

    def estimate_error(self, degree, u):
        """estimate_error: per cell sqrt(int |grad p - v_t|^2)."""
        out = np.zeros(self.n_cells)
        lib().ox_estimate_error(self._m, degree, _p(u), _p(out))
        return out

    def post_pressure(self, degree, u, rel_tol=1e-11, abs_tol=1e-20, max_iter_factor=6,
                      iterations=False, mean_free=False):
        """compute_post_pressure: n_cells x (k+2)^dim Gauss-Lobatto nodal values
        (and each cell's CG iteration count if `iterations`); mean_free solves
        the same local problem on the zero-sum subspace instead of fixing the
        first unknown."""
        out = np.zeros(self.n_cells * (degree + 2) ** self.dim)
        its = np.zeros(self.n_cells, dtype=np.int32)
        L = lib()
        L.ox_post_set_solver(abs_tol, rel_tol, max_iter_factor, its.ctypes.data, int(mean_free))
        try:
            L.ox_post_pressure(self._m, degree, _p(u), _p(out))
        finally:
            L.ox_post_set_solver(1e-20, 1e-11, 6, None, 0)
        return (out, its) if iterations else out

    def post_error(self, degree, exact, t, post):
        return lib().ox_post_error(self._m, degree, C.byref(exact), t, _p(post))

    def source_cells(self, degree, src):
        """(cells holding the source point, their nodal vectors b [n_s, (k+1)^dim])."""
        nd = (degree + 1) ** self.dim
        cells = np.zeros(8, dtype=np.int32)
        b = np.zeros(8 * nd)
        n = lib().ox_source_cells(self._m, degree, C.byref(src),
                                  cells.ctypes.data_as(C.POINTER(C.c_int)), _p(b))
        return cells[:n].copy(), b[:n * nd].reshape(n, nd).copy()

    def source_add(self, degree, src, t1, t2, u):
        lib().ox_source_add(self._m, degree, C.byref(src), t1, t2, _p(u))


class Lts:
    """ox_lts: ADER with local time stepping on an oracle mesh."""

    def __init__(self, mesh, degree, cfl, fastest_dt, max_clusters, max_diff, time0=0.0,
                 recon=0, speed_aware=False, source=None):
        L = lib()
        self.mesh = mesh
        self.h = L.ox_lts_create2(mesh._m, degree, cfl, fastest_dt, max_clusters, max_diff,
                                  time0, int(recon), LTS_SPEED_AWARE if speed_aware else 0)
        assert self.h, "ox_lts_create2 failed"
        if source is not None:
            self._src = source
            L.ox_lts_set_source(self.h, C.byref(source))
        self.n_clusters = L.ox_lts_n_clusters(self.h)
        self.cluster_diff = L.ox_lts_cluster_diff(self.h)
        self.global_dt = L.ox_lts_global_dt(self.h)

    def cells(self):
        cl = np.zeros(self.mesh.n_cells, np.int32)
        fl = np.zeros(self.mesh.n_cells, np.int32)
        ip = C.POINTER(C.c_int)
        lib().ox_lts_cells(self.h, cl.ctypes.data_as(ip), fl.ctypes.data_as(ip))
        return cl, fl

    def step(self, time_now, u):
        out = np.zeros_like(u)
        lib().ox_lts_step(self.h, time_now, _p(np.ascontiguousarray(u)), _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ox_lts_free(self.h)
            self.h = None


def run(problem: Problem):
    """Run a whole solve; returns (n_steps, [(step, time, err_p, err_v, mag_p)], energy);
    with problem.post_pressure set the rows carry a sixth entry, error p post."""
    buf = C.create_string_buffer(1 << 16)
    en = C.c_double(0.0)
    steps = lib().ox_run(C.byref(problem), buf, len(buf), C.byref(en))
    rows = []
    for line in buf.value.decode().strip().splitlines():
        s, t, ep, ev, mp, epp = line.split()
        row = (int(s), float(t), float(ep), float(ev), float(mp))
        rows.append(row + (float(epp),) if problem.post_pressure else row)
    return steps, rows, en.value


def gauss_lobatto(n):
    x = np.zeros(n)
    lib().ox_gauss_lobatto(n, _p(x))
    return x


def gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().ox_gauss(n, _p(x), _p(w))
    return x, w
