"""ctypes binding of the C-ABI in include/exwave_amd.h.

The library is built in-tree by ``python -m exwave_amd.build`` into
``exwave_amd/_lib/libexwave_amd.so``.  There is no fallback: if the library
is missing every entry point raises.  Device vectors are passed as raw
pointers; torch CUDA/HIP tensors are the usual owners (``tensor.data_ptr()``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libexwave_amd.so")

EXW_OK, EXW_EINVAL, EXW_EDEVICE, EXW_ENOMEM, EXW_EUNSUP = 0, -1, -2, -3, -4
FP64, FP32, PRECISION_AUTO = 0, 1, 2
(EXPL_EULER, CLRK4, LSRK45R2, LSRK33R2, LSRK59R2, ADER, ADER_LTS, LSRK45R3, SSPRK,
 ADER_ADCONFULL) = range(10)
INTEGRATORS = {"ExplEuler": EXPL_EULER, "clRK4": CLRK4, "LSRK45R2": LSRK45R2,
               "LSRK33R2": LSRK33R2, "LSRK59R2": LSRK59R2, "ADER": ADER, "ADERLTS": ADER_LTS,
               "LSRK45R3": LSRK45R3, "SSPRK": SSPRK, "ADERADCONFULL": ADER_ADCONFULL}
TRANSFORM_NONE, TRANSFORM_REFERENCE_SINE, TRANSFORM_AXIS_2PI, TRANSFORM_PRODUCT_2PI = range(4)


class ExwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"exwave_amd error {code}: {msg}")
        self.code = code


class Modes(C.Structure):
    _fields_ = [("n_modes", C.c_int), ("amp", C.c_double * 8), ("m", (C.c_int * 3) * 8),
                ("rho", C.c_double), ("c", C.c_double)]

    @staticmethod
    def make(dim, modes, amps=None, rho=1.0, c=1.0):
        e = Modes()
        e.n_modes = len(modes)
        for k, mm in enumerate(modes):
            e.amp[k] = 1.0 if amps is None else float(amps[k])
            for d in range(dim):
                e.m[k][d] = int(mm[d])
        e.rho, e.c = float(rho), float(c)
        return e

    @staticmethod
    def reference(dim, membrane_modes):
        """ExactSolution of include/input_parameters.h:94-128."""
        return Modes.make(dim, [[membrane_modes] * dim])


class Epilogue(C.Structure):
    _fields_ = [("b", C.c_void_p), ("o1", C.c_void_p), ("a1", C.c_double), ("c", C.c_void_p),
                ("o2", C.c_void_p), ("a2", C.c_double), ("o2_from_o1", C.c_int),
                ("cell_scale", C.c_void_p)]


class Source(C.Structure):
    """exw_source: Ricker point source of pressure at x (t0 defaults to 1/f0)."""
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


LTS_SPEED_AWARE = 1


class Request(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("n_refinements", C.c_int),
                ("rho", C.c_double), ("c", C.c_double), ("modes", Modes),
                ("final_time", C.c_double), ("cfl_number", C.c_double), ("precision", C.c_int)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("n_steps", C.c_int), ("time", C.c_double),
                ("err_p", C.c_double), ("err_v", C.c_double), ("mag_p", C.c_double),
                ("precision", C.c_int)]


class Problem(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("n_intervals", C.c_int),
                ("n_refinements", C.c_int), ("transform", C.c_int),
                ("transform_factor", C.c_double), ("boundary_id", C.c_int),
                ("integrator", C.c_int), ("cfl_number", C.c_double),
                ("max_time_steps", C.c_int), ("final_time", C.c_double),
                ("output_every_time", C.c_double), ("ader_recon", C.c_int),
                ("precision", C.c_int), ("modes", Modes),
                ("lts_max_clusters", C.c_int), ("lts_max_diff", C.c_int),
                ("post_pressure", C.c_int), ("n_adaptive_refinements", C.c_int),
                ("adaptive_refinement_interval", C.c_int)]


# (name, restype, argtypes) of every exported entry point
_VP, _D, _I, _DP, _IP = C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)
SIGNATURES = [
    ("exw_last_error", C.c_char_p, []),
    ("exw_version", C.c_char_p, []),
    ("exw_ctx_create", _I, [_I, _VP, C.POINTER(_VP)]),
    ("exw_ctx_set_stream", _I, [_VP, _VP]),
    ("exw_ctx_destroy", _I, [_VP]),
    ("exw_sync", _I, [_VP]),
    ("exw_alloc", _I, [_VP, C.c_size_t, C.POINTER(_VP)]),
    ("exw_free", _I, [_VP, _VP]),
    ("exw_memcpy_h2d", _I, [_VP, _VP, _VP, C.c_size_t]),
    ("exw_memcpy_d2h", _I, [_VP, _VP, _VP, C.c_size_t]),
    ("exw_op_create", _I, [_VP, _I, _I, _I, _DP, _IP, _DP, _DP, C.POINTER(_VP)]),
    ("exw_op_destroy", _I, [_VP]),
    ("exw_op_n_dofs", C.c_int64, [_VP]),
    ("exw_op_n_cells", _I, [_VP]),
    ("exw_op_is_cartesian", _I, [_VP]),
    ("exw_op_min_vertex_distance_over_c", _D, [_VP]),
    ("exw_op_apply", _I, [_VP, _VP, _VP, _I]),
    ("exw_op_apply_fused", _I, [_VP, _VP, C.POINTER(Epilogue), _I]),
    ("exw_rk_stage", _I, [_VP, _VP, _VP, _VP, _VP, _D, _D, _I, _I]),
    ("exw_rk_update", _I, [_VP, C.c_int64, _D, _D, _I, _VP, _VP, _VP]),
    ("exw_ader_ck", _I, [_VP, _VP, _VP, _D, _D, _D, _VP, _I]),
    ("exw_ader_step", _I, [_VP, _VP, _VP, _VP, _D, _I, _I]),
    ("exw_ader_steps", _I, [_VP, _I, _D, _VP, _VP, _I]),
    ("exw_time_step", _I, [_VP, _I, _D, _VP, _VP, _VP, _I, _I]),
    ("exw_time_step_work_vectors", _I, [_I]),
    ("exw_project", _I, [_VP, C.POINTER(Modes), _D, _VP]),
    ("exw_l2_errors", _I, [_VP, C.POINTER(Modes), _D, _VP, _DP]),
    ("exw_l2_errors_noalloc", _I, [_VP, C.POINTER(Modes), _D, _VP, _VP]),
    ("exw_energy", _I, [_VP, _VP, _DP]),
    ("exw_post_n_values", C.c_int64, [_VP]),
    ("exw_estimate_error", _I, [_VP, _VP, _VP, _VP]),
    ("exw_refinement_flags", _I, [_I, _DP, _IP, _I, _I, _D, _IP]),
    ("exw_adapt_box_refine", _I, [_I, _I, _IP, _IP, _IP]),
    ("exw_adapt_box_transfer", _I, [_VP, _I, _I, _I, _IP, _IP, _VP, _VP]),
    ("exw_write_vtu", _I, [_I, _I, _I, _DP, _DP, _DP, _DP, _DP, _DP, _DP, C.c_char_p]),
    ("exw_op_create_nc", _I, [_VP, _I, _I, _I, _DP, _IP, _I, _IP, _DP, _DP, C.POINTER(_VP)]),
    ("exw_mesh_check", _I, [_I, _I, _DP, _IP, _I, _IP]),
    ("exw_mesh_box_refined", _I, [_I, _I, _I, _D, _I, _IP, _IP, _DP, _IP, _IP, _IP]),
    ("exw_mesh_forest", _I, [_I, _I, _I, _D, _I, _I, _IP, _IP, _DP, _IP, _IP, _IP, _IP]),
    ("exw_adapt_forest", _I, [_I, _I, _I, _IP, _IP, _IP, _IP, _I]),
    ("exw_adapt_forest_transfer", _I, [_VP, _I, _I, _I, _I, _IP, _I, _IP, _VP, _VP]),
    ("exw_post_pressure", _I, [_VP, _VP, _VP, _VP, _VP]),
    ("exw_post_error", _I, [_VP, C.POINTER(Modes), _D, _VP, _DP]),
    ("exw_mesh_hypercube", _I, [_I, _I, _I, _I, _D, _I, _IP, _DP, _IP]),
    ("exw_run", _I, [_VP, C.POINTER(Problem), C.c_char_p, _I, _DP]),
    ("exw_server_create", _I, [_VP, C.c_size_t, C.POINTER(_VP)]),
    ("exw_server_destroy", _I, [_VP]),
    ("exw_server_submit", _I, [_VP, C.POINTER(Request), C.POINTER(C.c_int64)]),
    ("exw_server_progress", _I, [_VP, _I]),
    ("exw_server_poll", _I, [_VP, C.POINTER(Result), _I]),
    ("exw_server_status", _I, [_VP, _IP, _IP, _IP, C.POINTER(C.c_size_t)]),
    ("exw_solve", _I, [_VP, C.POINTER(Request), C.POINTER(Result)]),
    ("exw_request_initial_error", _I, [C.POINTER(Request), _DP, _DP, _DP, _IP]),
    ("exw_lts_categorize", _I, [_I, _I, _DP, _IP, _D, _I, _I, _IP, _IP, _IP, _IP, _IP]),
    ("exw_lts_create", _I, [_VP, _D, _D, _I, _I, _D, _I, C.POINTER(_VP)]),
    ("exw_lts_create2", _I, [_VP, _D, _D, _I, _I, _D, _I, _I, C.POINTER(_VP)]),
    ("exw_op_set_source", _I, [_VP, C.POINTER(Source)]),
    ("exw_source_add", _I, [_VP, _D, _D, _VP]),
    ("exw_lts_destroy", _I, [_VP]),
    ("exw_lts_info", _I, [_VP, _IP, _IP, _IP, _DP]),
    ("exw_lts_cells", _I, [_VP, _IP, _IP]),
    ("exw_lts_step", _I, [_VP, _D, _VP, _VP]),
]

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m exwave_amd.build` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc < 0:
        raise ExwError(rc, load().exw_last_error().decode())
    return rc


NC_FACE = -1000  # EXW_NC_FACE: neighbour entry of a hanging (2:1) face


def write_vtu(filename, dim, degree, vertices, solution, error=None, error_estimate=None,
              post_pressure=None, cluster_id=None, time_step=None):
    """exw_write_vtu from host arrays (numpy)."""
    import numpy as np

    def p(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        return a.ctypes.data_as(C.POINTER(C.c_double))
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    keep = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
            for x in (solution, error, error_estimate, post_pressure, cluster_id, time_step)]
    check(load().exw_write_vtu(dim, degree, V.shape[0], p(V), *[p(x) for x in keep],
                               str(filename).encode()))


def mesh_box_refined(dim, n_intervals, refine, transform=TRANSFORM_NONE, factor=0.0,
                     boundary_id=2):
    """exw_mesh_box_refined: (vertices, neighbor, hanging records) of the box
    with the flagged base cells (lexicographic, xi_0 fastest) refined once."""
    import numpy as np
    L = load()
    r = np.ascontiguousarray(refine, dtype=np.int32).ravel()
    ip = C.POINTER(C.c_int)
    nc, nh = C.c_int(0), C.c_int(0)
    check(L.exw_mesh_box_refined(dim, n_intervals, transform, factor, boundary_id,
                                 r.ctypes.data_as(ip), C.byref(nc), None, None, C.byref(nh),
                                 None))
    V = np.zeros((nc.value, 1 << dim, dim))
    nb = np.zeros((nc.value, 2 * dim), dtype=np.int32)
    H = np.zeros((max(nh.value, 1), 2 + (1 << (dim - 1))), dtype=np.int32)
    check(L.exw_mesh_box_refined(dim, n_intervals, transform, factor, boundary_id,
                                 r.ctypes.data_as(ip), C.byref(nc),
                                 V.ctypes.data_as(C.POINTER(C.c_double)), nb.ctypes.data_as(ip),
                                 C.byref(nh), H.ctypes.data_as(ip)))
    return V, nb, H[: nh.value]


def mesh_forest(dim, n_intervals, tree, transform=TRANSFORM_NONE, factor=0.0, boundary_id=2):
    """exw_mesh_forest: (vertices, neighbor, level, hanging records) of the box
    refined by the depth-first trees in `tree` (1 refined, 0 active)."""
    import numpy as np
    L = load()
    t = np.ascontiguousarray(tree, dtype=np.int32).ravel()
    ip = C.POINTER(C.c_int)
    nc, nh = C.c_int(0), C.c_int(0)
    check(L.exw_mesh_forest(dim, n_intervals, transform, factor, boundary_id, t.size,
                            t.ctypes.data_as(ip), C.byref(nc), None, None, None, C.byref(nh),
                            None))
    V = np.zeros((nc.value, 1 << dim, dim))
    nb = np.zeros((nc.value, 2 * dim), dtype=np.int32)
    lev = np.zeros(nc.value, dtype=np.int32)
    H = np.zeros((max(nh.value, 1), 2 + (1 << (dim - 1))), dtype=np.int32)
    check(L.exw_mesh_forest(dim, n_intervals, transform, factor, boundary_id, t.size,
                            t.ctypes.data_as(ip), C.byref(nc),
                            V.ctypes.data_as(C.POINTER(C.c_double)), nb.ctypes.data_as(ip),
                            lev.ctypes.data_as(ip), C.byref(nh), H.ctypes.data_as(ip)))
    return V, nb, lev, H[: nh.value]


def adapt_forest(dim, n_intervals, tree, flags):
    """exw_adapt_forest: the depth-first trees after adapt_mesh's flags."""
    import numpy as np
    L = load()
    t = np.ascontiguousarray(tree, dtype=np.int32).ravel()
    f = np.ascontiguousarray(flags, dtype=np.int32).ravel()
    ip = C.POINTER(C.c_int)
    n = C.c_int(0)
    check(L.exw_adapt_forest(dim, n_intervals, t.size, t.ctypes.data_as(ip),
                             f.ctypes.data_as(ip), C.byref(n), None, 0))
    out = np.zeros(n.value, dtype=np.int32)
    check(L.exw_adapt_forest(dim, n_intervals, t.size, t.ctypes.data_as(ip),
                             f.ctypes.data_as(ip), C.byref(n), out.ctypes.data_as(ip), out.size))
    return out


def mesh_hypercube(dim, n_intervals, n_refinements=0, transform=TRANSFORM_NONE, factor=0.0,
                   boundary_id=2):
    """input_geometry_description: returns (vertices[n,2^dim,dim], neighbor[n,2*dim])."""
    import numpy as np
    L = load()
    n = C.c_int(0)
    check(L.exw_mesh_hypercube(dim, n_intervals, n_refinements, transform, factor, boundary_id,
                               C.byref(n), None, None))
    V = np.zeros((n.value, 1 << dim, dim))
    nb = np.zeros((n.value, 2 * dim), dtype=np.int32)
    check(L.exw_mesh_hypercube(dim, n_intervals, n_refinements, transform, factor, boundary_id,
                               C.byref(n), V.ctypes.data_as(_DP), nb.ctypes.data_as(_IP)))
    return V, nb


def tile_order(V, nb, n1d, tile, blocks=1):
    """Renumber the cells of the lexicographic n1d^3 mesh of mesh_hypercube for
    locality, as MatrixFree renumbers cells: tiles of tile^3 cells (tiles in
    lexicographic order, cells lexicographic inside a tile), so that the z
    neighbours of a cell are tile^2 cells away instead of n1d^2.  With
    blocks = b > 1 the mesh is first split into b^3 blocks of (n1d/b)^3 cells
    numbered one block after the other (tiles inside each): with b^3 equal to
    the GPU's 8 XCDs, the contiguous cell range each XCD sweeps is one compact
    block, whose neighbours mostly sit in the same L2.  Returns (vertices,
    neighbor) in the new order."""
    import numpy as np
    idx = np.arange(n1d ** 3)
    x, y, z = idx % n1d, (idx // n1d) % n1d, idx // (n1d * n1d)
    nb1 = n1d // blocks if blocks > 1 and n1d % blocks == 0 else n1d
    nbk = n1d // nb1
    bx, by, bz = x // nb1, y // nb1, z // nb1
    lx, ly, lz = x % nb1, y % nb1, z % nb1
    nt = (nb1 + tile - 1) // tile
    tx, ty, tz = lx // tile, ly // tile, lz // tile
    key = ((((bz * nbk + by) * nbk + bx) * nt + tz) * nt + ty) * nt + tx
    key = ((key * tile + lz % tile) * tile + ly % tile) * tile + lx % tile
    order = np.argsort(key, kind="stable")  # new -> old
    new_of_old = np.empty_like(order)
    new_of_old[order] = np.arange(order.size)
    V2 = V[order]
    nb2 = nb[order].copy()
    m = nb2 >= 0
    nb2[m] = new_of_old[nb2[m]]
    return V2, nb2
