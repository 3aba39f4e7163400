"""exwave_amd: MI355X-native matrix-free DG acoustic operator and explicit time
stepping (the hot path of kronbichler/exwave) behind a C-ABI.

Python API (thin, over the C-ABI of include/exwave_amd.h):

    ctx = Context()                          # device 0, torch's current stream
    op  = WaveOperator(ctx, dim, degree, vertices, neighbor, rho, speed)
    op.apply(src, dst)                       # dst = M^-1 A src  (torch float64 tensors)
    op.time_step(LSRK45R2, dt, vec_n, vec_np, work)   # ExplicitIntegrator::perform_time_step

The C++ host mirror of the reference interface (WaveEquationOperationBase,
ExplicitIntegrator, LowStorageRK45Reg2, ...) lives in include/exwave_amd.hpp.
"""
from __future__ import annotations

import ctypes as C

from .capi import (ADER, ADER_ADCONFULL, ADER_LTS, CLRK4, EXPL_EULER, FP32, FP64,  # noqa
                   PRECISION_AUTO, INTEGRATORS, LSRK33R2, LSRK45R2, LSRK45R3, SSPRK,
                   LSRK59R2, TRANSFORM_AXIS_2PI, TRANSFORM_NONE, TRANSFORM_PRODUCT_2PI,
                   TRANSFORM_REFERENCE_SINE, NC_FACE, Epilogue, ExwError, Modes, Problem, check,
                   load, mesh_box_refined, mesh_forest, adapt_forest, mesh_hypercube,
                   tile_order, write_vtu, Source,
                   LTS_SPEED_AWARE)


def _ptr(t):
    if t is None:
        return None
    import torch
    assert isinstance(t, torch.Tensor) and t.dtype == torch.float64 and t.is_cuda, \
        "device vectors are float64 torch tensors on the GPU"
    assert t.is_contiguous()
    return C.c_void_p(t.data_ptr())


class Context:
    """exw_ctx bound to a device and (by default) torch's current stream."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        L = load()
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = C.c_void_p()
        check(L.exw_ctx_create(device, C.c_void_p(stream), C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        check(load().exw_sync(self.h))

    def __del__(self):
        if getattr(self, "h", None) is not None:
            try:
                load().exw_ctx_destroy(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None


class WaveOperator:
    """WaveEquationOperation<dim,k> on the device (setup + apply + integrators)."""

    def __init__(self, ctx: Context, dim: int, degree: int, vertices, neighbor, rho, speed,
                 hanging=None):
        """hanging: records of the 2:1 hanging faces (exw_op_create_nc), or None."""
        import numpy as np
        L = load()
        V = np.ascontiguousarray(vertices, dtype=np.float64)
        nb = np.ascontiguousarray(neighbor, dtype=np.int32)
        r = np.ascontiguousarray(rho, dtype=np.float64)
        c = np.ascontiguousarray(speed, dtype=np.float64)
        n_cells = V.shape[0]
        assert V.shape == (n_cells, 1 << dim, dim) and nb.shape == (n_cells, 2 * dim)
        assert r.shape == (n_cells,) and c.shape == (n_cells,)
        h = C.c_void_p()
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        if hanging is not None and len(hanging) > 0:
            H = np.ascontiguousarray(hanging, dtype=np.int32)
            assert H.shape[1] == 2 + (1 << (dim - 1))
            check(L.exw_op_create_nc(ctx.h, dim, degree, n_cells, V.ctypes.data_as(dp),
                                     nb.ctypes.data_as(ip), H.shape[0], H.ctypes.data_as(ip),
                                     r.ctypes.data_as(dp), c.ctypes.data_as(dp), C.byref(h)))
        else:
            check(L.exw_op_create(ctx.h, dim, degree, n_cells, V.ctypes.data_as(dp),
                                  nb.ctypes.data_as(ip), r.ctypes.data_as(dp),
                                  c.ctypes.data_as(dp), C.byref(h)))
        self.h, self.ctx, self.dim, self.degree, self.n_cells = h, ctx, dim, degree, n_cells
        self.n_dofs = int(L.exw_op_n_dofs(h))
        self.is_cartesian = bool(L.exw_op_is_cartesian(h))
        self.min_vertex_distance_over_c = float(L.exw_op_min_vertex_distance_over_c(h))

    @classmethod
    def hypercube(cls, ctx, dim, degree, n_intervals, n_refinements=0,
                  transform=TRANSFORM_NONE, factor=0.0, boundary_id=2, rho=1.0, speed=1.0):
        import numpy as np
        V, nb = mesh_hypercube(dim, n_intervals, n_refinements, transform, factor, boundary_id)
        n = V.shape[0]
        r = np.broadcast_to(np.asarray(rho, dtype=np.float64), (n,)).copy()
        c = np.broadcast_to(np.asarray(speed, dtype=np.float64), (n,)).copy()
        op = cls(ctx, dim, degree, V, nb, r, c)
        op.vertices, op.neighbor, op.rho, op.speed = V, nb, r, c
        return op

    def __del__(self):
        if getattr(self, "h", None) is not None:
            try:
                load().exw_op_destroy(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def new_vector(self, n=1):
        import torch
        return torch.zeros(n * self.n_dofs, dtype=torch.float64, device=f"cuda:{self.ctx.device}")

    def apply(self, src, dst, precision=FP64):
        check(load().exw_op_apply(self.h, _ptr(src), _ptr(dst), precision))

    def apply_fused(self, src, b=None, o1=None, a1=1.0, c=None, o2=None, a2=0.0,
                    o2_from_o1=False, cell_scale=None, precision=FP64):
        e = Epilogue(_ptr(b), _ptr(o1), a1, _ptr(c), _ptr(o2), a2, int(o2_from_o1),
                     _ptr(cell_scale))
        check(load().exw_op_apply_fused(self.h, _ptr(src), C.byref(e), precision))

    def rk_stage(self, src, b, y_out, b_out, f1, f2, is_last, precision=FP64):
        check(load().exw_rk_stage(self.h, _ptr(src), _ptr(b), _ptr(y_out), _ptr(b_out), f1, f2,
                                  int(is_last), precision))

    def ader_ck(self, src, recon, t2, t1, te, out, precision=FP64):
        check(load().exw_ader_ck(self.h, _ptr(src), _ptr(recon), t2, t1, te, _ptr(out),
                                 precision))

    def time_step(self, integrator, dt, vec_n, vec_np, work=None, ader_recon=False,
                  precision=FP64):
        check(load().exw_time_step(self.h, integrator, dt, _ptr(vec_n), _ptr(vec_np), _ptr(work),
                                   int(ader_recon), precision))

    def ader_steps(self, n_steps, dt, vec, work, precision=FP64):
        """exw_ader_steps: n_steps global ADER steps on vec in place (work: 2 vectors)."""
        check(load().exw_ader_steps(self.h, int(n_steps), dt, _ptr(vec), _ptr(work), precision))

    @staticmethod
    def work_vectors(integrator):
        return int(load().exw_time_step_work_vectors(integrator))

    def set_source(self, src):
        """exw_op_set_source: a Source, or None to remove it."""
        self._src = src
        check(load().exw_op_set_source(self.h, C.byref(src) if src is not None else None))

    def source_add(self, t1, t2, u):
        check(load().exw_source_add(self.h, t1, t2, _ptr(u)))

    def project(self, modes: Modes, t, dst):
        check(load().exw_project(self.h, C.byref(modes), t, _ptr(dst)))

    def l2_errors(self, modes: Modes, t, u):
        out = (C.c_double * 3)()
        check(load().exw_l2_errors(self.h, C.byref(modes), t, _ptr(u), out))
        return list(out)

    def energy(self, u):
        out = C.c_double(0.0)
        check(load().exw_energy(self.h, _ptr(u), C.byref(out)))
        return out.value

    def estimate_error(self, solution, tmp, error_per_cell):
        """estimate_error: tmp <- apply(solution), error_per_cell (n_cells doubles on
        the device) <- sqrt(int_K |grad p - v_t|^2)."""
        check(load().exw_estimate_error(self.h, _ptr(solution), _ptr(tmp),
                                        _ptr(error_per_cell)))

    @property
    def n_post(self):
        return load().exw_post_n_values(self.h)

    def post_pressure(self, solution, tmp, post, iterations=None):
        """compute_post_pressure: tmp <- apply(solution), post <- p* (n_post values,
        (k+2)^dim Gauss-Lobatto nodal values per cell); iterations: optional
        int32 device tensor of n_cells CG iteration counts."""
        it = None
        if iterations is not None:
            import torch
            assert iterations.dtype == torch.int32 and iterations.is_cuda
            assert iterations.is_contiguous() and iterations.numel() >= self.n_cells
            it = C.c_void_p(iterations.data_ptr())
        check(load().exw_post_pressure(self.h, _ptr(solution), _ptr(tmp), _ptr(post), it))

    def post_error(self, modes: Modes, t, post):
        out = C.c_double(0.0)
        check(load().exw_post_error(self.h, C.byref(modes), t, _ptr(post), C.byref(out)))
        return out.value


def refinement_flags(error_per_cell, level, min_level, max_level, max_error_init):
    """exw_refinement_flags: adapt_mesh's marking; 1 refine, -1 coarsen, 0 keep."""
    import numpy as np
    e = np.ascontiguousarray(error_per_cell, dtype=np.float64)
    lv = np.ascontiguousarray(level, dtype=np.int32)
    fl = np.zeros(e.size, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    check(load().exw_refinement_flags(e.size, e.ctypes.data_as(C.POINTER(C.c_double)),
                                      lv.ctypes.data_as(ip), min_level, max_level,
                                      max_error_init, fl.ctypes.data_as(ip)))
    return fl


def adapt_box_refine(dim, n_intervals, refine_old, flags):
    """exw_adapt_box_refine: the new base-cell refinement after the marking."""
    import numpy as np
    ro = np.ascontiguousarray(refine_old, dtype=np.int32).ravel()
    fl = np.ascontiguousarray(flags, dtype=np.int32)
    rn = np.zeros_like(ro)
    ip = C.POINTER(C.c_int)
    check(load().exw_adapt_box_refine(dim, n_intervals, ro.ctypes.data_as(ip),
                                      fl.ctypes.data_as(ip), rn.ctypes.data_as(ip)))
    return rn


def adapt_box_transfer(ctx, dim, degree, n_intervals, refine_old, refine_new, old, new):
    """exw_adapt_box_transfer between two exw_mesh_box_refined meshes (device vectors)."""
    import numpy as np
    ro = np.ascontiguousarray(refine_old, dtype=np.int32).ravel()
    rn = np.ascontiguousarray(refine_new, dtype=np.int32).ravel()
    ip = C.POINTER(C.c_int)
    check(load().exw_adapt_box_transfer(ctx.h, dim, degree, n_intervals, ro.ctypes.data_as(ip),
                                        rn.ctypes.data_as(ip), _ptr(old), _ptr(new)))


def adapt_forest_transfer(ctx, dim, degree, n_intervals, tree_old, tree_new, old, new):
    """exw_adapt_forest_transfer between two exw_mesh_forest meshes (device vectors)."""
    import numpy as np
    to = np.ascontiguousarray(tree_old, dtype=np.int32).ravel()
    tn = np.ascontiguousarray(tree_new, dtype=np.int32).ravel()
    ip = C.POINTER(C.c_int)
    check(load().exw_adapt_forest_transfer(ctx.h, dim, degree, n_intervals, to.size,
                                           to.ctypes.data_as(ip), tn.size, tn.ctypes.data_as(ip),
                                           _ptr(old), _ptr(new)))


def lts_categorize(dim, vertices, neighbor, cfl, max_clusters, max_diff):
    """exw_lts_categorize (host only): (cluster ids, flags, n_clusters,
    cluster_diff, iterations); flags 1 = slower neighbour, 2 = faster."""
    import numpy as np
    V = np.ascontiguousarray(vertices, dtype=np.float64)
    nb = np.ascontiguousarray(neighbor, dtype=np.int32)
    n = V.shape[0]
    cl = np.zeros(n, np.int32)
    fl = np.zeros(n, np.int32)
    nc, diff, it = C.c_int(), C.c_int(), C.c_int()
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    check(load().exw_lts_categorize(dim, n, V.ctypes.data_as(dp), nb.ctypes.data_as(ip), cfl,
                                    max_clusters, max_diff, cl.ctypes.data_as(ip),
                                    fl.ctypes.data_as(ip), C.byref(nc), C.byref(diff),
                                    C.byref(it)))
    return cl, fl, nc.value, diff.value, it.value


class LocalTimeStepping:
    """ADER with local time stepping on a WaveOperator (exw_lts_*)."""

    def __init__(self, op: WaveOperator, cfl, fastest_dt, max_clusters, max_diff, time0=0.0,
                 recon=True, speed_aware=False):
        h = C.c_void_p()
        check(load().exw_lts_create2(op.h, cfl, fastest_dt, max_clusters, max_diff, time0,
                                     int(recon), LTS_SPEED_AWARE if speed_aware else 0,
                                     C.byref(h)))
        self.h, self.op = h, op
        nc, diff, nu, gdt = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(load().exw_lts_info(h, C.byref(nc), C.byref(diff), C.byref(nu), C.byref(gdt)))
        self.n_clusters, self.cluster_diff = nc.value, diff.value
        self.n_updates, self.global_dt = nu.value, gdt.value

    def cells(self):
        import numpy as np
        cl = np.zeros(self.op.n_cells, np.int32)
        fl = np.zeros(self.op.n_cells, np.int32)
        ip = C.POINTER(C.c_int)
        check(load().exw_lts_cells(self.h, cl.ctypes.data_as(ip), fl.ctypes.data_as(ip)))
        return cl, fl

    def step(self, time_now, src, dst):
        check(load().exw_lts_step(self.h, time_now, _ptr(src), _ptr(dst)))

    def __del__(self):
        if getattr(self, "h", None) is not None:
            try:
                load().exw_lts_destroy(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None


def run(ctx: Context, problem: Problem):
    """exw_run: returns (n_steps, [(step, time, err_p, err_v, mag_p)], wall_seconds);
    with problem.post_pressure set the rows carry a sixth entry, error p post."""
    buf = C.create_string_buffer(1 << 16)
    wall = C.c_double(0.0)
    steps = check(load().exw_run(ctx.h, C.byref(problem), buf, len(buf), C.byref(wall)))
    rows = []
    for line in buf.value.decode().strip().splitlines():
        s, t, ep, ev, mp, epp = line.split()
        row = (int(s), float(t), float(ep), float(ev), float(mp))
        rows.append(row + (float(epp),) if problem.post_pressure else row)
    return steps, rows, wall.value
