#!/usr/bin/env python3
"""Headline benchmark: GDoF/s per RK stage of the fused FP64 DG operator + LSRK45
stage on one MI355X.

Workload (the reference's 3D membrane, scaled to one GPU's share): unit cube,
64^3 Cartesian hexes, degree k=4 (125 nodes/cell, fields v_x, v_y, v_z, p:
131,072,000 DoFs), rho = c = 1, boundary id 2 (p = 0), initial pressure
sin(pi x)sin(pi y)sin(pi z) + sin(2 pi x)sin(3 pi y)sin(pi z), v = 0,
LowStorageRK45Reg2 at CFL 0.2 (dt = 0.2/4^1.5 * 1/64).  A bench "step" is one
LSRK45 time step = 5 fused operator+update sweeps.  The cells are numbered in
8^3 tiles (--tile; the mesh and the results are the same as with the
lexicographic numbering, only the order of the cells in memory differs, so
that face neighbours share the L2 of an XCD).

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]

Prints one JSON line (see the keys below).  `roofline` is computed for the
fused stage kernel from algorithmic bytes (28.8 B/DoF per stage on average:
18 FP64 vector passes per LSRK45 step, plus 64 B of per-cell data) over its
average launch time measured with HIP events on the launch stream; `traffic`
is the measured HBM bytes per launch from the committed rocprofv3 counter
summary of this workload (profiles/), when present.  `cpu_baseline` times the
CPU oracle (a C/OpenMP restatement, not deal.II) on a bounded sample of the
same workload.

`serving` (unless --no-serving): the 512-request serving stream (seeded: dim
2/3, k 2..6, refinements 3..6 / 2..4, rho and c log-uniform on [0.5, 2],
modes 1..4, T in [0.5, 2], Poisson arrivals at 200/s), replayed in real time
through the continuously batched serving loop with the automatic FP32/FP64
choice, against the plain FP64 one-solve-at-a-time path (exw_solve per
request) in the same run: solves/s, p50/p99 latency (completion - arrival),
and how many results stay within 1.1x of the one-at-a-time FP64 L2 errors.

    python bench.py --copy-calibration    # plain copy kernel of known size,
                                          # for scaling FETCH_SIZE/WRITE_SIZE
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (the FP64 matrix peak is the same)
N_CELLS_1D, DEGREE, DIM = 64, 4, 3
STAGES = 5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--n", type=int, default=N_CELLS_1D, help="cells per direction")
    ap.add_argument("--degree", type=int, default=DEGREE)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle sample")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--tile", type=int, default=8,
                    help="number the cells in tiles of tile^3 (0: lexicographic)")
    ap.add_argument("--blocks", type=int, default=1,
                    help="with --tile: blocks^3 blocks numbered one after the other")
    ap.add_argument("--serving-cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--nonaffine", action="store_true",
                    help="non-affine 32^3/48^3, k=4/6, FP32 vs FP64 operator (errors + GDoF/s)")
    ap.add_argument("--nonaffine-steps", type=int, default=40)
    ap.add_argument("--nonaffine-case", action="append", default=None,
                    help="degree,cells_per_dir,map (axis|product): only this case (repeatable)")
    ap.add_argument("--ader", action="store_true",
                    help="the CK kernel and the ADER step on the membrane workloads")
    ap.add_argument("--ader-case", action="append", default=None,
                    help="dim,degree,cells_per_dir: only this --ader case (repeatable)")
    ap.add_argument("--refined", action="store_true",
                    help="the two-material box with a locally refined half (hanging faces)")
    ap.add_argument("--refined-time", type=float, default=0.5)
    return ap.parse_args()


TRAFFIC_FILES = [os.path.join(HERE, "profiles", f"dg_operator_{kind}_3d_k4_64cube_counters.md")
                 for kind in ("pipe", "cart")]


def measured_traffic(kernel_key):
    """HBM bytes per launch of `kernel_key` and the committed counter summary
    that holds it (a line 'traffic_bytes_per_launch <kernel_key> <bytes>')."""
    for path in TRAFFIC_FILES:
        try:
            for line in open(path):
                f = line.split()
                if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel_key:
                    return int(float(f[2])), path
        except OSError:
            pass
    return None, None


def stage_kernel(k, fp64):
    """(family, name) of the kernel that runs the fused LSRK stage on Cartesian
    3D cells: the pipelined FP64 kernel for k = 3..5 (unless EXW_CART_PIPE=0),
    the gathered-trace kernel for k = 6..9 (unless EXW_CART_GATHER=0), the
    big-cell kernel from k = 10, else dg_operator_cart_kernel."""
    prec = "double" if fp64 else "float"
    if fp64 and 3 <= k <= 5 and os.environ.get("EXW_CART_PIPE", "1")[:1] != "0":
        return "pipe", f"dg_operator_pipe_kernel<3,{k},2>"
    if 6 <= k <= 9 and os.environ.get("EXW_CART_GATHER", "1")[:1] != "0":
        return "cartg", f"dg_operator_cartg_kernel<3,{k},{prec}>"
    if k >= 10:
        return "big", f"dg_operator_big_kernel<3,{k},{prec}>"
    return "cart", f"dg_operator_cart_kernel<3,{k},{prec}>"


def stream_ceiling(reps=10, ctx=None):
    """Measured HBM stream rates on this GPU (1 GiB FP64 vectors, HIP events on
    the launch stream, best of `reps`): torch's copy (1 read + 1 write) and add
    (2 reads + 1 write), and, with `ctx`, the library's unfused RKVectorUpdater
    (exw_rk_update: reads a, b, writes b, c -- the fused stage's own 2 + 2
    pattern without the operator).  The best torch rate is the practical
    ceiling of a bandwidth-bound kernel here; the updater's is the ceiling of
    the stage's access pattern."""
    import torch
    n = 1 << 27
    stream = torch.cuda.current_stream()
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.rand(n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)

    def best(fn, nbytes):
        fn()
        rates = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            rates.append(nbytes / (e0.elapsed_time(e1) / 1e3) / 1e9)
        return round(max(rates), 1)
    out = {"copy_gbs": best(lambda: c.copy_(a), 16 * n),
           "add_gbs": best(lambda: torch.add(a, b, alpha=0.5, out=c), 24 * n)}
    out["stream_ceiling_gbs"] = max(out["copy_gbs"], out["add_gbs"])
    out["method"] = "torch copy_ / add on 1 GiB FP64 vectors, best of %d, HIP events" % reps
    if ctx is not None:
        import exwave_amd as X
        L = X.load()

        def upd():
            X.check(L.exw_rk_update(ctx.h, n, 1e-3, 2e-3, 0, C.c_void_p(a.data_ptr()),
                                    C.c_void_p(b.data_ptr()), C.c_void_p(c.data_ptr())))
        out["stage_pattern_gbs"] = best(upd, 32 * n)
        out["stage_pattern_method"] = ("exw_rk_update (RKVectorUpdater, not last: reads 2 and "
                                       "writes 2 FP64 vectors, the fused stage's pattern)")
    del a, b, c
    torch.cuda.empty_cache()
    return out


def stage_flops_per_cell(dim, k, two_outputs=True):
    """Floating-point operations of one fused Cartesian stage per cell, counted
    from the kernel's arithmetic (FMA = 2): per line of the dim*N^(dim-1) lines,
    the collocation and weak derivatives (2 N^2 FMA), both face traces of own
    and neighbour (4 N FMA per side), the flux maps (2 x 7 per side), the
    lifting (2 N FMA per side) and the 1/h scalings (2 N); per node the
    pressure-share sum (dim - 1) and the epilogue (an FMA per output)."""
    N = k + 1
    per_line = 4 * N * N + 2 * (8 * N + 14 + 4 * N) + 2 * N
    nodes = N ** dim
    outputs = 2 if two_outputs else 1
    return dim * N ** (dim - 1) * per_line + nodes * ((dim - 1) + (dim + 1) * 2 * outputs)


def copy_calibration():
    """A plain device copy of 2^27 doubles (1 GiB read, 1 GiB write), five
    times, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": 8 * n, "bytes_written": 8 * n,
                                           "launches": 5}}), flush=True)


def nonaffine_measure(steps, cases=None, tile=8):
    """Non-affine 3D membranes at 32^3 and 48^3 cells, k = 4 and 6: LSRK45 with
    the FP64 operator and with the FP32 operator (FP64 vectors, RK updates and
    norms) from the same projected initial field.

    map "axis": x_i += 0.05 prod_{e != i} sin(2 pi x_e) (the literal node map;
      it moves the boundary, so no closed-form solution): FP32 vs FP64 relative
      L2 difference of the final states, and for the 32^3 k=4 case the FP64 and
      FP32 states against the CPU oracle after 2 steps;
    map "product": x_i += 0.05 prod_e sin(2 pi x_e) (keeps the unit cube):
      L2 errors of both against the analytical membrane solution.
    GDoF/s per RK stage of the general-geometry fused kernel for each."""
    import numpy as np
    import torch

    import exwave_amd as X
    ctx = X.Context(0)
    stream = torch.cuda.current_stream()
    modes = X.Modes.make(DIM, [[1, 1, 1], [2, 3, 1]])
    out = []
    maps = {"axis": X.TRANSFORM_AXIS_2PI, "product": X.TRANSFORM_PRODUCT_2PI}
    todo = cases or [(k, n1d, t) for k in (4, 6) for n1d in (32, 48) for t in ("axis", "product")]
    for k, n1d, tname in todo:
        tr = maps[tname]
        # cells in 8^3 tiles, as for the headline (a caller's renumbering for
        # locality; lexicographic order puts the z neighbours n1d^2 cells away)
        V, nb = X.mesh_hypercube(DIM, n1d, 0, tr, 0.05, 2)
        if tile > 1:
            V, nb = X.tile_order(V, nb, n1d, tile)
        ones = np.ones(V.shape[0])
        op = X.WaveOperator(ctx, DIM, k, V, nb, ones, ones)
        dt = 0.2 / k ** 1.5 * op.min_vertex_distance_over_c
        rec = {"degree": k, "cells_per_dir": n1d, "map": tname, "n_dofs": op.n_dofs,
               "steps": steps, "dt": dt,
               "cell_order": f"tiles of {tile}^3" if tile > 1 else "lexicographic"}
        finals = {}
        for pname, prec in (("fp64", X.FP64), ("fp32", X.FP32)):
            a, b, w = op.new_vector(), op.new_vector(), op.new_vector()
            op.project(modes, 0.0, b)
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(steps):
                a, b = b, a
                op.time_step(X.LSRK45R2, dt, a, b, w, precision=prec)
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1)
            rec[f"gdofs_{pname}"] = round(op.n_dofs * 5 * steps / (ms / 1e3) / 1e9, 2)
            # algorithmic bytes per stage: 28.8 B/DoF of FP64 vectors plus the
            # cell's 8 vertices, neighbour row and material (248 B)
            stage_bytes = 28.8 * op.n_dofs + 248 * op.n_cells
            gbs = stage_bytes * 5 * steps / (ms / 1e3) / 1e9
            rec[f"GBps_{pname}"] = round(gbs, 1)
            rec[f"frac_hbm_peak_{pname}"] = round(gbs / PEAK_HBM_GBS, 4)
            if tname == "product":
                e = op.l2_errors(modes, steps * dt, b)
                rec[f"err_p_{pname}"], rec[f"err_v_{pname}"] = e[0], e[1]
            finals[pname] = b
            del a, w
        d = finals["fp32"] - finals["fp64"]
        rec["fp32_vs_fp64_rel_l2"] = float(torch.linalg.vector_norm(d) /
                                           torch.linalg.vector_norm(finals["fp64"]))
        del finals, d
        torch.cuda.empty_cache()
        out.append(rec)
        print(json.dumps(rec), file=sys.stderr, flush=True)
    if cases:
        print(json.dumps({"nonaffine": out}), flush=True)
        return
    # oracle cross-check of the literal map at 32^3, k=4 (two steps)
    import oracle as O
    k, n1d, st = 4, 32, 2
    op = X.WaveOperator.hypercube(ctx, DIM, k, n1d, 0, X.TRANSFORM_AXIS_2PI, 0.05, 2)
    m = O.Mesh(DIM, n1d, 0, O.TRANSFORM_AXIS_2PI, 0.05, 2)
    dt = 0.2 / k ** 1.5 * m.min_vertex_distance_over_c
    e = O.Exact.make(DIM, [[1, 1, 1], [2, 3, 1]])
    u = m.project(k, e, 0.0)
    for _ in range(st):
        vnp = np.zeros_like(u)
        m.time_step(k, O.LSRK45R2, dt, u, vnp)
        u = vnp
    chk = {"case": "oracle check, map axis, 32^3, k=4, 2 steps"}
    for pname, prec in (("fp64", X.FP64), ("fp32", X.FP32)):
        a, b, w = op.new_vector(), op.new_vector(), op.new_vector()
        op.project(modes, 0.0, b)
        for _ in range(st):
            a, b = b, a
            op.time_step(X.LSRK45R2, dt, a, b, w, precision=prec)
        g = b.cpu().numpy()
        chk[f"{pname}_rel_max_vs_oracle"] = float(np.max(np.abs(g - u)) / np.max(np.abs(u)))
        chk[f"{pname}_rel_l2_vs_oracle"] = float(np.linalg.norm(g - u) / np.linalg.norm(u))
    print(json.dumps({"nonaffine": out, "oracle_check": chk}), flush=True)


def ck_flops_per_cell(dim, k):
    """Floating-point operations of the CK time integral per Cartesian cell,
    counted from the algorithm (FMA = 2): every derivative (the first and the
    max(k-1, 1) steps, on level l with n = k+1-2l points) is two line
    derivatives of N^2 FMA per line plus 3 dim + 1 node operations and the
    contribution update (an FMA per field and node); every projection and
    interpolation pass is one FMA per field, output and input point."""
    nsteps = max(k - 1, 1)

    def n_of(lev):
        return max(k + 1 - 2 * lev, 1)
    total = 0
    for lev in [0] + [s // 2 for s in range(nsteps)]:
        n = n_of(lev)
        total += 2 * 2 * dim * n ** (dim + 1) + n ** dim * (3 * dim + 1) + 2 * (dim + 1) * n ** dim

    def passes(n_in, n_out):  # dim tensor passes n_in -> n_out, outputs per pass
        f, sizes = 0, [n_in] * dim
        for d in range(dim):
            sizes[d] = n_out
            outs = 1
            for v in sizes:
                outs *= v
            f += 2 * (dim + 1) * outs * n_in
        return f
    for s in range(nsteps):
        if s % 2 == 1 and s < k - 2:
            n = n_of(s // 2)
            total += passes(n, n - 2)  # projection
            total += passes(n - 2, n)  # interpolation back
    return total


def ader_measure(cases=None):
    """The ADER step on the membrane workloads (Cartesian, rho = c = 1,
    Dirichlet walls): the Cauchy-Kowalevski kernel alone (exw_ader_ck: reads the
    state, writes the time-integrated field, 16 B/DoF algorithmic), the whole
    step (CK + the sweep with the update epilogue) one call at a time, and 20
    steps in one exw_ader_steps call (the update sweep fused with the next
    step's CK where that kernel exists), FP64 and FP32 operator."""
    import torch

    import exwave_amd as X
    ctx = X.Context(0)
    stream = torch.cuda.current_stream()
    out = []
    import numpy as np
    for dim, k, n1d in cases or ((3, 3, 64), (3, 4, 64), (3, 6, 32), (2, 5, 512)):
        # 3D: cells in 8^3 tiles, as for the headline (lexicographic order puts
        # the z neighbours n1d^2 cells away: the sweep then reads 1.5x its bytes)
        V, nb = X.mesh_hypercube(dim, n1d, 0, X.TRANSFORM_NONE, 0.0, 2)
        if dim == 3:
            V, nb = X.tile_order(V, nb, n1d, 8)
        ones = np.ones(V.shape[0])
        op = X.WaveOperator(ctx, dim, k, V, nb, ones, ones)
        modes = X.Modes.make(dim, [[1] * dim])
        u, t = op.new_vector(), op.new_vector()
        op.project(modes, 0.0, u)
        w = op.new_vector(max(op.work_vectors(X.ADER), 1))
        dt = 0.1 / k ** 1.5 * op.min_vertex_distance_over_c
        rec = {"dim": dim, "degree": k, "cells_per_dir": n1d, "n_dofs": op.n_dofs,
               "cell_order": "tiles of 8^3" if dim == 3 else "lexicographic"}
        for pname, prec in (("fp64", X.FP64), ("fp32", X.FP32)):
            for what in ("ck", "step"):
                def once():
                    if what == "ck":
                        op.ader_ck(u, None, dt, 0.0, 0.0, t, prec)
                    else:
                        op.time_step(X.ADER, dt, u, t, w, precision=prec)
                once()
                torch.cuda.synchronize()
                R = 20
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
                ev0.record(stream)
                for _ in range(R):
                    once()
                ev1.record(stream)
                torch.cuda.synchronize()
                ms = ev0.elapsed_time(ev1) / R
                rec[f"{what}_ms_{pname}"] = round(ms, 4)
                if what == "ck":
                    rec[f"ck_GBps_{pname}"] = round(16 * op.n_dofs / (ms / 1e3) / 1e9, 1)
                    n_cells = op.n_dofs // ((dim + 1) * (k + 1) ** dim)
                    rec[f"ck_TFLOPs_{pname}"] = round(
                        n_cells * ck_flops_per_cell(dim, k) / (ms / 1e3) / 1e12, 2)
                else:
                    rec[f"step_gdofs_{pname}"] = round(op.n_dofs / (ms / 1e3) / 1e9, 2)
            # R steps in one call (exw_ader_steps: on box meshes of 3D k = 3..5 each
            # update sweep runs fused with the next step's CK), per step
            v = u.clone()
            w2 = op.new_vector(2)
            op.ader_steps(1, dt, v, w2, prec)
            torch.cuda.synchronize()
            R = 20
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            op.ader_steps(R, dt, v, w2, prec)
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / R
            rec[f"steps_ms_{pname}"] = round(ms, 4)
            rec[f"steps_gdofs_{pname}"] = round(op.n_dofs / (ms / 1e3) / 1e9, 2)
            del v, w2
        out.append(rec)
        print(json.dumps(rec), file=sys.stderr, flush=True)
        del u, t, w, op
        torch.cuda.empty_cache()
    print(json.dumps({"ader": out}), flush=True)


def refined_measure(final_time, cfl_number=0.15):
    """The two-material box with a locally refined half as the workload states
    it: 32^3 base cells on [0,1]^3, the x < 1/2 half refined once (2:1 hanging
    faces at x = 1/2), rho = c = 1 for x < 1/2 and rho = 2, c = 4 beyond, k = 3,
    ADER of order 4 with local time stepping: cells clustered by h/c, so the
    coarse c = 4 half is the fast cluster and the refined half steps twice as
    long (2 clusters, step ratio 2); p = 0 walls; a Ricker pressure source at
    (0.25, 0.5, 0.5), f0 = 10, t0 = 1/f0, time-integrated into every update of
    its cells; from rest to final_time, FP64.  CFL 0.15 instead of 0.2: at 0.2
    this scheme's energy grows by orders of magnitude within 40 steps (shown on
    the CPU oracle by the test suite), at 0.15 it decays.  Reported: ms and
    GDoF/s per global LTS step, the same time span with global ADER at the
    fastest step (the gain of LTS), energy and pressure norm at T, and two LTS
    steps on the 8^3-base version of the mesh against the CPU oracle."""
    import numpy as np
    import torch

    import exwave_amd as X
    import oracle as O
    ctx = X.Context(0)
    stream = torch.cuda.current_stream()
    k, bid, x0 = 3, 2, [0.25, 0.5, 0.5]

    def build(n):
        ref = np.zeros((n, n, n), dtype=np.int32)
        ref[:, :, : n // 2] = 1
        V, nb, H = X.mesh_box_refined(3, n, ref, X.TRANSFORM_NONE, 0.0, bid)
        cx_ = V[:, :, 0].mean(axis=1)
        rho = np.where(cx_ < 0.5, 1.0, 2.0)
        c = np.where(cx_ < 0.5, 1.0, 4.0)
        op = X.WaveOperator(ctx, 3, k, V, nb, rho, c, hanging=H)
        op.set_source(X.Source.make(x0, 1.0, 10.0))
        return op, ref, rho, c, H

    op, ref, rho, c, H = build(32)
    cfl = cfl_number / k ** 1.5
    fastest = cfl * op.min_vertex_distance_over_c
    lts = X.LocalTimeStepping(op, cfl, fastest, 2, 1, 0.0, False, speed_aware=True)
    cl, _ = lts.cells()
    steps = int(np.ceil(final_time / lts.global_dt - 1e-9))
    rec = {"case": "32^3 base, x<1/2 refined once, rho,c = 1,1 | 2,4, k=3, ADER order 4 with "
                   "LTS (clusters by h/c), p=0 walls, Ricker source f0=10 at (0.25,0.5,0.5)",
           "n_cells": op.n_cells, "n_hanging_faces": int(len(H)), "n_dofs": op.n_dofs,
           "cfl_number": cfl_number, "n_clusters": lts.n_clusters,
           "cluster_step_ratio": lts.global_dt / fastest,
           "cells_per_cluster": [int((cl == i).sum()) for i in range(lts.n_clusters)],
           "fastest_dt": fastest, "global_dt": lts.global_dt, "steps": steps,
           "final_time": steps * lts.global_dt}
    u0 = op.new_vector()
    u0.zero_()
    a, b = op.new_vector(), op.new_vector()
    b.copy_(u0)
    t = 0.0
    for _ in range(2):  # warm-up (from rest, discarded)
        t += lts.global_dt
        lts.step(t, b, a)
    del lts
    lts = X.LocalTimeStepping(op, cfl, fastest, 2, 1, 0.0, False, speed_aware=True)
    b.copy_(u0)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    t = 0.0
    for _ in range(steps):
        t += lts.global_dt
        lts.step(t, b, b)  # in place
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    rec["ms_per_global_step"] = round(ms / steps, 4)
    rec["gdofs_per_global_step"] = round(op.n_dofs * steps / (ms / 1e3) / 1e9, 2)
    rec["energy_T"] = op.energy(b)
    rec["p_l2_T"] = float(torch.linalg.vector_norm(b.view(op.n_cells, 4, -1)[:, 3, :]))
    # the same time span with global ADER at the fastest step (what LTS saves)
    nw = max(op.work_vectors(X.ADER), 1)
    w = op.new_vector(nw)
    g, h = op.new_vector(), op.new_vector()
    g.zero_()
    nglob = int(round(steps * lts.global_dt / fastest))
    torch.cuda.synchronize()
    ev0.record(stream)
    tg = 0.0
    for _ in range(nglob):
        op.time_step(X.ADER, fastest, g, h, w)
        op.source_add(tg, tg + fastest, h)
        tg += fastest
        g, h = h, g
    ev1.record(stream)
    torch.cuda.synchronize()
    msg = ev0.elapsed_time(ev1)
    rec["global_ader_ms_same_span"] = round(msg, 2)
    rec["lts_speedup_vs_global_ader"] = round(msg / ms, 3)
    d = b - g
    rec["lts_vs_global_ader_rel_l2"] = float(torch.linalg.vector_norm(d) /
                                             torch.linalg.vector_norm(g))
    del a, b, g, h, w, d, lts, op
    # two LTS steps on the 8^3-base mesh against the CPU oracle
    ops, refs, rhos, cs, _ = build(8)
    fs = cfl * ops.min_vertex_distance_over_c
    lts = X.LocalTimeStepping(ops, cfl, fs, 2, 1, 0.0, False, speed_aware=True)
    m = O.Mesh.box_refined(3, 8, refs, O.TRANSFORM_NONE, 0.0, bid)
    m.set_material(rhos, cs)
    ol = O.Lts(m, k, cfl, fs, 2, 1, 0.0, 0, speed_aware=True, source=O.Source.make(x0, 1.0, 10.0))
    e = O.Exact.make(3, [[1, 1, 1]])
    u = m.project(k, e, 0.0)
    s_ = torch.from_numpy(u.copy()).cuda()
    d_ = torch.zeros_like(s_)
    t = 0.0
    for _ in range(2):
        t += lts.global_dt
        lts.step(t, s_, d_)
        u = ol.step(t, u)
        s_, d_ = d_, s_
    rec["fp64_rel_max_vs_oracle_2_lts_steps_8cube_base"] = float(
        np.max(np.abs(s_.cpu().numpy() - u)) / np.max(np.abs(u)))
    print(json.dumps({"locally_refined": rec}), flush=True)


def serving_cpu_baseline(reqs, target_s):
    """The CPU oracle solving requests of the same stream (in arrival order,
    LSRK45, FP64) for about target_s; the stream's whole work divided by the
    measured DoF-stage rate gives the CPU solves/s."""
    import numpy as np

    import oracle as O
    from exwave_amd import serving_stream as SS
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    done_work, solved, steps_done = 0.0, 0, 0
    for q in reqs:
        m = O.Mesh(q["dim"], 1 << q["n_refinements"], 0, O.TRANSFORM_NONE, 0.0, 2)
        m.set_material(np.full(m.n_cells, q["rho"]), np.full(m.n_cells, q["c"]))
        e = O.Exact.make(q["dim"], [q["modes"]], rho=q["rho"], c=q["c"])
        u = m.project(q["degree"], e, 0.0)
        h = 1.0 / (1 << q["n_refinements"])
        dt = q["cfl"] / q["degree"] ** 1.5 * h / q["c"]
        n = SS.n_steps(q)
        for s in range(n):
            vnp = np.zeros_like(u)
            m.time_step(q["degree"], O.LSRK45R2, dt, u, vnp)
            u = vnp
            done_work += 5 * SS.n_dofs(q)
            steps_done += 1
            if time.perf_counter() - t0 > target_s:
                break
        else:
            solved += 1
            continue
        break
    el = time.perf_counter() - t0
    rate = done_work / el
    total = sum(5.0 * SS.n_dofs(q) * SS.n_steps(q) for q in reqs)
    return {"solves_per_s": round(len(reqs) / (total / rate), 4), "unit": "solves/s",
            "cores": threads, "kind": "port",
            "sample": f"oracle LSRK45 FP64 on the stream's first requests in arrival order: "
                      f"{solved} solved + {steps_done} steps in {el:.1f} s = "
                      f"{rate / 1e9:.3f} GDoF-stages/s, extrapolated to the stream's "
                      f"{total:.3e} DoF-stages; OpenMP {threads} threads; CPU restatement, "
                      "not deal.II"}


def serving_measure(ctx, cpu_seconds, want_cpu):
    import numpy as np

    import exwave_amd as X
    from exwave_amd import serving as S
    from exwave_amd import serving_stream as SS
    reqs = SS.make_stream()
    # warm-up: one tiny solve per (dim, degree) pool and precision (builds graphs,
    # loads kernels), outside the measurement
    warm = [dict(id=10000 + 10 * d + k, arrival=0.0, dim=d, degree=k, n_refinements=2, rho=1.0,
                 c=1.0, modes=[1] * d, final_time=0.02, cfl=0.2)
            for d in (2, 3) for k in range(2, 7)]
    S.run_one_at_a_time(ctx, warm, X.FP64)
    S.run_server(ctx, warm, lambda q: X.FP64, realtime=False)
    S.run_server(ctx, warm, lambda q: X.FP32, realtime=False)
    solo, t_solo = S.run_one_at_a_time(ctx, reqs, X.FP64)
    res, makespan = S.run_server(ctx, reqs, lambda q: X.PRECISION_AUTO, realtime=True)
    lat = np.array([res[q["id"]]["latency"] for q in reqs])
    within = sum(res[q["id"]]["err_p"] <= 1.1 * solo[q["id"]]["err_p"] and
                 res[q["id"]]["err_v"] <= 1.1 * solo[q["id"]]["err_v"] for q in reqs)
    worst = max(max(res[q["id"]]["err_p"] / solo[q["id"]]["err_p"],
                    res[q["id"]]["err_v"] / solo[q["id"]]["err_v"]) for q in reqs)
    n32 = sum(res[q["id"]]["precision"] == X.FP32 for q in reqs)
    out = {"requests": len(reqs), "arrival_rate_per_s": 200.0,
           "solves_per_s": round(len(reqs) / makespan, 3), "makespan_s": round(makespan, 3),
           "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
           "one_at_a_time_fp64_solves_per_s": round(len(reqs) / t_solo, 3),
           "one_at_a_time_fp64_s": round(t_solo, 3),
           "speedup_vs_one_at_a_time": round(t_solo / makespan, 3),
           "within_1p1x_fp64_error": f"{within}/{len(reqs)}",
           "worst_error_ratio": round(worst, 4), "fp32_solves": n32,
           "precision": "auto (FP32 operator where its round-off stays far below the "
                        "discretisation error, else FP64; vectors and RK in FP64)"}
    if want_cpu:
        out["cpu_baseline"] = serving_cpu_baseline(reqs, cpu_seconds)
    return out


def host_cpu():
    """CPU model (/proc/cpuinfo) and the logical CPUs this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(n1d, k, target_s):
    """Time the CPU oracle's apply + RKVectorUpdater on the same mesh and data
    (stage by stage, as the reference executes it) until ~target_s elapsed."""
    import ctypes as C

    import numpy as np

    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = O.Mesh(DIM, n1d, 0, O.TRANSFORM_NONE, 0.0, 2)
    N = m.n_dofs(k)
    e = O.Exact.make(DIM, [[1, 1, 1], [2, 3, 1]])
    vec_n = m.project(k, e, 0.0)
    vec_np = np.zeros(N)
    tmp = np.zeros(N)
    L = O.lib()
    a = [970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
         2251764453980. / 15575788980749., 26877169314380. / 34165994151039.]
    b = [1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
         -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
         5198255086312. / 14908931495163.]
    dt = 0.2 / k ** 1.5 / n1d
    regs = [(vec_n, vec_n, vec_np), (vec_n, vec_np, vec_n), (vec_np, vec_n, vec_np),
            (vec_n, vec_np, vec_n), (vec_np, vec_np, vec_n)]
    stages = 0
    t0 = time.perf_counter()
    while True:
        s = stages % STAGES
        src, v2, v3 = regs[s]
        L.ox_apply(m._m, k, O._p(src), O._p(tmp), 1.0)
        if s < STAGES - 1:
            L.ox_rk_update(N, a[s] * dt, (b[s] - a[s]) * dt, 0, O._p(tmp), O._p(v2), O._p(v3))
        else:
            L.ox_rk_update(N, b[s] * dt, 0.0, 1, O._p(tmp), O._p(v2), O._p(v3))
        stages += 1
        el = time.perf_counter() - t0
        if el >= target_s and stages >= 1:
            break
    return {"value": round(N * stages / el / 1e9, 4), "unit": "GDoF/s per RK stage",
            "cores": threads, "kind": "port", **host_cpu(),
            "sample": f"{stages} LSRK45 stages (oracle apply + RKVectorUpdater, FP64) of the "
                      f"{n1d}^3 k={k} workload, {el:.1f} s, OpenMP {threads} threads; CPU "
                      "restatement, not deal.II"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.nonaffine:
        nonaffine_measure(args.nonaffine_steps,
                          [(int(c.split(",")[0]), int(c.split(",")[1]), c.split(",")[2])
                           for c in args.nonaffine_case] if args.nonaffine_case else None,
                          args.tile)
        return
    if args.refined:
        refined_measure(args.refined_time)
        return
    if args.ader:
        ader_measure([tuple(int(v) for v in c.split(",")) for c in args.ader_case]
                     if args.ader_case else None)
        return
    if args.gpus != 1:
        print(f"bench.py: one GPU only (got --gpus {args.gpus}); the path runs on a single "
              "MI355X and has no multi-GPU sharding", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import exwave_amd as X
    ctx = X.Context(0)
    stream = torch.cuda.current_stream()
    prec = X.FP64 if args.precision == "fp64" else X.FP32
    n1d, k = args.n, args.degree
    if args.tile:
        V, nb = X.mesh_hypercube(DIM, n1d, 0, X.TRANSFORM_NONE, 0.0, 2)
        V, nb = X.tile_order(V, nb, n1d, args.tile, args.blocks)
        ones = np.ones(V.shape[0])
        op = X.WaveOperator(ctx, DIM, k, V, nb, ones, ones)
    else:
        op = X.WaveOperator.hypercube(ctx, DIM, k, n1d, 0, X.TRANSFORM_NONE, 0.0, 2)
    N = op.n_dofs
    assert op.is_cartesian
    modes = X.Modes.make(DIM, [[1, 1, 1], [2, 3, 1]])
    vec_n, vec_np, work = op.new_vector(), op.new_vector(), op.new_vector()
    op.project(modes, 0.0, vec_np)
    dt = 0.2 / k ** 1.5 * op.min_vertex_distance_over_c
    E0 = op.energy(vec_np)

    def step():
        nonlocal vec_n, vec_np
        vec_n, vec_np = vec_np, vec_n  # tmp_solutions.swap(solutions), explicit_wave.cc:784
        op.time_step(X.LSRK45R2, dt, vec_n, vec_np, work, precision=prec)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    t_sec = gpu_ms / 1e3
    launches = args.steps * STAGES
    stage_ms = gpu_ms / launches
    value = N * launches / t_sec / 1e9
    # algorithmic bytes per launch (average over the 5 stages: 18 vectors/step)
    n_cells = op.n_cells
    family, kernel_name = stage_kernel(k, prec == X.FP64)
    bytes_per_launch = 18 * 8 * N / STAGES + 64 * n_cells
    achieved = bytes_per_launch / (stage_ms / 1e3) / 1e9
    # accuracy of the result (outside the timed region)
    t_end = (args.warmup + args.steps) * dt
    err = op.l2_errors(modes, t_end, vec_np)
    E1 = op.energy(vec_np)
    out = {
        "metric": "GDoF/s per RK stage (3D k=4); batched solves/s + p99 latency at equal L2 err",
        "value": round(value, 3),
        "unit": "GDoF/s per RK stage",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == X.FP64 else "f32 operator, f64 vectors",
        "data": "synthetic (analytic membrane modes (1,1,1)+(2,3,1), projected)",
        "config": {"workload": f"3D membrane, unit cube, {n1d}^3 Cartesian hexes, k={k}, "
                               f"{N} DoFs, LSRK45R2 at CFL 0.2, rho=c=1, p=0 boundary",
                   "cell_order": (f"tiles of {args.tile}^3" +
                                  (f" in {args.blocks}^3 blocks" if args.blocks > 1 else "")
                                  if args.tile else "lexicographic"),
                   "n_dofs": N, "n_cells": n_cells, "degree": k, "integrator": "LSRK45R2",
                   "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None,
                     "kernel": kernel_name,
                     "algorithmic_bytes_per_launch": int(bytes_per_launch),
                     "avg_launch_ms": round(stage_ms, 5),
                     # 4 of 5 LSRK45 stages have two outputs
                     "flops_per_launch": int(n_cells * (4 * stage_flops_per_cell(DIM, k, True) +
                                                        stage_flops_per_cell(DIM, k, False)) / 5),
                     "fp64_vector_peak_tflops": PEAK_FP64_TFLOPS},
        "wall_s": round(wall, 4),
        "l2_error_after": {"t": t_end, "p": err[0], "v": err[1], "mag_p": err[2]},
        "energy_ratio": E1 / E0,
    }
    rl = out["roofline"]
    rl["achieved_tflops"] = round(rl["flops_per_launch"] / (stage_ms / 1e3) / 1e12, 2)
    rl["flop_frac"] = round(rl["achieved_tflops"] / PEAK_FP64_TFLOPS, 4)
    try:
        sc = stream_ceiling(ctx=ctx)
        rl["stream_ceiling_gbs"] = sc["stream_ceiling_gbs"]
        rl["stream"] = sc
        rl["frac_of_stream_ceiling"] = round(achieved / sc["stream_ceiling_gbs"], 4)
        if sc.get("stage_pattern_gbs"):
            rl["frac_of_stage_pattern_ceiling"] = round(achieved / sc["stage_pattern_gbs"], 4)
    except RuntimeError as ex:  # out of memory next to the workload: report none
        rl["stream_ceiling_gbs"] = None
        rl["stream_error"] = str(ex)[:200]
    key = f"{family}_3d_k{k}_n{n1d}_t{args.tile}_{'fp64' if prec == X.FP64 else 'fp32'}"
    traffic, source = measured_traffic(key)
    out["roofline"]["traffic"] = traffic
    if traffic:
        out["roofline"]["traffic_source"] = os.path.relpath(source, HERE)
    del vec_n, vec_np, work
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n1d, k, args.cpu_seconds)
    if not args.no_serving:
        out["serving"] = serving_measure(ctx, args.serving_cpu_seconds, not args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
