"""Python driver of the serving loop (exw_server_* in include/exwave_amd.h).

`run_server(reqs)` replays a request stream in real time: each request is
submitted when the wall clock reaches its arrival time, the loop keeps calling
exw_server_progress, and completions are timestamped when polled.
`run_one_at_a_time(reqs)` solves the same requests one after the other with
exw_solve (the plain path the serving loop is compared with).
"""
from __future__ import annotations

import ctypes as C
import time

from .capi import FP32, FP64, PRECISION_AUTO, Modes, Request, Result, check, load


def make_request(q, precision=FP64):
    r = Request()
    r.dim, r.degree, r.n_refinements = q["dim"], q["degree"], q["n_refinements"]
    r.rho, r.c = q["rho"], q["c"]
    r.modes = Modes.make(q["dim"], [q["modes"]], rho=q["rho"], c=q["c"])
    r.final_time, r.cfl_number = q["final_time"], q["cfl"]
    r.precision = precision
    return r


class Server:
    def __init__(self, ctx, hbm_budget_bytes=64 << 30):
        self.L = load()
        self.h = C.c_void_p()
        check(self.L.exw_server_create(ctx.h, hbm_budget_bytes, C.byref(self.h)))
        self._buf = (Result * 256)()

    def submit(self, req: Request) -> int:
        i = C.c_int64(0)
        check(self.L.exw_server_submit(self.h, C.byref(req), C.byref(i)))
        return i.value

    def progress(self, max_steps=4) -> int:
        return check(self.L.exw_server_progress(self.h, max_steps))

    def poll(self):
        n = check(self.L.exw_server_poll(self.h, self._buf, len(self._buf)))
        return [self._buf[i] for i in range(n)]

    def status(self):
        q, a, f = C.c_int(), C.c_int(), C.c_int()
        b = C.c_size_t()
        check(self.L.exw_server_status(self.h, C.byref(q), C.byref(a), C.byref(f), C.byref(b)))
        return q.value, a.value, f.value, b.value

    def close(self):
        if self.h:
            self.L.exw_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _res(r):
    return dict(n_steps=r.n_steps, time=r.time, err_p=r.err_p, err_v=r.err_v, mag_p=r.mag_p,
                precision=r.precision)


def initial_error(q):
    """exw_request_initial_error: (E0, ||p0||, FP32 score, selected precision)."""
    r = make_request(q, PRECISION_AUTO)
    e0, n0, sc, pr = C.c_double(), C.c_double(), C.c_double(), C.c_int()
    check(load().exw_request_initial_error(C.byref(r), C.byref(e0), C.byref(n0), C.byref(sc),
                                           C.byref(pr)))
    return e0.value, n0.value, sc.value, pr.value


def run_server(ctx, reqs, precision_of=lambda q: PRECISION_AUTO, hbm_budget_bytes=64 << 30,
               realtime=True, steps_per_call=4, stall_timeout=120.0):
    """Returns ({request id: result dict with 'latency'}, makespan seconds)."""
    srv = Server(ctx, hbm_budget_bytes)
    order = sorted(reqs, key=lambda q: q["arrival"])
    ids = {}
    out = {}
    t0 = time.perf_counter()
    nxt = 0
    last_change, n_done = t0, 0
    while len(out) < len(order):
        if len(out) != n_done:
            n_done, last_change = len(out), time.perf_counter()
        elif time.perf_counter() - last_change > stall_timeout:
            srv.close()
            raise RuntimeError(f"serving loop made no progress for {stall_timeout} s")
        now = time.perf_counter() - t0
        while nxt < len(order) and (not realtime or order[nxt]["arrival"] <= now):
            q = order[nxt]
            ids[srv.submit(make_request(q, precision_of(q)))] = q
            nxt += 1
        n_active = srv.progress(steps_per_call)
        t_poll = time.perf_counter() - t0
        for r in srv.poll():
            q = ids[r.id]
            d = _res(r)
            d["latency"] = t_poll - (q["arrival"] if realtime else 0.0)
            d["done_at"] = t_poll
            out[q["id"]] = d
        if n_active == 0 and nxt < len(order) and realtime:
            qd, ac, fi, _ = srv.status()
            if qd == 0 and ac == 0 and fi == 0:
                time.sleep(max(0.0, min(0.002, order[nxt]["arrival"] - (time.perf_counter() - t0))))
    makespan = time.perf_counter() - t0
    srv.close()
    return out, makespan


def run_one_at_a_time(ctx, reqs, precision=FP64):
    L = load()
    out = {}
    res = Result()
    t0 = time.perf_counter()
    for q in reqs:
        check(L.exw_solve(ctx.h, C.byref(make_request(q, precision)), C.byref(res)))
        out[q["id"]] = _res(res)
    return out, time.perf_counter() - t0
