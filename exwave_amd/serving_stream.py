"""The serving workload: a seeded stream of independent membrane solves.

512 requests; dim uniform in {2,3}; degree k uniform in {2..6}; refinement r
uniform in 3..6 (2D) or 2..4 (3D) of the unit hypercube (2^r cells per
direction); rho, c log-uniform on [0.5, 2]; mode numbers (m, n[, l]) uniform
in 1..4; final time T uniform in [0.5, 2]; Poisson arrivals at 200 requests/s
(exponential gaps, mean 5 ms); LSRK45 at CFL 0.2 (dt = 0.2/k^1.5 * h/c, the
reference's rule, source/parameters.cc:195 and source/explicit_wave.cc:
529-545); boundary id 2 (p = 0), Cartesian cells.
"""
from __future__ import annotations

import math

import numpy as np


def make_stream(n=512, seed=20260202, rate=200.0):
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        dim = int(rng.integers(2, 4))
        k = int(rng.integers(2, 7))
        r = int(rng.integers(3, 7)) if dim == 2 else int(rng.integers(2, 5))
        rho = float(math.exp(rng.uniform(math.log(0.5), math.log(2.0))))
        c = float(math.exp(rng.uniform(math.log(0.5), math.log(2.0))))
        modes = [int(x) for x in rng.integers(1, 5, size=dim)]
        T = float(rng.uniform(0.5, 2.0))
        reqs.append(dict(id=i, arrival=t, dim=dim, degree=k, n_refinements=r, rho=rho, c=c,
                         modes=modes, final_time=T, cfl=0.2))
    return reqs


def n_steps(req):
    """Steps of the reference's time loop: while time < T: time += dt."""
    h = 1.0 / (1 << req["n_refinements"])
    dt = req["cfl"] / req["degree"] ** 1.5 * h / req["c"]
    t, s = 0.0, 0
    while t < req["final_time"]:
        t += dt
        s += 1
    return s


def n_dofs(req):
    n = (1 << req["n_refinements"]) ** req["dim"]
    return n * (req["dim"] + 1) * (req["degree"] + 1) ** req["dim"]


if __name__ == "__main__":
    reqs = make_stream()
    tot = 0.0
    rows = []
    for q in reqs:
        w = n_dofs(q) * n_steps(q) * 5
        tot += w
        rows.append((w, q))
    rows.sort(key=lambda x: -x[0])
    print("total DoF-stages %.3e  -> %.1f s at 100 GDoF/s" % (tot, tot / 100e9))
    print("arrival span %.2f s" % reqs[-1]["arrival"])
    for w, q in rows[:10]:
        print("%.3e" % w, q["dim"], q["degree"], q["n_refinements"], n_dofs(q), n_steps(q))
    small = sum(1 for w, q in rows if n_dofs(q) < 100000)
    print("requests with < 1e5 DoFs:", small, " steps total", sum(n_steps(q) for q in reqs))
