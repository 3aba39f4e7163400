"""Extract the golden vectors of the reference's regression tests.

Reads ``tests/*.prm`` and ``tests/*.output`` of the reference checkout (default
/root/reference) and writes ``reference_fixtures.json``: for every test the
parameters the oracle needs and the numbers of every 'error' line
(time, error p, error p post, error v, solution mag p), plus the logged
number of time steps and time step size.  Only numbers are kept; no reference
source or script text is copied.

    python tests/golden/make_fixtures.py [/path/to/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
tdir = os.path.join(ref, "tests")
out = {}
keys = {"dimension": int, "fe_degree": int, "n_initial_intervals": int, "n_refinements": int,
        "grid_transform_factor": float, "time_integrator": str, "cfl_number": float,
        "max_time_steps": int, "final_time": float, "output_every_time": float,
        "use_ader_post": lambda s: s.strip() == "true", "spectral_evaluation":
        lambda s: s.strip() == "true", "initital_cases": int, "membrane_modes": int,
        "max_n_clusters": int, "max_diff_clusters": int, "boundary_id": int}
num = r"([-+0-9.eE]+)"
for fn in sorted(os.listdir(tdir)):
    if not fn.endswith(".prm"):
        continue
    name = fn[:-4]
    params = {}
    for line in open(os.path.join(tdir, fn)):
        m = re.match(r"\s*set\s+(\w+)\s*=\s*(\S+)", line)
        if m and m.group(1) in keys:
            params[m.group(1)] = keys[m.group(1)](m.group(2))
        m = re.match(r"#\s*mpirun:\s*(\d+)", line)
        if m:
            params["mpirun"] = int(m.group(1))
    rows, steps, dts, clusters, lts = [], None, [], [], {}
    for line in open(os.path.join(tdir, name + ".output")):
        m = re.match(r"\s*Time:\s*" + num + r"\s*, error p:\s*" + num + r"\s*, error p post:\s*"
                     + num + r"\s*, error v:\s*" + num + r"\s*, solution mag p:\s*" + num, line)
        if m:
            rows.append([m.group(i) for i in range(1, 6)])
        m = re.match(r"\s*Performed (\d+) time steps", line)
        if m:
            steps = int(m.group(1))
        m = re.match(r"\s*Time step size:\s*" + num, line)
        if m:
            dts.append(m.group(1))
        # ADER-LTS cluster report (cells are counted in batches of the
        # VectorizedArray width the run printed)
        m = re.match(r"\s*cluster\s+(\d+) contains (\d+) cells and works with time step\s*"
                     + num, line)
        if m:
            clusters.append([int(m.group(1)), int(m.group(2)), m.group(3)])
        m = re.match(r"\s*n_clusters (\d+) clusterdiff (\d+) levelmax (\d+)", line)
        if m:
            lts.update(n_clusters=int(m.group(1)), cluster_diff=int(m.group(2)),
                       level_max=int(m.group(3)))
        m = re.match(r"\s*needed to iterate (\d+) times", line)
        if m:
            lts["iterations"] = int(m.group(1))
        m = re.match(r"\s*Vectorization over (\d+) doubles", line)
        if m:
            params["n_vect"] = int(m.group(1))
        m = re.match(r"\s*Number of global active cells:\s*(\d+)", line)
        if m:
            params["n_cells"] = int(m.group(1))
        m = re.match(r"\s*Number of degrees of freedom DG system:\s*(\d+)", line)
        if m:
            params["n_dofs"] = int(m.group(1))
    out[name] = {"params": params, "error_lines": rows, "n_steps": steps,
                 "time_step_sizes": dts}
    if clusters:
        lts["clusters"] = clusters
        out[name]["lts"] = lts
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
with open(dst, "w") as f:
    json.dump({"source": "kronbichler/exwave tests/*.prm + tests/*.output (numbers only)",
               "columns": ["time", "error_p", "error_p_post", "error_v", "solution_mag_p"],
               "tests": out}, f, indent=1)
print("wrote", dst, len(out), "tests")
