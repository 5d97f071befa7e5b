#!/usr/bin/env python3
"""Benchmark of the MPM3D explicit step on one MI355X.

Workload: the single 3D soil slope (Drucker-Prager, quadratic B-spline, USF,
ramped gravity load), 40 m profile with a 45-degree face, 0.3 m thick, cell
0.1 m, 2x2x2 particles per cell: 1,289,970 particles, 500,526 nodes.

One step = one explicit UL step of every particle (MPM3D.f90:69-113).
Reported:
* value: particle-steps/s of the FP32-bulk path (FP32 particle/grid
  arithmetic; FP64 positions, volume, internal energy and grid mass), inputs
  resident in HBM, K timed steps after W warm-up steps;
* fp64_plain: the same workload on the plain FP64 path;
* parity: both paths against the CPU FP64 oracle after a few steps at full
  size (normwise per physical group);
* roofline: the dominant kernel's algorithmic bytes per launch over its
  average launch time (HIP events on the engine stream, timed region);
* cpu_baseline: the reference's own Fortran build (oracle/_ref/ref_driver)
  timed on the host for a bounded number of steps of the same deck, 1 core;
  the CPU oracle is timed instead when that binary cannot run.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def load_package():
    import importlib.util
    if "mpm3d_f90_amd" in sys.modules:
        return sys.modules["mpm3d_f90_amd"]
    pkg = os.path.join(ROOT, "mpm3d-f90_amd")
    spec = importlib.util.spec_from_file_location("mpm3d_f90_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mpm3d_f90_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


# Algorithmic bytes per particle / per touched node of each kernel class of
# the USF step (B-spline, Drucker-Prager), by bulk type size t (4 or 8):
# every field the kernel must read plus every field it must write, once.
# Particle: x,y,z 8 B each; volume and internal energy 8 B; cell and body
# ids 4 B; every other particle field t B.  Grid: velocity V, acceleration
# A, momentum and force 3t B each, mass 8 B.
def kernel_bytes(kid, t, loads=True):
    if kid == "p2g":        # read x, m, v; write icell | write Mg, PXg
        return 24 + t + 3 * t + 4, 8 + 3 * t
    if kid == "stress":     # read x, cell, body, vol, ie, sm, seqv, sd(6), sig_y, epeff, cp, loads
        rd = 24 + 4 + 4 + 8 + 8 + 11 * t + (3 * t if loads else 0)
        wr = 8 + 8 + 12 * t  # write vol, ie, sm, seqv, sd(6), sig_y, epeff, q, cp
        return rd + wr, 6 * t            # read V | write FXg
    if kid == "g2p":        # read x, cell, v, m, ie; write v, vbar, x | read V, A
        return 24 + 4 + 3 * t + t + 8 + 3 * t + 3 * t + 24, 6 * t
    return 0, 0


def measured_traffic(kernel, prec):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE, calibrated against device copies
    of known size); None when absent."""
    path = os.path.join(ROOT, "profiles", "slope_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*%s\s*\|\s*%s\s*\|.*\|\s*([0-9.e+]+)\s*\|\s*$" % (kernel, prec), line)
        if m:
            return float(m.group(1))
    return None


def time_steps(ctx, steps, warmup, profile):
    ctx.step(warmup)
    ctx.sync()
    if profile:
        ctx.profile(True)
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    kt = ctx.kernel_times() if profile else {}
    if profile:
        ctx.profile(False)
    return dt, kt


def group_err(mpm, a, b, groups):
    idx = {n: i for i, n in enumerate(mpm.PF_NAMES)}
    out = {}
    for g, fields in groups.items():
        A = a[[idx[f] for f in fields]]
        B = b[[idx[f] for f in fields]]
        ok = np.isfinite(A) & np.isfinite(B)
        den = np.max(np.abs(B[ok])) if ok.any() else 0.0
        num = np.max(np.abs(A[ok] - B[ok])) if ok.any() else 0.0
        out[g] = float(num / den) if den > 0 else float(num)
    return out


GROUPS = {"x": ["x", "y", "z"], "v": ["vx", "vy", "vz"],
          "stress": ["sm", "seqv", "sdxx", "sdyy", "sdzz", "sdyz", "sdxz", "sdxy", "q"],
          "vol": ["vol"], "ie": ["ie"], "epeff": ["epeff"]}


def cpu_baseline(deck_text, prob, mpm, nsteps, log):
    """Time the reference Fortran build on the same deck (1 core); fall back
    to the CPU oracle (a C restatement of the same step)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if os.path.exists(ref):
        d = tempfile.mkdtemp(prefix="mpmref_")
        try:
            with open(os.path.join(d, "slope.mpm"), "w") as f:
                f.write(deck_text)
            r = subprocess.run(["taskset", "-c", "0", ref, "slope.mpm", str(nsteps), os.path.join(d, "o"), "-1"],
                               cwd=d, capture_output=True, text=True, timeout=600)
            m = re.search(r"STEPLOOP\s+(\S+)\s+(\d+)\s+(\d+)", r.stdout)
            if r.returncode == 0 and m:
                sec, n, k = float(m.group(1)), int(m.group(2)), int(m.group(3))
                return {"value": n * k / sec, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
                        "sample": "%d steps of the full %d-particle slope deck, Fortran reference "
                                  "(flang -O2 -ffp-contract=off), step loop only" % (k, n)}
            log("reference binary did not run (rc=%s); timing the oracle instead" % r.returncode)
        except (OSError, subprocess.SubprocessError) as e:
            log("reference binary unavailable (%s); timing the oracle instead" % e)
        finally:
            shutil.rmtree(d, ignore_errors=True)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(prob)
    t0 = time.perf_counter()
    o.step(nsteps)
    sec = time.perf_counter() - t0
    return {"value": prob.n * nsteps / sec, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the full %d-particle slope deck, C oracle" % (nsteps, prob.n)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--check-steps", type=int, default=2, help="steps compared with the CPU oracle")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip oracle parity and CPU baseline")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (instances shard "
              "across GPUs without communication, see DESIGN.md)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    log = lambda *m: print("[bench]", *m, file=sys.stderr, flush=True)
    mpm = load_package()
    from mpm3d_f90_amd import decks
    text, n = decks.slope()
    prob = decks.load(text, mpm, "slope")
    log("workload: slope, %d particles, %d nodes" % (prob.n, prob.nb_node))

    res = {}
    for prec in ("fp64", "fp32"):
        ctx = mpm.Context(prob, prec)
        sec, kt = time_steps(ctx, a.steps, a.warmup, profile=True)
        g = ctx.grid()
        touched = int(np.count_nonzero(g[0]))
        res[prec] = {"sec": sec, "kt": kt, "touched": touched}
        log("%s: %.3f ms/step, %.4g particle-steps/s" % (prec, sec / a.steps * 1e3, prob.n * a.steps / sec))
        ctx.close()

    parity = {}
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(prob)
        o.step(a.check_steps)
        ref = o.particles()
        for prec in ("fp64", "fp32"):
            ctx = mpm.Context(prob, prec)
            ctx.step(a.check_steps)
            parity[prec + "_vs_fp64_oracle"] = group_err(mpm, ctx.particles(), ref, GROUPS)
            ctx.close()
        parity["steps"] = a.check_steps
        log("parity:", json.dumps(parity))
        cpu = cpu_baseline(text, prob, mpm, a.cpu_steps, log)
        log("cpu baseline:", json.dumps(cpu))

    r32 = res["fp32"]
    kt = r32["kt"]
    dom = max(("p2g", "stress", "g2p"), key=lambda k: kt[k][0])
    ms_avg = kt[dom][0] / max(kt[dom][1], 1)
    bp, bn = kernel_bytes(dom, 4)
    alg = bp * prob.n + bn * r32["touched"]
    achieved = alg / (ms_avg * 1e-3) / 1e9
    value = prob.n * a.steps / r32["sec"]
    v64 = prob.n * a.steps / res["fp64"]["sec"]
    out = {
        "metric": "particle-steps/s FP32-batched vs FP64-plain @<=1e-5 rel err; p50/p99 latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["sec"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "slope: Drucker-Prager soil, quadratic B-spline, USF, 40 m x 0.3 m x 20 m "
                               "profile with 45-degree face, h=0.1 m, 2x2x2 particles/cell, gravity ramp",
                   "particles": prob.n, "nodes": prob.nb_node, "nodes_touched": r32["touched"],
                   "precision": "FP32 bulk; FP64 positions, volume, internal energy, grid mass"},
        "fp64_plain": {"value": v64, "ms_per_step": res["fp64"]["sec"] / a.steps * 1e3},
        "speedup_fp32_vs_fp64": value / v64,
        "kernels_ms_avg": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kt.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, "fp32"),
                     "algorithmic_bytes": alg, "avg_launch_ms": ms_avg},
        "parity": parity or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
