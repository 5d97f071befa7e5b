#!/usr/bin/env python3
"""Benchmark of the MPM3D explicit step on one MI355X.

One step = one explicit UL step of every particle (MPM3D.f90:69-113).

Workloads:
* batched ensemble (value): 256 seeded Drucker-Prager slope instances
  (particles log-uniform in 2e4-4e5, 3.27e7 in total; slope angle, friction
  and cohesion varied), quadratic B-spline, USF, ramped gravity load, all in
  one device pool: one launch per phase for all instances, steps replayed
  from a captured HIP graph, launch shape picked by the engine's autotuner;
* single slope (FP64-plain comparison): 40 m profile with a 45-degree face,
  0.3 m thick, cell 0.1 m, 2x2x2 particles per cell: 1,289,970 particles,
  500,526 nodes, one instance on the plain fused-step path.

Reported:
* value: particle-steps/s of the FP32-bulk batched ensemble (FP32
  particle/grid arithmetic; FP64 positions, volume, internal energy, F and
  grid mass), inputs resident in HBM, K timed steps after W warm-up steps,
  at the steady-state step time: the window's time without the re-sorts
  that fell in it, plus one separately timed re-sort per re-sort period
  (timed_ms_per_step keeps the raw window);
* fp64_plain (FP64, scatter sums in atomic order), fp64_exact_plain (FP64 in
  the reference's summation order, bit-identical to it), fp32_plain,
  fp64_batched, and the speedups over both FP64 paths; every leg timed with
  profiling off, per-kernel times from a separate profiled pass;
* parity: the single slope after a few steps and three ensemble instances
  after every step the engine ran, against the CPU FP64 oracle (normwise per
  physical group);
* roofline: the dominant kernel's algorithmic bytes per launch over its
  average launch time (HIP events around every launch of a profiled run of
  the same engine right after the timed region);
* cpu_baseline: the reference's own Fortran build (oracle/_ref/ref_driver)
  timed on the host on the first three ensemble instances for a bounded
  number of steps, 1 core; the CPU oracle is timed instead when that binary
  cannot run; cpu_baseline_all_cores: one such run per host core at once;
* taylor_bar and sgmp lines with their parity against the FP64 oracle.
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
# ids 4 B; every other particle field t B.  Grid: momentum and force 3t B
# each, mass 8 B, fixity 1 B.  The gathers form the nodal velocity (and
# acceleration) from momentum, force and mass as they load their tiles.
def kernel_bytes(kid, t, loads=True):
    if kid == "p2g":        # read x, m, v; write icell | write Mg, PXg
        return 24 + t + 3 * t + 4, 8 + 3 * t
    if kid == "stress":     # read x, cell, body, vol, ie, sm, seqv, sd(6), sig_y, epeff, cp, loads
        rd = 24 + 4 + 4 + 8 + 8 + 11 * t + (3 * t if loads else 0)
        wr = 8 + 8 + 12 * t  # write vol, ie, sm, seqv, sd(6), sig_y, epeff, q, cp
        return rd + wr, 3 * t + 8 + 1 + 3 * t   # read PXg, Mg, fixity | write FXg
    if kid == "g2p":        # read x, cell, v, m, ie; write v, x | read PXg, FXg, Mg, fixity (integrated in the load)
        return 24 + 4 + 3 * t + t + 8 + 3 * t + 24, 6 * t + 8 + 1
    return 0, 0


def measured_traffic(kernel, prec, workload):
    """HBM bytes per launch of a kernel class on a workload ("ensemble" or
    "slope") from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE, calibrated against device copies of known size); None when
    absent."""
    path = os.path.join(ROOT, "profiles", "slope_ensemble_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*%s\s*\|\s*%s\s*\|\s*%s\s*\|.*\|\s*([0-9.e+]+)\s*\|\s*$" % (kernel, prec, workload),
                     line)
        if m:
            return float(m.group(1))
    return None


def time_sort(obj, reps=3):
    """mean milliseconds of one whole-pool re-sort (synchronised)"""
    obj.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        obj.sort()
    obj.sync()
    return (time.perf_counter() - t0) * 1e3 / reps


def timed_window(obj, steps, step, period):
    """Exactly `steps` steps timed with the host clock between
    synchronisations (profiling off), then the steady-state step time: the
    re-sorts that fell in the window are counted, one re-sort is timed on its
    own, and the step time without re-sorts gets one re-sort per period
    added, whatever the window length."""
    s0 = obj.counters()[1]
    obj.sync()
    t0 = time.perf_counter()
    step(steps)
    obj.sync()
    sec = time.perf_counter() - t0
    nsort = obj.counters()[1] - s0
    sort_ms = time_sort(obj)
    step_ms = max(sec * 1e3 - nsort * sort_ms, 0.0) / steps
    amort = step_ms + (sort_ms / period if period > 0 else 0.0)
    return {"sec": sec, "timed_ms_per_step": sec * 1e3 / steps, "sorts_in_window": nsort, "sort_ms": sort_ms,
            "sort_every": period, "step_ms_without_sort": step_ms, "ms_per_step": amort}


def profiled_kernels(obj, steps, step):
    """per-kernel (total ms, launches) from HIP events around every launch of
    a separate run of `steps` steps (not the timed window)"""
    obj.sync()
    obj.profile(True)
    step(steps)
    kt = obj.kernel_times()
    obj.profile(False)
    return kt


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


def cpu_baseline(sample, mpm, decks, nsteps, log):
    """Time the reference Fortran build on a few instances of the workload
    (1 core, nsteps each, step loop only); fall back to the CPU oracle (a C
    restatement of the same step)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    sizes = [n for _, n, _ in sample]
    if os.path.exists(ref):
        d = tempfile.mkdtemp(prefix="mpmref_")
        try:
            tot_sec, tot_ps = 0.0, 0
            for k, (text, n, _) in enumerate(sample):
                with open(os.path.join(d, "s%d.mpm" % k), "w") as f:
                    f.write(text)
                r = subprocess.run(["taskset", "-c", "0", ref, "s%d.mpm" % k, str(nsteps), os.path.join(d, "o"),
                                    "-1"], cwd=d, capture_output=True, text=True, timeout=600)
                m = re.search(r"STEPLOOP\s+(\S+)\s+(\d+)\s+(\d+)", r.stdout)
                if r.returncode != 0 or not m:
                    raise OSError("reference binary rc=%s" % r.returncode)
                tot_sec += float(m.group(1))
                tot_ps += int(m.group(2)) * int(m.group(3))
            return {"value": tot_ps / tot_sec, "unit": "particle-steps/s", "cores": 1, "kind": "reference",
                    "sample": "%d steps of each of the first %d ensemble instances (%s particles), Fortran "
                              "reference (flang -O2 -ffp-contract=off), step loop only, %.1f s"
                              % (nsteps, len(sample), "+".join(map(str, sizes)), tot_sec)}
        except (OSError, subprocess.SubprocessError) as e:
            log("reference binary unavailable (%s); timing the oracle instead" % e)
        finally:
            shutil.rmtree(d, ignore_errors=True)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    tot_sec, tot_ps = 0.0, 0
    for k, (text, n, _) in enumerate(sample):
        o = pyoracle.Oracle(decks.load(text, mpm, "s%d" % k))
        t0 = time.perf_counter()
        o.step(nsteps)
        tot_sec += time.perf_counter() - t0
        tot_ps += n * nsteps
    return {"value": tot_ps / tot_sec, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of each of the first %d ensemble instances (%s particles), C oracle"
                      % (nsteps, len(sample), "+".join(map(str, sizes)))}


def cpu_baseline_all_cores(sample, nsteps, log):
    """The reference Fortran build as k independent single-threaded runs at
    once, one per host core of this process (up to 16): aggregate
    particle-steps/s of the k runs over the wall time of the slowest."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(ref):
        return None
    cores = sorted(os.sched_getaffinity(0))[:16]
    d = tempfile.mkdtemp(prefix="mpmrefk_")
    try:
        for k, (text, n, _) in enumerate(sample):
            with open(os.path.join(d, "s%d.mpm" % k), "w") as f:
                f.write(text)
        procs = []
        t0 = time.perf_counter()
        for j, c in enumerate(cores):
            k = j % len(sample)
            procs.append((k, subprocess.Popen(["taskset", "-c", str(c), ref, "s%d.mpm" % k, str(nsteps),
                                               os.path.join(d, "o%d" % j), "-1"], cwd=d,
                                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        tot_ps, loops = 0, []
        for k, p in procs:
            so, _ = p.communicate(timeout=900)
            m = re.search(r"STEPLOOP\s+(\S+)\s+(\d+)\s+(\d+)", so)
            if p.returncode != 0 or not m:
                return None
            tot_ps += int(m.group(2)) * int(m.group(3))
            loops.append(float(m.group(1)))
        wall = time.perf_counter() - t0
    except (OSError, subprocess.SubprocessError) as e:
        log("all-core reference run failed: %s" % e)
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"value": tot_ps / max(loops), "unit": "particle-steps/s", "cores": len(cores), "kind": "reference",
            "sample": "%d concurrent single-threaded reference runs (one per core), %d steps of one of the first %d "
                      "ensemble instances each; particle-steps of all runs over the slowest run's step loop "
                      "(%.1f s; %.1f s wall with input)" % (len(cores), nsteps, len(sample), max(loops), wall)}


def largest_divisor_upto(p, cap):
    return max(d for d in range(1, min(p, cap) + 1) if p % d == 0)


def touched_nodes(eng, ids):
    return int(sum(np.count_nonzero(eng.grid(i)[0]) for i in ids))


def roofline(kt, n_particles, n_touched, tsz, traffic_of):
    dom = max(("p2g", "stress", "g2p"), key=lambda k: kt[k][0])
    ms_avg = kt[dom][0] / max(kt[dom][1], 1)
    bp, bn = kernel_bytes(dom, tsz)
    alg = bp * n_particles + bn * n_touched
    achieved = alg / (ms_avg * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(dom), "algorithmic_bytes": alg,
            "avg_launch_ms": ms_avg,
            "traffic_source": "committed rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, "
                              "profiles/slope_ensemble_pmc_traffic.md"}


PLAIN_SORT_EVERY = 64   # the engine's default re-sort period (contexts never autotune)


def run_plain(mpm, prob, a, log, precs=("fp64", "fp32", "fp64exact")):
    """The single slope on the plain (one-instance, fused-step) path:
    FP64 (atomic scatter order), FP32 bulk, and the reference-order FP64
    mode (bit-identical to the reference)."""
    res = {}
    for prec in precs:
        ctx = mpm.Context(prob, prec)
        ctx.step(a.warmup)
        w = timed_window(ctx, a.steps, ctx.step, PLAIN_SORT_EVERY)
        kt = profiled_kernels(ctx, a.profile_steps, ctx.step)
        touched = int(np.count_nonzero(ctx.grid()[0]))
        res[prec] = dict(w, kt=kt, touched=touched, value=prob.n / (w["ms_per_step"] * 1e-3))
        log("plain %s: %.3f ms/step (timed %.3f, %d re-sorts of %.2f ms in the window), %.4g particle-steps/s"
            % (prec, w["ms_per_step"], w["timed_ms_per_step"], w["sorts_in_window"], w["sort_ms"],
               res[prec]["value"]))
        ctx.close()
    return res


def run_ensemble(mpm, decks, a, log):
    """The seeded slope ensemble in one engine per precision: warm-up,
    autotune, then K graph-replayed steps timed with the host clock around
    synchronisations; a separate plain run of a few steps with HIP events
    around every launch gives the per-kernel averages."""
    ens = decks.slope_ensemble(a.ensemble, seed=a.seed)
    total_p = sum(n for _, n, _ in ens)
    total_n = sum(m["nodes"] for _, _, m in ens)
    slack = (a.ensemble + 4) * 256
    total_b = sum(m["bodies"] for _, _, m in ens)
    engs = {prec: mpm.Engine(max_particles=total_p + slack, max_nodes=total_n + slack,
                             max_instances=a.ensemble, max_materials=a.ensemble, max_bodies=total_b,
                             precision=prec) for prec in ("fp32", "fp64")}
    ids = {prec: [] for prec in engs}
    t0 = time.perf_counter()
    for k, (text, n, meta) in enumerate(ens):
        prob = decks.load(text, mpm, "slope%d" % k)
        for prec, eng in engs.items():
            ids[prec].append(eng.add(prob))
        del prob
        if (k + 1) % 64 == 0:
            log("ensemble: %d/%d instances resident (%.1f s)" % (k + 1, a.ensemble, time.perf_counter() - t0))
    log("ensemble: %d instances, %d particles, %d nodes" % (a.ensemble, total_p, total_n))
    out = {}
    for prec, eng in engs.items():
        eng.step(a.warmup)
        eng.sync()
        tune = eng.autotune(trial_steps=4) if a.autotune else eng.tuning()
        per = largest_divisor_upto(tune["sort_every"], 10) if tune["sort_every"] > 0 else 10
        log("ensemble %s tuning: %s, graph of %d steps" % (prec, json.dumps(tune), per))
        eng.step(per, graph=a.graph, graph_steps=per)   # capture outside the timed region
        eng.sync()
        w = timed_window(eng, a.steps, lambda n: eng.step(n, graph=a.graph, graph_steps=per), tune["sort_every"])
        kt = profiled_kernels(eng, a.profile_steps, eng.step)
        out[prec] = dict(w, kt=kt, tuning=tune, graph_steps=per, value=total_p / (w["ms_per_step"] * 1e-3))
        log("ensemble %s: %.3f ms/step (timed %.3f, %d re-sorts of %.2f ms in the window), %.4g particle-steps/s"
            % (prec, w["ms_per_step"], w["timed_ms_per_step"], w["sorts_in_window"], w["sort_ms"],
               out[prec]["value"]))
    out["touched"] = touched_nodes(engs["fp32"], ids["fp32"])
    out["particles"] = total_p
    out["nodes"] = total_n
    out["ens"] = ens
    out["engs"] = engs
    out["ids"] = ids
    return out


def ensemble_parity(mpm, decks, E, a, log):
    """FP32 and FP64 batched results of a few instances (smallest, second
    smallest, median size) against the CPU FP64 oracle after the same steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    sizes = [n for _, n, _ in E["ens"]]
    order = np.argsort(sizes)
    pick = [int(order[0]), int(order[1]), int(order[len(order) // 2])]
    res = []
    for k in pick:
        text, n, meta = E["ens"][k]
        prob = decks.load(text, mpm, "slope%d" % k)
        steps = int(E["engs"]["fp32"].scalars(E["ids"]["fp32"][k])[0])
        o = pyoracle.Oracle(prob)
        o.step(steps)
        ref = o.particles()
        r = {"instance": k, "particles": n, "steps": steps}
        for prec in ("fp32", "fp64"):
            assert int(E["engs"][prec].scalars(E["ids"][prec][k])[0]) == steps
            r[prec + "_vs_fp64_oracle"] = group_err(mpm, E["engs"][prec].particles(E["ids"][prec][k]), ref, GROUPS)
        res.append(r)
        log("ensemble parity:", json.dumps(r))
    return res


def run_bar(mpm, decks, a, log):
    """The Taylor bar (GIMP, J2, MUSL; the reference's CPU-runnable case) on
    the plain path, FP64 and FP32, and the reference's CPU step on it."""
    text, n = decks.taylor_bar()
    prob = decks.load(text, mpm, "taylor")
    out = {"workload": "Taylor bar R=3.2 mm L=32.4 mm, 190 m/s, copper J2 (400 MPa, H=100 MPa), GIMP, MUSL, "
                       "h=0.4 mm, 2x2x2 particles/cell", "particles": prob.n}
    for prec in ("fp64", "fp32", "fp64exact"):
        ctx = mpm.Context(prob, prec)
        ctx.step(a.warmup)
        w = timed_window(ctx, a.steps, ctx.step, PLAIN_SORT_EVERY)
        out[prec] = {"value": prob.n / (w["ms_per_step"] * 1e-3), "ms_per_step": w["ms_per_step"],
                     "timed_ms_per_step": w["timed_ms_per_step"]}
        ctx.close()
    if not a.no_cpu:   # every path against the FP64 oracle after a few full-size steps
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(prob)
        o.step(a.check_steps)
        ref = o.particles()
        par = {"steps": a.check_steps}
        for prec in ("fp64", "fp32", "fp64exact"):
            ctx = mpm.Context(prob, prec)
            ctx.step(a.check_steps)
            ps = ctx.particles()
            par[prec + "_vs_fp64_oracle"] = group_err(mpm, ps, ref, GROUPS)
            if prec == "fp64exact":
                par["fp64exact_bitwise"] = bool(np.array_equal(ps, ref, equal_nan=True)
                                                and np.array_equal(ctx.grid(), o.grid(), equal_nan=True))
            ctx.close()
        out["parity"] = par
    log("taylor bar: %s" % json.dumps(out))
    if not a.no_cpu:
        ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
        if os.path.exists(ref):
            d = tempfile.mkdtemp(prefix="mpmbar_")
            try:
                with open(os.path.join(d, "bar.mpm"), "w") as f:
                    f.write(text)
                r = subprocess.run(["taskset", "-c", "0", ref, "bar.mpm", "20", os.path.join(d, "o"), "-1"], cwd=d,
                                   capture_output=True, text=True, timeout=600)
                m = re.search(r"STEPLOOP\s+(\S+)\s+(\d+)\s+(\d+)", r.stdout)
                if r.returncode == 0 and m:
                    out["cpu_reference"] = {"value": int(m.group(2)) * int(m.group(3)) / float(m.group(1)),
                                            "cores": 1, "sample": "20 steps, step loop only"}
            finally:
                shutil.rmtree(d, ignore_errors=True)
    return out


def run_dropin(mpm, decks, a, log):
    """The reference's own main program over the device step (the Fortran
    drop-in: MPM3D.f90 and its input phase linked with update_step_hip.f90,
    oracle/_ref/mpm3d_dropin) on the Taylor bar and the single slope, with
    the reference program itself (oracle/_ref/mpm3d_ref, one core) on the same
    decks.  Reported: the programs' own "Time per particle cycle" line (CPU
    time of the step loop per particle and step, MPM3D.f90:259-279; the
    drop-in's host waits for the device inside that time), the loop's wall
    time as the program prints it (whole seconds) and the process's wall
    time (input phase included); steps from the curve file.  The drop-in downloads per step only the energies and the curve
    particles; these decks write no frames.  The slope runs without its
    gravity ramp (QuLo): with it the reference's loop only ends after
    CurrentTime > 100 (MPM3D.f90:152-155); the work per step is unchanged."""
    progs = {"dropin": os.path.join(ROOT, "oracle", "_ref", "mpm3d_dropin"),
             "reference": os.path.join(ROOT, "oracle", "_ref", "mpm3d_ref")}
    if not all(os.path.exists(p) for p in progs.values()):
        return None
    out = {}
    for key, (text, _) in (("taylor_bar", decks.taylor_bar()), ("single_slope", decks.slope())):
        prob = decks.load(text, mpm, key)
        res = {"particles": prob.n}
        runs = [("dropin_fp64", "dropin", 100, {"MPM3D_PRECISION": "fp64"}),
                ("dropin_fp32", "dropin", 100, {"MPM3D_PRECISION": "fp32"}),
                ("reference_1core", "reference", 3 if prob.n > 500000 else 20, {})]
        for name, prog, steps, env in runs:
            deck = re.sub(r"(?m)^endtime .*$", "endtime %.9e" % ((steps - 0.5) * prob.dt), text)
            deck = re.sub(r"(?m)^QuLo .*\n", "", deck)
            d = tempfile.mkdtemp(prefix="mpmdropin_")
            try:
                with open(os.path.join(d, "run.mpm"), "w") as f:
                    f.write(deck)
                cmd = ([progs[prog], "run"] if prog == "dropin" else ["taskset", "-c", "0", progs[prog], "run"])
                t0 = time.time()
                r = subprocess.run(cmd, cwd=d, env=dict(os.environ, **env), capture_output=True, text=True,
                                   timeout=600, stdin=subprocess.DEVNULL)
                t1 = time.time() - t0
                tp = re.search(r"Time per particle cycle:\s+(\S+)", r.stdout)
                wall = re.search(r"Total elasped time :\s+(\S+)", r.stdout)
                nst = max(0, len(open(os.path.join(d, "run_curv.dat")).read().splitlines()) - 1)
                if r.returncode == 0 and tp and wall and nst:
                    res[name] = {"ns_per_particle_cycle": float(tp.group(1)),
                                 "particle_steps_per_s_cpu_time": 1e9 / float(tp.group(1)),
                                 "steps": nst, "loop_wall_s": float(wall.group(1)),
                                 "process_wall_s": t1}
                else:
                    res[name] = {"error": (r.stdout + r.stderr)[-300:]}
                log("drop-in %s %s: %s" % (key, name, json.dumps(res[name])))
            finally:
                shutil.rmtree(d, ignore_errors=True)
        out[key] = res
    log("drop-in: %s" % json.dumps(out))
    return out


def run_sgmp(mpm, decks, a, log):
    """SGMP workloads: 64 Cook-membrane plates (nearly incompressible
    neo-Hookean) and one 3D dam break in one engine, FP32 bulk (FP64 F, det F,
    volume), graph-replayed steps after autotuning; the Cook plate's FP32
    result against the CPU FP64 oracle after a few steps."""
    cook_text, n_cook = decks.cook_membrane()
    dam_text, n_dam = decks.dam_break()
    cook = decks.load(cook_text, mpm, "cook")
    dam = decks.load(dam_text, mpm, "dam")
    nc = a.sgmp_cooks
    tot_p = nc * cook.n + dam.n
    eng = mpm.Engine(max_particles=tot_p + (nc + 2) * 256, max_nodes=nc * cook.nb_node + dam.nb_node + (nc + 2) * 256,
                     max_instances=nc + 1, max_materials=nc + 1, max_bodies=nc * len(cook.bodies) + 4,
                     precision="fp32")
    ids = [eng.add(cook) for _ in range(nc)] + [eng.add(dam)]
    log("sgmp: %d Cook plates (%d particles each) + dam break (%d particles): %d particles"
        % (nc, cook.n, dam.n, tot_p))
    eng.step(2)
    eng.sync()
    tune = eng.autotune(trial_steps=3)
    per = largest_divisor_upto(tune["sort_every"], 10) if tune["sort_every"] > 0 else 10
    eng.step(per, graph=a.graph, graph_steps=per)
    eng.sync()
    w = timed_window(eng, a.sgmp_steps, lambda n: eng.step(n, graph=a.graph, graph_steps=per), tune["sort_every"])
    kt = profiled_kernels(eng, 3, eng.step)
    par = None
    if not a.no_cpu:
        # a batched Cook plate against the CPU FP64 oracle after the same steps
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        steps = int(eng.scalars(ids[0])[0])
        o = pyoracle.Oracle(cook)
        o.step(steps)
        par = {"instance": "Cook plate 0 of the batch", "steps": steps,
               "fp32_vs_fp64_oracle": group_err(mpm, eng.particles(ids[0]), o.particles(), GROUPS)}
        log("sgmp parity:", json.dumps(par))
    eng.close()
    out = {"workload": "%d Cook-membrane plates 48x44(->16)x10 mm, h=0.5 mm, 2x2x2 particles/cell, neo-Hookean "
                       "E=240 MPa nu=0.4999 rho=1e-3 g/mm^3, shear traction ramp, + 3D dam break 0.6x0.3x0.3 m, "
                       "h=5 mm, c=50 m/s; SGMP, MUSL, FP32 bulk" % nc,
           "particles": tot_p, "steps": a.sgmp_steps, "ms_per_step": w["ms_per_step"],
           "timed_ms_per_step": w["timed_ms_per_step"], "sort_ms": w["sort_ms"],
           "value": tot_p / (w["ms_per_step"] * 1e-3), "unit": "particle-steps/s", "tuning": tune,
           "kernels_ms_avg": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kt.items()}, "parity": par}
    log("sgmp: %.3f ms/step, %.4g particle-steps/s" % (out["ms_per_step"], out["value"]))
    return out


def run_serving(mpm, decks, a, log):
    """Continuous-batching serving of a 70/30 mix of slope and Taylor-bar
    runs (100-300 steps each, FP32 bulk) under Poisson arrivals, one pool
    sized by the HBM budget; p50/p99 time to solution per arrival rate."""
    from mpm3d_f90_amd import serve
    t0 = time.perf_counter()
    pool = serve.request_pool(mpm, decks, count=48, seed=a.seed)
    log("serving: pool of %d requests (%d slopes, %d bars, %d-%d particles) built in %.1f s"
        % (len(pool), sum(k == "slope" for _, _, k in pool), sum(k == "bar" for _, _, k in pool),
           min(p.n for p, _, _ in pool), max(p.n for p, _, _ in pool), time.perf_counter() - t0))
    eng = mpm.Engine(hbm_bytes=a.serve_budget, max_instances=512, precision="fp32")
    srv = serve.Server(eng, chunk=8, graph="lazy")
    # warm-up: one request through the whole path (capture, admission, retire)
    srv.submit(serve.Request(-1, pool[0][0], 16, arrival=0.0, kind="warm"))
    srv.admit(0.0)
    while not srv.idle():
        srv.advance()
        srv.retire(lambda: 0.0)
    runs = []
    for k, rate in enumerate(float(r) for r in a.serve_rates.split(",")):
        srv.done.clear()
        srv.max_running = 0
        st = serve.run_poisson(srv, pool, rate, a.serve_window, seed=a.seed + k, max_wall=a.serve_window + 60)
        log("serving %.0f req/s: %s" % (rate, json.dumps(st)))
        runs.append(st)
    out = {"workload": "70%% slopes (ensemble generator), 30%% Taylor bars (v U(100,250) m/s, R U(2,5) mm), "
                       "100-300 steps per request, FP32 bulk, HBM budget %.0f GB, chunks of 8 graph-replayed steps"
                       % (a.serve_budget / 1e9),
           "pool_device_bytes": eng.device_bytes(), "runs": runs}
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="timed steps (default: one re-sort period of the tuned shape)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ensemble", type=int, default=256, help="instances of the batched slope ensemble")
    ap.add_argument("--seed", type=int, default=0, help="seed of the ensemble generator")
    ap.add_argument("--profile-steps", type=int, default=5, help="steps with per-launch HIP events")
    ap.add_argument("--check-steps", type=int, default=2, help="steps of the single slope compared with the oracle")
    ap.add_argument("--cpu-steps", type=int, default=30, help="steps of the CPU baseline sample per instance")
    ap.add_argument("--no-cpu", action="store_true", help="skip oracle parity and CPU baseline")
    ap.add_argument("--no-dropin", dest="dropin", action="store_false",
                    help="skip the Fortran drop-in timing on the Taylor bar and the single slope")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="timed steps as plain launches instead of graph replays (for tracing tools)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false", help="keep the default launch shape")
    ap.add_argument("--sgmp-cooks", type=int, default=64, help="Cook plates of the SGMP batch (0: skip)")
    ap.add_argument("--sgmp-steps", type=int, default=10, help="timed steps of the SGMP batch")
    ap.add_argument("--serve-rates", default="10,50,200", help="Poisson arrival rates (req/s) of the serving runs")
    ap.add_argument("--serve-window", type=float, default=8.0, help="seconds of arrivals per serving run (0: skip)")
    ap.add_argument("--serve-budget", type=float, default=200e9, help="HBM budget (bytes) of the serving pool")
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
    log("single slope: %d particles, %d nodes" % (prob.n, prob.nb_node))

    plain = run_plain(mpm, prob, a, log)
    E = run_ensemble(mpm, decks, a, log)

    parity = {}
    cpu = cpu_all = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.Oracle(prob)
        o.step(a.check_steps)
        ref = o.particles()
        single = {}
        for prec in ("fp64", "fp32", "fp64exact"):
            ctx = mpm.Context(prob, prec)
            ctx.step(a.check_steps)
            ps = ctx.particles()
            single[prec + "_vs_fp64_oracle"] = group_err(mpm, ps, ref, GROUPS)
            if prec == "fp64exact":   # the reference-order mode: bit for bit
                single["fp64exact_bitwise"] = bool(np.array_equal(ps, ref, equal_nan=True)
                                                   and np.array_equal(ctx.grid(), o.grid(), equal_nan=True))
            ctx.close()
        single["steps"] = a.check_steps
        parity["single_slope"] = single
        log("single slope parity:", json.dumps(single))
        parity["ensemble"] = ensemble_parity(mpm, decks, E, a, log)
        parity["max_fp32_err"] = max(max(r["fp32_vs_fp64_oracle"].values()) for r in parity["ensemble"])
    for eng in E["engs"].values():
        eng.close()
    bar = run_bar(mpm, decks, a, log)
    dropin = run_dropin(mpm, decks, a, log) if a.dropin else None
    sgmp = run_sgmp(mpm, decks, a, log) if a.sgmp_cooks > 0 else None
    serving = run_serving(mpm, decks, a, log) if a.serve_window > 0 else None
    if not a.no_cpu:
        sample = [E["ens"][k] for k in range(3)]
        cpu = cpu_baseline(sample, mpm, decks, a.cpu_steps, log)
        log("cpu baseline:", json.dumps(cpu))
        cpu_all = cpu_baseline_all_cores(sample, a.cpu_steps, log)
        log("cpu baseline, all cores:", json.dumps(cpu_all))

    e32 = E["fp32"]
    kt = e32["kt"]
    traffic = lambda k: measured_traffic(k, "fp32", "ensemble")
    v64p = plain["fp64"]["value"]
    v64x = plain["fp64exact"]["value"]
    leg = lambda r: {"value": r["value"], "ms_per_step": r["ms_per_step"], "timed_ms_per_step": r["timed_ms_per_step"],
                     "sort_ms": r["sort_ms"], "sort_every": r["sort_every"], "sorts_in_window": r["sorts_in_window"]}
    out = {
        "metric": "particle-steps/s FP32-batched vs FP64-plain @<=1e-5 rel err; p50/p99 latency",
        "value": e32["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": e32["ms_per_step"],
        "timed_ms_per_step": e32["timed_ms_per_step"],
        "resort": {"sort_ms": e32["sort_ms"], "sort_every": e32["sort_every"],
                   "sorts_in_window": e32["sorts_in_window"],
                   "note": "value and ms_per_step: step time of the timed window without its re-sorts plus one "
                           "measured re-sort per sort_every steps (steady state, whatever --steps is); "
                           "timed_ms_per_step: the raw window"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched slope ensemble: %d Drucker-Prager slope instances (seed %d; particles "
                               "log-uniform 2e4-4e5, angle U(30,60), friction U(20,40), cohesion lognormal "
                               "median 10 kPa), quadratic B-spline, USF, h=0.1 m, 2x2x2 particles/cell, "
                               "one launch per phase, graph-replayed steps, autotuned launch shape"
                               % (a.ensemble, a.seed),
                   "instances": a.ensemble, "particles": E["particles"], "nodes": E["nodes"],
                   "nodes_touched": E["touched"],
                   "precision": "FP32 bulk; FP64 positions, volume, internal energy, F, grid mass"},
        "fp64_plain": dict(leg(plain["fp64"]), workload="single slope, %d particles, %d nodes, FP64 (scatter "
                                                           "sums in atomic order)" % (prob.n, prob.nb_node)),
        "fp64_exact_plain": dict(leg(plain["fp64exact"]), workload="single slope, FP64 in the reference's "
                                                                    "summation order (bit-identical to it)"),
        "speedup_vs_fp64_plain": e32["value"] / v64p,
        "speedup_vs_fp64_exact_plain": e32["value"] / v64x,
        "fp32_plain": leg(plain["fp32"]),
        "fp64_batched": leg(E["fp64"]),
        "tuning": {"fp32": e32["tuning"], "fp64": E["fp64"]["tuning"], "graph_steps": e32["graph_steps"]},
        "kernels_ms_avg": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kt.items()},
        "roofline": roofline(kt, E["particles"], E["touched"], 4, traffic),
        "roofline_plain_fp32": roofline(plain["fp32"]["kt"], prob.n, plain["fp32"]["touched"], 4,
                                        lambda k: measured_traffic(k, "fp32", "slope")),
        "parity": parity or None,
        "taylor_bar": bar,
        "dropin": dropin,
        "sgmp": sgmp,
        "serving": serving,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all if not a.no_cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
