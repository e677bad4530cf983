#!/usr/bin/env python3
"""Benchmark of the liquid step on one MI355X.

Workload (default): faucet into an obstacle field, 256x128x128 MAC grid (h = 1/256 m, 1.0 x 0.5 x
0.5 m tank) initially half full at 8 particles per cell (16.8 M particles), a sphere (r 5 cm), a
cylinder (r 3 cm, h 20 cm), a torus (R 6 cm, r 1.5 cm) and a drain (r 2 cm); a 0.5 L/s faucet at
1.5 m/s angled 30 degrees down; red-black SOR omega 1.8 with 60 iterations; dt = 1/240 s, one
substep per step; FLIP 0.95; whitewater on; ripple layer M = 1024 (R = 512) with its mask.

A step = one FluidSim.step (emission, P2G, normalise+forces, divergence, 60 SOR iterations,
projection, G2P, compaction, whitewater, ripple frame). Inputs are resident in HBM before timing.
Prints one JSON line: particle-steps/s, the SOR kernel's HBM roofline and a CPU baseline (the fp32
C restatement of the reference, OpenMP on the host cores, on a bounded sample of the same workload).
Beside the headline: the projection residual the reference's 60 sweeps leave, the same workload over
its stated 1000 steps (timed in windows of 200), the same workload with the converged pressure solve
(multigrid-preconditioned CG to a residual tolerance), the ripple-and-droplet workload, the batched
tanks and the serving loop.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
suiri = importlib.import_module("hahenn-_amd")
scene = importlib.import_module("hahenn-_amd.scene")
workloads = importlib.import_module("hahenn-_amd.workloads")
residual = importlib.import_module("hahenn-_amd.residual")

HBM_PEAK_GBS = 8000.0


def progress(msg):
    """One line on stderr as each leg ends (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def build_sim(wl, quality="high"):
    g = suiri.Sim(suiri.config(quality, **wl["cfg"]))
    workloads.setup(g, wl)
    g.reorder()
    g.set_substeps(1)
    return g


def step(g, wl):
    """One frame; the workload's sources emit through the library's emitters (sl_step_items)."""
    g.step_items(wl["frame_dt"], wl["items"])


def sor_bytes_survey_model(nx, ny, nz):
    """The per-pass byte model of a reference-style pass that re-reads every face fraction:
    4C (p read) + 2C (p write, half) + 12N (fractions) + 2C (rhs, half) + 1C (flags) = 12N + 9C.
    the colour pass does not move these bytes (it reads stencil flags, and a cut cell's record of its
    fractions); reported beside the kernel's own count for comparison."""
    N = (nx + 1) * (ny + 1) * (nz + 1)
    C = nx * ny * nz
    return 12 * N + 9 * C


def sor_pass_bytes(g):
    """Bytes one SOR colour pass (k_sor_rec) moves, counted as the kernel reads and writes them, from the tank's
    state (cell counts and face fractions read back after the timed region): every slot of the colour
    loads its u16 stencil flags (2 B); each liquid cell of the colour reads its pressure and
    right-hand side and writes its pressure (12 B); the six neighbour loads of those cells touch
    other-colour pressures (4 B each, counted once); cut cells (an open face with 0 < f < 1) read
    their 32-byte record of face fractions and their sum. Returns (mean bytes of the two colours,
    parts)."""
    nx, ny, nz = g.nx, g.ny, g.nz
    slots = ((nx + 1) // 2) * ny * nz
    liquid = g.readback(suiri.F_CELLCOUNT).reshape(nz, ny, nx) > 0
    fr = g.readback(suiri.F_FRAC).reshape(nz + 1, ny + 1, nx + 1, 4)
    k, j, i = np.indices((nz, ny, nx), sparse=True)
    colour = (i + j + k) & 1
    faces = [fr[:nz, :ny, :nx, 0], fr[:nz, :ny, 1:, 0], fr[:nz, :ny, :nx, 1], fr[:nz, 1:, :nx, 1],
             fr[:nz, :ny, :nx, 2], fr[1:, :ny, :nx, 2]]
    cut = np.zeros_like(liquid)
    for f in faces:
        cut |= (f > 0) & (f != 1)
    cut &= liquid
    out = []
    for c in (0, 1):
        lc = liquid & (colour == c)
        near = np.zeros_like(lc)
        near[1:] |= lc[:-1]; near[:-1] |= lc[1:]
        near[:, 1:] |= lc[:, :-1]; near[:, :-1] |= lc[:, 1:]
        near[:, :, 1:] |= lc[:, :, :-1]; near[:, :, :-1] |= lc[:, :, 1:]
        cc = cut & (colour == c)
        parts = {"flags": 2 * slots, "liquid": 12 * int(lc.sum()), "neighbours": 4 * int(near.sum()),
                 "cut_records": 32 * int(cc.sum())}
        out.append(parts)
    mean = {k2: (out[0][k2] + out[1][k2]) / 2 for k2 in out[0]}
    return sum(mean.values()), {"per_pass_mean": mean, "per_colour": out, "liquid_cells": int(liquid.sum()),
                                "cut_cells": int(cut.sum())}


def calib_sor(passes):
    """Counter calibration on the SOR colour pass's own access pattern: the headline tank after a few steps, then
    `passes` times a 1 GiB device copy (which evicts the L2 and the Infinity Cache) followed by one
    colour-0 pass, so that every byte the pass reads comes from HBM. Prints the pass's byte count from
    sor_pass_bytes for rocprofv3's FETCH_SIZE / WRITE_SIZE of the same dispatches to be set against."""
    import torch
    wl = workloads.faucet()
    g = build_sim(wl)
    for _ in range(3):
        step(g, wl)
    g.sync()
    _, det = sor_pass_bytes(g)
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(passes):
        dst.copy_(src)
        torch.cuda.synchronize()
        g.run_stage(suiri.S_PRESSURE, 0, wl["frame_dt"])
        g.sync()
    c0 = det["per_colour"][0]
    print(json.dumps({"passes": passes, "colour": 0, "bytes": sum(c0.values()), "parts": c0,
                      "read_bytes": c0["flags"] + c0["liquid"] * 2 // 3 + c0["neighbours"] + c0["cut_records"],
                      "write_bytes": c0["liquid"] // 3}))


def cpu_baseline(wl, steps, threads):
    """The oracle (fp32 C restatement, OpenMP) on the same workload for `steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    c = wl["cfg"]
    q = O.quality("high", h=c["h"], iters=c["iters"], ripple_R=c["ripple_R"], cap=c["cap"], ww_cap=c["ww_cap"],
                  flip=c["flip"], omega=c["omega"])
    o = O.Oracle((c["tank_w"], c["tank_h"], c["tank_d"]), q, threads=threads)
    sc = scene.Scene(suiri.config("high", **c))   # host-side scene state of the same tank
    rows, _ = sc.pack(wl["items"], 0)
    o.set_objects(rows)
    inf = o.info()
    o.emit_now([scene.cell_block_emit(inf["h"], [inf["ox"], inf["oy"], inf["oz"]], *blk, seed=wl["seed"] + i)
                for i, blk in enumerate(wl["blocks"])])
    em = scene.Emitters(wl["items"], sc)
    o.step(wl["frame_dt"], nsub=1, emit_fn=em)   # first step includes GEOM: untimed
    t0 = time.perf_counter()
    psteps = 0
    for _ in range(steps):
        o.step(wl["frame_dt"], nsub=1, emit_fn=em)
        psteps += o.info()["count"]
    dt = time.perf_counter() - t0
    o.close()
    return psteps / dt, dt


TRAFFIC_KERNELS = ("k_sor_rec", "k_g2p", "k_p2g_tiles", "k_bin_count", "k_bin_scatter", "k_normforce")


def pmc_traffic(out_dir, timeout=150):
    """HBM bytes per launch of the SOR colour pass k_sor_rec (and of the G2P, P2G and NORMALIZE kernels,
    under "kernels") from two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) of a short bench, each
    scaled by what the same counter reads for a plain 256 MiB device copy. Child processes; returns None when rocprofv3 is
    missing or a pass fails."""
    import csv
    import shutil
    import statistics
    import subprocess
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    env = dict(os.environ, TMPDIR="/tmp")
    short = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "3", "--warmup", "1", "--no-cpu", "--no-pmc",
             "--only", "headline"]
    calib = [sys.executable, os.path.join(ROOT, "bench.py"), "--calib-copy-mb", "256"]
    regex = "|".join(TRAFFIC_KERNELS)
    runs = {"fetch": ("FETCH_SIZE", short, regex), "write": ("WRITE_SIZE", short, regex),
            "cfetch": ("FETCH_SIZE", calib, None), "cwrite": ("WRITE_SIZE", calib, None)}
    vals = {}
    for key, (counter, cmd, regex) in runs.items():
        d = os.path.join(out_dir, "pmc_" + key)
        args = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run"]
        if regex:
            args += ["--kernel-include-regex", regex]
        try:
            subprocess.run(args + ["--"] + cmd, env=env, cwd="/tmp", timeout=timeout, capture_output=True)
            rows = list(csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))))
        except Exception:
            return None
        if regex:
            per = {}
            for r in rows:
                name = r["Kernel_Name"]
                for k in TRAFFIC_KERNELS:
                    if name.startswith(k + "(") or name.startswith("void " + k + "<"):
                        per.setdefault(k, {}).setdefault(name, []).append(float(r["Counter_Value"]))
            if "k_sor_rec" not in per:
                return None
            # of several template variants (the tuner's trials of k_g2p<MINW>), the one the timed steps
            # ran: the most dispatches (a spilling trial variant writes scratch on top of the step's bytes)
            vals[key] = {k: statistics.mean(max(v.values(), key=len)) for k, v in per.items()}
        else:
            v = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
            if not v:
                return None
            vals[key] = statistics.mean(v)
    kib = 256 * 1024
    ff, wf = kib / vals["cfetch"], kib / vals["cwrite"]
    kern = {}
    for k in vals["fetch"]:
        if k in vals["write"]:
            fb, wb = vals["fetch"][k] * 1024 * ff, vals["write"][k] * 1024 * wf
            kern[k] = {"bytes": fb + wb, "fetch_bytes": fb, "write_bytes": wb}
    s0 = kern["k_sor_rec"]
    return {"bytes": s0["bytes"], "fetch_bytes": s0["fetch_bytes"], "write_bytes": s0["write_bytes"],
            "fetch_factor": ff, "write_factor": wf, "kernels": kern}


def projection_residual(g, dt):
    """The residual the pressure solve leaves after a step: the DIVERGENCE stage run again on the
    projected field gives, per liquid cell, its divergence minus the drift source the solve aimed at
    (src/js/30-wgsl-sim.js:240-264), in 1/s. Taken over liquid cells with an open face; a liquid cell
    whose six faces are all closed (inside an obstacle) has no pressure equation, so it is counted
    apart. `f64` evaluates the same formula in double precision from the stored fields (the
    divergence of the field itself, without the rounding of the f32 formula's subtractions);
    `above_f32_floor` is what remains of it above the per-cell bound of storing each face velocity
    in f32 (f * ulp(v) / 2 per face)."""
    g.run_stage(suiri.S_DIVERGENCE, 0, dt)
    div, liquid, has_open, floor = residual.divergence_f64(g)
    sel = (liquid & has_open).ravel()

    def stats(r):
        return {"max": float(r.max()) if r.size else 0.0, "p99": float(np.percentile(r, 99)) if r.size else 0.0,
                "rms": float(np.sqrt((r ** 2).mean())) if r.size else 0.0}

    out = {"unit": "1/s", "cells": int(sel.sum()), "closed_liquid_cells": int((liquid & ~has_open).sum())}
    out.update(stats(np.abs(g.readback(suiri.F_DIVG)[sel].astype(np.float64))))
    d = np.abs(div.ravel()[sel])
    fl = floor.ravel()[sel]
    out["f64"] = stats(d)
    out["f64"]["above_f32_floor"] = stats(np.maximum(d - fl, 0.0))
    out["f64"]["f32_floor_max"] = float(fl.max()) if fl.size else 0.0
    return out


def flow_leg(wl, steps=1000, window=200, warmup=5):
    """The faucet over its stated length (1000 steps after `warmup` untimed ones and the tuner), timed
    in windows of `window` steps: the step time grows as the stream fills the tank and spreads."""
    g = build_sim(wl)
    for _ in range(warmup):
        step(g, wl)
    g.autotune(reps=2)
    g.sync()
    wins = []
    t_all = time.perf_counter()
    for w0 in range(0, steps, window):
        n0 = g.stats()["count"]
        t0 = time.perf_counter()
        for _ in range(min(window, steps - w0)):
            step(g, wl)
        g.sync()
        dt = time.perf_counter() - t0
        k = min(window, steps - w0)
        n1 = g.stats()["count"]
        wins.append({"steps": [w0 + 1, w0 + k], "steps_per_s": k / dt, "particles": [n0, n1],
                     "particle_steps_per_s": (n0 + n1) / 2 * k / dt})
    total = time.perf_counter() - t_all
    st = g.stats()
    g.close()
    return {"steps": steps, "warmup": warmup, "seconds": total, "steps_per_s": steps / total,
            "gpu_errors": st["gpuErrors"], "windows": wins}


def converged_leg(wl, steps, warmup, tol, max_iters):
    """The workload with the converged pressure solve (multigrid-preconditioned CG to `tol` 1/s)."""
    g = build_sim(wl)
    for _ in range(warmup):
        step(g, wl)
    g.autotune(reps=2)
    g.set_pressure_solver(suiri.SOLVER_MGPCG, tol, max_iters)
    for _ in range(5):
        step(g, wl)
    c0 = g.stats()["count"]
    g.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(g, wl)
    g.sync()
    el = time.perf_counter() - t0
    c1 = g.stats()["count"]
    # after the timed region: iterations of a few more solves, then the residual of the last one
    its = []
    for _ in range(5):
        step(g, wl)
        its.append(g.pressure_solver_stats()["iters"])
    st = g.pressure_solver_stats()
    res = projection_residual(g, wl["frame_dt"])
    g.close()
    mean_p = (c0 + c1) / 2
    return {"solver": "mgpcg", "tol_per_s": tol, "max_iters": max_iters, "steps": steps,
            "ms_per_step": el / steps * 1e3, "steps_per_s": steps / el,
            "particle_steps_per_s": mean_p * steps / el, "iters_sampled": its,
            "last_solve": st, "divergence_after_projection": res}


def droplet_leg(steps, warmup):
    """The droplet layer (workloads.ripple_droplets): a 128^3 tank 0.1 m deep with 4096 sub-grid
    droplets per step on a 1024 x 1024 ripple layer of fixed depth 0.1 m (sigma 0.072, rho 1000),
    whitewater at K > 60, SOR omega 1.8 x 60, dt 1/240. Reports the coupled step, the droplet and
    ripple stages per frame, and the ripple frame's fraction of the HBM peak at its algorithmic
    176 M^2 + 16 (R+1)^2 bytes."""
    wl = workloads.ripple_droplets()
    g = build_sim(wl)
    for _ in range(warmup):
        step(g, wl)
    tuned = g.autotune(reps=2)
    g.timing(True)
    g.timing_read(reset=True)
    for _ in range(steps):
        step(g, wl)
    g.sync()
    tm = g.timing_read(reset=True)
    g.timing(False)
    c0 = g.stats()["count"]
    t0 = time.perf_counter()
    for _ in range(steps):
        step(g, wl)
    g.sync()
    el = time.perf_counter() - t0
    st = g.stats()
    g.close()
    R = wl["cfg"]["ripple_R"]
    M = 2 * R
    rb = 176 * M * M + 16 * (R + 1) ** 2
    rip_ms = tm["ripple"][0] / steps
    drop_ms = tm["droplets"][0] / steps
    return {"workload": "128^3 tank 0.5 m, 0.1 m deep, sphere; ripple M = 1024 (fixed depth 0.1 m, sigma 0.072, "
                        "rho 1000); 4096 droplets per step, lognormal D (median 1.5 mm, sigma_ln 0.4), v 1-4 m/s, "
                        "whitewater at K > 60; SOR 1.8 x 60; dt 1/240",
            "steps": steps, "ms_per_step": el / steps * 1e3, "steps_per_s": steps / el,
            "particles": int((c0 + st["count"]) / 2), "whitewater": st["ww"],
            "particle_steps_per_s": (c0 + st["count"]) / 2 * steps / el,
            "droplets_per_s": wl["cfg"]["droplets"] * steps / el,
            "stage_ms_per_step": {k: v[0] / steps for k, v in tm.items()},
            "droplet_ms_per_frame": drop_ms, "ripple_ms_per_frame": rip_ms,
            "ripple_roofline": {"bytes": rb, "GBps": rb / (rip_ms / 1e3) / 1e9,
                                "frac": rb / (rip_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
            "tuned_variants": tuned, "gpu_errors": st["gpuErrors"]}


def batched_leg(n_tanks, steps, warmup, seq_steps, graph_steps=50, seed=7):
    """64 independent tanks (grids 32^3 .. 128x64x64, fill 0.2 .. 0.7, 0 .. 3 primitives, 40 .. 80
    SOR iterations, dt = 1/120 s, re-orderings staggered over 16 frames) stepped as one batch, one
    launch per stage for all tanks, against the same tanks solved one after another on the plain
    path. Per-step device latency from hipEvent pairs. A block replayed from the captured hipGraph
    is measured after the eager block and reported beside it."""
    mix = workloads.tank_mix(n_tanks, seed=seed)
    dt = 1.0 / 120
    b = suiri.Batch([suiri.config("standard", **wl["cfg"]) for wl in mix])
    for t, wl in zip(b.tanks, mix):
        workloads.setup(t, wl)
    b.reorder()
    for _ in range(warmup):
        b.step(dt)
    b.sync()
    tuned = b.autotune(reps=1)

    def block(k):
        b.sync()
        p0 = sum(t.stats()["count"] for t in b.tanks)
        b.timing(True)
        t0 = time.perf_counter()
        for _ in range(k):
            b.step(dt)
        b.sync()
        el = time.perf_counter() - t0
        lat = b.step_times()
        b.timing(False)
        p1 = sum(t.stats()["count"] for t in b.tanks)
        return el, lat, (p0 + p1) / 2

    elapsed, lat, parts = block(steps)
    gbytes = sum(workloads.step_bytes(*wl["grid"], t.stats()["count"], wl["cfg"]["iters"])
                 for t, wl in zip(b.tanks, mix)) / 1e9
    graph = None
    if graph_steps:
        b.capture(dt)
        g_el, g_lat, g_parts = block(graph_steps)
        graph = {"steps": graph_steps, "ms_per_step": g_el / graph_steps * 1e3,
                 "step_latency_ms_p50": float(np.percentile(g_lat, 50)),
                 "value": g_parts * graph_steps / g_el,
                 "note": "measured after the eager block (later frames of the same scenes)"}
    b.close()
    value = parts * steps / elapsed
    # the same tanks, one plain solve after another (each tank: warm-up, then seq_steps timed steps)
    seq_time, seq_ps = 0.0, 0.0
    for wl in (mix if seq_steps else []):
        g = suiri.Sim(suiri.config("standard", **wl["cfg"]))
        workloads.setup(g, wl)
        g.reorder()
        g.set_substeps(1)
        for _ in range(2):
            g.step(dt)
        g.sync()
        c0 = g.stats()["count"]
        t1 = time.perf_counter()
        for _ in range(seq_steps):
            g.step(dt)
        g.sync()
        seq_time += time.perf_counter() - t1
        seq_ps += (c0 + g.stats()["count"]) / 2 * seq_steps
        g.close()
    seq_value = seq_ps / seq_time if seq_steps else None
    return {
        "workload": f"{n_tanks} tanks, grids 32^3..128x64x64, fill 0.2-0.7, 0-3 primitives, SOR 40-80 iters, "
                    f"dt 1/120, whitewater on, ripple M=256",
        "tanks": n_tanks, "particles": int(parts), "steps": steps, "launch": "eager, one launch per stage",
        "value": value, "unit": "particle-steps/s",
        "ms_per_step": elapsed / steps * 1e3,
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                            "max": float(lat.max()), "source": "hipEvent pairs around each batch step"},
        "algorithmic_GB_per_step": gbytes,
        "achieved_GBps": gbytes / (elapsed / steps),
        "graph_replay": graph,
        "sequential": {"value": seq_value, "unit": "particle-steps/s", "steps_per_tank": seq_steps,
                       "method": "each tank alone on the plain path, one after another"},
        "speedup_vs_sequential": value / seq_value if seq_value else None,
        "tuned_variants": tuned,
    }


KIND_NAMES = {0: "dam-break 64x32x32", 1: "faucet 256x128x128", 3: "batched-mix tanks"}


def serve_leg(rate=20.0, window=3.0, limit=45.0, steps=(100, 2000), kinds=(0, 1, 3), seed=11, budget_gb=200,
              max_jobs=128, frames_per_tick=4):
    """Serving stream: scene-bake requests arrive as a Poisson process (`rate` per second for `window`
    seconds), each scene drawn from the 64x32x32 dam break (0), the 256x128x128 faucet (1) and the
    batched-tank mix (3), each with 100..2000 frames; admitted against a `budget_gb` device-memory
    budget, stepped together as a continuous batch. The run stops when the stream drains or at
    `limit` seconds. Request latency = submit -> done on the server clock; per-frame latency from
    hipEvents around each batched frame."""
    rng = np.random.default_rng(seed)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window:
            break
        arrivals.append(t)
    reqs = []
    for i, ta in enumerate(arrivals):
        kind = int(rng.choice(kinds))
        n = int(rng.integers(steps[0], steps[1] + 1))
        wl = workloads.dam_break() if kind == 0 else (workloads.faucet() if kind == 1 else
                                                        workloads.mix_tank(seed * 7919 + i))
        reqs.append(dict(t=ta, kind=kind, steps=n, wl=wl))
    srv = suiri.Server(budget=int(budget_gb * (1 << 30)), max_jobs=max_jobs, frames_per_tick=frames_per_tick)
    ids = {}
    t0 = time.perf_counter()
    i = 0
    while True:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i]["t"] <= now:
            r = reqs[i]
            items, fill = workloads.job_inputs(r["wl"])
            q = "high" if r["kind"] == 1 else "standard"
            ids[srv.submit(suiri.config(q, **r["wl"]["cfg"]), items=items, fill=fill, steps=r["steps"],
                           dt=r["wl"]["frame_dt"])] = i
            i += 1
        running = srv.tick()
        now = time.perf_counter() - t0
        st = srv.stats()
        if now > limit or (i == len(reqs) and st["queued"] == 0 and running == 0):
            break
        if running == 0 and st["queued"] == 0 and i < len(reqs):
            time.sleep(max(0.0, reqs[i]["t"] - now))
    elapsed = time.perf_counter() - t0
    st = srv.stats()
    infos = [srv.info(j) for j in ids]
    srv.close()
    done = [x for x in infos if x["status"] == "done"]
    lat = np.array([x["t_done"] - x["t_submit"] for x in done]) if done else np.zeros(1)
    # every request: an unfinished one counts with the time it had waited at the cut-off, a lower bound
    # of its latency (right-censored)
    lat_all = np.array([(x["t_done"] if x["status"] == "done" else st["uptime"]) - x["t_submit"] for x in infos])
    # sequential plain solves of the same requests: per-kind time per frame measured alone on the plain
    # path (faucet and dam break at their sizes, batched-mix tanks on a sample), applied to each request
    def plain_rate(wl, q, k=8):
        g = build_sim(wl, q)
        for _ in range(2):
            step(g, wl)
        g.sync()
        c0 = g.stats()["count"]
        t1 = time.perf_counter()
        for _ in range(k):
            step(g, wl)
        g.sync()
        dtm = (time.perf_counter() - t1) / k
        c1 = g.stats()["count"]
        g.close()
        return dtm, (c0 + c1) / 2
    per_kind = {}
    if 0 in kinds:
        per_kind[0] = plain_rate(workloads.dam_break(), "standard", 20)
    if 1 in kinds:
        per_kind[1] = plain_rate(workloads.faucet(), "high", 8)
    if 3 in kinds:
        sample = [plain_rate(workloads.mix_tank(seed * 7919 + j), "standard", 10) for j in range(6)]
        per_kind[3] = (sum(x[0] for x in sample) / len(sample), sum(x[1] for x in sample) / len(sample))
    seq_t = sum(r["steps"] * per_kind[r["kind"]][0] for r in reqs)
    seq_ps = sum(r["steps"] * per_kind[r["kind"]][1] for r in reqs)
    served = st["particle_steps"] / elapsed
    return {
        "workload": f"Poisson {rate}/s for {window} s, scenes from "
                    + " / ".join(KIND_NAMES[k] for k in kinds)
                    + f", {steps[0]}-{steps[1]} frames each, {budget_gb} GB budget, up to {frames_per_tick} "
                    f"frames per tick, stopped at {limit} s or when drained",
        "requests": len(reqs), "completed": len(done),
        "unfinished": len(reqs) - len(done) - sum(x["status"] in ("cancelled", "refused") for x in infos),
        "refused": sum(x["status"] == "refused" for x in infos),
        "elapsed_s": elapsed,
        "value": served, "unit": "particle-steps/s",
        "frames": st["frames"],
        "frame_latency_ms": {"p50": st["step_ms_p50"], "p99": st["step_ms_p99"]},
        "device_busy": st["step_ms_sum"] / 1e3 / elapsed,
        "request_latency_s": {"p50": float(np.percentile(lat_all, 50)), "p99": float(np.percentile(lat_all, 99)),
                              "of": "all requests; unfinished ones censored at the cut-off (lower bounds)",
                              "censored": int(sum(x["status"] != "done" for x in infos)),
                              "completed_p50": float(np.percentile(lat, 50)),
                              "completed_p99": float(np.percentile(lat, 99))},
        "sequential_estimate": {"value": seq_ps / seq_t, "unit": "particle-steps/s",
                                "method": "each request's frames at the per-frame time of its scene run alone "
                                          "on the plain path"},
        "speedup_vs_sequential": served / (seq_ps / seq_t),
        "offered_load_x_capacity": (seq_ps / max(served, 1.0)) / window,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", default="256x128x128")
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle steps timed for cpu_baseline (~1 s each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ripple", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter runs for `traffic`")
    ap.add_argument("--only", choices=["all", "headline", "batched", "serve", "converged", "droplets", "flow"], default="all",
                    help="run one leg (its JSON object alone) instead of the full line")
    ap.add_argument("--conv-tol", type=float, default=1e-4, help="converged leg: residual tolerance (1/s)")
    ap.add_argument("--conv-iters", type=int, default=24, help="converged leg: CG iterations enqueued per solve")
    ap.add_argument("--conv-steps", type=int, default=60)
    ap.add_argument("--drop-steps", type=int, default=100, help="droplet leg: timed steps")
    ap.add_argument("--batch-tanks", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--seq-steps", type=int, default=10)
    ap.add_argument("--graph-steps", type=int, default=50)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-window", type=float, default=3.0)
    ap.add_argument("--serve-limit", type=float, default=45.0)
    ap.add_argument("--serve-kinds", default="0,1,3")
    ap.add_argument("--serve-fpt", type=int, default=4, help="serving: frames per tick")
    ap.add_argument("--calib-sor", type=int, default=0,
                    help="only run this many HBM-cold SOR colour passes of the headline tank (counter calibration) and exit")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.calib_copy_mb:
        import torch
        n = args.calib_copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_read": n * 4, "calib_copy_bytes_written": n * 4, "launches": 3}))
        return
    if args.calib_sor:
        calib_sor(args.calib_sor)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only (see DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.only == "serve":
        print(json.dumps(serve_leg(args.serve_rate, args.serve_window, args.serve_limit,
                                   kinds=tuple(int(k) for k in args.serve_kinds.split(",")),
                                  frames_per_tick=args.serve_fpt)))
        return
    if args.only == "batched":
        print(json.dumps(batched_leg(args.batch_tanks, args.batch_steps, 5, args.seq_steps, args.graph_steps)))
        return
    if args.only == "droplets":
        print(json.dumps(droplet_leg(args.drop_steps, args.warmup)))
        return
    grid = tuple(int(x) for x in args.grid.split("x"))
    wl = workloads.faucet(grid=grid, R=0 if args.no_ripple else 512)
    if args.only == "flow":
        print(json.dumps(flow_leg(wl)))
        return
    if args.only == "converged":
        print(json.dumps(converged_leg(wl, args.conv_steps, args.warmup, args.conv_tol, args.conv_iters)))
        return
    g = build_sim(wl)
    for _ in range(args.warmup):
        step(g, wl)
    tuned = g.autotune(reps=2)
    st0 = g.stats()
    g.timing(True)
    g.timing_read(reset=True)
    g.sync()
    t0 = time.perf_counter()
    counts = []
    for i in range(args.steps):
        step(g, wl)
        if i % 25 == 24:
            counts.append(g.stats()["count"])
    g.sync()
    elapsed = time.perf_counter() - t0
    st1 = g.stats()
    tm = g.timing_read(reset=True)
    g.timing(False)
    # timed region without event overhead
    g.sync()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step(g, wl)
    g.sync()
    elapsed_plain = time.perf_counter() - t1
    st2 = g.stats()
    # after the timed region: the residual the pressure solve left in the last step, i.e. the divergence
    # of the projected field minus the drift source it was solved for, over the liquid cells (1/s)
    resid = projection_residual(g, wl["frame_dt"])
    mean_p = (st1["count"] + st2["count"]) / 2
    ms_per_step = elapsed_plain / args.steps * 1e3
    value = mean_p * args.steps / elapsed_plain
    nx, ny, nz = g.nx, g.ny, g.nz
    iters = wl["cfg"]["iters"]
    sor_ms, sor_launches = tm["sor"]
    sor_launch_s = sor_ms / 1e3 / max(sor_launches, 1)
    bpl, bpl_detail = sor_pass_bytes(g)
    achieved = bpl / sor_launch_s / 1e9
    # per-stage fractions of the HBM peak with the stage byte counts of the step's roofline table
    # (P2G 28P + 24N + 4C, its binning another 21P; G2P 64P + 24N + 4C; a ripple frame
    # 176 M^2 + 16 (R+1)^2)
    N = (nx + 1) * (ny + 1) * (nz + 1)
    C = nx * ny * nz
    M, R = 2 * wl["cfg"]["ripple_R"], wl["cfg"]["ripple_R"]
    stage_bytes = {"p2g": 28 * mean_p + 24 * N + 4 * C, "g2p": 64 * mean_p + 24 * N + 4 * C,
                   "ripple": 176 * M * M + 16 * (R + 1) ** 2}
    stage_frac = {}
    for k, b in stage_bytes.items():
        ms = tm[k][0] / args.steps
        stage_frac[k] = {"bytes": b, "ms_per_step": ms, "GBps": b / (ms / 1e3) / 1e9 if ms else None,
                         "frac": b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS if ms else None}
    stage_frac["p2g"]["bytes_with_binning"] = stage_bytes["p2g"] + 21 * mean_p
    stage_frac["sor"] = {"bytes_per_pass": bpl, "us_per_pass": sor_launch_s * 1e6, "frac": achieved / HBM_PEAK_GBS}
    out = {
        "metric": "particle-steps/s at 256x128x128; SOR HBM GB/s; batched p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "steps_per_s": 1e3 / ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (half-full tank from the reference's lattice emitter, seeded)",
        "config": {"workload": wl["name"], "grid": list(grid), "particles": int(mean_p), "sor_iters": iters,
                   "omega": 1.8, "dt": wl["frame_dt"], "ripple_M": 2 * wl["cfg"]["ripple_R"],
                   "whitewater": True},
        "roofline": {"bound": "hbm", "kernel": "k_sor_rec (one colour pass)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": bpl, "bytes_detail": bpl_detail,
                     "avg_launch_us": sor_launch_s * 1e6,
                     "survey_model_bytes_per_launch": sor_bytes_survey_model(nx, ny, nz)},
        "sor_GBps": iters * 2 * bpl / (sor_ms / 1e3 / args.steps) / 1e9,
        "stage_roofline": stage_frac,
        "stage_ms_per_step": {k: v[0] / args.steps for k, v in tm.items()},
        "step_algorithmic_GB": workloads.step_bytes(nx, ny, nz, mean_p, iters) / 1e9,
        "divergence_after_projection": resid,
        "gpu_errors": st2["gpuErrors"],
        "tuned_variants": tuned,
    }
    g.close()
    progress(f"headline done: {out['steps_per_s']:.1f} steps/s")
    if args.only == "all":
        out["faucet_1000_steps"] = flow_leg(wl)
        progress("1000-step faucet done")
        out["converged"] = converged_leg(wl, args.conv_steps, args.warmup, args.conv_tol, args.conv_iters)
        progress("converged solve done")
        out["ripple_droplets"] = droplet_leg(args.drop_steps, args.warmup)
        progress("droplet layer done")
        out["batched"] = batched_leg(args.batch_tanks, args.batch_steps, 5, args.seq_steps, args.graph_steps)
        out["batched_p99_step_ms"] = out["batched"]["step_latency_ms"]["p99"]
        progress("batched tanks done")
        out["served"] = serve_leg(args.serve_rate, args.serve_window, args.serve_limit,
                                  kinds=tuple(int(k) for k in args.serve_kinds.split(",")),
                                  frames_per_tick=args.serve_fpt)
        progress("served stream done")
        # the same stream without the 256x128x128 faucet jobs, each of which fills the GPU on its own:
        # the scenes that a continuous batch can run side by side
        out["served_small_scenes"] = serve_leg(args.serve_rate, args.serve_window, args.serve_limit, kinds=(0, 3),
                                               frames_per_tick=args.serve_fpt)
        progress("served small-scene stream done")
    if not args.no_pmc:
        t = pmc_traffic(os.path.join(ROOT, "gpurun_out", "bench_pmc"))
        if t:
            out["roofline"]["traffic"] = t["bytes"]
            out["roofline"]["traffic_detail"] = t
            # the bandwidth the pass draws past L2 (counter bytes over the same launch time)
            rl = out["roofline"]
            rl["traffic_GBps"] = t["bytes"] / (rl["avg_launch_us"] * 1e-6) / 1e9
            rl["traffic_frac"] = rl["traffic_GBps"] / HBM_PEAK_GBS
            # HBM bytes per launch of the P2G and G2P kernels beside their algorithmic bytes
            kk = t.get("kernels", {})
            sr = out["stage_roofline"]
            if "k_g2p" in kk:
                sr["g2p"]["traffic"] = kk["k_g2p"]["bytes"]
            p2g = [kk[k]["bytes"] for k in ("k_bin_count", "k_bin_scatter", "k_p2g_tiles") if k in kk]
            if len(p2g) == 3:
                sr["p2g"]["traffic"] = sum(p2g)
                sr["p2g"]["traffic_detail"] = {k: kk[k]["bytes"] for k in ("k_bin_count", "k_bin_scatter", "k_p2g_tiles")}
        progress("counter runs done")
    if not args.no_cpu:
        # every core the process may use: the host's share for one GPU sets OMP_NUM_THREADS, else nproc
        nproc = os.cpu_count() or 1
        threads = int(os.environ.get("OMP_NUM_THREADS") or nproc)
        v, secs = cpu_baseline(wl, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": v, "unit": "particle-steps/s", "cores": threads, "nproc": nproc,
                               "kind": "port",
                               "sample": f"{args.cpu_steps} steps of the same workload after 1 untimed step "
                                         f"({secs:.1f} s), oracle/suiri_oracle.c with OpenMP on {threads} threads "
                                         f"(OMP_NUM_THREADS: the cores granted to this GPU's host share; nproc "
                                         f"{nproc})"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
