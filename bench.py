#!/usr/bin/env python3
"""Benchmark of the gfx950 WaterLily time step (one GPU).

A step is one sim_step! (mom_step!: predictor, two pressure projections with the multigrid to
Σr² < 1e-4, corrector, CFL) of the 3D flow past a sphere, 512x256x256 interior cells (the largest
single-GPU configuration), Float32, U=(1,0,0), radius 32 at (128,128,128), Re=1e4 (ν = U·64/Re),
static body (remeasure=false, which gives the same fields as remeasuring a static body), timed
after warm-up steps past the impulsive start.  Inputs are resident in HBM before timing.

Prints one JSON line: Mcell-steps/s (interior cells x steps / s), the per-kernel roofline of the
dominant kernel from HIP events recorded on the library's stream during the timed steps, and the
oracle's CPU rate on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from __graft_entry__ import ROOT, load_package

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-steps/s & sim_step!/s, plain vs learned-warm-start MG at equal residual"


def sphere_sim(wl, n, R, center, Re, L):
    return wl.Simulation(n, (1, 0, 0), L, ν=L / Re, body=wl.Sphere(center, R), T=np.float32)


def run_gpu(wl, n, R, center, Re, L, steps, warmup, profile_steps=None, mode=None, tune=False, tiles=None):
    """Warm up, choose the tile targets with the autotuner (set-up, before the timed region), time
    `steps` sim_step!s, then run `profile_steps` more with per-kernel HIP events recorded on the
    library's stream (the events add launch gaps, so they stay out of the timed steps)."""
    sim = sphere_sim(wl, n, R, center, Re, L)
    if mode is not None:
        sim.pois.set_solver(mode=mode)
    if tiles is not None:  # the targets of an earlier tuned run, set without timing (for profiling)
        sim.flow.set_tiles(*tiles)
    elif tune:
        tiles = sim.flow.autotune(5)
    for _ in range(warmup):
        wl.sim_step(sim, remeasure=False)
    n0 = len(sim.pois.n)
    t0 = time.perf_counter()
    for _ in range(steps):
        wl.sim_step(sim, remeasure=False)
    t1 = time.perf_counter()  # every sim_step returns after its CFL reduction reached the host
    nv = sim.pois.n[n0:]
    stats, psec = [], None
    ps = steps if profile_steps is None else profile_steps
    if ps > 0:
        sim.flow.profile(True)
        p0 = time.perf_counter()
        for _ in range(ps):
            wl.sim_step(sim, remeasure=False)
        psec = (time.perf_counter() - p0) / ps
        stats = sim.flow.profile_read()
        sim.flow.profile(False)
    cells = int(np.prod(n))
    sec = t1 - t0
    return dict(sec=sec, ms_per_step=1e3 * sec / steps, mcells=cells * steps / sec / 1e6, steps_per_s=steps / sec,
                vcycles_per_step=sum(nv) / steps, stats=stats, sim=sim, tiles=tiles,
                profiled_ms_per_step=None if psec is None else 1e3 * psec)


def c5_specs(n=64, seed0=0):
    """64 uneven 2D flows (seeds 0..63): grid from {64x64, 128x64, 256x128, 512x256}, one circle or
    ellipse (radius U[m/16, m/6], m the smaller side, centred at (N0/4, N1/2)), Re U[100, 1000] on
    the diameter (nu = 2R/Re), U = (1, 0), Float32.  Each spec: grid, body (kind 1 circle with
    radius, or kind 2 ellipse with axes), R and nu."""
    sizes = [(64, 64), (128, 64), (256, 128), (512, 256)]
    specs = []
    for k in range(n):
        rng = np.random.default_rng(seed0 + k)
        N = sizes[int(rng.integers(0, 4))]
        m = min(N)
        R = float(rng.uniform(m / 16, m / 6))
        c = (N[0] / 4, N[1] / 2)
        if rng.uniform() < 0.5:
            body = {"kind": 1, "center": c, "radius": R}
        else:
            body = {"kind": 2, "center": c, "axes": (R, float(rng.uniform(0.4, 1.0)) * R)}
        Re = float(rng.uniform(100, 1000))
        specs.append(dict(N=N, body=body, R=R, nu=2 * R / Re))
    return specs


def c5_members(wl, n=64, seed0=0):
    """The flows of c5_specs as Simulations."""
    sims = []
    for sp in c5_specs(n, seed0):
        b = sp["body"]
        body = wl.Sphere(b["center"], b["radius"]) if b["kind"] == 1 else wl.Ellipsoid(b["center"], b["axes"])
        sims.append(wl.Simulation(sp["N"], (1, 0), 2 * sp["R"], ν=sp["nu"], body=body, T=np.float32))
    return sims


def run_c5(wl, steps, warmup):
    """Aggregate throughput of the 64 uneven flows stepped as one batch, and stepped one at a time."""
    sims = c5_members(wl)
    cells = sum(int(np.prod(s.flow.N)) for s in sims)
    for s in sims:  # tile targets per member (set-up), used by the batch's block tables
        s.flow.autotune(3)
    b = wl.Batch(sims)
    for _ in range(warmup):
        b.step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        b.step()
        lat.append(time.perf_counter() - a)
    sec = time.perf_counter() - t0
    nv = np.mean([np.mean(s.pois.n[-2 * steps:]) for s in sims])
    del b
    solo = c5_members(wl)
    for s in solo:
        s.flow.autotune(3)
        for _ in range(warmup):
            wl.sim_step(s, remeasure=False)
    s0 = time.perf_counter()
    for s in solo:
        for _ in range(steps):
            wl.sim_step(s, remeasure=False)
    ssec = time.perf_counter() - s0
    lat_ms = np.array(lat) * 1e3
    return dict(workload="64 uneven 2D flows (64x64..512x256, circles/ellipses, Re 100-1000), Float32, one batch",
                sims=len(sims), cells=cells, steps=steps, mcells=round(cells * steps / sec / 1e6, 2),
                ms_per_batch_step=round(1e3 * sec / steps, 3),
                step_latency_ms_p50=round(float(np.percentile(lat_ms, 50)), 3),
                step_latency_ms_p99=round(float(np.percentile(lat_ms, 99)), 3),
                vcycles_per_projection=round(float(nv), 3),
                one_at_a_time_mcells=round(cells * steps / ssec / 1e6, 2),
                batch_speedup=round(ssec / sec, 2))


def roofline(stats, traffic_file):
    rows = []
    for s in stats:
        if s["launches"] == 0:
            continue
        avg = s["total_ms"] / s["launches"]
        gbs = s["bytes_per_launch"] / (avg * 1e-3) / 1e9 if s["bytes_per_launch"] > 0 else None
        rows.append(dict(kernel=f'{s["name"]}@L{s["level"]}', launches=s["launches"], total_ms=round(s["total_ms"], 3),
                         avg_us=round(avg * 1e3, 2), alg_MB=round(s["bytes_per_launch"] / 1e6, 3),
                         GBs=None if gbs is None else round(gbs, 1)))
    rows.sort(key=lambda r: -r["total_ms"])
    dom = next(r for r in rows if r["GBs"] is not None)
    traffic = None
    if traffic_file and os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tt = json.load(fh)
        traffic = tt.get(dom["kernel"])
        traffic = None if traffic is None else round(traffic)
    total_b = sum(s["bytes_per_launch"] * s["launches"] for s in stats)
    total_ms = sum(s["total_ms"] for s in stats)
    return dict(bound="hbm", kernel=dom["kernel"], achieved=dom["GBs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(dom["GBs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                step_alg_GBs=round(total_b / (total_ms * 1e-3) / 1e9, 1) if total_ms > 0 else None), rows


def run_cpu(n, R, center, Re, L, budget_s=15.0, label="512x256x256"):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import wl_oracle as O
    threads = os.cpu_count() or 1
    O.lib().wlo_threads(threads)
    sim = O.OSimulation(n, (1, 0, 0), L, nu=L / Re, bodies=[{"kind": 1, "center": center, "radius": R}])
    sim.step(remeasure=False)  # warm-up
    k, t0 = 0, time.perf_counter()
    while True:
        sim.step(remeasure=False)
        k += 1
        el = time.perf_counter() - t0
        if (el > budget_s and k >= 2) or k >= 200:
            break
    return dict(value=round(int(np.prod(n)) * k / el / 1e6, 3), unit="Mcell-steps/s", cores=threads, kind="port",
                sample=f"{k} sim_step! of the same {label} sphere flow after 1 warm-up step "
                       f"({el:.1f} s), oracle built -O3 -fopenmp on {threads} threads")


def run_c1(wl, steps=200):
    """The 2D circle the reference runs on its CPU path (192x128, radius 16 at (63,63), Re 250 on
    the diameter, Float32, from the default initial condition u = uBC): `steps` sim_step!s on the
    GPU and on the oracle's CPU path, which must give the same trajectory (V-cycle counts and the
    final pressure force compared)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import wl_oracle as O
    n, cells = (192, 128), 192 * 128
    sim = wl.Simulation(n, (1, 0), 32, ν=32 / 250, body=wl.Sphere((63, 63), 16), T=np.float32)
    t0 = time.perf_counter()
    for _ in range(steps):
        wl.sim_step(sim, remeasure=False)
    g = time.perf_counter() - t0
    threads = os.cpu_count() or 1
    O.lib().wlo_threads(threads)
    osim = O.OSimulation(n, (1, 0), 32, nu=32 / 250, bodies=[{"kind": 1, "center": (63, 63), "radius": 16}])
    t0 = time.perf_counter()
    for _ in range(steps):
        osim.step(remeasure=False)
    c = time.perf_counter() - t0
    F, Fo = wl.pressure_force(sim), osim.pressure_force()
    return dict(workload="2D circle 192x128, r=16 at (63,63), Re=250, Float32, 200 sim_step! from u = uBC",
                steps=steps, gpu_mcells=round(cells * steps / g / 1e6, 2), gpu_ms_per_step=round(1e3 * g / steps, 3),
                cpu_mcells=round(cells * steps / c / 1e6, 2), cpu_ms_per_step=round(1e3 * c / steps, 3),
                cpu_threads=threads, same_vcycle_counts=sim.pois.n == osim.n,
                pressure_force_gpu=[round(float(v), 6) for v in F], pressure_force_cpu=[round(float(v), 6) for v in Fo])


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


# rocprofv3 kernel names -> the library's profiler keys (level 0 = the largest dispatch of a name)
PROF_KEYS = {"k_inc_restrict_full": "inc_restrict", "k_mg_tail": "mg_tail", "k_mg_tail_lds": "mg_tail", "k_bdim2": "bdim2", "k_residual": "residual", "k_submean_l2": "submean_l2",
             "k_submean_l2_cond": "submean_l2",
             "k_inc_restrict": "inc_restrict", "k_gs_half": "gs_half", "k_increment": "increment",
             "k_prolong_inc": "prolong_inc", "k_eps_init": "eps_init", "k_correct": "correct", "k_cfl": "cfl",
             # the x-pair multigrid kernels (two cells per lane)
             "k_gs_half2": "gs_half", "k_increment2": "increment", "k_inc_restrict2": "inc_restrict",
             "k_prolong_inc2": "prolong_inc", "k_gs_half_s": "gs_half", "k_gs_dbl": "gs_dbl", "k_smooth4": "smooth4",
             "k_residual64": "residual64",
             # BDIM!'s rows the convection kernel leaves (a list of rows), and their flags
             "k_bdim2_rows": "bdim2", "k_conv_rows": "conv_rows", "k_copy_exit_plane": "exit_plane"}


def _key_of(kname):
    for k in ("k_conv_lds", "k_conv_bdim1"):  # template arguments <T, scheme, PRED, ...>
        if f"wl::{k}<" in kname:
            args = kname.split(f"wl::{k}<")[1].split(">")[0].split(",")
            return "conv_pred" if args[2].strip() == "true" else "conv_corr"
    for k, v in PROF_KEYS.items():
        if f"wl::{k}<" in kname or f"wl::{k}(" in kname:
            return v
    return None


def summarize(stats_csv, fetch_csv, write_csv, out_md, out_json, cmd, calib_bytes=1 << 30):
    """Write the kernel-time summary and the HBM traffic per level-0 launch (FETCH_SIZE/WRITE_SIZE
    scaled by what they read for a device copy of calib_bytes)."""
    import collections
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    lines = [f"# rocprofv3 kernel statistics", "", f"Command: `{cmd}`", "",
             "| kernel | calls | avg us | total ms | % |", "|---|---|---|---|---|"]
    for r in rows[:24]:
        nm = r["Name"].split("(")[0].replace("void ", "")
        lines.append(f"| {nm} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {r['Percentage']} |")

    for tag, title in (("wl::nn::", "warm-start network kernels"), ("wl::kb_", "batched kernels")):
        sel = [r for r in rows if tag in r["Name"]]
        if sel:
            lines += ["", f"## All {title}", "", "| kernel | calls | avg us | total ms |", "|---|---|---|---|"]
            for r in sel:
                nm = r["Name"].split("(")[0].replace("void ", "")
                lines.append(f"| {nm} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                             f"{float(r['TotalDurationNs']) / 1e6:.3f} |")

    trace_csv = stats_csv.replace("_kernel_stats.csv", "_kernel_trace.csv")
    if os.path.exists(trace_csv):
        grids = collections.defaultdict(list)
        for r in csv.DictReader(open(trace_csv)):
            k = _key_of(r["Kernel_Name"])
            if k:
                grids[(k, r["Kernel_Name"].split("(")[0].replace("void ", ""))].append(
                    (int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]),
                     (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
        lines += ["", "## Level-0 dispatches (the largest grid of each kernel, from the kernel trace)", "",
                  "| profiler key | kernel | dispatches | avg us |", "|---|---|---|---|"]
        # a key's level 0 is its largest grid over all its kernels (an instance that runs on coarser
        # levels only is listed as such)
        kmax = collections.defaultdict(int)
        for (k, kn), v in grids.items():
            kmax[k] = max(kmax[k], max(g for g, _ in v))
        # of the kernels of a key that reach its largest grid, the longest-running one is level 0's,
        # and so is any other within 3x of it (the z-marching tiles can give a coarser level's kernel,
        # at least 4x cheaper, the same grid)
        klead = {}
        for (k, kn), v in grids.items():
            d = np.mean([dd for g, dd in v if g == kmax[k]]) if any(g == kmax[k] for g, _ in v) else -1.0
            if d > klead.get(k, (None, -1.0))[1]:
                klead[k] = (kn, d)
        for (k, kn), v in sorted(grids.items(), key=lambda kv: -sum(d for g, d in kv[1])):
            gmax = max(g for g, _ in v)
            d0 = sorted(d for g, d in v if g == gmax)
            note = ""
            # the z-marching tiles aim every level at the same block count, so two levels can share
            # the largest grid: split at the widest gap (log scale) when the two groups differ 3x+
            if len(d0) >= 4:
                gaps = [np.log(d0[i + 1] / max(d0[i], 1e-9)) for i in range(len(d0) - 1)]
                i = int(np.argmax(gaps))
                lo, hi = d0[:i + 1], d0[i + 1:]
                if np.mean(hi) > 3 * np.mean(lo):
                    note = f" (split from {len(lo)} same-grid dispatches of a coarser level, avg {np.mean(lo):.2f} us)"
                    d0 = hi
            lvl0 = gmax == kmax[k] and (klead[k][0] == kn or 3 * np.mean(d0) >= klead[k][1])
            label = f"{k}@L0" if lvl0 else f"{k} (coarser levels only)"
            if k == "mg_tail":  # one workgroup that runs the coarse levels, never level 0
                label = "mg_tail (coarse levels, one workgroup)"
            lines.append(f"| {label} | {kn} | {len(d0)} | {np.mean(d0):.2f}{note} |")

    def counters(fn, cname):
        per = collections.defaultdict(list)
        copies = []
        for r in csv.DictReader(open(fn)):
            if r["Counter_Name"] != cname:
                continue
            v = float(r["Counter_Value"])
            if "copyBuffer" in r["Kernel_Name"] and v > 1000:
                copies.append(v)
            k = _key_of(r["Kernel_Name"])
            if k:
                per[(k, r["Kernel_Name"])].append(v)
        return per, copies

    if fetch_csv == "-" or write_csv == "-":  # kernel-trace summary only
        with open(out_md, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        return
    fetch, fc = counters(fetch_csv, "FETCH_SIZE")
    write, wc = counters(write_csv, "WRITE_SIZE")
    f_scale = calib_bytes / (float(np.median(fc)) * 1024) if fc else None
    w_scale = calib_bytes / (float(np.median(wc)) * 1024) if wc else None
    traffic = collections.defaultdict(list)
    lines += ["", f"## HBM traffic per level-0 launch (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", "",
              f"Copy calibration: a {calib_bytes >> 20} MiB device copy reads FETCH_SIZE = "
              f"{np.median(fc) if fc else float('nan'):.0f} KiB and WRITE_SIZE = {np.median(wc) if wc else float('nan'):.0f} KiB, "
              f"so bytes = FETCH_SIZE x 1024 x {f_scale:.3f} and WRITE_SIZE x 1024 x {w_scale:.3f}.", "",
              "| kernel | fetch MB | write MB | traffic MB |", "|---|---|---|---|"]
    for (k, kn), vals in sorted(fetch.items()):
        fv = max(vals) * 1024 * f_scale
        wv = max(write.get((k, kn), [0.0])) * 1024 * w_scale
        traffic[k].append(fv + wv)
        lines.append(f"| {kn.split('(')[0].replace('void ', '')} | {fv / 1e6:.1f} | {wv / 1e6:.1f} | {(fv + wv) / 1e6:.1f} |")
    # the level-0 launch of a key moves the most bytes of all its kernels' launches
    tj = {f"{k}@L0": float(np.max(v)) for k, v in traffic.items()}
    with open(out_md, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(out_json, "w") as fh:
        json.dump(dict(tj, _fetch_scale=f_scale, _write_scale=w_scale), fh, indent=1)


WARMNET = os.path.join(ROOT, "waterlily.jl_amd", "data", "warmnet_c4.npz")
# algorithmic flops of one network forward per level-1 voxel: 2·27·Σ ci·co over the real channels
# of each layer, weighted by the layer's voxel share (1, 1/8, 1/64, 1/512)
NET_FLOPS_PER_VOXEL = 2 * 27 * (3 * 16 + 16 * 16 + 16 * 16 + 16 * 1 + (16 * 32 + 32 * 16) / 8 + (32 * 64 + 64 * 32) / 64
                                + 64 * 64 / 512)


def train_warmstart(wl, out_path, nseeds, steps):
    """Train the warm-start network on pairs generated by random-ellipsoid flows (seeds 100..)."""
    n = (256, 128, 128)
    net = wl.WarmNet((128, 64, 64), replay_pairs=512)
    t0 = time.perf_counter()
    losses = wl.learned.train_on_flows(net, list(range(100, 100 + nseeds)), n=n, steps=steps,
                                       log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    net.save(out_path)
    print(json.dumps(dict(trained_pairs=2 * steps * nseeds, seconds=round(time.perf_counter() - t0, 1),
                          first_loss=losses[0], last_loss=losses[-1], weights=out_path)), flush=True)


def run_c4(wl, steps, warmup, seeds=(7, 8, 9)):
    """Learned warm start against the plain method at equal residual on held-out random-ellipsoid
    flows (256x128x128): both runs step the same flow from the same state; every projection of
    both iterates to Σr² < 1e-4."""
    if not os.path.exists(WARMNET):
        return dict(skipped=f"no trained weights at {os.path.relpath(WARMNET, ROOT)}")
    net = wl.WarmNet.load(WARMNET, grid=(128, 64, 64), replay_pairs=1)
    # plain; the guess in both projections or the predictor's only; the mixed-precision solver alone
    # and with the guess (the combination whose target is 1.5x the plain sim_step!/s)
    res = {"plain": [], "learned": [], "learned_predictor": [], "mixed": [], "learned_mixed": []}
    fdiff, vc = [], {k: [] for k in res}
    fwd_ms = None
    for seed in seeds:
        for mode in res:
            sim = wl.learned.random_ellipsoid_simulation(seed)
            for _ in range(warmup):
                wl.sim_step(sim, remeasure=False)
            if mode in ("mixed", "learned_mixed"):
                sim.pois.set_solver(mode=wl.WL_SOLVER_MIXED)
            if mode.startswith("learned"):
                wl.set_warmstart(sim, net, guess=True, predictor_only=(mode == "learned_predictor"))
            n0 = len(sim.pois.n)
            t0 = time.perf_counter()
            for _ in range(steps):
                wl.sim_step(sim, remeasure=False)
            res[mode].append((time.perf_counter() - t0) / steps)
            vc[mode].append(sum(sim.pois.n[n0:]) / steps)
            F = wl.pressure_force(sim)
            if mode == "plain":
                F0 = F
            else:
                fdiff.append(float(np.linalg.norm(F - F0) / max(np.linalg.norm(F0), 1e-30)))
            if mode == "learned" and seed == seeds[0]:
                sim.flow.profile(True)
                for _ in range(5):
                    wl.sim_step(sim, remeasure=False)
                st = {k["name"]: k for k in sim.flow.profile_read()}
                sim.flow.profile(False)
                fw = st.get("nn_forward")
                fwd_ms = fw["total_ms"] / fw["launches"] if fw else None
            wl.set_warmstart(sim, None)
            del sim
    cells = 256 * 128 * 128
    ms = {k: 1e3 * float(np.mean(v)) for k, v in res.items()}
    flops = NET_FLOPS_PER_VOXEL * 128 * 64 * 64
    return dict(workload="held-out random-ellipsoid flows 256x128x128 (seeds %s), Float32, tol Σr²<1e-4" % list(seeds),
                steps=steps, ms_per_step={k: round(v, 3) for k, v in ms.items()},
                mcells={k: round(cells / v / 1e3, 2) for k, v in ms.items()},
                vcycles_per_step={k: round(float(np.mean(v)), 3) for k, v in vc.items()},
                speedup_learned=round(ms["plain"] / ms["learned"], 3),
                speedup_learned_predictor=round(ms["plain"] / ms["learned_predictor"], 3),
                speedup_mixed=round(ms["plain"] / ms["mixed"], 3),
                speedup_learned_mixed=round(ms["plain"] / ms["learned_mixed"], 3),
                pressure_force_rel_diff_max=round(max(fdiff), 6),
                net_forward_ms=None if fwd_ms is None else round(fwd_ms, 4),
                net_roofline=None if not fwd_ms else dict(bound="mfma", achieved=round(flops / fwd_ms / 1e9, 2),
                                                          peak=2500.0, unit="TFLOP/s",
                                                          frac=round(flops / fwd_ms / 1e9 / 2500.0, 5)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the mixed-precision 512x256x256 line")
    ap.add_argument("--c3-steps", type=int, default=50)
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=100)
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--c5-steps", type=int, default=20)
    ap.add_argument("--c2-main", action="store_true", help="time the 256x128x128 case as the main line")
    ap.add_argument("--tiles", help="GS,RP tile targets to set (main line)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile targets on the main line (it tunes them during set-up)")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--c4-steps", type=int, default=30)
    ap.add_argument("--train-warmstart", metavar="OUT", help="train the warm-start network and save its weights")
    ap.add_argument("--train-seeds", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=160)
    ap.add_argument("--copy-calib", action="store_true")
    ap.add_argument("--quiet-kernels", action="store_true")
    ap.add_argument("--summarize", nargs=6, metavar=("STATS", "FETCH", "WRITE", "OUT_MD", "OUT_JSON", "CMD"),
                    help="turn rocprofv3 CSVs into the profile summary and the traffic table ('-' for FETCH and "
                         "WRITE: kernel times only; OUT_JSON is then not written)")
    a = ap.parse_args()
    if a.summarize:
        summarize(*a.summarize)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (the flow is stepped on a single "
              f"MI355X; see the sharding design note)", file=sys.stderr)
        sys.exit(2)
    def stage(msg):  # one line on stderr as each measurement starts (the JSON line goes to stdout)
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    if a.copy_calib:
        copy_calibration()
    wl = load_package()
    wl.load_library()
    if a.train_warmstart:
        train_warmstart(wl, a.train_warmstart, a.train_seeds, a.train_steps)
        return
    n, R, center, Re, L = (512, 256, 256), 32, (128, 128, 128), 1e4, 64
    traffic_file, label = "hbm_traffic_c3.json", "3D flow past a sphere 512x256x256, Re=1e4"
    n2, R2, c2, Re2, L2 = (256, 128, 128), 16, (64, 64, 64), 3700.0, 32
    if a.c2_main:  # the 256x128x128 case as the main line (for profiling it alone)
        n, R, center, Re, L = n2, R2, c2, Re2, L2
        traffic_file, label = "hbm_traffic.json", "3D flow past a sphere 256x128x128, Re=3700"
        a.no_c2 = a.no_c3 = True
    stage(f"main line {'x'.join(map(str, n))}: {a.warmup} warm-up + {a.steps} timed + {a.steps} profiled steps")
    r = run_gpu(wl, n, R, center, Re, L, a.steps, a.warmup, tune=not a.no_autotune and not a.tiles,
                tiles=tuple(int(v) for v in a.tiles.split(",")) if a.tiles else None)
    roof, rows = roofline(r["stats"], os.path.join(ROOT, "profiles", traffic_file))
    out = dict(metric=METRIC, value=round(r["mcells"], 2), unit="Mcell-steps/s", n_gpus=1, steps=a.steps,
               warmup=a.warmup, ms_per_step=round(r["ms_per_step"], 4), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload=label + ", plain MultiLevelPoisson, Float32, tol Σr²<1e-4", grid=list(n),
                           sphere_radius=R, Re=Re,
                           method="plain", remeasure=False),
               sim_steps_per_s=round(r["steps_per_s"], 3), vcycles_per_step=round(r["vcycles_per_step"], 3),
               tiles=dict(autotuned=r["tiles"] is not None and not a.tiles,
                          gs_target=r["tiles"][0] if r["tiles"] else "default",
                          restrict_prolong_target=r["tiles"][1] if r["tiles"] else "default"),
               roofline=roof, profiled_ms_per_step=round(r["profiled_ms_per_step"], 4),
               timing="value from the timed steps with no per-kernel events; per-kernel HIP-event times "
                      "from a second pass of the same number of steps")
    if not a.quiet_kernels:
        out["kernels"] = rows[:14]
    del r
    if not a.no_c3:
        stage("512x256x256 with the mixed-precision solver")
        r3m = run_gpu(wl, n, R, center, Re, L, a.c3_steps, a.warmup, profile_steps=0, mode=wl.WL_SOLVER_MIXED,
                      tune=True)
        out["c3_mixed"] = dict(workload="the same 512x256x256 sphere with FP32 V-cycles inside an FP64 "
                                        "defect-correction loop", steps=a.c3_steps, warmup=a.warmup,
                               mcells=round(r3m["mcells"], 2), ms_per_step=round(r3m["ms_per_step"], 3),
                               vcycles_per_step=round(r3m["vcycles_per_step"], 3),
                               vs_plain=round(r3m["mcells"] / out["value"], 3))
        del r3m
    if not a.no_c2:
        stage("256x128x128 sphere")
        r2 = run_gpu(wl, n2, R2, c2, Re2, L2, a.c2_steps, a.warmup, tune=True)
        roof2, rows2 = roofline(r2["stats"], os.path.join(ROOT, "profiles", "hbm_traffic.json"))
        out["c2"] = dict(workload="3D sphere 256x128x128, r=16, Re=3700, plain FP32 MG", steps=a.c2_steps,
                         warmup=a.warmup, tiles=list(r2["tiles"]),
                         mcells=round(r2["mcells"], 2), ms_per_step=round(r2["ms_per_step"], 3),
                         vcycles_per_step=round(r2["vcycles_per_step"], 3), roofline=roof2,
                         kernels_L0=[k for k in rows2 if k["kernel"].endswith("@L0")][:10])
        del r2
    if not a.no_c4:
        stage("learned warm start and mixed precision on held-out 256x128x128 flows")
        out["c4"] = run_c4(wl, a.c4_steps, 20)
    if not a.no_c5:
        stage("batch of 64 uneven 2D flows")
        out["c5"] = run_c5(wl, a.c5_steps, 3)
    if not a.no_cpu:
        stage("2D circle on the GPU and the oracle; the oracle's CPU rate")
        out["c1"] = run_c1(wl)
        out["cpu_baseline"] = run_cpu(n, R, center, Re, L, label="x".join(str(v) for v in n))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
