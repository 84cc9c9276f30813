"""Benchmark: grid points per second at standard error <= 1e-3 on one MI355X,
for the five workloads of the walk-on-spheres hot path.

Headline (`value`; one "step" = one full solve: classify, walk, merge):
  3D screened Poisson on the unit icosphere at subdivision 4 (2562 vertices,
  5120 triangles), alpha = 10, constant source f = 1 sampled uniformly in the
  ball, g = 0, 32^3 grid over [-1, 1]^3, 4096 walks per inside point,
  epsilon = 1e-4, seed 7, FP32 persistent walker.
`configs` holds every workload, each with its value, a `roofline` object for
its dominant kernel and a `cpu_baseline` (the reference core on the host):
  annulus   2D Laplace annulus (2 x 128-gon), 64^2, 1024 walks, eps 1e-3, seed 12345
  icosphere the headline workload
  stars     16 batched star domains, 256^2 each, Green sampling, adaptive 64..16384 walks
  torus     3D Laplace on a 40 000-triangle torus, 10 000 probes, 2048 walks, adjoint
  plate     screened Poisson on the 200k-triangle plate, 64 x 32 x 16, 8192 walks, gradients

Grid pts/s at SE <= 1e-3 = #inside / sum_p n_p * t_walk,p with
n_p = ceil(var_p / 1e-6) from the FP64 run (shared by every form) and
t_walk,p = (T / total steps) * mean_steps_p of the form being measured.

Roofline.  No stage is a dense contraction, so MFMA does not apply.  VALU
issue: VALU wave64 instructions per walk step (SQ_INSTS_VALU of the same
kernel on the same workload, profiles/walker_counters.json) x the steps of
the live launch / its live duration (HIP events on the launch stream),
against 256 CUs x 2.4 GHz x 1 wave64 VALU instruction per CU per cycle (the
issue rate behind the 78.6 TFLOP/s FP64 / 157.3 TFLOP/s packed-FP32 vector
peaks).  HBM: the algorithmic bytes of the records the query algorithm must
read (from the launch's own query-unit counts) per second against 8 TB/s,
with the counters' traffic (FETCH_SIZE + WRITE_SIZE, calibrated) beside it.
`bound` is whichever the counters show binding: HBM where the traffic's
share of 8 TB/s exceeds the VALU issue share (the torus), else VALU.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       python bench.py --workload annulus|icosphere|stars|torus|plate|welding
       python bench.py --workload X --fp32-only --steps 1 --warmup 0
           (the FP32 walker solves of workload X only, for counter runs)
       python bench.py --copy-calibration 1024   (one 1024 MiB fp32 read+write
           kernel, the scale for the HBM counters)
       python bench.py --grid-sweep   (distance-grid size against build time,
           memory and walker time, a markdown table)
"""
import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "walk-on-spheres_amd"))

METRIC = "grid pts/s at SE≤1e-3, FP32 persistent vs FP64 plain; adjoint walks/s"
WORKLOAD = ("3D screened Poisson, unit icosphere s4 (2562 v / 5120 t), alpha=10, f=1 uniform "
            "ball sampling, g=0, 32^3 grid, N=4096 walks/point, eps=1e-4, seed 7")
HBM_PEAK_GBS = 8000.0
# 256 CUs x 2.4 GHz x one wave64 VALU instruction per CU per cycle
VALU_PEAK_GINST = 256 * 2.4
SE_TARGET = 1e-3
COUNTERS = os.path.join(ROOT, "profiles", "walker_counters.json")
KERNEL_SOURCES = ("wos_engine.hip", "wos_device.hpp", "wos_common.hpp")
# the GPU box allots this many host CPUs per GPU; worker pools stay within it
BOX_CPU_SHARE = 16


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def source_sha():
    """Hash of the kernel sources, to match counter profiles to this build."""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "walk-on-spheres_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def host_cpus():
    """Host cores this process may use: its affinity mask, the cgroup CPU
    quota (cpu.max) and the box's CPU share per GPU; threads = the least."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            if q != "max":
                quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    threads = aff
    if quota is not None:
        threads = min(threads, max(1, int(quota)))
    threads = min(threads, BOX_CPU_SHARE)
    return dict(threads=threads, affinity=aff, total=os.cpu_count(), cgroup_quota=quota)


def pts_per_s_at_se(n_inside, var_ref, msteps, seconds, total_steps):
    n_p = np.maximum(1.0, np.ceil(var_ref / (SE_TARGET ** 2)))
    t_step = seconds / max(1, total_steps)
    return n_inside / float(np.sum(n_p * t_step * msteps))


def copy_calibration(mib):
    """One elementwise fp32 kernel that reads and writes `mib` MiB each: the
    known-size transfer the FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    assert float(dst[-1]) == 2.0
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))


# ---------------------------------------------------------------------------
# roofline
# ---------------------------------------------------------------------------
# Minimum records the FP32 walker's query algorithm reads per query unit
# (bytes): every query its grid cell (GridCellF 16); a list scan each entry
# read (GridEntry 8) and each primitive tested (TriF 48 / SegF 16); a tree
# query its start primitive, each internal node opened (NodeF3 64 / NodeF2
# 48) and each leaf primitive tested; per walk its start copy (StartF 32)
# and its value + steps written (12), plus, when g is not zero, the FP64
# terminal primitive (TriD 128 / SegD 32) and its slot -> primitive id (4).
RECORDS = {3: dict(cell=16, entry=8, prim=48, node=64, walk=32 + 12, terminal=128 + 4),
           2: dict(cell=16, entry=8, prim=16, node=48, walk=32 + 12, terminal=32 + 4)}


def walker_bytes(s, dim, terminal=True):
    r = RECORDS[dim]
    tree_q = s["q_queries"] - s["q_grid_bound"] - s["q_list_scans"]
    return (s["q_queries"] * r["cell"] + s["q_list_entries"] * r["entry"]
            + s["q_list_prims"] * r["prim"]
            + s["q_nodes"] * r["node"] + (s["q_leaf_prims"] + tree_q) * r["prim"]
            + s["total_walks"] * (r["walk"] + (r["terminal"] if terminal else 0)))


def roofline(key, kernel, steps, kernel_s, units=None, dim=3, terminal=False):
    """Roofline of the workload's dominant kernel: VALU issue and HBM, the
    bound the one the counters show binding.  `steps`/`kernel_s`: the live
    launch(es); `units`: the summary of a count_units solve of the same
    problem."""
    out = {"bound": "valu", "unit": "G VALU inst/s", "peak": VALU_PEAK_GINST, "kernel": kernel,
           "kernel_ms": kernel_s * 1e3, "steps": steps, "achieved": None, "frac": None,
           "traffic": None}
    prof = None
    if os.path.exists(COUNTERS):
        with open(COUNTERS) as f:
            prof = json.load(f).get(key)
    if prof:
        per_step = prof["valu_insts_per_step"]
        ach = per_step * steps / kernel_s / 1e9
        out.update(achieved=ach, frac=ach / VALU_PEAK_GINST,
                   valu_insts_per_step=per_step,
                   lane_utilisation=prof["valu_lane_utilisation"],
                   useful_lane_frac=ach / VALU_PEAK_GINST * prof["valu_lane_utilisation"],
                   traffic=prof.get("hbm_bytes_per_step", 0.0) * steps or None,
                   profile={"file": "profiles/walker_counters.json", "workload": key,
                            "source_sha": prof.get("source_sha"),
                            "matches_this_build": prof.get("source_sha") == source_sha(),
                            "profiled_kernel_ms": prof.get("kernel_ms"),
                            "profiled_valu_insts_per_cu_cycle": prof.get("valu_insts_per_cu_cycle")})
    if units is not None:
        b = walker_bytes(units, dim, terminal)
        out["hbm"] = {"algorithmic_bytes": b, "bytes_per_step": b / max(1, steps),
                      "achieved_gbs": b / kernel_s / 1e9,
                      "frac": b / kernel_s / 1e9 / HBM_PEAK_GBS, "peak_gbs": HBM_PEAK_GBS,
                      "units_per_step": {k: units[k] / max(1, steps) for k in
                                         ("q_queries", "q_grid_bound", "q_list_scans",
                                          "q_list_entries", "q_list_prims", "q_nodes",
                                          "q_leaf_prims")}}
    # The bound is the limiter the counters show: HBM when the measured
    # traffic takes a larger share of its peak than the issued VALU
    # instructions take of theirs.  `achieved` then stays the algorithmic
    # bytes (the records the queries must read) and the VALU figures move to
    # `valu`; traffic far above the algorithmic bytes is cache-line waste.
    if out["traffic"] and out["frac"] is not None and "hbm" in out:
        traffic_frac = out["traffic"] / kernel_s / 1e9 / HBM_PEAK_GBS
        out["traffic_frac"] = traffic_frac
        if traffic_frac > out["frac"]:
            valu = {k: out[k] for k in ("unit", "peak", "achieved", "frac", "valu_insts_per_step",
                                        "lane_utilisation", "useful_lane_frac")}
            h = out["hbm"]
            out.update(bound="hbm", unit="GB/s", peak=HBM_PEAK_GBS, achieved=h["achieved_gbs"],
                       frac=h["frac"], valu=valu)
    return out


# ---------------------------------------------------------------------------
# CPU baseline: the reference core (oracle/_ref) on the host cores, timed over
# the reference's window (classification + walks of the sweep; mesh load,
# BVH build and output excluded, runner.hpp:153-155, :240-243)
# ---------------------------------------------------------------------------
def _ref_lib():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    try:
        return pyoracle, pyoracle.Reference(), "reference"
    except FileNotFoundError:
        return pyoracle, pyoracle.Oracle(), "port"


def _cpu_run(kind, rmesh, eq, pts, walks, eps, seed, threads, seed_index):
    t0 = time.perf_counter()
    if kind == "reference":
        r = rmesh.solve_points(eq, pts, walks, eps, seed=seed, n_threads=threads,
                               seed_index=seed_index)
    else:
        r = rmesh.solve_points(eq, pts, walks, eps, seed=seed, rng_mode=0, n_threads=threads,
                               seed_index=seed_index)
    dt = time.perf_counter() - t0
    return dt, r


def cpu_sample(mesh, eq, pts, seed_index, eps, seed, budget_s, walks, min_points=None):
    """Time the reference on a prefix of `pts` (every point at the workload's
    own `walks`), as many points as fit about budget_s (at least
    `min_points`, all of them if they fit); a short pilot sizes it."""
    pyoracle, lib, kind = _ref_lib()
    cpus = host_cpus()
    rmesh = lib.mesh(*mesh)
    n = len(pts)
    pilot = min(n, 4 * cpus["threads"])
    pw = max(1, min(walks, 64))
    dt, _ = _cpu_run(kind, rmesh, eq, pts[:pilot], pw, eps, seed, cpus["threads"],
                     None if seed_index is None else seed_index[:pilot])
    per_point = max(dt, 1e-4) / pilot * walks / pw
    m = int(min(n, max(min_points or cpus["threads"], budget_s / per_point)))
    si = None if seed_index is None else seed_index[:m]
    dt, r = _cpu_run(kind, rmesh, eq, pts[:m], walks, eps, seed, cpus["threads"], si)
    ins = r["location"] == 1
    steps = float(np.sum(r["mean_steps"][ins]) * walks)
    return dict(kind=kind, cpus=cpus, walks=walks, seconds=dt, steps=steps, result=r,
                points=m, inside=int(ins.sum()), t_step=dt / max(steps, 1.0))


def cpu_baseline_grid(mesh, eq, pts_all, inside_idx, var_ref, eps, seed, walks, budget_s=12.0):
    """Grid pts/s at SE <= 1e-3 of the reference on a sample of the inside
    points (spread over the grid: a fixed random order), at the workload's
    walks per point, by the same cost model as the GPU forms."""
    cand = inside_idx[np.random.default_rng(0).permutation(len(inside_idx))]
    s = cpu_sample(mesh, eq, pts_all[cand], cand, eps, seed, budget_s, walks)
    pick = cand[:s["points"]]
    ms = s["result"]["mean_steps"]
    n_p = np.maximum(1.0, np.ceil(var_ref[np.searchsorted(inside_idx, pick)] / SE_TARGET ** 2))
    pts_s = len(pick) / float(np.sum(n_p * s["t_step"] * ms))
    c = s["cpus"]
    return dict(value=pts_s, unit="grid pts/s", cores=c["threads"], kind=s["kind"],
                host_cores_affinity=c["affinity"], host_cores_total=c["total"],
                cgroup_cpu_quota=c["cgroup_quota"],
                steps_per_s=s["steps"] / s["seconds"],
                walks_per_s=len(pick) * s["walks"] / s["seconds"],
                sample=f"{len(pick)} of {len(inside_idx)} inside points (random), "
                       f"{s['walks']} walks each, {s['seconds']:.1f} s on {c['threads']} threads")


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def _grid_points(ctx_summary, grid, dim):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    return pyoracle.grid_points(ctx_summary["grid_bounds"], *grid, dim)


def _fp32_vs_fp64(r32, r64, ins, floor=1e-300):
    z = np.abs(r32["mean"][ins] - r64["mean"][ins]) / np.maximum(
        np.sqrt(r32["standard_error"][ins] ** 2 + r64["standard_error"][ins] ** 2), floor)
    d = np.abs(r32["mean"][ins] - r64["mean"][ins])
    return {"max_abs_diff": float(d.max()), "mean_abs_diff": float(d.mean()),
            "max_z": float(z.max()), "mean_z2": float(np.mean(z ** 2)),
            "mean_steps_ratio": float(np.mean(r32["mean_steps"][ins]) /
                                      np.mean(r64["mean_steps"][ins]))}


def grid_workload(key, Context, capi, mesh, eq, grid, nw, eps, seed, steps, warmup, kernel,
                  want_cpu, terminal=False, fp64=True, graph=False, floor=1e-300, tree=False):
    """FP32 persistent and FP64 plain grid solves of one problem: grid pts/s
    at SE <= 1e-3 of each, the roofline of the FP32 walker, the host CPU
    baseline; returns (entry, FP32 result).  tree: also the FP32 walker
    without the distance grid (every query searches the BVH, as the FP64
    walker does), which splits FP32 / FP64 into the precision-and-launch
    part and the distance-grid part."""
    V, F, B = mesh
    dim = F.shape[1]
    ctx = Context(0)
    ctx.set_mesh(V, F, B)
    p32 = Context.params(nw, eps, seed, precision=capi.WOS_FP32_PERSISTENT)
    p64 = Context.params(nw, eps, seed, precision=capi.WOS_FP64_PLAIN)
    t0 = time.perf_counter()
    first = ctx.solve_grid(*grid, eq, p32)
    t_first = time.perf_counter() - t0
    for _ in range(max(0, warmup - 1)):
        ctx.solve_grid(*grid, eq, p32)
    walk_k = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r32 = ctx.solve_grid(*grid, eq, p32)
        walk_k.append(r32["summary"]["walk_kernel_seconds"])
    t32 = (time.perf_counter() - t0) / steps
    s32 = r32["summary"]
    entry = {"workload": key, "grid": list(grid), "walks_per_point": nw, "epsilon": eps,
             "seed": seed, "fp32_ms_per_solve": t32 * 1e3,
             "setup": {"first_solve_ms": t_first * 1e3,
                       "grid_build_ms": first["summary"]["grid_build_seconds"] * 1e3,
                       "grid_bytes": first["summary"]["grid_bytes"],
                       "grid_cells_setting": ctx.get_tuning()["grid_cells"]},
             "steps_per_s": s32["total_steps"] / t32, "walks_per_s": s32["total_walks"] / t32,
             "failed_walks": s32["failed_walks"]}
    if graph:
        pg = Context.params(nw, eps, seed, precision=capi.WOS_FP32_PERSISTENT, use_graph=1)
        rg = ctx.solve_grid(*grid, eq, pg)
        t0 = time.perf_counter()
        for _ in range(steps):
            rg = ctx.solve_grid(*grid, eq, pg)
        entry["fp32_graph"] = {"ms_per_solve": (time.perf_counter() - t0) / steps * 1e3,
                               "results_equal_eager": bool(np.array_equal(
                                   rg["mean"], r32["mean"], equal_nan=True))}
    pc = Context.params(nw, eps, seed, precision=capi.WOS_FP32_PERSISTENT, count_units=1)
    sc = ctx.solve_grid(*grid, eq, pc)["summary"]
    assert sc["total_steps"] == s32["total_steps"]
    kern_s = float(np.mean(walk_k))
    entry["roofline"] = roofline(key, kernel, s32["total_steps"], kern_s, sc, dim, terminal)
    ins = r32["location"] == 1
    if fp64:
        t0 = time.perf_counter()
        r64 = ctx.solve_grid(*grid, eq, p64)
        t64 = time.perf_counter() - t0
        s64 = r64["summary"]
        assert np.array_equal(r64["location"], r32["location"])
        var_ref = r64["variance"][ins]
        v32 = pts_per_s_at_se(ins.sum(), var_ref, r32["mean_steps"][ins], t32, s32["total_steps"])
        v64 = pts_per_s_at_se(ins.sum(), var_ref, r64["mean_steps"][ins], t64, s64["total_steps"])
        if "fp32_graph" in entry:
            # the same solves replayed from a captured hipGraph (use_graph)
            entry["fp32_graph"]["value"] = pts_per_s_at_se(
                ins.sum(), var_ref, r32["mean_steps"][ins],
                entry["fp32_graph"]["ms_per_solve"] * 1e-3, s32["total_steps"])
        entry.update(value=v32, unit="grid pts/s", inside=int(ins.sum()),
                     fp64_plain={"value": v64, "ms_per_solve": t64 * 1e3,
                                 "walk_kernel_ms": s64["walk_kernel_seconds"] * 1e3,
                                 "steps_per_s": s64["total_steps"] / t64},
                     fp32_over_fp64=v32 / v64,
                     fp32_vs_fp64=_fp32_vs_fp64(r32, r64, ins, floor))
        if tree:
            ctx.set_tuning(grid_cells=0)
            ctx.solve_grid(*grid, eq, p32)
            t0 = time.perf_counter()
            rt = ctx.solve_grid(*grid, eq, p32)
            tt = time.perf_counter() - t0
            ctx.set_tuning()
            vt = pts_per_s_at_se(ins.sum(), var_ref, rt["mean_steps"][ins], tt,
                                 rt["summary"]["total_steps"])
            entry["fp32_tree_only"] = {
                "value": vt, "ms_per_solve": tt * 1e3,
                "walk_kernel_ms": rt["summary"]["walk_kernel_seconds"] * 1e3,
                "over_fp64_plain": vt / v64, "grid_speedup": v32 / vt,
                "fp32_vs_fp64": _fp32_vs_fp64(rt, r64, ins, floor)}
        if want_cpu:
            inside_idx = np.nonzero(ins)[0]
            pts_all = _grid_points(s32, grid, dim)
            entry["cpu_baseline"] = cpu_baseline_grid((V, F, B), eq, pts_all, inside_idx, var_ref,
                                                      eps, seed, nw)
            entry["vs_cpu"] = v32 / entry["cpu_baseline"]["value"]
    ctx.close()
    return entry, r32


def workload_annulus(Context, meshgen, problems, capi, args, want_cpu):
    """2D Laplace on the annulus r_in = 0.25, r_out = 1 (two 128-gons, inner
    loop first and clockwise), g = 1 inside / 0 outside (u = ln r / ln 0.25),
    64^2 grid over [-1, 1]^2, 1024 walks, eps 1e-3, seed 12345: FP32 and FP64
    grid pts/s at SE <= 1e-3"""
    e, r32 = grid_workload("annulus", Context, capi, meshgen.annulus(128), problems.annulus_laplace(),
                           (64, 64, 1), 1024, 1e-3, 12345, args.steps, args.warmup,
                           "k_walk_fp32<2,0,false,false,2,false>", want_cpu, terminal=True,
                           graph=True)
    ins = r32["location"] == 1
    xs = np.linspace(-1, 1, 64)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    rr = np.sqrt(X ** 2 + Y ** 2).ravel()[ins]
    e["max_err_vs_exact"] = float(np.max(np.abs(r32["mean"][ins] - np.log(rr) / np.log(0.25))))
    e["description"] = workload_annulus.__doc__.split(":")[0].strip()
    return e


def workload_icosphere(Context, meshgen, problems, capi, args, want_cpu):
    """3D screened Poisson on the unit icosphere s4 (5120 triangles), alpha 10,
    f = 1 uniform-ball sampling, g = 0, 32^3 grid, 4096 walks, eps 1e-4, seed 7"""
    e, _ = grid_workload("icosphere", Context, capi, meshgen.icosphere(4),
                         problems.sphere_screened_constant(alpha=10.0, f=1.0), (32, 32, 32),
                         args.walks, 1e-4, 7, args.steps, args.warmup,
                         "k_walk_fp32<3,1,true,false,2,false>", want_cpu, graph=True, tree=True)
    e["description"] = workload_icosphere.__doc__.strip()
    return e


def workload_plate(Context, meshgen, problems, capi, args, want_cpu):
    """Welding-style stationary screened Poisson: box plate 120 x 60 x 10 mm,
    200 316 triangles, alpha = 0.05, f = Q exp(-|x - x0|^2 / sigma^2) with
    Q = 1, sigma = 3 mm, x0 = (40, 30, 10), uniform ball sampling, g = 0,
    64 x 32 x 16 grid, 8192 walks per point, eps 0.01, seed 5: FP32 and FP64
    grid pts/s at SE <= 1e-3 and the FP32 adjoint of L = sum u (gradients
    with respect to Q, sigma, x0, alpha)"""
    mesh = meshgen.plate_200k()
    eq = problems.plate_gaussian(alpha=0.05, q=1.0, sigma=3.0, x0=(40.0, 30.0, 10.0))
    grid = (64, 32, 16)
    # a floor for points where the source underflows FP32 (f ~ 1e-170 far
    # from x0: both estimates are 0 to 1e-160 and their SEs vanish)
    e, r32 = grid_workload("plate", Context, capi, mesh, eq, grid, 8192, 0.01, 5,
                           max(1, args.steps // 4), min(args.warmup, 1),
                           "k_walk_fp32<3,1,true,false,2,true>", want_cpu, floor=1e-12)
    ctx = Context(0)
    ctx.set_mesh(*mesh)
    p32 = Context.params(8192, 0.01, 5, precision=capi.WOS_FP32_PERSISTENT)
    ctx.solve_grid(*grid, eq, p32)
    ins = r32["location"] == 1
    dl = np.where(ins, 1.0, 0.0)
    t0 = time.perf_counter()
    _, _, gp = ctx.adjoint_grid(*grid, eq, p32, dl)
    tadj = time.perf_counter() - t0
    e.update(adjoint_ms=tadj * 1e3, adjoint_over_forward=tadj * 1e3 / e["fp32_ms_per_solve"],
             grad={"Q": gp[capi.WOS_GRAD_SRC_VALUE], "sigma": gp[capi.WOS_GRAD_SRC_RADIUS],
                   "x0": list(gp[capi.WOS_GRAD_SRC_CENTER:capi.WOS_GRAD_SRC_CENTER + 3]),
                   "alpha": gp[capi.WOS_GRAD_ALPHA]},
             description=workload_plate.__doc__.split(":")[0].strip())
    ctx.close()
    return e


def torus_probes(n, seed):
    """n points uniform inside the torus R = 1, r = 0.35 (rejection in the box)."""
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(o) for o in out) < n:
        p = rng.uniform([-1.35, -1.35, -0.35], [1.35, 1.35, 0.35], size=(4 * n, 3))
        q = (np.hypot(p[:, 0], p[:, 1]) - 1.0) ** 2 + p[:, 2] ** 2 < 0.35 ** 2
        out.append(p[q])
    return np.concatenate(out)[:n]


def workload_torus(Context, meshgen, problems, capi, args, want_cpu, walks=2048):
    """Differentiable 3D Laplace on the torus R = 1, r = 0.35 (200 x 100 quads,
    40 000 triangles), per-vertex values v ~ N(0, 1) (seed 3), 10 000 probes,
    2048 walks, eps 1e-3, loss 0.5 sum (u - t)^2 with t from v + 0.1 N(0, 1):
    forward solve and adjoint replay (dLoss/dv) at equal walks per point"""
    V, F, _ = meshgen.torus(1.0, 0.35, 200, 100)
    rng = np.random.default_rng(3)
    v = rng.normal(size=len(V))
    vt = v + 0.1 * rng.normal(size=len(V))
    pts = torus_probes(10000, 3)
    ctx = Context(0)
    ctx.set_mesh(V, F)
    p = Context.params(walks, 1e-3, 3, precision=capi.WOS_FP32_PERSISTENT)
    t0 = time.perf_counter()
    target = ctx.solve_points(pts, problems.laplace_vertex_values(vt), p)["mean"]
    t_first = time.perf_counter() - t0
    eq = problems.laplace_vertex_values(v)
    ctx.solve_points(pts, eq, p)  # warm-up
    fwd_t, adj_t = [], []
    for _ in range(max(1, args.steps // 2)):
        t0 = time.perf_counter()
        fwd = ctx.solve_points(pts, eq, p)
        fwd_t.append(time.perf_counter() - t0)
        dl = np.nan_to_num(fwd["mean"] - target)
        t0 = time.perf_counter()
        out, gv, _ = ctx.adjoint_points(pts, eq, p, dl)
        adj_t.append(time.perf_counter() - t0)
    t_fwd, t_adj = float(np.median(fwd_t)), float(np.median(adj_t))
    n_walks = out["summary"]["total_walks"]
    sa = out["summary"]
    sf = fwd["summary"]
    units = ctx.solve_points(pts, eq, Context.params(walks, 1e-3, 3, precision=capi.WOS_FP32_PERSISTENT,
                                                     count_units=1))["summary"]
    e = dict(workload="torus", description=workload_torus.__doc__.split(":")[0].strip(),
             metric="adjoint walks/s (walk replay); adjoint time / forward time",
             value=n_walks / t_adj, unit="replayed walks/s",
             loss=float(0.5 * np.sum(dl ** 2)), probes=len(pts), walks_per_point=walks,
             forward_ms=t_fwd * 1e3, adjoint_ms=t_adj * 1e3,
             adjoint_over_forward=t_adj / t_fwd,
             forward_walks_per_s=sf["total_walks"] / t_fwd,
             adjoint_kernel_walks_per_s=n_walks / sa["walk_kernel_seconds"],
             grad_norm=float(np.linalg.norm(gv)),
             setup={"first_solve_ms": t_first * 1e3},
             roofline=roofline("torus_adjoint", "k_walk_fp32<3,0,false,true,2,false>", sa["total_steps"],
                               sa["walk_kernel_seconds"], units, 3, True),
             forward_roofline=roofline("torus_forward", "k_walk_fp32<3,0,false,false,2,false>",
                                       sf["total_steps"], sf["walk_kernel_seconds"], units, 3,
                                       True))
    # FP64 forward at the same walks
    p64 = Context.params(walks, 1e-3, 3, precision=capi.WOS_FP64_PLAIN)
    t0 = time.perf_counter()
    f64 = ctx.solve_points(pts, eq, p64)
    e["fp64_forward_walks_per_s"] = f64["summary"]["total_walks"] / (time.perf_counter() - t0)
    if want_cpu:
        s = cpu_sample((V, F, None), eq, pts, None, 1e-3, 3, 10.0, walks)
        c = s["cpus"]
        e["cpu_baseline"] = dict(
            value=s["inside"] * s["walks"] / s["seconds"], unit="forward walks/s",
            cores=c["threads"], kind=s["kind"], host_cores_affinity=c["affinity"],
            host_cores_total=c["total"], cgroup_cpu_quota=c["cgroup_quota"],
            steps_per_s=s["steps"] / s["seconds"],
            note="the reference has neither an adjoint nor per-vertex boundary data: its "
                 "forward walks/s on the same mesh, probes and walks (g = 0 at the exit)",
            sample=f"{s['points']} of 10000 probes ({s['inside']} inside), {s['walks']} walks each, "
                   f"{s['seconds']:.1f} s on {c['threads']} threads")
        e["vs_cpu_forward"] = e["forward_walks_per_s"] / e["cpu_baseline"]["value"]
        e["vs_cpu_adjoint"] = e["value"] / e["cpu_baseline"]["value"]
    ctx.close()
    return e


def workload_stars(Context, meshgen, problems, capi, args, want_cpu):
    """Batched uneven 2D Poisson, Green-function source sampling: 16 star
    domains r = 1 + a cos(k theta) (a ~ U(0.1, 0.4), k ~ U{3..9}, 512..16384
    segments, meshgen seed 2024), f = exp(-|x|^2/0.1), g = sin 3x cos 2y,
    256^2 grid each, walks per point adaptive from 64 to 16384 until
    SE <= 1e-3, FP32, eps 1e-3, seed 11: every round is one walker launch over
    all domains.  Points reaching SE <= 1e-3 per second, capped points apart"""
    fam = meshgen.star_family(16, seed=2024, n_min=512, n_max=16384)
    ctx = Context(0)
    ctx.set_domains([f["mesh"] for f in fam])
    eq = problems.star_poisson_gaussian(1)
    p = Context.params(64, 1e-3, 11, precision=capi.WOS_FP32_PERSISTENT, target_se=SE_TARGET,
                       max_walks=16384)
    t0 = time.perf_counter()
    _, first = ctx.solve_grid_batch(16, 16, 1, eq, Context.params(64, 1e-3, 11,
                                                                  precision=capi.WOS_FP32_PERSISTENT))
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    res, tot = ctx.solve_grid_batch(256, 256, 1, eq, p)
    dt = time.perf_counter() - t0
    inside = reached = capped = 0
    per = []
    for f, r in zip(fam, res):
        ins = r["location"] == 1
        ok = ins & (r["standard_error"] <= SE_TARGET)
        cap = ins & ~ok
        inside += int(ins.sum()); reached += int(ok.sum()); capped += int(cap.sum())
        per.append(dict(segments=f["n"], a=round(f["a"], 4), k=f["k"], inside=int(ins.sum()),
                        reached=int(ok.sum()), capped=int(cap.sum()),
                        mean_walks=float(r["walks"][ins].mean()),
                        total_walks=int(r["walks"][ins].sum())))
    pc = Context.params(64, 1e-3, 11, precision=capi.WOS_FP32_PERSISTENT, target_se=SE_TARGET,
                        max_walks=16384, count_units=1)
    _, units = ctx.solve_grid_batch(256, 256, 1, eq, pc)
    e = dict(workload="stars", description=workload_stars.__doc__.split(":")[0].strip(),
             metric="points reaching SE<=1e-3 per s (FP32, batched, adaptive)",
             value=reached / dt, unit="pts/s", seconds=dt, inside=inside, reached=reached,
             capped=capped, rounds=tot["rounds"], total_walks=tot["total_walks"],
             total_steps=tot["total_steps"], steps_per_s=tot["total_steps"] / dt,
             walk_kernel_s=tot["walk_kernel_seconds"], failed_walks=tot["failed_walks"],
             setup={"grid_build_ms": first["grid_build_seconds"] * 1e3,
                    "grid_bytes": first["grid_bytes"], "first_solve_ms": t_setup * 1e3},
             roofline=roofline("stars", "k_walk_fp32<2,2,false,false,2,false>", tot["total_steps"],
                               tot["walk_kernel_seconds"], units, 2, True),
             domains=per)
    if want_cpu:
        # the reference, one domain after the other, on random inside points
        # of each at 1024 walks: its time per walk step there, times the
        # walks the adaptive run took per domain and the reference's own
        # mean steps
        est, walks_used, secs, npts = 0.0, None, 0.0, 0
        c = None
        for f, r, d in zip(fam, res, per):
            ins = np.nonzero(r["location"] == 1)[0]
            pick = ins[np.random.default_rng(0).permutation(len(ins))]
            pts = _grid_points(r["summary"], (256, 256, 1), 2)[pick]
            s = cpu_sample(f["mesh"], eq, pts, pick, 1e-3, 11, 0.75, 1024)
            c = s["cpus"]
            ms = float(np.mean(s["result"]["mean_steps"][s["result"]["location"] == 1]))
            est += s["t_step"] * ms * d["total_walks"]
            secs += s["seconds"]
            walks_used = s["walks"]
            npts += s["points"]
        e["cpu_baseline"] = dict(
            value=reached / est, unit="pts/s", cores=c["threads"], kind="reference",
            host_cores_affinity=c["affinity"], host_cores_total=c["total"],
            cgroup_cpu_quota=c["cgroup_quota"], estimated_seconds=est,
            sample=f"{npts} inside points over the 16 domains (random), {walks_used} walks "
                   f"each, {secs:.1f} s on {c['threads']} threads; time per step x the "
                   f"adaptive run's walks per domain x the reference's mean steps")
        e["vs_cpu"] = e["value"] / e["cpu_baseline"]["value"]
    ctx.close()
    return e


def workload_welding(Context, capi):
    """Welding transients, one screened WoS solve per implicit-Euler step:
    level 2 (plate 120 x 80 mm, 49 x 33 grid, 500 walks, 100 steps of 0.1 s,
    cubic history, stationary Gaussian for 3 s, seed 20260819) in FP32 and
    FP64, level 3 (97 x 65 grid, 500 walks, 150 steps, source moving for 5 s)
    in FP32; the compiled reference on the host cores times step 21 of level 2
    on a sample of its points (the reference's own seeds), scaled to the run:
    whole-run seconds"""
    from wos_amd import welding as W
    ctx = Context(0)
    out = {}
    for label, prec in (("fp32", capi.WOS_FP32_PERSISTENT), ("fp64", capi.WOS_FP64_PLAIN)):
        W.run_level2(ctx, W.Level2Config(), prec, steps=2)   # warm-up
        t0 = time.perf_counter()
        rec, _ = W.run_level2(ctx, W.Level2Config(), prec)
        out[f"level2_{label}_s"] = time.perf_counter() - t0
        out[f"level2_{label}_ms_per_step"] = 1e3 * out[f"level2_{label}_s"] / len(rec)
        out[f"level2_{label}_peak_rise"] = max(r["peak_rise"] for r in rec)
    t0 = time.perf_counter()
    rec3, _, _ = W.run_level3(ctx, W.Level3Config(), capi.WOS_FP32_PERSISTENT)
    out["level3_fp32_s"] = time.perf_counter() - t0
    out["level3_fp32_ms_per_step"] = 1e3 * out["level3_fp32_s"] / len(rec3)
    out["level3_fp32_peak_rise"] = max(r["peak_rise"] for r in rec3)
    c = W.Level2Config()
    _, prev = W.run_level2(ctx, c, capi.WOS_FP32_PERSISTENT, steps=20)
    cpu = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        from wos_amd import meshgen
        ref = pyoracle.Reference()
        x0, x1, y0, y1 = -0.06, 0.06, -0.04, 0.04
        kappa = c.conductivity / (c.density * c.specific_heat)
        alpha = 1.0 / np.sqrt(kappa * c.dt)
        ix = W.gaussian_integral(x0, x1, 0.0, c.sigma_x)
        iy = W.gaussian_integral(y0, y1, 0.0, c.sigma_y)
        term = (0.0, 0.0, c.sigma_x, c.sigma_y, c.electrical_power * c.efficiency,
                c.thickness * ix * iy)
        frac = W.heating_fraction(2.0, 2.1, c.heat_on_time)
        fs = W.field_source(prev, (c.nx, c.ny, x0, x1, y0, y1), W.WOS_INTERP_CUBIC, True,
                            alpha * alpha, frac, c.conductivity, [term])
        eq = W.field_equation(alpha, fs)
        idx = np.array([i * c.ny + j for i in range(1, c.nx - 1) for j in range(1, c.ny - 1)])
        pick = idx[::4]
        pts = np.array([(W.grid_coordinate(x0, x1, k // c.ny, c.nx),
                         W.grid_coordinate(y0, y1, k % c.ny, c.ny)) for k in pick])
        cpus = host_cpus()
        rm = ref.mesh(*meshgen.rectangle(x0, x1, y0, y1))
        t0 = time.perf_counter()
        rm.solve_points(eq, pts, c.walks, c.epsilon, seed=W.step_seed(c.seed, 21),
                        n_threads=cpus["threads"], seed_index=pick)
        dt = time.perf_counter() - t0
        run_s = dt * len(idx) / len(pick) * 100
        cpu = dict(value=run_s, unit="s (level-2 run, estimated)", cores=cpus["threads"],
                   kind="reference", host_cores_affinity=cpus["affinity"],
                   host_cores_total=cpus["total"],
                   sample=f"step 21 of level 2, {len(pick)} of {len(idx)} interior points, "
                          f"{c.walks} walks each, {dt:.2f} s on {cpus['threads']} threads, "
                          f"x 100 steps")
        out["level2_fp32_speedup_vs_cpu"] = run_s / out["level2_fp32_s"]
    except (FileNotFoundError, OSError):
        pass
    return dict(workload="welding", description=workload_welding.__doc__.split(":")[0].strip(),
                metric="whole-run seconds of the level-2 and level-3 transients",
                value=out["level2_fp32_s"], unit="s", higher_is_better=False, **out,
                cpu_baseline=cpu)


def fp32_only(name, Context, meshgen, problems, capi, args):
    """The FP32 walker solves of one workload (counter runs): prints the
    walker's steps and kernel time for the profile summary."""
    if name == "icosphere":
        mesh, eq, grid, nw, eps, seed = (meshgen.icosphere(4), problems.sphere_screened_constant(),
                                         (32, 32, 32), args.walks, 1e-4, 7)
    elif name == "annulus":
        mesh, eq, grid, nw, eps, seed = (meshgen.annulus(128), problems.annulus_laplace(),
                                         (64, 64, 1), 1024, 1e-3, 12345)
    elif name == "plate":
        mesh, eq, grid, nw, eps, seed = (meshgen.plate_200k(),
                                         problems.plate_gaussian(0.05, 1.0, 3.0, (40.0, 30.0, 10.0)),
                                         (64, 32, 16), 8192, 0.01, 5)
    else:
        mesh = None
    out = {"workload": name, "source_sha": source_sha()}
    ctx = Context(0)
    if mesh is not None:
        ctx.set_mesh(*mesh)
        p = Context.params(nw, eps, seed, precision=capi.WOS_FP32_PERSISTENT)
        if args.warmup:
            ctx.solve_grid(*grid, eq, p)
        ks, st = [], 0
        for _ in range(args.steps):
            s = ctx.solve_grid(*grid, eq, p)["summary"]
            ks.append(s["walk_kernel_seconds"])
            st += s["total_steps"]
        out.update(kernel_ms=1e3 * float(np.mean(ks)), total_steps=st, launches=args.steps)
    elif name == "stars":
        fam = meshgen.star_family(16, seed=2024, n_min=512, n_max=16384)
        ctx.set_domains([f["mesh"] for f in fam])
        p = Context.params(64, 1e-3, 11, precision=capi.WOS_FP32_PERSISTENT, target_se=SE_TARGET,
                           max_walks=16384)
        _, tot = ctx.solve_grid_batch(256, 256, 1, problems.star_poisson_gaussian(1), p)
        out.update(kernel_ms=1e3 * tot["walk_kernel_seconds"], total_steps=tot["total_steps"],
                   launches=tot["rounds"])
    elif name in ("torus_forward", "torus_adjoint"):
        V, F, _ = meshgen.torus(1.0, 0.35, 200, 100)
        rng = np.random.default_rng(3)
        v = rng.normal(size=len(V))
        pts = torus_probes(10000, 3)
        ctx.set_mesh(V, F)
        p = Context.params(2048, 1e-3, 3, precision=capi.WOS_FP32_PERSISTENT)
        eq = problems.laplace_vertex_values(v)
        if name == "torus_forward":
            s = ctx.solve_points(pts, eq, p)["summary"]
        else:
            s = ctx.adjoint_points(pts, eq, p, np.ones(len(pts)))[0]["summary"]
        out.update(kernel_ms=1e3 * s["walk_kernel_seconds"], total_steps=s["total_steps"],
                   launches=1)
    else:
        raise SystemExit(f"--fp32-only: unknown workload {name}")
    ctx.close()
    print(json.dumps(out))


WORKLOADS = ("annulus", "icosphere", "stars", "torus", "plate")


def grid_sweep(Context, meshgen, problems, capi):
    """Distance-grid size against build time, memory and walker time: for each
    workload and grid_cells setting (-1: the default sized from the primitive
    count) a fresh context, one first solve (the grid build included) and the
    best of two more solves.  FP32 walker, default relax; the stars with a
    fixed 1024 walks.  Prints a markdown table."""
    def setup(name, ctx):
        if name in ("icosphere", "plate", "annulus"):
            if name == "icosphere":
                mesh, eq, grid, nw, eps, seed = (meshgen.icosphere(4), problems.sphere_screened_constant(),
                                                 (32, 32, 32), 4096, 1e-4, 7)
            elif name == "plate":
                mesh, eq, grid, nw, eps, seed = (
                    meshgen.plate_200k(), problems.plate_gaussian(0.05, 1.0, 3.0, (40.0, 30.0, 10.0)),
                    (64, 32, 16), 8192, 0.01, 5)
            else:
                mesh, eq, grid, nw, eps, seed = (meshgen.annulus(128), problems.annulus_laplace(),
                                                 (64, 64, 1), 1024, 1e-3, 12345)
            ctx.set_mesh(*mesh)
            p = Context.params(nw, eps, seed, precision=capi.WOS_FP32_PERSISTENT)
            return lambda: ctx.solve_grid(*grid, eq, p)["summary"]
        if name == "torus":
            V, F, _ = meshgen.torus(1.0, 0.35, 200, 100)
            ctx.set_mesh(V, F)
            pts = torus_probes(10000, 3)
            eq = problems.laplace_vertex_values(np.random.default_rng(3).normal(size=len(V)))
            p = Context.params(2048, 1e-3, 3, precision=capi.WOS_FP32_PERSISTENT)
            return lambda: ctx.solve_points(pts, eq, p)["summary"]
        fam = meshgen.star_family(16, seed=2024, n_min=512, n_max=16384)
        ctx.set_domains([f["mesh"] for f in fam])
        p = Context.params(1024, 1e-3, 11, precision=capi.WOS_FP32_PERSISTENT)
        return lambda: ctx.solve_grid_batch(256, 256, 1, problems.star_poisson_gaussian(1), p)[1]
    settings = {"icosphere": (-1, 1 << 20, 1 << 22, 1 << 23, 1 << 25),
                "plate": (-1, 1 << 21, 1 << 22, 1 << 23, 1 << 25),
                "torus": (-1, 1 << 21, 1 << 22, 1 << 23),
                "stars": (-1, 1 << 18, 1 << 20, 1 << 22),
                "annulus": (-1, 1 << 16, 1 << 18, 1 << 20)}
    lines = ["| workload | grid_cells | build ms | first solve ms | grid MB | solve ms | walker kernel ms | walk steps |",
             "|---|---|---|---|---|---|---|---|"]
    for name, cells in settings.items():
        for g in cells:
            log(f"grid sweep: {name} {g}")
            ctx = Context(0)
            run = setup(name, ctx)
            ctx.set_tuning(grid_cells=g)
            t0 = time.perf_counter()
            first = run()
            t_first = time.perf_counter() - t0
            best = None
            for _ in range(2):
                t0 = time.perf_counter()
                s = run()
                t = time.perf_counter() - t0
                if best is None or t < best[0]:
                    best = (t, s)
            t, s = best
            label = "default" if g < 0 else f"2^{g.bit_length() - 1}"
            lines.append(f"| {name} | {label} | {first['grid_build_seconds'] * 1e3:.1f} | "
                         f"{t_first * 1e3:.1f} | {first['grid_bytes'] / 2 ** 20:.1f} | {t * 1e3:.2f} | "
                         f"{s['walk_kernel_seconds'] * 1e3:.2f} | {s['total_steps']} |")
            ctx.close()
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--walks", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32-only", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--grid-sweep", action="store_true",
                    help="distance-grid size against build time, memory and walker time "
                         "(markdown table)")
    ap.add_argument("--workload", default="all",
                    choices=("all",) + WORKLOADS + ("welding", "torus_forward", "torus_adjoint"),
                    help="all (default): the headline line with every workload under "
                         "`configs`; one name: that workload's entry alone")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the single-GPU kernel at full shape is the measured unit).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from wos_amd import Context, meshgen, problems
    from wos_amd import capi
    if args.grid_sweep:
        print(f"Kernel sources: `source_sha` {source_sha()}.\n")
        grid_sweep(Context, meshgen, problems, capi)
        return
    if args.fp32_only:
        fp32_only("icosphere" if args.workload == "all" else args.workload, Context, meshgen,
                  problems, capi, args)
        return
    if args.workload == "welding":
        print(json.dumps(workload_welding(Context, capi)))
        return
    fns = {"annulus": workload_annulus, "icosphere": workload_icosphere, "stars": workload_stars,
           "torus": workload_torus, "plate": workload_plate}
    want_cpu = not args.no_cpu
    if args.workload != "all":
        print(json.dumps(fns[args.workload](Context, meshgen, problems, capi, args, want_cpu)))
        return
    configs = {}
    for name in ("icosphere", "annulus", "torus", "stars", "plate"):
        log(f"workload {name}")
        configs[name] = fns[name](Context, meshgen, problems, capi, args, want_cpu)
        log(f"workload {name} done: {configs[name]['value']:.4g} {configs[name]['unit']}")
    ico = configs["icosphere"]
    line = {
        "metric": METRIC,
        "value": ico["value"],
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ico["fp32_ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "inside_points": ico["inside"],
                   "walks_per_point": args.walks,
                   "accumulation": "f64 (Welford, terminal projection, boundary value)",
                   "parallelism": "single GPU"},
        "roofline": ico["roofline"],
        "cpu_baseline": ico.get("cpu_baseline"),
        "vs_cpu": ico.get("vs_cpu"),
        "fp64_plain": ico["fp64_plain"],
        "fp32_over_fp64": ico["fp32_over_fp64"],
        "fp32_tree_only": ico.get("fp32_tree_only"),
        "source_sha": source_sha(),
        "configs": configs,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
